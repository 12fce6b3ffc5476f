"""Networks of LTSs in EXP form (synchronisation vectors over Aldebaran
.aut files) converted to gpf.

The reference's models are produced from EXP networks by a conversion script
(README.md line 22) that is not in the snapshot; this module provides the
conversion for the subset such networks use:

    [hide L1, L2, ... in]
    par using
        "a" * "b" * _ -> "c",
        "i" * _ * _,
        ...
    in
        "p0.aut" || "p1.aut" || "p2.aut"
    end par
    [end hide]

Semantics (one gpf label per synchronisation vector):
  * a vector with two or more participants becomes a synchronising label
    whose single rule is the set of participants; each participant's
    transitions carrying its component label are copied under that label;
  * a vector with one participant, and the internal action ("i" or "tau"),
    make that process's transitions local;
  * a transition whose label no vector covers cannot fire;
  * hiding only changes observable labels and is ignored.
Every process is renumbered so that its initial state is local state 0
(cu:673-676).  Vectors are numbered in file order, so the output is
deterministic.

    python -m gpuexplore_amd.exp2gpf model.exp out/model
"""

from __future__ import annotations

import os
import re
import sys
from typing import Dict, List, Optional, Sequence, Tuple

from gpuexplore_amd.gen import Network

INTERNAL = ("i", "tau")


class ExpError(ValueError):
    pass


def parse_aut(text: str) -> Tuple[int, int, List[Tuple[int, str, int]]]:
    """(initial state, number of states, [(source, label, target)])."""
    lines = [ln.strip() for ln in text.splitlines() if ln.strip()]
    if not lines:
        raise ExpError("empty .aut file")
    m = re.match(r"des\s*\(\s*(\d+)\s*,\s*(\d+)\s*,\s*(\d+)\s*\)$", lines[0])
    if not m:
        raise ExpError(f"bad .aut header: {lines[0]!r}")
    init, ntrans, nstates = (int(g) for g in m.groups())
    trans = []
    pat = re.compile(r'\(\s*(\d+)\s*,\s*(?:"((?:[^"\\]|\\.)*)"|([^,]*?))\s*,\s*(\d+)\s*\)$')
    for ln in lines[1:]:
        t = pat.match(ln)
        if not t:
            raise ExpError(f"bad .aut transition: {ln!r}")
        lab = t.group(2) if t.group(2) is not None else t.group(3).strip()
        src, dst = int(t.group(1)), int(t.group(4))
        if src >= nstates or dst >= nstates:
            raise ExpError(f"state out of range in {ln!r}")
        trans.append((src, lab, dst))
    if len(trans) != ntrans:
        raise ExpError(f".aut header announces {ntrans} transitions, file has {len(trans)}")
    if init >= nstates:
        raise ExpError("initial state out of range")
    return init, nstates, trans


_TOKEN = re.compile(r'\s*(?:(?P<str>"(?:[^"\\]|\\.)*")|(?P<arrow>->)|(?P<par>\|\|)|(?P<sym>[*,_])|(?P<word>[A-Za-z_][A-Za-z0-9_]*))')


def _tokens(text: str) -> List[str]:
    text = re.sub(r"\(\*.*?\*\)", " ", text, flags=re.S)  # comments
    out, pos = [], 0
    while pos < len(text):
        if text[pos:].strip() == "":
            break
        m = _TOKEN.match(text, pos)
        if not m or m.end() == pos:
            raise ExpError(f"unexpected text near {text[pos:pos + 30]!r}")
        out.append(m.group(0).strip())
        pos = m.end()
    return out


def parse_exp(text: str) -> Tuple[List[Tuple[List[Optional[str]], Optional[str]]], List[str]]:
    """([(components, result label)], [.aut file names])."""
    toks = _tokens(text)
    i = 0

    def expect(t: str) -> None:
        nonlocal i
        if i >= len(toks) or toks[i] != t:
            raise ExpError(f"expected {t!r} at token {i}: {toks[i] if i < len(toks) else 'end'}")
        i += 1

    if toks and toks[0] == "hide":
        i = 1
        while i < len(toks) and toks[i] != "in":
            i += 1
        expect("in")
    expect("par")
    expect("using")
    vectors = []
    while True:
        comps: List[Optional[str]] = []
        while True:
            if i >= len(toks):
                raise ExpError("unterminated vector")
            t = toks[i]
            if t == "_":
                comps.append(None)
            elif t.startswith('"'):
                comps.append(t[1:-1])
            else:
                raise ExpError(f"bad vector component {t!r}")
            i += 1
            if i < len(toks) and toks[i] == "*":
                i += 1
                continue
            break
        result = None
        if i < len(toks) and toks[i] == "->":
            i += 1
            if i >= len(toks) or not toks[i].startswith('"'):
                raise ExpError("bad vector result")
            result = toks[i][1:-1]
            i += 1
        vectors.append((comps, result))
        if i < len(toks) and toks[i] == ",":
            i += 1
            continue
        break
    expect("in")
    files = []
    while True:
        if i >= len(toks) or not toks[i].startswith('"'):
            raise ExpError("expected a quoted .aut file name")
        files.append(toks[i][1:-1])
        i += 1
        if i < len(toks) and toks[i] == "||":
            i += 1
            continue
        break
    expect("end")
    expect("par")
    if i < len(toks) and toks[i] == "end":
        i += 1
        expect("hide")
    if i != len(toks):
        raise ExpError(f"trailing text: {toks[i:]}")
    for comps, _ in vectors:
        if len(comps) != len(files):
            raise ExpError(f"vector {comps} has {len(comps)} components for {len(files)} processes")
    return vectors, files


def network_from(vectors: Sequence[Tuple[List[Optional[str]], Optional[str]]],
                 auts: Sequence[Tuple[int, int, List[Tuple[int, str, int]]]]) -> Network:
    """The gpf network of the parsed EXP vectors and .aut processes."""
    n = len(auts)
    if not 1 <= n <= 32:
        raise ExpError("1..32 processes supported by gpf")
    # renumber: the initial state becomes 0 (swapped with state 0)
    ren = []
    for init, ns, _ in auts:
        r = list(range(ns))
        r[0], r[init] = r[init], r[0]
        ren.append(r)
    nstates = [ns for _, ns, _ in auts]
    local: List[List[List[Tuple[int, List[int]]]]] = [[[] for _ in range(ns)] for ns in nstates]
    sync: List[List[Dict[int, List[int]]]] = [[{} for _ in range(ns)] for ns in nstates]
    rules: Dict[int, List[int]] = {}
    seen = set()
    # local moves: the internal action, and single-participant vectors
    local_labels: List[set] = [set(INTERNAL) for _ in range(n)]
    multi = []
    for comps, _ in vectors:
        parts = [p for p, c in enumerate(comps) if c is not None]
        key = tuple(comps)
        if key in seen or not parts:
            continue
        seen.add(key)
        if len(parts) == 1:
            local_labels[parts[0]].add(comps[parts[0]])
        else:
            multi.append((parts, comps))
    for p, (_, _, trans) in enumerate(auts):
        by_src: Dict[int, Dict[str, List[int]]] = {}
        for s, lab, t in trans:
            if lab in local_labels[p]:
                ts = by_src.setdefault(ren[p][s], {}).setdefault(lab, [])
                if ren[p][t] not in ts:
                    ts.append(ren[p][t])
        for s, labs in by_src.items():
            for lab in sorted(labs):
                local[p][s].append((0, labs[lab]))
    for v, (parts, comps) in enumerate(multi):
        lab = v + 1
        mask = 0
        for p in parts:
            mask |= 1 << p
            for s, l, t in auts[p][2]:
                if l == comps[p]:
                    ts = sync[p][ren[p][s]].setdefault(lab, [])
                    if ren[p][t] not in ts:
                        ts.append(ren[p][t])
        rules[lab] = [mask]
    net = Network(nstates, len(multi) + 1, local, sync, rules)
    net.check()
    return net


def convert(exp_path: str) -> Network:
    with open(exp_path) as f:
        vectors, files = parse_exp(f.read())
    base = os.path.dirname(os.path.abspath(exp_path))
    auts = []
    for name in files:
        with open(os.path.join(base, name)) as f:
            auts.append(parse_aut(f.read()))
    return network_from(vectors, auts)


def main(argv=None) -> int:
    argv = sys.argv[1:] if argv is None else argv
    if len(argv) != 2:
        print("usage: python -m gpuexplore_amd.exp2gpf <model.exp> <output base name>", file=sys.stderr)
        return 2
    try:
        net = convert(argv[0])
    except (OSError, ExpError) as e:
        print(f"ERROR: {e}", file=sys.stderr)
        return 1
    net.write(argv[1])
    print(f"{argv[1]}.gpf: {net.nprocs} processes, {net.nlabels - 1} synchronising labels")
    return 0


if __name__ == "__main__":
    sys.exit(main())
