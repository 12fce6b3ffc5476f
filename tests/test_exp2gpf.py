"""EXP networks to gpf: the converted model's state space (oracle and naive
enumerator) equals the product computed directly from the .aut processes and
the synchronisation vectors."""

import os
import random
import subprocess
import sys
from itertools import product

import pytest

from gpuexplore_amd import exp2gpf
from naive_enum import NaiveModel

SENDER = """des (0, 3, 3)
(0, "send", 1)
(1, "i", 2)
(2, "ack", 0)
"""
CHANNEL = """des (1, 4, 3)
(1, "send", 0)
(0, "deliver", 2)
(0, "lose", 1)
(2, i, 1)
"""
RECEIVER = """des (0, 4, 2)
(0, "deliver", 1)
(1, "ack", 0)
(1, "junk", 1)
(1, "ack", 0)
"""
EXP = """(* sender, lossy channel, receiver *)
hide send, deliver in
  par using
    "send" * "send" * _ -> "send",
    _ * "deliver" * "deliver" -> "deliver",
    "ack" * _ * "ack",
    _ * "lose" * _,
    _ * "deliver" * "deliver"
  in
    "sender.aut" || "channel.aut" || "receiver.aut"
  end par
end hide
"""


def direct(vectors, auts):
    """BFS over tuples of the processes' own state ids, straight from the
    definitions: internal and single-participant moves per process, and for
    every distinct multi-participant vector the product of its participants'
    matching moves; duplicate (label, target) pairs count once."""
    n = len(auts)
    moves = []  # per process: {(state, label): set(targets)}
    for init, ns, trans in auts:
        d = {}
        for s, lab, t in trans:
            d.setdefault((s, lab), set()).add(t)
        moves.append(d)
    singles = [set(exp2gpf.INTERNAL) for _ in range(n)]
    multis, seen = [], set()
    for comps, _ in vectors:
        key = tuple(comps)
        parts = [p for p, c in enumerate(comps) if c is not None]
        if key in seen or not parts:
            continue
        seen.add(key)
        if len(parts) == 1:
            singles[parts[0]].add(comps[parts[0]])
        else:
            multis.append((parts, comps))
    s0 = tuple(a[0] for a in auts)
    seen_states, frontier, trans_count, dead = {s0}, [s0], 0, False
    levels = 0
    while frontier:
        nxt = []
        for s in frontier:
            succ = []
            for p in range(n):
                for lab in sorted(singles[p]):
                    for t in sorted(moves[p].get((s[p], lab), ())):
                        succ.append(s[:p] + (t,) + s[p + 1:])
            for parts, comps in multis:
                lists = [sorted(moves[p].get((s[p], comps[p]), ())) for p in parts]
                for combo in product(*lists):
                    t = list(s)
                    for p, v in zip(parts, combo):
                        t[p] = v
                    succ.append(tuple(t))
            if not succ:
                dead = True
            trans_count += len(succ)
            for t in succ:
                if t not in seen_states:
                    seen_states.add(t)
                    nxt.append(t)
        frontier = nxt
        levels += 1
    return len(seen_states), trans_count, levels, dead


def check_network(tmp_path, vectors, auts, oracle_mod):
    net = exp2gpf.network_from(vectors, auts)
    base = str(tmp_path / "conv")
    net.write(base)
    states, trans, levels, dead = direct(vectors, auts)
    o = oracle_mod.explore(base, 0)
    assert (o["states"], o["transitions"], o["levels"]) == (states, trans, levels)
    nv = NaiveModel(open(base + ".gpf").read()).explore("")
    assert (nv["states"], nv["transitions"]) == (states, trans)
    assert oracle_mod.explore(base, 1)["violation"] == dead
    return states, trans


def test_handwritten_network(tmp_path, oracle_mod):
    for name, text in [("sender", SENDER), ("channel", CHANNEL), ("receiver", RECEIVER)]:
        (tmp_path / f"{name}.aut").write_text(text)
    (tmp_path / "net.exp").write_text(EXP)
    vectors, files = exp2gpf.parse_exp(EXP)
    assert files == ["sender.aut", "channel.aut", "receiver.aut"]
    assert len(vectors) == 5 and vectors[0] == (["send", "send", None], "send")
    net = exp2gpf.convert(str(tmp_path / "net.exp"))
    assert net.nprocs == 3 and net.nlabels == 1 + 3  # send, deliver, ack (the duplicate counts once)
    auts = [exp2gpf.parse_aut(t) for t in (SENDER, CHANNEL, RECEIVER)]
    states, trans = check_network(tmp_path, vectors, auts, oracle_mod)
    assert states > 3 and trans > states


def random_network(seed):
    rng = random.Random(seed)
    n = rng.randint(2, 5)
    labels = ["a", "b", "c", "d", "i"]
    auts = []
    for _ in range(n):
        ns = rng.randint(2, 6)
        trans = []
        for s in range(ns):
            for _ in range(rng.randint(0, 3)):
                trans.append((s, rng.choice(labels), rng.randrange(ns)))
        auts.append((rng.randrange(ns), ns, trans))
    vectors = []
    for _ in range(rng.randint(1, 6)):
        k = rng.randint(1, min(3, n))
        parts = rng.sample(range(n), k)
        comps = [rng.choice(labels[:4]) if p in parts else None for p in range(n)]
        vectors.append((comps, None))
    return vectors, auts


@pytest.mark.parametrize("seed", range(25))
def test_random_networks(tmp_path, oracle_mod, seed):
    vectors, auts = random_network(seed)
    check_network(tmp_path, vectors, auts, oracle_mod)


def test_cli_and_errors(tmp_path):
    for name, text in [("sender", SENDER), ("channel", CHANNEL), ("receiver", RECEIVER)]:
        (tmp_path / f"{name}.aut").write_text(text)
    (tmp_path / "net.exp").write_text(EXP)
    out = str(tmp_path / "out" / "net")
    os.makedirs(os.path.dirname(out))
    r = subprocess.run([sys.executable, "-m", "gpuexplore_amd.exp2gpf", str(tmp_path / "net.exp"), out],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert os.path.exists(out + ".gpf")
    (tmp_path / "bad.exp").write_text('par using "a" * "a" in "sender.aut" end par')
    r = subprocess.run([sys.executable, "-m", "gpuexplore_amd.exp2gpf", str(tmp_path / "bad.exp"), out],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "ERROR" in r.stderr
    with pytest.raises(exp2gpf.ExpError):
        exp2gpf.parse_aut("des (0, 2, 1)\n(0, \"a\", 0)\n")
