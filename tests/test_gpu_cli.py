"""The command line's state listings on the GPU: --dump (cu:2174-2185),
-v 3 (cu:2136-2139) and -v 2 (cu:2132-2135), in the reference's format
(cu:544-568, 594-612, 631-651), checked against an independent BFS of the
naive enumerator (own gpf parser, bit-by-bit packing)."""

import os
import re
import subprocess
import sys

import pytest

from conftest import GOLDEN
from gpuexplore_amd import gen
from naive_enum import NaiveModel, pack

pytestmark = pytest.mark.gpu

SUMMARY = re.compile(r"^nr\. of states in hash table: (\d+) \((\d+) unexplored states\)$")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run(gx, *args):
    return subprocess.run([gx.CLI_PATH, *args], capture_output=True, text=True, timeout=120)


def levels(model: NaiveModel, deadlock: bool = False):
    """BFS level sets; with deadlock=True stops after the level in which a
    state without successors is expanded (level-granular early stop)."""
    s0 = tuple([0] * model.n)
    seen, lv, out = {s0}, [s0], [[s0]]
    while lv:
        nxt, dead = [], False
        for s in lv:
            succ, dl, *_ = model.analyse(s)
            dead |= dl
            for t, _ in succ:
                if t not in seen:
                    seen.add(t)
                    nxt.append(t)
        out.append(nxt)
        if deadlock and dead:
            break
        lv = nxt
    return out


def parse_listing(lines, model: NaiveModel, words: int):
    """(set of states, set of unexplored states, summary (count, new)) of one
    listing; checks every line's packed words against the local states."""
    states, fresh = set(), set()
    for ln in lines:
        m = SUMMARY.match(ln)
        if m:
            return states, fresh, (int(m.group(1)), int(m.group(2)))
        new = ln.startswith("new: ")
        body = ln[5:] if new else ln
        locs, _, rest = body.partition(" ")
        s = tuple(int(x) for x in locs.split(","))
        w = [int(x) & 0xFFFFFFFF for x in rest.split()]
        assert len(w) == words and w == pack(model, s, words), (ln, pack(model, s, words))
        assert s not in states
        states.add(s)
        if new:
            fresh.add(s)
    raise AssertionError("listing without a summary line")


def listings(stdout: str):
    """Splits the -v 3 output into listings (each ends with its summary)."""
    cur, out = [], []
    for ln in stdout.splitlines():
        if re.match(r"^(new: )?\d+(,\d+)* ", ln) or SUMMARY.match(ln):
            cur.append(ln)
            if SUMMARY.match(ln):
                out.append(cur)
                cur = []
    return out


@pytest.mark.parametrize("name,flags", [("tiny2", []), ("tiny2", ["-d"]), ("config1", [])])
def test_dump_equals_naive_bfs(gpu, tmp_path, name, flags):
    if name == "tiny2":
        base = os.path.join(GOLDEN, "tiny2")
    else:
        base = str(tmp_path / "config1")
        gen.config1(1).write(base)
    text = open(base + ".gpf").read()
    model = NaiveModel(text)
    words = (model_bits(text) + 31) // 32
    out = str(tmp_path / "dump.txt")
    say("dump", name, flags)
    r = run(gpu, base, *flags, "--dump", out)
    assert r.returncode == 0, r.stderr
    assert "Dumping state space to file..." in r.stdout
    lv = levels(model, deadlock="-d" in flags)
    states, fresh, (count, new) = parse_listing(open(out).read().splitlines(), model, words)
    assert states == set().union(*map(set, lv))
    assert count == len(states) and new == len(fresh)
    # unexplored at the end: nothing after a complete run, the last level's
    # states after an early stop
    assert fresh == (set(lv[-1]) if "-d" in flags else set())


def model_bits(text: str) -> int:
    vals = [int(ln.split()[0]) for ln in text.strip().splitlines()]
    return vals[4]


def test_v3_listing_per_level(gpu):
    base = os.path.join(GOLDEN, "tiny2")
    text = open(base + ".gpf").read()
    model = NaiveModel(text)
    words = (model_bits(text) + 31) // 32
    say("-v 3 tiny2")
    r = run(gpu, base, "-v", "3")
    assert r.returncode == 0, r.stderr
    lv = levels(model)
    got = listings(r.stdout)
    # one listing per expanded level; after level i: levels 0..i+1 stored,
    # level i+1 unexplored
    assert len(got) == len(lv) - 1
    for i, lines in enumerate(got):
        states, fresh, (count, new) = parse_listing(lines, model, words)
        assert states == set().union(*map(set, lv[: i + 2]))
        assert fresh == set(lv[i + 1]) and new == len(fresh) and count == len(states)


def test_v2_counts_per_level(gpu, tmp_path):
    base = str(tmp_path / "config1")
    gen.config1(2).write(base)
    model = NaiveModel(open(base + ".gpf").read())
    say("-v 2 config1")
    r = run(gpu, base, "-v", "2")
    assert r.returncode == 0, r.stderr
    lv = levels(model)
    got = [SUMMARY.match(ln) for ln in r.stdout.splitlines()]
    got = [(int(m.group(1)), int(m.group(2))) for m in got if m]
    want = [(sum(len(x) for x in lv[: i + 2]), len(lv[i + 1])) for i in range(len(lv) - 1)]
    assert got == want


def test_dump_with_proviso_stamps(gpu, tmp_path, oracle_mod):
    """slots carry a level stamp under the cycle proviso: the listing shows
    states only"""
    base = str(tmp_path / "m")
    gen.config1(3).write(base)
    model = NaiveModel(open(base + ".gpf").read())
    words = (model_bits(open(base + ".gpf").read()) + 1 + 31) // 32  # +1 bit under POR (cu:1912)
    out = str(tmp_path / "dump.txt")
    r = run(gpu, base, "--por", "--cycle-proviso", "--dump", out)
    assert r.returncode == 0, r.stderr
    o = oracle_mod.explore(base, 0, por=True, proviso=True)
    states, fresh, (count, new) = parse_listing(open(out).read().splitlines(), model, words)
    full = set().union(*map(set, levels(model)))
    assert count == o["states"] == len(states) and states <= full and new == 0


def test_dump_unwritable_path(gpu, tmp_path):
    base = os.path.join(GOLDEN, "tiny2")
    r = run(gpu, base, "--dump", str(tmp_path / "no" / "such" / "dir" / "f"))
    assert r.returncode == 0
    assert "Could not open file to dump the state space" in r.stderr


def final_count(stdout: str) -> int:
    for ln in reversed(stdout.splitlines()):
        m = re.match(r"^nr\. of states in hash table: (\d+)$", ln)
        if m:
            return int(m.group(1))
    raise AssertionError("no count line")


@pytest.mark.parametrize("por", [False, True])
def test_liveness_model_with_p(gpu, tmp_path, oracle_mod, por):
    """gpf line 1 = 1 with -p: the property becomes LIVENESS, which no kernel
    checks (cu:1888-1898, 2158-2160), so the run is plain reachability with
    the cycle proviso that -p switches on (cu:1857-1861)"""
    net = gen.config1(3)
    net.liveness = 1
    base = str(tmp_path / "live")
    net.write(base)
    args = [base, "-p"] + (["--por"] if por else [])
    say("liveness", args)
    r = run(gpu, *args)
    assert r.returncode == 0, r.stderr
    assert "Property to check is a liveness property" in r.stdout
    assert "violation detected" not in r.stdout
    want = oracle_mod.explore(base, 0, por=por, proviso=por)
    assert final_count(r.stdout) == want["states"]
