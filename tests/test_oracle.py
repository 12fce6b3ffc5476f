"""Pins the CPU oracle: hand-computed fixture, closed forms, codec brute
force, and an independent pure-Python enumerator on small models."""

import json
import os
import subprocess
import random

import pytest

from conftest import GOLDEN
from gpuexplore_amd import gen
from naive_enum import NaiveModel, pack, unpack

PROPS = {"": 0, "d": 1, "p": 2}


def write(net, d, name):
    path = os.path.join(d, name)
    net.write(path)
    return path


def test_tiny2_parse(oracle_mod):
    exp = json.load(open(os.path.join(GOLDEN, "tiny2.json")))
    m = oracle_mod.OracleModel(os.path.join(GOLDEN, "tiny2"))
    arr = m.arrays()
    for k, v in exp["parse"].items():
        assert arr[k] == v, k


def test_tiny2_successors(oracle_mod):
    exp = json.load(open(os.path.join(GOLDEN, "tiny2.json")))
    m = oracle_mod.OracleModel(os.path.join(GOLDEN, "tiny2"))
    for s, succ in exp["successors"].items():
        got, dead = m.successors([int(s)])
        assert sorted(w[0] for w in got) == sorted(succ)
        assert dead == exp["deadlock"][s]


def test_tiny2_explore(oracle_mod):
    exp = json.load(open(os.path.join(GOLDEN, "tiny2.json")))["explore"]
    p = os.path.join(GOLDEN, "tiny2")
    for name, prop in [("none", 0), ("deadlock", 1), ("safety", 2)]:
        r = oracle_mod.explore(p, prop)
        for k in ("states", "transitions", "levels", "violation"):
            assert r[k] == exp[name][k], (name, k)
    r = oracle_mod.explore(p, 1, por=True)
    for k in ("states", "transitions", "levels", "violation"):
        assert r[k] == exp["por"][k], k


def test_tiny2_por_clusters(oracle_mod):
    exp = json.load(open(os.path.join(GOLDEN, "tiny2.json")))["por_cluster"]
    m = oracle_mod.OracleModel(os.path.join(GOLDEN, "tiny2"), por=True)
    for s, c in exp.items():
        assert m.por_cluster([int(s)]) == c


@pytest.mark.parametrize("lengths", [[3], [2, 3], [3, 4, 5], [2, 2, 2, 2, 2]])
def test_independent_cycles(oracle_mod, model_dir, lengths):
    p = write(gen.independent_cycles(lengths), model_dir, "cyc")
    r = oracle_mod.explore(p, 1)
    prod = 1
    for n in lengths:
        prod *= n
    assert r["states"] == prod
    assert r["transitions"] == len(lengths) * prod
    assert not r["violation"]


@pytest.mark.parametrize("n,k", [(1, 4), (3, 2), (4, 3), (5, 1)])
def test_independent_chains(oracle_mod, model_dir, n, k):
    p = write(gen.independent_chains(n, k), model_dir, "chain")
    r = oracle_mod.explore(p, 1)
    assert r["states"] == (k + 1) ** n
    assert r["transitions"] == n * k * (k + 1) ** (n - 1)
    assert r["violation"]
    assert r["levels"] == n * k + 1
    # cample POR without proviso: the lowest movable process is selected
    r = oracle_mod.explore(p, 1, por=True)
    assert r["states"] == n * k + 1
    assert r["transitions"] == n * k
    assert r["violation"]


@pytest.mark.parametrize("n", [2, 3, 5, 8])
def test_barrier(oracle_mod, model_dir, n):
    p = write(gen.barrier_flip(n), model_dir, "bar")
    r = oracle_mod.explore(p, 1)
    assert r["states"] == 2 ** n
    assert r["transitions"] == n * 2 ** (n - 1) + 1
    assert not r["violation"]


def test_codec_bruteforce(oracle_mod, model_dir):
    """orc_get/orc_set against a bit-by-bit reference on layouts with fields
    straddling word boundaries."""
    rng = random.Random(7)
    for trial in range(6):
        widths = [rng.randint(1, 13) for _ in range(rng.randint(2, 12))]
        net = gen.independent_cycles([2] * len(widths))
        net.field_widths = widths
        p = write(net, model_dir, f"codec{trial}")
        om = oracle_mod.OracleModel(p)
        nm = NaiveModel(open(p + ".gpf").read())
        words = om.words
        for _ in range(50):
            s = tuple(rng.randrange(1 << w) for w in widths)
            v = pack(nm, s, words)
            for q in range(len(widths)):
                assert om.get(v, q) == s[q]
                x = rng.randrange(1 << widths[q])
                v2 = om.set(v, q, x)
                s2 = s[:q] + (x,) + s[q + 1:]
                assert unpack(nm, v2) == s2
                assert v2 == pack(nm, s2, words)


def small_models():
    out = []
    for seed in range(1, 4):
        out.append((f"ring{seed}", gen.config1(seed)))
        out.append((f"thr{seed}", gen.throttled(4, 8, 2, 2, (0.6, 0.3, 0.1), 2, seed)))
        out.append((f"mon{seed}", gen.throttled(4, 6, 2, 2, (0.5, 0.4, 0.1), 2, seed, monitor=(0, 1, 2, 3))))
        out.append((f"tr3{seed}", gen.throttled(3, 12, 3, 3, (0.7, 0.2, 0.1), 3, seed)))
        wide = gen.throttled(3, 6, 2, 2, (0.5, 0.4, 0.1), 2, seed)
        wide.field_widths = [17, 16, 31]  # straddles both word boundaries
        out.append((f"wide{seed}", wide))
    return out


@pytest.mark.parametrize("name,net", small_models(), ids=[n for n, _ in small_models()])
def test_oracle_vs_naive(oracle_mod, model_dir, name, net):
    p = write(net, model_dir, name)
    nm = NaiveModel(open(p + ".gpf").read())
    for prop in ("", "d", "p"):
        for por, prov in ((False, False), (True, False), (True, True)):
            a = nm.explore(prop, por, prov)
            b = oracle_mod.explore(p, PROPS[prop], por, prov)
            for k in ("states", "transitions", "levels", "violation"):
                assert a[k] == b[k], (prop, por, prov, k, a, b)


def test_por_subset_and_verdict(oracle_mod, model_dir):
    for seed in range(1, 6):
        p = write(gen.config1(seed), model_dir, f"c1_{seed}")
        full = oracle_mod.explore(p, 1)
        red = oracle_mod.explore(p, 1, por=True)
        prov = oracle_mod.explore(p, 1, por=True, proviso=True)
        assert red["states"] <= full["states"]
        assert prov["states"] <= full["states"]
        assert red["violation"] == full["violation"]


def test_seed_invariance(oracle_mod, model_dir):
    p = write(gen.config1(3), model_dir, "c1")
    a = oracle_mod.explore(p, 1, seed=1)
    b = oracle_mod.explore(p, 1, seed=987654321)
    assert (a["states"], a["transitions"], a["levels"]) == (b["states"], b["transitions"], b["levels"])


@pytest.mark.parametrize("prop", ["", "deadlock", "safety"])
@pytest.mark.parametrize("por,proviso", [(False, False), (True, False), (True, True)])
def test_empty_sync_entry_oracle_vs_naive(oracle_mod, tmp_path, prop, por, proviso):
    """an entry that keeps its label but has no targets: the rule stays
    enabled and contributes no successors"""
    from quirk_models import empty_sync_entry_gpf
    text = empty_sync_entry_gpf()
    p = str(tmp_path / "q")
    open(p + ".gpf", "w").write(text)
    want = NaiveModel(text).explore({"": "", "deadlock": "d", "safety": "p"}[prop], por=por, proviso=proviso)
    code = {"": 0, "deadlock": 1, "safety": 2}[prop]
    got = oracle_mod.explore(p, code, por=por, proviso=proviso)
    assert (got["states"], got["transitions"], got["levels"], got["violation"]) == \
        (want["states"], want["transitions"], want["levels"], want["violation"])


@pytest.mark.parametrize("threads", [1, 3, 8])
def test_multithreaded_oracle_equals_sequential(oracle_mod, model_dir, threads):
    """orc_explore_mt (the multi-core CPU baseline) gives orc_explore's counts"""
    for name, net in small_models()[:8] + [("c1", gen.config1(2))]:
        p = os.path.join(model_dir, name)
        net.write(p)
        for prop in (0, 1, 2):
            a = oracle_mod.explore(p, prop)
            b = oracle_mod.explore_mt(p, threads, prop)
            assert (a["states"], a["transitions"], a["levels"], a["violation"]) == \
                (b["states"], b["transitions"], b["levels"], b["violation"]), (name, prop, a, b)
    a = oracle_mod.explore(p, 0, max_states=5000)
    b = oracle_mod.explore_mt(p, threads, 0, max_states=5000)
    assert (a["states"], a["transitions"], a["levels"], a["truncated"]) == \
        (b["states"], b["transitions"], b["levels"], b["truncated"])


@pytest.mark.parametrize("expect", [1000, 100000])
def test_presized_level_free_set(oracle_mod, model_dir, expect):
    """the visited set presized (no growth) and without a level array (plain
    and POR without the proviso), in the sequential and the multi-threaded
    BFS, gives the growing set's counts; the proviso keeps its levels"""
    for name, net in small_models()[:6] + [("c1", gen.config1(3))]:
        p = os.path.join(model_dir, name)
        net.write(p)
        for prop in (0, 1, 2):
            for por, prov in ((False, False), (True, False), (True, True)):
                a = oracle_mod.explore(p, prop, por, prov)
                b = oracle_mod.explore(p, prop, por, prov, expected_states=expect)
                assert (a["states"], a["transitions"], a["levels"], a["violation"]) == \
                    (b["states"], b["transitions"], b["levels"], b["violation"]), (name, prop, por, prov)
            a = oracle_mod.explore(p, prop)
            c = oracle_mod.explore_mt(p, 4, prop, expected_states=expect)
            assert (a["states"], a["transitions"], a["levels"], a["violation"]) == \
                (c["states"], c["transitions"], c["levels"], c["violation"]), (name, prop)


def test_multithreaded_oracle_under_tsan(tmp_path):
    """the multi-threaded BFS (sharded visited set behind mutexes, frontier
    chunks taken from an atomic counter, per-thread next levels joined
    between levels) has no data race ThreadSanitizer can see"""
    oracle_dir = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")
    r = subprocess.run(["make", "-s", "-C", oracle_dir, "tsan"], capture_output=True, text=True)
    if r.returncode != 0:
        pytest.skip("no ThreadSanitizer runtime for gcc here: " + r.stderr[-200:])
    cli = os.path.join(oracle_dir, "_build", "gx_oracle_cli_tsan")
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1 exitcode=66")
    for name, net, flags in (("c1", gen.config1(1), []), ("c1d", gen.config1(2), ["-d"]),
                             ("mon", small_models()[2][1], ["-p"])):
        p = str(tmp_path / name)
        net.write(p)
        out = subprocess.run([cli, p, "--threads", "8", "--expect", "5000"] + flags, capture_output=True, text=True,
                             env=env, timeout=600)
        assert out.returncode == 0 and "ThreadSanitizer" not in out.stderr, (name, out.stderr[-2000:])
        seq = oracle_mod_explore(p, flags)
        got = dict(tok.split("=") for tok in out.stdout.split())
        assert (int(got["states"]), int(got["transitions"]), int(got["levels"])) == seq, name


def oracle_mod_explore(p, flags):
    from oracle import oracle
    prop = 1 if "-d" in flags else (2 if "-p" in flags else 0)
    o = oracle.explore(p, prop)
    return (o["states"], o["transitions"], o["levels"])


@pytest.mark.parametrize("name", [n for n, _ in __import__("edge_models").edge_models()])
def test_edge_models_oracle_vs_naive(oracle_mod, model_dir, name):
    """3-5 key words, 32 processes, more than 64 labels"""
    from edge_models import edge_models
    net = dict(edge_models())[name]
    p = write(net, model_dir, name)
    nm = NaiveModel(open(p + ".gpf").read())
    for prop in ("", "d", "p"):
        for por, prov in ((False, False), (True, False), (True, True)):
            if name == "wide5" and (prop, por) != ("", False):
                continue  # the largest one once: the naive enumerator is slow
            a = nm.explore(prop, por, prov)
            b = oracle_mod.explore(p, PROPS[prop], por, prov)
            for k in ("states", "transitions", "levels", "violation"):
                assert a[k] == b[k], (prop, por, prov, k, a, b)
