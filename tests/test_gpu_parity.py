"""HIP path against the oracle: identical state, transition and level counts
and identical verdicts (bit-exact: integer work), plus visited-set parity
against a Python set and invariance to launch geometry, hash seed and
bucket width."""

import json
import os
import random
import sys

import numpy as np
import pytest

from conftest import GOLDEN
from gpuexplore_amd import gen
from test_oracle import small_models

pytestmark = pytest.mark.gpu

KEYS = ("states", "transitions", "levels", "violation")
MODES = [(False, False), (True, False), (True, True)]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def gpu_run(gx, path, prop=0, por=False, proviso=False, **kw):
    m = gx.Model.load(path, por=por)
    kw.setdefault("q_ints", 1 << 22)
    r = gx.explore(m, property=prop, por=int(por), cycle_proviso=int(proviso), **kw)
    r["violation"] = bool(r["violation"])
    return r


def same(a, b):
    return all(a[k] == b[k] for k in KEYS)


def test_tiny2_golden(gpu):
    exp = json.load(open(os.path.join(GOLDEN, "tiny2.json")))["explore"]
    p = os.path.join(GOLDEN, "tiny2")
    for name, prop in [("none", 0), ("deadlock", 1), ("safety", 2)]:
        say("tiny2", name)
        r = gpu_run(gpu, p, prop)
        assert all(r[k] == exp[name][k] for k in KEYS), (name, r)
    r = gpu_run(gpu, p, 1, por=True)
    assert all(r[k] == exp["por"][k] for k in KEYS), r


def test_closed_forms(gpu, model_dir):
    cases = [
        (gen.independent_cycles([3, 4, 5]), 60, 180, False),
        (gen.independent_chains(4, 3), 256, 768, True),
        (gen.barrier_flip(6), 64, 6 * 32 + 1, False),
    ]
    for i, (net, st, tr, dead) in enumerate(cases):
        p = os.path.join(model_dir, f"cf{i}")
        net.write(p)
        say("closed form", i)
        r = gpu_run(gpu, p, 1)
        assert (r["states"], r["transitions"], r["violation"]) == (st, tr, dead)
    p = os.path.join(model_dir, "chainpor")
    gen.independent_chains(5, 2).write(p)
    r = gpu_run(gpu, p, 1, por=True)
    assert (r["states"], r["transitions"], r["violation"]) == (11, 10, True)


@pytest.mark.parametrize("name,net", small_models(), ids=[n for n, _ in small_models()])
def test_small_models_all_modes(gpu, oracle_mod, model_dir, name, net):
    p = os.path.join(model_dir, name)
    net.write(p)
    for prop in (0, 1, 2):
        for por, prov in MODES:
            say(name, prop, por, prov)
            a = oracle_mod.explore(p, prop, por, prov)
            b = gpu_run(gpu, p, prop, por, prov)
            assert same(a, b), (prop, por, prov, a, b)


@pytest.mark.parametrize("name,net", small_models(), ids=[n for n, _ in small_models()])
def test_small_models_group_kernel(gpu, oracle_mod, model_dir, name, net):
    """plain mode through the process-group kernel (the default picks the
    one-state-per-lane kernel for these models)"""
    p = os.path.join(model_dir, name)
    net.write(p)
    for prop in (0, 1, 2):
        say(name, prop, "group kernel")
        a = oracle_mod.explore(p, prop)
        b = gpu_run(gpu, p, prop, expand_kernel=1)
        assert same(a, b), (prop, a, b)


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_config1(gpu, oracle_mod, model_dir, seed):
    p = os.path.join(model_dir, "c1")
    gen.config1(seed).write(p)
    for prop, por, prov in [(1, False, False), (0, True, False), (2, True, True)]:
        say("config1", seed, prop, por, prov)
        a = oracle_mod.explore(p, prop, por, prov)
        b = gpu_run(gpu, p, prop, por, prov)
        assert same(a, b), (a, b)


def test_invariance_launch_seed_bucket(gpu, oracle_mod, model_dir):
    p = os.path.join(model_dir, "c1")
    gen.config1(2).write(p)
    a = oracle_mod.explore(p, 1)
    for kw in [dict(blocks=1, threads=64), dict(blocks=3, threads=128, kernel_iters=4),
               dict(threads=512, kernel_iters=16), dict(hash_seed=12345),
               dict(bucket_bytes=64), dict(bucket_bytes=256), dict(cand_capacity=4096)]:
        say("invariance", kw)
        b = gpu_run(gpu, p, 1, **kw)
        assert same(a, b), (kw, a, b)


def test_retry_path_por_proviso(gpu, oracle_mod, model_dir):
    """a tiny successor buffer forces deferred states; counts are unchanged"""
    p = os.path.join(model_dir, "c1")
    gen.config1(4).write(p)
    a = oracle_mod.explore(p, 2, True, True)
    say("retry proviso")
    b = gpu_run(gpu, p, 2, True, True, cand_capacity=2048)
    assert same(a, b), (a, b)


@pytest.mark.parametrize("W", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("bb", [64, 128, 256])
def test_table_vs_set(gpu, W, bb):
    if bb // 8 < W:
        pytest.skip("slot wider than bucket")
    rng = np.random.default_rng(W * 1000 + bb)
    n = 20000
    base = rng.integers(0, 1 << 62, size=(n // 4, W), dtype=np.uint64)
    keys = base[rng.integers(0, len(base), size=n)]  # ~4x duplication
    say("table", W, bb)
    t = gpu.Table(1 << 20, W, bb, seed=W)
    flags, ins, rc = t.insert(keys.ravel())
    assert rc == 0
    uniq = {tuple(r) for r in base}
    assert ins == len({tuple(r) for r in keys})
    # exactly one copy of every distinct key was new
    seen = {}
    for i, row in enumerate(map(tuple, keys)):
        seen.setdefault(row, 0)
        seen[row] += int(flags[i])
    assert all(v == 1 for v in seen.values())
    # a second pass finds everything
    flags2, ins2, rc = t.insert(keys.ravel())
    assert ins2 == 0 and not flags2.any()
    assert t.count() == ins
    assert len(uniq) >= ins


@pytest.mark.parametrize("W,bb", [(2, 64), (1, 128), (3, 128), (5, 128)])
def test_table_vs_set_runs(gpu, W, bb):
    """passes large enough that each wave takes runs in steps of 64
    candidates (a key hashed once and handed to its lane group), with
    duplicates a few candidates apart: in one step, across groups and across
    steps; exactly one copy of every distinct key is new"""
    rng = np.random.default_rng(7 + W)
    n = 9_000_000  # above 8 iterations of every wave of a 256-CU grid at 64 B buckets
    base = rng.integers(0, 1 << 62, size=(n // 4, W), dtype=np.uint64)
    # random 62-bit rows: distinct rows are distinct keys (a collision among
    # 2.25 M rows has probability below 1e-6), so a key is its row index
    assert len(np.unique(base[:, 0])) == len(base) or W > 1
    idx = np.minimum(np.arange(n) // 4 + rng.integers(0, 8, size=n), len(base) - 1)
    keys = np.ascontiguousarray(base[idx])
    say("table runs", W, bb, n)
    t = gpu.Table(1 << 27, W, bb, seed=3 * W)
    flags, ins, rc = t.insert(keys.ravel())
    assert rc == 0
    n_distinct = len(np.unique(idx))
    assert ins == n_distinct
    per_key = np.bincount(idx, weights=flags.astype(np.float64), minlength=len(base))
    assert (per_key[np.unique(idx)] == 1).all()
    flags2, ins2, rc = t.insert(keys.ravel())
    assert rc == 0 and ins2 == 0 and not flags2.any()
    assert t.count() == n_distinct


def test_table_full_reports(gpu):
    t = gpu.Table(256, 1, 128)  # 8 buckets x 16 slots
    keys = np.arange(1, 1001, dtype=np.uint64)
    say("table full")
    _, ins, rc = t.insert(keys)
    assert rc == -10
    assert ins <= 128


def test_tiny_table_explore_reports_full(gpu, model_dir):
    p = os.path.join(model_dir, "c1")
    gen.config1(2).write(p)
    say("tiny table explore")
    r = gpu_run(gpu, p, 1, q_ints=1024)
    assert r["table_full"] == 1


def test_empty_sync_entry(gpu, oracle_mod, model_dir):
    """a synchronising entry without targets (outside the generator's
    constraints): no successor from its rules, in every kernel and mode"""
    from quirk_models import empty_sync_entry_gpf
    p = os.path.join(model_dir, "quirk_empty")
    open(p + ".gpf", "w").write(empty_sync_entry_gpf())
    for prop in (0, 1, 2):
        for por, prov in MODES:
            a = oracle_mod.explore(p, prop, por, prov)
            kernels = (0, 1) if not por else (0,)
            for ek in kernels:
                say("quirk empty entry", prop, por, prov, ek)
                b = gpu_run(gpu, p, prop, por, prov, expand_kernel=ek)
                assert same(a, b), (prop, por, prov, ek, a, b)


@pytest.mark.parametrize("seed", range(8))
def test_converted_exp_networks(gpu, oracle_mod, model_dir, seed):
    """networks converted from EXP synchronisation vectors over .aut processes"""
    from test_exp2gpf import random_network
    from gpuexplore_amd import exp2gpf
    vectors, auts = random_network(seed)
    p = os.path.join(model_dir, f"exp{seed}")
    exp2gpf.network_from(vectors, auts).write(p)
    for prop in (0, 1):
        for por, prov in MODES:
            say("exp network", seed, prop, por, prov)
            a = oracle_mod.explore(p, prop, por, prov)
            b = gpu_run(gpu, p, prop, por, prov)
            assert same(a, b), (seed, prop, por, prov, a, b)


@pytest.mark.parametrize("name", [n for n, _ in __import__("edge_models").edge_models()])
def test_edge_models(gpu, oracle_mod, model_dir, name):
    """3-5 key words, 32 processes, more than 64 labels, in every mode and
    with every plain-mode kernel that accepts the model"""
    from edge_models import edge_models
    p = os.path.join(model_dir, name)
    dict(edge_models())[name].write(p)
    for prop in (0, 1, 2):
        for por, prov in MODES:
            a = oracle_mod.explore(p, prop, por, prov)
            for ek in ((0, 1) if not por else (0,)):
                say("edge", name, prop, por, prov, ek)
                b = gpu_run(gpu, p, prop, por, prov, expand_kernel=ek)
                assert same(a, b), (name, prop, por, prov, ek, a, b)


@pytest.mark.parametrize("kiters", [1, 2, 3, 5])
def test_stop_with_levels_enqueued_ahead(gpu, oracle_mod, model_dir, kiters):
    """the host keeps levels enqueued beyond the one it reads; a run that
    stops (deadlock found, state bound reached, frontier empty) still
    reports exactly the oracle's level and counts"""
    p = os.path.join(model_dir, "chain")
    gen.independent_chains(3, 3).write(p)
    for prop in (0, 1):
        a = oracle_mod.explore(p, prop)
        say("ahead chain", kiters, prop)
        b = gpu_run(gpu, p, prop, kernel_iters=kiters)
        assert same(a, b), (kiters, prop, a, b)
    p = os.path.join(model_dir, "c1")
    gen.config1(3).write(p)
    a = oracle_mod.explore(p, 0, max_states=2000)
    say("ahead truncated", kiters)
    b = gpu_run(gpu, p, 0, kernel_iters=kiters, max_states=2000)
    assert same(a, b) and a["truncated"] and b["truncated"], (kiters, a, b)


def test_dedup_filter_invariance(gpu, oracle_mod, model_dir):
    """The FIND-OR-PUT duplicate filter changes no count (the reference's
    LDS cache is semantics-neutral, cu:210-265): on and off give the same
    states, transitions, levels and verdicts in every mode, and with it on
    fewer candidates reach the visited set."""
    from test_oracle import small_models
    nets = small_models()[:5] + [("c1", gen.config1(1)), ("c3por", gen.config3(3))]
    for name, net in nets:
        p = os.path.join(model_dir, name)
        net.write(p)
        modes = [] if name == "c3por" else [(0, False, False), (1, False, False), (2, False, False)]
        modes += [(1, True, False), (0, True, True)]  # config 3 unreduced does not fit this table
        for prop, por, prov in modes:
            say("dedup", name, prop, por, prov)
            runs = [gpu.explore(gpu.Model.load(p, por=por), property=prop, por=int(por), cycle_proviso=int(prov),
                                q_ints=1 << 26, dedup=d) for d in (0, 1)]
            keys = ("states", "transitions", "levels", "violation")
            assert runs[0]["table_full"] == runs[1]["table_full"] == 0, name
            assert [runs[0][k] for k in keys] == [runs[1][k] for k in keys], (name, prop, por, prov)
            assert runs[1]["lookups"] <= runs[0]["lookups"]
    # a large plain level: the filter takes out a share of the lookups
    p = os.path.join(model_dir, "c2")
    gen.config2(1).write(p)
    runs = [gpu.explore(gpu.Model.load(p), q_ints=1 << 30, max_states=20_000_000, dedup=d) for d in (0, 1)]
    assert [runs[0][k] for k in ("states", "transitions", "levels")] == \
        [runs[1][k] for k in ("states", "transitions", "levels")]
    assert runs[0]["lookups"] == runs[0]["transitions"] + 1  # + the initial state's
    assert runs[1]["lookups"] < 0.97 * runs[0]["lookups"], (runs[0]["lookups"], runs[1]["lookups"])


# Two-word keys around the 12-byte candidate limit (95 state bits): plain
# explorations of such models hand successors over packed, wider ones in
# 16-byte form.  Fields straddle the 32-bit and the 63-bit word boundaries.
@pytest.mark.parametrize("widths", [[20, 32, 32, 11], [31, 32, 32], [32, 32, 32], [31, 32, 32, 31]],
                         ids=["95_straddle", "95", "96", "126"])
def test_two_word_keys_around_the_packed_limit(gpu, oracle_mod, model_dir, widths):
    net = gen.throttled(len(widths), 6, 2, 2, (0.5, 0.4, 0.1), 2, 7)
    net.field_widths = widths
    p = os.path.join(model_dir, "w2")
    net.write(p)
    info = gpu.Model.load(p).info()
    assert info["key_words"] == 2 and info["state_bits"] == sum(widths)
    for prop in (0, 1, 2):
        for por, prov in ((False, False), (True, False), (True, True)):
            say("two-word keys", widths, prop, por, prov)
            a = oracle_mod.explore(p, prop, por, prov)
            b = gpu_run(gpu, p, prop, por, prov)
            assert same(a, b), (widths, prop, por, prov, a, b)
            if not por:  # the same counts through the duplicate filter (16-byte candidates)
                assert same(a, gpu_run(gpu, p, prop, dedup=1)), (widths, prop)


@pytest.mark.parametrize("cap", [2048, 4096])
def test_packed_candidates_retry_path(gpu, oracle_mod, model_dir, cap):
    """12-byte candidates when a level's successors do not fit the candidate
    buffer: the 6-process ring of 64,512 states with its fields widened to 95
    state bits, levels of thousands of states, a buffer of a few batches"""
    net = gen.config1(2)
    net.field_widths = [15, 16, 16, 16, 16, 16]
    p = os.path.join(model_dir, "ring95")
    net.write(p)
    assert gpu.Model.load(p).info()["state_bits"] == 95
    for prop in (0, 1):
        say("packed retry", cap, prop)
        a = oracle_mod.explore(p, prop)
        b = gpu_run(gpu, p, prop, cand_capacity=cap)
        assert same(a, b), (cap, prop, a, b)
