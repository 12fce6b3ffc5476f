"""Batched exploration (many models in one persistent round structure):
every instance's counts and verdict equal the oracle's and the single-model
path's, for any launch geometry, graph use and deferral pattern."""

import os
import sys

import pytest

from gpuexplore_amd import gen

pytestmark = pytest.mark.gpu

KEYS = ("states", "transitions", "levels", "violation")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def batch_models(tmp_path_factory, oracle_mod):
    d = tmp_path_factory.mktemp("batch")
    out = []
    for i, (net, flag) in enumerate(gen.config4_models(1000, 64)):
        p = os.path.join(str(d), f"m{i}")
        net.write(p)
        prop = 1 if flag == "-d" else 2
        o = oracle_mod.explore(p, prop)
        out.append((p, prop, o))
    return out


def run_batch(gx, batch_models, **kw):
    models = [gx.Model.load(p) for p, _, _ in batch_models]
    props = [prop for _, prop, _ in batch_models]
    kw.setdefault("q_ints", 64 << 24)  # 64 MiB per instance
    res, stats, rc = gx.explore_batch(models, props, **kw)
    return res, stats, rc


def check(batch_models, res):
    for (p, prop, o), r in zip(batch_models, res):
        assert r["table_full"] == 0, (p, "partition too small")
        got = (r["states"], r["transitions"], r["levels"], bool(r["violation"]))
        want = (o["states"], o["transitions"], o["levels"], o["violation"])
        assert got == want, (p, prop, got, want)


def test_batch_matches_oracle(gpu, batch_models):
    say("batch default")
    res, stats, rc = run_batch(gpu, batch_models)
    assert rc == 0
    check(batch_models, res)
    assert stats["states"] == sum(o["states"] for _, _, o in batch_models)


@pytest.mark.parametrize("kw", [dict(use_graph=0, levels_per_sync=1),
                                dict(levels_per_sync=32, threads=512),
                                dict(levels_per_sync=3, threads=128, blocks=17, bucket_bytes=128),
                                dict(cand_capacity=20000, levels_per_sync=4)],
                         ids=["nograph", "k32", "small-grid", "deferred"])
def test_batch_invariance(gpu, batch_models, kw):
    say("batch", kw)
    res, stats, rc = run_batch(gpu, batch_models, **kw)
    assert rc == 0
    check(batch_models, res)


def test_batch_equals_single(gpu, batch_models):
    res, _, rc = run_batch(gpu, batch_models[:8])
    assert rc == 0
    for (p, prop, _), r in zip(batch_models[:8], res):
        say("single", p)
        s = gpu.explore(gpu.Model.load(p), property=prop, q_ints=1 << 22)
        assert all(s[k] == r[k] for k in ("states", "transitions", "levels", "violation"))


def test_batch_weights_and_full_partition(gpu, batch_models):
    """a partition too small for its instance reports table_full on that
    instance only"""
    models = [gpu.Model.load(p) for p, _, _ in batch_models[:4]]
    props = [prop for _, prop, _ in batch_models[:4]]
    sizes = [o["states"] for _, _, o in batch_models[:4]]
    big = max(range(4), key=lambda i: sizes[i])
    w = [1000 if i != big else 1 for i in range(4)]
    res, _, rc = gpu.explore_batch(models, props, weights=w, q_ints=4 * 4096)
    assert res[big]["table_full"] == 1 or sizes[big] < 100
    for i in range(4):
        if i != big and not res[i]["table_full"]:
            o = batch_models[i][2]
            assert (res[i]["states"], res[i]["transitions"]) == (o["states"], o["transitions"])


def test_batch_kernel_timing(gpu, batch_models):
    """time_kernels: rounds launched one by one with events around the
    successor and FIND-OR-PUT launches; counts equal the oracle's, one timed
    FIND-OR-PUT launch per round"""
    say("batch time_kernels")
    res, stats, rc = run_batch(gpu, batch_models, time_kernels=1)
    assert rc == 0
    check(batch_models, res)
    assert stats["insert_launches"] == stats["rounds"] == stats["graph_launches"]
    assert 0 < stats["t_insert_s"] < stats["t_explore_s"]
    assert 0 < stats["t_expand_s"] < stats["t_explore_s"]
