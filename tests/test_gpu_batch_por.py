"""Batched exploration with partial-order reduction: instances of every POR
method, with and without the cycle proviso, run in one batch beside plain
instances; each instance's counts, levels and verdict equal the oracle's
with the same options, for any deferral pattern."""

import os
import sys

import pytest

from gpuexplore_amd import gen
from test_oracle import small_models

pytestmark = pytest.mark.gpu

MODES = [(False, False, 0), (True, False, 0), (True, True, 0), (True, False, 1), (True, True, 1),
         (True, False, 2), (True, True, 2)]  # (por, proviso, method)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def cases(oracle_mod, d, nets):
    out = []
    for name, net in nets:
        p = os.path.join(d, name)
        net.write(p)
        for prop in (0, 1, 2):
            for por, prov, method in MODES:
                o = oracle_mod.explore(p, prop, por, prov, method=method)
                out.append((p, prop, por, prov, method, o))
    return out


@pytest.fixture(scope="module")
def small_cases(tmp_path_factory, oracle_mod):
    return cases(oracle_mod, str(tmp_path_factory.mktemp("bpor")), small_models())


def run(gx, cs, **kw):
    models = [gx.Model.load(p, por=por) for p, _, por, _, _, _ in cs]
    jobs = [gx.job_options(prop, por, prov, method) for _, prop, por, prov, method, _ in cs]
    kw.setdefault("q_ints", len(cs) << 20)
    return gx.explore_batch(models, jobs=jobs, **kw)


def check(cs, res):
    for (p, prop, por, prov, method, o), r in zip(cs, res):
        assert r["table_full"] == 0, p
        got = (r["states"], r["transitions"], r["levels"], bool(r["violation"]))
        want = (o["states"], o["transitions"], o["levels"], o["violation"])
        assert got == want, (p, prop, por, prov, method, got, want)


def test_batch_por_all_modes(gpu, small_cases):
    say("batch POR, all modes,", len(small_cases), "instances")
    res, stats, rc = run(gpu, small_cases)
    assert rc == 0
    check(small_cases, res)


@pytest.mark.parametrize("kw", [dict(cand_capacity=3000, levels_per_sync=2),
                                dict(use_graph=0, levels_per_sync=1, threads=128, blocks=7)],
                         ids=["deferred", "nograph-small-grid"])
def test_batch_por_invariance(gpu, small_cases, kw):
    say("batch POR", kw)
    res, _, rc = run(gpu, small_cases, **kw)
    assert rc == 0
    check(small_cases, res)


def test_batch_por_edge_models(gpu, oracle_mod, tmp_path):
    """3-4 key words and 32 processes under POR in one batch"""
    from edge_models import edge_models
    nets = [(n, net) for n, net in edge_models() if n in ("wide3_1", "wide4", "procs32_1", "procs32_2")]
    cs = cases(oracle_mod, str(tmp_path), nets)
    say("batch POR edge models", len(cs))
    res, _, rc = run(gpu, cs)
    assert rc == 0
    check(cs, res)


def test_batch_config3_por(gpu, oracle_mod, tmp_path):
    """the 24-process POR workload in every POR mode, with -d and without,
    in one batch beside plain config-1 rings"""
    p3 = str(tmp_path / "config3")
    gen.config3(3).write(p3)
    cs = []
    for prop in (0, 1):
        for por, prov, method in MODES[1:]:
            say("oracle config3", prop, prov, method)
            cs.append((p3, prop, True, prov, method, oracle_mod.explore(p3, prop, True, prov, method=method)))
    for seed in (1, 2):
        p1 = str(tmp_path / f"ring{seed}")
        gen.config1(seed).write(p1)
        cs.append((p1, 1, False, False, 0, oracle_mod.explore(p1, 1)))
    res, _, rc = run(gpu, cs, q_ints=len(cs) << 22)
    assert rc == 0
    check(cs, res)
