"""Serving: jobs submitted while others run are admitted as slots and pages
free up; every job's result equals its result alone (gx_explore) and the
oracle's, whatever its footprint estimate."""

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
def jobs(tmp_path_factory, oracle_mod):
    d = tmp_path_factory.mktemp("serve")
    out = []
    for seed in range(40):
        net, flag, est = gen.config5_job(seed)
        p = os.path.join(str(d), f"j{seed}")
        net.write(p)
        o = oracle_mod.explore(p, 1)
        out.append((p, est, o))
    return out


def serve(gx, jobs, estimate=None, per_step=3, **kw):
    srv = gx.Server(**kw)
    ids = {}
    results = {}
    pending = list(enumerate(jobs))
    steps = 0
    while pending or len(results) < len(jobs):
        for _ in range(min(per_step, len(pending))):  # arrivals interleaved with steps
            i, (p, est, _) = pending.pop(0)
            jid = srv.submit(gx.Model.load(p), 1, est if estimate is None else estimate)
            ids[jid] = i
        srv.step()
        for r in srv.poll():
            results[ids[r["job_id"]]] = r
        steps += 1
        assert steps < 100000
    return results


def check(jobs, results):
    for i, (p, _, o) in enumerate(jobs):
        r = results[i]
        assert r["status"] == 0 and r["table_full"] == 0, (p, r)
        got = (r["states"], r["transitions"], r["levels"], bool(r["violation"]))
        assert got == (o["states"], o["transitions"], o["levels"], o["violation"]), (p, got, o)


def test_serve_mix_equals_oracle(gpu, jobs):
    say("serve mix")
    res = serve(gpu, jobs, slots=16, table_bytes=1 << 30, page_bytes=1 << 20)
    check(jobs, res)


def test_serve_alone_equals_mix(gpu, jobs):
    for p, est, o in jobs[:6]:
        say("alone", p)
        r = gpu.explore(gpu.Model.load(p), property=1, q_ints=1 << 24)
        assert (r["states"], r["transitions"], r["levels"]) == (o["states"], o["transitions"], o["levels"])


def test_serve_restarts_on_tiny_estimates(gpu, jobs):
    """estimates of one state: partitions fill, jobs restart with more pages"""
    say("serve tiny estimates")
    res = serve(gpu, jobs[:12], estimate=1, slots=8, table_bytes=1 << 28, page_bytes=1 << 16)
    check(jobs[:12], res)
    assert any(r["restarts"] > 0 for r in res.values())


def test_serve_chunk_beyond_candidate_buffer(gpu, jobs, tmp_path):
    """A job one of whose 64-state chunks emits more successors than the
    whole candidate buffer fails alone with GX_E_CAPACITY; the server keeps
    running, and a job that fits completes with its exact counts."""
    say("serve chunk beyond the candidate buffer")
    small = str(tmp_path / "cycle5")
    gen.independent_cycles([5]).write(small)  # one successor per state
    big_path = str(tmp_path / "ring6")
    gen.config1(1).write(big_path)  # 13,440 states: wide levels of out-degree 2
    srv = gpu.Server(slots=4, table_bytes=1 << 26, page_bytes=1 << 16, cand_capacity=64)
    j_big = srv.submit(gpu.Model.load(big_path), 0, 20000)
    j_small = srv.submit(gpu.Model.load(small), 0, 5)
    done = {}
    for _ in range(1000):
        srv.step()
        for r in srv.poll():
            done[r["job_id"]] = r
        if len(done) == 2:
            break
    assert set(done) == {j_big, j_small}
    assert done[j_big]["status"] == -10 and done[j_big]["restarts"] == 0
    s = done[j_small]
    alone = gpu.explore(gpu.Model.load(small), q_ints=1 << 16)
    assert s["status"] == 0 and (s["states"], s["transitions"]) == (5, 5)  # 1 cycle of 5
    assert s["levels"] == alone["levels"]
    # the server still admits and completes work afterwards
    j3 = srv.submit(gpu.Model.load(small), 0, 5)
    for _ in range(1000):
        srv.step()
        got = {r["job_id"]: r for r in srv.poll()}
        if j3 in got:
            assert got[j3]["status"] == 0 and got[j3]["states"] == 5
            break
    else:
        raise AssertionError("server stuck after a candidate overflow")
