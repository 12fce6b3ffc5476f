"""Serving: jobs submitted while others run are admitted as slots and pages
free up; every job's result equals its result alone (gx_explore) and the
oracle's (or the committed full-size counts), whatever its footprint
estimate.  The stream's mix: 24-process POR jobs (--por -d) beside small
plain jobs, and a 20-process job at the large plain workload's full scale."""

import json
import os
import sys

import pytest

from gpuexplore_amd import gen

pytestmark = pytest.mark.gpu

KEYS = ("states", "transitions", "levels", "violation")
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
# config-3-generator seeds whose POR-reduced spaces are small (7e4..2e5
# states), so the oracle checks each in well under a second
POR_SEEDS = (3, 4, 7, 20, 21, 22, 37, 1)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def small_job(seed):
    """a small plain job: phased rounds of 6..12 processes with 8 states"""
    rng = gen.Rng(0x5A11 + seed)
    n = 6 + rng.below(7)
    return gen.rounds(n, 8, 4, [1] * n, 4, seed, mix=(0.7, 0.2, 0.1), phased=True), int(3.0 ** n)


@pytest.fixture(scope="module")
def jobs(tmp_path_factory, oracle_mod):
    """(path, property, est, por, oracle result): the POR jobs and 32 small
    plain jobs, interleaved"""
    d = str(tmp_path_factory.mktemp("serve"))
    out = []
    for k in range(40):
        if k % 5 == 0:
            seed = POR_SEEDS[k // 5]
            net = gen.rounds(24, 8, 4, [1] * 24, 4, seed, mix=(0.7, 0.2, 0.1), phased=True)
            p = os.path.join(d, f"por{seed}")
            net.write(p)
            out.append((p, 1, 2_000_000, True, oracle_mod.explore(p, 1, True, False)))
        else:
            net, est = small_job(k)
            p = os.path.join(d, f"j{k}")
            net.write(p)
            out.append((p, 1, est, False, oracle_mod.explore(p, 1)))
    return out


def serve(gx, jobs, estimate=None, per_step=3, **kw):
    srv = gx.Server(**kw)
    ids = {}
    results = {}
    pending = list(enumerate(jobs))
    steps = 0
    while pending or len(results) < len(jobs):
        for _ in range(min(per_step, len(pending))):  # arrivals interleaved with steps
            i, (p, prop, est, por, _) = pending.pop(0)
            jid = srv.submit(gx.Model.load(p, por=por), prop, est if estimate is None else estimate, por=por)
            ids[jid] = i
        srv.step()
        for r in srv.poll():
            results[ids[r["job_id"]]] = r
        steps += 1
        assert steps < 100000
    return results


def check(jobs, results):
    for i, (p, _, _, _, o) in enumerate(jobs):
        r = results[i]
        assert r["status"] == 0 and r["table_full"] == 0, (p, r)
        got = (r["states"], r["transitions"], r["levels"], bool(r["violation"]))
        assert got == (o["states"], o["transitions"], o["levels"], o["violation"]), (p, got, o)


def test_serve_mix_equals_oracle(gpu, jobs):
    say("serve mix")
    res = serve(gpu, jobs, slots=16, table_bytes=1 << 30, page_bytes=1 << 20)
    check(jobs, res)


def test_serve_alone_equals_mix(gpu, jobs):
    for p, prop, est, por, o in jobs[:11]:
        say("alone", p)
        r = gpu.explore(gpu.Model.load(p, por=por), property=prop, por=int(por), q_ints=1 << 24)
        assert (r["states"], r["transitions"], r["levels"]) == (o["states"], o["transitions"], o["levels"])


def test_serve_restarts_on_tiny_estimates(gpu, jobs):
    """estimates of one state: partitions fill, jobs restart with more pages"""
    say("serve tiny estimates")
    res = serve(gpu, jobs[:12], estimate=1, slots=8, table_bytes=1 << 28, page_bytes=1 << 16)
    check(jobs[:12], res)
    assert any(r["restarts"] > 0 for r in res.values())


def test_serve_por_proviso_and_methods(gpu, oracle_mod, tmp_path):
    """POR jobs of every method, with and without the cycle proviso, served
    together"""
    cases = []
    for seed in (3, 20):
        p = str(tmp_path / f"por{seed}")
        gen.rounds(24, 8, 4, [1] * 24, 4, seed, mix=(0.7, 0.2, 0.1), phased=True).write(p)
        for method in (0, 1, 2):
            for prov in (False, True):
                cases.append((p, method, prov, oracle_mod.explore(p, 0, True, prov, method=method)))
    srv = gpu.Server(slots=16, table_bytes=1 << 30, page_bytes=1 << 20)
    ids = {srv.submit(gpu.Model.load(p, por=True), 0, 300000, por=True, cycle_proviso=prov, por_method=m): k
           for k, (p, m, prov, _) in enumerate(cases)}
    got = {}
    for _ in range(100000):
        srv.step()
        for r in srv.poll():
            got[ids[r["job_id"]]] = r
        if len(got) == len(cases):
            break
    for k, (p, m, prov, o) in enumerate(cases):
        r = got[k]
        assert r["status"] == 0, (p, m, prov, r)
        assert (r["states"], r["transitions"], r["levels"], bool(r["violation"])) == \
            (o["states"], o["transitions"], o["levels"], o["violation"]), (p, m, prov)


def test_serve_config2_scale_job_in_the_mix(gpu, jobs, tmp_path):
    """the 20-process, 3e8-state job at full scale, served beside POR and
    small jobs, against the oracle's committed full-size counts"""
    g = json.load(open(os.path.join(GOLDEN, "config2_counts.json")))["runs"]["plain"]
    p2 = str(tmp_path / "config2")
    gen.config2(1).write(p2)
    srv = gpu.Server(slots=32, table_bytes=16 << 30, page_bytes=4 << 20)
    say("serve config2 scale + mix")
    big = srv.submit(gpu.Model.load(p2), 0, 350_000_000)
    ids = {big: -1}
    for i, (p, prop, est, por, _) in enumerate(jobs[:10]):
        ids[srv.submit(gpu.Model.load(p, por=por), prop, est, por=por)] = i
    got = {}
    for _ in range(100000):
        srv.step()
        for r in srv.poll():
            got[ids[r["job_id"]]] = r
        if len(got) == len(ids):
            break
    r = got[-1]
    assert r["status"] == 0 and r["table_full"] == 0, r
    assert (r["states"], r["transitions"], r["levels"], bool(r["violation"])) == \
        (g["states"], g["transitions"], g["levels"], g["violation"])
    check(jobs[:10], {i: got[i] for i in range(10)})


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


def test_serve_workload_short_stream(gpu):
    """the benchmark's serving leg end to end on a short stream (ten jobs, two
    of them 20-process 3e8-state ones): every served verdict and count equals
    the job's plain-path result, in the stream and with every job submitted
    at once"""
    from gpuexplore_amd import serve
    out = serve.serve_workload(20.0, 0.5, seed=0, log=say)
    say("short stream", out["jobs"], out["states"], out["p50_time_to_verdict_s"])
    assert out["jobs"] == 10
    assert out["verdict_mismatches"] == 0
    assert out["p50_time_to_verdict_s"] > 0
