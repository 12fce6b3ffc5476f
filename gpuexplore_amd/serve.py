"""Serving workload: a Poisson stream of model-checking jobs through gx_server.

Jobs come from :func:`gpuexplore_amd.gen.config5_job`: 90% are 24-process
phased-rounds models checked with ``--por -d`` (cluster-based POR), 10% are
20-process token-budget models at the large plain workload's scale (about
3e8 states each) checked with -d; each comes with a footprint estimate.
Arrival times are drawn up front; the driver submits every job whose arrival
time has passed, steps the server, and polls results.  Time-to-verdict is
measured from a job's scheduled arrival, so time a job spends waiting for the
driver loop counts against the server.

Every verdict and count is checked against the plain path (gx_explore on the
job alone).  The plain path's per-job wall times also drive a one-at-a-time
FIFO simulation on the same arrival times, which is what serving the stream
without continuous batching would give.

    python -m gpuexplore_amd.serve --rate 20 --seconds 120
"""

from __future__ import annotations

import argparse
import json
import math
import sys
import time
from typing import Dict, List, Optional

import gpuexplore_amd as gx
from gpuexplore_amd import gen


def make_jobs(count: int, seed0: int = 0) -> List[dict]:
    jobs = []
    for k in range(count):
        net, flag, est, por = gen.config5_job(seed0 + k)
        jobs.append({"seed": seed0 + k, "gpf": net.to_gpf(), "prop": 1 if flag == "-d" else 2,
                     "est": est, "procs": len(net.nstates), "por": por})
    return jobs


def arrivals(count: int, rate: float, seed: int) -> List[float]:
    rng = gen.Rng(0xA77 + seed)
    t, out = 0.0, []
    for _ in range(count):
        t += -math.log(1.0 - rng.uniform()) / rate
        out.append(t)
    return out


def pct(xs: List[float], q: float) -> float:
    s = sorted(xs)
    if not s:
        return float("nan")
    return s[min(len(s) - 1, max(0, int(math.ceil(q * len(s))) - 1))]


# Server pools for the stream: a 96 GiB visited set (frontier pools of half
# its slots each), enough for six of the 5e8-state jobs at their estimate
# beside the POR jobs.
SERVER_KW = {"table_bytes": 96 << 30, "page_bytes": 4 << 20}


def run_stream(jobs: List[dict], times: Optional[List[float]], server_kw: Optional[dict] = None,
               log=None) -> Dict:
    """Runs the jobs through one server.  times=None submits everything at
    once (saturation throughput).  Returns per-job results and the stream's
    figures."""
    models = [gx.Model.from_text(j["gpf"], por=j["por"]) for j in jobs]
    srv = gx.Server(**(server_kw or SERVER_KW))
    n = len(jobs)
    sched = times if times is not None else [0.0] * n
    jid_of: Dict[int, int] = {}
    late: Dict[int, float] = {}
    res: Dict[int, dict] = {}
    nxt = 0
    t0 = time.monotonic()
    last_log = t0
    while len(res) < n:
        now = time.monotonic() - t0
        while nxt < n and sched[nxt] <= now:
            jid = srv.submit(models[nxt], jobs[nxt]["prop"], jobs[nxt]["est"], por=jobs[nxt]["por"])
            jid_of[jid] = nxt
            late[nxt] = (time.monotonic() - t0) - sched[nxt]  # driver lag before submit
            nxt += 1
        st = srv.step()
        for r in srv.poll():
            i = jid_of[r["job_id"]]
            r["latency_s"] = late[i] + (r["t_done"] - r["t_submit"])
            res[i] = r
        if st["running"] == 0 and st["waiting"] == 0 and nxt < n and len(res) == nxt:
            time.sleep(max(0.0, min(0.01, sched[nxt] - (time.monotonic() - t0))))
        if log and time.monotonic() - last_log > 10:
            last_log = time.monotonic()
            log(f"serve t={time.monotonic() - t0:.0f}s submitted {nxt} done {len(res)} running {st['running']}")
    wall = time.monotonic() - t0
    first = sched[0] if times is not None else 0.0
    lat = [res[i]["latency_s"] for i in range(n)]
    states = sum(res[i]["states"] for i in range(n))
    return {"results": [res[i] for i in range(n)], "wall_s": wall, "span_s": wall - first,
            "states": states, "p50_s": pct(lat, 0.50), "p99_s": pct(lat, 0.99),
            "mean_s": sum(lat) / n, "restarts": sum(res[i]["restarts"] for i in range(n))}


def plain_path(jobs: List[dict], log=None) -> List[dict]:
    """Each job alone through gx_explore: the reference for every verdict and
    the service time of a one-at-a-time server.  The table is sized from the
    estimate like the server's partition and grown 4x after a full table; the
    service time includes every attempt."""
    out = []
    t_log = time.monotonic()
    for k, j in enumerate(jobs):
        if log and time.monotonic() - t_log > 10:
            t_log = time.monotonic()
            log(f"plain path: {k} of {len(jobs)} jobs done")
        m = gx.Model.from_text(j["gpf"], por=j["por"])
        q = 1 << max(20, (8 * j["est"]).bit_length())  # 32-bit ints: >= 4 per estimated state
        # buffers sized from the estimate, not from free memory
        front = max(1 << 20, j["est"])
        cand = min(1 << 30, max(1 << 22, 4 * j["est"]))
        t = time.monotonic()
        while True:
            r = gx.explore(m, property=j["prop"], por=int(j["por"]), q_ints=q, frontier_capacity=front,
                           cand_capacity=cand)
            if r["status"] == -10:  # a level larger than the frontier buffer
                front *= 4
                continue
            if not r["table_full"] or q >= (1 << 34):
                break
            q = min(q * 4, 1 << 34)
        r["wall_s"] = time.monotonic() - t
        out.append(r)
    return out


def fifo_latencies(times: List[float], service: List[float]) -> List[float]:
    free, lat = 0.0, []
    for a, s in zip(times, service):
        start = max(a, free)
        free = start + s
        lat.append(free - a)
    return lat


def count_mismatches(jobs: List[dict], served: List[dict], plain: List[dict]) -> int:
    bad = 0
    for j, s, p in zip(jobs, served, plain):
        a = (s["states"], s["transitions"], s["levels"], s["violation"], s["table_full"], s["status"])
        b = (p["states"], p["transitions"], p["levels"], p["violation"], p["table_full"], 0)
        if a != b:
            bad += 1
            print(f"MISMATCH seed {j['seed']}: served {a} plain {b}", file=sys.stderr, flush=True)
    return bad


def serve_workload(rate: float, seconds: float, seed: int = 0, server_kw: Optional[dict] = None,
                   log=None, check: bool = True) -> Dict:
    """The stream through one server, then (with ``check``) every job alone
    through the plain path, as the verdict reference and the one-at-a-time
    baseline, and the whole stream submitted at once (saturation)."""
    count = max(1, int(round(rate * seconds)))
    jobs = make_jobs(count, seed * 1000003)
    times = arrivals(count, rate, seed)
    if log:
        log(f"serve: {count} jobs over {times[-1]:.1f}s (rate {rate}/s)")
    run = run_stream(jobs, times, server_kw, log)
    if not check:
        return {"jobs": count, "rate_per_s": rate, "seconds": seconds, "states": run["states"],
                "p50_time_to_verdict_s": run["p50_s"], "p99_time_to_verdict_s": run["p99_s"],
                "stream_states_per_s": run["states"] / run["span_s"], "restarts": run["restarts"],
                "by_kind": {k: {"p50_s": pct([run["results"][i]["latency_s"] for i, j in enumerate(jobs)
                                              if j["por"] == (k == "por_jobs")], 0.5),
                                "p99_s": pct([run["results"][i]["latency_s"] for i, j in enumerate(jobs)
                                              if j["por"] == (k == "por_jobs")], 0.99),
                                "restarts": sum(run["results"][i]["restarts"] for i, j in enumerate(jobs)
                                                if j["por"] == (k == "por_jobs"))}
                            for k in ("por_jobs", "large_plain_jobs")}}
    if log:
        log("serve: plain path for the verdict check")
    plain = plain_path(jobs, log)
    mismatches = count_mismatches(jobs, run["results"], plain)
    fifo = fifo_latencies(times, [p["wall_s"] for p in plain])
    burst = run_stream(jobs, None, server_kw, log)
    mismatches += count_mismatches(jobs, burst["results"], plain)
    plain_total = sum(p["wall_s"] for p in plain)
    by_kind = {}
    for kind, sel in (("por_jobs", lambda j: j["por"]), ("large_plain_jobs", lambda j: not j["por"])):
        idx = [i for i, j in enumerate(jobs) if sel(j)]
        lat = [run["results"][i]["latency_s"] for i in idx]
        by_kind[kind] = {"jobs": len(idx), "p50_s": pct(lat, 0.50), "p99_s": pct(lat, 0.99),
                         "states": sum(run["results"][i]["states"] for i in idx),
                         "restarts": sum(run["results"][i]["restarts"] for i in idx)}
    return {
        "by_kind": by_kind,
        "jobs": count, "rate_per_s": rate, "seconds": seconds,
        "states": run["states"],
        "p50_time_to_verdict_s": run["p50_s"], "p99_time_to_verdict_s": run["p99_s"],
        "mean_time_to_verdict_s": run["mean_s"],
        "stream_states_per_s": run["states"] / run["span_s"],
        "restarts": run["restarts"],
        "saturation_states_per_s": burst["states"] / burst["wall_s"],
        "saturation_wall_s": burst["wall_s"],
        "plain_sequential_states_per_s": run["states"] / plain_total,
        "plain_fifo_p50_s": pct(fifo, 0.50), "plain_fifo_p99_s": pct(fifo, 0.99),
        "verdict_mismatches": mismatches,
        "violations": sum(1 for r in run["results"] if r["violation"]),
    }


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--rate", type=float, default=20.0)
    ap.add_argument("--seconds", type=float, default=120.0)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--slots", type=int, default=256)
    ap.add_argument("--rounds-per-step", type=int, default=16)
    ap.add_argument("--table-gib", type=float, default=96.0)
    ap.add_argument("--large-cus", type=int, default=0, help="compute units of the large-job lane (0: all)")
    ap.add_argument("--no-check", action="store_true", help="the stream only: no plain-path check, no burst")
    args = ap.parse_args(argv)
    kw = dict(SERVER_KW, slots=args.slots, rounds_per_step=args.rounds_per_step,
              table_bytes=int(args.table_gib * (1 << 30)), large_cus=args.large_cus)
    out = serve_workload(args.rate, args.seconds, args.seed, kw,
                         log=lambda s: print(s, file=sys.stderr, flush=True), check=not args.no_check)
    print(json.dumps(out))
    return 1 if out.get("verdict_mismatches") else 0


if __name__ == "__main__":
    sys.exit(main())
