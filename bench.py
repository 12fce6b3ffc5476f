#!/usr/bin/env python3
"""Benchmark of the gather/explore path on one MI355X.

Workload (default): the large plain-method model -- a ring of 20 seeded
random LTSs with 16 local states each (80-bit state vectors), out-degree 3,
4 local labels, a neighbour rendezvous and a 20-way barrier, token-throttled
(gpuexplore_amd.gen.config2, seed 1) -- explored without a property in a
visited set of 2^34 32-bit ints (64 GiB).  One step is one full exploration
from the initial state (the visited set is reset at the start of the step;
device buffers are allocated once, before the timed steps).

Prints one JSON line: new states/s (whole job), the FIND-OR-PUT kernel's
achieved algorithmic bandwidth against the HBM peak (and, from the committed
rocprofv3 counter summary of the same command, the bytes it moved to and
from memory and its atomic rate), and a CPU baseline: the oracle's BFS of the
whole workload on the host cores one GPU's share allows (16), and a bounded
sample of it on one core.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBPS = 8000.0  # MI355X HBM3E, published
BYTES_PER_FINDORPUT = 128  # one 128-byte bucket line per lookup, one per insert
GOLDEN = os.path.join(ROOT, "tests", "golden", "config2_counts.json")
CPU_CORES = 16  # the host share of one GPU


def batched_vs_plain(gx, gen, tmp: str) -> dict:
    """64 models of uneven size (gen.config4_models, seeds 1000-1063; half
    -d, half -p): one gx_explore_batch with autotuned settings against one
    exploration per model (explorers created before timing)."""
    from gpuexplore_amd import autotune

    paths, props = [], []
    for i, (net, flag) in enumerate(gen.config4_models(1000, 64)):
        p = os.path.join(tmp, f"c4_{i}")
        net.write(p)
        paths.append(p)
        props.append(gx.PROP_DEADLOCK if flag == "-d" else gx.PROP_SAFETY)
    models = [gx.Model.load(p) for p in paths]
    q_batch = 64 << 24  # 64 MiB of visited set per instance
    tuned = autotune.tune_batch(models, props, q_batch, repeats=2)
    best = tuned["best"]
    runs = []
    for _ in range(3):
        res, stats, rc = gx.explore_batch(models, props, q_ints=q_batch, **best)
        runs.append(stats)
    t_batch = min(s["t_explore_s"] for s in runs)
    states = runs[0]["states"]
    if any(r["table_full"] for r in res):
        raise RuntimeError("a batch partition filled up")
    explorers = [gx.Explorer(m, property=pr, q_ints=1 << 24) for m, pr in zip(models, props)]
    plain_t = []
    for _ in range(2):
        t = 0.0
        for i, ex in enumerate(explorers):
            r = ex.run()
            if r["table_full"] or (r["states"], r["transitions"]) != (res[i]["states"], res[i]["transitions"]):
                raise RuntimeError(f"batched and plain counts differ for model {i}")
            t += r["t_explore_s"]
        plain_t.append(t)
    t_plain = min(plain_t)
    return {
        "workload": "config4: 64 models, N 4..14, 4..32 states, 32 -d + 32 -p",
        "states": states,
        "batched_new_states_per_s": states / t_batch,
        "plain_new_states_per_s": states / t_plain,
        "ratio": t_plain / t_batch,
        "t_batched_s": t_batch,
        "t_plain_s": t_plain,
        "autotuned": best,
        "rounds": runs[0]["rounds"],
        "graph_launches": runs[0]["graph_launches"],
    }


def por_reduction(gx, gen, tmp: str) -> dict:
    """config 3 (24 processes, 8 states, phased rounds): the full state space
    once, then every POR method with no property, with -d and with the cycle
    proviso; each reduced run is the best of three."""
    p = os.path.join(tmp, "config3")
    gen.config3(3).write(p)
    full_model = gx.Model.load(p)
    ex = gx.Explorer(full_model, q_ints=1 << 34, property=gx.PROP_NONE)
    full = ex.run()
    del ex
    por_model = gx.Model.load(p, por=True)
    out = {"workload": "config3: 24 LTSs x 8 states, 72-bit vectors, phased rounds (70/20/10 local/rendezvous/barrier)",
           "full": {"states": full["states"], "transitions": full["transitions"], "t_s": full["t_explore_s"]},
           "methods": {}}
    for name, method in (("cample", gx.POR_CAMPLE), ("ample", gx.POR_AMPLE), ("stubborn", gx.POR_STUBBORN)):
        row = {}
        for leg, prop, prov in (("none", gx.PROP_NONE, 0), ("deadlock", gx.PROP_DEADLOCK, 0), ("proviso", gx.PROP_NONE, 1)):
            ex = gx.Explorer(por_model, q_ints=1 << 26, property=prop, por=1, cycle_proviso=prov, por_method=method)
            rs = [ex.run() for _ in range(3)]
            r = min(rs, key=lambda x: x["t_explore_s"])
            row[leg] = {"states": r["states"], "transitions": r["transitions"], "levels": r["levels"],
                        "violation": bool(r["violation"]), "t_s": r["t_explore_s"]}
            del ex
        row["reduction"] = full["states"] / row["none"]["states"]
        out["methods"][name] = row
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--q-log2", type=int, default=34, help="visited set size: 2^q 32-bit ints")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-sample-states", type=int, default=0,
                    help="multi-threaded CPU baseline: stop after the level reaching this many states (0: all)")
    ap.add_argument("--cpu-single-states", type=int, default=10_000_000,
                    help="one-thread CPU baseline: stop after the level reaching this many states")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--expand-kernel", type=int, default=0, help="0 auto, 1 process groups, 2 one state per lane")
    ap.add_argument("--bucket-bytes", type=int, default=0, help="0 auto (64 for 2-word slots)")
    ap.add_argument("--threads", type=int, default=0, help="successor-kernel workgroup size, 0 = library default")
    ap.add_argument("--insert-blocks", type=int, default=0, help="FIND-OR-PUT workgroups per launch, 0 = sized per level")
    ap.add_argument("--dedup", type=int, default=0, help="FIND-OR-PUT duplicate filter (1 on, 0 off)")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched vs plain comparison")
    ap.add_argument("--no-por", action="store_true", help="skip the config 3 partial-order reduction runs")
    ap.add_argument("--serve-seconds", type=float, default=20.0,
                    help="length of the Poisson job stream through the server (0 skips it)")
    ap.add_argument("--serve-rate", type=float, default=20.0, help="jobs per second of the stream")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              f"(asked for --gpus {args.gpus})", file=sys.stderr)
        return 2

    import gpuexplore_amd as gx
    from gpuexplore_amd import gen

    if gx.device_count() == 0:
        print("no HIP device visible", file=sys.stderr)
        return 1
    tmp = tempfile.mkdtemp(prefix="gx_bench_")
    path = os.path.join(tmp, "config2")
    gen.config2(args.seed).write(path)
    model = gx.Model.load(path)
    info = model.info()

    import ctypes
    peak = ctypes.c_double(0)
    gx.lib().gx_bench_copy(4 << 30, 20, ctypes.byref(peak))
    peak_read = ctypes.c_double(0)
    gx.lib().gx_bench_read(4 << 30, 20, ctypes.byref(peak_read))
    rand64 = ctypes.c_double(0)
    gx.lib().gx_bench_random(32 << 30, 1 << 29, 5, ctypes.byref(rand64))

    ex = gx.Explorer(model, q_ints=1 << args.q_log2, property=gx.PROP_NONE,
                     expand_kernel=args.expand_kernel, bucket_bytes=args.bucket_bytes, threads=args.threads,
                     insert_blocks=args.insert_blocks, dedup=args.dedup)
    for i in range(args.warmup):
        r = ex.run()
        print(f"warmup {i}: states={r['states']} t_explore={r['t_explore_s']:.3f}s", file=sys.stderr, flush=True)
    runs = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        r = ex.run()
        runs.append(r)
        print(f"step {i}: states={r['states']} t_explore={r['t_explore_s']:.3f}s "
              f"t_hash={r['t_hash_s']:.3f}s", file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    states = runs[0]["states"]
    if any(r["states"] != states or r["transitions"] != runs[0]["transitions"] for r in runs):
        print("state counts differ between steps", file=sys.stderr)
        return 1
    golden = None
    if args.seed == 1 and args.q_log2 == 34:
        # the oracle's committed full-size counts of this exact model
        g = json.load(open(GOLDEN))["runs"]["plain"]
        got = (states, runs[0]["transitions"], runs[0]["levels"], bool(runs[0]["violation"]))
        want = (g["states"], g["transitions"], g["levels"], g["violation"])
        if got != want:
            print(f"counts {got} differ from the oracle's {want}", file=sys.stderr)
            return 1
        golden = "equal to tests/golden/config2_counts.json (oracle)"
    t_exp = sum(r["t_explore_s"] for r in runs)
    t_hash = sum(r["t_hash_s"] for r in runs)
    lookups = sum(r["lookups"] for r in runs)
    inserts = sum(r["inserts"] for r in runs)
    launches = sum(r["hash_launches"] for r in runs)
    new_per_s = inserts / t_exp
    algo_bytes = BYTES_PER_FINDORPUT * (lookups + inserts)
    achieved = algo_bytes / t_hash / 1e9

    # HBM traffic per k_insert launch from the committed rocprofv3 --pmc runs of
    # this same command (FETCH_SIZE and WRITE_SIZE, scaled by the counters'
    # reading of a streaming copy of known size), when the workload matches
    traffic = physical = atomics = None
    prof = os.path.join(ROOT, "profiles", "k_insert_config2_traffic.json")
    if os.path.exists(prof) and args.q_log2 == 34 and args.seed == 1 and not args.bucket_bytes \
            and not args.expand_kernel:
        pj = json.load(open(prof))
        for k, v in pj["kernels"].items():
            if "k_insert" in k:
                traffic = v["bytes_per_launch"]
                physical = v.get("dram_bytes_per_launch")
                atomics = v.get("atomics_per_launch")

    batch = None if args.no_batch else batched_vs_plain(gx, gen, tmp)
    por = None
    if not args.no_por:
        del ex  # the config 2 table (2^34 ints) is not needed any more
        por = por_reduction(gx, gen, tmp)

    serving = None
    if args.serve_seconds > 0:
        from gpuexplore_amd import serve
        serving = serve.serve_workload(args.serve_rate, args.serve_seconds, seed=0,
                                       log=lambda m: print(m, file=sys.stderr, flush=True))
        serving["workload"] = ("config5: Poisson stream of gen.config5_job models: 90% 24-process --por -d, "
                               "10% 20-process 3e8-state -d (the large plain workload's generator)")

    cpu = None
    if not args.no_cpu:
        # the oracle's BFS on this host: the whole workload on the cores one
        # GPU's share of the host allows (16, whatever the machine has), and a
        # bounded sample of it (the first levels) on one core
        from oracle import oracle
        threads = max(1, min(CPU_CORES, len(os.sched_getaffinity(0))))
        om = oracle.OracleModel(path, False)
        full_cpu = args.cpu_sample_states == 0
        mt = om.explore_mt(threads, oracle.PROP_NONE, max_states=args.cpu_sample_states,
                           expected_states=states if full_cpu else args.cpu_sample_states)
        st = om.explore(oracle.PROP_NONE, max_states=args.cpu_single_states)
        if full_cpu and (mt["states"], mt["transitions"]) != (states, runs[0]["transitions"]):
            print("CPU baseline counts differ from the GPU's", file=sys.stderr)
            return 1
        what = "the whole workload" if full_cpu else f"its first {mt['levels']} levels"
        cpu = {
            "value": mt["states"] / mt["seconds"],
            "unit": "new states/s",
            "cores": threads,
            "kind": "port",
            "sample": (f"oracle BFS (orc_explore_mt, sharded presized visited set) of {what}: "
                       f"{mt['states']} states, {mt['transitions']} transitions, {mt['levels']} levels in "
                       f"{mt['seconds']:.1f} s on {threads} threads (host cores {os.cpu_count()}); "
                       f"one thread: first {st['levels']} levels ({st['states']} states) in {st['seconds']:.1f} s"),
            "single_thread": {"value": st["states"] / st["seconds"], "cores": 1, "seconds": st["seconds"],
                              "states": st["states"]},
        }

    line = {
        "metric": "new states/sec + hash-phase HBM GB/s vs peak, 1 GPU; batched vs plain",
        "value": new_per_s,
        "unit": "new states/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * t_exp / args.steps,
        "ms_per_step_wall": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {
            "workload": "config2: ring of 20 LTSs x 16 states, 80-bit vectors, plain method, no property",
            "generator": f"gpuexplore_amd.gen.config2(seed={args.seed})",
            "table_ints": 1 << args.q_log2,
            "states": states,
            "transitions": runs[0]["transitions"],
            "levels": runs[0]["levels"],
            "counts_check": golden,
            "key_words": info["key_words"],
            "bucket_bytes": args.bucket_bytes,
            "threads": args.threads,
            "insert_blocks": args.insert_blocks,
            "dedup": args.dedup,
            "expand_kernel": ["auto", "process-groups", "state-per-lane"][args.expand_kernel],
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": PEAK_HBM_GBPS,
            "unit": "GB/s",
            "frac": achieved / PEAK_HBM_GBPS,
            "traffic": traffic,
            "traffic_source": "profiles/k_insert_config2_traffic.json" if traffic else None,
            # bytes the kernel's L2 misses and write-backs moved to and from
            # memory (MALL or HBM), from the counters, over its launch time
            "physical_gbps": physical / (t_hash / launches) / 1e9 if physical else None,
            "physical_frac": physical / (t_hash / launches) / 1e9 / PEAK_HBM_GBPS if physical else None,
            "kernel": "k_insert (FIND-OR-PUT)",
            "algorithmic_bytes_per_launch": algo_bytes / launches,
            "avg_launch_us": 1e6 * t_hash / launches,
            "measured_copy_peak_gbps": peak.value,
            "measured_read_peak_gbps": peak_read.value,
            "measured_random64_lookups_per_s": rand64.value,
            "lookups_per_s": lookups / t_hash,
            "lookups_vs_random64": lookups / t_hash / rand64.value if rand64.value else None,
        },
        "hash_phase": {
            "candidates": sum(r["transitions"] for r in runs) // args.steps,
            "lookups": lookups // args.steps,
            "dedup_share": 1.0 - lookups / max(1, sum(r["transitions"] for r in runs)),
            "lookups_per_s": lookups / t_hash,
            "inserts_per_s": inserts / t_hash,
            # each insert claims its slot with one 64-bit atomicCAS; attempts
            # lost to a racing insert into the same slot are not counted
            "cas_per_s": inserts / t_hash,
            # every L2 atomic of the kernel (CAS claims, frontier and counter
            # adds), TCC_ATOMIC from the committed counter summary
            "atomics_per_s": atomics / (t_hash / launches) if atomics else None,
            "t_hash_s_per_step": t_hash / args.steps,
            "t_expand_s_per_step": sum(r["t_expand_s"] for r in runs) / args.steps,
        },
        # the hash bytes plus each new state's frontier entry written once and
        # read once, counted at the reference's 4 * sv_nints bytes per vector
        # (the frontier here holds 8 * key_words bytes per state)
        "whole_iteration": {
            "bytes_per_step": (algo_bytes + 2 * 4 * info["sv_nints"] * inserts) / args.steps,
            "gbps": (algo_bytes + 2 * 4 * info["sv_nints"] * inserts) / t_exp / 1e9,
        },
        "cpu_baseline": cpu,
        "batched_vs_plain": batch,
        "por": por,
        "serving": serving,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
