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


def insert_roofline(stats: dict, n_inst: int, opts: dict, prof_kernel: str = None, prof_file: str = None) -> dict:
    """The batched FIND-OR-PUT's roofline from a time_kernels run: algorithmic
    bytes (128 B per lookup and per insert, inserts = states found by the
    kernel, i.e. without the instances' initial states) over its event-timed
    launches; bytes to and from memory per launch from a committed counter
    summary of the same workload run with the same options, when there is
    one."""
    inserts = stats["states"] - n_inst
    algo = BYTES_PER_FINDORPUT * (stats["lookups"] + inserts)
    t = stats["t_insert_s"]
    out = {"kernel": "k_batch_insert (FIND-OR-PUT)", "bound": "hbm", "unit": "GB/s", "peak": PEAK_HBM_GBPS,
           "achieved": algo / t / 1e9, "frac": algo / t / 1e9 / PEAK_HBM_GBPS,
           "launches": stats["insert_launches"], "avg_launch_us": 1e6 * t / max(1, stats["insert_launches"]),
           "algorithmic_bytes_per_launch": algo / max(1, stats["insert_launches"]),
           "t_insert_s": t, "t_expand_s": stats["t_expand_s"], "lookups": stats["lookups"],
           "lookups_per_s": stats["lookups"] / t, "physical_gbps": None, "physical_frac": None, "traffic": None}
    path = os.path.join(ROOT, "profiles", prof_file) if prof_file else None
    if path and os.path.exists(path) and json.load(open(path)).get("options") == opts:
        pj = json.load(open(path))
        for k, v in pj["kernels"].items():
            if prof_kernel and prof_kernel in k:
                # bytes to and from memory over the kernel's time, every
                # launch of the profiled command (rounds that find every
                # instance finished move nothing and take microseconds)
                out["physical_gbps"] = v["gbps"]
                out["physical_frac"] = v["gbps"] / PEAK_HBM_GBPS
                out["traffic"] = v["gbps"] * 1e9 * (t / max(1, stats["insert_launches"]))
                out["traffic_source"] = "profiles/" + prof_file
    return out


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def plain_runs(gx, models, props, plain_q: int, **kw) -> list:
    """Each model alone through a gx_explorer (created outside the timing,
    one at a time): the second of two runs, whose counts must equal the
    first's."""
    out = []
    for i, (m, pr) in enumerate(zip(models, props)):
        ex = gx.Explorer(m, property=pr, q_ints=plain_q, **kw)
        r0, r = ex.run(), ex.run()
        if r["table_full"] or (r["states"], r["transitions"]) != (r0["states"], r0["transitions"]):
            raise RuntimeError(f"plain runs of model {i} differ or filled the table")
        out.append(r)
        del ex
        if (i + 1) % 16 == 0:
            log(f"plain path: {i + 1} of {len(models)} models")
    return out


def batch_leg(gx, models, props, q_batch: int, opts: dict, plain: list, weights=None,
              prof_file: str = None) -> dict:
    """One batch of models through gx_explore_batch (best of three, with the
    given options) and once more with kernel timing (rounds launched one by
    one, HIP events around the FIND-OR-PUT launches), against the models'
    plain runs (plain_runs).  Counts must be equal."""
    runs = []
    for _ in range(3):
        res, stats, rc = gx.explore_batch(models, props, weights=weights, q_ints=q_batch, **opts)
        if rc != 0 or any(r["table_full"] for r in res):
            raise RuntimeError(f"batch failed (status {rc}) or a partition filled up")
        runs.append(stats)
    t_batch = min(s["t_explore_s"] for s in runs)
    states = runs[0]["states"]
    tk = {k: v for k, v in opts.items() if k != "use_graph"}
    tres, tstats, rc = gx.explore_batch(models, props, weights=weights, q_ints=q_batch, time_kernels=1, **tk)
    if rc != 0 or [(r["states"], r["transitions"]) for r in tres] != [(r["states"], r["transitions"]) for r in res]:
        raise RuntimeError("the kernel-timed batch run differs")
    for i, r in enumerate(plain):
        if (r["states"], r["transitions"]) != (res[i]["states"], res[i]["transitions"]):
            raise RuntimeError(f"batched and plain counts differ for model {i}")
    t_plain = sum(r["t_explore_s"] for r in plain)
    log(f"batch: {states} states, batched {t_batch * 1e3:.1f} ms, plain {t_plain * 1e3:.1f} ms")
    return {
        "states": states,
        "transitions": runs[0]["transitions"],
        "batched_new_states_per_s": states / t_batch,
        "plain_new_states_per_s": states / t_plain,
        "ratio": t_plain / t_batch,
        "t_batched_s": t_batch,
        "t_plain_s": t_plain,
        "options": opts,
        "rounds": runs[0]["rounds"],
        "graph_launches": runs[0]["graph_launches"],
        "batch_insert_roofline": insert_roofline(tstats, len(models), opts, "k_batch_insert", prof_file),
    }


def batched_vs_plain(gx, gen, tmp: str, legs: str = "both", tune: bool = True) -> dict:
    """config 4 (64 models of uneven size, gen.config4_models seeds 1000-1063,
    half -d and half -p; options autotuned) and a batch at the large plain
    workload's scale (gen.config2_batch: 64 config-2-generator models of
    14-17 processes, about 1e9 states together, no property; partitions
    sized from the plain runs' counts) against one exploration per model.
    legs: "both", "config4" or "large"; tune=False takes the library's
    default options for config 4 and 16384 FIND-OR-PUT workgroups for the
    large batch (for profiling runs)."""
    from gpuexplore_amd import autotune

    out = {}
    if legs in ("both", "config4"):
        out.update(config4_leg(gx, gen, tmp, tune))
    if legs in ("both", "large"):
        out["large"] = large_leg(gx, gen, tmp, tune)
    return out


def config4_leg(gx, gen, tmp: str, tune: bool) -> dict:
    from gpuexplore_amd import autotune

    paths, props = [], []
    for i, (net, flag) in enumerate(gen.config4_models(1000, 64)):
        p = os.path.join(tmp, f"c4_{i}")
        net.write(p)
        paths.append(p)
        props.append(gx.PROP_DEADLOCK if flag == "-d" else gx.PROP_SAFETY)
    models = [gx.Model.load(p) for p in paths]
    q_batch = 64 << 24  # 64 MiB of visited set per instance
    log("config 4 batch: autotuning" if tune else "config 4 batch")
    best = autotune.tune_batch(models, props, q_batch, repeats=2)["best"] if tune else {"levels_per_sync": 16}
    plain = plain_runs(gx, models, props, 1 << 24)
    out = {"workload": "config4: 64 models, N 4..14, 4..32 states, 32 -d + 32 -p", "autotuned": tune}
    out.update(batch_leg(gx, models, props, q_batch, best, plain,
                         prof_file="k_batch_insert_config4_traffic.json"))
    return out


def large_leg(gx, gen, tmp: str, tune: bool) -> dict:
    from gpuexplore_amd import autotune

    nets = gen.config2_batch()
    lm = []
    for i, net in enumerate(nets):
        p = os.path.join(tmp, f"c2b_{i}")
        net.write(p)
        lm.append(gx.Model.load(p))
    lprops = [gx.PROP_NONE] * len(lm)
    log("large batch: plain path")
    # plain explorers of a 2 GiB visited set each, with bounded frontier and
    # successor buffers (a larger level is deferred, as in any exploration)
    plain = plain_runs(gx, lm, lprops, 1 << 29, frontier_capacity=1 << 26, cand_capacity=1 << 28)
    # batch partitions in proportion to each model's state count, at most half full
    counts = [r["states"] for r in plain]
    q_large = max(1 << 28, 2 * 4 * sum(counts))  # 16-byte slots (4 ints), load <= 1/2
    log("large batch: autotuning" if tune else "large batch")
    if tune:
        best = autotune.tune_batch(lm, lprops, q_large, space={"insert_blocks": [0, 16384], "levels_per_sync": [16]},
                                   repeats=1, weights=counts)["best"]
    else:
        best = {"insert_blocks": 16384, "levels_per_sync": 16}
    leg = batch_leg(gx, lm, lprops, q_large, best, plain, weights=counts,
                    prof_file="k_batch_insert_config2batch_traffic.json")
    leg["workload"] = ("64 config-2-generator models (gen.config2_batch: 14-17 processes x 16 states, "
                       "token budgets, no property)")
    leg["table_bytes"] = 4 * q_large
    return leg


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
    ap.add_argument("--no-batch", action="store_true", help="skip the batched vs plain comparisons")
    ap.add_argument("--batch", choices=("both", "config4", "large"), default="both",
                    help="batched legs: config 4 and/or 64 config-2-generator models (large)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="batched legs with fixed options (profiling runs)")
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
        # the level passes' instantiation (the initial state's one-wave pass
        # may be another one)
        ins = [v for k, v in pj["kernels"].items() if "k_insert" in k]
        if ins:
            v = max(ins, key=lambda v: v["level_launches"])
            traffic = v["bytes_per_launch"]
            physical = v.get("dram_bytes_per_launch")
            atomics = v.get("atomics_per_launch")

    del ex  # the config 2 buffers (a 2^34-int visited set) are not needed any more
    batch = None if args.no_batch else batched_vs_plain(gx, gen, tmp, args.batch, not args.no_autotune)
    por = None if args.no_por else por_reduction(gx, gen, tmp)

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
        print(f"CPU baseline: oracle BFS on {threads} threads", file=sys.stderr, flush=True)
        mt = om.explore_mt(threads, oracle.PROP_NONE, max_states=args.cpu_sample_states,
                           expected_states=states if full_cpu else args.cpu_sample_states, progress=True)
        print("CPU baseline: one thread, first levels", file=sys.stderr, flush=True)
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
