"""rocprofv3 collection and the summaries kept under profiles/.

Collects separate runs of one command -- a kernel trace with statistics,
then two counter passes (counters are never combined with tracing, and each
pass stays within the four TCC counters one run can hold) -- and summarises
them:

    python -m gpuexplore_amd.profsum collect --out /abs/dir -- \\
        python3 $PWD/bench.py --steps 1 --warmup 0 --no-cpu --no-batch --no-por --serve-seconds 0
    python -m gpuexplore_amd.profsum summarize --dir /abs/dir --name k_insert_config2

Bytes moved between the L2 (TCC) and memory are counted from the L2's
memory-side requests by size, with no scale factor:

    read  = 32 x TCC_EA0_RDREQ_32B + 64 x TCC_EA0_RDREQ_64B + 128 x TCC_EA0_RDREQ_128B
    write = 64 x TCC_EA0_WRREQ_64B + 32 x (TCC_EA0_WRREQ - TCC_EA0_WRREQ_64B)

The same process runs kernels of known bytes per dispatch (bench.py's
streaming copy and read, 4 GiB each, and 2^29 random 64-byte bucket reads);
their readings are reported beside the known bytes, which validates the
accounting.  (The derived FETCH_SIZE counts a 128-byte request as 64 bytes
unless the L2 also counts it in TCC_BUBBLE, so it read half of a streaming
copy's reads.)  Requests to memory may be served by the MALL (Infinity
Cache) or by HBM; the L2 cannot tell them apart.

TCC_ATOMIC counts every atomic request the L2 serves (CAS claims, frontier
and counter additions).  Algorithmic bytes per FIND-OR-PUT launch come from
the bench line the kernel-trace run printed (128 B x (lookups + inserts)).
"""

from __future__ import annotations

import argparse
import csv
import glob
import json
import os
import subprocess
import sys
from collections import defaultdict
from typing import Dict, List

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RD = ["TCC_EA0_RDREQ_sum", "TCC_EA0_RDREQ_32B_sum", "TCC_EA0_RDREQ_64B_sum", "TCC_EA0_RDREQ_128B_sum"]
WR = ["TCC_EA0_WRREQ_sum", "TCC_EA0_WRREQ_64B_sum", "TCC_ATOMIC_sum", "TCC_BUBBLE_sum"]
PASSES = [("kt", ["--kernel-trace", "--stats"]), ("rd", ["--pmc", *RD]), ("wr", ["--pmc", *WR])]
# known bytes per dispatch of bench.py's measurement kernels: (read, written)
KNOWN = {"k_copy": (4 << 30, 4 << 30), "k_read": (4 << 30, 0), "k_random_read": (64 << 29, 0)}
BENCH_CMD = "python3 $PWD/bench.py --steps 1 --warmup 0 --no-cpu --no-batch --no-por --serve-seconds 0"


def collect(out: str, cmd: List[str], limit: int) -> int:
    env = dict(os.environ, TMPDIR="/tmp")
    os.makedirs(out, exist_ok=True)
    for name, opts in PASSES:
        print(f"profile pass {name}", flush=True)
        full = ["timeout", "-s", "KILL", str(limit), "rocprofv3", *opts, "-d", os.path.join(out, name), "-o", name,
                "--output-format", "csv", "--", *cmd]
        with open(os.path.join(out, f"{name}.stdout"), "w") as so, open(os.path.join(out, f"{name}.stderr"), "w") as se:
            rc = subprocess.run(full, cwd="/tmp", env=env, stdout=so, stderr=se).returncode
        lines = open(os.path.join(out, f"{name}.stdout")).read().splitlines()
        done = any(ln.startswith("{") for ln in lines)
        if rc != 0 and not (rc in (139, -11) and done):  # the profiler may segfault after a complete run
            print(f"pass {name} failed with {rc}", file=sys.stderr)
            return rc or 1
        print(f"profile pass {name} done", flush=True)
    return 0


def _one(pattern: str) -> str:
    hits = sorted(glob.glob(pattern, recursive=True))
    if not hits:
        raise FileNotFoundError(pattern)
    return hits[0]


def _short(name: str) -> str:
    """kernel name without its argument list and return type"""
    n = name.split("(")[0].strip()
    return n[5:] if n.startswith("void ") else n


def _base(name: str) -> str:
    """kernel name without namespace and template arguments"""
    return _short(name).split("<")[0].split("::")[-1]


def _counters(path: str):
    """per kernel: {counter: total}, and dispatches per kernel"""
    per = defaultdict(lambda: defaultdict(float))
    disp = defaultdict(set)
    for r in csv.DictReader(open(path)):
        k = _short(r["Kernel_Name"])
        per[k][r["Counter_Name"]] += float(r["Counter_Value"])
        disp[k].add(r.get("Dispatch_Id", r.get("Correlation_Id", "")))
    return per, {k: len(v) for k, v in disp.items()}


def traffic(rd: Dict[str, float], wr: Dict[str, float]) -> Dict[str, float]:
    read = 32 * rd["TCC_EA0_RDREQ_32B_sum"] + 64 * rd["TCC_EA0_RDREQ_64B_sum"] + 128 * rd["TCC_EA0_RDREQ_128B_sum"]
    write = 64 * wr["TCC_EA0_WRREQ_64B_sum"] + 32 * (wr["TCC_EA0_WRREQ_sum"] - wr["TCC_EA0_WRREQ_64B_sum"])
    unsized = rd["TCC_EA0_RDREQ_sum"] - (rd["TCC_EA0_RDREQ_32B_sum"] + rd["TCC_EA0_RDREQ_64B_sum"] +
                                         rd["TCC_EA0_RDREQ_128B_sum"])
    fetch_size = (wr["TCC_BUBBLE_sum"] * 128 + 32 * rd["TCC_EA0_RDREQ_32B_sum"] +
                  64 * (rd["TCC_EA0_RDREQ_sum"] - wr["TCC_BUBBLE_sum"] - rd["TCC_EA0_RDREQ_32B_sum"]))
    return {"read_bytes": read, "write_bytes": write, "atomics": wr["TCC_ATOMIC_sum"],
            "read_requests": rd["TCC_EA0_RDREQ_sum"], "read_requests_unsized": unsized,
            "req_32b": rd["TCC_EA0_RDREQ_32B_sum"], "req_64b": rd["TCC_EA0_RDREQ_64B_sum"],
            "req_128b": rd["TCC_EA0_RDREQ_128B_sum"], "fetch_size_expr_bytes": fetch_size}


def _get(obj: dict, path: str):
    for k in path.split("."):
        obj = obj[k]
    return obj


def summarize(d: str, name: str, kernel: str, out_dir: str, roof_path: str = "roofline",
              cmd: str = BENCH_CMD) -> Dict:
    """Summary of one collected directory for `kernel`: its kernel-trace
    average, its bytes to and from memory, and the algorithmic bytes per
    launch from the bench line's roofline object at `roof_path` (the
    single-model FIND-OR-PUT's by default; a batch leg's
    batch_insert_roofline for the batched one)."""
    bench = None
    for ln in open(os.path.join(d, "kt.stdout")).read().splitlines():
        if ln.startswith("{"):
            bench = json.loads(ln)
    if bench is None:
        raise RuntimeError("no bench line in the kernel-trace run")
    roof = _get(bench, roof_path)
    workload = bench["config"]["workload"] if roof_path == "roofline" else \
        _get(bench, roof_path.rsplit(".", 1)[0])["workload"]
    stats = list(csv.DictReader(open(_one(os.path.join(d, "kt", "**", "*kernel_stats.csv")))))
    trace = list(csv.DictReader(open(_one(os.path.join(d, "kt", "**", "*kernel_trace.csv")))))

    def grid(r):
        return int(r.get("Grid_Size", r.get("Grid_Size_X", "0")) or 0)

    # level launches (the single-model initial FIND-OR-PUT of s0 is one wave)
    level = [r for r in trace if kernel in r["Kernel_Name"] and grid(r) > 64]
    durs = [int(r["End_Timestamp"]) - int(r["Start_Timestamp"]) for r in level]
    rd, rdisp = _counters(_one(os.path.join(d, "rd", "**", "*counter_collection.csv")))
    wr, wdisp = _counters(_one(os.path.join(d, "wr", "**", "*counter_collection.csv")))
    kern, check = {}, {}
    for k in rd:
        if k not in wr:
            continue
        t = traffic(rd[k], wr[k])
        n = rdisp[k]
        base = _base(k)
        if base in KNOWN:
            kr, kw = KNOWN[base]
            check[k] = {"dispatches": n, "known_read_bytes": kr, "known_write_bytes": kw,
                        "read_bytes": t["read_bytes"] / n, "write_bytes": t["write_bytes"] / n,
                        "read_ratio": t["read_bytes"] / n / kr,
                        "write_ratio": (t["write_bytes"] / n / kw) if kw else None,
                        "fetch_size_expr_ratio": t["fetch_size_expr_bytes"] / n / kr,
                        "req_64b_per_dispatch": t["req_64b"] / n, "req_128b_per_dispatch": t["req_128b"] / n}
        if "expand" in k or "insert" in k:
            n_level = sum(1 for r in trace if _short(r["Kernel_Name"]) == k and grid(r) > 64) or n
            t_ns = sum(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]) for r in trace
                       if _short(r["Kernel_Name"]) == k)
            kern[k] = {"dispatches": n, "level_launches": n_level,
                       # every dispatch's bytes over every dispatch's time (the
                       # trace run's): launches that find no work weigh nothing
                       "gbps": (t["read_bytes"] + t["write_bytes"]) / t_ns if t_ns else None,
                       "kernel_ms": t_ns / 1e6,
                       "read_bytes": t["read_bytes"], "write_bytes": t["write_bytes"],
                       "bytes_per_launch": (t["read_bytes"] + t["write_bytes"]) / n_level,
                       "dram_bytes_per_launch": (t["read_bytes"] + t["write_bytes"]) / n_level,
                       "atomics_per_launch": t["atomics"] / n_level,
                       "read_requests_per_launch": t["read_requests"] / n_level,
                       "req_32b": t["req_32b"], "req_64b": t["req_64b"], "req_128b": t["req_128b"],
                       "read_requests_unsized": t["read_requests_unsized"]}
    avg_us = sum(durs) / len(durs) / 1e3 if durs else None
    data = {"workload": workload,
            "command": f"rocprofv3 --pmc {' '.join(RD)} | --pmc {' '.join(WR)} -- {cmd}",
            "accounting": "read = 32 RDREQ_32B + 64 RDREQ_64B + 128 RDREQ_128B; "
                          "write = 64 WRREQ_64B + 32 (WRREQ - WRREQ_64B); no scale factor",
            "validation": check, "kernels": kern,
            "options": None if roof_path == "roofline" else _get(bench, roof_path.rsplit(".", 1)[0]).get("options"),
            "algorithmic_bytes_per_launch": roof["algorithmic_bytes_per_launch"],
            "trace_avg_launch_us": avg_us, "bench_avg_launch_us": roof["avg_launch_us"]}
    os.makedirs(out_dir, exist_ok=True)
    with open(os.path.join(out_dir, f"{name}_traffic.json"), "w") as f:
        json.dump(data, f, indent=1)
    extra = "" if kernel == "k_insert" else f" --kernel {kernel}"
    extra += "" if roof_path == "roofline" else f" --roof {roof_path} --cmd '{cmd}'"
    md = [f"# {kernel} on {workload}", "",
          "From the repository root (each profiler pass runs in /tmp with TMPDIR=/tmp; <dir> absolute):", "",
          "```", f"python3 -m gpuexplore_amd.profsum collect --out <dir> -- {cmd}",
          f"python3 -m gpuexplore_amd.profsum summarize --dir <dir> --name {name}{extra}", "```", "",
          "Kernel time, `rocprofv3 --kernel-trace --stats`, the whole command:", "",
          "| kernel | calls | total ms | average us |", "|---|---|---|---|"]
    for r in sorted(stats, key=lambda r: -float(r["TotalDurationNs"]))[:10]:
        md.append(f"| `{r['Name'][:60]}` | {r['Calls']} | {float(r['TotalDurationNs']) / 1e6:.2f} | "
                  f"{float(r['AverageNs']) / 1e3:.1f} |")
    md += ["", f"{kernel} average over its {len(durs)} launches (a one-wave initial launch excluded): "
               f"{avg_us:.0f} us (kernel trace); bench.py's HIP-event average: "
               f"{roof['avg_launch_us']:.0f} us.", "",
           "Bytes between L2 and memory, from two counter runs of the same command "
           f"(`--pmc {' '.join(RD)}`, `--pmc {' '.join(WR)}`): read = 32 x RDREQ_32B + 64 x RDREQ_64B + "
           "128 x RDREQ_128B, write = 64 x WRREQ_64B + 32 x the other write requests, no scale factor.", "",
           "Validation on kernels of known bytes in the same process:", "",
           "| kernel | dispatches | known read GB | counted read GB | ratio | known write GB | counted write GB | "
           "FETCH_SIZE expression / known read |", "|---|---|---|---|---|---|---|---|"]
    for k, v in sorted(check.items()):
        md.append(f"| `{k}` | {v['dispatches']} | {v['known_read_bytes'] / 1e9:.2f} | {v['read_bytes'] / 1e9:.2f} | "
                  f"{v['read_ratio']:.3f} | {v['known_write_bytes'] / 1e9:.2f} | {v['write_bytes'] / 1e9:.2f} | "
                  f"{v['fetch_size_expr_ratio']:.3f} |")
    md += ["", "| kernel | launches | read GB | write GB | GB per launch | 64 B / 128 B read requests | "
               "atomics per launch |", "|---|---|---|---|---|---|---|"]
    for k, v in kern.items():
        md.append(f"| `{k}` | {v['level_launches']} | {v['read_bytes'] / 1e9:.1f} | {v['write_bytes'] / 1e9:.1f} | "
                  f"{v['bytes_per_launch'] / 1e9:.2f} | {v['req_64b'] / 1e9:.2f} G / {v['req_128b'] / 1e9:.2f} G | "
                  f"{v['atomics_per_launch'] / 1e6:.1f} M |")
    # the level passes' instantiation: the one with the most level launches
    # (the single-model initial state goes through another, one-wave one)
    ki = max((v for k, v in kern.items() if kernel in k), key=lambda v: v["level_launches"], default=None)
    if ki and avg_us:
        phys = ki["gbps"]
        md += ["", f"Algorithmic bytes per {kernel} launch (128 B x (lookups + inserts)): "
                   f"{roof['algorithmic_bytes_per_launch'] / 1e9:.2f} GB, {roof['achieved']:.0f} GB/s = "
                   f"{100 * roof['frac']:.1f} % of {roof['peak']:.0f} GB/s.  Counted bytes to and from memory: "
                   f"{ki['bytes_per_launch'] / 1e9:.2f} GB per launch; over the kernel's own time (every launch's "
                   f"bytes over every launch's duration, {ki['kernel_ms']:.1f} ms) {phys:.0f} GB/s = "
                   f"{100 * phys / roof['peak']:.1f} % of the peak; atomics "
                   f"{ki['atomics_per_launch'] * ki['level_launches'] / (ki['kernel_ms'] * 1e-3) / 1e9:.2f} G/s.", ""]
    with open(os.path.join(out_dir, f"{name}.md"), "w") as f:
        f.write("\n".join(md))
    return data


# SQ instruction and cycle counters of the level kernels, in passes of at
# most eight SQ counters each
SQ_PASSES = [
    ["SQ_WAVES", "SQ_INSTS_VALU", "SQ_INSTS_LDS", "SQ_INSTS_SALU", "SQ_INSTS_VALU_INT64", "SQ_INSTS_VALU_INT32",
     "SQ_ACTIVE_INST_VALU", "SQ_BUSY_CYCLES"],
    ["SQ_THREAD_CYCLES_VALU", "SQ_WAVE_CYCLES", "SQ_WAIT_INST_LDS", "SQ_ACTIVE_INST_LDS", "SQ_INSTS_BRANCH",
     "SQ_INSTS_VMEM_WR", "SQ_INSTS_VMEM_RD", "SQ_CYCLES"],
    ["SQ_LDS_BANK_CONFLICT", "SQ_LDS_IDX_ACTIVE", "SQ_WAIT_ANY", "SQ_INSTS_SMEM"],
]


def collect_sq(out: str, cmd: List[str], limit: int) -> int:
    env = dict(os.environ, TMPDIR="/tmp")
    os.makedirs(out, exist_ok=True)
    for i, counters in enumerate(SQ_PASSES):
        name = f"sq{i + 1}"
        print(f"profile pass {name}", flush=True)
        full = ["timeout", "-s", "KILL", str(limit), "rocprofv3", "--pmc", *counters, "-d", os.path.join(out, name),
                "-o", name, "--output-format", "csv", "--", *cmd]
        with open(os.path.join(out, f"{name}.stdout"), "w") as so, open(os.path.join(out, f"{name}.stderr"), "w") as se:
            rc = subprocess.run(full, cwd="/tmp", env=env, stdout=so, stderr=se).returncode
        done = any(ln.startswith("{") for ln in open(os.path.join(out, f"{name}.stdout")).read().splitlines())
        if rc != 0 and not (rc in (139, -11) and done):
            print(f"pass {name} failed with {rc}", file=sys.stderr)
            return rc or 1
    return 0


def summarize_sq(d: str, name: str, out_dir: str, kernels=("k_expand_tps", "k_insert"), cmd: str = BENCH_CMD) -> Dict:
    """SQ counters of the two level kernels summed over their launches, with
    per-unit figures: per 64-state batch (successor kernel) and per
    16-candidate iteration (FIND-OR-PUT), from the bench line's counts."""
    bench = None
    for ln in open(os.path.join(d, "sq1.stdout")).read().splitlines():
        if ln.startswith("{"):
            bench = json.loads(ln)
    tot = defaultdict(lambda: defaultdict(float))
    for i in range(len(SQ_PASSES)):
        per, _ = _counters(_one(os.path.join(d, f"sq{i + 1}", "**", "*counter_collection.csv")))
        for k, v in per.items():
            for c, x in v.items():
                tot[k][c] += x
    # per kernel, the instantiation with the most waves (the level passes)
    names = {kk: max((k for k in tot if kk in k and "<" in k), key=lambda k: tot[k].get("SQ_WAVES", 0), default=None)
             for kk in kernels}
    states = bench["config"]["states"]
    cand = bench["hash_phase"]["candidates"]
    batches = states / 64.0
    iters = cand / 16.0
    cus, clk = 256, 2.4e9
    md = [f"# SQ counters: {' and '.join(kernels)} on {bench['config']['workload']}", "",
          f"{len(SQ_PASSES)} `rocprofv3 --pmc` passes of one exploration (counters summed over each kernel's "
          "launches), from the repository root:", "", "```",
          f"python3 -m gpuexplore_amd.profsum collect-sq --out <dir> -- {cmd}",
          f"python3 -m gpuexplore_amd.profsum summarize-sq --dir <dir> --name {name}", "```", "",
          "| counter | " + " | ".join(f"`{names[k]}`" for k in kernels) + " |",
          "|---|" + "---|" * len(kernels)]
    allc = sorted({c for k in kernels if names[k] for c in tot[names[k]]})
    for c in allc:
        md.append(f"| {c} | " + " | ".join(f"{tot[names[k]].get(c, 0):.4g}" if names[k] else "-" for k in kernels) + " |")
    derived = {}
    e = tot[names[kernels[0]]] if names[kernels[0]] else {}
    f = tot[names[kernels[1]]] if names[kernels[1]] else {}
    if e:
        derived["expand_per_batch"] = {c: e.get(c, 0) / batches for c in
                                       ("SQ_INSTS_VALU", "SQ_INSTS_SALU", "SQ_INSTS_LDS", "SQ_INSTS_BRANCH")}
        derived["expand_valu_ms"] = e.get("SQ_INSTS_VALU", 0) / (cus * clk) * 1e3
        derived["expand_lanes_active"] = (e.get("SQ_THREAD_CYCLES_VALU", 0) / e["SQ_ACTIVE_INST_VALU"]
                                          if e.get("SQ_ACTIVE_INST_VALU") else None)
        derived["expand_lds_bank_conflict_share"] = (e.get("SQ_LDS_BANK_CONFLICT", 0) / e["SQ_LDS_IDX_ACTIVE"]
                                                     if e.get("SQ_LDS_IDX_ACTIVE") else None)
    if f:
        derived["insert_per_iteration"] = {c: f.get(c, 0) / iters for c in
                                           ("SQ_INSTS_VALU", "SQ_INSTS_SALU", "SQ_INSTS_LDS")}
        derived["insert_valu_ms"] = f.get("SQ_INSTS_VALU", 0) / (cus * clk) * 1e3
    md += ["", f"Derived ({states:,} states, {cand:,} successors; a wave instruction issues at one per CU cycle, "
               f"{cus} CUs x {clk / 1e9:.1f} GHz):", ""]
    if e:
        pb = derived["expand_per_batch"]
        md.append(f"- {kernels[0]} per 64-state batch: {pb['SQ_INSTS_VALU']:,.0f} VALU, {pb['SQ_INSTS_SALU']:,.0f} SALU, "
                  f"{pb['SQ_INSTS_LDS']:,.0f} LDS, {pb['SQ_INSTS_BRANCH']:,.0f} branch instructions; "
                  f"{derived['expand_valu_ms']:.1f} ms of VALU issue; "
                  + (f"{derived['expand_lanes_active']:.1f} of 64 lanes active per VALU instruction; "
                     if derived["expand_lanes_active"] else "")
                  + (f"LDS bank conflict cycles {100 * derived['expand_lds_bank_conflict_share']:.0f} % of LDS "
                     "cycles (SQ_LDS_BANK_CONFLICT / SQ_LDS_IDX_ACTIVE)." if derived["expand_lds_bank_conflict_share"]
                     is not None else ""))
    if f:
        pi = derived["insert_per_iteration"]
        md.append(f"- {kernels[1]} per 16-candidate iteration (4 lanes per candidate): {pi['SQ_INSTS_VALU']:.0f} VALU, "
                  f"{pi['SQ_INSTS_SALU']:.0f} SALU, {pi['SQ_INSTS_LDS']:.1f} LDS instructions; "
                  f"{derived['insert_valu_ms']:.1f} ms of VALU issue.")
    md.append(f"- Kernel times without counters (the bench line): {kernels[0]} "
              f"{1e3 * bench['hash_phase']['t_expand_s_per_step']:.1f} ms, {kernels[1]} "
              f"{1e3 * bench['hash_phase']['t_hash_s_per_step']:.1f} ms per exploration.")
    os.makedirs(out_dir, exist_ok=True)
    with open(os.path.join(out_dir, f"{name}.md"), "w") as fh:
        fh.write("\n".join(md) + "\n")
    return {"counters": {k: dict(tot[names[k]]) for k in kernels if names[k]}, "derived": derived}


def gaps(d: str, prefix: str = "k_batch", split_ns: int = 2_000_000) -> List[Dict]:
    """Idle time between the batch's kernels in the kernel-trace run: the
    launches whose names start with `prefix`, split into explorations where
    consecutive launches are more than `split_ns` apart (host work between
    explorations).  Per exploration: kernel busy time, span from the first
    start to the last end, and the gaps (what one persistent kernel could
    at most save)."""
    trace = list(csv.DictReader(open(_one(os.path.join(d, "kt", "**", "*kernel_trace.csv")))))
    rows = sorted((r for r in trace if _base(r["Kernel_Name"]).startswith(prefix)),
                  key=lambda r: int(r["Start_Timestamp"]))
    runs, cur = [], []
    for r in rows:
        if cur and int(r["Start_Timestamp"]) - int(cur[-1]["End_Timestamp"]) > split_ns:
            runs.append(cur)
            cur = []
        cur.append(r)
    if cur:
        runs.append(cur)
    out = []
    for run in runs:
        busy = sum(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]) for r in run)
        span = int(run[-1]["End_Timestamp"]) - int(run[0]["Start_Timestamp"])
        g = [int(b["Start_Timestamp"]) - int(a["End_Timestamp"]) for a, b in zip(run, run[1:])]
        out.append({"launches": len(run), "busy_ms": busy / 1e6, "span_ms": span / 1e6,
                    "idle_ms": (span - busy) / 1e6, "idle_frac": (span - busy) / span if span else 0.0,
                    "gaps_over_20us": sum(1 for x in g if x > 20000),
                    "median_gap_us": sorted(g)[len(g) // 2] / 1e3 if g else 0.0})
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("collect")
    c.add_argument("--out", required=True)
    c.add_argument("--limit", type=int, default=300)
    c.add_argument("rest", nargs=argparse.REMAINDER)
    s = sub.add_parser("summarize")
    s.add_argument("--dir", required=True)
    s.add_argument("--name", required=True)
    s.add_argument("--kernel", default="k_insert")
    s.add_argument("--profiles", default=os.path.join(ROOT, "profiles"))
    s.add_argument("--roof", default="roofline", help="path of the roofline object in the bench line")
    s.add_argument("--cmd", default=BENCH_CMD, help="the profiled command, as recorded in the summary")
    cs = sub.add_parser("collect-sq")
    cs.add_argument("--out", required=True)
    cs.add_argument("--limit", type=int, default=300)
    cs.add_argument("rest", nargs=argparse.REMAINDER)
    ss = sub.add_parser("summarize-sq")
    ss.add_argument("--dir", required=True)
    ss.add_argument("--name", required=True)
    ss.add_argument("--profiles", default=os.path.join(ROOT, "profiles"))
    g = sub.add_parser("gaps")
    g.add_argument("--dir", required=True)
    g.add_argument("--prefix", default="k_batch")
    a = ap.parse_args(argv)
    if a.cmd == "collect-sq":
        rest = a.rest[1:] if a.rest and a.rest[0] == "--" else a.rest
        return collect_sq(os.path.abspath(a.out), rest, a.limit)
    if a.cmd == "summarize-sq":
        print(json.dumps(summarize_sq(os.path.abspath(a.dir), a.name, a.profiles)["derived"], indent=1))
        return 0
    if a.cmd == "gaps":
        for r in gaps(os.path.abspath(a.dir), a.prefix):
            print(json.dumps(r))
        return 0
    if a.cmd == "collect":
        rest = a.rest[1:] if a.rest and a.rest[0] == "--" else a.rest
        return collect(os.path.abspath(a.out), rest, a.limit)
    print(json.dumps(summarize(os.path.abspath(a.dir), a.name, a.kernel, a.profiles, a.roof, a.cmd), indent=1))
    return 0


if __name__ == "__main__":
    sys.exit(main())
