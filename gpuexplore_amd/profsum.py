"""rocprofv3 collection and the summaries kept under profiles/.

Collect three separate runs of one command (kernel trace with statistics,
then the FETCH_SIZE counter, then WRITE_SIZE; counters are never combined
with tracing), then summarise them:

    python -m gpuexplore_amd.profsum collect --out /abs/dir -- \\
        python3 $PWD/bench.py --steps 1 --warmup 0 --no-cpu --no-batch --no-por --serve-seconds 0
    python -m gpuexplore_amd.profsum summarize --dir /abs/dir --name k_insert_config2

The counters are scaled by what they read for bench.py's streaming copy
(k_copy, 4 GiB read and 4 GiB written per dispatch) in the same process.
Algorithmic bytes per FIND-OR-PUT launch come from the bench line the
kernel-trace run printed (128 B x (lookups + inserts) per launch).
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
COPY_BYTES = 4 << 30  # per k_copy dispatch, read and written (bench.py)
PASSES = [("kt", ["--kernel-trace", "--stats"]), ("fetch", ["--pmc", "FETCH_SIZE"]),
          ("write", ["--pmc", "WRITE_SIZE"])]


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
    return 0


def _one(pattern: str) -> str:
    hits = sorted(glob.glob(pattern, recursive=True))
    if not hits:
        raise FileNotFoundError(pattern)
    return hits[0]


def _short(name: str) -> str:
    return name.split("(")[0].strip()


def _counters(path: str):
    per = defaultdict(float)
    calls = defaultdict(int)
    for r in csv.DictReader(open(path)):
        k = _short(r["Kernel_Name"])
        per[k] += float(r["Counter_Value"])
        calls[k] += 1
    return per, calls


def summarize(d: str, name: str, kernel: str, out_dir: str) -> Dict:
    bench = None
    for ln in open(os.path.join(d, "kt.stdout")).read().splitlines():
        if ln.startswith("{"):
            bench = json.loads(ln)
    if bench is None:
        raise RuntimeError("no bench line in the kernel-trace run")
    stats = list(csv.DictReader(open(_one(os.path.join(d, "kt", "**", "*kernel_stats.csv")))))
    trace = list(csv.DictReader(open(_one(os.path.join(d, "kt", "**", "*kernel_trace.csv")))))
    level = [r for r in trace if kernel in r["Kernel_Name"] and int(r.get("Grid_Size", r.get("Grid_Size_X", "0")) or 0) > 64]
    durs = [int(r["End_Timestamp"]) - int(r["Start_Timestamp"]) for r in level]
    fetch, fcalls = _counters(_one(os.path.join(d, "fetch", "**", "*counter_collection.csv")))
    write, wcalls = _counters(_one(os.path.join(d, "write", "**", "*counter_collection.csv")))
    copy = next(k for k in fetch if "k_copy" in k)
    fscale = COPY_BYTES / (fetch[copy] / fcalls[copy])
    wscale = COPY_BYTES / (write[copy] / wcalls[copy])
    kern = {}
    for k in fetch:
        if "k_expand" in k or "k_insert" in k or "k_fused" in k:
            n_level = sum(1 for r in trace if _short(r["Kernel_Name"]) == k and
                          int(r.get("Grid_Size", r.get("Grid_Size_X", "0")) or 0) > 64)
            n_level = n_level or fcalls[k]
            kern[k] = {"fetch_bytes": fetch[k] * fscale, "write_bytes": write.get(k, 0.0) * wscale,
                       "dispatches": fcalls[k], "level_launches": n_level,
                       "bytes_per_launch": (fetch[k] * fscale + write.get(k, 0.0) * wscale) / n_level}
    roof = bench["roofline"]
    cmd = "python3 $PWD/bench.py --steps 1 --warmup 0 --no-cpu --no-batch --no-por --serve-seconds 0"
    data = {"workload": bench["config"]["workload"], "command": f"rocprofv3 --pmc FETCH_SIZE|WRITE_SIZE -- {cmd}",
            "fetch_scale": fscale, "write_scale": wscale, "kernels": kern,
            "algorithmic_bytes_per_launch": roof["algorithmic_bytes_per_launch"],
            "trace_avg_launch_us": sum(durs) / len(durs) / 1e3 if durs else None,
            "bench_avg_launch_us": roof["avg_launch_us"]}
    os.makedirs(out_dir, exist_ok=True)
    with open(os.path.join(out_dir, f"{name}_traffic.json"), "w") as f:
        json.dump(data, f, indent=1)
    md = [f"# {kernel} on {bench['config']['workload']}", "",
          "Kernel time, `rocprofv3 --kernel-trace --stats`, one exploration:", "",
          "From the repository root (each profiler pass runs in /tmp with TMPDIR=/tmp; <dir> absolute):", "",
          "```", f"python3 -m gpuexplore_amd.profsum collect --out <dir> -- {cmd}",
          f"python3 -m gpuexplore_amd.profsum summarize --dir <dir> --name {name}", "```", "",
          "| kernel | calls | total ms | average us |", "|---|---|---|---|"]
    for r in sorted(stats, key=lambda r: -float(r["TotalDurationNs"]))[:8]:
        md.append(f"| `{r['Name'][:60]}` | {r['Calls']} | {float(r['TotalDurationNs']) / 1e6:.2f} | "
                  f"{float(r['AverageNs']) / 1e3:.1f} |")
    md += ["", f"{kernel} average over its {len(durs)} level launches (the 1-candidate initial launch excluded): "
               f"{data['trace_avg_launch_us']:.0f} us (kernel trace); bench.py's HIP-event average for the same "
               f"launches: {roof['avg_launch_us']:.0f} us.", "",
           "HBM traffic, `rocprofv3 --pmc FETCH_SIZE` and `--pmc WRITE_SIZE` runs of the same command, scaled by "
           "the counters' reading of bench.py's 4 GiB streaming copy (k_copy) in the same process:", "",
           f"- counters read KiB: FETCH_SIZE x {fscale / 1024:.3f} KiB, WRITE_SIZE x {wscale / 1024:.3f} KiB per "
           "count gives bytes (FETCH_SIZE reads half of what the copy fetches).",
           f"- measured streaming copy (bench.py): {bench['roofline'].get('measured_copy_peak_gbps', 0):.0f} GB/s.", "",
           "| kernel | level launches | fetch GB | write GB | GB per level launch |", "|---|---|---|---|---|"]
    for k, v in kern.items():
        md.append(f"| `{k}` | {v['level_launches']} | {v['fetch_bytes'] / 1e9:.1f} | {v['write_bytes'] / 1e9:.1f} | "
                  f"{v['bytes_per_launch'] / 1e9:.2f} |")
    md += ["", f"Algorithmic bytes per {kernel} launch (128 B x (lookups + inserts) per launch): "
               f"{roof['algorithmic_bytes_per_launch'] / 1e9:.2f} GB; achieved {roof['achieved']:.0f} GB/s = "
               f"{100 * roof['frac']:.1f} % of {roof['peak']:.0f} GB/s.", ""]
    with open(os.path.join(out_dir, f"{name}.md"), "w") as f:
        f.write("\n".join(md))
    return data


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
    a = ap.parse_args(argv)
    if a.cmd == "collect":
        rest = a.rest[1:] if a.rest and a.rest[0] == "--" else a.rest
        return collect(os.path.abspath(a.out), rest, a.limit)
    print(json.dumps(summarize(os.path.abspath(a.dir), a.name, a.kernel, a.profiles), indent=1))
    return 0


if __name__ == "__main__":
    sys.exit(main())
