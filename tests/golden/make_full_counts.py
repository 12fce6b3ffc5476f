"""Writes tests/golden/<workload>_counts.json: the oracle's counts for a
full-size generated workload, with the SHA-256 of the model text so a change
of the generator is caught before a count comparison is made.

    python tests/golden/make_full_counts.py config2 [--flags -d ...] [--threads T --expect N]

A full-size run of the oracle takes tens of minutes (config2: ~35 min and
~10 GB of host memory on one core; config3 unreduced, 945 M states: ~26 GB
and about 20 min on 8 threads with --expect presizing the level-free visited
set); the GPU tests compare against the JSON.
"""

import argparse
import hashlib
import json
import os
import subprocess
import sys
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from gpuexplore_amd import gen  # noqa: E402

WORKLOADS = {"config2": lambda: gen.config2(1), "config3": lambda: gen.config3(3)}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("workload", choices=sorted(WORKLOADS))
    ap.add_argument("--flags", nargs="*", default=[])
    ap.add_argument("--max-states", type=int, default=0,
                    help="stop after the first level that reaches this many states")
    ap.add_argument("--threads", type=int, default=1, help="orc_explore_mt threads (plain method only)")
    ap.add_argument("--expect", type=int, default=0, help="presize the visited set for this many states")
    args = ap.parse_args()
    net = WORKLOADS[args.workload]()
    text = net.to_gpf()
    cli = os.path.join(ROOT, "oracle", "_build", "gx_oracle_cli")
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True)
    with tempfile.TemporaryDirectory() as d:
        base = os.path.join(d, args.workload)
        with open(base + ".gpf", "w") as f:
            f.write(text)
        extra = ["--max-states", str(args.max_states)] if args.max_states else []
        if args.threads > 1:
            extra += ["--threads", str(args.threads)]
        if args.expect:
            extra += ["--expect", str(args.expect)]
        out = subprocess.run([cli, base] + args.flags + extra, check=True, capture_output=True,
                             text=True).stdout
    kv = dict(tok.split("=") for tok in out.split())
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), f"{args.workload}_counts.json")
    data = json.load(open(path)) if os.path.exists(path) else {}
    data["gpf_sha256"] = hashlib.sha256(text.encode()).hexdigest()
    name = " ".join(args.flags) or "plain"
    if args.max_states:
        name += f" max_states={args.max_states}"
    data.setdefault("runs", {})[name] = {
        "states": int(kv["states"]), "transitions": int(kv["transitions"]),
        "levels": int(kv["levels"]), "violation": bool(int(kv["violation"])),
        "truncated": bool(int(kv["truncated"])), "max_states": args.max_states,
        "oracle_seconds": float(kv["seconds"]), "oracle_threads": args.threads}
    with open(path, "w") as f:
        json.dump(data, f, indent=1)
    print(path, data)
    return 0


if __name__ == "__main__":
    sys.exit(main())
