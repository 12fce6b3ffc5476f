"""Autotuner for the batched path.

Sweeps threads per workgroup, rounds captured per graph replay (the
reference's kernel_iters), hash-bucket width and the FIND-OR-PUT grid on a
given batch, times
each setting with gx_explore_batch, and returns the fastest.  Every setting
must produce identical per-instance counts: a setting that does not is a
bug, and the sweep raises instead of picking it.
"""

from __future__ import annotations

import itertools
import json
import os
from typing import Dict, List, Optional, Sequence

import gpuexplore_amd as gx

DEFAULT_SPACE = {
    "threads": [128, 256, 512],
    "levels_per_sync": [1, 4, 16],
    "bucket_bytes": [64, 128],
    "insert_blocks": [0, 1024, 4096],  # 0: 8 per CU
}


def _signature(res: List[dict]) -> tuple:
    return tuple((r["states"], r["transitions"], r["levels"], r["violation"]) for r in res)


def tune_batch(models: Sequence["gx.Model"], properties: Sequence[int], q_ints: int,
               space: Optional[Dict[str, list]] = None, repeats: int = 2,
               cache_path: Optional[str] = None, key: Optional[str] = None, weights=None) -> dict:
    """Returns {"best": {option: value}, "t_explore_s": t, "table": [...]}.
    With cache_path and key, a previous result for the same key is reused.
    ``weights`` sizes the instances' visited-set partitions (gx_explore_batch)."""
    if cache_path and key and os.path.exists(cache_path):
        cached = json.load(open(cache_path))
        if key in cached:
            return cached[key]
    space = space or DEFAULT_SPACE
    names = list(space)
    table = []
    ref = None
    for values in itertools.product(*(space[n] for n in names)):
        opts = dict(zip(names, values))
        times = []
        for _ in range(repeats):
            res, stats, rc = gx.explore_batch(list(models), list(properties), weights=weights, q_ints=q_ints,
                                              **opts)
            if rc != 0:
                raise RuntimeError(f"batch failed with {opts}: {rc}")
            sig = _signature(res)
            if ref is None:
                ref = sig
            elif sig != ref:
                raise RuntimeError(f"counts differ under {opts}")
            times.append(stats["t_explore_s"])
        table.append({**opts, "t_explore_s": min(times)})
    best = min(table, key=lambda r: r["t_explore_s"])
    out = {"best": {n: best[n] for n in names}, "t_explore_s": best["t_explore_s"], "table": table}
    if cache_path and key:
        cached = json.load(open(cache_path)) if os.path.exists(cache_path) else {}
        cached[key] = out
        with open(cache_path, "w") as f:
            json.dump(cached, f, indent=1)
    return out
