"""Full-size workloads on the GPU: config2 against the oracle's committed
counts, config3 under POR (with and without the cycle proviso) against the
oracle run live."""

import json
import os
import sys

import pytest

from gpuexplore_amd import gen

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_config2_full_size(gpu, tmp_path):
    g = json.load(open(os.path.join(GOLDEN, "config2_counts.json")))["runs"]["plain"]
    p = str(tmp_path / "config2")
    gen.config2(1).write(p)
    say("config2 full size")
    r = gpu.explore(gpu.Model.load(p), q_ints=1 << 32)
    assert r["table_full"] == 0
    assert (r["states"], r["transitions"], r["levels"], bool(r["violation"])) == \
        (g["states"], g["transitions"], g["levels"], g["violation"])


@pytest.mark.parametrize("prop,proviso", [(0, False), (0, True), (1, False), (1, True)],
                         ids=["por", "por-proviso", "por-d", "por-proviso-d"])
def test_config3_por(gpu, oracle_mod, tmp_path, prop, proviso):
    p = str(tmp_path / "config3")
    gen.config3(3).write(p)
    o = oracle_mod.explore(p, prop, por=True, proviso=proviso)
    say("config3", prop, proviso, o["states"])
    r = gpu.explore(gpu.Model.load(p, por=True), property=prop, por=1,
                    cycle_proviso=int(proviso), q_ints=1 << 26)
    assert (r["states"], r["transitions"], r["levels"], bool(r["violation"])) == \
        (o["states"], o["transitions"], o["levels"], o["violation"])


def test_config3_no_por(gpu, tmp_path):
    """the unreduced space (945M states) against the oracle's counts of its
    first levels: the oracle cannot hold the whole set in host memory here"""
    path = os.path.join(GOLDEN, "config3_counts.json")
    if not os.path.exists(path):
        pytest.skip("no full-size counts for config3")
    runs = json.load(open(path))["runs"]
    p = str(tmp_path / "config3")
    gen.config3(3).write(p)
    for name, g in runs.items():
        if "--por" in name:
            continue
        prop = 1 if "-d" in name.split() else 0
        say("config3 no POR", name)
        r = gpu.explore(gpu.Model.load(p), property=prop, q_ints=1 << 34,
                        max_states=g.get("max_states", 0))
        assert r["table_full"] == 0
        assert bool(r["truncated"]) == g.get("truncated", False) or not g.get("max_states")
        assert (r["states"], r["transitions"], r["levels"], bool(r["violation"])) == \
            (g["states"], g["transitions"], g["levels"], g["violation"])
