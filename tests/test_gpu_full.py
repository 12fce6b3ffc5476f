"""Full-size workloads on the GPU: config2 and config3 unreduced against the
oracle's committed full counts, config3 under POR (with and without the
cycle proviso) against the oracle run live."""

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
    say("config2 full size, 2^34-int visited set")
    r = gpu.explore(gpu.Model.load(p), q_ints=1 << 34)
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


@pytest.mark.parametrize("name", ["plain", "-d"])
def test_config3_no_por(gpu, tmp_path, name):
    """the unreduced space (944.9 M states, 15.25 G transitions) with and
    without the deadlock check, against the oracle's committed full counts
    (tests/golden/make_full_counts.py, the multi-threaded oracle with a
    presized visited set)"""
    g = json.load(open(os.path.join(GOLDEN, "config3_counts.json")))["runs"][name]
    assert not g.get("truncated")
    p = str(tmp_path / "config3")
    gen.config3(3).write(p)
    prop = 1 if name == "-d" else 0
    say("config3 no POR", name)
    r = gpu.explore(gpu.Model.load(p), property=prop, q_ints=1 << 34)
    assert r["table_full"] == 0 and r["truncated"] == 0
    assert (r["states"], r["transitions"], r["levels"], bool(r["violation"])) == \
        (g["states"], g["transitions"], g["levels"], g["violation"])
