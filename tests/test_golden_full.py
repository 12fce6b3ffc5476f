"""The committed full-size counts belong to the models the generator writes
today: a changed generator must regenerate them (tests/golden/make_full_counts.py)."""

import hashlib
import json
import os

import pytest

from gpuexplore_amd import gen

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
WORKLOADS = {"config2": lambda: gen.config2(1), "config3": lambda: gen.config3(3)}


@pytest.mark.parametrize("name", sorted(WORKLOADS))
def test_full_golden_matches_generator(name):
    path = os.path.join(GOLDEN, f"{name}_counts.json")
    if not os.path.exists(path):
        pytest.skip(f"no full-size counts for {name}")
    g = json.load(open(path))
    text = WORKLOADS[name]().to_gpf()
    assert hashlib.sha256(text.encode()).hexdigest() == g["gpf_sha256"]
    for run in g["runs"].values():
        assert run["states"] >= run["levels"] > 0
        assert run["transitions"] >= run["states"] - 1
