"""Seeded random networks on the GPU against the oracle in every mode: the
plain method and each POR method (cample, ample sets, stubborn sets) with
and without the cycle proviso, for no property, deadlock and safety.
Bit-exact counts, levels and verdicts."""

import os

import pytest

from gpuexplore_amd import gen
from test_gpu_parity import gpu_run, same, say

pytestmark = pytest.mark.gpu

MODES = [(False, False, 0)] + [(True, prov, m) for m in (0, 1, 2) for prov in (False, True)]


def random_net(seed: int):
    """4-9 processes from the throttled-ring and phased-rounds families,
    with a safety monitor on every third seed"""
    import random
    rng = random.Random(seed)
    n = rng.randint(4, 9)
    if seed % 2:
        mix = (rng.uniform(0.3, 0.8), rng.uniform(0.1, 0.4), rng.uniform(0.0, 0.2))
        mon = (0, 1, n - 1, 2) if seed % 3 == 0 else None
        return gen.throttled(n, rng.choice([4, 6, 8]), 2, rng.randint(1, 3), mix, 2, seed, monitor=mon)
    return gen.rounds(n, 8, 2, [1] * n, 3, seed, mix=(0.7, 0.2, 0.1), phased=True)


@pytest.mark.parametrize("seed", range(1, 25))
def test_random_network_all_modes(gpu, oracle_mod, model_dir, seed):
    p = os.path.join(model_dir, f"r{seed}")
    random_net(seed).write(p)
    for prop in (0, 1, 2):
        for por, prov, method in MODES:
            a = oracle_mod.explore(p, prop, por, prov, method=method)
            say("random", seed, prop, por, prov, method)
            b = gpu_run(gpu, p, prop, por, prov, por_method=method, q_ints=1 << 26)
            assert same(a, b), (seed, prop, por, prov, method, a, b)
