"""Ample-set and stubborn-set partial-order reduction on the GPU against the
oracle (the methods' own pins are in test_por_methods.py): identical
counts, levels and verdicts for every property, with and without the cycle
proviso, on the small models, the edge models, config 3 and through the
command line."""

import os
import subprocess
import sys

import pytest

from gpuexplore_amd import gen
from test_gpu_parity import KEYS, gpu_run, same, say
from test_oracle import small_models
from test_por_methods import METHODS, triad

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("method", ["ample", "stubborn"])
@pytest.mark.parametrize("name,net", small_models(), ids=[n for n, _ in small_models()])
def test_small_models(gpu, oracle_mod, model_dir, method, name, net):
    p = os.path.join(model_dir, name)
    net.write(p)
    for prop in (0, 1, 2):
        for prov in (False, True):
            say(method, name, prop, prov)
            a = oracle_mod.explore(p, prop, True, prov, method=METHODS[method])
            b = gpu_run(gpu, p, prop, True, prov, por_method=METHODS[method])
            assert same(a, b), (method, prop, prov, a, b)


@pytest.mark.parametrize("method", list(METHODS))
def test_closed_forms(gpu, model_dir, method):
    m = METHODS[method]
    p = os.path.join(model_dir, "triad")
    triad().write(p)
    say("triad", method)
    r = gpu_run(gpu, p, 0, True, False, por_method=m)
    assert (r["states"], r["transitions"]) == ((2, 2) if method == "stubborn" else (4, 6))
    for n in (3, 5):
        p = os.path.join(model_dir, f"bar{n}")
        gen.barrier_flip(n).write(p)
        for prov in (False, True):
            r = gpu_run(gpu, p, 0, True, prov, por_method=m)
            assert (r["states"], r["transitions"]) == (n + 1, n + 1), (n, prov, r)
    p = os.path.join(model_dir, "chain")
    gen.independent_chains(4, 3).write(p)
    r = gpu_run(gpu, p, 1, True, False, por_method=m)
    assert (r["states"], r["transitions"], r["violation"]) == (13, 12, True)


@pytest.mark.parametrize("method", ["ample", "stubborn"])
@pytest.mark.parametrize("name", ["wide3_1", "wide4", "procs32_1", "procs32_2"])
def test_edge_models(gpu, oracle_mod, model_dir, method, name):
    """3-4 key words and 32 processes (full 32-bit process masks)"""
    from edge_models import edge_models
    p = os.path.join(model_dir, name)
    dict(edge_models())[name].write(p)
    for prop in (0, 1, 2):
        for prov in (False, True):
            say("edge", method, name, prop, prov)
            a = oracle_mod.explore(p, prop, True, prov, method=METHODS[method])
            b = gpu_run(gpu, p, prop, True, prov, por_method=METHODS[method])
            assert same(a, b), (method, name, prop, prov, a, b)


@pytest.mark.parametrize("method", ["ample", "stubborn"])
def test_config3(gpu, oracle_mod, model_dir, method):
    """the POR workload (945 M states unreduced): 107,677 states under ample
    sets, 41,169 under stubborn sets (77,841 under cample)"""
    p = os.path.join(model_dir, "c3")
    gen.config3(3).write(p)
    for prop, prov in ((0, False), (1, False), (0, True), (2, True)):
        a = oracle_mod.explore(p, prop, True, prov, method=METHODS[method])
        say("config3", method, prop, prov, a["states"])
        b = gpu_run(gpu, p, prop, True, prov, por_method=METHODS[method], q_ints=1 << 26)
        assert same(a, b), (method, prop, prov, a, b)
    assert a["states"] == {"ample": 107677, "stubborn": 41169}[method]


def test_retry_path(gpu, oracle_mod, model_dir):
    """a tiny successor buffer forces deferred states under stubborn sets
    with the proviso; counts are unchanged"""
    p = os.path.join(model_dir, "c1")
    gen.config1(4).write(p)
    a = oracle_mod.explore(p, 2, True, True, method=METHODS["stubborn"])
    say("retry stubborn")
    b = gpu_run(gpu, p, 2, True, True, por_method=METHODS["stubborn"], cand_capacity=2048)
    assert same(a, b), (a, b)


def test_cli_por_method(gpu, model_dir):
    p = os.path.join(model_dir, "triad")
    triad().write(p)
    for method, want in (("cample", 4), ("ample", 4), ("stubborn", 2)):
        say("cli", method)
        r = subprocess.run([gpu.CLI_PATH, p, "--por-method", method], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        assert f"nr. of states in hash table: {want}\n" in r.stdout, r.stdout
    r = subprocess.run([gpu.CLI_PATH, p, "--por-method", "sleep"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 1 and "unknown partial-order reduction method" in r.stderr


def test_bad_method_rejected(gpu, model_dir):
    p = os.path.join(model_dir, "triad")
    triad().write(p)
    with pytest.raises(gpu.GxError) as e:
        gpu_run(gpu, p, 0, True, False, por_method=3)
    assert e.value.status == -7  # GX_E_ARG


def test_bad_property_rejected(gpu, model_dir):
    """property codes outside GX_PROP_* are errors at every entry point"""
    p = os.path.join(model_dir, "triad")
    triad().write(p)
    m = gpu.Model.load(p)
    with pytest.raises(gpu.GxError) as e:
        gpu.explore(m, property=7, q_ints=1 << 20)
    assert e.value.status == -7
    with pytest.raises(gpu.GxError) as e:
        gpu.explore_batch([m], [-1], q_ints=1 << 20)
    assert e.value.status == -7
    srv = gpu.Server(slots=4, table_bytes=1 << 26, page_bytes=1 << 16)
    with pytest.raises(gpu.GxError) as e:
        srv.submit(m, 9, est_states=10)
    assert e.value.status == -7


# config-1 rings on which the cycle proviso changes the reduced count 5-30x
# (cample, no property: seed 4 4,404 -> 57,989 states, seed 5 11,639 ->
# 127,048, seed 7 7,848 -> 95,576, seed 8 903 -> 12,925)
@pytest.mark.parametrize("seed", [4, 5, 7, 8])
def test_rings_where_the_proviso_matters(gpu, oracle_mod, model_dir, seed):
    p = os.path.join(model_dir, f"ring{seed}")
    gen.config1(seed).write(p)
    for method in METHODS:
        for prop in (0, 1):
            counts = {}
            for prov in (False, True):
                say("ring", seed, method, prop, prov)
                a = oracle_mod.explore(p, prop, True, prov, method=METHODS[method])
                b = gpu_run(gpu, p, prop, True, prov, por_method=METHODS[method])
                assert same(a, b), (seed, method, prop, prov, a, b)
                counts[prov] = b["states"]
            if prop == 0:
                assert counts[True] > counts[False], (seed, method, counts)
