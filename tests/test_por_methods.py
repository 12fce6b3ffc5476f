"""Ample-set and stubborn-set partial-order reduction in the oracle.

The reference keeps these methods on branches absent from its snapshot
(README.md lines 24-26), so nothing of the reference pins them: they are
pinned by an independent restatement in the naive enumerator (own gpf
parser, own stubborn relation from the entry tables), hand-computed counts,
closed forms, and the invariants of a sound reduction (a subset of the full
state space; deadlock verdicts preserved)."""

import pytest

from gpuexplore_amd import gen
from gpuexplore_amd.gen import Network, _add_local, _add_sync, _empty
from naive_enum import NaiveModel
from test_oracle import PROPS, small_models, write

METHODS = {"cample": 0, "ample": 1, "stubborn": 2}


def triad() -> Network:
    """A flips 0<->1 and holds label x in state 0; B never moves; C flips
    0<->1 and holds x in both states; x's only rule is {A, B, C}, blocked by
    B for ever.  Full: 4 states, 8 transitions.  While A is in 0 every
    cample cluster is all three processes and no process is an ample
    candidate, so such states are expanded fully (4 states, 6 transitions);
    the stubborn closure of A is {A, B} (B blocks x), valid through A's local
    move, so A's move alone is taken (2 states, 2 transitions)."""
    local, sync = _empty([2, 1, 2])
    X = 1
    for s in (0, 1):
        _add_local(local, 0, s, 0, 1 - s)
        _add_local(local, 2, s, 0, 1 - s)
        _add_sync(sync, 2, s, X, s)
    _add_sync(sync, 0, 0, X, 0)
    net = Network([2, 1, 2], 2, local, sync, {X: [0b111]})
    net.check()
    return net


@pytest.mark.parametrize("method,want", [("cample", (4, 6)), ("ample", (4, 6)), ("stubborn", (2, 2))])
def test_triad_hand_counts(oracle_mod, model_dir, method, want):
    p = write(triad(), model_dir, "triad")
    full = oracle_mod.explore(p, 0)
    assert (full["states"], full["transitions"]) == (4, 8)
    r = oracle_mod.explore(p, 0, por=True, method=METHODS[method])
    assert (r["states"], r["transitions"]) == want
    nm = NaiveModel(open(p + ".gpf").read()).explore("", por=True, method=method)
    assert (nm["states"], nm["transitions"]) == want


@pytest.mark.parametrize("method", ["ample", "stubborn"])
@pytest.mark.parametrize("n,k", [(1, 4), (3, 2), (4, 3)])
def test_independent_chains(oracle_mod, model_dir, method, n, k):
    """every process is its own ample / stubborn set; the lowest movable one
    runs to its end first: N*k + 1 states, N*k transitions, the deadlock kept"""
    p = write(gen.independent_chains(n, k), model_dir, "chain")
    r = oracle_mod.explore(p, 1, por=True, method=METHODS[method])
    assert (r["states"], r["transitions"], r["violation"]) == (n * k + 1, n * k, True)


@pytest.mark.parametrize("method", list(METHODS))
@pytest.mark.parametrize("proviso", [False, True])
@pytest.mark.parametrize("n", [2, 3, 5])
def test_barrier_flip(oracle_mod, model_dir, method, proviso, n):
    """all of 0^n: one process at a time takes its local step (a barrier
    holder's closure also holds the lowest process still in 0), then the
    n-way barrier from 1^n is the only move: n + 1 states, n + 1 transitions"""
    p = write(gen.barrier_flip(n), model_dir, "bar")
    r = oracle_mod.explore(p, 0, por=True, proviso=proviso, method=METHODS[method])
    assert (r["states"], r["transitions"]) == (n + 1, n + 1)


@pytest.mark.parametrize("name,net", small_models(), ids=[n for n, _ in small_models()])
def test_oracle_vs_naive(oracle_mod, model_dir, name, net):
    p = write(net, model_dir, name)
    nm = NaiveModel(open(p + ".gpf").read())
    for method in ("ample", "stubborn"):
        for prop in ("", "d", "p"):
            for prov in (False, True):
                a = nm.explore(prop, True, prov, method=method)
                b = oracle_mod.explore(p, PROPS[prop], True, prov, method=METHODS[method])
                for k in ("states", "transitions", "levels", "violation"):
                    assert a[k] == b[k], (method, prop, prov, k, a, b)


@pytest.mark.parametrize("method", ["ample", "stubborn"])
def test_subset_and_deadlock_verdict(oracle_mod, model_dir, method):
    """a reduced exploration visits no more states than the full one and
    finds a deadlock exactly when one is reachable"""
    nets = [gen.config1(s) for s in range(1, 6)]
    nets += [gen.throttled(5, 6, 2, 2, (0.4, 0.4, 0.2), 2, s) for s in range(1, 6)]
    nets += [gen.independent_chains(3, 3), gen.barrier_flip(4), triad()]
    for i, net in enumerate(nets):
        p = write(net, model_dir, f"m{i}")
        full = oracle_mod.explore(p, 1)
        for prov in (False, True):
            red = oracle_mod.explore(p, 1, por=True, proviso=prov, method=METHODS[method])
            assert red["states"] <= full["states"], (i, prov)
            assert red["violation"] == full["violation"], (i, prov)


def test_selected_sets(oracle_mod, model_dir):
    """orc_por_select on the triad's initial state: cample and ample expand
    fully, stubborn takes {A, B}"""
    p = write(triad(), model_dir, "triad")
    om = oracle_mod.OracleModel(p, por=True)
    s0 = [0] * om.words
    assert om.por_select(s0, METHODS["cample"]) == om.por_cluster(s0)
    assert om.por_select(s0, METHODS["cample"]) & 0b111 == 0b111
    assert om.por_select(s0, METHODS["ample"]) & 0b111 == 0b111
    assert om.por_select(s0, METHODS["stubborn"]) == 0b011
