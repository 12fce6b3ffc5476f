"""Models at the edges of the supported shapes, within the gpf constraints:
wide state vectors (3 to 5 key words), 32 processes (full rule masks, a
32-party barrier) and more than 64 labels (past the 64-bit row label masks
of the state-per-lane kernels)."""

from gpuexplore_amd import gen
from gpuexplore_amd.gen import Network, _add_local, _add_sync, _empty


def wide(n: int, width: int, seed: int) -> Network:
    """throttled ring of n processes with `width`-bit state fields"""
    net = gen.throttled(n, 6, 2, 2, (0.5, 0.4, 0.1), 2, seed)
    net.field_widths = [width] * n
    net.check()
    return net


def procs32(seed: int) -> Network:
    """32 processes: 0..3 run a small random ring with locals, 4..31 sit in
    their only state; rendezvous rules pair every process with its right
    neighbour (only 0..3 can use them) and a 32-party barrier needs every
    process (4..31 always agree with a self-loop)."""
    n = 32
    rng = gen.Rng(seed)
    sizes = [4] * 4 + [1] * (n - 4)
    local, sync = _empty(sizes)
    RDV, BAR = 2, 3
    for p in range(4):
        for s in range(4):
            _add_local(local, p, s, rng.below(2), rng.below(4))
            if rng.below(2):
                _add_sync(sync, p, s, RDV, rng.below(4))
            if s == 3:
                _add_sync(sync, p, s, BAR, 0)
    for p in range(4, n):
        _add_sync(sync, p, 0, BAR, 0)
    rules = {RDV: [(1 << p) | (1 << ((p + 1) % n)) for p in range(n)], BAR: [(1 << n) - 1]}
    net = Network(sizes, 4, local, sync, rules)
    net.check()
    return net


def labels70(seed: int) -> Network:
    """3 processes with 70 synchronising labels (each a 2-party rule over a
    random pair) plus locals"""
    n, S, NL = 3, 5, 70
    rng = gen.Rng(seed)
    local, sync = _empty([S] * n)
    rules = {}
    for lab in range(1, NL + 1):
        a = rng.below(n)
        b = (a + 1 + rng.below(n - 1)) % n
        rules[lab] = [(1 << a) | (1 << b)]
        for p in (a, b):
            s = rng.below(S)
            _add_sync(sync, p, s, lab, rng.below(S))
    for p in range(n):
        for s in range(S):
            if rng.below(3) == 0:
                _add_local(local, p, s, 0, rng.below(S))
    net = Network([S] * n, NL + 1, local, sync, rules)
    net.check()
    return net


def edge_models():
    return [("wide3_1", wide(5, 30, 1)), ("wide3_2", wide(4, 32, 2)), ("wide4", wide(7, 32, 3)),
            ("wide5", wide(9, 31, 4)), ("procs32_1", procs32(1)), ("procs32_2", procs32(2)),
            ("labels70_1", labels70(1)), ("labels70_2", labels70(2))]
