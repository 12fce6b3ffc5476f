"""Seeded generator of networks of LTSs, written in the gpf format.

The gpf format is the reference's model input (GPUexplore/src/GPUexplore.cu
lines 1885-1983; field order in :func:`Network.to_gpf`).  The reference's own
generator script (GPUnetltsgen.py, named in README.md line 22) is not part of
the reference snapshot, so models here come from seeded synthetic families.

Every model this module writes respects the constraints under which the
reference kernels compute the intended semantics:

* local entries come before synchronising entries, and synchronising entries
  are grouped by label in ascending order (cu:904-949, 954-991);
* a (process, state, label) has at most two synchronising entries and the
  first one is full (cu:964-977, 1084-1113), its targets are distinct;
* every synchronisation rule has at least two participants and a label has at
  most N rules (cu:980-997, 1011-1028);
* the state-vector fields leave bit 31 (and bit 30 under POR) of the last
  word free (cu:182-204).
"""

from __future__ import annotations

import dataclasses
from typing import Dict, List, Optional, Sequence, Tuple

MASK64 = (1 << 64) - 1


class Rng:
    """splitmix64: a small, fully specified PRNG so model bytes never depend on
    the Python or numpy version."""

    def __init__(self, seed: int):
        self.s = (seed * 0x9E3779B97F4A7C15 + 0x632BE59BD9B4E019) & MASK64

    def next(self) -> int:
        self.s = (self.s + 0x9E3779B97F4A7C15) & MASK64
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def below(self, n: int) -> int:
        return self.next() % n

    def uniform(self) -> float:
        return (self.next() >> 11) / float(1 << 53)

    def sample(self, n: int, k: int) -> List[int]:
        pool = list(range(n))
        out = []
        for _ in range(min(k, n)):
            i = self.below(len(pool))
            out.append(pool.pop(i))
        return out


@dataclasses.dataclass
class Network:
    """A network of LTSs.

    ``local[p][s]`` is a list of ``(label, [targets])``; ``sync[p][s]`` maps a
    synchronising label to its target list; ``rules[label]`` lists process
    bitmasks.  All processes start in local state 0 (cu:673-676).
    """

    nstates: List[int]
    nlabels: int
    local: List[List[List[Tuple[int, List[int]]]]]
    sync: List[List[Dict[int, List[int]]]]
    rules: Dict[int, List[int]]
    liveness: int = 0
    field_widths: Optional[List[int]] = None  # wider state-vector fields than needed

    @property
    def nprocs(self) -> int:
        return len(self.nstates)

    def check(self) -> None:
        n = self.nprocs
        if not 1 <= n <= 32:
            raise ValueError("1..32 processes supported by the gpf rule packing")
        for a, rs in self.rules.items():
            if len(rs) > n:
                raise ValueError(f"label {a}: more rules than processes")
            for r in rs:
                if bin(r).count("1") < 2 or r >> n:
                    raise ValueError(f"label {a}: bad rule {r:#x}")
        for p in range(n):
            for s in range(self.nstates[p]):
                for lab, ts in self.sync[p][s].items():
                    if len(set(ts)) != len(ts) or not ts:
                        raise ValueError("sync targets must be distinct and non-empty")
                    if lab not in self.rules:
                        raise ValueError(f"sync label {lab} has no rules")
                for lab, ts in self.local[p][s]:
                    if not ts:
                        raise ValueError("local entry without targets")
                for ts in [t for _, t in self.local[p][s]] + list(self.sync[p][s].values()):
                    for t in ts:
                        if not 0 <= t < self.nstates[p]:
                            raise ValueError("target out of range")

    # ------------------------------------------------------------------
    def to_gpf(self) -> str:
        """Encode as gpf text (cu:1885-1983 parse order)."""
        self.check()
        n = self.nprocs
        bits_act = max(1, (self.nlabels - 1).bit_length())
        fw = [max(1, (s - 1).bit_length()) for s in self.nstates]
        if self.field_widths is not None:
            if len(self.field_widths) != n or any(a < b or a > 32 for a, b in zip(self.field_widths, fw)):
                raise ValueError("field widths too small or too large")
            fw = list(self.field_widths)
        fb = [0]
        for w in fw:
            fb.append(fb[-1] + w)
        if fb[-1] + 2 > 32 * 9:
            raise ValueError("state vector wider than 9 words")
        bits_sv = fb[-1] + 1  # includes the new-state flag bit (cu:182)
        bs = [max(1, s.bit_length()) for s in self.nstates]  # values up to S (id+1)
        cap = [(31 - bits_act) // b for b in bs]
        if min(cap) < 1:
            raise ValueError("targets do not fit in a transition entry")

        def entry(sync: int, lab: int, targets: Sequence[int], p: int) -> int:
            e = sync | (lab << 1)
            for i, t in enumerate(targets):
                e |= (t + 1) << (1 + bits_act + i * bs[p])
            assert e < (1 << 32)
            return e

        trans: List[int] = []
        starts: List[List[int]] = []
        max_buf = 1
        for p in range(n):
            st = []
            for s in range(self.nstates[p]):
                st.append(len(trans))
                for lab, ts in self.local[p][s]:
                    for i in range(0, len(ts), cap[p]):
                        trans.append(entry(0, lab, ts[i:i + cap[p]], p))
                for lab in sorted(self.sync[p][s]):
                    ts = self.sync[p][s][lab]
                    chunks = [ts[i:i + cap[p]] for i in range(0, len(ts), cap[p])]
                    if len(chunks) > 2:
                        raise ValueError("more than two entries for one (state, label)")
                    max_buf = max(max_buf, len(chunks))
                    for c in chunks:
                        trans.append(entry(1, lab, c, p))
            st.append(len(trans))
            starts.append(st)
        nr_trans = len(trans)
        nbits_offset = max(1, nr_trans.bit_length())
        if nbits_offset >= 32:
            raise ValueError("too many transition entries")
        k = 32 // nbits_offset
        proc_offsets: List[int] = []
        offsets_start = [0]
        for p in range(n):
            fields = starts[p]
            for i in range(0, len(fields), k):
                word = 0
                for j, f in enumerate(fields[i:i + k]):
                    word |= f << (j * nbits_offset)
                proc_offsets.append(word)
            offsets_start.append(len(proc_offsets))

        per = 32 // n
        syncbits: List[int] = []
        label_start = []
        for a in range(1 << bits_act):
            label_start.append(len(syncbits))
            rs = self.rules.get(a, [])
            for i in range(0, len(rs), per):
                word = 0
                for j, r in enumerate(rs[i:i + per]):
                    word |= r << (j * n)
                syncbits.append(word)
        label_start.append(len(syncbits))
        nbits_so = max(1, len(syncbits).bit_length())
        k2 = 32 // nbits_so
        syncbits_offsets = []
        for i in range(0, len(label_start), k2):
            word = 0
            for j, f in enumerate(label_start[i:i + k2]):
                word |= f << (j * nbits_so)
            syncbits_offsets.append(word)

        lines = [self.liveness, n, bits_act, max(self.nstates), bits_sv]
        lines += fb
        lines += bs
        lines += [nbits_offset, max_buf]
        lines += offsets_start
        lines += proc_offsets
        lines += [nr_trans] + trans
        lines += [nbits_so, len(syncbits_offsets)] + syncbits_offsets
        lines += [len(syncbits)] + syncbits
        return "\n".join(str(int(x)) for x in lines) + "\n"

    def write(self, path_noext: str) -> str:
        path = path_noext + ".gpf"
        with open(path, "w") as f:
            f.write(self.to_gpf())
        return path


def _empty(nstates: Sequence[int]):
    local = [[[] for _ in range(s)] for s in nstates]
    sync = [[{} for _ in range(s)] for s in nstates]
    return local, sync


def _add_local(local, p, s, lab, t):
    for i, (l, ts) in enumerate(local[p][s]):
        if l == lab:
            if t not in ts:
                ts.append(t)
            return
    local[p][s].append((lab, [t]))


def _add_sync(sync, p, s, lab, t):
    ts = sync[p][s].setdefault(lab, [])
    if t not in ts:
        ts.append(t)


def ring_rules(n: int) -> List[int]:
    """2-party rendezvous of every process with its right neighbour."""
    return [(1 << p) | (1 << ((p + 1) % n)) for p in range(n)]


# ----------------------------------------------------------------------
# Families
# ----------------------------------------------------------------------
def throttled(n: int, nstates: int, tokens: int, outdeg: int,
              mix: Tuple[float, float, float], n_local_labels: int, seed: int,
              rdv_grant: int = 1, monitor: Tuple[int, int, int, int] | None = None,
              guard_locals: bool = True, rdv_spends: bool = False,
              barrier_home: bool = False) -> Network:
    """Ring of ``n`` seeded random LTSs.

    Each process has ``nstates = control * tokens`` local states; state id
    ``c * tokens + u`` pairs a control state ``c`` with ``u`` tokens spent,
    so the initial state 0 has a full budget of ``tokens - 1`` tokens.  A
    random control LTS with out-degree ``outdeg`` draws every transition's
    kind from ``mix = (local, rendezvous, barrier)`` and is lifted to the
    product: a local move needs and spends one token (when ``guard_locals``),
    a rendezvous with a neighbour gives back ``rdv_grant`` tokens (or, with
    ``rdv_spends``, needs and spends one like a local move) and the global
    barrier restores the full budget (with ``barrier_home`` it also sends the
    process back to control state 0, so every barrier starts a new round
    from the initial state).  Labels: ``0 ..
    n_local_labels-1`` local, then one neighbour-rendezvous label and one
    n-way barrier label.  ``monitor = (i, x, j, y)`` appends a two-state
    monitor process that moves to state 1 exactly when process i is in x and
    process j in y (a safety property "last process in state 1").
    """
    if nstates % tokens:
        raise ValueError("nstates must be a multiple of tokens")
    ctrl = nstates // tokens
    rng = Rng(seed)
    L = n_local_labels
    RDV, BAR = L, L + 1
    nproc_total = n + (1 if monitor else 0)
    sizes = [nstates] * n + ([2] if monitor else [])
    local, sync = _empty(sizes)
    top = tokens - 1
    for p in range(n):
        for c in range(ctrl):
            for _ in range(outdeg):
                u = rng.uniform()
                c2 = rng.below(ctrl)
                if u < mix[0]:
                    lab = rng.below(L)
                    for k in range(tokens):
                        if guard_locals and tokens > 1:
                            if k < top:
                                _add_local(local, p, c * tokens + k, lab, c2 * tokens + k + 1)
                        else:
                            _add_local(local, p, c * tokens + k, lab, c2 * tokens + k)
                elif u < mix[0] + mix[1]:
                    for k in range(tokens):
                        if rdv_spends:
                            if k < top:
                                _add_sync(sync, p, c * tokens + k, RDV, c2 * tokens + k + 1)
                        else:
                            _add_sync(sync, p, c * tokens + k, RDV, c2 * tokens + max(0, k - rdv_grant))
                else:
                    for k in range(tokens):
                        _add_sync(sync, p, c * tokens + k, BAR, 0 if barrier_home else c2 * tokens)
    rules = {RDV: ring_rules(n) if n >= 3 else [3], BAR: [(1 << n) - 1]}
    nlabels = L + 2
    if monitor:
        i, x, j, y = monitor
        BAD = L + 2
        nlabels = L + 3
        m = n
        _add_sync(sync, i, x, BAD, x)
        _add_sync(sync, j, y, BAD, y)
        _add_sync(sync, m, 0, BAD, 1)
        rules[BAD] = [(1 << i) | (1 << j) | (1 << m)]
    # keep only labels that are actually used by some rule member
    net = Network(sizes, nlabels, local, sync, rules)
    net.check()
    return net


def rounds(n: int, nstates: int, tokens: int, budgets: Sequence[int], n_local_labels: int,
           seed: int, monitor: Tuple[int, int, int, int] | None = None,
           mix: Tuple[float, float, float] | None = None, outdeg: int = 3,
           phased: bool = False) -> Network:
    """Ring of ``n`` seeded random LTSs that run in rounds.

    Local state ``c * tokens + u`` pairs a control state ``c`` with ``u``
    tokens spent.  Every control state has out-degree 3: one local move
    (random local label, random target), one rendezvous with a neighbour
    (random target) and one n-way barrier; with ``mix = (local,
    rendezvous, barrier)`` it has ``outdeg`` transitions whose kinds are drawn
    from the mix instead.  With ``phased`` the round has two phases: local
    moves start from unspent states only, while rendezvous and barrier moves
    start from spent states only, so states at the start of a round carry no
    synchronising labels (the situation cluster-based POR reduces).  Local and rendezvous moves need
    and spend a token (process p may spend ``budgets[p] <= tokens-1`` per
    round); the barrier sends every process home to state 0.  The reachable
    set is therefore what one round can reach, and the per-process budgets
    set its size.
    """
    if nstates % tokens or len(budgets) != n or max(budgets) >= tokens:
        raise ValueError("bad rounds parameters")
    ctrl = nstates // tokens
    rng = Rng(seed)
    L = n_local_labels
    RDV, BAR = L, L + 1
    sizes = [nstates] * n + ([2] if monitor else [])
    local, sync = _empty(sizes)
    for p in range(n):
        for c in range(ctrl):
            if mix is None:
                kinds = [0, 1, 2]
            else:
                kinds = []
                for _ in range(outdeg):
                    x = rng.uniform()
                    kinds.append(0 if x < mix[0] else (1 if x < mix[0] + mix[1] else 2))
            for kind in kinds:
                if kind == 0:
                    lab, cl = rng.below(L), rng.below(ctrl)
                    for u in range(budgets[p]):
                        _add_local(local, p, c * tokens + u, lab, cl * tokens + u + 1)
                elif kind == 1:
                    cr = rng.below(ctrl)
                    lo = 1 if phased else 0
                    for u in range(lo, budgets[p] + lo):
                        _add_sync(sync, p, c * tokens + u, RDV, cr * tokens + min(u + 1, tokens - 1))
                else:
                    for u in range(1 if phased else 0, tokens):
                        _add_sync(sync, p, c * tokens + u, BAR, 0)
    rules = {RDV: ring_rules(n) if n >= 3 else [3], BAR: [(1 << n) - 1]}
    nlabels = L + 2
    if monitor:
        i, x, j, y = monitor
        BAD = L + 2
        nlabels = L + 3
        _add_sync(sync, i, x, BAD, x)
        _add_sync(sync, j, y, BAD, y)
        _add_sync(sync, n, 0, BAD, 1)
        rules[BAD] = [(1 << i) | (1 << j) | (1 << n)]
    net = Network(sizes, nlabels, local, sync, rules)
    net.check()
    return net


def config1(seed: int = 1) -> Network:
    """Ring of 6 processes, 8 states, out-degree 2, labels 0-2 local, label 3
    a 2-party rendezvous with the right neighbour; one 32-bit word."""
    n, S = 6, 8
    rng = Rng(seed)
    local, sync = _empty([S] * n)
    for p in range(n):
        for s in range(S):
            for _ in range(2):
                lab = rng.below(4)
                t = rng.below(S)
                if lab < 3:
                    _add_local(local, p, s, lab, t)
                else:
                    _add_sync(sync, p, s, 3, t)
    net = Network([S] * n, 4, local, sync, {3: ring_rules(n)})
    net.check()
    return net


def config2(seed: int = 1) -> Network:
    """Ring of 20 processes with 16 local states each (4 control states x a
    token variable of domain 4: 4 bits per process, 80 bits, 3 words),
    out-degree 3, labels 0-3 local, 4 neighbour rendezvous, 5 global 20-way
    barrier.  Local and rendezvous moves spend tokens; the barrier sends
    everyone home.  Processes 0, 5, 10, 15 may spend two tokens per round,
    the others one."""
    budgets = [2 if p % 5 == 0 else 1 for p in range(20)]
    return rounds(20, 16, 4, budgets, 4, seed)


def config3(seed: int = 3) -> Network:
    """24 processes with 8 local states each (4 control states x 2 tokens
    spent; 3 bits per process, 72 bits, 3 words), out-degree 3 drawn 70%
    local, 20% neighbour rendezvous, 10% barrier.  Rounds are phased: local
    moves from unspent states, rendezvous and barrier moves from spent
    states, so a round starts with independent local moves only (the POR
    workload)."""
    return rounds(24, 8, 4, [1] * 24, 4, seed, mix=(0.7, 0.2, 0.1), phased=True)


def config4_models(seed0: int = 1000, count: int = 64) -> List[Tuple[Network, str]]:
    """Batch of ``count`` models of uneven size: N in 4..14 processes, 4..32
    local states (control states x 2-4 tokens), out-degree 1..3 with kinds
    drawn from a mixed local / rendezvous / barrier distribution, 1-2 tokens
    spent per round (1 in models of 12 or more processes).  Even indices run -d; odd ones run -p with a seeded
    bad-state predicate over two processes, encoded as a two-state monitor in
    last position."""
    out = []
    for i in range(count):
        rng = Rng(seed0 + i)
        n = 4 + rng.below(11)
        tokens = 2 + rng.below(3)
        ctrl = max(2, (4 + rng.below(29)) // tokens)
        S = ctrl * tokens
        outdeg = 1 + rng.below(3)
        budgets = [1 + (rng.below(4) == 0 and tokens > 2 and n < 12) for _ in range(n)]
        mix = (0.5, 0.35, 0.15)
        if i % 2 == 0:
            net = rounds(n, S, tokens, budgets, 3, seed0 + i, mix=mix, outdeg=outdeg)
            out.append((net, "-d"))
        else:
            a, b = rng.sample(n, 2)
            mon = (a, rng.below(S), b, rng.below(S))
            net = rounds(n, S, tokens, budgets, 3, seed0 + i, monitor=mon, mix=mix, outdeg=outdeg)
            out.append((net, "-p"))
    return out


def config2_batch(seed0: int = 2000, count: int = 64, nmin: int = 14, nmax: int = 17) -> List[Network]:
    """A batch at the large plain workload's scale: ``count`` models from the
    config-2 generator (16 local states = 4 control states x 4 tokens, token
    budgets of 2 for every fifth process and 1 otherwise, out-degree 3: a
    local move, a neighbour rendezvous and an n-way barrier) with n uniform
    in nmin..nmax processes.  At 14..17 processes a model has 1e6..5e7
    states and the batch about 1e9; it is explored without a property."""
    out = []
    for i in range(count):
        rng = Rng(0xB16 * 1000003 + seed0 + i)
        n = nmin + rng.below(nmax - nmin + 1)
        budgets = [2 if p % 5 == 0 else 1 for p in range(n)]
        out.append(rounds(n, 16, 4, budgets, 4, seed0 + i))
    return out


def config5_job(seed: int) -> Tuple[Network, str, int, bool]:
    """One job of the serving stream: 90% drawn from the config-3 generator
    (24 processes, 8 local states, phased rounds) and checked with
    ``--por -d`` -- cluster-based POR without the cycle proviso, since -d
    switches it off (cu:1851-1874) -- and 10% from the config-2 generator at
    full scale (20 processes, 16 local states, token budgets, -d).  Returns
    (network, flag, estimated states, por).  The POR jobs' reduced spaces
    hold 7e4..4e6 states (mean about 1e6, 40 seeds), the config-2 jobs
    3e8..6e8; the estimates are what a client would guess, and the server
    restarts a job whose partition fills."""
    rng = Rng(0xC5 * 1000003 + seed)
    if rng.below(10) == 0:
        return config2(seed), "-d", 500_000_000, False
    return rounds(24, 8, 4, [1] * 24, 4, seed, mix=(0.7, 0.2, 0.1), phased=True), "-d", 2_000_000, True


# ----------------------------------------------------------------------
# Closed-form families (used to pin the oracle)
# ----------------------------------------------------------------------
def independent_cycles(lengths: Sequence[int]) -> Network:
    """N independent cycles 0->1->...->n_i-1->0: prod(n_i) states,
    N*prod(n_i) transitions, no deadlock."""
    local, sync = _empty(lengths)
    for p, n in enumerate(lengths):
        for s in range(n):
            _add_local(local, p, s, 0, (s + 1) % n)
    net = Network(list(lengths), 1, local, sync, {})
    net.check()
    return net


def independent_chains(n: int, k: int) -> Network:
    """N independent chains 0->...->k: (k+1)^N states,
    N*k*(k+1)^(N-1) transitions, one deadlock (all at k)."""
    local, sync = _empty([k + 1] * n)
    for p in range(n):
        for s in range(k):
            _add_local(local, p, s, 0, s + 1)
    net = Network([k + 1] * n, 1, local, sync, {})
    net.check()
    return net


def barrier_flip(n: int) -> Network:
    """Every process: local 0->1, and an n-way barrier 1->0.  2^N states,
    N*2^(N-1) + 1 transitions, no deadlock."""
    local, sync = _empty([2] * n)
    for p in range(n):
        _add_local(local, p, 0, 0, 1)
        _add_sync(sync, p, 1, 1, 0)
    net = Network([2] * n, 2, local, sync, {1: [(1 << n) - 1]})
    net.check()
    return net
