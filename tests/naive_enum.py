"""Naive enumerator of gpf models, written straight from the definitions.

This is deliberately independent of the C oracle: it parses the gpf text on
its own, represents states as tuples of local states (no bit packing), reads
the per-process offset stream as one flat list of fields, and computes the
cample clusters by fixpoint iteration over sets.  Pure Python loops: only
for tiny models.  It pins the oracle (tests/test_oracle.py).

Semantics (reference cu:N = GPUexplore/src/GPUexplore.cu):
* local move of p: every target of every local entry of p at s_p
  (cu:904-949);
* rule r of label a fires iff every p in r has an a-entry at s_p; successors
  are the product of the target lists (cu:951-1130);
* deadlock: no local entry and no enabled rule (cu:1134-1151);
* safety: a generated successor has the last process in state 1
  (cu:926-935, 1063-1069); under POR over the full successor set;
* cample POR clusters (cu:1616-1672), proviso per level tags.
"""

from __future__ import annotations

from itertools import product
from typing import Dict, List, Tuple


def _bits(x: int, start: int, length: int) -> int:
    return (x >> start) & ((1 << length) - 1)


class NaiveModel:
    def __init__(self, text: str):
        vals = [int(line.split()[0]) & 0xFFFFFFFF for line in text.strip().splitlines()]
        it = iter(vals)
        nx = lambda: next(it)  # noqa: E731
        self.liveness = nx()
        n = self.n = nx()
        self.bits_act = nx()
        nx()  # proc_nrstates, informational
        nx()  # bits_statevector
        self.fb = [nx() for _ in range(n + 1)]
        self.bs = [nx() for _ in range(n)]
        self.nbits_offset = nx()
        nx()  # max_buf_ints
        starts = [nx() for _ in range(n + 1)]
        offs = [nx() for _ in range(starts[n])]
        nr_trans = nx()
        trans = [nx() for _ in range(nr_trans)]
        nb_so = nx()
        nso = nx()
        so = [nx() for _ in range(nso)]
        nsb = nx()
        sb = [nx() for _ in range(nsb)]
        k = 32 // self.nbits_offset
        # flat field stream per process
        self.fields: List[List[int]] = []
        for p in range(n):
            f = []
            for w in offs[starts[p]:starts[p + 1]]:
                for j in range(k):
                    f.append(_bits(w, j * self.nbits_offset, self.nbits_offset))
            self.fields.append(f)
        self.trans = trans
        k2 = 32 // nb_so
        so_fields = []
        for w in so:
            for j in range(k2):
                so_fields.append(_bits(w, j * nb_so, nb_so))
        self.rules: Dict[int, List[int]] = {}
        per = 32 // n
        for a in range(1 << self.bits_act):
            lo, hi = so_fields[a], so_fields[a + 1]
            rs = []
            for w in sb[lo:hi]:
                for j in range(per):
                    r = _bits(w, j * n, n)
                    if r:
                        rs.append(r)
            self.rules[a] = rs
        self._cache: Dict[Tuple[int, int], Tuple[List[List[int]], Dict[int, List[int]]]] = {}

    def entries(self, p: int, s: int):
        key = (p, s)
        if key in self._cache:
            return self._cache[key]
        lo, hi = self.fields[p][s], self.fields[p][s + 1]
        cap = (31 - self.bits_act) // self.bs[p]
        locs: List[List[int]] = []
        syn: Dict[int, List[int]] = {}
        for e in self.trans[lo:hi]:
            tg = []
            for i in range(cap):
                v = _bits(e, 1 + self.bits_act + i * self.bs[p], self.bs[p])
                if v == 0:
                    break
                tg.append(v - 1)
            if e & 1:
                syn.setdefault(_bits(e, 1, self.bits_act), []).extend(tg)
            else:
                locs.append(tg)
        self._cache[key] = (locs, syn)
        return locs, syn

    def width(self, p: int) -> int:
        return self.fb[p + 1] - self.fb[p]

    def analyse(self, s: Tuple[int, ...]):
        """Returns (succ list of (state, origin)), deadlock flag, per-process
        cluster0, local-successor set, enabled-rule memberships."""
        n = self.n
        succ = []
        any_local = False
        cluster0 = [{p} for p in range(n)]
        has_local = set()
        enabled_rules = []  # (label, rule-set)
        ent = [self.entries(p, s[p]) for p in range(n)]
        for p in range(n):
            locs, _ = ent[p]
            for tg in locs:
                any_local = True
                for t in tg:
                    has_local.add(p)
                    mask = (1 << self.width(p)) - 1
                    succ.append((s[:p] + (t & mask,) + s[p + 1:], ("L", p)))
        labels = sorted({a for p in range(n) for a in ent[p][1]})
        for a in labels:
            en = {p for p in range(n) if a in ent[p][1]}
            for r in self.rules[a]:
                members = [p for p in range(n) if (r >> p) & 1]
                for p in members:
                    if p in en:
                        cluster0[p] |= set(members)
                if not set(members) <= en:
                    continue
                enabled_rules.append((a, frozenset(members)))
                lists = [ent[p][1][a] for p in members]
                for combo in product(*lists):
                    t = list(s)
                    for p, v in zip(members, combo):
                        t[p] = v & ((1 << self.width(p)) - 1)
                    succ.append((tuple(t), ("R", frozenset(members))))
        deadlock = not (any_local or enabled_rules)
        return succ, deadlock, cluster0, has_local, enabled_rules

    @staticmethod
    def select(n, cluster0, prov):
        c = [set(cluster0[p]) if p in prov else set(cluster0[p]) - {p} for p in range(n)]
        fin = []
        for p in range(n):
            K = set(c[p])
            while True:
                K2 = set(K)
                for i in K:
                    K2 |= c[i]
                if K2 == K:
                    break
                K = K2
            ok = any(i in c[i] for i in K)
            fin.append(K if ok else set())
        best, C = n, set(range(n))
        for p in range(n):
            if fin[p] and len(fin[p]) < best:
                best, C = len(fin[p]), fin[p]
        return C

    def stub_relation(self, s: Tuple[int, ...]):
        """d(p) of the stubborn-set method: p, plus for each rule r containing
        p of a label p has at s_p, r itself when every member has the label,
        else r's lowest member without it."""
        n = self.n
        ent = [self.entries(p, s[p]) for p in range(n)]
        d = [{p} for p in range(n)]
        for p in range(n):
            for a in ent[p][1]:
                for r in self.rules[a]:
                    if not (r >> p) & 1:
                        continue
                    members = [q for q in range(n) if (r >> q) & 1]
                    lacking = [q for q in members if a not in ent[q][1]]
                    d[p] |= set(members) if not lacking else {min(lacking)}
        return d

    @staticmethod
    def select_ample(n, cluster0, nloc, cand):
        best, C = None, set(range(n))
        for p in range(n):
            if p in cand and cluster0[p] == {p} and (best is None or nloc[p] < best):
                best, C = nloc[p], {p}
        return C

    @staticmethod
    def select_stubborn(n, d, target):
        best, C = n, set(range(n))
        for p in range(n):
            K = {p}
            while True:
                K2 = set(K)
                for i in K:
                    K2 |= d[i]
                if K2 == K:
                    break
                K = K2
            if K & target and len(K) < best:
                best, C = len(K), K
        return C

    def explore(self, prop: str = "", por: bool = False, proviso: bool = False, method: str = "cample"):
        n = self.n
        s0 = tuple([0] * n)
        level = {s0: 0}
        frontier = [s0]
        trans = 0
        lv = 0
        violation = False
        while frontier:
            nxt = []
            for s in frontier:
                succ, dead, cl0, has_local, enabled = self.analyse(s)
                if prop == "d" and dead:
                    violation = True
                if prop == "p" and any(t[n - 1] == 1 for t, _ in succ):
                    violation = True
                if por:
                    lnew = set()
                    if proviso:
                        prov = set()
                        for t, (kind, who) in succ:
                            if not (t in level and level[t] <= lv):
                                prov |= {who} if kind == "L" else set(who)
                                if kind == "L":
                                    lnew.add(who)
                    else:
                        prov = set(has_local)
                        for _, members in enabled:
                            prov |= members
                    if method == "ample":
                        nloc = [sum(1 for _, (k, w) in succ if k == "L" and w == p) for p in range(n)]
                        C = self.select_ample(n, cl0, nloc, lnew if proviso else set(has_local))
                    elif method == "stubborn":
                        C = self.select_stubborn(n, self.stub_relation(s), prov)
                    else:
                        C = self.select(n, cl0, prov)
                    amp = [t for t, (kind, who) in succ
                           if (kind == "L" and who in C) or (kind == "R" and who <= C)]
                else:
                    amp = [t for t, _ in succ]
                for t in amp:
                    trans += 1
                    if t not in level:
                        level[t] = lv + 1
                        nxt.append(t)
            lv += 1
            frontier = nxt
            if violation:
                break
        return dict(states=len(level), transitions=trans, levels=lv, violation=violation)


def pack(model: NaiveModel, s: Tuple[int, ...], words: int) -> List[int]:
    """Pack local states into little-endian 32-bit words per the firstbit
    table (a plain bit-by-bit loop, independent of any codec)."""
    bits = [0] * (32 * words)
    for p, v in enumerate(s):
        for i in range(model.width(p)):
            bits[model.fb[p] + i] = (v >> i) & 1
    return [sum(bits[32 * w + i] << i for i in range(32)) for w in range(words)]


def unpack(model: NaiveModel, words: List[int]) -> Tuple[int, ...]:
    out = []
    for p in range(model.n):
        v = 0
        for i in range(model.width(p)):
            b = model.fb[p] + i
            v |= ((words[b // 32] >> (b % 32)) & 1) << i
        out.append(v)
    return tuple(out)
