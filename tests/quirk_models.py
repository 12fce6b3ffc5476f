"""Models outside the generator's constraints, written by editing gpf text."""

from gpuexplore_amd import gen


def empty_sync_entry_gpf() -> str:
    """A 3-process ring with rendezvous where process 0's entry for the
    rendezvous label at local state 0 keeps its label but loses its target
    (all target fields 0).  The label still counts as present (the rule is
    enabled, so no deadlock there), and the rule's product is 0."""
    net = gen.config1(5)
    net.nstates = net.nstates[:3]
    net.local = net.local[:3]
    net.sync = net.sync[:3]
    net.rules = {3: gen.ring_rules(3)}
    net.check()
    text = net.to_gpf()
    vals = [int(x.split()[0]) for x in text.strip().splitlines()]
    bits_act = vals[2]
    lab = 3
    n = len(net.nstates)
    bs = vals[5 + n + 1:5 + n + 1 + n]
    lines = text.strip().splitlines()
    for p in range(n):
        for st in range(net.nstates[p]):
            ts = net.sync[p][st].get(lab)
            if not ts:
                continue
            full = 1 | (lab << 1)
            for i, t in enumerate(ts[: (31 - bits_act) // bs[p]]):
                full |= (t + 1) << (1 + bits_act + i * bs[p])
            hits = [i for i, v in enumerate(vals) if v == full and i > 5 + 2 * n]
            if len(hits) == 1:
                lines[hits[0]] = str(1 | (lab << 1))
                return "\n".join(lines) + "\n"
    raise AssertionError("no uniquely identifiable rendezvous entry")
