"""Seeded table scenarios run identically through any C-ABI table (the
device library, the reference library) and the oracle, after the matrix of
test/capi.cpp:514-638: each operation and width, plain and oriented,
presized exactly, to a power of 2, or from a tiny table that grows, then
updated; contents, queries and per-call results recorded.
"""
import numpy as np

import seqgen as G

KEEP, ASSIGN, ADD, OR, MIN, MAX = range(6)
M_FWD, M_REV = 0x5555555555555555, 0xA3A3A3A3A3A3A3A3


def operand(op, b, bits):
    """test/capi.cpp:139-149: some sums saturate, ORs set one bit each."""
    mx = (1 << bits) - 1
    h = ((b + 1) * 0x9E3779B97F4A7C15) & ((1 << 64) - 1)
    if op == ADD:
        return mx // 2 + 1 if b % 5 == 0 else 1 + b % 7
    if op == OR:
        return 1 << ((h >> 32) % bits)
    return (h >> 7) & mx


def sc(name, width, op, oriented=False, sizing="exact", update=None, groups=False, genome=60_000, seed=11):
    return dict(name=name, width=width, op=op, oriented=oriented, sizing=sizing, update=update, groups=groups,
                genome=genome, seed=seed)


SCENARIOS = [
    sc("set_exact", 0, KEEP),
    sc("set_tiny", 0, KEEP, sizing="tiny"),
    sc("set_groups_exact", 0, KEEP, groups=True),
    sc("map8_add_tiny", 1, ADD, sizing="tiny", update=ADD),
    sc("map16_or_pow2", 2, OR, sizing="pow2", update=OR),
    sc("map32_min_tiny", 4, MIN, sizing="tiny", update=MAX),
    sc("map64_max_exact", 8, MAX, update=ADD),
    sc("map32_keep_exact", 4, KEEP, update=ASSIGN),
    sc("map16_assign_tiny", 2, ASSIGN, sizing="tiny", update=KEEP),
    sc("map8_oriented_or_tiny", 1, OR, oriented=True, sizing="tiny", groups=True),
    sc("map64_oriented_add_exact", 8, ADD, oriented=True, groups=True),
    sc("map32_oriented_max_pow2", 4, MAX, oriented=True, sizing="pow2"),
]


def inputs(s):
    g = G.genome(s["genome"], s["seed"])
    S = G.mixed_inserts(g, s["seed"] + 1, s["genome"])
    if s["groups"]:
        for gi, grp in enumerate(G.minimizer_groups(6, 160, 42)):
            for x in grp:
                if gi % 3 != 1:
                    S.append(x)
                if gi % 3 != 0:
                    S.append(G.revcomp(x))
    neg = G.pieces(G.genome(12_000, s["seed"] + 2), 1024)
    Q = [S[i] for i in range(0, len(S), 4)] + [G.revcomp(S[i]) for i in range(0, len(S), 9)] + neg + G.edge_queries(S[0])
    return S, Q


def distinct(S):
    import oracle as O
    t = O.OTable(0)
    for b in G.ragged_batches(S):
        t.insert(b)
    return t.size()


def reductions(t, w):
    """test/capi.cpp:276-286: kache_reduce over the masks {max, 0x5A..5A & max}
    and each reduction a table of width w has (KEYS only for a set)."""
    mx = (1 << (8 * w)) - 1 if w else (1 << 64) - 1
    out = []
    for m in (mx, 0x5A5A5A5A5A5A5A5A & mx):
        for what in range(3 if w else 1):
            out.append(t.reduce(what, m, 1 + what))
    return out


def first_kmers(t, Q):
    """test/capi.cpp:361-369: kache_get of each query's first k-mer (queries
    of at least 31 characters), as (found, value) arrays."""
    r = [t.get(q[:31]) for q in Q if len(q) >= 31]
    return np.array([f for f, _ in r], np.uint8), np.array([v for _, v in r], np.uint64)


def run(table_cls, api, s, S=None, Q=None):
    """Runs scenario `s` through a table of class `table_cls` on `api`."""
    if S is None:
        S, Q = inputs(s)
    w, bits = s["width"], max(8 * s["width"], 1)
    n = 10 if s["sizing"] == "tiny" else distinct(S)
    t = table_cls(api, w, n, 0.8, s["sizing"] == "pow2", 4)
    u = t.register()
    ins = []
    for b, batch in enumerate(G.ragged_batches(S)):
        x, xr = operand(s["op"], b, bits), operand(s["op"], b + 977, bits)
        if s["oriented"]:
            ins.append(t.insert_oriented(batch, s["op"], x, xr, u))
        else:
            ins.append(t.insert(batch, s["op"], x if w else 0, u))
    r = {"inserted": ins}
    r["keys"], r["values"], _ = t.for_each()
    st = t.stats()
    r["capacity"], r["size"] = st["capacity"], t.size()
    f, v, fw = t.lookup(Q)
    r["lookup"] = (f, v if v is not None else np.zeros(0, np.uint64), fw)
    r["counts"] = t.count(Q)
    if w:
        r["sums"], r["masked"] = t.sum(Q), t.count_masked(Q, M_FWD, M_REV)
    r["reduce"] = reductions(t, w)
    r["get"] = first_kmers(t, Q)
    if w and s["update"] is not None:
        pres = []
        for b, batch in enumerate(G.ragged_batches(Q)):
            pres.append(t.update(batch, s["update"], operand(s["update"], b + 5000, bits)))
        r["present"] = pres
        r["keys2"], r["values2"], _ = t.for_each()
    r["stats"] = st
    t.close()
    return r


class OracleTable:
    """The oracle behind the Table interface used by `run`."""

    def __init__(self, api, width, n, lf, pow2, workers):
        import oracle as O
        self.t = O.OTable(width)
        self.width = width
        self.sched = O.Sched(n, lf, pow2)

    def register(self):
        return 0

    def insert(self, b, op, x, u):
        c = self.t.insert(b, op, x)
        self.sched.call(u, c)
        return c

    def insert_oriented(self, b, op, x, xr, u):
        c = self.t.insert_oriented(b, op, x, xr)
        self.sched.call(u, c)
        return c

    def update(self, b, op, x):
        return self.t.update(b, op, x)

    def for_each(self):
        k, v = self.t.dump()
        return k, v, 1

    def stats(self):
        return {"capacity": self.sched.cap * 32}

    def size(self):
        return self.t.size()

    def lookup(self, Q):
        f, v, fw = self.t.lookup(Q)
        return f, (v if self.width else None), fw

    def count(self, Q):
        return self.t.counts(Q)[0]

    def sum(self, Q):
        return self.t.counts(Q)[1]

    def count_masked(self, Q, mf, mr):
        m = (1 << (8 * self.width)) - 1
        return self.t.counts(Q, mf & m, mr & m)[2]

    def reduce(self, what, mask, workers=1):
        return self.t.reduce(what, mask)

    def get(self, kmer):
        import oracle as O
        valid, key, _, _, _ = O.describe(bytes(kmer[:31]))
        if not valid[0]:
            return False, 0
        return self.t.get(int(key[0]))

    def close(self):
        pass
