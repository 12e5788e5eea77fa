"""Seeded random programs of C-ABI calls, run through a table of any C-ABI
library (the device library, the reference library built from its sources)
and the oracle, with equality asserted at every step.

Each case draws a width, a sizing (from a tiny table that doubles many times
to an exact one, plain or power-of-2 capacity), one to three users, and a
program of calls over random sequences: pieces of a genome in either strand,
lengths 0-600, runs of N, mixed case, tandem repeats (the same k-mer many
times in one call) and minimizer groups (keys that overflow their buckets
into the secondaries and the overflow table).  Calls: inserts with every
operation and random operands (saturating ones included), oriented inserts,
updates, lookups, counts, sums, masked counts, reductions, get, size.  Every
return value, every query output, the capacity after every insert (against
the reference's growth schedule, include/Streaming_Kmer_Hash_Table.hpp:
224-237, 2246-2286) and the final contents must equal the oracle's.
Oriented ASSIGN is left out: within one call its result may be either
operand (test/capi.cpp:550-564)."""
import numpy as np

import oracle as O
import seqgen as G
from kache_hash_amd import capi

CASES = 128
DEVICE_CASES = 64
KEEP, ASSIGN, ADD, OR, MIN, MAX = range(6)


def _noise(rng, s):
    """Runs of N and lowercase stretches over a sequence."""
    b = bytearray(s)
    for _ in range(int(rng.integers(0, 3))):
        if len(b) > 2:
            i = int(rng.integers(0, len(b)))
            j = min(len(b), i + int(rng.integers(1, 4)))
            b[i:j] = b"N" * (j - i)
    if len(b) and rng.random() < 0.3:
        i = int(rng.integers(0, len(b)))
        j = min(len(b), i + int(rng.integers(1, 80)))
        b[i:j] = bytes(b[i:j]).lower()
    return bytes(b)


def sequences(rng, genome, groups, n):
    out = []
    for _ in range(n):
        r = rng.random()
        if r < 0.08:
            out.append(b"" if rng.random() < 0.5 else b"ACGTN"[: int(rng.integers(0, 5))] * int(rng.integers(1, 6)))
        elif r < 0.16:
            unit = genome[int(rng.integers(0, len(genome) - 50)):][: int(rng.integers(20, 45))]
            out.append(unit * int(rng.integers(3, 40)))
        elif r < 0.24 and groups:
            g = groups[int(rng.integers(0, len(groups)))]
            x = g[int(rng.integers(0, len(g)))]
            out.append(G.revcomp(x) if rng.random() < 0.5 else x)
        else:
            ln = int(rng.integers(1, 600))
            a = int(rng.integers(0, max(1, len(genome) - ln)))
            s = genome[a:a + ln]
            if rng.random() < 0.5:
                s = G.revcomp(s)
            out.append(_noise(rng, s))
    return out


def operand(rng, op, vmax):
    if op == ADD:
        return int(rng.choice([1, 1 + int(rng.integers(0, 7)), vmax // 2 + 1, vmax]))
    if op == OR:
        return 1 << int(rng.integers(0, max(1, vmax.bit_length())))
    return int(rng.integers(0, vmax, endpoint=True, dtype=np.uint64))


def same_contents(t, o):
    k, v, _ = t.for_each()
    ko, vo = o.dump()
    assert np.array_equal(k, ko)
    if t.width:
        assert np.array_equal(v.astype(np.uint64), vo.astype(np.uint64))


def set_growth(api, t, case):
    """On the device library, how case `case`'s table doubles (the random
    stream is left alone): automatically, in place at every size (every
    other such case with an away list of 256 entries, which must grow), or
    out of place."""
    if not hasattr(api.lib, "kache_amd_set_growth"):
        return
    mode = case % 3
    t.set_growth(mode, 256 if mode == capi.GROW_IN_PLACE and case % 2 else 0)


def run_program(api, case):
    """Runs case `case` through a table of the library `api` (a capi.Api)
    and the oracle, asserting equality at every step."""
    rng = np.random.default_rng(0xC0FFEE + case)
    width = int(rng.choice([0, 1, 2, 4, 8]))
    vmax = (1 << (8 * width)) - 1 if width else 0
    expected = int(rng.choice([10, 300, 5_000, 40_000]))
    pow2 = bool(rng.random() < 0.3)
    users = int(rng.integers(1, 4))
    genome = G.genome(int(rng.integers(3_000, 40_000)), 7000 + case)
    groups = G.minimizer_groups(3, 120, 8000 + case) if rng.random() < 0.5 else []
    calls = int(rng.integers(6, 14))
    print(f"random case {case}: width {width}, expected {expected}, pow2 {pow2}, users {users}, "
          f"genome {len(genome)}, groups {bool(groups)}, calls {calls}", flush=True)
    with capi.Table(api, width, expected, 0.8, pow2, 2) as t:
        set_growth(api, t, case)
        o = O.OTable(width)
        sched = O.Sched(expected, 0.8, pow2)
        us = [t.register() for _ in range(users)]
        for step in range(calls):
            seqs = sequences(rng, genome, groups, int(rng.integers(1, 60)))
            kind = rng.random()
            u = us[int(rng.integers(0, users))]
            if kind < 0.45 or (width == 0 and kind < 0.7):
                op = int(rng.integers(0, 6)) if width else KEEP
                x = operand(rng, op, vmax) if width else 0
                got = t.insert(seqs, op, x, u)
                want = o.insert(seqs, op, x)
                sched.call(u, want)
                assert got == want, (step, "insert", op, x)
                assert t.stats()["capacity"] == sched.cap * 32, (step, "capacity")
            elif kind < 0.6 and width:
                op = int(rng.choice([ADD, OR, MAX]))
                xf, xr = operand(rng, op, vmax), operand(rng, op, vmax)
                got = t.insert_oriented(seqs, op, xf, xr, u)
                want = o.insert_oriented(seqs, op, xf, xr)
                sched.call(u, want)
                assert got == want, (step, "oriented", op, xf, xr)
                assert t.stats()["capacity"] == sched.cap * 32, (step, "capacity")
            elif kind < 0.7 and width:
                op = int(rng.integers(0, 6))
                x = operand(rng, op, vmax)
                assert t.update(seqs, op, x) == o.update(seqs, op, x), (step, "update", op, x)
            else:
                f, v, fw = t.lookup(seqs)
                f2, v2, fw2 = o.lookup(seqs)
                assert np.array_equal(f, f2) and np.array_equal(fw, fw2), (step, "lookup")
                if width:
                    assert np.array_equal(v.astype(np.uint64), v2.astype(np.uint64)), (step, "lookup values")
                c, sm, m = o.counts(seqs, 0x5555555555555555 & vmax, 0xA3A3A3A3A3A3A3A3 & vmax)
                assert np.array_equal(t.count(seqs), c), (step, "count")
                if width:
                    assert np.array_equal(t.sum(seqs), sm), (step, "sum")
                    assert np.array_equal(t.count_masked(seqs, 0x5555555555555555, 0xA3A3A3A3A3A3A3A3), m), (step, "masked")
            assert t.size() == o.size(), (step, "size")
        same_contents(t, o)
        for what in ((0, 1, 2) if width else (0,)):
            for mask in ((vmax, 0x5A5A5A5A5A5A5A5A & vmax) if width else ((1 << 64) - 1,)):
                assert t.reduce(what, mask) == o.reduce(what, mask), ("reduce", what, mask)
        for s in sequences(rng, genome, groups, 20):
            if len(s) >= 31:
                valid, key, _, _, _ = O.describe(s[:31])
                want = o.get(int(key[0])) if valid[0] else (False, 0)
                assert t.get(s[:31]) == want, ("get", s[:31])


def _offsets(seqs):
    off = np.zeros(len(seqs) + 1, np.uint64)
    np.cumsum([len(s) + 1 for s in seqs], out=off[1:])
    return off


def run_device_program(api, case):
    """A seeded random program through the device library's own entry points
    (include/kache_amd.h): device-text inserts, oriented inserts, pipelined
    batches, updates and counts (totals, per batch, per sequence), and host
    texts; each against the oracle as the calls of the C ABI it stands for (a
    batch is one call of the growth schedule)."""
    import ctypes as C
    rng = np.random.default_rng(0xD5EED + case)
    width = int(rng.choice([0, 1, 2, 4, 8]))
    vmax = (1 << (8 * width)) - 1 if width else 0
    expected = int(rng.choice([10, 300, 5_000, 40_000]))
    pow2 = bool(rng.random() < 0.3)
    genome = G.genome(int(rng.integers(3_000, 40_000)), 9000 + case)
    groups = G.minimizer_groups(3, 120, 9500 + case) if rng.random() < 0.5 else []
    calls = int(rng.integers(6, 14))
    print(f"device case {case}: width {width}, expected {expected}, pow2 {pow2}, genome {len(genome)}, "
          f"groups {bool(groups)}, calls {calls}", flush=True)
    L = api.lib
    with capi.Table(api, width, expected, 0.8, pow2, 2) as t:
        set_growth(api, t, case)
        o = O.OTable(width)
        sched = O.Sched(expected, 0.8, pow2)
        u = t.register()
        for step in range(calls):
            seqs = sequences(rng, genome, groups, int(rng.integers(1, 60)))
            buf, nbytes = capi.device_text(api, seqs)
            off = _offsets(seqs)
            kind = rng.random()
            if kind < 0.2:
                op = int(rng.integers(0, 6)) if width else KEEP
                x = operand(rng, op, vmax) if width else 0
                want = o.insert(seqs, op, x)
                sched.call(u, want)
                assert t.insert_device(buf.ptr, nbytes, op, x, u) == want, (step, "insert_device")
            elif kind < 0.35:
                # Pipelined batches: bounds at random sequence starts.
                cuts = np.unique(rng.integers(0, len(seqs) + 1, size=int(rng.integers(0, 6))))
                idx = [0] + [int(c) for c in cuts if 0 < c < len(seqs)] + [len(seqs)]
                bounds = off[idx]
                op = int(rng.integers(0, 6)) if width else KEEP
                x = operand(rng, op, vmax) if width else 0
                got = t.insert_batches(buf.ptr, bounds, op, x, u)
                for i in range(len(idx) - 1):
                    want = o.insert(seqs[idx[i]:idx[i + 1]], op, x)
                    sched.call(u, want)
                    assert int(got[i]) == want, (step, "insert_batches", i)
            elif kind < 0.45 and width:
                op = int(rng.choice([ADD, OR, MAX]))
                xf, xr = operand(rng, op, vmax), operand(rng, op, vmax)
                want = o.insert_oriented(seqs, op, xf, xr)
                sched.call(u, want)
                assert t.insert_oriented_device(buf.ptr, nbytes, op, xf, xr, u) == want, (step, "oriented_device")
            elif kind < 0.55:
                op = int(rng.integers(0, 6)) if width else KEEP
                x = operand(rng, op, vmax) if width else 0
                text = b"".join(s + b"\n" for s in seqs)
                want = o.insert(seqs, op, x)
                sched.call(u, want)
                assert t.insert_text(text, op, x, u) == want, (step, "insert_text")
            elif kind < 0.65 and width:
                op = int(rng.integers(0, 6))
                x = operand(rng, op, vmax)
                want = o.update(seqs, op, x)
                if rng.random() < 0.5:
                    assert t.update_device(buf.ptr, nbytes, op, x) == want, (step, "update_device")
                else:
                    assert t.update_text(b"".join(s + b"\n" for s in seqs), op, x) == want, (step, "update_text")
            else:
                counts = o.counts(seqs)[0]
                assert t.count_device(buf.ptr, nbytes) == int(counts.sum()), (step, "count_device")
                per = np.zeros(max(len(seqs), 1), np.uint64)
                api.check(L.kache_amd_count_seqs_device(t.h, buf.ptr, int(off[-1]), off.ctypes.data, len(seqs), per.ctypes.data),
                          "kache_amd_count_seqs_device")
                assert np.array_equal(per[:len(seqs)], counts), (step, "count_seqs_device")
                h, p = t.count_batches(buf.ptr, np.array([0, off[-1]], np.uint64))
                assert int(h[0]) == int(counts.sum()), (step, "count_batches")
            buf.free()
            assert t.stats()["capacity"] == sched.cap * 32, (step, "capacity")
            assert t.size() == o.size(), (step, "size")
        same_contents(t, o)
