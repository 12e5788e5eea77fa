"""Parity of the MI355X table with the oracle and the reference, through the
C ABI (and the device-text extensions), on seeded inputs.  Bar: bit-exact
keys, values, orientations, per-call return values, per-position lookups,
per-sequence counts/sums, and capacities."""
import hashlib
import json
import os
import threading

import numpy as np
import pytest

import kache_hash_amd as KA
import oracle as O
import scenarios
import seqgen as G
from kache_hash_amd import capi

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


@pytest.fixture(scope="module")
def api():
    return KA.api()


def digest(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def test_device_describer_matches_reference_vectors(api):
    d = np.load(os.path.join(GOLD, "describer.npz"), allow_pickle=False)
    lens, text = d["lens"], d["text"].tobytes()
    seqs, o = [], 0
    for l in lens:
        seqs.append(text[o:o + int(l)])
        o += int(l)
    starts = np.cumsum([0] + [len(s) + 1 for s in seqs])[:-1]
    dt, n = capi.device_text(api, seqs)
    bk, bh = capi.DeviceBuffer(api, 8 * n), capi.DeviceBuffer(api, 8 * n)
    bc, bf, bv = (capi.DeviceBuffer(api, n) for _ in range(3))
    api.check(api.lib.kache_amd_describe_device(dt.ptr, n, bk.ptr, bh.ptr, bc.ptr, bf.ptr, bv.ptr), "describe")
    key, h = bk.download(dtype=np.uint64), bh.download(dtype=np.uint64)
    c, fwd, valid = bc.download(), bf.download(), bv.download()
    pos = starts[d["q"]] + d["i"].astype(np.int64)
    assert int(valid.sum()) == len(pos) and valid[pos].all()
    assert (key[pos] == d["key"]).all() and (h[pos] == d["h"]).all()
    assert (c[pos] == d["c_wang"]).all() and (fwd[pos] == d["fwd"]).all()


def device_hashes(api, x, cap):
    """kache_amd_hashes_device over x: one row of 10 words per input."""
    x = np.ascontiguousarray(x, np.uint64)
    bx, bo = capi.DeviceBuffer(api, 8 * len(x)), capi.DeviceBuffer(api, 80 * len(x))
    bx.upload(x)
    api.check(api.lib.kache_amd_hashes_device(bx.ptr, len(x), cap, bo.ptr), "kache_amd_hashes_device")
    return bo.download(dtype=np.uint64).reshape(-1, 10)


def test_device_hashes_match_reference_vectors(api):
    """The device builds of the Wang hash, the checksum, the minimizer hash,
    XXH3's 8-byte path with both seeds, bucket_of, filter_pos, pick_bucket
    and the secondaries, bit for bit against vectors the reference's own code
    wrote (tests/golden/gen_golden.cpp)."""
    s = np.load(os.path.join(GOLD, "scalars.npz"), allow_pickle=False)
    p = np.load(os.path.join(GOLD, "placement.npz"), allow_pickle=False)
    print("device hashes:", len(s["x"]), "scalars,", len(p["h"]), "hashes x", len(p["caps"]), "capacities", flush=True)
    o = device_hashes(api, s["x"], 1000)
    assert (o[:, 0] == s["wang"]).all()
    assert (o[:, 1] == s["c_wang"]).all()
    assert (o[:, 2] == s["minimizer"]).all()
    assert (o[:, 3] == s["xxh3_s0"]).all() and (o[:, 3] == s["xxh3"]).all()
    assert (o[:, 4] == s["xxh3_s63"]).all()
    for j, cap in enumerate(p["caps"]):
        o = device_hashes(api, p["h"], int(cap))
        assert (o[:, 5] == p["bucket_of"][:, j]).all(), int(cap)
        assert (o[:, 6] == p["filter_pos"]).all()
        assert (o[:, 7] == p["pick_bucket"][:, j]).all(), int(cap)
    for t in range(len(p["tab_n"])):
        sec = p["sec"][t]
        o = device_hashes(api, sec[:, 0], int(p["tab_buckets"][t]))
        assert (o[:, 8] == sec[:, 1]).all() and (o[:, 9] == sec[:, 2]).all()


GOLDEN = {g["name"]: g for g in json.load(open(os.path.join(GOLD, "ref_tables.json")))["scenarios"]}


@pytest.mark.parametrize("name", [s["name"] for s in scenarios.SCENARIOS])
def test_scenario_matches_oracle_and_reference(api, name):
    s = next(x for x in scenarios.SCENARIOS if x["name"] == name)
    S, Q = scenarios.inputs(s)
    print(f"scenario {name}: {len(S)} inserts, {len(Q)} queries", flush=True)
    a = scenarios.run(capi.Table, api, s, S, Q)
    b = scenarios.run(scenarios.OracleTable, None, s, S, Q)
    g = GOLDEN[name]
    assert a["inserted"] == b["inserted"] == g["inserted"]
    assert a["capacity"] == b["capacity"] == g["capacity"]
    assert a["size"] == b["size"] == g["size"]
    assert (a["keys"] == b["keys"]).all() and (a["values"] == b["values"]).all()
    assert digest(a["keys"], a["values"]) == g["dump"]
    for x, y in zip(a["lookup"], b["lookup"]):
        assert (x == y).all()
    assert digest(*a["lookup"]) == g["lookup"]
    assert (a["counts"] == b["counts"]).all() and digest(a["counts"]) == g["counts"]
    if s["width"]:
        assert (a["sums"] == b["sums"]).all() and (a["masked"] == b["masked"]).all()
    assert a["reduce"] == b["reduce"] == g["reduce"]
    assert (a["get"][0] == b["get"][0]).all() and (a["get"][1] == b["get"][1]).all()
    assert digest(*a["get"]) == g["get"]
    if g["present"] is not None:
        assert a["present"] == b["present"] == g["present"]
        assert (a["values2"] == b["values2"]).all() and digest(a["keys2"], a["values2"]) == g["dump_after_update"]
    st = a["stats"]
    assert abs(st["primary_pct"] + st["secondary_pct"] + st["overflow_pct"] - 100) < 1e-6
    if s["groups"] and s["sizing"] != "tiny":
        assert st["overflow_keys"] > 0      # the minimizer groups reached the overflow table


def test_known_answers(api):
    """test/capi.cpp:429-480."""
    for pow2, slots in ((False, 1280), (True, 2048)):
        with capi.Table(api, 0, 1000, 0.8, pow2, 2) as t:
            assert t.stats()["capacity"] == slots
    with capi.Table(api, 1, 1000, 0.8, False, 2) as t:
        assert [t.register() for _ in range(128)] == list(range(128))
        with pytest.raises(capi.KacheError) as e:
            t.register()
        assert e.value.status == capi.KACHE_EUSERS
        one = [b"ATGCGTACCTTAGGCATCCAGTTAGACGGATCA"]
        L = api.lib
        import ctypes as C
        b = capi.Batch(one)
        n = C.c_uint64()
        assert L.kache_insert(t.h, 128, b.c_seqs, b.c_lens, 1, capi.ADD, 1, None) == capi.KACHE_EINVAL
        assert L.kache_insert(t.h, 0, b.c_seqs, b.c_lens, 1, 6, 1, None) == capi.KACHE_EINVAL
        assert L.kache_insert(t.h, 0, b.c_seqs, b.c_lens, 1, capi.ADD, 256, None) == capi.KACHE_EINVAL
        assert L.kache_insert(t.h, 0, None, b.c_lens, 1, capi.ADD, 1, None) == capi.KACHE_EINVAL
        nulls = (C.c_char_p * 1)(None)
        five, zero = (C.c_size_t * 1)(5), (C.c_size_t * 1)(0)
        assert L.kache_insert(t.h, 0, nulls, five, 1, capi.ADD, 1, None) == capi.KACHE_EINVAL
        assert L.kache_insert(t.h, 0, nulls, zero, 1, capi.ADD, 1, C.byref(n)) == capi.KACHE_OK and n.value == 0
        for op in (capi.KEEP, capi.MIN):
            assert L.kache_insert_oriented(t.h, 0, b.c_seqs, b.c_lens, 1, op, 1, 2, None) == capi.KACHE_EINVAL
        assert L.kache_insert_oriented(t.h, 0, b.c_seqs, b.c_lens, 1, capi.OR, 1, 256, None) == capi.KACHE_EINVAL
        assert t.insert(one, capi.ADD, 255) == 3
        assert t.update(one, capi.ADD, 1) == 3
        assert t.get(one[0]) == (True, 255)       # saturated
        assert L.kache_count(t.h, b.c_seqs, b.c_lens, 1, None) == capi.KACHE_EINVAL
        assert L.kache_lookup(t.h, b.c_seqs, b.c_lens, 1, None, None, None) == capi.KACHE_EINVAL
        assert L.kache_reduce(t.h, 3, (1 << 64) - 1, 1, C.byref(n)) == capi.KACHE_EINVAL
        assert L.kache_reduce(t.h, 2, 1, 0, C.byref(n)) == capi.KACHE_EINVAL
        assert L.kache_for_each(t.h, 2, 2, capi.KEYS_FN(lambda *a: 0), None) == capi.KACHE_EINVAL
    with capi.Table(api, 0, 1000, 0.8, False, 2) as t:
        u = t.register()
        one = [b"ATGCGTACCTTAGGCATCCAGTTAGACGGATCA"]
        assert t.insert(one, 13, (1 << 64) - 1, u) == 3       # a set ignores the operation
        with pytest.raises(capi.KacheError):
            t.insert_oriented(one, capi.OR, 1, 2, u)
        with pytest.raises(capi.KacheError):
            t.update(one, capi.ADD, 1)
        with pytest.raises(capi.KacheError):
            t.sum(one)
        assert t.reduce(capi.REDUCE_KEYS, 0) == 3
        with pytest.raises(capi.KacheError):
            t.reduce(capi.REDUCE_POPCOUNT, 1)


def test_oriented_assign_order_alternatives(api):
    """test/capi.cpp:536-564: a key whose last call had it in both
    orientations may keep either operand."""
    g = G.genome(40_000, 21)
    S = G.mixed_inserts(g, 22, 40_000)
    bits = 16
    with capi.Table(api, 2, 10, 0.8, False, 2) as t:
        u = t.register()
        ot = O.OTable(2)
        last = {}
        for b, batch in enumerate(G.ragged_batches(S)):
            x, xr = scenarios.operand(capi.ASSIGN, b, bits), scenarios.operand(capi.ASSIGN, b + 977, bits)
            assert t.insert_oriented(batch, capi.ASSIGN, x, xr, u) == ot.insert_oriented(batch, capi.ASSIGN, x, xr)
            for s in batch:
                v, key, h, c, fwd = O.describe(s)
                for k, f in zip(key[v == 1], fwd[v == 1]):
                    lb, lo = last.get(int(k), (-1, 0))
                    if lb != b:
                        lb, lo = b, 0
                    last[int(k)] = (lb, lo | (1 if f else 2))
        k1, v1, _ = t.for_each()
        k2, v2 = ot.dump()
        assert (k1 == k2).all()
        for k, a, e in zip(k1, v1, v2):
            if a != e:
                lb, lo = last[int(k)]
                assert lo == 3
                alt = {scenarios.operand(capi.ASSIGN, lb, bits), scenarios.operand(capi.ASSIGN, lb + 977, bits)}
                assert int(a) in alt


def test_forward_strand_membership(api):
    """test/capi.cpp:640-706 (K2Rmini): OR 1 / 2 by orientation from a tiny table."""
    g = G.genome(80_000, 31)
    P = G.pieces(g, 1024)
    P = [G.revcomp(p) if i % 3 == 0 else p for i, p in enumerate(P)]
    F = set()
    for s in P:
        v, key, h, c, fwd = O.describe(s)
        for k, f in zip(key[v == 1], fwd[v == 1]):
            F.add(int(k) if f else O.revcomp31(int(k)))
    with capi.Table(api, 1, 10, 0.8, False, 4) as t:
        u = t.register()
        for batch in G.ragged_batches(P):
            t.insert_oriented(batch, capi.OR, 1, 2, u)
        assert t.reduce(capi.REDUCE_POPCOUNT, 3) == len(F)
        M = P + [G.revcomp(P[i]) for i in range(0, len(P), 5)] + G.pieces(G.genome(20_000, 32), 1024) + [b"", b"ACGT", b"N" * 40]
        counts = t.count_masked(M, 1, 2)
        for s, c in zip(M, counts):
            v, key, h, cc, fwd = O.describe(s)
            ref = sum((int(k) if f else O.revcomp31(int(k))) in F for k, f in zip(key[v == 1], fwd[v == 1]))
            assert c == ref
        st = t.stats()
        assert 0 < st["keys"] <= len(F)


def test_device_entry_points_match_host_entry_points(api):
    g = G.genome(100_000, 41)
    R = G.reads(g, 20_000, 150, 42, sub=0.01, nrate=0.001, lower=0.1)
    text, n = capi.device_text(api, R)
    with capi.Table(api, 8, 1000, 0.8, False, 4) as a, capi.Table(api, 8, 1000, 0.8, False, 4) as b:
        ua, ub = a.register(), b.register()
        ia = a.insert(R[:10000], capi.ADD, 1, ua) + a.insert(R[10000:], capi.ADD, 1, ua)
        ib = b.insert_device(text.ptr, n, capi.ADD, 1, ub)
        assert ia == ib
        ka, va, _ = a.for_each()
        kb, vb, _ = b.for_each()
        assert (ka == kb).all() and (va == vb).all()
        assert a.stats()["capacity"] == b.stats()["capacity"]
        assert int(a.count(R).sum()) == b.count_device(text.ptr, n)
        assert a.update(R, capi.ADD, 2) == b.update_device(text.ptr, n, capi.ADD, 2)
        ka, va, _ = a.for_each()
        kb, vb, _ = b.for_each()
        assert (va == vb).all()


def test_concurrent_users_equal_sequential(api):
    """Inserts from several host threads, each its own user (test/capi.cpp:156-182)."""
    g = G.genome(120_000, 51)
    S = G.mixed_inserts(g, 52, 120_000)
    B = G.ragged_batches(S)
    T = 6
    with capi.Table(api, 4, 10, 0.8, False, T) as t:
        users = [t.register() for _ in range(T)]
        tot = [0] * T

        def work(i):
            for b in range(i, len(B), T):
                tot[i] += t.insert(B[b], capi.ADD, 3, users[i])

        th = [threading.Thread(target=work, args=(i,)) for i in range(T)]
        [x.start() for x in th]
        [x.join() for x in th]
        ot = O.OTable(4)
        for b in B:
            ot.insert(b, capi.ADD, 3)
        k1, v1, _ = t.for_each()
        k2, v2 = ot.dump()
        assert sum(tot) == len(k2) == t.size()
        assert (k1 == k2).all() and (v1 == v2).all()


def test_shares_partition_the_table(api):
    g = G.genome(60_000, 61)
    S = G.mixed_inserts(g, 62, 60_000)
    for s in G.minimizer_groups(3, 160, 7):
        S += s
    with capi.Table(api, 4, 10, 0.8, False, 2) as t:
        u = t.register()
        for b in G.ragged_batches(S):
            t.insert(b, capi.OR, 5, u)
        k, v, _ = t.for_each()
        parts = [t.for_each(7, w)[0] for w in range(7)]
        allk = np.sort(np.concatenate(parts))
        assert (allk == k).all() and len(np.unique(allk)) == len(allk) == t.size()
        _, _, chunks = t.for_each(1, 0, stop_after=1)
        assert chunks == 1
        st = t.stats()
        assert st["overflow_keys"] >= 0 and abs(st["primary_pct"] + st["secondary_pct"] + st["overflow_pct"] - 100) < 1e-6


def test_count_map_at_scale_properties(api):
    """A count map over 1M reads (the count-map workload's read model), by
    device texts in batches of 65,536 reads: sum of counts = valid
    positions, keys = the oracle's distinct count on the same reads, every
    key present, and the growth schedule's capacity."""
    g = G.genome(2_000_000, 71)
    R = G.reads(g, 1_000_000, 150, 72, sub=0.005, nrate=0.0005, lower=0.05)
    ot = O.OTable(0)
    sched = O.Sched(1_000_000, 0.8, False)
    with capi.Table(api, 8, 1_000_000, 0.8, False, 4) as t:
        u = t.register()
        total_pos = 0
        for i in range(0, len(R), 65536):
            batch = R[i:i + 65536]
            text, n = capi.device_text(api, batch)
            a = t.insert_device(text.ptr, n, capi.ADD, 1, u)
            b = ot.insert(batch)
            assert a == b
            sched.call(0, b)
            hits, pos = t.count_device(text.ptr, n, positions=True)
            assert hits == pos
            total_pos += pos
        k, v, _ = t.for_each()
        assert len(k) == ot.size() == t.size()
        assert int(v.sum()) == total_pos
        assert (np.diff(k.astype(np.float64)) > 0).all() or len(np.unique(k)) == len(k)
        assert t.stats()["capacity"] == sched.cap * 32
        assert t.reduce(capi.REDUCE_SUM, (1 << 64) - 1) == total_pos


def test_synthetic_reads_match_host_generator(api):
    """The device read generator against its host run, read for read."""
    g = capi.DeviceBuffer(api, 1_000_000)
    api.check(api.lib.kache_amd_synth_genome(g.ptr, 1_000_000, 99), "genome")
    n, L = 5000, 150
    r = capi.DeviceBuffer(api, n * (L + 1))
    api.check(api.lib.kache_amd_synth_reads(r.ptr, n, L, 99, 1_000_000, 7, 0.01, 0.001, 0.5), "reads")
    dev = r.download()
    host = np.zeros(300 * (L + 1), np.uint8)
    api.check(api.lib.kache_amd_host_synth_reads(host.ctypes.data, 1234, 300, L, 99, 1_000_000, 7, 0.01, 0.001, 0.5), "host reads")
    assert (dev[1234 * (L + 1):1534 * (L + 1)] == host).all()
    rows = dev.reshape(n, L + 1)
    assert (rows[:, L] == ord("\n")).all()
    upper = rows[:, :L] & 0xDF
    assert np.isin(upper, np.frombuffer(b"ACGTN", np.uint8)).all()
    frac_n = (upper == ord("N")).mean()
    assert 0.0005 < frac_n < 0.002


@pytest.mark.parametrize("expected", [10, 2_000_000])
def test_pipelined_batches_equal_one_call_per_batch(api, expected):
    """kache_amd_insert_batches against one kache_amd_insert_device call per
    batch: per-batch counts, contents and capacity, from a tiny table that
    doubles many times (the pipeline drains before each doubling) and a
    presized one; batch starts at any byte offset."""
    g = G.genome(300_000, 81)
    R = G.reads(g, 40_000, 150, 82, sub=0.01, nrate=0.001, lower=0.1)
    rng = np.random.default_rng(83)
    cuts = np.sort(rng.choice(np.arange(1, len(R)), 60, replace=False))
    batches = np.split(np.arange(len(R)), cuts)
    text = b"".join(r + b"\n" for r in R)
    ends = np.cumsum([len(r) + 1 for r in R])
    bounds = np.array([0] + [int(ends[b[-1]]) for b in batches], np.uint64)
    buf, n = capi.device_text(api, R)
    with capi.Table(api, 4, expected, 0.8, False, 4) as a, capi.Table(api, 4, expected, 0.8, False, 4) as b:
        ua, ub = a.register(), b.register()
        got = a.insert_batches(buf.ptr, bounds, capi.ADD, 2, ua)
        ref = []
        for i in range(len(bounds) - 1):
            lo, hi = int(bounds[i]), int(bounds[i + 1])
            sub, m = capi.device_text(api, [text[lo:hi]])
            ref.append(b.insert_device(sub.ptr, m, capi.ADD, 2, ub))
        assert got.tolist() == ref
        ka, va, _ = a.for_each()
        kb, vb, _ = b.for_each()
        assert (ka == kb).all() and (va == vb).all()
        assert a.stats()["capacity"] == b.stats()["capacity"]
        hits, pos = a.count_batches(buf.ptr, bounds)
        for i in range(len(bounds) - 1):
            lo, hi = int(bounds[i]), int(bounds[i + 1])
            sub, m = capi.device_text(api, [text[lo:hi]])
            h2, p2 = b.count_device(sub.ptr, m, positions=True)
            assert (int(hits[i]), int(pos[i])) == (h2, p2) and h2 == p2


def test_count_batches_with_bounds_anywhere(api):
    """kache_amd_count_batches runs up to 256 batches in one launch; each
    batch's hits and k-mer positions must equal the oracle's count of its
    own slice of the text (a k-mer counts only within one batch), with bounds
    at arbitrary byte offsets (k-mers cut by them), empty batches, batches
    shorter than a k-mer or a 64-position wave, a first bound neither 0 nor
    16-aligned, and several launch groups (more than 256 batches)."""
    g = G.genome(400_000, 91)
    R = G.reads(g, 20_000, 150, 92, sub=0.01, nrate=0.001, lower=0.2)
    Q = G.reads(g, 6_000, 150, 93, sub=0.02, nrate=0.002, lower=0.2) + G.pieces(G.genome(20_000, 94), 300)
    text = b"".join(q + b"\n" for q in Q)
    rng = np.random.default_rng(95)
    start = 13
    cuts = rng.integers(start, len(text), 600)
    tiny = start + 50_000 + np.cumsum(rng.integers(0, 40, 200))
    bounds = np.sort(np.concatenate([[start, len(text)], cuts, cuts[:20], tiny])).astype(np.uint64)
    print(f"count_batches: {len(bounds) - 1} batches over {len(text)} bytes", flush=True)
    arr = np.frombuffer(text, np.uint8)
    base = np.isin(arr | 0x20, np.frombuffer(b"acgt", np.uint8)).astype(np.int64)
    run = np.concatenate([[0], np.cumsum(base)])
    ok = np.zeros(len(arr) + 1, np.int64)
    ok[:len(arr) - 30] = (run[31:] - run[:-31]) == 31            # a whole k-mer starts at p
    P = np.concatenate([[0], np.cumsum(ok)])
    t = capi.Table(api, 8, expected_keys=1_000_000)
    ot = O.OTable(8)
    u = t.register()
    assert t.insert(R, capi.ADD, 1, u) == ot.insert(R, capi.ADD, 1)
    buf, _ = capi.device_text(api, [text])
    hits, pos = t.count_batches(buf.ptr, bounds)
    for i in range(len(bounds) - 1):
        lo, hi = int(bounds[i]), int(bounds[i + 1])
        want_p = int(P[hi - 30] - P[lo]) if hi - lo >= 31 else 0
        want_h = int(ot.counts([text[lo:hi]])[0][0]) if hi > lo else 0
        assert (int(hits[i]), int(pos[i])) == (want_h, want_p), (i, lo, hi)
    assert int(pos.sum()) > 0 and 0 < int(hits.sum()) < int(pos.sum())
    h1, p1 = t.count_batches(buf.ptr, np.array([0, len(text)], np.uint64))
    assert int(h1[0]) == int(ot.counts([text])[0][0]) and int(p1[0]) == int(P[len(text) - 30])
    h0, p0 = t.count_batches(buf.ptr, np.array([5, 5, 5], np.uint64))
    assert h0.tolist() == [0, 0] and p0.tolist() == [0, 0]
    t.close()
