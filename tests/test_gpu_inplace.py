"""Doubling in place (kache_amd_set_growth): the store's address range is
mapped to twice its size and split from the top, the keys away from their
primary bucket go through the away list.  Every dump, value, lookup, per-call
return and capacity must equal the oracle's (and its growth schedule's), and
equal what the same calls give with out-of-place doublings; an away list too
small for the keys it must hold is grown mid-doubling without changing a
result; an in-place doubling needs less HBM beside the table than an
out-of-place one (include/Streaming_Kmer_Hash_Table.hpp:1473-1681 doubles its
store in place too)."""
import numpy as np
import pytest

import kache_hash_amd as KA
import oracle as O
import seqgen as G
from kache_hash_amd import capi

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def api():
    return KA.api()


def run(api, width, reads, queries, mode, away=0, expected=20_000, users=2, op=None, x=1):
    """Inserts `reads` in calls of 2,000 by alternating users; returns the
    per-call returns, capacities after each call, dump, lookups and info."""
    op = op if op is not None else (capi.ADD if width else capi.KEEP)
    t = capi.Table(api, width, expected_keys=expected)
    t.set_growth(mode, away)
    us = [t.register() for _ in range(users)]
    rets, caps = [], []
    for i in range(0, len(reads), 2000):
        rets.append(t.insert(reads[i:i + 2000], op, x, us[(i // 2000) % users]))
        caps.append(t.stats()["capacity"])
    k, v, _ = t.for_each()
    look = t.lookup(queries)
    counts = t.count(queries)
    info = t.info()
    t.close()
    return rets, caps, (k, v), look, counts, info


def oracle_run(width, reads, queries, expected=20_000, users=2, op=None, x=1):
    op = op if op is not None else (capi.ADD if width else capi.KEEP)
    ot, sched = O.OTable(width), O.Sched(expected)
    rets, caps = [], []
    for i in range(0, len(reads), 2000):
        r = ot.insert(reads[i:i + 2000], op, x)
        sched.call((i // 2000) % users, r)
        rets.append(r)
        caps.append(sched.cap * 32)
    return rets, caps, ot.dump(), ot.lookup(queries), ot.counts(queries)[0]


def same(a, b, width):
    assert a[0] == b[0], "per-call returns"
    assert a[1] == b[1], "capacities"
    assert np.array_equal(a[2][0], b[2][0]) and np.array_equal(a[2][1], b[2][1]), "dump"
    f1, v1, w1 = a[3]
    f2, v2, w2 = b[3]
    assert np.array_equal(f1, f2) and np.array_equal(w1, w2), "lookups"
    if width:
        assert np.array_equal(v1, v2), "lookup values"
    assert np.array_equal(a[4], b[4]), "counts"


@pytest.mark.parametrize("width", [0, 1, 4, 8])
def test_in_place_doublings_match_the_oracle(api, width):
    g = G.genome(600_000, 301 + width)
    reads = G.reads(g, 40_000, 150, 302 + width, sub=0.01, nrate=0.001, lower=0.2)
    groups = [s for grp in G.minimizer_groups(4, 150, 303 + width) for s in grp]
    reads = reads[:20_000] + groups + reads[20_000:]
    queries = G.reads(g, 600, 150, 304, sub=0.01) + groups[::7] + G.pieces(G.genome(4_000, 305), 300)
    print(f"in place, width {width}: {len(reads)} sequences from a 20k-key table", flush=True)
    got = run(api, width, reads, queries, capi.GROW_IN_PLACE)
    info = got[5]
    assert info["resizes"] >= 4 and info["in_place_doublings"] >= 4, info
    # A range reserves room for four doublings; when that is used up the next
    # doubling moves the store, out of place, into a range with room again.
    assert info["store_mapped_bytes"] >= info["buckets"] * info["bucket_bytes"]
    assert info["store_reserved_bytes"] >= info["store_mapped_bytes"]
    same(got, oracle_run(width, reads, queries), width)


@pytest.mark.parametrize("width", [0, 8])
def test_in_place_equals_out_of_place(api, width):
    g = G.genome(800_000, 311 + width)
    reads = G.reads(g, 40_000, 150, 312, sub=0.01, nrate=0.001, lower=0.1)
    queries = G.reads(g, 500, 150, 313, sub=0.01)
    print(f"in place vs out of place, width {width}", flush=True)
    a = run(api, width, reads, queries, capi.GROW_IN_PLACE)
    b = run(api, width, reads, queries, capi.GROW_OUT_OF_PLACE)
    # The first doubling moves the small (hipMalloc'd) store into a range that
    # can grow; a range's reservation holds a few doublings, then the store
    # moves again, out of place.
    assert a[5]["in_place_doublings"] >= max(3, a[5]["resizes"] // 2)
    assert b[5]["in_place_doublings"] == 0 and b[5]["resizes"] == a[5]["resizes"]
    same(a, b, width)


@pytest.mark.parametrize("width", [0, 4, 8])
def test_away_list_grows_mid_doubling(api, width):
    """An away list of 256 entries (one per region) fills in every doubling;
    each full phase is run again on a grown list, and every result stays
    exact."""
    g = G.genome(500_000, 321 + width)
    reads = G.reads(g, 30_000, 150, 322, sub=0.02, nrate=0.001, lower=0.1)
    groups = [s for grp in G.minimizer_groups(6, 140, 323 + width) for s in grp]
    reads = groups + reads
    queries = G.reads(g, 400, 150, 324, sub=0.02) + groups[::5]
    print(f"away list regrowth, width {width}", flush=True)
    got = run(api, width, reads, queries, capi.GROW_IN_PLACE, away=256, expected=8_000)
    assert got[5]["in_place_doublings"] >= 4 and got[5]["away_list_regrows"] >= 1, got[5]
    same(got, oracle_run(width, reads, queries, expected=8_000), width)


def test_in_place_oriented_and_values(api):
    """Oriented OR inserts (u8) and saturating adds (u16) through in-place doublings."""
    g = G.genome(400_000, 331)
    reads = G.reads(g, 24_000, 150, 332, sub=0.01, nrate=0.001)
    queries = G.reads(g, 300, 150, 333)
    print("in place: oriented OR u8, saturating ADD u16", flush=True)
    t = capi.Table(api, 1, expected_keys=10_000)
    t.set_growth(capi.GROW_IN_PLACE)
    ot = O.OTable(1)
    u = t.register()
    for i in range(0, len(reads), 3000):
        part = reads[i:i + 3000]
        assert t.insert_oriented(part, capi.OR, 1, 2, u) == ot.insert_oriented(part, capi.OR, 1, 2)
    k, v, _ = t.for_each()
    ko, vo = ot.dump()
    assert np.array_equal(k, ko) and np.array_equal(v, vo)
    assert t.info()["in_place_doublings"] >= 3
    t.close()
    t = capi.Table(api, 2, expected_keys=10_000)
    t.set_growth(capi.GROW_IN_PLACE)
    ot = O.OTable(2)
    u = t.register()
    for i in range(0, len(reads), 3000):
        part = reads[i:i + 3000]
        assert t.insert(part, capi.ADD, 9000, u) == ot.insert(part, capi.ADD, 9000)
    k, v, _ = t.for_each()
    ko, vo = ot.dump()
    assert np.array_equal(k, ko) and np.array_equal(v, vo)
    f1, v1, w1 = t.lookup(queries)
    f2, v2, w2 = ot.lookup(queries)
    assert (f1 == f2).all() and (v1 == v2).all() and (w1 == w2).all()
    t.close()


def test_auto_mode_doubles_in_place_when_out_of_place_does_not_fit(api):
    """A 3M-key u64 table (a store of ~54 MB).  In place, its next doubling
    needs about half the HBM an out-of-place one needs.  Automatically it
    doubles out of place while the device has room, and in place once a
    budget admits the in-place doubling but not the out-of-place one; every
    result stays equal to the oracle's."""
    g = G.genome(3_000_000, 341)
    reads = G.reads(g, 240_000, 150, 342, sub=0.005, nrate=0.0005, lower=0.1)
    queries = G.reads(g, 1_000, 150, 343, sub=0.005)
    print("auto mode on a 3M-key u64 map", flush=True)
    t = capi.Table(api, 8, expected_keys=3_000_000)
    info = t.info()
    assert info["growth_mode"] == capi.GROW_AUTO and info["store_reserved_bytes"] >= 2 * info["buckets"] * info["bucket_bytes"]
    t.set_growth(capi.GROW_IN_PLACE)
    in_need = t.info()["next_doubling_bytes"]
    t.set_growth(capi.GROW_OUT_OF_PLACE)
    out_need = t.info()["next_doubling_bytes"]
    assert in_need < 0.6 * out_need, (in_need, out_need)
    t.set_growth(capi.GROW_AUTO)
    assert t.info()["next_doubling_bytes"] == out_need     # the device has room: out of place
    ot, sched = O.OTable(8), O.Sched(3_000_000)
    u = t.register()
    doubled_in_place = False
    for i in range(0, len(reads), 5_000):
        part = reads[i:i + 5_000]
        if t.info()["resizes"] == 1 and not doubled_in_place:
            # Budget for the next doubling: in place fits, out of place does
            # not (half way between what the two need).
            inf = t.info()
            t.set_growth(capi.GROW_IN_PLACE)
            need_in = t.info()["next_doubling_bytes"]
            t.set_growth(capi.GROW_OUT_OF_PLACE)
            need_out = t.info()["next_doubling_bytes"]
            t.set_growth(capi.GROW_AUTO)
            t.set_hbm_budget(inf["device_bytes"] + (need_in + need_out) // 2)
            doubled_in_place = True
        r = t.insert(part, capi.ADD, 1, u)
        assert r == ot.insert(part, capi.ADD, 1)
        sched.call(0, r)
    inf = t.info()
    assert t.stats()["capacity"] == sched.cap * 32
    assert inf["resizes"] >= 2 and inf["in_place_doublings"] >= 1 and inf["resizes"] - inf["in_place_doublings"] >= 1, inf
    k, v, _ = t.for_each()
    ko, vo = ot.dump()
    assert np.array_equal(k, ko) and np.array_equal(v, vo)
    assert (t.count(queries) == ot.counts(queries)[0]).all()
    t.close()


def test_budget_returns_room_a_kept_range_brought(api):
    """A table grown in place and released leaves its range, mapped to its
    grown size, for the next table of a similar size; a budget set on that
    table unmaps what lies past its buckets (to within one 1-GiB chunk), and
    the table still grows in place and stays exact."""
    g = G.genome(1_000_000, 351)
    reads = G.reads(g, 60_000, 150, 352, sub=0.01, nrate=0.001)
    queries = G.reads(g, 500, 150, 353, sub=0.01)
    print("kept range, then a budget", flush=True)
    a = capi.Table(api, 8, expected_keys=200_000)
    a.set_growth(capi.GROW_IN_PLACE)
    ua = a.register()
    for i in range(0, len(reads), 10_000):
        a.insert(reads[i:i + 10_000], capi.ADD, 1, ua)
    assert a.info()["in_place_doublings"] >= 2
    a.close()
    t = capi.Table(api, 8, expected_keys=200_000)
    t.set_growth(capi.GROW_IN_PLACE)
    inf = t.info()
    logical = inf["buckets"] * inf["bucket_bytes"]
    print(f"store: {logical} bytes of buckets, {inf['store_mapped_bytes']} mapped", flush=True)
    assert inf["store_mapped_bytes"] >= logical
    t.set_hbm_budget(1 << 40)
    assert logical <= t.info()["store_mapped_bytes"] < logical + (1 << 30)
    ot, sched = O.OTable(8), O.Sched(200_000)
    u = t.register()
    for i in range(0, len(reads), 10_000):
        part = reads[i:i + 10_000]
        r = t.insert(part, capi.ADD, 1, u)
        assert r == ot.insert(part, capi.ADD, 1)
        sched.call(0, r)
    assert t.stats()["capacity"] == sched.cap * 32 and t.info()["in_place_doublings"] >= 2
    k, v, _ = t.for_each()
    ko, vo = ot.dump()
    assert np.array_equal(k, ko) and np.array_equal(v, vo)
    assert (t.count(queries) == ot.counts(queries)[0]).all()
    t.close()



def test_a_doubling_that_fails_part_way(api, monkeypatch):
    """A doubling whose away list cannot grow (refused from the host through
    KACHE_AMD_TEST_FAIL_AWAY_REGROW, as if the device had no memory left).
    Out of place the call fails with KACHE_ENOMEM and the table stays as it
    was: the same calls again give the oracle's keys.  In place the split has
    overwritten old buckets: the call fails and every later call on the
    table's keys is refused with KACHE_EINTERNAL; its info and release still
    work, and a new table is unaffected."""
    g = G.genome(300_000, 361)
    reads = G.reads(g, 24_000, 150, 362, sub=0.01, nrate=0.001)
    queries = G.reads(g, 300, 150, 363, sub=0.01)
    chunks = [reads[i:i + 2000] for i in range(0, len(reads), 2000)]
    print("a doubling that fails part way: out of place", flush=True)
    t = capi.Table(api, 0, expected_keys=8_000)
    t.set_growth(capi.GROW_OUT_OF_PLACE, 256)   # a list of 256 entries always has to grow
    u = t.register()
    ot = O.OTable(0)
    monkeypatch.setenv("KACHE_AMD_TEST_FAIL_AWAY_REGROW", "1")
    with pytest.raises(capi.KacheError) as e:
        t.insert(chunks[0], capi.KEEP, 0, u)
    monkeypatch.delenv("KACHE_AMD_TEST_FAIL_AWAY_REGROW")
    assert e.value.status == capi.KACHE_ENOMEM and "away list" in str(e.value), e.value
    assert t.info()["resizes"] == 0
    # The failed call's launches before the doubling stay applied; a set's
    # inserts are idempotent, so the same calls again complete the keys.
    for part in chunks[:4]:
        t.insert(part, capi.KEEP, 0, u)
        ot.insert(part, capi.KEEP, 0)
    info = t.info()
    assert info["resizes"] >= 2 and info["in_place_doublings"] == 0 and info["away_list_regrows"] >= 1, info
    k, _, _ = t.for_each()
    assert np.array_equal(k, ot.dump()[0])
    assert (t.count(queries) == ot.counts(queries)[0]).all()
    t.close()

    print("a doubling that fails part way: in place", flush=True)
    t = capi.Table(api, 0, expected_keys=8_000)
    t.set_growth(capi.GROW_IN_PLACE, 256)
    u = t.register()
    # The small first store cannot grow in place: the first doubling moves it
    # out of place into a range with room for more.
    i = 0
    while t.info()["resizes"] == 0:
        t.insert(chunks[i], capi.KEEP, 0, u)
        i += 1
    info = t.info()
    grain = 2 << 20
    room = -(-2 * info["buckets"] * info["bucket_bytes"] // grain) * grain
    assert info["store_reserved_bytes"] >= room, info          # so the next doubling runs in place
    resizes, in_place = info["resizes"], info["in_place_doublings"]
    monkeypatch.setenv("KACHE_AMD_TEST_FAIL_AWAY_REGROW", "1")
    with pytest.raises(capi.KacheError) as e:
        for part in chunks[i:]:
            t.insert(part, capi.KEEP, 0, u)
    monkeypatch.delenv("KACHE_AMD_TEST_FAIL_AWAY_REGROW")
    assert e.value.status == capi.KACHE_ENOMEM, e.value
    info = t.info()
    assert info["resizes"] == resizes and info["in_place_doublings"] == in_place, info
    for what, call in (("insert", lambda: t.insert(chunks[0], capi.KEEP, 0, u)), ("count", lambda: t.count(queries)),
                       ("lookup", lambda: t.lookup(queries)), ("for_each", t.for_each), ("size", t.size),
                       ("stats", t.stats)):
        with pytest.raises(capi.KacheError) as e:
            call()
        assert e.value.status == capi.KACHE_EINTERNAL and "in-place doubling failed" in str(e.value), (what, e.value)
    t.close()

    # A new table, in place as well, is unaffected.
    t = capi.Table(api, 0, expected_keys=8_000)
    t.set_growth(capi.GROW_IN_PLACE)
    u = t.register()
    ot = O.OTable(0)
    for part in chunks[:6]:
        assert t.insert(part, capi.KEEP, 0, u) == ot.insert(part, capi.KEEP, 0)
    assert t.info()["in_place_doublings"] >= 1
    k, _, _ = t.for_each()
    assert np.array_equal(k, ot.dump()[0])
    t.close()
