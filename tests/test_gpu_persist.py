"""Table persistence (kache_amd_save / kache_amd_load): a loaded table equals
the saved one in its dump, size, statistics and capacity, and both evolve
identically under the same later calls (the growth schedule and registered
users are part of the file); every dump, return value and capacity also
equals the oracle's table and growth schedule fed the same calls; a damaged
file is refused."""
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


def geometry(t):
    """The table's geometry and growth history from kache_amd_get_info, less
    how its store's memory was obtained (mapped and reserved bytes, hence
    whether and with how much HBM the next doubling runs in place; in-place
    doublings and away list regrowths belong to the process that grew it)."""
    skip = ("store_mapped_bytes", "store_reserved_bytes", "in_place_doublings", "away_list_regrows", "last_away_keys", "device_free_bytes",
            "next_doubling_bytes")
    return {k: v for k, v in t.info().items() if k not in skip}


@pytest.mark.parametrize("width", [0, 2, 8])
def test_save_load_round_trip(api, tmp_path, width):
    g = G.genome(400_000, 81 + width)
    reads = G.reads(g, 20_000, 150, 82, sub=0.01, nrate=0.001, lower=0.2)
    more = G.reads(g, 20_000, 150, 83, sub=0.01, nrate=0.001, lower=0.2)
    print(f"persist width {width}: 20000 + 20000 reads", flush=True)
    t = capi.Table(api, width, expected_keys=30_000)
    ot, sched = O.OTable(width), O.Sched(30_000)
    u1, u2 = t.register(), t.register()
    for part, u in ((reads[:10_000], u1), (reads[10_000:], u2)):
        r = t.insert(part, capi.ADD, 1, u)
        assert r == ot.insert(part, capi.ADD, 1)
        sched.call(u, r)
    assert t.info()["resizes"] >= 2 and t.stats()["capacity"] == sched.cap * 32
    path = tmp_path / "table.kamd"
    t.save(path)
    s = capi.Table.load(api, path)
    ka, va, _ = t.for_each()
    kb, vb, _ = s.for_each()
    ko, vo = ot.dump()
    assert np.array_equal(ka, kb) and np.array_equal(va, vb)
    assert np.array_equal(kb, ko) and np.array_equal(vb, vo)
    assert t.stats() == s.stats() and t.size() == s.size() == ot.size() and geometry(t) == geometry(s)
    assert s.stats()["capacity"] == sched.cap * 32
    # Same later calls, same results (users u1, u2 keep their pending counts),
    # on both tables and the oracle.
    for tab in (t, s):
        tab.r1 = tab.insert(more[:10_000], capi.ADD, 2, u2)
        tab.r2 = tab.insert(more[10_000:], capi.ADD, 2, u1)
    o1 = ot.insert(more[:10_000], capi.ADD, 2)
    sched.call(u2, o1)
    o2 = ot.insert(more[10_000:], capi.ADD, 2)
    sched.call(u1, o2)
    assert (t.r1, t.r2) == (s.r1, s.r2) == (o1, o2)
    ka, va, _ = t.for_each()
    kb, vb, _ = s.for_each()
    ko, vo = ot.dump()
    assert np.array_equal(ka, kb) and np.array_equal(va, vb)
    assert np.array_equal(kb, ko) and np.array_equal(vb, vo)
    assert t.stats()["capacity"] == s.stats()["capacity"] == sched.cap * 32
    assert np.array_equal(s.count(more[:500]), ot.counts(more[:500])[0])
    # A damaged image is refused.
    raw = bytearray(path.read_bytes())
    raw[len(raw) // 2] ^= 0xFF
    bad = tmp_path / "bad.kamd"
    bad.write_bytes(bytes(raw))
    with pytest.raises(capi.KacheError) as e:
        capi.Table.load(api, bad)
    assert e.value.status == capi.KACHE_EINVAL
    trunc = tmp_path / "trunc.kamd"
    trunc.write_bytes(bytes(raw[:len(raw) // 3]))
    with pytest.raises(capi.KacheError):
        capi.Table.load(api, trunc)
