"""Table persistence (kache_amd_save / kache_amd_load): a loaded table equals
the saved one in its dump, size, statistics and capacity, and both evolve
identically under the same later calls (the growth schedule and registered
users are part of the file); a damaged file is refused."""
import numpy as np
import pytest

import kache_hash_amd as KA
import seqgen as G
from kache_hash_amd import capi

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def api():
    return KA.api()


@pytest.mark.parametrize("width", [0, 2, 8])
def test_save_load_round_trip(api, tmp_path, width):
    g = G.genome(400_000, 81 + width)
    reads = G.reads(g, 20_000, 150, 82, sub=0.01, nrate=0.001, lower=0.2)
    more = G.reads(g, 20_000, 150, 83, sub=0.01, nrate=0.001, lower=0.2)
    t = capi.Table(api, width, expected_keys=30_000)
    u1, u2 = t.register(), t.register()
    t.insert(reads[:10_000], capi.ADD, 1, u1)
    t.insert(reads[10_000:], capi.ADD, 1, u2)
    assert t.info()["resizes"] >= 2
    path = tmp_path / "table.kamd"
    t.save(path)
    s = capi.Table.load(api, path)
    ka, va, _ = t.for_each()
    kb, vb, _ = s.for_each()
    assert np.array_equal(ka, kb) and np.array_equal(va, vb)
    assert t.stats() == s.stats() and t.size() == s.size() and t.info() == s.info()
    # Same later calls, same results (users u1, u2 keep their pending counts).
    for tab in (t, s):
        tab.r1 = tab.insert(more[:10_000], capi.ADD, 2, u2)
        tab.r2 = tab.insert(more[10_000:], capi.ADD, 2, u1)
    assert (t.r1, t.r2) == (s.r1, s.r2)
    ka, va, _ = t.for_each()
    kb, vb, _ = s.for_each()
    assert np.array_equal(ka, kb) and np.array_equal(va, vb)
    assert t.stats()["capacity"] == s.stats()["capacity"]
    assert np.array_equal(t.count(more[:500]), s.count(more[:500]))
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
