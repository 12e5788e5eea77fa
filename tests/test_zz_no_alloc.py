"""The allocation-free per-sequence count (kache_amd_count_seqs_device_static)
through the C ABI: equal to kache_amd_count_seqs_device and to the oracle on
two calls in a row, the second with fewer sequences over less text (the
scratch keeps the first call's contents), and its limits refused on the host
before any device work."""
import numpy as np
import pytest

import kache_hash_amd as KA
import oracle as O
import seqgen as G
from kache_hash_amd import capi

pytestmark = pytest.mark.gpu


def device_text(api, seqs):
    text = b"".join(bytes(x) + b"\n" for x in seqs)
    off = np.concatenate([[0], np.cumsum([len(x) + 1 for x in seqs])]).astype(np.uint64)
    d = capi.DeviceBuffer(api, len(text) + capi.TEXT_PAD)
    d.upload(text + b"\n" * capi.TEXT_PAD)
    return d, len(text), off


def counts(api, fn, t, d, nbytes, off, n):
    out = np.full(max(n, 1), 7, np.uint64)
    api.check(fn(t.h, d.ptr, nbytes, off.ctypes.data, n, out.ctypes.data), "count")
    return out[:n]


@pytest.mark.parametrize("width", [0, 4])
def test_static_count_equals_count_seqs_device(width):
    api = KA.api()
    g = G.genome(400_000, 301 + width)
    reads = G.reads(g, 30_000, 150, 302, sub=0.01)
    t = capi.Table(api, width, expected_keys=200_000)
    ot = O.OTable(width)
    u = t.register()
    assert t.insert(reads, capi.ADD, 1, u) == ot.insert(reads, capi.ADD, 1)
    rng = np.random.default_rng(303)
    big = G.reads(g, 20_000, 150, 304, sub=0.02)
    big += [g[int(i):int(i) + int(rng.integers(1, 3000))] for i in rng.integers(0, 390_000, 500)]
    big += [b"", b"ACG", b"N" * 40, g[:31], g[1000:1031].lower()]
    small = big[:777:3]
    print(f"no-alloc count, width {width}: {len(big)} then {len(small)} sequences", flush=True)
    for seqs in (big, small):
        d, nb, off = device_text(api, seqs)
        a = counts(api, api.lib.kache_amd_count_seqs_device, t, d, nb, off, len(seqs))
        b = counts(api, api.lib.kache_amd_count_seqs_device_static, t, d, nb, off, len(seqs))
        assert np.array_equal(a, b)
        assert np.array_equal(b, ot.counts(seqs)[0])
        d.free()
    t.close()


def test_static_count_limits():
    api = KA.api()
    t = capi.Table(api, 4, expected_keys=1000)
    n = (1 << 20) + 1                   # one sequence over the most the scratch holds
    d = capi.DeviceBuffer(api, n + capi.TEXT_PAD)
    api.check(api.lib.kache_amd_device_fill(d.ptr, ord("\n"), n + capi.TEXT_PAD), "fill")
    off = np.arange(n + 1, dtype=np.uint64)
    out = np.zeros(n, np.uint64)
    assert api.lib.kache_amd_count_seqs_device_static(t.h, d.ptr, n, off.ctypes.data, n, out.ctypes.data) == capi.KACHE_EINVAL
    assert "scratch" in api.last_error()
    # One sequence over more text than the scratch covers: refused on the host.
    big = np.array([0, (1 << 30) + 1], np.uint64)
    assert api.lib.kache_amd_count_seqs_device_static(t.h, d.ptr, (1 << 30) + 1, big.ctypes.data, 1, out.ctypes.data) == capi.KACHE_EINVAL
    # At the limit of sequences it runs, and agrees with the allocating call.
    m = 1 << 20
    a = counts(api, api.lib.kache_amd_count_seqs_device, t, d, m, off, m)
    b = counts(api, api.lib.kache_amd_count_seqs_device_static, t, d, m, off, m)
    assert np.array_equal(a, b) and not b.any()
    d.free()
    t.close()
