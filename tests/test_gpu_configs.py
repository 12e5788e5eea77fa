"""The benchmark configurations at their full sizes, against the reference's
own C ABI built from its sources (oracle/_ref) and the oracle:

- the CPU set configuration: a 10M-key set, 200k reads of 150 bp (1%
  substitutions, 0.1% N) from a 5 Mbp genome, inserted; 20k fresh reads
  counted; the insert total, the sorted for_each dump, per-read counts and
  the capacity after its one doubling;
- the count map configuration: a u64 map sized for 100M keys, 20M reads of
  150 bp from a 100 Mbp genome upserted +1 in batches of 65,536 reads; at
  this size the tables are compared through an order-free digest of every
  (key, value) pair (a sum of hashes mod 2^64), the key count, the capacity
  and lookups of 2M fresh reads;
- the batched configuration at 1/16 of its size (0.5 Gbp of lognormal-length
  sequences, u32 count map and u8 oriented map, 4 doublings, autotuned) and
  the counter configuration at a reduced size (prime 30 Mbp, count 2M reads,
  histogram): totals, capacities, digests and the histogram.
"""
import ctypes as C
import os
import threading

import numpy as np
import pytest

import kache_hash_amd as KA
import oracle as O
from kache_hash_amd import capi

pytestmark = pytest.mark.gpu
L1 = 151


@pytest.fixture(scope="module")
def api():
    return KA.api()


@pytest.fixture(scope="module")
def ref():
    path = O.ref_library()
    if path is None:
        pytest.skip("the reference library (oracle/_ref) was not built")
    return capi.Api(path)


def device_reads(api, n, genome_seed, genome_len, seed, sub, nrate, lower):
    buf = capi.DeviceBuffer(api, n * L1 + capi.TEXT_PAD)
    api.check(api.lib.kache_amd_synth_reads(buf.ptr, n, 150, genome_seed, genome_len, seed, sub, nrate, lower), "synth_reads")
    api.check(api.lib.kache_amd_device_fill(buf.ptr + n * L1, ord("\n"), capi.TEXT_PAD), "fill")
    return buf


def batch_of(host, first, count):
    base = np.uint64(host.ctypes.data)
    ptrs = (np.arange(first, first + count, dtype=np.uint64) * np.uint64(L1) + base).astype(np.uint64)
    return ptrs, np.full(count, 150, np.uint64)


def threaded_inserts(ref, t, host, n_reads, per_call, threads, op, x):
    users = [t.register() for _ in range(threads)]
    calls = [(a, min(per_call, n_reads - a)) for a in range(0, n_reads, per_call)]
    totals = [0] * threads

    def work(w):
        n = C.c_uint64()
        for j in range(w, len(calls), threads):
            p, l = batch_of(host, *calls[j])
            assert ref.lib.kache_insert(t.h, users[w], p.ctypes.data, l.ctypes.data, len(l), op, x, C.byref(n)) == 0
            totals[w] += n.value

    th = [threading.Thread(target=work, args=(w,)) for w in range(threads)]
    for x_ in th:
        x_.start()
    for x_ in th:
        x_.join()
    return sum(totals)


def mix(z):
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def digest(api, t, width):
    """(keys, sum of mix(key), sum of mix(key + mix(value))) over every pair, mod 2^64."""
    acc = [0, np.uint64(0), np.uint64(0)]

    def cb(ctx, kp, vp, n):
        k = np.ctypeslib.as_array(kp, shape=(n,))
        if width:
            raw = np.ctypeslib.as_array(C.cast(vp, C.POINTER(C.c_uint8)), shape=(n * width,))
            v = raw.view({1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}[width]).astype(np.uint64)
        else:
            v = np.zeros(n, np.uint64)
        with np.errstate(over="ignore"):
            acc[0] += n
            acc[1] = acc[1] + mix(k).sum(dtype=np.uint64)
            acc[2] = acc[2] + mix(k + mix(v + np.uint64(1))).sum(dtype=np.uint64)
        return 0

    fn = capi.KEYS_FN(cb)
    api.check(api.lib.kache_for_each(t.h, 1, 0, fn, None), "kache_for_each")
    return acc[0], int(acc[1]), int(acc[2])


def test_cpu_set_configuration(api, ref):
    """10M-key set, 200k x 150 bp reads of a 5 Mbp genome; count of 20k fresh reads."""
    n, seed = 200_000, 0xC1
    reads = device_reads(api, n, seed, 5_000_000, seed + 1, 0.01, 0.001, 0.0)
    fresh = device_reads(api, 20_000, seed, 5_000_000, seed + 2, 0.01, 0.001, 0.0)
    host, fhost = reads.download(n * L1), fresh.download(20_000 * L1)
    print("set configuration: 200k reads", flush=True)
    g = capi.Table(api, 0, expected_keys=10_000_000)
    u = g.register()
    # Inserts in calls of 64 reads, as the reference's benchmark does.
    tot = 0
    for a in range(0, n, 64):
        p, l = batch_of(host, a, min(64, n - a))
        nn = C.c_uint64()
        api.check(api.lib.kache_insert(g.h, u, p.ctypes.data, l.ctypes.data, len(l), 0, 0, C.byref(nn)), "insert")
        tot += nn.value
    r = capi.Table(ref, 0, 10_000_000, 0.8, False, 16)
    rtot = threaded_inserts(ref, r, host, n, 64, 16, 0, 0)
    o = O.OTable(0)
    seqs = [host[i * L1:i * L1 + 150].tobytes() for i in range(n)]
    otot = o.insert(seqs)
    assert tot == rtot == otot == g.size() == r.size()
    kg, _, _ = g.for_each()
    kr, _, _ = r.for_each()
    ko, _ = o.dump()
    assert np.array_equal(kg, kr) and np.array_equal(kg, np.sort(ko))
    assert g.stats()["capacity"] == r.stats()["capacity"] == 25_000_064          # one doubling of 12,500,032
    fseqs = [fhost[i * L1:i * L1 + 150].tobytes() for i in range(20_000)]
    cg, cr = g.count(fseqs), r.count(fseqs)
    assert np.array_equal(cg, cr) and np.array_equal(cg, o.counts(fseqs)[0])


def test_count_map_configuration(api, ref):
    """u64 map sized 100M, 20M reads upserted +1 in batches of 65,536; the
    reference inserts the same reads from 16 threads."""
    n, seed, batch = 20_000_000, 0xC2, 65_536
    reads = device_reads(api, n, seed, 100_000_000, seed + 1, 0.005, 0.0005, 0.5)
    g = capi.Table(api, 8, expected_keys=100_000_000)
    u = g.register()
    bounds = np.array([min(b, n) * L1 for b in range(0, n + batch, batch)], np.uint64)
    bounds = np.unique(bounds)
    print("count map configuration: 20M reads on the device", flush=True)
    ins = g.insert_batches(reads.ptr, bounds, capi.ADD, 1, u)
    host = reads.download(n * L1)
    del reads
    print("count map configuration: the reference, 16 threads", flush=True)
    r = capi.Table(ref, 8, 100_000_000, 0.8, False, 16)
    rtot = threaded_inserts(ref, r, host, n, batch, 16, capi.ADD, 1)
    assert int(ins.sum()) == rtot == g.size() == r.size()
    assert g.stats()["capacity"] == r.stats()["capacity"]
    print("count map configuration: digests", flush=True)
    dg, dr = digest(api, g, 8), digest(ref, r, 8)
    assert dg == dr
    look = device_reads(api, 2_000_000, seed, 100_000_000, seed + 2, 0.005, 0.0005, 0.5)
    lh = look.download(2_000_000 * L1)
    p, l = batch_of(lh, 0, 2_000_000)
    cg, cr = np.zeros(2_000_000, np.uint64), np.zeros(2_000_000, np.uint64)
    api.check(api.lib.kache_sum(g.h, p.ctypes.data, l.ctypes.data, 2_000_000, cg.ctypes.data), "sum")
    assert ref.lib.kache_sum(r.h, p.ctypes.data, l.ctypes.data, 2_000_000, cr.ctypes.data) == 0
    assert np.array_equal(cg, cr)


def values_of(api, t, width):
    """Every value of the table (in for_each order), as u64."""
    out = []

    def cb(ctx, kp, vp, n):
        raw = np.ctypeslib.as_array(C.cast(vp, C.POINTER(C.c_uint8)), shape=(n * width,))
        out.append(raw.view({1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}[width]).astype(np.uint64))
        return 0

    fn = capi.KEYS_FN(cb)
    api.check(api.lib.kache_for_each(t.h, 1, 0, fn, None), "kache_for_each")
    return np.concatenate(out) if out else np.zeros(0, np.uint64)


def uneven_text(api, gbp, genome_len, seed):
    """Sequences of lognormal lengths (median 2 kbp, sigma 1.3, 50 bp-100 kbp)
    totalling `gbp` Gbp of a seeded genome, 1% substitutions, as a device
    text; returns (buffer, text bytes)."""
    rng = np.random.default_rng(seed)
    lens = np.clip(np.rint(rng.lognormal(np.log(2000), 1.3, int(gbp * 1e9 / 3000) + 1000)), 50, 100_000).astype(np.uint64)
    lens = lens[:int(np.searchsorted(np.cumsum(lens), int(gbp * 1e9))) + 1]
    off = np.zeros(len(lens) + 1, np.uint64)
    np.cumsum(lens + 1, out=off[1:])
    buf = capi.DeviceBuffer(api, int(off[-1]) + capi.TEXT_PAD)
    d_off = capi.DeviceBuffer(api, 8 * len(off))
    d_off.upload(off.view(np.uint8))
    api.check(api.lib.kache_amd_synth_seqs(buf.ptr, d_off.ptr, len(lens), int(lens.max()), seed, genome_len, seed + 1, 0.01, 0.0005, 0.2),
              "synth_seqs")
    api.check(api.lib.kache_amd_device_fill(buf.ptr + int(off[-1]), ord("\n"), capi.TEXT_PAD), "fill")
    d_off.free()
    return buf, int(off[-1])


def ref_calls(host, max_bytes=1 << 20):
    """The text's sequences in calls of about 1 Mbp: (pointers, lengths) each."""
    ends = np.flatnonzero(host == 10)
    starts = np.concatenate([[0], ends[:-1] + 1])
    calls, a = [], 0
    csum = np.cumsum(ends - starts)
    while a < len(ends):
        b = min(max(int(np.searchsorted(csum, (csum[a - 1] if a else 0) + max_bytes)) + 1, a + 1), len(ends))
        calls.append(((starts[a:b] + np.uint64(host.ctypes.data)).astype(np.uint64), (ends[a:b] - starts[a:b]).astype(np.uint64)))
        a = b
    return calls


def ref_parallel(calls, threads, fn):
    tot = [0] * threads

    def work(w):
        n = C.c_uint64()
        for j in range(w, len(calls), threads):
            p, l = calls[j]
            assert fn(w, p, l, n) == 0
            tot[w] += n.value

    th = [threading.Thread(target=work, args=(w,)) for w in range(threads)]
    for x in th:
        x.start()
    for x in th:
        x.join()
    return sum(tot)


@pytest.mark.parametrize("width,oriented", [(4, False), (1, True)])
def test_batched_configuration_midsize(api, ref, width, oriented):
    """The batched configuration at 1/16 of its size: 0.5 Gbp of
    lognormal-length sequences (median 2 kbp, 50 bp-100 kbp) of a 30 Mbp
    genome with 1% substitutions, into a u32 count map (+1) or a u8 oriented
    map (OR 1 forward, 2 reverse) that starts at 1/16 of its final size and
    doubles 4 times, the device inserting the whole text as one call with an
    autotuned launch shape; the reference inserts the same sequences from 16
    threads in calls of about 1 Mbp.  Key counts, capacities and an
    order-free digest of every (key, value) pair are equal."""
    print(f"batched mid-size, width {width}, oriented {oriented}", flush=True)
    d, nbytes = uneven_text(api, 0.5, 30_000_000, 0xC3)
    expected = 8_000_000
    g = capi.Table(api, width, expected_keys=expected)
    u = g.register()
    g.autotune(1)
    if oriented:
        ins = g.insert_oriented_device(d.ptr, nbytes, capi.OR, 1, 2, u)
    else:
        ins = g.insert_device(d.ptr, nbytes, capi.ADD, 1, u)
    assert g.info()["resizes"] >= 4
    host = d.download(nbytes)
    d.free()
    r = capi.Table(ref, width, expected, 0.8, False, 16)
    users = [r.register() for _ in range(16)]
    calls = ref_calls(host)
    if oriented:
        rtot = ref_parallel(calls, 16, lambda w, p, l, n: ref.lib.kache_insert_oriented(r.h, users[w], p.ctypes.data, l.ctypes.data,
                                                                                          len(l), capi.OR, 1, 2, C.byref(n)))
    else:
        rtot = ref_parallel(calls, 16, lambda w, p, l, n: ref.lib.kache_insert(r.h, users[w], p.ctypes.data, l.ctypes.data, len(l),
                                                                                 capi.ADD, 1, C.byref(n)))
    assert ins == rtot == g.size() == r.size()
    assert g.stats()["capacity"] == r.stats()["capacity"]
    assert digest(api, g, width) == digest(ref, r, width)
    r.close()
    g.close()


def test_counter_configuration_midsize(api, ref):
    """The counter configuration at a reduced size: prime a pow2 u32 table at
    load 0.9 with a 30 Mbp genome in segments of 1-50 kbp (KEEP 0), count
    2M reads of 150 bp (0.5% substitutions) from host memory through the
    pipelined host-text path (+1), then histogram(10000); the reference
    primes and counts the same sequences from 16 threads through its C ABI
    (insert KEEP 0, update ADD 1).  Prime and count totals, size, capacity,
    the digest of every pair and the histogram are equal."""
    gl, seed = 30_000_000, 0xC4
    gbuf = capi.DeviceBuffer(api, gl)
    api.check(api.lib.kache_amd_synth_genome(gbuf.ptr, gl, seed), "genome")
    genome = gbuf.download(gl)
    gbuf.free()
    rng = np.random.default_rng(seed)
    cuts = np.concatenate([[0], np.cumsum(rng.integers(1, 50_000, gl // 20_000))])
    cuts = cuts[cuts < gl]
    prime = bytearray()
    for a, b in zip(cuts[:-1], cuts[1:]):
        prime += bytes(genome[a:b]) + b"\n"
    prime = np.frombuffer(bytes(prime), np.uint8)
    n_reads = 2_000_000
    reads = device_reads(api, n_reads, seed, gl, seed + 1, 0.005, 0.0005, 0.3)
    rhost = reads.download(n_reads * L1)
    reads.free()
    print(f"counter mid-size: prime {len(prime)} bytes, count {n_reads} reads", flush=True)
    g = capi.Table(api, 4, expected_keys=gl, load_factor=0.9, pow2=True)
    u = g.register()
    primed = g.insert_text(prime, capi.KEEP, 0, u)
    counted = g.update_text(rhost, capi.ADD_WRAP, 1)
    r = capi.Table(ref, 4, gl, 0.9, True, 16)
    users = [r.register() for _ in range(16)]
    rprimed = ref_parallel(ref_calls(prime), 16, lambda w, p, l, n: ref.lib.kache_insert(r.h, users[w], p.ctypes.data, l.ctypes.data,
                                                                                          len(l), capi.KEEP, 0, C.byref(n)))
    rcounted = ref_parallel([batch_of(rhost, a, min(1024, n_reads - a)) for a in range(0, n_reads, 1024)], 16,
                            lambda w, p, l, n: ref.lib.kache_update(r.h, p.ctypes.data, l.ctypes.data, len(l), capi.ADD, 1, C.byref(n)))
    assert primed == rprimed == g.size() == r.size()
    assert counted == rcounted
    assert g.stats()["capacity"] == r.stats()["capacity"]
    assert digest(api, g, 4) == digest(ref, r, 4)
    rv = values_of(ref, r, 4)
    h = np.bincount(rv[(rv >= 1) & (rv <= 10_000)].astype(np.int64), minlength=10_001).astype(np.uint64)
    h[0] = 0
    assert (rv == 0).any() and np.array_equal(g.histogram(10_000), h)
    r.close()
    g.close()
