"""Host-text streaming (kache_amd_insert_text / kache_amd_update_text) and the
value histogram (kache_amd_histogram), the pieces of the reference's
Kmer_Set_Counter (counter/Kmer_Set_Counter.cpp:64-127): prime = upsert with
KEEP 0, count = update with wrapping +1, histogram over [1, max].  Bit-exact
against the oracle on seeded inputs, across the 16 MiB chunk boundaries of
the double-buffered pipeline."""
import numpy as np
import pytest

import kache_hash_amd as KA
import oracle as O
import seqgen as G
from kache_hash_amd import capi

pytestmark = pytest.mark.gpu
MiB = 1 << 20


@pytest.fixture(scope="module")
def api():
    return KA.api()


def reads_text(g, n, length, seed, sub=0.005, nrate=0.0005):
    return b"\n".join(G.reads(g, n, length, seed, sub=sub, nrate=nrate, lower=0.3)) + b"\n"


def oracle_hist(o, max_count):
    _, v = o.dump()
    v = v[(v >= 1) & (v <= max_count)].astype(np.int64)
    h = np.bincount(v, minlength=max_count + 1).astype(np.uint64)
    h[0] = 0
    return h


def test_counter_prime_count_histogram_small(api):
    """Prime with segments (KEEP 0), count reads (wrapping +1), histogram and
    count_of, as the counter does."""
    g = G.genome(300_000, 41)
    segs = [g[a:a + int(l)] for a, l in zip(range(0, 280_000, 7_000), np.random.default_rng(1).integers(1, 7_000, 40))]
    prime = b"\n".join(segs) + b"\n"
    text = reads_text(g[:180_000], 20_000, 150, 42)
    t = capi.Table(api, 4, expected_keys=200_000, load_factor=0.9, pow2=True)
    u = t.register()
    o = O.OTable(4)
    assert t.insert_text(prime, capi.KEEP, 0, u) == o.insert([prime], capi.KEEP, 0)
    assert t.update_text(text, capi.ADD_WRAP, 1) == o.update([text], capi.ADD_WRAP, 1)
    ka, va, _ = t.for_each()
    kb, vb = o.dump()
    ob = np.argsort(kb)
    assert np.array_equal(ka, kb[ob]) and np.array_equal(va, vb[ob])
    assert (va == 0).any() and (va > 0).any()
    for m in (1, 10, 10_000, 20_000):
        assert np.array_equal(t.histogram(m), oracle_hist(o, m)), m
    # count_of: the value of a k-mer, 0 when absent or not a k-mer.
    for a in (0, 1234, 150_000, 250_000):
        kmer = g[a:a + 31]
        f, v = t.get(kmer)
        fo, vo = o.get(int(O.describe(kmer)[1][0]))
        assert f == fo and v == (vo if fo else 0)
    with pytest.raises(capi.KacheError):
        capi.Table(api, 0, expected_keys=1000).histogram(10)


def test_update_text_wraps_at_the_width(api):
    g = G.genome(50_000, 43)
    text = reads_text(g, 2_000, 150, 44)
    t = capi.Table(api, 2, expected_keys=100_000)
    u = t.register()
    o = O.OTable(2)
    assert t.insert_text(text, capi.ASSIGN, 0xFFFF, u) == o.insert([text], capi.ASSIGN, 0xFFFF)
    assert t.update_text(text, capi.ADD_WRAP, 3) == o.update([text], capi.ADD_WRAP, 3)
    ka, va, _ = t.for_each()
    kb, vb = o.dump()
    ob = np.argsort(kb)
    assert np.array_equal(ka, kb[ob]) and np.array_equal(va, vb[ob])
    assert int(va.max()) < 0xFFFF


def test_text_pipeline_chunks_and_long_runs(api):
    """A text over several 16 MiB chunks, ending without a separator, with a
    run of bases longer than a chunk (cut with k - 1 bases of overlap)."""
    g = G.genome(40 * MiB, 45)
    text = reads_text(g[:2_000_000], 150_000, 150, 46)            # ~22.6 MiB of reads
    long_run = g[1_000_000:1_000_000 + 17 * MiB + 77]               # one sequence longer than a chunk
    tail = g[5_000:5_500]                                           # no trailing separator
    full = text + long_run + b"N" + tail
    print(f"text pipeline: {len(full)} bytes", flush=True)
    t = capi.Table(api, 8, expected_keys=1 << 20)
    u = t.register()
    o = O.OTable(8)
    assert t.insert_text(full, capi.ADD, 1, u) == o.insert([full], capi.ADD, 1)
    assert t.update_text(full, capi.ADD, 2) == o.update([full], capi.ADD, 2)
    ka, va, _ = t.for_each()
    kb, vb = o.dump()
    ob = np.argsort(kb)
    assert np.array_equal(ka, kb[ob]) and np.array_equal(va, vb[ob])
    # The same through the C ABI's host-sequence path (also double-buffered).
    t2 = capi.Table(api, 8, expected_keys=1 << 20)
    u2 = t2.register()
    seqs = [text, long_run, tail]
    assert t2.insert(seqs, capi.ADD, 1, u2) == len(ka)
    assert t2.update(seqs, capi.ADD, 2) == o.counts([full])[0][0]
    kc, vc, _ = t2.for_each()
    assert np.array_equal(kc, ka) and np.array_equal(vc, va)
    assert t.stats()["capacity"] == t2.stats()["capacity"]


def test_text_entry_points_validate(api):
    t = capi.Table(api, 4, expected_keys=1000)
    u = t.register()
    L = api.lib
    E = capi.KACHE_EINVAL
    assert L.kache_amd_insert_text(t.h, 99, b"ACGT", 4, 0, 0, None) == E          # unregistered user
    assert L.kache_amd_insert_text(t.h, u, None, 4, 0, 0, None) == E
    assert L.kache_amd_insert_text(t.h, u, b"ACGT", 4, 7, 0, None) == E           # unknown op
    assert L.kache_amd_update_text(t.h, b"ACGT", 4, capi.ADD, 1 << 32, None) == E # operand too wide
    assert L.kache_amd_insert_text(t.h, u, None, 0, 0, 0, None) == 0              # empty text
    s = capi.Table(api, 0, expected_keys=1000)
    assert L.kache_amd_update_text(s.h, b"ACGT", 4, capi.ADD, 1, None) == E       # update of a set
    assert L.kache_insert(t.h, u, None, None, 0, capi.ADD_WRAP, 1, None) == E      # the C ABI refuses op 6
