"""The C++ API headers over the device library (include/kache_amd.hpp: the
reference's Streaming_Kmer_Hash_Table template API with a fixed set of value
operations; include/kache_amd_counter.hpp: the reference's Kmer_Set_Counter),
driven by tests/cpp/cpp_api_check on FASTA / FASTQ files, against the oracle."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
import seqgen as G

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CHECK = os.path.join(ROOT, "tests", "cpp", "cpp_api_check")


def u64(path):
    return np.fromfile(path, dtype="<u8")


def test_cpp_api_matches_oracle(tmp_path):
    assert os.path.exists(CHECK), "build it with __graft_entry__.build()"
    g = G.genome(120_000, 61)
    reads = G.reads(g, 3_000, 150, 62, sub=0.01, nrate=0.002, lower=0.3)
    segs = [g[a:a + 5_000] for a in range(0, 100_000, 5_000)]
    # Reads as FASTQ with CRLF line ends; segments as multi-line FASTA.
    (tmp_path / "reads.fq").write_bytes(b"".join(b"@r%d\r\n%s\r\n+\r\n%s\r\n" % (i, r, b"I" * len(r)) for i, r in enumerate(reads)))
    (tmp_path / "segs.fa").write_bytes(b"".join(b">s%d\n%s\n%s\n" % (i, s[:2_000], s[2_000:]) for i, s in enumerate(segs)))
    print("cpp_api_check on 3,000 reads and 20 segments", flush=True)
    subprocess.run([CHECK, str(tmp_path / "reads.fq"), str(tmp_path / "segs.fa"), str(tmp_path)], check=True, timeout=120)
    sc = dict(line.split() for line in (tmp_path / "scalars.txt").read_text().splitlines())
    sc = {k: int(v) for k, v in sc.items()}

    # Set.
    o = O.OTable(0)
    assert sc["set_inserted"] == o.insert(reads) == sc["set_size"]
    assert sc["set_count"] == int(o.counts(segs)[0].sum())
    k, _ = o.dump()
    assert np.array_equal(u64(tmp_path / "set_keys.u64"), np.sort(k))
    assert sc["set_find_first"] == int(o.get(int(O.describe(segs[0][:31])[1][0]))[0])

    # u32 map: v + 1 from 1 per read; max(v, 7) over the segments; lookups.
    o = O.OTable(4)
    assert sc["map_inserted"] == o.insert(reads, 6, 1) == sc["map_size"]
    assert sc["map_present"] == o.update(segs, 5, 7)
    kv = u64(tmp_path / "map_kv.u64").reshape(-1, 2)
    kb, vb = o.dump()
    ob = np.argsort(kb)
    ok = np.argsort(kv[:, 0])
    assert np.array_equal(kv[ok, 0], kb[ob]) and np.array_equal(kv[ok, 1], vb[ob])
    # find over all segments (one batch): g(q, i, r, fwd) per k-mer position, in order.
    found, vals, fwd = o.lookup(segs)
    q_of = np.concatenate([np.full(max(len(x) - 30, 0), q, np.uint64) for q, x in enumerate(segs)])
    i_of = np.concatenate([np.arange(max(len(x) - 30, 0), dtype=np.uint64) for x in segs])
    f = u64(tmp_path / "map_find.u64").reshape(-1, 4)
    pos = np.nonzero(found != 2)[0]
    assert np.array_equal(f[:, 0], q_of[pos]) and np.array_equal(f[:, 1], i_of[pos])
    assert np.array_equal(f[:, 2], np.where(found[pos] == 1, vals[pos] + 1, 0))
    assert np.array_equal(f[:, 3], fwd[pos])
    assert sc["map_val_mismatch_refused"] == 1

    # u8 oriented OR 1 / 2.
    o = O.OTable(1)
    assert sc["oriented_inserted"] == o.insert_oriented(reads, 3, 1, 2)
    kv = u64(tmp_path / "oriented_kv.u64").reshape(-1, 2)
    kb, vb = o.dump()
    ob, ok = np.argsort(kb), np.argsort(kv[:, 0])
    assert np.array_equal(kv[ok, 0], kb[ob]) and np.array_equal(kv[ok, 1], vb[ob])

    # Counter: prime with the segments (KEEP 0), count the reads (+1 wrapping).
    o = O.OTable(4)
    o.insert(segs, 0, 0)
    assert sc["counter_found"] == o.update(reads, 6, 1)
    assert sc["counter_size"] == o.size()
    _, v = o.dump()
    h = np.bincount(v[(v >= 1) & (v <= 50)].astype(np.int64), minlength=51).astype(np.uint64)
    h[0] = 0
    assert np.array_equal(u64(tmp_path / "counter_hist.u64"), h)
    fo, vo = o.get(int(O.describe(segs[0][:31])[1][0]))
    assert sc["counter_count_of_first"] == (vo if fo else 0)
    assert sc["counter_absent_throws"] == 1
