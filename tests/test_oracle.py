"""Pins the oracle (oracle/kache_oracle.c) before anything trusts it:
golden vectors written by the reference's own code, an independent XXH3
implementation (the `xxhash` package), brute force from the definitions, the
C ABI's known answers, and the reference library's recorded table outputs.
CPU only."""
import hashlib
import json
import os

import numpy as np
import pytest
import xxhash

import oracle as O
import scenarios
import seqgen as G

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


def golden_seqs():
    d = load("describer.npz")
    lens, text = d["lens"], d["text"].tobytes()
    out, o = [], 0
    for l in lens:
        out.append(text[o:o + int(l)])
        o += int(l)
    return out, d


def test_scalars_against_reference_vectors():
    s = load("scalars.npz")
    for i in range(len(s["x"])):
        x = int(s["x"][i])
        assert O.wang(x) == int(s["wang"][i])
        assert O.checksum_mul(x) == int(s["c_mul"][i])
        assert O.checksum_wang(x) == int(s["c_wang"][i])
        assert O.minimizer_hash(int(s["canon"][i])) == int(s["minimizer"][i])
        assert O.xxh3_8(x, 0) == int(s["xxh3_s0"][i]) == int(s["xxh3"][i])
        assert O.xxh3_8(x, 1 << 63) == int(s["xxh3_s63"][i])


def test_xxh3_against_independent_implementation():
    rng = np.random.default_rng(5)
    for x in [0, 1, (1 << 64) - 1, 1 << 63] + [int(v) for v in rng.integers(0, 2**63, 2000, dtype=np.uint64)]:
        b = x.to_bytes(8, "little")
        assert O.xxh3_8(x, 0) == xxhash.xxh3_64_intdigest(b)
        for seed in (1 << 63, 12345, 0xDEADBEEFCAFEF00D):
            assert O.xxh3_8(x, seed) == xxhash.xxh3_64_intdigest(b, seed=seed)


def _brute(s):
    """Describer quantities from the definitions, with strings."""
    comp = {ord("A"): "T", ord("C"): "G", ord("G"): "C", ord("T"): "A"}
    enc = {"A": 0, "C": 1, "G": 2, "T": 3}
    out = []
    for i in range(len(s) - 30):
        w = s[i:i + 31].upper()
        if any(ch not in "ACGT" for ch in w):
            out.append(None)
            continue
        rc = "".join(comp[ord(ch)] for ch in reversed(w))
        val = lambda t: sum(enc[ch] << (2 * (len(t) - 1 - j)) for j, ch in enumerate(t))
        f, r = val(w), val(rc)
        hs = []
        for o in range(10):
            l = w[o:o + 22]
            lr = "".join(comp[ord(ch)] for ch in reversed(l))
            hs.append(O.wang(min(val(l), val(lr))))
        out.append((min(f, r), min(hs), f <= r))
    return out


def test_describer_bruteforce():
    rng = np.random.default_rng(3)
    for t in range(20):
        n = int(rng.integers(25, 160))
        s = "".join("ACGTacgtN"[int(v)] if rng.random() < 0.2 else "ACGT"[int(v) % 4] for v in rng.integers(0, 9, n))
        valid, key, h, c, fwd = O.describe(s)
        br = _brute(s)
        assert len(br) == len(valid)
        for i, b in enumerate(br):
            assert bool(valid[i]) == (b is not None)
            if b is not None:
                assert (int(key[i]), int(h[i]), bool(fwd[i])) == b
                assert int(c[i]) == O.checksum_wang(b[0])


def test_describer_against_reference_vectors():
    seqs, d = golden_seqs()
    for q, s in enumerate(seqs):
        sel = d["q"] == q
        idx = d["i"][sel]
        v, key, h, c, fwd = O.describe(s, True)
        _, _, _, cm, _ = O.describe(s, False)
        assert int(v.sum()) == len(idx)
        assert np.nonzero(v)[0].tolist() == idx.tolist()
        assert (key[idx] == d["key"][sel]).all()
        assert (h[idx] == d["h"][sel]).all()
        assert (fwd[idx] == d["fwd"][sel]).all()
        assert (c[idx] == d["c_wang"][sel]).all()
        assert (cm[idx] == d["c_mul"][sel]).all()


def test_minimizer_hash_is_orientation_invariant_and_matches_describer():
    seqs, d = golden_seqs()
    for k, h in zip(d["key"][::37], d["h"][::37]):
        assert O.minimizer_hash(int(k)) == int(h)
        assert O.revcomp31(O.revcomp31(int(k))) == int(k)


def test_sizing_known_answers():
    # test/capi.cpp:429-436: 1000 keys -> 1280 slots exact, 2048 with pow2.
    assert O.cap_for(1000, 0.8, False) * 32 == 1280
    assert O.cap_for(1000, 0.8, True) * 32 == 2048
    # test/capi.cpp:423-428: beyond the most buckets.
    assert not O.fits(1 << 37, 0.8, False)
    assert not O.fits(55 << 30, 0.8, True)
    assert not O.fits((1 << 64) - 1, 0.8, False)
    assert O.fits(1 << 22, 0.8, False)
    # The threshold exceeds the expected keys, and one bucket fewer would not.
    for n in [0, 1, 10, 31, 32, 1000, 12345, 10_000_000, 100_000_000, 2_000_000_000]:
        c = O.cap_for(n, 0.8, False)
        assert int(c * 32 * 0.8) > n
        assert c == 1 or int((c - 1) * 32 * 0.8) <= n


def test_placement_against_reference_vectors():
    """bucket_of, filter_pos and pick_bucket as the reference's own (private)
    helpers compute them, over seeded hashes and capacities from 1 to 2^32 - 1
    (include/Streaming_Kmer_Hash_Table.hpp:1014-1019, 1055-1068; written by
    tests/golden/gen_golden.cpp)."""
    p = load("placement.npz")
    caps = [int(c) for c in p["caps"]]
    for i in range(len(p["h"])):
        h = int(p["h"][i])
        assert O.filter_pos(h) == int(p["filter_pos"][i])
        for j, cap in enumerate(caps):
            assert O.bucket_of(h, cap) == int(p["bucket_of"][i, j])
            assert O.pick_bucket(h, cap) == int(p["pick_bucket"][i, j])


def test_secondaries_and_overflow_sizes_against_reference_tables():
    """Secondaries of live reference tables (:2111-2115), their bucket counts
    (:1003-1011) and the overflow capacities the reference gives an overflow
    table alone (include/Concurrent_Hash_Table.hpp:158-164) and as a table
    sizes it (include/Streaming_Kmer_Hash_Table.hpp:775)."""
    p = load("placement.npz")
    for n, slots in zip(p["ov_n"], p["ov_slots"]):
        assert O.ov_slots_for(int(n)) == int(slots)
    for t in range(len(p["tab_n"])):
        cap = O.cap_for(int(p["tab_n"][t]), 0.8, False)
        assert cap == int(p["tab_buckets"][t])
        assert O.ov_slots_for(max(1024, cap)) == int(p["tab_ov_slots"][t])
        for h, ps, qs in p["sec"][t]:
            assert O.secondaries(int(h), cap) == (int(ps), int(qs))


def test_placement_doubling_and_planted_mistakes():
    """Doubling maps bucket b to 2b or 2b + 1 (the property the in-order
    split relies on); a transposed multiply or a dropped mask fails the
    reference vectors."""
    p = load("placement.npz")
    h = [int(x) for x in p["h"][:200]]
    for x in h:
        for cap in (1, 7, 1000, 3906251, (1 << 31) - 1):
            b = O.bucket_of(x, cap)
            assert b < cap and O.bucket_of(x, 2 * cap) in (2 * b, 2 * b + 1)
    j = [int(c) for c in p["caps"]].index(3906251)
    assert any((((x >> 32) * 3906251) >> 32) != int(p["bucket_of"][i, j]) for i, x in enumerate(h))
    assert any(((x * 3906251) >> 32) & ((1 << 64) - 1) != int(p["bucket_of"][i, j]) for i, x in enumerate(h))
    assert any(((x >> 32) & 0xFF) != int(p["filter_pos"][i]) for i, x in enumerate(h))    # the 0 -> 1 rule


def test_schedule_plain_cases():
    # No growth below the threshold; one doubling once a flush reaches it.
    s = O.Sched(1000, 0.8, False)
    assert s.cap == 40 and s.resize_th == 1024 and s.step == 1
    s.call(0, 1023)
    assert s.cap == 40
    s.call(0, 1)
    assert s.cap == 80 and s.resizes == 1


def digest(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


@pytest.mark.parametrize("name", [s["name"] for s in scenarios.SCENARIOS])
def test_oracle_matches_reference_tables(name):
    """The oracle against the reference library's recorded outputs."""
    gold = {g["name"]: g for g in json.load(open(os.path.join(GOLD, "ref_tables.json")))["scenarios"]}[name]
    s = next(x for x in scenarios.SCENARIOS if x["name"] == name)
    r = scenarios.run(scenarios.OracleTable, None, s)
    assert r["inserted"] == gold["inserted"]
    assert r["size"] == gold["size"]
    assert r["capacity"] == gold["capacity"]
    assert digest(r["keys"], r["values"]) == gold["dump"]
    assert digest(*r["lookup"]) == gold["lookup"]
    assert digest(r["counts"]) == gold["counts"]
    assert r["reduce"] == gold["reduce"]
    assert digest(*r["get"]) == gold["get"]
    if gold["present"] is not None:
        assert r["present"] == gold["present"]
        assert digest(r["keys2"], r["values2"]) == gold["dump_after_update"]


def test_planted_mistakes_fail():
    """Mistakes planted in the oracle's inputs change what the checks see:
    a transposed checksum variant, a dropped seed, a reversed key."""
    s = load("scalars.npz")
    x = [int(v) for v in s["x"][:64]]
    assert any(O.checksum_mul(v) != int(c) for v, c in zip(x, s["c_wang"][:64]))
    assert any(O.xxh3_8(v, 0) != int(c) for v, c in zip(x, s["xxh3_s63"][:64]))
    seqs, d = golden_seqs()
    k = [int(v) for v in d["key"][:64]]
    assert any(O.minimizer_hash(O.revcomp31(v) ^ 1) != int(h) for v, h in zip(k, d["h"][:64]))


def test_oracle_matches_reference_library_live():
    """When the reference library is built here, the oracle against it live,
    on a fresh seed (the recorded digests cover the fixed seeds)."""
    path = O.ref_library()
    if path is None:
        pytest.skip("reference library not built (make -C oracle ref)")
    from kache_hash_amd import capi
    api = capi.Api(path)
    s = dict(scenarios.SCENARIOS[3], seed=77, genome=30_000)
    S, Q = scenarios.inputs(s)
    a = scenarios.run(capi.Table, api, s, S, Q)
    b = scenarios.run(scenarios.OracleTable, None, s, S, Q)
    assert a["inserted"] == b["inserted"] and a["capacity"] == b["capacity"]
    assert (a["keys"] == b["keys"]).all() and (a["values"] == b["values"]).all()
    for x, y in zip(a["lookup"], b["lookup"]):
        assert (x == y).all()
    assert (a["counts"] == b["counts"]).all() and (a["sums"] == b["sums"]).all() and (a["masked"] == b["masked"]).all()
    assert a["present"] == b["present"] and (a["values2"] == b["values2"]).all()
    assert a["reduce"] == b["reduce"]
    assert (a["get"][0] == b["get"][0]).all() and (a["get"][1] == b["get"][1]).all()


def test_random_programs_reference_vs_oracle():
    """The seeded random programs (tests/programs.py) through the reference
    library, against the oracle: every call's result, capacity and query."""
    path = O.ref_library()
    if path is None:
        pytest.skip("reference library not built (make -C oracle ref)")
    import programs
    from kache_hash_amd import capi
    api = capi.Api(path)
    for case in range(programs.CASES):
        programs.run_program(api, case)
