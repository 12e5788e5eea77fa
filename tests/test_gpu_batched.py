"""Batched inserts of uneven sequences in persistent launches: the launch
shape (blocks per CU, positions per launch, hipGraph replay) and the online
autotuner change speed only.  Every shape must give the table, per-call
counts and capacity of the plain launch path and of the oracle, across
doublings; the graph is replayed many times with resizes (eager work) in
between, and each replay's counts are checked through the call totals."""
import numpy as np
import pytest

import kache_hash_amd as KA
import oracle as O
from kache_hash_amd import capi

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def api():
    return KA.api()


def uneven_text(api, n, seed, genome_len=3_000_000, median=2000, sigma=1.3, lo=50, hi=100_000):
    """n sequences of lognormal lengths as a device text (kache_amd_synth_seqs)."""
    rng = np.random.default_rng(seed)
    lens = np.clip(np.rint(rng.lognormal(np.log(median), sigma, n)), lo, hi).astype(np.uint64)
    off = np.zeros(n + 1, np.uint64)
    np.cumsum(lens + 1, out=off[1:])
    total = int(off[-1])
    buf = capi.DeviceBuffer(api, total + capi.TEXT_PAD)
    d_off = capi.DeviceBuffer(api, 8 * (n + 1))
    d_off.upload(off.view(np.uint8))
    api.check(api.lib.kache_amd_synth_seqs(buf.ptr, d_off.ptr, n, int(lens.max()), seed, genome_len, seed + 1, 0.01, 0.0005, 0.2), "synth_seqs")
    api.check(api.lib.kache_amd_device_fill(buf.ptr + total, ord("\n"), capi.TEXT_PAD), "fill")
    return buf, total, off, lens


def test_synth_seqs_layout(api):
    buf, total, off, lens = uneven_text(api, 500, 71)
    host = buf.download(total)
    assert (host[off[1:].astype(np.int64) - 1] == ord("\n")).all()
    assert int((host == ord("\n")).sum()) == 500
    ok = np.isin(host, np.frombuffer(b"ACGTNacgtn\n", np.uint8))
    assert ok.all()
    again, _, _, _ = uneven_text(api, 500, 71)
    assert np.array_equal(again.download(total), host)


def aligned_text(api, n, seed, calls=4):
    """An uneven text cut into `calls` pieces of whole sequences, each padded
    with separators to a multiple of 16 bytes (device texts start aligned).
    Returns (device buffer, host text, cut offsets)."""
    buf, total, off, lens = uneven_text(api, n, seed)
    host = buf.download(total)
    ends = off[1:].astype(np.int64)
    pieces, a = [], 0
    for g in np.array_split(np.arange(n), calls):
        e = int(ends[g[-1]])
        piece = host[a:e]
        pieces.append(np.concatenate([piece, np.full((-len(piece)) % 16, ord("\n"), np.uint8)]))
        a = e
    text = np.concatenate(pieces)
    cuts = np.concatenate([[0], np.cumsum([len(p) for p in pieces])]).astype(np.int64)
    d = capi.DeviceBuffer(api, len(text) + capi.TEXT_PAD)
    d.upload(np.concatenate([text, np.full(capi.TEXT_PAD, ord("\n"), np.uint8)]))
    return d, text, cuts


def run_shape(api, d, cuts, width, shape, expected, op, oriented):
    t = capi.Table(api, width, expected_keys=expected)
    u = t.register()
    if shape == "autotune":
        t.autotune(1)
    elif shape is not None:
        t.set_tuning(**shape)
    counts = []
    for a, b in zip(cuts[:-1], cuts[1:]):
        if oriented:
            counts.append(t.insert_oriented_device(d.ptr + int(a), int(b - a), op, 1, 2, u))
        else:
            counts.append(t.insert_device(d.ptr + int(a), int(b - a), op, 1, u))
    return t, counts


SHAPES = [
    None,
    {"blocks_per_cu": 0, "launch_log2": 16, "graph": False},
    {"blocks_per_cu": 3, "launch_log2": 17, "graph": False, "waves_per_simd": 6},
    {"blocks_per_cu": 0, "launch_log2": 16, "graph": True, "waves_per_simd": 8},
    {"blocks_per_cu": 2, "launch_log2": 18, "graph": True},
    {"blocks_per_cu": 0, "launch_log2": 16, "graph": False, "threads_per_block": 64},
    {"blocks_per_cu": 0, "launch_log2": 17, "graph": True, "waves_per_simd": 6, "threads_per_block": 128},
    {"blocks_per_cu": 5, "launch_log2": 16, "graph": False, "waves_per_simd": 8, "threads_per_block": 128},
    "autotune",
]


@pytest.mark.parametrize("width,oriented", [(4, False), (1, True), (0, False)])
def test_launch_shapes_give_the_same_table(api, width, oriented):
    d, text, cuts = aligned_text(api, 1500, 72 + width)
    print(f"uneven text: {len(text)} bytes, width {width}, oriented {oriented}", flush=True)
    op = capi.OR if oriented else capi.ADD
    o = O.OTable(width)
    exp = []
    for a, b in zip(cuts[:-1], cuts[1:]):
        piece = text[a:b].tobytes()
        exp.append(o.insert_oriented([piece], op, 1, 2) if oriented else o.insert([piece], op if width else 0, 1 if width else 0))
    kb, vb = o.dump()
    ob = np.argsort(kb)
    ref = None
    for shape in SHAPES:
        t, counts = run_shape(api, d, cuts, width, shape, 50_000, op, oriented)
        k, v, _ = t.for_each()
        assert counts == exp, shape
        assert np.array_equal(k, kb[ob]) and (not width or np.array_equal(v, vb[ob])), shape
        cap = t.stats()["capacity"]
        if ref is None:
            ref = cap
            assert t.info()["resizes"] >= 2, "the text must make the table double"
        assert cap == ref, shape
        if shape == "autotune":
            trials = t.tuning_trials()
            assert len(trials) >= 3 and {tr["waves_per_simd"] for tr in trials[:3]} == {6, 7, 8}
            if len(trials) >= 6:
                assert {tr["threads_per_block"] for tr in trials[3:6]} == {64, 128, 256}
        t.close()


def test_tuning_validation(api):
    t = capi.Table(api, 4, expected_keys=1000)
    with pytest.raises(capi.KacheError):
        t.set_tuning(0, 15, False)
    with pytest.raises(capi.KacheError):
        t.set_tuning(0, 31, False)
    with pytest.raises(capi.KacheError):
        t.autotune(17)
    with pytest.raises(capi.KacheError):
        t.set_tuning(0, 20, False, 5)
    with pytest.raises(capi.KacheError):
        t.set_tuning(0, 20, False, 9)
    for bad_threads in (32, 96, 512):
        with pytest.raises(capi.KacheError):
            t.set_tuning(0, 20, False, 7, bad_threads)
    t.set_tuning(4, 20, True)
    assert t.tuning() == {"blocks_per_cu": 4, "launch_log2": 20, "graph": 1, "tuning": 0, "waves_per_simd": 7, "threads_per_block": 256}
    t.set_tuning(0, 22, False, 8, 128)
    assert t.tuning() == {"blocks_per_cu": 0, "launch_log2": 22, "graph": 0, "tuning": 0, "waves_per_simd": 8, "threads_per_block": 128}
    # Batch calls check their bounds before reading them.
    u = t.register()
    d = capi.DeviceBuffer(api, 64 + capi.TEXT_PAD)
    L, E = api.lib, capi.KACHE_EINVAL
    assert L.kache_amd_insert_batches(t.h, u, d.ptr, None, 3, capi.ADD, 1, None) == E
    h = np.zeros(3, np.uint64)
    assert L.kache_amd_count_batches(t.h, d.ptr, None, 3, h.ctypes.data, None) == E
    bad = np.array([0, 32, 16], np.uint64)
    assert L.kache_amd_insert_batches(t.h, u, d.ptr, bad.ctypes.data, 2, capi.ADD, 1, None) == E
