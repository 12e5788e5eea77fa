"""Values in their words under heavy concurrency, bit-exact against the oracle.

Maps of widths 1, 2 and 4 keep values in 4-byte words (448-byte buckets),
width 8 in 8-byte words (576 bytes), sets none (320 bytes).  A tandem repeat
makes every one of its few k-mers occur tens of thousands of times in one
launch, so each value takes that many concurrent atomics: sums that cross
the width's maximum (and, in a 4-byte word, the word's own 2^32) must end
saturated (capi/Kmer_Table.hpp:38), wrapping adds must end modulo 2^(8 *
width), and MIN / MAX / OR must compare unsigned at the word's top bit."""
import numpy as np
import pytest

import kache_hash_amd as KA
import oracle as O
import seqgen as G
from kache_hash_amd import capi

pytestmark = pytest.mark.gpu

WIDTHS = (1, 2, 4, 8)


@pytest.fixture(scope="module")
def api():
    return KA.api()


def repeat_text(unit_len, copies, seed):
    """A tandem repeat of a random unit (unit_len distinct k-mers at most,
    each about `copies` times), then a random sequence of new k-mers."""
    unit = G.genome(unit_len, seed)
    return unit * copies + b"\n" + G.genome(5_000, seed + 1) + b"\n"


def check_same(t, o):
    ka, va, _ = t.for_each()
    kb, vb = o.dump()
    order = np.argsort(kb)
    assert np.array_equal(ka, kb[order])
    assert np.array_equal(va.astype(np.uint64), vb[order].astype(np.uint64))
    return va


def test_layout_sizes(api):
    for width, bucket in ((0, 320), (1, 448), (2, 448), (4, 448), (8, 576)):
        with capi.Table(api, width, expected_keys=10_000) as t:
            i = t.info()
            assert i["bucket_bytes"] == bucket
            assert i["device_bytes"] == i["buckets"] * bucket + i["overflow_slots"] * (8 + (0 if width == 0 else 4 if width <= 4 else 8))


@pytest.mark.parametrize("width", WIDTHS)
def test_concurrent_adds_saturate(api, width):
    mx = (1 << (8 * width)) - 1
    text = repeat_text(41, 60_000, 500 + width)
    print(f"values width {width}: saturating adds, {len(text)} bytes", flush=True)
    with capi.Table(api, width, expected_keys=20_000) as t:
        u = t.register()
        o = O.OTable(width)
        # A third of the maximum: the repeated k-mers pass it within their
        # first few of ~60k adds, and a 4-byte word's sums pass 2^32.
        x = mx // 3 + 1
        assert t.insert_text(text, capi.ADD, x, u) == o.insert([text], capi.ADD, x)
        va = check_same(t, o)
        assert int(va.max()) == mx
        assert t.update_text(text, capi.ADD, 1) == o.update([text], capi.ADD, 1)
        check_same(t, o)


@pytest.mark.parametrize("width", WIDTHS)
def test_concurrent_counts_and_wrapping_adds(api, width):
    mx = (1 << (8 * width)) - 1
    text = repeat_text(37, 50_000, 600 + width)
    print(f"values width {width}: counts and wrapping adds", flush=True)
    with capi.Table(api, width, expected_keys=20_000) as t:
        u = t.register()
        o = O.OTable(width)
        # Counts of +1: exact below the maximum (widths 2-8), saturated at 255.
        assert t.insert_text(text, capi.ADD, 1, u) == o.insert([text], capi.ADD, 1)
        check_same(t, o)
        # Wrapping adds of a large odd operand: the sum of ~50k of them modulo
        # 2^(8 * width), carried out of the width many times over.
        x = 0x9E3779B97F4A7C15 & mx
        assert t.update_text(text, capi.ADD_WRAP, x) == o.update([text], capi.ADD_WRAP, x)
        va = check_same(t, o)
        assert int(va.max()) > 0
        # Wrapping upserts (new keys take the operand).
        more = text + G.genome(3_000, 700 + width) + b"\n"
        assert t.insert_text(more, capi.ADD_WRAP, x, u) == o.insert([more], capi.ADD_WRAP, x)
        check_same(t, o)


@pytest.mark.parametrize("width", (4, 8))
def test_unsigned_extremes(api, width):
    top = 1 << (8 * width - 1)
    mx = (1 << (8 * width)) - 1
    text = repeat_text(29, 20_000, 800 + width)
    print(f"values width {width}: min/max/or at the top bit", flush=True)
    for op, first, then in ((capi.MAX, top - 1, top + 1), (capi.MIN, top + 1, top - 1), (capi.MIN, top - 1, mx),
                            (capi.OR, top, 1), (capi.ASSIGN, top + 5, mx)):
        with capi.Table(api, width, expected_keys=20_000) as t:
            u = t.register()
            o = O.OTable(width)
            assert t.insert_text(text, op, first, u) == o.insert([text], op, first)
            assert t.insert_text(text, op, then, u) == o.insert([text], op, then)
            check_same(t, o)
            assert t.update_text(text, capi.MAX, top) == o.update([text], capi.MAX, top)
            check_same(t, o)
