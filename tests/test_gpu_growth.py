"""Growth under an HBM budget (kache_amd_set_hbm_budget): a doubling that
would pass the budget fails before it allocates, with KACHE_ENOMEM; the
call's keys stay applied, the table keeps its capacity and every result stays
equal to the oracle's; once the budget allows it, the next call doubles."""
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


def same_as_oracle(t, ot, queries):
    k, v, _ = t.for_each()
    ko, vo = ot.dump()
    assert np.array_equal(k, ko) and np.array_equal(v, vo)
    f, val, fw = t.lookup(queries)
    f2, val2, fw2 = ot.lookup(queries)
    assert (f == f2).all() and (fw == fw2).all()
    if t.width:
        assert (val == val2).all()


@pytest.mark.parametrize("width", [0, 4])
def test_budget_refuses_a_doubling_and_the_table_stays_exact(api, width):
    # A table of 2M keys (25-45 MB) whose overflow table may grow for a launch
    # of 4,000 reads but which may not double.
    g = G.genome(1_500_000, 91 + width)
    reads = G.reads(g, 48_000, 150, 92, sub=0.01, nrate=0.001, lower=0.1)
    queries = G.reads(g, 500, 150, 93, sub=0.01) + G.pieces(G.genome(5_000, 94), 200)
    print(f"growth budget, width {width}: {len(reads)} reads in calls of 4000", flush=True)
    t = capi.Table(api, width, expected_keys=2_000_000)
    ot = O.OTable(width)
    u = t.register()
    cap0 = t.stats()["capacity"]
    # Room for overflow growth (<= 16 MB), not for a doubling: half what an
    # in-place doubling needs beside the table (the store's second half,
    # >= 25 MB, and the new overflow table; out of place needs twice that,
    # so under this budget the table would double in place).
    t.set_growth(capi.GROW_IN_PLACE)
    info = t.info()
    t.set_growth(capi.GROW_AUTO)
    t.set_hbm_budget(info["device_bytes"] + info["next_doubling_bytes"] // 2)
    refused = None
    for i in range(0, len(reads), 4000):
        part = reads[i:i + 4000]
        expect = ot.insert(part, capi.ADD, 1)
        try:
            assert t.insert(part, capi.ADD, 1, u) == expect
        except capi.KacheError as e:
            assert e.status == capi.KACHE_ENOMEM
            assert "doubling" in api.last_error() and "budget" in api.last_error()
            refused = i
            break
    assert refused is not None, "the table never needed to double"
    # The refused call's keys are in; capacity and results are exact.
    assert t.stats()["capacity"] == cap0 and t.info()["resizes"] == 0
    assert t.size() == ot.size()
    same_as_oracle(t, ot, queries)
    # Still refused while the budget holds; the keys of that call are applied too.
    part = reads[refused + 4000:refused + 8000]
    ot.insert(part, capi.ADD, 1)
    with pytest.raises(capi.KacheError) as e:
        t.insert(part, capi.ADD, 1, u)
    assert e.value.status == capi.KACHE_ENOMEM
    same_as_oracle(t, ot, queries)
    # Without the budget the next call doubles, and the table stays exact.
    t.set_hbm_budget(0)
    part = reads[refused + 8000:refused + 12000]
    assert t.insert(part, capi.ADD, 1, u) == ot.insert(part, capi.ADD, 1)
    assert t.stats()["capacity"] >= 2 * cap0 and t.info()["resizes"] >= 1
    same_as_oracle(t, ot, queries)
    t.close()


@pytest.mark.parametrize("width", [0, 4, 8])
def test_table_larger_than_the_device_is_enomem(api, width):
    """kache_new of a table within the reference's limit of 2^32 - 1 buckets
    (include/Streaming_Kmer_Hash_Table.hpp:84) but beyond the device's memory
    (100G keys: 3.9G buckets, 1.3-2.3 TB) returns KACHE_ENOMEM and leaves the
    handle NULL; the next table works."""
    import ctypes as C
    L = api.lib
    o = L.kache_default_options(31, width)
    o.expected_keys = 100_000_000_000
    h = C.c_void_p(1234)
    assert L.kache_new(C.byref(o), C.byref(h)) == capi.KACHE_ENOMEM and not h.value
    g = G.genome(20_000, 91 + width)
    with capi.Table(api, width, expected_keys=30_000) as t:
        u = t.register()
        ot = O.OTable(width)
        reads = G.reads(g, 500, 150, 92)
        assert t.insert(reads, capi.ADD if width else capi.KEEP, 1, u) == ot.insert(reads, capi.ADD if width else capi.KEEP, 1)
        same_as_oracle(t, ot, reads[:50])
