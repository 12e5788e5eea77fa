"""A table that can only double in place: a u64 map sized for 4.6G keys (a
104-GB bucket store) fed a 4.8-Gbp uniform genome in one device-text call, so
it doubles once.  Out of place the doubling would need a new store twice the
old one's size and a new overflow table beside the table (~216 GB, more than
the ~199 GB a 288-GiB device has left beside the store and the text); in
place it needs the second half, the new overflow table, 1/64 of the store
and the away list (~142 GB).  Checked through properties that do not need
the oracle at this size: the call's count of new
keys equals the table's size, the capacity follows the reference's growth
schedule, every k-mer of the genome is found afterwards, and the k-mers of
an unrelated sequence are not."""
import pytest

import kache_hash_amd as KA
import oracle as O
from kache_hash_amd import capi

pytestmark = pytest.mark.gpu

EXPECTED = 4_600_000_000
GENOME = 4_800_000_000


@pytest.fixture(scope="module")
def api():
    return KA.api()


def test_104_gb_table_doubles_in_place(api):
    L = api.lib
    assert L.kache_amd_trim_cache(None) == 0
    with capi.Table(api, 8, expected_keys=1000) as probe:
        free = probe.info()["device_free_bytes"]
    if free < (270 << 30):
        pytest.skip(f"needs ~270 GiB of free device memory, {free >> 30} GiB free")
    print(f"large table: {EXPECTED} expected keys, {GENOME}-bp genome, {free >> 30} GB free", flush=True)
    text = capi.DeviceBuffer(api, GENOME + 1 + capi.TEXT_PAD)
    api.check(L.kache_amd_synth_genome(text.ptr, GENOME, 0x1A7E), "synth_genome")
    api.check(L.kache_amd_device_fill(text.ptr + GENOME, ord("\n"), 1 + capi.TEXT_PAD), "fill")
    with capi.Table(api, 8, expected_keys=EXPECTED) as t:
        info = t.info()
        store = info["buckets"] * info["bucket_bytes"]
        assert store > 100_000_000_000
        t.set_growth(capi.GROW_OUT_OF_PLACE)
        out_need = t.info()["next_doubling_bytes"]
        t.set_growth(capi.GROW_AUTO)
        in_need = t.info()["next_doubling_bytes"]
        print(f"store {store >> 30} GB; next doubling needs {in_need >> 30} GB in place, {out_need >> 30} GB out of place", flush=True)
        assert out_need > free - store - GENOME        # out of place it could not double
        u = t.register()
        ins = t.insert_device(text.ptr, GENOME + 1, capi.ADD, 1, u)
        inf = t.info()
        print(f"inserted {ins} keys; {inf['resizes']} doubling(s), {inf['in_place_doublings']} in place, "
              f"{inf['last_away_keys']} keys listed away", flush=True)
        assert inf["resizes"] == 1 and inf["in_place_doublings"] == 1
        assert ins == t.size() and ins > 0.99 * (GENOME - 30)
        sched = O.Sched(EXPECTED)
        sched.call(0, ins)
        assert t.stats()["capacity"] == sched.cap * 32
        hits, positions = t.count_device(text.ptr, GENOME + 1, positions=True)
        assert positions == GENOME - 30 and hits == positions
        other = capi.DeviceBuffer(api, 100_000 + 1 + capi.TEXT_PAD)
        api.check(L.kache_amd_synth_genome(other.ptr, 100_000, 0xB0B), "synth_genome")
        api.check(L.kache_amd_device_fill(other.ptr + 100_000, ord("\n"), 1 + capi.TEXT_PAD), "fill")
        hits2, positions2 = t.count_device(other.ptr, 100_001, positions=True)
        assert positions2 == 100_000 - 30 and hits2 == 0
    assert L.kache_amd_trim_cache(None) == 0
