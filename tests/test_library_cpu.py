"""The C-ABI library on the CPU: it loads, exports every symbol the headers
declare, rejects invalid arguments before touching a device, and its device
code's hash and describer functions, run on the host, agree with the oracle
and the reference's vectors."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import kache_hash_amd as KA
import oracle as O
from kache_hash_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


@pytest.fixture(scope="module")
def api():
    return KA.api()


def declared_functions():
    names = set()
    for h in ("kache_capi.h", "kache_amd.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        names |= set(re.findall(r"\b(kache_\w+)\s*\(", text))
    return sorted(names)


def test_exports_every_declared_symbol(api):
    names = declared_functions()
    assert len([n for n in names if not n.startswith("kache_amd")]) == 18
    missing = [n for n in names if not hasattr(api.lib, n)]
    assert not missing, missing


def test_supports_and_defaults(api):
    L = api.lib
    assert all(L.kache_supports(31, w) for w in (0, 1, 2, 4, 8))
    assert not any(L.kache_supports(k, w) for k, w in ((25, 0), (31, 3), (31, 16), (32, 0)))
    o = L.kache_default_options(31, 4)
    assert (o.k, o.value_bytes, o.expected_keys, o.load_factor, o.resize_workers, o.pow2_capacity) == (31, 4, 1 << 22, 0.8, 16, 0)
    assert L.kache_strerror(0) == b"success" and L.kache_strerror(3) == b"out of memory"
    assert L.kache_strerror(99) == b"unknown status"


def test_new_rejects_before_allocating(api):
    """test/capi.cpp:406-428: every rejection returns EINVAL and leaves NULL."""
    L = api.lib
    h = C.c_void_p(1234)
    o = L.kache_default_options(31, 4)
    assert L.kache_new(C.byref(o), None) == capi.KACHE_EINVAL
    assert L.kache_new(None, C.byref(h)) == capi.KACHE_EINVAL and not h.value

    def rejects(**kw):
        x = L.kache_default_options(31, 4)
        for k, v in kw.items():
            setattr(x, k, v)
        h = C.c_void_p(1234)
        return L.kache_new(C.byref(x), C.byref(h)) == capi.KACHE_EINVAL and not h.value

    assert rejects(k=25) and rejects(value_bytes=3)
    for lf in (0.0, 1.0, 1.5, float("nan")):
        assert rejects(load_factor=lf)
    assert rejects(resize_workers=0) and rejects(pow2_capacity=2)
    assert rejects(expected_keys=1 << 37) and rejects(expected_keys=55 << 30, pow2_capacity=1)
    assert rejects(expected_keys=(1 << 64) - 1)


def test_null_table_calls_are_rejected(api):
    L = api.lib
    n = C.c_uint64()
    u = C.c_uint32()
    one = C.c_char_p(b"A" * 31)
    ln = (C.c_size_t * 1)(31)
    seqs = (C.c_char_p * 1)(b"A" * 31)
    E = capi.KACHE_EINVAL
    assert L.kache_register(None, C.byref(u)) == E
    assert L.kache_insert(None, 0, seqs, ln, 1, 0, 0, C.byref(n)) == E
    assert L.kache_insert_oriented(None, 0, seqs, ln, 1, 1, 0, 0, C.byref(n)) == E
    assert L.kache_update(None, seqs, ln, 1, 0, 0, C.byref(n)) == E
    buf = np.zeros(8, np.uint64)
    assert L.kache_count(None, seqs, ln, 1, buf.ctypes.data) == E
    assert L.kache_count_masked(None, seqs, ln, 1, 1, 2, buf.ctypes.data) == E
    assert L.kache_sum(None, seqs, ln, 1, buf.ctypes.data) == E
    assert L.kache_lookup(None, seqs, ln, 1, buf.ctypes.data, None, None) == E
    f = C.c_uint8()
    assert L.kache_get(None, one, C.byref(f), C.byref(n)) == E
    assert L.kache_size(None, C.byref(n)) == E
    assert L.kache_reduce(None, 0, 1, 1, C.byref(n)) == E
    assert L.kache_for_each(None, 1, 0, capi.KEYS_FN(lambda *a: 0), None) == E
    assert L.kache_get_stats(None, C.byref(capi.Stats())) == E
    L.kache_free(None)
    assert L.kache_amd_insert_device(None, 0, None, 0, 0, 0, None) == E
    assert L.kache_amd_count_device(None, None, 0, C.byref(n), None) == E
    assert L.kache_amd_get_info(None, None) == E


def test_device_hashes_on_host_match_oracle_and_reference(api):
    s = np.load(os.path.join(GOLD, "scalars.npz"), allow_pickle=False)
    L = api.lib
    for i in range(0, len(s["x"]), 3):
        x = int(s["x"][i])
        assert L.kache_amd_host_minimizer_hash(int(s["canon"][i])) == int(s["minimizer"][i])
        assert L.kache_amd_host_checksum(x) == int(s["c_wang"][i]) == O.checksum_wang(x)
        assert L.kache_amd_host_xxh3(x, 0) == int(s["xxh3_s0"][i])
        assert L.kache_amd_host_xxh3(x, 1 << 63) == int(s["xxh3_s63"][i])
    for n in (0, 1, 10, 1000, 10**6, 10**8, 2 * 10**9):
        for pow2 in (0, 1):
            assert L.kache_amd_host_cap_for(n, 0.8, pow2) == O.cap_for(n, 0.8, bool(pow2))
            assert L.kache_amd_host_cap_for(n, 0.9, pow2) == O.cap_for(n, 0.9, bool(pow2))


def test_device_describer_on_host_matches_reference_vectors(api):
    d = np.load(os.path.join(GOLD, "describer.npz"), allow_pickle=False)
    lens, text = d["lens"], d["text"].tobytes()
    # All sequences in one text, as the device stages them: each followed by a separator.
    parts, starts, o = [], [], 0
    for l in lens:
        starts.append(o)
        parts.append(text[sum(int(x) for x in lens[:len(starts) - 1]):][:int(l)] + b"\n")
        o += int(l) + 1
    t = b"".join(parts)
    n = len(t)
    key, h = np.zeros(n, np.uint64), np.zeros(n, np.uint64)
    c, fwd, valid = np.zeros(n, np.uint8), np.zeros(n, np.uint8), np.zeros(n, np.uint8)
    buf = C.create_string_buffer(t, n)
    assert api.lib.kache_amd_host_describe(buf, n, key.ctypes.data, h.ctypes.data, c.ctypes.data, fwd.ctypes.data, valid.ctypes.data) == 0
    pos = np.array(starts, np.int64)[d["q"]] + d["i"].astype(np.int64)
    assert int(valid.sum()) == len(pos)
    assert valid[pos].all()
    assert (key[pos] == d["key"]).all() and (h[pos] == d["h"]).all()
    assert (c[pos] == d["c_wang"]).all() and (fwd[pos] == d["fwd"]).all()


def test_server_rejects_invalid_arguments(api):
    """The serving loop validates before it starts a worker or touches a device."""
    L = api.lib
    E = capi.KACHE_EINVAL
    h = C.c_void_p()
    tk = C.c_uint64()
    r = capi.Request(kind=capi.REQ_COUNT)
    assert L.kache_amd_server_new(None, 0, 1 << 20, 16, C.byref(h)) == E and not h.value
    assert L.kache_amd_server_submit(None, C.byref(r), C.byref(tk)) == E
    assert L.kache_amd_server_wait(None, 1, None) == E
    assert L.kache_amd_server_get_stats(None, None) == E
    assert L.kache_amd_server_set_hbm_budget(None, 1) == E
    assert L.kache_amd_server_replay(None, None, None, 0, None) == E
    assert L.kache_amd_clock() > 0


def test_cpp_headers_compile(tmp_path):
    """The C++ API headers compile on their own (C++20, -Wall -Werror) and the
    C headers as C."""
    inc = os.path.join(ROOT, "include")
    src = tmp_path / "t.cpp"
    src.write_text('#include "kache_amd.hpp"\n#include "kache_amd_counter.hpp"\n'
                   'int main(){ using M = kache_amd::Streaming_Kmer_Hash_Table<31, true, uint16_t, 22>;'
                   ' using S = kache_amd::Streaming_Kmer_Hash_Table<31, true, void, 22>;'
                   ' return M::fits(1000, 0.8, false) && S::fits(1000, 0.8, true) ? 0 : 1; }\n')
    import subprocess
    subprocess.run(["g++", "-std=c++20", "-Wall", "-Werror", "-fsyntax-only", "-I", inc, str(src)], check=True)
    c = tmp_path / "t.c"
    c.write_text('#include "kache_amd.h"\nint main(void){ return kache_supports(31, 4) ? 0 : 1; }\n')
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-fsyntax-only", "-I", inc, str(c)], check=True)


def test_host_fits_matches_oracle(api):
    for n in (0, 1000, 10**9, 137 * 10**9, 2**40):
        for lf in (0.5, 0.8, 0.9):
            for pow2 in (0, 1):
                assert bool(api.lib.kache_amd_host_fits(n, lf, pow2)) == O.fits(n, lf, bool(pow2)), (n, lf, pow2)


def test_load_rejects_bad_files_before_touching_a_device(api, tmp_path):
    h = C.c_void_p(1234)
    assert api.lib.kache_amd_load(str(tmp_path / "absent").encode(), C.byref(h)) == capi.KACHE_EINVAL and not h.value
    junk = tmp_path / "junk"
    junk.write_bytes(b"KAMDTBL1" + b"\0" * 2000)
    assert api.lib.kache_amd_load(str(junk).encode(), C.byref(h)) == capi.KACHE_EINVAL
    assert "header" in api.last_error() or "version" in api.last_error()
    assert api.lib.kache_amd_load(None, C.byref(h)) == capi.KACHE_EINVAL
    assert api.lib.kache_amd_save(None, b"x") == capi.KACHE_EINVAL
