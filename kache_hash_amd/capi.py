"""ctypes bindings of the kache-hash C ABI (include/kache_capi.h).

The same bindings drive any library exporting that ABI: this package's
libkache_amd.so (the MI355X table) or a build of the reference's
capi/kache_capi.cpp.  `Table` mirrors the reference's C interface one call
per method, with its argument meanings, status codes and threading contract
(capi/kache_capi.h:111-202); the Rust crate's Builder/Inserter/Table phases
(rust/kache-hash-sys/src/lib.rs:264-587) map onto it the same way.
"""
import ctypes as C
import os

import numpy as np

KACHE_OK, KACHE_EINVAL, KACHE_EUSERS, KACHE_ENOMEM, KACHE_EINTERNAL = range(5)
KACHE_MAX_USERS = 128
KEEP, ASSIGN, ADD, OR, MIN, MAX = range(6)
ADD_WRAP = 6     # KACHE_AMD_ADD_WRAP: the C++ API's wrapping v + x (device library only)
REDUCE_KEYS, REDUCE_POPCOUNT, REDUCE_SUM = range(3)
ABSENT, PRESENT, NOT_KMER = range(3)
K = 31
TEXT_PAD = 512   # KACHE_AMD_TEXT_PAD


class Options(C.Structure):
    _fields_ = [("k", C.c_uint32), ("value_bytes", C.c_uint32), ("expected_keys", C.c_uint64),
                ("load_factor", C.c_double), ("resize_workers", C.c_uint32), ("pow2_capacity", C.c_uint32)]


class Stats(C.Structure):
    _fields_ = [("capacity", C.c_uint64), ("bytes", C.c_uint64), ("keys", C.c_uint64), ("overflow_keys", C.c_uint64),
                ("primary_pct", C.c_double), ("secondary_pct", C.c_double), ("overflow_pct", C.c_double)]


class Info(C.Structure):
    _fields_ = [("buckets", C.c_uint64), ("bucket_bytes", C.c_uint64), ("overflow_slots", C.c_uint64),
                ("overflow_keys", C.c_uint64), ("resizes", C.c_uint64), ("device_bytes", C.c_uint64),
                ("value_bytes", C.c_uint64), ("next_doubling_bytes", C.c_uint64), ("device_free_bytes", C.c_uint64),
                ("hbm_budget", C.c_uint64), ("store_mapped_bytes", C.c_uint64), ("store_reserved_bytes", C.c_uint64),
                ("in_place_doublings", C.c_uint64), ("away_list_regrows", C.c_uint64), ("last_away_keys", C.c_uint64),
                ("growth_mode", C.c_uint32),
                ("reserved0", C.c_uint32)]

GROW_AUTO, GROW_IN_PLACE, GROW_OUT_OF_PLACE = range(3)   # KACHE_AMD_GROW_*


REQ_COUNT, REQ_LOOKUP, REQ_UPDATE, REQ_INSERT = range(4)   # KACHE_AMD_REQ_*


class Tuning(C.Structure):
    _fields_ = [("blocks_per_cu", C.c_uint32), ("launch_log2", C.c_uint32), ("graph", C.c_uint32), ("tuning", C.c_uint32),
                ("waves_per_simd", C.c_uint32), ("threads_per_block", C.c_uint32)]


class TuningTrial(C.Structure):
    _fields_ = [("cfg", Tuning), ("launches", C.c_uint32), ("median_ns_per_position", C.c_double),
                ("median_vs_reference", C.c_double)]


class Request(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("op", C.c_uint32), ("x", C.c_uint64), ("seqs", C.c_void_p), ("lens", C.c_void_p),
                ("n", C.c_size_t), ("counts", C.c_void_p), ("found", C.c_void_p), ("values", C.c_void_p), ("fwd", C.c_void_p)]


class RequestResult(C.Structure):
    _fields_ = [("status", C.c_int), ("phase", C.c_uint32), ("ticket", C.c_uint64), ("wave", C.c_uint64),
                ("submit_s", C.c_double), ("start_s", C.c_double), ("done_s", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [("waves", C.c_uint64), ("requests", C.c_uint64), ("max_wave_requests", C.c_uint64),
                ("max_wave_bytes", C.c_uint64), ("direct_waves", C.c_uint64), ("refused", C.c_uint64)]


KEYS_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(C.c_uint64), C.c_void_p, C.c_size_t)
TEXT_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(C.c_char), C.c_uint64)

_P = C.c_void_p
_SEQS = C.c_void_p   # const char* const*
_LENS = C.c_void_p   # const size_t*

CAPI_SYMBOLS = {
    "kache_supports": (C.c_int, [C.c_uint32, C.c_uint32]),
    "kache_default_options": (Options, [C.c_uint32, C.c_uint32]),
    "kache_new": (C.c_int, [C.POINTER(Options), C.POINTER(_P)]),
    "kache_free": (None, [_P]),
    "kache_register": (C.c_int, [_P, C.POINTER(C.c_uint32)]),
    "kache_insert": (C.c_int, [_P, C.c_uint32, _SEQS, _LENS, C.c_size_t, C.c_uint32, C.c_uint64, C.POINTER(C.c_uint64)]),
    "kache_insert_oriented": (C.c_int, [_P, C.c_uint32, _SEQS, _LENS, C.c_size_t, C.c_uint32, C.c_uint64, C.c_uint64, C.POINTER(C.c_uint64)]),
    "kache_update": (C.c_int, [_P, _SEQS, _LENS, C.c_size_t, C.c_uint32, C.c_uint64, C.POINTER(C.c_uint64)]),
    "kache_count": (C.c_int, [_P, _SEQS, _LENS, C.c_size_t, _P]),
    "kache_count_masked": (C.c_int, [_P, _SEQS, _LENS, C.c_size_t, C.c_uint64, C.c_uint64, _P]),
    "kache_sum": (C.c_int, [_P, _SEQS, _LENS, C.c_size_t, _P]),
    "kache_lookup": (C.c_int, [_P, _SEQS, _LENS, C.c_size_t, _P, _P, _P]),
    "kache_get": (C.c_int, [_P, C.c_char_p, C.POINTER(C.c_uint8), C.POINTER(C.c_uint64)]),
    "kache_size": (C.c_int, [_P, C.POINTER(C.c_uint64)]),
    "kache_reduce": (C.c_int, [_P, C.c_uint32, C.c_uint64, C.c_uint32, C.POINTER(C.c_uint64)]),
    "kache_for_each": (C.c_int, [_P, C.c_uint32, C.c_uint32, KEYS_FN, _P]),
    "kache_get_stats": (C.c_int, [_P, C.POINTER(Stats)]),
    "kache_strerror": (C.c_char_p, [C.c_int]),
}

AMD_SYMBOLS = {
    "kache_amd_insert_device": (C.c_int, [_P, C.c_uint32, _P, C.c_uint64, C.c_uint32, C.c_uint64, C.POINTER(C.c_uint64)]),
    "kache_amd_insert_oriented_device": (C.c_int, [_P, C.c_uint32, _P, C.c_uint64, C.c_uint32, C.c_uint64, C.c_uint64, C.POINTER(C.c_uint64)]),
    "kache_amd_update_device": (C.c_int, [_P, _P, C.c_uint64, C.c_uint32, C.c_uint64, C.POINTER(C.c_uint64)]),
    "kache_amd_count_device": (C.c_int, [_P, _P, C.c_uint64, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "kache_amd_insert_batches": (C.c_int, [_P, C.c_uint32, _P, _P, C.c_uint64, C.c_uint32, C.c_uint64, _P]),
    "kache_amd_count_batches": (C.c_int, [_P, _P, _P, C.c_uint64, _P, _P]),
    "kache_amd_device_alloc": (C.c_int, [C.c_uint64, C.POINTER(_P)]),
    "kache_amd_device_free": (C.c_int, [_P]),
    "kache_amd_copy": (C.c_int, [_P, _P, C.c_uint64, C.c_int]),
    "kache_amd_device_fill": (C.c_int, [_P, C.c_uint8, C.c_uint64]),
    "kache_amd_device_sync": (C.c_int, []),
    "kache_amd_host_alloc": (C.c_int, [C.c_uint64, C.POINTER(_P)]),
    "kache_amd_host_free": (C.c_int, [_P]),
    "kache_amd_lookup_groups": (C.c_int, [_P, _SEQS, _LENS, C.c_size_t, _P, C.c_size_t, _P, _P, _P]),
    "kache_amd_count_seqs_device": (C.c_int, [_P, _P, C.c_uint64, _P, C.c_size_t, _P]),
    "kache_amd_count_seqs_device_static": (C.c_int, [_P, _P, C.c_uint64, _P, C.c_size_t, _P]),
    "kache_amd_lookup_groups_device": (C.c_int, [_P, _P, C.c_uint64, _P, C.c_size_t, _P, C.c_size_t, _P, _P, _P]),
    "kache_amd_insert_text": (C.c_int, [_P, C.c_uint32, _P, C.c_uint64, C.c_uint32, C.c_uint64, C.POINTER(C.c_uint64)]),
    "kache_amd_update_text": (C.c_int, [_P, _P, C.c_uint64, C.c_uint32, C.c_uint64, C.POINTER(C.c_uint64)]),
    "kache_amd_histogram": (C.c_int, [_P, C.c_uint64, _P]),
    "kache_amd_save": (C.c_int, [_P, C.c_char_p]),
    "kache_amd_load": (C.c_int, [C.c_char_p, C.POINTER(_P)]),
    "kache_amd_insert_seqs": (C.c_int, [_P, C.c_uint32, _SEQS, _LENS, C.c_size_t, C.c_uint32, C.c_uint64, C.POINTER(C.c_uint64)]),
    "kache_amd_update_seqs": (C.c_int, [_P, _SEQS, _LENS, C.c_size_t, C.c_uint32, C.c_uint64, C.POINTER(C.c_uint64)]),
    "kache_amd_read_fastx": (C.c_int, [C.c_char_p, TEXT_FN, _P]),
    "kache_amd_insert_fastx": (C.c_int, [_P, C.c_uint32, C.c_char_p, C.c_uint32, C.c_uint64, C.POINTER(C.c_uint64)]),
    "kache_amd_update_fastx": (C.c_int, [_P, C.c_char_p, C.c_uint32, C.c_uint64, C.POINTER(C.c_uint64)]),
    "kache_amd_synth_genome": (C.c_int, [_P, C.c_uint64, C.c_uint64]),
    "kache_amd_synth_reads": (C.c_int, [_P, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64, C.c_double, C.c_double, C.c_double]),
    "kache_amd_host_synth_reads": (C.c_int, [_P, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64, C.c_double, C.c_double, C.c_double]),
    "kache_amd_synth_error": (C.c_char_p, []),
    "kache_amd_synth_seqs": (C.c_int, [_P, _P, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64, C.c_double, C.c_double, C.c_double]),
    "kache_amd_set_tuning": (C.c_int, [_P, C.POINTER(Tuning)]),
    "kache_amd_get_tuning": (C.c_int, [_P, C.POINTER(Tuning)]),
    "kache_amd_autotune": (C.c_int, [_P, C.c_uint32]),
    "kache_amd_tuning_trials": (C.c_int, [_P, _P, C.c_size_t, C.POINTER(C.c_size_t)]),
    "kache_amd_describe_device": (C.c_int, [_P, C.c_uint64, _P, _P, _P, _P, _P]),
    "kache_amd_hashes_device": (C.c_int, [_P, C.c_uint64, C.c_uint64, _P]),
    "kache_amd_profile": (C.c_int, [_P, C.c_int]),
    "kache_amd_kernel_time": (C.c_int, [_P, C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_uint64)]),
    "kache_amd_set_blocks_per_cu": (C.c_int, [_P, C.c_uint32]),
    "kache_amd_set_hbm_budget": (C.c_int, [_P, C.c_uint64]),
    "kache_amd_set_growth": (C.c_int, [_P, C.c_uint32, C.c_uint64]),
    "kache_amd_trim_cache": (C.c_int, [_P]),
    "kache_amd_get_info": (C.c_int, [_P, C.POINTER(Info)]),
    "kache_amd_last_error": (C.c_char_p, []),
    "kache_amd_server_new": (C.c_int, [_P, C.c_uint32, C.c_uint64, C.c_uint32, C.POINTER(_P)]),
    "kache_amd_server_free": (None, [_P]),
    "kache_amd_server_submit": (C.c_int, [_P, C.POINTER(Request), C.POINTER(C.c_uint64)]),
    "kache_amd_server_wait": (C.c_int, [_P, C.c_uint64, C.POINTER(RequestResult)]),
    "kache_amd_server_get_stats": (C.c_int, [_P, C.POINTER(ServerStats)]),
    "kache_amd_server_set_hbm_budget": (C.c_int, [_P, C.c_uint64]),
    "kache_amd_server_replay": (C.c_int, [_P, C.POINTER(Request), _P, C.c_size_t, C.POINTER(RequestResult)]),
    "kache_amd_clock": (C.c_double, []),
    "kache_amd_host_minimizer_hash": (C.c_uint64, [C.c_uint64]),
    "kache_amd_host_checksum": (C.c_uint8, [C.c_uint64]),
    "kache_amd_host_xxh3": (C.c_uint64, [C.c_uint64, C.c_uint64]),
    "kache_amd_host_cap_for": (C.c_uint64, [C.c_uint64, C.c_double, C.c_uint32]),
    "kache_amd_host_fits": (C.c_int, [C.c_uint64, C.c_double, C.c_uint32]),
    "kache_amd_host_describe": (C.c_int, [_P, C.c_uint64, _P, _P, _P, _P, _P]),
}


class KacheError(RuntimeError):
    def __init__(self, status, what, detail=""):
        super().__init__(f"{what}: status {status}" + (f" ({detail})" if detail else ""))
        self.status = status


class Api:
    """A loaded library exporting the kache C ABI (and, for ours, the extensions)."""

    def __init__(self, path):
        self.path = path
        self.lib = C.CDLL(path)
        for name, (res, args) in CAPI_SYMBOLS.items():
            f = getattr(self.lib, name)
            f.restype, f.argtypes = res, args
        self.amd = hasattr(self.lib, "kache_amd_insert_device")
        if self.amd:
            for name, (res, args) in AMD_SYMBOLS.items():
                f = getattr(self.lib, name)
                f.restype, f.argtypes = res, args

    def last_error(self):
        return self.lib.kache_amd_last_error().decode() if self.amd else ""

    def check(self, status, what):
        if status != KACHE_OK:
            raise KacheError(status, what, self.last_error())


class Batch:
    """Sequences as the C ABI takes them: an array of pointers and lengths,
    here into one contiguous buffer (kept alive by the batch)."""

    def __init__(self, seqs):
        seqs = [s.encode() if isinstance(s, str) else bytes(s) for s in seqs]
        self.n = len(seqs)
        self.lens = np.array([len(s) for s in seqs], dtype=np.uint64)
        starts = np.zeros(self.n + 1, dtype=np.uint64)
        if self.n:
            np.cumsum(self.lens, out=starts[1:])
        self.buf = np.frombuffer(b"".join(seqs) + b"\0", dtype=np.uint8).copy()
        self.ptrs = (starts[:-1] + np.uint64(self.buf.ctypes.data)).astype(np.uint64)
        self.seqs = seqs

    @classmethod
    def from_buffer(cls, buf, starts, lens):
        b = cls.__new__(cls)
        b.buf = buf
        b.n = len(lens)
        b.lens = np.ascontiguousarray(lens, dtype=np.uint64)
        b.ptrs = (np.asarray(starts, dtype=np.uint64) + np.uint64(buf.ctypes.data)).astype(np.uint64)
        b.seqs = None
        return b

    def positions(self):
        l = self.lens.astype(np.int64)
        return int(np.maximum(l - (K - 1), 0).sum())

    @property
    def c_seqs(self):
        return self.ptrs.ctypes.data if self.n else None

    @property
    def c_lens(self):
        return self.lens.ctypes.data if self.n else None


def as_batch(seqs):
    return seqs if isinstance(seqs, Batch) else Batch(seqs)


class Table:
    """A table behind the C ABI: a set (width 0) or a map to values of
    `width` bytes, sized for `expected_keys` at `load_factor`."""

    def __init__(self, api, width=0, expected_keys=1 << 22, load_factor=0.8, pow2=False, workers=16):
        self.api = api
        self.width = width
        o = api.lib.kache_default_options(K, width)
        o.expected_keys, o.load_factor, o.resize_workers, o.pow2_capacity = expected_keys, load_factor, workers, int(pow2)
        h = C.c_void_p()
        api.check(api.lib.kache_new(C.byref(o), C.byref(h)), "kache_new")
        self.h = h

    def close(self):
        if self.h:
            self.api.lib.kache_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def register(self):
        u = C.c_uint32()
        self.api.check(self.api.lib.kache_register(self.h, C.byref(u)), "kache_register")
        return u.value

    def insert(self, seqs, op=KEEP, x=0, user=0):
        b = as_batch(seqs)
        n = C.c_uint64()
        self.api.check(self.api.lib.kache_insert(self.h, user, b.c_seqs, b.c_lens, b.n, op, x, C.byref(n)), "kache_insert")
        return n.value

    def insert_oriented(self, seqs, op, x_fwd, x_rev, user=0):
        b = as_batch(seqs)
        n = C.c_uint64()
        self.api.check(self.api.lib.kache_insert_oriented(self.h, user, b.c_seqs, b.c_lens, b.n, op, x_fwd, x_rev, C.byref(n)), "kache_insert_oriented")
        return n.value

    def update(self, seqs, op, x):
        b = as_batch(seqs)
        n = C.c_uint64()
        self.api.check(self.api.lib.kache_update(self.h, b.c_seqs, b.c_lens, b.n, op, x, C.byref(n)), "kache_update")
        return n.value

    def _per_seq(self, fn, b, *extra):
        out = np.zeros(max(b.n, 1), dtype=np.uint64)
        self.api.check(fn(self.h, b.c_seqs, b.c_lens, b.n, *extra, out.ctypes.data), fn.__name__)
        return out[:b.n]

    def count(self, seqs):
        return self._per_seq(self.api.lib.kache_count, as_batch(seqs))

    def count_masked(self, seqs, mask_fwd, mask_rev):
        return self._per_seq(self.api.lib.kache_count_masked, as_batch(seqs), mask_fwd, mask_rev)

    def sum(self, seqs):
        return self._per_seq(self.api.lib.kache_sum, as_batch(seqs))

    def lookup(self, seqs, values=True, fwd=True):
        b = as_batch(seqs)
        p = b.positions()
        found = np.zeros(max(p, 1), dtype=np.uint8)
        vals = np.zeros(max(p, 1) * max(self.width, 1), dtype=np.uint8) if (values and self.width) else None
        fw = np.zeros(max(p, 1), dtype=np.uint8) if fwd else None
        self.api.check(self.api.lib.kache_lookup(self.h, b.c_seqs, b.c_lens, b.n, found.ctypes.data,
                                                 vals.ctypes.data if vals is not None else None,
                                                 fw.ctypes.data if fw is not None else None), "kache_lookup")
        v = None
        if vals is not None:
            v = vals[:p * self.width].view({1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}[self.width]).astype(np.uint64)
        return found[:p], v, (fw[:p] if fw is not None else None)

    def get(self, kmer):
        f = C.c_uint8()
        v = C.c_uint64()
        kmer = kmer.encode() if isinstance(kmer, str) else bytes(kmer)
        self.api.check(self.api.lib.kache_get(self.h, kmer, C.byref(f), C.byref(v)), "kache_get")
        return bool(f.value), v.value

    def size(self):
        n = C.c_uint64()
        self.api.check(self.api.lib.kache_size(self.h, C.byref(n)), "kache_size")
        return n.value

    def reduce(self, what, mask=~0 & (2**64 - 1), workers=1):
        n = C.c_uint64()
        self.api.check(self.api.lib.kache_reduce(self.h, what, mask, workers, C.byref(n)), "kache_reduce")
        return n.value

    def for_each(self, shards=1, shard=0, stop_after=None):
        """Returns (keys, values, chunks) of the share, sorted by key."""
        keys, vals = [], []
        chunks = [0]
        w = self.width

        def cb(ctx, kp, vp, n):
            keys.append(np.ctypeslib.as_array(kp, shape=(n,)).copy())
            if vp and w:
                raw = np.ctypeslib.as_array(C.cast(vp, C.POINTER(C.c_uint8)), shape=(n * w,)).copy()
                vals.append(raw.view({1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}[w]).astype(np.uint64))
            chunks[0] += 1
            return 1 if (stop_after is not None and chunks[0] >= stop_after) else 0

        fn = KEYS_FN(cb)
        self.api.check(self.api.lib.kache_for_each(self.h, shards, shard, fn, None), "kache_for_each")
        k = np.concatenate(keys) if keys else np.zeros(0, np.uint64)
        v = np.concatenate(vals) if vals else np.zeros(len(k), np.uint64)
        o = np.argsort(k, kind="stable")
        return k[o], v[o], chunks[0]

    def stats(self):
        s = Stats()
        self.api.check(self.api.lib.kache_get_stats(self.h, C.byref(s)), "kache_get_stats")
        return {f: getattr(s, f) for f, _ in Stats._fields_}

    # --- device extensions (this package's library only) ---

    def info(self):
        i = Info()
        self.api.check(self.api.lib.kache_amd_get_info(self.h, C.byref(i)), "kache_amd_get_info")
        return {f: getattr(i, f) for f, _ in Info._fields_}

    def insert_device(self, text_ptr, text_bytes, op=KEEP, x=0, user=0):
        n = C.c_uint64()
        self.api.check(self.api.lib.kache_amd_insert_device(self.h, user, text_ptr, text_bytes, op, x, C.byref(n)), "kache_amd_insert_device")
        return n.value

    def insert_oriented_device(self, text_ptr, text_bytes, op, x_fwd, x_rev, user=0):
        n = C.c_uint64()
        self.api.check(self.api.lib.kache_amd_insert_oriented_device(self.h, user, text_ptr, text_bytes, op, x_fwd, x_rev, C.byref(n)), "kache_amd_insert_oriented_device")
        return n.value

    def insert_batches(self, text_ptr, bounds, op=KEEP, x=0, user=0):
        """Pipelined inserts of the batches [bounds[i], bounds[i+1]) of a device text; per-batch new-key counts."""
        b = np.ascontiguousarray(bounds, dtype=np.uint64)
        n = len(b) - 1
        out = np.zeros(max(n, 1), np.uint64)
        self.api.check(self.api.lib.kache_amd_insert_batches(self.h, user, text_ptr, b.ctypes.data, n, op, x, out.ctypes.data), "kache_amd_insert_batches")
        return out[:n]

    def count_batches(self, text_ptr, bounds):
        """(hits, positions) per batch of a device text."""
        b = np.ascontiguousarray(bounds, dtype=np.uint64)
        n = len(b) - 1
        h, p = np.zeros(max(n, 1), np.uint64), np.zeros(max(n, 1), np.uint64)
        self.api.check(self.api.lib.kache_amd_count_batches(self.h, text_ptr, b.ctypes.data, n, h.ctypes.data, p.ctypes.data), "kache_amd_count_batches")
        return h[:n], p[:n]

    def update_device(self, text_ptr, text_bytes, op, x):
        n = C.c_uint64()
        self.api.check(self.api.lib.kache_amd_update_device(self.h, text_ptr, text_bytes, op, x, C.byref(n)), "kache_amd_update_device")
        return n.value

    def count_device(self, text_ptr, text_bytes, positions=False):
        n, p = C.c_uint64(), C.c_uint64()
        self.api.check(self.api.lib.kache_amd_count_device(self.h, text_ptr, text_bytes, C.byref(n), C.byref(p)), "kache_amd_count_device")
        return (n.value, p.value) if positions else n.value

    def insert_text(self, text, op=KEEP, x=0, user=0):
        """Inserts a host text (sequences separated by non-ACGT bytes): new keys."""
        buf = np.frombuffer(text, np.uint8) if isinstance(text, (bytes, bytearray)) else np.ascontiguousarray(text, np.uint8)
        n = C.c_uint64()
        self.api.check(self.api.lib.kache_amd_insert_text(self.h, user, buf.ctypes.data if buf.size else None, buf.size, op, x, C.byref(n)), "kache_amd_insert_text")
        return n.value

    def update_text(self, text, op, x):
        """Updates the present k-mers of a host text: present positions."""
        buf = np.frombuffer(text, np.uint8) if isinstance(text, (bytes, bytearray)) else np.ascontiguousarray(text, np.uint8)
        n = C.c_uint64()
        self.api.check(self.api.lib.kache_amd_update_text(self.h, buf.ctypes.data if buf.size else None, buf.size, op, x, C.byref(n)), "kache_amd_update_text")
        return n.value

    def insert_fastx(self, path, op=KEEP, x=0, user=0):
        n = C.c_uint64()
        self.api.check(self.api.lib.kache_amd_insert_fastx(self.h, user, os.fsencode(path), op, x, C.byref(n)), "kache_amd_insert_fastx")
        return n.value

    def update_fastx(self, path, op, x):
        n = C.c_uint64()
        self.api.check(self.api.lib.kache_amd_update_fastx(self.h, os.fsencode(path), op, x, C.byref(n)), "kache_amd_update_fastx")
        return n.value

    def save(self, path):
        self.api.check(self.api.lib.kache_amd_save(self.h, os.fsencode(path)), "kache_amd_save")

    @classmethod
    def load(cls, api, path):
        """A table made from a file written by save()."""
        h = C.c_void_p()
        api.check(api.lib.kache_amd_load(os.fsencode(path), C.byref(h)), "kache_amd_load")
        t = cls.__new__(cls)
        t.api, t.h = api, h
        i = Info()
        api.check(api.lib.kache_amd_get_info(h, C.byref(i)), "kache_amd_get_info")
        t.width = i.value_bytes
        return t

    def histogram(self, max_count):
        """out[c] = keys with value c, c in [1, max_count]; out[0] = 0."""
        out = np.zeros(max_count + 1, np.uint64)
        self.api.check(self.api.lib.kache_amd_histogram(self.h, max_count, out.ctypes.data), "kache_amd_histogram")
        return out

    def profile(self, on=True):
        self.api.check(self.api.lib.kache_amd_profile(self.h, int(on)), "kache_amd_profile")

    def kernel_time(self, name):
        ms = C.c_double()
        n = C.c_uint64()
        self.api.check(self.api.lib.kache_amd_kernel_time(self.h, name.encode(), C.byref(ms), C.byref(n)), "kache_amd_kernel_time")
        return ms.value, n.value

    def set_tuning(self, blocks_per_cu=0, launch_log2=24, graph=False, waves_per_simd=0, threads_per_block=0):
        cfg = Tuning(blocks_per_cu, launch_log2, int(graph), 0, waves_per_simd, threads_per_block)
        self.api.check(self.api.lib.kache_amd_set_tuning(self.h, C.byref(cfg)), "kache_amd_set_tuning")

    def tuning(self):
        cfg = Tuning()
        self.api.check(self.api.lib.kache_amd_get_tuning(self.h, C.byref(cfg)), "kache_amd_get_tuning")
        return {f: getattr(cfg, f) for f, _ in Tuning._fields_}

    def autotune(self, rounds=2):
        self.api.check(self.api.lib.kache_amd_autotune(self.h, rounds), "kache_amd_autotune")

    def tuning_trials(self):
        n = C.c_size_t()
        self.api.check(self.api.lib.kache_amd_tuning_trials(self.h, None, 0, C.byref(n)), "kache_amd_tuning_trials")
        out = (TuningTrial * max(n.value, 1))()
        self.api.check(self.api.lib.kache_amd_tuning_trials(self.h, out, n.value, C.byref(n)), "kache_amd_tuning_trials")
        return [{"blocks_per_cu": t.cfg.blocks_per_cu, "launch_log2": t.cfg.launch_log2, "graph": t.cfg.graph,
                 "waves_per_simd": t.cfg.waves_per_simd, "threads_per_block": t.cfg.threads_per_block, "launches": t.launches,
                 "median_ns_per_position": t.median_ns_per_position,
                 "median_vs_reference": t.median_vs_reference} for t in out[:n.value]]

    def set_hbm_budget(self, nbytes):
        """Most bytes of HBM the table may hold (0: no limit but the device's)."""
        self.api.check(self.api.lib.kache_amd_set_hbm_budget(self.h, nbytes), "kache_amd_set_hbm_budget")

    def set_growth(self, mode, away_entries=0):
        """How the table doubles: GROW_AUTO, GROW_IN_PLACE or GROW_OUT_OF_PLACE;
        away_entries sets the first size of an in-place doubling's list of
        away keys (0: the default)."""
        self.api.check(self.api.lib.kache_amd_set_growth(self.h, mode, away_entries), "kache_amd_set_growth")

    def set_blocks_per_cu(self, blocks):
        self.api.check(self.api.lib.kache_amd_set_blocks_per_cu(self.h, blocks), "kache_amd_set_blocks_per_cu")


class DeviceBuffer:
    """Device memory from the library (hipMalloc), freed with the object."""

    def __init__(self, api, nbytes):
        self.api, self.nbytes = api, nbytes
        p = C.c_void_p()
        api.check(api.lib.kache_amd_device_alloc(nbytes, C.byref(p)), "kache_amd_device_alloc")
        self.ptr = p.value

    def free(self):
        if self.ptr:
            self.api.lib.kache_amd_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def upload(self, data, offset=0):
        b = np.ascontiguousarray(np.frombuffer(bytes(data), np.uint8) if not isinstance(data, np.ndarray) else data)
        self.api.check(self.api.lib.kache_amd_copy(self.ptr + offset, b.ctypes.data, b.nbytes, 1), "kache_amd_copy")

    def download(self, nbytes=None, offset=0, dtype=np.uint8):
        n = self.nbytes - offset if nbytes is None else nbytes
        out = np.empty(n, np.uint8)
        self.api.check(self.api.lib.kache_amd_copy(out.ctypes.data, self.ptr + offset, n, 0), "kache_amd_copy")
        return out.view(dtype)


class HostArray:
    """A zeroed numpy array in page-locked host memory from the library
    (kache_amd_host_alloc), which the device's copy engines write directly;
    the memory is freed with this object (keep it while the array is used)."""

    def __init__(self, api, n, dtype):
        self.api = api
        nbytes = max(n, 1) * np.dtype(dtype).itemsize
        p = C.c_void_p()
        api.check(api.lib.kache_amd_host_alloc(nbytes, C.byref(p)), "kache_amd_host_alloc")
        self.ptr = p.value
        self.array = np.frombuffer((C.c_uint8 * nbytes).from_address(self.ptr), np.uint8).view(dtype)[:max(n, 1)]
        self.array.fill(0)

    def __del__(self):
        if getattr(self, "ptr", None):
            self.api.lib.kache_amd_host_free(self.ptr)
            self.ptr = None


def make_request(width, kind, seqs, op=KEEP, x=0, values=True, fwd=True, api=None):
    """A serving request over a table of `width`-byte values and the arrays it
    writes: (Request, outputs, batch); reusable across servers and runs.  With
    `api`, the output arrays are page-locked (HostArray, kept in
    outputs["_host"]), so lookups are copied into them by DMA directly."""
    b = as_batch(seqs)
    r = Request(kind=kind, op=op, x=x, seqs=b.c_seqs, lens=b.c_lens, n=b.n)
    out = {}
    held = []

    def arr(n, dtype):
        if api is None:
            return np.zeros(n, dtype)
        h = HostArray(api, n, dtype)
        held.append(h)
        return h.array

    if kind == REQ_COUNT:
        out["counts"] = arr(max(b.n, 1), np.uint64)
        r.counts = out["counts"].ctypes.data
    elif kind == REQ_LOOKUP:
        p = max(b.positions(), 1)
        out["found"] = arr(p, np.uint8)
        r.found = out["found"].ctypes.data
        if values and width:
            out["values"] = arr(p * width, np.uint8)
            r.values = out["values"].ctypes.data
        if fwd:
            out["fwd"] = arr(p, np.uint8)
            r.fwd = out["fwd"].ctypes.data
    else:
        out["counts"] = arr(1, np.uint64)
        r.counts = out["counts"].ctypes.data
    if held:
        out["_host"] = held
    return r, out, b


class Server:
    """The serving loop of a table (kache_amd_server_*): requests submitted
    from any thread run in waves on one worker; `user` runs the upserts."""

    def __init__(self, table, user, wave_bytes=64 << 20, max_wave_requests=1 << 20):
        self.table, self.api = table, table.api
        h = C.c_void_p()
        self.api.check(self.api.lib.kache_amd_server_new(table.h, user, wave_bytes, max_wave_requests, C.byref(h)), "kache_amd_server_new")
        self.h = h
        self.keep = {}

    def close(self):
        if self.h:
            self.api.lib.kache_amd_server_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def request(self, kind, seqs, op=KEEP, x=0, values=True, fwd=True, pinned=False):
        """A request and the arrays it writes: (Request, outputs, batch); with
        `pinned`, the outputs are page-locked (see make_request)."""
        return make_request(self.table.width, kind, seqs, op, x, values, fwd, api=self.api if pinned else None)

    def submit(self, req):
        """Submits (Request, outputs, batch) from request(); returns the ticket."""
        t = C.c_uint64()
        self.api.check(self.api.lib.kache_amd_server_submit(self.h, C.byref(req[0]), C.byref(t)), "kache_amd_server_submit")
        self.keep[t.value] = req
        return t.value

    def wait(self, ticket):
        """Waits for a ticket; returns (result dict, outputs)."""
        res = RequestResult()
        self.api.check(self.api.lib.kache_amd_server_wait(self.h, ticket, C.byref(res)), "kache_amd_server_wait")
        req = self.keep.pop(ticket)
        return {f: getattr(res, f) for f, _ in RequestResult._fields_}, req[1]

    def replay(self, reqs, arrivals):
        """Open-loop load: request i submitted at arrivals[i] s; returns the results (a list of dicts)."""
        n = len(reqs)
        arr = (Request * max(n, 1))(*[r[0] for r in reqs])
        times = np.ascontiguousarray(arrivals, dtype=np.float64)
        out = (RequestResult * max(n, 1))()
        self.api.check(self.api.lib.kache_amd_server_replay(self.h, arr, times.ctypes.data, n, out), "kache_amd_server_replay")
        return [{f: getattr(out[i], f) for f, _ in RequestResult._fields_} for i in range(n)]

    def set_hbm_budget(self, nbytes):
        """Staging bytes the server's slots may hold together (0: no budget of its own)."""
        self.api.check(self.api.lib.kache_amd_server_set_hbm_budget(self.h, nbytes), "kache_amd_server_set_hbm_budget")

    def stats(self):
        s = ServerStats()
        self.api.check(self.api.lib.kache_amd_server_get_stats(self.h, C.byref(s)), "kache_amd_server_get_stats")
        return {f: getattr(s, f) for f, _ in ServerStats._fields_}


def read_fastx(api, path):
    """The sequences of a FASTA/FASTQ file as the library reads them: a list
    of bytes.  Raises KacheError on malformed input."""
    blocks = []

    def cb(ctx, p, n):
        blocks.append(C.string_at(p, n))
        return 0

    fn = TEXT_FN(cb)
    api.check(api.lib.kache_amd_read_fastx(os.fsencode(path), fn, None), "kache_amd_read_fastx")
    text = b"".join(blocks)
    return text.split(b"\n")[:-1] if text else []


def device_text(api, seqs):
    """Sequences as a device text: each followed by a separator, padded.
    Returns (buffer, text_bytes)."""
    t = b"".join(bytes(s) + b"\n" for s in seqs)
    buf = DeviceBuffer(api, len(t) + TEXT_PAD)
    buf.upload(t + b"\n" * TEXT_PAD)
    return buf, len(t)


def library_path():
    return os.path.join(os.path.dirname(os.path.abspath(__file__)), "libkache_amd.so")
