"""ctypes binding of the oracle (oracle/liboracle.so, built from kache_oracle.c).

Test infrastructure only: tests/, the smoke check and the CPU-baseline leg of
bench.py use it as a checker.  The product never imports it.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

_lib = None


def build():
    import subprocess
    subprocess.run(["make", "-C", HERE, "oracle"], check=True, stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        l = C.CDLL(LIB)
        u64, u8, u32, d, P = C.c_uint64, C.c_uint8, C.c_uint32, C.c_double, C.c_void_p
        sig = {
            "ko_wang_u64": (u64, [u64]), "ko_checksum_mul": (u8, [u64]), "ko_checksum_wang": (u8, [u64]),
            "ko_revcomp31": (u64, [u64]), "ko_minimizer_hash": (u64, [u64]), "ko_xxh3_8": (u64, [u64, u64]),
            "ko_describe": (C.c_size_t, [P, C.c_size_t, C.c_int, P, P, P, P, P]),
            "ko_cap_for": (u64, [u64, d, C.c_int]), "ko_fits": (C.c_int, [u64, d, C.c_int]),
            "ko_bucket_of": (u64, [u64, u64]), "ko_filter_pos": (u32, [u64]), "ko_pick_bucket": (u64, [u64, u64]),
            "ko_secondaries": (None, [u64, u64, P, P]), "ko_ov_slots_for": (u64, [u64]),
            "ko_sched_init": (None, [P, u64, d, C.c_int]), "ko_sched_call": (None, [P, u32, u64]),
            "ko_tab_new": (P, [u32]), "ko_tab_free": (None, [P]), "ko_tab_size": (u64, [P]),
            "ko_tab_insert": (u64, [P, P, P, C.c_size_t, u32, u64]),
            "ko_tab_insert_oriented": (u64, [P, P, P, C.c_size_t, u32, u64, u64]),
            "ko_tab_update": (u64, [P, P, P, C.c_size_t, u32, u64]),
            "ko_tab_lookup": (None, [P, P, P, C.c_size_t, P, P, P]),
            "ko_tab_count": (None, [P, P, P, C.c_size_t, P, P, u64, u64, P]),
            "ko_tab_dump": (u64, [P, P, P, u64]), "ko_tab_reduce": (u64, [P, u32, u64]),
            "ko_tab_get": (C.c_int, [P, u64, P]),
        }
        for name, (res, args) in sig.items():
            f = getattr(l, name)
            f.restype, f.argtypes = res, args
        _lib = l
    return _lib


K = 31


def wang(x): return lib().ko_wang_u64(x)
def checksum_mul(key): return lib().ko_checksum_mul(key)
def checksum_wang(key): return lib().ko_checksum_wang(key)
def revcomp31(x): return lib().ko_revcomp31(x)
def minimizer_hash(key): return lib().ko_minimizer_hash(key)
def xxh3_8(x, seed=0): return lib().ko_xxh3_8(x, seed)
def cap_for(n, lf=0.8, pow2=False): return lib().ko_cap_for(n, lf, int(pow2))
def fits(n, lf=0.8, pow2=False): return bool(lib().ko_fits(n, lf, int(pow2)))
def bucket_of(h, cap): return lib().ko_bucket_of(h, cap)
def filter_pos(h): return lib().ko_filter_pos(h)
def pick_bucket(x, cap): return lib().ko_pick_bucket(x, cap)
def ov_slots_for(n): return lib().ko_ov_slots_for(n)


def secondaries(h, cap):
    p, q = C.c_uint64(), C.c_uint64()
    lib().ko_secondaries(h, cap, C.byref(p), C.byref(q))
    return p.value, q.value


def describe(seq, wang_checksum=True):
    """Per position of `seq`: (valid, key, h, c, fwd) arrays."""
    s = seq.encode() if isinstance(seq, str) else bytes(seq)
    n = max(len(s) - (K - 1), 0)
    valid = np.zeros(max(n, 1), np.uint8)
    key = np.zeros(max(n, 1), np.uint64)
    h = np.zeros(max(n, 1), np.uint64)
    c = np.zeros(max(n, 1), np.uint8)
    fwd = np.zeros(max(n, 1), np.uint8)
    buf = C.create_string_buffer(s, len(s) + 1)
    lib().ko_describe(buf, len(s), int(wang_checksum), valid.ctypes.data, key.ctypes.data, h.ctypes.data, c.ctypes.data, fwd.ctypes.data)
    return valid[:n], key[:n], h[:n], c[:n], fwd[:n]


class Sched(C.Structure):
    """The reference's growth schedule: capacity in buckets after calls."""
    _fields_ = [("cap", C.c_uint64), ("resize_th", C.c_uint64), ("step", C.c_uint64), ("approx_sz", C.c_uint64),
                ("checkpoint", C.c_uint64 * 128), ("lf", C.c_double), ("resizes", C.c_uint32)]

    def __init__(self, n, lf=0.8, pow2=False):
        super().__init__()
        lib().ko_sched_init(C.byref(self), n, lf, int(pow2))

    def call(self, user, inserted):
        lib().ko_sched_call(C.byref(self), user, inserted)


class _Seqs:
    def __init__(self, seqs):
        seqs = [s.encode() if isinstance(s, str) else bytes(s) for s in seqs]
        self.n = len(seqs)
        self.lens = np.array([len(s) for s in seqs] or [0], dtype=np.uint64)
        starts = np.zeros(self.n + 1, np.uint64)
        if self.n:
            np.cumsum(self.lens[:self.n], out=starts[1:])
        self.buf = np.frombuffer(b"".join(seqs) + b"\0", np.uint8).copy()
        self.ptrs = (starts[:max(self.n, 1)] + np.uint64(self.buf.ctypes.data)).astype(np.uint64)

    def positions(self):
        l = self.lens[:self.n].astype(np.int64)
        return int(np.maximum(l - (K - 1), 0).sum())


class OTable:
    """The semantic table: canonical 31-mers to values of `width` bytes."""

    def __init__(self, width=0):
        self.width = width
        self.h = lib().ko_tab_new(width)

    def __del__(self):
        if getattr(self, "h", None):
            lib().ko_tab_free(self.h)
            self.h = None

    def size(self): return lib().ko_tab_size(self.h)

    def insert(self, seqs, op=0, x=0):
        s = _Seqs(seqs)
        return lib().ko_tab_insert(self.h, s.ptrs.ctypes.data, s.lens.ctypes.data, s.n, op, x)

    def insert_oriented(self, seqs, op, x_fwd, x_rev):
        s = _Seqs(seqs)
        return lib().ko_tab_insert_oriented(self.h, s.ptrs.ctypes.data, s.lens.ctypes.data, s.n, op, x_fwd, x_rev)

    def update(self, seqs, op, x):
        s = _Seqs(seqs)
        return lib().ko_tab_update(self.h, s.ptrs.ctypes.data, s.lens.ctypes.data, s.n, op, x)

    def lookup(self, seqs):
        s = _Seqs(seqs)
        p = s.positions()
        found = np.zeros(max(p, 1), np.uint8)
        vals = np.zeros(max(p, 1), np.uint64)
        fwd = np.zeros(max(p, 1), np.uint8)
        lib().ko_tab_lookup(self.h, s.ptrs.ctypes.data, s.lens.ctypes.data, s.n, found.ctypes.data, vals.ctypes.data, fwd.ctypes.data)
        return found[:p], vals[:p], fwd[:p]

    def counts(self, seqs, mask_fwd=0, mask_rev=0):
        """(counts, sums, masked counts) per sequence."""
        s = _Seqs(seqs)
        n = max(s.n, 1)
        c, sm, m = np.zeros(n, np.uint64), np.zeros(n, np.uint64), np.zeros(n, np.uint64)
        lib().ko_tab_count(self.h, s.ptrs.ctypes.data, s.lens.ctypes.data, s.n, c.ctypes.data, sm.ctypes.data, mask_fwd, mask_rev, m.ctypes.data)
        return c[:s.n], sm[:s.n], m[:s.n]

    def dump(self):
        n = self.size()
        k = np.zeros(max(n, 1), np.uint64)
        v = np.zeros(max(n, 1), np.uint64)
        lib().ko_tab_dump(self.h, k.ctypes.data, v.ctypes.data, n)
        return k[:n], v[:n]

    def reduce(self, what, mask):
        return lib().ko_tab_reduce(self.h, what, mask)

    def get(self, key):
        v = C.c_uint64()
        f = lib().ko_tab_get(self.h, key, C.byref(v))
        return bool(f), v.value


def ref_library():
    """Path of the reference C ABI built from its sources (oracle/_ref), for
    this host's ISA, or None if it was not built."""
    flags = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    flags = line
                    break
    except OSError:
        pass
    names = ["libkache_ref_avx512.so", "libkache_ref_avx2.so"] if all(
        t in flags.split() for t in ("avx512f", "avx512bw", "avx512dq", "avx512vl")) else ["libkache_ref_avx2.so"]
    for n in names:
        p = os.path.join(REF_DIR, n)
        if os.path.exists(p):
            return p
    return None
