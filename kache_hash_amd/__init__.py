"""MI355X-native streaming k-mer table (kache-hash's Streaming_Kmer_Hash_Table
path and C ABI, k = 31, l = 22).

The product is libkache_amd.so, built in-tree from csrc/ by build.py; this
package loads it and exposes its C ABI through `capi.Table`.  There is no
fallback: without the built library (or without a HIP device, for the calls
that need one) the calls fail loudly.
"""
import os

from . import capi
from .capi import (ABSENT, ADD, ADD_WRAP, ASSIGN, KEEP, MAX, MIN, NOT_KMER, OR, PRESENT, REDUCE_KEYS,
                   REDUCE_POPCOUNT, REDUCE_SUM, Batch, KacheError, Table)

_api = None


def library_path():
    """The built library; KACHE_AMD_LIB names another build of it (A/B runs)."""
    return os.environ.get("KACHE_AMD_LIB") or capi.library_path()


def api():
    """The loaded device library; raises if it has not been built."""
    global _api
    if _api is None:
        path = library_path()
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: build it with `python kache_hash_amd/build.py` "
                               "or __graft_entry__.build()")
        _api = capi.Api(path)
    return _api


def table(width=0, expected_keys=1 << 22, load_factor=0.8, pow2=False, workers=16):
    """A new device table of 31-mers: a set (width 0) or a map to `width`-byte values."""
    return Table(api(), width, expected_keys, load_factor, pow2, workers)


__all__ = ["api", "table", "library_path", "Table", "Batch", "KacheError", "KEEP", "ASSIGN", "ADD", "OR", "MIN", "MAX",
           "ADD_WRAP", "REDUCE_KEYS", "REDUCE_POPCOUNT", "REDUCE_SUM", "ABSENT", "PRESENT", "NOT_KMER"]
