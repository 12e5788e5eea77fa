"""Writes the golden fixtures of tests/golden/ from the reference's own code.

1. describer.npz / scalars.npz / placement.npz: runs the fixture writer of
   gen_golden.cpp, compiled against the reference headers in both ISA
   variants (`make -C oracle golden`), and packs its records.
2. ref_tables.json: runs seeded scenarios through the reference's C ABI
   (oracle/_ref/libkache_ref_*.so, `make -C oracle ref`) and records each
   call's return value, the final capacity and size, kache_reduce over two
   masks and the three reductions, kache_get of each query's first k-mer, and
   SHA-256 digests of the sorted (key, value) dump and of the query outputs.

Run from the repository root in the build container (it needs the reference
sources): python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

REC = np.dtype([("q", "<u4"), ("i", "<u4"), ("key", "<u8"), ("h", "<u8"), ("c", "u1"), ("o", "u1"), ("p1", "<u2"), ("p2", "<u4")])


def run_writers():
    with tempfile.TemporaryDirectory() as d:
        for tag in ("avx2", "avx512"):
            subprocess.run([os.path.join(ROOT, "oracle", "_ref", f"gen_golden_{tag}"), d, tag], check=True)
        raw = open(os.path.join(d, "seqs.bin"), "rb").read()
        n = int(np.frombuffer(raw[:8], np.uint64)[0])
        lens = np.frombuffer(raw[8:8 + 8 * n], np.uint64).copy()
        text = np.frombuffer(raw[8 + 8 * n:], np.uint8).copy()
        r2 = np.frombuffer(open(os.path.join(d, "describe_avx2.bin"), "rb").read(), REC)
        r5 = np.frombuffer(open(os.path.join(d, "describe_avx512.bin"), "rb").read(), REC)
        assert (r2[["q", "i", "key", "h", "o"]] == r5[["q", "i", "key", "h", "o"]]).all()
        np.savez_compressed(os.path.join(HERE, "describer.npz"), lens=lens, text=text, q=r2["q"], i=r2["i"],
                            key=r2["key"], h=r2["h"], fwd=r2["o"], c_mul=r2["c"], c_wang=r5["c"])
        s2 = np.frombuffer(open(os.path.join(d, "scalars_avx2.bin"), "rb").read()[8:], np.uint64).reshape(-1, 8)
        s5 = np.frombuffer(open(os.path.join(d, "scalars_avx512.bin"), "rb").read()[8:], np.uint64).reshape(-1, 8)
        np.savez_compressed(os.path.join(HERE, "scalars.npz"), x=s2[:, 0], wang=s2[:, 1], c_mul=s2[:, 2].astype(np.uint8),
                            c_wang=s5[:, 2].astype(np.uint8), canon=s2[:, 3], minimizer=s2[:, 4], xxh3_s0=s2[:, 5],
                            xxh3_s63=s2[:, 6], xxh3=s2[:, 7])
        p2 = open(os.path.join(d, "placement_avx2.bin"), "rb").read()
        assert p2 == open(os.path.join(d, "placement_avx512.bin"), "rb").read()
        np.savez_compressed(os.path.join(HERE, "placement.npz"), **parse_placement(np.frombuffer(p2, np.uint64)))


def parse_placement(w):
    """Unpacks placement_<tag>.bin (layout in gen_golden.cpp's header)."""
    i = 0

    def take(n):
        nonlocal i
        a = w[i:i + n].copy()
        i += n
        return a

    n_cap = int(take(1)[0])
    caps = take(n_cap)
    n_h = int(take(1)[0])
    rows = take(n_h * (2 + 2 * n_cap)).reshape(n_h, 2 + 2 * n_cap)
    out = dict(caps=caps, h=rows[:, 0], filter_pos=rows[:, 1], bucket_of=rows[:, 2::2], pick_bucket=rows[:, 3::2])
    tab_n, tab_buckets, tab_ov, sec = [], [], [], []
    for _ in range(int(take(1)[0])):
        n, buckets, ov = take(3)
        tab_n.append(n), tab_buckets.append(buckets), tab_ov.append(ov)
        sec.append(take(3 * int(take(1)[0])).reshape(-1, 3))
    out.update(tab_n=np.array(tab_n, np.uint64), tab_buckets=np.array(tab_buckets, np.uint64),
               tab_ov_slots=np.array(tab_ov, np.uint64), sec=np.stack(sec))
    ov = take(2 * int(take(1)[0])).reshape(-1, 2)
    out.update(ov_n=ov[:, 0], ov_slots=ov[:, 1])
    assert i == len(w)
    return out


def digest(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def ref_tables():
    import oracle as O
    import scenarios
    from kache_hash_amd import capi
    path = O.ref_library()
    assert path, "build the reference first: make -C oracle ref"
    api = capi.Api(path)
    out = {"library": os.path.basename(path), "scenarios": []}
    for sc in scenarios.SCENARIOS:
        r = scenarios.run(capi.Table, api, sc)
        out["scenarios"].append({"name": sc["name"], "inserted": r["inserted"], "capacity": r["capacity"], "size": r["size"],
                                 "dump": digest(r["keys"], r["values"]), "lookup": digest(*r["lookup"]),
                                 "counts": digest(r["counts"]), "present": r.get("present"),
                                 "reduce": r["reduce"], "get": digest(*r["get"]),
                                 "dump_after_update": digest(r["keys2"], r["values2"]) if "keys2" in r else None})
        print(sc["name"], "ok", flush=True)
    with open(os.path.join(HERE, "ref_tables.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    run_writers()
    ref_tables()
