"""Build recipe of the MI355X k-mer table library (libkache_amd.so).

Compiles the HIP sources under csrc/ for gfx950 with hipcc into an in-tree
shared library exporting the C ABI of include/kache_capi.h and the extensions
of include/kache_amd.h.  Packed FP32 instructions are disabled for the device
code (the code is integer-only; the flag keeps any float code off them).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libkache_amd.so")
SOURCES = [os.path.join(HERE, "csrc", f) for f in ("kache_table.hip", "kache_synth.hip", "kache_serve.cpp", "kache_fastx.cpp")]
HEADERS = [os.path.join(HERE, "csrc", f) for f in ("kache_device.hpp", "kache_kernels.hpp")] + [
    os.path.join(ROOT, "include", f) for f in ("kache_capi.h", "kache_amd.h")
]
ARCH = os.environ.get("KACHE_AMD_ARCH", "gfx950")


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.isabs(c) and os.path.exists(c) or not os.path.isabs(c)):
            return c
    return "hipcc"


def command(out=LIB, extra=()):
    return [
        hipcc(), "-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC", "-shared",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
        "-I", os.path.join(ROOT, "include"), "-I", os.path.join(HERE, "csrc"),
        *extra, *SOURCES, "-o", out,
    ]


ROOF = os.path.join(HERE, "memory_roof")
CPP_CHECK_SRC = os.path.join(ROOT, "tests", "cpp", "cpp_api_check.cpp")
CPP_CHECK = os.path.join(ROOT, "tests", "cpp", "cpp_api_check")


def build(verbose=True):
    """Compiles the library and the memory-roof probe (always); returns the
    library's path."""
    tmp = LIB + ".tmp"
    cmd = command(tmp)
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=ROOT)
    os.replace(tmp, LIB)
    roof = [hipcc(), "-O3", "-std=c++17", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
            "-packed-fp32-ops", os.path.join(HERE, "csrc", "memory_roof.hip"), "-o", ROOF + ".tmp"]
    if verbose:
        print(" ".join(roof), flush=True)
    subprocess.run(roof, check=True, cwd=ROOT)
    os.replace(ROOF + ".tmp", ROOF)
    # The C++ API's check program (host code over the library).
    cpp = ["g++", "-std=c++20", "-O2", "-Wall", "-I", os.path.join(ROOT, "include"), CPP_CHECK_SRC,
           "-L", HERE, "-lkache_amd", "-Wl,-rpath,$ORIGIN/../../kache_hash_amd", "-o", CPP_CHECK + ".tmp"]
    if verbose:
        print(" ".join(cpp), flush=True)
    subprocess.run(cpp, check=True, cwd=ROOT)
    os.replace(CPP_CHECK + ".tmp", CPP_CHECK)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
