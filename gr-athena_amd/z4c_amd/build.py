"""Build recipe of the gfx950 shared library (in-tree, no JIT cache).

``python -m z4c_amd.build`` (with gr-athena_amd/ on sys.path) or
``__graft_entry__.build()`` compiles csrc/z4c_kernels.hip into
z4c_amd/libz4c_mi355x.so with hipcc for gfx950 only.
"""
from __future__ import annotations

import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(PKG))
CSRC = os.path.join(os.path.dirname(PKG), "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(PKG, "libz4c_mi355x.so")
SOURCES = [os.path.join(CSRC, "z4c_kernels.hip")]
HEADERS = [os.path.join(CSRC, f) for f in ("z4c_device.hpp", "z4c_rhs_point.hpp", "z4c_stencil.hpp", "z4c_adjoint.hpp",
                                           "z4c_refine.hpp")] + [
    os.path.join(INCLUDE, "z4c_mi355x.h")]

# -ffp-contract=off: no fused multiply-add, so every product and sum rounds
# as in the reference build (configure.py:1145) and the CPU oracle.
# -packed-fp32-ops off: no packed FP32 instructions anywhere in the library.
HIPCC_FLAGS = [
    "-O3", "--offload-arch=gfx950", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + INCLUDE, "-I" + CSRC,
]


def hipcc() -> str:
    for c in ("/opt/rocm/bin/hipcc", "hipcc"):
        if os.path.isabs(c) and os.path.exists(c):
            return c
    return "hipcc"


def build(verbose: bool = False) -> str:
    """Compile the library (always) and return its path."""
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + ["-o", tmp] + SOURCES
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    return LIB


HOST = os.path.join(os.path.dirname(PKG), "host")
HOST_DEMO = os.path.join(HOST, "z4c_host_demo")


def build_host(verbose: bool = False) -> str:
    """Compile the C++ host driver (host/z4c_host_demo.cpp over
    host/z4c_host.hpp) against include/z4c_mi355x.h, linked to the library
    with an rpath to its directory; no device code of its own."""
    tmp = HOST_DEMO + ".tmp"
    cmd = [hipcc(), "-O2", "-std=c++17", "-I" + INCLUDE, "-I" + HOST, "-o", tmp,
           os.path.join(HOST, "z4c_host_demo.cpp"), "-L" + PKG, "-lz4c_mi355x",
           "-Wl,-rpath,$ORIGIN/../z4c_amd"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("host build failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, HOST_DEMO)
    return HOST_DEMO


def is_stale() -> bool:
    if not os.path.exists(LIB):
        return True
    t = os.path.getmtime(LIB)
    return any(os.path.getmtime(f) > t for f in SOURCES + HEADERS)


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_host(verbose=True))
    sys.exit(0)
