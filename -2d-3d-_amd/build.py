"""Build the gfx950 HIP library librblbm.so (the C ABI of include/lbm_rb.h) in-tree.

Usage: python ./-2d-3d-_amd/build.py   (or __graft_entry__.build()).  Always recompiles.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "librblbm.so")
# FP64 kernels and host code: no contraction, so FP64 results round exactly like the
# reference's CPU build.
SOURCES = {"lbm_kernels.hip": ["-ffp-contract=off"], "lbm_api.hip": ["-ffp-contract=off"],
           # reduced-precision forms: FP32 arithmetic with fused multiply-add inside each source
           # expression only, so that every kernel instance (layout, kernel form, tile) fuses the
           # same products and the forms give bitwise the same results
           "lbm_kernels_lp.hip": ["-ffp-contract=on"]}
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("RBLBM_ARCH", "gfx950")

# Packed FP32 instructions are switched off for every HIP translation unit.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-function", "-I", os.path.join(ROOT, "include")]


PROBE = os.path.join(LIBDIR, "libhbmprobe.so")


def build_probe(verbose: bool = False) -> str:
    """The bench's streaming copy / read kernels (hbm_probe.hip), a library of their own."""
    os.makedirs(LIBDIR, exist_ok=True)
    cmd = [HIPCC, *HIP_FLAGS, "-shared", os.path.join(CSRC, "hbm_probe.hip"), "-o", PROBE + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(PROBE + ".tmp", PROBE)
    return PROBE


def build(verbose: bool = False) -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    objs, procs = [], []
    for src, extra in SOURCES.items():      # the translation units compile in parallel
        obj = os.path.join(LIBDIR, os.path.splitext(src)[0] + ".o")
        cmd = [HIPCC, *HIP_FLAGS, *extra, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd)))
        objs.append(obj)
    for cmd, p in procs:
        if p.wait() != 0:
            raise subprocess.CalledProcessError(p.returncode, cmd)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp, *objs]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    build_probe(verbose)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
