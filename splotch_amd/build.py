"""Build the gfx950 HIP library in-tree: splotch_amd/libsplotch_hip.so, and the
memory microbenchmarks the bench measures peaks with: splotch_amd/libmembench.so.

    python -m splotch_amd.build        (or __graft_entry__.build())

Flags: -ffp-contract=off (no FMA contraction, host and device: the parity
contract with the reference's float arithmetic), packed FP32 instructions
disabled, gfx950 only.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libsplotch_hip.so")
BENCH_LIB = os.path.join(HERE, "libmembench.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
SOURCES = ["scan_sort.hip", "kernels.hip", "splat8.hip", "adjoint.hip", "post.hip",
           "normalize.hip", "ingest.hip", "capi.hip"]
FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
         "-fno-fast-math", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-result"]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, " ".join(cmd), r.stderr))
    return obj


def _link(objs, lib):
    tmp = lib + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s\n%s" % (" ".join(cmd), r.stderr))
    os.replace(tmp, lib)


def build(verbose: bool = True) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(len(SOURCES) + 1) as ex:
        objs = list(ex.map(_compile, SOURCES + ["membench.hip"]))
    _link(objs[-1:], BENCH_LIB)
    objs = objs[:-1]
    _link(objs, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
