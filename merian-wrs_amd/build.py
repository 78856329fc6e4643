#!/usr/bin/env python3
"""Builds merian-wrs_amd/libwrs_amd.so for gfx950 with hipcc (in-tree).

Each HIP/C++ translation unit under csrc/ is compiled to an object in
_build/ (in parallel), then linked into one shared library exporting the C ABI
of include/wrs.h and include/wrs_runtime.h.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libwrs_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("WRS_OFFLOAD_ARCH", "gfx950")

SOURCES = ["reduce.hip", "partition.hip", "scan.hip", "splitpack.hip", "mergepack.hip", "streampack.hip", "sample.hip",
           "stats.hip", "batched.cpp", "wrs_host.cpp", "tune.cpp"]

# -ffp-contract=off: no FMA contraction, so split/pack arithmetic is the
# sequence of IEEE operations written (as in the CPU oracle).
# -packed-fp32-ops disabled: packed FP32 forms are avoided on this hardware.
FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-function", "-I", os.path.join(HERE, "..", "include")]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    warn = "\n".join(l for l in r.stderr.splitlines()
                     if "packed-fp32-ops" not in l and "not a recognized feature" not in l
                     and l.strip())
    if warn:
        print(f"[{src}] {warn}", file=sys.stderr)
    return obj


def build(jobs: int = 8) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(jobs, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", f"--offload-arch={ARCH}", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build())
