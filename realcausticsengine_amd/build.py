"""Builds libRealCaustics.so (gfx950) in-tree with hipcc.

Every HIP translation unit is compiled with -ffp-contract=off so that no
multiply-add is fused: the kernels then reproduce, bit for bit, the float
arithmetic order of the reference and of the CPU oracle.  Packed FP32
instructions are disabled (-packed-fp32-ops).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libRealCaustics.so")
BUILD_DIR = os.path.join(HERE, "_build")

SOURCES = ["rc_kernels.hip", "rc_guide.hip", "rc_bvh.hip", "rc_ingest.hip", "rc_engine.cpp", "rc_guide_host.cpp"]
HEADERS = ["rc_device.h", "rc_engine.h", "rc_guide.h", "rc_guide_model.h", "rc_host.h", "rc_vec.h"]

COMMON = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wall", "-Wno-unused-command-line-argument", "-Wno-main",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build libRealCaustics.so)")


def build(verbose: bool = False) -> str:
    """Compile every source (always) and link libRealCaustics.so; returns its path."""
    os.makedirs(BUILD_DIR, exist_ok=True)
    cc = hipcc()
    objs, procs = [], []
    for src in SOURCES:  # the translation units compile in parallel
        obj = os.path.join(BUILD_DIR, os.path.splitext(src)[0] + ".o")
        cmd = [cc, *COMMON, "-I", os.path.join(HERE, "..", "include"), "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd)))
        objs.append(obj)
    failed = [src for src, p in procs if p.wait() != 0]
    if failed:
        raise RuntimeError("hipcc failed on " + ", ".join(failed))
    tmp = LIB + ".tmp"
    cmd = [cc, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-lz", "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
