"""Builds libnss_hip.so (gfx950) in-tree with hipcc.

Every HIP translation unit is compiled with packed-FP32 instructions disabled
and without fast-math (IEEE division and compares in the threshold code).
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libnss_hip.so")
ARCH = os.environ.get("NSS_HIP_ARCH", "gfx950")

SOURCES = ["bm3d_kernels.hip", "driver.cpp", "bm3d_args.cpp", "server.cpp", "wnnm.cpp", "noalloc.hip", "device_mem.cpp"]

HIPFLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # MFMA results in VGPRs (gfx950's unified register file) rather than
    # AGPRs: no v_accvgpr_read before the epilogues, fewer registers in all.
    "-mllvm", "-amdgpu-mfma-vgpr-form",
    "-Wall",
]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build(jobs: int = 4) -> str:
    """Compile every source (always; this is the build check) and link."""
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    objs = []
    procs = []
    for src in SOURCES:
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        lang = ["-x", "hip"] if src.endswith((".hip",)) else []
        cmd = [hipcc, *HIPFLAGS, *lang, "-c", os.path.join(CSRC, src), "-o", obj]
        print(" ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
        if len(procs) >= jobs:
            if procs.pop(0).wait() != 0:
                raise subprocess.CalledProcessError(1, "hipcc")
    for p in procs:
        if p.wait() != 0:
            raise subprocess.CalledProcessError(1, "hipcc")
    tmp = LIB + ".tmp"
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
