"""Builds tests/asan/capi_stress: the library's sources and the C ABI driver
with AddressSanitizer on the host code only (the device code of every .hip
file is built as in the product), linked into one executable.

    python tests/asan/build_asan.py

Run it with ASAN_OPTIONS=detect_leaks=0 (the HIP runtime keeps allocations
until exit); any heap error aborts it with a report."""
from __future__ import annotations

import importlib.util
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
PKG = os.path.join(ROOT, "vapoursynth-nssfactory_amd")
EXE = os.path.join(HERE, "capi_stress")

_spec = importlib.util.spec_from_file_location("nss_build", os.path.join(PKG, "build.py"))
_pkg_build = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(_pkg_build)

# host-side sanitizer flags (each -fsanitize after -Xarch_host: the device
# passes of hipcc take none)
HOST_ASAN = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fno-omit-frame-pointer"]
LINK_ASAN = ["-fsanitize=address"]


def build(jobs: int = 4) -> str:
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    flags = [f for f in _pkg_build.HIPFLAGS if f != "-O3"] + ["-O1", "-g"]
    units = [(os.path.join(PKG, "csrc", s), s.endswith(".hip")) for s in _pkg_build.SOURCES]
    units.append((os.path.join(HERE, "capi_stress.cpp"), False))
    objs, procs = [], []
    for src, is_hip in units:
        obj = os.path.join(objdir, os.path.splitext(os.path.basename(src))[0] + ".o")
        objs.append(obj)
        cmd = [hipcc, *flags, *HOST_ASAN, *(["-x", "hip"] if is_hip else []), "-c", src, "-o", obj]
        print(" ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
        if len(procs) >= jobs and procs.pop(0).wait() != 0:
            raise subprocess.CalledProcessError(1, "hipcc")
    for p in procs:
        if p.wait() != 0:
            raise subprocess.CalledProcessError(1, "hipcc")
    clang = "/opt/rocm/llvm/bin/clang++"
    link = [clang, *LINK_ASAN, "-o", EXE + ".tmp", *objs, "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib",
            "-lpthread"]
    print(" ".join(link), flush=True)
    subprocess.check_call(link)
    os.replace(EXE + ".tmp", EXE)
    return EXE


if __name__ == "__main__":
    build()
    sys.exit(0)
