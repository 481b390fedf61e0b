"""Build of the in-tree HIP library liblsm_hip.so for gfx950 (hipcc, no JIT cache).

``python -m laminarsmoke_amd.build`` compiles the kernels and the C-ABI host code
into ``laminarsmoke_amd/liblsm_hip.so``.
"""
from __future__ import annotations

import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "liblsm_hip.so")
ARCH = os.environ.get("LSM_OFFLOAD_ARCH", "gfx950")

SOURCES = ["lsm_kernels.hip", "lsm_api.cpp", "lsm_serve.cpp"]
HEADERS = ["lsm_bdf.hpp", "lsm_internal.hpp"]

# Packed-FP32 instructions are switched off for every HIP translation unit.
HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-parameter",
    "-I", os.path.join(ROOT, "include"),
]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.sep not in c or os.path.exists(c)):
            return c
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False, extra=(), out: str = LIB) -> str:
    """Compile every source (always; the build check must compile) into `out`."""
    objs = []
    for src in SOURCES:
        obj = os.path.join(os.path.dirname(out), os.path.basename(out) + "." + src + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else ["-x", "hip"]
        cmd = [hipcc(), *HIPCC_FLAGS, *extra, *lang, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.check_call(cmd)
        objs.append(obj)
    tmp = out + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, out)
    for o in objs:
        os.remove(o)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
