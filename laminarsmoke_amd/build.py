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

# (source, extra defines): lsm_kernels.hip is compiled as eight units (LSM_PART 1-8): the team
# kernel's register-array sizes (1-5), the rest (6) and the packed kernels' segment widths
# (7, 8) apart, so that they build in parallel.
SOURCES = [("lsm_kernels.hip", [f"-DLSM_PART={p}"]) for p in range(1, 9)] + [("lsm_api.cpp", []), ("lsm_serve.cpp", [])]
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
    """Compile every source (always; the build check must compile) into `out`, the translation
    units in parallel."""
    objs, procs = [], []
    for src, defs in SOURCES:
        obj = os.path.join(os.path.dirname(out), os.path.basename(out) + "." + src + "".join(defs) + ".o")
        cmd = [hipcc(), *HIPCC_FLAGS, *defs, *extra, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((subprocess.Popen(cmd), cmd))
        objs.append(obj)
    failed = [cmd for p, cmd in procs if p.wait() != 0]
    if failed:
        raise subprocess.CalledProcessError(1, failed[0])
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
