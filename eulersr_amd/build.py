"""Build libeulersr.so in-tree for gfx950.

Host sources (mesh preparation, orchestration, C ABI) are compiled by g++
without floating-point contraction, so host-side float32 precomputations
round exactly like the reference's arrays; device sources by hipcc for
gfx950 with packed-FP32 instructions disabled.

Usage: python -m eulersr_amd.build [-v]
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libeulersr.so"
ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))
ARCH = os.environ.get("ESR_OFFLOAD_ARCH", "gfx950")

HOST_SRC = ["esr_mesh.cpp", "esr_api.cpp", "esr_mlp_api.cpp"]
DEV_SRC = ["esr_kernels.hip", "esr_mlp.hip", "esr_noalloc.hip"]
HEADERS = ["esr_internal.h", "esr_launch.h", "esr_mlp.h"]

HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-Wall",
              f"-I{ROCM}/include", "-D__HIP_PLATFORM_AMD__"]
DEV_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-function"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(str(c) for c in cmd), flush=True)
    r = subprocess.run([str(c) for c in cmd], capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {' '.join(str(c) for c in cmd[:3])} ...")
    elif verbose and (r.stdout.strip() or r.stderr.strip()):
        print(r.stdout + r.stderr)


def build(verbose: bool = False) -> Path:
    hipcc = shutil.which("hipcc") or str(ROCM / "bin" / "hipcc")
    gxx = shutil.which("g++") or "g++"
    BUILD.mkdir(exist_ok=True)
    objs = []
    for src in HOST_SRC:
        obj = BUILD / (Path(src).stem + ".o")
        _run([gxx, *HOST_FLAGS, "-c", CSRC / src, "-o", obj], verbose)
        objs.append(obj)
    for src in DEV_SRC:
        obj = BUILD / (Path(src).stem + ".o")
        _run([hipcc, *DEV_FLAGS, "-c", CSRC / src, "-o", obj], verbose)
        objs.append(obj)
    tmp = LIB.with_suffix(".so.tmp")
    _run([hipcc, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp], verbose)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
