"""Build libcorrct_hip.so (HIP, gfx950) in-tree.

``python -m pycorrectedemissionct_amd.build`` compiles every source under
``csrc/`` with hipcc and links ``_lib/libcorrct_hip.so`` next to this file.
The library never links torch: the Python layer hands it raw device
pointers and the HIP stream torch is using.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "_lib")
LIBNAME = "libcorrct_hip.so"
ARCH = os.environ.get("CORRCT_HIP_ARCH", "gfx950")

SOURCES = ["plan.cpp", "projector.hip", "fp_stream.hip", "bp_adjoint.hip", "attenuation.hip", "solver_ops.hip", "batch.hip", "unet.hip", "grad_reg.hip", "bp_skimage_tile.hip", "fp_band.hip", "fp_tile.hip", "noalloc.hip"]

COMMON_FLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    f"--offload-arch={ARCH}",
    # float32 expressions keep the reference's rounding sequence; FMAs are
    # written explicitly (fmaf) where a kernel wants them
    "-ffp-contract=off",
    # no packed-FP32 instruction selection
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
]


def hipcc() -> str:
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found: the HIP toolchain (ROCm) is required to build libcorrct_hip")
    return path


def library_path() -> str:
    return os.path.join(LIBDIR, LIBNAME)


def build(verbose: bool = False, jobs: int = 4) -> str:
    """Compile all sources and link the shared library; returns its path."""
    os.makedirs(LIBDIR, exist_ok=True)
    objdir = os.path.join(LIBDIR, "obj")
    os.makedirs(objdir, exist_ok=True)
    cc = hipcc()

    def compile_one(src: str) -> str:
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        cmd = [cc, *COMMON_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        if src.endswith(".cpp"):
            cmd = [cc, *COMMON_FLAGS, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"hipcc failed for {src}:\n{res.stdout}\n{res.stderr}")
        if verbose and res.stderr.strip():
            print(res.stderr, file=sys.stderr)
        return obj

    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    out = library_path()
    tmp = out + ".tmp"
    cmd = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, out)
    if verbose:
        print("built", out)
    return out


if __name__ == "__main__":
    build(verbose=True)
