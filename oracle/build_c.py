"""Build the oracle's C restatement (``oracle/c``) into ``oracle/_ref/liboracle_c.so``.

Test infrastructure only. corrct itself is pure Python (no compiled sources
to build), so ``oracle/_ref`` holds only this C restatement of its
skimage-backend path, used as the timed CPU baseline and cross-checked
against the NumPy oracle by the tests.
"""

from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "c", "corrct_oracle.c")
OUT_DIR = os.path.join(HERE, "_ref")
LIB = os.path.join(OUT_DIR, "liboracle_c.so")


def build(verbose: bool = False) -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    cmd = ["gcc", "-O2", "-fopenmp", "-ffp-contract=off", "-fno-fast-math", "-shared", "-fPIC", SRC, "-o", LIB, "-lm"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"oracle C build failed:\n{res.stderr}")
    if verbose:
        print("built", LIB)
    return LIB


def load():
    import ctypes

    if not os.path.exists(LIB):
        build()
    lib = ctypes.CDLL(LIB)
    vp = ctypes.c_void_p
    lib.oracle_fp.argtypes = [vp, vp, vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp]
    lib.oracle_fp.restype = None
    lib.oracle_bp_skimage.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp]
    lib.oracle_bp_skimage.restype = None
    lib.oracle_bp_adjoint.argtypes = [vp, vp, vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp]
    lib.oracle_bp_adjoint.restype = None
    lib.oracle_num_threads.restype = ctypes.c_int
    return lib


if __name__ == "__main__":
    build(verbose=True)
