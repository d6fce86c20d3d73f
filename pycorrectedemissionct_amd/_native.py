"""ctypes binding of libcorrct_hip.so (the C ABI declared in include/corrct_hip.h).

torch is imported first so the library binds to the HIP runtime torch has
already loaded (one runtime per process). The library is never optional:
if it is missing or fails to load, every projector/solver call raises.
"""

from __future__ import annotations

import ctypes
import os
import threading

import numpy as np
import torch  # noqa: F401  (must be loaded before the HIP library)

from . import build as _build

c_float_p = ctypes.POINTER(ctypes.c_float)
c_double_p = ctypes.POINTER(ctypes.c_double)
c_void_p = ctypes.c_void_p

CHIP_BP_SKIMAGE = 0
CHIP_BP_SKIMAGE_SYM = 1
CHIP_BP_ADJOINT = 2

# enum chip_row_op
CHIP_ROW_PDHG_L2 = 0
CHIP_ROW_PDHG_KL = 1
CHIP_ROW_SIRT = 2
CHIP_ROW_MLEM = 3

_lock = threading.Lock()
_lib = None

# name -> argtypes (all return int status unless listed in _RESTYPES)
_SIGNATURES = {
    "chip_version": [],
    "chip_last_error": [],
    "chip_num_symbols": [],
    "chip_symbol_name": [ctypes.c_int],
    "chip_plan_create": [
        ctypes.POINTER(c_void_p), ctypes.c_int, ctypes.c_int, c_void_p, ctypes.c_int, c_void_p, c_void_p, c_void_p,
    ],
    "chip_plan_destroy": [c_void_p],
    "chip_plan_set_subangles": [c_void_p, ctypes.c_int, c_void_p, c_void_p, c_void_p, c_void_p],
    "chip_plan_set_radon_trig": [c_void_p, c_void_p],
    "chip_batch_set_radon_trig": [c_void_p, c_void_p],
    "chip_grad_dual": [c_void_p, c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_int, c_void_p, c_void_p,
                       ctypes.c_int, ctypes.c_int, ctypes.c_float, c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_int,
                       c_void_p, c_void_p],
    "chip_grad_update_primal": [c_void_p, c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_int, c_void_p, c_void_p,
                                ctypes.c_int, ctypes.c_float, c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_int,
                                c_void_p, ctypes.c_int64, c_void_p],
    "chip_l1_dual": [c_void_p, c_void_p, ctypes.c_int64, ctypes.c_float, c_void_p, ctypes.c_int64, c_void_p],
    "chip_l1_update_primal": [c_void_p, c_void_p, ctypes.c_int64, c_void_p, ctypes.c_int64, ctypes.c_int, c_void_p],
    "chip_axpy": [c_void_p, ctypes.c_float, c_void_p, ctypes.c_int64, ctypes.c_int64, c_void_p],
    "chip_rdiv": [c_void_p, ctypes.c_float, ctypes.c_int64, c_void_p],
    "chip_att_maps": [c_void_p, c_void_p, c_void_p, c_void_p, ctypes.c_int, c_void_p],
    "chip_bin_images": [c_void_p, c_void_p, ctypes.c_int64, ctypes.c_int, c_void_p, ctypes.c_int, c_void_p],
    "chip_batch_otf_fp_row_bp": [c_void_p, c_void_p, c_void_p, c_void_p, ctypes.c_int, c_void_p, c_void_p, c_void_p,
                                 c_void_p, c_void_p, ctypes.c_int, c_void_p],
    "chip_otf_fp_row_bp": [c_void_p, c_void_p, c_void_p, c_void_p, ctypes.c_int, ctypes.c_int, c_void_p, c_void_p,
                           c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, ctypes.c_int, c_void_p],
    "chip_fp": [c_void_p, c_void_p, c_void_p, ctypes.c_int64, c_void_p, ctypes.c_int, c_void_p],
    "chip_bp": [c_void_p, c_void_p, c_void_p, ctypes.c_int64, c_void_p, ctypes.c_int, ctypes.c_int, c_void_p],
    "chip_bp_noalloc": [c_void_p, c_void_p, c_void_p, ctypes.c_int64, c_void_p, ctypes.c_int, ctypes.c_int, c_void_p],
    "chip_fp_angle": [c_void_p, c_void_p, c_void_p, c_void_p, ctypes.c_int, c_void_p],
    "chip_bp_angle": [c_void_p, c_void_p, c_void_p, c_void_p, ctypes.c_int, ctypes.c_int, c_void_p],
    "chip_absmax": [c_void_p, ctypes.c_int64, c_void_p, c_void_p],
    "chip_precond_fix": [c_void_p, ctypes.c_int64, ctypes.c_float, c_void_p, ctypes.c_float, c_void_p],
    "chip_pdhg_dual_l2": [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, ctypes.c_int64, c_void_p],
    "chip_pdhg_dual_kl": [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, ctypes.c_int64, c_void_p],
    "chip_pdhg_dual_init": [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, ctypes.c_int64, ctypes.c_int, c_void_p],
    "chip_tv_dual": [c_void_p, c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_float, c_void_p],
    "chip_pdhg_primal": [
        c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_int, ctypes.c_float,
        c_void_p, ctypes.c_int, ctypes.c_int, c_void_p,
    ],
    "chip_sirt_residual": [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, ctypes.c_int64, c_void_p],
    "chip_pdhg_dual_norm": [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, ctypes.c_int64, ctypes.c_int,
                            c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, c_void_p],
    "chip_data_residual": [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, ctypes.c_int64,
                           ctypes.c_int, c_void_p, ctypes.c_int64, c_void_p],
    "chip_sirt_update": [
        c_void_p, c_void_p, c_void_p, c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_int, ctypes.c_float, c_void_p,
        ctypes.c_int, ctypes.c_int, c_void_p,
    ],
    "chip_filter_lines": [c_void_p, c_void_p, c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_int, c_void_p],
    "chip_psf_lines": [c_void_p, c_void_p, c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_int, c_void_p],
    "chip_fista_update": [
        c_void_p, c_void_p, c_void_p, c_void_p, ctypes.c_float, ctypes.c_float, c_void_p, ctypes.c_int64, c_void_p,
        ctypes.c_int64, ctypes.c_int, ctypes.c_float, ctypes.c_int, ctypes.c_float, c_void_p, ctypes.c_float,
        ctypes.c_int64, c_void_p,
    ],
    "chip_mlem_ratio": [c_void_p, c_void_p, c_void_p, c_void_p, ctypes.c_int64, c_void_p],
    "chip_clip_relax": [c_void_p, c_void_p, c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_int, ctypes.c_float,
                        c_void_p, c_void_p, c_void_p, ctypes.c_float, ctypes.c_int64, c_void_p],
    "chip_mlem_update": [
        c_void_p, c_void_p, c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_int, ctypes.c_float, c_void_p,
        ctypes.c_int64, c_void_p,
    ],
    "chip_sumsq": [c_void_p, ctypes.c_int64, c_void_p, c_void_p],
    "chip_plan_tune": [c_void_p, ctypes.c_char_p, ctypes.c_int],
    "chip_sumabs": [c_void_p, ctypes.c_int64, c_void_p, c_void_p],
    "chip_fp_otf": [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, ctypes.c_int, c_void_p],
    "chip_bp_otf": [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, ctypes.c_int, ctypes.c_int, c_void_p],
    "chip_batch_plan_create": [
        ctypes.POINTER(c_void_p), ctypes.c_int, c_void_p, ctypes.c_int, c_void_p, ctypes.c_int, c_void_p, c_void_p,
        c_void_p,
    ],
    "chip_batch_plan_destroy": [c_void_p],
    "chip_batch_info": [c_void_p, c_void_p, c_void_p, c_void_p],
    "chip_batch_tune": [c_void_p, ctypes.c_char_p, ctypes.c_int],
    "chip_batch_att_maps": [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p],
    "chip_batch_fp": [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p],
    "chip_batch_bp": [c_void_p, c_void_p, c_void_p, c_void_p, ctypes.c_int, c_void_p],
    "chip_batch_fp_otf": [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p],
    "chip_batch_bp_otf": [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, ctypes.c_int, c_void_p],
    "chip_batch_tv_dual": [c_void_p, c_void_p, c_void_p, ctypes.c_float, c_void_p],
    "chip_batch_pdhg_primal": [
        c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_int,
        ctypes.c_float, c_void_p, c_void_p,
    ],
    "chip_nn_conv3x3": [
        c_void_p, ctypes.c_int, c_void_p, ctypes.c_int, c_void_p, c_void_p, c_void_p, ctypes.c_int, c_void_p,
        ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, c_void_p,
    ],
    "chip_nn_conv3x3_wgrad": [
        c_void_p, c_void_p, ctypes.c_int, c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
        ctypes.c_int, c_void_p, ctypes.c_int64, c_void_p, c_void_p,
    ],
    "chip_nn_bias_grad": [c_void_p, ctypes.c_int64, ctypes.c_int, c_void_p, ctypes.c_int64, c_void_p, c_void_p],
    "chip_nn_flip_weights": [c_void_p, ctypes.c_int, ctypes.c_int, c_void_p, c_void_p],
    "chip_nn_relu_mask": [c_void_p, c_void_p, ctypes.c_int64, c_void_p],
    "chip_nn_avgpool2": [c_void_p, c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                         c_void_p],
    "chip_nn_upsample2": [c_void_p, c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                          c_void_p],
    "chip_nn_head_fwd": [c_void_p, ctypes.c_int, c_void_p, c_void_p, c_void_p, ctypes.c_int64, c_void_p],
    "chip_nn_head_bwd": [c_void_p, c_void_p, ctypes.c_int, c_void_p, c_void_p, c_void_p, ctypes.c_int64, c_void_p,
                         ctypes.c_int64, c_void_p],
    "chip_nn_mse": [c_void_p, c_void_p, ctypes.c_int64, c_void_p, c_void_p, ctypes.c_int, c_void_p, c_void_p],
    "chip_nn_adam": [c_void_p, c_void_p, c_void_p, c_void_p, ctypes.c_int64, ctypes.c_float, ctypes.c_float,
                     ctypes.c_float, ctypes.c_float, ctypes.c_int, c_void_p, c_void_p],
    "chip_nn_to_bf16": [c_void_p, c_void_p, ctypes.c_int64, c_void_p],
    "chip_nn_to_f32": [c_void_p, c_void_p, ctypes.c_int64, c_void_p],
    "chip_nn_add_bf16": [c_void_p, c_void_p, ctypes.c_int64, c_void_p],
    "chip_batch_sirt_update": [
        c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_int, ctypes.c_float,
        c_void_p, c_void_p,
    ],
}
_RESTYPES = {"chip_last_error": ctypes.c_char_p, "chip_symbol_name": ctypes.c_char_p}


class NativeError(RuntimeError):
    """A libcorrct_hip call returned a non-zero status."""


def library_path() -> str:
    return os.environ.get("CORRCT_HIP_LIB", _build.library_path())


def lib():
    """The loaded library (loads on first use; raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        path = library_path()
        if not os.path.exists(path):
            raise NativeError(
                f"libcorrct_hip.so not found at {path}: build it with `python -m pycorrectedemissionct_amd.build`"
            )
        handle = ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)
        for name, argtypes in _SIGNATURES.items():
            fn = getattr(handle, name)
            fn.argtypes = argtypes
            fn.restype = _RESTYPES.get(name, ctypes.c_int)
        _lib = handle
    return _lib


def check(status: int, what: str) -> None:
    if status != 0:
        msg = lib().chip_last_error().decode(errors="replace")
        raise NativeError(f"{what} failed (status {status}): {msg}")


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args), name)


def exported_symbols() -> list[str]:
    L = lib()
    return [L.chip_symbol_name(i).decode() for i in range(L.chip_num_symbols())]


def radon_trig(angles_rad_f32) -> np.ndarray:
    """[A, 2] float32 (cos, sin) as scikit-image's radon / iradon evaluate them: numpy's float32
    ``cos`` / ``sin`` of ``deg2rad(rad2deg(angle))`` (``_projector_backends.py:148, 299``;
    skimage ``radon_transform.py:104-112, 295-308``)."""
    ang = np.asarray(angles_rad_f32, dtype=np.float32).reshape(-1)
    th = np.deg2rad(np.rad2deg(ang))
    return np.ascontiguousarray(np.stack([np.cos(th), np.sin(th)], axis=1).astype(np.float32))


def ptr(t) -> int | None:
    """Device pointer of a contiguous float32 CUDA tensor (None -> NULL)."""
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("expected a device tensor")
    if not t.is_contiguous():
        raise ValueError("expected a contiguous tensor")
    return t.data_ptr()


def stream() -> int:
    """The HIP stream torch is currently using on this device."""
    return torch.cuda.current_stream().cuda_stream
