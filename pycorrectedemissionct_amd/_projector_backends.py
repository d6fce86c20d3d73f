"""Projector backends: the plugin boundary (corrct ``_projector_backends.py`` mirror).

* ``ProjectorBackend``     -- the ABC of ``_projector_backends.py:102-248``.
* ``ProjectorBackendHIP``  -- the MI355X backend. It reproduces the
  scikit-image backend's discretisation (``ProjectorBackendSKimage``,
  ``_projector_backends.py:251-359``) with hand-written HIP kernels reached
  through ``libcorrct_hip.so``; 3-D volumes are stacks of independent 2-D
  slices (parallel beam, rotation axis = Z).
* ``HipPlan``              -- owner of one ``chip_plan`` (geometry tables and
  attenuation scratch on the device).
* ``compute_attenuation``  -- ``_projector_backends.py:43-99`` on the GPU.
"""

from __future__ import annotations

import ctypes
import threading
from abc import ABC, abstractmethod

import numpy as np
import torch

from . import _arrays, _native
from .models import VolumeGeometry


class HipPlan:
    """A ``chip_plan``: N, float32 angles, detectors and weights (``chip_plan_create``)."""

    def __init__(self, n: int, angles_rad_f32, det_angles_rad=(np.pi / 2,), w_det=None, w_ang=None):
        self.n = int(n)
        # side of compute_attenuation's padded frame: N + 2 ceil((ceil(sqrt(2) N) - N) / 2)
        # (_projector_backends.py:75-77)
        self.n_pad = self.n + 2 * int(np.ceil((np.ceil(np.sqrt(2.0 * self.n * self.n)) - self.n) / 2.0))
        self.angles = np.ascontiguousarray(np.asarray(angles_rad_f32, dtype=np.float32).reshape(-1))
        self.det_angles = np.ascontiguousarray(np.asarray(det_angles_rad, dtype=np.float64).reshape(-1))
        self.n_angles = len(self.angles)
        self.n_det = len(self.det_angles)
        self.w_det = None if w_det is None else np.ascontiguousarray(
            np.broadcast_to(np.asarray(w_det, dtype=np.float64).reshape(-1), (self.n_det,))
        )
        self.w_ang = None if w_ang is None else np.ascontiguousarray(
            np.asarray(w_ang, dtype=np.float64).reshape(self.n_angles, self.n_det)
        )
        _arrays.device()  # fail early without a GPU
        handle = ctypes.c_void_p()
        _native.call(
            "chip_plan_create",
            ctypes.byref(handle),
            self.n,
            self.n_angles,
            self.angles.ctypes.data,
            self.n_det,
            self.det_angles.ctypes.data,
            None if self.w_det is None else self.w_det.ctypes.data,
            None if self.w_ang is None else self.w_ang.ctypes.data,
        )
        self._h = handle
        self._trig = _native.radon_trig(self.angles)
        _native.call("chip_plan_set_radon_trig", handle, self._trig.ctypes.data)

    @property
    def handle(self):
        return self._h

    def close(self):
        h = getattr(self, "_h", None)
        self._h = None
        if h is not None and h.value:
            try:
                _native.lib().chip_plan_destroy(h)
            except Exception:  # interpreter shutdown: module globals may already be gone
                pass

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def tune(self, key: str, value: int) -> None:
        """Kernel selection / tile parameter (``chip_plan_tune``)."""
        _native.call("chip_plan_tune", self._h, key.encode(), int(value))

    # ---- launches (device tensors in / out) -------------------------------
    def fp(self, vol: torch.Tensor, maps: torch.Tensor | None, out: torch.Tensor, batch: int, maps_stride: int = 0):
        _native.call("chip_fp", self._h, _native.ptr(vol), _native.ptr(maps), int(maps_stride), _native.ptr(out),
                     int(batch), _native.stream())
        return out

    def bp(self, sino: torch.Tensor, maps: torch.Tensor | None, out: torch.Tensor, batch: int, mode: int,
           maps_stride: int = 0):
        _native.call("chip_bp", self._h, _native.ptr(sino), _native.ptr(maps), int(maps_stride), _native.ptr(out),
                     int(batch), int(mode), _native.stream())
        return out

    def fp_angle(self, vol, maps, line, angle: int):
        _native.call("chip_fp_angle", self._h, _native.ptr(vol), _native.ptr(maps), _native.ptr(line), int(angle),
                     _native.stream())
        return line

    def bp_angle(self, line, maps, vol, angle: int, mode: int):
        _native.call("chip_bp_angle", self._h, _native.ptr(line), _native.ptr(maps), _native.ptr(vol), int(angle),
                     int(mode), _native.stream())
        return vol

    def set_subangles(self, sub_angles) -> None:
        """Per-detector outgoing sub-angle offsets (a sequence of 1-D arrays, ``chip_plan_set_subangles``)."""
        subs = [np.asarray(s, dtype=np.float64).reshape(-1) for s in sub_angles]
        if len(subs) != self.n_det:
            raise ValueError(f"Expected {self.n_det} sub-angle sets, got {len(subs)}")
        s_max = max(len(s) for s in subs)
        counts = np.ascontiguousarray([len(s) for s in subs], dtype=np.int32)
        table = np.zeros((self.n_det, s_max), dtype=np.float64)
        for d, s in enumerate(subs):
            table[d, : len(s)] = s
        _native.call("chip_plan_set_subangles", self._h, s_max, counts.ctypes.data, table.ctypes.data,
                     self.angles.ctypes.data, self.det_angles.ctypes.data)

    def att_maps(self, mu_in, mu_out, maps, batch: int = 1):
        _native.call("chip_att_maps", self._h, _native.ptr(mu_in), _native.ptr(mu_out), _native.ptr(maps),
                     int(batch), _native.stream())
        return maps

    def fp_otf(self, vol, mu_in, mu_out, out, batch: int = 1):
        """fp with the maps built per angle block from mu_in / mu_out (``chip_fp_otf``)."""
        _native.call("chip_fp_otf", self._h, _native.ptr(vol), _native.ptr(mu_in), _native.ptr(mu_out),
                     _native.ptr(out), int(batch), _native.stream())
        return out

    def bp_otf(self, sino, mu_in, mu_out, out, batch: int, mode: int):
        """bp (symmetric or exact adjoint) with the maps built per angle block (``chip_bp_otf``)."""
        _native.call("chip_bp_otf", self._h, _native.ptr(sino), _native.ptr(mu_in), _native.ptr(mu_out),
                     _native.ptr(out), int(batch), int(mode), _native.stream())
        return out


class ProjectorBackend(ABC):
    """Projector backend interface (``_projector_backends.py:102-248``)."""

    def __init__(self):
        self.is_initialized = False
        self.is_ready = False

    def initialize_geometry(self, vol_geom: VolumeGeometry, angles_rot_rad, rot_axis_shift_pix=None, prj_geom=None,
                            create_single_projs: bool = False):
        self.vol_geom = vol_geom
        self.vol_shape_zxy = self.vol_geom.shape_zxy
        self.angles_w_rad = np.array(angles_rot_rad, ndmin=1, dtype=np.float32)
        self.vol_eff_shape_zxy = np.floor(self.vol_shape_zxy * self.vol_geom.vox_size).astype(int)
        self.prj_shape_vwu = np.array(
            [*self.vol_eff_shape_zxy[:-2], len(self.angles_w_rad), self.vol_eff_shape_zxy[-1]], dtype=int
        )
        self.prj_shape_vu = np.array([*self.vol_eff_shape_zxy[:-2], 1, self.vol_eff_shape_zxy[-1]], dtype=int)
        self.has_individual_projs = create_single_projs
        self.is_initialized = True

    def get_vol_shape(self) -> np.ndarray:
        return self.vol_shape_zxy

    def get_prj_shape(self) -> np.ndarray:
        return self.prj_shape_vwu

    def make_ready(self) -> None:
        self.is_ready = True

    def dispose(self) -> None:
        self.is_ready = False

    def __del__(self):
        if getattr(self, "is_ready", False):
            self.dispose()

    def __repr__(self) -> str:
        name = f"{self.__class__.__name__}: "
        if self.is_initialized:
            return (name + "{\n" + f"  Shape vol ZXY: {self.get_vol_shape()}\n"
                    + f"  Shape prj VWU: {self.get_prj_shape()}\n"
                    + f"  Angles (deg): {np.rad2deg(self.angles_w_rad)}\n" + "}")
        return name + "{ Not initialized! }"

    @abstractmethod
    def fp(self, vol, angle_ind: int | None = None): ...

    @abstractmethod
    def bp(self, prj, angle_ind: int | None = None): ...


BP_MODES = {"skimage": _native.CHIP_BP_SKIMAGE, "adjoint": _native.CHIP_BP_ADJOINT}


class ProjectorBackendHIP(ProjectorBackend):
    """MI355X projector backend with scikit-image discretisation.

    ``bp_mode="skimage"`` (default) back-projects like
    ``ProjectorBackendSKimage.bp`` (iradon without filter, times 2/pi);
    ``bp_mode="adjoint"`` uses the exact transpose of ``fp`` instead.
    """

    def __init__(self, bp_mode: str = "skimage") -> None:
        super().__init__()
        if bp_mode not in BP_MODES:
            raise ValueError(f"Unknown bp_mode: {bp_mode}. Available options are: {list(BP_MODES)}")
        self.bp_mode = bp_mode
        self.is_ready = True
        self._plans: dict = {}
        # corrct's projectors call fp / bp per angle from a thread pool
        # (projectors.py:585-589, 612-616): plans are looked up and created under a lock
        self._plans_lock = threading.Lock()

    def initialize_geometry(self, vol_geom: VolumeGeometry, angles_rot_rad, rot_axis_shift_pix=None, prj_geom=None,
                            create_single_projs: bool = False):
        if rot_axis_shift_pix is not None:
            raise ValueError("With the HIP backend, rotation axis shifts are not supported!")
        if prj_geom is not None:
            raise ValueError("With the HIP backend, `ProjectionGeometry` is not supported!")
        super().initialize_geometry(vol_geom, angles_rot_rad, create_single_projs=create_single_projs)
        self.angles_w_deg = np.rad2deg(self.angles_w_rad)
        if self.vol_shape_zxy[-1] != self.vol_shape_zxy[-2]:
            raise ValueError("Only square slices are supported")

    @property
    def n(self) -> int:
        return int(self.vol_shape_zxy[-1])

    @property
    def n_slices(self) -> int:
        return int(np.prod(self.vol_shape_zxy[:-2])) if len(self.vol_shape_zxy) > 2 else 1

    def plan(self, det_angles_rad=(np.pi / 2,), w_det=None, w_ang=None) -> HipPlan:
        det = np.asarray(det_angles_rad, dtype=np.float64).reshape(-1)
        key = (det.tobytes(), None if w_det is None else np.asarray(w_det, np.float64).tobytes(),
               None if w_ang is None else np.asarray(w_ang, np.float64).tobytes())
        with self._plans_lock:
            p = self._plans.get(key)
            if p is None:
                p = HipPlan(self.n, self.angles_w_rad, det, w_det, w_ang)
                self._plans[key] = p
        return p

    def dispose(self) -> None:
        with self._plans_lock:
            plans, self._plans = self._plans, {}
        for p in plans.values():
            p.close()
        super().dispose()

    def make_ready(self) -> None:
        super().make_ready()

    def fp(self, vol, angle_ind: int | None = None):
        """Uncorrected forward projection ``[..., A, N]`` (or one line ``[..., N]``)."""
        t, host = _arrays.to_dev(vol)
        n = self.n
        batch = t.numel() // (n * n)
        p = self.plan()
        if angle_ind is None:
            out = torch.empty((batch, len(self.angles_w_rad), n), dtype=torch.float32, device=t.device)
            p.fp(t, None, out, batch)
            out = out.reshape(*t.shape[:-2], len(self.angles_w_rad), n)
        else:
            lines = []
            for b in range(batch):
                line = torch.empty((1, n), dtype=torch.float32, device=t.device)
                p.fp_angle(t.reshape(batch, n, n)[b], None, line, int(angle_ind))
                lines.append(line[0])
            out = torch.stack(lines).reshape(*t.shape[:-2], n)
        return _arrays.from_dev(out, host)

    def bp(self, prj, angle_ind: int | None = None):
        """Back-projection (``bp_mode``) of ``[..., A, N]`` (or one line ``[..., N]``)."""
        t, host = _arrays.to_dev(prj)
        n = self.n
        p = self.plan()
        mode = BP_MODES[self.bp_mode]
        if angle_ind is None:
            a = len(self.angles_w_rad)
            batch = t.numel() // (a * n)
            out = torch.empty((batch, n, n), dtype=torch.float32, device=t.device)
            p.bp(t, None, out, batch, mode)
            out = out.reshape(*t.shape[:-2], n, n)
        else:
            batch = t.numel() // n
            lines = t.reshape(batch, n)
            out = torch.empty((batch, n, n), dtype=torch.float32, device=t.device)
            for b in range(batch):
                p.bp_angle(lines[b].contiguous(), None, out[b], int(angle_ind), mode)
            out = out.reshape(*t.shape[:-1], n, n)
        return _arrays.from_dev(out, host)


def compute_attenuation(vol, angle_rad: float, invert: bool = False):
    """exp(-cumulative mu) along the beam at ``angle_rad`` (``_projector_backends.py:43-99``).

    ``vol`` is [..., N, N] (a stack of slices). The in-beam angle is used as a
    float32 value, as the projectors always pass it; ``invert`` adds pi in
    float64 (outgoing photons).
    """
    t, host = _arrays.to_dev(vol)
    n = int(t.shape[-1])
    if t.shape[-2] != n:
        raise ValueError("Only square slices are supported")
    batch = t.numel() // (n * n)
    a32 = np.float32(angle_rad)
    if invert:
        # theta + phi_det + pi with theta = float32(angle), phi_det = the float64 remainder
        plan = HipPlan(n, [a32], det_angles_rad=[float(np.float64(angle_rad) - np.float64(a32))])
        maps = torch.empty((batch, 1, 1, n, n), dtype=torch.float32, device=t.device)
        plan.att_maps(None, t, maps, batch)
    else:
        plan = HipPlan(n, [a32])
        maps = torch.empty((batch, 1, 1, n, n), dtype=torch.float32, device=t.device)
        plan.att_maps(t, None, maps, batch)
    torch.cuda.current_stream().synchronize()
    plan.close()
    return _arrays.from_dev(maps.reshape(t.shape), host)
