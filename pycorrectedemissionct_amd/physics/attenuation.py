"""Attenuation maps (corrct ``physics/attenuation.py`` mirror, GPU-built).

``AttenuationVolume`` follows ``attenuation.py:32-331``: it takes the local
linear attenuation of the incoming beam and of the emitted photons, the
(float32) rotation angles and the detector angles, and ``compute_maps``
fills ``maps[A, D, *vol]`` with ``E_in(theta_a) * E_out(theta_a + phi_d + pi)``
-- here in two HIP kernels (``chip_att_maps``) instead of a thread pool over
angles. The maps stay in device memory; ``get_maps`` returns NumPy when the
inputs were NumPy (as the reference does) and the device tensor otherwise.

Detectors with a non-zero surface average the outgoing factor over
``emitted_sub_sampling`` sub-angles spanning the detector's angular range
(``attenuation.py:110-119``), on the device. Not supported (raise): ``roi``
extraction (also unimplemented in the reference) and ``binning > 1``.
"""

from __future__ import annotations

from collections.abc import Sequence

import numpy as np
import torch

from .. import _arrays, _native
from .._projector_backends import HipPlan


class DetectorXRF:
    """Minimal XRF detector description (``physics/xrf.py:218-258``): angle and surface."""

    def __init__(self, surface_mm2: float = 0.0, distance_mm: float = 1.0, angle_rad: float = np.pi / 2):
        self.surface_mm2 = float(surface_mm2)
        self.distance_mm = float(distance_mm)
        self.angle_rad = float(angle_rad)

    @property
    def angle_range_rad(self) -> tuple[float, float]:
        """Half-opening angle of the detector surface seen from the sample (``physics/xrf.py:238-246``)."""
        half_edge = np.sqrt(self.surface_mm2) / 2
        half_angle = np.arctan2(half_edge, self.distance_mm)
        return (-half_angle, half_angle)


class AttenuationVolume:
    """Attenuation volume computation class (``attenuation.py:32-331``)."""

    def __init__(self, incident_local, emitted_local, angles_rot_rad, angles_det_rad=np.pi / 2,
                 emitted_sub_sampling: int = 1, dtype=np.float32):
        self.incident_local = incident_local
        self.emitted_local = emitted_local
        self.angles_rot_rad = np.array(angles_rot_rad, ndmin=1)
        self.detectors: list[DetectorXRF] = []
        if isinstance(angles_det_rad, (int, float, DetectorXRF)):
            angles_det_rad = [angles_det_rad]
        for a in angles_det_rad:
            if isinstance(a, DetectorXRF):
                self.detectors.append(a)
            elif isinstance(a, (float, np.floating)):
                self.detectors.append(DetectorXRF(surface_mm2=0, distance_mm=1, angle_rad=float(a)))
            else:
                raise ValueError(
                    f"Input parameter angles_det_rad = {angles_det_rad} should be one of: float | NDArray | "
                    "Sequence[float | DetectorXRF]"
                )
        self.dtype = dtype
        if self.incident_local is not None:
            self.vol_shape_zyx = np.array(np.shape(self.incident_local))
            if self.emitted_local is not None and np.any(self.vol_shape_zyx != np.shape(self.emitted_local)):
                raise ValueError(
                    f"Incident volume shape ({np.shape(self.incident_local)}) does not"
                    + f" match the emitted volume shape ({np.shape(self.emitted_local)})"
                )
        elif self.emitted_local is not None:
            self.vol_shape_zyx = np.array(np.shape(self.emitted_local))
        else:
            raise ValueError("No attenuation volumes were given.")
        self.emitted_sub_sampling = emitted_sub_sampling
        self.vol_shape_zyx = np.array(self.vol_shape_zyx, ndmin=1)
        num_dims = len(self.vol_shape_zyx)
        if num_dims not in (2, 3):
            raise ValueError(f"Maps can only be 2D or 3D Arrays. A {num_dims}-dimensional was passed ({self.vol_shape_zyx}).")
        self._host = _arrays.is_host(incident_local if incident_local is not None else emitted_local)
        self.maps_device: torch.Tensor | None = None

    def _get_detector_angles(self) -> np.ndarray:
        return np.array([det.angle_rad for det in self.detectors], ndmin=1)

    def _get_sub_angles(self) -> list[np.ndarray]:
        """Outgoing sub-angle offsets per detector: midpoints of ``emitted_sub_sampling`` equal
        parts of the detector's angular range, or a single 0 for a point detector
        (``attenuation.py:111-116``)."""
        subs = []
        for det in self.detectors:
            if det.surface_mm2 > 0.0:
                edges = np.linspace(*det.angle_range_rad, self.emitted_sub_sampling + 1)
                subs.append((edges[:-1] + edges[1:]) / 2)
            else:
                subs.append(np.zeros(1))
        return subs

    def compute_maps(self, use_multithreading: bool = True, verbose: bool = True) -> None:
        """Fill the maps ``[A, D, *vol]`` on the device (``attenuation.py:122-180``)."""
        n = int(self.vol_shape_zyx[-1])
        if self.vol_shape_zyx[-2] != n:
            raise ValueError("Only square slices are supported")
        n_slices = int(np.prod(self.vol_shape_zyx[:-2])) if len(self.vol_shape_zyx) > 2 else 1
        mu_in = None if self.incident_local is None else _arrays.to_dev(self.incident_local)[0]
        mu_out = None if self.emitted_local is None else _arrays.to_dev(self.emitted_local)[0]
        plan = HipPlan(n, np.asarray(self.angles_rot_rad, dtype=np.float32), self._get_detector_angles())
        subs = self._get_sub_angles()
        if mu_out is not None and any(len(sa) > 1 or sa[0] != 0.0 for sa in subs):
            plan.set_subangles(subs)
        n_a, n_d = plan.n_angles, plan.n_det
        maps = torch.empty((n_slices, n_a, n_d, n, n), dtype=torch.float32, device=_arrays.device())
        plan.att_maps(mu_in, mu_out, maps, n_slices)
        torch.cuda.current_stream().synchronize()
        plan.close()
        if n_slices == 1 and len(self.vol_shape_zyx) == 2:
            self.maps_device = maps[0]
        else:
            # [A, D, Z, N, N] as the reference lays out 3-D maps: a view of the slice-major
            # [Z, A, D, N, N] buffer the projector kernels read
            self.maps_device = maps.permute(1, 2, 0, 3, 4)

    @property
    def maps(self):
        if self.maps_device is None:
            raise ValueError("Maps not computed: call `compute_maps` first")
        return _arrays.from_dev(self.maps_device.contiguous() if self._host else self.maps_device, self._host)

    def get_maps(self, roi=None, rot_ind=None, det_ind=None, binning: int = 1):
        """The maps, optionally restricted to angles / detectors (``attenuation.py:253-297``)."""
        if self.maps_device is None:
            raise ValueError("Maps not computed: call `compute_maps` first")
        maps = self.maps_device
        if rot_ind is not None:
            if isinstance(rot_ind, int):
                rot_ind = slice(rot_ind, rot_ind + 1, 1)
            maps = maps[rot_ind, ...]
        if det_ind is not None:
            if isinstance(det_ind, int):
                det_ind = slice(det_ind, det_ind + 1, 1)
            maps = maps[:, det_ind, ...]
        if roi is not None:
            raise NotImplementedError("Extracting a region of interest is not supported, yet.")
        if binning > 1:
            maps = self._bin_maps(maps.contiguous(), binning)
        return _arrays.from_dev(maps.contiguous(), self._host)

    def _bin_maps(self, maps, binning):
        """``bin_imgs(maps, binning, axes=<volume axes>, auto_crop=True)`` (``processing/pre.py:180-220``):
        block means over the volume axes after cropping each to a multiple of ``binning``
        (``chip_bin_images``)."""
        if not float(binning).is_integer():
            raise NotImplementedError("the device binning takes an integer factor (fractional binning rescales)")
        binning = int(binning)
        nd = len(self.vol_shape_zyx)
        dims = [int(v) for v in maps.shape[-nd:]]
        lead = tuple(int(v) for v in maps.shape[:-nd])
        out = torch.empty((*lead, *[d // binning for d in dims]), dtype=torch.float32, device=maps.device)
        dims_c = np.ascontiguousarray(dims, dtype=np.int32)
        _native.call("chip_bin_images", maps.data_ptr(), out.data_ptr(), int(np.prod(lead, dtype=np.int64)), nd,
                     dims_c.ctypes.data, binning, _native.stream())
        return out

    def get_projector_args(self, roi=None, rot_ind=None, det_ind=None, binning: int = 1) -> dict:
        """``att_maps`` and ``angles_detectors_rad`` for the projector (``attenuation.py:299-331``)."""
        if det_ind is None:
            det_angles = self._get_detector_angles()
        else:
            det_angles = self.detectors[det_ind].angle_rad
        return dict(att_maps=self.get_maps(roi=roi, rot_ind=rot_ind, det_ind=det_ind, binning=binning),
                    angles_detectors_rad=det_angles)
