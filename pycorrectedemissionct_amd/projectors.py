"""Tomographic projectors (corrct ``projectors.py`` mirror) on the HIP backend.

* ``ProjectorMatrix``          -- ``projectors.py:24-100``: dense explicit matrix
  (the fake backend of corrct's solver tests), applied on the device.
* ``ProjectorUncorrected``     -- ``projectors.py:103-341``.
* ``ProjectorAttenuationXRF``  -- ``projectors.py:344-622``: the whole-sinogram
  ``fp``/``bp`` run as single fused launches over all angles and detectors
  (maps and weights applied inside the kernels) instead of a thread pool of
  per-angle scikit-image calls.

Same constructor signatures, shapes, error messages and quirks as the
reference: the default (non-symmetric) XRF ``bp`` ignores maps and detector
weights (``projectors.py:619-622``); ``prj_shape`` is the single-detector
``[A, N]`` shape; fp/bp return fresh arrays. ``backend`` accepts ``"hip"``,
``"skimage"`` (same discretisation, computed by the HIP backend), a
``ProjectorBackendHIP`` instance, or ``"astra"`` (not available here: falls
back with a warning, as the reference does without ASTRA).

Build-only extension: ``bp_mode="adjoint"`` makes ``bp`` the exact transpose
of ``fp`` (maps and weights included).
"""

from __future__ import annotations

import numpy as np
import torch

from . import _arrays, _native, models, operators
from . import _projector_backends as prj_backends
from .physics.attenuation import AttenuationVolume

num_threads = 1
astra_available = False


class ProjectorMatrix(operators.ProjectorOperator):
    """Projector that uses an explicit projection matrix (``projectors.py:24-100``)."""

    def __init__(self, A, vol_shape, prj_shape, _dev_mat: torch.Tensor | None = None) -> None:
        self.vol_shape = np.array(vol_shape, ndmin=1, dtype=int)
        self.prj_shape = np.array(prj_shape, ndmin=1, dtype=int)
        self.A = A
        self._dev = _dev_mat
        super().__init__()

    def _mat(self) -> torch.Tensor:
        if self._dev is None:
            self._dev = _arrays.to_dev(np.asarray(self.A, dtype=np.float32))[0]
        return self._dev

    def _transpose(self) -> "ProjectorMatrix":
        return ProjectorMatrix(np.asarray(self.A).transpose(), self.prj_shape, self.vol_shape, self._mat().t())

    def absolute(self) -> "ProjectorMatrix":
        return ProjectorMatrix(np.abs(np.asarray(self.A)), self.vol_shape, self.prj_shape, torch.abs(self._mat()))

    def fp(self, x):
        t, host = _arrays.to_dev(x)
        out = torch.mv(self._mat(), t.reshape(-1)).reshape(tuple(self.prj_shape))
        return _arrays.from_dev(out, host)

    def bp(self, x):
        t, host = _arrays.to_dev(x)
        out = torch.mv(self._mat().t(), t.reshape(-1)).reshape(tuple(self.vol_shape))
        return _arrays.from_dev(out, host)


def _resolve_backend(backend) -> prj_backends.ProjectorBackend:
    if isinstance(backend, prj_backends.ProjectorBackend):
        return backend
    if isinstance(backend, str):
        if backend == "astra":
            print("WARNING: ASTRA backend requested but not available. Falling back to the HIP backend.")
            return prj_backends.ProjectorBackendHIP()
        if backend in ("hip", "skimage"):
            return prj_backends.ProjectorBackendHIP()
        raise ValueError(f"Unknown backend: {backend}. Available options are: 'hip', 'skimage', 'astra'.")
    raise ValueError(f"Unknown backend: {backend}")


class ProjectorUncorrected(operators.ProjectorOperator):
    """Base projection class (``projectors.py:103-341``)."""

    def __init__(self, vol_geom, angles_rot_rad, rot_axis_shift_pix=None, *, prj_geom=None, prj_intensities=None,
                 psf=None, backend="hip", create_single_projs: bool = False):
        if not isinstance(vol_geom, models.VolumeGeometry):
            vol_geom = models.VolumeGeometry(_vol_shape_xyz=np.array(vol_geom))
        self.vol_geom = vol_geom
        self.prj_geom = prj_geom
        if len(self.vol_geom.shape_xyz) not in (2, 3):
            raise ValueError(f"Only 2D or 3D volumes are valid, following volume geometry passed instead: {self.vol_geom}")
        if not self.vol_geom.is_square():
            raise ValueError(f"Only square volumes are valid, following volume geometry passed instead: {self.vol_geom}")
        self.projector_backend = _resolve_backend(backend)
        if prj_geom is not None:
            raise ValueError("Using class `ProjectionGeometry` requires using astra-toolbox.")
        self.projector_backend.initialize_geometry(
            vol_geom=vol_geom, angles_rot_rad=angles_rot_rad, rot_axis_shift_pix=rot_axis_shift_pix,
            prj_geom=prj_geom, create_single_projs=create_single_projs,
        )
        if prj_intensities is not None:
            prj_intensities = np.array(prj_intensities, dtype=np.float64)
        self.prj_intensities = prj_intensities
        self._set_psf(psf)
        self.vol_shape = np.array(self.projector_backend.get_vol_shape(), ndmin=1)
        self.prj_shape = np.array(self.projector_backend.get_prj_shape(), ndmin=1)
        super().__init__()

    @property
    def angles_rot_rad(self) -> np.ndarray:
        return self.projector_backend.angles_w_rad

    def _set_psf(self, psf) -> None:
        """Detector PSF (``projectors.py:229-244``): a 1-D kernel convolved along u after every fp
        with edge-replicated borders. As in the reference (``is_conv_symm=False``), its adjoint is
        the identity, so the back-projections ignore it."""
        if psf is None:
            self.psf_vu = self.psf_vwu = None
            return
        psf = np.squeeze(np.array(psf, dtype=np.float32))
        if psf.ndim >= len(self.vol_geom.shape_xyz) or psf.ndim != 1:
            raise ValueError(
                "PSF should either be 1D (for 2D and 3D reconstructions) or 2D (for 3D reconstructions)."
                + f" Passed PSF has shape: {psf.shape}, and the reconstruction is {len(self.vol_geom.shape_xyz)}D."
            )
        self.psf_vu = self.psf_vwu = psf

    def _apply_psf(self, prj):
        if self.psf_vu is None:
            return prj
        t, host = _arrays.to_dev(prj)
        t = t.contiguous()
        n = int(t.shape[-1])
        kern = torch.as_tensor(self.psf_vu, device=t.device)
        out = torch.empty_like(t)
        _native.call("chip_psf_lines", t.data_ptr(), out.data_ptr(), kern.data_ptr(), int(kern.numel()),
                     t.numel() // n, n, _native.stream())
        return _arrays.from_dev(out, host)

    def __enter__(self):
        self.projector_backend.make_ready()
        return self

    def __exit__(self, *args):
        self.projector_backend.dispose()

    def get_pre_weights(self):
        return None

    def _intensities_dev(self, device) -> torch.Tensor:
        return torch.as_tensor(self.prj_intensities, dtype=torch.float32, device=device)

    def fp_angle(self, vol, angle_ind: int):
        prj_vu = self.projector_backend.fp(vol, angle_ind)
        if self.prj_intensities is not None:
            prj_vu = prj_vu * np.float32(self.prj_intensities[angle_ind])
        return self._apply_psf(prj_vu)

    def bp_angle(self, prj_vu, angle_ind: int):
        if self.prj_intensities is not None:
            prj_vu = prj_vu * np.float32(self.prj_intensities[angle_ind])
        return self.projector_backend.bp(prj_vu, angle_ind)

    def fp(self, vol):
        prj_vwu = self.projector_backend.fp(vol)
        if self.prj_intensities is not None:
            if isinstance(prj_vwu, torch.Tensor):
                prj_vwu = prj_vwu * self._intensities_dev(prj_vwu.device)[:, None]
            else:
                prj_vwu *= self.prj_intensities[:, None].astype(np.float32)
        return self._apply_psf(prj_vwu)

    def bp(self, prj_vwu):
        if self.prj_intensities is not None:
            if isinstance(prj_vwu, torch.Tensor):
                prj_vwu = prj_vwu * self._intensities_dev(prj_vwu.device)[:, None]
            else:
                prj_vwu = prj_vwu * self.prj_intensities[:, None].astype(np.float32)
        return self.projector_backend.bp(prj_vwu)


class ProjectorAttenuationXRF(ProjectorUncorrected):
    """Attenuation-corrected XRF projector with multi-detector support (``projectors.py:344-622``)."""

    def __init__(self, vol_geom, angles_rot_rad, rot_axis_shift_pix=None, *, prj_geom=None, prj_intensities=None,
                 backend="hip", att_maps=None, att_in=None, att_out=None, angles_detectors_rad=(np.pi / 2),
                 weights_detectors=None, psf=None, is_symmetric: bool = False, weights_angles=None,
                 use_multithreading: bool = True, data_type=np.float32, verbose: bool = True,
                 bp_mode: str = "skimage", maps_mode: str = "precomputed"):
        ProjectorUncorrected.__init__(
            self, vol_geom=vol_geom, angles_rot_rad=angles_rot_rad, rot_axis_shift_pix=rot_axis_shift_pix,
            prj_geom=prj_geom, psf=psf, prj_intensities=prj_intensities, backend=backend, create_single_projs=True,
        )
        if bp_mode not in ("skimage", "adjoint"):
            raise ValueError(f"Unknown bp_mode: {bp_mode}")
        if maps_mode not in ("precomputed", "otf"):
            raise ValueError(f"Unknown maps_mode: {maps_mode}")
        self.bp_mode = bp_mode
        self.maps_mode = maps_mode
        self.data_type = data_type
        self.use_multithreading = use_multithreading
        self.verbose = verbose
        self.is_symmetric = is_symmetric

        self.angles_det_rad = np.array(angles_detectors_rad, ndmin=1)
        num_det_angles = len(self.angles_det_rad)
        if weights_detectors is None:
            weights_detectors = np.ones_like(self.angles_det_rad)
        self.weights_det = np.array(weights_detectors, ndmin=1)
        num_det_weights = len(self.weights_det)
        if num_det_angles > 1:
            if num_det_weights == 1:
                self.weights_det = np.tile(self.weights_det, [num_det_angles])
            elif num_det_weights > 1 and not num_det_weights == num_det_angles:
                raise ValueError(
                    "Number of detector weights differs from number of"
                    f" detector angles: {num_det_weights} vs {num_det_angles}"
                )
        if weights_angles is None:
            weights_angles = np.ones((len(self.angles_rot_rad), num_det_angles))
        else:
            weights_angles = np.array(weights_angles)
        self.weights_angles = weights_angles

        self.mu_in_dev = self.mu_out_dev = None
        if maps_mode == "otf":
            # maps are rebuilt per angle block inside every fp / bp (chip_fp_otf, chip_bp_otf)
            if att_maps is not None:
                raise ValueError("maps_mode='otf' builds the maps from att_in / att_out: do not pass att_maps")
            if att_in is None and att_out is None:
                raise ValueError("No attenuation volumes were given.")
            for mu in (att_in, att_out):
                if mu is not None and tuple(np.shape(mu)) != tuple(self.vol_shape):
                    raise ValueError(f"Mismatching attenuation volume shape {np.shape(mu)} with vol_shape {self.vol_shape}")
            self.mu_in_dev = None if att_in is None else _arrays.to_dev(att_in)[0].contiguous()
            self.mu_out_dev = None if att_out is None else _arrays.to_dev(att_out)[0].contiguous()
            self.att_vol_angles = None
        elif att_maps is None:
            m = AttenuationVolume(att_in, att_out, self.angles_rot_rad, self.angles_det_rad)
            m.compute_maps(use_multithreading=self.use_multithreading, verbose=self.verbose)
            self.att_vol_angles = m.maps_device
        else:
            if att_maps.shape[1] != num_det_angles:
                raise ValueError(f"Number of maps ({att_maps.shape[1]}) differs from number of detectors ({num_det_angles})")
            if not np.all(np.equal(att_maps.shape[-2:], self.vol_shape[-2:])):
                raise ValueError(
                    f"Mismatching volume shape of input volume ({att_maps.shape[-2:]})"
                    + f" with vol_shape ({self.vol_shape} in 2D -> {self.vol_shape[-2:]})"
                )
            if tuple(att_maps.shape[2:]) != tuple(self.vol_shape):
                raise ValueError(f"Maps of shape {tuple(att_maps.shape)} do not match the volume {tuple(self.vol_shape)}")
            self.att_vol_angles = _arrays.to_dev(att_maps)[0]
        self.executor = None

    # ---- plan with the current weights ------------------------------------
    def _plan(self, with_intensities: bool = True) -> prj_backends.HipPlan:
        w_ang = np.asarray(self.weights_angles, dtype=np.float64).reshape(len(self.angles_rot_rad), -1)
        if w_ang.shape[1] != len(self.angles_det_rad):
            w_ang = np.broadcast_to(w_ang, (len(self.angles_rot_rad), len(self.angles_det_rad)))
        if with_intensities and self.prj_intensities is not None:
            w_ang = w_ang * self.prj_intensities[:, None]
        return self.projector_backend.plan(self.angles_det_rad, self.weights_det, np.ascontiguousarray(w_ang))

    @property
    def _n_slices(self) -> int:
        return int(np.prod(self.vol_shape[:-2])) if len(self.vol_shape) > 2 else 1

    def _maps(self, device) -> torch.Tensor | None:
        """The maps in the kernels' layout: [A, D, N, N] for a slice, [Z, A, D, N, N] for a
        volume (``att_vol_angles`` keeps the reference's [A, D, Z, N, N] view of it)."""
        m = self.att_vol_angles
        if m is None:
            return None
        if m.device != device:
            m = m.to(device)
        if len(self.vol_shape) > 2:
            z = m.permute(2, 0, 1, 3, 4)
            if not z.is_contiguous():
                z = z.contiguous()
                self.att_vol_angles = z.permute(1, 2, 0, 3, 4)
            return z
        if not m.is_contiguous():
            m = m.contiguous()
            self.att_vol_angles = m
        return m

    def _maps_stride(self) -> int:
        n = int(self.vol_shape[-1])
        return len(self.angles_rot_rad) * len(self.angles_det_rad) * n * n if len(self.vol_shape) > 2 else 0

    def _require_maps(self, what: str) -> None:
        if self.maps_mode == "otf":
            raise NotImplementedError(f"{what} needs precomputed maps (maps_mode='precomputed')")

    def collapse_detectors(self) -> None:
        """Convert multi-detector configurations into single-detector (``projectors.py:492-500``)."""
        self._require_maps("collapse_detectors")
        # the detector axis is 1 of [A, D, N, N] and of [A, D, Z, N, N]
        weights = np.reshape(self.weights_det, [1, -1] + [1] * (self.att_vol_angles.ndim - 2)) / np.sum(self.weights_det)
        w_dev = torch.as_tensor(weights, dtype=torch.float32, device=self.att_vol_angles.device)
        self.att_vol_angles = torch.sum(self.att_vol_angles * w_dev, dim=1, keepdim=True)
        weights = np.squeeze(weights)
        self.angles_det_rad = np.sum(self.angles_det_rad * weights, keepdims=True)
        self.weights_angles = np.sum(self.weights_angles * weights, axis=1, keepdims=True)
        self.weights_det = np.sum(self.weights_det, keepdims=True)

    # A volume [Z, N, N] is Z slices, each with its own maps; its sinogram is [D, Z, A, N]
    # (projectors.py:570-596 stacks [D, Z, N] lines over the angle axis -2), squeezed to
    # [Z, A, N] for one detector. The kernels run slice-major ([Z, D, A, N]).
    def fp_angle(self, vol, angle_ind: int):
        self._require_maps("fp_angle")
        t, host = _arrays.to_dev(vol)
        n = int(self.vol_shape[-1])
        d = len(self.angles_det_rad)
        z = self._n_slices
        t = t.reshape(z, n, n)
        maps = self._maps(t.device)
        plan = self._plan()
        lines = torch.empty((z, d, n), dtype=torch.float32, device=t.device)
        for iz in range(z):
            plan.fp_angle(t[iz], maps if z == 1 else maps[iz], lines[iz], int(angle_ind))
        line = lines[0] if len(self.vol_shape) == 2 else lines.permute(1, 0, 2)
        line = self._apply_psf(line)
        if d == 1:
            line = line[0]
        return _arrays.from_dev(line, host)

    def bp_angle(self, sino, angle_ind: int, single_line: bool = False):
        self._require_maps("bp_angle")
        t, host = _arrays.to_dev(sino)
        n = int(self.vol_shape[-1])
        d = len(self.weights_det)
        z = self._n_slices
        line = t if single_line else t[..., angle_ind, :]
        line = line.reshape(d, z, n).permute(1, 0, 2).contiguous()
        vol = torch.empty((z, n, n), dtype=torch.float32, device=t.device)
        plan = self._plan()
        maps = self._maps(t.device) if (self.bp_mode == "adjoint" or self.is_symmetric) else None
        mode = _native.CHIP_BP_ADJOINT if self.bp_mode == "adjoint" else _native.CHIP_BP_SKIMAGE_SYM
        for iz in range(z):
            plan.bp_angle(line[iz], None if maps is None else (maps if z == 1 else maps[iz]), vol[iz],
                          int(angle_ind), mode)
        return _arrays.from_dev(vol.reshape(tuple(self.vol_shape)), host)

    def fp(self, vol):
        t, host = _arrays.to_dev(vol)
        n = int(self.vol_shape[-1])
        a = len(self.angles_rot_rad)
        d = len(self.angles_det_rad)
        z = self._n_slices
        if t.numel() != z * n * n:
            raise ValueError(f"Expected a volume of shape {tuple(self.vol_shape)}, got {tuple(t.shape)}")
        out = torch.empty((z, d, a, n), dtype=torch.float32, device=t.device)
        if self.maps_mode == "otf":
            self._plan().fp_otf(t.contiguous(), self.mu_in_dev, self.mu_out_dev, out, z)
        else:
            self._plan().fp(t, self._maps(t.device), out, z, self._maps_stride())
        out = out[0] if len(self.vol_shape) == 2 else out.permute(1, 0, 2, 3)
        out = self._apply_psf(out)
        if d == 1:
            out = out[0]
        return _arrays.from_dev(out.contiguous(), host)

    def supports_fused_data_step(self) -> bool:
        """Whether ``fp_row_bp`` applies: on-the-fly maps, one slice, no detector PSF, and a bp that reads
        the maps (the exact adjoint or the symmetric bp)."""
        return (self.maps_mode == "otf" and self._n_slices == 1 and self.psf_vu is None
                and (self.bp_mode == "adjoint" or self.is_symmetric))

    def fp_row_bp(self, x, row_op: int, p, b, sigma=None, bg=None, b_mask=None, ax=None):
        """One solver iteration's projector work with on-the-fly maps (``chip_otf_fp_row_bp``): per angle
        block, the block's maps are built once, ``ax = fp(x)`` over the block, the row op updates the
        block's rows of ``p`` (``enum chip_row_op``: the PDHG l2 / KL dual step, the SIRT residual or the
        MLEM ratio), and the block's rows of ``p`` are back-projected. Equals ``fp``, the elementwise
        kernel and ``bp`` in turn, with half the map building. Device tensors in the sinogram shape the
        projector returns; ``p`` is updated in place. Returns (bp of p, ax)."""
        if not self.supports_fused_data_step():
            raise ValueError("fp_row_bp needs maps_mode='otf', a single slice, no PSF and a bp reading the maps")
        n = int(self.vol_shape[-1])
        xt = _arrays.to_dev(x)[0].contiguous()
        if ax is None:
            ax = torch.empty_like(p)
        for t in (p, b, sigma, bg, b_mask, ax):
            if t is not None and (not t.is_contiguous() or t.numel() != p.numel()):
                raise ValueError("fp_row_bp takes contiguous device arrays of the sinogram's size")
        out = torch.empty((n, n), dtype=torch.float32, device=xt.device)
        mode = _native.CHIP_BP_ADJOINT if self.bp_mode == "adjoint" else _native.CHIP_BP_SKIMAGE_SYM
        _native.call("chip_otf_fp_row_bp", self._plan()._h, xt.data_ptr(), _native.ptr(self.mu_in_dev),
                     _native.ptr(self.mu_out_dev), 1, int(row_op), p.data_ptr(), _native.ptr(sigma), b.data_ptr(),
                     _native.ptr(bg), _native.ptr(b_mask), ax.data_ptr(), out.data_ptr(), mode, _native.stream())
        return out.reshape(tuple(self.vol_shape)), ax

    def bp(self, sino):
        t, host = _arrays.to_dev(sino)
        n = int(self.vol_shape[-1])
        a = len(self.angles_rot_rad)
        d = len(self.weights_det)
        z = self._n_slices
        if t.numel() != d * z * a * n:
            raise ValueError(f"Expected a sinogram with {d * z * a * n} elements, got {tuple(t.shape)}")
        if z > 1:
            t = t.reshape(d, z, a, n).permute(1, 0, 2, 3)
        t = t.contiguous()
        out = torch.empty((z, n, n), dtype=torch.float32, device=t.device)
        mode = _native.CHIP_BP_ADJOINT if self.bp_mode == "adjoint" else _native.CHIP_BP_SKIMAGE_SYM
        if self.maps_mode == "otf" and (self.bp_mode == "adjoint" or self.is_symmetric):
            self._plan().bp_otf(t, self.mu_in_dev, self.mu_out_dev, out, z, mode)
        elif self.bp_mode == "adjoint" or self.is_symmetric:
            self._plan().bp(t, self._maps(t.device), out, z, mode, self._maps_stride())
        else:
            if self.prj_intensities is not None:
                t = (t.reshape(z, d, a, n) * self._intensities_dev(t.device)[None, None, :, None]).contiguous()
            self._plan(with_intensities=False).bp(t, None, out, z, _native.CHIP_BP_SKIMAGE)
        return _arrays.from_dev(out.reshape(tuple(self.vol_shape)), host)
