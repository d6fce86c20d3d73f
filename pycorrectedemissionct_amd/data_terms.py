"""Data-fidelity terms used by the solvers (corrct ``data_terms.py`` subset).

The classes carry the configuration the fused HIP solver kernels need; the
arithmetic of ``update_dual`` + ``apply_proximal_dual`` runs inside
``chip_pdhg_dual_l2`` / ``chip_pdhg_dual_kl`` / ``chip_pdhg_dual_norm`` (one HBM
pass per iteration), and SIRT / FISTA take ``compute_residual`` of the l2-based
terms and KL from ``chip_data_residual``.

* ``DataFidelity_l2``    -- ``data_terms.py:283-331``
* ``DataFidelity_wl2``   -- ``data_terms.py:334-395``
* ``DataFidelity_l2b``   -- ``data_terms.py:398-464``
* ``DataFidelity_Huber`` -- ``data_terms.py:467-532``
* ``DataFidelity_l1`` / ``DataFidelity_l12`` / ``DataFidelity_l1b`` --
  ``data_terms.py:535-674``; l1 / l12 are also the norms of the gradient
  regularizers (their prox is fused into ``chip_tv_dual`` / ``chip_grad_dual``).
* ``DataFidelity_KL``    -- ``data_terms.py:677-757``

``device_term(like)`` gives the ``enum chip_data_term`` code, the parameter
array (weights or local error, broadcast to the data shape unless scalar) and
the group layout of the l2 axis for the kernels.

``compute_residual`` / ``compute_residual_norm`` mirror ``data_terms.py:136-161``
and the norms of each class, evaluated on the device.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _arrays

EPS = float(np.finfo(np.float32).eps)

# enum chip_data_term
TERM_L2, TERM_KL, TERM_WL2, TERM_L2B, TERM_HUBER, TERM_L1, TERM_L12, TERM_L1B = range(8)


def _param_dev(value, like: torch.Tensor) -> tuple[torch.Tensor, int]:
    """A term parameter for the kernels: (device array, period); a scalar stays one element."""
    arr = np.asarray(value, dtype=np.float32)
    t = torch.as_tensor(arr, device=like.device)
    if t.numel() == 1:
        return t.reshape(1).contiguous(), 1
    t = torch.broadcast_to(t, like.shape).contiguous()
    return t, int(t.numel())


def _param_like(value, like: torch.Tensor) -> torch.Tensor:
    return torch.broadcast_to(torch.as_tensor(np.asarray(value, dtype=np.float32), device=like.device), like.shape)


def _group_layout(shape, axis: int | None) -> tuple[int, int]:
    """(group_len, group_stride) of the l2 axis ``axis`` of a C-contiguous ``shape`` (1, 1: no groups)."""
    if axis is None:
        return 1, 1
    ax = axis % len(shape)
    return int(shape[ax]), int(np.prod(shape[ax + 1:], dtype=np.int64))


def _soft_threshold(values: torch.Tensor, threshold) -> torch.Tensor:
    """sign(v) max(|v| - t, 0) (``data_terms.py:25-28``)."""
    return torch.sign(values) * torch.clamp(torch.abs(values) - threshold, min=0.0)


def _norm1(t: torch.Tensor) -> float:
    """``np.linalg.norm(t, ord=1)`` without an axis, as the reference calls it on un-flattened arrays
    (``data_terms.py:580-584``): the sum of |t| for 1-D, the largest column sum of |t| for 2-D, and
    numpy's ValueError for more dimensions."""
    t = t.double()
    if t.ndim == 1:
        return float(torch.sum(torch.abs(t)).item())
    if t.ndim == 2:
        return float(torch.max(torch.sum(torch.abs(t), dim=0)).item())
    raise ValueError("Improper number of dimensions to norm.")


class DataFidelityBase:
    __data_fidelity_name__ = ""
    _term: int | None = None

    def __init__(self, background=None) -> None:
        self.background = np.array(background) if background is not None else None
        self.data = None
        self.sigma = 1.0
        self.sigma_data = None

    def info(self) -> str:
        if self.background is not None:
            if np.array(self.background).size > 1:
                bckgrnd_str = "(B:<array>)"
            else:
                bckgrnd_str = "(B:%g)" % self.background
        else:
            bckgrnd_str = ""
        return self.__data_fidelity_name__ + bckgrnd_str

    def upper(self) -> str:
        return self.info().upper()

    def lower(self) -> str:
        return self.info().lower()

    def assign_data(self, data=None, sigma=1.0) -> None:
        self.data = data
        self.sigma = sigma

    def background_dev(self, like: torch.Tensor) -> torch.Tensor | None:
        if self.background is None:
            return None
        bg = torch.as_tensor(np.asarray(self.background, dtype=np.float32), device=like.device)
        return torch.broadcast_to(bg, like.shape).contiguous()

    def device_term(self, like: torch.Tensor) -> tuple[int, torch.Tensor | None, int, int, int]:
        """(``enum chip_data_term`` code, parameter array or None, its period, group_len, group_stride)
        for data shaped like ``like``."""
        if self._term is None:
            raise NotImplementedError(f"data term {self.info()} is not supported by the fused HIP kernels")
        return self._term, None, 1, 1, 1

    def compute_residual(self, proj_primal: torch.Tensor, mask=None) -> torch.Tensor:
        if self.background is not None:
            proj_primal = proj_primal + self.background_dev(proj_primal)
        residual = self.data - proj_primal if self.data is not None else proj_primal.clone()
        if mask is not None:
            residual = residual * mask
        return residual

    def compute_residual_norm(self, dual: torch.Tensor) -> float:
        raise NotImplementedError


class DataFidelity_l2(DataFidelityBase):
    """l2-norm data fidelity (``data_terms.py:283-331``)."""

    __data_fidelity_name__ = "l2"
    _term = TERM_L2

    def compute_residual_norm(self, dual: torch.Tensor) -> float:
        return float(torch.sum(dual.double() ** 2).item())


class DataFidelity_wl2(DataFidelity_l2):
    """Weighted l2-norm data fidelity (``data_terms.py:334-395``): residual ``(b - A x) w``; the dual
    step ``(v - sigma b) (1 - z) / (1 + sigma / (w + z))`` with ``z = (w == 0)``."""

    __data_fidelity_name__ = "wl2"
    _term = TERM_WL2

    def __init__(self, weights, background=None) -> None:
        super().__init__(background=background)
        self.weights = np.array(weights)

    def device_term(self, like):
        w, period = _param_dev(self.weights, like)
        return self._term, w, period, 1, 1

    def compute_residual(self, proj_primal: torch.Tensor, mask=None) -> torch.Tensor:
        if self.background is not None:
            proj_primal = proj_primal + self.background_dev(proj_primal)
        w = _param_like(self.weights, proj_primal)
        residual = (self.data - proj_primal) * w if self.data is not None else proj_primal * w
        if mask is not None:
            residual = residual * mask
        return residual

    def compute_residual_norm(self, dual: torch.Tensor) -> float:
        w = _param_like(self.weights, dual)
        valid = w != 0
        return float(torch.sum((dual[valid].double() / torch.sqrt(w[valid].double())) ** 2).item())


class DataFidelity_l2b(DataFidelity_l2):
    """l2-norm ball data fidelity (``data_terms.py:398-464``): residual and dual soft-thresholded by
    ``sigma sqrt(local_error)``, dual scaled by ``1 / (1 + sigma local_error)``."""

    __data_fidelity_name__ = "l2b"
    _term = TERM_L2B

    def __init__(self, local_error, background=None):
        super().__init__(background=background)
        self.local_error = local_error

    def device_term(self, like):
        e, period = _param_dev(self.local_error, like)
        return self._term, e, period, 1, 1

    def compute_residual(self, proj_primal: torch.Tensor, mask=None) -> torch.Tensor:
        residual = super().compute_residual(proj_primal, mask)
        sqrt_e = torch.sqrt(_param_like(self.local_error, residual))
        return _soft_threshold(residual, self.sigma * sqrt_e)


class DataFidelity_Huber(DataFidelityBase):
    """Huber-norm data fidelity (``data_terms.py:467-532``): dual ``u = (v - sigma b) / (1 + sigma e)``
    projected onto the unit ball, per element or over ``l2_axis``."""

    __data_fidelity_name__ = "Hub"
    _term = TERM_HUBER

    def __init__(self, local_error, background=None, l2_axis: int | None = None):
        super().__init__(background=background)
        self.local_error = local_error
        self.l2_axis = l2_axis

    def device_term(self, like):
        e, period = _param_dev(self.local_error, like)
        return (self._term, e, period, *_group_layout(like.shape, self.l2_axis))

    def compute_residual_norm(self, dual: torch.Tensor) -> float:
        """``data_terms.py:486-489`` as written: ``1 - l2_points`` is an integer array, so the l1 part
        indexes ``dual`` along its first axis with 0 / 1 rather than masking it."""
        l2_points = dual <= _param_like(self.local_error, dual)
        l1_points = 1 - l2_points.to(torch.int64)
        d = dual.double()
        return float((torch.sum(d[l2_points] ** 2) + torch.sum(torch.abs(d[l1_points]))).item())


class DataFidelity_l1(DataFidelityBase):
    """l1-norm data fidelity (``data_terms.py:535-597``); TV dual norm. Dual step: ``v / max(|v|, 1)``."""

    __data_fidelity_name__ = "l1"
    _term = TERM_L1

    def _inner_norm(self, dual: torch.Tensor) -> torch.Tensor:
        return torch.abs(dual)

    def _threshold(self, dual: torch.Tensor) -> torch.Tensor:
        return dual

    def compute_residual_norm(self, dual: torch.Tensor) -> float:
        return _norm1(self._inner_norm(self._threshold(dual)))


class DataFidelity_l12(DataFidelity_l1):
    """l12-norm (``data_terms.py:598-625``): l2 over ``l2_axis``, l1 over the rest."""

    __data_fidelity_name__ = "l12"
    _term = TERM_L12

    def __init__(self, background=None, l2_axis: int = 0):
        super().__init__(background=background)
        self.l2_axis = l2_axis

    def _inner_norm(self, dual: torch.Tensor) -> torch.Tensor:
        return torch.linalg.vector_norm(dual, ord=2, dim=self.l2_axis, keepdim=True)

    def device_term(self, like):
        return (self._term, None, 1, *_group_layout(like.shape, self.l2_axis))


class DataFidelity_l1b(DataFidelity_l1):
    """l1-norm ball data fidelity (``data_terms.py:628-674``): the dual is soft-thresholded by
    ``local_error`` (not ``sigma local_error``, as the reference's ``_apply_threshold`` reads it)
    before the l1 projection."""

    __data_fidelity_name__ = "l1b"
    _term = TERM_L1B

    def __init__(self, local_error, background=None) -> None:
        super().__init__(background=background)
        self.local_error = local_error

    def _threshold(self, dual: torch.Tensor) -> torch.Tensor:
        return _soft_threshold(dual, _param_like(self.local_error, dual))

    def device_term(self, like):
        e, period = _param_dev(self.local_error, like)
        return self._term, e, period, 1, 1


class DataFidelity_ln(DataFidelityBase):
    """Nuclear-norm data fidelity (``data_terms.py:760-842``): the TNV dual norm. Its dual proximal
    (singular values of each ``ln_axes`` matrix through ``spectral_norm``'s dual proximal) runs
    fused in ``chip_grad_dual`` for the ``Regularizer_TNV`` layout."""

    __data_fidelity_name__ = "ln"

    def __init__(self, background=None, ln_axes=(1, -1), spectral_norm: DataFidelityBase | None = None):
        super().__init__(background=background)
        self.ln_axes = ln_axes
        self.spectral_norm = DataFidelity_l1() if spectral_norm is None else spectral_norm
        self.use_fallback = False


class DataFidelity_KL(DataFidelityBase):
    """Kullback-Leibler data fidelity (``data_terms.py:677-757``)."""

    __data_fidelity_name__ = "KL"
    _term = TERM_KL

    def compute_residual(self, proj_primal: torch.Tensor, mask=None, use_proximal: bool = True) -> torch.Tensor:
        if self.background is not None:
            proj_primal = proj_primal + self.background_dev(proj_primal)
        proj_primal = torch.clamp(proj_primal, min=EPS)
        if use_proximal:
            residual = torch.clamp(proj_primal, min=EPS) * self.sigma
            sd = 4 * self.sigma * torch.clamp(self.data, min=0.0) if self.data is not None else 0.0
            residual = (1 + residual - torch.sqrt((residual - 1) ** 2 + sd)) / 2
        else:
            residual = 1.0 - torch.clamp(self.data, min=0.0) / proj_primal if self.data is not None else torch.ones_like(proj_primal)
        if mask is not None:
            residual = residual * mask
        return -residual

    def compute_residual_norm(self, dual: torch.Tensor) -> float:
        return float(torch.sum(torch.abs(dual.double())).item())


def _as_dev(x, like: torch.Tensor):
    if x is None:
        return None
    return _arrays.to_dev(x)[0].reshape(like.shape) if not isinstance(x, torch.Tensor) else x
