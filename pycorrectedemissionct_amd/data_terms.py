"""Data-fidelity terms used by the solvers (corrct ``data_terms.py`` subset).

The classes carry the configuration the fused HIP solver kernels need; the
arithmetic of ``update_dual`` + ``apply_proximal_dual`` runs inside
``chip_pdhg_dual_l2`` / ``chip_pdhg_dual_kl`` (one HBM pass per iteration).

* ``DataFidelity_l2``   -- ``data_terms.py:283-331``
* ``DataFidelity_KL``   -- ``data_terms.py:677-757``
* ``DataFidelity_l1`` / ``DataFidelity_l12`` -- ``data_terms.py:535-625``, the
  norm of the TV regularizer (its prox is fused into ``chip_tv_dual``).

``compute_residual`` / ``compute_residual_norm`` mirror ``data_terms.py:136-161``
and the norms of each class, evaluated on the device.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _arrays

EPS = float(np.finfo(np.float32).eps)


class DataFidelityBase:
    __data_fidelity_name__ = ""

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

    def compute_residual_norm(self, dual: torch.Tensor) -> float:
        return float(torch.sum(dual.double() ** 2).item())


class DataFidelity_l1(DataFidelityBase):
    """l1-norm data fidelity (``data_terms.py:535-597``); TV dual norm."""

    __data_fidelity_name__ = "l1"

    def compute_residual_norm(self, dual: torch.Tensor) -> float:
        return float(torch.sum(torch.abs(dual.double())).item())


class DataFidelity_l12(DataFidelity_l1):
    """l12-norm (``data_terms.py:600-625``): l2 over ``l2_axis``, l1 over the rest."""

    __data_fidelity_name__ = "l12"

    def __init__(self, background=None, l2_axis: int = 0):
        super().__init__(background=background)
        self.l2_axis = l2_axis


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
