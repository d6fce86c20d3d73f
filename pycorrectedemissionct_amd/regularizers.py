"""Regularizers used by the solvers (corrct ``regularizers.py`` subset).

``Regularizer_l1`` (``regularizers.py:531-585``): sparsity, identity transform;
its primal proximal (soft thresholding at tau * weight) is fused into
``chip_fista_update``.

``Regularizer_TV2D`` (``regularizers.py:308-321``) over ``Regularizer_Grad``
(``regularizers.py:218-289``): isotropic total variation with the l12 norm,
edge-padded forward gradient, ``sigma = 1/2`` and ``tau = 2 * ndims``. Its dual
update and prox run fused in ``chip_tv_dual``; its primal contribution
(``-div q``) is fused into the SIRT / PDHG primal kernels.
"""

from __future__ import annotations

import numpy as np

from . import data_terms as dt
from . import operators


class BaseRegularizer:
    __reg_name__ = ""

    def __init__(self, weight, norm: dt.DataFidelityBase, upd_mask=None, dtype=np.float32):
        self.weight = np.array(weight)
        self.dtype = dtype
        self.op = None
        self.norm = norm
        self.upd_mask = upd_mask

    def info(self) -> str:
        return self.__reg_name__ + f"(w:{self.weight.max():.3e})"

    def upper(self) -> str:
        return self.__reg_name__.upper()

    def lower(self) -> str:
        return self.__reg_name__.lower()


class Regularizer_Grad(BaseRegularizer):
    """Gradient regularizer (``regularizers.py:218-289``)."""

    __reg_name__ = "grad"

    def __init__(self, weight, ndims: int = 2, axes=None, pad_mode: str = "edge", upd_mask=None,
                 norm: dt.DataFidelityBase | None = None):
        super().__init__(weight=weight, norm=dt.DataFidelity_l12() if norm is None else norm, upd_mask=upd_mask)
        if axes is None:
            axes = np.arange(-ndims, 0, dtype=int)
        elif not ndims == len(axes):
            print("WARNING - Number of axes different from number of dimensions. Updating dimensions accordingly.")
            ndims = len(axes)
        self.ndims = ndims
        self.axes = axes
        self.pad_mode = pad_mode.lower()
        self.sigma = 0.5

    def initialize_sigma_tau(self, primal) -> float:
        """Sets up the gradient operator; returns the tau contribution (``regularizers.py:260-273``)."""
        self.op = operators.TransformGradient(np.shape(primal), axes=self.axes, pad_mode=self.pad_mode)
        self.sigma = 0.5
        tau = 2 * self.ndims
        if not isinstance(self.norm, dt.DataFidelity_l1):
            tau *= self.weight
        return tau

    def check_fused(self) -> float:
        """The weight, after checking the fused TV2D kernels implement this configuration."""
        if self.ndims != 2 or list(np.asarray(self.axes) % 100) != list(np.array([-2, -1]) % 100):
            raise NotImplementedError("the fused HIP TV kernels implement 2-D TV over the last two axes")
        if self.pad_mode != "edge":
            raise NotImplementedError("the fused HIP TV kernels implement edge padding")
        if not isinstance(self.norm, dt.DataFidelity_l12) or getattr(self.norm, "l2_axis", 0) != 0:
            raise NotImplementedError("the fused HIP TV kernels implement the isotropic l12 norm")
        if self.upd_mask is not None:
            raise NotImplementedError("upd_mask is not supported by the fused HIP TV kernels")
        if np.asarray(self.weight).size != 1:
            raise NotImplementedError("the fused HIP TV kernels take a scalar weight")
        return float(self.weight)


class Regularizer_TV2D(Regularizer_Grad):
    """Total Variation in 2D (``regularizers.py:308-321``)."""

    __reg_name__ = "TV2D"

    def __init__(self, weight, axes=None, pad_mode: str = "edge", upd_mask=None, norm=None):
        super().__init__(weight=weight, ndims=2, axes=axes, pad_mode=pad_mode, norm=norm, upd_mask=upd_mask)


class Regularizer_l1(BaseRegularizer):
    """l1-norm regularizer: promotes sparse reconstructions (``regularizers.py:531-585``)."""

    __reg_name__ = "l1"

    def __init__(self, weight, upd_mask=None, norm=None):
        super().__init__(weight=weight, norm=dt.DataFidelity_l1() if norm is None else norm, upd_mask=upd_mask)

    def initialize_sigma_tau(self, primal) -> float:
        self.op = None  # identity transform
        return 1.0

    def primal_prox_weight(self) -> float:
        """The weight of the fused soft threshold (prox of tau * weight * ||.||_1)."""
        if self.upd_mask is not None:
            raise NotImplementedError("an update mask is not supported by the fused l1 proximal")
        if np.ndim(self.weight) != 0:
            raise NotImplementedError("the fused l1 proximal takes a scalar weight")
        return float(self.weight)
