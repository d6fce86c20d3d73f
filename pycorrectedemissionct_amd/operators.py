"""Linear-operator protocol of the projectors (corrct ``operators.py`` mirror).

* ``BaseTransform``      -- ``operators.py:34-205``: a scipy ``LinearOperator``
  whose ``__call__`` applies the direct or adjoint operator depending on
  ``is_dir_operator``; ``.T`` is a shallow copy with the flag flipped;
  ``absolute()`` returns ``self``; ``explicit()`` builds the dense matrix.
* ``ProjectorOperator``  -- ``operators.py:283-362``: ``fp``/``bp`` interface,
  ``vol_shape``/``prj_shape`` aliases of the direct/adjoint shapes.
* ``TransformGradient``  -- ``operators.py:767-844``: forward differences with
  edge padding along the last ``ndims`` axes and its (negated divergence)
  adjoint. Inside the solvers the TV terms use the fused HIP kernels; this
  operator is the standalone API, evaluated on the device.

Arrays may be NumPy (results come back as NumPy) or torch device tensors
(results stay on the device).
"""

from __future__ import annotations

import copy as cp
from abc import ABC, abstractmethod

import numpy as np
import torch
from scipy.sparse.linalg import LinearOperator

from . import _arrays


class BaseTransform(LinearOperator, ABC):
    """Base operator class (``operators.py:34-205``)."""

    dir_shape: np.ndarray
    adj_shape: np.ndarray

    def __init__(self) -> None:
        num_cols = int(np.prod(self.dir_shape))
        num_rows = int(np.prod(self.adj_shape))
        super().__init__(np.float32, [num_rows, num_cols])
        self.is_dir_operator = True

    def _matvec(self, x):
        if self.is_dir_operator:
            x = np.asarray(x).reshape(self.dir_shape)
            return np.asarray(self._op_direct(x)).flatten()
        x = np.asarray(x).reshape(self.adj_shape)
        return np.asarray(self._op_adjoint(x)).flatten()

    def rmatvec(self, x):
        if self.is_dir_operator:
            x = np.asarray(x).reshape(self.adj_shape)
            return np.asarray(self._op_adjoint(x)).flatten()
        x = np.asarray(x).reshape(self.dir_shape)
        return np.asarray(self._op_direct(x)).flatten()

    def _transpose(self) -> "BaseTransform":
        op_t = cp.copy(self)
        op_t.shape = (self.shape[1], self.shape[0])
        op_t.is_dir_operator = not self.is_dir_operator
        return op_t

    def _adjoint(self) -> "BaseTransform":
        return self._transpose()

    @property
    def T(self) -> "BaseTransform":  # noqa: N802 (reference name)
        return self._transpose()

    def absolute(self) -> "BaseTransform":
        return self

    def explicit(self) -> np.ndarray:
        """Dense matrix, one unit vector per column (``operators.py:140-157``)."""
        he = np.empty(self.shape, dtype=self.dtype)
        dim_shape = self.dir_shape if self.is_dir_operator else self.adj_shape
        dim_size = int(np.prod(dim_shape))
        for ii in range(dim_size):
            xii = np.zeros((dim_size,), dtype=np.float32)
            xii[ii] = 1
            he[:, ii] = np.asarray(self(xii.reshape(dim_shape))).ravel()
        return he

    def __call__(self, x):
        if self.is_dir_operator:
            return self._op_direct(x)
        return self._op_adjoint(x)

    @abstractmethod
    def _op_direct(self, x): ...

    @abstractmethod
    def _op_adjoint(self, x): ...


class ProjectorOperator(BaseTransform):
    """Projector interface (``operators.py:283-362``)."""

    @property
    def vol_shape(self) -> np.ndarray:
        return self.dir_shape

    @vol_shape.setter
    def vol_shape(self, new_shape) -> None:
        self.dir_shape = np.array(new_shape, ndmin=1, dtype=int)

    @property
    def prj_shape(self) -> np.ndarray:
        return self.adj_shape

    @prj_shape.setter
    def prj_shape(self, new_shape) -> None:
        self.adj_shape = np.array(new_shape, ndmin=1, dtype=int)

    @abstractmethod
    def fp(self, x): ...

    @abstractmethod
    def bp(self, x): ...

    def _op_direct(self, x):
        return self.fp(x)

    def _op_adjoint(self, x):
        return self.bp(x)

    def get_pre_weights(self):
        return None


class TransformIdentity(BaseTransform):
    """Identity operator (``operators.py:365-384``)."""

    def __init__(self, x_shape):
        self.dir_shape = np.array(x_shape, ndmin=1, dtype=int)
        self.adj_shape = np.array(x_shape, ndmin=1, dtype=int)
        super().__init__()

    def _op_direct(self, x):
        return x

    def _op_adjoint(self, x):
        return x


class TransformGradient(BaseTransform):
    """Gradient with np.pad padding (edge, constant, wrap, reflect, symmetric) and its adjoint ``-div``
    (``operators.py:767-844``)."""

    def __init__(self, x_shape, axes=None, pad_mode: str = "edge"):
        x_shape = np.array(x_shape, ndmin=1, dtype=int)
        if axes is None:
            axes = np.arange(-len(x_shape), 0, dtype=int)
        self.axes = np.array(axes, ndmin=1, dtype=int)
        self.ndims = len(x_shape)
        self.pad_mode = pad_mode.lower()
        if self.pad_mode not in ("edge", "constant", "wrap", "reflect", "symmetric"):
            raise ValueError(f"Unsupported padding mode: {pad_mode}")
        self.dir_shape = x_shape
        self.adj_shape = np.array((len(self.axes), *self.dir_shape), ndmin=1, dtype=int)
        super().__init__()

    def gradient(self, x):
        t, host = _arrays.to_dev(x)
        out = []
        for ax in self.axes:
            ax = int(ax) % t.ndim
            d = torch.zeros_like(t)
            n = t.shape[ax]
            hi = t.narrow(ax, 1, n - 1)
            lo = t.narrow(ax, 0, n - 1)
            d.narrow(ax, 0, n - 1).copy_(hi - lo)
            # the last difference: the np.pad value past the end minus the last sample
            last = t.narrow(ax, n - 1, 1)
            if self.pad_mode == "constant":
                d.narrow(ax, n - 1, 1).copy_(-last)
            elif self.pad_mode == "wrap":
                d.narrow(ax, n - 1, 1).copy_(t.narrow(ax, 0, 1) - last)
            elif self.pad_mode == "reflect" and n > 1:
                d.narrow(ax, n - 1, 1).copy_(t.narrow(ax, n - 2, 1) - last)
            out.append(d)
        return _arrays.from_dev(torch.stack(out, dim=0), host)

    def divergence(self, x):
        t, host = _arrays.to_dev(x)
        acc = torch.zeros(t.shape[1:], dtype=t.dtype, device=t.device)
        for ii, ax in enumerate(self.axes):
            q = t[ii]
            ax = int(ax) % q.ndim
            n = q.shape[ax]
            d = torch.zeros_like(q)
            d.narrow(ax, 1, n - 1).copy_(q.narrow(ax, 1, n - 1) - q.narrow(ax, 0, n - 1))
            # the first difference: the first sample minus the np.pad value before it
            first = q.narrow(ax, 0, 1)
            if self.pad_mode == "constant":
                d.narrow(ax, 0, 1).copy_(first)
            elif self.pad_mode == "wrap":
                d.narrow(ax, 0, 1).copy_(first - q.narrow(ax, n - 1, 1))
            elif self.pad_mode == "reflect" and n > 1:
                d.narrow(ax, 0, 1).copy_(first - q.narrow(ax, 1, 1))
            acc = acc + d
        return _arrays.from_dev(acc, host)

    def _op_direct(self, x):
        return self.gradient(x)

    def _op_adjoint(self, x):
        d = self.divergence(x)
        return -d
