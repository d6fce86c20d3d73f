"""CPU restatement of corrct's gradient-family regularizers (float32 numpy).

Test infrastructure only (see ``oracle/__init__.py``).

* ``grad_nd`` / ``div_nd`` -- ``TransformGradient`` over the last ``nd`` axes
  (``operators.py:796-844``) with its ``np.pad`` mode (edge by default: the
  last forward difference is 0, and so is the divergence's first backward
  difference), summed over the axes in order. ``-div_nd`` is the operator's
  adjoint for the constant mode only; for edge it drops ``-q[0]`` at the first
  sample.
* ``prox_l12`` / ``prox_l1`` -- ``DataFidelity_l12`` / ``DataFidelity_l1``
  dual proximals with a weight (``data_terms.py:549-555, 598-608``).
* ``prox_vtv`` -- ``Regularizer_VTV.apply_proximal_dual``
  (``regularizers.py:1394-1430``), channels on axis 1 of the dual.
* ``prox_tnv`` -- ``DataFidelity_ln.apply_proximal_dual`` with ``ln_axes=(1, 0)``
  and the l1 spectral norm (``data_terms.py:781-820``): numpy's SVD of each
  (channels x derivatives) matrix, singular values ``s / max(s, 1)``.
* ``L1Reg`` -- ``Regularizer_l1`` with its identity transform in PDHG / SIRT.
* ``GradReg`` -- ``Regularizer_Grad`` bookkeeping (``regularizers.py:134-215,
  260-273``): tau contribution ``2 nd`` (times the weight outside the l1
  family, times ``upd_mask`` when given), dual update ``q += 0.5 grad`` then
  the prox, primal contribution ``-div q`` (times ``upd_mask``, then times the
  weight outside the l1 family).
"""

from __future__ import annotations

import numpy as np

F32 = np.float32


def _axes(ndim: int, nd: int, axes) -> list:
    return [ndim - nd + k for k in range(nd)] if axes is None else [int(a) % ndim for a in axes]


def grad_nd(x: np.ndarray, nd: int, pad_mode: str = "edge", axes=None) -> np.ndarray:
    """``np.diff(np.pad(x, (0, 1) along the axis, mode=pad_mode))`` per axis (``operators.py:796-816``);
    ``axes`` in their given order (default: the last ``nd``)."""
    x = np.asarray(x)
    g = np.zeros((nd, *x.shape), dtype=x.dtype)
    for k, ax in enumerate(_axes(x.ndim, nd, axes)):
        padding = [(0, 0)] * x.ndim
        padding[ax] = (0, 1)
        g[k] = np.diff(np.pad(x, padding, mode=pad_mode), n=1, axis=ax)
    return g


def div_nd(q: np.ndarray, pad_mode: str = "edge", axes=None) -> np.ndarray:
    """``sum_k np.diff(np.pad(q[k], (1, 0) along the axis, mode=pad_mode))`` (``operators.py:818-835``)."""
    nd = q.shape[0]
    ndim = q.ndim - 1
    out = None
    for k, ax in enumerate(_axes(ndim, nd, axes)):
        padding = [(0, 0)] * ndim
        padding[ax] = (1, 0)
        d = np.diff(np.pad(q[k], padding, mode=pad_mode), n=1, axis=ax)
        out = d if out is None else out + d
    return out


def _w(w, dtype):
    """A weight as the dual's dtype: a scalar, or an array broadcast against the primal's trailing axes."""
    return np.asarray(w, dtype=dtype)


def prox_l12(q: np.ndarray, w) -> np.ndarray:
    nrm = np.sqrt(np.sum(q * q, axis=0, keepdims=True, dtype=q.dtype))
    w = _w(w, q.dtype)
    return (q / np.fmax(nrm, w)) * w


def prox_l1(q: np.ndarray, w) -> np.ndarray:
    w = _w(w, q.dtype)
    return (q / np.fmax(np.abs(q), w)) * w


def prox_vtv(q: np.ndarray, w: float, der=2, chan=np.inf, nd: int | None = None, q_ref=None) -> np.ndarray:
    """``q`` is [nd, C, ...] (or [nd, ...] for one channel when ``q.ndim == nd + 1``); ``q_ref``
    ([nd, ...], the reference volume's dual) joins the norms as one more channel after the C
    channels and is dropped afterwards (``regularizers.py:1410-1411, 1434-1435``)."""
    nd = q.shape[0] if nd is None else nd
    scalar = q.ndim == nd + 1
    t = q[:, None] if scalar else q
    if q_ref is not None:
        t = np.concatenate((t, np.asarray(q_ref, dtype=t.dtype)[:, None]), axis=1)
    if der == 1:
        gn = np.abs(t)
    elif der == 2:
        gn = np.linalg.norm(t, axis=0, ord=2, keepdims=True)
    else:
        gn = np.linalg.norm(t, axis=0, ord=1, keepdims=True)
    if chan == 1:
        dn = gn
    elif chan == 2:
        dn = np.linalg.norm(gn, axis=1, ord=2, keepdims=True)
    else:
        dn = np.linalg.norm(gn, axis=1, ord=1, keepdims=True)
    if q_ref is not None:  # the weight applies to the C channels; the reference channel is dropped
        t = t[:, : t.shape[1] - 1]
        if dn.shape[1] > 1:
            dn = dn[:, : dn.shape[1] - 1]
    w = _w(w, t.dtype)
    t = (t / np.fmax(dn, w)) * w
    return t[:, 0] if scalar else t


def prox_tnv(q: np.ndarray) -> np.ndarray:
    """``q`` is [nd, C, ...]: each (C x nd) matrix keeps its singular vectors, s -> s / max(s, 1)."""
    nd, c = q.shape[:2]
    m = np.moveaxis(q.reshape(nd, c, -1), -1, 0)  # [P, nd, C]
    m = np.swapaxes(m, 1, 2).astype(np.float64)  # [P, C, nd]
    u, s, vt = np.linalg.svd(m, full_matrices=False)
    s = s / np.fmax(s, 1.0)
    r = np.matmul(u, s[..., None] * vt)  # [P, C, nd]
    r = np.moveaxis(np.swapaxes(r, 1, 2), 0, -1).reshape(q.shape)
    return r.astype(q.dtype)


class GradReg:
    """One gradient-family regularizer: kind in {"l12", "l1", "vtv", "tnv", "l2", "hub", "hub_el"}.

    "l2" is ``Regularizer_smooth1D/2D/3D`` (``regularizers.py:386-431``): the l2 norm assigned
    sigma = 1/2 without data scales the dual by ``1 / (1 + 1/2)`` (``data_terms.py:294-304``).
    "hub" is ``Regularizer_HubTV2D/3D`` (``regularizers.py:340-383``): Huber over the gradient
    axis, the dual scaled by ``1 / (1 + huber_size / 2)`` (a Python float, rounded to float32 when
    numpy multiplies, ``data_terms.py:482-495``) then divided by ``max(1, |q|_2)``
    (``data_terms.py:500-501``); "hub_el" is the same norm without an l2 axis, each component divided by
    ``max(1, |q|)`` (``data_terms.py:497-498``). All three weight the primal update (outside the l1
    family)."""

    def __init__(self, nd: int, weight, kind: str = "l12", der=2, chan=np.inf, upd_mask=None, x_ref=None,
                 huber_size: float = 0.0, pad_mode: str = "edge", axes=None):
        self.pad_mode = pad_mode
        self.axes = axes
        # an array weight broadcasts against the primal (and the dual's trailing axes)
        weight = float(weight) if np.ndim(weight) == 0 else np.asarray(weight, dtype=F32)
        self.nd, self.weight, self.kind, self.der, self.chan = nd, weight, kind, der, chan
        self.huber_size = float(huber_size)
        self.upd_mask = None if upd_mask is None else np.asarray(upd_mask, dtype=F32)
        # VTV's reference dual: 0.5 * grad(x_ref) from a zero dual (regularizers.py:1386-1392)
        self.q_ref = None if x_ref is None else F32(0.5) * grad_nd(np.asarray(x_ref, dtype=F32), nd)

    @property
    def weighted(self) -> bool:
        return self.kind in ("tnv", "l2", "hub", "hub_el")  # outside the l1 family

    def tau(self):
        t = 2.0 * self.nd
        t = t * self.weight if self.weighted else t
        return t if self.upd_mask is None else t * self.upd_mask

    def dual(self, q: np.ndarray, x: np.ndarray) -> np.ndarray:
        q = q + F32(0.5) * grad_nd(x, self.nd, self.pad_mode, self.axes)
        if self.kind == "l12":
            return prox_l12(q, self.weight)
        if self.kind == "l1":
            return prox_l1(q, self.weight)
        if self.kind == "vtv":
            return prox_vtv(q, self.weight, self.der, self.chan, self.nd, self.q_ref)
        if self.kind == "l2":
            return (q * F32(1 / (1 + 0.5))).astype(F32)
        if self.kind in ("hub", "hub_el"):
            q = (q * F32(1.0 / (1.0 + 0.5 * self.huber_size))).astype(F32)
            nrm = np.sqrt(np.sum(q * q, axis=0, keepdims=True, dtype=F32)) if self.kind == "hub" else np.abs(q)
            return (q / np.fmax(F32(1), nrm)).astype(F32)
        return prox_tnv(q)

    def zeros(self, x_shape) -> np.ndarray:
        return np.zeros((self.nd, *x_shape), dtype=F32)

    def primal(self, q: np.ndarray) -> np.ndarray:
        u = -div_nd(q, self.pad_mode, self.axes)
        if self.upd_mask is not None:
            u = u * self.upd_mask
        return u * np.asarray(self.weight, dtype=F32) if self.weighted else u


class L1Reg:
    """``Regularizer_l1`` in PDHG / SIRT (``regularizers.py:531-557`` with the base class's steps,
    ``regularizers.py:134-210``): identity transform, ``q += x`` then the l1 prox with the weight
    (scalar or array), primal contribution ``q`` (times ``upd_mask``), tau contribution 1 (times
    ``upd_mask``)."""

    def __init__(self, weight, upd_mask=None):
        self.weight = float(weight) if np.ndim(weight) == 0 else np.asarray(weight, dtype=F32)
        self.upd_mask = None if upd_mask is None else np.asarray(upd_mask, dtype=F32)

    def zeros(self, x_shape) -> np.ndarray:
        return np.zeros(x_shape, dtype=F32)

    def tau(self):
        return 1.0 if self.upd_mask is None else F32(1.0) * self.upd_mask

    def dual(self, q: np.ndarray, x: np.ndarray) -> np.ndarray:
        return prox_l1((q + x).astype(F32), self.weight)

    def primal(self, q: np.ndarray) -> np.ndarray:
        return q * self.upd_mask if self.upd_mask is not None else q
