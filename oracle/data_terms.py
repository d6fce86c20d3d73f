"""CPU restatement of corrct's data-fidelity terms (float32 numpy): the PDHG data dual step and
the residual SIRT / FISTA take from the term.

Test infrastructure only (see ``oracle/__init__.py``).

``DataTerm(kind, local_error=None, weights=None, l2_axis=None)`` with ``kind`` one of

* ``"l2"``  -- ``DataFidelity_l2``    (``data_terms.py:283-331``)
* ``"wl2"`` -- ``DataFidelity_wl2``   (``data_terms.py:334-395``)
* ``"l2b"`` -- ``DataFidelity_l2b``   (``data_terms.py:398-464``)
* ``"hub"`` -- ``DataFidelity_Huber`` (``data_terms.py:467-532``)
* ``"l1"``  -- ``DataFidelity_l1``    (``data_terms.py:535-597``)
* ``"l12"`` -- ``DataFidelity_l12``   (``data_terms.py:598-625``)
* ``"l1b"`` -- ``DataFidelity_l1b``   (``data_terms.py:628-674``)
* ``"kl"``  -- ``DataFidelity_KL``    (``data_terms.py:677-757``)

``assign(b, sigma)`` is ``assign_data`` (``data_terms.py:120-135``), ``dual(p, ax)`` is
``update_dual`` + ``apply_proximal_dual`` (``data_terms.py:217-230`` and each class), and
``residual(ax, mask)`` is ``compute_residual`` (``data_terms.py:136-161`` and the overrides).
Every product and sum is float32, in the reference's order; the reference computes
``sigma * np.sqrt(local_error)`` in float64 when ``local_error`` is a Python float (a numpy
float64 scalar after ``np.sqrt``), which moves the l2b threshold by at most one float32 ulp.

Pinned in ``tests/test_oracle_data_terms.py`` by what the mathematics fixes, not by re-typing
these lines: each elementwise dual step is the proximal map of ``sigma f*`` for the term's convex
conjugate ``f*``, found there by a brute-force 1-D minimisation; the reductions (wl2 with unit
weights is l2, l2b with zero error is l2, l1b with zero error and Huber with zero size are l1,
l12 with groups of one is l1) and the group projection's ball property.
"""

from __future__ import annotations

import numpy as np

F32 = np.float32
EPS = np.finfo(np.float32).eps


def soft_threshold(v: np.ndarray, t) -> np.ndarray:
    """``sign(v) max(|v| - t, 0)`` (``data_terms.py:25-28``)."""
    return (np.sign(v) * np.fmax(np.abs(v) - t, F32(0))).astype(F32)


class DataTerm:
    def __init__(self, kind: str, local_error=None, weights=None, l2_axis=None):
        if kind not in ("l2", "wl2", "l2b", "hub", "l1", "l12", "l1b", "kl"):
            raise ValueError(kind)
        self.kind = kind
        self.e = None if local_error is None else np.asarray(local_error, dtype=F32)
        self.w = None if weights is None else np.asarray(weights, dtype=F32)
        self.l2_axis = 0 if (kind == "l12" and l2_axis is None) else l2_axis
        self.b = None
        self.sigma = F32(1)

    def assign(self, b, sigma=1.0) -> None:
        """``assign_data`` (``data_terms.py:120-135``) and each class's extra factors."""
        self.b = np.asarray(b, dtype=F32)
        self.sigma = np.asarray(sigma, dtype=F32)
        s = self.sigma
        if self.kind == "kl":
            self.sigma_data = (F32(4) * s * np.fmax(self.b, F32(0))).astype(F32)  # data_terms.py:682-686
        else:
            self.sigma_data = (s * self.b).astype(F32)  # data_terms.py:178-182
        if self.kind == "l2":
            self.sigma1 = (F32(1) / (F32(1) + s)).astype(F32)  # data_terms.py:294-296
        elif self.kind == "wl2":
            z = (self.w == 0).astype(F32)  # data_terms.py:347-354
            self.sigma1 = ((F32(1) / (F32(1) + s / (self.w + z))) * (F32(1) - z)).astype(F32)
        elif self.kind == "l2b":
            self.sigma_sqrt_error = (s * np.sqrt(self.e)).astype(F32)  # data_terms.py:414-419
            self.sigma1 = (F32(1) / (F32(1) + s * self.e)).astype(F32)
        elif self.kind == "hub":
            self.one_sigma_error = (F32(1) / (F32(1) + s * self.e)).astype(F32)  # data_terms.py:482-483

    def _group_norm(self, v: np.ndarray) -> np.ndarray:
        return np.sqrt(np.sum(v * v, axis=self.l2_axis, keepdims=True, dtype=F32)).astype(F32)

    def prox(self, v: np.ndarray) -> np.ndarray:
        """``apply_proximal_dual`` on ``v`` (after the dual update)."""
        v = np.asarray(v, dtype=F32)
        if self.kind == "kl":  # data_terms.py:688-692
            return ((F32(1) + v - np.sqrt((v - F32(1)) ** 2 + self.sigma_data)) / F32(2)).astype(F32)
        v = (v - self.sigma_data).astype(F32)
        if self.kind in ("l2", "wl2"):  # data_terms.py:301-304
            return (v * self.sigma1).astype(F32)
        if self.kind == "l2b":  # data_terms.py:425-429
            return (soft_threshold(v, self.sigma_sqrt_error) * self.sigma1).astype(F32)
        if self.kind == "hub":  # data_terms.py:491-501
            v = (v * self.one_sigma_error).astype(F32)
            nrm = np.abs(v) if self.l2_axis is None else self._group_norm(v)
            return (v / np.fmax(F32(1), nrm)).astype(F32)
        if self.kind == "l1b":  # data_terms.py:644-645
            v = soft_threshold(v, self.e)
        nrm = self._group_norm(v) if self.kind == "l12" else np.abs(v)  # data_terms.py:549-555, 607-608
        return ((v / np.fmax(nrm, F32(1))) * F32(1)).astype(F32)

    def dual(self, p: np.ndarray, ax: np.ndarray, bg=None) -> np.ndarray:
        """``update_dual`` (``data_terms.py:217-230``) then the proximal."""
        a = np.asarray(ax, dtype=F32) if bg is None else (np.asarray(ax, dtype=F32) + np.asarray(bg, dtype=F32))
        return self.prox((p + a * self.sigma).astype(F32))

    def residual(self, ax: np.ndarray, mask=None, bg=None) -> np.ndarray:
        """``compute_residual`` for the terms SIRT / FISTA accept (l2-based and KL)."""
        a = np.asarray(ax, dtype=F32)
        if bg is not None:
            a = (a + np.asarray(bg, dtype=F32)).astype(F32)
        if self.kind == "kl":  # data_terms.py:721-740, use_proximal=True
            r = self.prox((np.fmax(np.fmax(a, EPS), EPS) * self.sigma).astype(F32))
            if mask is not None:
                r = r * mask
            return (-r).astype(F32)
        r = (self.b - a).astype(F32)
        if self.kind == "wl2":  # data_terms.py:356-365
            r = (r * self.w).astype(F32)
        if mask is not None:
            r = (r * mask).astype(F32)
        if self.kind == "l2b":  # data_terms.py:419-423
            r = soft_threshold(r, self.sigma_sqrt_error)
        elif self.kind not in ("l2", "wl2"):
            raise ValueError(f"the solvers take compute_residual from l2-based terms and KL, not {self.kind}")
        return r
