"""CPU restatement of corrct's SIRT / MLEM / PDHG loops (float32 numpy).

Test infrastructure only (see ``oracle/__init__.py``).

Operators are plain callables ``A(x)`` and ``At(y)`` on numpy arrays, so the
same loops run over the oracle projector, a dense matrix, or (in tests) the
GPU projector's explicit matrix.

* ``diagonal_scaling`` -- ``compute_diagonal_scaling`` (``solvers.py:60-117``)
* ``grad`` / ``div``    -- ``TransformGradient`` edge mode (``operators.py:796-844``)
* ``tv_dual``           -- ``Regularizer_Grad.update_dual`` + l12 prox
                           (``regularizers.py:134-162, 260-273``; ``data_terms.py:549-555, 607-608``)
* ``sirt``              -- ``solvers.py:867-984``
* ``mlem``              -- ``solvers.py:688-805``
* ``sart``              -- ``solvers.py:473-631`` (per-angle callables; explicit row order)
* ``power_method``      -- ``solvers.py:26-57`` (explicit start vector)
* ``fista``             -- ``solvers.py:1348-1513`` (l2 data term, optional l1 prox)
* ``pdhg``              -- ``solvers.py:1060-1206`` (gradient-family regularizers
                           from ``oracle/regularizers.py`` via ``regs``) with ``DataFidelity_l2``
                           (``data_terms.py:283-331``) or ``DataFidelity_KL``
                           (``data_terms.py:677-692``), or any ``oracle.data_terms.DataTerm``

``sirt`` and ``fista`` take an optional ``oracle.data_terms.DataTerm`` whose ``compute_residual``
replaces the plain l2 residual (``solvers.py:951-975, 1470-1471``).
"""

from __future__ import annotations

import numpy as np

F32 = np.float32
EPS = np.finfo(np.float32).eps


def _fix(v: np.ndarray) -> np.ndarray:
    v = v.copy()
    v[(v / np.max(v)) < 1e-5] = 1
    return v


def diagonal_scaling(A, At, b, x_shape, tau_reg=0.0, rho_sigma=1.0, rho_tau=1.0, x_mask=None, b_mask=None):
    """``sigma = rho_s / fix(|A 1_x|)``, ``tau = rho_t / fix(|At 1_b| + tau_reg)``."""
    b_ones = np.ones_like(b, dtype=F32)
    if b_mask is not None:
        b_ones = b_ones * b_mask
    tau = np.abs(At(b_ones)).astype(F32)
    if np.ndim(tau_reg):  # update masks make the regularizers' contribution an array
        tau = tau + np.asarray(tau_reg, dtype=F32)
    elif tau_reg:
        tau = tau + F32(tau_reg)
    tau = _fix(tau)
    tau = (F32(rho_tau) / tau).astype(F32)
    x_ones = np.ones(x_shape, dtype=F32)
    if x_mask is not None:
        x_ones = x_ones * x_mask
    sigma = np.abs(A(x_ones)).astype(F32)
    sigma = _fix(sigma)
    sigma = (F32(rho_sigma) / sigma).astype(F32)
    return sigma, tau


def grad(x: np.ndarray) -> np.ndarray:
    """Forward differences along (rows, cols) with edge padding (last difference 0)."""
    g = np.zeros((2, *x.shape), dtype=x.dtype)
    g[0, :-1, :] = x[1:, :] - x[:-1, :]
    g[1, :, :-1] = x[:, 1:] - x[:, :-1]
    return g


def div(q: np.ndarray) -> np.ndarray:
    """Backward differences with edge padding: ``d[k] = q[k] - q[k-1]``, ``d[0] = 0``."""
    d0 = np.zeros(q.shape[1:], dtype=q.dtype)
    d1 = np.zeros(q.shape[1:], dtype=q.dtype)
    d0[1:, :] = q[0, 1:, :] - q[0, :-1, :]
    d1[:, 1:] = q[1, :, 1:] - q[1, :, :-1]
    return d0 + d1


def tv_dual(q: np.ndarray, xbar: np.ndarray, weight: float) -> np.ndarray:
    """``q += 0.5 grad(xbar)``; ``q <- w q / max(|q|_2, w)``."""
    q = q + F32(0.5) * grad(xbar)
    nrm = np.sqrt(np.sum(q * q, axis=0, keepdims=True, dtype=q.dtype))
    q = q / np.fmax(nrm, F32(weight))
    return q * F32(weight)


def sirt(A, At, b, x_shape, iterations, x0=None, lower=None, upper=None, x_mask=None, b_mask=None, tv_weight=None,
         relaxation=1.95, regs=(), data_term=None):
    """``regs``: ``oracle.regularizers.GradReg`` objects, duals re-initialised every iteration."""
    b = np.asarray(b, dtype=F32)
    tau_reg = (4.0 if tv_weight is not None else 0.0) + sum(r.tau() for r in regs)
    sigma, tau = diagonal_scaling(A, At, b, x_shape, tau_reg=tau_reg, rho_tau=relaxation, x_mask=x_mask, b_mask=b_mask)
    x = np.zeros(x_shape, dtype=F32) if x0 is None else np.array(x0, dtype=F32)
    if data_term is not None:
        data_term.assign(b, sigma)
    for _ in range(iterations):
        if data_term is not None:
            res = data_term.residual(A(x), b_mask)
        else:
            res = b - A(x)
            if b_mask is not None:
                res = res * b_mask
        upd = At(res * sigma)
        if tv_weight is not None:
            q = tv_dual(np.zeros((2, *x_shape), dtype=F32), x, tv_weight)
            upd = upd - (-div(q))
        for r in regs:
            upd = upd - r.primal(r.dual(r.zeros(x_shape), x))
        x = x + upd * tau
        if lower is not None or upper is not None:
            x = x.clip(lower, upper)
        if x_mask is not None:
            x = x * x_mask
    return x


def sart(A_angle, At_angle, b, x_shape, n_angles, iterations, rows, x0=None, lower=None, upper=None, x_mask=None,
         b_mask=None, relaxation=1.0):
    """SART: per-angle diagonal scalings, then for each angle a in ``rows``:
    x -= At_a((A_a x - b_a) * mask_a * sigma_a) * tau_a, clipped and masked (``solvers.py:565-629``).
    b is [..., A, N]; A_angle(x, a) -> [..., N]; At_angle(line, a) -> x-shaped."""
    b = np.asarray(b, dtype=F32)
    b_ones = np.ones_like(b)
    if b_mask is not None:
        b_ones = b_ones * b_mask
    tau = np.abs(np.stack([At_angle(b_ones[..., a, :], a) for a in range(n_angles)]))  # [A, *x_shape]
    tau[(tau / np.max(tau)) < 1e-5] = 1
    tau = (F32(relaxation) / tau).astype(F32)
    x_ones = np.ones(x_shape, dtype=F32)
    if x_mask is not None:
        x_ones = x_ones * x_mask
    sigma = np.abs(np.stack([A_angle(x_ones, a) for a in range(n_angles)], axis=-2))  # [..., A, N]
    sigma[(sigma / np.max(sigma)) < 1e-5] = 1
    sigma = (1 / sigma).astype(F32)
    x = np.zeros(x_shape, dtype=F32) if x0 is None else np.array(x0, dtype=F32)
    for _ in range(iterations):
        for a in rows:
            res = A_angle(x, a) - b[..., a, :]
            if b_mask is not None:
                res = res * b_mask[..., a, :]
            x = x - At_angle(res * sigma[..., a, :], a) * tau[a]
            if lower is not None:
                x = np.fmax(x, F32(lower))
            if upper is not None:
                x = np.fmin(x, F32(upper))
            if x_mask is not None:
                x = x * x_mask
    return x


def power_method(A, At, x_start, iterations: int = 5) -> float:
    """||A||_2 by power iterations on A^T A from the data-space vector x_start."""
    x = np.asarray(x_start, dtype=np.float64)
    x = x / np.linalg.norm(x)
    x = At(x)
    x_norm = np.linalg.norm(x)
    L = x_norm
    for _ in range(iterations):
        x = x / x_norm
        x = At(A(x))
        x_norm = np.linalg.norm(x)
        L = np.sqrt(x_norm)
    return float(L)


def fista(A, At, b, x_shape, iterations, precondition=True, L=None, relaxation=1.0, l1=0.0, lower=None, upper=None,
          x_mask=None, b_mask=None, restart_period=None, data_term=None, l1_mask=None):
    """FISTA: y-gradient step with diagonal (SIRT-style) or 1/L scaling, soft threshold at
    tau * l1 (``l1`` a scalar or an array), clip, mask, Beck-Teboulle momentum
    (``solvers.py:1391-1490``). ``l1_mask``: ``Regularizer_l1``'s update mask, thresholding
    ``x m`` and keeping ``x (1 - m)`` (``regularizers.py:578-583``)."""
    b = np.asarray(b, dtype=F32)
    if precondition:
        sigma, tau = diagonal_scaling(A, At, b, x_shape, rho_tau=relaxation, x_mask=x_mask, b_mask=b_mask)
    else:
        sigma, tau = F32(1.0 / L), F32(relaxation / L)
    x = np.zeros(x_shape, dtype=F32)
    if x_mask is not None:
        x = x * x_mask
    y = x.copy()
    t = 1.0
    if data_term is not None:
        data_term.assign(b)
    for ii in range(iterations):
        if data_term is not None:
            res = data_term.residual(A(y), b_mask)
        else:
            res = b - A(y)
            if b_mask is not None:
                res = res * b_mask
        xn = y + tau * At(res * sigma)
        if np.ndim(l1) or l1 > 0:
            thr = (tau * np.asarray(l1, dtype=F32)).astype(F32)
            if l1_mask is not None:
                xm = (xn * l1_mask).astype(F32)
                xm = np.sign(xm) * np.fmax(np.abs(xm) - thr, 0)
                xn = (xn * (F32(1) - l1_mask)).astype(F32) + xm
            else:
                xn = np.sign(xn) * np.fmax(np.abs(xn) - thr, 0)
        if lower is not None or upper is not None:
            xn = xn.clip(lower, upper)
        if x_mask is not None:
            xn = xn * x_mask
        t_new = (1.0 + np.sqrt(1.0 + 4.0 * t * t)) / 2.0
        if restart_period is not None and (ii + 1) % restart_period == 0:
            t_new = 1.0
        y = xn + F32((t - 1.0) / t_new) * (xn - x)
        x = xn
        t = t_new
    return x


def mlem(A, At, b, x_shape, iterations, x0=None, lower=None, upper=None, x_mask=None, b_mask=None, background=None):
    b = np.asarray(b, dtype=F32)
    b_ones = np.ones_like(b)
    if b_mask is not None:
        b_ones = b_ones * b_mask
    tau = At(b_ones).astype(F32)
    x = np.ones(x_shape, dtype=F32) if x0 is None else np.array(x0, dtype=F32)
    if x_mask is not None:
        x = x * x_mask
    for _ in range(iterations):
        ax = A(x)
        if background is not None:
            ax = ax + F32(background)
        ax = ax.clip(EPS, None)
        upd = At(b / ax)
        x = x * (upd / tau)
        if lower is not None or upper is not None:
            x = x.clip(lower, upper)
        if x_mask is not None:
            x = x * x_mask
    return x


def pdhg(A, At, b, x_shape, iterations, data_term="l2", tv_weight=None, x0=None, lower=None, upper=None,
         x_mask=None, b_mask=None, relaxation=0.95, regs=(), precondition=True, L=None, b_test_mask=None,
         tolerance=None, track=False):
    """``regs``: ``oracle.regularizers.GradReg`` objects beside (or instead of) the 2-D TV ``tv_weight``.

    ``precondition=False``: scalar steps sigma = rho / L, tau = rho / (L + tau_reg) from the operator
    norm ``L`` (``compute_Lipschitz_scaling``, ``solvers.py:120-158``). ``b_test_mask`` splits the
    data mask as ``_initialize_b_masks`` does (``solvers.py:241-249``). With ``track`` (l2 data term)
    returns ``(x, info)``: residual0 / residuals of ``||(b - A x) mask||^2`` per iteration and the
    same for the test mask (``solvers.py:1145-1158, 1192-1204``); ``tolerance`` stops the loop as
    the reference does."""
    b = np.asarray(b, dtype=F32)
    if b_test_mask is not None:
        bm = np.ones_like(b) if b_mask is None else np.asarray(b_mask, dtype=F32)
        b_mask, b_test_mask = bm * (F32(1) - b_test_mask), b_test_mask * bm
    tau_reg = (4.0 if tv_weight is not None else 0.0) + sum(r.tau() for r in regs)
    if precondition:
        sigma, tau = diagonal_scaling(
            A, At, b, x_shape, tau_reg=tau_reg, rho_sigma=relaxation, rho_tau=relaxation,
            x_mask=x_mask, b_mask=b_mask,
        )
    else:
        sigma = np.full(b.shape, F32(relaxation / L), dtype=F32)
        if np.ndim(tau_reg):
            tau = (F32(relaxation) / (F32(L) + np.asarray(tau_reg, dtype=F32))).astype(F32)
        else:
            tau = np.full(x_shape, F32(relaxation / (L + tau_reg)), dtype=F32)
    x = np.zeros(x_shape, dtype=F32) if x0 is None else np.array(x0, dtype=F32)
    xr = x.copy()
    term = None if isinstance(data_term, str) else data_term
    if term is not None:
        term.assign(b, sigma)
    elif data_term == "l2":
        sigma_data = sigma * b
        sigma1 = (F32(1) / (F32(1) + sigma)).astype(F32)
    elif data_term == "kl":
        sigma_data = F32(4) * sigma * np.fmax(b, F32(0))
    else:
        raise ValueError(data_term)
    p = np.zeros_like(b)
    q = np.zeros((2, *x_shape), dtype=F32) if tv_weight is not None else None
    qs = [r.zeros(x_shape) for r in regs]

    def res_norm(ax, mask):
        r = (b - ax).astype(F32)
        if mask is not None:
            r = r * mask
        return float(np.linalg.norm(r.ravel().astype(np.float64)) ** 2)

    info = {"residuals": [], "residuals_cv": [], "iterations": 0}
    if track:
        ax0 = A(x)
        info["residual0"] = res_norm(ax0, b_mask)
        if b_test_mask is not None:
            info["residual0_cv"] = res_norm(ax0, b_test_mask)
    for _ in range(iterations):
        info["iterations"] += 1
        if term is not None:
            p = term.dual(p, A(xr))
        else:
            p = p + A(xr) * sigma
            if data_term == "l2":
                p = (p - sigma_data) * sigma1
            else:
                p = (F32(1) + p - np.sqrt((p - F32(1)) ** 2 + sigma_data)) / F32(2)
        if b_mask is not None:
            p = p * b_mask
        qs = [r.dual(q_r, xr) for r, q_r in zip(regs, qs)]
        upd = At(p)
        if q is not None:
            q = tv_dual(q, xr, tv_weight)
            upd = upd + (-div(q))
        for r, q_r in zip(regs, qs):
            upd = upd + r.primal(q_r)
        x_new = x - upd * tau
        if lower is not None or upper is not None:
            x_new = x_new.clip(lower, upper)
        if x_mask is not None:
            x_new = x_new * x_mask
        xr = x_new + (x_new - x)
        x = x_new
        if track:
            ax = A(x)
            info["residuals"].append(res_norm(ax, b_mask))
            if b_test_mask is not None:
                info["residuals_cv"].append(res_norm(ax, b_test_mask))
            if tolerance is not None and tolerance > info["residuals"][-1]:
                break
    return (x, info) if track else x
