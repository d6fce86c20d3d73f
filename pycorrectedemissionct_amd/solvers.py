"""SIRT / MLEM / PDHG on the device (corrct ``solvers.py`` mirror).

Same classes, constructor arguments, ``__call__`` signature and return value
``(x, SolutionInfo)`` as ``solvers.py:161-1206``. All iterates live in device
memory; each iteration is one projector ``fp`` + one ``bp`` plus one or two
fused HIP kernels for the elementwise work (dual prox, TV dual, primal
update), so nothing returns to the host unless a tolerance or a test mask
asks for residual norms.

* ``compute_diagonal_scaling`` -- ``solvers.py:60-117``
* ``SIRT``  -- ``solvers.py:808-984``
* ``MLEM``  -- ``solvers.py:634-805``
* ``FBP``   -- ``solvers.py:346-470`` (filters in ``filters.py``)
* ``SART``  -- ``solvers.py:473-631``
* ``FISTA`` -- ``solvers.py:1209-1513`` (l2 data term, optional l1 prox;
  ``power_method`` / ``compute_Lipschitz_scaling``: ``solvers.py:26-158``)
* ``PDHG``  -- ``solvers.py:987-1206`` (preconditioned; l2 or KL data term,
  optional TV2D regularizer)

``b`` may be a NumPy array (the result comes back as NumPy) or a device
tensor (the result stays on the device). The operator ``A`` is any
``BaseTransform``; it receives and returns device tensors.
"""

from __future__ import annotations

import copy as cp
from collections.abc import Sequence

import numpy as np
import torch

from . import _arrays, _native
from . import data_terms, filters, regularizers

eps = np.finfo(np.float32).eps


def _dev(x, like: torch.Tensor | None = None):
    if x is None:
        return None
    t = _arrays.to_dev(x)[0]
    if like is not None and t.shape != like.shape:
        t = torch.broadcast_to(t, like.shape).contiguous()
    return t


def _fused_step(A, *arrays) -> bool:
    """Whether a solver iteration can run its projector work and data step as one on-the-fly
    ``fp_row_bp`` call: an on-the-fly XRF projector and contiguous, full-size data arrays."""
    ok = getattr(A, "supports_fused_data_step", None)
    if ok is None or not ok():
        return False
    ref = arrays[0]
    return all(t is None or (t.is_contiguous() and t.numel() == ref.numel()) for t in arrays)


def _residual(res, b, ax, bg, b_mask, asig, scale, term, st) -> None:
    """``res = data_term.compute_residual(ax, b_mask) * scale`` on the device, for the l2-based terms and
    KL (``asig``: the sigma the term was assigned, None = 1)."""
    code, prm, period, _, _ = term
    if code == data_terms.TERM_L2:
        _native.call("chip_sirt_residual", res.data_ptr(), b.data_ptr(), ax.data_ptr(), _native.ptr(bg),
                     _native.ptr(b_mask), _native.ptr(scale), b.numel(), st)
    else:
        _native.call("chip_data_residual", res.data_ptr(), b.data_ptr(), ax.data_ptr(), _native.ptr(bg),
                     _native.ptr(b_mask), _native.ptr(asig), _native.ptr(scale), b.numel(), code, _native.ptr(prm),
                     period, st)


class _Limits:
    """Clipping limits of a solve (``x.clip(lower_limit, upper_limit)``, ``solvers.py:797-803,
    975-982, 1185-1190``). Scalars go into the fused update kernels; with an array-valued limit
    (broadcast to x) the fused kernel runs without limits and mask and ``chip_clip_relax`` then
    clips, masks and (PDHG, FISTA) extrapolates, in the reference's order."""

    def __init__(self, lower, upper, like: torch.Tensor):
        arr_lo = lower is not None and np.ndim(lower) != 0
        arr_hi = upper is not None and np.ndim(upper) != 0
        self.fused = not (arr_lo or arr_hi)
        self.hl, self.lo = (0, 0.0) if (lower is None or arr_lo) else (1, float(lower))
        self.hu, self.hi = (0, 0.0) if (upper is None or arr_hi) else (1, float(upper))
        self.lo_t = _dev(lower, like).contiguous() if arr_lo else None
        self.hi_t = _dev(upper, like).contiguous() if arr_hi else None

    def kernel(self):
        """(has_lower, lower, has_upper, upper) for a fused update kernel."""
        return (self.hl, self.lo, self.hu, self.hi) if self.fused else (0, 0.0, 0, 0.0)

    def kernel_mask(self, mask):
        """The mask a fused update kernel applies (none when chip_clip_relax does)."""
        return mask if self.fused else None

    def finish(self, x: torch.Tensor, mask, xbar=None, x_old=None, momentum: float = 0.0) -> None:
        if self.fused:
            return
        _native.call("chip_clip_relax", x.data_ptr(), _native.ptr(self.lo_t), _native.ptr(self.hi_t), self.hl,
                     self.lo, self.hu, self.hi, _native.ptr(mask), _native.ptr(xbar), _native.ptr(x_old),
                     float(momentum), x.numel(), _native.stream())


def compute_diagonal_scaling(A_abs, At_abs, b: torch.Tensor, regs: Sequence, relaxation_sigma: float = 1.0,
                             relaxation_tau: float = 1.0, x_mask=None, b_mask=None):
    """``sigma = rho_s / fix(|A 1|)``, ``tau = rho_t / fix(|A^T 1| + sum reg_tau)`` (``solvers.py:60-117``)."""
    st = _native.stream()
    b_ones = torch.ones_like(b)
    if b_mask is not None:
        b_ones = b_ones * b_mask
    tau = _arrays.to_dev(At_abs(b_ones))[0].clone()
    tau_reg = 0.0
    for reg in regs:
        t_reg, mask = _reg_tau(reg, tau)
        if mask is None:
            tau_reg += t_reg
        else:  # tau += t_reg * upd_mask (regularizers.py:271-272)
            _native.call("chip_axpy", tau.data_ptr(), t_reg, mask.data_ptr(), mask.numel(), tau.numel(), st)
    mx = torch.empty(1, dtype=torch.float32, device=tau.device)
    _native.call("chip_absmax", tau.data_ptr(), tau.numel(), mx.data_ptr(), st)
    _native.call("chip_precond_fix", tau.data_ptr(), tau.numel(), float(tau_reg), mx.data_ptr(), float(relaxation_tau), st)
    x_ones = torch.ones_like(tau)
    if x_mask is not None:
        x_ones = x_ones * x_mask
    sigma = _arrays.to_dev(A_abs(x_ones))[0].clone()
    _native.call("chip_absmax", sigma.data_ptr(), sigma.numel(), mx.data_ptr(), st)
    _native.call("chip_precond_fix", sigma.data_ptr(), sigma.numel(), 0.0, mx.data_ptr(), float(relaxation_sigma), st)
    return sigma, tau, tuple(tau.shape), tau.dtype


def power_method(A, b, iterations: int = 5):
    """l2-norm of A by the power method (``solvers.py:26-57``); the start vector is
    ``np.random.rand`` of b's shape, as in the reference. Returns (L, x_shape, x_dtype)."""
    b_t = _arrays.to_dev(b)[0]
    x = torch.as_tensor(np.random.rand(*b_t.shape).astype(np.float32), device=b_t.device)
    x = x / torch.linalg.vector_norm(x)
    x = _arrays.to_dev(A.T(x))[0]
    x_norm = torch.linalg.vector_norm(x)
    L = x_norm
    for _ in range(iterations):
        x = x / x_norm
        x = _arrays.to_dev(A.T(_arrays.to_dev(A(x))[0]))[0]
        x_norm = torch.linalg.vector_norm(x)
        L = torch.sqrt(x_norm)
    return float(L), tuple(x.shape), np.float32


def _reg_tau(reg, primal):
    """(scalar, flat device update mask or None) of a regularizer's tau contribution; sets the
    regularizer up as ``initialize_sigma_tau`` does."""
    out = reg.initialize_sigma_tau(primal)
    if hasattr(reg, "tau_terms"):
        return reg.tau_terms(tuple(np.shape(primal)))
    return float(out), None


def compute_Lipschitz_scaling(A, b, regs: Sequence, relaxation_sigma: float = 1.0, relaxation_tau: float = 1.0):
    """``sigma = rho_s / L``, ``tau = rho_t / (L + sum reg_tau)`` (``solvers.py:120-158``); tau is a
    float, or a float32 device tensor of the primal's shape when a regularizer has an update mask."""
    L, x_shape, x_dtype = power_method(A, b)
    tau = L
    dummy_x = np.empty(x_shape, dtype=x_dtype)
    masked = []
    for reg in regs:
        t_reg, mask = _reg_tau(reg, dummy_x)
        if mask is None:
            tau += t_reg
        else:
            masked.append((t_reg, mask))
    if not masked:
        return relaxation_sigma / L, relaxation_tau / tau, x_shape, x_dtype
    st = _native.stream()
    tau_t = torch.full(x_shape, float(tau), dtype=torch.float32, device=_arrays.device())
    for t_reg, mask in masked:
        _native.call("chip_axpy", tau_t.data_ptr(), t_reg, mask.data_ptr(), mask.numel(), tau_t.numel(), st)
    _native.call("chip_rdiv", tau_t.data_ptr(), float(relaxation_tau), tau_t.numel(), st)
    return relaxation_sigma / L, tau_t, x_shape, x_dtype


class SolutionInfo:
    """Reconstruction info (``solvers.py:161-201``)."""

    def __init__(self, method: str, max_iterations: int, tolerance, residual0: float = np.inf,
                 residual0_cv: float = np.inf) -> None:
        self.method = method
        self.max_iterations = max_iterations
        self.tolerance = tolerance
        self.residual0 = residual0
        self.residual0_cv = residual0_cv
        self.residuals = np.zeros(max_iterations)
        self.residuals_cv = np.zeros(max_iterations)
        self.iterations = 0

    @property
    def residuals_rel(self) -> np.ndarray:
        return self.residuals / self.residual0

    @property
    def residuals_cv_rel(self) -> np.ndarray:
        return self.residuals_cv / self.residual0_cv


class Solver:
    """Base solver (``solvers.py:204-343``)."""

    def __init__(self, verbose: bool = False, leave_progress: bool = True, relaxation: float = 1.0, tolerance=None,
                 data_term="l2", data_term_test=None):
        self.verbose = verbose
        self.leave_progress = leave_progress
        self.relaxation = relaxation
        self.tolerance = tolerance
        self.data_term = self._initialize_data_fidelity_function(data_term)
        if data_term_test is None:
            data_term_test = self.data_term
        else:
            data_term_test = self._initialize_data_fidelity_function(data_term_test)
        self.data_term_test = cp.deepcopy(data_term_test)

    def info(self) -> str:
        return type(self).__name__

    def upper(self) -> str:
        return type(self).__name__.upper()

    def lower(self) -> str:
        return type(self).__name__.lower()

    @staticmethod
    def _initialize_data_fidelity_function(data_term):
        if isinstance(data_term, str):
            if data_term.lower() == "l2":
                return data_terms.DataFidelity_l2()
            if data_term.lower() == "kl":
                return data_terms.DataFidelity_KL()
            raise ValueError(f"Unknown data term: '{data_term}', only accepted terms are: 'l2' | 'kl'.")
        if isinstance(data_term, (data_terms.DataFidelity_l2, data_terms.DataFidelity_KL)):
            return cp.deepcopy(data_term)
        raise ValueError(f"Unsupported data term: '{data_term.info()}', only accepted terms are 'kl' and 'l2'-based.")

    @staticmethod
    def _initialize_regularizer(regularizer):
        if regularizer is None:
            return []
        if isinstance(regularizer, regularizers.BaseRegularizer):
            return [regularizer]
        if isinstance(regularizer, (list, tuple)):
            if not all(isinstance(r, regularizers.BaseRegularizer) for r in regularizer):
                raise ValueError("Some regularizers are not derived from the regularizers.BaseRegularizer class")
            return list(regularizer)
        raise ValueError("Unknown regularizer type.")

    @staticmethod
    def _initialize_b_masks(b, b_mask, b_test_mask):
        if b_test_mask is not None:
            if b_mask is None:
                b_mask = torch.ones_like(b)
            b_mask, b_test_mask = b_mask * (1 - b_test_mask), b_test_mask * b_mask
        return b_mask, b_test_mask

    def _tv_weight(self) -> float | None:
        """The weight of a single 2-D TV regularizer that the fused TV2D kernels implement, else None."""
        if len(self.regularizer) != 1:
            return None
        reg = self.regularizer[0]
        if type(reg) not in (regularizers.Regularizer_Grad, regularizers.Regularizer_TV2D):
            return None
        try:
            return reg.check_fused()
        except NotImplementedError:
            return None

    def _grad_regs(self, x: torch.Tensor) -> list:
        """(kernel parameters, zero dual) of every regularizer for the HIP gradient kernels."""
        out = []
        for reg in self.regularizer:
            if not isinstance(reg, (regularizers.Regularizer_Grad, regularizers.Regularizer_l1)):
                raise NotImplementedError(f"regularizer {reg.info()} is not supported by the HIP kernels")
            cfg = reg.device_config(tuple(x.shape))
            out.append((cfg, torch.zeros(cfg["q_shape"], dtype=torch.float32, device=x.device)))
        return out

    @staticmethod
    def _grad_dual(cfg: dict, q: torch.Tensor, x: torch.Tensor, st) -> None:
        if cfg.get("identity"):
            _native.call("chip_l1_dual", q.data_ptr(), x.data_ptr(), cfg["n"], cfg["weight"],
                         _native.ptr(cfg["weights"]), cfg["weights_period"], st)
            return
        strides = cfg.get("strides")
        _native.call("chip_grad_dual", q.data_ptr(), x.data_ptr(), cfg["n_chan"], cfg["n_batch"], cfg["nd"],
                     cfg["dims"].ctypes.data, None if strides is None else strides.ctypes.data, cfg.get("pad", 0),
                     cfg["mode"], cfg["weight"],
                     _native.ptr(cfg.get("weights")), cfg.get("weights_period", 1), cfg["der"], cfg["chan"],
                     _native.ptr(cfg.get("qref")), st)

    @staticmethod
    def _grad_primal(cfg: dict, upd: torch.Tensor, q: torch.Tensor, subtract: bool, st) -> None:
        if cfg.get("identity"):
            _native.call("chip_l1_update_primal", upd.data_ptr(), q.data_ptr(), cfg["n"], _native.ptr(cfg["mask"]),
                         cfg["mask_period"], int(subtract), st)
            return
        strides = cfg.get("strides")
        _native.call("chip_grad_update_primal", upd.data_ptr(), q.data_ptr(), cfg["n_chan"], cfg["n_batch"],
                     cfg["nd"], cfg["dims"].ctypes.data, None if strides is None else strides.ctypes.data,
                     cfg.get("pad", 0), cfg["scale"], _native.ptr(cfg.get("scales")),
                     cfg.get("scales_period", 1), cfg["apply_scale"], int(subtract), _native.ptr(cfg.get("mask")),
                     cfg.get("mask_period", 1), st)

    def _track(self, info, ii, Ax, b, b_mask, b_test_mask) -> bool:
        """Residual bookkeeping of one iteration; True when the tolerance is reached."""
        res = self.data_term.compute_residual(Ax, mask=b_mask)
        info.residuals[ii] = self.data_term.compute_residual_norm(res)
        if b_test_mask is not None:
            res_test = self.data_term_test.compute_residual(Ax, mask=b_test_mask)
            info.residuals_cv[ii] = self.data_term_test.compute_residual_norm(res_test)
        if self.tolerance is not None and self.tolerance > info.residuals[ii]:
            if self.verbose:
                print(f"Residual reached the desired tolerance of {self.tolerance}. Ending iterations..")
            return True
        return False

    def _init_residuals(self, info, A, x, b, b_mask, b_test_mask, sigma=None):
        Ax = A(x)
        res_0 = self.data_term.compute_residual(Ax, mask=b_mask)
        info.residual0 = self.data_term.compute_residual_norm(res_0)
        if b_test_mask is not None:
            if self.data_term_test.background != self.data_term.background:
                print("WARNING - the data_term and and data_term_test should have the same background. Making them equal.")
                self.data_term_test.background = self.data_term.background
            self.data_term_test.assign_data(b, sigma if sigma is not None else 1.0)
            res_test_0 = self.data_term_test.compute_residual(Ax, mask=b_test_mask)
            info.residual0_cv = self.data_term_test.compute_residual_norm(res_test_0)


class FBP(Solver):
    """Filtered back-projection (``solvers.py:346-470``): ``A.T(filter(b))``, clipped and masked."""

    def __init__(self, verbose: bool = False, leave_progress: bool = False, regularizer=None, data_term="l2",
                 fbp_filter="ramp", pad_mode: str = "constant"):
        super().__init__(verbose=verbose)
        if isinstance(fbp_filter, str):
            fbp_filter = fbp_filter.lower()
        self.fbp_filter = fbp_filter
        self.pad_mode = pad_mode

    def info(self) -> str:
        if isinstance(self.fbp_filter, str):
            return super().info() + "(F:" + self.fbp_filter.upper() + ")"
        if isinstance(self.fbp_filter, np.ndarray):
            return super().info() + "(F:" + filters.FilterCustom.__name__.upper() + ")"
        return super().info() + "(F:" + type(self.fbp_filter).__name__.upper() + ")"

    def __call__(self, A, b, iterations: int = 0, x0=None, lower_limit=None, upper_limit=None, x_mask=None,
                 b_mask=None):
        if len(np.shape(b)) < 2:
            raise ValueError(f"Data should be at least 2-dimensional (b.shape = {np.shape(b)})")
        info = SolutionInfo(self.info(), max_iterations=0, tolerance=0.0)
        if isinstance(self.fbp_filter, str):
            if self.fbp_filter in ("mr", "data"):
                raise NotImplementedError("the data-driven (MR) FBP filter is not supported")
            local_filter = filters.FilterFBP(filter_name=self.fbp_filter, pad_mode=self.pad_mode)
        elif isinstance(self.fbp_filter, np.ndarray):
            local_filter = filters.FilterCustom(self.fbp_filter, pad_mode=self.pad_mode)
        else:
            local_filter = self.fbp_filter
            local_filter.pad_mode = self.pad_mode
        b_t, host = _arrays.to_dev(b)
        pre_weights = A.get_pre_weights() if hasattr(A, "get_pre_weights") else None
        if pre_weights is not None:
            b_t = b_t * _arrays.to_dev(pre_weights)[0]
        b_f = local_filter(b_t)
        x = _arrays.to_dev(A.T(b_f))[0]
        if lower_limit is not None or upper_limit is not None:
            x = x.clip(lower_limit, upper_limit)
        if x_mask is not None:
            x = x * _arrays.to_dev(x_mask)[0]
        return _arrays.from_dev(x, host), info


class SART(Solver):
    """SART (``solvers.py:473-631``): one angle at a time, on the per-angle projector entry points.

    ``rows_sequence`` (an extension) fixes the angle order; by default it is
    ``np.random.permutation(A_num_rows)`` drawn once per call, as in the reference.
    """

    def compute_residual(self, A, b, x, A_num_rows: int, b_mask):
        """Stack of per-angle residuals A_a x - b_a (``solvers.py:476-510``)."""
        fp = torch.stack([_arrays.to_dev(A(x, ii))[0] for ii in range(A_num_rows)], dim=-2)
        res = fp - _arrays.to_dev(b)[0]
        if b_mask is not None:
            res = res * _arrays.to_dev(b_mask)[0]
        return res

    def __call__(self, A, b, iterations: int, A_num_rows: int | None = None, At=None, x0=None, lower_limit=None,
                 upper_limit=None, x_mask=None, b_mask=None, rows_sequence=None):
        from . import projectors  # local: projectors imports this module's siblings

        if isinstance(A, projectors.ProjectorUncorrected):
            p = A
            if not p.projector_backend.has_individual_projs:
                raise ValueError("The projector needs to have enabled single projections.")
            A = lambda x, ii: p.fp_angle(x, ii)  # noqa: E731
            if isinstance(p, projectors.ProjectorAttenuationXRF):
                At = lambda y, ii: p.bp_angle(y, ii, single_line=True)  # noqa: E731
            else:
                At = lambda y, ii: p.bp_angle(y, ii)  # noqa: E731
            A_num_rows = len(p.angles_rot_rad)
        elif At is None:
            raise ValueError("Parameter `At` is required, if `A` is not a projector.")
        elif A_num_rows is None:
            raise ValueError("Parameter `A_num_rows` is required, if `A` is not a projector.")
        st = _native.stream()
        b_t, host = _arrays.to_dev(b)
        b_t = b_t.contiguous()
        # per-angle rows as contiguous [A, ..., N] blocks
        b_rows = b_t.movedim(-2, 0).contiguous()
        mask_rows = None if b_mask is None else _arrays.to_dev(b_mask)[0].movedim(-2, 0).contiguous()
        ones = torch.ones_like(b_rows[0])
        tau = torch.stack([_arrays.to_dev(At(ones if mask_rows is None else ones * mask_rows[ii], ii))[0]
                           for ii in range(A_num_rows)]).contiguous()
        mx = torch.empty(1, dtype=torch.float32, device=tau.device)
        _native.call("chip_absmax", tau.data_ptr(), tau.numel(), mx.data_ptr(), st)
        _native.call("chip_precond_fix", tau.data_ptr(), tau.numel(), 0.0, mx.data_ptr(), float(self.relaxation), st)
        x_shape = tuple(tau.shape[1:])
        x_ones = torch.ones(x_shape, dtype=torch.float32, device=tau.device)
        x_mask_t = None
        if x_mask is not None:
            x_mask_t = _arrays.to_dev(x_mask)[0].contiguous()
            x_ones = x_ones * x_mask_t
        sigma = torch.stack([_arrays.to_dev(A(x_ones, ii))[0] for ii in range(A_num_rows)]).contiguous()
        _native.call("chip_absmax", sigma.data_ptr(), sigma.numel(), mx.data_ptr(), st)
        _native.call("chip_precond_fix", sigma.data_ptr(), sigma.numel(), 0.0, mx.data_ptr(), 1.0, st)
        x = torch.zeros(x_shape, dtype=torch.float32, device=tau.device) if x0 is None else \
            _arrays.to_dev(x0)[0].clone().contiguous()
        info = SolutionInfo(self.info(), max_iterations=iterations, tolerance=self.tolerance)
        if self.tolerance is not None:
            info.residual0 = float(torch.linalg.vector_norm(self.compute_residual(A, b_t, x, A_num_rows, b_mask)))
        rows = np.random.permutation(A_num_rows) if rows_sequence is None else np.asarray(rows_sequence)
        lim = _Limits(lower_limit, upper_limit, x)
        hl, lo, hu, hi = lim.kernel()
        n = int(x.shape[-1])
        batch = x.numel() // (n * n)
        res = torch.empty_like(b_rows[0])
        for ii in range(iterations):
            info.iterations += 1
            for a in rows:
                a = int(a)
                line = _arrays.to_dev(A(x, a))[0].contiguous()
                # r = (b_a - A_a x) * mask_a * sigma_a = -(A_a x - b_a) * mask_a * sigma_a
                _native.call("chip_sirt_residual", res.data_ptr(), b_rows[a].data_ptr(), line.data_ptr(), None,
                             _native.ptr(None if mask_rows is None else mask_rows[a]), sigma[a].data_ptr(),
                             res.numel(), st)
                upd = _arrays.to_dev(At(res, a))[0].contiguous()
                # x = clip(x - At_a(res) tau_a) * mask, with res of the opposite sign
                _native.call("chip_sirt_update", x.data_ptr(), upd.data_ptr(), None, tau[a].data_ptr(), hl, lo, hu,
                             hi, _native.ptr(lim.kernel_mask(x_mask_t)), n, batch, st)
                lim.finish(x, x_mask_t)
            if self.tolerance is not None:
                info.residuals[ii] = float(torch.linalg.vector_norm(
                    self.compute_residual(A, b_t, x, A_num_rows, b_mask)))
                if self.tolerance > info.residuals[ii]:
                    break
        return _arrays.from_dev(x, host), info


class FISTA(Solver):
    """FISTA (``solvers.py:1209-1513``): accelerated proximal gradient with an l2 data term
    and at most one regularizer with a primal proximal (``Regularizer_l1``); the gradient step,
    proximal, clipping, mask and momentum run fused in ``chip_fista_update``."""

    def __init__(self, verbose: bool = False, leave_progress: bool = True, tolerance=None, relaxation: float = 1.0,
                 regularizer=None, data_term="l2", data_term_test=None, restart_period: int | None = None):
        super().__init__(verbose=verbose, leave_progress=leave_progress, relaxation=relaxation, tolerance=tolerance,
                         data_term=data_term, data_term_test=data_term_test)
        regs = self._initialize_regularizer(regularizer)
        if len(regs) > 1:
            raise ValueError(
                f"FISTA supports at most one regularizer, but {len(regs)} were provided. "
                "To combine multiple regularizers use PDHG instead."
            )
        for reg in regs:
            if not isinstance(reg, regularizers.Regularizer_l1):
                raise ValueError(
                    f"Regularizer '{reg.info()}' does not support a primal proximal and cannot be used with FISTA. "
                    "Switch to PDHG for this regularizer."
                )
        self.regularizer = list(regs)
        self.restart_period = restart_period

    def info(self) -> str:
        reg_info = "".join(["-" + r.info().upper() for r in self.regularizer])
        return Solver.info(self) + "-" + self.data_term.info() + reg_info

    @staticmethod
    def _initialize_data_fidelity_function(data_term):
        if isinstance(data_term, str):
            if data_term.lower() == "l2":
                return data_terms.DataFidelity_l2()
            raise ValueError(f"Unknown data term: '{data_term}'. FISTA only accepts 'l2' and derivatives.")
        if isinstance(data_term, data_terms.DataFidelity_l2):
            return cp.deepcopy(data_term)
        raise ValueError(f"Unsupported data term type: {type(data_term)}. FISTA only accepts 'l2' and derivatives.")

    def __call__(self, A, b, iterations: int, x0=None, lower_limit=None, upper_limit=None, x_mask=None, b_mask=None,
                 b_test_mask=None, precondition: bool = True):
        st = _native.stream()
        b_t, host = _arrays.to_dev(b)
        b_t = b_t.clone()
        b_mask = _dev(b_mask, b_t)
        b_test_mask = _dev(b_test_mask, b_t)
        if precondition:
            try:
                At_abs = A.T.absolute()
                A_abs = A.absolute()
            except AttributeError:
                print("WARNING: operator does not support absolute(); falling back to un-preconditioned Lipschitz "
                      "step size.")
                precondition = False
        b_mask, b_test_mask = self._initialize_b_masks(b_t, b_mask, b_test_mask)
        x_mask_t = None if x_mask is None else _arrays.to_dev(x_mask)[0]
        if precondition:
            sigma, tau, x_shape, _ = compute_diagonal_scaling(A_abs, At_abs, b_t, regs=[],
                                                              relaxation_tau=self.relaxation, x_mask=x_mask_t,
                                                              b_mask=b_mask)
            tau_ptr, tau_s = tau.data_ptr(), 0.0
            sigma_t = sigma
        else:
            sig, tau_v, x_shape, _ = compute_Lipschitz_scaling(A, b_t, regs=[], relaxation_tau=self.relaxation)
            tau_ptr, tau_s = None, float(tau_v)
            sigma_t = torch.full_like(b_t, float(sig))
        prox = dict(l1w=0.0, l1wa=None, l1w_period=1, l1m=None, l1m_period=1)
        if self.regularizer:
            self.regularizer[0].initialize_sigma_tau(np.zeros(x_shape, np.float32))
            prox = self.regularizer[0].fista_prox(x_shape)
        self.data_term.assign_data(b_t)
        bg = self.data_term.background_dev(b_t)
        term = self.data_term.device_term(b_t)
        dev = b_t.device
        x = torch.zeros(x_shape, dtype=torch.float32, device=dev) if x0 is None else _dev(x0, b_t).clone()
        x = x.reshape(x_shape).contiguous()
        x_mask_t = _dev(x_mask_t, x)
        if x_mask_t is not None:
            x = x * x_mask_t
        y = x.clone()
        t = 1.0
        lim = _Limits(lower_limit, upper_limit, x)
        hl, lo, hu, hi = lim.kernel()
        x_old = None if lim.fused else torch.empty_like(x)
        info = SolutionInfo(self.info(), max_iterations=iterations, tolerance=self.tolerance)
        track = b_test_mask is not None or self.tolerance is not None
        if track:
            self._init_residuals(info, A, x, b_t, b_mask, b_test_mask, sigma_t)
        res = torch.empty_like(b_t)
        for ii in range(iterations):
            info.iterations += 1
            Ay = _arrays.to_dev(A(y))[0]
            # residual = compute_residual(A y) * sigma, the data term assigned with sigma = 1
            _residual(res, b_t, Ay, bg, b_mask, None, sigma_t, term, st)
            grad = _arrays.to_dev(A.T(res))[0].contiguous()
            t_new = float(1.0 + np.sqrt(1.0 + 4.0 * t * t)) / 2.0
            if self.restart_period is not None and (ii + 1) % self.restart_period == 0:
                t_new = 1.0
            momentum = (t - 1.0) / t_new
            if x_old is not None:
                x_old.copy_(x)
            _native.call("chip_fista_update", x.data_ptr(), y.data_ptr(), grad.data_ptr(), tau_ptr, tau_s,
                         prox["l1w"], _native.ptr(prox["l1wa"]), prox["l1w_period"], _native.ptr(prox["l1m"]),
                         prox["l1m_period"], hl, lo, hu, hi, _native.ptr(lim.kernel_mask(x_mask_t)), momentum,
                         x.numel(), st)
            lim.finish(x, x_mask_t, y, x_old, momentum)
            t = t_new
            if track:
                Ax = _arrays.to_dev(A(x))[0]
                if self._track(info, ii, Ax, b_t, b_mask, b_test_mask):
                    break
        return _arrays.from_dev(x, host), info


class MLEM(Solver):
    """MLEM (``solvers.py:634-805``)."""

    def __init__(self, verbose: bool = False, leave_progress: bool = True, tolerance=None, regularizer=None,
                 data_term="kl", data_term_test=None):
        super().__init__(verbose=verbose, leave_progress=leave_progress, tolerance=tolerance, data_term=data_term,
                         data_term_test=data_term_test)
        self.regularizer = self._initialize_regularizer(regularizer)

    def info(self) -> str:
        return Solver.info(self) + f"(B:{self.data_term.background:g})" if self.data_term.background is not None else ""

    def __call__(self, A, b, iterations: int, x0=None, lower_limit=None, upper_limit=None, x_mask=None, b_mask=None,
                 b_test_mask=None):
        b_t, host = _arrays.to_dev(b)
        b_t = b_t.clone()
        b_mask = _dev(b_mask, b_t)
        b_test_mask = _dev(b_test_mask, b_t)
        b_mask, b_test_mask = self._initialize_b_masks(b_t, b_mask, b_test_mask)
        st = _native.stream()
        b_ones = torch.ones_like(b_t)
        if b_mask is not None:
            b_ones = b_ones * b_mask
        tau = _arrays.to_dev(A.T(b_ones))[0].contiguous()
        x_mask = _dev(x_mask, tau)
        x = torch.ones_like(tau) if x0 is None else _dev(x0, tau).clone()
        if x_mask is not None:
            x = x * x_mask
        x = x.contiguous()
        self.data_term.assign_data(b_t)
        bg = self.data_term.background_dev(b_t)
        lim = _Limits(lower_limit, upper_limit, x)
        hl, lo, hu, hi = lim.kernel()
        info = SolutionInfo(self.info(), max_iterations=iterations, tolerance=self.tolerance)
        track = b_test_mask is not None or self.tolerance is not None
        if track:
            self._init_residuals(info, A, x, b_t, b_mask, b_test_mask)
        ratio = torch.empty_like(b_t)
        # on-the-fly maps: fp, ratio and bp per angle block with the maps built once
        fused = _fused_step(A, b_t, bg)
        ax_buf = torch.empty_like(b_t) if fused else None
        for ii in range(iterations):
            info.iterations += 1
            if fused:
                upd, Ax = A.fp_row_bp(x, _native.CHIP_ROW_MLEM, ratio, b_t, None, bg, None, ax_buf)
            else:
                Ax = _arrays.to_dev(A(x))[0]
            if track and self._track(info, ii, Ax, b_t, b_mask, b_test_mask):
                break
            if not fused:
                _native.call("chip_mlem_ratio", ratio.data_ptr(), b_t.data_ptr(), Ax.data_ptr(), _native.ptr(bg),
                             b_t.numel(), st)
                upd = _arrays.to_dev(A.T(ratio))[0]
            _native.call("chip_mlem_update", x.data_ptr(), upd.data_ptr(), tau.data_ptr(), hl, lo, hu, hi,
                         _native.ptr(lim.kernel_mask(x_mask)), x.numel(), st)
            lim.finish(x, x_mask)
        return _arrays.from_dev(x, host), info


class SIRT(Solver):
    """SIRT (``solvers.py:808-984``)."""

    def __init__(self, verbose: bool = False, leave_progress: bool = True, relaxation: float = 1.95, tolerance=None,
                 regularizer=None, data_term="l2", data_term_test=None):
        super().__init__(verbose=verbose, leave_progress=leave_progress, relaxation=relaxation, tolerance=tolerance,
                         data_term=data_term, data_term_test=data_term_test)
        self.regularizer = self._initialize_regularizer(regularizer)

    def info(self) -> str:
        reg_info = "".join(["-" + r.info().upper() for r in self.regularizer])
        return Solver.info(self) + "-" + self.data_term.info() + reg_info

    def __call__(self, A, b, iterations: int, x0=None, lower_limit=None, upper_limit=None, x_mask=None, b_mask=None,
                 b_test_mask=None):
        b_t, host = _arrays.to_dev(b)
        b_t = b_t.clone()
        b_mask = _dev(b_mask, b_t)
        b_test_mask = _dev(b_test_mask, b_t)
        b_mask, b_test_mask = self._initialize_b_masks(b_t, b_mask, b_test_mask)
        tv_w = self._tv_weight()
        try:
            At_abs = A.T.absolute()
            A_abs = A.absolute()
        except AttributeError:
            print("WARNING: operator does not support absolute(); Using the operator itself.")
            A_abs, At_abs = A, A.T
        x_mask_t = None
        if x_mask is not None:
            x_mask_t = _arrays.to_dev(x_mask)[0]
        sigma, tau, x_shape, _ = compute_diagonal_scaling(A_abs, At_abs, b_t, regs=self.regularizer,
                                                          relaxation_tau=self.relaxation, x_mask=x_mask_t,
                                                          b_mask=b_mask)
        x_mask_t = _dev(x_mask_t, tau)
        x = torch.zeros_like(tau) if x0 is None else _dev(x0, tau).clone().contiguous()
        self.data_term.assign_data(b_t, sigma)
        bg = self.data_term.background_dev(b_t)
        term = self.data_term.device_term(b_t)
        lim = _Limits(lower_limit, upper_limit, x)
        hl, lo, hu, hi = lim.kernel()
        info = SolutionInfo(self.info(), max_iterations=iterations, tolerance=self.tolerance)
        track = b_test_mask is not None or self.tolerance is not None
        if track:
            self._init_residuals(info, A, x, b_t, b_mask, b_test_mask, sigma)
        st = _native.stream()
        res = torch.empty_like(b_t)
        n = int(x.shape[-1])
        batch = x.numel() // (n * n) if x.ndim >= 2 else 1
        q = torch.empty((batch, 2, n, n), dtype=torch.float32, device=x.device) if tv_w is not None else None
        grad_regs = self._grad_regs(x) if (tv_w is None and self.regularizer) else []
        # on-the-fly maps with the l2 term: fp, residual and bp per angle block with the maps built once
        fused = term[0] == data_terms.TERM_L2 and _fused_step(A, b_t, sigma, bg, b_mask)
        ax_buf = torch.empty_like(b_t) if fused else None
        for ii in range(iterations):
            info.iterations += 1
            if fused:
                upd, Ax = A.fp_row_bp(x, _native.CHIP_ROW_SIRT, res, b_t, sigma, bg, b_mask, ax_buf)
            else:
                Ax = _arrays.to_dev(A(x))[0]
            if track and self._track(info, ii, Ax, b_t, b_mask, b_test_mask):
                break
            if not fused:
                _residual(res, b_t, Ax, bg, b_mask, sigma, sigma, term, st)
            # the reference re-initialises the regularizer duals every iteration (solvers.py:950-958)
            for cfg, q_r in grad_regs:
                q_r.zero_()
                self._grad_dual(cfg, q_r, x, st)
            if not fused:
                upd = _arrays.to_dev(A.T(res))[0]
            if q is not None:
                q.zero_()
                _native.call("chip_tv_dual", q.data_ptr(), x.data_ptr(), n, batch, tv_w, st)
            for cfg, q_r in grad_regs:
                self._grad_primal(cfg, upd, q_r, True, st)
            _native.call("chip_sirt_update", x.data_ptr(), upd.data_ptr(), _native.ptr(q), tau.data_ptr(), hl, lo, hu,
                         hi, _native.ptr(lim.kernel_mask(x_mask_t)), n, batch, st)
            lim.finish(x, x_mask_t)
        return _arrays.from_dev(x, host), info


class PDHG(Solver):
    """Preconditioned primal-dual hybrid gradient (Chambolle-Pock), ``solvers.py:987-1206``."""

    def __init__(self, verbose: bool = False, leave_progress: bool = True, tolerance=None, relaxation: float = 0.95,
                 regularizer=None, data_term="l2", data_term_test=None):
        super().__init__(verbose=verbose, leave_progress=leave_progress, relaxation=relaxation, tolerance=tolerance,
                         data_term=data_term, data_term_test=data_term_test)
        self.regularizer = self._initialize_regularizer(regularizer)

    def info(self) -> str:
        reg_info = "".join(["-" + r.info().upper() for r in self.regularizer])
        return Solver.info(self) + "-" + self.data_term.info() + reg_info

    @staticmethod
    def _initialize_data_fidelity_function(data_term):
        if isinstance(data_term, str):
            if data_term.lower() == "l2":
                return data_terms.DataFidelity_l2()
            if data_term.lower() == "kl":
                return data_terms.DataFidelity_KL()
            if data_term.lower() == "l1":
                return data_terms.DataFidelity_l1()
            raise ValueError(f'Unknown data term: "{data_term}", accepted terms are: "l2" | "l1" | "kl".')
        return cp.deepcopy(data_term)

    def prepare(self, A, b, x0=None, x_mask=None, b_mask=None, b_test_mask=None, precondition: bool = True):
        """Device state of a solve: dict with b, masks, sigma, tau, x, x_relax, p, q, background.

        ``precondition=False`` (or an operator without ``absolute()``) takes the scalar step
        sizes of ``compute_Lipschitz_scaling`` (``solvers.py:120-158, 1127-1131``), held as
        constant arrays for the fused kernels."""
        b_t, host = _arrays.to_dev(b)
        b_t = b_t.clone()
        b_mask = _dev(b_mask, b_t)
        b_test_mask = _dev(b_test_mask, b_t)
        b_mask, b_test_mask = self._initialize_b_masks(b_t, b_mask, b_test_mask)
        tv_w = self._tv_weight()
        term = self.data_term.device_term(b_t)
        if precondition:
            try:
                At_abs = A.T.absolute()
                A_abs = A.absolute()
            except AttributeError:
                print(A)
                print("WARNING: Turning off preconditioning because system matrix does not support absolute")
                precondition = False
        x_mask_t = None if x_mask is None else _arrays.to_dev(x_mask)[0]
        if precondition:
            sigma, tau, x_shape, _ = compute_diagonal_scaling(A_abs, At_abs, b_t, regs=self.regularizer,
                                                              relaxation_tau=self.relaxation,
                                                              relaxation_sigma=self.relaxation, x_mask=x_mask_t,
                                                              b_mask=b_mask)
        else:
            sig, tau_s, x_shape, _ = compute_Lipschitz_scaling(A, b_t, regs=self.regularizer,
                                                               relaxation_sigma=self.relaxation,
                                                               relaxation_tau=self.relaxation)
            sigma = torch.full_like(b_t, float(sig))
            tau = tau_s if isinstance(tau_s, torch.Tensor) else \
                torch.full(x_shape, float(tau_s), dtype=torch.float32, device=b_t.device)
        x_mask_t = _dev(x_mask_t, tau)
        x = torch.zeros_like(tau) if x0 is None else _dev(x0, tau).clone().contiguous()
        n = int(x.shape[-1])
        batch = x.numel() // (n * n)
        self.data_term.assign_data(b_t, sigma)
        return dict(
            b=b_t, host=host, b_mask=b_mask, b_test_mask=b_test_mask, sigma=sigma, tau=tau, x_mask=x_mask_t, x=x,
            x_relax=x.clone(), p=torch.zeros_like(b_t), q=None if tv_w is None else torch.zeros(
                (batch, 2, n, n), dtype=torch.float32, device=x.device), tv_w=tv_w, n=n, batch=batch,
            bg=self.data_term.background_dev(b_t), term=term,
            grad_regs=self._grad_regs(x) if (tv_w is None and self.regularizer) else [],
        )

    def iterate(self, A, s: dict, iterations: int, lower_limit=None, upper_limit=None, info=None):
        """Run ``iterations`` PDHG iterations on the device state ``s`` (no host synchronisation)."""
        st = _native.stream()
        b, p, q, x, xr = s["b"], s["p"], s["q"], s["x"], s["x_relax"]
        lim = _Limits(lower_limit, upper_limit, x)
        hl, lo, hu, hi = lim.kernel()
        x_old = None if lim.fused else torch.empty_like(x)
        code, prm, period, g_len, g_stride = s["term"]
        dual = {data_terms.TERM_L2: "chip_pdhg_dual_l2", data_terms.TERM_KL: "chip_pdhg_dual_kl"}.get(code)
        track = info is not None and (s["b_test_mask"] is not None or self.tolerance is not None)
        # on-the-fly maps with the l2 / KL term: fp, dual step and bp per angle block, maps built once
        fused = dual is not None and _fused_step(A, b, s["sigma"], s["bg"], s["b_mask"], p)
        row_op = _native.CHIP_ROW_PDHG_KL if code == data_terms.TERM_KL else _native.CHIP_ROW_PDHG_L2
        ax_buf = torch.empty_like(b) if fused else None
        for ii in range(iterations):
            if info is not None:
                info.iterations += 1
            if fused:
                upd, _ = A.fp_row_bp(xr, row_op, p, b, s["sigma"], s["bg"], s["b_mask"], ax_buf)
            else:
                Ax = _arrays.to_dev(A(xr))[0]
                if dual is not None:
                    _native.call(dual, p.data_ptr(), Ax.data_ptr(), s["sigma"].data_ptr(), b.data_ptr(),
                                 _native.ptr(s["bg"]), _native.ptr(s["b_mask"]), p.numel(), st)
                else:
                    _native.call("chip_pdhg_dual_norm", p.data_ptr(), Ax.data_ptr(), s["sigma"].data_ptr(),
                                 b.data_ptr(), _native.ptr(s["bg"]), _native.ptr(s["b_mask"]), p.numel(), code,
                                 _native.ptr(prm), period, g_len, g_stride, st)
            if q is not None:
                _native.call("chip_tv_dual", q.data_ptr(), xr.data_ptr(), s["n"], s["batch"], s["tv_w"], st)
            for cfg, q_r in s["grad_regs"]:
                self._grad_dual(cfg, q_r, xr, st)
            if not fused:
                upd = _arrays.to_dev(A.T(p))[0]
            for cfg, q_r in s["grad_regs"]:
                self._grad_primal(cfg, upd, q_r, False, st)
            if x_old is not None:
                x_old.copy_(x)
            _native.call("chip_pdhg_primal", x.data_ptr(), xr.data_ptr(), upd.data_ptr(), _native.ptr(q),
                         s["tau"].data_ptr(), hl, lo, hu, hi, _native.ptr(lim.kernel_mask(s["x_mask"])), s["n"],
                         s["batch"], st)
            lim.finish(x, s["x_mask"], xr, x_old, 1.0)
            if track:
                Axn = _arrays.to_dev(A(x))[0]
                if self._track(info, ii, Axn, b, s["b_mask"], s["b_test_mask"]):
                    break
        return s

    def __call__(self, A, b, iterations: int, x0=None, lower_limit=None, upper_limit=None, x_mask=None, b_mask=None,
                 b_test_mask=None, precondition: bool = True):
        s = self.prepare(A, b, x0=x0, x_mask=x_mask, b_mask=b_mask, b_test_mask=b_test_mask,
                         precondition=precondition)
        info = SolutionInfo(self.info(), max_iterations=iterations, tolerance=self.tolerance)
        if s["b_test_mask"] is not None or self.tolerance is not None:
            self._init_residuals(info, A, s["x"], s["b"], s["b_mask"], s["b_test_mask"], s["sigma"])
        self.iterate(A, s, iterations, lower_limit, upper_limit, info)
        return _arrays.from_dev(s["x"], s["host"]), info
