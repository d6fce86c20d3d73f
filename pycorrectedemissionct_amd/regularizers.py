"""Regularizers used by the solvers (corrct ``regularizers.py`` subset).

``Regularizer_l1`` (``regularizers.py:531-585``): sparsity, identity transform;
its primal proximal (soft thresholding at tau * weight) is fused into
``chip_fista_update``, its dual steps in PDHG / SIRT run on ``chip_l1_dual`` /
``chip_l1_update_primal``.

``Regularizer_TV2D`` (``regularizers.py:308-321``) over ``Regularizer_Grad``
(``regularizers.py:218-289``): isotropic total variation with the l12 norm,
edge-padded forward gradient, ``sigma = 1/2`` and ``tau = 2 * ndims``. Its dual
update and prox run fused in ``chip_tv_dual``; its primal contribution
(``-div q``) is fused into the SIRT / PDHG primal kernels.

``Regularizer_TV1D`` / ``TV3D`` (``regularizers.py:292-337``), ``HubTV2D`` /
``HubTV3D`` (Huber norm, ``regularizers.py:340-383``), ``smooth1D/2D/3D`` (l2
norm, ``regularizers.py:386-431``), any other ``Regularizer_Grad`` over 1 to 3
axes (in any order and position, with any ``np.pad`` mode) with the l12, l1,
l2 or Huber (``l2_axis=0``) norm, and a scalar or array weight,
``Regularizer_VTV`` (``regularizers.py:1362-1430``) and ``Regularizer_TNV``
(``regularizers.py:1324-1359``) run on ``chip_grad_dual`` (dual update and
prox in one pass) and ``chip_grad_update_primal`` (``-div q``, weighted for
norms outside the l1 family), with the dual in the reference's
``[ndims, *primal]`` layout.
"""

from __future__ import annotations

import numpy as np

import torch

from . import _arrays, _native
from . import data_terms as dt
from . import operators


# TransformGradient's np.pad modes (operators.py:796-835) as the kernels' pad_mode codes
_PAD_CODE = {"edge": 0, "constant": 1, "wrap": 2, "reflect": 3, "symmetric": 4}


def _product(arrs, shape) -> torch.Tensor | None:
    """The elementwise product of flat arrays that each broadcast over ``shape`` (None if there are none)."""
    if not arrs:
        return None
    if len(arrs) == 1:
        return arrs[0]
    total = int(np.prod(shape))
    out = arrs[0].repeat(total // arrs[0].numel())
    for a in arrs[1:]:
        out = out * a.repeat(total // a.numel())
    return out.contiguous()


class BaseRegularizer:
    """Regularizer object API of ``regularizers.py:36-215``: ``initialize_sigma_tau``,
    ``initialize_dual``, ``update_dual``, ``apply_proximal_dual`` and ``compute_update_primal`` on
    NumPy arrays or device tensors (the solvers call the fused kernels directly)."""

    __reg_name__ = ""

    def __init__(self, weight, norm: dt.DataFidelityBase, upd_mask=None, dtype=np.float32):
        self.weight = np.array(weight)
        self.dtype = dtype
        self.op = None
        self.norm = norm
        self.upd_mask = upd_mask
        self.sigma = 1.0

    def _check_op(self) -> None:
        if self.op is None:
            raise ValueError("Regularizer not initialized! Please use method: `initialize_sigma_tau`.")

    def initialize_dual(self):
        """Zero dual of the operator's adjoint shape (``regularizers.py:120-132``)."""
        self._check_op()
        return np.zeros(self.op.adj_shape, dtype=self.dtype)

    def update_dual(self, dual, primal) -> None:
        """``dual += sigma * op(primal)`` in place (``regularizers.py:134-148``)."""
        self._check_op()
        dual += self.sigma * self.op(primal)

    def compute_update_primal(self, dual):
        """``op^T(dual)``, times the weight outside the l1 family (``regularizers.py:194-210``)."""
        self._check_op()
        upd = self.op.T(dual)
        if self.upd_mask is not None:
            upd = upd * self.upd_mask
        if not isinstance(self.norm, dt.DataFidelity_l1):
            w = float(self.weight) if self.weight.ndim == 0 else self.weight.astype(np.float32)
            if isinstance(upd, torch.Tensor) and not np.isscalar(w):
                w = torch.as_tensor(w, device=upd.device)
            upd = upd * w
        return upd

    @staticmethod
    def _bcast_dev(value, shape, what: str) -> torch.Tensor:
        """``value`` as a flat contiguous float32 device tensor that broadcasts over a primal of ``shape``:
        its shape (leading 1s dropped) must be a trailing part of ``shape``, so element e of the primal
        reads entry ``e % numel``, as NumPy broadcasts it in the reference."""
        t = value if isinstance(value, torch.Tensor) else torch.as_tensor(np.ascontiguousarray(value, dtype=np.float32))
        t = t.to(device=_arrays.device(), dtype=torch.float32).contiguous()
        vshape = tuple(int(v) for v in t.shape)
        while vshape and vshape[0] == 1:
            vshape = vshape[1:]
        shape = tuple(int(v) for v in shape)
        if len(vshape) > len(shape) or (vshape and shape[len(shape) - len(vshape):] != vshape):
            raise ValueError(f"{what} of shape {tuple(t.shape)} does not broadcast over a primal of shape {shape} "
                             "along its leading axes")
        return t.reshape(-1)

    def _weight_dev(self, shape) -> torch.Tensor | None:
        """The weight as a device array over the primal, or None for a scalar weight."""
        return None if self.weight.ndim == 0 else self._bcast_dev(self.weight, shape, "weight")

    def _prox_on_device(self, dual, cfg: dict) -> None:
        """The dual proximal of ``cfg`` in place: ``chip_grad_dual`` (``chip_l1_dual`` for the identity
        transform) with a zero primal."""
        t, host = _arrays.to_dev(dual)
        t = t.contiguous()
        if cfg.get("identity"):
            zero = torch.zeros_like(t)
            _native.call("chip_l1_dual", t.data_ptr(), zero.data_ptr(), t.numel(), cfg["weight"],
                         _native.ptr(cfg.get("weights")), cfg.get("weights_period", 1), _native.stream())
        else:
            zero = torch.zeros(cfg["q_shape"][1:], dtype=torch.float32, device=t.device)
            strides = cfg.get("strides")
            _native.call("chip_grad_dual", t.data_ptr(), zero.data_ptr(), cfg["n_chan"], cfg["n_batch"], cfg["nd"],
                         cfg["dims"].ctypes.data, None if strides is None else strides.ctypes.data,
                         cfg.get("pad", 0), cfg["mode"], cfg["weight"],
                         _native.ptr(cfg.get("weights")),
                         cfg.get("weights_period", 1), cfg["der"], cfg["chan"], _native.ptr(cfg.get("qref")),
                         _native.stream())
        if host:
            dual[...] = t.reshape(np.shape(dual)).cpu().numpy()
        elif t.data_ptr() != dual.data_ptr():
            dual.copy_(t.reshape(dual.shape))

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

    def initialize_sigma_tau(self, primal):
        """Sets up the gradient operator; returns the tau contribution (``regularizers.py:260-273``):
        a float, or ``tau * upd_mask`` when an update mask is set."""
        self.op = operators.TransformGradient(np.shape(primal), axes=self.axes, pad_mode=self.pad_mode)
        self.sigma = 0.5
        tau = 2 * self.ndims
        if not isinstance(self.norm, dt.DataFidelity_l1):
            tau *= self.weight
        if self.upd_mask is not None:
            return tau * self.upd_mask
        return tau

    def tau_terms(self, shape):
        """(scalar, flat device array or None) whose product is ``initialize_sigma_tau``'s result,
        for the solvers' device kernels: ``2 ndims`` (times a scalar weight outside the l1 family),
        and the update mask, an array weight outside the l1 family, or their product."""
        tau = 2 * self.ndims
        arrs = [] if self.upd_mask is None else [self._device_mask(shape)]
        if not isinstance(self.norm, dt.DataFidelity_l1):
            w = self._weight_dev(shape)
            if w is None:
                tau *= float(self.weight)
            else:
                arrs.insert(0, w)
        return float(tau), _product(arrs, shape)

    def _device_mask(self, shape) -> torch.Tensor:
        """The update mask as a flat device tensor broadcast over the primal's leading axes."""
        return self._bcast_dev(self.upd_mask, shape, "upd_mask")

    def apply_proximal_dual(self, dual) -> None:
        """The norm's dual proximal in place (``regularizers.py:150-162``; VTV / TNV override the
        norm as in the reference), on the device."""
        self._prox_on_device(dual, self.device_config(tuple(np.shape(dual))[1:]))

    def _layout(self, shape, multichannel: bool = False, what: str = "") -> tuple:
        """(channels, batch, gradient-axis sizes, strides or None) of a primal of ``shape``: the gradient
        axes are ``self.axes`` in their given order; with ``multichannel`` axis 0 holds the coupled
        channels (``_multichannel_layout``). Trailing axes in increasing order keep the batched layout
        (strides None); any other choice runs over one block per channel at the axes' strides."""
        nd_all = len(shape)
        axes = [int(a) % nd_all for a in np.asarray(self.axes).reshape(-1)]
        if len(set(axes)) != len(axes):
            raise ValueError(f"the gradient axes {list(self.axes)} repeat an axis")
        if not 1 <= len(axes) <= 3:
            raise NotImplementedError("the HIP gradient kernels take 1 to 3 gradient axes")
        if self.pad_mode not in _PAD_CODE:
            raise ValueError(f"Unsupported padding mode: {self.pad_mode}")
        n_chan, c0 = 1, 0
        if multichannel and nd_all > len(axes):
            n_chan, _ = _multichannel_layout(shape, len(axes), what)
            c0 = 1
            if 0 in axes:
                raise NotImplementedError(f"{what} takes the channels on axis 0, which cannot be a gradient axis")
        if axes == list(range(nd_all - len(axes), nd_all)):
            return n_chan, int(np.prod(shape[c0: nd_all - len(axes)])), [shape[a] for a in axes], None
        strides = [int(np.prod(shape[a + 1:], dtype=np.int64)) for a in axes]
        return n_chan, int(np.prod(shape[c0:], dtype=np.int64)), [shape[a] for a in axes], strides

    def device_config(self, shape) -> dict:
        """Kernel parameters of ``chip_grad_dual`` / ``chip_grad_update_primal`` for a primal of ``shape``."""
        shape = tuple(int(v) for v in shape)
        layout = self._layout(shape)
        if isinstance(self.norm, dt.DataFidelity_l12):
            if getattr(self.norm, "l2_axis", 0) != 0:
                raise NotImplementedError("the l12 norm is supported over the gradient axis (l2_axis=0)")
            mode = 0
        elif type(self.norm) is dt.DataFidelity_l1:
            mode = 1
        elif type(self.norm) is dt.DataFidelity_l2:
            # assign_data(None, sigma=0.5) leaves the dual step q * 1 / (1 + 0.5) (data_terms.py:294-304)
            cfg = self._config(shape, 4, layout, 0, 0)
            cfg["weight"] = float(np.float32(1 / (1 + self.sigma)))
            return cfg
        elif type(self.norm) is dt.DataFidelity_Huber:
            if self.norm.l2_axis not in (0, None):
                raise NotImplementedError("the Huber norm is supported over the gradient axis (l2_axis=0) or "
                                          "per component (l2_axis=None)")
            if np.asarray(self.norm.local_error).size != 1:
                raise NotImplementedError("the HIP gradient kernels take a scalar Huber size")
            cfg = self._config(shape, 5 if self.norm.l2_axis == 0 else 6, layout, 0, 0)
            # one_sigma_error = 1 / (1 + sigma huber_size) (data_terms.py:482-483)
            cfg["weight"] = float(np.float32(1.0 / (1.0 + self.sigma * float(self.norm.local_error))))
            return cfg
        else:
            raise NotImplementedError(f"norm {self.norm.info()} is not supported by the HIP gradient kernels")
        return self._config(shape, mode, layout, 0, 0)

    def _config(self, shape, mode, layout, der, chan) -> dict:
        n_chan, n_batch, dims, strides = layout
        l1_family = isinstance(self.norm, dt.DataFidelity_l1)
        mask = None if self.upd_mask is None else self._device_mask(shape)
        w_arr = self._weight_dev(shape)
        w = 0.0 if w_arr is not None else float(self.weight)
        # an array weight enters the prox of the l12 / l1 / VTV modes and the primal scale outside the
        # l1 family; TNV, l2 and Huber do not read it in the prox
        dual_w = w_arr if mode in (0, 1, 2) else None
        scales = None if l1_family else w_arr
        return dict(mode=mode, n_chan=int(n_chan), n_batch=int(n_batch), nd=len(dims),
                    dims=np.ascontiguousarray(dims, dtype=np.int32),
                    strides=None if strides is None else np.ascontiguousarray(strides, dtype=np.int64),
                    pad=_PAD_CODE[self.pad_mode], weight=w, weights=dual_w,
                    weights_period=1 if dual_w is None else int(dual_w.numel()),
                    der=der, chan=chan, scale=w, scales=scales,
                    scales_period=1 if scales is None else int(scales.numel()),
                    apply_scale=0 if l1_family else 1, q_shape=(self.ndims, *shape),
                    mask=mask, mask_period=1 if mask is None else int(mask.numel()), qref=None)

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


class Regularizer_TV1D(Regularizer_Grad):
    """Total Variation in 1D (``regularizers.py:292-305``)."""

    __reg_name__ = "TV1D"

    def __init__(self, weight, axes=None, pad_mode: str = "edge", upd_mask=None, norm=None):
        super().__init__(weight=weight, ndims=1, axes=axes, pad_mode=pad_mode, norm=norm, upd_mask=upd_mask)


class Regularizer_TV3D(Regularizer_Grad):
    """Total Variation in 3D (``regularizers.py:324-337``)."""

    __reg_name__ = "TV3D"

    def __init__(self, weight, axes=None, pad_mode: str = "edge", upd_mask=None, norm=None):
        super().__init__(weight=weight, ndims=3, axes=axes, pad_mode=pad_mode, norm=norm, upd_mask=upd_mask)


class Regularizer_HubTV2D(Regularizer_Grad):
    """Huber total variation in 2D (``regularizers.py:340-360``): the gradient's dual scaled by
    ``1 / (1 + 0.5 huber_size)`` then projected onto the unit l2 ball over the gradient axis."""

    __reg_name__ = "HubTV2D"

    def __init__(self, weight, huber_size: float, axes=None, pad_mode: str = "edge", upd_mask=None):
        super().__init__(weight=weight, ndims=2, axes=axes, pad_mode=pad_mode, upd_mask=upd_mask,
                         norm=dt.DataFidelity_Huber(huber_size, l2_axis=0))


class Regularizer_HubTV3D(Regularizer_Grad):
    """Huber total variation in 3D (``regularizers.py:363-383``)."""

    __reg_name__ = "HubTV3D"

    def __init__(self, weight, huber_size: float, axes=None, pad_mode: str = "edge", upd_mask=None):
        super().__init__(weight=weight, ndims=3, axes=axes, pad_mode=pad_mode, upd_mask=upd_mask,
                         norm=dt.DataFidelity_Huber(huber_size, l2_axis=0))


class Regularizer_smooth1D(Regularizer_Grad):
    """Smoothness (l2-norm of the gradient) in 1D (``regularizers.py:386-399``)."""

    __reg_name__ = "smooth1D"

    def __init__(self, weight, axes=None, pad_mode: str = "edge", upd_mask=None, norm=None):
        super().__init__(weight=weight, ndims=1, axes=axes, pad_mode=pad_mode, upd_mask=upd_mask,
                         norm=dt.DataFidelity_l2() if norm is None else norm)


class Regularizer_smooth2D(Regularizer_Grad):
    """Smoothness in 2D (``regularizers.py:402-415``)."""

    __reg_name__ = "smooth2D"

    def __init__(self, weight, axes=None, pad_mode: str = "edge", upd_mask=None, norm=None):
        super().__init__(weight=weight, ndims=2, axes=axes, pad_mode=pad_mode, upd_mask=upd_mask,
                         norm=dt.DataFidelity_l2() if norm is None else norm)


class Regularizer_smooth3D(Regularizer_Grad):
    """Smoothness in 3D (``regularizers.py:418-431``)."""

    __reg_name__ = "smooth3D"

    def __init__(self, weight, axes=None, pad_mode: str = "edge", upd_mask=None, norm=None):
        super().__init__(weight=weight, ndims=3, axes=axes, pad_mode=pad_mode, upd_mask=upd_mask,
                         norm=dt.DataFidelity_l2() if norm is None else norm)


_PWISE_CODE = {1: 1, 2: 2, np.inf: 0}


def _grad_edge(x: np.ndarray, nd: int) -> np.ndarray:
    """Forward differences over the last ``nd`` axes with edge padding (the last difference 0),
    stacked first: ``TransformGradient`` (``operators.py:796-819``) on the host, float32."""
    x = np.asarray(x, dtype=np.float32)
    out = np.zeros((nd, *x.shape), dtype=np.float32)
    for k in range(nd):
        ax = x.ndim - nd + k
        hi = [slice(None)] * x.ndim
        lo = [slice(None)] * x.ndim
        hi[ax], lo[ax] = slice(1, None), slice(None, -1)
        out[(k, *lo)] = x[tuple(hi)] - x[tuple(lo)]
    return out


def _multichannel_layout(shape, ndims: int, what: str) -> tuple[int, int]:
    """(channels, independent batch) of a primal whose axis 0 holds the channels."""
    if len(shape) <= ndims:
        return 1, 1
    n_chan = int(shape[0])
    if n_chan > 16:
        raise NotImplementedError(f"{what} couples at most 16 channels in the HIP kernels")
    return n_chan, int(np.prod(shape[1: len(shape) - ndims]))


class Regularizer_VTV(Regularizer_Grad):
    """Vectorial Total Variation (``regularizers.py:1362-1430``) for multi-channel volumes
    ``[C, ..., *gradient axes]``: the dual is divided by ``max(dn, weight)`` and multiplied by
    ``weight``, with ``dn`` the point-wise derivative norm (1, 2, inf) followed by the norm across
    channels (1, 2, inf; the dual norms of inf are 1)."""

    __reg_name__ = "VTV"

    def __init__(self, weight, ndims: int = 2, pwise_der_norm=2, pwise_chan_norm=np.inf, x_ref=None,
                 upd_mask=None):
        super().__init__(weight=weight, ndims=ndims, upd_mask=upd_mask)
        self.pwise_der_norm = pwise_der_norm
        if self.pwise_der_norm not in [1, 2, np.inf]:
            self._raise_pwise_norm_error()
        self.pwise_chan_norm = pwise_chan_norm
        if self.pwise_chan_norm not in [1, 2, np.inf]:
            self._raise_pwise_norm_error()
        # the reference volume's dual sigma * grad(x_ref), a fixed extra channel of the prox
        # (regularizers.py:1386-1392): computed once here, on the host, as the reference does
        if isinstance(x_ref, torch.Tensor):
            x_ref = x_ref.detach().cpu().numpy()
        self.x_ref = None if x_ref is None else np.asarray(x_ref, dtype=np.float32)
        self.q_ref = None if x_ref is None else \
            np.expand_dims((np.float32(0.5) * _grad_edge(self.x_ref, ndims)).astype(np.float32), axis=1)
        self._q_ref_dev = None

    def _raise_pwise_norm_error(self):
        raise ValueError(
            "The only supported point-wise norm exponents are: 1, 2, and Inf."
            + f" Provided the following instead: derivatives={self.pwise_der_norm}, channel={self.pwise_chan_norm}"
        )

    def device_config(self, shape) -> dict:
        shape = tuple(int(v) for v in shape)
        cfg = self._config(shape, 2, self._layout(shape, True, "VTV"), _PWISE_CODE[self.pwise_der_norm],
                           _PWISE_CODE[self.pwise_chan_norm])
        if self.q_ref is not None:
            want = shape[1:] if len(shape) > self.ndims else shape  # the primal without its channel axis
            if tuple(self.x_ref.shape) != want:
                raise ValueError(f"x_ref of shape {self.x_ref.shape} does not match the primal's {want}")
            if self._q_ref_dev is None:
                self._q_ref_dev = torch.as_tensor(np.ascontiguousarray(self.q_ref)).to(_arrays.device())
            cfg["qref"] = self._q_ref_dev
        return cfg


class Regularizer_TNV(Regularizer_Grad):
    """Total Nuclear Variation (``regularizers.py:1324-1359``) for multi-channel volumes
    ``[C, ..., *gradient axes]``: the nuclear norm of each (channels x derivatives) Jacobian,
    through ``DataFidelity_ln(ln_axes=(1, 0))``; its dual proximal clips the singular values at 1
    (the l1 spectral norm), and the weight scales tau and the primal update."""

    __reg_name__ = "TNV"

    def __init__(self, weight, ndims: int = 2, axes=None, pad_mode: str = "edge", upd_mask=None,
                 spectral_norm=None, x_ref=None):
        super().__init__(weight=weight, ndims=ndims, axes=axes, pad_mode=pad_mode, upd_mask=upd_mask)
        self.norm = dt.DataFidelity_ln(ln_axes=(1, 0), spectral_norm=spectral_norm)
        # kept as in the reference, whose TNV computes the reference dual (regularizers.py:1353-1356)
        # but never reads it: its dual proximal is the norm's (regularizers.py:150-162)
        self.x_ref = x_ref
        self.x_ref = None

    def device_config(self, shape) -> dict:
        shape = tuple(int(v) for v in shape)
        if type(self.norm.spectral_norm) is not dt.DataFidelity_l1:
            raise NotImplementedError("the HIP TNV kernel implements the l1 spectral norm")
        if len(shape) <= self.ndims:
            raise NotImplementedError("TNV needs a channel axis in front of the gradient axes")
        return self._config(shape, 3, self._layout(shape, True, "TNV"), 0, 0)


class Regularizer_l1(BaseRegularizer):
    """l1-norm regularizer: promotes sparse reconstructions (``regularizers.py:531-585``). In PDHG and
    SIRT its identity transform runs on ``chip_l1_dual`` / ``chip_l1_update_primal``; in FISTA its
    primal proximal is fused into ``chip_fista_update``. The weight may be an array and the update
    mask any array, each broadcast over the primal's leading axes."""

    __reg_name__ = "l1"

    def __init__(self, weight, upd_mask=None, norm=None):
        super().__init__(weight=weight, norm=dt.DataFidelity_l1() if norm is None else norm, upd_mask=upd_mask)

    def initialize_sigma_tau(self, primal):
        """Identity transform, sigma = 1, tau = 1 times the update mask (``regularizers.py:544-553``)."""
        self.op = operators.TransformIdentity(np.shape(primal))
        self.sigma = 1.0
        if self.upd_mask is not None:
            return 1.0 * self.upd_mask
        return 1.0

    def tau_terms(self, shape):
        """(1, update mask as a flat device array or None), as ``Regularizer_Grad.tau_terms``."""
        return 1.0, (None if self.upd_mask is None else self._bcast_dev(self.upd_mask, shape, "upd_mask"))

    def update_dual(self, dual, primal) -> None:
        dual += primal

    def _check_norm(self) -> None:
        if type(self.norm) is not dt.DataFidelity_l1:
            raise NotImplementedError(f"norm {self.norm.info()} of Regularizer_l1 is not supported by the HIP kernels")

    def device_config(self, shape) -> dict:
        """Kernel parameters of ``chip_l1_dual`` / ``chip_l1_update_primal`` for a primal of ``shape``."""
        self._check_norm()
        shape = tuple(int(v) for v in shape)
        w_arr = self._weight_dev(shape)
        mask = None if self.upd_mask is None else self._bcast_dev(self.upd_mask, shape, "upd_mask")
        return dict(identity=True, n=int(np.prod(shape)), q_shape=shape,
                    weight=0.0 if w_arr is not None else float(self.weight), weights=w_arr,
                    weights_period=1 if w_arr is None else int(w_arr.numel()),
                    mask=mask, mask_period=1 if mask is None else int(mask.numel()))

    def apply_proximal_dual(self, dual) -> None:
        """Clip to [-weight, weight] (``data_terms.py:549-555``), on the device."""
        self._prox_on_device(dual, self.device_config(tuple(np.shape(dual))))

    def fista_prox(self, shape) -> dict:
        """Parameters of the proximal fused into ``chip_fista_update``: the prox of tau * weight * |.|_1,
        inside the update mask when there is one (``regularizers.py:559-585``)."""
        self._check_norm()
        cfg = self.device_config(shape)
        return dict(l1w=cfg["weight"], l1wa=cfg["weights"], l1w_period=cfg["weights_period"], l1m=cfg["mask"],
                    l1m_period=cfg["mask_period"])
