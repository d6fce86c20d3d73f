"""Many independent XRF slices of different sizes as one operator.

corrct reconstructs a stack one slice at a time (``ProjectorAttenuationXRF``
over a 3-D volume; ``_projector_backends.py:86-92`` loops the slices for the
maps). Here every slice of a batch -- its own size, phantom and attenuation --
lives back to back in packed device buffers, and each fp / bp is one
persistent launch over all slices (``chip_batch_fp`` / ``chip_batch_bp``).

* ``BatchPlan``            -- the ``chip_batch_plan`` handle.
* ``BatchedProjectorXRF``  -- maps, fp, bp and pack / unpack of per-slice arrays.
* ``BatchedSIRT``, ``BatchedPDHG`` -- the SIRT (``solvers.py:867-984``) and
  PDHG-TV (``solvers.py:1060-1206``) updates of every slice at once, with the
  iterations captured in a HIP graph and replayed; results equal per-slice
  ``SIRT`` / ``PDHG`` runs.
* ``autotune``             -- times the persistent-grid sizes of the fp and bp
  on the actual batch and keeps the fastest.
"""

from __future__ import annotations

import ctypes
import time

import numpy as np
import torch

from . import _arrays, _native
from .solvers import Solver

_BP_MODES = {"skimage": _native.CHIP_BP_SKIMAGE, "symmetric": _native.CHIP_BP_SKIMAGE_SYM,
             "adjoint": _native.CHIP_BP_ADJOINT}


class BatchPlan:
    """A ``chip_batch_plan`` over slices of sides ``sizes`` (each a multiple of 4)."""

    def __init__(self, sizes, angles_rad_f32, det_angles_rad=(np.pi / 2,), w_det=None, w_ang=None):
        self.sizes = np.ascontiguousarray(np.asarray(sizes, dtype=np.int32).reshape(-1))
        self.angles = np.ascontiguousarray(np.asarray(angles_rad_f32, dtype=np.float32).reshape(-1))
        self.det_angles = np.ascontiguousarray(np.asarray(det_angles_rad, dtype=np.float64).reshape(-1))
        self.n_slices, self.n_angles, self.n_det = len(self.sizes), len(self.angles), len(self.det_angles)
        self.w_det = None if w_det is None else np.ascontiguousarray(
            np.broadcast_to(np.asarray(w_det, dtype=np.float64).reshape(-1), (self.n_det,)))
        self.w_ang = None if w_ang is None else np.ascontiguousarray(
            np.asarray(w_ang, dtype=np.float64).reshape(self.n_angles, self.n_det))
        _arrays.device()
        h = ctypes.c_void_p()
        _native.call("chip_batch_plan_create", ctypes.byref(h), self.n_slices, self.sizes.ctypes.data, self.n_angles,
                     self.angles.ctypes.data, self.n_det, self.det_angles.ctypes.data,
                     None if self.w_det is None else self.w_det.ctypes.data,
                     None if self.w_ang is None else self.w_ang.ctypes.data)
        self._h = h
        self._trig = _native.radon_trig(self.angles)
        _native.call("chip_batch_set_radon_trig", h, self._trig.ctypes.data)
        tot = (ctypes.c_int64 * 3)()
        _native.call("chip_batch_info", h, ctypes.addressof(tot), ctypes.addressof(tot) + 8,
                     ctypes.addressof(tot) + 16)
        self.vol_elems, self.sino_elems, self.maps_elems = int(tot[0]), int(tot[1]), int(tot[2])
        n2 = self.sizes.astype(np.int64) ** 2
        self.vol_start = np.concatenate([[0], np.cumsum(n2)])
        self.sino_start = np.concatenate([[0], np.cumsum(self.n_det * self.n_angles * self.sizes.astype(np.int64))])
        self.maps_start = np.concatenate([[0], np.cumsum(self.n_angles * self.n_det * n2)])
        self.tuning: dict[str, int] = {}

    def close(self) -> None:
        h = getattr(self, "_h", None)
        self._h = None
        if h is not None and h.value:
            try:
                _native.lib().chip_batch_plan_destroy(h)
            except Exception:  # interpreter shutdown
                pass

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def tune(self, key: str, value: int) -> None:
        _native.call("chip_batch_tune", self._h, key.encode(), int(value))
        self.tuning[key] = int(value)

    def att_maps(self, mu_in, mu_out, maps):
        _native.call("chip_batch_att_maps", self._h, _native.ptr(mu_in), _native.ptr(mu_out), _native.ptr(maps),
                     _native.stream())
        return maps

    def fp(self, vol, maps, sino):
        _native.call("chip_batch_fp", self._h, _native.ptr(vol), _native.ptr(maps), _native.ptr(sino),
                     _native.stream())
        return sino

    def bp(self, sino, maps, vol, mode: int):
        _native.call("chip_batch_bp", self._h, _native.ptr(sino), _native.ptr(maps), _native.ptr(vol), int(mode),
                     _native.stream())
        return vol


class BatchedProjectorXRF:
    """Attenuation-corrected XRF projector over a batch of slices of sizes ``sizes``.

    ``att_in`` / ``att_out`` are per-slice lists of [N_s, N_s] arrays (or
    None); ``att_maps`` may instead give the packed maps directly. Slices share
    the angles, detectors and weights. ``bp_mode``: "skimage" (corrct's
    default, maps and weights ignored), "symmetric" (``is_symmetric=True``) or
    "adjoint" (exact transpose of the fp). ``maps_mode="otf"`` keeps only the packed
    attenuation volumes and rebuilds each slice's maps per angle block inside every fp / bp
    (``chip_batch_fp_otf`` / ``chip_batch_bp_otf``): for batches whose maps do not fit.
    """

    def __init__(self, sizes, angles_rot_rad, att_in=None, att_out=None, *, att_maps=None,
                 angles_detectors_rad=(np.pi / 2,), weights_detectors=None, weights_angles=None,
                 bp_mode: str = "adjoint", maps_mode: str = "precomputed"):
        if bp_mode not in _BP_MODES:
            raise ValueError(f"Unknown bp_mode: {bp_mode}")
        if maps_mode not in ("precomputed", "otf"):
            raise ValueError(f"Unknown maps_mode: {maps_mode}")
        if maps_mode == "otf" and (att_maps is not None or (att_in is None and att_out is None)):
            raise ValueError("maps_mode='otf' builds the maps from att_in / att_out: give those, not att_maps")
        self.bp_mode = bp_mode
        self.maps_mode = maps_mode
        self.plan = BatchPlan(sizes, np.asarray(angles_rot_rad, dtype=np.float32), angles_detectors_rad,
                              weights_detectors, weights_angles)
        self.sizes = [int(s) for s in self.plan.sizes]
        self.n_angles, self.n_det = self.plan.n_angles, self.plan.n_det
        self.device = _arrays.device()
        if att_maps is not None:
            self.maps = _arrays.to_dev(att_maps)[0].reshape(-1).contiguous()
            if self.maps.numel() != self.plan.maps_elems:
                raise ValueError(f"att_maps holds {self.maps.numel()} elements, the batch needs {self.plan.maps_elems}")
        elif att_in is not None or att_out is not None:
            mu_in = None if att_in is None else self.pack_vol(att_in)
            mu_out = None if att_out is None else self.pack_vol(att_out)
            if maps_mode == "otf":
                self.mu_in, self.mu_out = mu_in, mu_out
                self.maps = None
            else:
                self.maps = torch.empty(self.plan.maps_elems, dtype=torch.float32, device=self.device)
                self.plan.att_maps(mu_in, mu_out, self.maps)
        else:
            self.maps = None

    # ---- packing -----------------------------------------------------------------------
    def _pack(self, arrays, start, shape_of) -> torch.Tensor:
        out = torch.empty(int(start[-1]), dtype=torch.float32, device=self.device)
        for s, a in enumerate(arrays):
            t = _arrays.to_dev(a)[0].reshape(-1)
            if t.numel() != int(start[s + 1] - start[s]):
                raise ValueError(f"slice {s}: expected shape {shape_of(s)}, got {tuple(np.shape(a))}")
            out[int(start[s]):int(start[s + 1])] = t
        return out

    def pack_vol(self, arrays) -> torch.Tensor:
        return self._pack(arrays, self.plan.vol_start, lambda s: (self.sizes[s],) * 2)

    def pack_sino(self, arrays) -> torch.Tensor:
        return self._pack(arrays, self.plan.sino_start, lambda s: (self.n_det, self.n_angles, self.sizes[s]))

    def unpack_vol(self, flat: torch.Tensor) -> list[torch.Tensor]:
        st = self.plan.vol_start
        return [flat[int(st[s]):int(st[s + 1])].view(n, n) for s, n in enumerate(self.sizes)]

    def unpack_sino(self, flat: torch.Tensor) -> list[torch.Tensor]:
        st = self.plan.sino_start
        return [flat[int(st[s]):int(st[s + 1])].view(self.n_det, self.n_angles, n) for s, n in enumerate(self.sizes)]

    def slice_maps(self, s: int) -> torch.Tensor:
        if self.maps_mode == "otf":
            raise NotImplementedError("maps_mode='otf' holds no maps")
        st = self.plan.maps_start
        n = self.sizes[s]
        return self.maps[int(st[s]):int(st[s + 1])].view(self.n_angles, self.n_det, n, n)

    # ---- operator ----------------------------------------------------------------------
    @property
    def vol_elems(self) -> int:
        return self.plan.vol_elems

    @property
    def sino_elems(self) -> int:
        return self.plan.sino_elems

    def fp(self, vol: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        if out is None:
            out = torch.empty(self.plan.sino_elems, dtype=torch.float32, device=self.device)
        if self.maps_mode == "otf":
            _native.call("chip_batch_fp_otf", self.plan._h, _native.ptr(vol), _native.ptr(self.mu_in),
                         _native.ptr(self.mu_out), _native.ptr(out), _native.stream())
            return out
        return self.plan.fp(vol, self.maps, out)

    def bp(self, sino: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        if out is None:
            out = torch.empty(self.plan.vol_elems, dtype=torch.float32, device=self.device)
        if self.maps_mode == "otf" and self.bp_mode != "skimage":
            _native.call("chip_batch_bp_otf", self.plan._h, _native.ptr(sino), _native.ptr(self.mu_in),
                         _native.ptr(self.mu_out), _native.ptr(out), _BP_MODES[self.bp_mode], _native.stream())
            return out
        maps = None if self.bp_mode == "skimage" else self.maps
        return self.plan.bp(sino, maps, out, _BP_MODES[self.bp_mode])

    def fused_data_step(self) -> bool:
        """Whether ``fp_row_bp`` applies: on-the-fly maps and a bp that reads them."""
        return self.maps_mode == "otf" and self.bp_mode != "skimage"

    def fp_row_bp(self, x: torch.Tensor, row_op: int, p: torch.Tensor, b: torch.Tensor, sigma, ax: torch.Tensor,
                  out: torch.Tensor) -> torch.Tensor:
        """One iteration's projector work per slice and angle block with the maps built once
        (``chip_batch_otf_fp_row_bp``): ax = fp(x), the row op on p, out = bp(p)."""
        _native.call("chip_batch_otf_fp_row_bp", self.plan._h, _native.ptr(x), _native.ptr(self.mu_in),
                     _native.ptr(self.mu_out), int(row_op), _native.ptr(p), _native.ptr(sigma), _native.ptr(b),
                     _native.ptr(ax), _native.ptr(out), _BP_MODES[self.bp_mode], _native.stream())
        return out


# fp tile shapes of the batched kernel that give the same bits: LDS staging through registers or
# gfx950's direct-to-LDS loads, each with the whole chunk box or only its footprint
# The batched fp's kernel shapes. Image tiles (the default family): a workgroup owns a tile of one
# or two slices and loops over a group of angles; its kinds differ only in when the map tiles are
# fetched and in the samples per loop trip, so every tile shape gives the same bits. The angle x
# bin-block stream shapes form a second family (the same bits among themselves, within 1e-6 of the
# tile shapes: their line sums run in another order).
TILE_SHAPES = {"tile": dict(fp_tile=1, fp_tile_kind=1), "tile_ahead2": dict(fp_tile=1, fp_tile_kind=0),
               "tile_staged": dict(fp_tile=1, fp_tile_kind=3)}
STREAM_SHAPES = {"stream": dict(fp_tile=0, fp_direct=0, fp_trim=0),
                 "stream_trim": dict(fp_tile=0, fp_direct=0, fp_trim=1),
                 "direct": dict(fp_tile=0, fp_direct=1, fp_trim=0),
                 "direct_trim": dict(fp_tile=0, fp_direct=1, fp_trim=1)}
FP_SHAPES = {**TILE_SHAPES, **STREAM_SHAPES}


def autotune(A: BatchedProjectorXRF, reps: int = 5, fp_candidates=(0, 2, 4, 5, 6), bp_candidates=(0, 2, 4, 8),
             fp_shapes=tuple(TILE_SHAPES)) -> dict:
    """Pick the batched fp's kernel shape (and, for the stream shapes, its grid size) and the bp's
    grid size by timing them on ``A``'s batch: k workgroups per CU of a persistent grid, or 0 for one
    workgroup per work item; the fp shapes named in ``fp_shapes`` (keys of ``FP_SHAPES``; the image
    tiles by default, whose grid follows from the batch).

    Returns ``{"fp_shape": s, "fp_wg_per_cu": k, "bp_wg_per_cu": k, "fp_us": {...}, "bp_us": {...}}``
    (``bp_wg_per_cu`` "pairs": the exact-adjoint bp runs two slices of a size per workgroup, four in
    one-detector passes);
    the chosen values are applied to ``A.plan``. The candidates of one family compute every work
    item the same way, so within the default candidates results do not depend on the choice (bit
    for bit), only the time does.
    """
    x = torch.rand(A.vol_elems, dtype=torch.float32, device=A.device)
    y = torch.empty(A.sino_elems, dtype=torch.float32, device=A.device)
    v = torch.empty_like(x)
    stream = torch.cuda.current_stream()

    def timed(fn) -> float:
        fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        e1.synchronize()
        return e0.elapsed_time(e1) * 1e3 / reps

    A.plan.tune("fp_patch", 0)
    fp_us, bp_us = {}, {}
    for shape in fp_shapes:
        for key, val in FP_SHAPES[shape].items():
            A.plan.tune(key, val)
        for k in (fp_candidates if shape in STREAM_SHAPES else (0,)):
            A.plan.tune("fp_wg_per_cu", k)
            fp_us[(shape, k)] = timed(lambda: A.fp(x, y))
    # exact-adjoint bp: two slices of a size per workgroup (one launch per size), or one slice per
    # workgroup in one launch over every slice with k workgroups per CU; the same bits either way
    A.plan.tune("bp_pairs", 1)
    bp_us["pairs"] = timed(lambda: A.bp(y, v))
    A.plan.tune("bp_pairs", 0)
    for k in bp_candidates:
        A.plan.tune("bp_wg_per_cu", k)
        bp_us[k] = timed(lambda: A.bp(y, v))
    best_shape, best_fp = min(fp_us, key=fp_us.get)
    best_bp = min(bp_us, key=bp_us.get)
    for key, val in FP_SHAPES[best_shape].items():
        A.plan.tune(key, val)
    A.plan.tune("fp_wg_per_cu", best_fp)
    A.plan.tune("bp_pairs", 1 if best_bp == "pairs" else 0)
    if best_bp != "pairs":
        A.plan.tune("bp_wg_per_cu", best_bp)
    return {"fp_shape": best_shape, "fp_wg_per_cu": best_fp, "bp_wg_per_cu": best_bp,
            "fp_us": {f"{s}/{k}": t for (s, k), t in fp_us.items()}, "bp_us": bp_us}


def _precondition(A: BatchedProjectorXRF, vec: torch.Tensor, start, add: float, rho: float) -> torch.Tensor:
    """Per slice: vec = rho / fix(|vec| + add), fix against the slice's own maximum (``solvers.py:98-116``)."""
    st = _native.stream()
    mx = torch.empty(len(start) - 1, dtype=torch.float32, device=vec.device)
    for s in range(len(start) - 1):
        a, b = int(start[s]), int(start[s + 1])
        seg = vec[a:b]
        _native.call("chip_absmax", seg.data_ptr(), b - a, mx[s:s + 1].data_ptr(), st)
        _native.call("chip_precond_fix", seg.data_ptr(), b - a, float(add), mx[s:s + 1].data_ptr(), float(rho), st)
    return vec


def _limit(v):
    return (0, 0.0) if v is None else (1, float(v))


class _GraphedLoop:
    """Runs ``step`` ``iterations`` times, replaying a HIP graph of ``chunk`` captured steps.

    The captured launches read and write only the solver's persistent buffers, so each
    replay continues where the previous one stopped. The first call runs one step eagerly
    (every kernel is loaded before the capture), captures ``chunk`` steps (a capture does
    not execute them), replays; the remainder runs eagerly.
    """

    def __init__(self, step, chunk: int, use_graph: bool):
        self.step, self.chunk, self.use_graph = step, int(chunk), bool(use_graph)
        self.graph = None

    def run(self, iterations: int) -> None:
        done = 0
        if self.use_graph and self.graph is None and iterations > self.chunk:
            self.step()
            done = 1
            side = torch.cuda.Stream()
            side.wait_stream(torch.cuda.current_stream())
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g, stream=side):
                for _ in range(self.chunk):
                    self.step()
            torch.cuda.current_stream().wait_stream(side)
            self.graph = g
        left = iterations - done
        if self.graph is not None:
            full, left = divmod(left, self.chunk)
            for _ in range(full):
                self.graph.replay()
        for _ in range(left):
            self.step()


class BatchedSIRT:
    """SIRT (``solvers.py:867-984``) on every slice of a batch: x += tau A^T(sigma (b - A x))."""

    def __init__(self, relaxation: float = 1.95, graph_chunk: int = 10, use_graph: bool = True):
        self.relaxation = relaxation
        self.graph_chunk = graph_chunk
        self.use_graph = use_graph

    def prepare(self, A: BatchedProjectorXRF, b: torch.Tensor, x0=None) -> dict:
        b = _arrays.to_dev(b)[0].reshape(-1).contiguous()
        tau = _precondition(A, A.bp(torch.ones_like(b)).clone(), A.plan.vol_start, 0.0, self.relaxation)
        sigma = _precondition(A, A.fp(torch.ones_like(tau)).clone(), A.plan.sino_start, 0.0, 1.0)
        x = torch.zeros_like(tau) if x0 is None else _arrays.to_dev(x0)[0].reshape(-1).clone()
        return dict(b=b, tau=tau, sigma=sigma, x=x, ax=torch.empty_like(b), res=torch.empty_like(b),
                    upd=torch.empty_like(tau))

    def iterate(self, A: BatchedProjectorXRF, s: dict, iterations: int, lower_limit=None, upper_limit=None):
        hl, lo = _limit(lower_limit)
        hu, hi = _limit(upper_limit)

        fused = A.fused_data_step()

        def step():
            st = _native.stream()
            if fused:  # on-the-fly maps: fp, residual and bp per angle block, the maps built once
                A.fp_row_bp(s["x"], _native.CHIP_ROW_SIRT, s["res"], s["b"], s["sigma"], s["ax"], s["upd"])
            else:
                A.fp(s["x"], s["ax"])
                _native.call("chip_sirt_residual", s["res"].data_ptr(), s["b"].data_ptr(), s["ax"].data_ptr(), None,
                             None, s["sigma"].data_ptr(), s["b"].numel(), st)
                A.bp(s["res"], s["upd"])
            _native.call("chip_batch_sirt_update", A.plan._h, s["x"].data_ptr(), s["upd"].data_ptr(), None,
                         s["tau"].data_ptr(), hl, lo, hu, hi, None, st)

        key = ("sirt", hl, lo, hu, hi)
        loop = s.get(key)
        if loop is None:
            loop = s[key] = _GraphedLoop(step, self.graph_chunk, self.use_graph)
        loop.run(iterations)
        return s

    def __call__(self, A, b, iterations: int, x0=None, lower_limit=None, upper_limit=None):
        s = self.prepare(A, b, x0)
        self.iterate(A, s, iterations, lower_limit, upper_limit)
        return s["x"]


class BatchedPDHG:
    """PDHG with an l2 or KL data term and isotropic TV(lambda) (``solvers.py:1060-1206``) on every slice.

    ``regularizer`` (instead of ``tv_lambda``) takes any gradient-family regularizer of
    ``regularizers.py`` -- TV2D over one slice, or VTV / TNV / TV3D coupling ``channels``
    consecutive slices of equal size (the emission channels of one slice) as one
    ``[channels, N, N]`` volume -- run by ``chip_grad_dual`` / ``chip_grad_update_primal`` per
    group inside the captured iteration.
    """

    def __init__(self, data_term: str = "l2", tv_lambda: float | None = None, relaxation: float = 0.95,
                 graph_chunk: int = 10, use_graph: bool = True, regularizer=None, channels: int = 1):
        if data_term not in ("l2", "kl"):
            raise ValueError(f"Unknown data term: {data_term}")
        if regularizer is not None and tv_lambda is not None:
            raise ValueError("Give either tv_lambda or regularizer")
        self.data_term = data_term
        self.tv_lambda = tv_lambda
        self.regularizer = regularizer
        self.channels = int(channels)
        self.relaxation = relaxation
        self.graph_chunk = graph_chunk
        self.use_graph = use_graph

    def _groups(self, A: BatchedProjectorXRF, x: torch.Tensor):
        """(kernel parameters, primal view, dual view) per channel group of the regularizer."""
        from .regularizers import Regularizer_Grad

        reg = self.regularizer
        if not isinstance(reg, Regularizer_Grad):
            raise NotImplementedError("the batched PDHG takes gradient-family regularizers")
        if reg.upd_mask is not None:
            raise NotImplementedError("the batched PDHG takes regularizers without an update mask")
        if np.ndim(reg.weight) != 0:
            raise NotImplementedError("the batched PDHG takes regularizers with a scalar weight")
        c = self.channels
        if len(A.sizes) % c:
            raise ValueError(f"{len(A.sizes)} slices do not split into groups of {c} channels")
        groups, q_elems = [], 0
        for g in range(len(A.sizes) // c):
            sizes = A.sizes[g * c:(g + 1) * c]
            if len(set(sizes)) != 1:
                raise ValueError(f"the channels of group {g} differ in size: {sizes}")
            shape = (c, sizes[0], sizes[0]) if c > 1 else (sizes[0], sizes[0])
            cfg = reg.device_config(shape)
            start = int(A.plan.vol_start[g * c])
            n_el = int(np.prod(shape))
            groups.append((cfg, start, n_el, q_elems))
            q_elems += cfg["nd"] * n_el
        q = torch.zeros(q_elems, dtype=torch.float32, device=x.device)
        tau_reg = float(reg.initialize_sigma_tau(np.zeros(groups[0][0]["q_shape"][1:], dtype=np.float32)))
        out = [(cfg, start, n_el, q[qo:qo + cfg["nd"] * n_el]) for cfg, start, n_el, qo in groups]
        return out, q, tau_reg

    def prepare(self, A: BatchedProjectorXRF, b: torch.Tensor, x0=None, dual_init: bool = False) -> dict:
        """Device state of the solve. With ``x0`` and ``dual_init``, the duals start consistent
        with x0 (data term: its gradient at A x0; TV: the projection of grad x0 onto the
        lambda-ball) instead of at 0, so a good primal estimate is not first undone."""
        b = _arrays.to_dev(b)[0].reshape(-1).contiguous()
        # TV's tau contribution: 2 * ndims, unweighted for its l12 (an l1-family) norm (regularizers.py:260-273)
        tau_reg = 4.0 if self.tv_lambda is not None else 0.0
        groups, q_groups = None, None
        if self.regularizer is not None:
            groups, q_groups, tau_reg = self._groups(A, torch.empty(0, device=A.device))
        tau = _precondition(A, A.bp(torch.ones_like(b)).clone(), A.plan.vol_start, tau_reg, self.relaxation)
        sigma = _precondition(A, A.fp(torch.ones_like(tau)).clone(), A.plan.sino_start, 0.0, self.relaxation)
        x = torch.zeros_like(tau) if x0 is None else _arrays.to_dev(x0)[0].reshape(-1).clone()
        q = None if self.tv_lambda is None else torch.zeros(2 * A.vol_elems, dtype=torch.float32, device=x.device)
        p = torch.zeros_like(b)
        if x0 is not None and dual_init:
            st = _native.stream()
            ax = A.fp(x)
            _native.call("chip_pdhg_dual_init", p.data_ptr(), ax.data_ptr(), b.data_ptr(), None, None, p.numel(),
                         int(self.data_term == "kl"), st)
            if q is not None:
                for _ in range(8):  # q <- proj(q + grad x / 2) -> lambda grad x / |grad x| where grad x != 0
                    _native.call("chip_batch_tv_dual", A.plan._h, q.data_ptr(), x.data_ptr(), float(self.tv_lambda), st)
        return dict(b=b, tau=tau, sigma=sigma, x=x, xbar=x.clone(), p=p, q=q,
                    ax=torch.empty_like(b), upd=torch.empty_like(tau), groups=groups, q_groups=q_groups)

    def iterate(self, A: BatchedProjectorXRF, s: dict, iterations: int, lower_limit=None, upper_limit=None):
        hl, lo = _limit(lower_limit)
        hu, hi = _limit(upper_limit)
        dual = "chip_pdhg_dual_kl" if self.data_term == "kl" else "chip_pdhg_dual_l2"
        row_op = _native.CHIP_ROW_PDHG_KL if self.data_term == "kl" else _native.CHIP_ROW_PDHG_L2
        fused = A.fused_data_step()

        def step():
            st = _native.stream()
            if fused:  # on-the-fly maps: fp, dual step and bp per angle block, the maps built once
                A.fp_row_bp(s["xbar"], row_op, s["p"], s["b"], s["sigma"], s["ax"], s["upd"])
            else:
                A.fp(s["xbar"], s["ax"])
                _native.call(dual, s["p"].data_ptr(), s["ax"].data_ptr(), s["sigma"].data_ptr(), s["b"].data_ptr(),
                             None, None, s["p"].numel(), st)
            if s["q"] is not None:
                _native.call("chip_batch_tv_dual", A.plan._h, s["q"].data_ptr(), s["xbar"].data_ptr(),
                             float(self.tv_lambda), st)
            if s["groups"] is not None:
                for cfg, start, n_el, qg in s["groups"]:
                    Solver._grad_dual(cfg, qg, s["xbar"][start:start + n_el], st)
            if not fused:
                A.bp(s["p"], s["upd"])
            if s["groups"] is not None:
                for cfg, start, n_el, qg in s["groups"]:
                    Solver._grad_primal(cfg, s["upd"][start:start + n_el], qg, False, st)
            _native.call("chip_batch_pdhg_primal", A.plan._h, s["x"].data_ptr(), s["xbar"].data_ptr(),
                         s["upd"].data_ptr(), _native.ptr(s["q"]), s["tau"].data_ptr(), hl, lo, hu, hi, None, st)

        key = ("pdhg", hl, lo, hu, hi)
        loop = s.get(key)
        if loop is None:
            loop = s[key] = _GraphedLoop(step, self.graph_chunk, self.use_graph)
        loop.run(iterations)
        return s

    def __call__(self, A, b, iterations: int, x0=None, lower_limit=None, upper_limit=None, dual_init: bool = False):
        s = self.prepare(A, b, x0, dual_init=dual_init)
        self.iterate(A, s, iterations, lower_limit, upper_limit)
        return s["x"]


def time_iterations(fn, iterations: int) -> float:
    """Seconds for ``fn(iterations)`` with the device drained on both sides."""
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fn(iterations)
    torch.cuda.synchronize()
    return time.perf_counter() - t0
