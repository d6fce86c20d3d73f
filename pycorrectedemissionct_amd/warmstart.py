"""Learned warm start for PDHG: a small U-Net trained by this library's own kernels.

No counterpart in the reference. A short SIRT reconstruction of a slice goes
in; the network predicts the PDHG-TV solution; PDHG starts from it (plain
PDHG remains the fallback and the judge of convergence).

Network (``levels`` = 4, ``base`` = 32 gives 2.1 M parameters): every
convolution is 3x3 with "same" zero padding, bias and ReLU.

* encoder level l: [2x2 average pool if l > 0] -> conv -> conv (e_l,
  ``base * 2**l`` channels)
* decoder level l = levels-2 .. 0: 2x nearest upsampling -> conv (u_l) ->
  conv of concat(u_l, e_l) -> conv (d_l)
* head: out = x + b + w . d_0 (1x1, float32, a correction on the input)

Activations and weights are bfloat16 with float32 accumulation on the matrix
cores (``chip_nn_conv3x3``: implicit GEMM on v_mfma_f32_16x16x16_bf16);
master weights, gradients and Adam moments are float32. Images are scaled by
the per-image maximum of |x_sirt| on the way in and back on the way out.

``make_pairs`` builds (SIRT-20, PDHG-TV reference) training pairs with the
batched projector and solvers; ``train`` runs Adam; ``warm_start`` seeds
batched PDHG.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _arrays, _native
from . import batch as _batch
from . import testing

BF16 = torch.int16  # bfloat16 bit patterns travel as int16 tensors


def layer_shapes(levels: int, base: int):
    """[(name, cin, cout)] of the 3x3 convolutions in execution order."""
    out = []
    cin = 1
    for l in range(levels):
        c = base * 2**l
        out += [(f"enc{l}a", cin, c), (f"enc{l}b", c, c)]
        cin = c
    for l in range(levels - 2, -1, -1):
        c = base * 2**l
        out += [(f"up{l}", 2 * c, c), (f"dec{l}a", 2 * c, c), (f"dec{l}b", c, c)]
    return out


def _p(t) -> int | None:
    return None if t is None else t.data_ptr()


class _Conv:
    def __init__(self, name, cin, cout, dev, gen):
        self.name, self.cin, self.cout = name, cin, cout
        std = float(np.sqrt(2.0 / (9 * cin)))
        self.w = (torch.randn((9, cout, cin), generator=gen, dtype=torch.float32) * std).to(dev)
        self.b = torch.zeros(cout, dtype=torch.float32, device=dev)
        self.w16 = torch.empty((9, cout, cin), dtype=BF16, device=dev)
        self.wt16 = torch.empty((9, cin, cout), dtype=BF16, device=dev)  # flipped, for the data gradient
        self.gw = torch.zeros_like(self.w)
        self.gb = torch.zeros_like(self.b)
        self.mw, self.vw = torch.zeros_like(self.w), torch.zeros_like(self.w)
        self.mb, self.vb = torch.zeros_like(self.b), torch.zeros_like(self.b)

    def refresh(self, st):
        _native.call("chip_nn_to_bf16", _p(self.w), _p(self.w16), self.w.numel(), st)
        _native.call("chip_nn_flip_weights", _p(self.w16), self.cout, self.cin, _p(self.wt16), st)


class UNet:
    """The warm-start network; parameters live on the current device."""

    def __init__(self, levels: int = 4, base: int = 32, seed: int = 0, head_std: float = 1e-3):
        self.levels, self.base = levels, base
        self.dev = _arrays.device()
        gen = torch.Generator().manual_seed(seed)
        self.convs = {name: _Conv(name, ci, co, self.dev, gen) for name, ci, co in layer_shapes(levels, base)}
        self.head = (torch.randn(base + 1, generator=gen, dtype=torch.float32) * head_std).to(self.dev)
        self.head[base] = 0.0
        self.ghead = torch.zeros_like(self.head)
        self.mh, self.vh = torch.zeros_like(self.head), torch.zeros_like(self.head)
        self.step_count = 0
        self._part = torch.empty(1 << 24, dtype=torch.float32, device=self.dev)  # reduction scratch
        self._dpart = torch.empty(1024, dtype=torch.float64, device=self.dev)
        self._loss = torch.zeros(1, dtype=torch.float64, device=self.dev)
        self.refresh()

    # ---- parameters ----------------------------------------------------------------
    def refresh(self) -> None:
        st = _native.stream()
        for c in self.convs.values():
            c.refresh(st)

    @property
    def n_params(self) -> int:
        return sum(c.w.numel() + c.b.numel() for c in self.convs.values()) + self.head.numel()

    def state_dict(self) -> dict:
        d = {f"{k}.w": c.w.cpu().numpy() for k, c in self.convs.items()}
        d.update({f"{k}.b": c.b.cpu().numpy() for k, c in self.convs.items()})
        d["head"] = self.head.cpu().numpy()
        d["meta"] = np.array([self.levels, self.base, self.step_count])
        return d

    def load_state_dict(self, d: dict) -> None:
        for k, c in self.convs.items():
            c.w.copy_(torch.as_tensor(np.asarray(d[f"{k}.w"], np.float32)))
            c.b.copy_(torch.as_tensor(np.asarray(d[f"{k}.b"], np.float32)))
        self.head.copy_(torch.as_tensor(np.asarray(d["head"], np.float32)))
        self.refresh()

    def save(self, path: str, inference_only: bool = False) -> None:
        """Weights as .npz; ``inference_only`` stores the conv weights rounded to the bfloat16
        values inference uses (exactly the same network output, about half the compressed size)."""
        d = self.state_dict()
        if inference_only:
            for k, c in self.convs.items():
                bits = c.w16.view(torch.int16).cpu().numpy().astype(np.uint16).astype(np.uint32) << 16
                d[f"{k}.w"] = bits.view(np.float32).reshape(d[f"{k}.w"].shape)
        np.savez_compressed(path, **d)

    @classmethod
    def load(cls, path: str) -> "UNet":
        d = dict(np.load(path, allow_pickle=False))
        levels, base = int(d["meta"][0]), int(d["meta"][1])
        net = cls(levels, base)
        net.load_state_dict(d)
        return net

    # ---- forward / backward ------------------------------------------------------------
    def _conv(self, name, x1, c1, x2, c2, n, h, w, relu=True, out=None):
        c = self.convs[name]
        y = torch.empty((n, h, w, c.cout), dtype=BF16, device=self.dev) if out is None else out
        _native.call("chip_nn_conv3x3", _p(x1), c1, _p(x2), c2, _p(c.w16), _p(c.b), _p(y), c.cout, None, n, h, w,
                     c.cout, int(relu), _native.stream())
        return y

    def forward(self, x: torch.Tensor, keep: bool = False) -> torch.Tensor:
        """x [N, H, W] float32 (H, W multiples of 2**(levels-1)) -> out [N, H, W] float32."""
        st = _native.stream()
        x = x.contiguous()
        n, H, W = x.shape
        if H % 2 ** (self.levels - 1) or W % 2 ** (self.levels - 1):
            raise ValueError(f"H and W must be multiples of {2 ** (self.levels - 1)}")
        t = {"x": x, "shape": (n, H, W)}
        x0 = torch.empty((n, H, W, 1), dtype=BF16, device=self.dev)
        _native.call("chip_nn_to_bf16", _p(x), _p(x0), x.numel(), st)
        h, hh, ww, ch = x0, H, W, 1
        for l in range(self.levels):
            c = self.base * 2**l
            if l > 0:
                p = torch.empty((n, hh // 2, ww // 2, ch), dtype=BF16, device=self.dev)
                _native.call("chip_nn_avgpool2", _p(h), _p(p), n, hh, ww, ch, 0, st)
                h, hh, ww = p, hh // 2, ww // 2
            t[f"in{l}"] = (h, ch)
            a = self._conv(f"enc{l}a", h, ch, None, 0, n, hh, ww)
            e = self._conv(f"enc{l}b", a, c, None, 0, n, hh, ww)
            t[f"a{l}"], t[f"e{l}"], t[f"hw{l}"] = a, e, (hh, ww)
            h, ch = e, c
        for l in range(self.levels - 2, -1, -1):
            c = self.base * 2**l
            U = torch.empty((n, 2 * hh, 2 * ww, ch), dtype=BF16, device=self.dev)
            _native.call("chip_nn_upsample2", _p(h), _p(U), n, hh, ww, ch, 0, st)
            hh, ww = 2 * hh, 2 * ww
            u = self._conv(f"up{l}", U, ch, None, 0, n, hh, ww)
            c1 = self._conv(f"dec{l}a", u, c, t[f"e{l}"], c, n, hh, ww)
            d = self._conv(f"dec{l}b", c1, c, None, 0, n, hh, ww)
            t[f"U{l}"], t[f"u{l}"], t[f"c{l}"], t[f"d{l}"] = U, u, c1, d
            h, ch = d, c
        out = torch.empty((n, H, W), dtype=torch.float32, device=self.dev)
        _native.call("chip_nn_head_fwd", _p(h), self.base, _p(self.head), _p(x), _p(out), n * H * W, st)
        t["feat"] = h
        if keep:
            self._t = t
        return out

    def _conv_bwd(self, name, x1, c1, x2, c2, y, dy, n, h, w, need_dx=True, split=None):
        """Gradient through ReLU(output y) and conv `name` of concat(x1, x2); fills the
        conv's gw / gb and returns dx (or (dx1, dx2) with `split`)."""
        st = _native.stream()
        c = self.convs[name]
        _native.call("chip_nn_relu_mask", _p(dy), _p(y), dy.numel(), st)
        _native.call("chip_nn_conv3x3_wgrad", _p(dy), _p(x1), c1, _p(x2), c2, n, h, w, c.cout, _p(self._part),
                     self._part.numel(), _p(c.gw), st)
        _native.call("chip_nn_bias_grad", _p(dy), n * h * w, c.cout, _p(self._part), self._part.numel(), _p(c.gb),
                     st)
        if not need_dx:
            return None
        if split is None:
            dx = torch.empty((n, h, w, c.cin), dtype=BF16, device=self.dev)
            _native.call("chip_nn_conv3x3", _p(dy), c.cout, None, 0, _p(c.wt16), None, _p(dx), c.cin, None, n, h, w,
                         c.cin, 0, st)
            return dx
        d1 = torch.empty((n, h, w, split), dtype=BF16, device=self.dev)
        d2 = torch.empty((n, h, w, c.cin - split), dtype=BF16, device=self.dev)
        _native.call("chip_nn_conv3x3", _p(dy), c.cout, None, 0, _p(c.wt16), None, _p(d1), split, _p(d2), n, h, w,
                     c.cin, 0, st)
        return d1, d2

    def backward(self, dout: torch.Tensor) -> None:
        """Gradients of all parameters for the upstream gradient dout [N, H, W] of the last forward(keep=True)."""
        st = _native.stream()
        t = self._t
        n, H, W = t["shape"]
        dout = dout.contiguous()
        dh = torch.empty((n, H, W, self.base), dtype=BF16, device=self.dev)
        _native.call("chip_nn_head_bwd", _p(dout), _p(t["feat"]), self.base, _p(self.head), _p(dh), _p(self._part),
                     self._part.numel(), _p(self.ghead), n * H * W, st)
        dskip = {}
        hh, ww = H, W
        for l in range(0, self.levels - 1):
            c = self.base * 2**l
            dc = self._conv_bwd(f"dec{l}b", t[f"c{l}"], c, None, 0, t[f"d{l}"], dh, n, hh, ww)
            du, dskip[l] = self._conv_bwd(f"dec{l}a", t[f"u{l}"], c, t[f"e{l}"], c, t[f"c{l}"], dc, n, hh, ww,
                                          split=c)
            dU = self._conv_bwd(f"up{l}", t[f"U{l}"], 2 * c, None, 0, t[f"u{l}"], du, n, hh, ww)
            dh = torch.empty((n, hh // 2, ww // 2, 2 * c), dtype=BF16, device=self.dev)
            _native.call("chip_nn_upsample2", _p(dU), _p(dh), n, hh // 2, ww // 2, 2 * c, 1, st)
            hh, ww = hh // 2, ww // 2
        for l in range(self.levels - 1, -1, -1):
            c = self.base * 2**l
            hh, ww = t[f"hw{l}"]
            de = dh
            if l < self.levels - 1:
                _native.call("chip_nn_add_bf16", _p(de), _p(dskip[l]), de.numel(), st)
            da = self._conv_bwd(f"enc{l}b", t[f"a{l}"], c, None, 0, t[f"e{l}"], de, n, hh, ww)
            hin, cin = t[f"in{l}"]
            dinp = self._conv_bwd(f"enc{l}a", hin, cin, None, 0, t[f"a{l}"], da, n, hh, ww, need_dx=l > 0)
            if l > 0:
                dh = torch.empty((n, 2 * hh, 2 * ww, cin), dtype=BF16, device=self.dev)
                _native.call("chip_nn_avgpool2", _p(dinp), _p(dh), n, 2 * hh, 2 * ww, cin, 1, st)

    def loss_and_grad(self, out: torch.Tensor, tgt: torch.Tensor):
        """MSE of out against tgt (device double scalar) and its gradient (float32)."""
        dout = torch.empty_like(out)
        _native.call("chip_nn_mse", _p(out), _p(tgt.contiguous()), out.numel(), _p(dout), _p(self._dpart),
                     self._dpart.numel(), _p(self._loss), _native.stream())
        return self._loss, dout

    def adam(self, lr: float = 1e-3, beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8) -> None:
        st = _native.stream()
        self.step_count += 1
        k = self.step_count
        for c in self.convs.values():
            _native.call("chip_nn_adam", _p(c.w), _p(c.gw), _p(c.mw), _p(c.vw), c.w.numel(), lr, beta1, beta2, eps,
                         k, _p(c.w16), st)
            _native.call("chip_nn_adam", _p(c.b), _p(c.gb), _p(c.mb), _p(c.vb), c.b.numel(), lr, beta1, beta2, eps,
                         k, None, st)
            _native.call("chip_nn_flip_weights", _p(c.w16), c.cout, c.cin, _p(c.wt16), st)
        _native.call("chip_nn_adam", _p(self.head), _p(self.ghead), _p(self.mh), _p(self.vh), self.head.numel(), lr,
                     beta1, beta2, eps, k, None, st)

    def train_step(self, x: torch.Tensor, tgt: torch.Tensor, lr: float = 1e-3) -> torch.Tensor:
        out = self.forward(x, keep=True)
        loss, dout = self.loss_and_grad(out, tgt)
        self.backward(dout)
        self.adam(lr)
        self._t = None
        return loss.clone()

    def predict(self, x: torch.Tensor, chunk: int = 8) -> torch.Tensor:
        outs = [self.forward(x[i:i + chunk]) for i in range(0, x.shape[0], chunk)]
        return torch.cat(outs, dim=0)


# ---- data ---------------------------------------------------------------------------------------

def phantom_voronoi(n: int, rng: np.random.Generator, cells=(20, 200), intensities=(0.2, 1.0)) -> np.ndarray:
    """Piecewise-constant Voronoi cells inside the inscribed disc; a random third of the cells empty."""
    k = int(rng.integers(cells[0], cells[1] + 1))
    c = n // 2
    pts = rng.uniform(0, n, size=(k, 2))
    vals = rng.uniform(*intensities, size=k) * (rng.random(k) > 1 / 3)
    yy, xx = np.mgrid[:n, :n].astype(np.float64)
    d = (yy[..., None] - pts[:, 0]) ** 2 + (xx[..., None] - pts[:, 1]) ** 2
    img = vals[np.argmin(d, axis=-1)]
    img[(yy - c) ** 2 + (xx - c) ** 2 > (c - 1) ** 2] = 0
    return img.astype(np.float32)


def training_phantom(n: int, rng: np.random.Generator) -> np.ndarray:
    """Ellipses + blobs, Voronoi cells, or both (seeded)."""
    kind = rng.integers(3)
    if kind == 0:
        return testing.phantom_ellipses(n, int(rng.integers(5, 20)), rng, semi_axes=(n // 40 + 2, n // 5),
                                        n_blobs=int(rng.integers(0, 4)), blob_sigma=(n / 80, n / 25))
    if kind == 1:
        return phantom_voronoi(n, rng)
    f = 0.5 * phantom_voronoi(n, rng) + testing.phantom_ellipses(n, 5, rng, semi_axes=(n // 40 + 2, n // 6))
    return f.astype(np.float32)


def make_pairs(n_images: int, n: int = 256, n_angles: int = 360, sirt_its: int = 20, pdhg_its: int = 2000,
               tv_lambda: float = 1e-2, peak_range=(1e2, 1e4), seed: int = 0, batch: int = 32,
               data_term: str = "kl", det_angles=(np.pi / 2,), phantom: str = "training", keep_chunks: bool = True,
               log=None) -> dict:
    """(SIRT-``sirt_its``, PDHG-TV-``pdhg_its``) pairs of seeded phantoms, made on the GPU in batches.

    Each phantom has its own attenuation (mu_in = 0.02 (f > 0) + 0.005, mu_out = 0.05 (f > 0)),
    detectors at ``det_angles``, Poisson counts at a peak drawn from ``peak_range``. ``phantom``:
    "training" (ellipses + blobs, Voronoi cells, or both) or "ellipses" (30 ellipses + 5 Gaussian
    blobs, the PDHG benchmark's family). Returns device tensors x_sirt, x_ref, phantom
    [n_images, n, n] and, with ``keep_chunks``, each batch's projector and data to re-solve (the
    projectors hold their maps: training sets leave them out).
    """
    rng = np.random.default_rng(seed)
    ang = testing.angles_uniform(n_angles)
    xs, xr, fs, items = [], [], [], []
    for i0 in range(0, n_images, batch):
        nb = min(batch, n_images - i0)
        if phantom == "ellipses":
            f = [testing.phantom_ellipses(n, 30, rng, n_blobs=5) for _ in range(nb)]
        else:
            f = [training_phantom(n, rng) for _ in range(nb)]
        mus = [testing.xrf_attenuation(fi) for fi in f]
        A = _batch.BatchedProjectorXRF([n] * nb, ang, [m[0] for m in mus], [m[1] for m in mus],
                                       angles_detectors_rad=list(det_angles), bp_mode="adjoint")
        clean = A.fp(A.pack_vol(f))
        sinos = A.unpack_sino(clean)
        b = A.pack_sino([testing.poisson_data(s.cpu().numpy(), float(rng.uniform(*peak_range)), rng) for s in sinos])
        x_s = _batch.BatchedSIRT()(A, b, sirt_its)
        x_r = _batch.BatchedPDHG(data_term=data_term, tv_lambda=tv_lambda)(A, b, pdhg_its)
        xs.append(torch.stack(A.unpack_vol(x_s)))
        xr.append(torch.stack(A.unpack_vol(x_r)))
        fs.append(torch.as_tensor(np.stack(f), device=x_s.device))
        if keep_chunks:
            items.append(dict(A=A, b=b, mus=mus))
        del A
        if log:
            log(f"pairs {i0 + nb}/{n_images}")
    return dict(x_sirt=torch.cat(xs), x_ref=torch.cat(xr), phantom=torch.cat(fs), chunks=items,
                angles=ang, n=n, tv_lambda=tv_lambda, data_term=data_term)


def _scale(x: torch.Tensor) -> torch.Tensor:
    """Per-image normalisation constant: max |x| (1 where the image is 0)."""
    s = x.abs().amax(dim=(1, 2), keepdim=True)
    return torch.where(s > 0, s, torch.ones_like(s))


def train(net: UNet, x_in: torch.Tensor, x_tgt: torch.Tensor, epochs: int = 10, batch: int = 8, lr: float = 1e-3,
          seed: int = 0, log=None) -> list[float]:
    """Adam on the MSE between net(x_in / s) and x_tgt / s, s = max |x_in| per image; returns epoch losses."""
    s = _scale(x_in)
    xi, xt = (x_in / s).contiguous(), (x_tgt / s).contiguous()
    g = torch.Generator().manual_seed(seed)
    losses = []
    for ep in range(epochs):
        perm = torch.randperm(xi.shape[0], generator=g).to(xi.device)
        tot = torch.zeros((), dtype=torch.float64, device=xi.device)
        nb = 0
        for i in range(0, xi.shape[0] - batch + 1, batch):
            idx = perm[i:i + batch]
            tot += net.train_step(xi[idx].contiguous(), xt[idx].contiguous(), lr)[0]
            nb += 1
        losses.append(float(tot) / max(nb, 1))
        if log:
            log(f"epoch {ep + 1}/{epochs}: loss {losses[-1]:.4e}")
    return losses


def warm_start(net: UNet, x_sirt: torch.Tensor) -> torch.Tensor:
    """The network's estimate of the PDHG solution for SIRT reconstructions x_sirt [N, H, W]."""
    s = _scale(x_sirt)
    return net.predict((x_sirt / s).contiguous()) * s


def iterations_to_tolerance(solver, A, b, x_ref_flat: torch.Tensor, x0, tols, max_its: int, check_every: int = 10,
                            dual_init: bool = False):
    """Batched PDHG from x0: per slice, the first multiple of ``check_every`` iterations at which
    ||x - x_ref|| / ||x_ref|| <= tol, for each tol (``max_its + 1`` where not reached)."""
    st = solver.prepare(A, b, x0=x0, dual_init=dual_init)
    vs = A.plan.vol_start
    n_sl = len(vs) - 1
    ref_norm = [float(torch.linalg.vector_norm(x_ref_flat[int(vs[i]):int(vs[i + 1])])) for i in range(n_sl)]
    hit = np.full((len(tols), n_sl), max_its + 1)
    done = 0
    errs0 = None
    while done < max_its and np.any(hit > max_its):
        solver.iterate(A, st, check_every)
        done += check_every
        d = st["x"] - x_ref_flat
        err = np.array([float(torch.linalg.vector_norm(d[int(vs[i]):int(vs[i + 1])])) / max(ref_norm[i], 1e-30)
                        for i in range(n_sl)])
        if errs0 is None:
            errs0 = err
        for k, tol in enumerate(tols):
            newly = (hit[k] > max_its) & (err <= tol)
            hit[k][newly] = done
    return hit


def evaluate(net: UNet, test: dict, tols=(1e-1, 5e-2, 2e-2, 1e-2, 5e-3, 1e-3), max_its: int = 2000,
             check_every: int = 10, variants=("cold", "warm", "warm_dual"), log=None) -> dict:
    """Cold vs warm-started batched PDHG on the held-out chunk(s) of ``test``: iterations to each
    tolerance against the reference ``test["x_ref"]``, for the warm start alone ("warm") and with
    duals initialised from it ("warm_dual"), and the slices/s ratios counting the warm start's
    SIRT-20 + network time."""
    runs = {k: [] for k in variants}
    t_it = t_sirt = t_net = 0.0
    n_sl = 0
    for ch in test["chunks"]:
        A, b = ch["A"], ch["b"]
        k = len(A.sizes)
        x_ref = A.pack_vol(list(test["x_ref"][n_sl:n_sl + k]))
        solver = _batch.BatchedPDHG(data_term=test["data_term"], tv_lambda=test["tv_lambda"])
        sirt = _batch.BatchedSIRT()
        t_sirt += _batch.time_iterations(lambda its: sirt(A, b, its), 20)
        xs = torch.stack(A.unpack_vol(sirt(A, b, 20)))
        torch.cuda.synchronize()
        t_net += _batch.time_iterations(lambda _: warm_start(net, xs), 1)
        x0 = A.pack_vol(list(warm_start(net, xs)))
        st = solver.prepare(A, b)
        t_it += _batch.time_iterations(lambda its: solver.iterate(A, st, its), 200) / 200
        for key in variants:
            runs[key].append(iterations_to_tolerance(solver, A, b, x_ref, None if key == "cold" else x0, tols,
                                                     max_its, check_every, dual_init=key == "warm_dual"))
            if log:
                log(f"evaluated {key} on slices {n_sl}..{n_sl + k - 1}")
        n_sl += k
    its = {key: np.concatenate(v, axis=1) for key, v in runs.items()}
    n_ch = len(test["chunks"])
    res = {"slices": n_sl, "ms_per_iteration_batch": 1e3 * t_it / n_ch, "ms_sirt20_batch": 1e3 * t_sirt / n_ch,
           "ms_unet_batch": 1e3 * t_net / n_ch, "max_its": max_its, "per_tol": []}
    for i, tol in enumerate(tols):
        row = {"tol": tol}
        c_med = float(np.median(its["cold"][i]))
        row["cold_median_its"] = c_med
        row["cold_reached"] = int(np.sum(its["cold"][i] <= max_its))
        for key in (v for v in variants if v != "cold"):
            w_med = float(np.median(its[key][i]))
            row[f"{key}_median_its"] = w_med
            row[f"{key}_reached"] = int(np.sum(its[key][i] <= max_its))
            ok = c_med <= max_its and w_med <= max_its
            row[f"{key}_speedup"] = (c_med * t_it) / (w_med * t_it + t_sirt + t_net) if ok else None
        res["per_tol"].append(row)
    return res


def main() -> int:
    import argparse
    import json
    import sys
    import time

    ap = argparse.ArgumentParser(description="Train the warm-start U-Net and evaluate it against plain PDHG.")
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--angles", type=int, default=360)
    ap.add_argument("--train", type=int, default=384)
    ap.add_argument("--test", type=int, default=32)
    ap.add_argument("--pdhg-its", type=int, default=2000, help="PDHG iterations of the training targets")
    ap.add_argument("--ref-its", type=int, default=6000, help="PDHG iterations of the held-out references")
    ap.add_argument("--eval-its", type=int, default=4000, help="iteration cap of the cold / warm runs")
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--load", default=None, help="trained weights (.npz) instead of training")
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--levels", type=int, default=4)
    ap.add_argument("--base", type=int, default=32)
    ap.add_argument("--save", default=None)
    ap.add_argument("--detectors", type=int, default=1, help="1: one detector at 90 deg; 2: at +-90 deg")
    ap.add_argument("--peak", type=float, nargs=2, default=(1e2, 1e4), help="Poisson peak range")
    ap.add_argument("--phantom", choices=["training", "ellipses"], default="training")
    ap.add_argument("--pair-batch", type=int, default=32)
    ap.add_argument("--no-dual-init", action="store_true", help="evaluate without the dual-initialised variant")
    args = ap.parse_args()

    def log(msg):
        print(f"[warmstart] {msg}", file=sys.stderr, flush=True)

    det = (np.pi / 2,) if args.detectors == 1 else (np.pi / 2, -np.pi / 2)
    kw = dict(n=args.n, n_angles=args.angles, batch=args.pair_batch, det_angles=det, peak_range=tuple(args.peak),
              phantom=args.phantom)
    t0 = time.perf_counter()
    losses = [float("nan")]
    if args.load:
        net = UNet.load(args.load)
        t1 = t2 = time.perf_counter()
    else:
        parts = []
        step = args.pair_batch
        for i0 in range(0, args.train, step):
            parts.append(make_pairs(min(step, args.train - i0), seed=1000 + i0, pdhg_its=args.pdhg_its,
                                    keep_chunks=False, **kw))
            log(f"training pairs {min(i0 + step, args.train)}/{args.train} ({time.perf_counter() - t0:.0f} s)")
        x_in = torch.cat([p["x_sirt"] for p in parts])
        x_tgt = torch.cat([p["x_ref"] for p in parts])
        del parts
        t1 = time.perf_counter()
        net = UNet(args.levels, args.base, seed=0)
        losses = train(net, x_in, x_tgt, epochs=args.epochs, batch=args.batch, lr=args.lr, log=log)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        if args.save:
            net.save(args.save, inference_only=True)
            log(f"saved {args.save}; loss {losses[0]:.4e} -> {losses[-1]:.4e}")
    if args.test <= 0:
        print(json.dumps({"loss_first_last": [losses[0], losses[-1]], "s_pairs": t1 - t0, "s_training": t2 - t1,
                          "params": net.n_params, "train_pairs": args.train}), flush=True)
        return 0
    test = make_pairs(args.test, seed=99991, pdhg_its=args.ref_its, log=log, **kw)
    log(f"evaluating cold vs warm start against {args.ref_its}-iteration references")
    variants = ("cold", "warm") if args.no_dual_init else ("cold", "warm", "warm_dual")
    res = evaluate(net, test, max_its=args.eval_its, variants=variants, log=log)
    res.update({"n": args.n, "angles": args.angles, "train_pairs": args.train, "test_slices": args.test,
                "train_target_its": args.pdhg_its, "test_reference_its": args.ref_its, "epochs": args.epochs,
                "params": net.n_params, "loss_first_last": [losses[0], losses[-1]], "s_pairs": t1 - t0,
                "s_training": t2 - t1})
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
