"""CPU restatement of the warm-start U-Net, forward and backward (test infrastructure).

The network (no counterpart in the reference; it seeds PDHG): ``levels``
resolution levels with ``base * 2**l`` channels; every convolution is 3x3 with
"same" zero padding, bias and ReLU.

* encoder level l: [2x2 average pool if l > 0] -> conv -> conv  (output e_l)
* decoder level l (from levels-2 down to 0): 2x nearest upsampling -> conv to
  base * 2**l channels (u_l) -> conv of concat(u_l, e_l) -> conv (output d_l)
* head: out = x + b + sum_c w[c] d_0[c] (1x1, float32, residual on the input)
* loss: mean squared error against the target.

Arithmetic is float64. With ``bf16=True`` every tensor the GPU stores in
bfloat16 (inputs, weights, activations after each layer, activation
gradients) is rounded to bfloat16 (round to nearest even) at the same points,
so the GPU's float32-accumulated results can be compared closely.

Pinned by ``tests/test_oracle_unet.py``: the backward pass against central
finite differences of the forward pass (float64, no rounding), the
convolution against a direct sum, and pooling / upsampling adjointness.
"""

from __future__ import annotations

import numpy as np


def bf16_round(x) -> np.ndarray:
    """float -> nearest bfloat16 (ties to even), returned as float64."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


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


def conv3x3(x, w, b):
    """x [N,H,W,Ci], w [9,Co,Ci], b [Co] -> [N,H,W,Co] (zero padding)."""
    n, h, wd, ci = x.shape
    xp = np.zeros((n, h + 2, wd + 2, ci))
    xp[:, 1:-1, 1:-1] = x
    y = np.zeros((n, h, wd, w.shape[1])) + b
    for tap in range(9):
        dh, dw = tap // 3, tap % 3
        y += xp[:, dh:dh + h, dw:dw + wd, :] @ w[tap].T
    return y


def conv3x3_bwd(x, w, dy):
    """Gradients of conv3x3 w.r.t. x, w and b for upstream dy."""
    n, h, wd, ci = x.shape
    xp = np.zeros((n, h + 2, wd + 2, ci))
    xp[:, 1:-1, 1:-1] = x
    dxp = np.zeros_like(xp)
    dw = np.zeros_like(w)
    for tap in range(9):
        dh, dwd = tap // 3, tap % 3
        dw[tap] = dy.reshape(-1, dy.shape[-1]).T @ xp[:, dh:dh + h, dwd:dwd + wd, :].reshape(-1, ci)
        dxp[:, dh:dh + h, dwd:dwd + wd, :] += dy @ w[tap]
    return dxp[:, 1:-1, 1:-1], dw, dy.sum(axis=(0, 1, 2))


def avgpool2(x):
    n, h, w, c = x.shape
    return x.reshape(n, h // 2, 2, w // 2, 2, c).mean(axis=(2, 4))


def avgpool2_bwd(dy):
    return np.repeat(np.repeat(dy, 2, axis=1), 2, axis=2) * 0.25


def upsample2(x):
    return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)


def upsample2_bwd(dy):
    n, h, w, c = dy.shape
    return dy.reshape(n, h // 2, 2, w // 2, 2, c).sum(axis=(2, 4))


class UNetOracle:
    """Parameters: dict name -> (w [9,Co,Ci], b [Co]) and head (wh [C], bh)."""

    def __init__(self, levels: int, base: int, params: dict, head_w, head_b, bf16: bool = True):
        self.levels, self.base = levels, base
        self.params = {k: (np.asarray(w, np.float64), np.asarray(b, np.float64)) for k, (w, b) in params.items()}
        self.head_w = np.asarray(head_w, np.float64)
        self.head_b = float(head_b)
        self.bf16 = bf16

    def _r(self, x):
        return bf16_round(x) if self.bf16 else x

    def _w(self, name):
        w, b = self.params[name]
        return (bf16_round(w) if self.bf16 else w), b

    def _conv(self, name, x):
        w, b = self._w(name)
        return self._r(np.maximum(conv3x3(x, w, b), 0.0))

    def forward(self, x_in):
        """x_in [N,H,W] -> out [N,H,W]; keeps what backward needs."""
        t = {}
        x0 = self._r(np.asarray(x_in, np.float64)[..., None])
        h = x0
        for l in range(self.levels):
            if l > 0:
                h = self._r(avgpool2(h))
            t[f"in{l}"] = h
            a = self._conv(f"enc{l}a", h)
            e = self._conv(f"enc{l}b", a)
            t[f"a{l}"], t[f"e{l}"] = a, e
            h = e
        for l in range(self.levels - 2, -1, -1):
            U = upsample2(h)
            u = self._conv(f"up{l}", U)
            cat = np.concatenate([u, t[f"e{l}"]], axis=-1)
            c1 = self._conv(f"dec{l}a", cat)
            d = self._conv(f"dec{l}b", c1)
            t[f"U{l}"], t[f"u{l}"], t[f"cat{l}"], t[f"c{l}"], t[f"d{l}"] = U, u, cat, c1, d
            h = d
        t["feat"] = h
        out = np.asarray(x_in, np.float64) + self.head_b + h @ self.head_w
        if self.bf16:
            out = out.astype(np.float32).astype(np.float64)
        self.t = t
        return out

    def loss(self, out, tgt):
        return float(np.mean((out - tgt) ** 2))

    def _conv_bwd(self, name, x, y, dy, grads):
        """Through ReLU (output y) and the conv of input x; returns dx (rounded)."""
        dpre = np.where(y > 0, dy, 0.0)
        w, _ = self._w(name)
        dx, dw, db = conv3x3_bwd(x, w, dpre)
        grads[name] = (dw, db)
        return self._r(dx)

    def backward(self, out, tgt):
        """Gradients of the MSE loss: dict name -> (dw, db), plus 'head' -> (dwh, dbh)."""
        t = self.t
        grads = {}
        dout = 2.0 * (out - tgt) / out.size
        if self.bf16:
            dout = dout.astype(np.float32).astype(np.float64)
        feat = t["feat"]
        grads["head"] = (np.einsum("nhwc,nhw->c", feat, dout), float(dout.sum()))
        dh = self._r(dout[..., None] * self.head_w)
        dskip = {}
        for l in range(0, self.levels - 1):
            dc = self._conv_bwd(f"dec{l}b", t[f"c{l}"], t[f"d{l}"], dh, grads)
            dcat = self._conv_bwd(f"dec{l}a", t[f"cat{l}"], t[f"c{l}"], dc, grads)
            c = self.base * 2**l
            du, dskip[l] = dcat[..., :c], dcat[..., c:]
            dU = self._conv_bwd(f"up{l}", t[f"U{l}"], t[f"u{l}"], du, grads)
            dh = self._r(upsample2_bwd(dU))
        for l in range(self.levels - 1, -1, -1):
            de = dh if l == self.levels - 1 else self._r(dh + dskip[l])
            da = self._conv_bwd(f"enc{l}b", t[f"a{l}"], t[f"e{l}"], de, grads)
            dinp = self._conv_bwd(f"enc{l}a", t[f"in{l}"], t[f"a{l}"], da, grads)
            if l > 0:
                dh = self._r(avgpool2_bwd(dinp))
        return grads


def adam_step(w, g, m, v, step, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (float64); returns new (w, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    w = w - lr * (m / (1 - b1**step)) / (np.sqrt(v / (1 - b2**step)) + eps)
    return w, m, v
