"""Pins the CPU restatement of the warm-start U-Net (oracle/unet.py).

* the backward pass equals central finite differences of the forward pass
  (float64, no bfloat16 rounding) for weights and biases of every layer kind
  and for the head;
* conv3x3 equals a direct quadruple sum; its data gradient is the adjoint of
  the convolution (<conv(x), y> = <x, conv_bwd(y)>);
* pooling and upsampling are each other's adjoints up to the factor 1/4;
* bfloat16 rounding: ties to even, exact values kept;
* the network has the documented size (2.1 M parameters at 4 levels x 32).
"""

import numpy as np

from oracle import unet as ou


def _params(levels, base, rng, scale=0.3):
    p = {}
    for name, ci, co in ou.layer_shapes(levels, base):
        p[name] = (rng.standard_normal((9, co, ci)) * scale, rng.standard_normal(co) * 0.1)
    return p


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2**-8, 1.0 + 2**-7, 1.0 + 3 * 2**-8, -2.5, 0.0], dtype=np.float32)
    r = ou.bf16_round(x)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: to even (1); 1 + 3 * 2^-8 ties up to 1 + 2^-6
    assert list(r) == [1.0, 1.0, 1.0 + 2**-7, 1.0 + 2**-6, -2.5, 0.0]


def test_conv3x3_direct_sum_and_adjoint():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2, 5, 6, 3))
    w = rng.standard_normal((9, 4, 3))
    b = rng.standard_normal(4)
    y = ou.conv3x3(x, w, b)
    n, h, wd, _ = x.shape
    ref = np.zeros((n, h, wd, 4)) + b
    for tap in range(9):
        dh, dw = tap // 3 - 1, tap % 3 - 1
        for i in range(h):
            for j in range(wd):
                if 0 <= i + dh < h and 0 <= j + dw < wd:
                    ref[:, i, j, :] += x[:, i + dh, j + dw, :] @ w[tap].T
    assert np.allclose(y, ref, rtol=0, atol=1e-12)
    dy = rng.standard_normal(y.shape)
    dx, dw_, db = ou.conv3x3_bwd(x, w, dy)
    lhs = np.sum((ou.conv3x3(x, w, np.zeros(4))) * dy)
    assert np.isclose(lhs, np.sum(x * dx), rtol=1e-12)
    assert np.isclose(lhs, np.sum(w * dw_), rtol=1e-12)
    assert np.allclose(db, dy.sum(axis=(0, 1, 2)))


def test_pool_upsample_adjoint():
    rng = np.random.default_rng(2)
    x = rng.standard_normal((2, 4, 6, 3))
    y = rng.standard_normal((2, 2, 3, 3))
    assert np.isclose(np.sum(ou.avgpool2(x) * y), np.sum(x * ou.avgpool2_bwd(y)))
    assert np.isclose(np.sum(ou.upsample2(y) * x), np.sum(y * ou.upsample2_bwd(x)))
    assert np.allclose(ou.avgpool2_bwd(y), 0.25 * ou.upsample2(y))


def test_backward_matches_finite_differences():
    levels, base = 3, 2
    rng = np.random.default_rng(3)
    params = _params(levels, base, rng)
    hw = rng.standard_normal(base) * 0.5
    net = ou.UNetOracle(levels, base, params, hw, 0.1, bf16=False)
    x = rng.standard_normal((2, 8, 8))
    tgt = rng.standard_normal((2, 8, 8))
    out = net.forward(x)
    grads = net.backward(out, tgt)

    def loss():
        return net.loss(net.forward(x), tgt)

    eps = 1e-6
    checked = 0
    for name in ["enc0a", "enc1b", "enc2a", "up1", "dec0a", "dec1b"]:
        w, b = net.params[name]
        for idx in [(0, 0, 0), (4, w.shape[1] - 1, w.shape[2] - 1), (8, 1 % w.shape[1], 0)]:
            old = w[idx]
            w[idx] = old + eps
            lp = loss()
            w[idx] = old - eps
            lm = loss()
            w[idx] = old
            fd = (lp - lm) / (2 * eps)
            assert abs(fd - grads[name][0][idx]) <= 1e-6 * max(1.0, abs(fd)), (name, idx, fd, grads[name][0][idx])
            checked += 1
        old = b[0]
        b[0] = old + eps
        lp = loss()
        b[0] = old - eps
        lm = loss()
        b[0] = old
        assert abs((lp - lm) / (2 * eps) - grads[name][1][0]) <= 1e-6
    old = net.head_w[1]
    net.head_w[1] = old + eps
    lp = loss()
    net.head_w[1] = old - eps
    lm = loss()
    net.head_w[1] = old
    assert abs((lp - lm) / (2 * eps) - grads["head"][0][1]) <= 1e-6
    assert checked == 18


def test_network_size():
    shapes = ou.layer_shapes(4, 32)
    n = sum(9 * ci * co + co for _, ci, co in shapes) + 33
    assert 2.0e6 < n < 2.3e6
