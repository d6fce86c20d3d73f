"""The warm-start U-Net's kernels against the CPU oracle (run with -m gpu).

Bars (bfloat16 storage, float32 accumulation on the matrix cores vs float64
with the same bfloat16 rounding points): single convolutions to one bfloat16
ulp (max-relative 8e-3); weight gradients of a convolution to 1e-4; the whole
network's output to 1e-3 and every parameter gradient to 2e-2 (relative l2:
a few bfloat16 rounding flips propagate); Adam to 1e-6. Training reduces the
loss, and a small trained network reaches a PDHG tolerance in fewer
iterations than the cold start on held-out phantoms.
"""

import numpy as np
import pytest

from oracle import unet as ou
from tests.conftest import rel_l2, rel_max

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pkg():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import pycorrectedemissionct_amd as p

    return p


def _bf16_bits(x):
    import torch

    r = np.ascontiguousarray(ou.bf16_round(x).astype(np.float32))
    return torch.as_tensor((r.view(np.uint32) >> 16).astype(np.uint16).view(np.int16)).cuda()


def _from_bits(t):
    b = t.cpu().numpy().astype(np.uint16).astype(np.uint32) << 16
    return b.view(np.float32).astype(np.float64)


@pytest.mark.parametrize("c1,c2,cout,o1,h,w", [(1, 0, 24, 24, 9, 13), (8, 8, 32, 32, 16, 16), (16, 8, 40, 24, 6, 10),
                                               # the input layer's own weight-gradient kernel (cout = 8 * 2^k)
                                               (1, 0, 32, 32, 37, 29),
                                               # channel counts of 32s: the 16x16x32 MFMA kernel, ragged tiles
                                               (32, 0, 64, 64, 9, 21), (64, 32, 32, 32, 8, 16),
                                               (32, 32, 64, 32, 5, 33), (128, 128, 256, 256, 4, 6),
                                               # several 8 x 16 tiles per weight-gradient workgroup, ragged edges
                                               (32, 0, 32, 32, 250, 200),
                                               # 16-row forward tiles (grids of >= 1024 tiles), ragged edges
                                               (32, 0, 64, 64, 520, 264), (64, 0, 32, 32, 264, 520)])
def test_conv3x3_and_wgrad(pkg, c1, c2, cout, o1, h, w):
    import torch

    rng = np.random.default_rng(900 + c1 + c2 + cout)
    n = 2
    x = rng.standard_normal((n, h, w, c1 + c2))
    wt = rng.standard_normal((9, cout, c1 + c2)) * 0.3
    b = rng.standard_normal(cout).astype(np.float32)
    xb, wb = ou.bf16_round(x), ou.bf16_round(wt)
    ref = ou.bf16_round(np.maximum(ou.conv3x3(xb, wb, b), 0))
    x1, x2 = _bf16_bits(x[..., :c1]), (_bf16_bits(x[..., c1:]) if c2 else None)
    y1 = torch.empty((n, h, w, o1), dtype=torch.int16, device="cuda")
    y2 = torch.empty((n, h, w, cout - o1), dtype=torch.int16, device="cuda") if o1 < cout else None
    L = pkg._native
    w16, bias = _bf16_bits(wt), torch.as_tensor(b).cuda()  # kept alive across the call
    L.call("chip_nn_conv3x3", x1.data_ptr(), c1, None if x2 is None else x2.data_ptr(), c2, w16.data_ptr(),
           bias.data_ptr(), y1.data_ptr(), o1, None if y2 is None else y2.data_ptr(), n, h, w, cout, 1, L.stream())
    got = _from_bits(y1) if y2 is None else np.concatenate([_from_bits(y1), _from_bits(y2)], axis=-1)
    e = rel_max(ref, got)
    print(f"conv c1={c1} c2={c2} cout={cout}: rel_max {e:.2e}")
    assert e < 8e-3
    # weight gradient against the oracle for an upstream gradient dy
    dy = ou.bf16_round(rng.standard_normal((n, h, w, cout)))
    _, dw_ref, db_ref = ou.conv3x3_bwd(xb, wb, dy)
    part = torch.empty(1 << 20, dtype=torch.float32, device="cuda")
    dw = torch.empty((9, cout, c1 + c2), dtype=torch.float32, device="cuda")
    db = torch.empty(cout, dtype=torch.float32, device="cuda")
    dyb = _bf16_bits(dy)
    L.call("chip_nn_conv3x3_wgrad", dyb.data_ptr(), x1.data_ptr(), c1, None if x2 is None else x2.data_ptr(), c2, n,
           h, w, cout, part.data_ptr(), part.numel(), dw.data_ptr(), L.stream())
    L.call("chip_nn_bias_grad", dyb.data_ptr(), n * h * w, cout, part.data_ptr(), part.numel(), db.data_ptr(),
           L.stream())
    assert rel_max(dw_ref, dw.cpu().numpy()) < 1e-4
    assert rel_max(db_ref, db.cpu().numpy()) < 1e-4


def _oracle_of(net):
    sd = net.state_dict()
    params = {k[:-2]: (sd[k], sd[k[:-2] + ".b"]) for k in sd if k.endswith(".w")}
    return ou.UNetOracle(net.levels, net.base, params, sd["head"][:-1], sd["head"][-1], bf16=True)


def test_unet_forward_backward_vs_oracle(pkg):
    import torch

    rng = np.random.default_rng(910)
    net = pkg.warmstart.UNet(levels=3, base=8, seed=1, head_std=0.3)
    x = rng.standard_normal((2, 16, 24)).astype(np.float32)
    tgt = rng.standard_normal((2, 16, 24)).astype(np.float32)
    xt, tt = torch.as_tensor(x).cuda(), torch.as_tensor(tgt).cuda()
    out = net.forward(xt, keep=True)
    loss, dout = net.loss_and_grad(out, tt)
    net.backward(dout)
    orc = _oracle_of(net)
    ref = orc.forward(x)
    e = rel_l2(ref, out.cpu().numpy())
    print(f"unet forward: rel_l2 {e:.2e}; loss {float(loss[0]):.6f} vs {orc.loss(ref, tgt):.6f}")
    assert e < 1e-3
    assert abs(float(loss[0]) - orc.loss(ref, tgt)) < 1e-3 * orc.loss(ref, tgt)
    grads = orc.backward(ref, tgt)
    worst = 0.0
    for name, c in net.convs.items():
        ew = rel_l2(grads[name][0], c.gw.cpu().numpy())
        eb = rel_l2(grads[name][1], c.gb.cpu().numpy())
        worst = max(worst, ew, eb)
        print(f"  {name}: dW rel_l2 {ew:.2e}, db rel_l2 {eb:.2e}")
    eh = rel_l2(np.append(grads["head"][0], grads["head"][1]), net.ghead.cpu().numpy())
    print(f"  head: rel_l2 {eh:.2e}; worst conv {worst:.2e}")
    assert worst < 2e-2 and eh < 2e-2


def test_adam_step_vs_oracle(pkg):
    import torch

    rng = np.random.default_rng(920)
    w, g = rng.standard_normal(1000), rng.standard_normal(1000)
    m, v = rng.standard_normal(1000) * 0.1, rng.random(1000) * 0.1
    wt, gt, mt, vt = (torch.as_tensor(a, dtype=torch.float32).cuda() for a in (w, g, m, v))
    L = pkg._native
    L.call("chip_nn_adam", wt.data_ptr(), gt.data_ptr(), mt.data_ptr(), vt.data_ptr(), 1000, 1e-3, 0.9, 0.999, 1e-8,
           3, None, L.stream())
    f32 = lambda a: float(np.float32(a))  # noqa: E731  (the kernel's float32 hyper-parameters)
    wr, mr, vr = ou.adam_step(w.astype(np.float32).astype(np.float64), g.astype(np.float32).astype(np.float64),
                              m.astype(np.float32).astype(np.float64), v.astype(np.float32).astype(np.float64), 3,
                              lr=f32(1e-3), b1=f32(0.9), b2=f32(0.999), eps=f32(1e-8))
    assert rel_max(wr, wt.cpu().numpy()) < 1e-6
    assert rel_max(mr, mt.cpu().numpy()) < 1e-6 and rel_max(vr, vt.cpu().numpy()) < 1e-6


def test_training_reduces_loss(pkg):
    import torch

    torch.manual_seed(0)
    rng = np.random.default_rng(930)
    n = 32
    xs = np.stack([pkg.testing.phantom_ellipses(n, 3, rng) for _ in range(16)])
    tgt = torch.as_tensor(xs).cuda()
    inp = tgt + 0.1 * torch.as_tensor(rng.standard_normal(xs.shape), dtype=torch.float32).cuda()
    net = pkg.warmstart.UNet(levels=3, base=8, seed=2)
    losses = pkg.warmstart.train(net, inp, tgt, epochs=40, batch=8, lr=3e-3)
    print(f"training loss: first {losses[0]:.4e}, last {losses[-1]:.4e}")
    assert losses[-1] < 0.5 * losses[0]


def test_warm_start_gain(pkg):
    """Trained on 64 pairs (SIRT-20 -> PDHG-TV-1000) of 64 x 64 phantoms, 90 angles; on 8
    held-out phantoms the warm-started PDHG reaches 5e-2 relative error against the reference
    in fewer iterations than the cold start (median over the slices)."""
    import torch

    ws = pkg.warmstart
    kw = dict(n=64, n_angles=90, sirt_its=20, pdhg_its=1000, batch=36, data_term="l2")
    train_set = ws.make_pairs(72, seed=11, **kw)
    net = ws.UNet(levels=3, base=16, seed=3)
    losses = ws.train(net, train_set["x_sirt"][:64], train_set["x_ref"][:64], epochs=30, batch=8, lr=2e-3)
    test_set = ws.make_pairs(8, seed=12, **kw)
    ch = test_set["chunks"][0]
    A, b = ch["A"], ch["b"]
    x_ref = A.pack_vol(list(test_set["x_ref"]))
    x0 = A.pack_vol(list(ws.warm_start(net, test_set["x_sirt"])))
    solver = pkg.batch.BatchedPDHG(data_term="l2", tv_lambda=test_set["tv_lambda"], use_graph=False)
    tols = [5e-2]
    cold = ws.iterations_to_tolerance(solver, A, b, x_ref, None, tols, max_its=600)
    warm = ws.iterations_to_tolerance(solver, A, b, x_ref, x0, tols, max_its=600)
    print(f"train loss {losses[0]:.3e} -> {losses[-1]:.3e}; its to 5e-2: cold {cold[0].tolist()}, "
          f"warm {warm[0].tolist()}")
    assert np.median(warm[0]) < np.median(cold[0])
    torch.cuda.synchronize()
