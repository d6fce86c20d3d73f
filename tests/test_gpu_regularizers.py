"""Gradient-family regularizers on the GPU against the oracle (run with -m gpu).

Bars: one dual update + prox and one primal contribution of TV1D / TV2D / TV3D, anisotropic
TV, VTV (all nine point-wise norm pairs) and TNV (singular values by a float64 Jacobi
eigen-solve on the GPU against numpy's SVD in the oracle) to 1e-5 relative; PDHG / SIRT with
TV3D, VTV and TNV on multi-slice and multi-channel volumes (uncorrected and XRF with per-channel
attenuation) to the solver bar 1e-4 after 30 iterations.
"""

import numpy as np
import pytest

from oracle import attenuation as oatt
from oracle import regularizers as oreg
from oracle import solvers as osol
from oracle import xrf as oxrf
from tests.conftest import rel_l2, rel_max

pytestmark = pytest.mark.gpu

TOL = 1e-4


@pytest.fixture(scope="module")
def pkg():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import pycorrectedemissionct_amd as p

    return p


def _cases(R, dt):
    inf = np.inf
    yield "tv1d", R.Regularizer_TV1D(0.3), oreg.GradReg(1, 0.3, "l12"), (4, 40)
    yield "tv2d", R.Regularizer_Grad(0.3, ndims=2), oreg.GradReg(2, 0.3, "l12"), (3, 17, 19)
    yield "tv3d", R.Regularizer_TV3D(0.3), oreg.GradReg(3, 0.3, "l12"), (6, 10, 12)
    yield "aniso", R.Regularizer_Grad(0.3, ndims=2, norm=dt.DataFidelity_l1()), oreg.GradReg(2, 0.3, "l1"), (2, 9, 11)
    for der in (1, 2, inf):
        for chan in (1, 2, inf):
            yield (f"vtv{der}{chan}", R.Regularizer_VTV(0.3, pwise_der_norm=der, pwise_chan_norm=chan),
                   oreg.GradReg(2, 0.3, "vtv", der, chan), (3, 12, 14))
    yield "vtv_c6_z", R.Regularizer_VTV(0.3), oreg.GradReg(2, 0.3, "vtv"), (6, 2, 9, 10)
    yield "vtv_3d", R.Regularizer_VTV(0.3, ndims=3), oreg.GradReg(3, 0.3, "vtv"), (5, 6, 7)
    yield "tnv_c3", R.Regularizer_TNV(0.3), oreg.GradReg(2, 0.3, "tnv"), (3, 12, 14)
    yield "tnv_c6", R.Regularizer_TNV(0.3), oreg.GradReg(2, 0.3, "tnv"), (6, 10, 11)
    yield "tnv_c1", R.Regularizer_TNV(0.3), oreg.GradReg(2, 0.3, "tnv"), (1, 10, 11)
    yield "tnv_3d", R.Regularizer_TNV(0.3, ndims=3), oreg.GradReg(3, 0.3, "tnv"), (4, 6, 7, 8)


def test_dual_and_primal_kernels(pkg):
    import torch

    R, dt = pkg.regularizers, pkg.data_terms
    st = pkg._native.stream()
    rng = np.random.default_rng(970)
    for name, reg, orc, shape in _cases(R, dt):
        x = rng.standard_normal(shape).astype(np.float32)
        q = (rng.standard_normal((orc.nd, *shape)) * 0.4).astype(np.float32)
        cfg = reg.device_config(shape)
        assert cfg["q_shape"] == q.shape
        xt, qt = torch.as_tensor(x).cuda(), torch.as_tensor(q).cuda()
        pkg.solvers.Solver._grad_dual(cfg, qt, xt, st)
        want = orc.dual(q, x)
        e = rel_max(want, qt.cpu().numpy())
        upd = rng.standard_normal(shape).astype(np.float32)
        for subtract in (False, True):
            ut = torch.as_tensor(upd).cuda()
            pkg.solvers.Solver._grad_primal(cfg, ut, qt, subtract, st)
            contrib = orc.primal(qt.cpu().numpy())
            want_u = upd - contrib if subtract else upd + contrib
            eu = rel_max(want_u, ut.cpu().numpy())
            assert eu < 1e-5, (name, subtract, eu)
        print(f"{name} {shape}: dual rel_max {e:.2e}")
        assert e < 1e-5, name
        assert orc.tau() == pytest.approx(float(reg.initialize_sigma_tau(np.zeros(shape, np.float32))))


def _stack_ops(ang, n, maps=None):
    def A(x):
        return np.stack([oxrf.fp(x[c], ang, None if maps is None else maps[c])[0] for c in range(x.shape[0])])

    def At(y):
        return np.stack([oxrf.bp_skimage(y[c], ang, n) for c in range(y.shape[0])])

    return A, At


def test_solvers_with_volume_regularizers(pkg):
    """PDHG (l2 and KL) and SIRT with TV3D, VTV and TNV on a 3-channel 24 x 24 stack, uncorrected."""
    n, a, c = 24, 16, 3
    rng = np.random.default_rng(980)
    ang = np.linspace(0, 2 * np.pi, a, endpoint=False).astype(np.float32)
    base = pkg.testing.phantom_ellipses(n, 3, rng, semi_axes=(3, 7))
    f = np.stack([base * s for s in (1.0, 0.6, 0.3)]).astype(np.float32)
    A, At = _stack_ops(ang, n)
    b = pkg.testing.poisson_data(A(f), 1e3, rng)
    prj = pkg.ProjectorUncorrected([n, n, c], ang)
    R = pkg.regularizers
    cases = [
        (lambda: R.Regularizer_TV3D(5e-2), oreg.GradReg(3, 5e-2, "l12")),
        (lambda: R.Regularizer_VTV(5e-2), oreg.GradReg(2, 5e-2, "vtv")),
        (lambda: R.Regularizer_TNV(5e-2), oreg.GradReg(2, 5e-2, "tnv")),
    ]
    for make, orc in cases:
        for term in ("l2", "kl"):
            x = pkg.PDHG(data_term=term, regularizer=make())(prj, b, iterations=30)[0]
            ref = osol.pdhg(A, At, b, f.shape, 30, data_term=term, regs=[orc])
            print(f"PDHG-{term}-{orc.kind}{orc.nd}: rel_max {rel_max(ref, x):.2e} rel_l2 {rel_l2(ref, x):.2e}")
            assert rel_max(ref, x) < TOL
        x = pkg.SIRT(regularizer=make())(prj, b, iterations=30)[0]
        ref = osol.sirt(A, At, b, f.shape, 30, regs=[orc])
        print(f"SIRT-{orc.kind}{orc.nd}: rel_max {rel_max(ref, x):.2e}")
        assert rel_max(ref, x) < TOL


def test_smooth_and_huber_tv_regularizers(pkg):
    """``Regularizer_smooth1D/2D/3D`` (l2 norm) and ``Regularizer_HubTV2D/3D`` (Huber over the gradient
    axis) in PDHG (l2, KL) and SIRT on the 3-channel stack, against the oracle."""
    n, a, c = 24, 16, 3
    rng = np.random.default_rng(981)
    ang = np.linspace(0, 2 * np.pi, a, endpoint=False).astype(np.float32)
    base = pkg.testing.phantom_ellipses(n, 3, rng, semi_axes=(3, 7))
    f = np.stack([base * s for s in (1.0, 0.6, 0.3)]).astype(np.float32)
    A, At = _stack_ops(ang, n)
    b = pkg.testing.poisson_data(A(f), 1e3, rng)
    prj = pkg.ProjectorUncorrected([n, n, c], ang)
    R = pkg.regularizers
    cases = [
        ("smooth1D", lambda: R.Regularizer_smooth1D(5e-2), oreg.GradReg(1, 5e-2, "l2")),
        ("smooth2D", lambda: R.Regularizer_smooth2D(5e-2), oreg.GradReg(2, 5e-2, "l2")),
        ("smooth3D", lambda: R.Regularizer_smooth3D(5e-2), oreg.GradReg(3, 5e-2, "l2")),
        ("HubTV2D", lambda: R.Regularizer_HubTV2D(5e-2, 0.1), oreg.GradReg(2, 5e-2, "hub", huber_size=0.1)),
        ("HubTV3D", lambda: R.Regularizer_HubTV3D(5e-2, 0.1), oreg.GradReg(3, 5e-2, "hub", huber_size=0.1)),
    ]
    for name, make, orc in cases:
        for term in ("l2", "kl"):
            x = pkg.PDHG(data_term=term, regularizer=make())(prj, b, iterations=30)[0]
            ref = osol.pdhg(A, At, b, f.shape, 30, data_term=term, regs=[orc])
            print(f"PDHG-{term}-{name}: rel_max {rel_max(ref, x):.2e}")
            assert rel_max(ref, x) < TOL, name
        x = pkg.SIRT(regularizer=make())(prj, b, iterations=30)[0]
        ref = osol.sirt(A, At, b, f.shape, 30, regs=[orc])
        print(f"SIRT-{name}: rel_max {rel_max(ref, x):.2e}")
        assert rel_max(ref, x) < TOL, name
    x = pkg.PDHG(regularizer=R.Regularizer_Grad(5e-2, norm=pkg.data_terms.DataFidelity_Huber(0.1)))(
        prj, b, iterations=30)[0]
    ref = osol.pdhg(A, At, b, f.shape, 30, regs=[oreg.GradReg(2, 5e-2, "hub_el", huber_size=0.1)])
    print(f"PDHG-l2-Huber per component: rel_max {rel_max(ref, x):.2e}")
    assert rel_max(ref, x) < TOL
    with pytest.raises(NotImplementedError):
        pkg.PDHG(regularizer=R.Regularizer_Grad(1e-2, norm=pkg.data_terms.DataFidelity_Huber(0.1, l2_axis=1)))(
            prj, b, iterations=1)


def test_array_weights_and_l1_regularizer(pkg):
    """Array-valued weights (``weight: float | NDArray``, broadcast over the primal) of the gradient
    regularizers and ``Regularizer_l1`` with its identity transform in PDHG (l2, KL, and Lipschitz
    steps) and SIRT, and FISTA's l1 prox with an array weight and an update mask, on the 3-channel
    stack against the oracle."""
    n, a, c = 24, 16, 3
    rng = np.random.default_rng(987)
    ang = np.linspace(0, 2 * np.pi, a, endpoint=False).astype(np.float32)
    base = pkg.testing.phantom_ellipses(n, 3, rng, semi_axes=(3, 7))
    f = np.stack([base * s for s in (1.0, 0.6, 0.3)]).astype(np.float32)
    A, At = _stack_ops(ang, n)
    b = pkg.testing.poisson_data(A(f), 1e3, rng)
    prj = pkg.ProjectorUncorrected([n, n, c], ang)
    R = pkg.regularizers
    w2 = rng.uniform(0.01, 0.1, (n, n)).astype(np.float32)  # broadcast over the channels
    w3 = rng.uniform(0.01, 0.1, (c, n, n)).astype(np.float32)
    m2 = (rng.random((n, n)) > 0.3).astype(np.float32)
    cases = [
        ("TV3D w2", lambda: R.Regularizer_TV3D(w2), oreg.GradReg(3, w2, "l12")),
        ("TV2D w3", lambda: R.Regularizer_TV2D(w3), oreg.GradReg(2, w3, "l12")),
        ("aniso w2", lambda: R.Regularizer_Grad(w2, norm=pkg.data_terms.DataFidelity_l1()),
         oreg.GradReg(2, w2, "l1")),
        ("VTV w3", lambda: R.Regularizer_VTV(w3), oreg.GradReg(2, w3, "vtv")),
        ("TNV w2", lambda: R.Regularizer_TNV(w2), oreg.GradReg(2, w2, "tnv")),
        ("smooth2D w3 mask", lambda: R.Regularizer_smooth2D(w3, upd_mask=m2),
         oreg.GradReg(2, w3, "l2", upd_mask=m2)),
        ("HubTV2D w2", lambda: R.Regularizer_HubTV2D(w2, 0.1), oreg.GradReg(2, w2, "hub", huber_size=0.1)),
        ("l1", lambda: R.Regularizer_l1(0.05), oreg.L1Reg(0.05)),
        ("l1 w3 mask", lambda: R.Regularizer_l1(w3, upd_mask=m2), oreg.L1Reg(w3, upd_mask=m2)),
    ]
    for name, make, orc in cases:
        for term in ("l2", "kl"):
            x = pkg.PDHG(data_term=term, regularizer=make())(prj, b, iterations=30)[0]
            ref = osol.pdhg(A, At, b, f.shape, 30, data_term=term, regs=[orc])
            print(f"PDHG-{term}-{name}: rel_max {rel_max(ref, x):.2e}")
            assert rel_max(ref, x) < TOL, name
        x = pkg.SIRT(regularizer=make())(prj, b, iterations=30)[0]
        ref = osol.sirt(A, At, b, f.shape, 30, regs=[orc])
        print(f"SIRT-{name}: rel_max {rel_max(ref, x):.2e}")
        assert rel_max(ref, x) < TOL, name
    # Lipschitz steps: the array tau terms go through compute_Lipschitz_scaling
    for name, make, orc in (cases[0], cases[-1]):
        np.random.seed(3)
        x = pkg.PDHG(regularizer=make())(prj, b, iterations=20, precondition=False)[0]
        np.random.seed(3)
        L = osol.power_method(A, At, np.random.rand(*b.shape).astype(np.float32))
        ref = osol.pdhg(A, At, b, f.shape, 20, regs=[orc], precondition=False, L=L)
        print(f"PDHG Lipschitz {name}: rel_max {rel_max(ref, x):.2e}")
        assert rel_max(ref, x) < 1e-3, name
    x = pkg.FISTA(regularizer=R.Regularizer_l1(w3 * 0.1, upd_mask=m2))(prj, b, iterations=30, lower_limit=0.0)[0]
    ref = osol.fista(A, At, b, f.shape, 30, l1=w3 * 0.1, l1_mask=m2, lower=0.0)
    print(f"FISTA l1 w3 mask: rel_max {rel_max(ref, x):.2e}")
    assert rel_max(ref, x) < TOL
    # the regularizer's own dual prox on host arrays, with an array weight
    q = rng.normal(0, 0.1, (c, n, n)).astype(np.float32)
    reg = R.Regularizer_l1(w3)
    reg.initialize_sigma_tau(np.zeros((c, n, n), np.float32))
    reg.apply_proximal_dual(q_dev := q.copy())
    assert np.array_equal(q_dev, oreg.prox_l1(q, w3))
    with pytest.raises(ValueError):
        pkg.PDHG(regularizer=R.Regularizer_TV2D(np.ones((5, 5), np.float32)))(prj, b, iterations=1)


def test_gradient_padding_modes(pkg):
    """TransformGradient's np.pad modes (``operators.py:796-835``) in the HIP gradient kernels and the
    host operator: PDHG-l2 and SIRT with TV3D / TV2D / smooth2D under constant, wrap, reflect and
    symmetric padding against the oracle, and the operator's gradient / divergence against it."""
    n, a, c = 24, 16, 3
    rng = np.random.default_rng(988)
    ang = np.linspace(0, 2 * np.pi, a, endpoint=False).astype(np.float32)
    base = pkg.testing.phantom_ellipses(n, 3, rng, semi_axes=(3, 7))
    f = np.stack([base * s for s in (1.0, 0.6, 0.3)]).astype(np.float32)
    A, At = _stack_ops(ang, n)
    b = pkg.testing.poisson_data(A(f), 1e3, rng)
    prj = pkg.ProjectorUncorrected([n, n, c], ang)
    R = pkg.regularizers
    for mode in ("constant", "wrap", "reflect", "symmetric"):
        cases = [
            ("TV3D", lambda: R.Regularizer_TV3D(5e-2, pad_mode=mode), oreg.GradReg(3, 5e-2, "l12", pad_mode=mode)),
            ("TV2D", lambda: R.Regularizer_TV2D(5e-2, pad_mode=mode), oreg.GradReg(2, 5e-2, "l12", pad_mode=mode)),
            ("smooth2D", lambda: R.Regularizer_smooth2D(5e-2, pad_mode=mode),
             oreg.GradReg(2, 5e-2, "l2", pad_mode=mode)),
        ]
        for name, make, orc in cases:
            x = pkg.PDHG(regularizer=make())(prj, b, iterations=30)[0]
            ref = osol.pdhg(A, At, b, f.shape, 30, regs=[orc])
            x2 = pkg.SIRT(regularizer=make())(prj, b, iterations=30)[0]
            ref2 = osol.sirt(A, At, b, f.shape, 30, regs=[orc])
            print(f"{mode} {name}: PDHG rel_max {rel_max(ref, x):.2e}, SIRT {rel_max(ref2, x2):.2e}")
            assert rel_max(ref, x) < TOL and rel_max(ref2, x2) < TOL, (mode, name)
        op = pkg.operators.TransformGradient(f.shape, pad_mode=mode)
        q = rng.normal(size=(3, *f.shape)).astype(np.float32)
        assert np.allclose(op(f), oreg.grad_nd(f, 3, mode), atol=1e-6), mode
        assert np.allclose(op.T(q), -oreg.div_nd(q, mode), atol=1e-6), mode
    with pytest.raises(ValueError):
        pkg.PDHG(regularizer=R.Regularizer_TV3D(5e-2, pad_mode="mean"))(prj, b, iterations=1)


def test_gradient_over_chosen_axes(pkg):
    """``Regularizer_Grad(axes=...)`` over axes that are not the trailing ones in order
    (``regularizers.py:247-258``): TV1D along the rows, TV2D with reversed axes and over
    (slices, rows), a permuted TV3D with wrap padding, smooth along (slices, columns) and TNV with
    reversed axes, in PDHG and SIRT on the 3-slice stack against the oracle."""
    n, a, c = 24, 16, 3
    rng = np.random.default_rng(989)
    ang = np.linspace(0, 2 * np.pi, a, endpoint=False).astype(np.float32)
    base = pkg.testing.phantom_ellipses(n, 3, rng, semi_axes=(3, 7))
    f = np.stack([base * s for s in (1.0, 0.6, 0.3)]).astype(np.float32)
    A, At = _stack_ops(ang, n)
    b = pkg.testing.poisson_data(A(f), 1e3, rng)
    prj = pkg.ProjectorUncorrected([n, n, c], ang)
    R = pkg.regularizers
    cases = [
        ("TV1D rows", lambda: R.Regularizer_TV1D(5e-2, axes=(-2,)), oreg.GradReg(1, 5e-2, "l12", axes=(-2,))),
        ("TV2D reversed", lambda: R.Regularizer_TV2D(5e-2, axes=(-1, -2)),
         oreg.GradReg(2, 5e-2, "l12", axes=(-1, -2))),
        ("TV2D slices-rows", lambda: R.Regularizer_TV2D(5e-2, axes=(0, 1)), oreg.GradReg(2, 5e-2, "l12", axes=(0, 1))),
        ("TV3D permuted wrap", lambda: R.Regularizer_TV3D(5e-2, axes=(2, 0, 1), pad_mode="wrap"),
         oreg.GradReg(3, 5e-2, "l12", axes=(2, 0, 1), pad_mode="wrap")),
        ("smooth slices-cols", lambda: R.Regularizer_smooth2D(5e-2, axes=(0, 2)),
         oreg.GradReg(2, 5e-2, "l2", axes=(0, 2))),
        ("TNV reversed", lambda: R.Regularizer_TNV(5e-2, axes=(-1, -2)), oreg.GradReg(2, 5e-2, "tnv", axes=(-1, -2))),
    ]
    for name, make, orc in cases:
        x = pkg.PDHG(data_term="kl", regularizer=make())(prj, b, iterations=30)[0]
        ref = osol.pdhg(A, At, b, f.shape, 30, data_term="kl", regs=[orc])
        x2 = pkg.SIRT(regularizer=make())(prj, b, iterations=30)[0]
        ref2 = osol.sirt(A, At, b, f.shape, 30, regs=[orc])
        print(f"{name}: PDHG-KL rel_max {rel_max(ref, x):.2e}, SIRT {rel_max(ref2, x2):.2e}")
        assert rel_max(ref, x) < TOL and rel_max(ref2, x2) < TOL, name
    with pytest.raises(NotImplementedError):  # the channel axis of a coupled regularizer
        pkg.PDHG(regularizer=R.Regularizer_TNV(5e-2, axes=(0, 1)))(prj, b, iterations=1)
    with pytest.raises(ValueError):
        pkg.PDHG(regularizer=R.Regularizer_TV2D(5e-2, axes=(1, -2)))(prj, b, iterations=1)


def test_pdhg_vtv_xrf_channels(pkg):
    """Emission channels as slices of one XRF projector, each with its own outgoing attenuation
    (mu_out scaled per channel), coupled by VTV: PDHG-KL against the oracle."""
    n, a = 24, 16
    rng = np.random.default_rng(990)
    ang = np.linspace(0, 2 * np.pi, a, endpoint=False).astype(np.float32)
    base = pkg.testing.phantom_ellipses(n, 3, rng, semi_axes=(3, 7))
    mu_in, mu_out = pkg.testing.xrf_attenuation(base)
    scales = (0.5, 1.0, 2.0, 4.0)
    c = len(scales)
    f = np.stack([base * rng.uniform(0.3, 1.0) for _ in scales]).astype(np.float32)
    m_in = np.stack([mu_in] * c).astype(np.float32)
    m_out = np.stack([mu_out * s for s in scales]).astype(np.float32)
    maps = [oatt.compute_maps(m_in[i], m_out[i], ang, [np.pi / 2]) for i in range(c)]
    A, At = _stack_ops(ang, n, maps)
    b = pkg.testing.poisson_data(A(f), 1e3, rng)
    prj = pkg.ProjectorAttenuationXRF([n, n, c], ang, att_in=m_in, att_out=m_out)
    x = pkg.PDHG(data_term="kl", regularizer=pkg.regularizers.Regularizer_VTV(2e-2))(prj, b, iterations=30)[0]
    ref = osol.pdhg(A, At, b, f.shape, 30, data_term="kl", regs=[oreg.GradReg(2, 2e-2, "vtv")])
    print(f"PDHG-KL-VTV XRF 4 channels: rel_max {rel_max(ref, x):.2e}")
    assert rel_max(ref, x) < TOL


def test_update_masks(pkg):
    """``upd_mask`` of the gradient-family regularizers (``regularizers.py:206-207, 271-272``): the
    mask scales the regularizer's tau term and its primal update; a spatial mask broadcasts over the
    channels. PDHG (l2, KL; preconditioned and with Lipschitz steps) and SIRT on the 3-channel stack
    against the oracle, and a 2-D TV with a mask (which leaves the fused TV kernels for the masked
    generic path) on one slice."""
    import torch

    n, a, c = 24, 16, 3
    rng = np.random.default_rng(985)
    ang = np.linspace(0, 2 * np.pi, a, endpoint=False).astype(np.float32)
    base = pkg.testing.phantom_ellipses(n, 3, rng, semi_axes=(3, 7))
    f = np.stack([base * s for s in (1.0, 0.6, 0.3)]).astype(np.float32)
    A, At = _stack_ops(ang, n)
    b = pkg.testing.poisson_data(A(f), 1e3, rng)
    prj = pkg.ProjectorUncorrected([n, n, c], ang)
    yy, xx = np.mgrid[:n, :n] - (n - 1) / 2
    mask = np.where(yy**2 + xx**2 < 64, 1.0, 0.25).astype(np.float32)
    R = pkg.regularizers
    cases = [
        (lambda m: R.Regularizer_TV3D(5e-2, upd_mask=m), oreg.GradReg(3, 5e-2, "l12", upd_mask=mask)),
        (lambda m: R.Regularizer_VTV(5e-2, upd_mask=m), oreg.GradReg(2, 5e-2, "vtv", upd_mask=mask)),
        (lambda m: R.Regularizer_TNV(5e-2, upd_mask=m), oreg.GradReg(2, 5e-2, "tnv", upd_mask=mask)),
    ]
    for make, orc in cases:
        for term in ("l2", "kl"):
            x = pkg.PDHG(data_term=term, regularizer=make(mask))(prj, b, iterations=30)[0]
            ref = osol.pdhg(A, At, b, f.shape, 30, data_term=term, regs=[orc])
            e = rel_max(ref, x)
            print(f"masked PDHG-{term}-{orc.kind}{orc.nd}: rel_max {e:.2e}")
            assert e < TOL
        x = pkg.SIRT(regularizer=make(torch.as_tensor(mask).cuda()))(prj, b, iterations=30)[0]
        ref = osol.sirt(A, At, b, f.shape, 30, regs=[orc])
        print(f"masked SIRT-{orc.kind}{orc.nd}: rel_max {rel_max(ref, x):.2e}")
        assert rel_max(ref, x) < TOL
    # Lipschitz steps: tau = rho / (L + tau_reg * mask) is an array
    np.random.seed(31)
    x, _ = pkg.PDHG(data_term="l2", regularizer=R.Regularizer_TV3D(5e-2, upd_mask=mask))(
        prj, b, iterations=30, precondition=False)
    np.random.seed(31)
    L = osol.power_method(A, At, np.random.rand(*np.shape(b)).astype(np.float32))
    ref = osol.pdhg(A, At, b, f.shape, 30, regs=[oreg.GradReg(3, 5e-2, "l12", upd_mask=mask)], precondition=False,
                    L=L)
    e = rel_max(ref, x)
    print(f"masked Lipschitz PDHG-l2-TV3D: rel_max {e:.2e} (L = {L:.4f})")
    assert e < 1e-3
    # one slice, 2-D TV with a mask
    A1 = lambda v: oxrf.fp(v, ang)[0]  # noqa: E731
    At1 = lambda y: oxrf.bp_skimage(y, ang, n)  # noqa: E731
    p1 = pkg.ProjectorUncorrected([n, n], ang)
    x = pkg.PDHG(data_term="kl", regularizer=R.Regularizer_TV2D(5e-2, upd_mask=mask))(p1, b[0], iterations=30)[0]
    ref = osol.pdhg(A1, At1, b[0], (n, n), 30, data_term="kl", regs=[oreg.GradReg(2, 5e-2, "l12", upd_mask=mask)])
    e = rel_max(ref, x)
    print(f"masked PDHG-kl-TV2D one slice: rel_max {e:.2e}")
    assert e < TOL
    with pytest.raises(ValueError, match="upd_mask"):
        pkg.PDHG(regularizer=R.Regularizer_TV3D(5e-2, upd_mask=np.ones((5, 5), np.float32)))(prj, b, iterations=2)


def test_vtv_reference_volume(pkg):
    """VTV's ``x_ref`` (``regularizers.py:1386-1392, 1410-1435``): the reference volume's dual joins
    every point-wise norm as an extra channel. The dual kernel against the oracle for all nine norm
    pairs, with a batch axis and with one channel; PDHG-KL with it against the oracle; TNV's x_ref,
    which the reference computes but its proximal never reads, changes nothing."""
    import torch

    R = pkg.regularizers
    st = pkg._native.stream()
    rng = np.random.default_rng(975)
    inf = np.inf
    cases = [((3, 12, 14), der, chan) for der in (1, 2, inf) for chan in (1, 2, inf)]
    cases += [((3, 2, 9, 10), 2, inf), ((12, 14), 2, 2)]
    for shape, der, chan in cases:
        x = rng.standard_normal(shape).astype(np.float32)
        q = (rng.standard_normal((2, *shape)) * 0.4).astype(np.float32)
        xr = rng.standard_normal(shape[1:] if len(shape) > 2 else shape).astype(np.float32)
        reg = R.Regularizer_VTV(0.3, pwise_der_norm=der, pwise_chan_norm=chan, x_ref=xr)
        cfg = reg.device_config(shape)
        qt = torch.as_tensor(q).cuda()
        pkg.solvers.Solver._grad_dual(cfg, qt, torch.as_tensor(x).cuda(), st)
        want = oreg.GradReg(2, 0.3, "vtv", der, chan, x_ref=xr).dual(q, x)
        e = rel_max(want, qt.cpu().numpy())
        print(f"vtv x_ref {shape} der={der} chan={chan}: rel_max {e:.2e}")
        assert e < 1e-5
    with pytest.raises(ValueError, match="x_ref"):
        R.Regularizer_VTV(0.3, x_ref=np.zeros((5, 5), np.float32)).device_config((3, 12, 14))

    n, a, c = 24, 16, 3
    ang = np.linspace(0, 2 * np.pi, a, endpoint=False).astype(np.float32)
    base = pkg.testing.phantom_ellipses(n, 3, rng, semi_axes=(3, 7))
    f = np.stack([base * s for s in (1.0, 0.6, 0.3)]).astype(np.float32)
    A, At = _stack_ops(ang, n)
    b = pkg.testing.poisson_data(A(f), 1e3, rng)
    prj = pkg.ProjectorUncorrected([n, n, c], ang)
    x_ref = (base > 0).astype(np.float32)
    x = pkg.PDHG(data_term="kl", regularizer=R.Regularizer_VTV(5e-2, x_ref=x_ref))(prj, b, iterations=30)[0]
    ref = osol.pdhg(A, At, b, f.shape, 30, data_term="kl", regs=[oreg.GradReg(2, 5e-2, "vtv", x_ref=x_ref)])
    e = rel_max(ref, x)
    plain = osol.pdhg(A, At, b, f.shape, 30, data_term="kl", regs=[oreg.GradReg(2, 5e-2, "vtv")])
    print(f"PDHG-KL-VTV with x_ref: rel_max {e:.2e} (x_ref moves the result by {rel_max(plain, ref):.2e})")
    assert e < TOL and rel_max(plain, ref) > 10 * TOL
    x1 = pkg.PDHG(data_term="kl", regularizer=R.Regularizer_TNV(5e-2, x_ref=x_ref))(prj, b, iterations=10)[0]
    x2 = pkg.PDHG(data_term="kl", regularizer=R.Regularizer_TNV(5e-2))(prj, b, iterations=10)[0]
    assert np.array_equal(x1, x2)
