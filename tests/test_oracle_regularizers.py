"""The oracle's gradient-family regularizers (CPU): known answers for the n-D gradient and
divergence, the VTV prox's reductions to the l12 / l1 / Frobenius projections, the TNV prox as
the Frobenius-nearest point of the spectral-norm unit ball, and the generic PDHG / SIRT paths
against the 2-D TV ones they generalise."""

import numpy as np

from oracle import regularizers as oreg
from oracle import solvers as osol

F32 = np.float32


def test_grad_div_nd_known_answers():
    x = np.array([[[0, 1], [3, 7]], [[2, 2], [5, 4]]], dtype=F32)  # [z, y, x]
    g = oreg.grad_nd(x, 3)
    assert np.array_equal(g[0], [[[2, 1], [2, -3]], [[0, 0], [0, 0]]])
    assert np.array_equal(g[1], [[[3, 6], [0, 0]], [[3, 2], [0, 0]]])
    assert np.array_equal(g[2], [[[1, 0], [4, 0]], [[0, 0], [-1, 0]]])
    q = np.arange(24, dtype=F32).reshape(3, 2, 2, 2) ** 2
    d = oreg.div_nd(q)
    # backward differences with the first one 0, summed over (z, y, x)
    want = np.zeros((2, 2, 2), F32)
    want[1] += q[0, 1] - q[0, 0]
    want[:, 1] += q[1, :, 1] - q[1, :, 0]
    want[:, :, 1] += q[2, :, :, 1] - q[2, :, :, 0]
    assert np.array_equal(d, want)
    # over the last two axes it is the 2-D TV operator pair, slice by slice
    rng = np.random.default_rng(1)
    x2 = rng.random((3, 6, 7)).astype(F32)
    g2 = oreg.grad_nd(x2, 2)
    for z in range(3):
        assert np.array_equal(g2[:, z], osol.grad(x2[z]))
    q2 = rng.random((2, 3, 6, 7)).astype(F32)
    for z in range(3):
        assert np.array_equal(oreg.div_nd(q2)[z], osol.div(q2[:, z]))
    # one axis: np.diff with a trailing 0
    x1 = rng.random((4, 9)).astype(F32)
    assert np.array_equal(oreg.grad_nd(x1, 1)[0], np.concatenate([np.diff(x1, axis=-1), np.zeros((4, 1), F32)], -1))


def test_vtv_prox_reductions():
    rng = np.random.default_rng(2)
    q = (rng.standard_normal((2, 3, 5, 6)) * 0.3).astype(F32)  # [nd, C, y, x]
    w = 0.25
    # one channel: derivative l2, channel l1 (inf's dual) is the isotropic TV prox
    q1 = q[:, :1]
    assert np.allclose(oreg.prox_vtv(q1, w, 2, np.inf), oreg.prox_l12(q1, w), rtol=1e-6)
    assert np.allclose(oreg.prox_vtv(q1[:, 0], w, 2, np.inf, nd=2), oreg.prox_l12(q1[:, 0], w), rtol=1e-6)
    # point-wise l1 in both: the anisotropic prox
    assert np.allclose(oreg.prox_vtv(q, w, 1, 1), oreg.prox_l1(q, w), rtol=1e-6)
    # l2 in both: projection of each (nd x C) block onto the Frobenius w-ball
    fro = np.sqrt((q.astype(np.float64) ** 2).sum(axis=(0, 1), keepdims=True))
    assert np.allclose(oreg.prox_vtv(q, w, 2, 2), q * w / np.fmax(fro, w), rtol=1e-5)
    # channel inf norm (dual l1): the blocks land on the l1-of-l2 ball
    r = oreg.prox_vtv(q, w, 2, np.inf)
    assert np.all(np.sqrt((r.astype(np.float64) ** 2).sum(axis=0)).sum(axis=0) <= w * (1 + 1e-5))


def test_tnv_prox_is_spectral_ball_projection():
    rng = np.random.default_rng(3)
    q = (rng.standard_normal((2, 4, 3, 3)) * 1.5).astype(F32)  # [nd, C, y, x]
    r = oreg.prox_tnv(q)
    mats_q = np.moveaxis(q.reshape(2, 4, -1), -1, 0).transpose(0, 2, 1).astype(np.float64)  # [P, C, nd]
    mats_r = np.moveaxis(r.reshape(2, 4, -1), -1, 0).transpose(0, 2, 1).astype(np.float64)
    for m, p in zip(mats_q, mats_r):
        assert np.linalg.norm(p, 2) <= 1 + 1e-5
        d0 = np.linalg.norm(m - p)
        for _ in range(50):  # no feasible matrix is closer in the Frobenius norm
            c = p + rng.standard_normal(p.shape) * 0.05
            c /= max(1.0, np.linalg.norm(c, 2))
            assert np.linalg.norm(m - c) >= d0 - 1e-6
    # one channel: the nuclear norm is the l2 norm, so the prox is q / max(|q|, 1)
    q1 = q[:, :1]
    nrm = np.sqrt((q1.astype(np.float64) ** 2).sum(axis=0, keepdims=True))
    assert np.allclose(oreg.prox_tnv(q1), q1 / np.fmax(nrm, 1.0), rtol=1e-5)


def _dense_problem(shape, m_rows, seed):
    rng = np.random.default_rng(seed)
    n = int(np.prod(shape))
    mat = rng.random((m_rows, n)).astype(F32)
    x = rng.random(shape).astype(F32)
    b = (mat @ x.reshape(-1)).astype(F32)
    A = lambda v: (mat @ v.reshape(-1)).astype(F32)  # noqa: E731
    At = lambda y: (mat.T @ y).astype(F32).reshape(shape)  # noqa: E731
    return A, At, b


def test_generic_pdhg_sirt_match_tv2d():
    shape = (6, 7)
    A, At, b = _dense_problem(shape, 60, 4)
    for term in ("l2", "kl"):
        ref = osol.pdhg(A, At, b, shape, 25, data_term=term, tv_weight=0.05)
        got = osol.pdhg(A, At, b, shape, 25, data_term=term, regs=[oreg.GradReg(2, 0.05, "l12")])
        assert np.array_equal(ref, got)
    ref = osol.sirt(A, At, b, shape, 25, tv_weight=0.05)
    got = osol.sirt(A, At, b, shape, 25, regs=[oreg.GradReg(2, 0.05, "l12")])
    assert np.array_equal(ref, got)


def test_generic_regularizers_reduce_objective():
    """PDHG with TV3D, VTV and TNV on a piecewise-constant 3-channel volume from noisy dense data
    ends closer to the truth than the unregularised run (the regularizers act as TV does)."""
    rng = np.random.default_rng(5)
    shape = (3, 8, 8)
    truth = np.zeros(shape, F32)
    truth[:, 2:6, 3:7] = np.array([1.0, 0.6, 0.3], F32)[:, None, None]
    n = truth.size
    mat = rng.random((n // 2, n)).astype(F32)
    b = (mat @ truth.reshape(-1)).astype(F32)
    b = b + (rng.standard_normal(b.shape) * 0.01 * b.max()).astype(F32)
    A = lambda v: (mat @ v.reshape(-1)).astype(F32)  # noqa: E731
    At = lambda y: (mat.T @ y).astype(F32).reshape(shape)  # noqa: E731
    plain = osol.pdhg(A, At, b, shape, 300)
    e0 = np.linalg.norm(plain - truth)
    for reg in (oreg.GradReg(3, 0.5, "l12"), oreg.GradReg(2, 0.5, "vtv"), oreg.GradReg(2, 0.5, "tnv")):
        x = osol.pdhg(A, At, b, shape, 300, regs=[reg])
        assert np.linalg.norm(x - truth) < e0, reg.kind


def test_update_mask_special_cases():
    """upd_mask (regularizers.py:206-207, 271-272): a mask of ones gives the unmasked run bit for bit;
    a mask of zeros removes the regularizer (its tau term and primal update vanish) bit for bit, for
    PDHG (preconditioned and Lipschitz) and SIRT; a half mask lies strictly between. A dropped mask
    on either term, or the mask applied after the weight of a weighted norm, breaks one of these."""
    shape = (6, 7)
    A, At, b = _dense_problem(shape, 60, 4)
    ones, zeros = np.ones(shape, F32), np.zeros(shape, F32)
    half = np.zeros(shape, F32)
    half[:, :4] = 1
    x = np.ones(shape, F32)
    for _ in range(60):  # the operator norm, for the Lipschitz steps
        x = At(A(x))
        x = x / np.linalg.norm(x)
    L = float(np.sqrt(np.linalg.norm(At(A(x)))))

    def same(u, v, exact):
        return np.array_equal(u, v) if exact else np.abs(u - v).max() <= 1e-6 * np.abs(v).max()

    for kind in ("l12", "tnv"):
        # the Lipschitz tau is rho / (L + terms): one float64 division for scalars, float32 for arrays
        for run, exact in ((lambda regs: osol.pdhg(A, At, b, shape, 25, data_term="kl", regs=regs), True),
                           (lambda regs: osol.pdhg(A, At, b, shape, 25, regs=regs, precondition=False, L=L), False),
                           (lambda regs: osol.sirt(A, At, b, shape, 25, regs=regs), True)):
            plain_reg = run([oreg.GradReg(2, 0.05, kind)])
            assert same(run([oreg.GradReg(2, 0.05, kind, upd_mask=ones)]), plain_reg, exact)
            assert same(run([oreg.GradReg(2, 0.05, kind, upd_mask=zeros)]), run([]), exact)
            mid = run([oreg.GradReg(2, 0.05, kind, upd_mask=half)])
            assert not same(mid, plain_reg, False) and not same(mid, run([]), False)
    # the primal term: -div q masked, then weighted (tnv) -- the order of the reference's two products
    rng = np.random.default_rng(11)
    q = rng.standard_normal((2, *shape)).astype(F32)
    r = oreg.GradReg(2, 0.3, "tnv", upd_mask=half)
    assert np.array_equal(r.primal(q), (-oreg.div_nd(q) * half) * F32(0.3))
    assert np.array_equal(r.tau(), F32(4 * 0.3) * half)


def test_vtv_reference_volume():
    """VTV's x_ref (regularizers.py:1386-1392, 1410-1435): its dual 0.5 grad(x_ref) joins every
    point-wise norm as one more channel. Hand-computed one-pixel case (q = (3, 4), reference dual
    (0, 12), weight 1): the channel 2-norm gives dn = 13, the channel inf-norm (summed) 17, the
    per-channel norm 5 (the reference cannot act); a zero reference changes nothing, bit for bit."""
    q = np.array([3.0, 4.0], F32).reshape(2, 1, 1, 1)  # [nd, C, 1 x 1 image]
    r = np.array([0.0, 12.0], F32).reshape(2, 1, 1)
    for chan, dn in ((2, 13.0), (np.inf, 17.0), (1, 5.0)):
        got = oreg.prox_vtv(q, 1.0, der=2, chan=chan, nd=2, q_ref=r)
        assert np.allclose(got[:, 0, 0, 0], np.array([3.0, 4.0]) / dn, rtol=1e-6), chan
    rng = np.random.default_rng(3)
    q = rng.standard_normal((2, 3, 5, 6)).astype(F32)
    for der in (1, 2, np.inf):
        for chan in (1, 2, np.inf):
            ref0 = oreg.prox_vtv(q, 0.4, der, chan, 2, q_ref=np.zeros((2, 5, 6), F32))
            assert np.array_equal(ref0, oreg.prox_vtv(q, 0.4, der, chan, 2)), (der, chan)
    x_ref = rng.random((5, 6)).astype(F32)
    r = oreg.GradReg(2, 0.4, "vtv", x_ref=x_ref)
    assert np.array_equal(r.q_ref, F32(0.5) * oreg.grad_nd(x_ref, 2))
    assert not np.array_equal(r.dual(np.zeros_like(q), rng.random((3, 5, 6)).astype(F32)),
                              oreg.GradReg(2, 0.4, "vtv").dual(np.zeros_like(q), rng.random((3, 5, 6)).astype(F32)))


def _grad_div_matrices(shape, nd):
    """Explicit G (edge-padded gradient) and D (the primal update's ``-div``) of ``grad_nd`` / ``div_nd``."""
    n = int(np.prod(shape))
    G = np.stack([oreg.grad_nd(np.eye(n)[i].reshape(shape), nd).reshape(-1) for i in range(n)], axis=1)
    D = np.stack([-oreg.div_nd(np.eye(nd * n)[i].reshape((nd, *shape))).reshape(-1) for i in range(nd * n)], axis=1)
    return G, D


def test_smooth_and_huber_tv_fixed_points():
    """PDHG's fixed point (p = M x - b, q the regularizer's dual at G x) solves
    M^T (M x - b) + w D psi'(G x) = 0, with psi'(g) = g for ``smooth2D`` (the normal equations
    (M^T M + w D G) x = M^T b) and psi'(g) = g / h where |g|_2 <= h, g / |g|_2 beyond, per pixel, for
    ``HubTV2D``. D is the reference's ``-div`` (operators.py:818-844), which is G^T except at the
    first sample of each axis (the edge-mode divergence drops -q[0] there)."""
    from scipy.optimize import root

    shape = (5, 6)
    rng = np.random.default_rng(12)
    n = int(np.prod(shape))
    mat = rng.random((20, n)).astype(F32)
    truth = np.zeros(shape, F32)
    truth[1:4, 2:5] = 1
    b = (mat @ truth.reshape(-1) + rng.normal(0, 0.05, 20)).astype(F32)
    A = lambda v: (mat @ v.reshape(-1)).astype(F32)  # noqa: E731
    At = lambda y: (mat.T @ y).astype(F32).reshape(shape)  # noqa: E731
    G, D = _grad_div_matrices(shape, 2)
    assert np.abs(D - G.T).max() == 1.0  # the quirk: not the transpose
    M = mat.astype(np.float64)
    w = 0.5
    want = np.linalg.solve(M.T @ M + w * D @ G, M.T @ b.astype(np.float64)).reshape(shape)
    got = osol.pdhg(A, At, b, shape, 3000, regs=[oreg.GradReg(2, w, "l2")])
    err = np.max(np.abs(got - want)) / np.max(np.abs(want))
    print(f"smooth2D fixed point vs normal equations: {err:.2e}")
    assert err < 1e-5

    w, h = 0.2, 0.05

    def fixed_point(x):
        g = (G @ x).reshape(2, -1)
        r = np.sqrt(np.sum(g * g, axis=0))
        dpsi = np.where(r <= h, 1 / h, 1 / np.maximum(r, 1e-300)) * g
        return M.T @ (M @ x - b) + w * D @ dpsi.reshape(-1)

    sol = root(fixed_point, want.reshape(-1), method="hybr", options=dict(xtol=1e-14))
    assert sol.success and np.max(np.abs(fixed_point(sol.x))) < 1e-10
    want = sol.x.reshape(shape)
    got = osol.pdhg(A, At, b, shape, 6000, regs=[oreg.GradReg(2, w, "hub", huber_size=h)])
    err = np.max(np.abs(got - want)) / np.max(np.abs(want))
    print(f"HubTV2D fixed point vs root of the optimality system: {err:.2e}")
    assert err < 1e-4


def test_array_weights_reduce_to_scalar_weights():
    """A constant weight array gives the scalar weight's iterates bit for bit, for every kind."""
    shape = (3, 6, 7)
    A, At, b = _dense_problem(shape, 70, 13)
    w = np.full(shape[1:], 0.05, F32)
    for kind, nd, kw in (("l12", 2, {}), ("l1", 3, {}), ("vtv", 2, {}), ("tnv", 2, {}), ("l2", 2, {}),
                         ("hub", 3, dict(huber_size=0.1))):
        ref = osol.pdhg(A, At, b, shape, 15, data_term="kl", regs=[oreg.GradReg(nd, 0.05, kind, **kw)])
        got = osol.pdhg(A, At, b, shape, 15, data_term="kl", regs=[oreg.GradReg(nd, w, kind, **kw)])
        assert np.array_equal(ref, got), kind
    ref = osol.sirt(A, At, b, shape, 15, regs=[oreg.L1Reg(0.05)])
    got = osol.sirt(A, At, b, shape, 15, regs=[oreg.L1Reg(w)])
    assert np.array_equal(ref, got)


def test_l1_regularizer_fixed_point_and_masked_fista_prox():
    """PDHG with the identity-transform l1 regularizer minimises 1/2 |M x - b|^2 + sum w |x|; FISTA
    with 1/L steps (gradient step tau sigma = 1/L^2, threshold tau w' = w' / L) minimises
    1/2 |M x - b|^2 + L sum w' |x|, so with w' = w / L the two algorithms meet at the same
    minimiser, with a scalar and with a spatially varying weight. The masked FISTA prox reduces to
    the unmasked one for a mask of ones and to no prox for a mask of zeros."""
    shape = (5, 6)
    rng = np.random.default_rng(14)
    n = int(np.prod(shape))
    mat = rng.random((40, n)).astype(F32)
    truth = np.where(rng.random(shape) < 0.3, rng.random(shape), 0).astype(F32)
    b = (mat @ truth.reshape(-1) + rng.normal(0, 0.02, 40)).astype(F32)
    A = lambda v: (mat @ v.reshape(-1)).astype(F32)  # noqa: E731
    At = lambda y: (mat.T @ y).astype(F32).reshape(shape)  # noqa: E731
    L = float(np.linalg.norm(mat.astype(np.float64), 2))
    for w in (0.3, np.linspace(0.1, 0.6, n, dtype=F32).reshape(shape)):
        pd = osol.pdhg(A, At, b, shape, 8000, regs=[oreg.L1Reg(w)])
        fi = osol.fista(A, At, b, shape, 8000, precondition=False, L=L, l1=(np.asarray(w) / F32(L)).astype(F32))
        err = np.max(np.abs(pd - fi)) / np.max(np.abs(fi))
        print(f"l1 fixed point, weight {'array' if np.ndim(w) else w}: PDHG vs FISTA {err:.2e}")
        assert err < 1e-4 and np.count_nonzero(fi == 0) > 0
    l1 = np.linspace(0.1, 0.6, n, dtype=F32).reshape(shape)
    plain = osol.fista(A, At, b, shape, 20, l1=l1)
    assert np.array_equal(osol.fista(A, At, b, shape, 20, l1=l1, l1_mask=np.ones(shape, F32)), plain)
    assert np.array_equal(osol.fista(A, At, b, shape, 20, l1=l1, l1_mask=np.zeros(shape, F32)),
                          osol.fista(A, At, b, shape, 20))


def test_gradient_padding_modes():
    """The np.pad modes of TransformGradient: -div is the exact adjoint of the gradient for the
    constant (zero) and wrap (periodic) modes, for any dual; reflect and symmetric by hand on a 1-D
    ramp (symmetric equals edge)."""
    rng = np.random.default_rng(15)
    for shape, nd in (((5, 7), 2), ((3, 4, 6), 3), ((2, 9), 1)):
        x = rng.normal(size=shape)
        q = rng.normal(size=(nd, *shape))
        for mode in ("constant", "wrap"):
            lhs = np.sum(oreg.grad_nd(x, nd, mode) * q)
            rhs = np.sum(x * -oreg.div_nd(q, mode))
            assert abs(lhs - rhs) < 1e-12 * max(1.0, abs(lhs)), (mode, shape)
        for mode in ("edge", "reflect", "symmetric"):  # not adjoint: the boundary terms differ
            lhs = np.sum(oreg.grad_nd(x, nd, mode) * q)
            assert abs(lhs - np.sum(x * -oreg.div_nd(q, mode))) > 1e-6
    x = np.array([1.0, 2.0, 4.0])
    q = np.array([[1.0, 2.0, 4.0]])
    assert np.array_equal(oreg.grad_nd(x, 1, "reflect")[0], [1.0, 2.0, -2.0])
    assert np.array_equal(oreg.div_nd(q, "reflect"), [-1.0, 1.0, 2.0])
    assert np.array_equal(oreg.grad_nd(x, 1, "symmetric"), oreg.grad_nd(x, 1, "edge"))
    assert np.array_equal(oreg.div_nd(q, "symmetric"), oreg.div_nd(q, "edge"))
    assert np.array_equal(oreg.grad_nd(x, 1, "edge")[0], [1.0, 2.0, 0.0])
    assert np.array_equal(oreg.div_nd(q, "edge"), [0.0, 1.0, 2.0])


def test_gradient_over_chosen_axes():
    """Gradient axes in any order and position are the trailing-axis operator applied to the
    array with those axes moved last: components follow the given order, and the divergence
    sums them back."""
    rng = np.random.default_rng(16)
    x = rng.normal(size=(3, 5, 7)).astype(F32)
    g = oreg.grad_nd(x, 2, axes=(-1, -2))
    assert np.array_equal(g, oreg.grad_nd(x, 2)[::-1])
    g0 = oreg.grad_nd(x, 1, "wrap", axes=(0,))
    want = np.moveaxis(oreg.grad_nd(np.moveaxis(x, 0, -1), 1, "wrap")[0], -1, 0)
    assert np.array_equal(g0[0], want)
    axes = (2, 0, 1)
    g3 = oreg.grad_nd(x, 3, "reflect", axes=axes)
    xt = np.transpose(x, axes)
    gt = oreg.grad_nd(xt, 3, "reflect")
    for k in range(3):
        assert np.array_equal(np.transpose(g3[k], axes), gt[k])
    q = rng.normal(size=(3, 3, 5, 7)).astype(F32)
    dt_ = oreg.div_nd(np.stack([np.transpose(q[k], axes) for k in range(3)]), "reflect")
    assert np.allclose(np.transpose(oreg.div_nd(q, "reflect", axes=axes), axes), dt_, atol=1e-6)
