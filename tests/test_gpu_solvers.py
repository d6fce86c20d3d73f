"""GPU parity of the solvers (run with -m gpu).

Bar (north star): solver outputs after a fixed number of iterations agree with
the oracle's float32 restatement of corrct's loops (fed by the oracle
projector) to 1e-4 relative. Plus the reference's own convergence tests
(tests/test_solvers.py:100-164) on the dense-matrix fake projector, seeded.
"""

import numpy as np
import pytest

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


@pytest.fixture(scope="module")
def xrf_problem(pkg):
    n, a = 32, 24
    rng = np.random.default_rng(600)
    f = pkg.testing.phantom_ellipses(n, 4, rng, semi_axes=(3, 8))
    mu_in, mu_out = pkg.testing.xrf_attenuation(f)
    ang = np.linspace(0, 2 * np.pi, a, endpoint=False).astype(np.float32)
    det = [np.pi / 2, -np.pi / 2]
    prj = pkg.ProjectorAttenuationXRF([n, n], ang, att_in=mu_in, att_out=mu_out, angles_detectors_rad=det)
    maps = prj.att_vol_angles.cpu().numpy()
    sino = oxrf.fp(f, ang, maps)
    b = pkg.testing.poisson_data(sino, 1e3, rng)
    A = lambda x: oxrf.fp(x, ang, maps)  # noqa: E731
    At = lambda y: oxrf.bp_skimage(y, ang, n)  # noqa: E731
    return dict(prj=prj, b=b, A=A, At=At, shape=(n, n), f=f)


def test_sirt_parity(pkg, xrf_problem):
    p = xrf_problem
    x = pkg.SIRT()(p["prj"], p["b"], iterations=30)[0]
    ref = osol.sirt(p["A"], p["At"], p["b"], p["shape"], 30)
    print(f"SIRT 30 it: rel_max {rel_max(ref, x):.2e} rel_l2 {rel_l2(ref, x):.2e}")
    assert rel_max(ref, x) < TOL


def test_mlem_parity(pkg, xrf_problem):
    p = xrf_problem
    x = pkg.MLEM()(p["prj"], p["b"], iterations=30)[0]
    ref = osol.mlem(p["A"], p["At"], p["b"], p["shape"], 30)
    print(f"MLEM 30 it: rel_max {rel_max(ref, x):.2e} rel_l2 {rel_l2(ref, x):.2e}")
    assert rel_max(ref, x) < TOL


@pytest.mark.parametrize("data_term", ["l2", "kl"])
def test_pdhg_tv_parity(pkg, xrf_problem, data_term):
    p = xrf_problem
    reg = pkg.regularizers.Regularizer_TV2D(1e-2)
    x = pkg.PDHG(data_term=data_term, regularizer=reg)(p["prj"], p["b"], iterations=30)[0]
    ref = osol.pdhg(p["A"], p["At"], p["b"], p["shape"], 30, data_term=data_term, tv_weight=1e-2)
    print(f"PDHG-{data_term}-TV 30 it: rel_max {rel_max(ref, x):.2e} rel_l2 {rel_l2(ref, x):.2e}")
    assert rel_max(ref, x) < TOL


def test_pdhg_limits_and_masks(pkg, xrf_problem):
    p = xrf_problem
    n = p["shape"][0]
    x_mask = np.zeros(p["shape"], np.float32)
    x_mask[2:-2, 2:-2] = 1
    b_mask = (np.random.default_rng(7).random(p["b"].shape) > 0.1).astype(np.float32)
    x = pkg.PDHG(data_term="kl")(p["prj"], p["b"], iterations=20, lower_limit=0, upper_limit=0.8, x_mask=x_mask,
                                 b_mask=b_mask)[0]
    ref = osol.pdhg(p["A"], p["At"], p["b"], p["shape"], 20, data_term="kl", lower=0, upper=0.8, x_mask=x_mask,
                    b_mask=b_mask)
    assert rel_max(ref, x) < TOL
    assert x.max() <= 0.8 and x.min() >= 0 and np.all(x[:2] == 0) and n == 32


# --- the reference's convergence tests on the dense fake projector ---------------


@pytest.fixture(scope="module")
def rand_2d():
    rng = np.random.default_rng(9)
    vol = np.fmin(rng.random((10, 10)) + np.finfo(np.float32).eps, 1).astype(np.float32)
    mat = (rng.random((500, 100)) > 0.5).astype(np.float32)
    data = (mat @ vol.ravel()).reshape(5, 100).astype(np.float32)
    return mat, vol, data


@pytest.mark.parametrize("algo,its", [("SIRT", 1000), ("PDHG", 500), ("MLEM", 10000)])
def test_algo_rand(pkg, rand_2d, algo, its):
    mat, vol, data = rand_2d
    prj = pkg.ProjectorMatrix(mat, vol.shape, data.shape)
    sol, _ = getattr(pkg, algo)()(prj, data, iterations=its)
    assert np.all(np.isclose(sol, vol, atol=1e-2))


def test_algo_flat_tv(pkg):
    rng = np.random.default_rng(11)
    yy, xx = np.mgrid[:10, :10] - 4.5
    vol = ((yy**2 + xx**2) <= 9).astype(np.float32)
    mat = (rng.random((500, 100)) > 0.5).astype(np.float32)
    data = ((mat @ vol.ravel()).reshape(5, 100) + rng.normal(size=(5, 100)) * 1e-3).astype(np.float32)
    prj = pkg.ProjectorMatrix(mat, vol.shape, data.shape)
    for algo, its, kw in ((pkg.SIRT, 2500, {}), (pkg.PDHG, 500, {}), (pkg.PDHG, 500, dict(lower_limit=0, upper_limit=1))):
        sol, _ = algo(regularizer=pkg.regularizers.Regularizer_TV2D(1e-4))(prj, data, iterations=its, **kw)
        assert np.all(np.isclose(sol, vol, atol=1e-2)), algo


def test_sart_parity_xrf_and_uncorrected(pkg):
    """SART (``solvers.py:473-631``) with a fixed angle order, 3 sweeps, against the oracle:
    two-detector XRF projector (weights on the per-angle bp, maps on the fp) with a lower
    limit, and the uncorrected projector."""
    n, a = 32, 12
    rng = np.random.default_rng(800)
    ang = np.linspace(0, 2 * np.pi, a, endpoint=False).astype(np.float32)
    maps = rng.uniform(0.3, 1.0, (a, 2, n, n)).astype(np.float32)
    w_det = np.array([1.0, 0.7])
    f = pkg.testing.phantom_ellipses(n, 4, rng)
    rows = rng.permutation(a)
    prj = pkg.ProjectorAttenuationXRF([n, n], ang, att_maps=maps, angles_detectors_rad=[np.pi / 2, -np.pi / 2],
                                      weights_detectors=w_det)
    b = oxrf.fp(f, ang, maps, w_det=w_det)
    x, info = pkg.SART()(prj, b, 3, rows_sequence=rows, lower_limit=0.0)
    ref = osol.sart(lambda v, i: oxrf.fp(v, ang[i:i + 1], maps[i:i + 1], w_det=w_det)[:, 0, :],
                    lambda y, i: oxrf.bp_skimage(y.reshape(2, 1, n), ang[i:i + 1], n, w_det=w_det, symmetric=True),
                    b, (n, n), a, 3, rows, lower=0.0)
    e = rel_max(ref, x)
    print(f"SART xrf: rel_max vs oracle {e:.2e}; error vs phantom {rel_max(f, x):.2e}")
    assert e < 1e-4 and info.iterations == 3
    pu = pkg.ProjectorUncorrected([n, n], ang, create_single_projs=True)
    bu = oxrf.fp(f, ang)[0]
    xu, _ = pkg.SART()(pu, bu, 3, rows_sequence=rows)
    refu = osol.sart(lambda v, i: oxrf.fp(v, ang[i:i + 1])[0, 0], lambda y, i: oxrf.bp_skimage(y.reshape(1, 1, n),
                     ang[i:i + 1], n), bu, (n, n), a, 3, rows)
    assert rel_max(refu, xu) < 1e-4
    with pytest.raises(ValueError):
        pkg.SART()(pkg.ProjectorUncorrected([n, n], ang), bu, 1)


def test_fista_parity_and_reference_convergence(pkg, xrf_problem):
    """FISTA (``solvers.py:1209-1513``) after 30 iterations against the oracle: diagonal
    preconditioning with an l1 prox and a lower limit; Lipschitz step from the power method
    (same ``np.random.rand`` start vector). Then the reference's own flat-object cases
    (``tests/test_solvers.py:124-134``: 2000 iterations without preconditioning, 1000 with
    limits) on a 500 x 100 random matrix, atol 1e-2."""
    p = xrf_problem
    x, info = pkg.FISTA(regularizer=pkg.regularizers.Regularizer_l1(1e-3))(p["prj"], p["b"], iterations=30,
                                                                         lower_limit=0.0)
    ref = osol.fista(p["A"], p["At"], p["b"], p["shape"], 30, l1=1e-3, lower=0.0)
    e = rel_max(ref, x)
    print(f"FISTA preconditioned 30 it: rel_max {e:.2e}")
    assert e < TOL and info.iterations == 30
    np.random.seed(5)
    x2, _ = pkg.FISTA()(p["prj"], p["b"], iterations=30, precondition=False)
    np.random.seed(5)
    L = osol.power_method(p["A"], p["At"], np.random.rand(*np.shape(p["b"])).astype(np.float32))
    ref2 = osol.fista(p["A"], p["At"], p["b"], p["shape"], 30, precondition=False, L=L)
    e2 = rel_max(ref2, x2)
    print(f"FISTA Lipschitz 30 it: rel_max {e2:.2e} (L = {L:.4f})")
    assert e2 < 1e-3
    with pytest.raises(ValueError):
        pkg.FISTA(regularizer=pkg.regularizers.Regularizer_TV2D(1e-2))
    rng = np.random.default_rng(11)
    yy, xx = np.mgrid[:10, :10] - 4.5
    vol = ((yy**2 + xx**2) <= 9).astype(np.float32)
    mat = (rng.random((500, 100)) > 0.5).astype(np.float32)
    data = ((mat @ vol.ravel()).reshape(5, 100) + rng.normal(size=(5, 100)) * 1e-3).astype(np.float32)
    prj = pkg.ProjectorMatrix(mat, vol.shape, data.shape)
    for its, kw in ((2000, dict(precondition=False)), (1000, dict(lower_limit=0, upper_limit=1))):
        sol, _ = pkg.FISTA()(prj, data, iterations=its, **kw)
        assert np.all(np.isclose(sol, vol, atol=1e-2)), kw


def test_fbp_parity(pkg):
    """FBP (``solvers.py:346-470``): the ramp filter on the uncorrected projector equals
    scikit-image's filtered iradon (golden vectors); shepp-logan / hann and a two-detector
    XRF sinogram equal the oracle; a custom half-spectrum filter equals the named one."""
    import os

    from oracle import filters as ofl

    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "skimage_fbp_golden.npz"))
    sino, rad = g["fbp_sino"], g["fbp_rad"]
    a, n = sino.shape
    pu = pkg.ProjectorUncorrected([n, n], rad)
    x, _ = pkg.FBP()(pu, sino)
    e = rel_max(g["fbp_ramp_vol"], x)
    print(f"FBP ramp vs scikit-image iradon: rel_max {e:.2e}")
    assert e < 1e-5
    for name in ("shepp-logan", "hann"):
        xf, _ = pkg.FBP(fbp_filter=name)(pu, sino)
        ref = ofl.fbp(lambda y: oxrf.bp_skimage(y[None].astype(np.float32), rad, n), sino, a, name)
        assert rel_max(ref, xf) < 1e-5, name
    half = ofl.fbp_filter_half(n, a, "ramp")
    xc, _ = pkg.FBP(fbp_filter=half)(pu, sino)
    assert rel_max(x, xc) < 1e-6
    rng = np.random.default_rng(820)
    maps = rng.uniform(0.3, 1.0, (a, 2, n, n)).astype(np.float32)
    prj = pkg.ProjectorAttenuationXRF([n, n], rad, att_maps=maps, angles_detectors_rad=[np.pi / 2, -np.pi / 2])
    s2 = rng.random((2, a, n)).astype(np.float32)
    x2, _ = pkg.FBP()(prj, s2, lower_limit=0.0)
    ref2 = np.clip(ofl.fbp(lambda y: oxrf.bp_skimage(y.astype(np.float32), rad, n), s2, a, "ramp"), 0.0, None)
    assert rel_max(ref2, x2) < 1e-5


def test_sirt_at_the_reference_cpu_configuration(pkg):
    """The reference's CPU-runnable case: a 128^2 slice of 10 ellipses, attenuation 0.02 on the
    support + 0.005 in and 0.05 out, 256 angles over 2 pi, one detector at pi/2, Poisson counts
    at peak 1e4. The HIP map builder against the oracle's, then 20 SIRT iterations against the
    oracle's SIRT on the oracle projector; then the configuration's full 200 iterations against
    the oracle loop on the C restatement of the projector."""
    from oracle import attenuation as oatt
    from oracle import cops

    n, a = 128, 256
    rng = np.random.default_rng(20260101)
    f = pkg.testing.phantom_ellipses(n, 10, rng, semi_axes=(5, 40))
    mu_in, mu_out = pkg.testing.xrf_attenuation(f)
    ang = np.linspace(0, 2 * np.pi, a, endpoint=False).astype(np.float32)
    prj = pkg.ProjectorAttenuationXRF([n, n], ang, att_in=mu_in, att_out=mu_out, angles_detectors_rad=np.pi / 2)
    maps = oatt.compute_maps(mu_in, mu_out, ang, np.array([np.pi / 2]))
    em = rel_max(maps, prj.att_vol_angles.cpu().numpy().reshape(maps.shape))
    print(f"maps 128^2 x 256: rel_max {em:.2e}")
    assert em < 1e-5
    b = pkg.testing.poisson_data(oxrf.fp(f, ang, maps)[0], 1e4, rng)
    x = pkg.SIRT()(prj, b, iterations=20)[0]
    ref = osol.sirt(lambda v: oxrf.fp(v, ang, maps)[0], lambda y: oxrf.bp_skimage(y[None], ang, n), b, (n, n), 20)
    print(f"SIRT-20 128^2 x 256: rel_max {rel_max(ref, x):.2e} rel_l2 {rel_l2(ref, x):.2e}")
    assert rel_max(ref, x) < TOL
    tabs = cops.tables(ang, n)
    x200 = pkg.SIRT()(prj, b, iterations=200)[0]
    ref200 = osol.sirt(lambda v: cops.fp(v, ang, maps, tabs=tabs)[0], lambda y: cops.bp_skimage(y, ang, n, tabs=tabs),
                       b, (n, n), 200)
    print(f"SIRT-200 128^2 x 256: rel_max {rel_max(ref200, x200):.2e} rel_l2 {rel_l2(ref200, x200):.2e}")
    assert rel_max(ref200, x200) < TOL


def test_pdhg_without_preconditioning(pkg, xrf_problem):
    """``precondition=False``: scalar steps from the power-method norm (``compute_Lipschitz_scaling``,
    ``solvers.py:120-158, 1127-1131``; the same ``np.random.rand`` start vector) against the oracle."""
    p = xrf_problem
    np.random.seed(21)
    x, info = pkg.PDHG(data_term="l2", regularizer=pkg.regularizers.Regularizer_TV2D(1e-2))(
        p["prj"], p["b"], iterations=30, precondition=False)
    np.random.seed(21)
    L = osol.power_method(p["A"], p["At"], np.random.rand(*np.shape(p["b"])).astype(np.float32))
    ref = osol.pdhg(p["A"], p["At"], p["b"], p["shape"], 30, data_term="l2", tv_weight=1e-2, precondition=False, L=L)
    e = rel_max(ref, x)
    print(f"PDHG-l2-TV Lipschitz 30 it: rel_max {e:.2e} (L = {L:.4f})")
    assert e < TOL and info.iterations == 30


def test_residual_tracking_test_mask_and_tolerance(pkg, xrf_problem):
    """``SolutionInfo`` bookkeeping (``solvers.py:1145-1158, 1192-1204``): residual0 and the per-iteration
    residuals of the l2 data term over the reconstruction mask and over ``b_test_mask`` (split from the
    data mask as ``_initialize_b_masks`` does, ``solvers.py:333-343``) against the oracle, and the
    tolerance stopping at the oracle's iteration."""
    p = xrf_problem
    b = p["b"]
    test = (np.random.default_rng(22).random(b.shape) < 0.1).astype(np.float32)
    its = 40
    ref_x, ref = osol.pdhg(p["A"], p["At"], b, p["shape"], its, data_term="l2", tv_weight=1e-2, b_test_mask=test,
                           track=True)
    x, info = pkg.PDHG(data_term="l2", regularizer=pkg.regularizers.Regularizer_TV2D(1e-2))(
        p["prj"], b, iterations=its, b_test_mask=test)
    res, res_cv = np.asarray(ref["residuals"]), np.asarray(ref["residuals_cv"])
    e_x = rel_max(ref_x, x)
    e_r = rel_max(res, info.residuals)
    e_cv = rel_max(res_cv, info.residuals_cv)
    e_0 = abs(info.residual0 - ref["residual0"]) / ref["residual0"]
    e_0cv = abs(info.residual0_cv - ref["residual0_cv"]) / ref["residual0_cv"]
    print(f"tracking: x {e_x:.2e}, residuals {e_r:.2e}, cv {e_cv:.2e}, residual0 {e_0:.2e}, residual0_cv {e_0cv:.2e}")
    assert e_x < TOL and e_r < TOL and e_cv < TOL and e_0 < TOL and e_0cv < TOL
    assert np.allclose(info.residuals_rel, info.residuals / info.residual0)
    # a tolerance between two of the oracle's residuals stops both at the same iteration
    k = next(i for i in range(15, its) if res[i - 1] > res[i] * (1 + 1e-3))
    tol = float(0.5 * (res[k - 1] + res[k]))
    _, ref_t = osol.pdhg(p["A"], p["At"], b, p["shape"], its, data_term="l2", tv_weight=1e-2, b_test_mask=test,
                         track=True, tolerance=tol)
    _, info_t = pkg.PDHG(data_term="l2", regularizer=pkg.regularizers.Regularizer_TV2D(1e-2), tolerance=tol)(
        p["prj"], b, iterations=its, b_test_mask=test)
    print(f"tolerance {tol:.4g}: GPU stops after {info_t.iterations}, oracle after {ref_t['iterations']} (k = {k})")
    assert info_t.iterations == ref_t["iterations"] == k + 1


def test_array_valued_limits(pkg, xrf_problem):
    """``x.clip(lower_limit, upper_limit)`` with array limits (``solvers.py:797-803, 975-982, 1185-1190``) in
    PDHG, SIRT, MLEM and FISTA, with a reconstruction mask, against the oracle."""
    p = xrf_problem
    rng = np.random.default_rng(23)
    lo = (rng.random(p["shape"]) * 0.05 - 0.02).astype(np.float32)
    hi = (0.5 + 0.3 * rng.random(p["shape"])).astype(np.float32)
    x_mask = np.ones(p["shape"], np.float32)
    x_mask[:3] = 0
    A, At, b, sh = p["A"], p["At"], p["b"], p["shape"]
    runs = {
        "PDHG-KL": (pkg.PDHG(data_term="kl", regularizer=pkg.regularizers.Regularizer_TV2D(1e-2))(
            p["prj"], b, 20, lower_limit=lo, upper_limit=hi, x_mask=x_mask)[0],
            osol.pdhg(A, At, b, sh, 20, data_term="kl", tv_weight=1e-2, lower=lo, upper=hi, x_mask=x_mask)),
        "SIRT": (pkg.SIRT()(p["prj"], b, 20, lower_limit=lo, upper_limit=0.6, x_mask=x_mask)[0],
                 osol.sirt(A, At, b, sh, 20, lower=lo, upper=0.6, x_mask=x_mask)),
        "MLEM": (pkg.MLEM()(p["prj"], b, 20, lower_limit=0.01, upper_limit=hi, x_mask=x_mask)[0],
                 osol.mlem(A, At, b, sh, 20, lower=0.01, upper=hi, x_mask=x_mask)),
        "FISTA": (pkg.FISTA()(p["prj"], b, 20, lower_limit=lo, upper_limit=hi, x_mask=x_mask)[0],
                  osol.fista(A, At, b, sh, 20, lower=lo, upper=hi, x_mask=x_mask)),
    }
    for name, (got, ref) in runs.items():
        e = rel_max(ref, got)
        print(f"{name} array limits 20 it: rel_max {e:.2e}")
        assert e < TOL, name
        assert np.all(got[3:] <= np.maximum(hi[3:], 0.6) + 1e-6) and np.all(got[:3] == 0)
