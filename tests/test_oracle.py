"""CPU tests that pin the oracle (no GPU).

The oracle restates corrct's skimage-backend path. It is pinned by
  * bit-level golden vectors made by scikit-image 0.18.3 itself
    (tests/golden/make_golden_skimage.py),
  * closed forms and invariants a plausible mistake would break (axis swap,
    sign of the angle, missing roll term, wrong centre, transposed weights),
  * the reference's own solver convergence tests (tests/test_solvers.py:100-164)
    and TV known answers (tests/test_operators.py:26-71,
    tests/test_regularizers.py:84-103), with seeds.
"""

import math

import os

import numpy as np
import pytest

from oracle import attenuation as oatt
from oracle import skimage_ops as sk
from oracle import solvers as osol
from oracle import xrf as oxrf
from tests.conftest import rel_max

F32 = np.float32


# ---------------------------------------------------------------------------
# golden vectors from scikit-image 0.18.3 (bit-exact)
# ---------------------------------------------------------------------------


def test_golden_radon_bit_exact(golden):
    for n in golden["radon_sizes"]:
        img = golden[f"radon_n{n}_img"]
        lines = np.stack([sk.radon_line(img, d) for d in sk.rad2deg_f32(golden[f"radon_n{n}_rad"])])
        assert np.array_equal(lines, golden[f"radon_n{n}_lines"]), n


def test_golden_iradon_bp_bit_exact(golden):
    for n in (8, 16, 17):
        vol = oxrf.bp_skimage(golden[f"bp_n{n}_prj"][None], golden[f"bp_n{n}_rad"], n)
        assert np.array_equal(vol, golden[f"bp_n{n}_vol"]), n


def test_golden_rotate_bit_exact(golden):
    img = golden["rot_img"]
    for d, o in zip(golden["rot_deg_f32"], golden["rot_out_f32"]):
        assert np.array_equal(sk.rotate(img, d), o), d
    for d, o in zip(golden["rot_deg_f64"], golden["rot_out_f64"]):
        assert np.array_equal(sk.rotate(img, np.float64(d)), o), d


def test_golden_attenuation_and_maps_bit_exact(golden):
    for n in (16, 21):
        mu = golden[f"att_n{n}_mu"]
        rad = golden[f"att_n{n}_rad"]
        det = golden[f"att_n{n}_det"]
        e_in = np.stack([oatt.compute_attenuation(mu, a) for a in rad])
        assert np.array_equal(e_in, golden[f"att_n{n}_ein"])
        e_out = np.stack([[oatt.compute_attenuation(mu, np.float64(a) + p, invert=True) for p in det] for a in rad])
        assert np.array_equal(e_out, golden[f"att_n{n}_eout"])
        maps = oatt.compute_maps(mu, mu, rad, det)
        assert np.array_equal(maps, golden[f"att_n{n}_maps"])


def test_golden_xrf_fp_bit_exact(golden):
    sino = oxrf.fp(golden["xrf_vol"], golden["xrf_rad"], golden["xrf_maps"], w_det=golden["xrf_wdet"])
    assert np.array_equal(sino, golden["xrf_sino"])


def test_golden_xrf_symmetric_bp_bit_exact(golden):
    """The symmetric XRF bp's composition (per angle: weighted one-angle iradon per detector,
    times that detector's map, summed over detectors; then over angles) against scikit-image
    composed as corrct does (projectors.py:556-565, 612-616); swapping the detectors' maps, a
    plausible indexing slip, misses by far."""
    n = golden["xrf_vol"].shape[-1]
    args = (golden["xrf_bp_sino"], golden["xrf_rad"], n)
    got = oxrf.bp_skimage(*args, golden["xrf_maps"], w_det=golden["xrf_wdet"], symmetric=True)
    assert np.array_equal(got, golden["xrf_bp_sym"])
    swapped = oxrf.bp_skimage(*args, golden["xrf_maps"][:, ::-1], w_det=golden["xrf_wdet"], symmetric=True)
    assert np.abs(swapped - golden["xrf_bp_sym"]).max() > 1e-2 * np.abs(golden["xrf_bp_sym"]).max()


def test_angle_conversions_match_numpy():
    x = (np.random.default_rng(0).random(10000) * 7).astype(F32)
    assert np.array_equal(sk.rad2deg_f32(x), np.rad2deg(x))
    d = np.rad2deg(x)
    assert np.array_equal(sk.deg2rad_f32(d), np.deg2rad(d))


# ---------------------------------------------------------------------------
# closed forms: radon
# ---------------------------------------------------------------------------


def test_radon_theta0_is_column_sum():
    rng = np.random.default_rng(1)
    img = rng.random((24, 24)).astype(F32)
    line = sk.radon_line(img, F32(0.0))
    ref = np.zeros(24, F32)
    for row in img:
        ref += row
    assert np.array_equal(line, ref)


@pytest.mark.parametrize("n", [16, 17])
def test_radon_theta90_reads_flipped_rows_fp64(n):
    rng = np.random.default_rng(2)
    img = rng.random((n, n))
    c = n // 2
    line = sk.radon_line(img, 90.0, fp32=False)
    ref = np.array([img[2 * c - u].sum() if 0 <= 2 * c - u < n else 0.0 for u in range(n)])
    assert np.max(np.abs(line - ref)) < 1e-9


def test_radon_impulse_centre_of_mass():
    """An impulse at (row i0, col j0) lands at u = c + cos(j0 - c) - sin(i0 - c): pins sign and axes."""
    n = 33
    c = n // 2
    for deg in (17.0, 63.0, 121.0, 250.0):
        for i0, j0 in ((10, 20), (20, 9)):
            img = np.zeros((n, n))
            img[i0, j0] = 1.0
            p = sk.radon_line(img, deg, fp32=False)
            th = math.radians(deg)
            u_ref = c + math.cos(th) * (j0 - c) - math.sin(th) * (i0 - c)
            assert abs(np.sum(np.arange(n) * p) / np.sum(p) - u_ref) < 0.25, (deg, i0, j0)


def test_fp_fp32_vs_fp64_error_small():
    rng = np.random.default_rng(3)
    n = 32
    img = rng.random((n, n)).astype(F32)
    yy, xx = np.mgrid[:n, :n] - n // 2
    img[(yy**2 + xx**2) > (n // 2) ** 2] = 0
    ang = np.linspace(0, 2 * np.pi, 16, endpoint=False).astype(F32)
    s32 = oxrf.fp(img, ang)
    s64 = oxrf.fp(img, ang, fp32=False)
    assert rel_max(s64, s32) < 1e-5


# ---------------------------------------------------------------------------
# closed forms: iradon bp
# ---------------------------------------------------------------------------


def test_iradon_theta0_replicates_column():
    rng = np.random.default_rng(4)
    n = 16
    col = rng.random(n).astype(F32)
    img = sk.iradon_line(col, F32(0.0), n) * F32(2) / F32(np.pi)
    assert np.max(np.abs(img - col[None, :])) < 1e-6


def test_iradon_constant_column_is_indicator():
    n = 17
    c = n // 2
    col = np.ones(n, dtype=F32)
    deg = F32(33.0)
    img = sk.iradon_line(col, deg, n, fp32=False) * (2 / np.pi)
    th = math.radians(float(deg))
    ii, jj = np.mgrid[:n, :n]
    t = (jj - c) * math.cos(th) - (ii - c) * math.sin(th)
    inside = (t >= -c) & (t <= n - 1 - c)
    assert np.allclose(img, inside.astype(float), atol=1e-12)


# ---------------------------------------------------------------------------
# adjoint pair
# ---------------------------------------------------------------------------


def test_bp_adjoint_is_transpose_of_explicit_fp():
    n = 8
    rng = np.random.default_rng(5)
    ang = np.linspace(0, 2 * np.pi, 6, endpoint=False).astype(F32)
    maps = rng.random((6, 2, n, n)).astype(F32)
    w_det = np.array([1.0, 0.7])
    cols = []
    for k in range(n * n):
        e = np.zeros(n * n)
        e[k] = 1
        cols.append(oxrf.fp(e.reshape(n, n), ang, maps, w_det=w_det, fp32=False).ravel())
    a_mat = np.stack(cols, axis=1)
    y = rng.random(a_mat.shape[0])
    bt = oxrf.bp_adjoint(y, ang, n, maps, w_det=w_det)
    # fp in FP64 uses float64 coordinates; bp_adjoint uses the float32 ones the GPU uses
    assert rel_max(a_mat.T @ y, bt.ravel()) < 1e-5
    x = rng.random(n * n)
    lhs = np.dot(oxrf.fp(x.reshape(n, n), ang, maps, w_det=w_det, fp32=False).ravel(), y)
    rhs = np.dot(x, bt.ravel())
    assert abs(lhs - rhs) / abs(lhs) < 1e-5


# ---------------------------------------------------------------------------
# attenuation invariants
# ---------------------------------------------------------------------------


def test_maps_zero_mu_is_one():
    mu = np.zeros((12, 12), F32)
    ang = np.linspace(0, 2 * np.pi, 5, endpoint=False).astype(F32)
    maps = oatt.compute_maps(mu, mu, ang, [np.pi / 2, -np.pi / 2])
    assert np.array_equal(maps, np.ones_like(maps))


def test_maps_multiplicative():
    rng = np.random.default_rng(6)
    mu_in = (rng.random((14, 14)) * 0.05).astype(F32)
    mu_out = (rng.random((14, 14)) * 0.05).astype(F32)
    ang = np.linspace(0, 2 * np.pi, 4, endpoint=False).astype(F32)
    det = [np.pi / 2]
    m = oatt.compute_maps(mu_in, mu_out, ang, det)
    m_in = oatt.compute_maps(mu_in, None, ang, det)
    m_out = oatt.compute_maps(None, mu_out, ang, det)
    assert np.array_equal(m, (m_in * m_out).astype(F32))


def test_detector_sub_angles_closed_form():
    # surface 4 mm^2 at 1 mm: half-opening atan(1) = pi / 4, so two parts have midpoints -+pi / 8
    assert np.allclose(oatt.detector_sub_angles(4.0, 1.0, 2), [-np.pi / 8, np.pi / 8], rtol=0, atol=1e-15)
    assert np.allclose(oatt.detector_sub_angles(4.0, 1.0, 4), np.array([-3, -1, 1, 3]) * np.pi / 16, atol=1e-15)
    assert np.array_equal(oatt.detector_sub_angles(4.0, 1.0, 1), [0.0])
    assert np.array_equal(oatt.detector_sub_angles(0.0, 1.0, 5), [0.0])


def test_sub_angle_maps_are_mean_of_shifted_detectors():
    """Averaging over offsets s_k equals the mean of point detectors at phi + s_k (to the
    rounding of (theta + phi) + s against theta + (phi + s)); asymmetric offsets so a sign or
    normalisation slip shows; the in-beam factor multiplies the mean."""
    rng = np.random.default_rng(8)
    mu_in = (rng.random((16, 16)) * 0.05).astype(F32)
    mu_out = (rng.random((16, 16)) * 0.1).astype(F32)
    ang = np.linspace(0, 2 * np.pi, 5, endpoint=False).astype(F32)
    phi, subs = np.pi / 2, np.array([0.1, 0.25, -0.05])
    m = oatt.compute_maps(None, mu_out, ang, [phi], sub_angles=[subs])
    parts = [oatt.compute_maps(None, mu_out, ang, [phi + s]) for s in subs]
    assert rel_max(np.mean(parts, axis=0), m) < 1e-5
    assert rel_max(parts[0], m) > 1e-3  # the offsets matter at this mu
    m_in = oatt.compute_maps(mu_in, None, ang, [phi])
    m_both = oatt.compute_maps(mu_in, mu_out, ang, [phi], sub_angles=[subs])
    assert np.array_equal(m_both, (m_in * m).astype(F32))
    # a single zero offset is the point detector, bit for bit
    assert np.array_equal(oatt.compute_maps(mu_in, mu_out, ang, [phi], sub_angles=[np.zeros(1)]),
                          oatt.compute_maps(mu_in, mu_out, ang, [phi]))


def test_log_map_linear_in_mu_fp64():
    rng = np.random.default_rng(7)
    mu = rng.random((10, 10)) * 0.05
    e1 = oatt.compute_attenuation(mu, np.float64(0.7), fp32=False)
    e3 = oatt.compute_attenuation(3.0 * mu, np.float64(0.7), fp32=False)
    assert np.allclose(-np.log(e3), 3.0 * -np.log(e1), rtol=1e-10, atol=1e-12)


def test_theta0_trapezoid_closed_form_fp64():
    """At angle 0 the rotations are the identity: E = exp(-trapezoid prefix sum down the rows)."""
    rng = np.random.default_rng(8)
    n = 10
    mu = rng.random((n, n)) * 0.1
    e = oatt.pad_edges(n)
    mp = np.pad(mu, e)
    npad = n + 2 * e
    cum = np.zeros_like(mp)
    for i in range(npad):
        prev = mp[i - 1] if i > 0 else mp[npad - 1]
        cum[i] = (cum[i - 1] if i > 0 else 0) + (mp[i] + prev) / 2
    ref = np.exp(-cum[e:-e, e:-e])
    got = oatt.compute_attenuation(mu, np.float64(0.0), fp32=False)
    assert np.allclose(got, ref, rtol=1e-12)


def test_pad_edges_matches_reference_formula():
    # padded sides Np = N + 2e for N = 128 ... 1024 (ceil(sqrt(2) N) rounded up to N + even)
    for n, npad in ((128, 182), (256, 364), (384, 544), (512, 726), (768, 1088), (1024, 1450)):
        assert n + 2 * oatt.pad_edges(n) == npad


# ---------------------------------------------------------------------------
# solvers: the reference's convergence tests (seeded) and TV known answers
# ---------------------------------------------------------------------------


@pytest.fixture(scope="module")
def rand_2d():
    rng = np.random.default_rng(9)
    vol = np.fmin(rng.random((10, 10)) + np.finfo(F32).eps, 1).astype(F32)
    mat = (rng.random((500, 100)) > 0.5).astype(F32)
    data = (mat @ vol.ravel()).reshape(5, 100).astype(F32)
    return mat, vol, data


def _ops(mat, x_shape, b_shape):
    return (lambda x: (mat @ x.ravel()).reshape(b_shape).astype(F32),
            lambda y: (mat.T @ y.ravel()).reshape(x_shape).astype(F32))


@pytest.mark.parametrize("algo,its", [("sirt", 1000), ("pdhg", 500), ("mlem", 10000)])
def test_oracle_solvers_converge_like_reference(rand_2d, algo, its):
    mat, vol, data = rand_2d
    A, At = _ops(mat, vol.shape, data.shape)
    if algo == "sirt":
        x = osol.sirt(A, At, data, vol.shape, its)
    elif algo == "pdhg":
        x = osol.pdhg(A, At, data, vol.shape, its)
    else:
        x = osol.mlem(A, At, data, vol.shape, its)
    assert np.all(np.isclose(x, vol, atol=1e-2))


def test_tv_known_answers():
    rng = np.random.default_rng(10)
    x = rng.random((6, 7)).astype(F32)
    g = osol.grad(x)
    assert g.shape == (2, 6, 7)
    assert np.all(g[0, -1, :] == 0) and np.all(g[1, :, -1] == 0)
    assert np.allclose(g[0, :-1], np.diff(x, axis=0)) and np.allclose(g[1, :, :-1], np.diff(x, axis=1))
    q = rng.random((2, 6, 7)).astype(F32)
    d = osol.div(q)
    assert np.all(d[0, :] == (q[1, 0, 1:] - q[1, 0, :-1]).sum() * 0 + np.concatenate([[0], q[1, 0, 1:] - q[1, 0, :-1]]))
    # the dual update with zero dual and a weight larger than |grad|/2 is grad/2 (tests/test_regularizers.py:84-103)
    q0 = osol.tv_dual(np.zeros((2, 6, 7), F32), x, weight=10.0)
    assert np.allclose(q0, 0.5 * g)
    # the l12 prox projects onto the lambda ball
    qb = osol.tv_dual(np.zeros((2, 6, 7), F32), 100 * x, weight=0.1)
    assert np.all(np.sqrt((qb**2).sum(0)) <= 0.1 + 1e-6)


def test_kl_prox_closed_form_minimises():
    """prox of sigma KL*: argmin_p (p - v)^2 / 2 + sigma f*(p/sigma) ... checked via the Moreau identity numerically."""
    from scipy.optimize import minimize_scalar

    for v, s, b in ((0.3, 0.5, 2.0), (-1.0, 2.0, 0.5), (2.5, 0.1, 3.0)):
        p = (1 + v - math.sqrt((v - 1) ** 2 + 4 * s * b)) / 2
        # dual of KL(b, .): f*(p) = -b log(1 - p) (p < 1); prox_{s f*}(v) minimises (p - v)^2 / (2 s) - b log(1 - p)
        res = minimize_scalar(lambda t: (t - v) ** 2 / (2 * s) - b * math.log(1 - t), bounds=(-50, 1 - 1e-12), method="bounded",
                              options={"xatol": 1e-12})
        assert abs(res.x - p) < 1e-6


def test_oracle_sart_solves_consistent_system():
    """SART (oracle) on a consistent non-negative system split into 8 'angles' of 12 rows:
    converges to the true solution; a single 'angle' reduces to one preconditioned
    Landweber step x - tau A^T(sigma (A x - b))."""
    from oracle import solvers as osol

    rng = np.random.default_rng(77)
    n_ang, rows_per, nx = 8, 12, 30
    M = rng.random((n_ang, rows_per, nx)).astype(np.float32)
    x_true = rng.random(nx).astype(np.float32)
    b = np.einsum("arx,x->ar", M, x_true).astype(np.float32)  # [A, N] layout

    def A_angle(x, a):
        return (M[a] @ x).astype(np.float32)

    def At_angle(y, a):
        return (M[a].T @ y).astype(np.float32)

    x = osol.sart(A_angle, At_angle, b, (nx,), n_ang, 400, rng.permutation(n_ang))
    assert np.max(np.abs(x - x_true)) < 1e-3
    x1 = osol.sart(lambda x, a: M[0] @ x, lambda y, a: M[0].T @ y, b[:1], (nx,), 1, 1, [0])
    tau = 1 / np.abs(M[0].T @ np.ones(rows_per))
    sigma = 1 / np.abs(M[0] @ np.ones(nx))
    assert np.allclose(x1, -tau * (M[0].T @ (sigma * (0 - b[0]))), rtol=1e-5)


def test_oracle_power_method_and_fista():
    """power_method -> the largest singular value (numpy SVD); FISTA on a consistent system
    converges to the solution with both step-size rules; the l1 prox is soft thresholding."""
    from oracle import solvers as osol

    rng = np.random.default_rng(78)
    M = rng.random((60, 25)).astype(np.float64)
    x_true = rng.random(25)
    b = (M @ x_true).astype(np.float32)
    L = osol.power_method(lambda x: M @ x, lambda y: M.T @ y, rng.random(60), iterations=60)
    assert abs(L - np.linalg.svd(M, compute_uv=False)[0]) < 1e-6 * L
    for pre in (True, False):
        x = osol.fista(lambda v: (M @ v).astype(np.float32), lambda y: (M.T @ y).astype(np.float32), b, (25,),
                       3000, precondition=pre, L=L)
        assert np.max(np.abs(x - x_true)) < 2e-3, pre
    # pure proximal step: with M = I, one step from 0 with tau = 1 is soft(b, l1)
    e = np.eye(4, dtype=np.float32)
    bb = np.array([0.5, -0.05, 0.2, -1.0], dtype=np.float32)
    x1 = osol.fista(lambda v: e @ v, lambda y: e @ y, bb, (4,), 1, precondition=False, L=1.0, l1=0.1)
    assert np.allclose(x1, np.sign(bb) * np.fmax(np.abs(bb) - 0.1, 0))


def test_oracle_fbp_filters_and_iradon_golden():
    """The oracle's FBP filters equal scikit-image's (all five names, sizes 64 and 128), and
    ramp filter + iradon-style back-projection equals scikit-image's filtered iradon. (For the
    windowed filters corrct multiplies the real FFT by the first half of the filter, which is not
    Hermitian-symmetric for hann / hamming / cosine, so it deliberately differs from iradon's
    full complex FFT there: 3 % for hann.)"""
    from oracle import filters as ofl
    from oracle import xrf as oxrf

    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "skimage_fbp_golden.npz"))
    for size in (64, 128):
        for name in ofl.NAMES:
            assert np.allclose(ofl.fourier_filter(size, name), g[f"filter_{name}_{size}"], rtol=0, atol=1e-12), name
    assert ofl.padding_size(24) == 64 and ofl.padding_size(100) == 256
    sino, rad = g["fbp_sino"], g["fbp_rad"]
    n = sino.shape[-1]
    got = ofl.fbp(lambda y: oxrf.bp_skimage(y[None].astype(np.float32), rad, n), sino, len(rad), "ramp")
    ref = g["fbp_ramp_vol"]
    assert np.max(np.abs(got - ref)) < 1e-5 * np.max(np.abs(ref))
    got = ofl.fbp(lambda y: oxrf.bp_skimage(y[None].astype(np.float32), rad, n), sino, len(rad), "hann")
    assert 1e-2 < np.max(np.abs(got - g["fbp_hann_vol"])) / np.max(np.abs(g["fbp_hann_vol"])) < 0.1


def test_oracle_psf_matches_scipy_edge_convolution():
    """The oracle's detector PSF equals the reference's recipe executed with scipy: edge padding
    by (K-1)//2, scipy.signal.convolve(mode="same"), crop (``operators.py:462-480``)."""
    import scipy.signal as spsig

    from oracle import xrf as oxrf

    rng = np.random.default_rng(79)
    x = rng.random((5, 37))
    for k_len in (1, 3, 4, 7):
        kern = rng.random(k_len)
        pw = (k_len - 1) // 2
        xp = np.pad(x, ((0, 0), (pw, pw)), mode="edge")
        ref = spsig.convolve(xp, kern[None, :], mode="same")
        ref = ref[:, pw:xp.shape[1] - pw] if pw else ref
        assert np.allclose(oxrf.psf_lines(x, kern), ref, rtol=1e-12, atol=1e-12), k_len


def test_bin_maps_known_answers():
    """get_maps(binning) = bin_imgs(auto_crop=True): a 5 x 5 image drops its first row and column
    (excess 1: 1 - 0 from the start, 0 from the end) and averages 2 x 2 blocks; a 3-D volume bins
    its slice axis too; constants stay constant."""
    from oracle import attenuation as oatt

    img = np.arange(25, dtype=np.float32).reshape(1, 5, 5)
    want = np.array([[[(6 + 7 + 11 + 12) / 4, (8 + 9 + 13 + 14) / 4],
                      [(16 + 17 + 21 + 22) / 4, (18 + 19 + 23 + 24) / 4]]], np.float32)
    assert np.array_equal(oatt.bin_maps(img, 2, 2), want)
    vol = np.arange(2 * 4 * 6, dtype=np.float32).reshape(1, 2, 4, 6)
    got = oatt.bin_maps(vol, 2, 3)
    assert got.shape == (1, 1, 2, 3)
    assert got[0, 0, 0, 0] == np.mean([0, 1, 6, 7, 24, 25, 30, 31])
    c = np.full((3, 2, 7, 9), 0.25, np.float32)
    assert np.array_equal(oatt.bin_maps(c, 3, 2), np.full((3, 2, 2, 3), 0.25, np.float32))
