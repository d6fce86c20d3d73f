"""The oracle's C restatement (CPU baseline) agrees with the NumPy oracle."""

import numpy as np

from oracle import build_c
from oracle import skimage_ops as sk
from oracle import xrf as oxrf
from tests.conftest import rel_max


def geom_tables(ang):
    degs = sk.rad2deg_f32(ang)
    n_a = len(ang)
    geom = np.zeros((n_a, 4), np.float32)
    trig = np.zeros((n_a, 2), np.float32)
    return degs, geom, trig


def c_tables(ang, n):
    degs = sk.rad2deg_f32(ang)
    geom = np.stack([sk.radon_matrix(d, n)[[0, 0, 0, 1], [0, 1, 2, 2]] for d in degs]).astype(np.float32)
    th = sk.deg2rad_f32(degs)
    trig = np.stack([np.cos(th), np.sin(th)], axis=1).astype(np.float32)
    return np.ascontiguousarray(geom), np.ascontiguousarray(trig)


def test_c_oracle_matches_numpy_oracle():
    lib = build_c.load()
    n, a, d = 24, 10, 2
    rng = np.random.default_rng(12)
    vol = rng.random((n, n)).astype(np.float32)
    maps = rng.random((a, d, n, n)).astype(np.float32)
    ang = np.linspace(0, 2 * np.pi, a, endpoint=False).astype(np.float32)
    w = rng.uniform(0.5, 1.5, (a, d)).astype(np.float32)
    geom, trig = c_tables(ang, n)
    sino = np.zeros((d, a, n), np.float32)
    lib.oracle_fp(vol.ctypes.data, maps.ctypes.data, geom.ctypes.data, w.ctypes.data, n, a, d, sino.ctypes.data)
    ref = oxrf.fp(vol, ang, maps, w_ang=w.astype(np.float64))
    assert rel_max(ref, sino) < 1e-6
    bp = np.zeros((n, n), np.float32)
    y = rng.random((d, a, n)).astype(np.float32)
    lib.oracle_bp_skimage(y.ctypes.data, trig.ctypes.data, n, a, d, bp.ctypes.data)
    assert rel_max(oxrf.bp_skimage(y, ang, n), bp) < 1e-6


def test_c_adjoint_matches_numpy_adjoint_and_dot_test():
    """The C scatter adjoint equals the NumPy oracle's transpose of fp (float64 scatter) and
    passes the dot test against the C fp; a dropped map factor or a transposed tap fails both."""
    from oracle import cops

    n, a, d = 20, 9, 2
    rng = np.random.default_rng(13)
    ang = np.linspace(0, 2 * np.pi, a, endpoint=False).astype(np.float32)
    maps = rng.uniform(0.2, 1.0, (a, d, n, n)).astype(np.float32)
    w_det = np.array([0.8, 1.2])
    w_ang = rng.uniform(0.5, 1.5, (a, d))
    y = rng.random((d, a, n)).astype(np.float32)
    got = cops.bp_adjoint(y, ang, n, maps, w_det, w_ang)
    ref = oxrf.bp_adjoint(y, ang, n, maps, w_det, w_ang)
    assert rel_max(ref, got) < 1e-6
    x = rng.random((n, n)).astype(np.float32)
    ax = cops.fp(x, ang, maps, w_det, w_ang)
    lhs = float(np.dot(ax.ravel().astype(np.float64), y.ravel().astype(np.float64)))
    rhs = float(np.dot(x.ravel().astype(np.float64), got.ravel().astype(np.float64)))
    assert abs(lhs - rhs) / abs(lhs) < 1e-6
    # the explicit matrix's transpose (FP64), independent of both scatters
    mat = oxrf.explicit_fp_matrix(n, ang, maps, w_det, w_ang)
    assert rel_max(mat.T @ y.ravel().astype(np.float64), got.ravel()) < 1e-5
    # planted mistakes the checks must catch: maps ignored, taps transposed
    wrong = cops.bp_adjoint(y, ang, n, None, w_det, w_ang)
    assert rel_max(ref, wrong) > 1e-2
    assert rel_max(ref, got.T) > 1e-2
