"""CPU restatement of corrct's XRF projector on the scikit-image backend.

Test infrastructure only (see ``oracle/__init__.py``).

Layouts: volume [N, N]; maps [A, D, N, N]; sinogram [D, A, N].

* ``fp``          -- ``ProjectorAttenuationXRF.fp`` / ``fp_angle``
                     (``projectors.py:502-531, 570-596``): per angle a and
                     detector d, ``w_det[d] w_ang[a, d] radon_a(vol * M[a, d])``.
* ``bp_skimage``  -- ``ProjectorAttenuationXRF.bp`` (``projectors.py:598-622``):
                     non-symmetric = ``sum_d iradon-bp(sino[d])`` (no maps, no
                     weights); symmetric = ``sum_a sum_d w M[a, d] bp_a(sino[d, a])``.
                     The single-detector bp is ``ProjectorBackendSKimage.bp``
                     (``_projector_backends.py:335-359``): per-angle iradon
                     images summed over angles, times ``2 / pi``.
* ``bp_adjoint``  -- the exact transpose of ``fp`` (no reference counterpart),
                     written as a scatter of each sample's bilinear weights,
                     a code path independent of ``fp``'s gather.
"""

from __future__ import annotations

import numpy as np

from . import skimage_ops as sk
from .skimage_ops import F32


def _weights(w_det, w_ang, n_angles: int, n_det: int) -> np.ndarray:
    """``weights_det * weights_angles[a, :]`` for every angle, [A, D] float64."""
    wd = np.ones(n_det) if w_det is None else np.broadcast_to(np.asarray(w_det, dtype=np.float64), (n_det,))
    wa = np.ones((n_angles, n_det)) if w_ang is None else np.asarray(w_ang, dtype=np.float64).reshape(n_angles, n_det)
    return wd[None, :] * wa


def fp(vol, angles_rad_f32, maps=None, w_det=None, w_ang=None, n_det: int = 1, fp32: bool = True) -> np.ndarray:
    """Attenuation-corrected forward projection, output [D, A, N]."""
    angles = np.asarray(angles_rad_f32, dtype=F32)
    degs = sk.rad2deg_f32(angles)
    n = vol.shape[-1]
    if maps is not None:
        n_det = maps.shape[1]
    w = _weights(w_det, w_ang, len(angles), n_det)
    dt = F32 if fp32 else np.float64
    out = np.zeros((n_det, len(angles), n), dtype=dt)
    v = vol.astype(dt)
    for ia in range(len(angles)):
        theta = degs[ia] if fp32 else np.float64(angles[ia]) * (180.0 / np.pi)
        for d in range(n_det):
            t = v if maps is None else v * maps[ia, d].astype(dt)
            line = sk.radon_line(t, theta, fp32=fp32)
            out[d, ia] = (line * dt(w[ia, d])) if fp32 else line * w[ia, d]
    return out


def _bp_single(sino_an, degs, angles, n, fp32: bool):
    """``ProjectorBackendSKimage.bp`` of one detector's [A, N] sinogram."""
    imgs = []
    for ia in range(sino_an.shape[0]):
        theta = degs[ia] if fp32 else np.float64(angles[ia]) * (180.0 / np.pi)
        imgs.append(sk.iradon_line(sino_an[ia], theta, n, fp32=fp32))
    acc = np.stack(imgs).sum(axis=0)
    if fp32:
        return (acc * F32(2)) / F32(np.pi)
    return acc * (2 / np.pi)


def bp_skimage(
    sino, angles_rad_f32, n: int, maps=None, w_det=None, w_ang=None, symmetric: bool = False, fp32: bool = True
) -> np.ndarray:
    """XRF back-projection with the skimage (iradon) discretisation, output [N, N]."""
    angles = np.asarray(angles_rad_f32, dtype=F32)
    degs = sk.rad2deg_f32(angles)
    n_angles = len(angles)
    sino = np.asarray(sino)
    n_det = sino.size // (n_angles * n)
    sino = sino.reshape(n_det, n_angles, n)
    dt = F32 if fp32 else np.float64
    if not symmetric:
        acc = np.zeros((n, n), dtype=dt)
        for d in range(n_det):
            acc += _bp_single(sino[d], degs, angles, n, fp32)
        return acc
    w = _weights(w_det, w_ang, n_angles, n_det)
    acc = np.zeros((n, n), dtype=dt)
    for ia in range(n_angles):
        theta = degs[ia] if fp32 else np.float64(angles[ia]) * (180.0 / np.pi)
        per_det = np.zeros((n, n), dtype=dt)
        for d in range(n_det):
            img = sk.iradon_line(sino[d, ia], theta, n, fp32=fp32)
            img = (img * F32(2)) / F32(np.pi) if fp32 else img * (2 / np.pi)
            img = img * dt(w[ia, d])
            if maps is not None:
                img = img * maps[ia, d].astype(dt)
            per_det += img
        acc += per_det
    return acc


def bilinear_weights(n: int, theta_deg, fp32: bool = True):
    """Sparse list of (sample u, pixel row i, pixel col j, weight) for one angle.

    Sample (y, u) of ``radon_line`` reads pixel (i, j) with the product of the
    1-D hat weights of its bilinear blend. Built by scattering every sample's
    four taps, independently of the gather in ``radon_line``.
    """
    m = sk.radon_matrix(theta_deg, n, fp32)
    yy, uu = np.mgrid[:n, :n]
    if fp32:
        m = m.astype(F32)
        x = (m[0, 0] * uu.astype(F32) + m[0, 1] * yy.astype(F32)) + m[0, 2]
        y = (m[1, 0] * uu.astype(F32) + m[1, 1] * yy.astype(F32)) + m[1, 2]
    else:
        x = (m[0, 0] * uu + m[0, 1] * yy) + m[0, 2]
        y = (m[1, 0] * uu + m[1, 1] * yy) + m[1, 2]
    x = x.astype(np.float64).ravel()
    y = y.astype(np.float64).ravel()
    u = uu.ravel()
    r0 = np.floor(y)
    c0 = np.floor(x)
    dr = y - r0
    dc = x - c0
    r0 = r0.astype(np.int64)
    c0 = c0.astype(np.int64)
    us, iis, jjs, ws = [], [], [], []
    for oi, oj, wt in ((0, 0, (1 - dr) * (1 - dc)), (0, 1, (1 - dr) * dc), (1, 0, dr * (1 - dc)), (1, 1, dr * dc)):
        ii = r0 + oi
        jj = c0 + oj
        ok = (ii >= 0) & (ii < n) & (jj >= 0) & (jj < n) & (wt != 0)
        us.append(u[ok])
        iis.append(ii[ok])
        jjs.append(jj[ok])
        ws.append(wt[ok])
    return np.concatenate(us), np.concatenate(iis), np.concatenate(jjs), np.concatenate(ws)


def bp_adjoint(sino, angles_rad_f32, n: int, maps=None, w_det=None, w_ang=None) -> np.ndarray:
    """Exact transpose of ``fp`` (float64), output [N, N]."""
    angles = np.asarray(angles_rad_f32, dtype=F32)
    degs = sk.rad2deg_f32(angles)
    n_angles = len(angles)
    sino = np.asarray(sino, dtype=np.float64)
    n_det = sino.size // (n_angles * n)
    sino = sino.reshape(n_det, n_angles, n)
    w = _weights(w_det, w_ang, n_angles, n_det)
    out = np.zeros((n, n), dtype=np.float64)
    for ia in range(n_angles):
        u, ii, jj, wt = bilinear_weights(n, degs[ia], fp32=True)
        for d in range(n_det):
            img = np.zeros((n, n), dtype=np.float64)
            np.add.at(img, (ii, jj), wt * sino[d, ia, u] * w[ia, d])
            if maps is not None:
                img *= maps[ia, d].astype(np.float64)
            out += img
    return out


def explicit_fp_matrix(n: int, angles_rad_f32, maps=None, w_det=None, w_ang=None, n_det: int = 1) -> np.ndarray:
    """Dense [D*A*N, N*N] matrix of ``fp`` built column by column from unit volumes."""
    cols = []
    for k in range(n * n):
        e = np.zeros(n * n, dtype=F32)
        e[k] = 1
        cols.append(fp(e.reshape(n, n), angles_rad_f32, maps, w_det, w_ang, n_det, fp32=False).ravel())
    return np.stack(cols, axis=1)


def psf_lines(prj, kern) -> np.ndarray:
    """Detector PSF along the last axis: ``TransformConvolution`` with edge padding
    (``operators.py:471-480``): out[i] = sum_k xe[i + (K-1)//2 - k] kern[k], with xe the line
    extended by (K-1)//2 edge copies on each side and zero beyond (an even-length kernel
    reaches one sample past the padding on the low side)."""
    x = np.asarray(prj, dtype=np.float64)
    kern = np.asarray(kern, dtype=np.float64).ravel()
    n, k_len = x.shape[-1], len(kern)
    pw = (k_len - 1) // 2
    out = np.zeros_like(x)
    for k in range(k_len):
        m = np.arange(n) + pw - k
        inside = (m >= -pw) & (m <= n - 1 + pw)
        out += np.where(inside, x[..., np.clip(m, 0, n - 1)], 0.0) * kern[k]
    return out
