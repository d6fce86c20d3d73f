"""CPU restatement of the scikit-image 0.18.3 operations corrct's skimage backend calls.

Test infrastructure only (see ``oracle/__init__.py``).

scikit-image is third-party to corrct (unpinned: ``pyproject.toml:35``). The
calls restated here are

* ``skt.radon(vol, [deg])``       -- ``_projector_backends.py:330,333``
* ``skt.iradon(col, [deg], circle=False, output_size=N, filter_name=None)``
                                   -- ``_projector_backends.py:355,358``
* ``skt.rotate(vol, deg, order=1, clip=False)``
                                   -- ``_projector_backends.py:68,73``

following the published scikit-image 0.18.3 algorithm (``radon_transform.py``
``radon``/``iradon``, ``_warps.py`` ``rotate``/``warp``, ``_geometric.py``
``SimilarityTransform``): an affine inverse map evaluated in the image dtype,
then order-1 interpolation with constant-0 boundary whose blend is done in
double precision, rounded to the image dtype per output pixel.

Two precisions:
  ``fp32=True``  reproduces the reference's float32 arithmetic op by op;
  ``fp32=False`` evaluates the same formulas in float64 (the FP64 evaluation
                 the float32 error is stated against).
"""

from __future__ import annotations

import math

import numpy as np

F32 = np.float32

# numpy's float32 rad2deg multiplies by (180f / pi_f) in float32
_R2D_F32 = F32(180.0) / F32(np.pi)
_D2R_F32 = F32(np.pi / 180.0)


def rad2deg_f32(x):
    """numpy ``rad2deg`` on float32 input (checked bit-exact against numpy)."""
    return (np.asarray(x, dtype=F32) * _R2D_F32).astype(F32)


def deg2rad_f32(x):
    """numpy ``deg2rad`` on float32 input (checked bit-exact against numpy)."""
    return (np.asarray(x, dtype=F32) * _D2R_F32).astype(F32)


# ---------------------------------------------------------------------------
# bilinear sampling (warp, order=1, mode='constant', cval=0)
# ---------------------------------------------------------------------------


def bilinear_sample(img: np.ndarray, r: np.ndarray, c: np.ndarray, fp32: bool = True) -> np.ndarray:
    """Order-1 sample of ``img`` at (row ``r``, col ``c``), 0 outside.

    Taps: (floor r, floor c), (floor r, ceil c), (ceil r, floor c),
    (ceil r, ceil c); ``dr = r - floor r``; blend
    ``((1-dc) TL + dc TR)(1-dr) + ((1-dc) BL + dc BR) dr``. With ``fp32`` the
    products ``dc TR`` and ``dc BR`` are float32, the rest float64, and the
    result is rounded to float32 (the output image dtype).
    """
    rows, cols = img.shape
    r64 = np.asarray(r, dtype=np.float64)
    c64 = np.asarray(c, dtype=np.float64)
    r0 = np.floor(r64)
    c0 = np.floor(c64)
    r1 = np.ceil(r64)
    c1 = np.ceil(c64)
    if fp32:
        # r - floor(r) is formed in float32 (inexact only for r in (-1, 0))
        dr = (r64.astype(F32) - r0.astype(F32)).astype(np.float64)
        dc = (c64.astype(F32) - c0.astype(F32)).astype(np.float64)
    else:
        dr = r64 - r0
        dc = c64 - c0
    img64 = img.astype(np.float64)

    def tap(ri, ci):
        ri = ri.astype(np.int64)
        ci = ci.astype(np.int64)
        ok = (ri >= 0) & (ri < rows) & (ci >= 0) & (ci < cols)
        out = np.zeros(ri.shape, dtype=np.float64)
        out[ok] = img64[ri[ok], ci[ok]]
        return out

    tl = tap(r0, c0)
    tr = tap(r0, c1)
    bl = tap(r1, c0)
    br = tap(r1, c1)
    if fp32:
        # scikit-image's compiled blend forms dc * right-neighbour in float32
        # and everything else in double (pinned bit-exact against
        # scikit-image 0.18.3 outputs, tests/golden/skimage_golden.npz)
        tr_w = (dc.astype(F32) * tr.astype(F32)).astype(np.float64)
        br_w = (dc.astype(F32) * br.astype(F32)).astype(np.float64)
    else:
        tr_w = dc * tr
        br_w = dc * br
    top = (1 - dc) * tl + tr_w
    bot = (1 - dc) * bl + br_w
    out = (1 - dr) * top + dr * bot
    return out.astype(F32) if fp32 else out


def warp_affine(img: np.ndarray, m: np.ndarray, out_shape=None, fp32: bool = True) -> np.ndarray:
    """``warp(img, M, order=1, clip=False)`` for a 3x3 affine inverse map M.

    Output pixel (r, c) samples the input at ``x = M00 c + M01 r + M02``,
    ``y = M10 c + M11 r + M12`` (x = column, y = row). With ``fp32`` the
    matrix is cast to float32 (``skimage:_warps.py:874``) and the map is
    evaluated in float32 as ``(M0 x + M1 y) + M2``.
    """
    rows, cols = img.shape if out_shape is None else out_shape
    rr, cc = np.mgrid[:rows, :cols]
    if fp32:
        m = np.asarray(m, dtype=np.float64).astype(F32)
        rr = rr.astype(F32)
        cc = cc.astype(F32)
        x = (m[0, 0] * cc + m[0, 1] * rr) + m[0, 2]
        y = (m[1, 0] * cc + m[1, 1] * rr) + m[1, 2]
    else:
        m = np.asarray(m, dtype=np.float64)
        x = (m[0, 0] * cc + m[0, 1] * rr) + m[0, 2]
        y = (m[1, 0] * cc + m[1, 1] * rr) + m[1, 2]
    return bilinear_sample(img, y, x, fp32=fp32)


# ---------------------------------------------------------------------------
# radon (one angle) -- skimage radon_transform.py:24-115, circle=True, square
# ---------------------------------------------------------------------------


def radon_matrix(theta_deg, n: int, fp32: bool = True) -> np.ndarray:
    """The inverse map ``R`` of ``radon`` for one angle in degrees.

    ``c = n // 2``; ``R = [[cos, sin, -c (cos + sin - 1)], [-sin, cos,
    -c (cos - sin - 1)]]`` with the angle ``deg2rad(theta)`` (float32 for a
    float32 theta, as corrct passes ``rad2deg`` of float32 angles,
    ``_projector_backends.py:148,299``).
    """
    c = n // 2
    if fp32:
        ang = deg2rad_f32(theta_deg)
        ca = np.cos(ang)
        sa = np.sin(ang)
        m02 = F32(-c) * ((ca + sa) - F32(1))
        m12 = F32(-c) * ((ca - sa) - F32(1))
        return np.array([[ca, sa, m02], [-sa, ca, m12], [0, 0, 1]], dtype=F32)
    ang = float(theta_deg) * (math.pi / 180.0)
    ca = math.cos(ang)
    sa = math.sin(ang)
    return np.array(
        [[ca, sa, -c * (ca + sa - 1)], [-sa, ca, -c * (ca - sa - 1)], [0, 0, 1]], dtype=np.float64
    )


def radon_line(img: np.ndarray, theta_deg, fp32: bool = True, matrix: np.ndarray | None = None) -> np.ndarray:
    """``np.squeeze(skt.radon(img, [theta_deg]))`` for a square image.

    The warped image is summed over rows (axis 0); numpy accumulates an
    axis-0 reduction of a C-ordered array row after row, so the float32 sum
    is sequential in the row index.
    """
    n = img.shape[0]
    m = radon_matrix(theta_deg, n, fp32) if matrix is None else matrix
    rot = warp_affine(img.astype(F32) if fp32 else img.astype(np.float64), m, fp32=fp32)
    acc = np.zeros(n, dtype=F32 if fp32 else np.float64)
    for row in rot:
        acc += row
    return acc


# ---------------------------------------------------------------------------
# iradon without filter (one angle) -- radon_transform.py:186-314
# ---------------------------------------------------------------------------


def _unfiltered_column(col: np.ndarray) -> np.ndarray:
    """The column after iradon's all-ones-filter FFT round trip.

    ``radon_transform.py:278-292``: zero-pad to ``max(64, 2^ceil(log2(2n)))``,
    single-precision ``scipy.fft.fft`` of the float32 column, times the
    float64 ones filter, double ``ifft``, real part of the first n entries.
    """
    import scipy.fft as sfft

    n = col.shape[0]
    size = max(64, int(2 ** np.ceil(np.log2(2 * n))))
    img = np.zeros((size, 1), dtype=F32)
    img[:n, 0] = col
    proj = sfft.fft(img, axis=0) * np.ones((size, 1))
    return np.real(sfft.ifft(proj, axis=0)[:n, 0])


def iradon_line(col: np.ndarray, theta_deg, n: int, fp32: bool = True) -> np.ndarray:
    """``skt.iradon(col[:, None], [theta], circle=False, output_size=n, filter_name=None)``.

    ``t = (j - c) cos - (i - c) sin`` (float64 from float32 cos/sin),
    ``np.interp(t, arange(n) - c, col, left=0, right=0)``, accumulated into a
    float32 image, then ``* pi / 2``. In fp32 mode the column first goes
    through iradon's FFT round trip (identity up to ~1e-7 of its norm).
    """
    c = n // 2
    xpr, ypr = np.mgrid[:n, :n] - c
    xp = np.arange(col.shape[0]) - col.shape[0] // 2
    if fp32:
        ang = deg2rad_f32(theta_deg)
        t = ypr * np.float64(np.cos(ang)) - xpr * np.float64(np.sin(ang))
        v = np.interp(t, xp, _unfiltered_column(np.asarray(col, dtype=F32)), left=0, right=0)
        rec = np.zeros((n, n), dtype=F32)
        rec += v
        return (rec * F32(np.pi)) / F32(2)
    ang = float(theta_deg) * (math.pi / 180.0)
    t = ypr * math.cos(ang) - xpr * math.sin(ang)
    v = np.interp(t, xp, col.astype(np.float64), left=0, right=0)
    return v * (math.pi / 2)


# ---------------------------------------------------------------------------
# rotate(order=1, clip=False) -- _warps.py:296-406
# ---------------------------------------------------------------------------


def rotate_matrix(img_shape, angle_deg) -> np.ndarray:
    """float64 inverse map of ``skt.rotate(img, angle_deg)``.

    centre ``(cols/2 - 0.5, rows/2 - 0.5)``; ``tform = T(-c) + R(r) + T(c)``
    with ``A + B = B.params @ A.params`` and ``R(r) = [[cos, -sin], [sin,
    cos]]`` on (x, y) = (col, row), ``r = deg2rad(angle_deg)`` (float32 when
    the angle is float32), ``cos``/``sin`` evaluated in double (``math``).
    """
    rows, cols = img_shape
    cx = cols / 2.0 - 0.5
    cy = rows / 2.0 - 0.5
    if isinstance(angle_deg, (np.float32,)) or (isinstance(angle_deg, np.ndarray) and angle_deg.dtype == F32):
        r = float(deg2rad_f32(angle_deg))
    else:
        r = float(angle_deg) * (math.pi / 180.0)
    t1 = np.array([[1.0, 0.0, cx], [0.0, 1.0, cy], [0.0, 0.0, 1.0]])
    rot = np.array([[math.cos(r), -math.sin(r), 0.0], [math.sin(r), math.cos(r), 0.0], [0.0, 0.0, 1.0]])
    t3 = np.array([[1.0, 0.0, -cx], [0.0, 1.0, -cy], [0.0, 0.0, 1.0]])
    m = t1 @ (rot @ t3)
    m[2] = (0, 0, 1)
    return m


def rotate(img: np.ndarray, angle_deg, fp32: bool = True) -> np.ndarray:
    """``skt.rotate(img, angle_deg, order=1, clip=False)``."""
    m = rotate_matrix(img.shape, angle_deg)
    if not fp32:
        # FP64 evaluation: the exact rotation by the real angle
        a = float(angle_deg) * (math.pi / 180.0)
        rows, cols = img.shape
        cx = cols / 2.0 - 0.5
        cy = rows / 2.0 - 0.5
        ca, sa = math.cos(a), math.sin(a)
        m = np.array(
            [[ca, -sa, cx - ca * cx + sa * cy], [sa, ca, cy - sa * cx - ca * cy], [0, 0, 1]], dtype=np.float64
        )
        return warp_affine(img.astype(np.float64), m, fp32=False)
    return warp_affine(img.astype(F32), m, fp32=True)
