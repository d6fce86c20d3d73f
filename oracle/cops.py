"""NumPy-facing wrappers of the oracle's C restatement (``oracle/c``).

Test infrastructure only (see ``oracle/__init__.py``): the tests use these
at the configurations' full sizes, where the NumPy oracle would take minutes,
and ``bench.py``'s CPU baseline times them. Same layouts and semantics as
``oracle.xrf``: volume [N, N], maps [A, D, N, N], sinogram [D, A, N].

* ``fp``          -- ``oracle.xrf.fp`` (``projectors.py:502-531, 570-596``)
* ``bp_skimage``  -- ``oracle.xrf.bp_skimage`` non-symmetric (``projectors.py:619-622``)
* ``bp_adjoint``  -- ``oracle.xrf.bp_adjoint``: the exact transpose of ``fp``
"""

from __future__ import annotations

import numpy as np

from . import build_c
from . import skimage_ops as sk

_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = build_c.load()
    return _lib


def tables(angles_rad_f32, n: int):
    """(geom [A, 4] = (cos, sin, m02, m12) of skimage's radon map, trig [A, 2] = (cos, sin)), float32."""
    ang = np.asarray(angles_rad_f32, dtype=np.float32)
    degs = sk.rad2deg_f32(ang)
    geom = np.stack([sk.radon_matrix(d, n)[[0, 0, 0, 1], [0, 1, 2, 2]] for d in degs]).astype(np.float32)
    th = sk.deg2rad_f32(degs)
    trig = np.stack([np.cos(th), np.sin(th)], axis=1).astype(np.float32)
    return np.ascontiguousarray(geom), np.ascontiguousarray(trig)


def _w(w_det, w_ang, a: int, d: int):
    wd = np.ones(d) if w_det is None else np.broadcast_to(np.asarray(w_det, np.float64), (d,))
    wa = np.ones((a, d)) if w_ang is None else np.asarray(w_ang, np.float64).reshape(a, d)
    return np.ascontiguousarray((wd[None, :] * wa).astype(np.float32))


def fp(vol, angles_rad_f32, maps=None, w_det=None, w_ang=None, n_det: int = 1, tabs=None) -> np.ndarray:
    vol = np.ascontiguousarray(vol, dtype=np.float32)
    n = vol.shape[-1]
    a = len(angles_rad_f32)
    d = maps.shape[1] if maps is not None else n_det
    geom, _ = tabs if tabs is not None else tables(angles_rad_f32, n)
    mp = None if maps is None else np.ascontiguousarray(maps, dtype=np.float32)
    w = _w(w_det, w_ang, a, d)
    out = np.zeros((d, a, n), np.float32)
    lib().oracle_fp(vol.ctypes.data, None if mp is None else mp.ctypes.data, geom.ctypes.data, w.ctypes.data, n, a, d,
                    out.ctypes.data)
    return out


def bp_skimage(sino, angles_rad_f32, n: int, tabs=None) -> np.ndarray:
    a = len(angles_rad_f32)
    y = np.ascontiguousarray(sino, dtype=np.float32).reshape(-1, a, n)
    _, trig = tabs if tabs is not None else tables(angles_rad_f32, n)
    out = np.zeros((n, n), np.float32)
    lib().oracle_bp_skimage(y.ctypes.data, trig.ctypes.data, n, a, y.shape[0], out.ctypes.data)
    return out


def bp_adjoint(sino, angles_rad_f32, n: int, maps=None, w_det=None, w_ang=None, tabs=None) -> np.ndarray:
    a = len(angles_rad_f32)
    y = np.ascontiguousarray(sino, dtype=np.float32).reshape(-1, a, n)
    d = y.shape[0]
    geom, _ = tabs if tabs is not None else tables(angles_rad_f32, n)
    mp = None if maps is None else np.ascontiguousarray(maps, dtype=np.float32)
    w = _w(w_det, w_ang, a, d)
    out = np.zeros((n, n), np.float32)
    lib().oracle_bp_adjoint(y.ctypes.data, None if mp is None else mp.ctypes.data, geom.ctypes.data, w.ctypes.data, n,
                            a, d, out.ctypes.data)
    return out
