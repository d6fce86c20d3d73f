"""CPU restatement of corrct's attenuation maps.

Test infrastructure only (see ``oracle/__init__.py``).

* ``compute_attenuation``  -- ``_projector_backends.py:43-99``
* ``compute_maps``         -- ``physics/attenuation.py:105-180``, with the
  outgoing factor of each detector averaged (float32, in order) over its
  sub-angle offsets (``attenuation.py:110-119``; one offset 0 for a point
  detector)
* ``bin_maps``             -- ``get_maps(binning=b)`` (``physics/attenuation.py:294-295``) over
  ``bin_imgs(auto_crop=True)`` (``processing/pre.py:180-220``): crop each volume
  axis to a multiple of b (``excess - excess // 2`` from the start), then block
  means (scikit-image ``block_reduce`` with ``np.mean``).
* ``detector_sub_angles``  -- the offsets: midpoints of S equal parts of
  ``[-atan2(sqrt(surface) / 2, distance), +...]`` (``physics/xrf.py:238-246``,
  ``attenuation.py:112-113``)
"""

from __future__ import annotations

import math

import numpy as np

from .skimage_ops import F32, rad2deg_f32, rotate


def pad_edges(n: int) -> int:
    """Pad width ``e`` of ``compute_attenuation`` for an n x n slice.

    ``min_size = ceil(sqrt(n^2 + n^2))``, ``e = ceil((min_size - n) / 2)``
    (``_projector_backends.py:75-77``).
    """
    min_size = math.ceil(math.sqrt(n * n + n * n))
    return int(math.ceil((min_size - n) / 2))


def _deg(angle_rad, fp32: bool):
    if fp32 and isinstance(angle_rad, np.float32):
        return rad2deg_f32(angle_rad)[()]
    return np.float64(angle_rad) * (180.0 / np.pi)


def compute_attenuation(mu: np.ndarray, angle_rad, invert: bool = False, fp32: bool = True) -> np.ndarray:
    """exp(-cumulative attenuation) along the beam at ``angle_rad``.

    pad by ``e`` -> ``rotate(-deg)`` -> ``v += roll(v, 1, rows)`` ->
    ``cumsum(v / 2, rows)`` -> ``rotate(+deg)`` -> crop -> ``exp(-.)``.
    ``invert`` adds pi to the angle (outgoing photons). A float32 angle stays
    float32 through ``rad2deg`` (in-beam angles); ``invert`` angles are
    float64 in the reference (``attenuation.py:109``).
    """
    n = mu.shape[-1]
    e = pad_edges(n)
    if invert:
        angle_rad = np.float64(angle_rad) + np.pi
    deg = _deg(angle_rad, fp32)
    dt = F32 if fp32 else np.float64
    vol = np.pad(mu.astype(dt), e, mode="constant")
    vol = rotate(vol, -deg, fp32=fp32)
    vol = vol + np.roll(vol, 1, axis=-2)
    vol = np.cumsum(vol / dt(2), axis=-2, dtype=dt)
    vol = rotate(vol, deg, fp32=fp32)
    vol = vol[e:-e, e:-e]
    return np.exp(-vol).astype(dt)


def compute_maps(
    mu_in: np.ndarray | None,
    mu_out: np.ndarray | None,
    angles_rot_rad_f32: np.ndarray,
    angles_det_rad: np.ndarray,
    fp32: bool = True,
    sub_angles=None,
) -> np.ndarray:
    """``maps[a, d] = E_in(theta_a) * E_out(theta_a + phi_d + pi)``, shape [A, D, N, N].

    ``sub_angles`` (one 1-D array of offsets per detector, None = a single 0)
    replaces E_out by the mean over the offsets s of E_out(theta_a + phi_d + s + pi).

    ``theta_a`` are float32 (the projector's ``angles_w_rad``); the outgoing
    angle ``theta_a + phi_d`` is float64 (``attenuation.py:109``). A missing
    mu contributes a factor 1.
    """
    angles = np.asarray(angles_rot_rad_f32, dtype=F32)
    dets = np.asarray(angles_det_rad, dtype=np.float64).reshape(-1)
    ref = mu_in if mu_in is not None else mu_out
    if ref is None:
        raise ValueError("No attenuation volumes were given.")
    n = ref.shape[-1]
    dt = F32 if fp32 else np.float64
    maps = np.ones((len(angles), len(dets), n, n), dtype=dt)
    for ia, a in enumerate(angles):
        if mu_in is not None:
            maps[ia] *= compute_attenuation(mu_in, a if fp32 else np.float64(a), invert=False, fp32=fp32)[None]
        if mu_out is not None:
            for idet, phi in enumerate(dets):
                subs = np.zeros(1) if sub_angles is None else np.asarray(sub_angles[idet], dtype=np.float64)
                acc = np.zeros((n, n), dtype=dt)
                for s in subs:
                    acc += compute_attenuation(mu_out, (np.float64(a) + phi) + s, invert=True, fp32=fp32)
                acc /= len(subs)
                maps[ia, idet] *= acc
    return maps


def detector_sub_angles(surface_mm2: float, distance_mm: float, n_sub: int) -> np.ndarray:
    """Sub-angle offsets of a detector of the given surface and distance."""
    if surface_mm2 <= 0.0:
        return np.zeros(1)
    half = math.atan2(math.sqrt(surface_mm2) / 2, distance_mm)
    edges = np.linspace(-half, half, n_sub + 1)
    return np.convolve(edges, np.ones(2), "valid") / 2


def bin_maps(maps: np.ndarray, binning: int, nd: int) -> np.ndarray:
    """Block means over the last ``nd`` axes after the auto-crop (float64 sums, rounded to float32)."""
    maps = np.asarray(maps)
    sl = [slice(None)] * maps.ndim
    for ax in range(maps.ndim - nd, maps.ndim):
        excess = maps.shape[ax] % binning
        sl[ax] = slice(excess - excess // 2, maps.shape[ax] - excess // 2)
    m = maps[tuple(sl)].astype(np.float64)
    lead = m.shape[: m.ndim - nd]
    shp = list(lead)
    for d in m.shape[m.ndim - nd:]:
        shp += [d // binning, binning]
    m = m.reshape(shp)
    axes = tuple(len(lead) + 2 * k + 1 for k in range(nd))
    return m.mean(axis=axes).astype(np.float32)
