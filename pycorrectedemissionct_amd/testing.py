"""Synthetic XRF workloads: seeded phantoms, attenuation and noise.

* ``phantom_ellipses``  -- piecewise-constant ellipses (seeded centres,
  semi-axes, rotations, intensities) kept inside the inscribed circle so the
  radon transform neither warns nor clips, plus optional Gaussian blobs.
* ``xrf_attenuation``   -- ``mu_in = a_in (f > 0) + a_bg``, ``mu_out = a_out (f > 0)``.
* ``add_noise``         -- corrct ``testing.add_noise`` (``testing.py:266-320``):
  scale by the photon count, optional Poisson draw, optional readout noise.
* ``workload_*``        -- the standard synthetic configurations used by the
  tests and the benchmark (sizes in the function names).
"""

from __future__ import annotations

import numpy as np


def _ellipse_mask(n: int, cy, cx, ay, ax, rot):
    yy, xx = np.mgrid[:n, :n].astype(np.float64)
    y = yy - cy
    x = xx - cx
    c, s = np.cos(rot), np.sin(rot)
    u = c * x + s * y
    v = -s * x + c * y
    return (u / ax) ** 2 + (v / ay) ** 2 <= 1.0


def phantom_ellipses(n: int, n_ellipses: int, rng: np.random.Generator, semi_axes=(5, 40), intensities=(0.2, 1.0),
                     n_blobs: int = 0, blob_sigma=(3, 15)) -> np.ndarray:
    """Seeded float32 phantom of ``n_ellipses`` ellipses (+ ``n_blobs`` Gaussian blobs) on zero."""
    img = np.zeros((n, n), dtype=np.float64)
    c = n // 2
    r_max = n // 2 - 1
    lo, hi = semi_axes
    hi = min(hi, r_max)
    lo = min(lo, hi)
    placed = 0
    tries = 0
    while placed < n_ellipses and tries < 100 * n_ellipses:
        tries += 1
        ay, ax = rng.uniform(lo, hi, size=2)
        rot = rng.uniform(0, np.pi)
        r = max(ay, ax)
        dist = rng.uniform(0, max(r_max - r, 0))
        ang = rng.uniform(0, 2 * np.pi)
        cy, cx = c + dist * np.sin(ang), c + dist * np.cos(ang)
        if np.hypot(cy - c, cx - c) + r > r_max:
            continue
        img[_ellipse_mask(n, cy, cx, ay, ax, rot)] = rng.uniform(*intensities)
        placed += 1
    if n_blobs:
        yy, xx = np.mgrid[:n, :n].astype(np.float64)
        for _ in range(n_blobs):
            sg = rng.uniform(*blob_sigma)
            dist = rng.uniform(0, max(r_max - 3 * sg, 0))
            ang = rng.uniform(0, 2 * np.pi)
            cy, cx = c + dist * np.sin(ang), c + dist * np.cos(ang)
            r2 = (yy - cy) ** 2 + (xx - cx) ** 2
            # truncated at 3 sigma, so the blob's support (and its attenuation) stays local
            img += np.where(r2 <= 9 * sg**2, rng.uniform(*intensities) * np.exp(-r2 / (2 * sg**2)), 0.0)
    yy, xx = np.mgrid[:n, :n] - c
    img[(yy**2 + xx**2) > (n // 2) ** 2] = 0
    return img.astype(np.float32)


def xrf_attenuation(f: np.ndarray, a_in: float = 0.02, a_bg: float = 0.005, a_out: float = 0.05):
    support = (f > 0).astype(np.float32)
    mu_in = (a_in * support + a_bg).astype(np.float32)
    mu_out = (a_out * support).astype(np.float32)
    return mu_in, mu_out


def add_noise(img_clean, num_photons, add_poisson: bool = False, readout_noise_std=None, background_avg=None,
              background_std=None, detection_efficiency: float = 1.0, dtype=np.float32, rng=None):
    """(noisy, clean, background) scaled by photons x efficiency (``testing.py:266-320``)."""
    rng = np.random.default_rng() if rng is None else rng
    img_clean = num_photons * detection_efficiency * np.asarray(img_clean).copy().astype(dtype)
    img_noise = img_clean.copy()
    if background_avg is not None:
        if background_std is None:
            background_std = background_avg * 5e-2
        img_noise += num_photons * detection_efficiency * np.abs(rng.normal(background_avg, background_std, img_clean.shape))
    background = float(np.mean((img_noise - img_clean).flatten()))
    if add_poisson:
        img_noise = rng.poisson(img_noise).astype(dtype)
    if readout_noise_std is not None:
        img_noise += rng.normal(0, readout_noise_std, img_clean.shape)
    return img_noise, img_clean, background


def angles_uniform(n_angles: int, span: float = 2 * np.pi) -> np.ndarray:
    return np.linspace(0, span, n_angles, endpoint=False).astype(np.float32)


def workload_sirt_128(seed: int = 20260101):
    """128 x 128, 10 ellipses, 256 angles over 2 pi, one detector at pi/2."""
    rng = np.random.default_rng(seed)
    f = phantom_ellipses(128, 10, rng)
    mu_in, mu_out = xrf_attenuation(f)
    return dict(f=f, mu_in=mu_in, mu_out=mu_out, angles=angles_uniform(256), det=np.array([np.pi / 2]),
                peak=1e4, rng=rng)


def workload_pdhg_512(seed: int = 20260102, n: int = 512, n_angles: int = 720):
    """512 x 512, 30 ellipses + 5 blobs, 720 angles over 2 pi, detectors at +-pi/2, Poisson peak 1e3."""
    rng = np.random.default_rng(seed)
    f = phantom_ellipses(n, 30, rng, n_blobs=5)
    mu_in, mu_out = xrf_attenuation(f)
    return dict(f=f, mu_in=mu_in, mu_out=mu_out, angles=angles_uniform(n_angles),
                det=np.array([np.pi / 2, -np.pi / 2]), peak=1e3, rng=rng)


def poisson_data(sino: np.ndarray, peak: float, rng: np.random.Generator) -> np.ndarray:
    """Counts drawn at ``peak`` counts for the sinogram maximum, rescaled to the sinogram's units."""
    scale = peak / float(np.max(sino))
    return (rng.poisson(np.clip(sino, 0, None) * scale) / scale).astype(np.float32)
