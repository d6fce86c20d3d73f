"""CPU restatement of corrct's FBP filtering (test infrastructure only).

* ``padding_size``   -- ``Filter.get_padding_size`` (``filters.py:284-298``):
  max(64, 2**ceil(log2(2 N))).
* ``fourier_filter`` -- the Fourier-domain filter scikit-image builds for
  ``FilterFBP.compute_filter`` (``filters.py:503-517``): the ramp from the
  FFT of its band-limited spatial form (Kak & Slaney, Principles of
  Computerized Tomographic Imaging, ch. 3 eq. 61: h[0] = 1/4, h[odd k] =
  -1/(pi k)^2 with k the circular distance, h[even] = 0), times the window
  (shepp-logan sinc, cosine, hamming, hann).
* ``apply_filter``   -- ``Filter.apply_filter`` (``filters.py:346-390``) with
  zero ("constant") padding: the line is padded to the FFT size with zeros
  after its end, multiplied by the half-spectrum filter, transformed back and
  cut to N samples -- a circular convolution of length P.
* ``fbp``            -- ``FBP.__call__`` (``solvers.py:396-470``): filter with
  pi / (2 A) folded into the filter, then back-project.

Pinned by ``tests/test_oracle.py`` against scikit-image's own filters and
``iradon`` (``tests/golden/skimage_fbp_golden.npz``).
"""

from __future__ import annotations

import numpy as np

NAMES = ("ramp", "shepp-logan", "cosine", "hamming", "hann")


def padding_size(n: int) -> int:
    return max(64, int(2 ** np.ceil(np.log2(2 * n))))


def fourier_filter(size: int, name: str) -> np.ndarray:
    """Full-length (size) real Fourier filter, float64."""
    k = np.arange(size)
    dist = np.minimum(k, size - k)  # circular distance to sample 0
    h = np.zeros(size)
    h[0] = 0.25
    odd = dist % 2 == 1
    h[odd] = -1.0 / (np.pi * dist[odd]) ** 2
    filt = 2 * np.real(np.fft.fft(h))
    if name == "ramp":
        pass
    elif name == "shepp-logan":
        omega = np.pi * np.fft.fftfreq(size)[1:]
        filt[1:] *= np.sin(omega) / omega
    elif name == "cosine":
        filt *= np.fft.fftshift(np.sin(np.linspace(0, np.pi, size, endpoint=False)))
    elif name == "hamming":
        filt *= np.fft.fftshift(np.hamming(size))
    elif name == "hann":
        filt *= np.fft.fftshift(np.hanning(size))
    else:
        raise ValueError(f"Unknown filter {name}")
    return filt


def apply_filter(data, filt_half: np.ndarray) -> np.ndarray:
    """Filter the last axis of data with the half spectrum ``filt_half`` (length P // 2 + 1)."""
    data = np.asarray(data, dtype=np.float64)
    n = data.shape[-1]
    p = 2 * (len(filt_half) - 1)
    pad = np.zeros((*data.shape[:-1], p))
    pad[..., :n] = data
    return np.fft.irfft(np.fft.rfft(pad, axis=-1) * filt_half, n=p, axis=-1)[..., :n]


def fbp_filter_half(n: int, n_angles: int, name: str = "ramp") -> np.ndarray:
    """``FilterFBP.compute_filter``: half spectrum of the padded-size filter times pi / (2 A)."""
    p = padding_size(n)
    return fourier_filter(p, name)[: p // 2 + 1] * np.pi / (2 * n_angles)


def fbp(At, sino, n_angles: int, name: str = "ramp"):
    """Filtered back-projection with the back-projector ``At``."""
    sino = np.asarray(sino)
    return At(apply_filter(sino, fbp_filter_half(sino.shape[-1], n_angles, name)))
