"""FBP filters (corrct ``filters.py`` subset: ``Filter``, ``FilterFBP``, ``FilterCustom``).

The Fourier-domain filter is built on the host once per data shape, as corrct
builds it (``FilterFBP.compute_filter``, ``filters.py:503-517``). Its
half spectrum is multiplied into the zero-padded line spectrum
(``Filter.apply_filter``, ``filters.py:346-390``). That product equals a
circular convolution of length P with the real-space kernel h = irfft(filter).
The device computes that convolution directly with float64 sums
(``chip_filter_lines``). Filters:

* ``ramp`` -- the band-limited ramp of Kak & Slaney, Principles of
  Computerized Tomographic Imaging, ch. 3 eq. 61: the FFT of h[0] = 1/4,
  h[k] = -1/(pi k)^2 for odd circular distance k, 0 for even.
* ``shepp-logan``, ``cosine``, ``hamming``, ``hann`` -- the ramp times a sinc,
  sine, Hamming or Hann window, as scikit-image arranges them, of which corrct
  keeps the first P/2 + 1 samples for its real FFT.

Only zero ("constant") padding is supported. The data-driven ``FilterMR``
(``filters.py:530-700``) is not built.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _arrays, _native


def padding_size(n: int) -> int:
    """``Filter.get_padding_size`` (``filters.py:284-298``)."""
    return max(64, int(2 ** np.ceil(np.log2(2 * n))))


def _ramp_window(size: int, name: str) -> np.ndarray:
    k = np.arange(size)
    dist = np.minimum(k, size - k)
    h = np.zeros(size)
    h[0] = 0.25
    odd = dist % 2 == 1
    h[odd] = -1.0 / (np.pi * dist[odd]) ** 2
    filt = 2 * np.real(np.fft.fft(h))
    if name == "shepp-logan":
        omega = np.pi * np.fft.fftfreq(size)[1:]
        filt[1:] *= np.sin(omega) / omega
    elif name == "cosine":
        filt *= np.fft.fftshift(np.sin(np.linspace(0, np.pi, size, endpoint=False)))
    elif name == "hamming":
        filt *= np.fft.fftshift(np.hamming(size))
    elif name == "hann":
        filt *= np.fft.fftshift(np.hanning(size))
    return filt


class Filter:
    """Base FBP filter: holds the half spectrum (``filters.py:247-420``)."""

    def __init__(self, fbp_filter=None, pad_mode: str = "constant", use_rfft: bool = True, dtype=np.float32):
        if pad_mode.lower() != "constant":
            raise NotImplementedError("only zero ('constant') padding is supported by the device filter")
        if not use_rfft:
            raise NotImplementedError("only the real-FFT form of the filters is supported")
        self.pad_mode = pad_mode.lower()
        self.use_rfft = True
        self.dtype = dtype
        self.fbp_filter = np.array([1.0]) if fbp_filter is None else np.array(np.real(fbp_filter), dtype=np.float64)

    def get_padding_size(self, data_wu_shape) -> int:
        return padding_size(int(data_wu_shape[-1]))

    def compute_filter(self, data_wu) -> None:  # pragma: no cover - overridden
        pass

    def kernel(self, n: int) -> np.ndarray:
        """Real-space circular kernel h (length P) of the current half-spectrum filter."""
        p = padding_size(n)
        f = np.asarray(self.fbp_filter, dtype=np.float64).reshape(-1)
        if f.size != p // 2 + 1:
            raise ValueError(f"filter has {f.size} samples; padded size {p} needs {p // 2 + 1}")
        return np.fft.irfft(f, n=p)

    def apply_filter(self, data_wu, fbp_filter=None):
        """Filter the last axis of data_wu (device or host array; returns the same kind)."""
        if fbp_filter is not None:
            self.fbp_filter = np.array(np.real(fbp_filter), dtype=np.float64)
        t, host = _arrays.to_dev(data_wu)
        t = t.contiguous()
        n = int(t.shape[-1])
        h = torch.as_tensor(self.kernel(n), dtype=torch.float64, device=t.device)
        out = torch.empty_like(t)
        _native.call("chip_filter_lines", t.data_ptr(), out.data_ptr(), h.data_ptr(), t.numel() // n, n,
                     padding_size(n), _native.stream())
        return _arrays.from_dev(out, host)

    def __call__(self, data_wu):
        self.compute_filter(data_wu)
        return self.apply_filter(data_wu)


class FilterCustom(Filter):
    """A user half-spectrum filter (``filters.py:446-472``)."""

    def __init__(self, fbp_filter, pad_mode: str = "constant", use_rfft: bool = True, dtype=np.float32):
        super().__init__(fbp_filter=fbp_filter, pad_mode=pad_mode, use_rfft=use_rfft, dtype=dtype)


class FilterFBP(Filter):
    """Traditional FBP filter (``filters.py:475-527``)."""

    FILTERS = ("ramp", "shepp-logan", "cosine", "hamming", "hann")

    def __init__(self, filter_name: str = "ramp", pad_mode: str = "constant", use_rfft: bool = True,
                 dtype=np.float32):
        if filter_name.lower() not in self.FILTERS:
            raise ValueError(f"Unknown filter {filter_name}. Available filters: {self.FILTERS}")
        super().__init__(fbp_filter=None, pad_mode=pad_mode, use_rfft=use_rfft, dtype=dtype)
        self.filter_name = filter_name.lower()

    def compute_filter(self, data_wu) -> None:
        shape = np.shape(data_wu)
        p = padding_size(int(shape[-1]))
        full = _ramp_window(p, self.filter_name) * np.pi / (2 * shape[-2])
        self.fbp_filter = full[: p // 2 + 1]

    def get_available_filters(self):
        return self.FILTERS
