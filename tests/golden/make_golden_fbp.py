"""Write golden vectors for corrct's FBP (``solvers.py:346-470``, ``filters.py:247-527``) from scikit-image.

Run with an interpreter that has scikit-image 0.18.3 (this container:
``/opt/conda/bin/python3.9 tests/golden/make_golden_fbp.py``); corrct itself is not imported.

* ``filter_<name>_<size>``: scikit-image's Fourier-domain FBP filter of that size
  (``skimage.transform.radon_transform._get_fourier_filter``, which corrct's
  ``FilterFBP.compute_filter`` calls at ``filters.py:511``).
* ``fbp_<name>_vol``: ``skimage.transform.iradon(sino.T, theta, filter_name=name,
  circle=False, output_size=N)`` of a seeded sinogram. corrct's FBP on the
  skimage backend computes the same quantity: it pads each line with zeros to
  max(64, 2**ceil(log2(2 N))) samples (``filters.py:284-298, 346-390``), filters with
  the same Fourier filter times pi / (2 A) (``filters.py:511-512``), and
  back-projects with iradon(filter None) times 2 / pi per angle
  (``_projector_backends.py:350-359``); iradon with a filter multiplies by
  pi / (2 A) once at the end.

Output: ``tests/golden/skimage_fbp_golden.npz`` (plain arrays).
"""

import os
import warnings

import numpy as np
import skimage
import skimage.transform as skt
from skimage.transform.radon_transform import _get_fourier_filter

warnings.filterwarnings("ignore")

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "skimage_fbp_golden.npz")
NAMES = ("ramp", "shepp-logan", "cosine", "hamming", "hann")


def main():
    rng = np.random.default_rng(20260110)
    out = {"skimage_version": np.array(skimage.__version__)}
    for size in (64, 128):
        for name in NAMES:
            out[f"filter_{name}_{size}"] = np.squeeze(_get_fourier_filter(size, name))
    n, a = 24, 30
    rad = np.linspace(0, np.pi, a, endpoint=False).astype(np.float32)
    deg = np.rad2deg(rad)
    sino = rng.random((a, n)).astype(np.float32)
    out["fbp_rad"] = rad
    out["fbp_sino"] = sino
    for name in ("ramp", "hann"):
        out[f"fbp_{name}_vol"] = skt.iradon(sino.T, deg, filter_name=name, circle=False, output_size=n)
    np.savez_compressed(OUT, **out)
    print("wrote", OUT, "with", len(out), "arrays; skimage", skimage.__version__)


if __name__ == "__main__":
    main()
