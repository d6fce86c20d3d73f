"""Write golden vectors for the skimage-backend semantics, from scikit-image itself.

Run with an interpreter that has scikit-image 0.18.3 (this container:
``/opt/conda/bin/python3.9 tests/golden/make_golden_skimage.py``). corrct
itself is NOT imported; the few lines that compose scikit-image calls the
way corrct's skimage backend does are restated here, each citing the corrct
line it follows:

* fp line      = ``np.squeeze(skt.radon(vol, [deg]))``          (``_projector_backends.py:333``)
* bp           = ``sum_a skt.iradon(prj[a, :, None], [deg_a], circle=False,
                  output_size=N, filter_name=None) * 2 / pi``     (``_projector_backends.py:350-359``)
* attenuation  = pad -> rotate(-deg) -> v += roll(v, 1, -2) -> cumsum(v / 2, -2)
                  -> rotate(+deg) -> crop -> exp(-.)              (``_projector_backends.py:63-97``)
* maps         = ones *= E_in(theta); *= E_out(theta + phi_d + pi) (``physics/attenuation.py:105-161``)
* XRF fp       = w_d * radon(vol * M[a, d])                       (``projectors.py:519-525``)
* XRF symmetric bp = sum_a sum_d (w_d * bp_a(sino[d, a])) * M[a, d], bp_a the one-angle
                  ``iradon(line[:, None], [deg_a], ...) * 2 / pi``   (``projectors.py:556-565, 612-616``,
                                                                     ``_projector_backends.py:357-359``)

Angles follow corrct's dtypes: rotation angles are float32 radians converted
with ``np.rad2deg`` (``_projector_backends.py:148, 299``); outgoing-beam
angles are float64 (``physics/attenuation.py:109``).

Output: ``tests/golden/skimage_golden.npz`` (plain arrays, no pickles).
"""

import os
import warnings

import numpy as np
import skimage
import skimage.transform as skt

warnings.filterwarnings("ignore")

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "skimage_golden.npz")


def phantom(rng, n):
    """Random non-negative image that is zero outside the inscribed circle."""
    img = rng.random((n, n)).astype(np.float32)
    yy, xx = np.mgrid[:n, :n] - n // 2
    img[(yy**2 + xx**2) > (n // 2) ** 2] = 0
    return img


def corrct_fp_line(vol, deg):
    return np.squeeze(skt.radon(vol, np.array([deg], dtype=np.float32)))


def corrct_bp(prj, degs, n):
    vol = np.empty([len(degs), n, n], dtype=prj.dtype)
    for ii, a in enumerate(degs):
        vol[ii] = skt.iradon(prj[ii, :, np.newaxis], [a], circle=False, output_size=n, filter_name=None)
    vol = vol.sum(axis=0)
    return vol * 2 / np.pi


def corrct_bp_angle(line, degs_one, n):
    return skt.iradon(line[:, np.newaxis], degs_one, circle=False, output_size=n, filter_name=None) * 2 / np.pi


def corrct_compute_attenuation(vol, angle_rad, invert=False):
    size_lims = np.array(vol.shape[-2:])
    min_size = np.ceil(np.sqrt(np.sum(size_lims**2)))
    edges = np.ceil((min_size - size_lims) / 2).astype(int)
    if invert:
        angle_rad += np.pi
    deg = np.rad2deg(angle_rad)
    arr = np.pad(vol, ((edges[0], edges[0]), (edges[1], edges[1])), mode="constant")
    arr = skt.rotate(arr, -deg, order=1, clip=False)
    arr += np.roll(arr, 1, axis=-2)
    arr = np.cumsum(arr / 2, axis=-2)
    arr = skt.rotate(arr, deg, order=1, clip=False)
    arr = arr[..., edges[0] : -edges[0], edges[1] : -edges[1]]
    return np.exp(-arr)


def main():
    rng = np.random.default_rng(20260101)
    out = {"skimage_version": np.array(skimage.__version__), "numpy_version": np.array(np.__version__)}

    # --- radon lines ---------------------------------------------------------
    cases = []
    for n in (8, 16, 17, 32):
        rad = np.concatenate(
            [np.array([0.0, np.pi / 2, np.pi, np.pi / 4, 3 * np.pi / 2], dtype=np.float32), rng.random(4).astype(np.float32) * np.float32(2 * np.pi)]
        ).astype(np.float32)
        img = phantom(rng, n)
        degs = np.rad2deg(rad)
        lines = np.stack([corrct_fp_line(img, d) for d in degs])
        cases.append(n)
        out[f"radon_n{n}_img"] = img
        out[f"radon_n{n}_rad"] = rad
        out[f"radon_n{n}_lines"] = lines
    out["radon_sizes"] = np.array(cases)

    # --- iradon-based bp -------------------------------------------------------
    for n in (8, 16, 17):
        rad = np.linspace(0, 2 * np.pi, 12, endpoint=False).astype(np.float32)
        degs = np.rad2deg(rad)
        prj = rng.random((len(rad), n)).astype(np.float32)
        out[f"bp_n{n}_rad"] = rad
        out[f"bp_n{n}_prj"] = prj
        out[f"bp_n{n}_vol"] = corrct_bp(prj, degs, n)

    # --- rotate -------------------------------------------------------------
    img = rng.random((23, 23)).astype(np.float32)
    out["rot_img"] = img
    out["rot_deg_f32"] = np.array([-37.5, 12.25, 90.0, -200.0], dtype=np.float32)
    out["rot_out_f32"] = np.stack([skt.rotate(img, d, order=1, clip=False) for d in out["rot_deg_f32"]])
    out["rot_deg_f64"] = np.array([-101.3, 33.3333, 271.0], dtype=np.float64)
    out["rot_out_f64"] = np.stack([skt.rotate(img, d, order=1, clip=False) for d in out["rot_deg_f64"]])

    # --- compute_attenuation & maps -----------------------------------------
    for n in (16, 21):
        mu = (rng.random((n, n)) * 0.1).astype(np.float32)
        rad = np.linspace(0, 2 * np.pi, 6, endpoint=False).astype(np.float32)
        det = np.array([np.pi / 2, -np.pi / 2], dtype=np.float64)
        e_in = np.stack([corrct_compute_attenuation(mu, a, invert=False) for a in rad])
        e_out = np.stack([[corrct_compute_attenuation(mu, a + p, invert=True) for p in det] for a in rad])
        maps = np.ones((len(rad), len(det), n, n), dtype=np.float32)
        for ia in range(len(rad)):
            maps[ia] *= e_in[ia][None]
            maps[ia] *= e_out[ia]
        out[f"att_n{n}_mu"] = mu
        out[f"att_n{n}_rad"] = rad
        out[f"att_n{n}_det"] = det
        out[f"att_n{n}_ein"] = e_in.astype(np.float32)
        out[f"att_n{n}_eout"] = e_out.astype(np.float32)
        out[f"att_n{n}_maps"] = maps

    # --- XRF fp with maps and weights ----------------------------------------
    n = 16
    vol = phantom(rng, n)
    rad = np.linspace(0, 2 * np.pi, 8, endpoint=False).astype(np.float32)
    degs = np.rad2deg(rad)
    maps = rng.random((len(rad), 2, n, n)).astype(np.float32)
    w_det = np.array([1.0, 0.5])
    sino = np.zeros((2, len(rad), n), dtype=np.float32)
    for ia in range(len(rad)):
        tmp = vol * maps[ia]
        for d in range(2):
            sino[d, ia] = w_det[d] * corrct_fp_line(tmp[d], degs[ia])
    out["xrf_vol"] = vol
    out["xrf_rad"] = rad
    out["xrf_maps"] = maps
    out["xrf_wdet"] = w_det
    out["xrf_sino"] = sino

    # --- XRF symmetric bp (is_symmetric=True): the weights are float64 (weights_det times the
    # float64 default angle weights), each detector's one-angle bp is scaled, stacked, multiplied
    # by the angle's maps in place and summed over detectors, then the angles are summed
    sino_b = rng.random((2, len(rad), n)).astype(np.float32)
    w_ang = np.ones((len(rad), 2))
    vols = []
    for ia in range(len(rad)):
        weights = w_det * w_ang[ia, :]
        per = [weights[d] * corrct_bp_angle(sino_b[d, ia], degs[ia:ia + 1], n) for d in range(2)]
        per = np.stack(per, axis=0)
        per *= maps[ia]
        vols.append(np.sum(per, axis=0))
    out["xrf_bp_sino"] = sino_b
    out["xrf_bp_sym"] = np.sum(vols, axis=0)

    np.savez_compressed(OUT, **out)
    print("wrote", OUT, "with", len(out), "arrays; skimage", skimage.__version__)


if __name__ == "__main__":
    main()
