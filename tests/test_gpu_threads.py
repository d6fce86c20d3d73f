"""The HIP backend driven the way unmodified corrct drives a backend (run with -m gpu).

corrct's ``ProjectorAttenuationXRF.fp`` maps ``fp_angle`` over the angles on a
``ThreadPoolExecutor(round(log2(ncpu + 1)))`` (``projectors.py:20, 585-587``);
``fp_angle`` multiplies the volume by the angle's maps and calls
``backend.fp(vol * M[a, d], a)`` once per detector, weighted by
``w_det[d] * w_ang[a, d]`` (``projectors.py:502-531``). The symmetric bp maps
``bp_angle`` the same way and sums (``projectors.py:533-568, 609-618``), and
the default bp calls ``backend.bp(sino[d])`` per detector (``:619-622``).
Here those calls hit one shared ``ProjectorBackendHIP`` from many threads at
once, its plan cache starting empty, and the assembled results must equal
the whole-sinogram HIP projector.
"""

import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from tests.conftest import rel_max

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pkg():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import pycorrectedemissionct_amd as p

    return p


def _workers():
    # corrct's pool size (projectors.py:20), at least 4 so that calls overlap
    return max(4, int(round(np.log2((os.cpu_count() or 1) + 1))))


def _setup(pkg, n=96, a=72, seed=17):
    rng = np.random.default_rng(seed)
    f = pkg.testing.phantom_ellipses(n, 6, rng, semi_axes=(4, 20))
    mu_in, mu_out = pkg.testing.xrf_attenuation(f)
    ang = np.linspace(0, 2 * np.pi, a, endpoint=False).astype(np.float32)
    det = [np.pi / 2, -np.pi / 2]
    w_det = np.array([0.7, 1.3])
    return f, mu_in, mu_out, ang, det, w_det


def test_per_angle_thread_pool_matches_whole_sinogram(pkg):
    f, mu_in, mu_out, ang, det, w_det = _setup(pkg)
    n, a = f.shape[0], len(ang)
    A = pkg.ProjectorAttenuationXRF([n, n], ang, att_in=mu_in, att_out=mu_out, angles_detectors_rad=det,
                                    weights_detectors=w_det)
    maps = A.att_vol_angles.cpu().numpy()  # [A, D, N, N]
    w_ang = np.ones((a, len(det)))
    whole_fp = A.fp(f)  # [D, A, N]
    sino = whole_fp.astype(np.float32)
    whole_bp = A.bp(sino)

    backend = pkg._projector_backends.ProjectorBackendHIP()
    backend.initialize_geometry(pkg.models.VolumeGeometry.get_default_from_volume(f), ang)

    def fp_angle(ai):
        temp = f * maps[ai]
        w = w_det * w_ang[ai]
        return np.stack([w[d] * backend.fp(np.ascontiguousarray(temp[d]), ai) for d in range(len(det))], axis=0)

    with ThreadPoolExecutor(_workers()) as ex:
        lines = list(ex.map(fp_angle, range(a)))
    per_angle_fp = np.stack(lines, axis=-2)  # [D, A, N]
    e_fp = rel_max(whole_fp, per_angle_fp)

    def bp_angle_sym(ai):
        w = w_det * w_ang[ai]
        vols = np.stack([w[d] * backend.bp(np.ascontiguousarray(sino[d, ai]), ai) for d in range(len(det))], axis=0)
        return np.sum(vols * maps[ai], axis=0)

    with ThreadPoolExecutor(_workers()) as ex:
        vols = list(ex.map(bp_angle_sym, range(a)))
    per_angle_bp_sym = np.sum(vols, axis=0)
    A_sym = pkg.ProjectorAttenuationXRF([n, n], ang, att_in=mu_in, att_out=mu_out, angles_detectors_rad=det,
                                        weights_detectors=w_det, is_symmetric=True)
    whole_bp_sym = A_sym.bp(sino)
    e_sym = rel_max(whole_bp_sym, per_angle_bp_sym)

    # default (non-symmetric) bp: one whole-sinogram backend call per detector, also from threads
    with ThreadPoolExecutor(_workers()) as ex:
        per_det = list(ex.map(lambda d: backend.bp(np.ascontiguousarray(sino[d])), range(len(det))))
    e_bp = rel_max(whole_bp, np.sum(per_det, axis=0))
    print(f"threads={_workers()}: fp {e_fp:.2e}, symmetric bp {e_sym:.2e}, default bp {e_bp:.2e}")
    assert e_fp < 1e-5 and e_sym < 1e-5 and e_bp < 1e-5
    assert len(backend._plans) == 1  # one plan, created once under the lock


def test_concurrent_whole_calls_on_shared_plan_and_streams(pkg):
    """Threads on separate torch streams share one plan's adjoint-bp part scratch."""
    import torch

    f, mu_in, mu_out, ang, det, w_det = _setup(pkg, n=128, a=160, seed=23)
    A = pkg.ProjectorAttenuationXRF([128, 128], ang, att_in=mu_in, att_out=mu_out, angles_detectors_rad=det,
                                    bp_mode="adjoint")
    dev = torch.device("cuda", 0)
    rng = np.random.default_rng(5)
    sinos = [torch.from_numpy(rng.random((len(det), len(ang), 128), dtype=np.float32)).to(dev) for _ in range(8)]
    ref = [A.bp(s).clone() for s in sinos]
    torch.cuda.synchronize()

    def run(k):
        st = torch.cuda.Stream(device=dev)
        with torch.cuda.stream(st):
            outs = [A.bp(sinos[(k + r) % len(sinos)]) for r in range(6)]
        st.synchronize()
        return [(k + r) % len(sinos) for r in range(6)], outs

    with ThreadPoolExecutor(4) as ex:
        results = list(ex.map(run, range(4)))
    worst = 0.0
    for idx, outs in results:
        for i, o in zip(idx, outs):
            worst = max(worst, float((o - ref[i]).abs().max()))
    print(f"max |concurrent - serial| = {worst:.3e}")
    assert worst == 0.0
