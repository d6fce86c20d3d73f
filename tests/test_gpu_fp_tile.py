"""GPU parity of the image-tile fp (fp_kernel 10, the default whole-sinogram fp) against the oracle.

The tile fp splits every line over the tiles its samples fall in; each sample must be counted by
exactly one tile (an interval of rows per bin and tile, refined with the samples' own float32
arithmetic). Cases: sides that are not a multiple of the 64 x 32 tile, a single partial tile,
axis-aligned and near-axis angles (where the row intervals are bounded by tiny sines / cosines),
odd batches (slice pairs plus a single), one to three detectors, no maps, angle ranges with a maps
block (on-the-fly maps), and the bin table rebuilt after the radon table changes.
"""

import numpy as np
import pytest

from oracle import xrf as oxrf
from tests.conftest import rel_max

pytestmark = pytest.mark.gpu

TOL = 1e-5


@pytest.fixture(scope="module")
def pkg():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import pycorrectedemissionct_amd as p

    return p


def _angles(a, extra=True):
    ang = np.linspace(0, 2 * np.pi, a, endpoint=False)
    if extra:  # axis-aligned and near-axis angles
        ang = np.concatenate([ang, [np.pi / 2, np.pi, 3 * np.pi / 2, 1e-4, np.pi - 1e-5, np.pi / 4, 2e-7]])
    return ang.astype(np.float32)


def _run(pkg, n, ang, d, batch, maps=True, shape=None, seed=0):
    import torch

    from pycorrectedemissionct_amd._projector_backends import HipPlan

    rng = np.random.default_rng(seed)
    det = [np.pi / 2, -np.pi / 2, 0.3][:d]
    w_det = rng.uniform(0.5, 1.5, d)
    plan = HipPlan(n, ang, det, w_det=w_det)
    for key, val in (shape or {}).items():
        plan.tune(key, val)
    a = len(ang)
    vol = rng.random((batch, n, n)).astype(np.float32)  # non-zero outside the inscribed circle too
    m = rng.uniform(0.2, 1.0, (batch, a, d, n, n)).astype(np.float32) if maps else None
    out = torch.zeros((batch, d, a, n), device="cuda")
    plan.fp(torch.as_tensor(vol, device="cuda"), None if m is None else torch.as_tensor(m, device="cuda"), out, batch,
            a * d * n * n)
    got = out.cpu().numpy()
    errs = []
    for b in range(batch):
        ref = oxrf.fp(vol[b], ang, None if m is None else m[b], w_det=w_det, n_det=d)
        errs.append(rel_max(ref, got[b]))
    return max(errs)


@pytest.mark.parametrize("n", [8, 12, 36, 68, 100, 132])
def test_tile_fp_sizes_and_near_axis_angles(pkg, n):
    """Sides below, between and above the tile size; 24 regular plus 7 (near-)axis angles."""
    err = _run(pkg, n, _angles(24), 2, 3, seed=n)
    print(f"tile fp n={n}: rel_max {err:.2e}")
    assert err < TOL


@pytest.mark.parametrize("d", [1, 2, 3])
@pytest.mark.parametrize("batch", [1, 2, 5])
def test_tile_fp_detectors_and_batches(pkg, d, batch):
    """One to three detectors (passes of two and one), single slices, pairs and pairs + one."""
    err = _run(pkg, 76, _angles(20), d, batch, seed=10 * d + batch)
    print(f"tile fp d={d} batch={batch}: rel_max {err:.2e}")
    assert err < TOL


@pytest.mark.parametrize("kind", [0, 1, 2, 3])
def test_tile_fp_kinds_give_the_same_bits(pkg, kind):
    """The kernel kinds, angle groups and slice pairing change the time, not the bits."""
    import torch

    from pycorrectedemissionct_amd._projector_backends import HipPlan

    n, ang, d, batch = 100, _angles(40), 2, 3
    rng = np.random.default_rng(5)
    vol = torch.as_tensor(rng.random((batch, n, n)).astype(np.float32), device="cuda")
    m = torch.as_tensor(rng.uniform(0.2, 1, (batch, len(ang), d, n, n)).astype(np.float32), device="cuda")
    plan = HipPlan(n, ang, [np.pi / 2, -np.pi / 2])
    outs = []
    for group, pairs in ((8, 1), (1, 0), (47, 1)):
        plan.tune("fp_tile_kind", kind)
        plan.tune("fp_tile_group", group)
        plan.tune("fp_tile_pairs", pairs)
        out = torch.zeros((batch, d, len(ang), n), device="cuda")
        plan.fp(vol, m, out, batch, len(ang) * d * n * n)
        outs.append(out)
    ref = oxrf.fp(vol[2].cpu().numpy(), ang, m[2].cpu().numpy())
    assert rel_max(ref, outs[0][2].cpu().numpy()) < TOL
    for o in outs[1:]:
        assert torch.equal(o, outs[0])


def test_tile_fp_without_maps(pkg):
    """The uncorrected projector's slices (no maps): one channel per slice, every detector's weight."""
    err = _run(pkg, 68, _angles(30), 2, 3, maps=False, seed=3)
    print(f"tile fp without maps: rel_max {err:.2e}")
    assert err < TOL
    prj = pkg.ProjectorUncorrected([52, 52], _angles(17))
    vol = np.random.default_rng(4).random((52, 52)).astype(np.float32)
    assert rel_max(oxrf.fp(vol, _angles(17))[0], prj.fp(vol)) < TOL


def test_tile_fp_matches_stream_fp_and_c_oracle_at_512(pkg):
    """The bench shape's geometry (512^2, 720 angles, 2 detectors) on two slices against the stream fp
    and the oracle's C restatement."""
    import torch

    from oracle import cops
    from pycorrectedemissionct_amd._projector_backends import HipPlan

    n, a, d, batch = 512, 720, 2, 2
    ang = np.linspace(0, 2 * np.pi, a, endpoint=False).astype(np.float32)
    plan = HipPlan(n, ang, [np.pi / 2, -np.pi / 2])
    gen = torch.Generator(device="cuda").manual_seed(9)
    vol = torch.rand((batch, n, n), device="cuda", generator=gen)
    maps = torch.rand((batch, a, d, n, n), device="cuda", generator=gen) * 0.8 + 0.2
    out = {}
    for k in (10, 5):
        plan.tune("fp_kernel", k)
        o = torch.zeros((batch, d, a, n), device="cuda")
        plan.fp(vol, maps, o, batch, a * d * n * n)
        out[k] = o
    assert float((out[10] - out[5]).abs().max() / out[5].abs().max()) < 2e-6
    ref = cops.fp(vol[1].cpu().numpy(), ang, maps[1].cpu().numpy(), n_det=2)
    err = rel_max(ref, out[10][1].cpu().numpy())
    print(f"tile fp 512 x 720 x 2 vs C oracle: {err:.2e}")
    assert err < TOL


def test_tile_fp_table_follows_radon_table(pkg):
    """A new radon table (chip_plan_set_radon_trig) rebuilds the bin table before the next fp."""
    import torch

    from pycorrectedemissionct_amd import _native
    from pycorrectedemissionct_amd._projector_backends import HipPlan

    n, d = 64, 1
    ang = _angles(16, extra=False)
    plan = HipPlan(n, ang, [np.pi / 2])
    vol = torch.as_tensor(np.random.default_rng(6).random((1, n, n)).astype(np.float32), device="cuda")
    out = torch.zeros((1, d, len(ang), n), device="cuda")
    plan.fp(vol, None, out, 1)
    ang2 = (ang + np.float32(0.37)).astype(np.float32)
    trig = _native.radon_trig(ang2)
    _native.call("chip_plan_set_radon_trig", plan.handle, trig.ctypes.data)
    plan.fp(vol, None, out, 1)
    ref = oxrf.fp(vol[0].cpu().numpy(), ang2)
    assert rel_max(ref[0], out[0, 0].cpu().numpy()) < TOL


def test_tile_fp_on_the_fly_angle_blocks(pkg):
    """On-the-fly maps run the tile fp over angle blocks of a maps block: the same bits as the
    precomputed maps, with several blocks (a small budget)."""
    import torch

    n, a = 96, 50
    rng = np.random.default_rng(8)
    ang = _angles(a, extra=False)
    f = pkg.testing.phantom_ellipses(n, 5, rng)
    mu_in, mu_out = pkg.testing.xrf_attenuation(f)
    pre = pkg.ProjectorAttenuationXRF([n, n], ang, att_in=mu_in, att_out=mu_out, angles_detectors_rad=[np.pi / 2])
    otf = pkg.ProjectorAttenuationXRF([n, n], ang, att_in=mu_in, att_out=mu_out, angles_detectors_rad=[np.pi / 2],
                                      maps_mode="otf")
    otf._plan().tune("otf_budget_mib", 16)
    x = torch.as_tensor(f, device="cuda")
    assert torch.equal(otf.fp(x), pre.fp(x))
