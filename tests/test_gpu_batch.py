"""Batches of slices of different sizes (run with -m gpu).

The persistent batched kernels compute every work item exactly as the
single-slice kernels do, so per slice the batched maps, fp and bp are
bit-identical to the single-slice projector (itself checked against the
oracle to 1e-5; these cases use fewer than 64 angles, below which a lone
slice's bp does not split its angle range into parts summed in order), and the
batched SIRT / PDHG match per-slice solver runs.
HIP-graph replays are checked against eager iterations, bit for bit, over
several replays with eager work in between.
"""

import numpy as np
import pytest

from oracle import attenuation as oatt
from oracle import xrf as oxrf
from tests.conftest import rel_max

pytestmark = pytest.mark.gpu

SIZES = [64, 128, 32, 96, 128, 40]
DET = [np.pi / 2, -np.pi / 2]
W_DET = [1.0, 0.7]


@pytest.fixture(scope="module")
def pkg():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import pycorrectedemissionct_amd as p

    return p


def _angles(a):
    return np.linspace(0, 2 * np.pi, a, endpoint=False).astype(np.float32)


def _mus(rng, sizes):
    mu_in = [(rng.random((n, n)) * 0.02).astype(np.float32) for n in sizes]
    mu_out = [(rng.random((n, n)) * 0.04).astype(np.float32) for n in sizes]
    return mu_in, mu_out


def _single(pkg, n, ang, mi, mo, bp_mode):
    return pkg.ProjectorAttenuationXRF([n, n], ang, att_in=mi, att_out=mo, angles_detectors_rad=DET,
                                       weights_detectors=W_DET, bp_mode="adjoint" if bp_mode == "adjoint" else "skimage",
                                       is_symmetric=bp_mode == "symmetric")


@pytest.mark.parametrize("bp_mode", ["adjoint", "skimage", "symmetric"])
def test_batched_projector_equals_single_slices(pkg, bp_mode):
    import torch

    a = 36
    rng = np.random.default_rng(700)
    ang = _angles(a)
    mu_in, mu_out = _mus(rng, SIZES)
    A = pkg.batch.BatchedProjectorXRF(SIZES, ang, mu_in, mu_out, angles_detectors_rad=DET, weights_detectors=W_DET,
                                      bp_mode=bp_mode)
    xs = [rng.random((n, n)).astype(np.float32) for n in SIZES]
    ys = [rng.random((2, a, n)).astype(np.float32) for n in SIZES]
    fx = A.unpack_sino(A.fp(A.pack_vol(xs)))
    by_auto = A.unpack_vol(A.bp(A.pack_sino(ys)))
    # the paired adjoint bp splits the angle range into parts by occupancy (another grouping of
    # the sum); with one part it adds in the single-slice order
    A.plan.tune("bp_pair_parts", 1)
    by = A.unpack_vol(A.bp(A.pack_sino(ys)))
    for s in range(len(SIZES)):
        assert float((by_auto[s] - by[s]).abs().max() / by[s].abs().max()) < 1e-6, f"bp parts of slice {s}"
    for s, n in enumerate(SIZES):
        single = _single(pkg, n, ang, mu_in[s], mu_out[s], bp_mode)
        assert torch.equal(A.slice_maps(s), single.att_vol_angles), f"maps of slice {s}"
        assert torch.equal(fx[s], single.fp(torch.as_tensor(xs[s], device="cuda"))), f"fp of slice {s}"
        assert torch.equal(by[s], single.bp(torch.as_tensor(ys[s], device="cuda"))), f"bp of slice {s}"
    # one slice against the oracle directly
    s = 2
    maps = oatt.compute_maps(mu_in[s], mu_out[s], ang, np.array(DET))
    assert rel_max(oxrf.fp(xs[s], ang, maps, w_det=W_DET), fx[s].cpu().numpy()) < 1e-5


def test_autotune_changes_time_not_results(pkg):
    import torch

    a = 24
    rng = np.random.default_rng(701)
    ang = _angles(a)
    mu_in, mu_out = _mus(rng, SIZES)
    A = pkg.batch.BatchedProjectorXRF(SIZES, ang, mu_in, mu_out, angles_detectors_rad=DET, bp_mode="adjoint")
    x = A.pack_vol([rng.random((n, n)).astype(np.float32) for n in SIZES])
    y = A.pack_sino([rng.random((2, a, n)).astype(np.float32) for n in SIZES])
    A.plan.tune("fp_wg_per_cu", 1)
    A.plan.tune("bp_wg_per_cu", 1)
    f1, b1 = A.fp(x).clone(), A.bp(y).clone()
    res = pkg.batch.autotune(A)
    print("autotune:", res)
    assert f"{res['fp_shape']}/{res['fp_wg_per_cu']}" in res["fp_us"] and res["bp_wg_per_cu"] in res["bp_us"]
    assert torch.equal(A.fp(x), f1) and torch.equal(A.bp(y), b1)
    for shape, keys in pkg.batch.TILE_SHAPES.items():  # every shape the tuner may pick: the same bits
        for key, val in keys.items():
            A.plan.tune(key, val)
        for group, pairs in ((8, 1), (3, 0), (64, 1)):  # angle groups and slice pairing: the same bits
            A.plan.tune("fp_tile_group", group)
            A.plan.tune("fp_tile_pairs", pairs)
            assert torch.equal(A.fp(x), f1), (shape, group, pairs)
    for shape, keys in pkg.batch.STREAM_SHAPES.items():  # the other family: the same bits among themselves
        for key, val in keys.items():
            A.plan.tune(key, val)
        if shape == "stream":
            f2 = A.fp(x).clone()
        assert torch.equal(A.fp(x), f2), shape
        assert float((A.fp(x) - f1).abs().max() / f1.abs().max()) < 1e-6
    for key, val in (("fp_patch", 1), ("fp_direct", 1), ("fp_wg_per_cu", 0), ("bp_wg_per_cu", 0)):
        A.plan.tune(key, val)  # other kernel shapes and flat grids: results equal to 2e-7 (patch: sum order)
        assert float((A.fp(x) - f1).abs().max() / f1.abs().max()) < 1e-6
        assert torch.equal(A.bp(y), b1)
    with pytest.raises(pkg._native.NativeError):
        A.plan.tune("fp_wg_per_cu", -1)


def test_batch_rejects_sizes_not_multiple_of_4(pkg):
    with pytest.raises(pkg._native.NativeError, match="N % 4"):
        pkg.batch.BatchPlan([64, 30], _angles(8))


@pytest.mark.parametrize("data_term", ["l2", "kl"])
def test_batched_pdhg_tv_equals_single_solves(pkg, data_term):
    import torch

    a, its = 32, 30
    rng = np.random.default_rng(702)
    ang = _angles(a)
    sizes = [64, 32, 96]
    mu_in, mu_out = _mus(rng, sizes)
    A = pkg.batch.BatchedProjectorXRF(sizes, ang, mu_in, mu_out, angles_detectors_rad=DET, weights_detectors=W_DET,
                                      bp_mode="adjoint")
    fs = [pkg.testing.phantom_ellipses(n, 4, rng) for n in sizes]
    b = A.fp(A.pack_vol(fs))
    b = torch.clamp(b + 0.01 * b.max() * torch.as_tensor(rng.standard_normal(b.numel()), dtype=torch.float32,
                                                         device="cuda"), min=0.0)
    solver = pkg.batch.BatchedPDHG(data_term=data_term, tv_lambda=1e-2, use_graph=True)
    xb = A.unpack_vol(solver(A, b, its))
    bs = A.unpack_sino(b)
    for s, n in enumerate(sizes):
        single = _single(pkg, n, ang, mu_in[s], mu_out[s], "adjoint")
        ref = pkg.PDHG(data_term=data_term, regularizer=pkg.regularizers.Regularizer_TV2D(1e-2))
        st = ref.prepare(single, bs[s].contiguous())
        ref.iterate(single, st, its)
        e = rel_max(st["x"].cpu().numpy(), xb[s].cpu().numpy())
        print(f"batched PDHG-{data_term} slice {s} (n={n}): rel_max vs single {e:.2e}")
        assert e < 1e-5


def test_batched_sirt_equals_single_solves(pkg):
    import torch

    a, its = 32, 20
    rng = np.random.default_rng(703)
    ang = _angles(a)
    sizes = [32, 64]
    mu_in, mu_out = _mus(rng, sizes)
    A = pkg.batch.BatchedProjectorXRF(sizes, ang, mu_in, mu_out, angles_detectors_rad=DET, weights_detectors=W_DET,
                                      bp_mode="adjoint")
    fs = [pkg.testing.phantom_ellipses(n, 3, rng) for n in sizes]
    b = A.fp(A.pack_vol(fs))
    xb = A.unpack_vol(pkg.batch.BatchedSIRT(use_graph=True)(A, b, its))
    bs = A.unpack_sino(b)
    for s, n in enumerate(sizes):
        single = _single(pkg, n, ang, mu_in[s], mu_out[s], "adjoint")
        x, _ = pkg.SIRT()(single, bs[s].contiguous(), its)
        e = rel_max(x.cpu().numpy(), xb[s].cpu().numpy())
        print(f"batched SIRT slice {s} (n={n}): rel_max vs single {e:.2e}")
        assert e < 1e-5


def test_graph_replays_equal_eager(pkg):
    """25 + 20 iterations: one eager warm-up step, a 10-step graph replayed 2 + 2 times and
    eager remainders in between, against 45 eager iterations; bit-identical."""
    import torch

    a = 24
    rng = np.random.default_rng(704)
    ang = _angles(a)
    sizes = [64, 32]
    mu_in, mu_out = _mus(rng, sizes)
    A = pkg.batch.BatchedProjectorXRF(sizes, ang, mu_in, mu_out, angles_detectors_rad=DET, bp_mode="adjoint")
    b = A.fp(A.pack_vol([pkg.testing.phantom_ellipses(n, 3, rng) for n in sizes]))
    runs = {}
    for use_graph in (True, False):
        solver = pkg.batch.BatchedPDHG(data_term="l2", tv_lambda=1e-2, graph_chunk=10, use_graph=use_graph)
        st = solver.prepare(A, b)
        solver.iterate(A, st, 25)
        torch.cuda.synchronize()
        mid = st["x"].clone()
        solver.iterate(A, st, 20)
        runs[use_graph] = (mid, st["x"].clone())
    assert torch.equal(runs[True][0], runs[False][0])
    assert torch.equal(runs[True][1], runs[False][1])


@pytest.mark.parametrize("reg_name", ["vtv", "tnv"])
def test_batched_pdhg_channel_coupling_equals_volume_solves(pkg, reg_name):
    """Emission channels of a slice as consecutive batch slices, coupled by VTV / TNV: the batched
    PDHG-KL equals the single-path PDHG on each [C, N, N] channel volume (per-channel attenuation)."""
    import torch

    a, its, c = 24, 30, 3
    rng = np.random.default_rng(703)
    ang = _angles(a)
    sides = [48, 64]
    sizes = [n for n in sides for _ in range(c)]
    mu_in, mu_out = [], []
    for n in sides:
        mi = (rng.random((n, n)) * 0.02).astype(np.float32)
        mo = (rng.random((n, n)) * 0.04).astype(np.float32)
        for k in range(c):
            mu_in.append(mi)
            mu_out.append((mo * (0.5 + k)).astype(np.float32))
    A = pkg.batch.BatchedProjectorXRF(sizes, ang, mu_in, mu_out, angles_detectors_rad=DET, weights_detectors=W_DET,
                                      bp_mode="adjoint")
    base = [pkg.testing.phantom_ellipses(n, 4, rng) for n in sides]
    fs = [base[g] * (0.3 + 0.3 * k) for g in range(len(sides)) for k in range(c)]
    b = A.fp(A.pack_vol(fs))
    b = torch.clamp(b + 0.01 * b.max() * torch.as_tensor(rng.standard_normal(b.numel()), dtype=torch.float32,
                                                         device="cuda"), min=0.0)
    make = {"vtv": lambda: pkg.regularizers.Regularizer_VTV(2e-2), "tnv": lambda: pkg.regularizers.Regularizer_TNV(2e-2)}
    solver = pkg.batch.BatchedPDHG(data_term="kl", regularizer=make[reg_name](), channels=c, use_graph=True)
    xb = A.unpack_vol(solver(A, b, its))
    bs = A.unpack_sino(b)
    for g, n in enumerate(sides):
        sl = range(g * c, (g + 1) * c)
        vol_prj = pkg.ProjectorAttenuationXRF([n, n, c], ang, att_in=np.stack([mu_in[s] for s in sl]),
                                              att_out=np.stack([mu_out[s] for s in sl]), angles_detectors_rad=DET,
                                              weights_detectors=W_DET, bp_mode="adjoint")
        # the volume sinogram is [D, C, A, N]
        b_vol = torch.stack([bs[s] for s in sl], dim=1).contiguous()
        ref = pkg.PDHG(data_term="kl", regularizer=make[reg_name]())
        st = ref.prepare(vol_prj, b_vol)
        ref.iterate(vol_prj, st, its)
        got = torch.stack([xb[s] for s in sl])
        e = rel_max(st["x"].cpu().numpy(), got.cpu().numpy())
        print(f"batched PDHG-KL-{reg_name} group {g} (n={n}, {c} channels): rel_max vs volume solve {e:.2e}")
        assert e < 1e-4


def test_batched_fused_otf_step_equals_separate_steps(pkg):
    """On-the-fly batches run the iteration's projector work as one fused step per slice and angle
    block (chip_batch_otf_fp_row_bp, the maps built once): graphed batched SIRT and PDHG (l2, KL, TV) on
    it equal the same solvers with the separate fp / elementwise / bp steps bit for bit, with several
    angle blocks per slice."""
    import torch

    a = 72
    rng = np.random.default_rng(706)
    ang = _angles(a)
    sizes = [96, 64, 128]
    mi, mo = _mus(rng, sizes)
    kw = dict(angles_detectors_rad=DET, weights_detectors=W_DET, bp_mode="adjoint", maps_mode="otf")
    fused = pkg.batch.BatchedProjectorXRF(sizes, ang, mi, mo, **kw)
    sep = pkg.batch.BatchedProjectorXRF(sizes, ang, mi, mo, **kw)
    for prj in (fused, sep):
        prj.plan.tune("otf_budget_mib", 16)
    sep.fused_data_step = lambda: False
    assert fused.fused_data_step()
    b = fused.fp(fused.pack_vol([pkg.testing.phantom_ellipses(n, 4, rng) for n in sizes]))
    xf = pkg.batch.BatchedSIRT(use_graph=True)(fused, b, 12)
    xs = pkg.batch.BatchedSIRT(use_graph=True)(sep, b, 12)
    assert torch.equal(xf, xs)
    for term in ("l2", "kl"):
        xf = pkg.batch.BatchedPDHG(data_term=term, tv_lambda=1e-2, use_graph=True)(fused, b, 12)
        xs = pkg.batch.BatchedPDHG(data_term=term, tv_lambda=1e-2, use_graph=True)(sep, b, 12)
        print(f"batched fused otf PDHG-{term}: equal to separate steps: {bool(torch.equal(xf, xs))}")
        assert torch.equal(xf, xs), term


def test_batched_otf_equals_precomputed(pkg):
    """maps_mode='otf' (each slice's maps rebuilt per angle block inside every fp / bp; a 16 MiB
    budget gives the 96^2 and 128^2 slices two blocks) equals the precomputed-maps batch: fp bit
    for bit, bp to rounding (blocks after the first add into the output), and the batched PDHG on
    it follows the precomputed one."""
    import torch

    a = 36
    rng = np.random.default_rng(705)
    ang = _angles(a)
    mu_in, mu_out = _mus(rng, SIZES)
    xs = [rng.random((n, n)).astype(np.float32) for n in SIZES]
    ys = [rng.random((2, a, n)).astype(np.float32) for n in SIZES]
    for bp_mode in ("adjoint", "symmetric"):
        kw = dict(angles_detectors_rad=DET, weights_detectors=W_DET, bp_mode=bp_mode)
        pre = pkg.batch.BatchedProjectorXRF(SIZES, ang, mu_in, mu_out, **kw)
        otf = pkg.batch.BatchedProjectorXRF(SIZES, ang, mu_in, mu_out, maps_mode="otf", **kw)
        assert otf.maps is None
        otf.plan.tune("otf_budget_mib", 16)
        x, y = pre.pack_vol(xs), pre.pack_sino(ys)
        assert torch.equal(otf.fp(x), pre.fp(x)), bp_mode
        e = rel_max(pre.bp(y).cpu().numpy(), otf.bp(y).cpu().numpy())
        print(f"batched otf bp ({bp_mode}) vs precomputed: rel_max {e:.2e}")
        assert e < 1e-5
    with pytest.raises(pkg._native.NativeError, match="otf_budget_mib"):
        otf.plan.tune("otf_budget_mib", 8)
    with pytest.raises(NotImplementedError):
        otf.slice_maps(0)
    with pytest.raises(ValueError, match="otf"):
        pkg.batch.BatchedProjectorXRF(SIZES, ang, maps_mode="otf")
    # a few graphed PDHG-TV iterations on each
    kw = dict(angles_detectors_rad=DET, weights_detectors=W_DET, bp_mode="adjoint")
    sizes = [64, 32, 96]
    mi, mo = _mus(rng, sizes)
    pre = pkg.batch.BatchedProjectorXRF(sizes, ang, mi, mo, **kw)
    otf = pkg.batch.BatchedProjectorXRF(sizes, ang, mi, mo, maps_mode="otf", **kw)
    b = pre.fp(pre.pack_vol([pkg.testing.phantom_ellipses(n, 4, rng) for n in sizes]))
    xp = pkg.batch.BatchedPDHG(data_term="l2", tv_lambda=1e-2, use_graph=True)(pre, b, 25)
    xo = pkg.batch.BatchedPDHG(data_term="l2", tv_lambda=1e-2, use_graph=True)(otf, b, 25)
    e = rel_max(xp.cpu().numpy(), xo.cpu().numpy())
    print(f"batched PDHG on otf vs precomputed: rel_max {e:.2e}")
    assert e < 1e-4


@pytest.mark.parametrize("det", [DET, [np.pi / 2]])
def test_batched_bp_slice_pairs_same_bits(pkg, det):
    """The exact-adjoint bp with four (one-detector passes) or two slices of a size per workgroup (the
    rest in pairs, an odd one alone) gives the bits of one slice per workgroup, with and without maps."""
    import torch

    a = 40
    rng = np.random.default_rng(707)
    sizes = [64, 32, 64, 96, 64, 32, 64, 64, 32, 32]  # 64: four + one, 32: four, 96: one
    ang = _angles(a)
    mu_in, mu_out = _mus(rng, sizes)
    A = pkg.batch.BatchedProjectorXRF(sizes, ang, mu_in, mu_out, angles_detectors_rad=det, bp_mode="adjoint")
    y = A.pack_sino([rng.random((len(det), a, n)).astype(np.float32) for n in sizes])
    A.plan.tune("bp_pairs", 0)
    ref = A.bp(y).clone()
    A.plan.tune("bp_pairs", 1)
    A.plan.tune("bp_pair_parts", 1)  # the whole angle range per workgroup: the same bits
    for quads in (1, 0):
        A.plan.tune("bp_quads", quads)
        assert torch.equal(A.bp(y), ref), quads
    A.plan.tune("bp_quads", 1)
    for parts in (2, 3, 0):  # angle parts summed in part order (0: from occupancy): the sum's grouping only
        A.plan.tune("bp_pair_parts", parts)
        assert float((A.bp(y) - ref).abs().max() / ref.abs().max()) < 1e-6, parts
    from oracle import xrf as oxrf

    for s in (3, 9):  # a slice alone and one of a group of four
        maps = A.slice_maps(s).cpu().numpy()
        got = A.unpack_vol(A.bp(y))[s].cpu().numpy()
        want = oxrf.bp_adjoint(A.unpack_sino(y)[s].cpu().numpy(), ang, sizes[s], maps)
        assert float(np.abs(got - want).max() / np.abs(want).max()) < 1e-5, s
    with pytest.raises(pkg._native.NativeError):
        A.plan.tune("bp_quads", 2)


@pytest.mark.parametrize("d", [1, 2])
def test_volume_adjoint_bp_slice_groups_equal_slices(pkg, d):
    """A [Z, N, N] volume's exact-adjoint bp runs its slices four (one detector) or two per workgroup,
    an odd last one alone: every slice has the bits of the single-slice projector."""
    n, a, z = 48, 40, 7
    rng = np.random.default_rng(708)
    ang = _angles(a)
    det = DET[:d]
    mu_in = (rng.random((z, n, n)) * 0.04).astype(np.float32)
    mu_out = (rng.random((z, n, n)) * 0.06).astype(np.float32)
    y = rng.random((d, z, a, n) if d > 1 else (z, a, n)).astype(np.float32)
    kw = dict(angles_detectors_rad=det, bp_mode="adjoint")
    p3 = pkg.ProjectorAttenuationXRF([n, n, z], ang, att_in=mu_in, att_out=mu_out, **kw)
    b3 = p3.bp(y)
    for i in range(z):
        p2 = pkg.ProjectorAttenuationXRF([n, n], ang, att_in=mu_in[i], att_out=mu_out[i], **kw)
        yi = np.ascontiguousarray(y[:, i] if d > 1 else y[i])
        assert np.array_equal(b3[i], p2.bp(yi)), i
