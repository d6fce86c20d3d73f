"""The benchmark configurations at their own shapes (run with -m gpu).

Each test states the configuration it exercises. Oracles: the NumPy
restatement (``oracle.xrf``, ``oracle.solvers``) where it finishes in seconds,
and the C restatement (``oracle.cops``: fp, skimage bp, exact-adjoint scatter;
pinned against the NumPy oracle in ``tests/test_oracle_c.py``) at full sizes.
Bars: fp / bp 1e-5 max-relative, dot test 1e-5, solvers after fixed iteration
counts 1e-4, the warm-start network's gradients 2e-2 and its output as the
bfloat16 storage allows (see the test).
"""

import numpy as np
import pytest

from oracle import cops
from oracle import solvers as osol
from oracle import xrf as oxrf
from tests.conftest import rel_l2, rel_max

pytestmark = pytest.mark.gpu

F32 = np.float32


@pytest.fixture(scope="module")
def pkg():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import pycorrectedemissionct_amd as p

    return p


def _angles(a):
    return np.linspace(0, 2 * np.pi, a, endpoint=False).astype(np.float32)


def _say(msg):
    print(msg, flush=True)


# ---- the benchmark's own solver path: PDHG-KL-TV / MLEM with the exact-adjoint bp ----------


def test_adjoint_bp_solvers_vs_oracle(pkg):
    """PDHG-KL-TV(1e-2), PDHG-l2, MLEM and SIRT with ``bp_mode="adjoint"`` (the bench's solver path)
    after 30 iterations against the oracle loops driven by the oracle fp and its exact transpose."""
    n, a = 32, 24
    rng = np.random.default_rng(1600)
    f = pkg.testing.phantom_ellipses(n, 4, rng, semi_axes=(3, 8))
    mu_in, mu_out = pkg.testing.xrf_attenuation(f)
    ang = _angles(a)
    det = [np.pi / 2, -np.pi / 2]
    prj = pkg.ProjectorAttenuationXRF([n, n], ang, att_in=mu_in, att_out=mu_out, angles_detectors_rad=det,
                                      bp_mode="adjoint")
    maps = prj.att_vol_angles.cpu().numpy()
    b = pkg.testing.poisson_data(oxrf.fp(f, ang, maps), 1e3, rng)
    A = lambda x: oxrf.fp(x, ang, maps)  # noqa: E731
    At = lambda y: oxrf.bp_adjoint(y, ang, n, maps).astype(F32)  # noqa: E731
    for dt in ("kl", "l2"):
        x = pkg.PDHG(data_term=dt, regularizer=pkg.regularizers.Regularizer_TV2D(1e-2))(prj, b, iterations=30)[0]
        ref = osol.pdhg(A, At, b, (n, n), 30, data_term=dt, tv_weight=1e-2)
        _say(f"PDHG-{dt}-TV adjoint 30 it: rel_max {rel_max(ref, x):.2e}")
        assert rel_max(ref, x) < 1e-4
    x = pkg.MLEM()(prj, b, iterations=30)[0]
    ref = osol.mlem(A, At, b, (n, n), 30)
    _say(f"MLEM adjoint 30 it: rel_max {rel_max(ref, x):.2e}")
    assert rel_max(ref, x) < 1e-4
    x = pkg.SIRT()(prj, b, iterations=30)[0]
    ref = osol.sirt(A, At, b, (n, n), 30)
    _say(f"SIRT adjoint 30 it: rel_max {rel_max(ref, x):.2e}")
    assert rel_max(ref, x) < 1e-4


def test_headline_pdhg_kl_tv_and_mlem_vs_c_oracle(pkg):
    """The bench workload itself: 512^2, 30 ellipses + 5 blobs, 720 angles, detectors at +-90 deg,
    Poisson peak 1e3, exact-adjoint bp; PDHG-KL-TV(1e-2) and MLEM for 5 iterations against the
    oracle loops on the C fp and C exact-adjoint scatter (same maps, built on the GPU)."""
    wl = pkg.testing.workload_pdhg_512()
    f, ang, det, rng = wl["f"], wl["angles"], wl["det"], wl["rng"]
    n = f.shape[0]
    prj = pkg.ProjectorAttenuationXRF([n, n], ang, att_in=wl["mu_in"], att_out=wl["mu_out"],
                                      angles_detectors_rad=det, bp_mode="adjoint")
    maps = prj.att_vol_angles.cpu().numpy()
    tabs = cops.tables(ang, n)
    b = pkg.testing.poisson_data(cops.fp(f, ang, maps, tabs=tabs), wl["peak"], rng)
    A = lambda x: cops.fp(x, ang, maps, tabs=tabs)  # noqa: E731
    At = lambda y: cops.bp_adjoint(y, ang, n, maps, tabs=tabs)  # noqa: E731
    its = 5
    x = pkg.PDHG(data_term="kl", regularizer=pkg.regularizers.Regularizer_TV2D(1e-2))(prj, b, iterations=its)[0]
    ref = osol.pdhg(A, At, b, (n, n), its, data_term="kl", tv_weight=1e-2)
    _say(f"PDHG-KL-TV 512 x 720 x 2, {its} it: rel_max {rel_max(ref, x):.2e}, rel_l2 {rel_l2(ref, x):.2e}")
    assert rel_max(ref, x) < 1e-4
    xm = pkg.MLEM()(prj, b, iterations=its)[0]
    refm = osol.mlem(A, At, b, (n, n), its)
    _say(f"MLEM 512 x 720 x 2, {its} it: rel_max {rel_max(refm, xm):.2e}")
    assert rel_max(refm, xm) < 1e-4


# ---- on-the-fly maps at the full-scale shape -------------------------------------------


def test_full_scale_otf_1024_1440(pkg):
    """1024^2, 1440 angles, one detector: the on-the-fly fp equals the fp over the 6 GB of
    precomputed maps bit for bit, the on-the-fly adjoint bp matches the precomputed one, the dot
    test holds (the adjoint's bar), and a 90-angle subset of the fp matches the C oracle to 1e-5.
    The adjoint against the C scatter: the scatter transposes the fp's float32-rounded sample
    coordinates, the kernel takes its tent weights from unrounded offsets, and float32 rounding of
    coordinates of magnitude up to N = 1024 is up to 3e-5 px, so the two agree to a few 1e-5
    (2e-6 ... 8e-6 at N = 128 ... 768, growing with N; see the batched test)."""
    import torch

    n, a = 1024, 1440
    rng = np.random.default_rng(1700)
    f = pkg.testing.phantom_ellipses(n, 30, rng, n_blobs=5)
    mu_in, mu_out = pkg.testing.xrf_attenuation(f)
    ang = _angles(a)
    kw = dict(att_in=mu_in, att_out=mu_out, angles_detectors_rad=[np.pi / 2], bp_mode="adjoint")
    pre = pkg.ProjectorAttenuationXRF([n, n], ang, **kw)
    otf = pkg.ProjectorAttenuationXRF([n, n], ang, maps_mode="otf", **kw)
    ft = torch.as_tensor(f, device="cuda")
    s_pre, s_otf = pre.fp(ft), otf.fp(ft)
    _say(f"otf fp 1024 x 1440: equal={torch.equal(s_pre, s_otf)}")
    assert torch.equal(s_pre, s_otf)
    y = torch.as_tensor(rng.random((a, n), dtype=np.float32), device="cuda")
    b_pre, b_otf = pre.bp(y), otf.bp(y)
    e_bp = rel_max(b_pre.cpu().numpy(), b_otf.cpu().numpy())
    x = torch.as_tensor(rng.random((n, n), dtype=np.float32), device="cuda")
    lhs = float(torch.dot(pre.fp(x).double().ravel(), y.double().ravel()))
    rhs = float(torch.dot(x.double().ravel(), b_pre.double().ravel()))
    e_dot = abs(lhs - rhs) / abs(lhs)
    _say(f"otf bp vs precomputed: rel_max {e_bp:.2e}; dot test {e_dot:.2e}")
    assert e_bp < 1e-5 and e_dot < 1e-5
    sub = slice(0, a, 16)
    maps_sub = pre.att_vol_angles[sub].cpu().numpy()  # [90, 1, n, n]
    ref_fp = cops.fp(f, ang[sub], maps_sub)[0]
    e_cfp = rel_max(ref_fp, s_pre[sub].cpu().numpy())
    y_sub = torch.zeros_like(y)
    y_sub[sub] = y[sub]
    ref_bp = cops.bp_adjoint(y[sub].cpu().numpy(), ang[sub], n, maps_sub)
    e_cbp = rel_max(ref_bp, pre.bp(y_sub).cpu().numpy())
    _say(f"90-angle subset vs C oracle: fp {e_cfp:.2e}, adjoint bp {e_cbp:.2e}")
    assert e_cfp < 1e-5 and e_cbp < 5e-5
    del pre
    torch.cuda.empty_cache()


def test_otf_volume_of_8_slices_equals_slices(pkg):
    """Eight 1024^2 slices with their own attenuation, 1440 angles, on-the-fly maps: the volume
    projector's fp and adjoint bp equal eight single-slice on-the-fly projectors."""
    import torch

    n, a, z = 1024, 1440, 8
    rng = np.random.default_rng(1701)
    fs = np.stack([pkg.testing.phantom_ellipses(n, 12, rng) for _ in range(z)])
    mus = [pkg.testing.xrf_attenuation(f) for f in fs]
    mu_in, mu_out = np.stack([m[0] for m in mus]), np.stack([m[1] for m in mus])
    ang = _angles(a)
    vol = pkg.ProjectorAttenuationXRF([n, n, z], ang, att_in=mu_in, att_out=mu_out, maps_mode="otf",
                                      bp_mode="adjoint")
    ft = torch.as_tensor(fs, device="cuda")
    s_vol = vol.fp(ft)  # [z, A, n]
    y = torch.as_tensor(rng.random((z, a, n), dtype=np.float32), device="cuda")
    b_vol = vol.bp(y)
    for k in range(z):
        one = pkg.ProjectorAttenuationXRF([n, n], ang, att_in=mu_in[k], att_out=mu_out[k], maps_mode="otf",
                                          bp_mode="adjoint")
        assert torch.equal(s_vol[k], one.fp(ft[k].contiguous())), f"fp slice {k}"
        assert torch.equal(b_vol[k], one.bp(y[k].contiguous())), f"bp slice {k}"
    _say(f"otf volume of {z} slices: fp and bp equal the slices")


def test_otf_bp_first_chunk_splits_into_parts(pkg):
    """128^2, 160 angles, two slices, the default on-the-fly budget: the first angle chunk of each
    slice has 64 or more angles, so the adjoint bp splits it into parts summed in the plan scratch,
    reused slice after slice. Against the precomputed-maps adjoint and the C scatter oracle."""
    import torch

    n, a, z = 128, 160, 2
    rng = np.random.default_rng(1702)
    fs = np.stack([pkg.testing.phantom_ellipses(n, 5, rng) for _ in range(z)])
    mus = [pkg.testing.xrf_attenuation(f) for f in fs]
    mu_in, mu_out = np.stack([m[0] for m in mus]), np.stack([m[1] for m in mus])
    ang = _angles(a)
    det = [np.pi / 2, -np.pi / 2]
    kw = dict(att_in=mu_in, att_out=mu_out, angles_detectors_rad=det, bp_mode="adjoint")
    otf = pkg.ProjectorAttenuationXRF([n, n, z], ang, maps_mode="otf", **kw)
    pre = pkg.ProjectorAttenuationXRF([n, n, z], ang, **kw)
    y = rng.random((2, z, a, n)).astype(np.float32)  # [D, Z, A, N]
    got, want = otf.bp(y), pre.bp(y)
    assert rel_max(want, got) < 1e-5
    for k in range(z):
        single = pkg.ProjectorAttenuationXRF([n, n], ang, att_in=mu_in[k], att_out=mu_out[k],
                                             angles_detectors_rad=det, bp_mode="adjoint")
        maps = single.att_vol_angles.cpu().numpy()
        ref = cops.bp_adjoint(y[:, k], ang, n, maps)
        e = rel_max(ref, got[k])
        _say(f"otf parts bp slice {k}: rel_max vs scatter oracle {e:.2e}, vs precomputed {rel_max(want, got):.2e}")
        assert e < 1e-5
    torch.cuda.synchronize()


# ---- batches of uneven slices at the batched configuration's sizes ----------------------

CFG3_SIZES = [128, 256, 384, 512, 768]
CFG3_SCALES = [0.02, 0.04, 0.08, 0.16]


def _cfg3_batch(pkg, rng, a):
    sizes, fs, mu_in, mu_out = [], [], [], []
    for n in CFG3_SIZES:
        f = pkg.testing.phantom_ellipses(n, 10, rng)
        support = (f > 0).astype(np.float32)
        mi = (0.02 * support + 0.005).astype(np.float32)
        for s in CFG3_SCALES:
            sizes.append(n)
            fs.append((support * rng.uniform(0, 1, (n, n))).astype(np.float32))
            mu_in.append(mi)
            mu_out.append((s * support).astype(np.float32))
    return sizes, fs, mu_in, mu_out


def test_batched_cfg3_sizes_and_channels(pkg):
    """One slice of every size 128 ... 768, each with 4 emission channels (mu_out scales 0.02 ...
    0.16), 360 angles, one detector: the batched fp equals the single-slice fp bit for bit and the
    batched adjoint bp matches it to 1e-5 (a lone slice splits its angle range), one channel per
    size matches the C oracle's fp and adjoint, and 5 batched PDHG-TV / SIRT iterations with 1 %
    Gaussian noise equal single-slice solves."""
    import torch

    a = 360
    rng = np.random.default_rng(1800)
    ang = _angles(a)
    sizes, fs, mu_in, mu_out = _cfg3_batch(pkg, rng, a)
    A = pkg.batch.BatchedProjectorXRF(sizes, ang, mu_in, mu_out, angles_detectors_rad=[np.pi / 2],
                                      bp_mode="adjoint")
    xs = A.pack_vol(fs)
    fx = A.unpack_sino(A.fp(xs))
    ys = [rng.random((1, a, n)).astype(np.float32) for n in sizes]
    by = A.unpack_vol(A.bp(A.pack_sino(ys)))
    for s, n in enumerate(sizes):
        single = pkg.ProjectorAttenuationXRF([n, n], ang, att_in=mu_in[s], att_out=mu_out[s],
                                             angles_detectors_rad=[np.pi / 2], bp_mode="adjoint")
        assert torch.equal(A.slice_maps(s), single.att_vol_angles), f"maps of slice {s}"
        assert torch.equal(fx[s].reshape(a, n), single.fp(torch.as_tensor(fs[s], device="cuda"))), f"fp {s}"
        e = rel_max(single.bp(ys[s][0]), by[s].cpu().numpy())
        assert e < 1e-5, f"bp of slice {s}: {e:.2e}"
        if s % len(CFG3_SCALES) == 0:
            maps = A.slice_maps(s).cpu().numpy()
            e_fp = rel_max(cops.fp(fs[s], ang, maps), fx[s].cpu().numpy())
            e_bp = rel_max(cops.bp_adjoint(ys[s], ang, n, maps), by[s].cpu().numpy())
            _say(f"cfg3 n={n}: batched fp vs C oracle {e_fp:.2e}, adjoint bp {e_bp:.2e}")
            assert e_fp < 1e-5 and e_bp < 1e-5
        del single
    b = A.fp(xs)
    b = b + 0.01 * b.max() * torch.as_tensor(rng.standard_normal(b.numel()), dtype=torch.float32, device="cuda")
    its = 5
    xb = A.unpack_vol(pkg.batch.BatchedPDHG(data_term="l2", tv_lambda=1e-2, use_graph=True)(A, b, its))
    xsb = A.unpack_vol(pkg.batch.BatchedSIRT(use_graph=True)(A, b, its))
    bs = A.unpack_sino(b)
    worst_p = worst_s = 0.0
    for s, n in enumerate(sizes):
        if s % len(CFG3_SCALES) not in (0, 3):
            continue
        single = pkg.ProjectorAttenuationXRF([n, n], ang, att_in=mu_in[s], att_out=mu_out[s],
                                             angles_detectors_rad=[np.pi / 2], bp_mode="adjoint")
        bsl = bs[s].reshape(a, n).contiguous()
        ref = pkg.PDHG(data_term="l2", regularizer=pkg.regularizers.Regularizer_TV2D(1e-2))(single, bsl, its)[0]
        worst_p = max(worst_p, rel_max(ref.cpu().numpy(), xb[s].cpu().numpy()))
        ref_s = pkg.SIRT()(single, bsl, its)[0]
        worst_s = max(worst_s, rel_max(ref_s.cpu().numpy(), xsb[s].cpu().numpy()))
        del single
    _say(f"cfg3 batched PDHG-TV {its} it vs single: {worst_p:.2e}; SIRT: {worst_s:.2e}")
    assert worst_p < 1e-5 and worst_s < 1e-5


# ---- the warm-start network at its configuration's size ----------------------------------


def test_unet_4_levels_32_channels_256(pkg):
    """The 4-level, 32-base-channel U-Net (about 2.1 M parameters) on a 256^2 input against the
    float64 oracle with the same bfloat16 rounding points: every parameter gradient within 2e-2
    relative l2. The forward output is checked through 17 layers of bfloat16 storage, where a
    float32 accumulation and a float64 one round a few activations to neighbouring bfloat16 values
    (one ulp = 2^-8 = 3.9e-3 relative): within 1e-2 of the rounded oracle, and no farther from the
    unrounded float64 network than the rounded oracle itself (times 1.5)."""
    import torch

    from oracle import unet as ou
    from tests.test_gpu_unet import _oracle_of

    rng = np.random.default_rng(1900)
    net = pkg.warmstart.UNet(levels=4, base=32, seed=4, head_std=0.3)
    n_params = sum(int(c.w.numel() + c.b.numel()) for c in net.convs.values())
    assert 2.0e6 < n_params < 2.3e6, n_params
    x = pkg.testing.phantom_ellipses(256, 10, rng)[None].astype(np.float32)
    tgt = (x + 0.05 * rng.standard_normal(x.shape)).astype(np.float32)
    xt, tt = torch.as_tensor(x).cuda(), torch.as_tensor(tgt).cuda()
    out = net.forward(xt, keep=True)
    loss, dout = net.loss_and_grad(out, tt)
    net.backward(dout)
    orc = _oracle_of(net)
    ref = orc.forward(x)
    e = rel_l2(ref, out.cpu().numpy())
    sd = net.state_dict()
    params = {k[:-2]: (sd[k], sd[k[:-2] + ".b"]) for k in sd if k.endswith(".w")}
    exact = ou.UNetOracle(net.levels, net.base, params, sd["head"][:-1], sd["head"][-1], bf16=False).forward(x)
    e_gpu64, e_orc64 = rel_l2(exact, out.cpu().numpy()), rel_l2(exact, ref)
    grads = orc.backward(ref, tgt)
    worst = 0.0
    for name, c in net.convs.items():
        worst = max(worst, rel_l2(grads[name][0], c.gw.cpu().numpy()), rel_l2(grads[name][1], c.gb.cpu().numpy()))
    eh = rel_l2(np.append(grads["head"][0], grads["head"][1]), net.ghead.cpu().numpy())
    _say(f"U-Net 4 x 32 at 256^2 ({n_params} params): forward rel_l2 {e:.2e} vs the rounded oracle; vs the "
         f"unrounded float64 network: GPU {e_gpu64:.2e}, rounded oracle {e_orc64:.2e}; worst gradient {worst:.2e}, "
         f"head {eh:.2e}")
    assert e < 1e-2 and e_gpu64 <= 1.5 * e_orc64 + 1e-4
    assert worst < 2e-2 and eh < 2e-2
