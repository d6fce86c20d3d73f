"""GPU parity of the data-fidelity terms beyond l2 / KL (run with -m gpu).

chip_pdhg_dual_norm and chip_data_residual against the oracle's restatement of corrct's
data_terms.py (one step, elementwise and grouped, scalar and array parameters, masks and
background), then PDHG, SIRT and FISTA solves with those terms against the oracle solvers
on the two-detector XRF problem. Bar: one kernel step within |got - want| <= 3e-6 max(|want|, 1)
of the oracle (the same float32 operations in the same order, the group norms summed in
another order); solves within 1e-4 relative after 30 iterations.
"""

import numpy as np
import pytest

from oracle import solvers as osol
from oracle import xrf as oxrf
from oracle.data_terms import DataTerm
from tests.conftest import rel_max

pytestmark = pytest.mark.gpu

TOL = 1e-4
F32 = np.float32


@pytest.fixture(scope="module")
def pkg():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import pycorrectedemissionct_amd as p

    return p


def _dev(x):
    import torch

    return torch.from_numpy(np.ascontiguousarray(x, dtype=F32)).cuda()


def _step_inputs(shape, seed):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 1, shape).astype(F32)
    ax = rng.uniform(0, 3, shape).astype(F32)
    sigma = rng.uniform(0.05, 1.5, shape).astype(F32)
    b = rng.uniform(-0.5, 3, shape).astype(F32)
    bg = rng.uniform(0, 0.2, shape).astype(F32)
    mask = (rng.random(shape) > 0.15).astype(F32)
    w = np.where(rng.random(shape) < 0.1, 0, rng.uniform(0.1, 2, shape)).astype(F32)
    e_arr = rng.uniform(0, 1.5, shape).astype(F32)
    return p, ax, sigma, b, bg, mask, w, e_arr


# (name, term class builder, oracle kind, oracle kwargs) for a data shape (D, A, N)
def _cases(pkg, w, e_arr):
    dt = pkg.data_terms
    return [
        ("wl2", dt.DataFidelity_wl2(w), "wl2", dict(weights=w)),
        ("wl2 scalar", dt.DataFidelity_wl2(0.7), "wl2", dict(weights=F32(0.7))),
        ("l2b", dt.DataFidelity_l2b(0.3), "l2b", dict(local_error=F32(0.3))),
        ("l2b array", dt.DataFidelity_l2b(e_arr), "l2b", dict(local_error=e_arr)),
        ("hub", dt.DataFidelity_Huber(0.4), "hub", dict(local_error=F32(0.4))),
        ("hub axis 0", dt.DataFidelity_Huber(0.4, l2_axis=0), "hub", dict(local_error=F32(0.4), l2_axis=0)),
        ("hub axis 2", dt.DataFidelity_Huber(e_arr, l2_axis=2), "hub", dict(local_error=e_arr, l2_axis=2)),
        ("l1", dt.DataFidelity_l1(), "l1", {}),
        ("l12 axis 0", dt.DataFidelity_l12(l2_axis=0), "l12", dict(l2_axis=0)),
        ("l12 axis 1", dt.DataFidelity_l12(l2_axis=1), "l12", dict(l2_axis=1)),
        ("l12 axis -1", dt.DataFidelity_l12(l2_axis=-1), "l12", dict(l2_axis=-1)),
        ("l1b", dt.DataFidelity_l1b(0.2), "l1b", dict(local_error=F32(0.2))),
        ("l1b array", dt.DataFidelity_l1b(e_arr), "l1b", dict(local_error=e_arr)),
    ]


def test_dual_norm_kernel_matches_oracle(pkg):
    shape = (3, 20, 37)
    p, ax, sigma, b, bg, mask, w, e_arr = _step_inputs(shape, 1)
    st = pkg._native.stream()
    for name, term, kind, kw in _cases(pkg, w, e_arr):
        for with_extras in (False, True):
            ref_term = DataTerm(kind, **kw)
            ref_term.assign(b, sigma)
            want = ref_term.dual(p, ax, bg if with_extras else None)
            if with_extras:
                want = (want * mask).astype(F32)
            # every operand stays referenced until the call returns (a temporary's memory would be
            # handed to the next allocation)
            pt, at, st_, bt = _dev(p), _dev(ax), _dev(sigma), _dev(b)
            bgt = _dev(bg) if with_extras else None
            mt = _dev(mask) if with_extras else None
            code, prm, period, g_len, g_stride = term.device_term(bt)
            pkg._native.call("chip_pdhg_dual_norm", pt.data_ptr(), at.data_ptr(), st_.data_ptr(), bt.data_ptr(),
                             pkg._native.ptr(bgt), pkg._native.ptr(mt), pt.numel(), code, pkg._native.ptr(prm),
                             period, g_len, g_stride, st)
            got = pt.cpu().numpy()
            err = np.max(np.abs(got - want) / np.fmax(np.abs(want), 1))
            print(f"dual {name}{' +bg +mask' if with_extras else ''}: max rel err {err:.2e}")
            assert err < 3e-6, name


def test_data_residual_kernel_matches_oracle(pkg):
    shape = (2, 16, 33)
    _, ax, sigma, b, bg, mask, w, e_arr = _step_inputs(shape, 2)
    scale = np.random.default_rng(3).uniform(0.1, 2, shape).astype(F32)
    dt = pkg.data_terms
    cases = [("l2", dt.DataFidelity_l2(), {}), ("wl2", dt.DataFidelity_wl2(w), dict(weights=w)),
             ("l2b", dt.DataFidelity_l2b(e_arr), dict(local_error=e_arr)), ("kl", dt.DataFidelity_KL(), {})]
    st = pkg._native.stream()
    for name, term, kw in cases:
        for asig in (None, sigma):
            ref_term = DataTerm(name, **kw)
            ref_term.assign(b, 1.0 if asig is None else asig)
            want = (ref_term.residual(ax, mask, bg) * scale).astype(F32)
            res, bt, at, bgt, mt, sct = (_dev(v) for v in (np.zeros(shape), b, ax, bg, mask, scale))
            ast = None if asig is None else _dev(asig)
            code, prm, period, _, _ = term.device_term(bt)
            pkg._native.call("chip_data_residual", res.data_ptr(), bt.data_ptr(), at.data_ptr(), bgt.data_ptr(),
                             mt.data_ptr(), pkg._native.ptr(ast), sct.data_ptr(), res.numel(), code,
                             pkg._native.ptr(prm), period, st)
            got = res.cpu().numpy()
            err = np.max(np.abs(got - want) / np.fmax(np.abs(want), 1))
            print(f"residual {name} asig={'sigma' if asig is not None else 1}: max rel err {err:.2e}")
            assert err < 3e-6, name


def test_kernels_reject_bad_configurations(pkg):
    import torch

    t = torch.zeros(12, device="cuda")
    st = pkg._native.stream()
    dt = pkg.data_terms
    bad = [
        (dt.TERM_KL, None, 1, 1, 1),  # KL has its own entry
        (dt.TERM_L1, None, 1, 3, 4),  # groups belong to L12 / HUBER
        (dt.TERM_L12, None, 1, 5, 1),  # 12 is not a multiple of 5
        (dt.TERM_WL2, None, 1, 1, 1),  # weights missing
        (dt.TERM_L2B, t, 5, 1, 1),  # period does not divide n
    ]
    for code, prm, period, g_len, g_stride in bad:
        with pytest.raises(RuntimeError):
            pkg._native.call("chip_pdhg_dual_norm", t.data_ptr(), t.data_ptr(), t.data_ptr(), t.data_ptr(), None,
                             None, 12, code, pkg._native.ptr(prm), period, g_len, g_stride, st)
    with pytest.raises(RuntimeError):
        pkg._native.call("chip_data_residual", t.data_ptr(), t.data_ptr(), t.data_ptr(), None, None, None, None, 12,
                         dt.TERM_L1, None, 1, st)
    with pytest.raises(NotImplementedError):
        pkg.PDHG(data_term=dt.DataFidelity_ln())(pkg.ProjectorMatrix(np.eye(4, dtype=F32), (2, 2), (4,)),
                                                 np.ones(4, F32), iterations=1)
    with pytest.raises(ValueError):
        pkg.SIRT(data_term=dt.DataFidelity_l1())


@pytest.fixture(scope="module")
def xrf_problem(pkg):
    n, a = 32, 24
    rng = np.random.default_rng(600)
    f = pkg.testing.phantom_ellipses(n, 4, rng, semi_axes=(3, 8))
    mu_in, mu_out = pkg.testing.xrf_attenuation(f)
    ang = np.linspace(0, 2 * np.pi, a, endpoint=False).astype(np.float32)
    det = [np.pi / 2, -np.pi / 2]
    prj = pkg.ProjectorAttenuationXRF([n, n], ang, att_in=mu_in, att_out=mu_out, angles_detectors_rad=det)
    maps = prj.att_vol_angles.cpu().numpy()
    sino = oxrf.fp(f, ang, maps)
    b = pkg.testing.poisson_data(sino, 1e3, rng)
    A = lambda x: oxrf.fp(x, ang, maps)  # noqa: E731
    At = lambda y: oxrf.bp_skimage(y, ang, n)  # noqa: E731
    w = np.where(rng.random(b.shape) < 0.05, 0, 1 / np.fmax(b, 1)).astype(F32)
    return dict(prj=prj, b=b, A=A, At=At, shape=(n, n), w=w)


def test_pdhg_with_each_data_term(pkg, xrf_problem):
    p = xrf_problem
    dt = pkg.data_terms
    e = float(np.mean(p["b"]) * 0.05)
    cases = [
        ("wl2", dt.DataFidelity_wl2(p["w"]), DataTerm("wl2", weights=p["w"])),
        ("l2b", dt.DataFidelity_l2b(e), DataTerm("l2b", local_error=F32(e))),
        ("hub", dt.DataFidelity_Huber(e), DataTerm("hub", local_error=F32(e))),
        ("hub axis 0", dt.DataFidelity_Huber(e, l2_axis=0), DataTerm("hub", local_error=F32(e), l2_axis=0)),
        ("l1", "l1", DataTerm("l1")),
        ("l12 axis 0", dt.DataFidelity_l12(l2_axis=0), DataTerm("l12", l2_axis=0)),
        ("l1b", dt.DataFidelity_l1b(e), DataTerm("l1b", local_error=F32(e))),
    ]
    for name, term, ref_term in cases:
        reg = pkg.regularizers.Regularizer_TV2D(1e-2)
        x = pkg.PDHG(data_term=term, regularizer=reg)(p["prj"], p["b"], iterations=30)[0]
        ref = osol.pdhg(p["A"], p["At"], p["b"], p["shape"], 30, data_term=ref_term, tv_weight=1e-2)
        err = rel_max(ref, x)
        print(f"PDHG-{name}-TV 30 it: rel_max {err:.2e}")
        assert err < TOL, name


def test_sirt_and_fista_with_l2_family_and_kl_terms(pkg, xrf_problem):
    p = xrf_problem
    dt = pkg.data_terms
    e = float(np.mean(p["b"]) * 0.05)
    cases = [
        ("wl2", lambda: dt.DataFidelity_wl2(p["w"]), lambda: DataTerm("wl2", weights=p["w"])),
        ("l2b", lambda: dt.DataFidelity_l2b(e), lambda: DataTerm("l2b", local_error=F32(e))),
        ("kl", lambda: "kl", lambda: DataTerm("kl")),
    ]
    for name, term, ref_term in cases:
        x = pkg.SIRT(data_term=term())(p["prj"], p["b"], iterations=30)[0]
        ref = osol.sirt(p["A"], p["At"], p["b"], p["shape"], 30, data_term=ref_term())
        err = rel_max(ref, x)
        print(f"SIRT-{name} 30 it: rel_max {err:.2e}")
        assert err < TOL, name
        if name == "kl":
            continue
        x = pkg.FISTA(data_term=term())(p["prj"], p["b"], iterations=30)[0]
        ref = osol.fista(p["A"], p["At"], p["b"], p["shape"], 30, data_term=ref_term())
        err = rel_max(ref, x)
        print(f"FISTA-{name} 30 it: rel_max {err:.2e}")
        assert err < TOL, name
