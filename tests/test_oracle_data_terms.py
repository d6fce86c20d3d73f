"""The oracle's data terms (oracle/data_terms.py) pinned by what the mathematics fixes.

Each elementwise PDHG data dual step of corrct (data_terms.py:283-757) is the proximal map
prox_{sigma f*}(v - ...) of the term's convex conjugate: here the minimiser of
sigma g(p) + (p - v)^2 / 2 is found by a brute-force ternary search in float64 and compared
with the oracle's closed form. Reductions between terms are bit-exact checks, and the group
projections are checked for the unit-ball projection property. Host logic of the product's
data-term classes (residual norms, group layout) runs on CPU tensors.
"""

import numpy as np
import pytest
import torch

from oracle.data_terms import DataTerm

F32 = np.float32


def _argmin_convex(phi, lo, hi, iters=200):
    """Vectorised ternary search of a convex phi on [lo, hi] (float64)."""
    lo = np.array(lo, dtype=np.float64)
    hi = np.array(hi, dtype=np.float64)
    for _ in range(iters):
        m1 = lo + (hi - lo) / 3
        m2 = hi - (hi - lo) / 3
        left = phi(m1) <= phi(m2)
        hi = np.where(left, m2, hi)
        lo = np.where(left, lo, m1)
    return (lo + hi) / 2


def _samples(kind, n=400, seed=0):
    rng = np.random.default_rng(seed)
    v = rng.uniform(-3, 3, n).astype(F32)
    s = rng.uniform(0.1, 3, n).astype(F32)
    b = (rng.uniform(0.1, 3, n) if kind == "kl" else rng.uniform(-2, 2, n)).astype(F32)
    e = rng.uniform(0.0, 2.0, n).astype(F32)
    w = rng.uniform(0.1, 3.0, n).astype(F32)
    return v, s, b, e, w


# sigma * (the term's conjugate), as a function of the dual p, and the dual's domain
CONJ = {
    "l2": (lambda p, s, b, e, w: s * (0.5 * p * p + p * b), None),
    "wl2": (lambda p, s, b, e, w: s * (p * p / (2 * w) + p * b), None),
    "l2b": (lambda p, s, b, e, w: s * (p * b + 0.5 * e * p * p + np.sqrt(e) * np.abs(p)), None),
    "hub": (lambda p, s, b, e, w: s * (p * b + 0.5 * e * p * p), 1.0),
    "l1": (lambda p, s, b, e, w: s * p * b, 1.0),
    # the reference thresholds by local_error itself, not sigma * local_error (data_terms.py:644-645)
    "l1b": (lambda p, s, b, e, w: s * p * b + e * np.abs(p), 1.0),
    "kl": (lambda p, s, b, e, w: -s * b * np.log1p(-p), "kl"),
}


@pytest.mark.parametrize("kind", list(CONJ))
def test_dual_step_is_the_conjugate_prox(kind):
    v, s, b, e, w = _samples(kind)
    term = DataTerm(kind, local_error=e, weights=w)
    term.assign(b, s)
    got = term.prox(v).astype(np.float64)
    g, dom = CONJ[kind]
    v64, s64, b64, e64, w64 = (x.astype(np.float64) for x in (v, s, b, e, w))

    def phi(p):
        return g(p, s64, b64, e64, w64) + 0.5 * (p - v64) ** 2

    if dom is None:
        r = np.abs(v64) + s64 * np.abs(b64) + 10
        lo, hi = -r, r
    elif dom == "kl":
        lo, hi = -(np.abs(v64) + 10), np.full_like(v64, 1 - 1e-12)
    else:
        lo, hi = np.full_like(v64, -dom), np.full_like(v64, dom)
    want = _argmin_convex(phi, lo, hi)
    err = np.max(np.abs(got - want) / (1 + np.abs(want)))
    print(f"{kind}: max err vs brute-force prox {err:.2e}")
    assert err < 2e-5


def test_reductions_between_terms():
    rng = np.random.default_rng(1)
    shape = (3, 7, 9)
    v = rng.normal(0, 2, shape).astype(F32)
    s = rng.uniform(0.1, 2, shape).astype(F32)
    b = rng.normal(0, 1, shape).astype(F32)

    def run(kind, **kw):
        t = DataTerm(kind, **kw)
        t.assign(b, s)
        return t.prox(v)

    l2, l1 = run("l2"), run("l1")
    assert np.array_equal(run("wl2", weights=np.ones(shape, F32)), l2)
    assert np.array_equal(run("wl2", weights=F32(1)), l2)
    assert np.array_equal(run("l1b", local_error=F32(0)), l1)
    assert np.array_equal(run("hub", local_error=F32(0)), l1)
    # l12 over a length-1 axis is l1 (sqrt(fl(v * v)) = |v| in binary floating point)
    t = DataTerm("l12", l2_axis=0)
    t.assign(b[None], s[None])
    assert np.array_equal(t.prox(v[None])[0], l1)
    # zero weights pin the dual at zero; l2b at zero radius leaves v - sigma b unscaled
    w = np.where(rng.random(shape) < 0.3, 0, 1).astype(F32)
    out = run("wl2", weights=w)
    assert np.all(out[w == 0] == 0) and np.array_equal(out[w == 1], l2[w == 1])
    assert np.array_equal(run("l2b", local_error=F32(0)), (v - s * b).astype(F32))


@pytest.mark.parametrize("kind,axis", [("l12", 0), ("l12", -1), ("l12", 1), ("hub", 0), ("hub", 2)])
def test_group_steps_project_onto_the_unit_ball(kind, axis):
    rng = np.random.default_rng(2)
    shape = (3, 6, 5)
    v = rng.normal(0, 1.0, shape).astype(F32)
    s = rng.uniform(0.1, 1, shape).astype(F32)
    b = rng.normal(0, 0.5, shape).astype(F32)
    e = F32(0.3)
    t = DataTerm(kind, local_error=e, l2_axis=axis)
    t.assign(b, s)
    out = t.prox(v).astype(np.float64)
    u = (v - s * b).astype(np.float64)
    if kind == "hub":
        u = u / (1 + s.astype(np.float64) * 0.3)
    un = np.linalg.norm(u, axis=axis, keepdims=True)
    on = np.linalg.norm(out, axis=axis, keepdims=True)
    inside = np.broadcast_to(un <= 1, shape)
    assert np.all(on <= 1 + 1e-6)
    np.testing.assert_allclose(out[inside], u[inside], rtol=2e-6, atol=1e-7)
    np.testing.assert_allclose(on[un > 1], 1.0, rtol=2e-6)
    cos = np.sum(out * u, axis=axis, keepdims=True) / np.maximum(on * un, 1e-30)
    np.testing.assert_allclose(cos[un > 1], 1.0, rtol=2e-6)


def test_residuals():
    rng = np.random.default_rng(3)
    shape = (4, 11)
    a = rng.normal(1, 1, shape).astype(F32)
    b = rng.uniform(0, 2, shape).astype(F32)
    s = rng.uniform(0.2, 2, shape).astype(F32)
    m = (rng.random(shape) < 0.8).astype(F32)

    def res(kind, **kw):
        t = DataTerm(kind, **kw)
        t.assign(b, s)
        return t.residual(a, m)

    plain = ((b - a) * m).astype(F32)
    assert np.array_equal(res("l2"), plain)
    assert np.array_equal(res("wl2", weights=F32(1)), plain)
    assert np.array_equal(res("l2b", local_error=F32(0)), plain)
    # l2b shrinks each residual towards 0 by sigma sqrt(e), never past it
    r = res("l2b", local_error=F32(0.25))
    np.testing.assert_allclose(np.abs(r), np.fmax(np.abs(plain) - s * 0.5, 0), rtol=1e-6, atol=1e-7)
    assert np.all(r * plain >= 0)
    # KL without data: the residual is -min(sigma a, 1) (the conjugate's domain is p < 1)
    t = DataTerm("kl")
    t.assign(np.zeros(shape, F32), s)
    want = -np.minimum(s * np.fmax(a, np.finfo(F32).eps), 1) * m
    np.testing.assert_allclose(t.residual(a, m), want, rtol=1e-6, atol=1e-7)


def test_host_residual_norms_and_group_layout():
    from pycorrectedemissionct_amd import data_terms as dt

    rng = np.random.default_rng(4)
    x1 = rng.normal(size=7).astype(F32)
    x2 = rng.normal(size=(5, 6)).astype(F32)
    # np.linalg.norm(x, ord=1) without an axis: vector 1-norm, then the largest column sum
    assert dt.DataFidelity_l1().compute_residual_norm(torch.from_numpy(x1)) == pytest.approx(np.abs(x1).sum())
    assert dt.DataFidelity_l1().compute_residual_norm(torch.from_numpy(x2)) == pytest.approx(
        np.abs(x2).sum(axis=0).max())
    with pytest.raises(ValueError):
        dt.DataFidelity_l1().compute_residual_norm(torch.zeros(2, 3, 4))
    # Huber as data_terms.py:486-489 evaluates it: the l1 part indexes rows 0 / 1
    d = torch.tensor([[0.5, 2.0], [3.0, -0.2]])
    assert dt.DataFidelity_Huber(1.0).compute_residual_norm(d) == pytest.approx(0.25 + 0.04 + 2 * (0.5 + 2 + 3 + 0.2))
    # wl2 divides by sqrt(w) where w != 0 and drops the rest
    w = np.array([[1.0, 4.0], [0.0, 0.25]], F32)
    assert dt.DataFidelity_wl2(w).compute_residual_norm(d) == pytest.approx(0.25 + 1.0 + 0.16)
    assert dt._group_layout((2, 5, 7), 0) == (2, 35)
    assert dt._group_layout((2, 5, 7), 1) == (5, 7)
    assert dt._group_layout((2, 5, 7), -1) == (7, 1)
    assert dt._group_layout((2, 5, 7), None) == (1, 1)
    like = torch.zeros(2, 3)
    t, period = dt._param_dev(0.5, like)
    assert period == 1 and t.shape == (1,)
    t, period = dt._param_dev(np.arange(3), like)
    assert period == 6 and torch.equal(t.reshape(-1), torch.tensor([0.0, 1, 2, 0, 1, 2]))
    with pytest.raises(NotImplementedError):
        dt.DataFidelity_ln().device_term(like)
