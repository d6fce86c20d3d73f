"""The oracle's solver step constants pinned by hand-computed values, not by convergence alone:
SIRT / PDHG relaxation 1.95 / 0.95 (solvers.py:839, 1019), the diagonal scalings' floor at 1e-5 of
the maximum (solvers.py:98-116), the gradient regularizers' sigma = 1/2 dual step and tau = 2 ndims
(regularizers.py:260-273). A slip in any of them (rho 1.0, a 1e-4 or 1e-6 floor, tau 4 for 3-D,
a dual step of 1) fails one of these cases. CPU only."""

import inspect

import numpy as np

from oracle import regularizers as oreg
from oracle import solvers as osol

F32 = np.float32


def _op(m):
    m = np.asarray(m, dtype=F32)
    return (lambda x: m @ x), (lambda y: m.T @ y)


def test_relaxation_defaults():
    assert inspect.signature(osol.sirt).parameters["relaxation"].default == 1.95
    assert inspect.signature(osol.pdhg).parameters["relaxation"].default == 0.95


def test_diagonal_scaling_floor_and_weights_by_hand():
    # |A 1| = [3, 3, 4e-6]: 4e-6 / 3 < 1e-5 of the maximum, so that entry becomes 1
    A, At = _op([[1, 2], [0, 3], [4e-6, 0]])
    sigma, tau = osol.diagonal_scaling(A, At, np.zeros(3, F32), (2,), tau_reg=4.0, rho_sigma=0.95, rho_tau=0.95)
    np.testing.assert_allclose(sigma, [0.95 / 3, 0.95 / 3, 0.95], rtol=1e-6)
    # |A^T 1| + tau_reg = [1 + 4e-6 + 4, 5 + 4]
    np.testing.assert_allclose(tau, [0.95 / 5.000004, 0.95 / 9], rtol=1e-6)
    # 4e-5 / 3 is above the floor: kept
    A, At = _op([[1, 2], [0, 3], [4e-5, 0]])
    sigma, _ = osol.diagonal_scaling(A, At, np.zeros(3, F32), (2,), rho_sigma=1.0, rho_tau=1.0)
    np.testing.assert_allclose(sigma, [1 / 3, 1 / 3, 1 / 4e-5], rtol=1e-6)


def test_pdhg_and_sirt_pass_the_regularizer_tau(monkeypatch):
    seen = []
    real = osol.diagonal_scaling

    def spy(*args, **kw):
        seen.append((kw.get("tau_reg"), kw.get("rho_sigma", 1.0), kw.get("rho_tau", 1.0)))
        return real(*args, **kw)

    monkeypatch.setattr(osol, "diagonal_scaling", spy)
    b = np.ones(4, F32)
    x_shape = (2, 2)
    A2 = lambda x: x.reshape(-1).astype(F32)
    At2 = lambda y: y.reshape(x_shape).astype(F32)
    osol.pdhg(A2, At2, b, x_shape, 1, data_term="l2", tv_weight=0.1)
    osol.pdhg(A2, At2, b, x_shape, 1, data_term="l2")
    osol.sirt(A2, At2, b, x_shape, 1, tv_weight=0.1)
    assert seen[0] == (4.0, 0.95, 0.95)  # 2-D TV: tau = 2 ndims, rho 0.95 on both steps
    assert seen[1] == (0.0, 0.95, 0.95)
    assert seen[2][0] == 4.0 and seen[2][2] == 1.95  # SIRT: rho_tau 1.95
    assert oreg.GradReg(3, 0.1).tau() == 6.0  # 3-D: 2 ndims
    assert oreg.GradReg(2, 0.1).tau() == 4.0


def test_tv_dual_step_is_one_half():
    # forward differences with the last one 0: rows [[3, 4], [0, 0]], columns [[1, 0], [2, 0]]
    xbar = np.array([[0, 1], [3, 5]], F32)
    q = osol.tv_dual(np.zeros((2, 2, 2), F32), xbar, 100.0)
    np.testing.assert_array_equal(q, F32(0.5) * np.array([[[3, 4], [0, 0]], [[1, 0], [2, 0]]], F32))
    # onto the weight ball per pixel: (1.5, 0.5) has norm sqrt(2.5) > 1
    q1 = osol.tv_dual(np.zeros((2, 2, 2), F32), xbar, 1.0)
    np.testing.assert_allclose(q1[:, 0, 0], np.array([1.5, 0.5]) / np.sqrt(2.5), rtol=1e-6)
    np.testing.assert_allclose(q1[:, 1, 0], [0.0, 1.0], rtol=1e-6)
