"""Oracle SU(3) exponential and RK3 Wilson flow.

Exponential (GLAC/math/exponentiation/su3exp.cpp:16-79): scipy's expm on
random su(3) over eight decades of norm, closed form for diagonal
arguments, both signs of c0, the |w| = 0.05 branch point, Q = 0.
Flow (GLAC/flow/flow.cpp:43-73): third-order convergence (a wrong RK3
coefficient drops the order), pure gauge stays pure gauge, E(t) decreasing,
gauge invariance of E(t) and Q(t)."""
import numpy as np
import pytest
import scipy.linalg

import oracle as O
from fieldtools import dag, gauge_transform, near_unit_field, random_field, random_su3


def random_su3_algebra(rng, scale):
    h = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
    X = (h - dag(h)) / 2
    X -= np.trace(X) / 3 * np.eye(3)
    return scale * X / np.linalg.norm(X)


@pytest.mark.parametrize("scale", [1e-7, 1e-4, 1e-2, 0.1, 0.5, 1.0, 3.0])
def test_exp_matches_scipy_expm(scale):
    rng = np.random.default_rng(int(scale * 1e7) % 1000 + 5)
    for _ in range(30):
        X = random_su3_algebra(rng, scale)
        got = O.to_complex(O.su3_exp(O.to_pairs(X)))
        assert np.abs(got - scipy.linalg.expm(X)).max() < 2e-14
        assert abs(np.linalg.det(got) - 1) < 1e-14
        assert np.abs(got @ dag(got) - np.eye(3)).max() < 1e-14


def test_exp_diagonal_closed_form_both_signs_of_c0():
    for a, b in [(0.3, 0.5), (-0.3, -0.5), (0.7, -0.1), (1e-3, 2e-3)]:
        X = np.diag([1j * a, 1j * b, -1j * (a + b)])
        got = O.to_complex(O.su3_exp(O.to_pairs(X)))
        assert np.abs(got - np.diag(np.exp(np.diag(X)))).max() < 1e-15 * 10


def test_exp_continuous_across_w_branch():
    """w = sqrt(c1) sin(theta/3) crosses 0.05; compare with expm on both sides."""
    for a in [0.0499, 0.05, 0.0501, 0.04, 0.06]:
        # eigenvalues of H = -iX: (u + w?) -> choose diag(a, -a, 0): theta = pi/2
        X = np.diag([1j * a, -1j * a, 0.0])
        # rotate to make it non-diagonal
        R = random_su3(np.random.default_rng(3))
        Xr = R @ X @ dag(R)
        got = O.to_complex(O.su3_exp(O.to_pairs(Xr)))
        assert np.abs(got - scipy.linalg.expm(Xr)).max() < 1e-14


def test_exp_of_zero_is_identity():
    got = O.su3_exp(np.zeros(18))
    assert np.array_equal(got, O.to_pairs(np.eye(3)))


def test_taylor_comparator_agrees():
    rng = np.random.default_rng(9)
    X = random_su3_algebra(rng, 0.8)
    assert np.abs(O.su3_exp_taylor(O.to_pairs(X), 30) - O.su3_exp(O.to_pairs(X))).max() < 1e-14


N4 = (4, 4, 4, 4)


def flowed(U0, n, eps, steps):
    U = U0.copy()
    for _ in range(steps):
        O.flow_step(U, n, eps)
    return U


def test_flow_is_third_order():
    rng = np.random.default_rng(21)
    U0 = near_unit_field(N4, rng, 0.6)
    T = 0.24
    ref = flowed(U0, N4, T / 96, 96)
    errs = []
    for steps in (6, 12):
        U = flowed(U0, N4, T / steps, steps)
        errs.append(np.abs(U - ref).max())
    ratio = errs[0] / errs[1]
    assert 6.0 < ratio < 10.5, (errs, ratio)  # 2^3 = 8 for a third-order scheme


def test_flow_time_scale_matches_an_independent_integrator():
    """The RK3 scheme integrates dV/dt = Z(V) V over exactly eps per step
    (GLAC/flow/flow.cpp:43-73): four steps of 0.0125 against 2000 Lie-Euler
    steps V <- exp(dt Z(V)) V of the same ODE, with Z from the oracle's
    action derivative and the exponential from its eigendecomposition.  The
    third-order test alone leaves the scale open (a flow of 2 eps per step, or
    of -eps, converges at the same order); here such a flow misses by ~0.1."""
    rng = np.random.default_rng(24)
    U0 = near_unit_field(N4, rng, 0.6)
    T, n_rk, n_euler = 0.05, 4, 2000
    U = flowed(U0, N4, T / n_rk, n_rk)
    V = O.to_complex(U0).reshape(4, -1, 3, 3)
    dt = T / n_euler
    for _ in range(n_euler):
        Z = np.stack([O.to_complex(O.action_derivative(O.to_pairs(V), N4, mu)).reshape(-1, 3, 3)
                      for mu in range(4)])
        # exp(dt Z), Z antihermitian: Z = i H with H hermitian
        w, v = np.linalg.eigh(-1j * dt * Z)
        V = (v * np.exp(1j * w)[..., None, :]) @ dag(v) @ V
    err = np.abs(O.to_complex(U).reshape(4, -1, 3, 3) - V).max()
    wrong = flowed(U0, N4, 2 * T / n_rk, n_rk)
    err_wrong = np.abs(O.to_complex(wrong).reshape(4, -1, 3, 3) - V).max()
    assert err < 3e-5, err  # Lie-Euler: 1e-5 at 2000 steps, halving with each doubling
    assert err_wrong > 100 * err, (err, err_wrong)


def test_flow_keeps_pure_gauge_and_unitarity():
    rng = np.random.default_rng(22)
    U = gauge_transform(O.unit_field(N4), N4, random_su3(rng, (4, 4, 4, 4)))
    for _ in range(3):
        O.flow_step(U, N4, 0.02)
    plaq, e, q, qt = O.measure(U, N4)
    assert abs(plaq - 1) < 1e-13 and abs(e) < 1e-12 and abs(q) < 1e-13


def test_flow_decreases_wilson_action_and_links_stay_su3():
    """The Wilson flow is the gradient flow of the Wilson action, so
    1 - plaquette decreases monotonically (the clover E need not on a rough
    field: the four leaves first align, which grows the clover)."""
    rng = np.random.default_rng(23)
    U = random_field(N4, rng)
    plaqs = [O.measure(U, N4)[0]]
    for _ in range(8):
        O.flow_step(U, N4, 0.02)
        plaqs.append(O.measure(U, N4)[0])
    assert np.all(np.diff(plaqs) > 0)
    M = O.to_complex(U)
    assert np.abs(M @ dag(M) - np.eye(3)).max() < 1e-12
    assert np.abs(np.linalg.det(M) - 1).max() < 1e-12


def test_flow_observables_are_gauge_invariant():
    rng = np.random.default_rng(24)
    U = random_field(N4, rng)
    Ug = gauge_transform(U, N4, random_su3(rng, (4, 4, 4, 4)))
    for _ in range(3):
        O.flow_step(U, N4, 0.02)
        O.flow_step(Ug, N4, 0.02)
    a, b = O.measure(U, N4), O.measure(Ug, N4)
    assert abs(a[0] - b[0]) < 1e-12 and abs(a[1] - b[1]) < 1e-10 and abs(a[2] - b[2]) < 1e-12
    assert np.abs(a[3] - b[3]).max() < 1e-12
