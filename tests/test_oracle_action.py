"""Oracle Wilson staple, dS and flow force Z.

Pins: GLAC's dS = -36 test (GLAC/tests/unit_tests/actiontests.cpp:67-114),
pure-gauge staple, gauge covariance, dS against the brute-force change of
the Wilson action, Z against GLAC's explicit Gell-Mann projection
(GLAC/actions/wilsonexplicitder.cpp:88-107) and against the finite-difference
derivative of the action."""
import numpy as np

import oracle as O
from fieldtools import (dag, gauge_transform, random_field, random_su3, to_grid, shift,
                        wilson_action)

N4 = (4, 4, 4, 4)


def test_delta_action_is_minus_36_on_unit_field():
    U = O.unit_field(N4)
    two = np.zeros(18)
    two[[0, 8, 16]] = 2.0
    for x, y, z, t, mu in [(0, 0, 0, 0, 0), (1, 2, 3, 0, 1), (3, 3, 3, 3, 3), (2, 0, 1, 3, 2)]:
        A = O.staple(U, N4, x, y, z, t, mu)
        site = x + 4 * (y + 4 * (z + 4 * t))
        assert abs(O.delta_action(U[mu, site], two, A, 6.0) + 36.0) <= 1e-15


def test_staple_of_pure_gauge_is_six_times_link_dagger():
    rng = np.random.default_rng(11)
    g = random_su3(rng, (4, 4, 4, 4))
    U = gauge_transform(O.unit_field(N4), N4, g)
    for mu in range(4):
        for (x, y, z, t) in [(0, 1, 2, 3), (3, 0, 0, 1)]:
            A = O.to_complex(O.staple(U, N4, x, y, z, t, mu))
            site = x + 4 * (y + 4 * (z + 4 * t))
            Ul = O.to_complex(U[mu, site])
            assert np.allclose(A, 6 * Ul.conj().T, atol=1e-13)


def test_staple_is_gauge_covariant():
    rng = np.random.default_rng(12)
    U = random_field(N4, rng)
    g = random_su3(rng, (4, 4, 4, 4))
    Ug = gauge_transform(U, N4, g)
    for mu in range(4):
        c = (1, 2, 3, 0)
        A = O.to_complex(O.staple(U, N4, *c, mu))
        Ag = O.to_complex(O.staple(Ug, N4, *c, mu))
        cp = list(c)
        cp[mu] = (cp[mu] + 1) % 4
        gx = g[c[3], c[2], c[1], c[0]]
        gxp = g[cp[3], cp[2], cp[1], cp[0]]
        # A_mu(x) -> g(x + mu) A_mu(x) g(x)^dagger
        assert np.allclose(Ag, gxp @ A @ gx.conj().T, atol=1e-12)


def test_delta_action_equals_brute_force_action_change():
    rng = np.random.default_rng(13)
    n = (4, 4, 4, 6)
    U = random_field(n, rng)
    beta = 5.7
    for mu, c in [(0, (1, 2, 3, 4)), (3, (0, 0, 0, 5)), (2, (3, 1, 0, 2))]:
        site = c[0] + n[0] * (c[1] + n[1] * (c[2] + n[2] * c[3]))
        A = O.staple(U, n, *c, mu)
        X = random_su3(rng)
        Unew = U.copy()
        Unew[mu, site] = O.to_pairs(X @ O.to_complex(U[mu, site]))
        dS = O.delta_action(U[mu, site], Unew[mu, site], A, beta)
        brute = wilson_action(Unew, n, beta) - wilson_action(U, n, beta)
        assert abs(dS - brute) < 1e-10


def explicit_derivative(Omega):
    """GLAC's WilsonExplicitDer projection (wilsonexplicitder.cpp:88-107),
    written over complex matrices: traceless antihermitian part with the
    opposite sign, element by element."""
    Z = np.zeros((3, 3), dtype=np.complex128)
    im = Omega.imag
    Z[0, 0] = 1j * (im[1, 1] - 2 * im[0, 0] + im[2, 2]) / 3
    Z[1, 1] = 1j * (im[0, 0] - 2 * im[1, 1] + im[2, 2]) / 3
    Z[2, 2] = 1j * (im[0, 0] + im[1, 1] - 2 * im[2, 2]) / 3
    for (i, j) in [(0, 1), (0, 2), (1, 2)]:
        Z[i, j] = 0.5 * (Omega[j, i].real - Omega[i, j].real) - 0.5j * (Omega[i, j].imag + Omega[j, i].imag)
        Z[j, i] = -Z[i, j].conjugate()
    return Z


def test_force_matches_explicit_gell_mann_projection():
    rng = np.random.default_rng(14)
    U = random_field(N4, rng)
    G = to_grid(U, N4)
    for mu in range(4):
        Z = O.to_complex(O.action_derivative(U, N4, mu)).reshape(4, 4, 4, 4, 3, 3)
        for c in [(0, 0, 0, 0), (1, 3, 2, 0), (3, 2, 1, 3)]:
            site = c[0] + 4 * (c[1] + 4 * (c[2] + 4 * c[3]))
            A = O.to_complex(O.staple(U, N4, *c, mu))
            Om = G[mu][c[3], c[2], c[1], c[0]] @ A
            assert np.allclose(Z[c[3], c[2], c[1], c[0]], explicit_derivative(Om), atol=1e-13)
            assert np.allclose(Z.reshape(-1, 3, 3)[site], explicit_derivative(Om), atol=1e-13)


def test_force_vanishes_on_pure_gauge():
    rng = np.random.default_rng(15)
    U = gauge_transform(O.unit_field(N4), N4, random_su3(rng, (4, 4, 4, 4)))
    for mu in range(4):
        assert np.abs(O.action_derivative(U, N4, mu)).max() < 1e-13


def test_force_is_derivative_of_wilson_action():
    """d/ds S_W(exp(s X) U_mu(x)) at s = 0 equals (beta/3) tr(X Z_mu(x)) for
    traceless antihermitian X; checked by central differences."""
    rng = np.random.default_rng(16)
    n = (4, 4, 4, 4)
    U = random_field(n, rng)
    beta = 3.0
    mu, c = 2, (1, 0, 3, 2)
    site = c[0] + 4 * (c[1] + 4 * (c[2] + 4 * c[3]))
    Z = O.to_complex(O.action_derivative(U, n, mu)[site])
    for _ in range(3):
        h = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
        X = (h - dag(h)) / 2
        X -= np.trace(X) / 3 * np.eye(3)
        s = 1e-5
        vals = []
        for sg in (+1, -1):
            Up = U.copy()
            w, v = np.linalg.eig(sg * s * X)
            E = v @ np.diag(np.exp(w)) @ np.linalg.inv(v)
            Up[mu, site] = O.to_pairs(E @ O.to_complex(U[mu, site]))
            vals.append(wilson_action(Up, n, beta))
        deriv = (vals[0] - vals[1]) / (2 * s)
        assert abs(deriv - (beta / 3) * np.trace(X @ Z).real) < 1e-6
