"""Oracle 3x3 algebra and GLAC's random-matrix generators against the known
answers of GLAC's own unit tests (GLAC/tests/unit_tests/testcore.cpp:31-166,
su3operations.cpp:48-51, 309-326, randommatrixtests.cpp:10-46) and closed
forms."""
import numpy as np
import pytest

import oracle as O


def cm(rows):
    """3x3 (or 2x2) complex matrix from rows of (re, im) pairs."""
    return np.array([[complex(*z) for z in r] for r in rows])


# Fixture values of testcore.cpp:31-166 (integer matrices).
U1 = cm([[(1, 1), (1, 2), (1, 3)], [(2, 1), (2, 2), (2, 3)], [(3, 1), (3, 2), (3, 3)]])
U2 = cm([[(4, 4), (4, 5), (4, 6)], [(5, 4), (5, 5), (5, 6)], [(6, 4), (6, 5), (6, 6)]])
UMUL = cm([[(-9, 44), (-15, 47), (-21, 50)], [(6, 56), (0, 62), (-6, 68)],
           [(21, 68), (15, 77), (9, 86)]])
UCT = cm([[(1, -1), (2, -1), (3, -1)], [(1, -2), (2, -2), (3, -2)], [(1, -3), (2, -3), (3, -3)]])
UTRACE = cm([[(4, 4), (5, 5), (2, 6)], [(6, 4), (5, 5), (1, 6)], [(6, 4), (9, 5), (2, 6)]])
U_RST = cm([[(-8, 9), (-123, 82), (-78, 83)], [(-20, 37), (-391, 342), (-234, 319)],
            [(2, 6), (48, 58), (44, 35)]])
S_R = cm([[(1, 2), (3, 4)], [(5, 6), (7, 8)]])
S_S = cm([[(2, 5), (4, 8)], [(2, 6), (10, 3)]])
S_T = cm([[(7, 2), (6, 1)], [(6, 4), (5, 2)]])


def su2_pairs(m):
    out = np.empty(8)
    out[0::2] = m.real.ravel()
    out[1::2] = m.imag.ravel()
    return out


def test_su3_multiplication_fixture():
    got = O.su3_mul(O.to_pairs(U1), O.to_pairs(U2))
    assert np.array_equal(got, O.to_pairs(UMUL))


def test_su3_dagger_fixture():
    assert np.array_equal(O.su3_dag(O.to_pairs(U1)), O.to_pairs(UCT))


def test_trace_real_multiplication_fixture():
    assert O.su3_retrace_mul(O.to_pairs(U1), O.to_pairs(UTRACE)) == -21.0


def test_determinant_matches_numpy():
    rng = np.random.default_rng(1)
    for _ in range(20):
        m = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
        assert abs(O.su3_det(O.to_pairs(m)) - np.linalg.det(m)) < 1e-12


def test_rst_multiplication_fixture():
    got = O.rst_mul(su2_pairs(S_R), su2_pairs(S_S), su2_pairs(S_T))
    assert np.array_equal(got, O.to_pairs(U_RST))


def test_rst_inverse_multiplication_fixture():
    got = O.rst_mul(su2_pairs(S_R), su2_pairs(S_S), su2_pairs(S_T), inverse=True)
    assert np.array_equal(got, O.to_pairs(U_RST.conj().T))


PAULI = [np.array([[0, 1], [1, 0]]), np.array([[0, -1j], [1j, 0]]), np.array([[1, 0], [0, -1]])]


@pytest.mark.parametrize("r", [[0.3, -0.1, 0.2, 0.4], [-0.2, 0.49, -0.3, 0.01], [0.0, 0.1, 0.1, 0.1]])
def test_generate_su2_closed_form(r):
    eps = 0.24
    got = O.generate_su2(r, eps)
    m = got[0::2] + 1j * got[1::2]
    m = m.reshape(2, 2)
    x0 = np.sqrt(1 - eps * eps) * (-1.0 if r[0] < 0 else 1.0)
    xv = eps * np.array(r[1:]) / np.linalg.norm(r[1:])
    expect = x0 * np.eye(2) + 1j * sum(x * s for x, s in zip(xv, PAULI))
    assert np.allclose(m, expect, atol=1e-15)
    assert np.allclose(m @ m.conj().T, np.eye(2), atol=1e-15)
    assert abs(np.linalg.det(m) - 1) < 1e-15


def test_generate_rst_is_su3_and_near_sign_matrices():
    rng = np.random.default_rng(2)
    eps = 0.24
    dets, dists = [], []
    for _ in range(400):
        u = rng.uniform(-0.5, 0.5, size=13)
        X = O.to_complex(O.generate_rst(u, eps))
        assert np.allclose(X @ X.conj().T, np.eye(3), atol=1e-14)
        dets.append(np.linalg.det(X))
        # X is close to a diagonal +-1 matrix: off-diagonals are O(eps)
        dists.append(np.abs(X - np.diag(np.diag(X))).max())
    assert np.allclose(dets, 1.0, atol=1e-14)
    assert max(dists) < 3 * eps
    # u[12] < 0 gives the inverse product: X(u with coin flipped) = X(u)^dagger
    u = rng.uniform(-0.5, 0.5, size=13)
    u[12] = 0.3
    a = O.to_complex(O.generate_rst(u, eps))
    u[12] = -0.3
    b = O.to_complex(O.generate_rst(u, eps))
    assert np.allclose(b, a.conj().T, atol=1e-15)


def test_generate_random_has_det_minus_one():
    rng = np.random.default_rng(3)
    for _ in range(50):
        u = rng.uniform(-0.24, 0.24, size=18)
        H = O.to_complex(O.generate_random(u))
        assert np.allclose(H @ H.conj().T, np.eye(3), atol=1e-14)
        assert abs(np.linalg.det(H) + 1.0) < 1e-14
        c0 = (u[0::2] + 1j * u[1::2]).reshape(3, 3)[:, 0]
        assert np.allclose(H[:, 0], c0 / np.linalg.norm(c0), atol=1e-15)


def test_reunitarize_gives_su3_and_is_idempotent():
    rng = np.random.default_rng(4)
    for _ in range(50):
        m = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
        R = O.to_complex(O.reunitarize(O.to_pairs(m)))
        assert np.allclose(R @ R.conj().T, np.eye(3), atol=1e-14)
        assert abs(np.linalg.det(R) - 1.0) < 1e-14
        # first row is the normalised first row of the input
        assert np.allclose(R[0], m[0] / np.linalg.norm(m[0]), atol=1e-14)
        R2 = O.to_complex(O.reunitarize(O.to_pairs(R)))
        assert np.allclose(R2, R, atol=1e-14)


def test_reunitarize_keep_det_sign():
    """det = -1 links (GLAC's generateRandom start) stay det = -1 and close
    to the input; det = +1 links get or_reunitarize's result exactly."""
    rng = np.random.default_rng(5)
    for _ in range(50):
        H = O.to_complex(O.generate_random(rng.uniform(-0.24, 0.24, size=18)))
        noisy = H + 1e-6 * (rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3)))
        R = O.to_complex(O.reunitarize_keep_det_sign(O.to_pairs(noisy)))
        assert np.allclose(R @ R.conj().T, np.eye(3), atol=1e-14)
        assert abs(np.linalg.det(R) + 1.0) < 1e-14
        assert np.abs(R - H).max() < 1e-5
        m = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
        S = O.to_complex(O.reunitarize(O.to_pairs(m)))
        S_noisy = O.to_pairs(S + 1e-6 * rng.normal(size=(3, 3)))
        assert np.array_equal(O.reunitarize_keep_det_sign(S_noisy), O.reunitarize(S_noisy))
