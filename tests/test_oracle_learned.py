"""Oracle of the learned proposal: the su(3) algebra map against the
Gell-Mann matrices, projection as its adjoint, gauge invariance of the staple
features, the training gradient against central differences, and the
learned-proposal sweep against the exact single-link density and GLAC's
recorded plaquette."""
import numpy as np
import pytest

import oracle as O
import oracle.training as T
from fieldtools import random_su3

GELLMANN = np.zeros((8, 3, 3), dtype=np.complex128)
GELLMANN[0][0, 1] = GELLMANN[0][1, 0] = 1
GELLMANN[1][0, 1], GELLMANN[1][1, 0] = -1j, 1j
GELLMANN[2][0, 0], GELLMANN[2][1, 1] = 1, -1
GELLMANN[3][0, 2] = GELLMANN[3][2, 0] = 1
GELLMANN[4][0, 2], GELLMANN[4][2, 0] = -1j, 1j
GELLMANN[5][1, 2] = GELLMANN[5][2, 1] = 1
GELLMANN[6][1, 2], GELLMANN[6][2, 1] = -1j, 1j
GELLMANN[7] = np.diag([1, 1, -2]) / np.sqrt(3)


def test_algebra_map_is_i_w_lambda_over_two():
    rng = np.random.default_rng(1)
    for _ in range(10):
        w = rng.normal(size=8)
        want = 0.5j * np.einsum("a,aij->ij", w, GELLMANN)
        assert np.allclose(O.to_complex(O.algebra_from_omega(w)), want, atol=1e-15)
        assert np.allclose(T.algebra(w), want, atol=1e-15)


def test_projection_is_adjoint_of_algebra_map():
    # Re tr(i T_a i T_b) = -delta_ab / 2, so project(Y(w)) = -w / 2
    rng = np.random.default_rng(2)
    w = rng.normal(size=8)
    assert np.allclose(O.project_algebra(O.algebra_from_omega(w)), -0.5 * w, atol=1e-15)
    N = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
    want = np.array([np.trace(0.5j * GELLMANN[a] @ N).real for a in range(8)])
    assert np.allclose(O.project_algebra(O.to_pairs(N)), want, atol=1e-14)
    assert np.allclose(T.project(N), want, atol=1e-14)


def test_features_are_gauge_invariant():
    rng = np.random.default_rng(3)
    A = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
    g, h = random_su3(rng), random_su3(rng)
    fa = O.staple_features(O.to_pairs(A), 6.1)
    fb = O.staple_features(O.to_pairs(g @ A @ h.conj().T), 6.1)
    assert np.allclose(fa, fb, rtol=1e-12, atol=1e-12)
    assert np.allclose(fa, T.features(A[None], 6.1)[0])


def test_training_gradient_matches_central_differences():
    rng = np.random.default_rng(4)
    S, K = 16, 3
    A = np.array([sum(random_su3(rng) for _ in range(6)) for _ in range(S)])
    U = random_su3(rng, (S,))
    Z = rng.normal(size=(S, K, 8))
    p = T.init_params(rng, scale0=0.3)
    p[T.SLICES["W3"]] += rng.normal(size=36 * 32) * 0.05
    f = T.features(A, 6.0)
    mean, std = f.mean(0), f.std(0) + 1.0
    loss, g = T.loss_and_grad(p, mean, std, A, U, 6.0, Z)
    idx = rng.choice(T.NPARAM, size=40, replace=False)
    for i in list(idx) + [T.SLICES["b3"].start, T.SLICES["b3"].start + 9, T.SLICES["W1"].start]:
        h = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        num = (T.loss_and_grad(pp, mean, std, A, U, 6.0, Z)[0] - T.loss_and_grad(pm, mean, std, A, U, 6.0, Z)[0]) / (2 * h)
        assert abs(num - g[i]) < 1e-6 + 1e-4 * abs(num), (i, num, g[i])


def test_learned_single_link_chain_samples_exact_density():
    """The learned proposal in a fixed staple A = kappa 1 (one link, nothing
    else moving) samples exp(beta kappa / 3 Re tr U) whatever the network."""
    rng = np.random.default_rng(5)
    beta, kappa = 6.0, 0.9
    H = random_su3(rng, (400000,))
    tr = np.trace(H, axis1=-2, axis2=-1).real
    wts = np.exp(beta * kappa / 3 * tr)
    exact = (wts * tr).sum() / wts.sum()
    p = T.init_params(rng, scale0=0.4)
    mean, std = np.zeros(6), np.ones(6)
    A = O.to_pairs(kappa * np.eye(3))
    L = O.cholesky_from_output(O.mlp_forward(p, mean, std, O.staple_features(A, beta)))
    U = O.to_pairs(np.eye(3))
    vals = []
    for it in range(40000):
        z = rng.uniform(-0.5, 0.5, 8) * np.sqrt(12)
        X = O.su3_exp(O.algebra_from_omega(L @ z))
        Up = O.su3_mul(X, U)
        if np.exp(-O.delta_action(U, Up, A, beta)) > rng.uniform():
            U = Up
        if it >= 1000 and it % 4 == 0:
            vals.append(U[0] + U[8] + U[16])
    vals = np.array(vals)
    blocks = vals[: len(vals) // 40 * 40].reshape(40, -1).mean(axis=1)
    err = blocks.std(ddof=1) / np.sqrt(len(blocks))
    assert abs(vals.mean() - exact) < 4 * err + 0.01, (vals.mean(), exact, err)


def test_learned_sweep_reproduces_recorded_plaquette():
    """A constant isotropic proposal (network weights zero, output bias only)
    through the full lattice sweep reproduces GLAC's recorded plaquette."""
    n = (4, 4, 4, 8)
    p = np.zeros(T.NPARAM)
    p[T.SLICES["b3"].start:T.SLICES["b3"].start + 8] = np.log(0.35)
    mean, std = np.zeros(6), np.ones(6)
    U = O.unit_field(n)
    for s in range(150):
        O.learned_cb_sweep(U, n, 6.0, 10, 9, s, p, mean, std)
    acc, pl = [], []
    for s in range(150, 450):
        acc.append(O.learned_cb_sweep(U, n, 6.0, 10, 9, s, p, mean, std))
        pl.append(O.measure(U, n)[0])
    assert abs(np.mean(pl) - 0.59586) < 0.008
    assert 0.05 < np.mean(acc) < 0.95


def test_mlp_rounds_layer_inputs_to_bf16():
    """An identity network (W1, W2, W3 = unit rows, biases 0) passes input 0
    through three bf16 roundings: 1 + 3/512 lies between the bf16 values 1
    and 1 + 1/128 and rounds to the nearer, 1 + 1/128 (truncation would give
    1, no rounding 1 + 3/512); a negative input goes through the leaky ReLU
    (slope 0.01) twice before each rounding."""
    P = np.zeros(2468)
    W1, b1 = 0, 192
    W2, b2 = 224, 1248
    W3, b3 = 1280, 2432
    for i in range(6):
        P[W1 + i * 6 + i] = 1.0
    for i in range(32):
        P[W2 + i * 32 + i] = 1.0
        P[W3 + i * 32 + i] = 1.0
    f = np.array([1 + 3 / 512, -1.0, 0.0, 0.0, 0.0, 0.0])
    o = O.mlp_forward(P, np.zeros(6), np.ones(6), f)
    assert o[0] == 1 + 1 / 128
    # -1 -> lrelu -0.01 -> bf16 -> lrelu -> bf16: compare with the two-step rounding
    def bf16(v):
        u = np.array([v], dtype=np.float32).view(np.uint32).astype(np.uint64)
        return float(((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32).view(np.float32)[0])
    assert o[1] == bf16(0.01 * bf16(-0.01))
    assert o[1] != 0.01 * -0.01
