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


def test_training_acceptance_model_predicts_the_plain_samplers_outcomes():
    """The training objective weights each step by alpha_2 = min(1,
    exp(-dS_2)), dS_2 the second-order expansion of dS in the step's
    generator (oracle/training.py).  Checked against the accept/reject
    outcomes the plain sampler records: GLAC's RST proposals (ε = 0.24, 10
    hits per link; GLAC/system.cpp:333-361) on a thermalised 4^4 field at
    beta = 6, each proposal X evaluated with its generator Y = log X and the
    link's W = U A.  The model's mean acceptance equals the recorded
    acceptance rate, and binned by alpha_2 the recorded rate follows the
    model: within 0.03 in every decile (about 0.02 below it where
    acceptance is near 1, the expansion's third-order error).  About seven in
    ten RST proposals lie far from the identity (GLAC keeps sign(r0), so an
    SU(2) factor may sit near -1): there dS and dS_2 are both large and both
    reject."""
    n = (4, 4, 4, 4)
    beta, eps = 6.0, 0.24
    U = O.hot_start_haar(n, 3)
    for s in range(30):
        O.cb_sweep(U, n, beta, eps, 10, 3, s)
    rng = np.random.default_rng(7)
    rec_A, rec_U, rec_X, acc = [], [], [], []
    # GLAC's order: x outermost, then y, z, t, the four links of a site
    # innermost, 10 hits each
    sites = [(x, y, z, t) for x in range(n[0]) for y in range(n[1]) for z in range(n[2]) for t in range(n[3])]
    for _sweep in range(4):
        for x, y, z, t in sites:
            for mu in range(4):
                s = x + n[0] * (y + n[1] * (z + n[2] * t))
                A = O.staple(U, n, x, y, z, t, mu)
                u = U[mu, s].copy()
                for _ in range(10):
                    X = O.generate_rst(rng.uniform(-0.5, 0.5, 13), eps)
                    up = O.su3_mul(X, u)
                    accepted = np.exp(-O.delta_action(u, up, A, beta)) > rng.uniform()
                    rec_A.append(A)
                    rec_U.append(u.copy())
                    rec_X.append(X)
                    acc.append(accepted)
                    if accepted:
                        u = up
                U[mu, s] = u
    A = O.to_complex(np.array(rec_A))
    Uc = O.to_complex(np.array(rec_U))
    X = O.to_complex(np.array(rec_X))
    acc = np.array(acc, dtype=float)
    # generator of each proposal: X = V diag(exp(i theta)) V^-1, Y = V diag(i theta) V^-1
    lam, vec = np.linalg.eig(X)
    Y = vec @ (1j * np.angle(lam)[..., None] * np.linalg.inv(vec))
    W = Uc @ A
    lin = np.trace(Y @ W, axis1=-2, axis2=-1).real
    quad = np.trace(Y @ Y @ W, axis1=-2, axis2=-1).real
    alpha2 = np.minimum(1.0, np.exp((beta / 3.0) * (lin + 0.5 * quad)))
    sigma = np.sqrt(acc.mean() * (1 - acc.mean()) / acc.size)
    assert abs(alpha2.mean() - acc.mean()) < 3 * sigma, (alpha2.mean(), acc.mean(), sigma)
    edges = np.quantile(alpha2, np.linspace(0, 1, 11))
    for i in range(10):
        m = (alpha2 >= edges[i]) & (alpha2 <= edges[i + 1])
        print(f"decile {i}: model {alpha2[m].mean():.4f}, recorded {acc[m].mean():.4f}")
        assert abs(alpha2[m].mean() - acc[m].mean()) < 0.03, (i, alpha2[m].mean(), acc[m].mean())
    print(f"proposals {acc.size}: recorded acceptance {acc.mean():.4f} +- {sigma:.4f}, model {alpha2.mean():.4f}; "
          f"alpha_2 < 1e-6 for {np.mean(alpha2 < 1e-6):.3f}")
    assert 0.6 < np.mean(alpha2 < 1e-6) < 0.8      # the proposals far from the identity
