"""Oracle Metropolis samplers (GLAC/system.cpp:321-361) and hot starts.

Pins: GLAC's own recorded run at 4^3x8, beta = 6, eps = 0.24, 10 hits,
cold start (examples/test_config/configExampleLoad.out:58-73: acceptance
0.1079, per-config plaquettes with mean 0.5959) for both the sequential
GLAC-order sampler and the checkerboard sampler; the strong-coupling
expansion <P> = beta/18 + beta^2/216; a single-link chain against the exact
density exp(beta/3 Re tr(U A)); Haar moments of the hot start."""
import numpy as np
import pytest

import oracle as O
from fieldtools import random_su3

GLAC_PLAQ = np.array([0.60353658, 0.58615566, 0.60581965, 0.60296651, 0.59396066,
                      0.59986059, 0.59454736, 0.59676180, 0.59096302, 0.58405070])
GLAC_ACCEPT = 0.1079086669921875
N48 = (4, 4, 4, 8)


def test_glac_order_sampler_reproduces_recorded_run():
    U = O.unit_field(N48)
    g = O.GlacRng(1001, 1000)
    for _ in range(150):
        O.glac_sweep(U, N48, g, 6.0, 0.24, 10)
    acc, pl = [], []
    for _ in range(300):
        acc.append(O.glac_sweep(U, N48, g, 6.0, 0.24, 10))
        pl.append(O.measure(U, N48)[0])
    assert abs(np.mean(acc) - GLAC_ACCEPT) < 0.004
    assert abs(np.mean(pl) - GLAC_PLAQ.mean()) < 0.008


def test_checkerboard_sampler_reproduces_recorded_run():
    U = O.unit_field(N48)
    for s in range(150):
        O.cb_sweep(U, N48, 6.0, 0.24, 10, 77, s)
    acc, pl = [], []
    for s in range(150, 450):
        acc.append(O.cb_sweep(U, N48, 6.0, 0.24, 10, 77, s))
        pl.append(O.measure(U, N48)[0])
    assert abs(np.mean(acc) - GLAC_ACCEPT) < 0.004
    assert abs(np.mean(pl) - GLAC_PLAQ.mean()) < 0.008


@pytest.mark.parametrize("sampler", ["glac", "cb"])
def test_strong_coupling_plaquette(sampler):
    beta = 0.6
    n = (4, 4, 4, 4)
    U = O.hot_start_haar(n, 5)
    g = O.GlacRng(3, 4)
    pl = []
    for s in range(260):
        if sampler == "glac":
            O.glac_sweep(U, n, g, beta, 0.24, 10)
        else:
            O.cb_sweep(U, n, beta, 0.24, 10, 9, s)
        if s >= 60:
            pl.append(O.measure(U, n)[0])
    expect = beta / 18 + beta ** 2 / 216
    assert abs(np.mean(pl) - expect) < 0.0025, (np.mean(pl), expect)


def test_single_link_chain_samples_exact_density():
    """Metropolis with GLAC's RST proposal on one link in a fixed staple
    A = kappa 1 samples exp(beta kappa / 3 Re tr U); compare <Re tr U> with
    the exact expectation by reweighted Haar samples."""
    rng = np.random.default_rng(41)
    beta, kappa, eps = 6.0, 0.9, 0.24
    A = O.to_pairs(kappa * np.eye(3))
    # exact expectation by importance weights on Haar samples
    H = random_su3(rng, (400000,))
    tr = np.trace(H, axis1=-2, axis2=-1).real
    w = np.exp(beta * kappa / 3 * tr)
    exact = (w * tr).sum() / w.sum()
    # chain
    U = O.to_pairs(np.eye(3))
    vals = []
    for it in range(60000):
        X = O.generate_rst(rng.uniform(-0.5, 0.5, size=13), eps)
        Up = O.su3_mul(X, U)
        if np.exp(-O.delta_action(U, Up, A, beta)) > rng.uniform():
            U = Up
        if it >= 2000 and it % 5 == 0:
            vals.append(U[0] + U[8] + U[16])
    vals = np.array(vals)
    # error with blocking over autocorrelation
    blocks = vals[: len(vals) // 50 * 50].reshape(50, -1).mean(axis=1)
    err = blocks.std(ddof=1) / np.sqrt(len(blocks))
    assert abs(vals.mean() - exact) < 4 * err + 0.01, (vals.mean(), exact, err)


def test_haar_hot_start_moments():
    n = (8, 8, 8, 8)
    U = O.hot_start_haar(n, 2024)
    M = O.to_complex(U).reshape(-1, 3, 3)
    assert np.abs(M @ np.conj(np.swapaxes(M, -1, -2)) - np.eye(3)).max() < 1e-14
    assert np.abs(np.linalg.det(M) - 1).max() < 1e-14
    tr = np.trace(M, axis1=-2, axis2=-1)
    N = len(tr)
    assert abs(tr.mean()) < 4 / np.sqrt(N)
    assert abs((np.abs(tr) ** 2).mean() - 1.0) < 5 / np.sqrt(N)      # <|tr U|^2> = 1
    assert abs((np.abs(tr) ** 4).mean() - 2.0) < 12 / np.sqrt(N)     # <|tr U|^4> = 2


def test_glac_random_hot_start_is_unitary_with_det_minus_one():
    n = (4, 4, 4, 4)
    U = O.new_field(n)
    O.glac_hot_start(U, n, O.GlacRng(1, 2), 0.24, rst=False)
    M = O.to_complex(U).reshape(-1, 3, 3)
    assert np.abs(M @ np.conj(np.swapaxes(M, -1, -2)) - np.eye(3)).max() < 1e-14
    assert np.abs(np.linalg.det(M) + 1).max() < 1e-14


def _moments(U):
    M = O.to_complex(U).reshape(-1, 3, 3)
    tr = np.trace(M, axis1=-2, axis2=-1)
    return M, np.array([tr.real.mean(), (np.abs(tr) ** 2).mean(), (M[:, 0, 0].real ** 2).mean()]), len(tr)


@pytest.mark.parametrize("rst", [False, True], ids=["glac_random", "rst"])
def test_stream_hot_starts_sample_glacs_distributions(rst):
    """The device's GLAC-kind hot starts (link streams) against GLAC's own
    generators on its mt19937_64 stream (or_glac_hot_start): same law, so the
    same moments within statistical error; unitarity and det = -1 / +1."""
    n = (8, 8, 8, 8)
    kind = O.HOT_RST if rst else O.HOT_GLAC_RANDOM
    A = O.hot_start(n, 2025, kind, 0.24)
    B = O.new_field(n)
    O.glac_hot_start(B, n, O.GlacRng(5, 6), 0.24, rst=rst)
    Ma, ma, N = _moments(A)
    Mb, mb, _ = _moments(B)
    assert np.abs(Ma @ np.conj(np.swapaxes(Ma, -1, -2)) - np.eye(3)).max() < 1e-14
    assert np.abs(np.linalg.det(Ma) - (1.0 if rst else -1.0)).max() < 1e-14
    assert np.all(np.abs(ma - mb) < 6 / np.sqrt(N)), (ma, mb)


def test_stream_haar_kind_is_the_haar_start():
    n = (4, 4, 4, 6)
    assert np.array_equal(O.hot_start(n, 9, O.HOT_HAAR), O.hot_start_haar(n, 9))
