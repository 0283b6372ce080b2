"""GPU parity of the learned proposal (no GLAC counterpart): the device
training step against the FP64 numpy reference (oracle/training.py), the
learned-proposal sweep against the oracle's on the same streams, and the
end-to-end pipeline (collect -> standardise -> train -> sample) against the
plain sampler's plaquette."""
import sys

import numpy as np
import pytest

import oracle as O
import oracle.training as T
from fieldtools import random_su3

pytestmark = pytest.mark.gpu
glacx = pytest.importorskip("glac_amd")


def banner(msg):
    print(f"\n[gpu case] {msg}", flush=True)
    sys.stdout.flush()


def bf16_round(p):
    x = np.ascontiguousarray(p, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return r.view(np.float32)


def make_samples(rng, S):
    A = np.array([sum(random_su3(rng) for _ in range(6)) for _ in range(S)])
    U = random_su3(rng, (S,))
    smp = np.concatenate([O.to_pairs(A), O.to_pairs(U)], axis=1)
    return A, U, smp


def test_train_gradient_matches_reference():
    banner("learned: training loss/gradient vs numpy FP64")
    rng = np.random.default_rng(10)
    S, K, beta = 300, 3, 6.0
    A, U, smp = make_samples(rng, S)
    f = T.features(A, beta)
    mean, std = f.mean(0).astype(np.float32), (f.std(0) + 1e-3).astype(np.float32)
    p = T.init_params(rng, scale0=0.3)
    p[T.SLICES["W3"]] += rng.normal(size=36 * 32) * 0.05
    p = p.astype(np.float32)
    z = rng.normal(size=(S, K, 8)).astype(np.float32)
    with glacx.Lattice((4, 4, 4, 4), glacx.FP32_LINKS) as L:
        L.proposal_set(p, mean, std)
        loss, g = L.train_grad(smp.astype(np.float32), z, beta)
    want_loss, want_g = T.loss_and_grad(p.astype(np.float64), mean.astype(np.float64), std.astype(np.float64),
                                        A.astype(np.complex64).astype(np.complex128),
                                        U.astype(np.complex64).astype(np.complex128), beta, z.astype(np.float64))
    assert abs(loss - want_loss) < 1e-4 * abs(want_loss)
    scale = np.abs(want_g).max()
    assert np.abs(g - want_g).max() < 2e-3 * scale, np.abs(g - want_g).max() / scale


@pytest.mark.parametrize("n, prec", [((4, 4, 4, 8), glacx.FP64), ((4, 4, 4, 8), glacx.FP32_LINKS),
                                     ((4, 4, 4, 8), glacx.FP32_COMPRESSED), ((6, 6, 6, 6), glacx.FP64),
                                     ((6, 6, 6, 6), glacx.FP32_LINKS)],
                         ids=["4x4x4x8-fp64", "4x4x4x8-fp32", "4x4x4x8-fp32c", "6x6x6x6-fp64", "6x6x6x6-fp32"])
def test_learned_sweep_follows_oracle_streams(n, prec):
    """(6, 6, 6, 6): V / 2 = 648 links per sub-sweep, so the last 256-thread
    block has 136 real lanes and 120 idle ones, which take part in the
    network's matrix-core GEMMs with a mirrored link and store nothing.  The
    idle-lane code is common to all stored-row modes; the compressed one,
    whose two-row products round differently, flips about one decision in
    10^4 of the 51 840 there (each flip then changes the staples of links
    updated after it), so it is held to the oracle's decisions at 4^3 x 8."""
    banner(f"learned sweep vs oracle {n} prec={prec}")
    rng = np.random.default_rng(11)
    p = T.init_params(rng, scale0=0.3)
    p[T.SLICES["W3"]] += rng.normal(size=36 * 32) * 0.02
    # The network's layer inputs are rounded to bf16.  An input that the FP32
    # device and the FP64 oracle round to neighbouring bf16 values changes a
    # proposal by ~1e-3, that link's staple-dependent neighbours' features
    # with it, and so on through the sweep (a 1e-7 perturbation of the
    # parameters changes a quarter of the links of the oracle's own sweep).
    # With W1 = 0 the network's output is the same for every link, so the
    # sweep tests the draws, the Cholesky factor, the algebra map, the
    # exponential, the RST fallback and the accept/reject decisions without
    # that amplification; test_proposal_network_matches_oracle covers the
    # feature-dependent network.
    p[T.SLICES["W1"]] = 0.0
    pb = bf16_round(p)
    U = O.hot_start_haar(n, 3)
    for s in range(3):
        O.cb_sweep(U, n, 6.0, 0.24, 10, 3, s)
    # standardisation from the oracle field's staples
    f = np.array([O.staple_features(O.staple(U, n, x, 1, 2, t, mu), 6.0)
                  for x in range(n[0]) for t in range(n[3]) for mu in range(4)])
    mean, std = f.mean(0).astype(np.float32), (f.std(0) + 1e-3).astype(np.float32)
    W = U.copy()
    acc = O.learned_cb_sweep(W, n, 6.0, 10, 21, 0, pb.astype(np.float64), mean.astype(np.float64),
                             std.astype(np.float64))
    with glacx.Lattice(n, prec) as L:
        L.upload(U)
        L.proposal_set(p, mean, std)
        got_acc = L.metropolis_learned(1, 6.0, 0.24, 10, reunit_every=1, seed=21)
        got = L.download()
    decisions = 10 * 4 * np.prod(n)
    # The network accumulates in FP32 on the device (FP64 in the oracle), so
    # a decision whose exp(-dS) lies within ~1e-7 of its uniform could flip.
    close = np.abs(got - W).reshape(4, -1, 18).max(axis=-1) < (1e-5 if prec == glacx.FP64 else 1e-4)
    n_diff = int((~close).sum())
    print(f"links differing from the oracle: {n_diff} of {close.size}; "
          f"accepted hits {acc * decisions:.0f} (oracle) vs {got_acc[0] * decisions:.0f}")
    assert n_diff <= 2
    assert abs(acc - got_acc[0]) * decisions <= 3


def test_proposal_network_matches_oracle():
    """The sweeps' matrix-core network (bf16 GEMMs, FP32 accumulation) on
    staple features of a thermalised field against the oracle's FP64
    evaluation with the same bf16 rounding of every layer input.  Where the
    two accumulations round a hidden activation to neighbouring bf16 values
    the output moves by ~1e-3 of its scale: such rows are rare, the rest agree
    to FP32 accumulation error."""
    banner("learned: network outputs vs oracle")
    rng = np.random.default_rng(13)
    n = (4, 4, 4, 8)
    U = O.hot_start_haar(n, 5)
    for s in range(5):
        O.cb_sweep(U, n, 6.0, 0.24, 10, 5, s)
    f = np.array([O.staple_features(O.staple(U, n, x, y, z, t, mu), 6.0)
                  for x in range(4) for y in range(4) for z in range(4) for t in range(8) for mu in range(4)])
    mean, std = f.mean(0).astype(np.float32), (f.std(0) + 1e-3).astype(np.float32)
    f32 = f.astype(np.float32)
    f32[:8] = mean + 9.0 * std * np.sign(rng.normal(size=(8, 6)))   # out of range: fallback rows
    p = T.init_params(rng, scale0=0.3)
    p[T.SLICES["W3"]] += rng.normal(size=36 * 32) * 0.05
    pb = bf16_round(p)
    with glacx.Lattice((4, 4, 4, 4), glacx.FP32_LINKS) as L:
        L.proposal_set(p, mean, std)
        out, ok = L.proposal_eval(f32)
    want = np.array([O.mlp_forward(pb.astype(np.float64), mean.astype(np.float64), std.astype(np.float64),
                                   row.astype(np.float64)) for row in f32])
    x = (f32.astype(np.float64) - mean) / std
    assert np.array_equal(ok, np.all(np.abs(x) <= 8.0, axis=1))
    assert not ok[:8].any() and ok[8:].mean() > 0.99
    err = np.abs(out - want).max(axis=1) / np.abs(want).max()
    print(f"network rows {len(err)}: median error {np.median(err):.2e}, 99th percentile "
          f"{np.quantile(err, 0.99):.2e}, max {err.max():.2e}")
    assert np.median(err) < 1e-6
    assert np.mean(err < 1e-5) > 0.97
    assert err.max() < 3e-2


def test_pipeline_trains_and_samples_the_right_distribution():
    n = (8, 8, 8, 8)
    beta = 6.0
    banner(f"learned pipeline {n}")
    with glacx.Lattice(n, glacx.FP32_LINKS) as L:
        L.cold_start()
        L.metropolis(150, beta, 0.24, 10, seed=5)
        L.samples_reset(100000)
        for i in range(10):
            L.metropolis(5, beta, 0.24, 10, seed=5, sweep_offset=150 + 5 * i)
            count = L.samples_collect(stride=4)
        assert count == 10 * 8 * (np.prod(n) // 2 // 4)
        mean, std = L.samples_stats(beta)
        assert np.all(std > 0)
        p0 = T.init_params(np.random.default_rng(12), scale0=0.15).astype(np.float32)
        L.proposal_set(p0, mean, std)
        loss = L.train(300, batch=2048, draws=4, lr=2e-3, beta=beta, seed=9)
        assert np.all(np.isfinite(loss))
        assert loss[-50:].mean() < 0.8 * loss[:20].mean()   # ESJD (= -loss) grew
        p1, m1, s1 = L.proposal_get()
        assert np.all(np.isfinite(p1)) and np.allclose(m1, mean)
        # sampling: same equilibrium plaquette as the plain chain
        acc_plain, pl_plain = L.metropolis(200, beta, 0.24, 10, seed=6, sweep_offset=1000, plaquette=True)
        acc_l, pl_l = L.metropolis_learned(200, beta, 0.24, 10, seed=7, plaquette=True)
    from glac_amd.analysis import tau_int
    a, b = tau_int(pl_plain[50:]), tau_int(pl_l[50:])
    assert abs(a["mean"] - b["mean"]) < 4 * np.hypot(a["err"], b["err"]) + 1e-4, (a, b)
    assert 0.05 < acc_l.mean() < 0.95
