"""Learned staple-conditioned Metropolis proposal: parameter initialisation
and the collect -> standardise -> train pipeline on a resident lattice.

No counterpart in GLAC: the proposal replaces the fixed-eps RST generator
(GLAC/math/matrices/su3matrixgenerator.h:170-321) inside the Metropolis
update (GLAC/system.cpp:321-361).  The network reads six gauge invariants of
the link's staple A (tr AA^dag, tr (AA^dag)^2, |det A|^2, Re det A, Im det A,
beta), standardised with the training samples' mean and deviation, and
outputs the Cholesky factor L of the covariance of an su(3) step
w = L z (z: 8 symmetric unit-variance numbers), X = exp(i w.lambda / 2).
The covariance depends on A only, which does not change while the link is
updated, so X and X^-1 are equally likely and exact accept/reject keeps
detailed balance.  Training maximises the expected squared jump distance
E[min(1, exp(-dS)) |w|^2] on (staple, link) pairs drawn from a chain.

Parameters: 2468 floats packed W1[32][6], b1[32], W2[32][32], b2[32],
W3[36][32], b3[36]; b3[0:8] are the logs of the diagonal of L.
"""
from __future__ import annotations

import numpy as np

N_IN, N_H1, N_H2, N_OUT = 6, 32, 32, 36
N_PARAM = N_H1 * N_IN + N_H1 + N_H2 * N_H1 + N_H2 + N_OUT * N_H2 + N_OUT


def init_params(seed: int = 12, step: float = 0.15) -> np.ndarray:
    """He-initialised hidden layers; a small output layer whose bias starts
    the proposal as an isotropic step of size `step` (L = step * 1)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(N_PARAM, np.float64)
    o = 0
    p[o:o + N_H1 * N_IN] = rng.normal(size=N_H1 * N_IN) * np.sqrt(2.0 / N_IN)
    o += N_H1 * N_IN + N_H1
    p[o:o + N_H2 * N_H1] = rng.normal(size=N_H2 * N_H1) * np.sqrt(2.0 / N_H1)
    o += N_H2 * N_H1 + N_H2
    p[o:o + N_OUT * N_H2] = rng.normal(size=N_OUT * N_H2) * 0.01
    o += N_OUT * N_H2
    p[o:o + 8] = np.log(step)
    return p.astype(np.float32)


def train_from_chain(L, beta: float, eps: float = 0.24, n_hit: int = 10, collect_sweeps: int = 20,
                     sweeps_between: int = 1, samples: int = 1_000_000, steps: int = 3000, batch: int = 4096,
                     draws: int = 4, lr: float = 2e-3, seed: int = 1, sweep_offset: int = 0,
                     init_seed: int = 12) -> np.ndarray:
    """Collect about `samples` (staple, link) pairs from `collect_sweeps`
    RST sweeps of the chain field (continuing it from `sweep_offset`), set
    the feature standardisation, initialise the network and train it for
    `steps` Adam steps.  Returns the per-step loss (negative ESJD) as an
    array whose attribute `samples` is the number of training samples.  The
    chain field ends `collect_sweeps * sweeps_between` sweeps further on."""
    vh = L.V // 2
    stride = max(1, (4 * L.V * collect_sweeps) // samples)     # 8 (mu, parity) passes of V/2 links
    L.samples_reset(8 * ((vh + stride - 1) // stride) * collect_sweeps)
    count = 0
    for i in range(collect_sweeps):
        L.metropolis(sweeps_between, beta, eps, n_hit, seed=seed,
                     sweep_offset=sweep_offset + i * sweeps_between)
        count = L.samples_collect(stride=stride)
    mean, std = L.samples_stats(beta)
    L.proposal_set(init_params(init_seed), mean, std)
    loss = L.train(steps, batch=batch, draws=draws, lr=lr, beta=beta, seed=seed).view(TrainLoss)
    loss.samples = count
    return loss


class TrainLoss(np.ndarray):
    """Per-step training loss; `samples` holds the training-set size."""
    samples = 0
