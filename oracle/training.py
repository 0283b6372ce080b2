"""FP64 numpy reference of the learned proposal's training objective and its
gradient.  TEST INFRASTRUCTURE ONLY (checker of the device training kernels).

Per training sample (staple A, link U) and per draw z (8 symmetric
unit-variance numbers), with W = U A and the network output L(A):
    w = L z,  Y = i sum_a w_a lambda_a / 2,
    dS = -(beta / 3) (Re tr(Y W) + Re tr(Y Y W) / 2)     (second order in Y),
    alpha = min(1, exp(-dS)),   loss = -alpha |w|^2,
averaged over samples and draws: the negative expected squared jump distance
of the proposal, which training minimises.  The sampler itself always uses
the exact exponential and the exact dS, so this approximation only shapes
the proposal, never the stationary distribution.
"""
from __future__ import annotations

import numpy as np

N1, N2, NOUT, NIN = 32, 32, 36, 6
SLICES = {}
_o = 0
for name, size in [("W1", N1 * NIN), ("b1", N1), ("W2", N2 * N1), ("b2", N2), ("W3", NOUT * N2), ("b3", NOUT)]:
    SLICES[name] = slice(_o, _o + size)
    _o += size
NPARAM = _o
assert NPARAM == 2468

R3 = 1.0 / np.sqrt(3.0)
TRIL = [(a, b) for a in range(8) for b in range(a)]  # off-diagonal order: 8 + a(a-1)/2 + b


def unpack(p):
    return (p[SLICES["W1"]].reshape(N1, NIN), p[SLICES["b1"]], p[SLICES["W2"]].reshape(N2, N1), p[SLICES["b2"]],
            p[SLICES["W3"]].reshape(NOUT, N2), p[SLICES["b3"]])


def features(A, beta):
    """A: complex (S, 3, 3) -> (S, 6)."""
    B = A @ np.conj(np.swapaxes(A, -1, -2))
    d = np.linalg.det(A)
    return np.stack([np.trace(B, axis1=-2, axis2=-1).real,
                     np.trace(B @ B, axis1=-2, axis2=-1).real,
                     np.abs(d) ** 2, d.real, d.imag, np.full(len(A), beta)], axis=-1)


def algebra(w):
    """w: (..., 8) -> Y: complex (..., 3, 3) = i sum w_a lambda_a / 2."""
    Y = np.zeros(w.shape[:-1] + (3, 3), dtype=np.complex128)
    Y[..., 0, 0] = 0.5j * (w[..., 2] + R3 * w[..., 7])
    Y[..., 1, 1] = 0.5j * (-w[..., 2] + R3 * w[..., 7])
    Y[..., 2, 2] = -1j * R3 * w[..., 7]
    for (i, j, re, im) in [(0, 1, 1, 0), (0, 2, 4, 3), (1, 2, 6, 5)]:
        Y[..., i, j] = 0.5 * w[..., re] + 0.5j * w[..., im]
        Y[..., j, i] = -0.5 * w[..., re] + 0.5j * w[..., im]
    return Y


def project(N):
    """N: complex (..., 3, 3) -> g: (..., 8), g_a = Re tr(i lambda_a / 2 N)."""
    g = np.empty(N.shape[:-2] + (8,))
    g[..., 0] = -0.5 * (N[..., 0, 1].imag + N[..., 1, 0].imag)
    g[..., 1] = 0.5 * (N[..., 1, 0].real - N[..., 0, 1].real)
    g[..., 2] = -0.5 * (N[..., 0, 0].imag - N[..., 1, 1].imag)
    g[..., 3] = -0.5 * (N[..., 0, 2].imag + N[..., 2, 0].imag)
    g[..., 4] = 0.5 * (N[..., 2, 0].real - N[..., 0, 2].real)
    g[..., 5] = -0.5 * (N[..., 1, 2].imag + N[..., 2, 1].imag)
    g[..., 6] = 0.5 * (N[..., 2, 1].real - N[..., 1, 2].real)
    g[..., 7] = -0.5 * R3 * (N[..., 0, 0].imag + N[..., 1, 1].imag - 2 * N[..., 2, 2].imag)
    return g


def lrelu(v):
    return np.where(v > 0, v, 0.01 * v)


def forward(p, mean, std, f):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    x = (f - mean) / std
    a1 = x @ W1.T + b1
    h1 = lrelu(a1)
    a2 = h1 @ W2.T + b2
    h2 = lrelu(a2)
    o = h2 @ W3.T + b3
    return x, a1, h1, a2, h2, o


def cholesky(o):
    S = o.shape[0]
    L = np.zeros((S, 8, 8))
    idx = np.arange(8)
    L[:, idx, idx] = np.exp(o[:, :8])
    for k, (a, b) in enumerate(TRIL):
        L[:, a, b] = o[:, 8 + k]
    return L


def loss_and_grad(p, mean, std, A, U, beta, Z):
    """A, U: complex (S, 3, 3); Z: (S, K, 8).  Returns (loss, grad[NPARAM])."""
    S, K, _ = Z.shape
    f = features(A, beta)
    x, a1, h1, a2, h2, o = forward(p, mean, std, f)
    L = cholesky(o)
    W = U @ A
    w = np.einsum("sab,skb->ska", L, Z)
    Y = algebra(w)
    Wb = W[:, None]
    YW = Y @ Wb
    lin = np.trace(YW, axis1=-2, axis2=-1).real
    quad = np.trace(Y @ YW, axis1=-2, axis2=-1).real
    dS = -(beta / 3.0) * (lin + 0.5 * quad)
    alpha = np.minimum(1.0, np.exp(-dS))
    w2 = (w * w).sum(-1)
    loss = -(alpha * w2).mean()
    # d loss / d w
    gW = project(Wb + 0 * Y)                            # (S, K, 8): d lin / d w
    gQ = project(YW + Wb @ Y)                           # d quad / d w
    ddS = -(beta / 3.0) * (gW + 0.5 * gQ)
    dalpha = np.where(dS > 0, -alpha, 0.0)[..., None] * ddS
    dl_dw = -(2.0 * alpha[..., None] * w + w2[..., None] * dalpha) / (S * K)
    dL = np.einsum("ska,skb->sab", dl_dw, Z)
    do = np.empty((S, NOUT))
    idx = np.arange(8)
    do[:, :8] = dL[:, idx, idx] * L[:, idx, idx]
    for k, (a, b) in enumerate(TRIL):
        do[:, 8 + k] = dL[:, a, b]
    W1, b1, W2, b2, W3, b3 = unpack(p)
    g = np.zeros(NPARAM)
    g[SLICES["W3"]] = (do.T @ h2).ravel()
    g[SLICES["b3"]] = do.sum(0)
    dh2 = do @ W3
    da2 = dh2 * np.where(a2 > 0, 1.0, 0.01)
    g[SLICES["W2"]] = (da2.T @ h1).ravel()
    g[SLICES["b2"]] = da2.sum(0)
    dh1 = da2 @ W2
    da1 = dh1 * np.where(a1 > 0, 1.0, 0.01)
    g[SLICES["W1"]] = (da1.T @ x).ravel()
    g[SLICES["b1"]] = da1.sum(0)
    return loss, g


def init_params(rng, scale0=0.15):
    """He-style initialisation; the output bias starts the proposal as
    scale0 times the identity covariance factor."""
    p = np.zeros(NPARAM)
    p[SLICES["W1"]] = rng.normal(size=N1 * NIN) * np.sqrt(2.0 / NIN)
    p[SLICES["W2"]] = rng.normal(size=N2 * N1) * np.sqrt(2.0 / N1)
    p[SLICES["W3"]] = rng.normal(size=NOUT * N2) * 0.01
    b3 = np.zeros(NOUT)
    b3[:8] = np.log(scale0)
    p[SLICES["b3"]] = b3
    return p
