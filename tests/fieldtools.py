"""numpy helpers for the tests: lattice fields in GLAC's host layout
(float64 (4, V, 18), x-fastest sites), random SU(3), gauge transforms,
reflections, brute-force Wilson action, constant abelian fields."""
import numpy as np

import oracle as O

AXIS = {0: 3, 1: 2, 2: 1, 3: 0}  # mu -> axis of the (Nt, Nz, Ny, Nx) site grid


def to_grid(U, n):
    """(4, V, 18) pairs -> complex (4, Nt, Nz, Ny, Nx, 3, 3)."""
    c = O.to_complex(U)
    return c.reshape(4, n[3], n[2], n[1], n[0], 3, 3)


def from_grid(G):
    return O.to_pairs(G.reshape(4, -1, 3, 3))


def shift(F, mu, s):
    """F(x + s mu) for a (Nt, Nz, Ny, Nx, ...) array."""
    return np.roll(F, -s, axis=AXIS[mu])


def dag(M):
    return np.conj(np.swapaxes(M, -1, -2))


def random_su3(rng, shape=()):
    z = (rng.normal(size=shape + (3, 3)) + 1j * rng.normal(size=shape + (3, 3))) / np.sqrt(2)
    q, r = np.linalg.qr(z)
    d = np.diagonal(r, axis1=-2, axis2=-1)
    q = q * (d / np.abs(d))[..., None, :]
    det = np.linalg.det(q)
    q[..., :, 2] /= det[..., None]
    return q


def random_field(n, rng):
    V = int(np.prod(n))
    return O.to_pairs(random_su3(rng, (4, V)))


def near_unit_field(n, rng, spread):
    """exp(i spread H) with H random hermitian traceless: a smooth-ish field."""
    V = int(np.prod(n))
    a = rng.normal(size=(4, V, 3, 3)) + 1j * rng.normal(size=(4, V, 3, 3))
    h = (a + dag(a)) / 2
    h -= np.trace(h, axis1=-2, axis2=-1)[..., None, None] * np.eye(3) / 3
    w, v = np.linalg.eigh(h)
    m = v @ (np.exp(1j * spread * w)[..., None] * dag(v))
    return O.to_pairs(m)


def gauge_transform(U, n, g):
    """U_mu(x) -> g(x) U_mu(x) g(x + mu)^dagger; g complex (Nt, Nz, Ny, Nx, 3, 3)."""
    G = to_grid(U, n)
    out = np.empty_like(G)
    for mu in range(4):
        out[mu] = g @ G[mu] @ dag(shift(g, mu, +1))
    return from_grid(out)


def reflect_t(U, n):
    """Time reflection t -> -t: U'_mu(x) = U_mu(Px) (mu < 3),
    U'_3(x) = U_3(Px - 3)^dagger, with P(t) = -t mod Nt."""
    G = to_grid(U, n)
    idx = (-np.arange(n[3])) % n[3]
    out = np.empty_like(G)
    for mu in range(3):
        out[mu] = G[mu][idx]
    out[3] = dag(G[3][(idx - 1) % n[3]])
    return from_grid(out)


def plaquettes(U, n):
    """Re tr P_{mu nu}(x) for the 6 planes, shape (6, Nt, Nz, Ny, Nx)."""
    G = to_grid(U, n)
    res = []
    for mu in range(4):
        for nu in range(mu + 1, 4):
            P = G[mu] @ shift(G[nu], mu, 1) @ dag(shift(G[mu], nu, 1)) @ dag(G[nu])
            res.append(np.trace(P, axis1=-2, axis2=-1).real)
    return np.array(res)


def wilson_action(U, n, beta):
    return beta / 3.0 * (3.0 - plaquettes(U, n)).sum()


def abelian_field(n, k01, k23):
    """Constant abelian field strength in planes (0,1) and (2,3), embedded as
    diag(e^{ia}, e^{-ia}, 1): every plaquette in plane (0,1) is
    diag(e^{i phi1}, e^{-i phi1}, 1) with phi1 = 2 pi k01 / (N0 N1), likewise
    (2,3); the twist sits on the last slice in mu = 0 and mu = 2."""
    N0, N1, N2, N3 = n
    phi1 = 2 * np.pi * k01 / (N0 * N1)
    phi2 = 2 * np.pi * k23 / (N2 * N3)
    t, z, y, x = np.meshgrid(np.arange(N3), np.arange(N2), np.arange(N1), np.arange(N0),
                             indexing="ij")
    ph = np.zeros((4, N3, N2, N1, N0))
    ph[1] = phi1 * x                      # U_1(x) = e^{i phi1 x0}
    ph[0] = np.where(x == N0 - 1, -phi1 * N0 * y, 0.0)
    ph[3] = phi2 * z                      # U_3(x) = e^{i phi2 x2}
    ph[2] = np.where(z == N2 - 1, -phi2 * N2 * t, 0.0)
    G = np.zeros((4, N3, N2, N1, N0, 3, 3), dtype=np.complex128)
    G[..., 0, 0] = np.exp(1j * ph)
    G[..., 1, 1] = np.exp(-1j * ph)
    G[..., 2, 2] = 1.0
    return from_grid(G), phi1, phi2
