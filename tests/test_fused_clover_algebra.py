"""The leaf bookkeeping of the flow stage that measures the clover
observables (glac_amd/csrc/k_flow.hip: plane_first, plane_slot and the phase
schedule kPhaseNu / kPhaseAdd / kPhaseL3), restated in numpy and checked
against the oracle's clover measurement
(GLAC/observables/mastersamplertopcxyz.cpp:236-384).

Per wavefront mu and staple term nu the stage holds
  up  = U_nu(x+mu) U_mu(x+nu)^dag U_nu(x)^dag,
  low = U_nu(x+mu-nu)^dag U_mu(x-nu)^dag U_nu(x-nu);
the first-index wavefront of a plane adds T(U_mu (up - low)), the other one
T(U_mu low), and the plane's L3 owner adds T(L3), T(M) = traceless(M - M^dag).
The two wavefronts of a plane must reach it in different phases (the earlier
stores, the later adds), every plane needs one L3 owner among its two
wavefronts.  A wrong slot, schedule entry, sign or leaf orientation changes
E, Q or Q(t) or breaks one of these conditions."""
import numpy as np
import pytest

import oracle as O
from fieldtools import dag, random_field, shift, to_grid

K_PHASE_NU = 0x210130302321
K_PHASE_ADD = 0x936
K_PHASE_L3 = 0x3A3


def T2(m):
    a = m - dag(m)
    return a - np.trace(a, axis1=-2, axis2=-1)[..., None, None] / 3 * np.eye(3)


def plane_first(mu, nu):
    return mu == 0 or nu == mu % 3 + 1


def plane_slot(d1, d2):
    return d2 - 1 if d1 == 0 else (5 if d1 == 1 else d1 + 1)


def phase(mu, k):
    """(nu, adds to the slot, owns L3) of wavefront mu in phase k."""
    e = 3 * mu + k
    return (K_PHASE_NU >> (4 * e)) & 0xF, bool((K_PHASE_ADD >> e) & 1), bool((K_PHASE_L3 >> e) & 1)


def fused_measure(U, n):
    G = to_grid(U, n)
    F = [None] * 6
    plaq = 0.0
    l3_count = [0] * 4
    for k in range(3):               # phases, in order, as the kernel's barriers impose
        for mu in range(4):
            nu, add, l3 = phase(mu, k)
            u = G[mu]
            up = shift(G[nu], mu, 1) @ dag(shift(G[mu], nu, 1)) @ dag(G[nu])
            low = dag(shift(shift(G[nu], mu, 1), nu, -1)) @ dag(shift(G[mu], nu, -1)) @ shift(G[nu], nu, -1)
            first = plane_first(mu, nu)
            d1, d2 = (mu, nu) if first else (nu, mu)
            slot = plane_slot(d1, d2)
            if first:
                plaq += np.trace(u @ up, axis1=-2, axis2=-1).real.sum()
                f = T2(u @ (up - low))
            else:
                f = T2(u @ low)
            if l3:
                l3_count[mu] += 1
                m1 = shift(shift(G[d1], d1, -1), d2, -1) @ shift(G[d2], d2, -1)   # U_d1(y) U_d2(x-d2)
                m2 = shift(shift(G[d2], d1, -1), d2, -1) @ shift(G[d1], d1, -1)   # U_d2(y) U_d1(x-d1)
                f = f + T2(dag(m2) @ m1)
            assert (F[slot] is not None) == add   # stored by the other wavefront in an earlier phase
            F[slot] = F[slot] + f if add else f
    V = int(np.prod(n))
    tr = lambda a, b: np.einsum('...ij,...ji->...', a, b).real
    e = sum(tr(F[k], F[k]).sum() for k in range(6)) / V
    qt = -sum(tr(F[k], F[k + 3]).sum(axis=(1, 2, 3)) for k in range(3)) / (256 * np.pi ** 2)
    return plaq / (18 * V), e, qt, l3_count


def test_slot_tables_cover_each_plane_once():
    seen = {}
    for mu in range(4):
        for nu in range(4):
            if mu != nu and plane_first(mu, nu):
                seen[plane_slot(mu, nu)] = (mu, nu)
    assert seen == {0: (0, 1), 1: (0, 2), 2: (0, 3), 3: (2, 3), 4: (3, 1), 5: (1, 2)}
    for mu in range(4):
        for nu in range(4):
            if mu != nu:
                assert plane_first(mu, nu) != plane_first(nu, mu)


def test_phase_schedule():
    when, owners = {}, {}
    per_phase_l3 = [0, 0, 0]
    for mu in range(4):
        assert sorted(phase(mu, k)[0] for k in range(3)) == sorted(set(range(4)) - {mu})
        for k in range(3):
            nu, add, l3 = phase(mu, k)
            when[(mu, nu)] = (k, add)
            if l3:
                owners.setdefault(frozenset((mu, nu)), []).append(mu)
                per_phase_l3[k] += 1
    for mu in range(4):
        for nu in range(mu + 1, 4):
            (ka, adda), (kb, addb) = when[(mu, nu)], when[(nu, mu)]
            assert ka != kb                                 # never the same phase
            assert adda == (ka > kb) and addb == (kb > ka)  # the later one adds
            assert len(owners[frozenset((mu, nu))]) == 1    # one L3 per plane
    assert per_phase_l3 == [2, 2, 2]


@pytest.mark.parametrize("n,seed", [((4, 6, 4, 6), 5), ((6, 4, 4, 8), 6)])
def test_fused_leaf_sums_equal_the_oracle_clover(n, seed):
    U = random_field(n, np.random.default_rng(seed))
    plaq, e, qt, l3_count = fused_measure(U, n)
    ref = O.measure(U, n)
    assert l3_count == [2, 1, 2, 1]
    assert abs(plaq - ref[0]) < 1e-13
    assert abs(e - ref[1]) < 1e-11 * abs(ref[1])
    assert np.abs(qt - ref[3]).max() < 1e-12
    assert abs(qt.sum() - ref[2]) < 1e-12
    assert abs(ref[2]) > 1e-3  # not vacuous
