"""Oracle clover observables (GLAC/observables/mastersamplertopcxyz.cpp:236-384).

Pins: unit field; constant abelian fields whose plaquette, energy and charge
have closed forms (every leaf equals diag(e^{i phi}, e^{-i phi}, 1), so
C - C^dagger = diag(8i sin phi, -8i sin phi, 0) and
Q = V sin(phi1) sin(phi2) / (2 pi^2) -> 2 k01 k23 as the lattice grows);
gauge invariance; time reflection flips Q and keeps E; sum_t Q(t) = Q
(scripts/topct_sanity_test/topct_topc_test.py:31-37)."""
import numpy as np
import pytest

import oracle as O
from fieldtools import abelian_field, dag, gauge_transform, plaquettes, random_field, random_su3, reflect_t


def test_unit_field():
    n = (4, 4, 4, 8)
    plaq, e, q, qt = O.measure(O.unit_field(n), n)
    assert plaq == 1.0 and e == 0.0 and q == 0.0 and np.all(qt == 0.0)


@pytest.mark.parametrize("n,k1,k2", [((6, 6, 4, 4), 1, 1), ((8, 8, 6, 6), 1, -2), ((8, 4, 6, 8), 2, 1)])
def test_constant_abelian_field_closed_forms(n, k1, k2):
    U, phi1, phi2 = abelian_field(n, k1, k2)
    V = int(np.prod(n))
    plaq, e, q, qt = O.measure(U, n)
    s1, s2 = np.sin(phi1), np.sin(phi2)
    assert abs(plaq - ((2 * np.cos(phi1) + 1) + (2 * np.cos(phi2) + 1) + 12) / 18) < 1e-14
    assert abs(e - (-128.0 * (s1 * s1 + s2 * s2))) < 1e-11
    assert abs(q - V * s1 * s2 / (2 * np.pi ** 2)) < 1e-12
    assert np.allclose(qt, q / n[3], atol=1e-13)
    # continuum limit: Q -> 2 k01 k23
    assert abs(q - 2 * k1 * k2) < 0.25 * abs(2 * k1 * k2)


def test_plaquette_matches_brute_force():
    rng = np.random.default_rng(31)
    n = (4, 4, 6, 4)
    U = random_field(n, rng)
    plaq = O.measure(U, n)[0]
    assert abs(plaq - plaquettes(U, n).sum() / (18 * np.prod(n))) < 1e-14


def test_gauge_invariance_and_sum_rule():
    rng = np.random.default_rng(32)
    n = (4, 4, 4, 6)
    U = random_field(n, rng)
    g = random_su3(rng, (n[3], n[2], n[1], n[0]))
    a = O.measure(U, n)
    b = O.measure(gauge_transform(U, n, g), n)
    assert abs(a[0] - b[0]) < 1e-14 and abs(a[1] - b[1]) < 1e-11 and abs(a[2] - b[2]) < 1e-13
    assert np.abs(a[3] - b[3]).max() < 1e-13
    assert abs(a[3].sum() - a[2]) < 1e-14


def test_time_reflection_flips_charge():
    rng = np.random.default_rng(33)
    n = (4, 4, 4, 6)
    U = random_field(n, rng)
    a = O.measure(U, n)
    b = O.measure(reflect_t(U, n), n)
    assert abs(a[0] - b[0]) < 1e-14 and abs(a[1] - b[1]) < 1e-11
    assert abs(a[2] + b[2]) < 1e-13
    assert abs(a[2]) > 1e-4  # the check is not vacuous


# ---- GLAC's SuperSampler: the same observables plus the Weinberg operator


def test_super_sampler_shares_the_clover_observables():
    rng = np.random.default_rng(41)
    n = (4, 4, 6, 4)
    U = random_field(n, rng)
    a = O.measure(U, n)
    b = O.measure_super(U, n)
    assert a[0] == b[0] and a[1] == b[1] and a[2] == b[2] and np.array_equal(a[3], b[4])
    assert abs(b[5].sum() - b[3]) < 1e-12 * max(1.0, abs(b[3]))


def test_weinberg_gauge_invariant_and_odd_under_time_reflection():
    rng = np.random.default_rng(42)
    n = (4, 4, 4, 6)
    U = random_field(n, rng)
    w = O.measure_super(U, n)[3]
    g = random_su3(rng, (n[3], n[2], n[1], n[0]))
    wg = O.measure_super(gauge_transform(U, n, g), n)[3]
    wr = O.measure_super(reflect_t(U, n), n)[3]
    assert abs(w) > 1e-3                      # not vacuous
    assert abs(w - wg) < 1e-12 * abs(w)
    assert abs(w + wr) < 1e-12 * abs(w)


def test_weinberg_vanishes_on_unit_and_abelian_fields():
    n = (4, 4, 4, 8)
    assert O.measure_super(O.unit_field(n), n)[3] == 0.0
    U, phi1, phi2 = abelian_field(n, 1, 1)
    _, _, q, w, _, wt = O.measure_super(U, n)
    assert abs(q) > 0.5 and abs(w) < 1e-12 and np.abs(wt).max() < 1e-12


def test_weinberg_is_cubic_and_energy_quadratic_in_the_field_strength():
    """Near the unit field, U = exp(i s H) with H fixed: E ~ s^2 and the
    Weinberg operator ~ s^3 (three field-strength tensors; a product of two
    would scale as s^2).  Halving s divides them by 4 and 8."""
    rng = np.random.default_rng(5)
    n = (4, 4, 4, 4)
    V = int(np.prod(n))
    a = rng.normal(size=(4, V, 3, 3)) + 1j * rng.normal(size=(4, V, 3, 3))
    h = (a + dag(a)) / 2
    h -= np.trace(h, axis1=-2, axis2=-1)[..., None, None] * np.eye(3) / 3
    w, v = np.linalg.eigh(h)
    r = [O.measure_super(O.to_pairs(v @ (np.exp(1j * s * w)[..., None] * dag(v))), n) for s in (0.02, 0.01)]
    assert 3.8 < r[0][1] / r[1][1] < 4.2
    assert 7.5 < r[0][3] / r[1][3] < 8.7, (r[0][3], r[1][3])


# ---- GLAC's LatticeActionChargeDensity: per-site densities


def test_densities_sum_to_the_totals_and_are_local_invariants():
    rng = np.random.default_rng(43)
    n = (4, 6, 4, 4)
    V = int(np.prod(n))
    U = random_field(n, rng)
    qd, ed = O.measure_density(U, n)
    _, e, q, _ = O.measure(U, n)
    assert abs(qd.sum() - q) < 1e-12 and abs(ed.sum() / V - e) < 1e-11 * abs(e)
    g = random_su3(rng, (n[3], n[2], n[1], n[0]))
    qg, eg = O.measure_density(gauge_transform(U, n, g), n)
    assert np.abs(qg - qd).max() < 1e-13 and np.abs(eg - ed).max() < 1e-10
    # translating the field by one site in y translates the densities
    from fieldtools import from_grid, to_grid
    Ut = from_grid(np.roll(to_grid(U, n), 1, axis=3))          # grid axes (mu, t, z, y, x, 3, 3)
    qt, et = O.measure_density(Ut, n)
    grid = lambda a: a.reshape(n[3], n[2], n[1], n[0])
    assert np.abs(grid(qt) - np.roll(grid(qd), 1, axis=2)).max() < 1e-14
    assert np.abs(grid(et) - np.roll(grid(ed), 1, axis=2)).max() < 1e-11
