"""GPU parity: libglacx.so through its C ABI against the FP64 oracle.

Bars: FP64 links agree with the oracle to rounding (1e-11 relative on
observables, 1e-10 on links after flow); FP32 links to the stated FP32
tolerances; Metropolis on the same counter-based streams gives the oracle's
accept/reject decisions (FP64) and GLAC's recorded acceptance and plaquette
(both precisions, statistically)."""
import os
import sys

import numpy as np
import pytest

import oracle as O
from fieldtools import abelian_field, gauge_transform, near_unit_field, random_field, random_su3

pytestmark = pytest.mark.gpu

glacx = pytest.importorskip("glac_amd")

PRECISIONS = [glacx.FP64, glacx.FP32_LINKS, glacx.FP32_COMPRESSED]
PNAME = {glacx.FP64: "fp64", glacx.FP32_LINKS: "fp32", glacx.FP32_COMPRESSED: "fp32c"}


def banner(msg):
    print(f"\n[gpu case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(params=PRECISIONS, ids=lambda p: PNAME[p])
def prec(request):
    return request.param


def test_upload_download_roundtrip(prec):
    n = (4, 4, 4, 8)
    banner(f"roundtrip {n} {PNAME[prec]}")
    U = random_field(n, np.random.default_rng(1))
    with glacx.Lattice(n, prec) as L:
        L.upload(U)
        back = L.download()
        L.copy(glacx.FLOW, glacx.MAIN)
        back2 = L.download(glacx.FLOW)
    # FP32 rounding of each stored entry (2^-24 relative); the compressed
    # mode's rebuilt third row carries a few such roundings
    tol = {glacx.FP64: 0.0, glacx.FP32_LINKS: 1e-7, glacx.FP32_COMPRESSED: 5e-7}[prec]
    assert np.abs(back - U).max() <= tol
    assert np.array_equal(back, back2)


def test_hot_start_matches_oracle(prec):
    n = (4, 6, 4, 8)
    banner(f"hot start {n} {PNAME[prec]}")
    with glacx.Lattice(n, prec) as L:
        L.hot_start(2024)
        got = L.download()
    want = O.hot_start_haar(n, 2024)
    tol = {glacx.FP64: 1e-13, glacx.FP32_LINKS: 1e-7, glacx.FP32_COMPRESSED: 5e-7}[prec]
    assert np.abs(got - want).max() < tol


@pytest.mark.parametrize("kind", [glacx.HOT_GLAC_RANDOM, glacx.HOT_RST], ids=["glac_random", "rst"])
def test_hot_start_kinds_match_oracle(prec, kind):
    """GLAC's two hot starts (generateRandom, det = -1; RSTHotStart) drawn
    from the link streams, against the oracle on the same streams."""
    n = (4, 6, 4, 8)
    banner(f"hot start kind {kind} {n} {PNAME[prec]}")
    with glacx.Lattice(n, prec) as L:
        if kind == glacx.HOT_GLAC_RANDOM and prec == glacx.FP32_COMPRESSED:
            with pytest.raises(glacx.GlacxError, match="det U = -1"):
                L.hot_start(77, kind=kind, eps=0.24)
            return
        L.hot_start(77, kind=kind, eps=0.24)
        got = L.download()
        p, _, _, _ = L.measure()
    want = O.hot_start(n, 77, kind, 0.24).reshape(4, -1, 18)
    tol = {glacx.FP64: 1e-13, glacx.FP32_LINKS: 1e-7, glacx.FP32_COMPRESSED: 5e-7}[prec]
    assert np.abs(got - want).max() < tol
    det = np.linalg.det(O.to_complex(got).reshape(-1, 3, 3))
    assert np.abs(det - (-1.0 if kind == glacx.HOT_GLAC_RANDOM else 1.0)).max() < 10 * tol + 1e-13
    assert abs(p - O.measure(want, n)[0]) < (1e-13 if prec == glacx.FP64 else 5e-7)


def test_cold_start_measures_unity(prec):
    n = (4, 4, 4, 4)
    banner(f"cold start {n} {PNAME[prec]}")
    with glacx.Lattice(n, prec) as L:
        L.cold_start()
        p, e, q, qt = L.measure()
        assert L.plaquette() == 1.0
    assert p == 1.0 and e == 0.0 and q == 0.0 and np.all(qt == 0)


@pytest.mark.parametrize("n", [(4, 4, 4, 4), (4, 4, 4, 8), (6, 4, 8, 6), (8, 8, 8, 16)])
def test_measure_matches_oracle(n, prec):
    banner(f"measure {n} {PNAME[prec]}")
    U = random_field(n, np.random.default_rng(sum(n)))
    want = O.measure(U, n)
    with glacx.Lattice(n, prec) as L:
        L.upload(U)
        got = L.measure()
        plaq = L.plaquette()
    if prec == glacx.FP64:
        assert abs(got[0] - want[0]) < 1e-13 and abs(plaq - want[0]) < 1e-13
        assert abs(got[1] - want[1]) < 1e-11 * abs(want[1])
        assert abs(got[2] - want[2]) < 1e-12
        assert np.abs(got[3] - want[3]).max() < 1e-12
    else:
        assert abs(got[0] - want[0]) < 2e-7 and abs(plaq - want[0]) < 2e-7
        assert abs(got[1] - want[1]) < 2e-6 * abs(want[1])
        assert abs(got[2] - want[2]) < 1e-5
        assert np.abs(got[3] - want[3]).max() < 1e-5


def test_measure_abelian_closed_form(prec):
    n = (8, 8, 6, 6)
    banner(f"abelian {n} {PNAME[prec]}")
    U, phi1, phi2 = abelian_field(n, 1, -2)
    V = np.prod(n)
    with glacx.Lattice(n, prec) as L:
        L.upload(U)
        p, e, q, qt = L.measure()
    s1, s2 = np.sin(phi1), np.sin(phi2)
    tol = 1e-11 if prec == glacx.FP64 else 2e-5
    assert abs(q - V * s1 * s2 / (2 * np.pi ** 2)) < tol * 10
    assert abs(e + 128 * (s1 * s1 + s2 * s2)) < tol * 100
    assert abs(p - ((2 * np.cos(phi1) + 1) + (2 * np.cos(phi2) + 1) + 12) / 18) < tol


def test_action_derivative_matches_oracle(prec):
    n = (4, 4, 6, 4)
    banner(f"action derivative {n} {PNAME[prec]}")
    U = random_field(n, np.random.default_rng(7))
    with glacx.Lattice(n, prec) as L:
        L.upload(U)
        for mu in range(4):
            got = L.action_derivative(mu)
            want = O.action_derivative(U, n, mu)
            tol = 1e-13 if prec == glacx.FP64 else 5e-6
            assert np.abs(got - want).max() < tol


def run_oracle_flow(U, n, eps, steps):
    W = U.copy()
    rows = [O.measure(W, n)]
    for _ in range(steps):
        O.flow_step(W, n, eps)
        rows.append(O.measure(W, n))
    return W, rows


@pytest.mark.parametrize("n,steps,spread", [((4, 4, 4, 4), 12, 2.0), ((4, 4, 4, 8), 20, 0.7)])
def test_flow_matches_oracle_per_step(n, steps, spread, prec):
    banner(f"flow {n} steps={steps} {PNAME[prec]}")
    U = near_unit_field(n, np.random.default_rng(steps), spread)
    W, rows = run_oracle_flow(U, n, 0.02, steps)
    with glacx.Lattice(n, prec) as L:
        L.upload(U, glacx.FLOW)
        p, e, q, qt = L.flow(steps, 0.02)
        got = L.download(glacx.FLOW)
    wp = np.array([r[0] for r in rows])
    we = np.array([r[1] for r in rows])
    wq = np.array([r[2] for r in rows])
    wqt = np.array([r[3] for r in rows])
    if prec == glacx.FP64:
        assert np.abs(got - W).max() < 1e-11
        assert np.abs(p - wp).max() < 1e-12
        assert np.abs(e / we - 1).max() < 1e-11
        assert np.abs(q - wq).max() < 1e-12
        assert np.abs(qt - wqt).max() < 1e-12
    else:
        de = np.abs(e / we - 1).max()
        print(f"fp32 flow {n}: max |dE|/E {de:.2e}, |dQ| {np.abs(q - wq).max():.2e}, "
              f"|dU| {np.abs(got - W).max():.2e}")
        assert np.abs(got - W).max() < 5e-6
        assert np.abs(p - wp).max() < 5e-7
        assert de < 1e-6
        assert np.abs(q - wq).max() < 1e-4


@pytest.fixture(scope="module")
def config_8x8x8x16():
    """8^3 x 16, beta = 6.0: Haar hot start and 200 FP64 device sweeps
    (eps = 0.24, 10 hits), then 100 RK3 steps of the oracle at eps = 0.01
    with the clover observables per step."""
    n = (8, 8, 8, 16)
    banner(f"config {n}: 200 fp64 device sweeps + oracle flow")
    with glacx.Lattice(n, glacx.FP64) as L:
        L.hot_start(99)
        acc, plaq = L.metropolis(200, 6.0, 0.24, 10, reunit_every=1, seed=99, plaquette=True)
        U = np.ascontiguousarray(L.download())
    W, rows = run_oracle_flow(U, n, 0.01, 100)
    return n, U, acc, plaq, rows


def test_flow_8x8x8x16_against_oracle(prec, config_8x8x8x16):
    """8^3 x 16 after 200 device sweeps, flowed 100 RK3 steps at eps = 0.01:
    plaquette, E(t) and Q(t) per step against the oracle."""
    n, U, acc, plaq, rows = config_8x8x8x16
    banner(f"flow {n} 100 steps {PNAME[prec]}")
    assert 0.55 < plaq[-50:].mean() < 0.62 and 0.05 < acc.mean() < 0.2
    with glacx.Lattice(n, prec) as L:
        L.upload(U)
        L.copy(glacx.FLOW, glacx.MAIN)
        p, e, q, qt = L.flow(100, 0.01)
    wp = np.array([r[0] for r in rows])
    we = np.array([r[1] for r in rows])
    wq = np.array([r[2] for r in rows])
    wqt = np.array([r[3] for r in rows])
    de, dq = np.abs(e / we - 1).max(), np.abs(q - wq).max()
    print(f"{PNAME[prec]} flow {n}: max |dE|/E {de:.2e}, |dQ| {dq:.2e}, |dplaq| {np.abs(p - wp).max():.2e}")
    if prec == glacx.FP64:
        assert de < 1e-11 and dq < 1e-11
        assert np.abs(p - wp).max() < 1e-12 and np.abs(qt - wqt).max() < 1e-11
    else:
        assert de < 1e-6 and dq < 1e-4
        assert np.abs(p - wp).max() < 5e-7


def test_flow_keeps_pure_gauge(prec):
    n = (4, 4, 4, 4)
    banner(f"flow pure gauge {n} {PNAME[prec]}")
    U = gauge_transform(O.unit_field(n), n, random_su3(np.random.default_rng(3), (4, 4, 4, 4)))
    with glacx.Lattice(n, prec) as L:
        L.upload(U, glacx.FLOW)
        p, e, q, qt = L.flow(5, 0.02)
    tol = 1e-12 if prec == glacx.FP64 else 1e-5
    assert np.abs(p - 1).max() < tol and np.abs(e).max() < tol * 100 and np.abs(q).max() < tol


def test_metropolis_fp64_follows_oracle_streams():
    """Same seed, same counter-based streams, same checkerboard order: the
    FP64 device sweeps make the oracle's accept/reject decisions."""
    n = (4, 4, 4, 8)
    banner(f"metropolis fp64 vs oracle {n}")
    U = O.hot_start_haar(n, 5)
    accs = [O.cb_sweep(U, n, 6.0, 0.24, 10, 31, s, True) for s in range(3)]
    with glacx.Lattice(n, glacx.FP64) as L:
        L.hot_start(5)
        got_acc = L.metropolis(3, 6.0, 0.24, 10, reunit_every=1, seed=31, sweep_offset=0)
        got = L.download()
    assert np.array_equal(np.round(np.array(accs) * 4 * 10 * np.prod(n)),
                          np.round(got_acc * 4 * 10 * np.prod(n)))
    assert np.abs(got - U).max() < 1e-10


def test_metropolis_fp32_single_sweep_close_to_oracle():
    n = (4, 4, 4, 8)
    banner(f"metropolis fp32 one sweep {n}")
    U = O.hot_start_haar(n, 6)
    acc = O.cb_sweep(U, n, 6.0, 0.24, 10, 32, 0, True)
    with glacx.Lattice(n, glacx.FP32_LINKS) as L:
        L.hot_start(6)
        got_acc = L.metropolis(1, 6.0, 0.24, 10, reunit_every=1, seed=32)
        got = L.download()
    # an FP32 rounding can flip a rare near-threshold decision; nearly all
    # links must match the FP64 chain
    close = np.abs(got - U).reshape(4, -1, 18).max(axis=-1) < 1e-4
    assert close.mean() > 0.99
    assert abs(got_acc[0] - acc) < 0.002


GLAC_PLAQ_MEAN = 0.59586  # examples/test_config/configExampleLoad.out:60-69
GLAC_ACCEPT = 0.1079086669921875  # :72


def test_metropolis_reproduces_glac_recorded_run(prec):
    n = (4, 4, 4, 8)
    banner(f"metropolis GLAC recorded run {n} {PNAME[prec]}")
    with glacx.Lattice(n, prec) as L:
        L.cold_start()
        L.metropolis(150, 6.0, 0.24, 10, seed=11)
        acc, plaq = L.metropolis(600, 6.0, 0.24, 10, seed=11, sweep_offset=150, plaquette=True)
        U = L.download()
    assert abs(acc.mean() - GLAC_ACCEPT) < 0.004
    assert abs(plaq.mean() - GLAC_PLAQ_MEAN) < 0.008
    M = O.to_complex(U)
    tol = 1e-13 if prec == glacx.FP64 else 2e-6
    assert np.abs(M @ np.conj(np.swapaxes(M, -1, -2)) - np.eye(3)).max() < tol * 10
    assert np.abs(np.linalg.det(M) - 1).max() < tol * 10


def test_chroma_load_is_big_endian_bin(tmp_path, prec):
    n = (4, 4, 4, 6)
    banner(f"chroma load {n} {PNAME[prec]}")
    U = random_field(n, np.random.default_rng(7))
    site_major = np.ascontiguousarray(U.reshape(4, -1, 18).transpose(1, 0, 2))
    path = str(tmp_path / "cfg.chroma")
    site_major.astype(">f8").tofile(path)
    with glacx.Lattice(n, prec) as L:
        L.load_chroma(path)
        got = L.download()
        p_chroma = L.plaquette()
        L.upload(U)
        p_direct = L.plaquette()
    tol = 0 if prec == glacx.FP64 else 1e-6
    assert np.abs(got - U).max() <= tol
    assert abs(p_chroma - p_direct) <= tol


def test_bin_roundtrip_glac_layout(tmp_path, prec):
    n = (4, 4, 6, 8)
    banner(f"bin io {n} {PNAME[prec]}")
    U = random_field(n, np.random.default_rng(5))
    path = str(tmp_path / "cfg.bin")
    with glacx.Lattice(n, prec) as L:
        L.upload(U)
        L.save_bin(path)
        raw = np.fromfile(path, dtype=np.float64)
        L.cold_start()
        L.load_bin(path)
        back = L.download()
    # GLAC layout: site-major, 4 links per site (GLAC/io/fieldio.cpp:95-107)
    assert raw.size == 72 * np.prod(n)
    site_major = raw.reshape(-1, 4, 18).transpose(1, 0, 2)
    tol = {glacx.FP64: 0.0, glacx.FP32_LINKS: 1e-7, glacx.FP32_COMPRESSED: 5e-7}[prec]
    assert np.abs(site_major - U).max() <= tol
    assert np.abs(back - U).max() <= tol


def test_super_sampler_weinberg_matches_oracle(prec):
    n = (4, 6, 4, 8)
    banner(f"super sampler {n} {PNAME[prec]}")
    U = near_unit_field(n, np.random.default_rng(21), 0.6)
    want = O.measure_super(U, n)
    with glacx.Lattice(n, prec) as L:
        L.upload(U)
        got = L.measure_super()
        plain = L.measure()
    assert abs(want[3]) > 1e-4            # the Weinberg value is not trivially zero
    if prec == glacx.FP64:
        for i in range(4):
            assert abs(got[i] - want[i]) < 1e-11 * max(1.0, abs(want[i])), i
        assert np.abs(got[4] - want[4]).max() < 1e-12 and np.abs(got[5] - want[5]).max() < 1e-12
    else:
        assert abs(got[0] - want[0]) < 5e-7 and abs(got[1] / want[1] - 1) < 1e-5
        assert abs(got[2] - want[2]) < 1e-4 and abs(got[3] - want[3]) < 1e-4 * max(1.0, abs(want[3]))
        assert np.abs(got[5] - want[5]).max() < 1e-4 * max(1.0, np.abs(want[5]).max())
    # the plain measurement agrees with the SuperSampler's shared observables
    assert plain[0] == got[0] and plain[1] == got[1] and plain[2] == got[2]
    assert abs(got[5].sum() - got[3]) < 1e-12 * max(1.0, abs(got[3]))


def test_flow_super_matches_oracle(prec):
    n = (4, 4, 4, 6)
    banner(f"flow super {n} {PNAME[prec]}")
    U = near_unit_field(n, np.random.default_rng(22), 0.5)
    W = U.copy()
    rows = [O.measure_super(W, n)]
    for _ in range(3):
        O.flow_step(W, n, 0.02)
        rows.append(O.measure_super(W, n))
    with glacx.Lattice(n, prec) as L:
        L.upload(U, field=glacx.FLOW)
        p, e, q, w, qt, wt = L.flow_super(3, 0.02)
    tol = 1e-11 if prec == glacx.FP64 else 1e-4
    for i, r in enumerate(rows):
        assert abs(w[i] - r[3]) < tol * max(1.0, abs(r[3])), i
        assert np.abs(wt[i] - r[5]).max() < tol * max(1.0, np.abs(r[5]).max()), i
        assert abs(q[i] - r[2]) < tol * 10


def test_measure_density_matches_oracle(prec):
    n = (4, 6, 4, 8)
    banner(f"density {n} {PNAME[prec]}")
    U = near_unit_field(n, np.random.default_rng(23), 0.6)
    qd, ed = O.measure_density(U, n)
    with glacx.Lattice(n, prec) as L:
        L.upload(U)
        gq, ge, p, e, q = L.measure_density()
        plain = L.measure()
    assert p == plain[0] and e == plain[1] and q == plain[2]
    if prec == glacx.FP64:
        assert np.abs(gq - qd).max() < 1e-14 and np.abs(ge - ed).max() < 1e-11 * np.abs(ed).max()
    else:
        assert np.abs(gq - qd).max() < 2e-6 and np.abs(ge - ed).max() < 2e-5 * np.abs(ed).max()
    assert abs(gq.sum() - q) < 1e-9 and abs(ge.sum() / np.prod(n) - e) < 1e-9 * abs(e)


def test_compressed_mode_refuses_non_su3_links(tmp_path):
    """Two stored rows rebuild the third only for SU(3) links: GLAC's
    generateRandom hot start (det = -1) must be refused, not mangled."""
    n = (4, 4, 4, 4)
    banner("compressed mode refuses det = -1")
    U = O.new_field(n)
    rng = O.GlacRng(3, 4)
    O.glac_hot_start(U, n, rng, 0.24)
    assert np.allclose(np.linalg.det(O.to_complex(U[0, :4])), -1.0)
    path = str(tmp_path / "glac_hot.bin")
    np.ascontiguousarray(U.reshape(4, -1, 18).transpose(1, 0, 2)).tofile(path)
    with glacx.Lattice(n, glacx.FP32_COMPRESSED) as L:
        with pytest.raises(glacx.GlacxError, match="SU\\(3\\)"):
            L.upload(U)
        with pytest.raises(glacx.GlacxError, match="SU\\(3\\)"):
            L.load_bin(path)
    with glacx.Lattice(n, glacx.FP32_LINKS) as L:     # the uncompressed mode takes them
        L.upload(U)
        assert np.abs(L.download() - U).max() < 1e-6


def test_errors_are_reported():
    with pytest.raises(glacx.GlacxError):
        glacx.Lattice((3, 4, 4, 4))
    with glacx.Lattice((4, 4, 4, 4)) as L:
        with pytest.raises(glacx.GlacxError, match="cannot open"):
            L.load_bin("/nonexistent/cfg.bin")
        with pytest.raises(glacx.GlacxError):
            L.metropolis(1, 6.0, eps=1.5)
