"""The full pipeline at its real size: 48^3 x 96 (10.6 M sites, 3 GB of
compressed FP32 links) on one GPU, beta = 6.45.

Start: cold start and THERM sweeps of GLAC's fixed-eps RST update in
compressed FP32, the configuration both chains continue from.  (A hot start
at this coupling leaves |Q| ~ 70 of dislocations and long-distance modes far
from equilibrium after a thousand sweeps, so two hot-started chains are not
comparable.)

Learned chain: network trained on samples of the RST chain, then LEARNED
learned-proposal sweeps (the spacing between measurements).  Its
configuration is flowed to t = 10 (500 RK3 steps of 0.02) in compressed FP32
and, from the same links, in FP64: at every flow step the two agree to the
flow tolerances (|dE|/E < 1e-6, |dQ| < 1e-4 absolute, plaquette 1e-6), E(t)
falls and the t0 scale (t^2 E = 0.3), when the flow reaches it, is the same.

Plain FP64 chain: the same number of sweeps of GLAC's RST update in FP64
from the shared start, flowed in FP64.  The two chains' plaquettes over their
last 500 sweeps, and E(t) of their final configurations at t = 1 and 2,
agree within statistical error, E's estimated per configuration by a
jackknife over blocks of time slices of the per-site energy density
(GLAC/observables/latticeactionchargedensity.cpp:169-306).  Longer flow times
probe modes that a few thousand local updates do not bring to equilibrium at
this lattice spacing (t^2 E(10) ~ 0.22 after 4000 sweeps from the cold
start, where the equilibrium value passes 0.3 near t = 7): there the faster
learned chain is measurably further along, so t >= 3 is printed, not
compared, and t0 is compared only between the FP32 and FP64 flows of one
configuration."""
import sys

import numpy as np
import pytest

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(900)]
glacx = pytest.importorskip("glac_amd")

N = (48, 48, 48, 96)
BETA, EPS, NHIT = 6.45, 0.24, 10
THERM, LEARNED = 3000, 1000
CHUNK, CHUNKS, DT = 50, 10, 0.02     # flow to t = 10 in 10 chunks of 50 steps
T_EQUIL = 2.0                         # flow times compared between the two chains
BLOCK_T = 8                           # time slices per jackknife block


def banner(msg):
    print(f"\n[gpu case] {msg}", flush=True)
    sys.stdout.flush()


def flow_in_chunks(L, field):
    """Flow CHUNKS x CHUNK steps measuring every step; after each chunk also
    the per-site energy density, summed per time slice.  Returns per-step
    (plaq, E_raw, Q) over the whole flow and the per-slice E at the chunk
    ends."""
    p, e, q, slices = [], [], [], []
    nt = N[3]
    for c in range(CHUNKS):
        pc, ec, qc, _ = L.flow(CHUNK, DT, field=field)
        sl = slice(0 if c == 0 else 1, None)
        p.append(pc[sl]), e.append(ec[sl]), q.append(qc[sl])
        _, ed, _, _, _ = L.measure_density(field)
        slices.append(ed.reshape(nt, -1).sum(axis=1) / (ed.size / nt))
    return np.concatenate(p), np.concatenate(e), np.concatenate(q), np.array(slices)


def jackknife_mean(per_slice):
    """Mean of the per-slice values and its jackknife error over blocks of
    BLOCK_T consecutive time slices."""
    blocks = per_slice.reshape(-1, BLOCK_T).mean(axis=1)
    nb = blocks.size
    loo = (blocks.sum() - blocks) / (nb - 1)
    return blocks.mean(), np.sqrt((nb - 1) / nb * ((loo - loo.mean()) ** 2).sum())


@pytest.fixture(scope="module")
def pipeline():
    from glac_amd.proposal import train_from_chain
    out = {}
    banner(f"pipeline {N}: generate + train (fp32c)")
    with glacx.Lattice(N, glacx.FP32_COMPRESSED) as L:
        assert L.device_bytes < 16 * 2 ** 30
        L.cold_start()
        acc0 = L.metropolis(THERM, BETA, EPS, NHIT, seed=45)
        U0 = L.download()
        loss = train_from_chain(L, BETA, seed=45, sweep_offset=THERM, collect_sweeps=4, samples=2_000_000,
                                steps=1000)
        banner(f"pipeline {N}: {LEARNED} learned sweeps")
        acc, plaq = L.metropolis_learned(LEARNED, BETA, EPS, NHIT, seed=46, sweep_offset=THERM + 4, plaquette=True)
        U = L.download()
        banner(f"pipeline {N}: flow fp32c to t = {CHUNK * CHUNKS * DT:g}")
        L.copy(glacx.FLOW, glacx.MAIN)
        out["fp32c"] = flow_in_chunks(L, glacx.FLOW)
    out.update(loss=loss, acc0=acc0, acc=acc, plaq=plaq)
    banner(f"pipeline {N}: flow fp64 of the same links")
    with glacx.Lattice(N, glacx.FP64) as L:
        L.upload(U, field=glacx.FLOW)
        del U
        out["fp64"] = flow_in_chunks(L, glacx.FLOW)
        banner(f"pipeline {N}: plain fp64 chain, {4 + LEARNED} RST sweeps from the shared start")
        L.upload(U0)
        del U0
        acc64, plaq64 = L.metropolis(4 + LEARNED, BETA, EPS, NHIT, seed=47, sweep_offset=THERM, plaquette=True)
        L.copy(glacx.FLOW, glacx.MAIN)
        banner(f"pipeline {N}: flow the plain fp64 configuration")
        out["plain64"] = flow_in_chunks(L, glacx.FLOW)
    out.update(acc64=acc64, plaq64=plaq64)
    return out


def test_generation_and_training(pipeline):
    loss, acc, acc0 = pipeline["loss"], pipeline["acc"], pipeline["acc0"]
    assert loss.samples >= 2_000_000
    assert np.all(np.isfinite(loss)) and loss[-50:].mean() < loss[:20].mean()
    assert 0.05 < acc.mean() < 0.95 and acc0.mean() < 0.5
    assert 0.5 < pipeline["plaq"][-1] < 1.0
    print(f"acceptance: RST {acc0[-100:].mean():.3f}, learned {acc.mean():.3f}; "
          f"plaquette learned {pipeline['plaq'][-1]:.6f}, plain fp64 {pipeline['plaq64'][-1]:.6f}")


def test_fp32_flow_matches_fp64_every_step(pipeline):
    from glac_amd.analysis import flow_scale
    p32, e32, q32, _ = pipeline["fp32c"]
    p64, e64, q64, _ = pipeline["fp64"]
    assert e32.size == CHUNK * CHUNKS + 1
    t = DT * np.arange(e32.size)
    E32, E64 = -e32 / 64, -e64 / 64
    assert np.all(np.diff(E64) < 0)
    de = float(np.abs(E32 / E64 - 1).max())
    dq = float(np.abs(q32 - q64).max())
    dp = float(np.abs(p32 - p64).max())
    t0_32, t0_64 = flow_scale(t, E32), flow_scale(t, E64)
    print(f"max |dE|/E {de:.2e}, max |dQ| {dq:.2e}, max |dplaq| {dp:.2e}, Q(t=10) {q64[-1]:.5f}, "
          f"t^2 E(10) {t[-1] ** 2 * E64[-1]:.4f}, t0 {t0_32:.5f} / {t0_64:.5f}")
    assert de < 1e-6, de
    assert dq < 1e-4, dq
    assert dp < 1e-6, dp
    assert np.isfinite(t0_32) == np.isfinite(t0_64)
    if np.isfinite(t0_64):
        assert abs(t0_32 - t0_64) < 1e-4 * t0_64


def test_learned_fp32_and_plain_fp64_configurations_agree(pipeline):
    from glac_amd.analysis import tau_int
    a, b = tau_int(pipeline["plaq"][-500:]), tau_int(pipeline["plaq64"][-500:])
    zp = abs(a["mean"] - b["mean"]) / np.hypot(a["err"], b["err"])
    print(f"plaquette over the last 500 sweeps: learned fp32 {a['mean']:.6f} +- {a['err']:.6f}, "
          f"plain fp64 {b['mean']:.6f} +- {b['err']:.6f}, z = {zp:.2f}")
    assert zp < 3.0
    _, e32, _, s32 = pipeline["fp32c"]
    _, e64p, _, s64p = pipeline["plain64"]
    for c in range(CHUNKS):
        t = DT * CHUNK * (c + 1)
        m1, d1 = jackknife_mean(s32[c])
        m2, d2 = jackknife_mean(s64p[c])
        # the per-slice means reproduce the global E_raw
        assert abs(m1 / e32[CHUNK * (c + 1)] - 1) < 1e-9 and abs(m2 / e64p[CHUNK * (c + 1)] - 1) < 1e-9
        z = abs(m1 - m2) / np.hypot(d1, d2)
        print(f"t = {t:g}: t^2 E learned fp32 {-t * t * m1 / 64:.5f} +- {t * t * d1 / 64:.5f}, "
              f"plain fp64 {-t * t * m2 / 64:.5f} +- {t * t * d2 / 64:.5f}, z = {z:.2f}")
        if t <= T_EQUIL + 1e-9:
            assert z < 3.0
