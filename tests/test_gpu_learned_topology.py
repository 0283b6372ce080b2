"""The learned proposal samples the topological charge as GLAC's fixed-eps
proposal does (no GLAC counterpart for the learned proposal; GLAC's update:
System::update, GLAC/system.cpp:333-361, RST proposal
GLAC/math/matrices/su3matrixgenerator.h:170-243).

16^4 at beta = 6.0, compressed FP32 links for both chains, where Q tunnels
within a few hundred sweeps of either chain: a Haar hot start thermalised by
2000 RST sweeps, 40 sweeps of (staple, link) samples (3.5e6) train the
network, then the same chain runs SWEEPS sweeps with GLAC's proposal and
SWEEPS with the learned one, Q and E flowed to t = 1 (50 RK3 steps of 0.02)
every NCOR sweeps.  <Q^2> and <E(t = 1)> of the two chains agree within
three autocorrelation-corrected standard errors (Wolff's Gamma method);
tau_int(Q^2) of both chains is reported (one MI355X run: 9.6 +- 4.1 for
the fixed-eps chain, 2.6 +- 0.7 for the learned one, in units of 50
sweeps)."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
glacx = pytest.importorskip("glac_amd")

N = (16, 16, 16, 16)
BETA, EPS, NHIT = 6.0, 0.24, 10
THERM, SWEEPS, NCOR = 2000, 40000, 50


def banner(msg):
    print(f"\n[gpu case] {msg}", flush=True)
    sys.stdout.flush()


def test_learned_chain_samples_the_topological_charge():
    from glac_amd.analysis import tau_int
    from glac_amd.proposal import train_from_chain
    from glac_amd.system import EnsembleConfig, run_ensemble
    runs = {}
    with glacx.Lattice(N, glacx.FP32_COMPRESSED) as L:
        banner(f"learned vs fixed eps topology {N}: thermalise and train")
        L.hot_start(7)
        L.metropolis(THERM, BETA, EPS, NHIT, seed=7)
        loss = train_from_chain(L, BETA, seed=7, sweep_offset=THERM, collect_sweeps=40, samples=3_000_000,
                                steps=3000)
        assert loss.samples >= 3_000_000 and loss[-100:].mean() < loss[:20].mean()
        off = THERM + 40
        for name, seed in (("rst", 8), ("learned", 9)):
            banner(f"learned vs fixed eps topology {N}: {SWEEPS} sweeps {name}")
            cfg = EnsembleConfig(dims=N, beta=BETA, NCf=SWEEPS // NCOR, NCor=NCOR, NUpdates=NHIT, SU3Eps=EPS,
                                 NFlows=50, flowEpsilon=0.02, seed=seed, proposal=name)
            runs[name] = run_ensemble(L, cfg, thermalise=False, sweep_offset=off)
            off += SWEEPS
    r, l = runs["rst"], runs["learned"]
    for what, a, b in (("Q^2", r.q ** 2, l.q ** 2), ("E(t = 1)", r.e, l.e)):
        s, t = tau_int(a), tau_int(b)
        z = abs(s["mean"] - t["mean"]) / np.hypot(s["err"], t["err"])
        print(f"{what}: fixed eps {s['mean']:.6g} +- {s['err']:.2g} (tau_int {s['tau_int']:.1f} +- {s['dtau']:.1f}), "
              f"learned {t['mean']:.6g} +- {t['err']:.2g} (tau_int {t['tau_int']:.1f} +- {t['dtau']:.1f}), z = {z:.2f}")
        assert z < 3.0, (what, s, t)
    for run in (r, l):                      # <Q> = 0: both chains visit both signs
        q = tau_int(run.q)
        assert abs(q["mean"]) < 3.0 * q["err"], q
    assert 0.05 < l.acceptance < 0.95
