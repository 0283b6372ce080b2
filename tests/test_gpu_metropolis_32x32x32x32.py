"""Plain Metropolis generation at 32^4, beta = 6.1, as GLAC's
System::thermalize / runMetropolis run it (GLAC/system.cpp:262-319, 333-361):
Haar hot start, 2000 checkerboard sweeps of GLAC's fixed-eps RST proposal
(eps = 0.24, 10 hits per link, acceptance as GLAC/system.cpp:359), links
reunitarised every sweep, plaquette measured after every sweep.  Three
precisions run their own chain (different streams): FP64, FP32 links with
FP64 sums, and compressed FP32 links.  The equilibrium plaquette and
acceptance of each FP32 chain agree with the FP64 chain within the
autocorrelation-corrected statistical error (Wolff's Gamma method), and the
links stay in SU(3)."""
import sys

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
glacx = pytest.importorskip("glac_amd")

N = (32, 32, 32, 32)
BETA, EPS, NHIT, SWEEPS, THERM = 6.1, 0.24, 10, 2000, 500
PRECS = {"fp64": glacx.FP64, "fp32": glacx.FP32_LINKS, "fp32c": glacx.FP32_COMPRESSED}


def banner(msg):
    print(f"\n[gpu case] {msg}", flush=True)
    sys.stdout.flush()


def su3_defects(U, sample=20000):
    M = O.to_complex(U.reshape(-1, 18)[::max(1, U.size // 18 // sample)])
    unit = np.abs(M @ np.conj(np.swapaxes(M, -1, -2)) - np.eye(3)).max()
    det = np.abs(np.linalg.det(M) - 1).max()
    return unit, det


@pytest.fixture(scope="module")
def chains():
    out = {}
    for k, (name, prec) in enumerate(PRECS.items()):
        banner(f"metropolis {N} beta {BETA}: {SWEEPS} sweeps {name}")
        with glacx.Lattice(N, prec) as L:
            L.hot_start(610 + k)
            acc, plaq = L.metropolis(SWEEPS, BETA, EPS, NHIT, reunit_every=1, seed=620 + k, plaquette=True)
            U = L.download()
        out[name] = (acc, plaq, su3_defects(U))
        del U
    return out


def test_series_are_sane(chains):
    from glac_amd.analysis import tau_int
    for name, (acc, plaq, (unit, det)) in chains.items():
        assert np.all(np.isfinite(acc)) and np.all(np.isfinite(plaq))
        # relaxation from the hot start: the plaquette climbs from ~0 and then stays put
        assert plaq[0] < 0.3 and plaq[-1] > 0.55
        t = tau_int(plaq[THERM:])
        print(f"{name}: <plaq> {t['mean']:.6f} +- {t['err']:.6f} (tau {t['tau_int']:.1f}), "
              f"<acc> {acc[THERM:].mean():.5f}, |UU^+-1| {unit:.2e}, |det-1| {det:.2e}")
        tol = 1e-13 if name == "fp64" else 2e-6
        assert unit < tol and det < tol


@pytest.mark.parametrize("name", ["fp32", "fp32c"])
def test_fp32_chains_match_fp64_within_errors(chains, name):
    from glac_amd.analysis import tau_int
    a64, p64, _ = chains["fp64"]
    a32, p32, _ = chains[name]
    for what, x, y in (("plaquette", p32, p64), ("acceptance", a32, a64)):
        s, r = tau_int(x[THERM:]), tau_int(y[THERM:])
        sigma = np.hypot(s["err"], r["err"])
        z = abs(s["mean"] - r["mean"]) / sigma
        print(f"{name} vs fp64 {what}: {s['mean']:.6f} vs {r['mean']:.6f}, sigma {sigma:.2e}, z = {z:.2f}")
        assert z < 3.0, (what, s, r)
