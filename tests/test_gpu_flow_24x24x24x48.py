"""Full-size flow parity, the gradient-flow workload at its real shape:
a 24^3 x 48 configuration thermalised from a seed (500 FP64 Metropolis
sweeps at beta = 6.0), flowed to t = 10 (1000 RK3 steps of 0.01) in FP64 and
in the two FP32 link modes; at every flow step E, t^2 E, Q and Q(t) of the
FP32 flows must agree with the FP64 flow within |dE|/E < 1e-6 and
|dQ| < 1e-4.  (The FP64 device flow itself follows the FP64 oracle to
1e-12 on smaller lattices, test_gpu_parity.py.)"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
glacx = pytest.importorskip("glac_amd")

N = (24, 24, 24, 48)


def banner(msg):
    print(f"\n[gpu case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def flows(tmp_path_factory):
    path = str(tmp_path_factory.mktemp("cfg") / "thermalised.bin")
    banner(f"thermalise {N}")
    with glacx.Lattice(N, glacx.FP64) as L:
        L.hot_start(6)
        L.metropolis(500, 6.0, 0.24, 10, seed=6)
        L.save_bin(path)
    out = {}
    for prec in (glacx.FP64, glacx.FP32_LINKS, glacx.FP32_COMPRESSED):
        banner(f"flow {N} 1000 steps precision {prec}")
        with glacx.Lattice(N, prec) as L:
            L.load_bin(path, field=glacx.FLOW)
            out[prec] = L.flow(1000, 0.01, field=glacx.FLOW)
    return out


@pytest.mark.parametrize("prec", [glacx.FP32_LINKS, glacx.FP32_COMPRESSED], ids=["fp32", "fp32c"])
def test_fp32_flow_tracks_fp64_to_t10(flows, prec):
    p0, e0, q0, qt0 = flows[glacx.FP64]
    p, e, q, qt = flows[prec]
    t = 0.01 * np.arange(1001)
    E0, E = -e0 / 64, -e / 64
    assert np.all(E0 > 0) and E0[-1] < E0[0]                 # flowed towards smoothness
    assert np.abs(E / E0 - 1).max() < 1e-6                    # |dE|/E, hence t^2 E too
    assert np.abs(t ** 2 * (E - E0)).max() < 1e-6 * np.abs(t ** 2 * E0).max()
    assert np.abs(q - q0).max() < 1e-4
    assert np.abs(qt - qt0).max() < 1e-4
    assert np.abs(p - p0).max() < 1e-6
    assert abs(q0[-1] - round(q0[-1])) < 0.1                  # Q near an integer at t = 10
