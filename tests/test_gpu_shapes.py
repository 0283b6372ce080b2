"""Lattice shapes at the edges of the layout: the smallest lattice (2^4),
extents whose runs are not multiples of the 32-site chunk (6^4, 2x6x4x10),
mixed block sizes (12x8x4x6).  Every kernel's partial-block guards run here.
FP64 against the oracle (measurement, flow, same-stream Metropolis), and the
compressed FP32 mode against the same within FP32 tolerances."""
import sys

import numpy as np
import pytest

import oracle as O
from fieldtools import near_unit_field

pytestmark = pytest.mark.gpu
glacx = pytest.importorskip("glac_amd")

SHAPES = [(2, 2, 2, 2), (6, 6, 6, 6), (2, 6, 4, 10), (12, 8, 4, 6)]


def banner(msg):
    print(f"\n[gpu case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("n", SHAPES, ids=lambda n: "x".join(map(str, n)))
@pytest.mark.parametrize("prec", [glacx.FP64, glacx.FP32_COMPRESSED], ids=["fp64", "fp32c"])
def test_shapes_measure_flow_metropolis(n, prec):
    banner(f"shape {n} precision {prec}")
    fp64 = prec == glacx.FP64
    U = near_unit_field(n, np.random.default_rng(sum(n)), 0.7)
    want = O.measure(U, n)
    W = U.copy()
    flow_rows = [want]
    for _ in range(2):
        O.flow_step(W, n, 0.02)
        flow_rows.append(O.measure(W, n))
    M = O.hot_start_haar(n, 7)
    acc_o = O.cb_sweep(M, n, 6.0, 0.24, 10, 8, 0, True)
    with glacx.Lattice(n, prec) as L:
        L.upload(U)
        got = L.measure()
        L.copy(glacx.FLOW, glacx.MAIN)
        p, e, q, qt = L.flow(2, 0.02)
        L.hot_start(7)
        acc = L.metropolis(1, 6.0, 0.24, 10, seed=8)[0]
        chain = L.download()
    tol = 1e-11 if fp64 else 2e-5
    assert abs(got[0] - want[0]) < tol and abs(got[1] - want[1]) < tol * max(1.0, abs(want[1]))
    assert abs(got[2] - want[2]) < (1e-12 if fp64 else 1e-4) and np.abs(got[3] - want[3]).max() < (1e-12 if fp64 else 1e-4)
    for i, r in enumerate(flow_rows):
        assert abs(p[i] - r[0]) < tol and abs(e[i] / r[1] - 1) < (1e-11 if fp64 else 1e-5)
    nlinks = 4 * 10 * np.prod(n)
    if fp64:
        assert round(acc * nlinks) == round(acc_o * nlinks)
        assert np.abs(chain - M).max() < 1e-10
    else:
        assert abs(acc - acc_o) * nlinks <= 3
