"""Clover observables measured inside each flow step's first RK3 stage
(lattices whose z-block runs are whole 32-site chunks, FP32 and FP64 links)
against the clover kernel, which measures between steps otherwise.

Row 0 of a measured flow is stage 0's measurement of the starting field, so
it sees the very links the clover kernel sees: only the order of the sums
differs.  Later rows follow fields that differ by the rounding of the staple
sum (the fused stage keeps the upper and lower staple terms apart)."""
import numpy as np
import pytest

import glac_amd as G

pytestmark = pytest.mark.gpu


def _rows_by_clover_kernel(L, steps, eps):
    rows = []
    for i in range(steps + 1):
        rows.append(L.measure(G.FLOW))
        if i < steps:
            L.flow(1, eps, field=G.FLOW, measure=False)
    return rows


@pytest.mark.parametrize("prec", [G.FP32_LINKS, G.FP32_COMPRESSED, G.FP64], ids=["fp32", "fp32c", "fp64"])
@pytest.mark.parametrize("dims", [(8, 8, 8, 16), (16, 8, 12, 8), (32, 16, 8, 4)], ids=str)
def test_fused_measurement_matches_clover_kernel(prec, dims):
    print(f"[case] fused measure {dims} prec {prec}", flush=True)
    steps, eps = 8, 0.02
    L = G.Lattice(dims, prec)
    L.hot_start(11)
    L.metropolis(15, 6.0, 0.24, 10, seed=2)
    L.copy(G.FLOW, G.MAIN)
    L.timing(True)
    L.timing_reset()
    plaq, e, q, qt = L.flow(steps, eps, field=G.FLOW)
    fused_launches, _ = L.timing_read(G.K_FLOW_STAGE_MEASURE)
    clover_launches, _ = L.timing_read(G.K_MEASURE)
    L.timing(False)
    assert fused_launches == steps and clover_launches == 1
    L.copy(G.FLOW, G.MAIN)
    ref = _rows_by_clover_kernel(L, steps, eps)
    L.close()
    # row 0: the same links
    p0, e0, q0, qt0 = ref[0]
    f = 1e-6 if prec == G.FP64 else 1.0   # FP64: the FP32 bounds scaled towards FP64 rounding
    assert abs(plaq[0] - p0) < 2e-7 * f
    assert abs(e[0] - e0) < 2e-6 * f * abs(e0)
    assert abs(q[0] - q0) < 2e-5 * f * max(1.0, np.abs(qt0).sum())
    assert np.abs(qt[0] - qt0).max() < 2e-5 * f * max(1.0, np.abs(qt0).max())
    for i in range(1, steps + 1):
        pi, ei, qi, qti = ref[i]
        assert abs(plaq[i] - pi) < 1e-6 * f
        assert abs(e[i] - ei) < 1e-5 * f * abs(ei)
        assert np.abs(qt[i] - qti).max() < 1e-4 * f * max(1.0, np.abs(qti).max())


def test_unfusable_shape_uses_the_clover_kernel():
    for dims, prec in [((6, 6, 6, 6), G.FP32_COMPRESSED), ((6, 6, 6, 6), G.FP64)]:
        print(f"[case] clover-kernel path {dims} prec {prec}", flush=True)
        L = G.Lattice(dims, prec)
        L.hot_start(3, field=G.FLOW)
        L.timing(True)
        L.timing_reset()
        L.flow(3, 0.02, field=G.FLOW)
        assert L.timing_read(G.K_FLOW_STAGE_MEASURE)[0] == 0
        assert L.timing_read(G.K_MEASURE)[0] == 4
        L.close()
