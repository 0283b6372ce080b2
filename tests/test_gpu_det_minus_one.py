"""GLAC's generateRandom hot start gives det U = -1 links
(GLAC/math/matrices/su3matrixgenerator.h:156-161), and GLAC's own Metropolis
(X U with X in SU(3)) keeps them there, so configurations GLAC wrote can hold
such links.  The modes that store three rows take them: the clover
measurement, the flow (with its FP32 reunitarisation, which must keep the
determinant) and the flow's fused measurement follow the oracle, which makes
no SU(3) assumption anywhere."""
import sys

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

glacx = pytest.importorskip("glac_amd")


def glac_started_field(n):
    U = O.new_field(n)
    O.glac_hot_start(U, n, O.GlacRng(3, 4), 0.24)
    for s in range(3):                       # GLAC's chain: no reunitarisation
        O.cb_sweep(U, n, 6.0, 0.24, 10, 7, s, False)
    return U.reshape(4, -1, 18)


@pytest.mark.parametrize("prec", [glacx.FP64, glacx.FP32_LINKS], ids=["fp64", "fp32"])
def test_det_minus_one_flow_and_measurement_follow_the_oracle(prec):
    n, steps, eps = (4, 4, 4, 8), 6, 0.02   # runs of 32 sites: FP32 measures inside stage 0
    print(f"\n[gpu case] det = -1 flow {n} prec {prec}", flush=True)
    sys.stdout.flush()
    U = glac_started_field(n)
    assert np.allclose(np.linalg.det(O.to_complex(U[:, :16]).reshape(-1, 3, 3)), -1.0, atol=1e-12)
    W = U.copy()
    rows = [O.measure(W, n)]
    for _ in range(steps):
        O.flow_step(W, n, eps)
        rows.append(O.measure(W, n))
    with glacx.Lattice(n, prec) as L:
        L.upload(U, glacx.FLOW)
        m0 = L.measure(glacx.FLOW)
        p, e, q, qt = L.flow(steps, eps)
        got = L.download(glacx.FLOW)
    wp = np.array([r[0] for r in rows])
    we = np.array([r[1] for r in rows])
    wq = np.array([r[2] for r in rows])
    det = np.linalg.det(O.to_complex(got).reshape(-1, 3, 3))
    if prec == glacx.FP64:
        tol_p, tol_e, tol_q, tol_u = 1e-12, 1e-11, 1e-12, 1e-11
    else:
        tol_p, tol_e, tol_q, tol_u = 5e-7, 1e-5, 1e-4, 5e-6
    assert abs(m0[0] - wp[0]) < tol_p and abs(m0[1] / we[0] - 1) < tol_e and abs(m0[2] - wq[0]) < tol_q
    assert np.abs(p - wp).max() < tol_p
    assert np.abs(e / we - 1).max() < tol_e
    assert np.abs(q - wq).max() < tol_q
    assert np.abs(got - W.reshape(4, -1, 18)).max() < tol_u
    assert np.abs(det + 1).max() < 10 * tol_u


@pytest.mark.parametrize("prec", [glacx.FP64, glacx.FP32_LINKS], ids=["fp64", "fp32"])
def test_det_minus_one_sweeps_with_reunitarisation_keep_det(prec):
    """Metropolis sweeps with reunitarisation on a det = -1 field keep
    det U = -1, and (FP64) make the oracle's accept/reject decisions on the
    same streams."""
    n = (4, 4, 4, 8)
    print(f"\n[gpu case] det = -1 sweeps with reunitarisation {n} prec {prec}", flush=True)
    sys.stdout.flush()
    U = glac_started_field(n)
    W = U.copy()
    accs = [O.cb_sweep(W, n, 6.0, 0.24, 10, 31, s, True) for s in range(3)]
    with glacx.Lattice(n, prec) as L:
        L.upload(U)
        got_acc = L.metropolis(3, 6.0, 0.24, 10, reunit_every=1, seed=31)
        got = L.download()
    det = np.linalg.det(O.to_complex(got).reshape(-1, 3, 3))
    want_det = np.linalg.det(O.to_complex(W).reshape(-1, 3, 3))
    assert np.abs(want_det + 1).max() < 1e-13
    if prec == glacx.FP64:
        assert np.abs(det + 1).max() < 1e-13
        nl = 4 * 10 * np.prod(n)
        assert np.array_equal(np.round(np.array(accs) * nl), np.round(got_acc * nl))
        assert np.abs(got - W).max() < 1e-10
    else:
        assert np.abs(det + 1).max() < 1e-5
        close = np.abs(got - W).reshape(4, -1, 18).max(axis=-1) < 1e-4
        assert close.mean() > 0.98
