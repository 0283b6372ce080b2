"""Host-side drivers without a GPU: GLAC's run loop (glac_amd/system.py) on
a recording stand-in of the Lattice binding, and the proposal parameter
initialisation (glac_amd/proposal.py)."""
import numpy as np
import pytest

from glac_amd import FLOW, MAIN
from glac_amd.proposal import N_PARAM, init_params
from glac_amd.system import EnsembleConfig, run_ensemble


class FakeLattice:
    """Records the calls run_ensemble makes; observables are scripted."""

    def __init__(self, q_series):
        self.calls = []
        self.q = list(q_series)

    def metropolis(self, n, beta, eps, n_hit, reunit_every=1, seed=1, sweep_offset=0, plaquette=False):
        self.calls.append(("rst", n, sweep_offset, seed))
        return np.full(n, 0.1)

    def metropolis_learned(self, n, beta, eps, n_hit, reunit_every=1, seed=1, sweep_offset=0, plaquette=False):
        self.calls.append(("learned", n, sweep_offset, seed))
        return np.full(n, 0.25)

    def synchronize(self):
        pass

    def copy(self, dst, src):
        self.calls.append(("copy", dst, src))

    def flow(self, n, eps, field=FLOW, measure=True):
        self.calls.append(("flow", n, eps, field, measure))

    def plaquette(self, field=MAIN):
        return 0.6

    def measure(self, field=MAIN):
        return 0.9, -64.0 * 0.01, self.q.pop(0), None


def test_run_loop_follows_glac_and_continues_the_streams():
    qs = [1.0, -2.0, 0.5, 3.0]
    L = FakeLattice(qs)
    cfg = EnsembleConfig(NTherm=7, NCf=4, NCor=5, NFlows=3, flowEpsilon=0.02, seed=11, proposal="learned")
    r = run_ensemble(L, cfg, sweep_offset=100)
    upd = [c for c in L.calls if c[0] in ("rst", "learned")]
    # thermalisation then NCf blocks of NCor sweeps, consecutive sweep indices
    assert upd[0] == ("learned", 7, 100, 11)
    assert [c[2] for c in upd[1:]] == [107, 112, 117, 122] and all(c[1] == 5 for c in upd[1:])
    # each configuration: copy chain -> flow field, flow NFlows steps, measure
    assert L.calls.count(("copy", FLOW, MAIN)) == 4
    assert L.calls.count(("flow", 3, 0.02, FLOW, False)) == 4
    assert np.array_equal(r.q, qs) and np.allclose(r.e, 0.01) and np.allclose(r.plaq, 0.6)
    assert r.acceptance == pytest.approx(0.25) and r.sweeps_done == 127


def test_run_loop_rejects_unknown_proposal():
    with pytest.raises(ValueError):
        run_ensemble(FakeLattice([]), EnsembleConfig(proposal="heatbath"))


def test_init_params_shape_and_isotropic_start():
    p = init_params(seed=3, step=0.2)
    assert p.shape == (N_PARAM,) and p.dtype == np.float32 and N_PARAM == 2468
    b3 = p[-36:]
    assert np.allclose(b3[:8], np.log(0.2)) and np.all(b3[8:] == 0.0)
    assert np.array_equal(init_params(seed=3, step=0.2), p)
