"""Ensemble driver: GLAC's System::runMetropolis loop on one resident lattice.

Mirrors GLAC/system.cpp:371-470 (thermalise, then NCf times: NCor updates,
copy the chain to the flow lattice, flow it, record the observables) and
System::flowConfiguration (GLAC/system.cpp:471-517), at the granularity of the
C ABI: a block of NCor sweeps is one call, a flow of NFlows steps is one call.
The update is either GLAC's fixed-eps RST proposal (System::update,
GLAC/system.cpp:333-361) or the learned staple-conditioned proposal.

Timing follows GLAC's own bookkeeping: update time and flow time are wall
clock around the calls, after a device synchronisation.
"""
import time
from dataclasses import dataclass, field

import numpy as np

from . import FLOW, MAIN, Lattice
from .analysis import tau_int


@dataclass
class EnsembleConfig:
    """Keys as in GLAC's JSON configuration (GLAC/config/configloader.cpp:55-181)."""
    dims: tuple = (32, 32, 32, 32)
    beta: float = 6.1
    NTherm: int = 200
    NCf: int = 100
    NCor: int = 20
    NUpdates: int = 10
    SU3Eps: float = 0.24
    NFlows: int = 100
    flowEpsilon: float = 0.01
    seed: int = 1
    proposal: str = "rst"          # "rst" (GLAC) or "learned"
    reunit_every: int = 1
    sweep_plaquette: bool = False  # also record the chain plaquette after every sweep


@dataclass
class EnsembleResult:
    plaq: np.ndarray = field(default_factory=lambda: np.empty(0))      # chain plaquette per config
    q: np.ndarray = field(default_factory=lambda: np.empty(0))         # Q at flow time NFlows * eps
    e: np.ndarray = field(default_factory=lambda: np.empty(0))         # E = -E_raw / 64 at that time
    acceptance: float = 0.0
    sweep_plaq: np.ndarray = field(default_factory=lambda: np.empty(0))  # per sweep (sweep_plaquette)
    sweep_acc: np.ndarray = field(default_factory=lambda: np.empty(0))   # acceptance per sweep
    update_seconds: float = 0.0
    flow_seconds: float = 0.0
    therm_seconds: float = 0.0

    @property
    def configs_per_second(self):
        return len(self.q) / (self.update_seconds + self.flow_seconds)

    def tau_int_q2(self):
        return tau_int(self.q ** 2)

    def independent_configs_per_second(self):
        """configs/s over 2 tau_int(Q^2), tau_int in units of the config spacing."""
        return self.configs_per_second / (2.0 * self.tau_int_q2()["tau_int"])


def run_ensemble(L: Lattice, cfg: EnsembleConfig, thermalise=True, sweep_offset=0, progress=None):
    """Generate cfg.NCf configurations on L's chain field; flow each one for
    cfg.NFlows steps on the flow field and record Q and E at the end.

    The chain field must already hold the starting configuration (hot or cold
    start, or a loaded .bin).  With a learned proposal the context's network
    must be set (Lattice.proposal_set)."""
    update = L.metropolis_learned if cfg.proposal == "learned" else L.metropolis
    if cfg.proposal not in ("rst", "learned"):
        raise ValueError(f"unknown proposal {cfg.proposal!r}")
    res = EnsembleResult()
    sweep = sweep_offset
    if thermalise and cfg.NTherm > 0:
        L.synchronize()
        t0 = time.perf_counter()
        update(cfg.NTherm, cfg.beta, cfg.SU3Eps, cfg.NUpdates, reunit_every=cfg.reunit_every,
               seed=cfg.seed, sweep_offset=sweep)
        L.synchronize()
        res.therm_seconds = time.perf_counter() - t0
        sweep += cfg.NTherm
    plaq, q, e, acc, splaq = [], [], [], [], []
    for i in range(cfg.NCf):
        t0 = time.perf_counter()
        a = update(cfg.NCor, cfg.beta, cfg.SU3Eps, cfg.NUpdates, reunit_every=cfg.reunit_every,
                   seed=cfg.seed, sweep_offset=sweep, plaquette=cfg.sweep_plaquette)
        if cfg.sweep_plaquette:
            a, p_sweeps = a
            splaq.append(p_sweeps)
        L.synchronize()
        t1 = time.perf_counter()
        sweep += cfg.NCor
        L.copy(FLOW, MAIN)
        L.flow(cfg.NFlows, cfg.flowEpsilon, field=FLOW, measure=False)
        p_chain = L.plaquette(MAIN)
        p, e_raw, qq, _ = L.measure(FLOW)
        t2 = time.perf_counter()
        res.update_seconds += t1 - t0
        res.flow_seconds += t2 - t1
        plaq.append(p_chain)
        q.append(qq)
        e.append(-e_raw / 64.0)
        acc.append(a)
        if progress is not None:
            progress(i, qq)
    res.plaq, res.q, res.e = np.array(plaq), np.array(q), np.array(e)
    res.sweep_acc = np.concatenate(acc) if acc else np.empty(0)
    res.sweep_plaq = np.concatenate(splaq) if splaq else np.empty(0)
    res.acceptance = float(res.sweep_acc.mean()) if acc else 0.0
    res.sweeps_done = sweep
    return res
