"""GLAC's output files written from a device run, read back and checked
against the oracle.

The run is config 1 of GLAC's flow study as GLAC's System would do it on one
GPU: 8^3 x 16, beta = 6.0, Haar hot start, 200 FP64 Metropolis sweeps
(eps = 0.24, 10 hits), the configuration saved as GLAC's .bin
(IO::FieldIO::writeFieldToFile, GLAC/io/fieldio.cpp:61-122), then flowed 100
RK3 steps of 0.01 with the clover observables measured at every step
(System::flowConfiguration, GLAC/system.cpp:471-517) and written as GLAC's
per-configuration flow files (IO::writeFlowObservableToFile /
writeMatrixToFile, GLAC/io/observablesio.cpp:80-193), plus the per-site
densities (GLAC/io/fieldio.cpp:137-186).  Parsed back, every number agrees
with the oracle's flow of the same saved configuration to FP64 rounding of
the printed 15 decimals."""
import os
import sys

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
glacx = pytest.importorskip("glac_amd")
from glac_amd import observables_io as io  # noqa: E402

N = (8, 8, 8, 16)
BETA, NFLOW, DT = 6.0, 100, 0.01


def banner(msg):
    print(f"\n[gpu case] {msg}", flush=True)
    sys.stdout.flush()


def read_flow_file(path):
    with open(path) as f:
        lines = f.read().splitlines()
    head = dict(line.split(" ", 1) for line in lines[:3])
    rows = np.array([[float(v) for v in line.split()] for line in lines[3:]])
    return head, rows


def test_config1_run_writes_glac_files(tmp_path):
    banner(f"output files {N}: 200 FP64 sweeps, 100 flow steps")
    out, batch = str(tmp_path), "gpuRun"
    with glacx.Lattice(N, glacx.FP64) as L:
        L.hot_start(2024)
        acc = L.metropolis(200, BETA, 0.24, 10, reunit_every=1, seed=2025)
        binname = io.config_filename(batch, BETA, N[0], N[3], 0)
        binpath = os.path.join(out, batch, "field_configurations", binname)
        os.makedirs(os.path.dirname(binpath))
        L.save_bin(binpath)
        U = L.download()
        L.copy(glacx.FLOW, glacx.MAIN)
        plaq, e_raw, q, qt = L.flow(NFLOW, DT, field=glacx.FLOW)
        qd, ed, _, _, _ = L.measure_density(glacx.FLOW)
    paths = io.write_flow_config(out, batch, 0, plaq, e_raw, q, qt, BETA, DT)
    paths.append(io.write_scalar_field(out, batch, "topc", 0, qd, BETA, N[0], N[3]))
    paths.append(io.write_scalar_field(out, batch, "energy", 0, ed, BETA, N[0], N[3]))
    paths.append(io.write_observables(out, batch, "plaq", [plaq[0]], float(acc.mean()), BETA, 200, 1, 0))
    assert 0.05 < acc.mean() < 0.3

    # the .bin is the downloaded field in GLAC's byte layout: site-major, then mu
    raw = np.fromfile(binpath, dtype=np.float64)
    assert raw.size == 4 * np.prod(N) * 18
    assert np.array_equal(raw.reshape(-1, 4, 18).transpose(1, 0, 2), U)

    # the oracle flows the same configuration, read back from the .bin
    W = np.ascontiguousarray(raw.reshape(-1, 4, 18).transpose(1, 0, 2))
    rows_o = [O.measure(W, N)]
    for _ in range(NFLOW):
        O.flow_step(W, N, DT)
        rows_o.append(O.measure(W, N))
    want = np.array([r[:3] for r in rows_o])
    want_qt = np.array([r[3] for r in rows_o])

    for name, col in (("plaq", 0), ("topc", 2), ("energy", 1)):
        head, rows = read_flow_file(io.flow_file_path(out, batch, name, 0))
        assert head == {"beta": "6", "NFlows": str(NFLOW), "FlowEpsilon": "0.01"}
        assert rows.shape == (NFLOW + 1, 2)
        assert np.allclose(rows[:, 0], DT * np.arange(NFLOW + 1), atol=1e-15, rtol=0)
        tol = 1e-11 * max(1.0, np.abs(want[:, col]).max())
        err = np.abs(rows[:, 1] - want[:, col]).max()
        print(f"{name}: max |file - oracle| {err:.2e}")
        assert err < tol, (name, err)
    head, rows = read_flow_file(io.flow_file_path(out, batch, "topct", 0))
    assert rows.shape == (NFLOW + 1, 1 + N[3])
    assert np.abs(rows[:, 1:] - want_qt).max() < 1e-12
    assert np.allclose(rows[:, 1:].sum(axis=1), want[:, 2], atol=1e-11, rtol=0)
    # per-site densities of the flowed field, in GLAC's site order
    qd_file = np.fromfile(os.path.join(out, batch, "scalar_fields", "topc", binname), dtype=np.float64)
    ed_file = np.fromfile(os.path.join(out, batch, "scalar_fields", "energy", binname), dtype=np.float64)
    oq, oe = O.measure_density(W, N)
    assert np.abs(qd_file - oq).max() < 1e-12 and np.abs(ed_file - oe).max() < 1e-11 * np.abs(oe).max()
    with open(os.path.join(out, batch, "observables", f"plaq_{batch}.dat")) as f:
        lines = f.read().splitlines()
    assert lines[0] == "beta 6" and lines[2] == "NCor 200" and abs(float(lines[-1]) - want[0, 0]) < 1e-14
