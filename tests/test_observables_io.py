"""GLAC's output files (flow observables, Q(t) matrix, per-run observables,
configuration names) as written by glac_amd.observables_io.  The number
formatting is checked against the C++ library's streams (a small program
built here with g++), the layout against the structure GLAC's writers
produce (GLAC/io/observablesio.cpp:29-193), and the statistics against
their definition (GLAC/observables/tools/observablestorer.cpp:77-98)."""
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

from glac_amd import observables_io as io

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
VALUES = ["6", "6.1", "0.01", "0.02", "1e-05", "0.5958634", "-0.123456789012345678", "123456.789",
          "1234567.0", "1e-20", "-3.25e+30", "0", "-0", "2.5", "0.000123456789", "1e+100", "0.1"]


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_number_formatting_matches_cpp_streams(tmp_path):
    exe = str(tmp_path / "fmt")
    subprocess.run(["g++", "-O1", "-std=c++11", os.path.join(ROOT, "tests", "stream_format_check.cpp"), "-o", exe],
                   check=True)
    subprocess.run([exe, str(tmp_path)] + VALUES, check=True)
    vals = [float(v) for v in VALUES]
    assert (tmp_path / "default.txt").read_text() == "".join(io._g(v) + "\n" for v in vals)
    assert (tmp_path / "fixed15.txt").read_text() == "".join(io._f15(v) + "\n" for v in vals)
    assert (tmp_path / "to_string.txt").read_text() == "".join("%f\n" % v for v in vals)


def test_flow_file_layout(tmp_path):
    plaq = [0.6, 0.7, 0.75]
    path = io.write_flow_observable(str(tmp_path), "b1", "plaq", 3, plaq, 6.1, 0.01, config_start=10)
    assert path.endswith(os.path.join("b1", "flow_observables", "plaq", "b1_plaq_flow_config00013.dat"))
    lines = open(path).read().splitlines()
    assert lines[:3] == ["beta 6.1", "NFlows 2", "FlowEpsilon 0.01"]
    assert lines[3:] == ["0.000000000000000 0.600000000000000", "0.010000000000000 0.700000000000000",
                         "0.020000000000000 0.750000000000000"]


def test_topct_matrix_and_config_set(tmp_path):
    qt = np.arange(12, dtype=float).reshape(3, 4) / 7.0
    paths = io.write_flow_config(str(tmp_path), "run", 0, [1, 2, 3], [4, 5, 6], [7, 8, 9], qt, 6.0, 0.02)
    names = [os.path.basename(p) for p in paths]
    assert names == ["run_plaq_flow_config00000.dat", "run_topc_flow_config00000.dat",
                     "run_energy_flow_config00000.dat", "run_topct_flow_config00000.dat"]
    lines = open(paths[3]).read().splitlines()
    assert lines[:3] == ["beta 6", "NFlows 2", "FlowEpsilon 0.02"]
    row = lines[4].split(" ")
    assert row[0] == "0.020000000000000" and len(row) == 5
    assert [float(v) for v in row[1:]] == pytest.approx(qt[1], abs=1e-15)


def test_run_observables_and_statistics(tmp_path):
    vals = [0.59, 0.6, 0.61, 0.605]
    avg, var, std = io.statistics(vals)
    n = len(vals)
    assert avg == pytest.approx(np.mean(vals), rel=1e-15)
    assert var == pytest.approx(np.var(vals) / n, rel=1e-9)
    assert std == pytest.approx(math.sqrt(np.var(vals) / n), rel=1e-9)
    path = io.write_observables(str(tmp_path), "run", "plaq", vals, 0.1079, 6.0, 200, 4, 10)
    assert path.endswith(os.path.join("run", "observables", "plaq_run.dat"))
    lines = open(path).read().splitlines()
    assert lines[:5] == ["beta 6", "acceptanceCounter 0.1079", "NCor 200", "NCf 4", "NTherm 10"]
    assert lines[5].startswith("Averageplaq ") and lines[6].startswith("Varianceplaq ")
    assert lines[7].startswith("stdplaq ") and len(lines) == 8 + n
    assert lines[8] == "0.590000000000000"


def test_config_filename():
    assert io.config_filename("b", 6.0, 8, 16, 7) == "b_b6.000000_N8_NT16_np1_config00007.bin"


def test_scalar_field_file(tmp_path):
    vals = np.arange(4 * 4 * 4 * 8, dtype=float) / 3.0
    path = io.write_scalar_field(str(tmp_path), "run", "topc", 2, vals, 6.0, 4, 8)
    assert path.endswith(os.path.join("run", "scalar_fields", "topc", "run_b6.000000_N4_NT8_np1_config00002.bin"))
    assert np.array_equal(np.fromfile(path, dtype=np.float64), vals)
