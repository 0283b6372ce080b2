"""GLAC's observable output files, written from the device results.

GLAC's downstream analysis reads per-configuration flow files and per-run
observable files; these functions write the same text, byte for byte, so a
GPU run can replace a GLAC run without touching that analysis:

  <out>/<batch>/flow_observables/<name>/<batch>_<name>_flow_config%05d.dat
      "beta <b>" / "NFlows <n>" / "FlowEpsilon <e>" then one row per flow
      time "t value" (IO::writeFlowObservableToFile,
      GLAC/io/observablesio.cpp:80-126), or "t v_0 ... v_{N-1}" for the
      Q(t) matrix (IO::writeMatrixToFile, :141-193);
  <out>/<batch>/observables/<name>_<batch>.dat
      beta, acceptance, NCor, NCf, NTherm, the average, variance and standard
      deviation, then one value per configuration (IO::writeObservablesToFile,
      :29-66, with ObservableStorer::runStatistics,
      GLAC/observables/tools/observablestorer.cpp:77-98).

  <out>/<batch>/scalar_fields/<observable>/<config name>.bin
      per-site densities (topc, energy) as raw doubles in GLAC's site order
      (IO::FieldIO::writeDoublesFieldToFile, GLAC/io/fieldio.cpp:137-186).

Number formatting follows the C++ streams GLAC writes with: header values
in the default format (6 significant digits, like printf's %g), rows in
std::fixed with 15 decimals (printf's %.15f).  Observable names are GLAC's
("plaq", "topc", "energy", "topct", GLAC/observables/mastersamplertopcxyz.cpp:
68-71); the energy written is E_raw / V as GLAC stores it (physical E is
-E_raw / 64).  Configuration .bin names follow IO::FieldIO::writeFieldToFile
(GLAC/io/fieldio.cpp:71-79).
"""
from __future__ import annotations

import math
import os
from typing import Sequence


def _g(x) -> str:
    """A double or integer as a default-formatted C++ stream prints it."""
    if isinstance(x, int):
        return str(x)
    return "%g" % x


def _f15(x: float) -> str:
    return "%.15f" % x


def _header(beta: float, nflows: int, flow_eps: float) -> str:
    return f"beta {_g(float(beta))}\nNFlows {int(nflows)}\nFlowEpsilon {_g(float(flow_eps))}\n"


def flow_file_path(out: str, batch: str, name: str, config: int, config_start: int = 0) -> str:
    return os.path.join(out, batch, "flow_observables", name,
                        f"{batch}_{name}_flow_config{config + config_start:05d}.dat")


def write_flow_observable(out: str, batch: str, name: str, config: int, values: Sequence[float], beta: float,
                          flow_eps: float, config_start: int = 0) -> str:
    """One flow observable of one configuration: values[i] at flow time
    i * flow_eps, i = 0..NFlows."""
    path = flow_file_path(out, batch, name, config, config_start)
    os.makedirs(os.path.dirname(path), exist_ok=True)
    nflows = len(values) - 1
    with open(path, "w") as f:
        f.write(_header(beta, nflows, flow_eps))
        for i, v in enumerate(values):
            f.write(f"{_f15(i * float(flow_eps))} {_f15(float(v))}\n")
    return path


def write_flow_matrix(out: str, batch: str, name: str, config: int, rows: Sequence[Sequence[float]], beta: float,
                      flow_eps: float, config_start: int = 0) -> str:
    """A matrix flow observable (Q(t): one row of Nt values per flow time)."""
    path = flow_file_path(out, batch, name, config, config_start)
    os.makedirs(os.path.dirname(path), exist_ok=True)
    nflows = len(rows) - 1
    with open(path, "w") as f:
        f.write(_header(beta, nflows, flow_eps))
        for i, row in enumerate(rows):
            f.write(_f15(i * float(flow_eps)) + "".join(" " + _f15(float(v)) for v in row) + "\n")
    return path


def write_flow_config(out: str, batch: str, config: int, plaq, e_raw, q, q_t, beta: float, flow_eps: float,
                      config_start: int = 0) -> list:
    """The four files GLAC writes per flowed configuration with the
    MasterSamplerTopcXYZ observables (GLAC/observables/mastersamplertopcxyz.cpp:
    75-94), from the arrays Lattice.flow returns."""
    return [write_flow_observable(out, batch, "plaq", config, plaq, beta, flow_eps, config_start),
            write_flow_observable(out, batch, "topc", config, q, beta, flow_eps, config_start),
            write_flow_observable(out, batch, "energy", config, e_raw, beta, flow_eps, config_start),
            write_flow_matrix(out, batch, "topct", config, q_t, beta, flow_eps, config_start)]


def statistics(values: Sequence[float]):
    """(average, variance of the average, standard deviation) as GLAC's
    ObservableStorer::runStatistics computes them: plain sequential sums,
    variance (<x^2> - <x>^2) / N."""
    n = len(values)
    s = s2 = 0.0
    for v in values:
        v = float(v)
        s += v
        s2 += v * v
    avg, avg2 = s / n, s2 / n
    var = (avg2 - avg * avg) / n
    return avg, var, math.sqrt(var) if var >= 0.0 else float("nan")


def write_observables(out: str, batch: str, name: str, values: Sequence[float], acceptance: float, beta: float,
                      ncor: int, ncf: int, ntherm: int) -> str:
    """A per-run observable file (one value per configuration)."""
    path = os.path.join(out, batch, "observables", f"{name}_{batch}.dat")
    os.makedirs(os.path.dirname(path), exist_ok=True)
    avg, var, std = statistics(values)
    with open(path, "w") as f:
        f.write(f"beta {_g(float(beta))}\nacceptanceCounter {_g(float(acceptance))}\n"
                f"NCor {int(ncor)}\nNCf {int(ncf)}\nNTherm {int(ntherm)}\n")
        f.write(f"Average{name} {_f15(avg)}\nVariance{name} {_f15(var)}\nstd{name} {_f15(std)}\n")
        for v in values:
            f.write(_f15(float(v)) + "\n")
    return path


def write_scalar_field(out: str, batch: str, observable: str, config: int, values, beta: float, n_spatial: int,
                       n_temporal: int, num_proc: int = 1, config_start: int = 0) -> str:
    """A per-site density (V doubles, GLAC's site order, native-endian, no
    header) as IO::FieldIO::writeDoublesFieldToFile writes it
    (GLAC/io/fieldio.cpp:137-186) under <out>/<batch>/scalar_fields/<observable>/."""
    import numpy as np
    path = os.path.join(out, batch, "scalar_fields", observable,
                        config_filename(batch, beta, n_spatial, n_temporal, config, num_proc, config_start))
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.ascontiguousarray(values, dtype=np.float64).tofile(path)
    return path


def config_filename(batch: str, beta: float, n_spatial: int, n_temporal: int, config: int, num_proc: int = 1,
                    config_start: int = 0) -> str:
    """GLAC's .bin name; beta as std::to_string prints a double (%f)."""
    return (f"{batch}_b{'%f' % beta}_N{n_spatial}_NT{n_temporal}_np{num_proc}"
            f"_config{config + config_start:05d}.bin")
