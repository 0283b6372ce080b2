"""Autocorrelation analysis of Markov-chain time series.

Integrated autocorrelation time by U. Wolff's Gamma method with automatic
windowing ("Monte Carlo errors with less errors", Comput. Phys. Commun. 156
(2004) 143-153, eqs. 31-35 and the window condition of section 3.3).  GLAC
left this analysis to its external LatticeAnalyser; the build needs it for
the headline metric, independent configurations per second =
(configurations per second) / (2 tau_int).
"""
from __future__ import annotations

import math

import numpy as np


def autocorrelation(a: np.ndarray, tmax: int | None = None) -> np.ndarray:
    """Normalised autocorrelation rho(t), t = 0..tmax, of a 1-D series
    (Gamma(t) = <(a_i - abar)(a_{i+t} - abar)> / Gamma(0)), via FFT."""
    a = np.asarray(a, dtype=np.float64)
    n = a.size
    if tmax is None:
        tmax = n // 2
    d = a - a.mean()
    m = 1 << (2 * n - 1).bit_length()
    f = np.fft.rfft(d, m)
    g = np.fft.irfft(f * np.conj(f), m)[: tmax + 1]
    g /= np.arange(n, n - tmax - 1, -1)
    return g / g[0]


def tau_int(a: np.ndarray, S: float = 1.5) -> dict:
    """Integrated autocorrelation time with Wolff's automatic window.

    Returns {"tau_int", "dtau", "window", "mean", "err"} where err is the
    autocorrelation-corrected error of the mean, sqrt(2 tau_int var / N),
    and dtau Wolff's error estimate tau_int sqrt(2 (2W + 1) / N)."""
    a = np.asarray(a, dtype=np.float64)
    n = a.size
    if n < 8:
        raise ValueError("need at least 8 samples")
    var = a.var()
    if var == 0.0:
        return {"tau_int": 0.5, "dtau": 0.0, "window": 0, "mean": float(a.mean()), "err": 0.0}
    rho = autocorrelation(a, n // 2)
    tint = 0.5
    window = n // 2
    for W in range(1, n // 2 + 1):
        tint += rho[W]
        if tint <= 0.5:
            # no positive correlation left: the chain is effectively independent here
            window = W
            tint = 0.5
            break
        tau = S / math.log((2.0 * tint + 1.0) / (2.0 * tint - 1.0))
        g = math.exp(-W / tau) - tau / math.sqrt(W * n)
        if g < 0.0:
            window = W
            break
    dtau = tint * math.sqrt(2.0 * (2 * window + 1) / n)
    err = math.sqrt(2.0 * tint * var / n)
    return {"tau_int": float(tint), "dtau": float(dtau), "window": int(window), "mean": float(a.mean()),
            "err": float(err)}


def effective_configs_per_second(configs_per_second: float, tau: float) -> float:
    """Independent configurations per second, (configs/s) / (2 tau_int)."""
    return configs_per_second / (2.0 * tau)


def flow_scale(t: np.ndarray, e: np.ndarray, target: float = 0.3) -> float:
    """The gradient-flow scale t0: the flow time at which t^2 E(t) first
    reaches `target` (0.3, Luscher, JHEP 1008:071), by linear interpolation
    of t^2 E between the bracketing flow times.  `e` is the physical energy
    density E = -E_raw / 64 at flow times `t`.  Returns nan when t^2 E stays
    below the target."""
    t = np.asarray(t, dtype=np.float64)
    f = t * t * np.asarray(e, dtype=np.float64)
    above = np.nonzero(f >= target)[0]
    if above.size == 0:
        return float("nan")
    i = int(above[0])
    if i == 0:
        return float(t[0])
    return float(t[i - 1] + (target - f[i - 1]) * (t[i] - t[i - 1]) / (f[i] - f[i - 1]))
