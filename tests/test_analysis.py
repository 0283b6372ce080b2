"""Gamma-method tau_int (glac_amd/analysis.py) against closed forms: an AR(1)
chain a_{i+1} = phi a_i + noise has rho(t) = phi^t and
tau_int = (1 + phi) / (2 (1 - phi)); white noise has tau_int = 1/2."""
import numpy as np
import pytest

from glac_amd.analysis import autocorrelation, effective_configs_per_second, tau_int


def ar1(phi, n, seed):
    rng = np.random.default_rng(seed)
    e = rng.normal(size=n)
    a = np.empty(n)
    a[0] = e[0] / np.sqrt(1 - phi * phi)
    for i in range(1, n):
        a[i] = phi * a[i - 1] + e[i]
    return a


@pytest.mark.parametrize("phi", [0.5, 0.8, 0.9, 0.95])
def test_tau_int_of_ar1(phi):
    a = ar1(phi, 200000, int(phi * 100))
    r = tau_int(a)
    exact = 0.5 * (1 + phi) / (1 - phi)
    assert abs(r["tau_int"] - exact) < 4 * r["dtau"] + 0.02 * exact, (r, exact)
    assert r["dtau"] < 0.1 * exact


def test_autocorrelation_of_ar1_is_geometric():
    a = ar1(0.7, 400000, 3)
    rho = autocorrelation(a, 10)
    assert np.allclose(rho, 0.7 ** np.arange(11), atol=0.01)


def test_white_noise_has_tau_one_half():
    a = np.random.default_rng(4).normal(size=100000)
    r = tau_int(a)
    assert abs(r["tau_int"] - 0.5) < 0.05
    assert abs(r["err"] - 1 / np.sqrt(100000)) < 0.1 / np.sqrt(100000)


def test_effective_configs():
    assert effective_configs_per_second(10.0, 2.5) == 2.0


def test_flow_scale_interpolates_t2e():
    from glac_amd.analysis import flow_scale
    t = np.linspace(0.0, 5.0, 501)
    # t^2 E = 0.1 t exactly: crosses 0.3 at t = 3
    e = np.where(t > 0, 0.1 / np.where(t > 0, t, 1.0), 0.0)
    assert abs(flow_scale(t, e) - 3.0) < 1e-12
    # quadratic t^2 E = 0.05 t^2: crosses 0.3 at sqrt(6); linear interpolation error < step^2
    e2 = np.full_like(t, 0.05)
    assert abs(flow_scale(t, e2) - np.sqrt(6.0)) < 1e-4
    assert np.isnan(flow_scale(t, np.zeros_like(t)))
