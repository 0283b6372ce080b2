"""FP64 CPU oracle for GLAC's gauge-generation and gradient-flow path.

TEST INFRASTRUCTURE ONLY.  ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` use it as the checker and as the timed
CPU baseline; the product (``glac_amd``) never imports it.

Thin ctypes wrapper over ``liboracle.so`` (built from ``glac_oracle.c`` by
``oracle/Makefile``).  Fields use GLAC's host layout: a float64 array of
shape ``(4, Nt, Nz, Ny, Nx, 3, 3)`` complex pairs flattened to
``(4, V, 18)`` -- mu-major, x-fastest sites, 18 doubles per link with
(re, im) interleaved, row-major (GLAC/math/matrices/su3.h:22,64;
GLAC/parallelization/index.h:32-35).  Function-level citations and the
checks that pin each function are in ``glac_oracle.c``.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_i4 = C.c_int * 4


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.or_num_threads.restype = C.c_int
        L.or_su3_mul.argtypes = [_dp, _dp, _dp]
        L.or_su3_dag.argtypes = [_dp, _dp]
        L.or_su3_retrace_mul.argtypes = [_dp, _dp]
        L.or_su3_retrace_mul.restype = C.c_double
        L.or_su3_det.argtypes = [_dp, _dp]
        L.or_mt64_seed.argtypes = [C.c_void_p, C.c_uint64]
        L.or_mt64_next.argtypes = [C.c_void_p]
        L.or_mt64_next.restype = C.c_uint64
        L.or_mt64_uniform.argtypes = [C.c_void_p, C.c_double, C.c_double]
        L.or_mt64_uniform.restype = C.c_double
        L.or_philox4x32_10.argtypes = [_u32p, _u32p, _u32p]
        L.or_xoshiro128ss_next.argtypes = [_u32p]
        L.or_xoshiro128ss_next.restype = C.c_uint32
        L.or_generate_su2.argtypes = [_dp, C.c_double, _dp]
        L.or_rst_mul.argtypes = [_dp, _dp, _dp, _dp]
        L.or_rst_mul_inv.argtypes = [_dp, _dp, _dp, _dp]
        L.or_generate_rst.argtypes = [_dp, C.c_double, _dp]
        L.or_generate_random.argtypes = [_dp, _dp]
        L.or_staple.argtypes = [_dp, _i4, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _dp]
        L.or_delta_action.argtypes = [_dp, _dp, _dp, C.c_double]
        L.or_delta_action.restype = C.c_double
        L.or_action_derivative.argtypes = [_dp, _i4, C.c_int, _dp]
        L.or_su3_exp.argtypes = [_dp, _dp]
        L.or_su3_exp_taylor.argtypes = [_dp, C.c_int, _dp]
        L.or_flow_step.argtypes = [_dp, _i4, C.c_double]
        L.or_measure.argtypes = [_dp, _i4, _dp, _dp]
        L.or_measure_super.argtypes = [_dp, _i4, _dp, _dp, _dp]
        L.or_measure_density.argtypes = [_dp, _i4, _dp, _dp]
        L.or_glac_rng_new.argtypes = [C.c_uint64, C.c_uint64]
        L.or_glac_rng_new.restype = C.c_void_p
        L.or_glac_rng_free.argtypes = [C.c_void_p]
        L.or_glac_hot_start.argtypes = [_dp, _i4, C.c_void_p, C.c_double, C.c_int]
        L.or_glac_sweep.argtypes = [_dp, _i4, C.c_void_p, C.c_double, C.c_double, C.c_int]
        L.or_glac_sweep.restype = C.c_double
        L.or_reunitarize.argtypes = [_dp]
        L.or_reunitarize_keep_det_sign.argtypes = [_dp]
        L.or_hot_start_haar.argtypes = [_dp, _i4, C.c_uint64]
        L.or_hot_start.argtypes = [_dp, _i4, C.c_uint64, C.c_int, C.c_double]
        L.or_cb_sweep.argtypes = [_dp, _i4, C.c_double, C.c_double, C.c_int, C.c_uint64,
                                  C.c_uint64, C.c_int]
        L.or_cb_sweep.restype = C.c_double
        L.or_staple_features.argtypes = [_dp, C.c_double, _dp]
        L.or_mlp_forward.argtypes = [_dp, _dp, _dp, _dp, _dp]
        L.or_cholesky_from_output.argtypes = [_dp, _dp]
        L.or_algebra_from_omega.argtypes = [_dp, _dp]
        L.or_project_algebra.argtypes = [_dp, _dp]
        L.or_learned_cb_sweep.argtypes = [_dp, _i4, C.c_double, C.c_double, C.c_int, C.c_uint64, C.c_uint64,
                                          C.c_int, _dp, _dp, _dp]
        L.or_learned_cb_sweep.restype = C.c_double
        _lib = L
    return _lib


def num_threads() -> int:
    """Threads of the oracle's parallel site loops (OpenMP)."""
    return int(lib().or_num_threads())


# ---------------------------------------------------------------- conversions
def to_pairs(m: np.ndarray) -> np.ndarray:
    """complex (..., 3, 3) -> float64 (..., 18) in GLAC's interleaved order."""
    m = np.asarray(m, dtype=np.complex128)
    out = np.empty(m.shape[:-2] + (18,), dtype=np.float64)
    out[..., 0::2] = m.real.reshape(m.shape[:-2] + (9,))
    out[..., 1::2] = m.imag.reshape(m.shape[:-2] + (9,))
    return out


def to_complex(p: np.ndarray) -> np.ndarray:
    """float64 (..., 18) -> complex (..., 3, 3)."""
    p = np.asarray(p, dtype=np.float64)
    return (p[..., 0::2] + 1j * p[..., 1::2]).reshape(p.shape[:-1] + (3, 3))


def dims4(n) -> "C.Array":
    return _i4(*[int(v) for v in n])


def volume(n) -> int:
    return int(n[0]) * int(n[1]) * int(n[2]) * int(n[3])


def new_field(n) -> np.ndarray:
    return np.zeros((4, volume(n), 18), dtype=np.float64)


def unit_field(n) -> np.ndarray:
    U = new_field(n)
    U[..., 0] = U[..., 8] = U[..., 16] = 1.0
    return U


# ------------------------------------------------------------- 3x3 algebra
def su3_mul(a, b):
    c = np.empty(18)
    lib().or_su3_mul(np.ascontiguousarray(a, dtype=np.float64),
                     np.ascontiguousarray(b, dtype=np.float64), c)
    return c


def su3_dag(a):
    c = np.empty(18)
    lib().or_su3_dag(np.ascontiguousarray(a, dtype=np.float64), c)
    return c


def su3_retrace_mul(a, b) -> float:
    return lib().or_su3_retrace_mul(np.ascontiguousarray(a, dtype=np.float64),
                                    np.ascontiguousarray(b, dtype=np.float64))


def su3_det(a) -> complex:
    d = np.empty(2)
    lib().or_su3_det(np.ascontiguousarray(a, dtype=np.float64), d)
    return complex(d[0], d[1])


# ------------------------------------------------------------- random numbers
class MT64:
    """mt19937_64 plus libstdc++'s uniform_real_distribution<double>."""

    def __init__(self, seed: int):
        self._buf = C.create_string_buffer(312 * 8 + 16)
        lib().or_mt64_seed(self._buf, C.c_uint64(seed))

    def next(self) -> int:
        return lib().or_mt64_next(self._buf)

    def uniform(self, a: float, b: float) -> float:
        return lib().or_mt64_uniform(self._buf, a, b)


def philox4x32_10(ctr, key) -> np.ndarray:
    out = np.zeros(4, dtype=np.uint32)
    lib().or_philox4x32_10(np.asarray(ctr, dtype=np.uint32), np.asarray(key, dtype=np.uint32), out)
    return out


def xoshiro128ss(state: np.ndarray, count: int) -> np.ndarray:
    st = np.ascontiguousarray(state, dtype=np.uint32).copy()
    out = np.array([lib().or_xoshiro128ss_next(st) for _ in range(count)], dtype=np.uint32)
    state[:] = st
    return out


# ------------------------------------------------------------- generators
def generate_su2(r, eps) -> np.ndarray:
    out = np.empty(8)
    lib().or_generate_su2(np.asarray(r, dtype=np.float64), eps, out)
    return out


def rst_mul(r, s, t, inverse=False) -> np.ndarray:
    X = np.empty(18)
    f = lib().or_rst_mul_inv if inverse else lib().or_rst_mul
    f(np.asarray(r, dtype=np.float64), np.asarray(s, dtype=np.float64),
      np.asarray(t, dtype=np.float64), X)
    return X


def generate_rst(u13, eps) -> np.ndarray:
    X = np.empty(18)
    lib().or_generate_rst(np.asarray(u13, dtype=np.float64), eps, X)
    return X


def generate_random(u18) -> np.ndarray:
    H = np.empty(18)
    lib().or_generate_random(np.asarray(u18, dtype=np.float64), H)
    return H


# ------------------------------------------------------------- action / flow
def staple(U, n, x, y, z, t, mu) -> np.ndarray:
    A = np.empty(18)
    lib().or_staple(np.ascontiguousarray(U), dims4(n), x, y, z, t, mu, A)
    return A


def delta_action(Uold, Unew, A, beta) -> float:
    return lib().or_delta_action(np.ascontiguousarray(Uold, dtype=np.float64),
                                 np.ascontiguousarray(Unew, dtype=np.float64),
                                 np.ascontiguousarray(A, dtype=np.float64), beta)


def action_derivative(U, n, mu) -> np.ndarray:
    Z = np.empty((volume(n), 18))
    lib().or_action_derivative(np.ascontiguousarray(U), dims4(n), mu, Z)
    return Z


def su3_exp(Q) -> np.ndarray:
    out = np.empty(18)
    lib().or_su3_exp(np.ascontiguousarray(Q, dtype=np.float64), out)
    return out


def su3_exp_taylor(Q, order=24) -> np.ndarray:
    out = np.empty(18)
    lib().or_su3_exp_taylor(np.ascontiguousarray(Q, dtype=np.float64), order, out)
    return out


def flow_step(U, n, eps) -> None:
    """One RK3 step, in place on the (4, V, 18) float64 field."""
    assert U.flags["C_CONTIGUOUS"] and U.dtype == np.float64
    lib().or_flow_step(U, dims4(n), eps)


def measure(U, n):
    """Return (plaquette, E_raw, Q, Q(t)) as GLAC's MasterSamplerTopcXYZ."""
    out = np.empty(3)
    qt = np.empty(int(n[3]))
    lib().or_measure(np.ascontiguousarray(U), dims4(n), out, qt)
    return float(out[0]), float(out[1]), float(out[2]), qt


def measure_density(U, n):
    """Per-site (charge, action) densities, GLAC's site order."""
    V = volume(n)
    qd, ed = np.empty(V), np.empty(V)
    lib().or_measure_density(np.ascontiguousarray(U), dims4(n), qd, ed)
    return qd, ed


def measure_super(U, n):
    """Return (plaquette, E_raw, Q, W, Q(t), W(t)) as GLAC's SuperSampler."""
    out = np.empty(4)
    qt = np.empty(int(n[3]))
    wt = np.empty(int(n[3]))
    lib().or_measure_super(np.ascontiguousarray(U), dims4(n), out, qt, wt)
    return float(out[0]), float(out[1]), float(out[2]), float(out[3]), qt, wt


# ------------------------------------------------------------- Metropolis
class GlacRng:
    """GLAC's two mt19937_64 streams (Metropolis accept, matrix generator)."""

    def __init__(self, metropolis_seed: int, matrix_seed: int):
        self.ptr = lib().or_glac_rng_new(metropolis_seed, matrix_seed)

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().or_glac_rng_free(self.ptr)
            self.ptr = None


def glac_hot_start(U, n, rng: GlacRng, eps, rst=False) -> None:
    lib().or_glac_hot_start(U, dims4(n), rng.ptr, eps, int(rst))


def glac_sweep(U, n, rng: GlacRng, beta, eps, nhit) -> float:
    return lib().or_glac_sweep(U, dims4(n), rng.ptr, beta, eps, nhit)


def reunitarize(m) -> np.ndarray:
    m = np.array(m, dtype=np.float64)
    lib().or_reunitarize(m)
    return m


def reunitarize_keep_det_sign(m) -> np.ndarray:
    m = np.array(m, dtype=np.float64)
    lib().or_reunitarize_keep_det_sign(m)
    return m


HOT_HAAR, HOT_GLAC_RANDOM, HOT_RST = 0, 1, 2


def hot_start(n, seed, kind=HOT_HAAR, eps=0.24) -> np.ndarray:
    """The device's hot starts on the same link streams: Haar SU(3), GLAC's
    generateRandom (det = -1) or GLAC's RST matrix with step eps."""
    U = new_field(n)
    lib().or_hot_start(U, dims4(n), seed, kind, eps)
    return U


def hot_start_haar(n, seed) -> np.ndarray:
    U = new_field(n)
    lib().or_hot_start_haar(U, dims4(n), seed)
    return U


def cb_sweep(U, n, beta, eps, nhit, seed, sweep, reunit=True) -> float:
    assert U.flags["C_CONTIGUOUS"] and U.dtype == np.float64
    return lib().or_cb_sweep(U, dims4(n), beta, eps, nhit, seed, sweep, int(reunit))


# ------------------------------------------------------------- learned proposal
MLP_NPARAM = 2468


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def staple_features(A, beta) -> np.ndarray:
    f = np.empty(6)
    lib().or_staple_features(_f64(A), beta, f)
    return f


def mlp_forward(params, mean, std, f) -> np.ndarray:
    o = np.empty(36)
    lib().or_mlp_forward(_f64(params), _f64(mean), _f64(std), _f64(f), o)
    return o


def cholesky_from_output(o) -> np.ndarray:
    L = np.empty(64)
    lib().or_cholesky_from_output(_f64(o), L)
    return L.reshape(8, 8)


def algebra_from_omega(w) -> np.ndarray:
    Y = np.empty(18)
    lib().or_algebra_from_omega(_f64(w), Y)
    return Y


def project_algebra(N) -> np.ndarray:
    g = np.empty(8)
    lib().or_project_algebra(_f64(N), g)
    return g


def learned_cb_sweep(U, n, beta, nhit, seed, sweep, params, mean, std, reunit=True, eps=0.24) -> float:
    """Checkerboard sweep with the learned proposal (links whose standardised
    features leave [-8, 8] use GLAC's fixed-eps proposal)."""
    assert U.flags["C_CONTIGUOUS"] and U.dtype == np.float64
    return lib().or_learned_cb_sweep(U, dims4(n), beta, eps, nhit, seed, sweep, int(reunit), _f64(params),
                                     _f64(mean), _f64(std))
