"""glac_amd -- MI355X-native implementation of GLAC's gauge-generation and
gradient-flow hot path.

The compute lives in ``libglacx.so`` (HIP kernels for gfx950 behind the C ABI
declared in ``include/glacx.h``).  This module is a thin ctypes binding used
by the tests and the benchmark; it raises if the library is missing rather
than falling back to anything else.

Host arrays use GLAC's layout: float64 ``(4, V, 18)`` -- one block of V links
per direction mu, sites in x-fastest order, 18 doubles per link with (re, im)
interleaved row-major (GLAC/math/matrices/su3.h:22,64;
GLAC/parallelization/index.h:32-35).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

__all__ = ["Lattice", "GlacxError", "library", "FP64", "FP32_LINKS", "FP32_COMPRESSED", "MAIN", "FLOW",
           "HOT_HAAR", "HOT_GLAC_RANDOM", "HOT_RST",
           "K_FLOW_STAGE", "K_MEASURE", "K_METROPOLIS", "K_PLAQUETTE", "K_METROPOLIS_LEARNED", "K_TRAIN",
           "K_FLOW_STAGE_MEASURE",
           "PROPOSAL_NPARAM", "LIB_PATH"]

FP64, FP32_LINKS, FP32_COMPRESSED = 0, 1, 2
MAIN, FLOW = 0, 1
HOT_HAAR, HOT_GLAC_RANDOM, HOT_RST = 0, 1, 2
K_FLOW_STAGE, K_MEASURE, K_METROPOLIS, K_PLAQUETTE, K_METROPOLIS_LEARNED, K_TRAIN = 0, 1, 2, 3, 4, 5
K_FLOW_STAGE_MEASURE = 6  # flow stage 0 with the clover measurement fused
PROPOSAL_NPARAM = 2468

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libglacx.so")
_lib = None

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_vp = C.c_void_p


class GlacxError(RuntimeError):
    pass


def library():
    """Load libglacx.so (built by ``make -C glac_amd/csrc``)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise GlacxError(f"{LIB_PATH} is not built; run __graft_entry__.build() "
                         "or make -C glac_amd/csrc")
    L = C.CDLL(LIB_PATH)
    i4 = C.c_int * 4
    sig = {
        "glacx_create": ([i4, C.c_int, C.c_int, C.POINTER(_vp)], C.c_int),
        "glacx_destroy": ([_vp], C.c_int),
        "glacx_last_error": ([_vp], C.c_char_p),
        "glacx_device_bytes": ([_vp, C.POINTER(C.c_size_t)], C.c_int),
        "glacx_upload_field": ([_vp, C.c_int, _dp * 4], C.c_int),
        "glacx_download_field": ([_vp, C.c_int, _dp * 4], C.c_int),
        "glacx_copy_field": ([_vp, C.c_int, C.c_int], C.c_int),
        "glacx_load_bin": ([_vp, C.c_int, C.c_char_p], C.c_int),
        "glacx_save_bin": ([_vp, C.c_int, C.c_char_p], C.c_int),
        "glacx_load_chroma": ([_vp, C.c_int, C.c_char_p], C.c_int),
        "glacx_measure_super": ([_vp, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp], C.c_int),
        "glacx_measure_density": ([_vp, C.c_int, _dp, _dp, _dp, _dp, _dp], C.c_int),
        "glacx_flow_super": ([_vp, C.c_int, C.c_int, C.c_double, _dp, _dp, _dp, _dp, _dp, _dp], C.c_int),
        "glacx_cold_start": ([_vp, C.c_int], C.c_int),
        "glacx_hot_start": ([_vp, C.c_int, C.c_uint64, C.c_int, C.c_double], C.c_int),
        "glacx_metropolis": ([_vp, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int, C.c_uint64,
                              C.c_uint64, _dp, _dp], C.c_int),
        "glacx_measure": ([_vp, C.c_int, _dp, _dp, _dp, _dp], C.c_int),
        "glacx_plaquette": ([_vp, C.c_int, _dp], C.c_int),
        "glacx_flow": ([_vp, C.c_int, C.c_int, C.c_double, _dp, _dp, _dp, _dp], C.c_int),
        "glacx_flow_noalloc": ([_vp, C.c_int, C.c_int, C.c_double, _dp, _dp, _dp, _dp], C.c_int),
        "glacx_action_derivative": ([_vp, C.c_int, C.c_int, _dp], C.c_int),
        "glacx_samples_reset": ([_vp, C.c_size_t], C.c_int),
        "glacx_samples_collect": ([_vp, C.c_int, C.POINTER(C.c_size_t)], C.c_int),
        "glacx_samples_stats": ([_vp, C.c_double, _fp, _fp], C.c_int),
        "glacx_proposal_set": ([_vp, _fp, _fp, _fp], C.c_int),
        "glacx_proposal_get": ([_vp, _fp, _fp, _fp], C.c_int),
        "glacx_train": ([_vp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_uint64, _dp], C.c_int),
        "glacx_train_grad": ([_vp, _fp, C.c_int, _fp, C.c_int, C.c_double, _dp, _fp], C.c_int),
        "glacx_proposal_eval": ([_vp, _fp, C.c_int, _fp, C.POINTER(C.c_int)], C.c_int),
        "glacx_metropolis_learned": ([_vp, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int, C.c_uint64,
                                      C.c_uint64, _dp, _dp], C.c_int),
        "glacx_synchronize": ([_vp], C.c_int),
        "glacx_timing_enable": ([_vp, C.c_int], C.c_int),
        "glacx_timing_reset": ([_vp], C.c_int),
        "glacx_timing_read": ([_vp, C.c_int, C.POINTER(C.c_longlong), _dp], C.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def _ptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def _fptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_fp)


class Lattice:
    """A device-resident lattice: two gauge fields (MAIN = GLAC's m_lattice,
    FLOW = m_flowLattice) plus the flow and reduction work buffers."""

    def __init__(self, dims, precision=FP64, device=0):
        self.dims = tuple(int(d) for d in dims)
        self.V = int(np.prod(self.dims))
        self.precision = precision
        self._L = library()
        h = _vp()
        rc = self._L.glacx_create((C.c_int * 4)(*self.dims), precision, device, C.byref(h))
        self._h = h
        if rc != 0:
            msg = self._L.glacx_last_error(h).decode() if h.value else "glacx_create failed"
            if h.value:
                self._L.glacx_destroy(h)
                self._h = _vp()
            raise GlacxError(f"glacx_create rc={rc}: {msg}")

    # ------------------------------------------------------------ plumbing
    def _check(self, rc, what):
        if rc != 0:
            raise GlacxError(f"{what} rc={rc}: {self._L.glacx_last_error(self._h).decode()}")

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._L.glacx_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def device_bytes(self) -> int:
        n = C.c_size_t()
        self._check(self._L.glacx_device_bytes(self._h, C.byref(n)), "device_bytes")
        return n.value

    # ------------------------------------------------------------ fields
    def upload(self, U: np.ndarray, field=MAIN):
        U = np.ascontiguousarray(U, dtype=np.float64).reshape(4, self.V, 18)
        arr = (_dp * 4)(*[U[mu].ctypes.data_as(_dp) for mu in range(4)])
        self._check(self._L.glacx_upload_field(self._h, field, arr), "upload")

    def download(self, field=MAIN) -> np.ndarray:
        U = np.empty((4, self.V, 18), dtype=np.float64)
        arr = (_dp * 4)(*[U[mu].ctypes.data_as(_dp) for mu in range(4)])
        self._check(self._L.glacx_download_field(self._h, field, arr), "download")
        return U

    def copy(self, dst=FLOW, src=MAIN):
        self._check(self._L.glacx_copy_field(self._h, dst, src), "copy")

    def load_bin(self, path, field=MAIN):
        self._check(self._L.glacx_load_bin(self._h, field, os.fsencode(path)), "load_bin")

    def load_chroma(self, path, field=MAIN):
        """Chroma configuration: GLAC's .bin order with big-endian doubles."""
        self._check(self._L.glacx_load_chroma(self._h, field, os.fsencode(path)), "load_chroma")

    def save_bin(self, path, field=MAIN):
        self._check(self._L.glacx_save_bin(self._h, field, os.fsencode(path)), "save_bin")

    def cold_start(self, field=MAIN):
        self._check(self._L.glacx_cold_start(self._h, field), "cold_start")

    def hot_start(self, seed, field=MAIN, kind=None, eps=0.24):
        """GLAC's hot start from the link streams of `seed`: kind HOT_HAAR
        (default), HOT_GLAC_RANDOM (GLAC's generateRandom, det = -1) or
        HOT_RST (GLAC's RSTHotStart with step eps)."""
        kind = HOT_HAAR if kind is None else kind
        self._check(self._L.glacx_hot_start(self._h, field, seed, kind, eps), "hot_start")

    # ------------------------------------------------------------ compute
    def metropolis(self, n_sweeps, beta, eps=0.24, n_hit=10, reunit_every=1, seed=1,
                   sweep_offset=0, plaquette=False):
        acc = np.empty(n_sweeps)
        plaq = np.empty(n_sweeps) if plaquette else None
        self._check(self._L.glacx_metropolis(self._h, n_sweeps, beta, eps, n_hit, reunit_every, seed,
                                             sweep_offset, _ptr(acc), _ptr(plaq)), "metropolis")
        return (acc, plaq) if plaquette else acc

    def measure(self, field=MAIN):
        p, e, q = C.c_double(), C.c_double(), C.c_double()
        qt = np.empty(self.dims[3])
        self._check(self._L.glacx_measure(self._h, field, C.byref(p), C.byref(e), C.byref(q), _ptr(qt)),
                    "measure")
        return p.value, e.value, q.value, qt

    def measure_super(self, field=MAIN):
        """GLAC's SuperSampler: (plaq, E_raw, Q, W, Q(t), W(t))."""
        p, e, q, w = C.c_double(), C.c_double(), C.c_double(), C.c_double()
        qt, wt = np.empty(self.dims[3]), np.empty(self.dims[3])
        self._check(self._L.glacx_measure_super(self._h, field, C.byref(p), C.byref(e), C.byref(q), C.byref(w),
                                                _ptr(qt), _ptr(wt)), "measure_super")
        return p.value, e.value, q.value, w.value, qt, wt

    def measure_density(self, field=MAIN):
        """Per-site (charge, action) densities in GLAC's site order, plus
        (plaq, E_raw, Q): GLAC's LatticeActionChargeDensity."""
        qd, ed = np.empty(self.V), np.empty(self.V)
        p, e, q = C.c_double(), C.c_double(), C.c_double()
        self._check(self._L.glacx_measure_density(self._h, field, _ptr(qd), _ptr(ed), C.byref(p), C.byref(e),
                                                  C.byref(q)), "measure_density")
        return qd, ed, p.value, e.value, q.value

    def plaquette(self, field=MAIN) -> float:
        p = C.c_double()
        self._check(self._L.glacx_plaquette(self._h, field, C.byref(p)), "plaquette")
        return p.value

    def flow(self, n_steps, eps=0.01, field=FLOW, measure=True):
        """RK3 flow in place; returns (plaq, E_raw, Q, Q(t)) per flow time
        (n_steps + 1 rows) when measure is true."""
        if not measure:
            self._check(self._L.glacx_flow(self._h, field, n_steps, eps, None, None, None, None), "flow")
            return None
        p, e, q = np.empty(n_steps + 1), np.empty(n_steps + 1), np.empty(n_steps + 1)
        qt = np.empty((n_steps + 1, self.dims[3]))
        self._check(self._L.glacx_flow(self._h, field, n_steps, eps, _ptr(p), _ptr(e), _ptr(q), _ptr(qt)),
                    "flow")
        return p, e, q, qt

    def flow_noalloc(self, n_steps, eps=0.01, field=FLOW, measure=True):
        """flow() through glacx_flow_noalloc, which allocates no device memory
        (its scratch is a fixed per-GPU array)."""
        if not measure:
            self._check(self._L.glacx_flow_noalloc(self._h, field, n_steps, eps, None, None, None, None),
                        "flow_noalloc")
            return None
        p, e, q = np.empty(n_steps + 1), np.empty(n_steps + 1), np.empty(n_steps + 1)
        qt = np.empty((n_steps + 1, self.dims[3]))
        self._check(self._L.glacx_flow_noalloc(self._h, field, n_steps, eps, _ptr(p), _ptr(e), _ptr(q), _ptr(qt)),
                    "flow_noalloc")
        return p, e, q, qt

    def flow_super(self, n_steps, eps=0.01, field=FLOW):
        """RK3 flow in place measuring the SuperSampler observables before
        the first and after every step: (plaq, E_raw, Q, W, Q(t), W(t))."""
        m = n_steps + 1
        p, e, q, w = np.empty(m), np.empty(m), np.empty(m), np.empty(m)
        qt, wt = np.empty((m, self.dims[3])), np.empty((m, self.dims[3]))
        self._check(self._L.glacx_flow_super(self._h, field, n_steps, eps, _ptr(p), _ptr(e), _ptr(q), _ptr(w),
                                             _ptr(qt), _ptr(wt)), "flow_super")
        return p, e, q, w, qt, wt

    def action_derivative(self, mu, field=MAIN) -> np.ndarray:
        Z = np.empty((self.V, 18))
        self._check(self._L.glacx_action_derivative(self._h, field, mu, _ptr(Z)), "action_derivative")
        return Z

    def synchronize(self):
        self._check(self._L.glacx_synchronize(self._h), "synchronize")

    # ------------------------------------------------------------ learned proposal
    def samples_reset(self, capacity):
        self._check(self._L.glacx_samples_reset(self._h, int(capacity)), "samples_reset")

    def samples_collect(self, stride=1) -> int:
        n = C.c_size_t()
        self._check(self._L.glacx_samples_collect(self._h, int(stride), C.byref(n)), "samples_collect")
        return n.value

    def samples_stats(self, beta):
        mean, std = np.empty(6, np.float32), np.empty(6, np.float32)
        self._check(self._L.glacx_samples_stats(self._h, beta, _fptr(mean), _fptr(std)), "samples_stats")
        return mean, std

    def proposal_set(self, params, mean, std):
        p = np.ascontiguousarray(params, dtype=np.float32)
        assert p.size == PROPOSAL_NPARAM
        self._check(self._L.glacx_proposal_set(self._h, _fptr(p), _fptr(np.ascontiguousarray(mean, np.float32)),
                                               _fptr(np.ascontiguousarray(std, np.float32))), "proposal_set")

    def proposal_get(self):
        p = np.empty(PROPOSAL_NPARAM, np.float32)
        mean, std = np.empty(6, np.float32), np.empty(6, np.float32)
        self._check(self._L.glacx_proposal_get(self._h, _fptr(p), _fptr(mean), _fptr(std)), "proposal_get")
        return p, mean, std

    def train(self, steps, batch=4096, draws=4, lr=1e-3, beta=6.0, seed=1):
        loss = np.empty(steps)
        self._check(self._L.glacx_train(self._h, steps, batch, draws, lr, beta, seed, _ptr(loss)), "train")
        return loss

    def train_grad(self, samples, z, beta):
        smp = np.ascontiguousarray(samples, dtype=np.float32)
        zz = np.ascontiguousarray(z, dtype=np.float32)
        n, draws = zz.shape[0], zz.shape[1]
        assert smp.shape == (n, 36)
        loss = C.c_double()
        grad = np.empty(PROPOSAL_NPARAM, np.float32)
        self._check(self._L.glacx_train_grad(self._h, _fptr(smp), n, _fptr(zz), draws, beta, C.byref(loss),
                                             _fptr(grad)), "train_grad")
        return loss.value, grad

    def proposal_eval(self, features):
        """The sweeps' network on rows of six staple features: (outputs
        (n, 36), in-range flags (n,))."""
        f = np.ascontiguousarray(features, dtype=np.float32)
        n = f.shape[0]
        assert f.shape == (n, 6)
        out = np.empty((n, 36), np.float32)
        ok = np.empty(n, np.int32)
        self._check(self._L.glacx_proposal_eval(self._h, _fptr(f), n, _fptr(out),
                                                ok.ctypes.data_as(C.POINTER(C.c_int))), "proposal_eval")
        return out, ok.astype(bool)

    def metropolis_learned(self, n_sweeps, beta, eps=0.24, n_hit=10, reunit_every=1, seed=1, sweep_offset=0,
                           plaquette=False):
        acc = np.empty(n_sweeps)
        plaq = np.empty(n_sweeps) if plaquette else None
        self._check(self._L.glacx_metropolis_learned(self._h, n_sweeps, beta, eps, n_hit, reunit_every, seed,
                                                     sweep_offset, _ptr(acc), _ptr(plaq)), "metropolis_learned")
        return (acc, plaq) if plaquette else acc

    # ------------------------------------------------------------ timing
    def timing(self, on=True):
        self._check(self._L.glacx_timing_enable(self._h, int(on)), "timing_enable")

    def timing_reset(self):
        self._check(self._L.glacx_timing_reset(self._h), "timing_reset")

    def timing_read(self, kernel):
        n, ms = C.c_longlong(), C.c_double()
        self._check(self._L.glacx_timing_read(self._h, kernel, C.byref(n), C.byref(ms)), "timing_read")
        return n.value, ms.value
