"""glacx_flow_noalloc, the flow entry point that allocates nothing (its
reduction partials and per-step rows live in a fixed per-GPU array), against
glacx_flow on the same field: identical observables and identical flowed
links, on two calls in a row -- the second on a smaller lattice with fewer
steps, so it runs over what the first left in the shared array -- in FP64 and
in compressed FP32 (the fused stage-0 measurement).  A flow too long for the
array is refused with GLACX_E_ARG before any work."""
import sys

import numpy as np
import pytest

from fieldtools import random_su3

pytestmark = pytest.mark.gpu
glacx = pytest.importorskip("glac_amd")


def banner(msg):
    print(f"\n[gpu case] {msg}", flush=True)
    sys.stdout.flush()


def field(n, seed):
    V = int(np.prod(n))
    return np.ascontiguousarray(random_su3(np.random.default_rng(seed), (4, V)).view(np.float64).reshape(4, V, 18))


@pytest.mark.parametrize("prec", [glacx.FP64, glacx.FP32_COMPRESSED], ids=["fp64", "fp32c"])
def test_noalloc_flow_matches_flow(prec):
    for n, steps, seed in (((8, 8, 8, 16), 12, 1), ((4, 4, 4, 8), 5, 2)):
        banner(f"flow_noalloc vs flow {n} steps {steps} prec {prec}")
        U = field(n, seed)
        with glacx.Lattice(n, prec) as L:
            L.upload(U, glacx.FLOW)
            want = L.flow(steps, 0.02, field=glacx.FLOW)
            W_want = L.download(glacx.FLOW)
            L.upload(U, glacx.FLOW)
            got = L.flow_noalloc(steps, 0.02, field=glacx.FLOW)
            W_got = L.download(glacx.FLOW)
        for name, a, b in zip(("plaq", "E_raw", "Q", "Q(t)"), want, got):
            assert a.shape == b.shape, name
            assert np.array_equal(a, b), (name, np.abs(a - b).max())
        assert np.array_equal(W_want, W_got)
        assert want[0][-1] > want[0][0]   # the flow smooths: the plaquette rises


def test_noalloc_flow_refuses_more_than_its_scratch():
    banner("flow_noalloc size limit")
    n = (4, 4, 4, 8)
    with glacx.Lattice(n, glacx.FP64) as L:
        L.cold_start(glacx.FLOW)
        with pytest.raises(glacx.GlacxError, match="scratch"):
            L.flow_noalloc(200_000, 0.01, field=glacx.FLOW)
        # refused before any work: the field is still the cold start
        assert np.array_equal(L.download(glacx.FLOW), field_cold(n))
        # without measurement nothing is kept in the array: any length runs
        L.flow_noalloc(3, 0.01, field=glacx.FLOW, measure=False)


def field_cold(n):
    V = int(np.prod(n))
    U = np.zeros((4, V, 18))
    U[:, :, [0, 8, 16]] = 1.0
    return U
