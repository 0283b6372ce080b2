"""Several contexts driven at once from separate host threads: each owns its
stream and buffers, so small lattices can share the GPU (GLAC runs one
lattice per process).  Four 8^3 x 16 chains generated and flowed
concurrently give exactly the numbers the same chains give one after
another (no state crosses contexts; every reduction has a fixed order)."""
import sys
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
glacx = pytest.importorskip("glac_amd")

N = (8, 8, 8, 16)
CHAINS, SWEEPS, STEPS = 4, 20, 60


def banner(msg):
    print(f"\n[gpu case] {msg}", flush=True)
    sys.stdout.flush()


def chain(L, i):
    L.hot_start(300 + i)
    acc = L.metropolis(SWEEPS, 6.0, 0.24, 10, seed=40 + i)
    L.copy(glacx.FLOW, glacx.MAIN)
    plaq, e, q, qt = L.flow(STEPS, 0.01, field=glacx.FLOW)
    return acc, plaq, e, q, qt, L.download()


@pytest.mark.parametrize("prec", [glacx.FP32_COMPRESSED, glacx.FP64], ids=["fp32c", "fp64"])
def test_concurrent_contexts_match_sequential_runs(prec):
    banner(f"{CHAINS} concurrent contexts {N} prec={prec}")
    lattices = [glacx.Lattice(N, prec) for _ in range(CHAINS)]
    results = [None] * CHAINS
    errors = []

    def run(i):
        try:
            results[i] = chain(lattices[i], i)
        except Exception as exc:  # reported below, in the test's thread
            errors.append(exc)

    threads = [threading.Thread(target=run, args=(i,)) for i in range(CHAINS)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    for L in lattices:
        L.close()
    assert not errors, errors
    with glacx.Lattice(N, prec) as L:
        for i in range(CHAINS):
            want = chain(L, i)
            for got, ref in zip(results[i], want):
                assert np.array_equal(np.asarray(got), np.asarray(ref)), i
    # the chains differ from each other (distinct seeds)
    assert not np.array_equal(results[0][1], results[1][1])
