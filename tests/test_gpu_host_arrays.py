"""Per-step outputs in page-locked blocks (radtranbook_amd/_lib.py
HostArrays over imc_host_alloc): the library writes them straight from its
kernels; results equal those written through its staging into pageable
arrays, and a block returns to its pool when its last array is gone."""
import gc
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

A = 0.01372
C = 29.98


@pytest.fixture(scope="module")
def L():
    from radtranbook_amd import _lib

    _lib.load()
    _lib.require_gpu()
    return _lib


def _wait_for(cond, timeout=10.0):
    t0 = time.time()
    while not cond() and time.time() - t0 < timeout:
        time.sleep(0.01)
    return cond()


def test_pinned_outputs_match_pageable(L):
    print("test_pinned_outputs_match_pageable: xy 96x96, 6e4 particles", flush=True)
    n = 96
    e = np.linspace(0.0, 1.0, n + 1)
    eng = L.Engine2D(e, e, "xy", "fp64", seed=5)
    dt = 0.01
    T = np.full((n, n), 0.05)
    eng.set_step(1)
    eng.set_material(T, 100.0 * T ** -3, np.full((n, n), 0.3), 1.0, dt)
    eng.equilibrium_sample(20000, T)
    eng.sample_boundary(0, 20000, A * C / 4.0 * 0.5 ** 4 * 1.0 * dt, dt)
    h = eng._host
    h.take()
    assert _wait_for(lambda: len(h._q) == h.depth)
    assert h._pinned and h.depth > 0
    Ec, tot, cnt = eng.sample_volume(L.SRC_EMISSION, 20000, None, dt, True, want_energy=True)
    assert not Ec.flags.owndata  # a page-locked block, not a NumPy allocation
    census = eng.get_census()
    dep1, si1, bl1, st1 = eng.transport(dt, (False,) * 4, 10**6)
    assert not dep1.flags.owndata and not si1.flags.owndata
    # the same pool and step again, outputs through staging into pageable arrays
    eng.put_census(*census)
    take = h.take
    h.take = lambda: np.empty(eng.shape)
    try:
        dep2, si2, bl2, st2 = eng.transport(dt, (False,) * 4, 10**6)
    finally:
        h.take = take
    assert dep2.flags.owndata
    np.testing.assert_allclose(dep1, dep2, rtol=1e-12, atol=0)
    np.testing.assert_allclose(si1, si2, rtol=1e-12, atol=0)
    assert bl1 == pytest.approx(bl2, rel=1e-12) and np.array_equal(st1, st2)
    assert st1[0] > 1000 and dep1.sum() > 0
    # a view keeps its block; dropping every array returns it to the pool
    addr = dep1.ctypes.data
    view = dep1[3:]
    del dep1
    gc.collect()
    assert addr not in h._free
    del view
    gc.collect()
    assert addr in h._free
    eng.close()
