"""imc2d_transport_noalloc (the gray 2-D tracker with the library's fixed
device scratch, no allocation) against imc2d_transport on identical pools.

Two contexts with the same seed, material and sources hold bit-identical
pools.  One runs the allocating entry point, the other the no-allocation
variant into caller-owned device buffers.  Every history is replayed from
its own Philox stream, so the census after transport and the event counters
must agree exactly; the per-cell tallies and the boundary loss are sums of
FP64 atomics whose order differs between launches, so they agree to 1e-12
relative.  Two calls in a row, the second on a smaller mesh and pool, check
that the shared scratch is cleared between calls."""
import ctypes as C
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _announce(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def L():
    from radtranbook_amd import _lib

    _lib.load()
    _lib.require_gpu()
    return _lib


class DeviceBuffers:
    """Caller-owned device memory from the HIP runtime the library uses (the
    test is the caller and may allocate; the entry point under test may not)."""

    def __init__(self):
        # the runtime instance libimc_hip.so is bound to (torch, when imported
        # first, brings its own copy under the same soname)
        self.hip = C.CDLL("libamdhip64.so.7", mode=os.RTLD_NOLOAD)
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.live = []

    def _ok(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed with hipError {rc}")

    def filled(self, n, dtype, value):
        """n elements of dtype on the device, filled with value."""
        host = np.full(n, value, dtype=dtype)
        p = C.c_void_p()
        self._ok(self.hip.hipMalloc(C.byref(p), host.nbytes), "hipMalloc")
        self.live.append(p)
        self._ok(self.hip.hipMemcpy(p, host.ctypes.data, host.nbytes, 1), "hipMemcpy H2D")
        return p.value

    def read(self, ptr, n, dtype):
        host = np.empty(n, dtype=dtype)
        self._ok(self.hip.hipMemcpy(host.ctypes.data, C.c_void_p(ptr), host.nbytes, 2), "hipMemcpy D2H")
        return host

    def free(self):
        for p in self.live:
            self.hip.hipFree(p)
        self.live = []


@pytest.fixture
def dev(L):
    d = DeviceBuffers()
    yield d
    d.free()


def _engine(L, nx, ny, geometry, precision, npart, seed):
    e1 = np.linspace(0.0, 1.0, nx + 1)
    e2 = np.linspace(0.0, 0.8, ny + 1)
    g = np.random.default_rng(seed).standard_normal((nx, ny))
    sig0 = np.exp(2.3 + 1.5 * (g - g.mean()) / g.std())
    T = np.exp(np.random.default_rng(seed + 1).uniform(np.log(0.05), np.log(0.6), (nx, ny)))
    dt = 0.01
    eng = L.Engine2D(e1, e2, geometry, precision, seed=seed)
    eng.set_step(1)
    eng.set_material(T, sig0 * T ** -3, np.full((nx, ny), 0.3), 1.0, dt)
    eng.sample_volume(L.SRC_EMISSION, npart, None, dt, True)
    side = 0 if geometry == "xy" else 2
    eng.sample_boundary(side, npart // 2, 0.05, dt)
    return eng, dt


def _compare_once(L, dev, nx, ny, geometry, precision, npart, seed):
    _announce(f"noalloc vs transport: {geometry} {precision} {nx}x{ny} cells, ~{npart + npart // 2} particles")
    a, dt = _engine(L, nx, ny, geometry, precision, npart, seed)
    b, _ = _engine(L, nx, ny, geometry, precision, npart, seed)
    pool_a, pool_b = a.get_census(), b.get_census()
    assert all(np.array_equal(x, y) for x, y in zip(pool_a, pool_b)), "identical set-up must give identical pools"
    reflect = (True, False, False, True)
    dep, si, bl, st = a.transport(dt, reflect, 10**6, -1.0)
    nc = nx * ny
    dep_d = dev.filled(nc, np.float64, np.nan)
    si_d = dev.filled(nc, np.float64, np.nan)
    bl_d = dev.filled(1, np.float64, np.nan)
    st_d = dev.filled(8, np.int64, -1)
    b.transport_noalloc(dt, reflect, 10**6, -1.0, dep_d, si_d, bl_d, st_d)
    dep2 = dev.read(dep_d, nc, np.float64).reshape(nx, ny)
    si2 = dev.read(si_d, nc, np.float64).reshape(nx, ny)
    bl2 = float(dev.read(bl_d, 1, np.float64)[0])
    st2 = dev.read(st_d, 8, np.int64)
    _announce(f"  events {st[0]} vs {st2[0]}; boundary loss {bl:.12e} vs {bl2:.12e}")
    np.testing.assert_array_equal(st2, st)
    after_a, after_b = a.get_census(), b.get_census()
    for x, y in zip(after_a, after_b):
        np.testing.assert_array_equal(y, x)
    np.testing.assert_allclose(dep2, dep, rtol=1e-12, atol=1e-300)
    np.testing.assert_allclose(si2, si, rtol=1e-12, atol=1e-300)
    assert bl2 == pytest.approx(bl, rel=1e-12, abs=1e-300)
    assert st[0] > len(pool_a[0])  # the pool did move
    a.close()
    b.close()


@pytest.mark.parametrize("geometry,precision", [("xy", "fp64"), ("xy", "fp32"), ("rz", "fp32")])
def test_noalloc_transport_matches_transport_twice(L, dev, geometry, precision):
    _compare_once(L, dev, 64, 48, geometry, precision, 120_000, seed=11)
    # second call on the same GPU: smaller mesh and pool over the stale scratch
    _compare_once(L, dev, 16, 12, geometry, precision, 9_000, seed=12)


def test_noalloc_transport_rejects_oversized_scratch(L, dev):
    """200 x 200 cells with 64 tally copies need 41 MB of tally partials,
    more than the fixed 32 MiB: a clear error, nothing launched."""
    _announce("noalloc rejects oversized scratch")
    nx = ny = 200
    e1 = np.linspace(0.0, 1.0, nx + 1)
    eng = L.Engine2D(e1, e1, "xy", "fp32", seed=3)
    eng.set_launch(0, 256, 64)
    eng.set_material(np.full((nx, ny), 0.1), np.full((nx, ny), 5.0), np.full((nx, ny), 0.3), 1.0, 0.01)
    dep_d = dev.filled(nx * ny, np.float64, 0.0)
    si_d = dev.filled(nx * ny, np.float64, 0.0)
    with pytest.raises(L.ImcError, match="bytes of scratch"):
        eng.transport_noalloc(0.01, (False,) * 4, 10**6, -1.0, dep_d, si_d)
    eng.close()
