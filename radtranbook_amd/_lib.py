"""ctypes binding of ``radtranbook_amd/lib/libimc_hip.so`` (C-ABI in include/imc_hip.h).

There is no CPU fallback: every entry point needs the HIP library and a GPU,
and raises ``ImcError`` when either is missing.
"""
from __future__ import annotations

import collections
import ctypes as C
import os
import queue
import threading
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libimc_hip.so")

GEOM_XY = 0
GEOM_RZ = 1
GEOM_SLAB = 2
GEOM_SPHERE = 3
FP64 = 0
FP32_TRACK = 1
SRC_EMISSION = 0
SRC_FIXED = 1
SRC_CF = 2

A_RAD = 0.01372
C_LIGHT = 29.98

_dp = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)
_i32p = C.POINTER(C.c_int32)
_vp = C.c_void_p

# name -> (restype, argtypes); mirrors include/imc_hip.h
SIGNATURES = {
    "imc_last_error": (C.c_char_p, []),
    "imc_device_count": (C.c_int, []),
    "imc_host_alloc": (C.c_int, [C.c_uint64, C.POINTER(_vp)]),
    "imc_host_free": (C.c_int, [_vp]),
    "imc2d_create": (C.c_int, [C.POINTER(_vp), C.c_int, C.c_int, C.c_int, _dp, _dp, C.c_int, C.c_uint64, C.c_int]),
    "imc2d_destroy": (None, [_vp]),
    "imc2d_set_step": (C.c_int, [_vp, C.c_uint32]),
    "imc2d_set_seed": (C.c_int, [_vp, C.c_uint64]),
    "imc2d_set_launch": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int]),
    "imc2d_set_tracker": (C.c_int, [_vp, C.c_int, C.c_int]),
    "imc2d_census_size": (C.c_int64, [_vp]),
    "imc2d_put_census": (C.c_int, [_vp, C.c_int64, _dp, _dp, _dp, _dp, _dp, _dp, _i64p, _i64p]),
    "imc2d_get_census": (C.c_int, [_vp, C.c_int64, _dp, _dp, _dp, _dp, _dp, _dp, _i64p, _i64p]),
    "imc2d_census_energy": (C.c_int, [_vp, _dp]),
    "imc2d_census_by_cell": (C.c_int, [_vp, _dp]),
    "imc2d_set_material": (C.c_int, [_vp, _dp, _dp, _dp, C.c_double, C.c_double, C.c_int, _dp, _dp]),
    "imc2d_set_material_powerlaw": (C.c_int, [_vp, _dp, C.c_double, C.c_double, _dp, _dp, _dp]),
    "imc2d_fleck_powerlaw": (C.c_int, [_vp, C.c_double, C.c_double, C.c_int]),
    "imc2d_update_material": (C.c_int, [_vp, _dp, _dp, _dp, _dp]),
    "imc2d_equilibrium_sample": (C.c_int, [_vp, C.c_int64, _dp]),
    "imc2d_sample_boundary": (C.c_int, [_vp, C.c_int, C.c_int64, C.c_double, C.c_double, _dp, C.c_int]),
    "imc2d_sample_volume": (C.c_int, [_vp, C.c_int, C.c_int64, _dp, C.c_double, C.c_int, _dp, _dp, _i64p]),
    "imc2d_transport": (C.c_int, [_vp, C.c_double, _i32p, C.c_int64, C.c_double, _dp, _dp, _dp, _i64p]),
    "imc2d_transport_noalloc": (C.c_int, [_vp, C.c_double, _i32p, C.c_int64, C.c_double, _vp, _vp, _vp, _vp]),
    "imc2d_set_material_cf": (C.c_int, [_vp, _dp, _dp, _dp]),
    "imc2d_transport_cf": (C.c_int, [_vp, C.c_double, _i32p, C.c_int64, C.c_double, C.c_double, C.c_double,
                                     C.c_int64, _dp, _dp, _dp, _i64p]),
    "imc2d_last_transport_ms": (C.c_float, [_vp]),
    "imc2d_comb": (C.c_int, [_vp, C.c_int64, _dp, _dp, _dp, _dp]),
    "imc2d_step": (C.c_int, [_vp, C.c_double, C.c_double, C.c_int64, C.c_int64, _dp, C.c_int64, _i32p, C.c_int64,
                             C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _i64p]),
    "imc2d_enable_adjoint": (C.c_int, [_vp, C.c_int]),
    "imc2d_adjoint": (C.c_int, [_vp, C.c_double, _i32p, C.c_int64, _dp, _dp, _dp]),
    "imc2d_emission_counts": (C.c_int, [_vp, _i64p, _i64p]),
    "imc2d_snapshot": (C.c_int, [_vp, C.c_int]),
    "imc1d_create": (C.c_int, [C.POINTER(_vp), C.c_int, C.c_int, _i32p, _dp, C.c_int, C.c_uint64, C.c_int]),
    "imc1d_destroy": (None, [_vp]),
    "imc1d_set_step": (C.c_int, [_vp, C.c_uint32]),
    "imc1d_set_seed": (C.c_int, [_vp, C.c_uint64]),
    "imc1d_set_launch": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int]),
    "imc1d_set_tracker": (C.c_int, [_vp, C.c_int, C.c_int]),
    "imc1d_census_counts": (C.c_int64, [_vp, _i64p]),
    "imc1d_census_energy": (C.c_int, [_vp, _dp, _dp]),
    "imc1d_put_census": (C.c_int, [_vp, _i64p, _dp, _dp, _dp, _dp, _i64p]),
    "imc1d_get_census": (C.c_int, [_vp, C.c_int64, _dp, _dp, _dp, _dp, _i64p, _i32p]),
    "imc1d_set_material": (C.c_int, [_vp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
    "imc1d_equilibrium_sample": (C.c_int, [_vp, _i64p, _dp, _dp]),
    "imc1d_sample_boundary": (C.c_int, [_vp, C.c_int, _i64p, _dp]),
    "imc1d_sample_volume": (C.c_int, [_vp, C.c_int, _i64p, _dp, _dp, _dp, _dp]),
    "imc1d_stage_pool": (C.c_int64, [_vp, _i64p]),
    "imc1d_get_pool": (C.c_int, [_vp, C.c_int64, _dp, _dp, _dp, _dp, _i64p, _i32p]),
    "imc1d_transport": (C.c_int, [_vp, _i32p, _dp, _dp, _dp, _dp, _dp, _i64p]),
    "imc1d_last_transport_ms": (C.c_float, [_vp]),
    "imc1d_set_material_cf": (C.c_int, [_vp, _dp, _dp, _dp, _dp]),
    "imc1d_transport_cf": (C.c_int, [_vp, _i32p, _dp, C.c_int64, C.c_double, _dp, _dp, _dp, _i64p, _i64p]),
    "imc1d_comb": (C.c_int, [_vp, _i64p, _dp, _dp, _dp, _dp]),
    "mgimc_create": (C.c_int, [C.POINTER(_vp), C.c_int, C.c_int, C.c_int, _dp, _dp, C.c_int, _dp, C.c_int,
                               C.c_uint64, C.c_int]),
    "mgimc_destroy": (None, [_vp]),
    "mgimc_set_step": (C.c_int, [_vp, C.c_uint32]),
    "mgimc_set_seed": (C.c_int, [_vp, C.c_uint64]),
    "mgimc_set_launch": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int]),
    "mgimc_set_tracker": (C.c_int, [_vp, C.c_int, C.c_int]),
    "mgimc_snapshot": (C.c_int, [_vp, C.c_int]),
    "mgimc_census_size": (C.c_int64, [_vp]),
    "mgimc_put_census": (C.c_int, [_vp, C.c_int64, _dp, _dp, _dp, _dp, _dp, _dp, _i64p, _i64p, _i32p]),
    "mgimc_get_census": (C.c_int, [_vp, C.c_int64, _dp, _dp, _dp, _dp, _dp, _dp, _i64p, _i64p, _i32p]),
    "mgimc_census_by_bin": (C.c_int, [_vp, _dp, _dp]),
    "mgimc_set_material": (C.c_int, [_vp, _dp, _dp, _dp, C.c_double, C.c_double, _dp, C.c_double, _dp, _dp, _dp,
                                     _dp, _dp]),
    "mgimc_equilibrium_sample": (C.c_int, [_vp, C.c_int64, _dp, C.c_double]),
    "mgimc_sample_boundary": (C.c_int, [_vp, C.c_int, C.c_int64, C.c_double, C.c_double, C.c_double, _dp, _dp,
                                        C.c_int, _dp]),
    "mgimc_sample_source": (C.c_int, [_vp, C.c_int64, _dp, C.c_double, _dp]),
    "mgimc_sample_emission": (C.c_int, [_vp, C.c_int64, C.c_double, _dp]),
    "mgimc_transport": (C.c_int, [_vp, C.c_double, _i32p, C.c_int64, C.c_double, _dp, _dp, _dp, _i64p]),
    "mgimc_last_transport_ms": (C.c_float, [_vp]),
    "mgimc_comb": (C.c_int, [_vp, C.c_int64, _dp, _dp, _dp, _dp, _dp]),
    "mgimc_group_sum": (C.c_int, [_vp, C.c_int, _dp]),
    "mgimc_create_sphere": (C.c_int, [C.POINTER(_vp), C.c_int, _dp, C.c_int, _dp, C.c_int, C.c_uint64, C.c_int]),
    "mgimc_sphere_equilibrium": (C.c_int, [_vp, C.c_int64, _dp, C.c_double]),
    "mgimc_sphere_face": (C.c_int, [_vp, C.c_int, C.c_int64, C.c_double, C.c_double, _dp]),
    "mgimc_sphere_volume": (C.c_int, [_vp, C.c_int, C.c_int64, _dp, C.c_double, _dp]),
    "mgimc_sphere_transport": (C.c_int, [_vp, C.c_double, _i32p, C.c_double, _dp, _dp, _dp, _i64p]),
    "mgimc_sphere_comb": (C.c_int, [_vp, C.c_int64, C.c_int, _dp, _i64p, C.POINTER(C.c_int)]),
    "mgimc_sphere_census": (C.c_int, [_vp, _dp, _dp, _dp]),
}


def quarter_root(x):
    """x ** 0.25 for x >= 0 as two square roots (several times faster than
    the generic power on the host, equal to it within 1 ulp)."""
    return np.sqrt(np.sqrt(x))


class ImcError(RuntimeError):
    pass


_lib = None


def load():
    """Load the HIP library (raises ImcError if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImcError(f"{LIB_PATH} is missing: build it with `python -m radtranbook_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        msg = load().imc_last_error()
        raise ImcError(f"libimc_hip error {rc}: {msg.decode() if msg else ''}")


def require_gpu():
    n = load().imc_device_count()
    if n < 1:
        raise ImcError("no GPU visible: the IMC engine runs only on the GPU")
    return n


def _p(a, ct=_dp):
    return a.ctypes.data_as(ct) if a is not None else None


def _cell(a, shape):
    return np.ascontiguousarray(np.broadcast_to(np.asarray(a, dtype=np.float64), shape), dtype=np.float64)


# ---------------------------------------------------------------------------
# per-step host outputs
# ---------------------------------------------------------------------------
_PF_WORK: "queue.SimpleQueue" = queue.SimpleQueue()
_PF_LOCK = threading.Lock()
_PF_THREAD = None


def _pf_worker():
    while True:
        pool = _PF_WORK.get()()
        if pool is not None:
            pool._refill()


class HostArrays:
    """Fresh float64 arrays of one shape for per-step outputs.

    Every step returns new per-cell arrays (temperatures, tallies, emission
    energies), which a caller may keep.  These arrays live in page-locked
    host blocks from ``imc_host_alloc``, into which the library's kernels
    write their results directly (no staging copy, no host memcpy after the
    stream synchronises; a fresh pageable array would also pay ~1 us per
    4 KiB page of first-touch faults).  One helper thread (shared by all
    engines) allocates blocks ahead of use -- the ctypes call releases the
    GIL, so this runs while the calling thread waits on the device -- and a
    block comes back to its pool when the last array that uses it is
    garbage-collected.  ``take`` never blocks: with nothing in reserve it
    returns a plain ``np.empty``, which the library handles through its
    staging.  Without a GPU (no page-locked memory) the reserve holds
    pageable arrays whose pages the helper thread has already touched."""

    def __init__(self, shape, depth=12, max_bytes=64 << 20, min_bytes=64 << 10):
        self.shape = tuple(int(x) for x in shape)
        self.count = int(np.prod(self.shape))
        self.nbytes = self.count * 8
        self.depth = 0 if self.nbytes < min_bytes else max(0, min(int(depth), max_bytes // max(self.nbytes, 1)))
        self._q = collections.deque()     # arrays ready to hand out
        self._free = collections.deque()  # addresses of blocks no array uses any more
        self._queued = False
        self._closed = False
        self._pinned = True  # until an allocation fails (no GPU): then pre-faulted pageable arrays
        self._ref = weakref.ref(self)

    # -- blocks ------------------------------------------------------------------
    def _alloc(self):
        p = C.c_void_p()
        try:
            rc = load().imc_host_alloc(C.c_uint64(self.nbytes), C.byref(p))
        except ImcError:
            return None
        return p.value if rc == 0 and p.value else None

    def _wrap(self, addr):
        buf = (C.c_double * self.count).from_address(addr)
        fin = weakref.finalize(buf, _host_block_released, self._ref, addr)
        fin.atexit = False  # the process's memory goes with it
        return np.frombuffer(buf, dtype=np.float64).reshape(self.shape)

    def _refill(self):
        self._queued = False
        while not self._closed and len(self._q) < self.depth:
            if self._free:
                self._q.append(self._wrap(self._free.popleft()))
                continue
            addr = self._alloc() if self._pinned else None
            if addr is None:
                self._pinned = False
                a = np.empty(self.shape)
                C.memset(a.ctypes.data, 0, a.nbytes)  # fault the pages in here, not in the caller's copy
                self._q.append(a)
            else:
                self._q.append(self._wrap(addr))

    def take(self):
        try:
            a = self._q.popleft()
        except IndexError:
            try:
                a = self._wrap(self._free.popleft())
            except IndexError:
                a = np.empty(self.shape)
        if self.depth and not self._queued and not self._closed:
            self._queued = True
            global _PF_THREAD
            with _PF_LOCK:
                if _PF_THREAD is None:
                    _PF_THREAD = threading.Thread(target=_pf_worker, name="imc-host-arrays", daemon=True)
                    _PF_THREAD.start()
            _PF_WORK.put(self._ref)
        return a

    def close(self):
        self._closed = True
        self._q.clear()  # their blocks come back through _host_block_released
        while self._free:
            load().imc_host_free(C.c_void_p(self._free.popleft()))


def _host_block_released(pool_ref, addr):
    pool = pool_ref()
    if pool is not None and not pool._closed:
        pool._free.append(addr)
    else:
        load().imc_host_free(C.c_void_p(addr))


class Engine2D:
    """Device context for one 2-D mesh (wraps ``imc2d_ctx``)."""

    def __init__(self, edges1, edges2, geometry="xy", precision="fp64", seed=1, device=0):
        L = load()
        require_gpu()
        self.e1 = np.ascontiguousarray(edges1, dtype=np.float64)
        self.e2 = np.ascontiguousarray(edges2, dtype=np.float64)
        self.nx = self.e1.size - 1
        self.ny = self.e2.size - 1
        self.shape = (self.nx, self.ny)
        self.geometry = geometry
        self.gcode = {"xy": GEOM_XY, "rz": GEOM_RZ}[geometry]
        self.precision = {"fp64": FP64, "fp32": FP32_TRACK}[precision]
        self.seed = int(seed)
        h = _vp()
        check(L.imc2d_create(C.byref(h), self.gcode, self.nx, self.ny, _p(self.e1), _p(self.e2), self.precision,
                             self.seed, device))
        self._h = h
        self._L = L
        self._host = HostArrays(self.shape)
        # whole steps of the device-held material replay a captured HIP graph
        self.use_graph = True

    def close(self):
        if getattr(self, "_host", None) is not None:
            self._host.close()
        if getattr(self, "_h", None) is not None and self._h.value:
            self._L.imc2d_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- configuration -------------------------------------------------------
    def set_step(self, step):
        check(self._L.imc2d_set_step(self._h, int(step) & 0xFFFFFFFF))

    def set_seed(self, seed):
        self.seed = int(seed)
        check(self._L.imc2d_set_seed(self._h, self.seed))

    def set_launch(self, grid=0, batch=256, tally_copies=8):
        check(self._L.imc2d_set_launch(self._h, int(grid), int(batch), int(tally_copies)))

    def set_tracker(self, block=256, refill_min=1):
        check(self._L.imc2d_set_tracker(self._h, int(block), int(refill_min)))

    # -- census ---------------------------------------------------------------
    @property
    def n(self):
        return int(self._L.imc2d_census_size(self._h))

    def put_census(self, w, d1, d2, t, p1, p2, ci, cj):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (w, d1, d2, t, p1, p2)]
        ints = [np.ascontiguousarray(a, dtype=np.int64) for a in (ci, cj)]
        n = arrs[0].size
        check(self._L.imc2d_put_census(self._h, n, *[_p(a) for a in arrs], *[_p(a, _i64p) for a in ints]))

    def get_census(self):
        n = self.n
        f = [np.empty(n) for _ in range(6)]
        i = [np.empty(n, dtype=np.int64) for _ in range(2)]
        check(self._L.imc2d_get_census(self._h, n, *[_p(a) for a in f], *[_p(a, _i64p) for a in i]))
        return tuple(f) + tuple(i)

    def census_energy(self):
        out = C.c_double()
        check(self._L.imc2d_census_energy(self._h, C.byref(out)))
        return out.value

    def census_by_cell(self):
        out = np.empty(self.shape)
        check(self._L.imc2d_census_by_cell(self._h, _p(out)))
        return out

    # -- step stages ------------------------------------------------------------
    def set_material(self, T, sigma_true, cv, theta, dt, clip=True, want_opacity=False):
        T = _cell(T, self.shape); s = _cell(sigma_true, self.shape); c = _cell(cv, self.shape)
        sa = np.empty(self.shape) if want_opacity else None
        ss = np.empty(self.shape) if want_opacity else None
        check(self._L.imc2d_set_material(self._h, _p(T), _p(s), _p(c), float(theta), float(dt), 1 if clip else 0,
                                         _p(sa), _p(ss)))
        return sa, ss

    def set_material_powerlaw(self, sigma0, power, T_floor, cv, T, u):
        """Device-held power-law material: sigma0, c_v, T and u per cell."""
        a = [_cell(x, self.shape) for x in (sigma0, cv, T, u)]
        check(self._L.imc2d_set_material_powerlaw(self._h, _p(a[0]), float(power), float(T_floor), _p(a[1]),
                                                  _p(a[2]), _p(a[3])))

    def fleck_powerlaw(self, theta, dt, clip=True):
        check(self._L.imc2d_fleck_powerlaw(self._h, float(theta), float(dt), 1 if clip else 0))

    def update_material(self, want=("T", "Tr", "u")):
        """u += dep - E_emit/V, T = u/c_v, T_r from si on the device; returns
        (T, T_r, u) (None where not wanted) and sum(u V)."""
        out = {k: (self._host.take() if k in want else None) for k in ("T", "Tr", "u")}
        tot = C.c_double()
        check(self._L.imc2d_update_material(self._h, _p(out["T"]), _p(out["Tr"]), _p(out["u"]), C.byref(tot)))
        return out["T"], out["Tr"], out["u"], tot.value

    def equilibrium_sample(self, N, Tr):
        Tr = _cell(Tr, self.shape)
        check(self._L.imc2d_equilibrium_sample(self._h, int(N), _p(Tr)))

    def sample_boundary(self, side, n, E_total, dt, prof_cum=None):
        pc = None if prof_cum is None else np.ascontiguousarray(prof_cum, dtype=np.float64)
        nface = 0 if pc is None else pc.size
        check(self._L.imc2d_sample_boundary(self._h, int(side), int(n), float(E_total), float(dt), _p(pc), nface))

    def sample_volume(self, kind, N, QV=None, dt=0.0, rz_linear=True, want_energy=False):
        q = None if QV is None else _cell(QV, self.shape)
        E = self._host.take() if want_energy else None
        tot = C.c_double()
        cnt = C.c_int64()
        check(self._L.imc2d_sample_volume(self._h, int(kind), int(N), _p(q), float(dt), 1 if rz_linear else 0,
                                          _p(E), C.byref(tot), C.byref(cnt)))
        return E, tot.value, cnt.value

    def transport(self, dt, reflect, max_events, weight_floor=-1.0, want_tallies=True):
        refl = np.array([1 if r else 0 for r in reflect], dtype=np.int32)
        dep = self._host.take() if want_tallies else None
        si = self._host.take() if want_tallies else None
        bl = C.c_double()
        st = np.zeros(8, dtype=np.int64)
        check(self._L.imc2d_transport(self._h, float(dt), _p(refl, _i32p), int(max_events), float(weight_floor),
                                      _p(dep), _p(si), C.byref(bl), _p(st, _i64p)))
        return dep, si, bl.value, st

    def transport_noalloc(self, dt, reflect, max_events, weight_floor, dep_dev, si_dev, bl_dev=0, stats_dev=0):
        """``imc2d_transport_noalloc``: the tracker with the library's fixed
        device scratch; the outputs are device addresses (ints) the caller
        owns: dep and si (ncell doubles), bl (1 double), stats (8 int64)."""
        refl = np.array([1 if r else 0 for r in reflect], dtype=np.int32)
        check(self._L.imc2d_transport_noalloc(self._h, float(dt), _p(refl, _i32p), int(max_events),
                                              float(weight_floor), _vp(dep_dev), _vp(si_dev), _vp(bl_dev or None),
                                              _vp(stats_dev or None)))

    def set_material_cf(self, T, sigma_a, beta):
        """Carter-Forest material: T, the true sigma_a and beta per cell."""
        check(self._L.imc2d_set_material_cf(self._h, _p(_cell(T, self.shape)), _p(_cell(sigma_a, self.shape)),
                                            _p(_cell(beta, self.shape))))

    def transport_cf(self, dt, reflect, max_events, weight_floor=-1.0, fastpath_threshold=0.0,
                     no_progress_rel_tol=1e-12, no_progress_max_streak=8, want_tallies=True):
        """Carter-Forest transport; returns (dep = captured energy / V,
        si = sum w s / (dt V), boundary loss, stats[8])."""
        refl = np.array([1 if r else 0 for r in reflect], dtype=np.int32)
        dep = self._host.take() if want_tallies else None
        si = self._host.take() if want_tallies else None
        bl = C.c_double()
        st = np.zeros(8, dtype=np.int64)
        check(self._L.imc2d_transport_cf(self._h, float(dt), _p(refl, _i32p), int(max_events), float(weight_floor),
                                         float(fastpath_threshold), float(no_progress_rel_tol),
                                         int(no_progress_max_streak), _p(dep), _p(si), C.byref(bl), _p(st, _i64p)))
        return dep, si, bl.value, st

    def last_transport_ms(self):
        return float(self._L.imc2d_last_transport_ms(self._h))

    def comb(self, Nmax, want_ecen=False, want_disc=False):
        """Returns (ecen per cell or None, disc per cell or None, pre-comb
        census energy, post-comb census energy)."""
        ecen = self._host.take() if want_ecen else None
        disc = self._host.take() if want_disc else None
        et = C.c_double()
        tr = C.c_double()
        check(self._L.imc2d_comb(self._h, int(Nmax), _p(ecen), _p(disc), C.byref(et), C.byref(tr)))
        return ecen, disc, et.value, tr.value

    def step_device(self, theta, dt, Ntarget, Nboundary, face_E, Nmax, reflect, max_events, rz_linear=True,
                    graph=None):
        """``imc2d_step``: one whole step on the device-held power-law
        material.  Returns a dict of the per-cell outputs (T, Tr, u, dep, si,
        E_emit), the scalars and the counts."""
        refl = np.array([1 if r else 0 for r in reflect], dtype=np.int32)
        fe = np.ascontiguousarray(face_E, dtype=np.float64)
        out = {k: self._host.take() for k in ("T", "Tr", "u", "dep", "si", "E_emit")}
        sc = np.zeros(6)
        cn = np.zeros(11, dtype=np.int64)
        use_graph = self.use_graph if graph is None else bool(graph)
        check(self._L.imc2d_step(self._h, float(theta), float(dt), int(Ntarget), int(Nboundary), _p(fe), int(Nmax),
                                 _p(refl, _i32p), int(max_events), 1 if rz_linear else 0, 1 if use_graph else 0,
                                 _p(out["T"]), _p(out["Tr"]), _p(out["u"]), _p(out["dep"]), _p(out["si"]),
                                 _p(out["E_emit"]), _p(sc), _p(cn, _i64p)))
        out.update(e_emit=sc[0], boundary_loss=sc[1], internal_total=sc[2], ecen_total=sc[3], rad_total=sc[4],
                   transport_ms=sc[5], n_transported=int(cn[0]), n=int(cn[1]), emit_count=int(cn[2]),
                   stats=cn[3:].copy())
        return out

    # -- differentiable step ----------------------------------------------------
    def enable_adjoint(self, on=True):
        check(self._L.imc2d_enable_adjoint(self._h, 1 if on else 0))

    def adjoint(self, dt, reflect, max_events, lam):
        """dL/dsigma_a, dL/dsigma_s per cell for lam = dL/dE_abs per cell."""
        refl = np.array([1 if r else 0 for r in reflect], dtype=np.int32)
        lam = _cell(lam, self.shape)
        gA = np.empty(self.shape)
        gS = np.empty(self.shape)
        check(self._L.imc2d_adjoint(self._h, float(dt), _p(refl, _i32p), int(max_events), _p(lam), _p(gA), _p(gS)))
        return gA, gS

    def emission_counts(self, fix=None):
        """Counts of the last emission sample; fix: pin counts (array) or
        unpin (False)."""
        if fix is False:
            check(self._L.imc2d_emission_counts(self._h, None, None))
            return None
        if fix is not None:
            f = np.ascontiguousarray(np.broadcast_to(np.asarray(fix, dtype=np.int64), self.shape))
            check(self._L.imc2d_emission_counts(self._h, None, _p(f, _i64p)))
            return None
        out = np.empty(self.shape, dtype=np.int64)
        check(self._L.imc2d_emission_counts(self._h, _p(out, _i64p), None))
        return out

    def snapshot(self, save=True):
        check(self._L.imc2d_snapshot(self._h, 1 if save else 0))


def _per(a, n, dtype=np.float64):
    return np.ascontiguousarray(np.broadcast_to(np.asarray(a, dtype=dtype), (n,)), dtype=dtype)


class Engine1D:
    """Device context for P independent 1-D problems (wraps ``imc1d_ctx``).

    ``meshes`` is a list of edge arrays (one per problem).  Per-problem
    arguments accept a scalar (broadcast) or P values; per-cell arguments are
    the problems' cell arrays concatenated in problem order."""

    def __init__(self, meshes, geometry="slab", precision="fp64", seed=1, device=0):
        L = load()
        require_gpu()
        self.edges = [np.ascontiguousarray(e, dtype=np.float64) for e in meshes]
        self.P = len(self.edges)
        self.ncells = np.array([e.size - 1 for e in self.edges], dtype=np.int32)
        self.coff = np.concatenate(([0], np.cumsum(self.ncells))).astype(np.int64)
        self.ncell = int(self.coff[-1])
        self.geometry = geometry
        self.gcode = {"slab": GEOM_SLAB, "spherical": GEOM_SPHERE, "sphere": GEOM_SPHERE}[geometry]
        self.precision = {"fp64": FP64, "fp32": FP32_TRACK}[precision]
        self.seed = int(seed)
        allE = np.concatenate(self.edges)
        h = _vp()
        check(L.imc1d_create(C.byref(h), self.gcode, self.P, _p(self.ncells, _i32p), _p(allE), self.precision,
                             self.seed, device))
        self._h = h
        self._L = L
        self._host = HostArrays((self.ncell,))

    def close(self):
        if getattr(self, "_host", None) is not None:
            self._host.close()
        if getattr(self, "_h", None) is not None and self._h.value:
            self._L.imc1d_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def split(self, a):
        """Per-cell array -> list of per-problem arrays."""
        return [a[self.coff[p]:self.coff[p + 1]] for p in range(self.P)]

    def set_step(self, step):
        check(self._L.imc1d_set_step(self._h, int(step) & 0xFFFFFFFF))

    def set_seed(self, seed):
        self.seed = int(seed)
        check(self._L.imc1d_set_seed(self._h, self.seed))

    def set_launch(self, grid=0, batch=256, tally_copies=8):
        check(self._L.imc1d_set_launch(self._h, int(grid), int(batch), int(tally_copies)))

    def set_tracker(self, block=256, refill_min=1):
        check(self._L.imc1d_set_tracker(self._h, int(block), int(refill_min)))

    def census_counts(self):
        per = np.zeros(self.P, dtype=np.int64)
        tot = self._L.imc1d_census_counts(self._h, _p(per, _i64p))
        return per, int(tot)

    def census_energy(self):
        """(energy per cell, energy per problem) of the census."""
        pc = np.empty(self.ncell)
        pp = np.empty(self.P)
        check(self._L.imc1d_census_energy(self._h, _p(pc), _p(pp)))
        return pc, pp

    def put_census(self, counts, w, mu, t, x, cell):
        cnt = _per(counts, self.P, np.int64)
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (w, mu, t, x)]
        ci = np.ascontiguousarray(cell, dtype=np.int64)
        assert arrs[0].size == int(cnt.sum())
        check(self._L.imc1d_put_census(self._h, _p(cnt, _i64p), *[_p(a) for a in arrs], _p(ci, _i64p)))

    def get_census(self):
        """Returns (w, mu, t, x, cell, prob) grouped by problem."""
        _, n = self.census_counts()
        f = [np.empty(n) for _ in range(4)]
        ci = np.empty(n, dtype=np.int64)
        pr = np.empty(n, dtype=np.int32)
        check(self._L.imc1d_get_census(self._h, n, *[_p(a) for a in f], _p(ci, _i64p), _p(pr, _i32p)))
        return tuple(f) + (ci, pr)

    def stage_pool(self):
        per = np.zeros(self.P, dtype=np.int64)
        n = int(self._L.imc1d_stage_pool(self._h, _p(per, _i64p)))
        if n < 0:
            check(n)
        return per, n

    def get_pool(self, n):
        """The staged or transported pool: (w, mu, t, x, cell, prob)."""
        f = [np.empty(n) for _ in range(4)]
        ci = np.empty(n, dtype=np.int64)
        pr = np.empty(n, dtype=np.int32)
        check(self._L.imc1d_get_pool(self._h, n, *[_p(a) for a in f], _p(ci, _i64p), _p(pr, _i32p)))
        return tuple(f) + (ci, pr)

    def set_material(self, T, sigma_true, cv, theta, dt, want_opacity=False):
        T = _per(T, self.ncell); s = _per(sigma_true, self.ncell); c = _per(cv, self.ncell)
        th = _per(theta, self.P); d = _per(dt, self.P)
        sa = np.empty(self.ncell) if want_opacity else None
        ss = np.empty(self.ncell) if want_opacity else None
        check(self._L.imc1d_set_material(self._h, _p(T), _p(s), _p(c), _p(th), _p(d), _p(sa), _p(ss)))
        return sa, ss

    def equilibrium_sample(self, N, Tr, T_emit_floor=0.0):
        n = _per(N, self.P, np.int64); tr = _per(Tr, self.ncell); tf = _per(T_emit_floor, self.P)
        check(self._L.imc1d_equilibrium_sample(self._h, _p(n, _i64p), _p(tr), _p(tf)))

    def sample_boundary(self, side, n, E):
        nn = _per(n, self.P, np.int64); e = _per(E, self.P)
        check(self._L.imc1d_sample_boundary(self._h, int(side), _p(nn, _i64p), _p(e)))

    def sample_volume(self, kind, N, QV=None, T_emit_floor=0.0, want_energy=False):
        n = _per(N, self.P, np.int64)
        q = None if QV is None else _per(QV, self.ncell)
        tf = _per(T_emit_floor, self.P)
        E = self._host.take() if want_energy else None
        tot = np.zeros(self.P)
        check(self._L.imc1d_sample_volume(self._h, int(kind), _p(n, _i64p), _p(q), _p(tf), _p(E), _p(tot)))
        return E, tot

    def transport(self, reflect, weight_floor=-1.0, want_tallies=True):
        """reflect: (left, right) or P such pairs.  Returns (dep, si, leak[P],
        floor_leak[P], stats[8])."""
        r = np.asarray(reflect, dtype=bool)
        r = np.ascontiguousarray(np.broadcast_to(r.reshape(-1, 2), (self.P, 2)), dtype=np.int32)
        wf = _per(weight_floor, self.P)
        dep = self._host.take() if want_tallies else None
        si = self._host.take() if want_tallies else None
        lk = np.zeros(self.P); fl = np.zeros(self.P)
        st = np.zeros(8, dtype=np.int64)
        check(self._L.imc1d_transport(self._h, _p(r, _i32p), _p(wf), _p(dep), _p(si), _p(lk), _p(fl),
                                      _p(st, _i64p)))
        return dep, si, lk, fl, st

    def last_transport_ms(self):
        return float(self._L.imc1d_last_transport_ms(self._h))

    def set_material_cf(self, T, sigma_a, beta, dt):
        """Carter-Forest material (FP64 contexts): T, true sigma_a and beta =
        4 a T^3 / c_v per cell, dt per problem; starts a new step."""
        check(self._L.imc1d_set_material_cf(self._h, _p(_per(T, self.ncell)), _p(_per(sigma_a, self.ncell)),
                                            _p(_per(beta, self.ncell)), _p(_per(dt, self.P))))

    def transport_cf(self, reflect, weight_floor=-1.0, max_events=0, fastpath_threshold=0.0, want_tallies=True):
        """Carter-Forest transport.  Returns (dep, si, leak[P], stats[8],
        particles transported)."""
        r = np.asarray(reflect, dtype=bool)
        r = np.ascontiguousarray(np.broadcast_to(r.reshape(-1, 2), (self.P, 2)), dtype=np.int32)
        wf = _per(weight_floor, self.P)
        dep = self._host.take() if want_tallies else None
        si = self._host.take() if want_tallies else None
        lk = np.zeros(self.P)
        st = np.zeros(8, dtype=np.int64)
        ntr = np.zeros(1, dtype=np.int64)
        check(self._L.imc1d_transport_cf(self._h, _p(r, _i32p), _p(wf), int(max_events), float(fastpath_threshold),
                                         _p(dep), _p(si), _p(lk), _p(st, _i64p), _p(ntr, _i64p)))
        return dep, si, lk, st, int(ntr[0])

    def comb(self, Nmax, want_ecen=False, want_disc=False):
        """Returns (ecen per cell or None, disc per cell or None, pre-comb
        census energy per problem, post-comb census energy per problem)."""
        nm = _per(Nmax, self.P, np.int64)
        ecen = self._host.take() if want_ecen else None
        disc = self._host.take() if want_disc else None
        et = np.zeros(self.P); tr = np.zeros(self.P)
        check(self._L.imc1d_comb(self._h, _p(nm, _i64p), _p(ecen), _p(disc), _p(et), _p(tr)))
        return ecen, disc, et, tr


class EngineMG:
    """Device context for one multigroup 2-D mesh (wraps ``mgimc_ctx``).
    Per-(group, cell) arrays are (G, nx, ny)."""

    def __init__(self, edges1, edges2, energy_edges, geometry="xy", precision="fp64", seed=1, device=0):
        L = load()
        require_gpu()
        self.e1 = np.ascontiguousarray(edges1, dtype=np.float64)
        self.e2 = np.ascontiguousarray(edges2, dtype=np.float64)
        self.edges = np.ascontiguousarray(energy_edges, dtype=np.float64)
        self.nx, self.ny = self.e1.size - 1, self.e2.size - 1
        self.G = self.edges.size - 1
        self.shape = (self.nx, self.ny)
        self.gshape = (self.G, self.nx, self.ny)
        self.geometry = geometry
        self.gcode = {"xy": GEOM_XY, "rz": GEOM_RZ}[geometry]
        self.precision = {"fp64": FP64, "fp32": FP32_TRACK}[precision]
        self.seed = int(seed)
        h = _vp()
        check(L.mgimc_create(C.byref(h), self.gcode, self.nx, self.ny, _p(self.e1), _p(self.e2), self.G,
                             _p(self.edges), self.precision, self.seed, device))
        self._h = h
        self._L = L
        self._host = HostArrays(self.shape)
        self._ghost = HostArrays(self.gshape)

    def close(self):
        for p in (getattr(self, "_host", None), getattr(self, "_ghost", None)):
            if p is not None:
                p.close()
        if getattr(self, "_h", None) is not None and self._h.value:
            self._L.mgimc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_step(self, step):
        check(self._L.mgimc_set_step(self._h, int(step) & 0xFFFFFFFF))

    def set_seed(self, seed):
        self.seed = int(seed)
        check(self._L.mgimc_set_seed(self._h, self.seed))

    def set_launch(self, grid=0, batch=256, tally_copies=8):
        check(self._L.mgimc_set_launch(self._h, int(grid), int(batch), int(tally_copies)))

    def set_tracker(self, block=256, refill_min=1):
        check(self._L.mgimc_set_tracker(self._h, int(block), int(refill_min)))

    def snapshot(self, save=True):
        check(self._L.mgimc_snapshot(self._h, 1 if save else 0))

    @property
    def n(self):
        return int(self._L.mgimc_census_size(self._h))

    def put_census(self, w, d1, d2, t, p1, p2, ci, cj, g):
        f = [np.ascontiguousarray(a, dtype=np.float64) for a in (w, d1, d2, t, p1, p2)]
        ii = [np.ascontiguousarray(a, dtype=np.int64) for a in (ci, cj)]
        gg = np.ascontiguousarray(g, dtype=np.int32)
        check(self._L.mgimc_put_census(self._h, f[0].size, *[_p(a) for a in f], *[_p(a, _i64p) for a in ii],
                                       _p(gg, _i32p)))

    def get_census(self):
        n = self.n
        f = [np.empty(n) for _ in range(6)]
        ii = [np.empty(n, dtype=np.int64) for _ in range(2)]
        gg = np.empty(n, dtype=np.int32)
        check(self._L.mgimc_get_census(self._h, n, *[_p(a) for a in f], *[_p(a, _i64p) for a in ii],
                                       _p(gg, _i32p)))
        return tuple(f) + tuple(ii) + (gg,)

    def census_by_bin(self):
        out = np.empty(self.gshape)
        tot = C.c_double()
        check(self._L.mgimc_census_by_bin(self._h, _p(out), C.byref(tot)))
        return out, tot.value

    def set_material(self, T, sigma, cv, theta, dt, emission_fractions=None, T_emit_floor=0.0, want=()):
        """sigma: (G, nx, ny).  want: any of 'sa', 'ss', 'E', 'Eplanck'.
        Returns (dict of requested arrays, emission energy total)."""
        T = _cell(T, self.shape)
        s = np.ascontiguousarray(np.broadcast_to(np.asarray(sigma, dtype=np.float64), self.gshape))
        c = _cell(cv, self.shape)
        fr = None if emission_fractions is None else np.ascontiguousarray(emission_fractions, dtype=np.float64)
        out = {k: (self._ghost.take() if k != "Eplanck" else self._host.take()) for k in want}
        et = C.c_double()
        check(self._L.mgimc_set_material(self._h, _p(T), _p(s), _p(c), float(theta), float(dt), _p(fr),
                                         float(T_emit_floor), _p(out.get("sa")), _p(out.get("ss")), _p(out.get("E")),
                                         C.byref(et), _p(out.get("Eplanck"))))
        return out, et.value

    def equilibrium_sample(self, N, Tr, T_emit_floor=0.0):
        check(self._L.mgimc_equilibrium_sample(self._h, int(N), _p(_cell(Tr, self.shape)), float(T_emit_floor)))

    def sample_boundary(self, side, n, E_total, dt, Tb=0.0, prof_cum=None, Tface=None, want_groups=False):
        pc = None if prof_cum is None else np.ascontiguousarray(prof_cum, dtype=np.float64)
        tf = None if Tface is None else np.ascontiguousarray(Tface, dtype=np.float64)
        nface = 0 if pc is None else pc.size
        eg = np.zeros(self.G) if want_groups else None
        check(self._L.mgimc_sample_boundary(self._h, int(side), int(n), float(E_total), float(dt), float(Tb), _p(pc),
                                            _p(tf), nface, _p(eg)))
        return eg

    def sample_source(self, N, S, dt):
        S = np.ascontiguousarray(np.broadcast_to(np.asarray(S, dtype=np.float64), self.gshape))
        tot = C.c_double()
        check(self._L.mgimc_sample_source(self._h, int(N), _p(S), float(dt), C.byref(tot)))
        return tot.value

    def sample_emission(self, N, dt):
        et = C.c_double()
        check(self._L.mgimc_sample_emission(self._h, int(N), float(dt), C.byref(et)))
        return et.value

    def transport(self, dt, reflect, max_events, weight_floor=-1.0, want_tallies=True, want_dep=None):
        """want_dep (default: want_tallies) downloads the (G, nx, ny)
        deposition; group_sum(1) gives its per-cell total without it."""
        refl = np.array([1 if r else 0 for r in reflect], dtype=np.int32)
        want_dep = want_tallies if want_dep is None else want_dep
        dep = self._ghost.take() if want_dep else None
        si = self._ghost.take() if want_tallies else None
        bl = np.zeros(1 + self.G + 4)
        st = np.zeros(8, dtype=np.int64)
        check(self._L.mgimc_transport(self._h, float(dt), _p(refl, _i32p), int(max_events), float(weight_floor),
                                      _p(dep), _p(si), _p(bl), _p(st, _i64p)))
        return dep, si, bl, st

    def last_transport_ms(self):
        return float(self._L.mgimc_last_transport_ms(self._h))

    def comb(self, Nmax, want_ecen=False, want_disc=False, want_post=False):
        ecen = self._ghost.take() if want_ecen else None
        disc = self._ghost.take() if want_disc else None
        post = self._ghost.take() if want_post else None
        et = C.c_double()
        tr = C.c_double()
        check(self._L.mgimc_comb(self._h, int(Nmax), _p(ecen), _p(disc), _p(post), C.byref(et), C.byref(tr)))
        return ecen, disc, post, et.value, tr.value

    def group_block(self):
        """A fresh (G, nx, ny) float64 array in a page-locked block (the
        library copies such arrays by DMA, with no staging copy)."""
        return self._ghost.take()

    _GROUP_SUM = {"emission": 0, "dep": 1, "disc": 2}

    def group_sum(self, what):
        """Per-cell sum over groups, (nx, ny), of 'emission' (last
        set_material), 'dep' (last transport) or 'disc' (last comb with
        want_disc or want_post), summed on the device in numpy's axis-0 order."""
        out = self._host.take()
        check(self._L.mgimc_group_sum(self._h, self._GROUP_SUM[what], _p(out)))
        return out


class EngineMG1D:
    """Device context for a multigroup 1-D spherical mesh (``mgimc_ctx`` made
    by ``mgimc_create_sphere``).  ``edges``: I+1 radii; per-(group, cell)
    arrays are (G, I)."""

    def __init__(self, edges, energy_edges, precision="fp64", seed=1, device=0):
        L = load()
        require_gpu()
        self.e = np.ascontiguousarray(edges, dtype=np.float64)
        self.edges = np.ascontiguousarray(energy_edges, dtype=np.float64)
        self.I = self.e.size - 1
        self.G = self.edges.size - 1
        self.gshape = (self.G, self.I)
        self.precision = {"fp64": FP64, "fp32": FP32_TRACK}[precision]
        self.seed = int(seed)
        h = _vp()
        check(L.mgimc_create_sphere(C.byref(h), self.I, _p(self.e), self.G, _p(self.edges), self.precision, self.seed,
                                    device))
        self._h = h
        self._L = L

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._L.mgimc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_step(self, step):
        check(self._L.mgimc_set_step(self._h, int(step) & 0xFFFFFFFF))

    def set_seed(self, seed):
        self.seed = int(seed)
        check(self._L.mgimc_set_seed(self._h, self.seed))

    def set_launch(self, grid=0, batch=256, tally_copies=8):
        check(self._L.mgimc_set_launch(self._h, int(grid), int(batch), int(tally_copies)))

    @property
    def n(self):
        return int(self._L.mgimc_census_size(self._h))

    def put_census(self, w, mu, t, x, cell, g):
        n = len(w)
        f = [np.ascontiguousarray(a, dtype=np.float64) for a in (w, mu, t, x)]
        z = np.zeros(n)
        ci = np.ascontiguousarray(cell, dtype=np.int64)
        cj = np.zeros(n, dtype=np.int64)
        gg = np.ascontiguousarray(g, dtype=np.int32)
        check(self._L.mgimc_put_census(self._h, n, _p(f[0]), _p(f[1]), _p(z), _p(f[2]), _p(f[3]), _p(z),
                                       _p(ci, _i64p), _p(cj, _i64p), _p(gg, _i32p)))

    def get_census(self):
        """(w, mu, t, x, cell int64, group int32)."""
        n = self.n
        f = [np.empty(n) for _ in range(6)]
        ii = [np.empty(n, dtype=np.int64) for _ in range(2)]
        gg = np.empty(n, dtype=np.int32)
        check(self._L.mgimc_get_census(self._h, n, *[_p(a) for a in f], *[_p(a, _i64p) for a in ii],
                                       _p(gg, _i32p)))
        return f[0], f[1], f[3], f[4], ii[0], gg

    def set_material(self, T, sigma, cv, theta, dt, T_emit_floor=0.0, want=()):
        """sigma: (G, I) true opacities.  want: any of 'sa', 'ss', 'E'.
        Returns (dict of requested (G, I) arrays, emission energy total)."""
        T = np.ascontiguousarray(np.broadcast_to(np.asarray(T, dtype=np.float64), (self.I,)))
        s = np.ascontiguousarray(np.broadcast_to(np.asarray(sigma, dtype=np.float64), self.gshape))
        c = np.ascontiguousarray(np.broadcast_to(np.asarray(cv, dtype=np.float64), (self.I,)))
        out = {k: np.empty(self.gshape) for k in want}
        et = C.c_double()
        check(self._L.mgimc_set_material(self._h, _p(T), _p(s), _p(c), float(theta), float(dt), None,
                                         float(T_emit_floor), _p(out.get("sa")), _p(out.get("ss")), _p(out.get("E")),
                                         C.byref(et), None))
        return out, et.value

    def equilibrium_sample(self, N, Tr, T_emit_floor=0.0):
        Tr = np.ascontiguousarray(np.broadcast_to(np.asarray(Tr, dtype=np.float64), (self.I,)))
        check(self._L.mgimc_sphere_equilibrium(self._h, int(N), _p(Tr), float(T_emit_floor)))

    def sample_face(self, side, n, T, dt):
        E = C.c_double()
        check(self._L.mgimc_sphere_face(self._h, int(side), int(n), float(T), float(dt), C.byref(E)))
        return E.value

    def sample_volume(self, kind, N, dt, QV=None):
        q = None if QV is None else np.ascontiguousarray(np.broadcast_to(np.asarray(QV, dtype=np.float64),
                                                                          self.gshape))
        tot = C.c_double()
        check(self._L.mgimc_sphere_volume(self._h, int(kind), int(N), _p(q), float(dt), C.byref(tot)))
        return tot.value

    def transport(self, dt, reflect, weight_floor=-1.0):
        refl = np.array([1 if r else 0 for r in reflect], dtype=np.int32)
        dep = np.empty(self.gshape)
        si = np.empty(self.gshape)
        bl = np.zeros(self.G)
        st = np.zeros(8, dtype=np.int64)
        check(self._L.mgimc_sphere_transport(self._h, float(dt), _p(refl, _i32p), float(weight_floor), _p(dep), _p(si),
                                             _p(bl), _p(st, _i64p)))
        return dep, si, bl, st

    def last_transport_ms(self):
        return float(self._L.mgimc_last_transport_ms(self._h))

    def comb(self, Nmax, radiation=False):
        """Returns (disc (G, I), n_alive, combed)."""
        disc = np.empty(self.gshape)
        alive = C.c_int64()
        combed = C.c_int()
        check(self._L.mgimc_sphere_comb(self._h, int(Nmax), 1 if radiation else 0, _p(disc), C.byref(alive),
                                        C.byref(combed)))
        return disc, alive.value, bool(combed.value)

    def census(self):
        """(energy (G, I), sum w mu (G, I), total energy) of the pool."""
        E = np.empty(self.gshape)
        F = np.empty(self.gshape)
        tot = C.c_double()
        check(self._L.mgimc_sphere_census(self._h, _p(E), _p(F), C.byref(tot)))
        return E, F, tot.value
