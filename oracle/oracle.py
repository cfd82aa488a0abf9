"""oracle.py -- TEST INFRASTRUCTURE ONLY (the checker and the timed CPU baseline).

ctypes bindings to ``oracle/_build/liboracle.so`` (a plain-C FP64 restatement of
the reference's Fleck-Cummings IMC step) plus a host composition of one time
step that follows the reference's ``step()`` line by line:

* gray 2-D: IMC/IMC2D.py:1591-1897 (``Oracle2D``)
* gray 1-D: IMC/IMC1D.py:908-1124 (``Oracle1D``)

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg
may import this module.  The shipped library never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

A_RAD = 0.01372
C_LIGHT = 29.98

TAG_TRANSPORT = 1
TAG_EMISSION = 2
TAG_SOURCE = 3
TAG_COMB = 4
TAG_IC = 5
TAG_BOUNDARY0 = 8

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

_dp = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)
_i32p = C.POINTER(C.c_int32)
_u32p = C.POINTER(C.c_uint32)


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc); returns the library path."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE] + (["-B"] if force else []), check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_philox.argtypes = [_u32p, _u32p, _u32p]
        L.orc_stream_u01.argtypes = [C.c_uint64, C.c_int32, C.c_uint64, C.c_uint32, C.c_uint32]
        L.orc_stream_u01.restype = C.c_double
        L.orc_canonical_sum.argtypes = [_dp, C.c_int64]
        L.orc_canonical_sum.restype = C.c_double
        L.orc_fleck.argtypes = [C.c_int64, _dp, _dp, _dp, C.c_double, C.c_double, C.c_int32, _dp, _dp, _dp]
        L.orc_fleck.restype = C.c_int
        L.orc_emission_energy.argtypes = [C.c_int64, _dp, _dp, _dp, C.c_double, _dp]
        L.orc_counts.argtypes = [C.c_int64, _dp, C.c_int64, C.c_int32, _i64p]
        L.orc_counts.restype = C.c_int64
        L.orc_rz_slopes.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp]
        L.orc_sample_volume_2d.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _i64p, C.c_int32, C.c_int32, _dp,
                                           C.c_double, C.c_uint64, C.c_uint32, C.c_int32,
                                           _dp, _dp, _dp, _dp, _dp, _dp, _i32p, _i32p]
        L.orc_boundary_energy.argtypes = [C.c_int32, C.c_int32, C.c_double, C.c_double, _dp, C.c_int, _dp, C.c_int]
        L.orc_boundary_energy.restype = C.c_double
        L.orc_sample_boundary_2d.argtypes = [C.c_int32, C.c_int32, C.c_int64, C.c_double, C.c_double, _dp, C.c_int,
                                             _dp, C.c_int, _dp, C.c_uint64, C.c_uint32,
                                             _dp, _dp, _dp, _dp, _dp, _dp, _i32p, _i32p]
        L.orc_transport_2d.argtypes = [C.c_int64, _dp, _dp, _dp, _dp, _dp, _dp, _i32p, _i32p, _dp, _dp, C.c_int,
                                       C.c_int, _dp, _dp, _dp, C.c_double, _i32p, C.c_int64, C.c_int32, C.c_double,
                                       C.c_uint64, C.c_uint32, _dp, _dp, _dp, _i64p, C.c_int32]
        L.orc_comb_counts.argtypes = [C.c_int64, _dp, _i64p, C.c_int64, C.c_int64, C.c_uint64, C.c_uint32,
                                      C.c_uint64, _dp, _dp, _i64p, _i32p]
        L.orc_comb_counts.restype = C.c_int64
        _lib = L
    return _lib


def _p(a, ct=_dp):
    return a.ctypes.data_as(ct)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# ----------------------------------------------------------------------------
# small exported helpers
# ----------------------------------------------------------------------------
def philox4x32_10(ctr, key):
    c = np.asarray(ctr, dtype=np.uint32).copy()
    k = np.asarray(key, dtype=np.uint32).copy()
    o = np.zeros(4, dtype=np.uint32)
    lib().orc_philox(_p(c, _u32p), _p(k, _u32p), _p(o, _u32p))
    return o


def stream_u01(seed, tag, idx, step, slot):
    return lib().orc_stream_u01(seed, tag, idx, step, slot)


def canonical_sum(x):
    x = _f64(x).ravel()
    return lib().orc_canonical_sum(_p(x), x.size)


def fleck(T, sigma_true, cv, theta, dt, clip=True):
    T = _f64(T).ravel(); s = _f64(sigma_true).ravel(); cvv = _f64(np.broadcast_to(np.ravel(cv), T.shape))
    sa = np.empty_like(T); ss = np.empty_like(T); f = np.empty_like(T)
    rc = lib().orc_fleck(T.size, _p(T), _p(s), _p(cvv), theta, dt, 1 if clip else 0, _p(sa), _p(ss), _p(f))
    if rc != 0:
        raise AssertionError("Fleck factor out of bounds")
    return sa, ss, f


def emission_energy(T, sigma_a, vol, dt):
    T = _f64(T).ravel(); sa = _f64(sigma_a).ravel(); v = _f64(vol).ravel()
    E = np.empty_like(T)
    lib().orc_emission_energy(T.size, _p(T), _p(sa), _p(v), dt, _p(E))
    return E


def counts(E, N, min_one):
    E = _f64(E).ravel()
    out = np.zeros(E.size, dtype=np.int64)
    tot = lib().orc_counts(E.size, _p(E), int(N), 1 if min_one else 0, _p(out, _i64p))
    return out, int(tot)


# ----------------------------------------------------------------------------
# particle pool (structure of arrays)
# ----------------------------------------------------------------------------
@dataclass
class Pool2D:
    w: np.ndarray
    d1: np.ndarray
    d2: np.ndarray
    t: np.ndarray
    p1: np.ndarray
    p2: np.ndarray
    ci: np.ndarray
    cj: np.ndarray

    @staticmethod
    def empty(n=0):
        z = lambda: np.zeros(n, dtype=np.float64)
        return Pool2D(z(), z(), z(), z(), z(), z(), np.zeros(n, np.int32), np.zeros(n, np.int32))

    def __len__(self):
        return self.w.size

    def concat(self, other):
        return Pool2D(*[np.concatenate((a, b)) for a, b in zip(self.arrays(), other.arrays())])

    def arrays(self):
        return (self.w, self.d1, self.d2, self.t, self.p1, self.p2, self.ci, self.cj)

    def take(self, idx):
        return Pool2D(*[a[idx] for a in self.arrays()])


def sample_volume_2d(e1, e2, QV, cnt, kind, geometry_code, T, dt, seed, step, tag):
    e1 = _f64(e1); e2 = _f64(e2); nx = e1.size - 1; ny = e2.size - 1
    QV = _f64(QV).ravel(); cnt = np.ascontiguousarray(cnt, dtype=np.int64).ravel()
    Tt = _f64(T).ravel() if T is not None else np.zeros(nx * ny)
    n = int(cnt.sum())
    p = Pool2D.empty(n)
    lib().orc_sample_volume_2d(nx, ny, _p(e1), _p(e2), _p(QV), _p(cnt, _i64p), kind, geometry_code, _p(Tt), dt,
                               seed, step, tag, _p(p.w), _p(p.d1), _p(p.d2), _p(p.t), _p(p.p1), _p(p.p2),
                               _p(p.ci, _i32p), _p(p.cj, _i32p))
    return p


def boundary_energy(geometry_code, side, Tb, dt, e1, e2):
    e1 = _f64(e1); e2 = _f64(e2)
    return lib().orc_boundary_energy(geometry_code, side, Tb, dt, _p(e1), e1.size - 1, _p(e2), e2.size - 1)


def sample_boundary_2d(geometry_code, side, n, E_total, dt, e1, e2, prof_cum, seed, step):
    e1 = _f64(e1); e2 = _f64(e2)
    p = Pool2D.empty(n)
    pc = _p(_f64(prof_cum)) if prof_cum is not None else None
    keep = prof_cum
    lib().orc_sample_boundary_2d(geometry_code, side, n, E_total, dt, _p(e1), e1.size - 1, _p(e2), e2.size - 1,
                                 pc, seed, step, _p(p.w), _p(p.d1), _p(p.d2), _p(p.t), _p(p.p1), _p(p.p2),
                                 _p(p.ci, _i32p), _p(p.cj, _i32p))
    del keep
    return p


def transport_2d(pool: Pool2D, e1, e2, sigma_a, sigma_s, vol, dt, reflect, max_events, geometry_code,
                 weight_floor, seed, step, nthreads=0):
    """Transport ``pool`` in place; returns (dep, si, boundary_loss, stats[8]).

    Follows IMC/IMC2D.py:938-1107 (``_transport_particles_2d``)."""
    e1 = _f64(e1); e2 = _f64(e2); nx = e1.size - 1; ny = e2.size - 1
    sa = _f64(sigma_a).ravel(); ss = _f64(sigma_s).ravel(); v = _f64(vol).ravel()
    refl = np.array([1 if r else 0 for r in reflect], dtype=np.int32)
    dep = np.zeros(nx * ny); si = np.zeros(nx * ny); bl = np.zeros(1); st = np.zeros(8, dtype=np.int64)
    lib().orc_transport_2d(len(pool), _p(pool.w), _p(pool.d1), _p(pool.d2), _p(pool.t), _p(pool.p1), _p(pool.p2),
                           _p(pool.ci, _i32p), _p(pool.cj, _i32p), _p(e1), _p(e2), nx, ny, _p(sa), _p(ss), _p(v),
                           dt, _p(refl, _i32p), int(max_events), geometry_code, weight_floor, seed, step,
                           _p(dep), _p(si), _p(bl), _p(st, _i64p), nthreads)
    return dep.reshape(nx, ny), si.reshape(nx, ny), float(bl[0]), st


def comb_counts(w, cell, ncell, Nmax, seed, step, base=0):
    w = _f64(w); cell = np.ascontiguousarray(cell, dtype=np.int64)
    ecen = np.zeros(ncell); ew = np.zeros(ncell); cp = np.zeros(w.size, dtype=np.int64); pt = np.zeros(1, np.int32)
    tot = lib().orc_comb_counts(w.size, _p(w), _p(cell, _i64p), ncell, int(Nmax), seed, step, base, _p(ecen),
                                _p(ew), _p(cp, _i64p), _p(pt, _i32p))
    return ecen, ew, cp, int(tot), bool(pt[0])


def comb_2d(pool: Pool2D, Nmax, nx, ny, seed, step):
    """Per-cell comb (IMC/IMC2D.py:1448-1511).  Returns (new_pool, disc (nx,ny))."""
    cell = pool.ci.astype(np.int64) * ny + pool.cj.astype(np.int64)
    ecen, ew, cp, tot, passthrough = comb_counts(pool.w, cell, nx * ny, Nmax, seed, step)
    idx = np.repeat(np.arange(len(pool)), cp)
    newp = pool.take(idx)
    if not passthrough:
        newp.w = ew[cell[idx]]
    ecen_after = np.bincount(cell[idx], weights=newp.w, minlength=nx * ny)
    return newp, (ecen - ecen_after).reshape(nx, ny)


# ----------------------------------------------------------------------------
# 2-D step composition (IMC/IMC2D.py:1514-1897)
# ----------------------------------------------------------------------------
def volumes_2d(e1, e2, geometry):
    e1 = _f64(e1); e2 = _f64(e2)
    if geometry == "xy":
        return np.diff(e1)[:, None] * np.diff(e2)[None, :]
    dr2 = e1[1:] * e1[1:] - e1[:-1] * e1[:-1]
    return np.pi * dr2[:, None] * np.diff(e2)[None, :]


@dataclass
class OracleState2D:
    pool: Pool2D
    internal_energy: np.ndarray
    temperature: np.ndarray
    radiation_temperature: np.ndarray
    time: float
    previous_total_energy: float
    count: int = 0
    extra: dict = field(default_factory=dict)


def _face_profile(geometry, side, e1, e2, dt, bfunc):
    """Running sum of per-face-cell boundary emission (IMC/IMC2D.py:182-281, 340-435)."""
    names = ("left", "right", "bottom", "top") if geometry == "xy" else ("rmin", "rmax", "zmin", "zmax")
    name = names[side]
    if side < 2:
        X = e1[0] if side == 0 else e1[-1]
        cen = 0.5 * (e2[:-1] + e2[1:])
        Tc = np.array([bfunc(X, c, name) for c in cen], dtype=np.float64)
        if geometry == "xy":
            size = np.diff(e2)
        else:
            size = 2.0 * np.pi * X * np.diff(e2)
    else:
        Y = e2[0] if side == 2 else e2[-1]
        cen = 0.5 * (e1[:-1] + e1[1:])
        Tc = np.array([bfunc(c, Y, name) for c in cen], dtype=np.float64)
        if geometry == "xy":
            size = np.diff(e1)
        else:
            size = np.pi * (e1[1:] ** 2 - e1[:-1] ** 2)
    em = np.where(Tc > 0.0, A_RAD * C_LIGHT * Tc ** 4 / 4.0 * size * dt, 0.0)
    if not np.any(Tc > 0.0):
        return None
    return np.cumsum(em)


class Oracle2D:
    """FP64 CPU restatement of the gray 2-D IMC step (IMC/IMC2D.py)."""

    def __init__(self, edges1, edges2, geometry="xy", seed=1, nthreads=0):
        self.e1 = _f64(edges1); self.e2 = _f64(edges2)
        self.nx = self.e1.size - 1; self.ny = self.e2.size - 1
        self.geometry = geometry
        self.gcode = 0 if geometry == "xy" else 1
        self.vol = volumes_2d(self.e1, self.e2, geometry)
        self.seed = int(seed)
        self.nthreads = nthreads

    def init_simulation(self, Ntarget, Tinit, Tr_init, eos, inv_eos, Ntarget_ic=None):
        u = eos(Tinit)
        assert np.allclose(inv_eos(u), Tinit), "Inverse EOS failed"
        N_ic = Ntarget if Ntarget_ic is None else Ntarget_ic
        tr = np.maximum(np.asarray(Tr_init, dtype=np.float64), 0.0)
        E = A_RAD * ((tr * tr) * (tr * tr)) * self.vol
        cnt, _ = counts(E, N_ic, True)
        pool = sample_volume_2d(self.e1, self.e2, E, cnt, 1, 0 if self.geometry == "xy" else 1, None, 0.0,
                                self.seed, 0, TAG_IC)
        rad = np.bincount(pool.ci.astype(np.int64) * self.ny + pool.cj, weights=pool.w,
                          minlength=self.nx * self.ny).reshape(self.nx, self.ny)
        Tr = (rad / self.vol / A_RAD) ** 0.25
        tot = float(np.sum(u * self.vol)) + float(np.sum(pool.w))
        return OracleState2D(pool, u, np.array(Tinit, dtype=np.float64, copy=True), Tr, 0.0, tot, 0)

    def step(self, state, Ntarget, Nboundary, Nsource, Nmax, T_boundary, dt, sigma_a_func, inv_eos, cv, source,
             reflect=(False, False, False, False), theta=1.0, use_scalar_intensity_Tr=True,
             conserve_comb_energy=False, rz_linear_source=True, max_events_per_particle=1_000_000,
             boundary_source_func=None):
        nx, ny = self.nx, self.ny
        step_id = state.count
        T = state.temperature
        sig = np.asarray(sigma_a_func(T), dtype=np.float64) * np.ones((nx, ny))
        cvv = np.asarray(cv(T), dtype=np.float64) * np.ones((nx, ny))
        sa, ss, _ = fleck(T, sig, cvv, theta, dt, clip=True)
        sa = sa.reshape(nx, ny); ss = ss.reshape(nx, ny)

        pool = state.pool
        boundary_emission = 0.0
        if Nboundary > 0:
            for side in range(4):
                Tb = T_boundary[side]
                Tb = Tb(state.time) if callable(Tb) else Tb
                prof = None
                if boundary_source_func is not None:
                    prof = _face_profile(self.geometry, side, self.e1, self.e2, dt, boundary_source_func)
                    if prof is None:
                        continue
                    Etot = float(prof[-1])
                else:
                    if Tb <= 0.0:
                        continue
                    Etot = boundary_energy(self.gcode, side, Tb, dt, self.e1, self.e2)
                b = sample_boundary_2d(self.gcode, side, int(Nboundary), Etot, dt, self.e1, self.e2, prof,
                                       self.seed, step_id)
                pool = pool.concat(b)
                boundary_emission += float(np.sum(b.w))

        source_emission = 0.0
        gsrc = 0 if self.geometry == "xy" else (2 if rz_linear_source else 1)
        if Nsource > 0 and np.max(source) > 0.0:
            QV = np.asarray(source, dtype=np.float64) * dt * self.vol
            cnt, tot = counts(QV, Nsource, False)
            if tot > 0:
                s = sample_volume_2d(self.e1, self.e2, QV, cnt, 0, gsrc, T, dt, self.seed, step_id, TAG_SOURCE)
                pool = pool.concat(s)
                source_emission = float(np.sum(s.w))

        E_c = emission_energy(T, sa, self.vol, dt).reshape(nx, ny)
        E_emit = canonical_sum(E_c)
        if E_emit > 0.0 and Ntarget > 0:
            cnt, tot = counts(E_c, Ntarget, False)
            if tot > 0:
                e = sample_volume_2d(self.e1, self.e2, E_c, cnt, 0, gsrc, T, dt, self.seed, step_id, TAG_EMISSION)
                pool = pool.concat(e)

        n_tr = len(pool)
        weight_floor = 1e-10 * canonical_sum(pool.w) / max(n_tr, 1)
        dep, si, bl, stats = transport_2d(pool, self.e1, self.e2, sa, ss, self.vol, dt, reflect,
                                          max_events_per_particle, self.gcode, weight_floor, self.seed, step_id,
                                          self.nthreads)
        u = state.internal_energy + dep - E_c / self.vol
        Tn = inv_eos(u)
        if use_scalar_intensity_Tr:
            Tr = (si / A_RAD / C_LIGHT) ** 0.25
        else:
            valid = pool.w > 0.0
            rad = np.bincount(pool.ci[valid].astype(np.int64) * ny + pool.cj[valid], weights=pool.w[valid],
                              minlength=nx * ny).reshape(nx, ny)
            Tr = (rad / self.vol / A_RAD) ** 0.25
        newp, disc = comb_2d(pool, Nmax, nx, ny, self.seed, step_id)
        if conserve_comb_energy:
            u = u + disc / self.vol
            Tn = inv_eos(u)
        newp.t = np.zeros(len(newp))
        total_internal = float(np.sum(u * self.vol))
        total_rad = float(np.sum(newp.w))
        total = total_internal + total_rad
        loss = total - state.previous_total_energy - boundary_emission + bl - source_emission
        comb_disc_total = float(np.sum(disc))
        state.pool = newp
        state.internal_energy = u
        state.temperature = Tn
        state.radiation_temperature = Tr
        state.time += dt
        state.previous_total_energy = total
        state.count += 1
        info = {
            "time": state.time, "temperature": Tn, "radiation_temperature": Tr, "N_particles": len(newp),
            "total_energy": total, "total_internal_energy": total_internal, "total_radiation_energy": total_rad,
            "boundary_emission": boundary_emission, "boundary_loss": bl, "source_emission": source_emission,
            "energy_loss": loss, "comb_discrepancy": 0.0 if conserve_comb_energy else comb_disc_total,
            "dep": dep, "si": si, "emitted_energies": E_c, "stats": stats, "n_transported": n_tr,
            "weight_floor": weight_floor,
        }
        return state, info


def transport_2d_adjoint(pool, e1, e2, geometry, sigma_a, sigma_s, dt, reflect, max_events, weight_floor, seed, step,
                         lam, emit_lo=0, emit_hi=0, nthreads=0):
    """Replay adjoint of transport_2d on the same input pool (unchanged):
    returns (dL/dsigma_a, dL/dsigma_s) per cell (nx, ny) and Lambda per
    history, for lam = dL/dE_abs per cell."""
    L = lib()
    if not hasattr(L, "_adj_ready"):
        L.orc_transport_2d_adjoint.argtypes = [C.c_int64, _dp, _dp, _dp, _dp, _dp, _dp, _i32p, _i32p, _dp, _dp,
                                               C.c_int, C.c_int, _dp, _dp, C.c_double, _i32p, C.c_int64, C.c_int32,
                                               C.c_double, C.c_uint64, C.c_uint32, _dp, C.c_int64, C.c_int64, _dp,
                                               _dp, _dp, C.c_int32]
        L._adj_ready = True
    e1 = _f64(e1); e2 = _f64(e2); nx, ny = e1.size - 1, e2.size - 1
    gcode = 0 if geometry in ("xy", 0) else 1
    ci = np.ascontiguousarray(pool.ci, dtype=np.int32); cj = np.ascontiguousarray(pool.cj, dtype=np.int32)
    gA = np.zeros(nx * ny); gS = np.zeros(nx * ny); Lam = np.zeros(len(pool))
    refl = np.array([1 if r else 0 for r in reflect], dtype=np.int32)
    L.orc_transport_2d_adjoint(len(pool), _p(_f64(pool.w)), _p(_f64(pool.d1)), _p(_f64(pool.d2)), _p(_f64(pool.t)),
                               _p(_f64(pool.p1)), _p(_f64(pool.p2)), _p(ci, _i32p), _p(cj, _i32p), _p(e1), _p(e2),
                               nx, ny, _p(_f64(sigma_a).ravel()), _p(_f64(sigma_s).ravel()), float(dt),
                               _p(refl, _i32p), int(max_events), gcode, float(weight_floor), int(seed), int(step),
                               _p(_f64(lam).ravel()), int(emit_lo), int(emit_hi), _p(gA), _p(gS), _p(Lam),
                               int(nthreads))
    return gA.reshape(nx, ny), gS.reshape(nx, ny), Lam


def transport_2d_lr(pool, e1, e2, geometry, sigma_a, sigma_s, dt, reflect, max_events, weight_floor, seed, step,
                    lam, vA, vS, emit_lo=0, emit_hi=0, nthreads=0):
    """Forward-mode likelihood-ratio estimate of the directional derivative of
    L = sum_c lam_c E_abs,c along (vA, vS), perturbations of (sigma_a,
    sigma_s) per cell, over one transport of ``pool`` (unchanged): returns
    (pathwise part, score-function part, sum of the histories' L)."""
    L = lib()
    if not hasattr(L, "_lr_ready"):
        L.orc_transport_2d_lr.argtypes = [C.c_int64, _dp, _dp, _dp, _dp, _dp, _dp, _i32p, _i32p, _dp, _dp, C.c_int,
                                          C.c_int, _dp, _dp, C.c_double, _i32p, C.c_int64, C.c_int32, C.c_double,
                                          C.c_uint64, C.c_uint32, _dp, _dp, _dp, C.c_int64, C.c_int64, _dp,
                                          C.c_int32]
        L._lr_ready = True
    e1 = _f64(e1); e2 = _f64(e2); nx, ny = e1.size - 1, e2.size - 1
    gcode = 0 if geometry in ("xy", 0) else 1
    ci = np.ascontiguousarray(pool.ci, dtype=np.int32); cj = np.ascontiguousarray(pool.cj, dtype=np.int32)
    refl = np.array([1 if r else 0 for r in reflect], dtype=np.int32)
    out = np.zeros(3)
    L.orc_transport_2d_lr(len(pool), _p(_f64(pool.w)), _p(_f64(pool.d1)), _p(_f64(pool.d2)), _p(_f64(pool.t)),
                          _p(_f64(pool.p1)), _p(_f64(pool.p2)), _p(ci, _i32p), _p(cj, _i32p), _p(e1), _p(e2), nx, ny,
                          _p(_f64(sigma_a).ravel()), _p(_f64(sigma_s).ravel()), float(dt), _p(refl, _i32p),
                          int(max_events), gcode, float(weight_floor), int(seed), int(step), _p(_f64(lam).ravel()),
                          _p(_f64(vA).ravel()), _p(_f64(vS).ravel()), int(emit_lo), int(emit_hi), _p(out),
                          int(nthreads))
    return float(out[0]), float(out[1]), float(out[2])


# ----------------------------------------------------------------------------
# gray 1-D (slab / sphere), IMC/IMC1D.py:859-1124
# ----------------------------------------------------------------------------
def _lib1d():
    L = lib()
    if not hasattr(L, "_o1_ready"):
        L.o1_sample_volume.argtypes = [C.c_int, C.c_int, _dp, _dp, _i64p, C.c_int32, _dp, C.c_double, C.c_uint64,
                                       C.c_uint32, C.c_int32, C.c_uint64, _dp, _dp, _dp, _dp, _i32p]
        L.o1_boundary_energy.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double]
        L.o1_boundary_energy.restype = C.c_double
        L.o1_sample_boundary.argtypes = [C.c_int, C.c_int, C.c_int64, C.c_double, C.c_double, C.c_int, _dp,
                                         C.c_uint64, C.c_uint32, C.c_uint64, _dp, _dp, _dp, _dp, _i32p]
        L.o1_transport.argtypes = [C.c_int, C.c_int64, _dp, _dp, _dp, _dp, _i32p, C.c_int, _dp, _dp, _dp,
                                   C.c_double, _i32p, C.c_double, C.c_uint64, C.c_uint32, C.c_uint64, _dp, _dp,
                                   _dp, _dp, _i64p, C.c_int32]
        L.o1_counts_ic.argtypes = [C.c_int, _dp, C.c_int64, _i64p]
        L.o1_counts_ic.restype = C.c_int64
        L._o1_ready = True
    return L


@dataclass
class Pool1D:
    w: np.ndarray
    mu: np.ndarray
    t: np.ndarray
    x: np.ndarray
    cell: np.ndarray

    @staticmethod
    def empty(n=0):
        z = lambda: np.zeros(n, dtype=np.float64)
        return Pool1D(z(), z(), z(), z(), np.zeros(n, np.int32))

    def __len__(self):
        return self.w.size

    def arrays(self):
        return (self.w, self.mu, self.t, self.x, self.cell)

    def concat(self, other):
        return Pool1D(*[np.concatenate((a, b)) for a, b in zip(self.arrays(), other.arrays())])

    def take(self, idx):
        return Pool1D(*[a[idx] for a in self.arrays()])


def volumes_1d(edges, geometry):
    e = _f64(edges)
    if geometry == "slab":
        return e[1:] - e[:-1]
    # r*r*r as o1_vol in imc_oracle_1d.c (IMC/IMC1D.py:856 writes r**3)
    return (4.0 / 3.0) * np.pi * (e[1:] * e[1:] * e[1:] - e[:-1] * e[:-1] * e[:-1])


def sample_volume_1d(sph, edges, QV, cnt, kind, T, dt, seed, step, tag, base):
    e = _f64(edges); I = e.size - 1
    QV = _f64(QV); cnt = np.ascontiguousarray(cnt, dtype=np.int64)
    Tt = _f64(T) if T is not None else np.ones(I)
    p = Pool1D.empty(int(cnt.sum()))
    _lib1d().o1_sample_volume(1 if sph else 0, I, _p(e), _p(QV), _p(cnt, _i64p), kind, _p(Tt), dt, seed, step, tag,
                              base, _p(p.w), _p(p.mu), _p(p.t), _p(p.x), _p(p.cell, _i32p))
    return p


def sample_boundary_1d(sph, side, n, E, dt, edges, seed, step, base):
    e = _f64(edges)
    p = Pool1D.empty(n)
    _lib1d().o1_sample_boundary(1 if sph else 0, side, n, E, dt, e.size - 1, _p(e), seed, step, base, _p(p.w),
                                _p(p.mu), _p(p.t), _p(p.x), _p(p.cell, _i32p))
    return p


def transport_1d(sph, pool, edges, sigma_a, sigma_s, dt, reflect, weight_floor, seed, step, base, nthreads=0):
    """Transport in place (IMC/IMC1D.py:133-200, 439-532).  Returns
    (dep, si, leak, floor_leak, stats[8]); stats = events, boundary, scatter,
    census, floor kills, reflections, assertion failures."""
    e = _f64(edges); I = e.size - 1
    sa = _f64(sigma_a); ss = _f64(sigma_s)
    refl = np.array([1 if r else 0 for r in reflect], dtype=np.int32)
    dep = np.zeros(I); si = np.zeros(I); lk = np.zeros(1); fl = np.zeros(1); st = np.zeros(8, np.int64)
    _lib1d().o1_transport(1 if sph else 0, len(pool), _p(pool.w), _p(pool.mu), _p(pool.t), _p(pool.x),
                          _p(pool.cell, _i32p), I, _p(e), _p(sa), _p(ss), dt, _p(refl, _i32p), weight_floor, seed,
                          step, base, _p(dep), _p(si), _p(lk), _p(fl), _p(st, _i64p), nthreads)
    return dep, si, float(lk[0]), float(fl[0]), st


def comb_1d(pool, Nmax, I, seed, step, base):
    """IMC/IMC1D.py:770-828 with the comb stream offset by the problem base."""
    w = pool.w
    cell = pool.cell.astype(np.int64)
    ecen, ew, cp, tot, passthrough = comb_counts(w, cell, I, Nmax, seed, step, base)
    idx = np.repeat(np.arange(len(pool)), cp)
    newp = pool.take(idx)
    if not passthrough:
        newp.w = ew[cell[idx]]
    ecen_after = np.bincount(cell[idx], weights=newp.w, minlength=I)
    return newp, ecen - ecen_after


@dataclass
class OracleState1D:
    pool: Pool1D
    internal_energy: np.ndarray
    temperature: np.ndarray
    radiation_temperature: np.ndarray
    time: float
    previous_total_energy: float
    count: int = 0


class Oracle1D:
    """FP64 CPU restatement of the gray 1-D IMC step (IMC/IMC1D.py).

    ``problem`` selects the Philox index base (problem << 32) so that problem
    p of a batched device run is reproduced exactly."""

    def __init__(self, mesh, geometry="slab", seed=1, problem=0, nthreads=0):
        mesh = np.asarray(mesh, dtype=np.float64)
        self.edges = np.concatenate((mesh[:, 0], mesh[-1:, 1]))
        self.I = mesh.shape[0]
        self.geometry = geometry
        self.sph = geometry != "slab"
        self.vol = volumes_1d(self.edges, geometry)
        self.seed = int(seed)
        self.base = int(problem) << 32
        self.nthreads = nthreads

    def init_simulation(self, Ntarget, Tinit, Tr_init, eos, inv_eos, Ntarget_ic=None, T_emit_floor=0.0):
        u = eos(Tinit)
        assert np.allclose(inv_eos(u), Tinit), "Inverse EOS failed"
        N_ic = Ntarget_ic if Ntarget_ic is not None else Ntarget
        tr = np.asarray(Tr_init, dtype=np.float64)
        E = A_RAD * ((tr * tr) * (tr * tr)) * self.vol
        if T_emit_floor > 0.0:
            E = np.where(tr < T_emit_floor, 0.0, E)
        cnt = np.zeros(self.I, np.int64)
        _lib1d().o1_counts_ic(self.I, _p(_f64(E)), int(N_ic), _p(cnt, _i64p))
        pool = sample_volume_1d(self.sph, self.edges, E, cnt, 1, None, 0.0, self.seed, 0, TAG_IC, self.base)
        Tr = (np.bincount(pool.cell, weights=pool.w, minlength=self.I) / self.vol / A_RAD) ** 0.25
        tot = float(np.sum(u * self.vol)) + float(np.sum(pool.w))
        return OracleState1D(pool, u, np.array(Tinit, dtype=np.float64, copy=True), Tr, 0.0, tot, 0)

    def step(self, state, Ntarget, Nboundary, Nsource, NMax, T_boundary, dt, sigma_a_func, inv_eos, cv, source,
             reflect=(False, False), theta=1.0, use_scalar_intensity_Tr=True, conserve_comb_energy=False,
             T_emit_floor=0.0):
        I = self.I
        sid = state.count
        T = state.temperature
        sig = np.asarray(sigma_a_func(T), dtype=np.float64) * np.ones(I)
        cvv = np.asarray(cv(T), dtype=np.float64) * np.ones(I)
        sa, ss, _ = fleck(T, sig, cvv, theta, dt, clip=False)
        pool = state.pool
        be = 0.0
        for side in (0, 1):
            Tb = T_boundary[side]
            Tb = Tb(state.time) if callable(Tb) else Tb
            if Tb > 0 and Nboundary > 0:
                R = self.edges[0] if side == 0 else self.edges[-1]
                E = _lib1d().o1_boundary_energy(1 if self.sph else 0, Tb, dt, R)
                b = sample_boundary_1d(self.sph, side, int(Nboundary), E, dt, self.edges, self.seed, sid,
                                       self.base)
                pool = pool.concat(b)
                be += float(np.sum(b.w))
        src_e = 0.0
        src = np.asarray(source, dtype=np.float64)
        if np.max(src) > 0 and Nsource > 0:
            QV = src * dt * self.vol
            cnt, tot = counts(QV, Nsource, False)
            s = sample_volume_1d(self.sph, self.edges, QV, cnt, 2, T, dt, self.seed, sid, TAG_SOURCE, self.base)
            pool = pool.concat(s)
            src_e = float(np.sum(s.w))
        t = np.maximum(T, 0.0)
        Ei = A_RAD * C_LIGHT * ((t * t) * (t * t)) * sa * dt * self.vol
        if T_emit_floor > 0.0:
            Ei = np.where(T < T_emit_floor, 0.0, Ei)
        cnt, tot = counts(Ei, Ntarget, False)
        if tot > 0:
            e = sample_volume_1d(self.sph, self.edges, Ei, cnt, 0, T, dt, self.seed, sid, TAG_EMISSION, self.base)
            pool = pool.concat(e)
        n_tr = len(pool)
        floor = 1e-10 * canonical_sum(pool.w) / max(n_tr, 1)
        dep, si, leak, fleak, stats = transport_1d(self.sph, pool, self.edges, sa, ss, dt, reflect, floor, self.seed,
                                                   sid, self.base, self.nthreads)
        if stats[6] > 0:
            raise AssertionError("Non-positive distance in move_particle")
        u = state.internal_energy + dep - Ei / self.vol
        Tn = inv_eos(u)
        if use_scalar_intensity_Tr:
            Tr = (si / A_RAD / C_LIGHT) ** 0.25
        else:
            valid = (pool.cell >= 0) & (pool.cell < I) & (pool.w > 0)
            Tr = (np.bincount(pool.cell[valid], weights=pool.w[valid], minlength=I) / self.vol / A_RAD) ** 0.25
        # escaped particles carry zero weight after transport and get no comb
        # copies; combing the whole transported pool keeps the comb stream
        # index equal to the transport index
        newp, disc = comb_1d(pool, NMax, I, self.seed, sid, self.base)
        if conserve_comb_energy:
            u = u + disc / self.vol
            Tn = inv_eos(u)
        newp.t = np.zeros(len(newp))
        ti = float(np.sum(u * self.vol))
        tr = float(np.sum(newp.w))
        total = ti + tr
        loss = total - state.previous_total_energy - be + leak - src_e
        state.pool = newp
        state.internal_energy = u
        state.temperature = Tn
        state.radiation_temperature = Tr
        state.time += dt
        state.previous_total_energy = total
        state.count += 1
        info = {"time": state.time, "temperature": Tn, "radiation_temperature": Tr, "N_particles": len(newp),
                "total_energy": total, "total_internal_energy": ti, "total_radiation_energy": tr,
                "boundary_emission": be, "boundary_loss": leak, "source_emission": src_e, "energy_loss": loss,
                "comb_discrepancy": 0.0 if conserve_comb_energy else float(np.sum(disc)),
                "floor_kill_leak": fleak, "dep": dep, "si": si, "emitted_energies": Ei, "stats": stats,
                "n_transported": n_tr, "weight_floor": floor}
        return state, info
