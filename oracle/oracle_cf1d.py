"""Oracle (test infrastructure only): FP64 CPU restatement of the
Carter-Forest gray 1-D IMC step, IMC/IMC1D_CarterForest.py (slab and sphere).

Only tests/, the smoke check and bench.py's CPU baseline call this module;
the product path never does.

Same Philox streams as the device engine: boundary particles, fixed sources
and the comb as the Fleck-Cummings 1-D oracle (oracle.py); emission = the
tilted volume sampler over the per-cell Carter-Forest emission energies
with max(1, round(N E_c / E)) particles per cell, each time mapped from slot
0 of its stream through the truncated exponential of
IMC/IMC1D_CarterForest.py:531-539; transport draws per event from slots
6e..6e+4 (imc_oracle_1d.c, o1_transport_cf).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .oracle import (A_RAD, C_LIGHT, TAG_EMISSION, TAG_SOURCE, Oracle1D, OracleState1D, Pool1D, _dp, _f64, _i32p,
                     _i64p, _lib1d, _p, canonical_sum, comb_1d, counts, sample_boundary_1d, sample_volume_1d)


def _cf1_lib():
    L = _lib1d()
    if not hasattr(L, "_cf1_ready"):
        L.o1_cf_time.argtypes = [C.c_double, C.c_double, C.c_double]
        L.o1_cf_time.restype = C.c_double
        L.o1_cf_emission_times.argtypes = [C.c_int64, _i32p, _dp, C.c_double, C.c_uint64, C.c_uint32, C.c_int32,
                                           C.c_uint64, _dp]
        L.o1_transport_cf.argtypes = [C.c_int, C.c_int64, _dp, _dp, _dp, _dp, _i32p, C.c_int, _dp, _dp, _dp,
                                      C.c_double, _i32p, C.c_double, C.c_int64, C.c_double, C.c_uint64, C.c_uint32,
                                      C.c_uint64, _dp, _dp, _dp, _i64p, C.c_int32]
        L._cf1_ready = True
    return L


def cf1_time(rate, dt, xi):
    """Truncated-exponential emission time (IMC/IMC1D_CarterForest.py:531-539)."""
    return _cf1_lib().o1_cf_time(float(rate), float(dt), float(xi))


def cf1_beta(T, cvv):
    """beta = 4 a T^3 / c_v (IMC/IMC1D_CarterForest.py:1180; no temperature floor)."""
    T = _f64(T)
    return 4.0 * A_RAD * T ** 3 / cvv


def cf1_emission(T, sigma_a, beta, vol, dt):
    """(rate, per-cell emission energies E_c = e_c V_c) of
    IMC/IMC1D_CarterForest.py:484-494 (slab) and 913-918 (sphere)."""
    T = _f64(T)
    rate = C_LIGHT * sigma_a * beta
    T4 = T ** 4
    e = np.empty_like(T)
    for i in range(T.size):
        if rate[i] > 1e-10:
            e[i] = A_RAD * T4[i] * (1.0 - np.exp(-rate[i] * dt)) / beta[i]
        else:
            e[i] = A_RAD * T4[i] * C_LIGHT * sigma_a[i] * dt
    return rate, e * vol


def cf1_counts(E, N, sph):
    """max(1, round(N E_c / E)) per cell, none when E < 1e-30: slab
    N * (E_c / E) (IMC/IMC1D_CarterForest.py:504-505), sphere N * E_c / E
    (IMC/IMC1D_CarterForest.py:927); E summed in canonical order."""
    E = _f64(E)
    Et = canonical_sum(E)
    if not Et >= 1e-30:
        return np.zeros(E.size, np.int64)
    q = (N * E) / Et if sph else N * (E / Et)
    return np.maximum(1, np.rint(q).astype(np.int64))


def cf1_emission_times(pool, rate, dt, seed, step, base, tag=TAG_EMISSION):
    """Replace pool.t (a batch from sample_volume_1d with this tag, step and
    stream base) by the Carter-Forest emission times."""
    cell = np.ascontiguousarray(pool.cell, dtype=np.int32)
    t = np.empty(len(pool))
    _cf1_lib().o1_cf_emission_times(len(pool), _p(cell, _i32p), _p(_f64(rate)), float(dt), int(seed), int(step),
                                    int(tag), int(base), _p(t))
    pool.t = t
    return pool


def transport_1d_cf(sph, pool, edges, sigma_a, beta, dt, reflect, weight_floor, max_events, fastpath, seed, step,
                    base, nthreads=0):
    """Transport the pool in place (IMC/IMC1D_CarterForest.py:210-332,
    723-819).  Returns (dep_raw, path_raw, leak, stats[8]): captured weight
    and sum of w s per cell, escaped weight."""
    e = _f64(edges); I = e.size - 1
    for name in ("w", "mu", "t", "x"):
        setattr(pool, name, _f64(getattr(pool, name)).copy())
    pool.cell = np.ascontiguousarray(pool.cell, dtype=np.int32)
    refl = np.array([1 if r else 0 for r in reflect], dtype=np.int32)
    dep = np.zeros(I); path = np.zeros(I); lk = np.zeros(1); st = np.zeros(8, np.int64)
    _cf1_lib().o1_transport_cf(1 if sph else 0, len(pool), _p(pool.w), _p(pool.mu), _p(pool.t), _p(pool.x),
                               _p(pool.cell, _i32p), I, _p(e), _p(_f64(sigma_a)), _p(_f64(beta)), float(dt),
                               _p(refl, _i32p), float(weight_floor), int(max_events), float(fastpath), int(seed),
                               int(step), int(base), _p(dep), _p(path), _p(lk), _p(st, _i64p), int(nthreads))
    return dep, path, float(lk[0]), st


class OracleCF1D(Oracle1D):
    """FP64 CPU restatement of IMC/IMC1D_CarterForest.py's step (1151-1378);
    init_simulation is the Fleck-Cummings 1-D one."""

    def step(self, state, Ntarget, Nboundary, Nsource, NMax, T_boundary, dt, sigma_a_func, inv_eos, cv, source,
             reflect=(False, False), use_scalar_intensity_Tr=True, conserve_comb_energy=False,
             event_cap_per_particle=0, fastpath_threshold=0.0):
        I = self.I
        sid = state.count
        T = state.temperature
        sa = np.asarray(sigma_a_func(T), dtype=np.float64) * np.ones(I)
        beta = cf1_beta(T, np.asarray(cv(T), dtype=np.float64) * np.ones(I))
        pool = state.pool
        be = 0.0
        for side in (0, 1):
            Tb = T_boundary[side]
            Tb = Tb(state.time) if callable(Tb) else Tb
            if Tb > 0 and Nboundary > 0:
                R = self.edges[0] if side == 0 else self.edges[-1]
                E = _lib1d().o1_boundary_energy(1 if self.sph else 0, Tb, dt, R)
                b = sample_boundary_1d(self.sph, side, int(Nboundary), E, dt, self.edges, self.seed, sid, self.base)
                pool = pool.concat(b)
                be += float(np.sum(b.w))
        src_e = 0.0
        src = np.asarray(source, dtype=np.float64)
        if np.max(src) > 0 and Nsource > 0:
            QV = src * dt * self.vol
            cnt, _ = counts(QV, Nsource, False)
            s = sample_volume_1d(self.sph, self.edges, QV, cnt, 2, T, dt, self.seed, sid, TAG_SOURCE, self.base)
            pool = pool.concat(s)
            src_e = float(np.sum(s.w))
        rate, Ec = cf1_emission(T, sa, beta, self.vol, dt)
        cnt = cf1_counts(Ec, Ntarget, self.sph) if Ntarget > 0 else np.zeros(I, np.int64)
        if cnt.sum() > 0:
            e = sample_volume_1d(self.sph, self.edges, Ec, cnt, 0, T, dt, self.seed, sid, TAG_EMISSION, self.base)
            cf1_emission_times(e, rate, dt, self.seed, sid, self.base)
            pool = pool.concat(e)
        n_tr = len(pool)
        floor = 1e-10 * canonical_sum(pool.w) / max(n_tr, 1)
        dep_raw, path_raw, leak, stats = transport_1d_cf(self.sph, pool, self.edges, sa, beta, dt, reflect, floor,
                                                         event_cap_per_particle, fastpath_threshold, self.seed, sid,
                                                         self.base, self.nthreads)
        u = state.internal_energy + dep_raw / self.vol - Ec / self.vol
        u = np.maximum(u, 1e-10)
        Tn = inv_eos(u)
        si = path_raw / self.vol / dt
        if use_scalar_intensity_Tr:
            Tr = (si / A_RAD / C_LIGHT) ** 0.25
        else:
            valid = (pool.cell >= 0) & (pool.cell < I) & (pool.w > 0)
            Tr = (np.bincount(pool.cell[valid], weights=pool.w[valid], minlength=I) / self.vol / A_RAD) ** 0.25
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
                "comb_discrepancy": 0.0 if conserve_comb_energy else float(np.sum(disc)), "dep": dep_raw / self.vol,
                "si": si, "emitted_energies": Ec, "stats": stats, "n_transported": n_tr, "weight_floor": floor,
                "beta": beta, "rate": rate}
        return state, info


__all__ = ["OracleCF1D", "OracleState1D", "Pool1D", "transport_1d_cf", "cf1_emission_times", "cf1_emission",
           "cf1_counts", "cf1_beta", "cf1_time"]
