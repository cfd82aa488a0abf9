"""Oracle (test infrastructure only): FP64 CPU restatement of the
Carter-Forest gray 2-D IMC step, IMC/IMC2D_CarterForest.py.

Only tests/, the smoke check and bench.py's CPU baseline call this module;
the product path never does.

Same Philox streams as the device engine: face particles and the comb as
the Fleck-Cummings oracle (oracle.py); emission = the fixed-source sampler
over the per-cell Carter-Forest emission energies, with each particle's time
mapped from slot 0 of its stream through the truncated exponential of
IMC/IMC2D_CarterForest.py:51-56; transport draws per event from slots
6e..6e+5 (imc_oracle.c, orc_transport_2d_cf).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .oracle import (A_RAD, C_LIGHT, TAG_EMISSION, TAG_SOURCE, OracleState2D, Oracle2D, _f64, _i32p, _i64p, _dp,
                     _p, _face_profile, boundary_energy, canonical_sum, comb_2d, counts, lib, sample_boundary_2d,
                     sample_volume_2d)

NO_PROGRESS_REL_TOL = 1e-12  # IMC/IMC2D_CarterForest.py:45
NO_PROGRESS_MAX_STREAK = 8  # IMC/IMC2D_CarterForest.py:46


def _cf_lib():
    L = lib()
    if not hasattr(L, "_cf_ready"):
        L.orc_cf_time.argtypes = [C.c_double, C.c_double, C.c_double]
        L.orc_cf_time.restype = C.c_double
        L.orc_cf_emission_times.argtypes = [C.c_int64, _i32p, _i32p, C.c_int, _dp, C.c_double, C.c_uint64, C.c_uint32,
                                            C.c_int32, _dp]
        L.orc_transport_2d_cf.argtypes = [C.c_int64, _dp, _dp, _dp, _dp, _dp, _dp, _i32p, _i32p, _dp, _dp, C.c_int,
                                          C.c_int, _dp, _dp, C.c_double, _i32p, C.c_int64, C.c_int32, C.c_double,
                                          C.c_double, C.c_double, C.c_int64, C.c_uint64, C.c_uint32, _dp, _dp, _dp,
                                          _i64p, C.c_int32]
        L._cf_ready = True
    return L


def cf_time(rate, dt, xi):
    """Truncated-exponential emission time for uniform xi
    (IMC/IMC2D_CarterForest.py:51-56)."""
    return _cf_lib().orc_cf_time(float(rate), float(dt), float(xi))


def cf_rates_and_energies(T, sigma_a, beta, vol, dt):
    """rate = c sigma_a beta and the per-cell emission energies of
    IMC/IMC2D_CarterForest.py:876-884."""
    rate = C_LIGHT * sigma_a * beta
    em = np.zeros_like(T)
    mask = (rate > 1e-12) & (beta > 1e-15)
    em[mask] = A_RAD * np.maximum(T[mask], 0.0) ** 4 * (1.0 - np.exp(-rate[mask] * dt)) / beta[mask]
    em[~mask] = A_RAD * np.maximum(T[~mask], 0.0) ** 4 * C_LIGHT * sigma_a[~mask] * dt
    return rate, em * vol


def cf_beta(T, cvv):
    """beta = 4 a max(T, 1e-12)^3 / c_v (IMC/IMC2D_CarterForest.py:949)."""
    return 4.0 * A_RAD * np.maximum(T, 1e-12) ** 3 / cvv


def cf_emission_times(pool, ny, rate, dt, seed, step, tag=TAG_EMISSION):
    """Replace pool.t (a batch from sample_volume_2d with this tag and step)
    by the Carter-Forest emission times."""
    ci = np.ascontiguousarray(pool.ci, dtype=np.int32)
    cj = np.ascontiguousarray(pool.cj, dtype=np.int32)
    t = np.empty(len(pool))
    _cf_lib().orc_cf_emission_times(len(pool), _p(ci, _i32p), _p(cj, _i32p), ny, _p(_f64(rate).ravel()), float(dt),
                                    int(seed), int(step), int(tag), _p(t))
    pool.t = t
    return pool


def transport_2d_cf(pool, e1, e2, sigma_a, beta, dt, reflect, max_events, geometry_code, weight_floor, seed, step,
                    fastpath=0.0, tol=NO_PROGRESS_REL_TOL, max_streak=NO_PROGRESS_MAX_STREAK, nthreads=0):
    """Transport the pool in place; returns (dep_raw, si_raw, boundary loss,
    stats[8]) with dep_raw the captured energy and si_raw the sum of w s per
    cell, (nx, ny)."""
    e1 = _f64(e1); e2 = _f64(e2); nx, ny = e1.size - 1, e2.size - 1
    for name in ("w", "d1", "d2", "t", "p1", "p2"):
        setattr(pool, name, _f64(getattr(pool, name)).copy())
    pool.ci = np.ascontiguousarray(pool.ci, dtype=np.int32)
    pool.cj = np.ascontiguousarray(pool.cj, dtype=np.int32)
    dep = np.zeros(nx * ny); si = np.zeros(nx * ny); bl = C.c_double(); st = np.zeros(8, dtype=np.int64)
    refl = np.array([1 if r else 0 for r in reflect], dtype=np.int32)
    _cf_lib().orc_transport_2d_cf(len(pool), _p(pool.w), _p(pool.d1), _p(pool.d2), _p(pool.t), _p(pool.p1),
                                  _p(pool.p2), _p(pool.ci, _i32p), _p(pool.cj, _i32p), _p(e1), _p(e2), nx, ny,
                                  _p(_f64(sigma_a).ravel()), _p(_f64(beta).ravel()), float(dt), _p(refl, _i32p),
                                  int(max_events), int(geometry_code), float(weight_floor), float(fastpath),
                                  float(tol), int(max_streak), int(seed), int(step), _p(dep), _p(si), C.byref(bl),
                                  _p(st, _i64p), int(nthreads))
    return dep.reshape(nx, ny), si.reshape(nx, ny), bl.value, st


class OracleCF2D(Oracle2D):
    """FP64 CPU restatement of IMC/IMC2D_CarterForest.py's step (909-1160);
    init_simulation is the Fleck-Cummings one, as in the reference."""

    def step(self, state, Ntarget, Nboundary, Nsource, Nmax, T_boundary, dt, sigma_a_func, inv_eos, cv, source,
             reflect=(False, False, False, False), theta=1.0, use_scalar_intensity_Tr=True,
             conserve_comb_energy=False, rz_linear_source=True, max_events_per_particle=1_000_000,
             boundary_source_func=None, fastpath_threshold=0.0, no_progress_rel_tol=NO_PROGRESS_REL_TOL,
             no_progress_max_streak=NO_PROGRESS_MAX_STREAK):
        nx, ny = self.nx, self.ny
        step_id = state.count
        T = state.temperature
        sig = np.asarray(sigma_a_func(T), dtype=np.float64) * np.ones((nx, ny))
        cvv = np.asarray(cv(T), dtype=np.float64) * np.ones((nx, ny))
        beta = cf_beta(T, cvv)

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

        rate, E_c = cf_rates_and_energies(T, sig, beta, self.vol, dt)
        if canonical_sum(E_c) > 0.0 and Ntarget > 0:
            cnt, tot = counts(E_c, Ntarget, False)
            if tot > 0:
                e = sample_volume_2d(self.e1, self.e2, E_c, cnt, 0, gsrc, T, dt, self.seed, step_id, TAG_EMISSION)
                cf_emission_times(e, ny, rate, dt, self.seed, step_id)
                pool = pool.concat(e)

        n_tr = len(pool)
        weight_floor = 1e-10 * canonical_sum(pool.w) / max(n_tr, 1)
        dep_raw, si_raw, bl, stats = transport_2d_cf(pool, self.e1, self.e2, sig, beta, dt, reflect,
                                                     max_events_per_particle, self.gcode, weight_floor, self.seed,
                                                     step_id, fastpath_threshold, no_progress_rel_tol,
                                                     no_progress_max_streak, self.nthreads)
        u = state.internal_energy + dep_raw / self.vol - E_c / self.vol
        u = np.maximum(u, 1e-12)
        Tn = inv_eos(u)
        si = si_raw / (dt * self.vol)
        if use_scalar_intensity_Tr:
            Tr = (si / A_RAD / C_LIGHT) ** 0.25
        else:
            valid = pool.w > 0.0
            rad = np.bincount(pool.ci[valid].astype(np.int64) * ny + pool.cj[valid], weights=pool.w[valid],
                              minlength=nx * ny).reshape(nx, ny)
            Tr = (rad / self.vol / A_RAD) ** 0.25
        newp, disc = comb_2d(pool, Nmax, nx, ny, self.seed, step_id)
        if conserve_comb_energy:
            u = np.maximum(u + disc / self.vol, 1e-12)
            Tn = inv_eos(u)
        newp.t = np.zeros(len(newp))
        total_internal = float(np.sum(u * self.vol))
        total_rad = float(np.sum(newp.w))
        total = total_internal + total_rad
        loss = total - state.previous_total_energy - boundary_emission + bl - source_emission
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
            "energy_loss": loss, "comb_discrepancy": 0.0 if conserve_comb_energy else float(np.sum(disc)),
            "dep": dep_raw / self.vol, "si": si, "emitted_energies": E_c, "stats": stats, "n_transported": n_tr,
            "weight_floor": weight_floor, "beta": beta, "rate": rate,
        }
        return state, info


__all__ = ["OracleCF2D", "OracleState2D", "transport_2d_cf", "cf_emission_times", "cf_rates_and_energies", "cf_beta",
           "cf_time"]
