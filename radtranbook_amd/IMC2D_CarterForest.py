"""IMC2D_CarterForest -- Carter-Forest (time-delayed re-emission) gray 2-D
implicit Monte Carlo, xy and rz, on one MI355X.

Drop-in for the reference module IMC/IMC2D_CarterForest.py: the same
``step`` and ``run_simulation`` signatures (909-934, 1162-1190) and ``info``
keys; ``init_simulation`` and the state are IMC2D's, as in the reference,
which initialises with the Fleck-Cummings module.  No Fleck factor: particles
are absorbed at the true sigma_a and re-emitted isotropically after a delay
-ln(xi)/(c sigma_a beta) with beta = 4 a T^3 / c_v, or captured when the
delay runs past census.  Particles stay in device memory; the tracker runs in
FP64 (create the state with ``precision="fp64"``, the default).

Deliberate differences from the reference, beyond the random numbers:
emission weights use the per-cell emission energies directly; an emission
time is drawn from the same uniform the fixed-source sampler uses for a
uniform time; a weight at or below the floor is deposited in its cell (the
reference drops it from the balance).  Added ``info`` keys:
``comb_discrepancy``, ``energy_residual`` (the balance with the comb
discrepancy booked), ``deposited``, ``scalar_intensity``,
``emitted_energies`` and ``profiling.transport_kernel_ms``.
"""
from __future__ import annotations

import time as _time

import numpy as np

from ._lib import A_RAD, C_LIGHT, FP64, SRC_CF, SRC_FIXED, quarter_root
from .IMC2D import (SimulationState2D, _cell_volumes, _face_energy, _face_profile, _shape_from_edges,
                    init_simulation)

__c = C_LIGHT
__a = A_RAD

VERBOSE = True

NO_PROGRESS_REL_TOL = 1e-12  # IMC/IMC2D_CarterForest.py:45
NO_PROGRESS_MAX_STREAK = 8  # IMC/IMC2D_CarterForest.py:46

__all__ = ["SimulationState2D", "init_simulation", "step", "run_simulation", "emission_energies"]


def emission_energies(temperature, sigma_a, beta, volumes, dt):
    """Per-cell Carter-Forest emission energies and rates c sigma_a beta
    (IMC/IMC2D_CarterForest.py:876-884)."""
    rate = __c * sigma_a * beta
    emissions = np.zeros_like(temperature)
    mask = (rate > 1e-12) & (beta > 1e-15)
    emissions[mask] = __a * np.maximum(temperature[mask], 0.0) ** 4 * (1.0 - np.exp(-rate[mask] * dt)) / beta[mask]
    emissions[~mask] = __a * np.maximum(temperature[~mask], 0.0) ** 4 * __c * sigma_a[~mask] * dt
    return emissions * volumes, rate


def step(state, Ntarget, Nboundary, Nsource, Nmax, T_boundary, dt, edges1, edges2, sigma_a_func, inv_eos, cv, source,
         reflect=(False, False, False, False), theta=1.0, use_scalar_intensity_Tr=True, conserve_comb_energy=False,
         geometry="xy", rz_linear_source=True, max_events_per_particle=1_000_000, boundary_source_func=None,
         fastpath_threshold=0.0, no_progress_rel_tol=NO_PROGRESS_REL_TOL,
         no_progress_max_streak=NO_PROGRESS_MAX_STREAK):
    """Advance one Carter-Forest 2-D step (IMC/IMC2D_CarterForest.py:909-1160).

    ``theta`` is accepted and ignored, as in the reference."""
    del theta
    eng = state._engine
    nx, ny = _shape_from_edges(edges1, edges2)
    if (nx, ny) != eng.shape or geometry != eng.geometry:
        raise ValueError("step(): mesh or geometry differs from init_simulation")
    if eng.precision != FP64:
        raise ValueError("step(): the Carter-Forest tracker runs in FP64; initialise with precision='fp64'")
    e1, e2 = eng.e1, eng.e2
    t_step_start = _time.perf_counter()
    volumes = getattr(eng, "_volumes", None)
    if volumes is None:
        volumes = eng._volumes = _cell_volumes(e1, e2, geometry)
    state._invalidate()
    eng.set_step(state.count)

    temperature = state.temperature
    sigma_a = np.asarray(sigma_a_func(temperature), dtype=np.float64) * np.ones((nx, ny))
    beta = 4.0 * __a * np.maximum(temperature, 1e-12) ** 3 / (np.asarray(cv(temperature), dtype=np.float64)
                                                              * np.ones((nx, ny)))
    eng.set_material_cf(temperature, sigma_a, beta)

    boundary_emission = 0.0
    if Nboundary > 0:
        for side in range(4):
            Tb = T_boundary[side]
            Tb = Tb(state.time) if callable(Tb) else Tb
            prof = None
            if boundary_source_func is not None:
                prof = _face_profile(geometry, side, e1, e2, dt, boundary_source_func)
                if prof is None:
                    continue
                E = float(prof[-1])
            else:
                if Tb <= 0.0:
                    continue
                E = _face_energy(geometry, side, float(Tb), dt, e1, e2)
            eng.sample_boundary(side, Nboundary, E, dt, prof)
            boundary_emission += E

    source_emission = 0.0
    src = np.asarray(source, dtype=np.float64)
    if Nsource > 0 and np.max(src) > 0.0:
        _, tot, cnt = eng.sample_volume(SRC_FIXED, Nsource, src * dt * volumes, dt, rz_linear_source)
        if cnt > 0:
            source_emission = tot

    emitted_energies, _ = emission_energies(temperature, sigma_a, beta, volumes, dt)
    if float(np.sum(emitted_energies)) > 0.0 and Ntarget > 0:
        eng.sample_volume(SRC_CF, Ntarget, emitted_energies, dt, rz_linear_source)

    t_transport_start = _time.perf_counter()
    n_transported = eng.n
    dep_cell, si_cell, boundary_loss, stats = eng.transport_cf(dt, reflect, max_events_per_particle, -1.0,
                                                               fastpath_threshold, no_progress_rel_tol,
                                                               no_progress_max_streak)
    t_post_start = _time.perf_counter()

    internal_energy = state.internal_energy + dep_cell - emitted_energies / volumes
    internal_energy = np.maximum(internal_energy, 1e-12)
    temperature = inv_eos(internal_energy)

    ecen, comb_disc, ecen_total, total_rad = eng.comb(Nmax, want_ecen=not use_scalar_intensity_Tr,
                                                      want_disc=conserve_comb_energy)
    if use_scalar_intensity_Tr:
        radiation_temperature = quarter_root(si_cell * (1.0 / (__a * __c)))
    else:
        radiation_temperature = quarter_root(ecen / volumes / __a)
    if conserve_comb_energy:
        internal_energy = np.maximum(internal_energy + comb_disc / volumes, 1e-12)
        temperature = inv_eos(internal_energy)

    total_internal = float(np.sum(internal_energy * volumes))
    total_energy = total_internal + total_rad
    energy_loss = total_energy - state.previous_total_energy - boundary_emission + boundary_loss - source_emission
    disc_total = 0.0 if conserve_comb_energy else ecen_total - total_rad

    state.internal_energy = internal_energy
    state.temperature = temperature
    state.radiation_temperature = radiation_temperature
    state.time += dt
    state.previous_total_energy = total_energy
    state.count += 1
    state._invalidate()

    t_end = _time.perf_counter()
    events_total = int(stats[0])
    n_tr = max(int(n_transported), 1)
    info = {
        "time": state.time,
        "temperature": temperature,
        "radiation_temperature": radiation_temperature,
        "N_particles": eng.n,
        "total_energy": total_energy,
        "total_internal_energy": total_internal,
        "total_radiation_energy": total_rad,
        "boundary_emission": boundary_emission,
        "boundary_loss": boundary_loss,
        "source_emission": source_emission,
        "energy_loss": energy_loss,
        "comb_discrepancy": disc_total,
        "energy_residual": energy_loss + disc_total,
        "deposited": dep_cell,
        "scalar_intensity": si_cell,
        "emitted_energies": emitted_energies,
        "profiling": {
            "phase_times_s": {
                "sampling": t_transport_start - t_step_start,
                "transport": t_post_start - t_transport_start,
                "postprocess": t_end - t_post_start,
                "total": t_end - t_step_start,
            },
            "transport_kernel_ms": eng.last_transport_ms(),
            "transport_events": {
                "total": events_total,
                "boundary_crossings": int(stats[1]),
                "absorption_continue_events": int(stats[2]),
                "census_events": int(stats[3]),
                "absorption_capture_events": int(stats[4]),
                "weight_floor_kills": int(stats[5]),
                "reflections": int(stats[6]),
                "event_cap_hits": int(stats[7]),
                "avg_events_per_particle": events_total / n_tr,
                "n_particles_transported": int(n_tr),
            },
        },
    }
    return state, info


def run_simulation(Ntarget, Nboundary, Nsource, Nmax, Tinit, Tr_init, T_boundary, dt, edges1, edges2, sigma_a_func,
                   eos, inv_eos, cv, source, final_time, reflect=(False, False, False, False), output_freq=1,
                   theta=1.0, use_scalar_intensity_Tr=True, Ntarget_ic=None, conserve_comb_energy=False,
                   geometry="xy", rz_linear_source=True, max_events_per_particle=1_000_000, fastpath_threshold=0.0,
                   no_progress_rel_tol=NO_PROGRESS_REL_TOL, no_progress_max_streak=NO_PROGRESS_MAX_STREAK, seed=1,
                   device=0):
    """Run a Carter-Forest 2-D simulation (IMC/IMC2D_CarterForest.py:1162-1251);
    returns (times, Tr history, T history)."""
    state = init_simulation(Ntarget, Tinit, Tr_init, edges1, edges2, eos, inv_eos, Ntarget_ic=Ntarget_ic,
                            geometry=geometry, precision="fp64", seed=seed, device=device)
    times = [0.0]
    Tr_hist = [state.radiation_temperature.copy()]
    T_hist = [state.temperature.copy()]
    while state.time < final_time - 1e-15:
        step_dt = min(dt, final_time - state.time)
        state, info = step(state, Ntarget, Nboundary, Nsource, Nmax, T_boundary, step_dt, edges1, edges2,
                           sigma_a_func, inv_eos, cv, source, reflect=reflect, theta=theta,
                           use_scalar_intensity_Tr=use_scalar_intensity_Tr, conserve_comb_energy=conserve_comb_energy,
                           geometry=geometry, rz_linear_source=rz_linear_source,
                           max_events_per_particle=max_events_per_particle, fastpath_threshold=fastpath_threshold,
                           no_progress_rel_tol=no_progress_rel_tol, no_progress_max_streak=no_progress_max_streak)
        if (state.count - 1) % output_freq == 0 or (final_time - state.time) < 1e-12:
            times.append(info["time"])
            Tr_hist.append(state.radiation_temperature.copy())
            T_hist.append(state.temperature.copy())
            if VERBOSE:
                print("{:.6f}".format(info["time"]), info["N_particles"], "{:.6f}".format(info["total_energy"]),
                      "{:.6f}".format(info["total_internal_energy"]), "{:.6f}".format(info["total_radiation_energy"]),
                      "{:.6f}".format(info["boundary_emission"]), "{:.6e}".format(info["energy_loss"]), sep="\t")
    return np.array(times), np.array(Tr_hist), np.array(T_hist)
