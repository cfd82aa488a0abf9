"""IMC1D_CarterForest -- Carter-Forest (time-delayed re-emission) gray 1-D
implicit Monte Carlo, slab and sphere, on one MI355X.

Drop-in for the reference module IMC/IMC1D_CarterForest.py: the same
``init_simulation``, ``step`` and ``run_simulation`` signatures (1103, 1151,
1381) and ``info`` keys.  No Fleck factor: particles are absorbed at the
true sigma_a and re-emitted isotropically after a delay -ln(xi)/(c sigma_a
beta) with beta = 4 a T^3 / c_v, or captured when the delay runs past
census.  Particles stay in device memory; the tracker runs in FP64.

Deliberate differences from the reference, beyond the random numbers: the
initial census and the boundary and fixed sources are the Fleck-Cummings
module's (IMC/IMC1D.py: ceil-allocated equilibrium census with uniform
positions, slab face particles at x = 1e-8); an emission time is drawn from
the same uniform the fixed-source sampler uses for a uniform time; a history
whose weight starts at or below the floor stays in census (the reference
drops it unbooked).  Added ``info`` keys: ``comb_discrepancy``,
``energy_residual`` (the balance with the comb discrepancy booked),
``deposited``, ``scalar_intensity``, ``emitted_energies`` and
``profiling.transport_kernel_ms``.
"""
from __future__ import annotations

import time as _time

import numpy as np

from ._lib import A_RAD, C_LIGHT, SRC_CF, SRC_FIXED, quarter_root
from .IMC1D import SimulationState, _boundary_energy, _cell_volumes
from .IMC1D import init_simulation as _init_fc

__c = C_LIGHT
__a = A_RAD

VERBOSE = True

__all__ = ["SimulationState", "init_simulation", "step", "run_simulation", "emission_energies"]


def emission_energies(temperature, sigma_a, beta, volumes, dt):
    """Per-cell Carter-Forest emission energies and rates c sigma_a beta
    (IMC/IMC1D_CarterForest.py:484-494, 913-918)."""
    T4 = np.asarray(temperature, dtype=np.float64) ** 4
    rate = __c * sigma_a * beta
    e = np.empty_like(T4)
    for i in range(T4.size):
        if rate[i] > 1e-10:
            e[i] = __a * T4[i] * (1.0 - np.exp(-rate[i] * dt)) / beta[i]
        else:
            e[i] = __a * T4[i] * __c * sigma_a[i] * dt
    return e * volumes, rate


def init_simulation(Ntarget, Tinit, Tr_init, mesh, eos, inv_eos, Ntarget_ic=None, geometry="slab", seed=1, device=0,
                    engine=None):
    """Initialise the material and the equilibrium census
    (IMC/IMC1D_CarterForest.py:1103-1148) on an FP64 device context."""
    return _init_fc(Ntarget, Tinit, Tr_init, mesh, eos, inv_eos, Ntarget_ic=Ntarget_ic, geometry=geometry,
                    precision="fp64", seed=seed, device=device, engine=engine)


def step(state, Ntarget, Nboundary, Nsource, NMax, T_boundary, dt, mesh, sigma_a_func, inv_eos, cv, source,
         reflect=(False, False), use_scalar_intensity_Tr=True, conserve_comb_energy=False, geometry="slab",
         event_cap_per_particle=0, fastpath_threshold=0.0):
    """Advance one Carter-Forest 1-D step (IMC/IMC1D_CarterForest.py:1151-1378)."""
    eng = state._engine
    mesh = np.asarray(mesh, dtype=np.float64)
    I = mesh.shape[0]
    if I != eng.ncell or geometry != eng.geometry:
        raise ValueError("step(): mesh or geometry differs from init_simulation")
    if eng.precision != 0:
        raise ValueError("step(): the Carter-Forest tracker runs in FP64; initialise with precision='fp64'")
    volumes = _cell_volumes(mesh, geometry)
    t_step_start = _time.perf_counter()
    state._invalidate()
    eng.set_step(state.count)

    temperature = state.temperature
    sigma_a = np.asarray(sigma_a_func(temperature), dtype=np.float64) * np.ones(I)
    beta = 4.0 * __a * temperature ** 3 / (np.asarray(cv(temperature), dtype=np.float64) * np.ones(I))
    eng.set_material_cf(temperature, sigma_a, beta, dt)

    t_sampling_start = _time.perf_counter()
    boundary_emission = 0.0
    edges = eng.edges[0]
    for side in (0, 1):
        Tb = T_boundary[side]
        Tb = Tb(state.time) if callable(Tb) else Tb
        if Tb > 0 and Nboundary > 0:
            R = edges[0] if side == 0 else edges[-1]
            E = _boundary_energy(geometry, float(Tb), dt, R)
            eng.sample_boundary(side, int(Nboundary), E)
            boundary_emission += E

    source_emission = 0.0
    src = np.asarray(source, dtype=np.float64) * np.ones(I)
    if np.max(src) > 0 and Nsource > 0:
        _, tot = eng.sample_volume(SRC_FIXED, int(Nsource), src * dt * volumes)
        source_emission = float(tot[0])

    emitted_energies, _ = emission_energies(temperature, sigma_a, beta, volumes, dt)
    if Ntarget > 0:
        eng.sample_volume(SRC_CF, int(Ntarget), emitted_energies)

    t_transport_start = _time.perf_counter()
    dep, si, leak, stats, n_transported = eng.transport_cf(reflect, -1.0, int(event_cap_per_particle),
                                                           float(fastpath_threshold))
    t_post_start = _time.perf_counter()

    internal_energy = np.maximum(state.internal_energy + dep - emitted_energies / volumes, 1e-10)
    temperature = inv_eos(internal_energy)
    boundary_loss = float(leak[0])

    ecen, comb_disc, ecen_total, total_rad = eng.comb(int(NMax), want_ecen=not use_scalar_intensity_Tr,
                                                      want_disc=conserve_comb_energy)
    if use_scalar_intensity_Tr:
        radiation_temperature = quarter_root(si / __a / __c)
    else:
        radiation_temperature = quarter_root(ecen / volumes / __a)
    if conserve_comb_energy:
        internal_energy = internal_energy + comb_disc / volumes
        temperature = inv_eos(internal_energy)

    total_internal = float(np.sum(internal_energy * volumes))
    total_rad = float(total_rad[0])
    total_energy = total_internal + total_rad
    energy_loss = total_energy - state.previous_total_energy - boundary_emission + boundary_loss - source_emission
    disc_total = 0.0 if conserve_comb_energy else float(ecen_total[0]) - total_rad

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
        "radiation_temperature": radiation_temperature,
        "temperature": temperature,
        "N_particles": eng.census_counts()[1],
        "total_energy": total_energy,
        "total_internal_energy": total_internal,
        "total_radiation_energy": total_rad,
        "boundary_emission": boundary_emission,
        "boundary_loss": boundary_loss,
        "source_emission": source_emission,
        "energy_loss": energy_loss,
        "comb_discrepancy": disc_total,
        "energy_residual": energy_loss + disc_total,
        "deposited": dep,
        "scalar_intensity": si,
        "emitted_energies": emitted_energies,
        "profiling": {
            "phase_times_s": {
                "sampling": t_transport_start - t_sampling_start,
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
    if VERBOSE:
        print("{:.6f}".format(state.time), info["N_particles"], "{:.6f}".format(total_energy),
              "{:.6f}".format(total_internal), "{:.6f}".format(total_rad), "{:.6f}".format(boundary_emission),
              "{:.6f}".format(energy_loss), sep="\t")
    return state, info


def run_simulation(Ntarget, Nboundary, Nsource, NMax, Tinit, Tr_init, T_boundary, dt, mesh, sigma_a_func, eos,
                   inv_eos, cv, source, final_time, reflect=(False, False), output_freq=1,
                   use_scalar_intensity_Tr=True, Ntarget_ic=None, conserve_comb_energy=False, geometry="slab",
                   event_cap_per_particle=0, fastpath_threshold=0.0, seed=1, device=0):
    """Run a Carter-Forest 1-D simulation (IMC/IMC1D_CarterForest.py:1381-1413);
    returns (times, Tr history, T history)."""
    state = init_simulation(Ntarget, Tinit, Tr_init, mesh, eos, inv_eos, Ntarget_ic=Ntarget_ic, geometry=geometry,
                            seed=seed, device=device)
    Tr_hist = [state.radiation_temperature.copy()]
    T_hist = [state.temperature.copy()]
    times = [0.0]
    while state.time < final_time:
        step_dt = min(dt, final_time - state.time)
        state, info = step(state, Ntarget, Nboundary, Nsource, NMax, T_boundary, step_dt, mesh, sigma_a_func,
                           inv_eos, cv, source, reflect=reflect, use_scalar_intensity_Tr=use_scalar_intensity_Tr,
                           conserve_comb_energy=conserve_comb_energy, geometry=geometry,
                           event_cap_per_particle=event_cap_per_particle, fastpath_threshold=fastpath_threshold)
        if state.count % output_freq == 0:
            Tr_hist.append(state.radiation_temperature.copy())
            T_hist.append(state.temperature.copy())
            times.append(state.time)
    return np.array(times), np.array(Tr_hist), np.array(T_hist)
