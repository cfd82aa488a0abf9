"""MG_IMC1D -- multigroup 1-D spherical implicit Monte Carlo on one MI355X.

Drop-in for the reference module MG_IMC/MG_IMC1D.py: the same
``init_simulation``, ``step`` and ``run_simulation`` signatures (850, 917,
1312), the same ``SimulationState1DMG`` attributes and the same ``info``
keys.  Particles live in device memory; host code evaluates the user
callables (``sigma_a_funcs``, ``cv``, ``inv_eos``) on cell arrays and keeps
the per-cell energy bookkeeping.

Deliberate differences from the reference, beyond the random numbers: B_g is
the Planck group integral (the reference needs the absent planck_integrals
module, whose gray fallback splits B equally over the groups).  Added
``info`` keys: ``comb_discrepancy`` (G, I; in the "radiation" mode the
energy of bins whose combed copies all vanished, which the reference drops
without booking), ``balance_residual`` (the residual with every unbooked
comb loss booked), ``deposited``,
``scalar_intensity``, ``emitted_by_group`` and
``profiling.transport_kernel_ms``.

Extra keyword arguments: ``precision`` ('fp64' or 'fp32' tracking), ``seed``,
``device``, ``engine`` (reuse the device context of an earlier state).
"""
from __future__ import annotations

import math
import time as _time

import numpy as np

from ._census import DeviceCensus

from ._lib import quarter_root, A_RAD, C_LIGHT, SRC_EMISSION, SRC_FIXED, EngineMG1D

__c = C_LIGHT
__a = A_RAD

VERBOSE = True


class SimulationState1DMG(DeviceCensus):
    """Mutable multigroup spherical state (MG_IMC/MG_IMC1D.py:116-138); the
    particle arrays are downloaded from the device on access."""

    _PARTICLE_FIELDS = ("weights", "mus", "times", "positions", "cell_indices", "groups")

    def __init__(self, engine, internal_energy, temperature, radiation_temperature, radiation_energy_by_group,
                 radiation_flux_by_group, time, previous_total_energy, count=0):
        DeviceCensus.__init__(self, engine)
        self.internal_energy = internal_energy
        self.temperature = temperature
        self.radiation_temperature = radiation_temperature
        self.radiation_energy_by_group = radiation_energy_by_group
        self.radiation_flux_by_group = radiation_flux_by_group
        self.radiation_energy_by_group_postcomb = None
        self.time = time
        self.previous_total_energy = previous_total_energy
        self.count = count


def _edges(mesh):
    mesh = np.asarray(mesh, dtype=np.float64)
    if mesh.ndim != 2 or mesh.shape[1] != 2:
        raise ValueError("mesh must be (I, 2) [[r_inner, r_outer], ...]")
    if np.any(mesh[1:, 0] != mesh[:-1, 1]):
        raise ValueError("mesh cells must be contiguous")
    return np.concatenate((mesh[:, 0], mesh[-1:, 1]))


def _volumes(mesh):
    mesh = np.asarray(mesh, dtype=np.float64)
    r0, r1 = mesh[:, 0], mesh[:, 1]
    return (4.0 / 3.0) * math.pi * (r1 * r1 * r1 - r0 * r0 * r0)


def _face_energy(T, dt, R):
    """a c T^4 / 4 dt 4 pi R^2 (MG_IMC1D.py:668, 1014-1017)."""
    t4 = (T * T) * (T * T)
    return __a * __c * t4 / 4.0 * dt * 4.0 * math.pi * R * R


def init_simulation(Ntarget, Tinit, Tr_init, mesh, energy_edges, eos, inv_eos, Ntarget_ic=None, T_emit_floor=0.0,
                    precision="fp64", seed=1, device=0, engine=None):
    """Initialise a multigroup spherical simulation (MG_IMC1D.py:850-914)."""
    e = _edges(mesh)
    I = e.size - 1
    energy_edges = np.asarray(energy_edges, dtype=np.float64)
    G = energy_edges.size - 1
    volumes = _volumes(mesh)
    internal_energy = eos(Tinit)
    temperature = np.array(Tinit, dtype=np.float64, copy=True)
    assert np.allclose(inv_eos(internal_energy), Tinit), "Inverse EOS failed"
    if engine is not None:
        eng = engine
        if eng.I != I or eng.G != G or eng.precision != {"fp64": 0, "fp32": 1}[precision]:
            raise ValueError("init_simulation(): engine was built for another mesh, group structure or precision")
        eng.set_seed(seed)
    else:
        eng = EngineMG1D(e, energy_edges, precision=precision, seed=seed, device=device)
    N_ic = Ntarget if Ntarget_ic is None else Ntarget_ic
    eng.equilibrium_sample(N_ic, np.asarray(Tr_init, dtype=np.float64), T_emit_floor)
    Eg, _, total_rad = eng.census()
    radiation_temperature = quarter_root(Eg.sum(axis=0) / volumes / __a)
    total_internal = float(np.sum(internal_energy * volumes))
    prev_total = total_internal + total_rad
    if VERBOSE:
        print("Time", "N", "Total Energy", "Total Internal", "Total Radiation", "Boundary Emission",
              "Energy Residual", sep="\t")
        print("=" * 120)
        print(f"{0.0:.6e}", eng.n, f"{prev_total:.6e}", f"{total_internal:.6e}", f"{total_rad:.6e}",
              f"{0.0:.6e}", f"{0.0:.6e}", sep="\t")
    return SimulationState1DMG(eng, internal_energy, temperature, radiation_temperature, Eg, np.zeros((G, I)), 0.0,
                               prev_total, 0)


def step(state, Ntarget, Nboundary, Nsource, Nmax, T_boundary, dt, mesh, energy_edges, sigma_a_funcs, inv_eos, cv,
         source, reflect=(False, False), theta=1.0, use_scalar_intensity_Tr=True, conserve_comb_energy=True,
         max_events_per_particle=1_000_000, T_emit_floor=0.0, particle_budget_fmin=0.0, Nmax_growth=0,
         Nmax_final=None):
    """Advance one multigroup spherical step (MG_IMC1D.py:917-1309).

    ``max_events_per_particle`` is accepted and unused, as in the reference."""
    t0 = _time.perf_counter()
    eng = state._engine
    e = _edges(mesh)
    I, G = eng.I, eng.G
    if e.size - 1 != I or len(energy_edges) - 1 != G:
        raise ValueError("step(): mesh or energy groups differ from init_simulation")
    volumes = _volumes(mesh)
    state._invalidate()
    eng.set_step(state.count)
    temperature = state.temperature
    sigma = np.array([np.asarray(f(temperature), dtype=np.float64) * np.ones(I) for f in sigma_a_funcs])
    if sigma.shape != (G, I):
        raise ValueError("sigma_a_funcs must give one opacity array per group")
    cvv = np.asarray(cv(temperature), dtype=np.float64) * np.ones(I)
    mats, E_emit_total = eng.set_material(temperature, sigma, cvv, theta, dt, T_emit_floor, want=("E",))
    Ebin = mats["E"]

    T_inner = T_boundary[0](state.time) if callable(T_boundary[0]) else T_boundary[0]
    T_outer = T_boundary[1](state.time) if callable(T_boundary[1]) else T_boundary[1]
    src = np.asarray(source, dtype=np.float64)
    if src.ndim == 1:
        src_mg = np.zeros((G, I))
        src_mg[0, :] = src  # gray source into group 0 (MG_IMC1D.py:695-698)
    elif src.ndim == 2:
        src_mg = src.copy()
    else:
        raise ValueError(f"source shape {src.shape} not supported")

    if particle_budget_fmin > 0.0:
        E_in = _face_energy(T_inner, dt, e[0]) if T_inner > 0.0 else 0.0
        E_out = _face_energy(T_outer, dt, e[-1]) if T_outer > 0.0 else 0.0
        E_src = float(np.sum(np.maximum(src_mg, 0.0) * volumes[np.newaxis, :] * dt))
        E_all = E_in + E_out + E_src + E_emit_total
        N_total = int(Ntarget)
        N_floor = max(1, int(round(particle_budget_fmin * N_total)))
        if E_all > 0.0:
            part = lambda E: max(int(round(N_total * E / E_all)), N_floor) if E > 0.0 else 0
            N_in, N_out, N_src, N_emit = part(E_in), part(E_out), part(E_src), part(E_emit_total)
        else:
            N_in, N_out, N_src, N_emit = N_total, 0, 0, 0
    else:
        N_in = N_out = int(Nboundary)
        N_src, N_emit = int(Nsource), int(Ntarget)

    boundary_emission = 0.0
    N_boundary = 0
    for side, Tb, Nb in ((0, T_inner, N_in), (1, T_outer, N_out)):
        if Nb > 0 and Tb > 0.0:
            boundary_emission += eng.sample_face(side, Nb, Tb, dt)
            N_boundary += Nb

    source_emission = 0.0
    if N_src > 0 and np.max(src_mg) > 0.0:
        QV = src_mg * volumes[np.newaxis, :] * dt
        source_emission = eng.sample_volume(SRC_FIXED, N_src, dt, QV)

    emitted_by_group = np.zeros((G, I))
    if N_emit > 0 and E_emit_total > 0.0:
        eng.sample_volume(SRC_EMISSION, N_emit, dt)
        emitted_by_group = Ebin

    t_transport_start = _time.perf_counter()
    n_transported = eng.n
    dep, si, bl_by_group, stats = eng.transport(dt, reflect, -1.0)
    if stats[7] > 0:
        raise AssertionError("Non-positive distance in spherical transport")
    t_post_start = _time.perf_counter()
    boundary_loss = float(np.sum(bl_by_group))

    internal_energy = state.internal_energy + dep.sum(axis=0) - emitted_by_group.sum(axis=0) / volumes
    temperature = inv_eos(internal_energy)
    if use_scalar_intensity_Tr:
        radiation_temperature_new = quarter_root(si.sum(axis=0) / __a / __c)
    else:
        Eg_live, _, _ = eng.census()
        radiation_temperature_new = quarter_root(Eg_live.sum(axis=0) / volumes / __a)

    radiation_mode = conserve_comb_energy == "radiation"
    comb_disc, _, combed = eng.comb(int(Nmax), radiation=radiation_mode)
    if combed and (conserve_comb_energy is True or conserve_comb_energy == "material") and np.any(comb_disc != 0.0):
        internal_energy = internal_energy + comb_disc.sum(axis=0) / volumes
        temperature = inv_eos(internal_energy)

    Eg, wmu, total_rad = eng.census()
    radiation_flux_by_group = __c * wmu / volumes[np.newaxis, :]
    total_internal = float(np.sum(internal_energy * volumes))
    total_energy = total_internal + total_rad
    energy_residual = total_energy - state.previous_total_energy - boundary_emission + boundary_loss - source_emission
    # booked: the material mode's discrepancy; unbooked: a discarded
    # discrepancy, or in radiation mode the energy of bins left without copies
    unbooked = 0.0 if (conserve_comb_energy is True or conserve_comb_energy == "material") \
        else float(np.sum(comb_disc))

    state.internal_energy = internal_energy
    state.temperature = temperature
    state.radiation_temperature = radiation_temperature_new
    state.radiation_energy_by_group = Eg
    state.radiation_flux_by_group = radiation_flux_by_group
    state.time += dt
    state.previous_total_energy = total_energy
    state.count += 1
    state._invalidate()

    t_end = _time.perf_counter()
    n_tp = max(int(n_transported), 1)
    if Nmax_growth > 0:
        Nmax_next = min(Nmax + Nmax_growth, Nmax_final) if Nmax_final is not None else Nmax + Nmax_growth
    else:
        Nmax_next = Nmax
    info = {
        "time": state.time,
        "temperature": temperature.copy(),
        "radiation_temperature": radiation_temperature_new.copy(),
        "radiation_energy_by_group": Eg.copy(),
        "radiation_flux_by_group": radiation_flux_by_group.copy(),
        "N_particles": eng.n,
        "total_energy": total_energy,
        "total_internal_energy": total_internal,
        "total_radiation_energy": total_rad,
        "boundary_emission": boundary_emission,
        "N_boundary": N_boundary,
        "boundary_loss": boundary_loss,
        "boundary_loss_by_group": bl_by_group.copy(),
        "source_emission": source_emission,
        "energy_residual": energy_residual,
        "Nmax_next": Nmax_next,
        "comb_discrepancy": comb_disc,
        "balance_residual": energy_residual + unbooked,
        "deposited": dep,
        "scalar_intensity": si,
        "emitted_by_group": emitted_by_group,
        "profiling": {
            "phase_times_s": {
                "sampling": t_transport_start - t0,
                "transport": t_post_start - t_transport_start,
                "postprocess": t_end - t_post_start,
                "total": t_end - t0,
            },
            "transport_kernel_ms": eng.last_transport_ms(),
            "transport_events": {
                "total": int(stats[0]),
                "boundary_crossings": int(stats[1]),
                "scatter_events": int(stats[2]),
                "census_events": int(stats[3]),
                "weight_floor_kills": int(stats[5]),
                "reflections": int(stats[6]),
                "avg_events_per_particle": int(stats[0]) / n_tp,
                "n_particles_transported": int(n_transported),
            },
        },
    }
    return state, info


def run_simulation(Ntarget, Nboundary, Nsource, Nmax, Tinit, Tr_init, T_boundary, dt, mesh, energy_edges,
                   sigma_a_funcs, eos, inv_eos, cv, source, final_time, reflect=(False, False), output_freq=1,
                   theta=1.0, use_scalar_intensity_Tr=True, Ntarget_ic=None, conserve_comb_energy=True,
                   Nmax_growth=0, Nmax_final=None, precision="fp64", seed=1, device=0):
    """Run from t=0 to final_time (MG_IMC1D.py:1312-1369).  Returns
    (time_values, Tr_history, T_history, state, history)."""
    state = init_simulation(Ntarget, Tinit, Tr_init, mesh, energy_edges, eos, inv_eos, Ntarget_ic=Ntarget_ic,
                            precision=precision, seed=seed, device=device)
    time_values = [0.0]
    Tr_history = [state.radiation_temperature.copy()]
    T_history = [state.temperature.copy()]
    history = []
    time_tol = max(1e-15, 1e-12 * max(final_time, 1.0))
    while state.time < final_time - time_tol:
        step_dt = min(dt, final_time - state.time)
        state, info = step(state, Ntarget, Nboundary, Nsource, Nmax, T_boundary, step_dt, mesh, energy_edges,
                           sigma_a_funcs, inv_eos, cv, source, reflect, theta=theta,
                           use_scalar_intensity_Tr=use_scalar_intensity_Tr, conserve_comb_energy=conserve_comb_energy,
                           Nmax_growth=Nmax_growth, Nmax_final=Nmax_final)
        Nmax = info["Nmax_next"]
        history.append(info)
        if (state.count - 1) % output_freq == 0:
            time_values.append(info["time"])
            Tr_history.append(state.radiation_temperature.copy())
            T_history.append(state.temperature.copy())
            if VERBOSE:
                print(f"{info['time']:.6e}", info["N_particles"], f"{info['total_energy']:.6e}",
                      f"{info['total_internal_energy']:.6e}", f"{info['total_radiation_energy']:.6e}",
                      f"{info['boundary_emission']:.6e}", f"{info['energy_residual']:.3e}", sep="\t")
    return np.array(time_values), np.array(Tr_history), np.array(T_history), state, history
