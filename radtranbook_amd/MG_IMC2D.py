"""MG_IMC2D -- multigroup 2-D implicit Monte Carlo (xy and rz) on one MI355X.

Drop-in for the reference module MG_IMC/MG_IMC2D.py: the same
``init_simulation``, ``step`` and ``run_simulation`` signatures (1431, 1635,
2125), the same ``SimulationState2DMG`` attributes and the same ``info`` keys.
Particles live in device memory; host code evaluates the user callables
(``sigma_a_funcs``, ``cv``, ``inv_eos``, ``boundary_source_func``) on cell
arrays and keeps the per-cell energy bookkeeping.

Deliberate differences from the reference, beyond the random numbers:
B_g is the Planck group integral (the reference needs the absent
planck_integrals module); ``boundary_source_func`` also works in xy (the
reference ignores it there); right/top and rmax/zmax face particles start in
the last cell (the reference loses or drops them); a reflecting left/rmin
face resets the cell index.  Added ``info`` keys: ``comb_discrepancy`` and
``balance_residual`` (the residual with the comb discrepancy booked),
``deposited`` (per cell, summed over groups), ``scalar_intensity`` (by
group), ``profiling.transport_kernel_ms``, and with ``diagnostics=True``
``deposited_by_group`` and ``emission_by_group``.

Extra keyword arguments: ``precision`` ('fp64' or 'fp32' tracking), ``seed``,
``device``.
"""
from __future__ import annotations

import time as _time

import numpy as np

from ._census import DeviceCensus

from ._lib import A_RAD, C_LIGHT, EngineMG, quarter_root

__c = C_LIGHT
__a = A_RAD

VERBOSE = True

# Called with the engine right before each transport (the launch-shape
# autotuner times its candidates there).
_TRANSPORT_HOOK = None

_SIDES = {"xy": ("left", "right", "bottom", "top"), "rz": ("rmin", "rmax", "zmin", "zmax")}


class SimulationState2DMG(DeviceCensus):
    """Mutable multigroup 2-D state (MG_IMC/MG_IMC2D.py:116-136); the
    particle arrays are downloaded from the device on access."""

    _PARTICLE_FIELDS = ("weights", "dir1", "dir2", "times", "pos1", "pos2", "cell_i", "cell_j", "groups")

    def __init__(self, engine, internal_energy, temperature, radiation_temperature, radiation_energy_by_group, time,
                 previous_total_energy, count=0):
        DeviceCensus.__init__(self, engine)
        self.internal_energy = internal_energy
        self.temperature = temperature
        self.radiation_temperature = radiation_temperature
        self.radiation_energy_by_group = radiation_energy_by_group
        self.radiation_energy_by_group_postcomb = None
        self.time = time
        self.previous_total_energy = previous_total_energy
        self.count = count


def _cell_volumes(edges1, edges2, geometry):
    e1 = np.asarray(edges1, dtype=np.float64)
    e2 = np.asarray(edges2, dtype=np.float64)
    if geometry == "xy":
        return np.diff(e1)[:, None] * np.diff(e2)[None, :]
    if geometry == "rz":
        return np.pi * (e1[1:] * e1[1:] - e1[:-1] * e1[:-1])[:, None] * np.diff(e2)[None, :]
    raise ValueError(f"Unknown geometry: {geometry}")


def _boundary_temperature_value(Tb, t):
    return Tb(t) if callable(Tb) else Tb


def _face_energy(geometry, side, Tb, dt, e1, e2):
    """a c T^4/4 * area * dt of a uniform face (MG_IMC2D.py:406-413, 465-470)."""
    if geometry == "xy":
        area = (e2[-1] - e2[0]) if side < 2 else (e1[-1] - e1[0])
    elif side < 2:
        area = 2.0 * np.pi * (e1[0] if side == 0 else e1[-1]) * (e2[-1] - e2[0])
    else:
        area = np.pi * (e1[-1] * e1[-1] - e1[0] * e1[0])
    return __a * __c * (Tb * Tb * Tb * Tb) / 4.0 * area * dt


def _face_profile(geometry, side, dt, e1, e2, bfunc):
    """(running sum of per-face-cell energy, face-cell temperatures) for
    boundary_source_func (MG_IMC2D.py:472-536); None if no face cell emits."""
    name = _SIDES[geometry][side]
    if side < 2:
        X = e1[0] if side == 0 else e1[-1]
        cen = 0.5 * (e2[:-1] + e2[1:])
        Tc = np.array([bfunc(X, c, name) for c in cen], dtype=np.float64)
        size = np.diff(e2) if geometry == "xy" else 2.0 * np.pi * X * np.diff(e2)
    else:
        Y = e2[0] if side == 2 else e2[-1]
        cen = 0.5 * (e1[:-1] + e1[1:])
        Tc = np.array([bfunc(c, Y, name) for c in cen], dtype=np.float64)
        size = np.diff(e1) if geometry == "xy" else np.pi * (e1[1:] * e1[1:] - e1[:-1] * e1[:-1])
    if not np.any(Tc > 0.0):
        return None
    em = np.where(Tc > 0.0, __a * __c * (Tc * Tc * Tc * Tc) / 4.0 * size * dt, 0.0)
    return np.cumsum(em), Tc


def _estimate_boundary_emission_energy(geometry, e1, e2, Tb4, dt, bfunc, profiles=None):
    """MG_IMC2D.py:1546-1632, including its xy left/right area of 2 (z1 - z0).
    profiles: the four _face_profile results when already evaluated."""
    E = 0.0
    for side in range(4):
        if bfunc is not None:
            prof = profiles[side] if profiles is not None else _face_profile(geometry, side, dt, e1, e2, bfunc)
            if prof is None:
                continue
            e = float(prof[0][-1])
        elif Tb4[side] > 0.0:
            e = _face_energy(geometry, side, Tb4[side], dt, e1, e2)
        else:
            continue
        if geometry == "xy" and side < 2:
            e *= 2.0
        E += e
    return E


def init_simulation(Ntarget, Tinit, Tr_init, edges1, edges2, energy_edges, eos, inv_eos, Ntarget_ic=None,
                    T_emit_floor=0.0, geometry="xy", precision="fp64", seed=1, device=0, engine=None):
    """Initialise material state and the multigroup equilibrium census
    (MG_IMC2D.py:1431-1543).  ``engine``: reuse an earlier state's device
    context (same mesh, groups, geometry and precision)."""
    nx, ny = len(edges1) - 1, len(edges2) - 1
    volumes = _cell_volumes(edges1, edges2, geometry)
    internal_energy = eos(Tinit)
    temperature = np.array(Tinit, dtype=np.float64, copy=True)
    assert np.allclose(inv_eos(internal_energy), Tinit), "Inverse EOS failed"
    if engine is not None:
        eng = engine
        if eng.shape != (nx, ny) or eng.geometry != geometry or eng.G != len(energy_edges) - 1:
            raise ValueError("init_simulation(): engine was built for another mesh, group set or geometry")
        eng.set_seed(seed)
    else:
        eng = EngineMG(edges1, edges2, energy_edges, geometry=geometry, precision=precision, seed=seed, device=device)
    N_ic = Ntarget if Ntarget_ic is None else Ntarget_ic
    eng.equilibrium_sample(int(N_ic), Tr_init, T_emit_floor)
    by_bin, total_rad = eng.census_by_bin()
    erg = by_bin / volumes[None, :, :]
    radiation_temperature = quarter_root(np.sum(erg, axis=0) / __a)
    total_internal = float(np.sum(internal_energy * volumes))
    previous_total = total_internal + total_rad
    if VERBOSE:
        print("Time", "N", "Total Energy", "Total Internal Energy", "Total Radiation Energy", "Boundary Emission",
              "Boundary Outgoing", "Source Emission", "Residual", sep="\t")
        print("=" * 158)
        print("{:.6e}".format(0.0), eng.n, "{:.6e}".format(previous_total), "{:.6e}".format(total_internal),
              "{:.6e}".format(total_rad), *["{:.6e}".format(0.0)] * 4, sep="\t")
    return SimulationState2DMG(eng, internal_energy, temperature, radiation_temperature, erg, 0.0, previous_total, 0)


def step(state, Ntarget, Nboundary, Nsource, Nmax, T_boundary, dt, edges1, edges2, energy_edges, sigma_a_funcs,
         inv_eos, cv, source, reflect=(False, False, False, False), theta=1.0, use_scalar_intensity_Tr=True,
         conserve_comb_energy=False, geometry="xy", max_events_per_particle=1_000_000, boundary_source_func=None,
         emission_fractions=None, Ntotal=0, Ntotal_T_floor=0.0, particle_budget_fmin=0.1, T_emit_floor=0.0,
         Nmax_growth=0, Nmax_final=None, diagnostics=False, _timing=False):
    """Advance one multigroup 2-D IMC step (MG_IMC2D.py:1635-2122).

    The per-cell totals over groups (deposition, emission, comb discrepancy)
    are summed on the device; ``diagnostics=True`` also returns the (G, nx, ny)
    arrays ``deposited_by_group`` and ``emission_by_group`` in ``info``."""
    eng = state._engine
    nx, ny = len(edges1) - 1, len(edges2) - 1
    G = len(energy_edges) - 1
    if (nx, ny) != eng.shape or geometry != eng.geometry or G != eng.G:
        raise ValueError("step(): mesh, group set or geometry differs from init_simulation")
    e1, e2 = eng.e1, eng.e2
    volumes = _cell_volumes(e1, e2, geometry)
    t_step_start = _time.perf_counter()
    state._invalidate()
    eng.set_step(state.count)

    temperature = state.temperature
    # the opacities go straight into a page-locked block (uploaded by one DMA
    # copy); a callable may return a scalar or an (nx, ny) array
    sigma = eng.group_block()
    for g in range(G):
        sigma[g] = sigma_a_funcs[g](temperature)
    cvv = np.asarray(cv(temperature), dtype=np.float64) * np.ones((nx, ny))
    want = ("Eplanck",) if Ntotal > 0 else ()
    if diagnostics:
        want = want + ("E",)
    mat, E_emit = eng.set_material(temperature, sigma, cvv, theta, dt, emission_fractions, T_emit_floor, want)
    emission_by_group = mat.get("E")

    Tb4 = [_boundary_temperature_value(T_boundary[s], state.time) for s in range(4)]
    profiles = None
    if boundary_source_func is not None:
        profiles = [_face_profile(geometry, side, dt, e1, e2, boundary_source_func) for side in range(4)]
    E_bc_est = float("nan")
    E_mat_est = float("nan")
    if Ntotal > 0:
        E_bc_est = _estimate_boundary_emission_energy(geometry, e1, e2, Tb4, dt, boundary_source_func, profiles)
        cell_em = mat["Eplanck"]
        if Ntotal_T_floor > 0.0:
            cell_em = np.where(temperature > Ntotal_T_floor, cell_em, 0.0)
        E_mat_est = float(np.sum(cell_em))
        E_total_est = E_bc_est + E_mat_est
        frac_bc = E_bc_est / E_total_est if E_total_est > 0.0 else 0.5
        fmin = float(np.clip(particle_budget_fmin, 0.0, 0.5))
        frac_bc = float(np.clip(frac_bc, fmin, 1.0 - fmin))
        Nboundary = int(round(Ntotal * frac_bc))
        Ntarget = Ntotal - Nboundary

    t_sampling_start = _time.perf_counter()
    boundary_emission = 0.0
    be_group = np.zeros(G)
    be_side = np.zeros(4)
    if Nboundary > 0:
        for side in range(4):
            prof = Tface = None
            if boundary_source_func is not None:
                pr = profiles[side]
                if pr is None:
                    continue
                prof, Tface = pr
                E = float(prof[-1])
            else:
                if not (Tb4[side] > 0.0):
                    continue
                E = _face_energy(geometry, side, float(Tb4[side]), dt, e1, e2)
            eg = eng.sample_boundary(side, int(Nboundary), E, dt, float(Tb4[side] or 0.0), prof, Tface,
                                     want_groups=True)
            boundary_emission += E
            be_side[side] += E
            be_group += eg

    source_emission = 0.0
    src = np.asarray(source, dtype=np.float64)
    if Nsource > 0 and np.max(src) > 0.0:
        if src.ndim == 3:
            Sg = src
        else:
            Sg = np.zeros((G, nx, ny))
            Sg[0] = src * np.ones((nx, ny))
        source_emission = eng.sample_source(int(Nsource), Sg * volumes[None, :, :] * dt, dt)

    if Ntarget > 0:
        eng.sample_emission(int(Ntarget), dt)
        emitted_cell = eng.group_sum("emission")
    else:
        emitted_cell = np.zeros((nx, ny))
        if diagnostics:
            emission_by_group = np.zeros((G, nx, ny))

    t_transport_start = _time.perf_counter()
    n_transported = eng.n
    if _TRANSPORT_HOOK is not None:
        _TRANSPORT_HOOK(eng)
    dep, si, bl, stats = eng.transport(dt, reflect, max_events_per_particle, -1.0, want_dep=diagnostics)
    t_post_start = _time.perf_counter()
    dep_cell = eng.group_sum("dep")
    internal_energy = state.internal_energy + dep_cell - emitted_cell / volumes
    temperature = inv_eos(internal_energy)

    ecen, _, post, ecen_total, total_rad = eng.comb(int(Nmax), want_ecen=not use_scalar_intensity_Tr,
                                                    want_post=True)
    # per-group arrays of the step written into page-locked blocks (no fresh
    # 8-byte x G x ncell allocations); post is not returned, so in place
    if use_scalar_intensity_Tr:
        erg_pre = np.divide(si, __c, out=eng.group_block())
    else:
        erg_pre = np.divide(ecen, volumes[None, :, :], out=ecen)
    if conserve_comb_energy:
        internal_energy = internal_energy + eng.group_sum("disc") / volumes
        temperature = inv_eos(internal_energy)
    erg_post = np.divide(post, volumes[None, :, :], out=post)
    radiation_temperature = quarter_root(np.sum(erg_pre, axis=0) / __a)

    total_internal = float(np.sum(internal_energy * volumes))
    total_energy = total_internal + total_rad
    boundary_loss = float(bl[0])
    dE_system = total_energy - state.previous_total_energy
    energy_residual = dE_system - boundary_emission + boundary_loss - source_emission
    disc_total = 0.0 if conserve_comb_energy else float(ecen_total - total_rad)

    state.internal_energy = internal_energy
    state.temperature = temperature
    state.radiation_temperature = radiation_temperature
    state.radiation_energy_by_group = erg_pre
    state.radiation_energy_by_group_postcomb = erg_post
    state.time += dt
    state.previous_total_energy = total_energy
    state.count += 1
    state._invalidate()

    t_end = _time.perf_counter()
    events_total = int(stats[0])
    info = {
        "time": state.time,
        "temperature": temperature,
        "radiation_temperature": radiation_temperature,
        "radiation_energy_by_group": erg_pre,
        "radiation_energy_by_group_postcomb": erg_post,
        "N_particles": eng.n,
        "N_boundary": Nboundary,
        "N_target": Ntarget,
        "particle_budget_fmin": float(np.clip(particle_budget_fmin, 0.0, 0.5)),
        "E_boundary_est": E_bc_est,
        "E_material_est": E_mat_est,
        "total_energy": total_energy,
        "total_internal_energy": total_internal,
        "total_radiation_energy": total_rad,
        "boundary_emission": boundary_emission,
        "boundary_emission_by_group": be_group,
        "boundary_emission_by_side": be_side,
        "boundary_loss": boundary_loss,
        "boundary_outgoing": boundary_loss,
        "boundary_outgoing_by_group": bl[1:1 + G].copy(),
        "boundary_outgoing_by_side": bl[1 + G:].copy(),
        "source_emission": source_emission,
        "incoming_energy_total": boundary_emission + source_emission,
        "outgoing_energy_total": boundary_loss,
        "dE_system": dE_system,
        "energy_expected_dE": boundary_emission + source_emission - boundary_loss,
        "energy_residual": energy_residual,
        "energy_loss": energy_residual,
        "comb_discrepancy": disc_total,
        "balance_residual": energy_residual + disc_total,
        "deposited": dep_cell,
        "scalar_intensity": si,
        "Nmax_next": ((min(Nmax + Nmax_growth, Nmax_final) if Nmax_final is not None else Nmax + Nmax_growth)
                      if Nmax_growth > 0 else Nmax),
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
                "avg_events_per_particle": events_total / max(n_transported, 1),
                "n_particles_transported": int(n_transported),
            },
        },
    }
    if diagnostics:
        info["deposited_by_group"] = dep
        info["emission_by_group"] = emission_by_group
    return state, info


def run_simulation(Ntarget, Nboundary, Nsource, Nmax, Tinit, Tr_init, T_boundary, dt, edges1, edges2, energy_edges,
                   sigma_a_funcs, eos, inv_eos, cv, source, final_time, reflect=(False, False, False, False),
                   output_freq=1, theta=1.0, use_scalar_intensity_Tr=True, Ntarget_ic=None, conserve_comb_energy=False,
                   geometry="xy", max_events_per_particle=1_000_000, emission_fractions=None, T_emit_floor=0.0,
                   Nmax_growth=0, Nmax_final=None, precision="fp64", seed=1, device=0):
    """Run a multigroup simulation (MG_IMC2D.py:2125-2254); returns (history, state)."""
    state = init_simulation(Ntarget, Tinit, Tr_init, edges1, edges2, energy_edges, eos, inv_eos,
                            Ntarget_ic=Ntarget_ic, geometry=geometry, precision=precision, seed=seed, device=device)
    history = []
    t = 0.0
    step_count = 0
    cumulative_residual = 0.0
    time_tol = max(1e-15, 1e-12 * max(final_time, 1.0))
    while t < final_time - time_tol:
        dt_step = min(dt, final_time - t)
        if dt_step <= time_tol:
            break
        state, info = step(state, Ntarget, Nboundary, Nsource, Nmax, T_boundary, dt_step, edges1, edges2,
                           energy_edges, sigma_a_funcs, inv_eos, cv, source, reflect=reflect, theta=theta,
                           use_scalar_intensity_Tr=use_scalar_intensity_Tr, conserve_comb_energy=conserve_comb_energy,
                           geometry=geometry, max_events_per_particle=max_events_per_particle,
                           emission_fractions=emission_fractions, T_emit_floor=T_emit_floor,
                           Nmax_growth=Nmax_growth, Nmax_final=Nmax_final)
        t = state.time
        Nmax = info["Nmax_next"]
        step_count += 1
        cumulative_residual += info["energy_residual"]
        if step_count % output_freq == 0 or (final_time - t) < time_tol:
            info["cumulative_energy_residual"] = cumulative_residual
            info["net_boundary_energy"] = info["boundary_emission"] - info["boundary_outgoing"]
            history.append(info)
            if VERBOSE:
                print("{:.6e}".format(t), info["N_particles"], "{:.6e}".format(info["total_energy"]),
                      "{:.6e}".format(info["total_internal_energy"]), "{:.6e}".format(info["total_radiation_energy"]),
                      "{:.6e}".format(info["boundary_emission"]), "{:.6e}".format(info["boundary_outgoing"]),
                      "{:.6e}".format(info["source_emission"]), "{:.6e}".format(info["energy_residual"]), sep="\t")
    return history, state
