"""IMC2D -- gray 2-D implicit Monte Carlo (xy and rz) on one MI355X.

Drop-in for the reference module IMC/IMC2D.py: the same ``init_simulation``,
``step`` and ``run_simulation`` signatures, the same ``SimulationState2D``
attributes and the same ``info`` keys.  Sourcing, transport and the comb run
as HIP kernels on particles that stay in device memory; host code here only
evaluates the user callables (``sigma_a_func``, ``cv``, ``inv_eos``,
``boundary_source_func``) on per-cell arrays and does per-cell bookkeeping.

Extra keyword arguments (all optional): ``precision`` ('fp64' reproduces the
reference arithmetic; 'fp32' tracks geometry in FP32 with FP64 weights and
tallies), ``seed`` (Philox key) and ``device``.
"""
from __future__ import annotations

import time as _time

import numpy as np

from ._census import DeviceCensus

from ._lib import quarter_root, A_RAD, C_LIGHT, SRC_EMISSION, SRC_FIXED, Engine2D
from .materials import device_material

__c = C_LIGHT
__a = A_RAD

VERBOSE = True

# Called with the engine right before each staged transport (the launch-shape
# autotuner times its candidates there); while set, steps take the staged
# path.
_TRANSPORT_HOOK = None

_SIDES = {"xy": ("left", "right", "bottom", "top"), "rz": ("rmin", "rmax", "zmin", "zmax")}


class SimulationState2D(DeviceCensus):
    """Mutable 2-D IMC state (IMC/IMC2D.py:65-84).

    The particle arrays live on the GPU; reading ``weights``, ``dir1``, ...
    downloads the census (cached until the next step), assigning uploads it.
    """

    _PARTICLE_FIELDS = ("weights", "dir1", "dir2", "times", "pos1", "pos2", "cell_i", "cell_j")

    def __init__(self, engine, internal_energy, temperature, radiation_temperature, time, previous_total_energy,
                 count=0):
        DeviceCensus.__init__(self, engine)
        self.internal_energy = internal_energy
        self.temperature = temperature
        self.radiation_temperature = radiation_temperature
        self.time = time
        self.previous_total_energy = previous_total_energy
        self.count = count


def _shape_from_edges(edges1, edges2):
    return len(edges1) - 1, len(edges2) - 1


def _cell_volumes(edges1, edges2, geometry):
    e1 = np.asarray(edges1, dtype=np.float64)
    e2 = np.asarray(edges2, dtype=np.float64)
    if geometry == "xy":
        return np.diff(e1)[:, None] * np.diff(e2)[None, :]
    if geometry == "rz":
        return np.pi * (e1[1:] * e1[1:] - e1[:-1] * e1[:-1])[:, None] * np.diff(e2)[None, :]
    raise ValueError(f"Unknown geometry: {geometry}")


def _face_energy(geometry, side, Tb, dt, e1, e2):
    """a c T^4/4 * area * dt (IMC/IMC2D.py:283-284, 357-361)."""
    if geometry == "xy":
        area = (e2[-1] - e2[0]) if side < 2 else (e1[-1] - e1[0])
    elif side < 2:
        R = e1[0] if side == 0 else e1[-1]
        area = 2.0 * np.pi * R * (e2[-1] - e2[0])
    else:
        area = np.pi * (e1[-1] * e1[-1] - e1[0] * e1[0])
    return __a * __c * (Tb * Tb * Tb * Tb) / 4.0 * area * dt


def _face_profile(geometry, side, e1, e2, dt, bfunc):
    """Running sum of per-face-cell emission for boundary_source_func
    (IMC/IMC2D.py:182-281, 340-435); None when no face cell emits."""
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
        size = np.diff(e1) if geometry == "xy" else np.pi * (e1[1:] ** 2 - e1[:-1] ** 2)
    if not np.any(Tc > 0.0):
        return None
    em = np.where(Tc > 0.0, __a * __c * Tc ** 4 / 4.0 * size * dt, 0.0)
    return np.cumsum(em)


def _device_material(eng, sigma_a_func, cv, inv_eos):
    """device_material() of the step's callables, checked once per set of
    callable objects (the check scans every cell of sigma0 and c_v)."""
    key = (sigma_a_func, getattr(cv, "__self__", cv), getattr(cv, "__func__", None),
           getattr(inv_eos, "__self__", inv_eos), getattr(inv_eos, "__func__", None))
    hit = getattr(eng, "_dm_check", None)
    if hit is not None and all(a is b for a, b in zip(hit[0], key)):
        return hit[1]
    dm = device_material(sigma_a_func, cv, inv_eos, eng.shape)
    eng._dm_check = (key, dm)
    return dm


def init_simulation(Ntarget, Tinit, Tr_init, edges1, edges2, eos, inv_eos, Ntarget_ic=None, geometry="xy",
                    precision="fp64", seed=1, device=0, engine=None):
    """Initialise material state and the equilibrium census (IMC/IMC2D.py:1514-1588).

    ``engine``: the device context of an earlier state on the same mesh,
    geometry and precision (``state._engine``) to reuse with its buffers."""
    nx, ny = _shape_from_edges(edges1, edges2)
    volumes = _cell_volumes(edges1, edges2, geometry)
    internal_energy = eos(Tinit)
    temperature = np.array(Tinit, dtype=np.float64, copy=True)
    assert np.allclose(inv_eos(internal_energy), Tinit), "Inverse EOS failed"
    if engine is not None:
        eng = engine
        if eng.shape != (nx, ny) or eng.geometry != geometry or eng.precision != {"fp64": 0, "fp32": 1}[precision]:
            raise ValueError("init_simulation(): engine was built for another mesh, geometry or precision")
        eng.set_seed(seed)
    else:
        eng = Engine2D(edges1, edges2, geometry=geometry, precision=precision, seed=seed, device=device)
    N_ic = Ntarget if Ntarget_ic is None else Ntarget_ic
    eng.equilibrium_sample(N_ic, Tr_init)
    rad_cell = eng.census_by_cell()
    radiation_temperature = quarter_root(rad_cell / volumes / __a)
    total_internal = float(np.sum(internal_energy * volumes))
    total_rad = eng.census_energy()
    previous_total = total_internal + total_rad
    if VERBOSE:
        print("Time", "N", "Total Energy", "Total Internal Energy", "Total Radiation Energy", "Boundary Emission",
              "Lost Energy", sep="\t")
        print("=" * 111)
        print("{:.6f}".format(0.0), eng.n, "{:.6f}".format(previous_total), "{:.6f}".format(total_internal),
              "{:.6f}".format(total_rad), "{:.6f}".format(0.0), "{:.6f}".format(0.0), sep="\t")
    return SimulationState2D(eng, internal_energy, temperature, radiation_temperature, 0.0, previous_total, 0)


def step(state, Ntarget, Nboundary, Nsource, Nmax, T_boundary, dt, edges1, edges2, sigma_a_func, inv_eos, cv, source,
         reflect=(False, False, False, False), theta=1.0, use_scalar_intensity_Tr=True, conserve_comb_energy=False,
         geometry="xy", rz_linear_source=True, max_events_per_particle=1_000_000, boundary_source_func=None,
         _timing=False):
    """Advance one 2-D IMC step (IMC/IMC2D.py:1591-1897).

    T_boundary order: (left/rmin, right/rmax, bottom/zmin, top/zmax)."""
    eng = state._engine
    nx, ny = _shape_from_edges(edges1, edges2)
    if (nx, ny) != eng.shape or geometry != eng.geometry:
        raise ValueError("step(): mesh or geometry differs from init_simulation")
    e1 = eng.e1
    e2 = eng.e2
    t_step_start = _time.perf_counter()
    volumes = getattr(eng, "_volumes", None)
    if volumes is None:
        volumes = eng._volumes = _cell_volumes(e1, e2, geometry)
    state._invalidate()
    eng.set_step(state.count)

    # a power-law opacity with a linear EOS runs on the device-held T and u
    dm = _device_material(eng, sigma_a_func, cv, inv_eos) if (use_scalar_intensity_Tr and
                                                             not conserve_comb_energy) else None
    if dm is not None:
        refs = getattr(eng, "_device_material", None)
        cur = (sigma_a_func, cv.__self__, state.temperature, state.internal_energy)
        if refs is None or any(a is not b for a, b in zip(refs, cur)):
            eng.set_material_powerlaw(dm[0], dm[1], dm[2], dm[3], state.temperature, state.internal_energy)
        src = np.asarray(source, dtype=np.float64)
        if boundary_source_func is None and not (Nsource > 0 and np.max(src) > 0.0) and _TRANSPORT_HOOK is None:
            # the whole step as one device-resident sequence (a replayed graph)
            return _step_device(state, eng, Ntarget, Nboundary, Nmax, T_boundary, dt, sigma_a_func, cv, reflect,
                                theta, geometry, rz_linear_source, max_events_per_particle, t_step_start)
        eng.fleck_powerlaw(theta, dt, clip=True)
    else:
        eng._device_material = None
        temperature = state.temperature
        sig = sigma_a_func(temperature)
        cvv = cv(temperature)
        eng.set_material(temperature, sig, cvv, theta, dt, clip=True)

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
        QV = src * dt * volumes
        _, tot, cnt = eng.sample_volume(SRC_FIXED, Nsource, QV, dt, rz_linear_source)
        if cnt > 0:
            source_emission = tot

    E_emit_cells, E_emit, _ = eng.sample_volume(SRC_EMISSION, Ntarget, None, dt, rz_linear_source,
                                                want_energy=True)

    if _TRANSPORT_HOOK is not None:
        _TRANSPORT_HOOK(eng)
    t_transport_start = _time.perf_counter()
    n_transported = eng.n
    dep_cell, si_cell, boundary_loss, stats = eng.transport(dt, reflect, max_events_per_particle, -1.0)
    t_post_start = _time.perf_counter()

    if dm is not None:
        # u += dep - E_emit / V, T = u / c_v, T_r from si, sum(u V) on the device
        temperature, radiation_temperature, internal_energy, total_internal = eng.update_material()
        eng._device_material = (sigma_a_func, cv.__self__, temperature, internal_energy)
        ecen, comb_disc, ecen_total, total_rad = eng.comb(Nmax, want_ecen=False, want_disc=False)
    else:
        internal_energy = state.internal_energy + dep_cell - E_emit_cells / volumes
        temperature = inv_eos(internal_energy)

        ecen, comb_disc, ecen_total, total_rad = eng.comb(Nmax, want_ecen=not use_scalar_intensity_Tr,
                                                          want_disc=conserve_comb_energy)
        if use_scalar_intensity_Tr:
            radiation_temperature = quarter_root(si_cell * (1.0 / (__a * __c)))
        else:
            radiation_temperature = quarter_root(ecen / volumes / __a)
        if conserve_comb_energy:
            internal_energy = internal_energy + comb_disc / volumes
            temperature = inv_eos(internal_energy)

        total_internal = float(np.sum(internal_energy * volumes))
    total_energy = total_internal + total_rad
    energy_loss = total_energy - state.previous_total_energy - boundary_emission + boundary_loss - source_emission
    # comb discrepancy booked as a sink unless it went back into the material
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
        "emitted_energies": E_emit_cells,
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


def _step_device(state, eng, Ntarget, Nboundary, Nmax, T_boundary, dt, sigma_a_func, cv, reflect, theta, geometry,
                 rz_linear_source, max_events_per_particle, t_step_start):
    """step() for the device-held power-law material with uniform face
    temperatures and no fixed source: ``imc2d_step`` runs the Fleck block,
    sourcing, transport, material update and comb on the device in one call
    (the same kernels and random streams as the staged path)."""
    e1, e2 = eng.e1, eng.e2
    face_E = np.zeros(4)
    if Nboundary > 0:
        for side in range(4):
            Tb = T_boundary[side]
            Tb = Tb(state.time) if callable(Tb) else Tb
            if Tb > 0.0:
                face_E[side] = _face_energy(geometry, side, float(Tb), dt, e1, e2)
    boundary_emission = float(np.sum(face_E))
    t_transport_start = _time.perf_counter()
    r = eng.step_device(theta, dt, Ntarget, Nboundary, face_E, Nmax, reflect, max_events_per_particle,
                        rz_linear=rz_linear_source)
    t_post_start = _time.perf_counter()
    temperature, radiation_temperature, internal_energy = r["T"], r["Tr"], r["u"]
    eng._device_material = (sigma_a_func, cv.__self__, temperature, internal_energy)
    total_internal = r["internal_total"]
    total_rad = r["rad_total"]
    boundary_loss = r["boundary_loss"]
    total_energy = total_internal + total_rad
    energy_loss = total_energy - state.previous_total_energy - boundary_emission + boundary_loss
    disc_total = r["ecen_total"] - total_rad

    state.internal_energy = internal_energy
    state.temperature = temperature
    state.radiation_temperature = radiation_temperature
    state.time += dt
    state.previous_total_energy = total_energy
    state.count += 1
    state._invalidate()

    t_end = _time.perf_counter()
    stats = r["stats"]
    events_total = int(stats[0])
    n_tr = max(int(r["n_transported"]), 1)
    info = {
        "time": state.time,
        "temperature": temperature,
        "radiation_temperature": radiation_temperature,
        "N_particles": r["n"],
        "total_energy": total_energy,
        "total_internal_energy": total_internal,
        "total_radiation_energy": total_rad,
        "boundary_emission": boundary_emission,
        "boundary_loss": boundary_loss,
        "source_emission": 0.0,
        "energy_loss": energy_loss,
        "comb_discrepancy": disc_total,
        "energy_residual": energy_loss + disc_total,
        "deposited": r["dep"],
        "scalar_intensity": r["si"],
        "emitted_energies": r["E_emit"],
        "profiling": {
            "phase_times_s": {
                "sampling": t_transport_start - t_step_start,
                "transport": t_post_start - t_transport_start,
                "postprocess": t_end - t_post_start,
                "total": t_end - t_step_start,
            },
            "transport_kernel_ms": r["transport_ms"],
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
                   geometry="xy", rz_linear_source=True, max_events_per_particle=1_000_000, precision="fp64", seed=1,
                   device=0):
    """Run a full 2-D simulation (IMC/IMC2D.py:1900-1984)."""
    state = init_simulation(Ntarget, Tinit, Tr_init, edges1, edges2, eos, inv_eos, Ntarget_ic=Ntarget_ic,
                            geometry=geometry, precision=precision, seed=seed, device=device)
    times = [0.0]
    Tr_hist = [state.radiation_temperature.copy()]
    T_hist = [state.temperature.copy()]
    while state.time < final_time - 1e-15:
        step_dt = min(dt, final_time - state.time)
        state, info = step(state, Ntarget, Nboundary, Nsource, Nmax, T_boundary, step_dt, edges1, edges2,
                           sigma_a_func, inv_eos, cv, source, reflect=reflect, theta=theta,
                           use_scalar_intensity_Tr=use_scalar_intensity_Tr, conserve_comb_energy=conserve_comb_energy,
                           geometry=geometry, rz_linear_source=rz_linear_source,
                           max_events_per_particle=max_events_per_particle)
        if (state.count - 1) % output_freq == 0 or (final_time - state.time) < 1e-12:
            times.append(info["time"])
            Tr_hist.append(state.radiation_temperature.copy())
            T_hist.append(state.temperature.copy())
            if VERBOSE:
                print("{:.6f}".format(info["time"]), info["N_particles"], "{:.6f}".format(info["total_energy"]),
                      "{:.6f}".format(info["total_internal_energy"]), "{:.6f}".format(info["total_radiation_energy"]),
                      "{:.6f}".format(info["boundary_emission"]), "{:.6e}".format(info["energy_loss"]), sep="\t")
    return np.array(times), np.array(Tr_hist), np.array(T_hist)
