"""IMC1D -- gray 1-D implicit Monte Carlo (slab and sphere) on one MI355X.

Drop-in for the reference module IMC/IMC1D.py: the same ``init_simulation``,
``step`` and ``run_simulation`` signatures (IMC/IMC1D.py:859, 908, 1127), the
same ``SimulationState`` attributes and the same ``info`` keys.  Sourcing,
transport and the comb run as HIP kernels on particles that stay in device
memory; host code evaluates the user callables on per-cell arrays.

``BatchIMC1D`` runs many independent slab or sphere problems of uneven cell
and particle counts as one batch: every stage (sourcing, the persistent
tracking kernel, the comb) is one launch over all problems.

Extra keyword arguments (all optional): ``precision`` ('fp64' reproduces the
reference arithmetic; 'fp32' tracks geometry in FP32 with FP64 weights and
tallies), ``seed`` (Philox key) and ``device``.
"""
from __future__ import annotations

import time as _time

import numpy as np

from ._census import DeviceCensus

from ._lib import quarter_root, A_RAD, C_LIGHT, SRC_EMISSION, SRC_FIXED, Engine1D, ImcError

__c = C_LIGHT
__a = A_RAD

VERBOSE = True


def _cell_volumes(mesh, geometry):
    """dx for slab, (4/3) pi (r1^3 - r0^3) for spherical (IMC/IMC1D.py:851-856)."""
    mesh = np.asarray(mesh, dtype=np.float64)
    if geometry == "slab":
        return mesh[:, 1] - mesh[:, 0]
    r0, r1 = mesh[:, 0], mesh[:, 1]
    return (4.0 / 3.0) * np.pi * (r1 * r1 * r1 - r0 * r0 * r0)


def _edges(mesh):
    mesh = np.asarray(mesh, dtype=np.float64)
    return np.concatenate((mesh[:, 0], mesh[-1:, 1]))


def _boundary_energy(geometry, Tb, dt, R):
    """a c T^4 / 4 dt, times 4 pi R^2 on a sphere (IMC/IMC1D.py:319-329, 744-763)."""
    e = __a * __c * (Tb * Tb * Tb * Tb) / 4.0 * dt
    if geometry != "slab":
        e = e * 4.0 * np.pi * R * R
    return e


class SimulationState(DeviceCensus):
    """Mutable 1-D IMC state (IMC/IMC1D.py:836-848).

    The particle arrays live on the GPU; reading ``weights``, ``mus``, ...
    downloads the census (cached until the next step), assigning uploads it.
    """

    _PARTICLE_FIELDS = ("weights", "mus", "times", "positions", "cell_indices")

    def __init__(self, engine, internal_energy, temperature, radiation_temperature, time, previous_total_energy,
                 count=0):
        DeviceCensus.__init__(self, engine)
        self.internal_energy = internal_energy
        self.temperature = temperature
        self.radiation_temperature = radiation_temperature
        self.time = time
        self.previous_total_energy = previous_total_energy
        self.count = count

    def _download(self):
        return tuple(self._engine.get_census()[:5])

    def _upload(self, arrs):
        self._engine.put_census([len(arrs[0])], *arrs)

    def _device_count(self):
        return self._engine.census_counts()[1]


def init_simulation(Ntarget, Tinit, Tr_init, mesh, eos, inv_eos, Ntarget_ic=None, geometry="slab", T_emit_floor=0.0,
                    precision="fp64", seed=1, device=0, engine=None):
    """Initialise material state and the equilibrium census (IMC/IMC1D.py:859-905).

    ``engine``: the device context of an earlier state on the same mesh,
    geometry and precision (``state._engine``) to reuse with its buffers."""
    if geometry not in ("slab", "spherical"):
        raise ValueError(f"Unknown geometry: {geometry}")
    volumes = _cell_volumes(mesh, geometry)
    internal_energy = eos(Tinit)
    temperature = np.array(Tinit, dtype=np.float64, copy=True)
    assert np.allclose(inv_eos(internal_energy), Tinit), "Inverse EOS failed"
    if engine is not None:
        eng = engine
        if (eng.P != 1 or eng.ncell != volumes.size or eng.geometry != geometry
                or eng.precision != {"fp64": 0, "fp32": 1}[precision]):
            raise ValueError("init_simulation(): engine was built for another mesh, geometry or precision")
        eng.set_seed(seed)
    else:
        eng = Engine1D([_edges(mesh)], geometry=geometry, precision=precision, seed=seed, device=device)
    N_ic = Ntarget if Ntarget_ic is None else Ntarget_ic
    eng.equilibrium_sample(int(N_ic), Tr_init, T_emit_floor)
    rad_cell, rad_tot = eng.census_energy()
    radiation_temperature = quarter_root(rad_cell / volumes / __a)
    total_internal = float(np.sum(internal_energy * volumes))
    total_rad = float(rad_tot[0])
    previous_total = total_internal + total_rad
    if VERBOSE:
        print("Time", "N", "Total Energy", "Total Internal Energy", "Total Radiation Energy", "Boundary Emission",
              "Lost Energy", sep="\t")
        print("=" * 111)
        print("{:.6f}".format(0.0), eng.census_counts()[1], "{:.6f}".format(previous_total), "{:.6f}".format(total_internal),
              "{:.6f}".format(total_rad), "{:.6f}".format(0.0), "{:.6f}".format(0.0), sep="\t")
    return SimulationState(eng, internal_energy, temperature, radiation_temperature, 0.0, previous_total, 0)


def _set_material(eng, T, sig, cvv, theta, dt):
    try:
        eng.set_material(T, sig, cvv, theta, dt)
    except ImcError as e:
        if "Fleck factor out of bounds" in str(e):
            raise AssertionError("Fleck factor out of bounds") from e
        raise


def step(state, Ntarget, Nboundary, Nsource, NMax, T_boundary, dt, mesh, sigma_a_func, inv_eos, cv, source,
         reflect=(False, False), theta=1.0, use_scalar_intensity_Tr=True, conserve_comb_energy=False,
         T_emit_floor=0.0, geometry="slab"):
    """Advance one 1-D IMC step (IMC/IMC1D.py:908-1124)."""
    eng = state._engine
    mesh = np.asarray(mesh, dtype=np.float64)
    I = mesh.shape[0]
    if I != eng.ncell or geometry != eng.geometry:
        raise ValueError("step(): mesh or geometry differs from init_simulation")
    volumes = _cell_volumes(mesh, geometry)
    t_step_start = _time.perf_counter()
    state._invalidate()
    eng.set_step(state.count)

    temperature = state.temperature
    sig = np.asarray(sigma_a_func(temperature), dtype=np.float64) * np.ones(I)
    cvv = np.asarray(cv(temperature), dtype=np.float64) * np.ones(I)
    _set_material(eng, temperature, sig, cvv, theta, dt)

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
        QV = src * dt * volumes
        _, tot = eng.sample_volume(SRC_FIXED, int(Nsource), QV)
        source_emission = float(tot[0])

    E_emit_cells, _ = eng.sample_volume(SRC_EMISSION, int(Ntarget), None, T_emit_floor, want_energy=True)

    t_transport_start = _time.perf_counter()
    dep, si, leak, fleak, stats = eng.transport(reflect, -1.0)
    n_transported = int(stats[7])
    if stats[6] > 0:
        raise AssertionError("non-positive distance in move_particle")
    t_post_start = _time.perf_counter()

    internal_energy = state.internal_energy + dep - E_emit_cells / volumes
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
        "floor_kill_leak": float(fleak[0]),
        "energy_residual": energy_loss + disc_total + float(fleak[0]),
        "deposited": dep,
        "scalar_intensity": si,
        "emitted_energies": E_emit_cells,
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
                "scatter_events": int(stats[2]),
                "census_events": int(stats[3]),
                "weight_floor_kills": int(stats[4]),
                "reflections": int(stats[5]),
                "avg_events_per_particle": events_total / n_tr,
                "n_particles_transported": int(n_transported),
            },
        },
    }
    return state, info


def run_simulation(Ntarget, Nboundary, Nsource, NMax, Tinit, Tr_init, T_boundary, dt, mesh, sigma_a_func, eos,
                   inv_eos, cv, source, final_time, reflect=(False, False), output_freq=1, theta=1.0,
                   use_scalar_intensity_Tr=True, Ntarget_ic=None, conserve_comb_energy=False, geometry="slab",
                   precision="fp64", seed=1, device=0):
    """Run a full 1-D simulation (IMC/IMC1D.py:1127-1170).  Returns
    (time_values, radiation_temperatures, temperatures)."""
    state = init_simulation(Ntarget, Tinit, Tr_init, mesh, eos, inv_eos, Ntarget_ic=Ntarget_ic, geometry=geometry,
                            precision=precision, seed=seed, device=device)
    Tr_hist = [state.radiation_temperature.copy()]
    T_hist = [state.temperature.copy()]
    times = [0.0]
    while state.time < final_time:
        step_dt = min(dt, final_time - state.time)
        state, info = step(state, Ntarget, Nboundary, Nsource, NMax, T_boundary, step_dt, mesh, sigma_a_func,
                           inv_eos, cv, source, reflect, theta=theta, use_scalar_intensity_Tr=use_scalar_intensity_Tr,
                           conserve_comb_energy=conserve_comb_energy, geometry=geometry)
        if (state.count - 1) % output_freq == 0 or (info["time"] - final_time) < step_dt:
            Tr_hist.append(state.radiation_temperature.copy())
            T_hist.append(state.temperature.copy())
            times.append(info["time"])
            if VERBOSE:
                print("{:.6f}".format(info["time"]), info["N_particles"], "{:.6f}".format(info["total_energy"]),
                      "{:.6f}".format(info["total_internal_energy"]), "{:.6f}".format(info["total_radiation_energy"]),
                      "{:.6f}".format(info["boundary_emission"]), "{:.6e}".format(info["energy_loss"]), sep="\t")
    return np.array(times), np.array(Tr_hist), np.array(T_hist)


# ---------------------------------------------------------------------------
# many independent problems in one batch
# ---------------------------------------------------------------------------
class BatchIMC1D:
    """P independent gray 1-D problems (same geometry) advanced together.

    Per-problem quantities (``Ntarget``, ``Nboundary``, ``NMax``, ``dt``,
    ``theta``, boundary temperatures) are arrays of P values or scalars;
    ``sigma_a_func``, ``cv``, ``eos`` and ``inv_eos`` are one callable or P
    callables, each applied to its problem's cell array.  Problem p draws the
    Philox streams a single-problem run with the same seed would draw if its
    index base were p << 32, so each problem can be checked on its own.
    """

    def __init__(self, meshes, Tinit, Tr_init, eos, inv_eos, Ntarget_ic, geometry="slab", T_emit_floor=0.0,
                 precision="fp32", seed=1, device=0):
        self.meshes = [np.asarray(m, dtype=np.float64) for m in meshes]
        self.P = len(self.meshes)
        self.geometry = geometry
        self.eng = Engine1D([_edges(m) for m in self.meshes], geometry=geometry, precision=precision, seed=seed,
                            device=device)
        self.volumes = [_cell_volumes(m, geometry) for m in self.meshes]
        self._vcat = np.concatenate(self.volumes)
        self.reset(Tinit, Tr_init, eos, inv_eos, Ntarget_ic, T_emit_floor, seed)

    def reset(self, Tinit, Tr_init, eos, inv_eos, Ntarget_ic, T_emit_floor=0.0, seed=None):
        """Start again from t = 0 (equilibrium census) on the same device
        buffers, optionally with another Philox seed."""
        if seed is not None:
            self.eng.set_seed(seed)
        self.T = [np.array(t, dtype=np.float64, copy=True) for t in Tinit]
        self.u = [self._fn(eos, p)(self.T[p]) for p in range(self.P)]
        self.eng.equilibrium_sample(self._per(Ntarget_ic, np.int64), np.concatenate(Tr_init),
                                    self._per(T_emit_floor))
        self.Tr = [np.asarray(t, dtype=np.float64).copy() for t in Tr_init]
        self.time = np.zeros(self.P)
        self.count = 0
        self._inv = inv_eos
        _, rad = self.eng.census_energy()
        self.previous_total = np.array([float(np.sum(self.u[p] * self.volumes[p])) for p in range(self.P)]) + rad

    def _per(self, v, dtype=np.float64):
        return np.ascontiguousarray(np.broadcast_to(np.asarray(v, dtype=dtype), (self.P,)), dtype=dtype)

    @staticmethod
    def _fn(f, p):
        return f[p] if isinstance(f, (list, tuple)) else f

    def step(self, Ntarget, Nboundary, NMax, T_boundary, dt, sigma_a_func, cv, reflect=(False, False), theta=1.0,
             T_emit_floor=0.0, tune=False):
        """One step of every problem.  T_boundary: (P, 2) array or pair.
        ``tune``: True times the transport launch shapes on this step first
        and keeps the fastest (results do not depend on the shape); a
        callable is called with the engine before the transport instead.
        Returns a dict of per-problem energy diagnostics and device timings."""
        eng = self.eng
        P = self.P
        eng.set_step(self.count)
        dtv = self._per(dt)
        # the callables are per problem; their values go straight into the
        # concatenated per-cell arrays (a scalar fills its problem's cells)
        coff = eng.coff
        sig = np.empty(int(coff[-1]))
        cvv = np.empty(int(coff[-1]))
        for p in range(P):
            sig[coff[p]:coff[p + 1]] = self._fn(sigma_a_func, p)(self.T[p])
            cvv[coff[p]:coff[p + 1]] = self._fn(cv, p)(self.T[p])
        _set_material(eng, np.concatenate(self.T), sig, cvv, self._per(theta), dtv)
        Tb = np.broadcast_to(np.asarray(T_boundary, dtype=np.float64).reshape(-1, 2), (P, 2))
        nb = self._per(Nboundary, np.int64)
        be = np.zeros(P)
        for side in (0, 1):
            E = np.array([_boundary_energy(self.geometry, float(Tb[p, side]), float(dtv[p]),
                                           self.meshes[p][0, 0] if side == 0 else self.meshes[p][-1, 1])
                          if Tb[p, side] > 0 else 0.0 for p in range(P)])
            n = np.where((Tb[:, side] > 0) & (nb > 0), nb, 0)
            if n.any():
                eng.sample_boundary(side, n, E)
                be += np.where(n > 0, E, 0.0)
        Ec, _ = eng.sample_volume(SRC_EMISSION, self._per(Ntarget, np.int64), None, self._per(T_emit_floor),
                                  want_energy=True)
        tuned = None
        if callable(tune):
            tune(eng)
        elif tune:
            from .autotune import tune_engine1d

            tuned = tune_engine1d(eng, reflect)
        dep, si, leak, fleak, stats = eng.transport(reflect, -1.0)
        ms = eng.last_transport_ms()
        if stats[6] > 0:
            raise AssertionError("non-positive distance in move_particle")
        _, _, ecen_total, total_rad = eng.comb(self._per(NMax, np.int64))
        # material update over all cells at once (IMC/IMC1D.py:1021-1035); only
        # inv_eos is called per problem.  Fresh arrays every step: the lists
        # handed out earlier keep their values.
        u = np.concatenate(self.u)
        u += dep
        u -= Ec / self._vcat
        self.u = eng.split(u)
        self.T = [np.asarray(self._fn(self._inv, p)(self.u[p]), dtype=np.float64) for p in range(P)]
        self.Tr = eng.split(quarter_root(si / A_RAD / C_LIGHT))
        total = np.add.reduceat(u * self._vcat, coff[:-1]) + total_rad
        loss = total - self.previous_total - be + leak
        disc = ecen_total - total_rad
        self.previous_total = total
        self.time += dtv
        self.count += 1
        return {"total_energy": total, "boundary_emission": be, "boundary_loss": leak, "energy_loss": loss,
                "comb_discrepancy": disc, "floor_kill_leak": fleak, "energy_residual": loss + disc + fleak,
                "stats": stats, "transport_kernel_ms": ms, "N_particles": eng.census_counts()[0], "tuned": tuned}
