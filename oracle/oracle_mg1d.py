"""oracle_mg1d.py -- TEST INFRASTRUCTURE ONLY (the checker and the timed CPU baseline).

ctypes bindings to the multigroup 1-D spherical part of
``oracle/_build/liboracle.so`` (``imc_oracle_mg1d.c``) and a host composition
of one step that follows the reference's ``MG_IMC/MG_IMC1D.py:step``
(917-1309) line by line: Planck-mean Fleck factor, the optional
energy-proportional particle budget, sphere face sources, fixed sources,
material emission by (group, cell) bin, transport, material update, the
(cell, group) comb with its three energy modes and threshold mode, and the
post-comb radiation diagnostics.

Per-(group, cell) arrays are (G, I); flat bin b = g*I + i.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg
may import this module.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import oracle as O
from . import oracle_mg as OM

A_RAD = O.A_RAD
C_LIGHT = O.C_LIGHT
_dp, _i64p, _i32p, _f64 = O._dp, O._i64p, O._i32p, O._f64
TAG_EMISSION, TAG_SOURCE = O.TAG_EMISSION, O.TAG_SOURCE


def _p(a, ct=_dp):
    return None if a is None else a.ctypes.data_as(ct)


def _lib():
    L = OM._lib()
    if not hasattr(L, "_m1_ready"):
        L.m1_volumes.argtypes = [C.c_int, _dp, _dp]
        L.m1_counts.argtypes = [C.c_int64, _dp, C.c_int64, C.c_int, _i64p]
        L.m1_counts.restype = C.c_int64
        L.m1_equilibrium_energy.argtypes = [C.c_int, _dp, _dp, C.c_double, _dp]
        L.m1_sample_ic.argtypes = [C.c_int, _dp, _dp, _dp, _i64p, C.c_int, _dp, _dp, C.c_uint64] + [_dp] * 4 + \
            [_i32p] * 2
        L.m1_sample_bins.argtypes = [C.c_int, C.c_int, _dp, _dp, _i64p, C.c_double, C.c_uint64, C.c_uint32,
                                     C.c_int32] + [_dp] * 4 + [_i32p] * 2
        L.m1_face_energy.argtypes = [C.c_double, C.c_double, C.c_double]
        L.m1_face_energy.restype = C.c_double
        L.m1_sample_face.argtypes = [C.c_int, C.c_int64, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int,
                                     C.c_int, _dp, _dp, C.c_uint64, C.c_uint32] + [_dp] * 4 + [_i32p] * 2
        L.m1_transport.argtypes = [C.c_int64, _dp, _dp, _dp, _dp, _i32p, _i32p, C.c_int, _dp, C.c_int, _dp, _dp,
                                   C.c_double, _i32p, C.c_double, C.c_uint64, C.c_uint32, _dp, _dp, _dp, _i64p,
                                   C.c_int32]
        L._m1_ready = True
    return L


@dataclass
class Pool1MG:
    w: np.ndarray
    mu: np.ndarray
    t: np.ndarray
    x: np.ndarray
    cell: np.ndarray
    g: np.ndarray

    @staticmethod
    def empty(n=0):
        z = lambda: np.zeros(n)
        return Pool1MG(z(), z(), z(), z(), np.zeros(n, np.int32), np.zeros(n, np.int32))

    def __len__(self):
        return self.w.size

    def arrays(self):
        return (self.w, self.mu, self.t, self.x, self.cell, self.g)

    def concat(self, o):
        return Pool1MG(*[np.concatenate((a, b)) for a, b in zip(self.arrays(), o.arrays())])

    def take(self, idx):
        return Pool1MG(*[a[idx] for a in self.arrays()])

    def _ptrs(self):
        return [_p(self.w), _p(self.mu), _p(self.t), _p(self.x), _p(self.cell, _i32p), _p(self.g, _i32p)]


def volumes(edges):
    e = _f64(edges)
    v = np.zeros(e.size - 1)
    _lib().m1_volumes(e.size - 1, _p(e), _p(v))
    return v


def counts(E, N, min_one):
    E = _f64(E).ravel()
    cnt = np.zeros(E.size, np.int64)
    tot = _lib().m1_counts(E.size, _p(E), int(N), int(min_one), _p(cnt, _i64p))
    return cnt, int(tot)


def equilibrium_energy(Tr, vol, T_emit_floor=0.0):
    Tr = _f64(Tr); E = np.zeros(Tr.size)
    _lib().m1_equilibrium_energy(Tr.size, _p(Tr), _p(_f64(vol)), float(T_emit_floor), _p(E))
    return E


def sample_ic(edges, Tr, E, cnt, energy_edges, seed):
    n = int(cnt.sum()); G = len(energy_edges) - 1
    p = Pool1MG.empty(n)
    _lib().m1_sample_ic(len(edges) - 1, _p(_f64(edges)), _p(_f64(Tr)), _p(_f64(E)),
                        _p(np.ascontiguousarray(cnt, np.int64), _i64p), G, _p(_f64(energy_edges)),
                        _p(OM.planck_cdf()), seed, *p._ptrs())
    return p


def sample_bins(edges, G, E, cnt, dt, seed, step, tag):
    n = int(cnt.sum())
    p = Pool1MG.empty(n)
    _lib().m1_sample_bins(len(edges) - 1, G, _p(_f64(edges)), _p(_f64(E).ravel()),
                          _p(np.ascontiguousarray(cnt, np.int64).ravel(), _i64p), float(dt), seed, step, tag,
                          *p._ptrs())
    return p


def face_energy(T, dt, R):
    return _lib().m1_face_energy(float(T), float(dt), float(R))


def sample_face(side, n, E, T, dt, R, I, energy_edges, seed, step):
    G = len(energy_edges) - 1
    p = Pool1MG.empty(n)
    _lib().m1_sample_face(side, n, float(E), float(T), float(dt), float(R), I, G, _p(_f64(energy_edges)),
                          _p(OM.planck_cdf()), seed, step, *p._ptrs())
    return p


def transport(pool, edges, G, sa, ss, dt, reflect, weight_floor, seed, step, nthreads=0):
    """In place.  Returns (dep_raw (G,I), path_raw (G,I), bl (G,), stats[8])."""
    I = len(edges) - 1
    dep = np.zeros(G * I); path = np.zeros(G * I); bl = np.zeros(G); st = np.zeros(8, np.int64)
    refl = np.array([1 if r else 0 for r in reflect], np.int32)
    _lib().m1_transport(len(pool), _p(pool.w), _p(pool.mu), _p(pool.t), _p(pool.x), _p(pool.cell, _i32p),
                        _p(pool.g, _i32p), I, _p(_f64(edges)), G, _p(_f64(sa).ravel()), _p(_f64(ss).ravel()),
                        float(dt), _p(refl, _i32p), float(weight_floor), seed, step, _p(dep), _p(path), _p(bl),
                        _p(st, _i64p), int(nthreads))
    return dep.reshape(G, I), path.reshape(G, I), bl, st


def tallies(dep_raw, path_raw, sa, vol, dt):
    """dep = sum dep / V; si = sum dep / (sigma_a dt V) where sigma_a > 1e-10,
    else sum w s / (dt V) (MG_IMC1D.py:341-347, 413-414, 455-458)."""
    thick = sa > 1e-10
    dep = dep_raw / vol[None, :]
    si = np.where(thick, dep_raw / np.where(thick, sa, 1.0), path_raw) / (dt * vol[None, :])
    return dep, si


def comb(pool, Nmax, I, G, seed, step, radiation):
    """(cell, group) comb (MG_IMC1D.py:746-843) over the transported pool (the
    comb stream of a particle is its pool position; zero weights get no
    copies).  radiation: rescale survivors per bin to the pre-comb energy.
    Returns (new pool, pre - post energy per bin): in radiation mode that is
    the energy of bins whose copies all vanished, which the reference
    reports as zero and drops unbooked."""
    b = pool.g.astype(np.int64) * I + np.clip(pool.cell, 0, I - 1)
    ecen, ew, cp, tot, passthrough = O.comb_counts(pool.w, b, G * I, Nmax, seed, step)
    idx = np.repeat(np.arange(len(pool)), cp)
    newp = pool.take(idx)
    newp.cell = np.clip(newp.cell, 0, I - 1).astype(np.int32)
    nbid = b[idx]
    if not passthrough:
        newp.w = ew[nbid]
    if radiation and len(newp):
        post = np.bincount(nbid, weights=newp.w, minlength=G * I)
        scale = np.where(post > 0.0, ecen / (post + 1e-300), 1.0)
        newp.w = newp.w * scale[nbid]
        after = np.where(post > 0.0, ecen, 0.0)
    else:
        after = np.bincount(nbid, weights=newp.w, minlength=G * I) if len(newp) else np.zeros(G * I)
    return newp, (ecen - after).reshape(G, I)


@dataclass
class StateMG1D:
    pool: Pool1MG
    internal_energy: np.ndarray
    temperature: np.ndarray
    radiation_temperature: np.ndarray
    radiation_energy_by_group: np.ndarray
    radiation_flux_by_group: np.ndarray
    time: float
    previous_total_energy: float
    count: int = 0


class OracleMG1D:
    """FP64 CPU restatement of ``MG_IMC/MG_IMC1D.py`` (spherical)."""

    def __init__(self, mesh, energy_edges, seed=1, nthreads=0, gray_bg=False):
        self.gray_bg = bool(gray_bg)  # the reference's B_g fallback (oracle_mg.fleck_mg_gray_bg)
        mesh = np.asarray(mesh, dtype=np.float64)
        self.edges = np.concatenate((mesh[:, 0], mesh[-1:, 1]))
        self.I = mesh.shape[0]
        self.energy_edges = _f64(energy_edges)
        self.G = self.energy_edges.size - 1
        self.vol = volumes(self.edges)
        self.seed = int(seed)
        self.nthreads = nthreads

    def _by_bin(self, pool):
        I, G = self.I, self.G
        ok = (pool.cell >= 0) & (pool.cell < I) & (pool.w > 0.0)
        b = pool.g[ok].astype(np.int64) * I + pool.cell[ok]
        E = np.bincount(b, weights=pool.w[ok], minlength=G * I).reshape(G, I)
        F = np.bincount(b, weights=pool.w[ok] * pool.mu[ok], minlength=G * I).reshape(G, I)
        return E, C_LIGHT * F / self.vol[None, :]

    def init_simulation(self, Ntarget, Tinit, Tr_init, eos, inv_eos, Ntarget_ic=None, T_emit_floor=0.0):
        u = eos(Tinit)
        assert np.allclose(inv_eos(u), Tinit), "Inverse EOS failed"
        N_ic = Ntarget if Ntarget_ic is None else Ntarget_ic
        E = equilibrium_energy(Tr_init, self.vol, T_emit_floor)
        if N_ic > 0 and np.sum(E) > 0:
            cnt, _ = counts(E, N_ic, 1)
            pool = sample_ic(self.edges, Tr_init, E, cnt, self.energy_edges, self.seed)
        else:
            pool = Pool1MG.empty(0)
        Eg, _ = self._by_bin(pool)
        Tr = (Eg.sum(axis=0) / self.vol / A_RAD) ** 0.25
        tot = float(np.sum(u * self.vol)) + float(np.sum(pool.w))
        return StateMG1D(pool, u, np.array(Tinit, dtype=np.float64, copy=True), Tr, Eg, np.zeros_like(Eg), 0.0,
                         tot, 0)

    def step(self, state, Ntarget, Nboundary, Nsource, Nmax, T_boundary, dt, sigma_a_funcs, inv_eos, cv, source,
             reflect=(False, False), theta=1.0, use_scalar_intensity_Tr=True, conserve_comb_energy=True,
             T_emit_floor=0.0, particle_budget_fmin=0.0):
        I, G, vol, edges = self.I, self.G, self.vol, self.edges
        sid = state.count
        T = state.temperature
        sig = np.array([np.asarray(f(T), dtype=np.float64) * np.ones(I) for f in sigma_a_funcs])
        cvv = np.asarray(cv(T), dtype=np.float64) * np.ones(I)
        sa, ss, bstar, f = OM.fleck_mg(T, sig, cvv, theta, dt, self.energy_edges, self.gray_bg)
        Ebin, Ecell = OM.emission_mg(T, sa, bstar, vol, dt, T_emit_floor)
        Tin = T_boundary[0](state.time) if callable(T_boundary[0]) else T_boundary[0]
        Tout = T_boundary[1](state.time) if callable(T_boundary[1]) else T_boundary[1]
        src = np.asarray(source, dtype=np.float64)
        src_mg = np.zeros((G, I))
        if src.ndim == 1:
            src_mg[0, :] = src  # gray source into group 0 (MG_IMC1D.py:695-698)
        else:
            src_mg[:] = src
        if particle_budget_fmin > 0.0:
            E_in = face_energy(Tin, dt, edges[0]) if Tin > 0.0 else 0.0
            E_out = face_energy(Tout, dt, edges[-1]) if Tout > 0.0 else 0.0
            E_src = float(np.sum(np.maximum(src_mg, 0.0) * vol[None, :] * dt))
            E_emit = float(np.sum(Ecell))
            E_all = E_in + E_out + E_src + E_emit
            Nf = max(1, int(round(particle_budget_fmin * int(Ntarget))))
            if E_all > 0.0:
                part = lambda E: max(int(round(int(Ntarget) * E / E_all)), Nf) if E > 0.0 else 0
                N_in, N_out, N_src, N_emit = part(E_in), part(E_out), part(E_src), part(E_emit)
            else:
                N_in, N_out, N_src, N_emit = int(Ntarget), 0, 0, 0
        else:
            N_in = N_out = int(Nboundary)
            N_src, N_emit = int(Nsource), int(Ntarget)
        pool = state.pool
        be = 0.0
        for side, Tb, Nb, R in ((0, Tin, N_in, edges[0]), (1, Tout, N_out, edges[-1])):
            if Nb > 0 and Tb > 0.0:
                E = face_energy(Tb, dt, R)
                b = sample_face(side, Nb, E, Tb, dt, R, I, self.energy_edges, self.seed, sid)
                pool = pool.concat(b)
                be += E
        se = 0.0
        if N_src > 0 and np.max(src_mg) > 0.0:
            QV = src_mg * vol[None, :] * dt
            cnt, tot = counts(QV, N_src, 0)
            if tot > 0:
                s = sample_bins(edges, G, QV, cnt, dt, self.seed, sid, TAG_SOURCE)
                pool = pool.concat(s)
                se = float(np.sum(s.w))
        emitted = np.zeros((G, I))
        if N_emit > 0 and np.sum(Ebin) > 0.0:
            cnt, tot = counts(Ebin, N_emit, 1)
            e = sample_bins(edges, G, Ebin, cnt, dt, self.seed, sid, TAG_EMISSION)
            pool = pool.concat(e)
            emitted = Ebin.reshape(G, I)
        n_tr = len(pool)
        floor = 1e-14 * O.canonical_sum(pool.w) / max(n_tr, 1)
        pool = Pool1MG(*[a.copy() for a in pool.arrays()])
        dep_raw, path_raw, bl_g, stats = transport(pool, edges, G, sa.reshape(G, I), ss.reshape(G, I), dt, reflect,
                                                   floor, self.seed, sid, self.nthreads)
        if stats[6] > 0:
            raise AssertionError("Non-positive distance in spherical transport")
        dep, si = tallies(dep_raw, path_raw, sa.reshape(G, I), vol, dt)
        u = state.internal_energy + dep.sum(axis=0) - emitted.sum(axis=0) / vol
        Tn = inv_eos(u)
        if use_scalar_intensity_Tr:
            Tr = (si.sum(axis=0) / A_RAD / C_LIGHT) ** 0.25
        else:
            ok = (pool.cell >= 0) & (pool.cell < I) & (pool.w > 0.0)
            Tr = (np.bincount(pool.cell[ok], weights=pool.w[ok], minlength=I) / vol / A_RAD) ** 0.25
        target = abs(Nmax)
        n_alive = int(np.sum(pool.w > 0.0))
        do_comb = Nmax > 0 or (Nmax < 0 and n_alive >= target)
        disc = np.zeros((G, I))
        if do_comb:
            newp, disc = comb(pool, target, I, G, self.seed, sid, conserve_comb_energy == "radiation")
            if (conserve_comb_energy is True or conserve_comb_energy == "material") and np.any(disc != 0.0):
                # the discrepancy goes into the material (MG_IMC1D.py:1217-1221)
                u = u + disc.sum(axis=0) / vol
                Tn = inv_eos(u)
        else:
            newp = pool.take(np.nonzero(pool.w > 0.0)[0])
        newp.t = np.zeros(len(newp))
        Eg, Fg = self._by_bin(newp)
        ti = float(np.sum(u * vol))
        tr = float(np.sum(newp.w))
        total = ti + tr
        resid = total - state.previous_total_energy - be + float(np.sum(bl_g)) - se
        state.pool = newp
        state.internal_energy = u
        state.temperature = Tn
        state.radiation_temperature = Tr
        state.radiation_energy_by_group = Eg
        state.radiation_flux_by_group = Fg
        state.time += dt
        state.previous_total_energy = total
        state.count += 1
        info = {"time": state.time, "temperature": Tn, "radiation_temperature": Tr, "N_particles": len(newp),
                "total_energy": total, "total_internal_energy": ti, "total_radiation_energy": tr,
                "boundary_emission": be, "boundary_loss": float(np.sum(bl_g)), "boundary_loss_by_group": bl_g,
                "source_emission": se, "energy_residual": resid, "comb_discrepancy": disc, "deposited": dep,
                "scalar_intensity": si, "emitted_by_group": emitted, "stats": stats,
                "radiation_energy_by_group": Eg, "radiation_flux_by_group": Fg, "n_transported": n_tr,
                "budget": (N_in, N_out, N_src, N_emit)}
        return state, info
