"""oracle_mg.py -- TEST INFRASTRUCTURE ONLY (the checker and the timed CPU baseline).

ctypes bindings to the multigroup part of ``oracle/_build/liboracle.so``
(``imc_oracle_mg.c``) and a host composition of one multigroup 2-D step that
follows the reference's ``MG_IMC/MG_IMC2D.py:step`` (1635-2122) line by line,
with the deliberate departures listed in ``imc_oracle_mg.c``.

Arrays: per-(group, cell) data are (G, nx, ny) on the outside and flat
(G, ncell) with cell c = i*ny + j inside.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg
may import this module.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import oracle as O

A_RAD = O.A_RAD
C_LIGHT = O.C_LIGHT
_dp, _i64p, _i32p, _f64 = O._dp, O._i64p, O._i32p, O._f64


def _p(a, ct=_dp):
    return None if a is None else a.ctypes.data_as(ct)


def _lib():
    L = O.lib()
    if not hasattr(L, "_om_ready"):
        L.om_planck_cdf.argtypes = [_dp]
        L.om_planck_fraction.argtypes = [C.c_double, C.c_double]
        L.om_planck_fraction.restype = C.c_double
        L.om_bg.argtypes = [C.c_double, C.c_double, C.c_double]
        L.om_bg.restype = C.c_double
        L.om_fleck_mg.argtypes = [C.c_int64, C.c_int, _dp, _dp, _dp, _dp, C.c_double, C.c_double, _dp, _dp, _dp, _dp]
        L.om_emission_mg.argtypes = [C.c_int64, C.c_int, _dp, _dp, _dp, _dp, C.c_double, C.c_double, _dp, _dp]
        L.om_prefix.argtypes = [_dp, C.c_int64, _dp]
        L.om_group_cdf.argtypes = [C.c_int, _dp, _dp]
        L.om_sample_emission.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp, C.c_int64,
                                         C.c_double, C.c_uint64, C.c_uint32] + [_dp] * 6 + [_i32p] * 3
        L.om_sample_ic.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, C.c_int64,
                                   C.c_uint64] + [_dp] * 6 + [_i32p] * 3
        L.om_sample_boundary.argtypes = [C.c_int, C.c_int, C.c_int64, C.c_double, C.c_double, C.c_double, C.c_int,
                                         C.c_int, _dp, _dp, C.c_int, _dp, _dp, _dp, _dp, C.c_uint64,
                                         C.c_uint32] + [_dp] * 6 + [_i32p] * 3
        L.om_sample_source.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp, C.c_double,
                                       C.c_int64, C.c_double, C.c_uint64, C.c_uint32] + [_dp] * 6 + [_i32p] * 3
        L.om_transport.argtypes = [C.c_int, C.c_int64, _dp, _dp, _dp, _dp, _dp, _dp, _i32p, _i32p, _i32p, C.c_int,
                                   C.c_int, C.c_int, _dp, _dp, _dp, _dp, C.c_double, _i32p, C.c_int64, C.c_double,
                                   C.c_uint64, C.c_uint32, _dp, _dp, _dp, _i64p]
        L.om_comb_counts.argtypes = [C.c_int64, _dp, _i32p, _i32p, _i32p, C.c_int, C.c_int, C.c_int, C.c_int64,
                                     C.c_uint64, C.c_uint32, _dp, _dp, _i64p, _i32p]
        L.om_comb_counts.restype = C.c_int64
        L._om_ready = True
    return L


@dataclass
class PoolMG:
    w: np.ndarray
    d1: np.ndarray
    d2: np.ndarray
    t: np.ndarray
    p1: np.ndarray
    p2: np.ndarray
    ci: np.ndarray
    cj: np.ndarray
    g: np.ndarray

    @staticmethod
    def empty(n=0):
        z = lambda: np.zeros(n)
        zi = lambda: np.zeros(n, np.int32)
        return PoolMG(z(), z(), z(), z(), z(), z(), zi(), zi(), zi())

    def __len__(self):
        return self.w.size

    def arrays(self):
        return (self.w, self.d1, self.d2, self.t, self.p1, self.p2, self.ci, self.cj, self.g)

    def concat(self, o):
        return PoolMG(*[np.concatenate((a, b)) for a, b in zip(self.arrays(), o.arrays())])

    def take(self, idx):
        return PoolMG(*[a[idx] for a in self.arrays()])

    def _ptrs(self):
        return [_p(a) for a in self.arrays()[:6]] + [_p(a, _i32p) for a in self.arrays()[6:]]


_PCDF = None


def planck_cdf():
    global _PCDF
    if _PCDF is None:
        _PCDF = np.zeros(200)
        _lib().om_planck_cdf(_p(_PCDF))
    return _PCDF


def planck_fraction(xlo, xhi):
    return _lib().om_planck_fraction(float(xlo), float(xhi))


def bg(T, elo, ehi):
    return _lib().om_bg(float(T), float(elo), float(ehi))


def volumes(e1, e2, geometry):
    e1 = _f64(e1); e2 = _f64(e2)
    if geometry == "xy":
        return (np.diff(e1)[:, None] * np.diff(e2)[None, :]).ravel()
    dr2 = e1[1:] * e1[1:] - e1[:-1] * e1[:-1]
    return (np.pi * dr2[:, None] * np.diff(e2)[None, :]).ravel()


def fleck_mg_gray_bg(T, sigma, cv, theta, dt):
    """Fleck block of MG_IMC/MG_IMC2D.py:1702-1719 and the b* of 1879-1885
    under the reference's gray B_g fallback (MG_IMC/MG_IMC2D.py:60-94,
    MG_IMC/MG_IMC1D.py:63-94, used when planck_integrals is absent): every
    group gets the same B = (c a / 4) T^4 / (4 pi), so sigma_P is the group
    mean and b*_g = 1/G (0 at T = 0).  Test infrastructure for the fixtures the
    reference writes in that mode (tests/golden/make_reference_fixtures.py)."""
    T = _f64(T).ravel(); sig = _f64(sigma).reshape(-1, T.size); cv = _f64(cv).ravel()
    G = sig.shape[0]
    B = np.broadcast_to((C_LIGHT * A_RAD / 4.0) * T ** 4 / (4.0 * np.pi), sig.shape)
    bs_sum = np.sum(B, axis=0) + 1e-300
    sP = np.sum(sig * B, axis=0) / bs_sum
    beta = 4.0 * A_RAD * T ** 3 / cv
    f = np.clip(1.0 / (1.0 + theta * beta * sP * C_LIGHT * dt), 0.0, 1.0)
    return sig * f[None, :], sig * (1.0 - f)[None, :], B / bs_sum[None, :], f


def fleck_mg(T, sigma, cv, theta, dt, edges, gray_bg=False):
    """sigma (G, ncell) -> (sa, ss, bstar, f)."""
    if gray_bg:
        return fleck_mg_gray_bg(T, sigma, cv, theta, dt)
    T = _f64(T).ravel(); sig = _f64(sigma).reshape(-1, T.size); cv = _f64(cv).ravel()
    G, n = sig.shape
    sa = np.zeros_like(sig); ss = np.zeros_like(sig); bs = np.zeros_like(sig); f = np.zeros(n)
    _lib().om_fleck_mg(n, G, _p(_f64(edges)), _p(T), _p(sig), _p(cv), float(theta), float(dt), _p(sa), _p(ss),
                       _p(bs), _p(f))
    return sa, ss, bs, f


def emission_mg(T, sa, bstar, vol, dt, T_emit_floor=0.0):
    T = _f64(T).ravel(); G, n = sa.shape
    E = np.zeros((G, n)); tot = np.zeros(n)
    _lib().om_emission_mg(n, G, _p(T), _p(_f64(sa)), _p(_f64(bstar)), _p(_f64(vol)), float(dt), float(T_emit_floor),
                          _p(E), _p(tot))
    return E, tot


def prefix(x):
    x = _f64(x).ravel(); out = np.zeros_like(x)
    _lib().om_prefix(_p(x), x.size, _p(out))
    return out


def group_cdf(tot):
    tot = _f64(tot); out = np.zeros_like(tot)
    _lib().om_group_cdf(tot.size, _p(tot), _p(out))
    return out


def sample_emission(geometry, e1, e2, E, cdf, N, dt, seed, step):
    """N emission particles (MG_IMC2D.py:1896-1943) in non-decreasing cell
    order: their cells are the order statistics of N uniforms (om_sample_emission)."""
    G, n = E.shape; nx, ny = len(e1) - 1, len(e2) - 1
    p = PoolMG.empty(int(N))
    _lib().om_sample_emission(0 if geometry == "xy" else 1, nx, ny, G, _p(_f64(e1)), _p(_f64(e2)), _p(_f64(E)),
                              _p(_f64(cdf)), int(N), float(dt), seed, step, *p._ptrs())
    return p


def sample_ic(geometry, e1, e2, edges, Tr, cdf, N, seed):
    nx, ny = len(e1) - 1, len(e2) - 1; G = len(edges) - 1
    p = PoolMG.empty(int(N))
    _lib().om_sample_ic(0 if geometry == "xy" else 1, nx, ny, G, _p(_f64(e1)), _p(_f64(e2)), _p(_f64(edges)),
                        _p(_f64(Tr).ravel()), _p(_f64(cdf)), _p(planck_cdf()), int(N), seed, *p._ptrs())
    return p


def sample_boundary(geometry, side, n, E_total, dt, Tb, e1, e2, edges, prof, Tface, seed, step):
    nx, ny = len(e1) - 1, len(e2) - 1; G = len(edges) - 1
    p = PoolMG.empty(int(n))
    pr = None if prof is None else _f64(prof)
    tf = None if Tface is None else _f64(Tface)
    _lib().om_sample_boundary(0 if geometry == "xy" else 1, side, int(n), float(E_total), float(dt), float(Tb), nx,
                              ny, _p(_f64(e1)), _p(_f64(e2)), G, _p(_f64(edges)), _p(planck_cdf()), _p(pr), _p(tf),
                              seed, step, *p._ptrs())
    return p


def sample_source(geometry, e1, e2, S, N, dt, seed, step):
    """S: (G, ncell) energy per cell this step."""
    G, n = S.shape; nx, ny = len(e1) - 1, len(e2) - 1
    cdfg = np.stack([prefix(S[g]) for g in range(G)])
    gcdf = group_cdf(cdfg[:, -1])
    total = 0.0
    for g in range(G):
        total += cdfg[g, -1]
    p = PoolMG.empty(int(N))
    _lib().om_sample_source(0 if geometry == "xy" else 1, nx, ny, G, _p(_f64(e1)), _p(_f64(e2)), _p(cdfg), _p(gcdf),
                            total, int(N), float(dt), seed, step, *p._ptrs())
    return p, total


def transport(geometry, pool, e1, e2, G, sa, ss, dt, reflect, max_events, weight_floor, seed, step):
    """In place.  Returns (dep_raw (G,ncell), path_raw (G,ncell), bl[1+G+4], stats[8])."""
    nx, ny = len(e1) - 1, len(e2) - 1
    dep = np.zeros((G, nx * ny)); path = np.zeros((G, nx * ny)); bl = np.zeros(1 + G + 4)
    st = np.zeros(8, np.int64)
    refl = np.array([1 if r else 0 for r in reflect], np.int32)
    _lib().om_transport(0 if geometry == "xy" else 1, len(pool), *pool._ptrs(), nx, ny, G, _p(_f64(e1)),
                        _p(_f64(e2)), _p(_f64(sa)), _p(_f64(ss)), float(dt), _p(refl, _i32p), int(max_events),
                        float(weight_floor), seed, step, _p(dep), _p(path), _p(bl), _p(st, _i64p))
    return dep, path, bl, st


def tallies(dep_raw, path_raw, sa, vol, dt):
    """dep = sum dep / V; si = sum dep / sigma_a / (dt V), or sum w s / (dt V)
    where sigma_a <= 1e-12 (MG_IMC2D.py:872-885)."""
    dep = dep_raw / vol[None, :]
    thick = sa > 1e-12
    si = np.where(thick, dep_raw / np.where(thick, sa, 1.0), path_raw) / (dt * vol[None, :])
    return dep, si


def comb(pool, Nmax, nx, ny, G, seed, step):
    n = len(pool)
    nb = nx * ny * G
    ecen = np.zeros(nb); ew = np.zeros(nb); cp = np.zeros(n, np.int64); pt = np.zeros(1, np.int32)
    _lib().om_comb_counts(n, _p(pool.w), _p(pool.ci, _i32p), _p(pool.cj, _i32p), _p(pool.g, _i32p), nx, ny, G,
                          int(Nmax), seed, step, _p(ecen), _p(ew), _p(cp, _i64p), _p(pt, _i32p))
    idx = np.repeat(np.arange(n), cp)
    newp = pool.take(idx)
    if not pt[0]:
        b = newp.g.astype(np.int64) * (nx * ny) + newp.ci.astype(np.int64) * ny + newp.cj
        newp.w = ew[b]
    newp.t = np.zeros(len(newp))
    b = newp.g.astype(np.int64) * (nx * ny) + newp.ci.astype(np.int64) * ny + newp.cj
    after = np.bincount(b, weights=newp.w, minlength=nb) if len(newp) else np.zeros(nb)
    return newp, (ecen - after).reshape(G, nx * ny), ecen.reshape(G, nx * ny)


_SIDES = {"xy": ("left", "right", "bottom", "top"), "rz": ("rmin", "rmax", "zmin", "zmax")}


def face_energy(geometry, side, Tb, dt, e1, e2):
    """Uniform face: a c T^4/4 * area * dt (MG_IMC2D.py:406-413, 465-470)."""
    if geometry == "xy":
        area = (e2[-1] - e2[0]) if side < 2 else (e1[-1] - e1[0])
    elif side < 2:
        area = 2.0 * np.pi * (e1[0] if side == 0 else e1[-1]) * (e2[-1] - e2[0])
    else:
        area = np.pi * (e1[-1] * e1[-1] - e1[0] * e1[0])
    return A_RAD * C_LIGHT * (Tb * Tb * Tb * Tb) / 4.0 * area * dt


def face_profile(geometry, side, dt, e1, e2, bfunc):
    """(running sum of per-face-cell energy, face-cell temperatures) for a
    position-dependent face temperature (MG_IMC2D.py:472-536); None if no
    face cell emits."""
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
    em = np.where(Tc > 0.0, A_RAD * C_LIGHT * (Tc * Tc * Tc * Tc) / 4.0 * size * dt, 0.0)
    return np.cumsum(em), Tc


def boundary_estimate(geometry, e1, e2, Tb4, dt, bfunc):
    """_estimate_boundary_emission_energy (MG_IMC2D.py:1546-1632), including
    its xy left/right area of 2 (z1 - z0)."""
    E = 0.0
    for side in range(4):
        if bfunc is not None:
            prof = face_profile(geometry, side, dt, e1, e2, bfunc)
            if prof is not None:
                e = float(prof[0][-1])
                if geometry == "xy" and side < 2:
                    e *= 2.0
                E += e
        elif Tb4[side] > 0.0:
            e = face_energy(geometry, side, Tb4[side], dt, e1, e2)
            if geometry == "xy" and side < 2:
                e *= 2.0
            E += e
    return E


@dataclass
class StateMG:
    pool: PoolMG
    internal_energy: np.ndarray
    temperature: np.ndarray
    radiation_temperature: np.ndarray
    time: float
    previous_total_energy: float
    count: int = 0


class OracleMG2D:
    """FP64 CPU restatement of the multigroup 2-D IMC step."""

    def __init__(self, edges1, edges2, energy_edges, geometry="xy", seed=1, gray_bg=False):
        self.gray_bg = bool(gray_bg)  # the reference's B_g fallback (fleck_mg_gray_bg)
        self.e1 = _f64(edges1); self.e2 = _f64(edges2); self.edges = _f64(energy_edges)
        self.nx, self.ny = self.e1.size - 1, self.e2.size - 1
        self.ncell = self.nx * self.ny
        self.G = self.edges.size - 1
        self.geometry = geometry
        self.vol = volumes(self.e1, self.e2, geometry)
        self.seed = int(seed)

    def init_simulation(self, Ntarget, Tinit, Tr_init, eos, inv_eos, Ntarget_ic=None, T_emit_floor=0.0):
        u = eos(Tinit)
        N = Ntarget if Ntarget_ic is None else Ntarget_ic
        tr = _f64(Tr_init).ravel()
        E = A_RAD * ((tr * tr) * (tr * tr)) * self.vol
        if T_emit_floor > 0.0:
            E = np.where(tr < T_emit_floor, 0.0, E)
        cdf = prefix(E)
        if N > 0 and cdf[-1] > 0.0:
            pool = sample_ic(self.geometry, self.e1, self.e2, self.edges, tr, cdf, N, self.seed)
        else:
            pool = PoolMG.empty(0)
        rad = np.bincount(pool.ci.astype(np.int64) * self.ny + pool.cj, weights=pool.w, minlength=self.ncell)
        Tr = (rad / self.vol / A_RAD) ** 0.25
        tot = float(np.sum(_f64(u).ravel() * self.vol)) + float(np.sum(pool.w))
        return StateMG(pool, _f64(u).reshape(self.nx, self.ny), _f64(Tinit).reshape(self.nx, self.ny).copy(),
                       Tr.reshape(self.nx, self.ny), 0.0, tot, 0)

    def step(self, state, Ntarget, Nboundary, Nsource, Nmax, T_boundary, dt, sigma_a_funcs, inv_eos, cv, source,
             reflect=(False, False, False, False), theta=1.0, use_scalar_intensity_Tr=True,
             conserve_comb_energy=False, max_events_per_particle=1_000_000, boundary_source_func=None,
             emission_fractions=None, Ntotal=0, Ntotal_T_floor=0.0, particle_budget_fmin=0.1, T_emit_floor=0.0):
        G, nx, ny, nc = self.G, self.nx, self.ny, self.ncell
        sid = state.count
        T = state.temperature
        sig = np.stack([np.asarray(sigma_a_funcs[g](T), dtype=np.float64) * np.ones((nx, ny)) for g in range(G)])
        cvv = np.asarray(cv(T), dtype=np.float64) * np.ones((nx, ny))
        sa, ss, bstar_planck, _ = fleck_mg(T, sig.reshape(G, nc), cvv, theta, dt, self.edges, self.gray_bg)
        Tb = [Tb_(state.time) if callable(Tb_) else Tb_ for Tb_ in T_boundary]
        if Ntotal > 0:
            E_bc = boundary_estimate(self.geometry, self.e1, self.e2, Tb, dt, boundary_source_func)
            _, cell_em = emission_mg(T, sa, bstar_planck, self.vol, dt, 0.0)
            if Ntotal_T_floor > 0.0:
                cell_em = np.where(T.ravel() > Ntotal_T_floor, cell_em, 0.0)
            E_mat = float(np.sum(cell_em))
            tot = E_bc + E_mat
            frac = E_bc / tot if tot > 0.0 else 0.5
            fmin = float(np.clip(particle_budget_fmin, 0.0, 0.5))
            frac = float(np.clip(frac, fmin, 1.0 - fmin))
            Nboundary = int(round(Ntotal * frac))
            Ntarget = Ntotal - Nboundary
        pool = state.pool
        be = 0.0
        be_g = np.zeros(G); be_s = np.zeros(4)
        if Nboundary > 0:
            for side in range(4):
                prof = Tface = None
                if boundary_source_func is not None:
                    pr = face_profile(self.geometry, side, dt, self.e1, self.e2, boundary_source_func)
                    if pr is None:
                        continue
                    prof, Tface = pr
                    E = float(prof[-1])
                else:
                    if not (Tb[side] > 0.0):
                        continue
                    E = face_energy(self.geometry, side, float(Tb[side]), dt, self.e1, self.e2)
                b = sample_boundary(self.geometry, side, int(Nboundary), E, dt, float(Tb[side] or 0.0), self.e1,
                                    self.e2, self.edges, prof, Tface, self.seed, sid)
                pool = pool.concat(b)
                s = float(np.sum(b.w))
                be += s
                be_s[side] += s
                be_g += np.bincount(b.g, weights=b.w, minlength=G)
        src_e = 0.0
        src = np.asarray(source, dtype=np.float64)
        if Nsource > 0 and np.max(src) > 0.0:
            if src.ndim == 3:
                Sg = src.reshape(G, nc)
            else:
                Sg = np.zeros((G, nc))
                Sg[0] = (src * np.ones((nx, ny))).ravel()
            S = Sg * self.vol[None, :] * dt
            s_pool, total = sample_source(self.geometry, self.e1, self.e2, S, Nsource, dt, self.seed, sid)
            if total > 0.0:
                pool = pool.concat(s_pool)
                src_e = float(np.sum(s_pool.w))
        if emission_fractions is not None:
            bstar = np.repeat(_f64(emission_fractions)[:, None], nc, axis=1)
        else:
            bstar = bstar_planck
        E_g, E_cell = emission_mg(T, sa, bstar, self.vol, dt, T_emit_floor)
        if Ntarget > 0:
            cdf = prefix(E_cell)
            if cdf[-1] > 0.0:
                pool = pool.concat(sample_emission(self.geometry, self.e1, self.e2, E_g, cdf, Ntarget, dt, self.seed,
                                                   sid))
        else:
            E_g = np.zeros_like(E_g)
        n_tr = len(pool)
        floor = 1e-10 * float(np.sum(pool.w)) / max(n_tr, 1)
        dep_raw, path_raw, bl, stats = transport(self.geometry, pool, self.e1, self.e2, G, sa, ss, dt, reflect,
                                                 max_events_per_particle, floor, self.seed, sid)
        dep, si = tallies(dep_raw, path_raw, sa, self.vol, dt)
        u = state.internal_energy.ravel() + dep.sum(0) - E_g.sum(0) / self.vol
        Tn = inv_eos(u.reshape(nx, ny))
        if use_scalar_intensity_Tr:
            erg_pre = si / C_LIGHT
        else:
            ok = (pool.w > 0) & (pool.ci >= 0) & (pool.ci < nx) & (pool.cj >= 0) & (pool.cj < ny)
            b = pool.g[ok].astype(np.int64) * nc + pool.ci[ok].astype(np.int64) * ny + pool.cj[ok]
            erg_pre = np.bincount(b, weights=pool.w[ok], minlength=G * nc).reshape(G, nc) / self.vol[None, :]
        newp, disc, ecen = comb(pool, Nmax, nx, ny, G, self.seed, sid)
        if conserve_comb_energy:
            u = u + disc.sum(0) / self.vol
            Tn = inv_eos(u.reshape(nx, ny))
        Tr = (erg_pre.sum(0) / A_RAD) ** 0.25
        ti = float(np.sum(u * self.vol))
        tr = float(np.sum(newp.w))
        total = ti + tr
        dE = total - state.previous_total_energy
        resid = dE - be + float(bl[0]) - src_e
        state.pool = newp
        state.internal_energy = u.reshape(nx, ny)
        state.temperature = Tn
        state.radiation_temperature = Tr.reshape(nx, ny)
        state.time += dt
        state.previous_total_energy = total
        state.count += 1
        disc_total = 0.0 if conserve_comb_energy else float(np.sum(disc))
        info = {"time": state.time, "temperature": Tn, "radiation_temperature": Tr.reshape(nx, ny),
                "radiation_energy_by_group": erg_pre.reshape(G, nx, ny), "N_particles": len(newp),
                "N_boundary": Nboundary, "N_target": Ntarget, "total_energy": total, "total_internal_energy": ti,
                "total_radiation_energy": tr, "boundary_emission": be, "boundary_emission_by_group": be_g,
                "boundary_emission_by_side": be_s, "boundary_loss": float(bl[0]),
                "boundary_outgoing_by_group": bl[1:1 + G].copy(), "boundary_outgoing_by_side": bl[1 + G:].copy(),
                "source_emission": src_e, "energy_residual": resid, "energy_loss": resid,
                "comb_discrepancy": disc_total, "balance_residual": resid + disc_total,
                "dep": dep.reshape(G, nx, ny), "si": si.reshape(G, nx, ny), "emission_by_group": E_g.reshape(G, nx, ny),
                "stats": stats, "n_transported": n_tr, "weight_floor": floor}
        return state, info
