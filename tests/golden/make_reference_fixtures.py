"""Write statistical golden fixtures from the reference implementation.

The reference has no golden vectors for the IMC step (its RNG is Numba's
hidden per-thread MT19937, so no single history can be reproduced).  What
only the reference itself can pin is the composition of its ``step()``:
sourcing, Fleck factor, transport, material update and comb together.  This
script imports the reference modules from /root/reference (Numba is absent in
this container, so they run through their documented pure-Python fallback,
IMC/IMC2D.py:26-46, IMC/IMC1D.py:46-66, MG_IMC/MG_IMC2D.py:38-58,
MG_IMC/MG_IMC1D.py:40-60), runs tiny problems over many seeds with both
``random`` and ``np.random`` seeded, and stores per-cell ensemble means and
standard errors of T and T_r plus energy totals.

The multigroup modules import here without the absent ``planck_integrals``
and fall back to a gray B_g, equal in every group (MG_IMC/MG_IMC2D.py:60-94,
MG_IMC/MG_IMC1D.py:63-94): the Planck-mean opacity of the Fleck factor is the
plain group mean and b*_g = 1/G.  The multigroup fixtures therefore pin the
composition of the multigroup steps (group-dependent opacities, sourcing,
group-aware transport, per-(cell, group) comb, budget split) under that B_g;
the oracle runs them with the same gray B_g.

Run from the repository root:
  python tests/golden/make_reference_fixtures.py [name ...]
Outputs: tests/golden/<name>.npz for name in ref_imc2d_xy, ref_imc2d_rz,
ref_imc1d_slab, ref_imc1d_sphere, ref_mgimc2d_xy, ref_mgimc1d_budget,
ref_mgimc1d_radcomb (all of them without arguments).
"""
from __future__ import annotations

import contextlib
import io
import os
import random
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference/IMC"
REF_MG = "/root/reference/MG_IMC"

# problem definitions shared with tests/test_reference_fixtures.py
XY = dict(nx=6, ny=5, Lx=0.3, Ly=0.25, N=600, Nb=600, Nmax=1200, dt=0.01, steps=3, Tb=0.6, cv=0.3, T0=0.05)
RZ = dict(nx=5, ny=6, R=0.25, Lz=0.3, N=600, Nb=600, Nmax=1200, dt=0.01, steps=3, Tb=0.6, cv=0.3, T0=0.05)
SLAB = dict(I=20, L=0.2, N=1500, Nb=1500, Nmax=3000, dt=0.01, steps=4, Tb=1.0, cv=0.3, T0=0.01)
SPH = dict(I=15, R=0.3, N=1500, Nb=1500, Nmax=3000, dt=0.01, steps=4, Tb=0.8, cv=0.3, T0=0.05, s0=1.0)
# multigroup: three groups, opacities sigma0(cell) * mult_g * max(T, 1e-3)^-3
MG_EDGES = (1e-3, 0.2, 1.0, 20.0)
MG_MULT = (4.0, 1.0, 0.25)
MG2D = dict(nx=5, ny=4, Lx=0.25, Ly=0.2, N=400, Nb=400, Nmax=800, dt=0.005, steps=3, Tb=0.5, cv=0.3, T0=0.2)
# sphere, budget split (particle_budget_fmin > 0: N is the total) and the comb
# discrepancy into the material
MG1D_BUDGET = dict(I=10, R=0.2, N=1200, Nb=0, Nmax=600, dt=0.005, steps=3, Tb=0.5, cv=0.3, T0=0.2, s0=5.0,
                   fmin=0.1, conserve=True)
# sphere, fixed counts, threshold comb (Nmax < 0) with radiation rescaling
MG1D_RADCOMB = dict(I=10, R=0.2, N=600, Nb=600, Nmax=-300, dt=0.005, steps=3, Tb=0.5, cv=0.3, T0=0.2, s0=5.0,
                    fmin=0.0, conserve="radiation")
SEEDS = 200


def sigma0_xy():
    g = np.random.default_rng(11).standard_normal((XY["nx"], XY["ny"]))
    return np.exp(1.5 + 0.8 * g)


def sigma0_rz():
    g = np.random.default_rng(12).standard_normal((RZ["nx"], RZ["ny"]))
    return np.exp(1.5 + 0.8 * g)


def run_xy(mod, seed):
    p = XY
    e1 = np.linspace(0, p["Lx"], p["nx"] + 1)
    e2 = np.linspace(0, p["Ly"], p["ny"] + 1)
    s0 = sigma0_xy()
    sf = lambda T: s0 * np.maximum(T, 1e-3) ** -3
    cv = p["cv"]
    T0 = np.full((p["nx"], p["ny"]), p["T0"])
    st = mod.init_simulation(p["N"], T0, T0, e1, e2, lambda T: cv * T, lambda u: u / cv, geometry="xy")
    out = []
    for _ in range(p["steps"]):
        st, info = mod.step(st, p["N"], p["Nb"], 0, p["Nmax"], (p["Tb"], 0.0, 0.0, 0.0), p["dt"], e1, e2, sf,
                            lambda u: u / cv, lambda T: cv + 0 * T, np.zeros_like(T0),
                            reflect=(False, False, True, False), geometry="xy")
        out.append((info["temperature"].copy(), info["radiation_temperature"].copy(), info["total_energy"],
                    info["total_radiation_energy"], info["boundary_loss"], info["N_particles"]))
    return out


def run_rz(mod, seed):
    p = RZ
    e1 = np.linspace(0, p["R"], p["nx"] + 1)
    e2 = np.linspace(0, p["Lz"], p["ny"] + 1)
    s0 = sigma0_rz()
    sf = lambda T: s0 * np.maximum(T, 1e-3) ** -3
    cv = p["cv"]
    T0 = np.full((p["nx"], p["ny"]), p["T0"])
    st = mod.init_simulation(p["N"], T0, T0, e1, e2, lambda T: cv * T, lambda u: u / cv, geometry="rz")
    out = []
    for _ in range(p["steps"]):
        st, info = mod.step(st, p["N"], p["Nb"], 0, p["Nmax"], (0.0, 0.0, p["Tb"], 0.0), p["dt"], e1, e2, sf,
                            lambda u: u / cv, lambda T: cv + 0 * T, np.zeros_like(T0),
                            reflect=(False, False, False, False), geometry="rz", rz_linear_source=True)
        out.append((info["temperature"].copy(), info["radiation_temperature"].copy(), info["total_energy"],
                    info["total_radiation_energy"], info["boundary_loss"], info["N_particles"]))
    return out


def run_slab(mod, seed):
    p = SLAB
    I = p["I"]
    edges = np.linspace(0, p["L"], I + 1)
    mesh = np.column_stack([edges[:-1], edges[1:]])
    cv = p["cv"]
    sf = lambda T: 300.0 * np.maximum(T, 1e-3) ** -3
    T0 = np.full(I, p["T0"])
    st = mod.init_simulation(p["N"], T0, T0, mesh, lambda T: cv * T, lambda u: u / cv, geometry="slab")
    out = []
    for _ in range(p["steps"]):
        st, info = mod.step(st, p["N"], p["Nb"], 0, p["Nmax"], (p["Tb"], 0.0), p["dt"], mesh, sf, lambda u: u / cv,
                            lambda T: cv + 0 * T, np.zeros(I), reflect=(False, False), geometry="slab")
        out.append((info["temperature"].copy(), info["radiation_temperature"].copy(), info["total_energy"],
                    info["total_radiation_energy"], info["boundary_loss"], info["N_particles"]))
    return out


def run_sphere(mod, seed):
    p = SPH
    I = p["I"]
    edges = np.linspace(0, p["R"], I + 1)
    mesh = np.column_stack([edges[:-1], edges[1:]])
    cv = p["cv"]
    s0 = p["s0"]
    sf = lambda T: s0 * np.maximum(T, 1e-3) ** -3
    T0 = np.full(I, p["T0"])
    st = mod.init_simulation(p["N"], T0, T0, mesh, lambda T: cv * T, lambda u: u / cv, geometry="spherical")
    out = []
    for _ in range(p["steps"]):
        st, info = mod.step(st, p["N"], p["Nb"], 0, p["Nmax"], (0.0, p["Tb"]), p["dt"], mesh, sf, lambda u: u / cv,
                            lambda T: cv + 0 * T, np.zeros(I), reflect=(False, False), geometry="spherical")
        out.append((info["temperature"].copy(), info["radiation_temperature"].copy(), info["total_energy"],
                    info["total_radiation_energy"], info["boundary_loss"], info["N_particles"]))
    return out


def sigma0_mg2d():
    g = np.random.default_rng(13).standard_normal((MG2D["nx"], MG2D["ny"]))
    return np.exp(0.5 + 0.6 * g)


def mg_sigma_funcs(s0):
    return [(lambda T, m=m: s0 * m * np.maximum(T, 1e-3) ** -3) for m in MG_MULT]


def run_mg2d(mod, seed):
    p = MG2D
    e1 = np.linspace(0, p["Lx"], p["nx"] + 1)
    e2 = np.linspace(0, p["Ly"], p["ny"] + 1)
    edges = np.array(MG_EDGES)
    sfs = mg_sigma_funcs(sigma0_mg2d())
    cv = p["cv"]
    T0 = np.full((p["nx"], p["ny"]), p["T0"])
    st = mod.init_simulation(p["N"], T0, T0, e1, e2, edges, lambda T: cv * T, lambda u: u / cv, geometry="xy")
    out = []
    for _ in range(p["steps"]):
        st, info = mod.step(st, p["N"], p["Nb"], 0, p["Nmax"], (p["Tb"], 0.0, 0.0, 0.0), p["dt"], e1, e2, edges, sfs,
                            lambda u: u / cv, lambda T: cv + 0 * T, np.zeros_like(T0),
                            reflect=(False, False, True, True), geometry="xy")
        out.append((info["temperature"].copy(), info["radiation_temperature"].copy(), info["total_energy"],
                    info["total_radiation_energy"], info["boundary_loss"], info["N_particles"]))
    return out


def _run_mg1d(mod, p):
    I = p["I"]
    edges = np.linspace(0, p["R"], I + 1)
    mesh = np.column_stack([edges[:-1], edges[1:]])
    G = len(MG_EDGES) - 1
    s0 = p["s0"] * np.linspace(1.0, 2.0, I)
    sfs = mg_sigma_funcs(s0)
    cv = p["cv"]
    T0 = np.full(I, p["T0"])
    st = mod.init_simulation(p["N"], T0, T0, mesh, np.array(MG_EDGES), lambda T: cv * T, lambda u: u / cv)
    out = []
    for _ in range(p["steps"]):
        st, info = mod.step(st, p["N"], p["Nb"], 0, p["Nmax"], (0.0, p["Tb"]), p["dt"], mesh, np.array(MG_EDGES), sfs,
                            lambda u: u / cv, lambda T: cv + 0 * T, np.zeros(I), reflect=(False, False),
                            conserve_comb_energy=p["conserve"], particle_budget_fmin=p["fmin"])
        out.append((info["temperature"].copy(), info["radiation_temperature"].copy(), info["total_energy"],
                    info["total_radiation_energy"], info["boundary_loss"], info["N_particles"]))
    return out


def run_mg1d_budget(mod, seed):
    return _run_mg1d(mod, MG1D_BUDGET)


def run_mg1d_radcomb(mod, seed):
    return _run_mg1d(mod, MG1D_RADCOMB)


def ensemble(runner, mod, seeds):
    rows = []
    for s in range(seeds):
        random.seed(1000 + s)
        np.random.seed(1000 + s)
        with contextlib.redirect_stdout(io.StringIO()):
            rows.append(runner(mod, s))
        print(f"  seed {s + 1}/{seeds}", flush=True)
    nsteps = len(rows[0])
    res = {}
    for k, name in enumerate(["T", "Tr", "E_total", "E_rad", "boundary_loss", "N"]):
        arr = np.array([[r[st][k] for st in range(nsteps)] for r in rows], dtype=np.float64)
        res[name + "_mean"] = arr.mean(0)
        res[name + "_sem"] = arr.std(0, ddof=1) / np.sqrt(arr.shape[0])
    res["seeds"] = np.array(seeds)
    return res


def main():
    sys.path.insert(0, REF)
    sys.path.insert(0, REF_MG)
    import IMC1D as ref1d  # noqa: E402
    import IMC2D as ref2d  # noqa: E402
    with contextlib.redirect_stdout(io.StringIO()):
        import MG_IMC1D as refmg1d  # noqa: E402
        import MG_IMC2D as refmg2d  # noqa: E402

    jobs = (("ref_imc2d_xy", run_xy, ref2d), ("ref_imc2d_rz", run_rz, ref2d), ("ref_imc1d_slab", run_slab, ref1d),
            ("ref_imc1d_sphere", run_sphere, ref1d), ("ref_mgimc2d_xy", run_mg2d, refmg2d),
            ("ref_mgimc1d_budget", run_mg1d_budget, refmg1d), ("ref_mgimc1d_radcomb", run_mg1d_radcomb, refmg1d))
    only = set(sys.argv[1:])
    for name, runner, mod in jobs:
        if only and name not in only:
            continue
        print(name, flush=True)
        res = ensemble(runner, mod, SEEDS)
        np.savez(os.path.join(HERE, name + ".npz"), **res)


if __name__ == "__main__":
    main()
