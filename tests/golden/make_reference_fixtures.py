"""Write statistical golden fixtures from the reference implementation.

The reference has no golden vectors for the IMC step (its RNG is Numba's
hidden per-thread MT19937, so no single history can be reproduced).  What
only the reference itself can pin is the composition of its ``step()``:
sourcing, Fleck factor, transport, material update and comb together.  This
script imports the reference modules from /root/reference (Numba is absent in
this container, so they run through their documented pure-Python fallback,
IMC/IMC2D.py:26-46, IMC/IMC1D.py:46-66), runs tiny problems over many seeds
with both ``random`` and ``np.random`` seeded, and stores per-cell ensemble
means and standard errors of T and T_r plus energy totals.

Run from the repository root:  python tests/golden/make_reference_fixtures.py
Outputs: tests/golden/ref_imc2d_xy.npz, ref_imc2d_rz.npz, ref_imc1d_slab.npz
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

# problem definitions shared with tests/test_reference_fixtures.py
XY = dict(nx=6, ny=5, Lx=0.3, Ly=0.25, N=600, Nb=600, Nmax=1200, dt=0.01, steps=3, Tb=0.6, cv=0.3, T0=0.05)
RZ = dict(nx=5, ny=6, R=0.25, Lz=0.3, N=600, Nb=600, Nmax=1200, dt=0.01, steps=3, Tb=0.6, cv=0.3, T0=0.05)
SLAB = dict(I=20, L=0.2, N=1500, Nb=1500, Nmax=3000, dt=0.01, steps=4, Tb=1.0, cv=0.3, T0=0.01)
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
    import IMC1D as ref1d  # noqa: E402
    import IMC2D as ref2d  # noqa: E402

    for name, runner, mod in (("ref_imc2d_xy", run_xy, ref2d), ("ref_imc2d_rz", run_rz, ref2d),
                              ("ref_imc1d_slab", run_slab, ref1d)):
        print(name, flush=True)
        res = ensemble(runner, mod, SEEDS)
        np.savez(os.path.join(HERE, name + ".npz"), **res)


if __name__ == "__main__":
    main()
