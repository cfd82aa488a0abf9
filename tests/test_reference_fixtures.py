"""The oracle's composed IMC step against statistical fixtures written by the
reference implementation itself (tests/golden/make_reference_fixtures.py).

The reference's RNG (Numba MT19937) cannot be reproduced history by history,
so the pin is statistical: per-cell ensemble means of T and T_r and the energy
totals after each step must agree within a few combined standard errors.
This checks the composition of sourcing, Fleck factor, transport, material
update and comb (IMC/IMC2D.py:1591-1897) beyond what the per-function pins
cover.
"""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_reference_fixtures as F  # noqa: E402

SEEDS = 200


def _oracle_ensemble(oracle, geometry):
    p = F.XY if geometry == "xy" else F.RZ
    if geometry == "xy":
        e1 = np.linspace(0, p["Lx"], p["nx"] + 1)
        e2 = np.linspace(0, p["Ly"], p["ny"] + 1)
        s0 = F.sigma0_xy()
        Tb = (p["Tb"], 0.0, 0.0, 0.0)
        refl = (False, False, True, False)
    else:
        e1 = np.linspace(0, p["R"], p["nx"] + 1)
        e2 = np.linspace(0, p["Lz"], p["ny"] + 1)
        s0 = F.sigma0_rz()
        Tb = (0.0, 0.0, p["Tb"], 0.0)
        refl = (False, False, False, False)
    cv = p["cv"]
    sf = lambda T: s0 * np.maximum(T, 1e-3) ** -3
    T0 = np.full((p["nx"], p["ny"]), p["T0"])
    rows = []
    for seed in range(SEEDS):
        orc = oracle.Oracle2D(e1, e2, geometry, seed=100 + seed, nthreads=1)
        st = orc.init_simulation(p["N"], T0, T0, lambda T: cv * T, lambda u: u / cv)
        out = []
        for _ in range(p["steps"]):
            st, info = orc.step(st, p["N"], p["Nb"], 0, p["Nmax"], Tb, p["dt"], sf, lambda u: u / cv,
                                lambda T: cv + 0 * T, np.zeros_like(T0), reflect=refl, rz_linear_source=True)
            out.append((info["temperature"], info["radiation_temperature"], info["total_energy"],
                        info["total_radiation_energy"], info["boundary_loss"], info["N_particles"]))
        rows.append(out)
    res = {}
    for k, name in enumerate(["T", "Tr", "E_total", "E_rad", "boundary_loss", "N"]):
        arr = np.array([[r[s][k] for s in range(p["steps"])] for r in rows], dtype=np.float64)
        res[name + "_mean"] = arr.mean(0)
        res[name + "_sem"] = arr.std(0, ddof=1) / np.sqrt(arr.shape[0])
    return res


def _compare(ref, got, name, tol_frac=0.02, zmax=4.5):
    m1, s1 = ref[name + "_mean"], ref[name + "_sem"]
    m2, s2 = got[name + "_mean"], got[name + "_sem"]
    se = np.sqrt(s1 ** 2 + s2 ** 2)
    ok = se > 1e-12 * np.maximum(np.abs(m1), 1e-300)
    z = np.abs(m1 - m2)[ok] / se[ok]
    # quantities with no spread must agree to rounding
    if (~ok).any():
        np.testing.assert_allclose(m2[~ok], m1[~ok], rtol=1e-9, atol=1e-300)
    return z


@pytest.mark.parametrize("geometry", ["xy", "rz"])
def test_oracle_matches_reference_ensemble(oracle, geometry):
    ref = dict(np.load(os.path.join(HERE, "golden", f"ref_imc2d_{geometry}.npz")))
    got = _oracle_ensemble(oracle, geometry)
    zs = {}
    for name in ("T", "Tr", "E_total", "E_rad", "boundary_loss"):
        zs[name] = _compare(ref, got, name)
    allz = np.concatenate([z.ravel() for z in zs.values()])
    frac3 = np.mean(allz > 3.0)
    print(geometry, {k: float(v.max()) for k, v in zs.items()}, "frac z>3", frac3)
    assert allz.max() < 5.0, {k: float(v.max()) for k, v in zs.items()}
    assert frac3 < 0.02
    # particle counts after the comb: Nmax plus one-per-occupied-cell effects
    assert np.all(np.abs(got["N_mean"] - ref["N_mean"]) < 5 * np.sqrt(ref["N_sem"] ** 2 + got["N_sem"] ** 2) + 2)


def _oracle_slab_ensemble(oracle):
    p = F.SLAB
    I = p["I"]
    edges = np.linspace(0, p["L"], I + 1)
    mesh = np.column_stack([edges[:-1], edges[1:]])
    cv = p["cv"]
    sf = lambda T: 300.0 * np.maximum(T, 1e-3) ** -3
    T0 = np.full(I, p["T0"])
    rows = []
    for seed in range(SEEDS):
        orc = oracle.Oracle1D(mesh, "slab", seed=500 + seed, nthreads=1)
        st = orc.init_simulation(p["N"], T0, T0, lambda T: cv * T, lambda u: u / cv)
        out = []
        for _ in range(p["steps"]):
            st, info = orc.step(st, p["N"], p["Nb"], 0, p["Nmax"], (p["Tb"], 0.0), p["dt"], sf, lambda u: u / cv,
                                lambda T: cv + 0 * T, np.zeros(I), reflect=(False, False))
            out.append((info["temperature"], info["radiation_temperature"], info["total_energy"],
                        info["total_radiation_energy"], info["boundary_loss"], info["N_particles"]))
        rows.append(out)
    res = {}
    for k, name in enumerate(["T", "Tr", "E_total", "E_rad", "boundary_loss", "N"]):
        arr = np.array([[r[s][k] for s in range(p["steps"])] for r in rows], dtype=np.float64)
        res[name + "_mean"] = arr.mean(0)
        res[name + "_sem"] = arr.std(0, ddof=1) / np.sqrt(arr.shape[0])
    return res


def test_oracle_slab_matches_reference_ensemble(oracle):
    ref = dict(np.load(os.path.join(HERE, "golden", "ref_imc1d_slab.npz")))
    got = _oracle_slab_ensemble(oracle)
    zs = {name: _compare(ref, got, name) for name in ("T", "Tr", "E_total", "E_rad", "boundary_loss")}
    allz = np.concatenate([z.ravel() for z in zs.values()])
    print("slab", {k: float(v.max()) for k, v in zs.items()}, "frac z>3", np.mean(allz > 3.0))
    assert allz.max() < 5.0, {k: float(v.max()) for k, v in zs.items()}
    assert np.mean(allz > 3.0) < 0.02
