"""GPU parity of the Carter-Forest gray 2-D engine (libimc_hip.so through the
C-ABI) against the Carter-Forest oracle on identical Philox streams, FP64.

Emission: positions and directions to device-libm ulps, times to 1e-14.
Transport: history by history (> 99.9% identical; a history whose event
decision flips on a libm ulp diverges), totals to 1e-5.  Full step: the
first step's temperatures to 1e-6 and every step's energy balance to 1e-12.
Also the one-cell two-state closed form on the device."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

A = 0.01372
C = 29.98


@pytest.fixture(scope="module")
def L():
    from radtranbook_amd import _lib

    _lib.load()
    _lib.require_gpu()
    return _lib


@pytest.fixture(scope="module")
def CF(oracle):
    from oracle import oracle_cf

    return oracle_cf


def _announce(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _mesh(n=10, m=8):
    return np.linspace(0.0, 1.0, n + 1), np.linspace(0.0, 0.8, m + 1)


def _material(n, m, seed=2):
    rng = np.random.default_rng(seed)
    T = np.exp(rng.uniform(np.log(0.1), np.log(0.8), (n, m)))
    sig = np.exp(rng.uniform(np.log(0.5), np.log(30.0), (n, m)))
    return T, sig, np.full((n, m), 0.05)


def _cmp(got, ref, what, pos_rtol=1e-12, allow_mismatch=0):
    names = ["w", "d1", "d2", "t", "p1", "p2", "ci", "cj"]
    ra = ref.arrays()
    assert got[0].shape == ra[0].shape, (what, got[0].shape, ra[0].shape)
    bad = (got[6] != ra[6]) | (got[7] != ra[7])
    assert bad.sum() <= allow_mismatch, (what, int(bad.sum()))
    ok = ~bad
    for nm, a, b in zip(names, got, ra):
        a, b = a[ok], np.asarray(b)[ok].astype(a.dtype)
        if nm in ("ci", "cj"):
            assert np.array_equal(a, b), (what, nm)
        elif nm in ("d1", "d2"):
            np.testing.assert_allclose(a, b, rtol=0, atol=2e-15, err_msg=f"{what} {nm}")
        else:
            np.testing.assert_allclose(a, b, rtol=pos_rtol, atol=1e-300, err_msg=f"{what} {nm}")


@pytest.mark.parametrize("geometry", ["xy", "rz"])
def test_cf_emission_parity(L, CF, geometry):
    _announce(f"test_cf_emission_parity {geometry}")
    from oracle import oracle as O

    e1, e2 = _mesh()
    n, m = e1.size - 1, e2.size - 1
    T, sig, cv = _material(n, m)
    dt = 0.004
    vol = O.volumes_2d(e1, e2, geometry)
    beta = CF.cf_beta(T, cv)
    rate, E = CF.cf_rates_and_energies(T, sig, beta, vol, dt)
    eng = L.Engine2D(e1, e2, geometry, "fp64", seed=31)
    eng.set_step(3)
    eng.set_material_cf(T, sig, beta)
    eng.sample_volume(L.SRC_CF, 5000, E, dt, True)
    got = eng.get_census()
    cnt, _ = O.counts(E, 5000, False)
    gcode = 0 if geometry == "xy" else 2
    ref = O.sample_volume_2d(e1, e2, E, cnt, 0, gcode, T, dt, 31, 3, O.TAG_EMISSION)
    CF.cf_emission_times(ref, m, rate, dt, 31, 3)
    _cmp(got, ref, f"emission {geometry}")
    assert np.all(got[3] >= 0.0) and np.all(got[3] <= dt * (1 + 1e-12))


def _transport_setup(L, CF, geometry, npart=20000, seed=7):
    from oracle import oracle as O

    e1, e2 = _mesh()
    n, m = e1.size - 1, e2.size - 1
    T, sig, cv = _material(n, m, seed=4)
    dt = 0.005
    vol = O.volumes_2d(e1, e2, geometry)
    beta = CF.cf_beta(T, cv)
    rate, E = CF.cf_rates_and_energies(T, sig, beta, vol, dt)
    cnt, _ = O.counts(E, npart, False)
    gcode = 0 if geometry == "xy" else 1
    pool = O.sample_volume_2d(e1, e2, E, cnt, 0, gcode, T, dt, seed, 1, O.TAG_EMISSION)
    CF.cf_emission_times(pool, m, rate, dt, seed, 1)
    side = 0 if geometry == "xy" else 2
    pool = pool.concat(O.sample_boundary_2d(gcode, side, npart // 2, O.boundary_energy(gcode, side, 0.7, dt, e1, e2),
                                            dt, e1, e2, None, seed, 1))
    eng = L.Engine2D(e1, e2, geometry, "fp64", seed=seed)
    eng.set_step(1)
    eng.set_material_cf(T, sig, beta)
    eng.put_census(*pool.arrays())
    return dict(e1=e1, e2=e2, dt=dt, vol=vol, sig=sig, beta=beta, pool=pool, eng=eng, gcode=gcode)


@pytest.mark.parametrize("geometry", ["xy", "rz"])
def test_cf_transport_history_parity(L, CF, geometry):
    _announce(f"test_cf_transport_history_parity {geometry}")
    from oracle import oracle as O

    r = _transport_setup(L, CF, geometry)
    eng, pool = r["eng"], r["pool"]
    floor = 1e-10 * float(np.sum(pool.w)) / len(pool)
    reflect = (False, True, True, False)
    dep_d, si_d, bl_d, st_d = eng.transport_cf(r["dt"], reflect, 10**6, floor)
    ref = O.Pool2D(*[a.copy() for a in pool.arrays()])
    dep_raw, si_raw, bl, st = CF.transport_2d_cf(ref, r["e1"], r["e2"], r["sig"], r["beta"], r["dt"], reflect, 10**6,
                                                 r["gcode"], floor, 7, 1)
    got = eng.get_census()
    same = (got[6] == ref.ci) & (got[7] == ref.cj) & np.isclose(got[0], ref.w, rtol=1e-12, atol=0) & \
        np.isclose(got[4], ref.p1, rtol=1e-10, atol=1e-12) & np.isclose(got[5], ref.p2, rtol=1e-10, atol=1e-12)
    _announce(f"  identical histories {same.mean():.6f}; events {st_d[0]} vs {st[0]}; captures {st_d[4]} vs {st[4]}")
    assert same.mean() > 0.999
    dep = dep_raw / r["vol"]
    si = si_raw / (r["dt"] * r["vol"])
    vol = r["vol"]
    assert np.sum(dep_d * vol) == pytest.approx(np.sum(dep * vol), rel=1e-5)
    assert np.sum(si_d * vol) == pytest.approx(np.sum(si * vol), rel=1e-5)
    assert bl_d == pytest.approx(bl, rel=1e-5, abs=1e-300)
    assert abs(int(st_d[0]) - int(st[0])) <= 1e-3 * st[0]
    # energy: census + captured + escaped = input, per engine
    assert np.sum(got[0]) + np.sum(dep_d * vol) + bl_d == pytest.approx(np.sum(pool.w), rel=1e-12)


def test_cf_one_cell_closed_form(L):
    """Two-state chain on the device: census fraction in one reflecting cell."""
    _announce("test_cf_one_cell_closed_form")
    e = np.array([0.0, 1.0])
    eng = L.Engine2D(e, e, "xy", "fp64", seed=13)
    n, dt, sigma, beta = 400_000, 0.01, 5.0, 0.5
    rng = np.random.default_rng(5)
    mu = rng.uniform(-1, 1, n)
    phi = rng.uniform(0, 2 * np.pi, n)
    z = np.zeros(n)
    eng.put_census(np.ones(n), np.sqrt(1 - mu * mu) * np.cos(phi), np.sqrt(1 - mu * mu) * np.sin(phi), z, z + 0.5,
                   z + 0.5, np.zeros(n, np.int64), np.zeros(n, np.int64))
    eng.set_step(0)
    eng.set_material_cf(np.array([[0.5]]), np.array([[sigma]]), np.array([[beta]]))
    dep, si, bl, st = eng.transport_cf(dt, (True,) * 4, 10**7, 0.0)
    w = eng.get_census()[0]
    l1, l2 = C * sigma, C * sigma * beta
    s = l1 + l2
    p_rad = l2 / s + l1 / s * np.exp(-s * dt)
    se = np.sqrt(p_rad * (1 - p_rad) / n)
    assert abs(w.sum() / n - p_rad) < 4 * se
    assert w.sum() + dep.sum() == pytest.approx(n, rel=1e-12)
    path = C * (l2 / s * dt + l1 / s**2 * (1 - np.exp(-s * dt)))
    assert si.sum() * dt / n == pytest.approx(path, rel=5e-3)


def _run_both(L, CF, nsteps, N, seed, geometry, **kw):
    from radtranbook_amd import IMC2D_CarterForest as CFM

    CFM.VERBOSE = False
    e1, e2 = _mesh()
    n, m = e1.size - 1, e2.size - 1
    rng = np.random.default_rng(1)
    sig0 = np.exp(rng.uniform(np.log(0.5), np.log(20.0), (n, m)))
    sf = lambda T: sig0 * np.maximum(T, 1e-3) ** -1.0
    inv = lambda u: u / 0.05
    cvf = lambda T: 0.05 + 0 * T
    T0 = np.full((n, m), 0.2)
    Tb = (0.6, 0, 0, 0) if geometry == "xy" else (0, 0, 0.6, 0)
    st = CFM.init_simulation(N // 2, T0, T0, e1, e2, lambda T: 0.05 * T, inv, geometry=geometry, seed=seed)
    orc = CF.OracleCF2D(e1, e2, geometry, seed=seed)
    so = orc.init_simulation(N // 2, T0, T0, lambda T: 0.05 * T, inv)
    a_all, b_all = [], []
    for _ in range(nsteps):
        st, a = CFM.step(st, N // 2, N // 2, 0, N, Tb, 0.003, e1, e2, sf, inv, cvf, np.zeros((n, m)),
                         geometry=geometry, **kw)
        so, b = orc.step(so, N // 2, N // 2, 0, N, Tb, 0.003, sf, inv, cvf, np.zeros((n, m)), **kw)
        a_all.append(a)
        b_all.append(b)
    return a_all, b_all


@pytest.mark.parametrize("geometry", ["xy", "rz"])
def test_cf_full_step_matches_oracle(L, CF, geometry):
    _announce(f"test_cf_full_step_matches_oracle {geometry}")
    a_all, b_all = _run_both(L, CF, 3, 16000, 11, geometry)
    a, b = a_all[0], b_all[0]
    np.testing.assert_allclose(a["emitted_energies"], b["emitted_energies"], rtol=1e-13, atol=1e-300)
    assert a["boundary_emission"] == pytest.approx(b["boundary_emission"], rel=1e-13)
    close = np.isclose(a["temperature"], b["temperature"], rtol=1e-6, atol=0)
    assert close.mean() > 0.98
    for a, b in zip(a_all, b_all):
        assert a["total_energy"] == pytest.approx(b["total_energy"], rel=1e-3)
        assert abs(a["energy_residual"]) <= 1e-12 * a["total_energy"]
        assert a["profiling"]["transport_events"]["absorption_capture_events"] > 0


def test_cf_rejects_fp32(L):
    from radtranbook_amd import IMC2D, IMC2D_CarterForest as CFM

    IMC2D.VERBOSE = False
    e = np.linspace(0, 1, 5)
    T0 = np.full((4, 4), 0.1)
    st = IMC2D.init_simulation(100, T0, T0, e, e, lambda T: 0.1 * T, lambda u: u / 0.1, precision="fp32")
    with pytest.raises(ValueError, match="FP64"):
        CFM.step(st, 100, 100, 0, 200, (0.5, 0, 0, 0), 0.01, e, e, lambda T: 1.0 + 0 * T, lambda u: u / 0.1,
                 lambda T: 0.1 + 0 * T, np.zeros((4, 4)))
