"""GPU parity of the multigroup 1-D spherical engine (libimc_hip.so through
the C-ABI) against the FP64 oracle (oracle/imc_oracle_mg1d.c) on identical
Philox streams.

FP64: sourcing bit for bit (radii to a pow ulp; Planck groups except where a
device-libm ulp moves a draw across an edge), transport history by history,
the comb bit for bit, composed steps to statistical agreement with every
energy sink booked.  FP32-track: the 0-D closed form to 5e-5 and ensemble
means within a few standard errors of the FP64 engine."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

A = 0.01372
C = 29.98
EDGES = np.array([1e-4, 0.03, 0.1, 0.3, 1.0, 10.0])
G = 5


@pytest.fixture(scope="module")
def L():
    from radtranbook_amd import _lib

    _lib.load()
    _lib.require_gpu()
    return _lib


@pytest.fixture(scope="module")
def M1(oracle):
    from oracle import oracle_mg1d

    return oracle_mg1d


def _announce(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _edges(I=16, R=0.6, r0=0.0):
    return np.linspace(r0, R, I + 1)


def _sigma(T, I):
    rng = np.random.default_rng(3)
    s0 = np.exp(rng.uniform(np.log(2.0), np.log(60.0), I))
    return np.array([s0 * (1.0 + 2.0 * g) * np.maximum(T, 1e-3) ** -1.0 for g in range(G)])


def _cmp(got, ref, what, allow_group=0):
    w, mu, t, x, cell, g = got
    assert w.shape == ref.w.shape, (what, w.shape, ref.w.shape)
    bad = g != ref.g
    assert bad.sum() <= allow_group, (what, int(bad.sum()))
    np.testing.assert_allclose(w, ref.w, rtol=1e-15, atol=0, err_msg=what)
    np.testing.assert_allclose(mu, ref.mu, rtol=0, atol=1e-15, err_msg=what)
    np.testing.assert_allclose(t, ref.t, rtol=1e-15, atol=1e-300, err_msg=what)
    np.testing.assert_allclose(x, ref.x, rtol=2e-15, atol=1e-300, err_msg=what)
    assert np.array_equal(cell, ref.cell), what


def test_mg1d_sourcing_bitwise(L, M1):
    _announce("test_mg1d_sourcing_bitwise")
    e = _edges()
    I = e.size - 1
    eng = L.EngineMG1D(e, EDGES, "fp64", seed=21)
    vol = M1.volumes(e)
    Tr = np.linspace(0.6, 0.1, I)
    # equilibrium census
    eng.equilibrium_sample(20000, Tr)
    E = M1.equilibrium_energy(Tr, vol)
    cnt, tot = M1.counts(E, 20000, 1)
    ref = M1.sample_ic(e, Tr, E, cnt, EDGES, 21)
    _cmp(eng.get_census(), ref, "ic", allow_group=max(2, len(ref) // 5000))
    # faces, fixed source and emission appended in the step's order
    T = np.linspace(0.5, 0.2, I)
    sig = _sigma(T, I)
    eng.set_step(3)
    mats, Et = eng.set_material(T, sig, np.full(I, 0.3), 1.0, 0.01, want=("sa", "ss", "E"))
    sa, ss, bstar, f = M1.OM.fleck_mg(T, sig, np.full(I, 0.3), 1.0, 0.01, EDGES)
    Eo, Ecell = M1.OM.emission_mg(T, sa, bstar, vol, 0.01)
    np.testing.assert_allclose(mats["sa"], sa, rtol=1e-13)
    np.testing.assert_allclose(mats["E"], Eo, rtol=1e-13, atol=1e-300)
    eng.put_census(*[np.zeros(0)] * 4, np.zeros(0, np.int64), np.zeros(0, np.int32))
    E_out = eng.sample_face(1, 7000, 0.8, 0.01)
    assert E_out == pytest.approx(M1.face_energy(0.8, 0.01, e[-1]), rel=1e-15)
    QV = np.zeros((G, I)); QV[0, 3] = 0.2; QV[2, 7] = 0.05
    se = eng.sample_volume(L.SRC_FIXED, 3000, 0.01, QV)
    eng.sample_volume(L.SRC_EMISSION, 9000, 0.01)
    got = eng.get_census()
    face = M1.sample_face(1, 7000, E_out, 0.8, 0.01, e[-1], I, EDGES, 21, 3)
    c_s, _ = M1.counts(QV, 3000, 0)
    src = M1.sample_bins(e, G, QV, c_s, 0.01, 21, 3, M1.TAG_SOURCE)
    c_e, _ = M1.counts(Eo, 9000, 1)
    emi = M1.sample_bins(e, G, Eo, c_e, 0.01, 21, 3, M1.TAG_EMISSION)
    assert se == pytest.approx(float(np.sum(src.w)), rel=1e-14)
    _cmp(got, face.concat(src).concat(emi), "face+source+emission", allow_group=3)


def _transport_case(L, M1, precision, npart=20000, refl=(False, False), r0=0.0):
    e = _edges(12, 0.5, r0)
    I = e.size - 1
    rng = np.random.default_rng(8)
    T = np.exp(rng.uniform(np.log(0.05), np.log(0.6), I))
    sig = _sigma(T, I)
    vol = M1.volumes(e)
    sa, ss, bstar, f = M1.OM.fleck_mg(T, sig, np.full(I, 0.2), 1.0, 0.01, EDGES)
    Eo, _ = M1.OM.emission_mg(T, sa, bstar, vol, 0.01)
    cnt, _ = M1.counts(Eo, npart, 1)
    pool = M1.sample_bins(e, G, Eo, cnt, 0.01, 5, 1, M1.TAG_EMISSION)
    pool = M1.sample_face(1, npart // 2, M1.face_energy(0.7, 0.01, e[-1]), 0.7, 0.01, e[-1], I, EDGES, 5,
                          1).concat(pool)
    eng = L.EngineMG1D(e, EDGES, precision, seed=5)
    eng.set_step(1)
    eng.set_material(T, sig, np.full(I, 0.2), 1.0, 0.01)
    eng.put_census(*pool.arrays())
    return dict(e=e, I=I, sa=sa.reshape(G, I), ss=ss.reshape(G, I), vol=vol, pool=pool, eng=eng)


@pytest.mark.parametrize("r0", [0.0, 0.1])
def test_mg1d_transport_fp64_history_parity(L, M1, r0):
    _announce(f"test_mg1d_transport_fp64_history_parity r0={r0}")
    refl = (True, False) if r0 > 0 else (False, False)
    r = _transport_case(L, M1, "fp64", refl=refl, r0=r0)
    eng, pool = r["eng"], r["pool"]
    floor = 1e-14 * M1.O.canonical_sum(pool.w) / len(pool)
    dep_d, si_d, bl_d, st_d = eng.transport(0.01, refl, floor)
    ref = M1.Pool1MG(*[a.copy() for a in pool.arrays()])
    dep_raw, path_raw, bl, st = M1.transport(ref, r["e"], G, r["sa"], r["ss"], 0.01, refl, floor, 5, 1)
    dep, si = M1.tallies(dep_raw, path_raw, r["sa"], r["vol"], 0.01)
    w, mu, t, x, cell, g = eng.get_census()
    same = (cell == ref.cell) & np.isclose(w, ref.w, rtol=1e-10, atol=0) & \
        np.isclose(x, ref.x, rtol=1e-9, atol=1e-12)
    _announce(f"  identical histories {same.mean():.6f}; events {st_d[0]} vs {st[0]}")
    assert same.mean() > 0.999
    vol = r["vol"][None, :]
    assert np.sum(dep_d * vol) == pytest.approx(np.sum(dep * vol), rel=1e-6)
    np.testing.assert_allclose(bl_d, bl, rtol=1e-6, atol=1e-12 * np.sum(pool.w))
    close = np.isclose(dep_d, dep, rtol=1e-6, atol=1e-9 * np.abs(dep).max())
    assert close.mean() > 0.98
    close_si = np.isclose(si_d, si, rtol=1e-6, atol=1e-9 * np.abs(si).max())
    assert close_si.mean() > 0.98
    # stats: device [ev, bnd, sc, cen, 0, floor, refl, bad]; oracle [ev, bnd, sc, cen, floor, refl, bad]
    np.testing.assert_allclose(st_d[[0, 1, 2, 3, 5, 6]], st[[0, 1, 2, 3, 4, 5]], rtol=2e-3, atol=2)
    assert st_d[7] == 0 and st[6] == 0
    assert np.sum(w) + np.sum(dep_d * vol) + np.sum(bl_d) == pytest.approx(np.sum(pool.w), rel=1e-12)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_mg1d_zero_d_closed_form(L, M1, precision):
    """One reflecting sphere: dep_g = sum w (1 - exp(-sigma_a,g c (dt - t))) / V."""
    _announce(f"test_mg1d_zero_d_closed_form {precision}")
    e = np.array([0.0, 0.4])
    eng = L.EngineMG1D(e, EDGES, precision, seed=2)
    T = np.array([0.3])
    sig = _sigma(T, 1)
    dt = 0.02
    mats, _ = eng.set_material(T, sig, np.array([0.02]), 1.0, dt, want=("sa",))
    eng.set_step(0)
    eng.sample_volume(L.SRC_EMISSION, 8000, dt)
    w0, mu0, t0, x0, c0, g0 = eng.get_census()
    dep, si, bl, st = eng.transport(dt, (False, True), 0.0)
    vol = M1.volumes(e)[0]
    sa = mats["sa"][:, 0]
    exact = np.array([np.sum(np.where(g0 == g, w0 * (1 - np.exp(-sa[g] * C * (dt - t0))), 0.0)) for g in range(G)])
    tol = 1e-10 if precision == "fp64" else 5e-5
    m = exact > 1e-6 * exact.max()
    np.testing.assert_allclose(dep[m, 0] * vol, exact[m], rtol=tol)
    assert np.all(bl == 0.0)
    w1 = eng.get_census()[0]
    assert np.sum(w1) + np.sum(dep) * vol == pytest.approx(np.sum(w0), rel=1e-12)


@pytest.mark.parametrize("radiation", [False, True])
def test_mg1d_comb_parity(L, M1, radiation):
    _announce(f"test_mg1d_comb_parity radiation={radiation}")
    r = _transport_case(L, M1, "fp64", npart=8000)
    ref = r["pool"]
    M1.transport(ref, r["e"], G, r["sa"], r["ss"], 0.01, (False, False), 0.0, 5, 1)
    eng = r["eng"]
    eng.put_census(*ref.arrays())
    eng.set_step(1)
    disc_d, alive, combed = eng.comb(5000, radiation=radiation)
    newp, disc = M1.comb(ref, 5000, r["I"], G, 5, 1, radiation)
    assert combed and alive == int(np.sum(ref.w > 0))
    w, mu, t, x, cell, g = eng.get_census()
    assert len(w) == len(newp)
    np.testing.assert_allclose(w, newp.w, rtol=1e-13 if radiation else 1e-14, atol=0)
    assert np.array_equal(cell, newp.cell) and np.array_equal(g, newp.g)
    np.testing.assert_allclose(mu, newp.mu, rtol=0, atol=0)
    assert np.all(t == 0.0)
    np.testing.assert_allclose(disc_d, disc, rtol=1e-9, atol=1e-12 * np.abs(ref.w).sum())
    # threshold far above: no comb, dead particles dropped
    eng.put_census(*ref.arrays())
    _, alive2, combed2 = eng.comb(-10 ** 9)
    assert not combed2 and eng.n == alive2 == alive


def _run_both(L, M1, precision, nsteps, N, seed, mode=True, fmin=0.0):
    from radtranbook_amd import MG_IMC1D

    MG_IMC1D.VERBOSE = False
    I = 16
    r = np.linspace(0.0, 0.6, I + 1)
    mesh = np.stack([r[:-1], r[1:]], axis=1)
    funcs = [lambda T, g=g: 20.0 * (1 + 2 * g) * np.maximum(T, 1e-3) ** -1.0 for g in range(G)]
    inv = lambda u: u / 0.2
    cv = lambda T: 0.2 + 0 * T
    T0 = np.full(I, 0.05)
    st = MG_IMC1D.init_simulation(N // 2, T0, T0, mesh, EDGES, lambda T: 0.2 * T, inv, precision=precision,
                                  seed=seed)
    orc = M1.OracleMG1D(mesh, EDGES, seed=seed, nthreads=4)
    so = orc.init_simulation(N // 2, T0, T0, lambda T: 0.2 * T, inv)
    a_all, b_all = [], []
    for _ in range(nsteps):
        st, a = MG_IMC1D.step(st, N // 2, N // 2, 0, N, (0.0, 0.8), 0.005, mesh, EDGES, funcs, inv, cv,
                              np.zeros(I), conserve_comb_energy=mode, particle_budget_fmin=fmin)
        so, b = orc.step(so, N // 2, N // 2, 0, N, (0.0, 0.8), 0.005, funcs, inv, cv, np.zeros(I),
                         conserve_comb_energy=mode, particle_budget_fmin=fmin)
        a_all.append(a)
        b_all.append(b)
    return a_all, b_all


@pytest.mark.parametrize("mode", [True, "radiation", False])
def test_mg1d_full_step_fp64_matches_oracle(L, M1, mode):
    _announce(f"test_mg1d_full_step_fp64_matches_oracle mode={mode}")
    a_all, b_all = _run_both(L, M1, "fp64", 3, 20000, 7, mode=mode)
    a, b = a_all[0], b_all[0]
    assert a["boundary_emission"] == pytest.approx(b["boundary_emission"], rel=1e-14)
    np.testing.assert_allclose(a["emitted_by_group"], b["emitted_by_group"], rtol=1e-12, atol=1e-300)
    assert a["profiling"]["transport_events"]["n_particles_transported"] == b["n_transported"]
    close = np.isclose(a["deposited"], b["deposited"], rtol=1e-6, atol=1e-9 * np.abs(b["deposited"]).max())
    assert close.mean() > 0.98
    for a, b in zip(a_all, b_all):
        assert a["total_energy"] == pytest.approx(b["total_energy"], rel=1e-3)
        assert abs(a["balance_residual"]) <= 1e-12 * a["total_energy"]
        if mode is False:
            assert a["energy_residual"] == pytest.approx(-a["comb_discrepancy"].sum(), abs=1e-12 * a["total_energy"])
        assert a["N_particles"] == pytest.approx(b["N_particles"], rel=0.02)


def test_mg1d_budget_split(L, M1):
    _announce("test_mg1d_budget_split")
    from radtranbook_amd import MG_IMC1D

    MG_IMC1D.VERBOSE = False
    a_all, b_all = _run_both(L, M1, "fp64", 2, 12000, 4, fmin=0.1)
    for a, b in zip(a_all, b_all):
        assert a["N_boundary"] == b["budget"][1]
        assert abs(a["balance_residual"]) <= 1e-12 * a["total_energy"]


def test_mg1d_fp32_vs_fp64_ensemble(L, M1):
    """Material temperature after 3 steps: FP32-track ensemble mean within a
    few standard errors of the FP64 ensemble mean."""
    _announce("test_mg1d_fp32_vs_fp64_ensemble")
    res = {}
    for prec in ("fp64", "fp32"):
        Ts = []
        for seed in range(12):
            from radtranbook_amd import MG_IMC1D

            MG_IMC1D.VERBOSE = False
            I = 16
            r = np.linspace(0.0, 0.6, I + 1)
            mesh = np.stack([r[:-1], r[1:]], axis=1)
            # moderately thin groups so the face heats several shells
            funcs = [lambda T, g=g: 3.0 * (1 + g) + 0 * T for g in range(G)]
            inv = lambda u: u / 0.2
            cv = lambda T: 0.2 + 0 * T
            T0 = np.full(I, 0.05)
            st = MG_IMC1D.init_simulation(4000, T0, T0, mesh, EDGES, lambda T: 0.2 * T, inv, precision=prec,
                                          seed=100 + seed)
            for _ in range(3):
                st, info = MG_IMC1D.step(st, 4000, 4000, 0, 8000, (0.0, 1.0), 0.005, mesh, EDGES, funcs, inv, cv,
                                         np.zeros(I))
                assert abs(info["balance_residual"]) <= 1e-12 * info["total_energy"]
            Ts.append(st.temperature.copy())
        res[prec] = np.array(Ts)
    m64, m32 = res["fp64"].mean(0), res["fp32"].mean(0)
    se = np.sqrt(res["fp64"].var(0, ddof=1) / 12 + res["fp32"].var(0, ddof=1) / 12)
    hot = (m64 - 0.05) > 5 * np.sqrt(res["fp64"].var(0, ddof=1) / 12)
    _announce(f"  heated shells {hot.sum()}: {np.round(m64, 4)}")
    assert hot.sum() >= 3
    z = np.abs(m64 - m32)[hot] / (se[hot] + 1e-300)
    _announce(f"  max |z| over {hot.sum()} heated shells: {z.max():.2f}")
    assert z.max() < 4.5
