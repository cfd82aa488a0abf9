"""GPU parity of the Carter-Forest gray 1-D engine (slab and sphere,
libimc_hip.so through the C-ABI) against the 1-D Carter-Forest oracle on
identical Philox streams, FP64, two problems per batch (index base p << 32).

Emission: counts, weights, cells and directions bit for bit, positions to
libm ulps, times to 1e-14.  Transport: history by history (> 99.9%
identical; a history whose event decision flips on a libm ulp diverges),
tallies and leaks to 1e-9.  The one-cell two-state closed form on the
device.  Full steps through the drop-in module against the oracle step:
emission energies to 1e-13, first-step temperatures to 1e-6, every step's
energy balance to 1e-12."""
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
    from oracle import oracle_cf1d

    return oracle_cf1d


def _announce(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _meshes(geometry):
    out = []
    for k, n in enumerate((24, 50)):
        inner = 0.05 if (geometry != "slab" and k == 1) else 0.0
        e = np.linspace(inner, 0.3 + 0.2 * k, n + 1)
        e[1:-1] += 0.2 * (e[1] - e[0]) * np.sin(np.arange(1, n))
        out.append(e)
    return out


def _material(meshes, seed=2):
    rng = np.random.default_rng(seed)
    T = [np.exp(rng.uniform(np.log(0.1), np.log(0.8), e.size - 1)) for e in meshes]
    sig = [np.exp(rng.uniform(np.log(0.5), np.log(60.0), e.size - 1)) for e in meshes]
    beta = [4 * A * t ** 3 / 0.05 for t in T]
    return T, sig, beta


@pytest.mark.parametrize("geometry", ["slab", "spherical"])
def test_cf1d_emission_parity(L, oracle, CF, geometry):
    _announce(f"test_cf1d_emission_parity {geometry}")
    sph = geometry != "slab"
    meshes = _meshes(geometry)
    T, sig, beta = _material(meshes)
    dts = np.array([0.004, 0.002])
    N = np.array([3000, 5000])
    Ec = []
    for p, e in enumerate(meshes):
        _, E = CF.cf1_emission(T[p], sig[p], beta[p], oracle.volumes_1d(e, geometry), dts[p])
        Ec.append(E)
    eng = L.Engine1D(meshes, geometry, "fp64", seed=31)
    eng.set_step(3)
    eng.set_material_cf(np.concatenate(T), np.concatenate(sig), np.concatenate(beta), dts)
    eng.sample_volume(L.SRC_CF, N, np.concatenate(Ec))
    per, n = eng.stage_pool()
    w, mu, t, x, cell, prob = eng.get_pool(n)
    for p in range(2):
        cnt = CF.cf1_counts(Ec[p], int(N[p]), sph)
        ref = oracle.sample_volume_1d(sph, meshes[p], Ec[p], cnt, 0, T[p], dts[p], 31, 3, oracle.TAG_EMISSION, p << 32)
        rate = C * sig[p] * beta[p]
        CF.cf1_emission_times(ref, rate, dts[p], 31, 3, p << 32)
        m = prob == p
        assert per[p] == len(ref) == m.sum()
        assert np.array_equal(w[m], ref.w) and np.array_equal(cell[m], ref.cell) and np.array_equal(mu[m], ref.mu)
        np.testing.assert_allclose(x[m], ref.x, rtol=1e-15 if not sph else 4.5e-16, atol=1e-300)
        np.testing.assert_allclose(t[m], ref.t, rtol=1e-14, atol=1e-300)
        assert np.all((t[m] >= 0) & (t[m] <= dts[p] * (1 + 1e-12)))


def _pools(oracle, CF, meshes, geometry, T, sig, beta, dts, seed, npart):
    sph = geometry != "slab"
    pools = []
    for p, e in enumerate(meshes):
        rate, Ec = CF.cf1_emission(T[p], sig[p], beta[p], oracle.volumes_1d(e, geometry), dts[p])
        cnt = CF.cf1_counts(Ec, npart, sph)
        pool = oracle.sample_volume_1d(sph, e, Ec, cnt, 0, T[p], dts[p], seed, 1, oracle.TAG_EMISSION, p << 32)
        CF.cf1_emission_times(pool, rate, dts[p], seed, 1, p << 32)
        side = 0 if not sph else 1
        Eb = oracle._lib1d().o1_boundary_energy(int(sph), 0.7, dts[p], e[0] if side == 0 else e[-1])
        pool = pool.concat(oracle.sample_boundary_1d(sph, side, npart // 2, Eb, dts[p], e, seed, 1, p << 32))
        pools.append(pool)
    return pools


@pytest.mark.parametrize("geometry", ["slab", "spherical"])
def test_cf1d_transport_history_parity(L, oracle, CF, geometry):
    _announce(f"test_cf1d_transport_history_parity {geometry}")
    sph = geometry != "slab"
    meshes = _meshes(geometry)
    T, sig, beta = _material(meshes, seed=4)
    dts = np.array([0.005, 0.003])
    pools = _pools(oracle, CF, meshes, geometry, T, sig, beta, dts, 7, 20000)
    eng = L.Engine1D(meshes, geometry, "fp64", seed=7)
    eng.set_step(1)
    cnt = np.array([len(q) for q in pools])
    eng.put_census(cnt, *[np.concatenate([q.arrays()[k] for q in pools]) for k in range(5)])
    eng.set_material_cf(np.concatenate(T), np.concatenate(sig), np.concatenate(beta), dts)
    floors = np.array([1e-10 * oracle.canonical_sum(q.w) / len(q) for q in pools])
    reflect = np.array([[True, False], [False, True]])
    dep_d, si_d, lk_d, st_d, ntr = eng.transport_cf(reflect, floors)
    assert ntr == cnt.sum()
    w, mu, t, x, cell, prob = eng.get_pool(ntr)
    st_ref = np.zeros(8, np.int64)
    for p in range(2):
        ref = oracle.Pool1D(*[a.copy() for a in pools[p].arrays()])
        dep, path, lk, st = CF.transport_1d_cf(sph, ref, meshes[p], sig[p], beta[p], dts[p], reflect[p], floors[p],
                                               0, 0.0, 7, 1, p << 32, nthreads=4)
        st_ref += st
        vol = oracle.volumes_1d(meshes[p], geometry)
        np.testing.assert_allclose(eng.split(dep_d)[p], dep / vol, rtol=1e-9, atol=1e-12 * np.abs(dep / vol).max())
        np.testing.assert_allclose(eng.split(si_d)[p], path / vol / dts[p], rtol=1e-9,
                                   atol=1e-12 * np.abs(path / vol).max() / dts[p])
        assert lk_d[p] == pytest.approx(lk, rel=1e-9, abs=1e-15)
        m = prob == p
        same = (cell[m] == ref.cell) & np.isclose(w[m], ref.w, rtol=1e-12, atol=0) & \
            np.isclose(x[m], ref.x, rtol=1e-10, atol=1e-12) & np.isclose(t[m], ref.t, rtol=1e-10, atol=1e-15)
        _announce(f"  p={p} identical histories {same.mean():.6f}")
        assert same.mean() > 0.999
        # energy: census + captured + escaped = input
        assert w[m].sum() + dep.sum() + lk == pytest.approx(pools[p].w.sum(), rel=1e-12)
    np.testing.assert_allclose(st_d, st_ref, rtol=1e-3)
    assert st_d[2] > 0 and st_d[4] > 0


@pytest.mark.parametrize("geometry", ["slab", "spherical"])
def test_cf1d_one_cell_closed_form(L, geometry):
    """Two-state chain on the device: census fraction in one reflecting cell."""
    _announce(f"test_cf1d_one_cell_closed_form {geometry}")
    eng = L.Engine1D([np.array([0.0, 1.0])], geometry, "fp64", seed=13)
    n, dt, sigma, beta = 400_000, 0.01, 5.0, 0.5
    mu = np.random.default_rng(5).uniform(-1, 1, n)
    z = np.zeros(n)
    eng.put_census([n], np.ones(n), mu, z, z + 0.5, np.zeros(n, np.int64))
    eng.set_step(0)
    eng.set_material_cf(np.array([0.5]), np.array([sigma]), np.array([beta]), dt)
    dep, si, lk, st, ntr = eng.transport_cf((True, True), 0.0)
    w = eng.get_pool(ntr)[0]
    l1, l2 = C * sigma, C * sigma * beta
    s = l1 + l2
    p_rad = l2 / s + l1 / s * np.exp(-s * dt)
    assert abs(w.sum() / n - p_rad) < 4 * np.sqrt(p_rad * (1 - p_rad) / n)
    vol = 1.0 if geometry == "slab" else 4.0 / 3.0 * np.pi
    assert w.sum() + dep.sum() * vol == pytest.approx(n, rel=1e-12)
    path = C * (l2 / s * dt + l1 / s ** 2 * (1 - np.exp(-s * dt)))
    assert si.sum() * vol * dt / n == pytest.approx(path, rel=5e-3)


def _run_both(L, CF, nsteps, N, seed, geometry):
    from radtranbook_amd import IMC1D, IMC1D_CarterForest as CFM

    CFM.VERBOSE = False
    IMC1D.VERBOSE = False
    I = 30
    e = np.linspace(0.0, 0.4, I + 1)
    mesh = np.stack([e[:-1], e[1:]], axis=1)
    sig0 = np.exp(np.random.default_rng(1).uniform(np.log(0.5), np.log(20.0), I))
    sf = lambda T: sig0 * np.maximum(T, 1e-3) ** -1.0
    inv = lambda u: u / 0.05
    cvf = lambda T: 0.05 + 0 * T
    T0 = np.full(I, 0.2)
    Tb = (0.6, 0.0) if geometry == "slab" else (0.0, 0.6)
    st = CFM.init_simulation(N // 2, T0, T0, mesh, lambda T: 0.05 * T, inv, geometry=geometry, seed=seed)
    orc = CF.OracleCF1D(mesh, geometry, seed=seed)
    so = orc.init_simulation(N // 2, T0, T0, lambda T: 0.05 * T, inv)
    a_all, b_all = [], []
    for _ in range(nsteps):
        st, a = CFM.step(st, N // 2, N // 2, 0, N, Tb, 0.003, mesh, sf, inv, cvf, np.zeros(I), geometry=geometry)
        so, b = orc.step(so, N // 2, N // 2, 0, N, Tb, 0.003, sf, inv, cvf, np.zeros(I))
        a_all.append(a)
        b_all.append(b)
    return a_all, b_all


@pytest.mark.parametrize("geometry", ["slab", "spherical"])
def test_cf1d_full_step_matches_oracle(L, CF, geometry):
    _announce(f"test_cf1d_full_step_matches_oracle {geometry}")
    a_all, b_all = _run_both(L, CF, 3, 16000, 11, geometry)
    a, b = a_all[0], b_all[0]
    np.testing.assert_allclose(a["emitted_energies"], b["emitted_energies"], rtol=1e-13, atol=1e-300)
    assert a["boundary_emission"] == pytest.approx(b["boundary_emission"], rel=1e-13)
    close = np.isclose(a["temperature"], b["temperature"], rtol=1e-6, atol=0)
    assert close.mean() > 0.9
    for a, b in zip(a_all, b_all):
        assert a["total_energy"] == pytest.approx(b["total_energy"], rel=1e-3)
        assert abs(a["energy_residual"]) <= 1e-12 * a["total_energy"]
        assert a["profiling"]["transport_events"]["absorption_capture_events"] > 0


def test_cf1d_rejects_fp32(L):
    from radtranbook_amd import IMC1D, IMC1D_CarterForest as CFM

    IMC1D.VERBOSE = False
    e = np.linspace(0, 1, 5)
    mesh = np.stack([e[:-1], e[1:]], axis=1)
    T0 = np.full(4, 0.1)
    st = IMC1D.init_simulation(100, T0, T0, mesh, lambda T: 0.1 * T, lambda u: u / 0.1, precision="fp32")
    with pytest.raises(ValueError, match="FP64"):
        CFM.step(st, 100, 100, 0, 200, (0.5, 0.0), 0.01, mesh, lambda T: 1.0 + 0 * T, lambda u: u / 0.1,
                 lambda T: 0.1 + 0 * T, np.zeros(4))
