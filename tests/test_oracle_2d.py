"""Pins for the gray 2-D oracle (oracle/imc_oracle.c) against what the
mathematics fixes, independent of the oracle's own formulas:

* Fleck factor closed form for sigma = sigma0 T^-3 (IMC/IMC2D.py:1646-1651)
* boundary-source energy a c T^4/4 A dt (IMC/_check_bc_energy.py:24-49)
* ceil particle allocation and exact per-cell energy sums (IMC/IMC2D.py:1191-1238)
* 0-D reflecting box: every history travels exactly c (dt - t_i), so the
  deposit is sum_i w_i (1 - exp(-sigma_a c (dt - t_i))) / V whatever the
  scattering does (IMC/IMC2D.py:938-1107)
* straight-line pure absorber: exit point and attenuation by hand
* rz chord lengths against a brute-force 3-D ray march (IMC/IMC2D.py:711-736)
* comb invariants (IMC/IMC2D.py:1448-1511)
* step energy balance closes against the comb discrepancy (IMC/IMC2D.py:1833-1842)
"""
import numpy as np
import pytest

A = 0.01372
C = 29.98


def test_fleck_closed_form(oracle):
    rng = np.random.default_rng(0)
    T = np.exp(rng.uniform(np.log(1e-3), np.log(2.0), 500))
    sig0 = np.exp(rng.normal(2.3, 1.5, 500))
    cv = 0.3
    dt = 0.005
    sa, ss, f = oracle.fleck(T, sig0 * T ** -3, np.full(500, cv), 1.0, dt)
    f_exact = 1.0 / (1.0 + 4 * A * sig0 * C * dt / cv)
    np.testing.assert_allclose(f, f_exact, rtol=1e-13)
    np.testing.assert_allclose(sa + ss, sig0 * T ** -3, rtol=1e-14)
    # theta = 0: f = 1, no effective scattering
    sa0, ss0, f0 = oracle.fleck(T, sig0, np.full(500, cv), 0.0, dt)
    assert np.all(f0 == 1.0) and np.all(ss0 == 0.0)


@pytest.mark.parametrize("geometry,side", [(0, 0), (0, 2), (1, 1), (1, 2), (1, 3)])
def test_boundary_energy_and_distribution(oracle, geometry, side):
    e1 = np.linspace(0.0 if geometry == 1 else -0.3, 0.5, 11)
    e2 = np.linspace(0.0, 1.0, 17)
    Tb, dt = 0.7, 0.01
    E = oracle.boundary_energy(geometry, side, Tb, dt, e1, e2)
    if geometry == 0:
        area = (e2[-1] - e2[0]) if side < 2 else (e1[-1] - e1[0])
    elif side < 2:
        area = 2 * np.pi * (e1[0] if side == 0 else e1[-1]) * (e2[-1] - e2[0])
    else:
        area = np.pi * (e1[-1] ** 2 - e1[0] ** 2)
    assert E == pytest.approx(A * C * Tb ** 4 / 4 * area * dt, rel=1e-14)
    n = 20000
    p = oracle.sample_boundary_2d(geometry, side, n, E, dt, e1, e2, None, 5, 0)
    assert p.w.sum() == pytest.approx(E, rel=1e-12)
    assert np.all((p.t >= 0) & (p.t < dt))
    # cells agree with the positions
    assert np.all(e1[p.ci] <= p.p1) and np.all(p.p1 <= e1[p.ci + 1])
    assert np.all(e2[p.cj] <= p.p2) and np.all(p.p2 <= e2[p.cj + 1])
    # Lambertian: normal cosine mu_n = sqrt(xi) -> E[mu_n] = 2/3
    if geometry == 0:
        mun = np.abs(p.d1) if side < 2 else np.abs(p.d2)
    else:
        mun = np.abs(p.d1) if side < 2 else np.abs(p.d2)
    assert abs(mun.mean() - 2 / 3) < 4 * np.sqrt(1 / 18 / n)
    # inward pointing
    sgn = 1 if side in (0, 2) else -1
    normal = p.d1 if side < 2 else p.d2
    assert np.all(sgn * normal >= 0)


def test_boundary_uniform_position_rz_zmin(oracle):
    # zmin face: r uniform in area -> r^2 uniform on [r0^2, r1^2]
    e1 = np.linspace(0.1, 0.6, 6)
    e2 = np.linspace(0, 1, 5)
    n = 40000
    p = oracle.sample_boundary_2d(1, 2, n, 1.0, 0.01, e1, e2, None, 11, 2)
    u = (p.p1 ** 2 - e1[0] ** 2) / (e1[-1] ** 2 - e1[0] ** 2)
    hist, _ = np.histogram(u, bins=20, range=(0, 1))
    chi2 = np.sum((hist - n / 20) ** 2 / (n / 20))
    assert chi2 < 50  # 19 dof, p ~ 1e-4


def test_counts_ceil_rule(oracle):
    rng = np.random.default_rng(2)
    E = rng.exponential(1.0, 1000)
    E[::7] = 0.0
    cnt, tot = oracle.counts(E, 5000, False)
    exp = np.ceil(5000 * E / oracle.canonical_sum(E)).astype(np.int64)
    assert np.array_equal(cnt, exp) and tot == exp.sum()
    assert np.all(cnt[E > 0] >= 1) and np.all(cnt[E == 0] == 0)
    cnt1, _ = oracle.counts(np.zeros(10), 50, True)
    assert np.all(cnt1 == 1)


def test_canonical_sum(oracle):
    rng = np.random.default_rng(4)
    for n in (0, 1, 1023, 1024, 1025, 300000):
        x = rng.standard_normal(n)
        assert oracle.canonical_sum(x) == pytest.approx(float(np.sum(x)), abs=1e-10 * max(1, n))
    x = rng.standard_normal(5000)
    chunks = [sum(x[i:i + 1024].tolist()) for i in range(0, 5000, 1024)]
    s = 0.0
    for c in chunks:
        s += c
    assert oracle.canonical_sum(x) == s


@pytest.mark.parametrize("geometry", ["xy", "rz"])
def test_volume_source_energy_and_cells(oracle, geometry):
    e1 = np.linspace(0.0, 0.4, 9)
    e2 = np.linspace(0.0, 0.3, 7)
    rng = np.random.default_rng(7)
    QV = rng.exponential(1.0, (8, 6))
    cnt, tot = oracle.counts(QV, 3000, False)
    gcode = 0 if geometry == "xy" else 2
    T = rng.uniform(0.1, 1.0, (8, 6))
    p = oracle.sample_volume_2d(e1, e2, QV, cnt, 0, gcode, T, 0.01, 3, 1, oracle.TAG_EMISSION)
    sums = np.bincount(p.ci * 6 + p.cj, weights=p.w, minlength=48)
    np.testing.assert_allclose(sums, QV.ravel(), rtol=1e-13)
    assert np.all(e1[p.ci] <= p.p1) and np.all(p.p1 <= e1[p.ci + 1])
    assert np.all(e2[p.cj] <= p.p2) and np.all(p.p2 <= e2[p.cj + 1])
    assert np.all((p.t >= 0) & (p.t < 0.01))


def test_rz_linear_source_distribution(oracle):
    """r density in one cell proportional to r (a + M (r - r0)) (Box 9.4),
    checked by chi^2 against the analytic pdf; z tilt by its branch mixture."""
    e1 = np.array([0.0, 0.2, 0.4, 0.6])
    e2 = np.array([0.0, 0.5, 1.0])
    T = np.array([[0.6, 0.9], [0.7, 0.5], [0.8, 0.2]])
    QV = np.zeros((3, 2)); QV[1, 0] = 1.0
    cnt = np.zeros(6, dtype=np.int64); cnt[2] = 60000
    p = oracle.sample_volume_2d(e1, e2, QV, cnt, 0, 2, T, 0.01, 9, 0, oracle.TAG_EMISSION)
    T4 = T ** 4
    rc = 0.5 * (e1[:-1] + e1[1:])
    M = (T4[2, 0] - T4[0, 0]) / (rc[2] - rc[0])
    a = T4[1, 0] - M * (rc[1] - e1[1])
    r0, r1 = e1[1], e1[2]
    assert a > 0 and a + M * (r1 - r0) > 0  # no slope clamp in this cell
    edges = np.linspace(r0, r1, 21)
    F = lambda r: a * 0.5 * (r ** 2 - r0 ** 2) + M / 3 * (r ** 3 - r0 ** 3)
    prob = np.diff(F(edges)) / F(r1)
    hist, _ = np.histogram(p.p1, bins=edges)
    chi2 = np.sum((hist - prob * p.p1.size) ** 2 / (prob * p.p1.size))
    assert chi2 < 50
    # z: mixture of the two sqrt branches with weight Pn toward z1
    zc = 0.5 * (e2[:-1] + e2[1:])
    Mz = (T4[1, 1] - T4[1, 0]) / (zc[1] - zc[0])
    az = T4[1, 0] - Mz * (zc[0] - e2[0])
    t0, t1 = max(az, 0), max(az + Mz * 0.5, 0)
    Pn = t0 / (t0 + t1)
    zz = (p.p2 - e2[0]) / 0.5
    # E[z/dz] = Pn * 1/3 + (1-Pn) * 2/3
    assert abs(zz.mean() - (Pn / 3 + (1 - Pn) * 2 / 3)) < 5 * np.sqrt(1 / 18 / zz.size)


def _zero_d(oracle, geometry, theta, n=4000, seed=3):
    """Single cell, all faces reflecting."""
    e1 = np.array([0.0 if geometry == 1 else 0.2, 0.7])
    e2 = np.array([-0.1, 0.4])
    vol = oracle.volumes_2d(e1, e2, "xy" if geometry == 0 else "rz")
    T = np.array([[0.4]])
    sig = np.array([[50.0]])
    dt = 0.02
    sa, ss, f = oracle.fleck(T, sig, np.array([[0.01]]), theta, dt)
    QV = np.array([[2.5]])
    cnt = np.array([n], dtype=np.int64)
    p = oracle.sample_volume_2d(e1, e2, QV, cnt, 0, 0 if geometry == 0 else 1, T, dt, seed, 0, oracle.TAG_EMISSION)
    t0 = p.t.copy(); w0 = p.w.copy()
    dep, si, bl, st = oracle.transport_2d(p, e1, e2, sa.reshape(1, 1), ss.reshape(1, 1), vol, dt,
                                          (True, True, True, True), 10**9, geometry, 0.0, seed, 0)
    return dict(p=p, t0=t0, w0=w0, dep=dep, si=si, bl=bl, st=st, sa=sa[0], vol=vol[0, 0], dt=dt, f=f[0])


@pytest.mark.parametrize("geometry", [0, 1])
@pytest.mark.parametrize("theta", [0.0, 1.0])
def test_zero_d_deposit_closed_form(oracle, geometry, theta):
    r = _zero_d(oracle, geometry, theta)
    path = C * (r["dt"] - r["t0"])
    dep_exact = np.sum(r["w0"] * (1 - np.exp(-r["sa"] * path))) / r["vol"]
    si_exact = np.sum(r["w0"] * (1 - np.exp(-r["sa"] * path)) / r["sa"]) / (r["dt"] * r["vol"])
    assert r["dep"][0, 0] == pytest.approx(dep_exact, rel=1e-10)
    assert r["si"][0, 0] == pytest.approx(si_exact, rel=1e-10)
    np.testing.assert_allclose(r["p"].w, r["w0"] * np.exp(-r["sa"] * path), rtol=1e-9)
    assert r["bl"] == 0.0
    st = r["st"]
    assert st[3] == len(r["w0"])  # every history ends at census
    if theta == 0.0:
        assert st[2] == 0
    else:
        assert st[2] > 0
    assert st[6] > 0 and st[6] <= st[1]


def test_straight_line_absorber_xy(oracle):
    e1 = np.linspace(0, 1, 6)
    e2 = np.linspace(0, 1, 4)
    nx, ny = 5, 3
    sa = np.full((nx, ny), 2.0)
    ss = np.zeros((nx, ny))
    vol = oracle.volumes_2d(e1, e2, "xy")
    pool = oracle.Pool2D.empty(3)
    ux, uy = 0.6, 0.3  # not unit: 2-D projection of a 3-D direction
    pool.w[:] = [1.0, 2.0, 1.0]
    pool.d1[:] = [ux, -1.0, 0.0]
    pool.d2[:] = [uy, 0.0, 0.0]
    pool.t[:] = [0.0, 0.0, 0.0]
    pool.p1[:] = [0.1, 0.95, 0.5]
    pool.p2[:] = [0.2, 0.5, 0.5]
    pool.ci[:] = [0, 4, 2]
    pool.cj[:] = [0, 1, 1]
    dt = 1.0 / C  # census distance 1 cm
    dep, si, bl, st = oracle.transport_2d(pool, e1, e2, sa, ss, vol, dt, (False,) * 4, 10**6, 0, 0.0, 1, 0)
    # particle 0 exits at x=1 after s = 0.9/0.6 = 1.5 > 1 -> census at s=1
    s0 = 1.0
    # particle 1 exits at x=0 after 0.95
    s1 = 0.95
    # particle 2 does not move (zero direction): census at s = 1 in place
    s2 = 1.0
    assert pool.w[0] == pytest.approx(np.exp(-2 * s0), rel=1e-12)
    assert pool.p1[0] == pytest.approx(0.1 + ux * s0, rel=1e-12)
    assert pool.p2[0] == pytest.approx(0.2 + uy * s0, rel=1e-12)
    assert pool.w[1] == 0.0
    assert bl == pytest.approx(2.0 * np.exp(-2 * s1), rel=1e-12)
    assert pool.w[2] == pytest.approx(np.exp(-2 * s2), rel=1e-12)
    dep_tot = np.sum(dep * vol)
    assert dep_tot == pytest.approx((1 - np.exp(-2 * s0)) + 2 * (1 - np.exp(-2 * s1)) + (1 - np.exp(-2 * s2)),
                                    rel=1e-12)
    # path-length estimator with sigma_a -> 0 limit: si*dt*V = sum w s
    sa0 = np.zeros((nx, ny))
    pool2 = oracle.Pool2D.empty(1)
    pool2.w[:] = 1.0; pool2.d1[:] = 1.0; pool2.p1[:] = 0.05; pool2.p2[:] = 0.5; pool2.cj[:] = 1
    dep2, si2, bl2, _ = oracle.transport_2d(pool2, e1, e2, sa0, ss, vol, dt, (False,) * 4, 10**6, 0, 0.0, 1, 0)
    np.testing.assert_allclose(np.sum(si2 * dt * vol), 0.95, rtol=1e-12)
    assert bl2 == 1.0


def _ray_exit_bruteforce(r, z, mu, eta, R, zlo, zhi, smax, ds=2e-6):
    """March the 3-D ray; return path length to leave r<=R, zlo<=z<=zhi."""
    st = np.sqrt(max(0.0, 1 - eta * eta))
    ox, oy, oz = st * mu, st * np.sqrt(max(0.0, 1 - mu * mu)), eta
    s = np.arange(0, smax, ds)
    x = r + ox * s; y = oy * s; zz = z + oz * s
    out = (x * x + y * y > R * R) | (zz < zlo) | (zz > zhi)
    idx = np.argmax(out) if np.any(out) else len(s) - 1
    return s[idx]


def test_rz_chord_bruteforce(oracle):
    """Single particle, vacuum, sigma=0, two radial cells: the total path to
    leak must match a brute-force 3-D ray march."""
    e1 = np.array([0.0, 0.25, 0.5])
    e2 = np.array([0.0, 1.0])
    vol = oracle.volumes_2d(e1, e2, "rz")
    rng = np.random.default_rng(12)
    zeros = np.zeros((2, 1))
    for _ in range(25):
        r = rng.uniform(0.01, 0.49); z = rng.uniform(0.1, 0.9)
        mu = rng.uniform(-1, 1); eta = rng.uniform(-0.9, 0.9)
        pool = oracle.Pool2D.empty(1)
        pool.w[:] = 1.0; pool.d1[:] = mu; pool.d2[:] = eta; pool.p1[:] = r; pool.p2[:] = z
        pool.ci[:] = 0 if r < 0.25 else 1
        dt = 10.0 / C
        dep, si, bl, st = oracle.transport_2d(pool, e1, e2, zeros, zeros, vol, dt, (False,) * 4, 10**6, 1, 0.0, 1, 0)
        # sigma_a = 0: si*dt*V sums w * s over the path
        s_tot = np.sum(si * dt * vol)
        s_bf = _ray_exit_bruteforce(r, z, mu, eta, 0.5, 0.0, 1.0, 3.0)
        assert bl == 1.0
        assert s_tot == pytest.approx(s_bf, abs=5e-6)


def test_comb_invariants(oracle):
    rng = np.random.default_rng(5)
    n, ncell = 3000, 40
    w = rng.exponential(1.0, n)
    w[::11] = 0.0
    cell = rng.integers(0, ncell, n)
    Nmax = 1000
    reps = []
    for s in range(40):
        ecen, ew, cp, tot, pt = oracle.comb_counts(w, cell, ncell, Nmax, 7, s)
        assert not pt
        np.testing.assert_allclose(ecen, np.bincount(cell, weights=np.where(w > 0, w, 0), minlength=ncell),
                                   rtol=1e-13)
        E = ecen.sum()
        desired = np.maximum(1, np.rint(Nmax * ecen / E))
        np.testing.assert_allclose(ew, ecen / desired, rtol=1e-14)
        assert np.all(cp[w == 0] == 0)
        frac = w / ew[cell]
        # floor(w/ew + u) is floor or ceil of w/ew
        assert np.all((cp == np.floor(frac)) | (cp == np.floor(frac) + 1))
        reps.append(cp)
    reps = np.array(reps)
    # unbiased: E[copies] = w / ew
    m = reps.mean(axis=0)
    assert np.all(np.abs(m - w / ew[cell]) <= 0.5 / np.sqrt(40) * 4 + 1e-12)
    assert abs(reps.sum(axis=1).mean() - Nmax) < 0.05 * Nmax


def _c2_like(oracle, N, nsteps, seed=1, n=24):
    e = np.linspace(0, 1, n + 1)
    g = np.random.default_rng(1).standard_normal((n, n))
    sig0 = np.exp(2.3 + 1.5 * (g - g.mean()) / g.std())
    cvv = 0.3
    orc = oracle.Oracle2D(e, e, "xy", seed=seed)
    st = orc.init_simulation(N // 2, np.full((n, n), 1e-3), np.full((n, n), 1e-3), lambda T: cvv * T,
                             lambda u: u / cvv)
    infos = []
    for _ in range(nsteps):
        st, info = orc.step(st, N // 2, N // 2, 0, N, (0.5, 0, 0, 0), 0.005,
                            lambda T: sig0 * np.maximum(T, 1e-3) ** -3, lambda u: u / cvv,
                            lambda T: cvv + 0 * T, np.zeros((n, n)))
        infos.append(info)
    return st, infos


def test_step_energy_balance(oracle):
    st, infos = _c2_like(oracle, 20000, 4)
    for info in infos:
        resid = info["energy_loss"] + info["comb_discrepancy"]
        assert abs(resid) <= 1e-12 * info["total_energy"]
        assert info["boundary_emission"] > 0
