"""Pins of the gray 1-D oracle (slab and sphere, IMC/IMC1D.py) against what
the mathematics fixes: face-source energies and the Lambertian law, the
sampled position densities (KS tests against CDFs derived here), 0-D
reflecting-cell closed forms, and straight-line attenuation through several
cells, with sphere chord lengths from a brute-force 3-D ray march.  The slab
step is also pinned by the reference's own ensemble fixture
(tests/test_reference_fixtures.py)."""
import numpy as np
import pytest
from scipy import stats

C = 29.98
A = 0.01372


def test_boundary_energy_and_lambertian(oracle):
    L = oracle._lib1d()
    for sph, R in ((False, 0.0), (True, 0.7)):
        E = L.o1_boundary_energy(int(sph), 0.8, 0.01, R)
        exact = A * C * 0.8 ** 4 / 4.0 * 0.01 * (4.0 * np.pi * R * R if sph else 1.0)
        assert E == pytest.approx(exact, rel=1e-14)
    edges = np.linspace(0.0, 0.7, 8)
    for sph in (False, True):
        for side in (0, 1):
            p = oracle.sample_boundary_1d(sph, side, 40000, 2.0, 0.01, edges, 3, 1, 0)
            # mu = +-sqrt(xi): mu^2 uniform on (0,1)
            assert stats.kstest(p.mu ** 2, "uniform").pvalue > 1e-3
            assert np.all(p.mu > 0) if side == 0 else np.all(p.mu < 0)
            assert np.sum(p.w) == pytest.approx(2.0, rel=1e-12)
            assert np.all((p.t >= 0.0) & (p.t < 0.01))
            assert np.all(p.cell == (0 if side == 0 else 6))


def test_sphere_uniform_volume_sampling(oracle):
    edges = np.array([0.0, 0.2, 0.45, 0.5])
    QV = np.array([1.0, 2.0, 0.5])
    cnt = np.array([20000, 20000, 20000])
    p = oracle.sample_volume_1d(True, edges, QV, cnt, 1, None, 0.0, 5, 0, oracle.TAG_IC, 0)
    for i in range(3):
        r = p.x[p.cell == i]
        r0, r1 = edges[i], edges[i + 1]
        u = (r ** 3 - r0 ** 3) / (r1 ** 3 - r0 ** 3)  # volume fraction is uniform
        assert stats.kstest(u, "uniform").pvalue > 1e-3
        assert np.sum(p.w[p.cell == i]) == pytest.approx(QV[i], rel=1e-12)
    assert stats.kstest((p.mu + 1.0) / 2.0, "uniform").pvalue > 1e-3
    assert np.all(p.t == 0.0)


def test_slab_linear_tilt_density(oracle):
    """Emission positions in a slab cell follow s x + (T - s dx/2) on
    [0, dx] (IMC/IMC1D.py:203-217) with the slope of IMC/IMC1D.py:254-262."""
    edges = np.array([0.0, 0.1, 0.25, 0.3, 0.5])
    T = np.array([0.5, 0.6, 0.9, 0.95])
    dx = np.diff(edges)
    s = np.zeros(4)
    s[1:-1] = (T[2:] - T[:-2]) / (dx[2:] + dx[:-2]) * 2
    s[0] = (T[1] - T[0]) / (dx[1] + dx[0]) * 2
    s[-1] = (T[-1] - T[-2]) / (dx[-1] + dx[-2]) * 2
    s[(T - s * dx / 2 < 0) | (T + s * dx / 2 < 0)] = 0.0
    QV = np.ones(4)
    cnt = np.full(4, 30000)
    p = oracle.sample_volume_1d(False, edges, QV, cnt, 0, T, 0.01, 9, 1, oracle.TAG_EMISSION, 0)
    for i in range(4):
        x = p.x[p.cell == i] - edges[i]
        si = float(np.clip(s[i], -2 * T[i] / dx[i] ** 2, 2 * T[i] / dx[i] ** 2))
        if abs(si) < 1e-6:
            si = 0.0
        b = T[i] - si * dx[i] / 2
        cdf = lambda y, b=b, si=si, d=dx[i], t=T[i]: (b * y + si * y * y / 2.0) / (t * d)
        assert stats.kstest(x, cdf).pvalue > 1e-3, i
    assert np.all((p.t >= 0.0) & (p.t < 0.01))


def test_sphere_emission_tilt_density(oracle):
    """Spherical emission follows r^2 (a + M r) with a the inner-edge T^4
    (the CDF stated at IMC/IMC1D.py:580-590)."""
    edges = np.array([0.1, 0.3, 0.4, 0.6])
    T = np.array([0.4, 0.7, 0.8])
    t4 = T ** 4
    rm = 0.5 * (edges[1:] + edges[:-1])
    M = np.zeros(3)
    M[1:-1] = (t4[2:] - t4[:-2]) / (rm[2:] - rm[:-2])
    M[0] = (t4[1] - t4[0]) / (rm[1] - rm[0])
    M[-1] = (t4[-1] - t4[-2]) / (rm[-1] - rm[-2])
    a = t4 - M * (rm - edges[:-1])
    b = t4 + M * (edges[1:] - rm)
    M[(a < 0) | (b < 0)] = 0.0
    a = t4 - M * (rm - edges[:-1])
    cnt = np.full(3, 30000)
    p = oracle.sample_volume_1d(True, edges, np.ones(3), cnt, 0, T, 0.01, 4, 1, oracle.TAG_EMISSION, 0)
    for i in range(3):
        r = p.x[p.cell == i]
        r0, r1 = edges[i], edges[i + 1]
        G = lambda y, i=i: a[i] / 3.0 * y ** 3 + M[i] / 4.0 * y ** 4
        cdf = lambda y, i=i, r0=r0, r1=r1, G=G: (G(y) - G(r0)) / (G(r1) - G(r0))
        assert stats.kstest(r, cdf).pvalue > 1e-3, i


@pytest.mark.parametrize("sph", [False, True])
def test_zero_d_reflecting_cell(oracle, sph):
    """One reflecting cell: every history stays to census, so the deposit is
    sum w (1 - exp(-sigma_a c (dt - t))) whatever the scattering."""
    edges = np.array([0.0, 0.3])
    rng = np.random.default_rng(2)
    n = 3000
    p = oracle.Pool1D.empty(n)
    p.w[:] = rng.uniform(0.5, 1.5, n)
    p.mu[:] = rng.uniform(-1, 1, n)
    p.t[:] = rng.uniform(0, 0.01, n)
    p.x[:] = rng.uniform(0.0, 0.3, n) if not sph else 0.3 * rng.uniform(0, 1, n) ** (1 / 3)
    w0, t0 = p.w.copy(), p.t.copy()
    vol = oracle.volumes_1d(edges, "sphere" if sph else "slab")
    for ss in (0.0, 30.0):
        q = oracle.Pool1D(*[x.copy() for x in p.arrays()])
        dep, si, leak, fl, st = oracle.transport_1d(sph, q, edges, np.array([6.0]), np.array([ss]), 0.01,
                                                    (True, True), 0.0, 7, 1, 0)
        exact = np.sum(w0 * (1 - np.exp(-6.0 * C * (0.01 - t0))))
        assert dep[0] * vol[0] == pytest.approx(exact, rel=1e-11)
        assert leak == 0.0
        assert np.sum(q.w) + dep[0] * vol[0] == pytest.approx(np.sum(w0), rel=1e-13)
        # path-length tally: si dt sigma_a == dep (sigma_a > 0)
        assert si[0] * 0.01 * 6.0 == pytest.approx(dep[0], rel=1e-12)
        if ss > 0:
            assert st[2] > 0


def _shell_lengths_march(r, mu, edges, nstep=400000):
    """Brute-force 3-D ray march from radius r with radial cosine mu until the
    ray leaves the outer sphere: path length spent in each shell."""
    pos = np.array([0.0, 0.0, r])
    d = np.array([np.sqrt(max(0.0, 1 - mu * mu)), 0.0, mu])
    R = edges[-1]
    # exit distance from the outer sphere, then midpoint sampling
    bq = np.dot(pos, d)
    smax = -bq + np.sqrt(bq * bq - (np.dot(pos, pos) - R * R))
    h = smax / nstep
    s = (np.arange(nstep) + 0.5) * h
    rr = np.sqrt((pos[0] + s * d[0]) ** 2 + (pos[2] + s * d[2]) ** 2)
    idx = np.clip(np.searchsorted(edges, rr, "right") - 1, 0, len(edges) - 2)
    inside = rr >= edges[0]
    return np.bincount(idx[inside], minlength=len(edges) - 1) * h


def _shell_lengths_march_to(r, mu, edges, smax, nstep=400000):
    h = smax / nstep
    s = (np.arange(nstep) + 0.5) * h
    d1 = np.sqrt(max(0.0, 1 - mu * mu))
    rr = np.sqrt((s * d1) ** 2 + (r + s * mu) ** 2)
    idx = np.clip(np.searchsorted(edges, rr, "right") - 1, 0, len(edges) - 2)
    return np.bincount(idx, minlength=len(edges) - 1) * h


def test_straight_line_attenuation(oracle):
    """Pure absorbers, vacuum faces, one particle at a time: the leaked
    weight is w exp(-sum sigma_i l_i); slab l_i = dx_i/|mu|; sphere l_i from a
    3-D ray march (chords that pass close to the inner face, which
    has vacuum inside, so a particle that reaches it escapes)."""
    sa = np.array([3.0, 0.5, 7.0, 1.0])
    for sph, edges in ((False, np.array([0.0, 0.1, 0.3, 0.35, 0.5])), (True, np.array([0.05, 0.1, 0.3, 0.35, 0.5]))):
        for (x0, mu) in ((0.12, 0.7), (0.33, -0.4), (0.2, -0.95), (0.45, 0.1), (0.2, -0.999)):
            p = oracle.Pool1D.empty(1)
            p.w[:] = 1.0; p.mu[:] = mu; p.t[:] = 0.0; p.x[:] = x0
            p.cell[:] = np.searchsorted(edges, x0, "right") - 1
            dep, si, leak, fl, st = oracle.transport_1d(sph, p, edges, sa, np.zeros(4), 1.0, (False, False), 0.0, 1,
                                                        1, 0)
            if not sph:
                c0 = np.searchsorted(edges, x0, "right") - 1
                ell = np.zeros(4)
                if mu > 0:
                    ell[c0] = (edges[c0 + 1] - x0) / mu
                    ell[c0 + 1:] = np.diff(edges)[c0 + 1:] / mu
                else:
                    ell[c0] = (x0 - edges[c0]) / -mu
                    ell[:c0] = np.diff(edges)[:c0] / -mu
            else:
                ell = _shell_lengths_march(x0, mu, edges)
                if x0 * np.sqrt(1 - mu * mu) < edges[0] and mu < 0:
                    # the chord meets the inner face: keep the inbound part only
                    b2 = x0 * x0 * (1 - mu * mu)
                    s_in = -x0 * mu - np.sqrt(edges[0] ** 2 - b2)
                    ell = _shell_lengths_march_to(x0, mu, edges, s_in)
            expect = np.exp(-np.sum(sa * ell))
            assert leak == pytest.approx(expect, rel=2e-4 if sph else 1e-12), (sph, x0, mu)
