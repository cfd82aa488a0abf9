"""Pins of the multigroup oracle (oracle/imc_oracle_mg.c, oracle/oracle_mg.py)
by closed forms, quadrature, distributions and invariants that a plausible
mistake (a dropped term, a wrong group index, a swapped slot) would break."""
import numpy as np
import pytest
from scipy import integrate, stats

A = 0.01372
C = 29.98
EDGES = np.logspace(-4, 1, 11)


@pytest.fixture(scope="module")
def M(oracle):
    from oracle import oracle_mg

    return oracle_mg


def test_planck_fraction_matches_quadrature(M):
    f = lambda x: x ** 3 / np.expm1(x) if x > 0 else 0.0
    for lo, hi in [(0.0, 0.3), (0.2, 3.0), (0.9, 1.1), (1.0, 2.0), (2.5, 40.0), (5.0, 6.0), (1e-4, 1e-3), (30, 80)]:
        ref = integrate.quad(f, lo, hi, limit=200, epsabs=0, epsrel=1e-13)[0] * 15 / np.pi ** 4
        assert M.planck_fraction(lo, hi) == pytest.approx(ref, rel=1e-12, abs=1e-300)
    # a partition of [0, inf) sums to one
    e = np.concatenate(([0.0], np.logspace(-3, 2.5, 37)))
    assert sum(M.planck_fraction(a, b) for a, b in zip(e[:-1], e[1:])) == pytest.approx(1.0, abs=1e-13)
    # B_g = a c T^4/(4 pi) * fraction
    assert M.bg(0.7, 0.1, 2.0) == pytest.approx(A * C / (4 * np.pi) * 0.7 ** 4 * M.planck_fraction(0.1 / 0.7, 2 / 0.7))
    assert M.bg(0.0, 0.1, 1.0) == 0.0


def test_mixture_of_gammas_reproduces_planck(M):
    """Boundary photons at T: group histogram against the exact Planck group
    fractions (chi-square), on groups fine enough to see the shape."""
    T = 0.8
    edges = np.concatenate(([0.0], T * np.linspace(0.25, 12.0, 30), [1e6]))
    e1 = np.array([0.0, 1.0]); e2 = np.array([0.0, 1.0])
    n = 200000
    p = M.sample_boundary("xy", 0, n, 1.0, 0.01, T, e1, e2, edges, None, None, 3, 0)
    cnt = np.bincount(p.g, minlength=edges.size - 1)
    frac = np.array([M.planck_fraction(a / T, b / T) for a, b in zip(edges[:-1], edges[1:])])
    assert frac.sum() == pytest.approx(1.0, abs=1e-12)
    exp = n * frac
    chi2 = np.sum((cnt - exp) ** 2 / exp)
    assert stats.chi2.sf(chi2, edges.size - 2) > 1e-3, chi2
    # Lambertian normal cosine: E[mu_n] = 2/3; positions on the face
    assert np.mean(p.d1) == pytest.approx(2.0 / 3.0, abs=4 * np.sqrt(1 / 18 / n))
    assert np.all(p.p1 == 0.0) and np.all((p.p2 >= 0) & (p.p2 < 1))


def test_fleck_mg_gray_limit_and_closed_form(M, oracle):
    """sigma_g = sigma for every group gives sigma_P = sigma: the gray Fleck
    factor; with sigma = sigma0 T^-3 it is the T-independent closed form."""
    rng = np.random.default_rng(2)
    T = rng.uniform(0.05, 1.0, 30)
    s0 = rng.uniform(1, 100, 30)
    sig = np.tile(s0 * T ** -3, (10, 1))
    cv = np.full(30, 0.3)
    sa, ss, bs, f = M.fleck_mg(T, sig, cv, 1.0, 0.01, EDGES)
    f_exact = 1.0 / (1.0 + 4 * A * s0 * C * 0.01 / 0.3)
    np.testing.assert_allclose(f, f_exact, rtol=1e-12)
    np.testing.assert_allclose(sa, sig * f_exact, rtol=1e-12)
    np.testing.assert_allclose(ss, sig * (1 - f_exact), rtol=1e-12)
    # b* is a normalised Planck split over the groups
    frac = np.array([[M.planck_fraction(a / t, b / t) for t in T] for a, b in zip(EDGES[:-1], EDGES[1:])])
    np.testing.assert_allclose(bs, frac / frac.sum(0), rtol=1e-12)
    # emission: total a c T^4 sigma_a dt V (sum b* = 1)
    vol = np.full(30, 0.02)
    E, tot = M.emission_mg(T, sa, bs, vol, 0.01)
    np.testing.assert_allclose(tot, A * C * T ** 4 * sa[0] * 0.01 * vol, rtol=1e-12)
    E2, _ = M.emission_mg(T, sa, bs, vol, 0.01, T_emit_floor=0.5)
    assert np.all(E2[:, T < 0.5] == 0.0) and np.all(E2[:, T >= 0.5] == E[:, T >= 0.5])


def test_emission_sampler_distribution(M):
    """Cells drawn in proportion to their emission, groups in proportion to
    the cell's group energies, weights E/N, t uniform in [0, dt)."""
    e1 = np.linspace(0, 1, 5); e2 = np.linspace(0, 1, 4)
    rng = np.random.default_rng(5)
    E = rng.uniform(0, 1, (10, 12)) * (rng.uniform(0, 1, (10, 12)) > 0.3)
    tot = E.sum(0)
    N = 120000
    cdf = M.prefix(tot)
    p = M.sample_emission("xy", e1, e2, E, cdf, N, 0.02, 7, 3)
    assert np.allclose(p.w, cdf[-1] / N, rtol=0)
    cell = p.ci * 3 + p.cj
    # cells from the order statistics of the uniforms: drawn in cell order
    assert np.all(np.diff(cell) >= 0)
    cnt = np.bincount(cell, minlength=12)
    exp = N * tot / tot.sum()
    assert stats.chi2.sf(np.sum((cnt - exp) ** 2 / exp), 11) > 1e-3
    c0 = int(np.argmax(tot))
    gc = np.bincount(p.g[cell == c0], minlength=10)
    pe = E[:, c0] / tot[c0]
    m = pe > 0
    assert np.all(gc[~m] == 0)
    assert stats.chi2.sf(np.sum((gc[m] - gc.sum() * pe[m]) ** 2 / (gc.sum() * pe[m])), m.sum() - 1) > 1e-3
    # positions inside their cells; isotropic 3-D directions projected to xy
    assert np.all((p.p1 >= e1[p.ci]) & (p.p1 < e1[p.ci + 1]) & (p.p2 >= e2[p.cj]) & (p.p2 < e2[p.cj + 1]))
    assert np.mean(p.d1 ** 2 + p.d2 ** 2) == pytest.approx(2.0 / 3.0, abs=0.01)
    assert np.all((p.t >= 0) & (p.t < 0.02))
    # the counts are multinomial (i.i.d. draws, not stratified): over seeds,
    # a cell's count has variance N p (1 - p)
    n_small, nseed = 3000, 120
    pc = tot[c0] / tot.sum()
    k = []
    for seed in range(100, 100 + nseed):
        q = M.sample_emission("xy", e1, e2, E, cdf, n_small, 0.02, seed, 3)
        k.append(np.sum(q.ci * 3 + q.cj == c0))
    k = np.array(k)
    stat = (nseed - 1) * k.var(ddof=1) / (n_small * pc * (1 - pc))
    assert 1e-3 < stats.chi2.cdf(stat, nseed - 1) < 1 - 1e-3
    assert k.mean() == pytest.approx(n_small * pc, abs=5 * np.sqrt(n_small * pc * (1 - pc) / nseed))


def test_source_sampler(M):
    e1 = np.linspace(0, 1, 4); e2 = np.linspace(0, 2, 3)
    S = np.zeros((10, 6))
    S[2, 1] = 1.0; S[2, 4] = 3.0; S[7, 0] = 4.0
    N = 80000
    p, total = M.sample_source("rz", e1, e2, S, N, 0.01, 3, 1)
    assert total == pytest.approx(8.0) and np.allclose(p.w, 8.0 / N)
    assert set(np.unique(p.g)) == {2, 7}
    assert np.mean(p.g == 7) == pytest.approx(0.5, abs=4 * np.sqrt(0.25 / N))
    c = p.ci * 2 + p.cj
    assert set(np.unique(c[p.g == 2])) == {1, 4}
    assert np.mean(c[p.g == 2] == 4) == pytest.approx(0.75, abs=0.01)
    # rz volume-uniform radius in [r0, r1]: r^2 uniform
    sel = (p.g == 7)
    r0, r1 = e1[0], e1[1]
    u = (p.p1[sel] ** 2 - r0 ** 2) / (r1 ** 2 - r0 ** 2)
    assert stats.kstest(u, "uniform").pvalue > 1e-3


def test_transport_zero_d_and_absorber(M):
    """(i) one reflecting cell: each history absorbs w(1 - e^{-sigma_a,g c (dt - t)});
    (ii) a pure absorber strip: a history crossing one cell of width L at
    normal incidence keeps w e^{-sigma_a L} and leaks."""
    e1 = np.array([0.0, 0.5]); e2 = np.array([0.0, 0.5])
    rng = np.random.default_rng(4)
    n = 500
    pool = M.PoolMG(rng.uniform(0.5, 1, n), *M.PoolMG.empty(n).arrays()[1:])
    pool.d1 = rng.uniform(-0.5, 0.5, n); pool.d2 = rng.uniform(-0.5, 0.5, n)
    pool.t = rng.uniform(0, 0.01, n); pool.p1 = rng.uniform(0, 0.5, n); pool.p2 = rng.uniform(0, 0.5, n)
    pool.g = rng.integers(0, 10, n).astype(np.int32)
    sa = np.linspace(1.0, 40.0, 10)[:, None]
    ss = np.linspace(30.0, 0.0, 10)[:, None]
    w0, t0, g0 = pool.w.copy(), pool.t.copy(), pool.g.copy()
    dep, path, bl, st = M.transport("xy", pool, e1, e2, 10, sa, ss, 0.01, (True,) * 4, 10**8, 0.0, 1, 0)
    exact = np.array([np.sum(np.where(g0 == g, w0 * (1 - np.exp(-sa[g, 0] * C * (0.01 - t0))), 0)) for g in range(10)])
    np.testing.assert_allclose(dep[:, 0], exact, rtol=1e-12)
    assert st[3] == n and bl[0] == 0.0
    np.testing.assert_allclose(np.sum(pool.w) + dep.sum(), np.sum(w0), rtol=1e-13)
    # absorber: x-directed histories through a 3-cell strip
    e1 = np.array([0.0, 0.1, 0.2, 0.3]); e2 = np.array([0.0, 1.0])
    k = 50
    p = M.PoolMG.empty(k)
    p.w[:] = 1.0; p.d1[:] = 1.0; p.p1[:] = 0.0; p.p2[:] = np.linspace(0.1, 0.9, k); p.g[:] = 3
    sa = np.tile(np.array([[2.0], [5.0], [1.0]]).T, (10, 1))
    ss = np.zeros_like(sa)
    dep, path, bl, st = M.transport("xy", p, e1, e2, 10, sa, ss, 1.0, (False,) * 4, 100, 0.0, 1, 0)
    assert bl[0] == pytest.approx(k * np.exp(-0.1 * (2 + 5 + 1)), rel=1e-12)
    assert bl[1 + 3] == bl[0] and bl[1 + 10 + 1] == bl[0]
    assert dep[3, 0] == pytest.approx(k * (1 - np.exp(-0.2)), rel=1e-12)


def test_comb_invariants(M):
    rng = np.random.default_rng(9)
    n = 20000
    p = M.PoolMG.empty(n)
    p.w = rng.exponential(1.0, n); p.ci = rng.integers(0, 4, n).astype(np.int32)
    p.cj = rng.integers(0, 3, n).astype(np.int32); p.g = rng.integers(0, 10, n).astype(np.int32)
    p.w[rng.uniform(size=n) < 0.1] = 0.0
    newp, disc, ecen = M.comb(p, 5000, 4, 3, 10, 2, 0)
    after = np.zeros_like(ecen)
    np.add.at(after, (newp.g, newp.ci * 3 + newp.cj), newp.w)
    np.testing.assert_allclose(ecen - disc, after, rtol=1e-12, atol=1e-12)
    occupied = ecen > 0
    assert np.all(np.bincount(newp.g * 12 + newp.ci * 3 + newp.cj, minlength=120).reshape(10, 12)[occupied] >= 0)
    assert abs(len(newp) - 5000) < 5 * np.sqrt(5000) + 120
    assert abs(disc.sum()) < 0.05 * ecen.sum()


def test_mg_step_balance_and_gray_limit(M, oracle):
    """The composed MG step closes its energy balance to 1e-12 with the comb
    discrepancy booked, and with sigma_g = sigma it reproduces the gray step
    in the ensemble mean (MG with one opacity is gray physics)."""
    e1 = np.linspace(0, 1, 7); e2 = np.linspace(0, 0.6, 5)
    nx, ny = 6, 4
    s0 = 8.0
    sf_g = [lambda T: s0 + 0 * T] * 10
    inv = lambda u: u / 0.3
    cvf = lambda T: 0.3 + 0 * T
    T0 = np.full((nx, ny), 0.05)
    mg, gr = [], []
    for seed in range(24):
        orc = M.OracleMG2D(e1, e2, EDGES, "xy", seed=100 + seed)
        st = orc.init_simulation(1500, T0, T0, lambda T: 0.3 * T, inv)
        og = oracle.Oracle2D(e1, e2, "xy", seed=500 + seed, nthreads=1)
        sg = og.init_simulation(1500, T0, T0, lambda T: 0.3 * T, inv)
        for _ in range(2):
            st, info = orc.step(st, 1500, 1500, 0, 3000, (0.4, 0, 0, 0), 0.004, sf_g, inv, cvf, np.zeros((nx, ny)))
            assert abs(info["balance_residual"]) <= 1e-12 * info["total_energy"]
            sg, ig = og.step(sg, 1500, 1500, 0, 3000, (0.4, 0, 0, 0), 0.004, lambda T: s0 + 0 * T, inv, cvf,
                             np.zeros((nx, ny)))
        mg.append(info["temperature"].ravel())
        gr.append(ig["temperature"].ravel())
    mg = np.array(mg); gr = np.array(gr)
    se = np.sqrt(mg.var(0, ddof=1) / len(mg) + gr.var(0, ddof=1) / len(gr))
    z = np.abs(mg.mean(0) - gr.mean(0)) / np.maximum(se, 1e-300)
    assert np.max(z) < 4.5, z


def _mgimc_10():
    import json
    import os

    with open(os.path.join(os.path.dirname(__file__), "golden", "mgimc_10_powerlaw.json")) as fh:
        return json.load(fh)


def test_planck_fractions_match_reference_printed_chi(M):
    """The reference's 10-group power-law Marshak run prints its Planck
    emission fractions chi_g at T = 0.05 keV, computed by its planck_integrals
    module (MG_IMC/problems/test_marshak_wave_multigroup_powerlaw.py:262-273;
    results/marshak_wave_powerlaw/imc_10g_publication/mgimc_10.out:15-25).
    The oracle's b*_g reproduces all ten to the printed 6 decimals."""
    ref = _mgimc_10()
    T = ref["T_chi"]
    frac = np.array([M.planck_fraction(a / T, b / T) for a, b in zip(EDGES[:-1], EDGES[1:])])
    chi = frac / (frac.sum() + 1e-300)
    np.testing.assert_allclose(chi, ref["chi"], rtol=0, atol=5.000001e-7)
    # the printed group edges (4 decimals) are the logspace(-4, 1, 11) of the run
    np.testing.assert_allclose(EDGES[:-1], ref["group_lo_printed"], rtol=0, atol=5.000001e-5)
    # b*_g through the Fleck/emission path: equal group opacities make the
    # per-group emission energies proportional to b*_g
    sa, ss, bs, f = M.fleck_mg(np.array([T]), np.ones((10, 1)), np.array([0.3]), 1.0, 0.01, EDGES)
    np.testing.assert_allclose(bs[:, 0], ref["chi"], rtol=0, atol=5.000001e-7)
    # a plausible slip -- x = E T instead of E / T -- misses the table
    bad = np.array([M.planck_fraction(a * T, b * T) for a, b in zip(EDGES[:-1], EDGES[1:])])
    assert np.max(np.abs(bad / bad.sum() - ref["chi"])) > 0.1


def test_powerlaw_group_opacities_match_reference_printed(M):
    """Geometric-mean group opacities of the power-law material at 0.05 keV
    (mgimc_10.out:16-25): the problem generator's opacity reproduces the
    printed values to their 4 digits."""
    from radtranbook_amd.problems import mg_powerlaw_marshak

    ref = _mgimc_10()
    prob = mg_powerlaw_marshak()
    T = np.full(prob.shape, ref["T_chi"])
    sig = np.array([float(np.mean(f(T))) for f in prob.sigma_a_funcs])
    np.testing.assert_allclose(sig, ref["sigma_a"], rtol=6e-4)
