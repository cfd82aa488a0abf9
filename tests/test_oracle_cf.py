"""Pins of the Carter-Forest oracle (oracle/oracle_cf.py, oracle/imc_oracle.c)
by closed forms, on the CPU.

- Emission time: the inverse of the truncated-exponential CDF.
- Transport in one reflecting cell with constant sigma_a and beta: a history
  is a two-state Markov chain (radiation -> material at rate l1 = c sigma_a,
  material -> radiation at rate l2 = c sigma_a beta), so the census fraction
  is l2/(l1+l2) + l1/(l1+l2) exp(-(l1+l2) dt) and the path-length tally per
  particle c [l2/(l1+l2) dt + l1/(l1+l2)^2 (1 - exp(-(l1+l2) dt))].
- Emission energies: a c sigma T^4 dt V (1 - e^{-x})/x with x = c sigma beta dt.
- Full step: energy balance to 1e-12, and one step of a closed cell against
  the exact solution of the linear two-state system it solves."""
import numpy as np
import pytest

C_LIGHT = 29.98
A_RAD = 0.01372


@pytest.fixture(scope="module")
def CF(oracle):
    from oracle import oracle_cf

    return oracle_cf


def test_emission_time_inverts_the_truncated_exponential(CF):
    dt = 0.01
    for rate in (0.0, 1e-13, 3.0, 80.0, 2000.0):
        for xi in np.linspace(0.0, 0.999999, 23):
            t = CF.cf_time(rate, dt, xi)
            assert 0.0 <= t <= dt * (1 + 1e-12)
            if rate <= 1e-12:
                assert t == xi * dt
            else:
                F = (1.0 - np.exp(-rate * t)) / (1.0 - np.exp(-rate * dt))
                assert F == pytest.approx(xi, abs=1e-12)


def _one_cell_pool(O, n, x0=0.5):
    p = O.Pool2D.empty(n)
    rng = np.random.default_rng(3)
    mu = rng.uniform(-1, 1, n)
    phi = rng.uniform(0, 2 * np.pi, n)
    p.w[:] = 1.0
    p.d1[:] = np.sqrt(1 - mu * mu) * np.cos(phi)
    p.d2[:] = np.sqrt(1 - mu * mu) * np.sin(phi)
    p.t[:] = 0.0
    p.p1[:] = x0
    p.p2[:] = x0
    p.ci[:] = 0
    p.cj[:] = 0
    return p


@pytest.mark.parametrize("sigma,beta", [(5.0, 0.5), (2.0, 3.0)])
def test_one_cell_markov_closed_form(oracle, CF, sigma, beta):
    from oracle import oracle as O

    n, dt = 200_000, 0.01
    e = np.array([0.0, 1.0])
    pool = _one_cell_pool(O, n)
    dep, si, bl, st = CF.transport_2d_cf(pool, e, e, np.array([[sigma]]), np.array([[beta]]), dt, (True,) * 4,
                                         10**7, 0, 0.0, seed=7, step=0)
    l1, l2 = C_LIGHT * sigma, C_LIGHT * sigma * beta
    s = l1 + l2
    p_rad = l2 / s + l1 / s * np.exp(-s * dt)
    se = np.sqrt(p_rad * (1 - p_rad) / n)
    census = pool.w.sum() / n
    assert abs(census - p_rad) < 4 * se, (census, p_rad, se)
    # every weight is either in census or captured, exactly
    assert pool.w.sum() + dep.sum() == pytest.approx(n, rel=1e-12)
    assert bl == 0.0 and st[4] == np.count_nonzero(pool.w == 0.0)
    path = C_LIGHT * (l2 / s * dt + l1 / s**2 * (1 - np.exp(-s * dt)))
    assert si.sum() / n == pytest.approx(path, rel=5e-3)


def test_emission_energy_closed_form(CF):
    T = np.array([[0.3, 1.0], [2.0, 0.0]])
    sig = np.array([[10.0, 1e-16], [0.5, 4.0]])
    cv = np.full((2, 2), 0.2)
    beta = CF.cf_beta(T, cv)
    vol = np.array([[1.0, 2.0], [0.5, 1.0]])
    dt = 0.003
    rate, E = CF.cf_rates_and_energies(T, sig, beta, vol, dt)
    x = rate * dt
    lin = A_RAD * C_LIGHT * sig * np.maximum(T, 0) ** 4 * dt * vol
    ratio = np.where(x > 1e-12 * dt, (1 - np.exp(-x)) / np.where(x > 0, x, 1), 1.0)
    np.testing.assert_allclose(E, lin * ratio, rtol=1e-12, atol=1e-300)
    np.testing.assert_allclose(beta, 4 * A_RAD * np.maximum(T, 1e-12) ** 3 / cv, rtol=1e-15)


def _run(CF, O, nsteps, N, seed, geometry="xy", T0=1e-3, Tb=(0.6, 0, 0, 0), reflect=(False,) * 4, n=8, cvc=0.3,
         sig=lambda T: 20.0 * np.maximum(T, 1e-3) ** -1, **kw):
    e1, e2 = np.linspace(0, 0.5, n + 1), np.linspace(0, 0.5, n + 1)
    orc = CF.OracleCF2D(e1, e2, geometry, seed=seed, nthreads=2)
    Ti = np.full((n, n), T0)
    st = orc.init_simulation(N // 2, Ti, Ti, lambda T: cvc * T, lambda u: u / cvc)
    out = []
    for _ in range(nsteps):
        st, info = orc.step(st, N // 2, N // 2, 0, N, Tb, 0.01, sig, lambda u: u / cvc, lambda T: cvc + 0 * T,
                            np.zeros((n, n)), reflect=reflect, **kw)
        out.append(info)
    return out, st


@pytest.mark.parametrize("geometry", ["xy", "rz"])
def test_full_step_balance(oracle, CF, geometry):
    from oracle import oracle as O

    # warm medium: beta = 4 a T^3 / c_v large enough for re-emission within a step
    infos, _ = _run(CF, O, 3, 20000, 5, geometry=geometry, T0=0.3, Tb=(0.8, 0, 0, 0))
    for info in infos:
        resid = info["energy_loss"] + info["comb_discrepancy"]
        assert abs(resid) <= 1e-12 * info["total_energy"], resid
        assert info["stats"][2] > 0 and info["stats"][4] > 0  # re-emissions and captures happen


def _one_step_radiation(CF, T0, Tr0, sigma, cvc, dt, N, seed):
    """Pre-comb radiation energy density after one step of a closed one-cell
    medium, and the closed form of the linear two-state system the
    Carter-Forest step solves exactly for frozen sigma and beta: with
    E0 = a Tr0^4, U0 = a T0^4, beta = 4 a T0^3 / c_v, s = c sigma (1 + beta),
    E(dt) = [beta E0 + U0 + (E0 - U0) exp(-s dt)] / (1 + beta)."""
    e = np.array([0.0, 1.0])
    orc = CF.OracleCF2D(e, e, "xy", seed=seed, nthreads=2)
    st = orc.init_simulation(N, np.array([[T0]]), np.array([[Tr0]]), lambda T: cvc * T, lambda u: u / cvc)
    # the comb target is small: the estimate is taken before the comb, and long
    # sequential sums of equal copy weights in one cell would lose ~n eps
    st, info = orc.step(st, N, 0, 0, N // 4, (0, 0, 0, 0), dt, lambda T: sigma + 0 * T, lambda u: u / cvc,
                        lambda T: cvc + 0 * T, np.zeros((1, 1)), reflect=(True,) * 4)
    E1 = info["total_radiation_energy"] + info["comb_discrepancy"]
    beta = 4 * A_RAD * T0**3 / cvc
    E0, U0 = A_RAD * Tr0**4, A_RAD * T0**4
    s = C_LIGHT * sigma * (1 + beta)
    return E1, (beta * E0 + U0 + (E0 - U0) * np.exp(-s * dt)) / (1 + beta), info


@pytest.mark.parametrize("T0,Tr0", [(1.0, 0.5), (0.4, 0.9)])
def test_one_step_linear_closed_form(oracle, CF, T0, Tr0):
    """Emission (energy and truncated-exponential times) and transport
    together against the exact solution of the frozen-coefficient system
    (4 standard errors of the Monte Carlo estimate; a 5 % error in the
    emission energy moves the result by about 9 of them)."""
    errs = []
    for seed in range(1, 9):
        E1, exact, info = _one_step_radiation(CF, T0, Tr0, 3.0, 0.01, 0.001, 40000, seed)
        errs.append(E1 / exact - 1.0)
        # energy balance: sums of ~1e5 terms whose two thread partials are
        # combined in either order, so a few 1e-14 of the total either way
        assert abs(info["energy_loss"] + info["comb_discrepancy"]) <= 1e-11 * info["total_energy"]
    errs = np.array(errs)
    se = errs.std(ddof=1) / np.sqrt(errs.size)
    assert abs(errs.mean()) < 4 * se + 1e-4, (errs.mean(), se)
