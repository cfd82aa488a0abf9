"""Pins of the replay adjoint of the gray 2-D transport (oracle side):
the absorption (pathwise) terms and the emission-weight term against finite
differences with common random numbers, which are exact for them because
sigma_a changes weights but not paths; the 0-D closed form; and the
scatter-sampling (score-function) terms against the sample-path derivative of
finite differences, which has the same expectation."""
import numpy as np
import pytest

C = 29.98


def _pool(oracle, n, e1, e2, rng, t_max=0.01):
    p = oracle.Pool2D.empty(n)
    p.w[:] = rng.uniform(0.5, 1.5, n)
    uz = rng.uniform(-1, 1, n)
    phi = rng.uniform(0, 2 * np.pi, n)
    r = np.sqrt(1 - uz * uz)
    p.d1[:] = r * np.cos(phi)
    p.d2[:] = r * np.sin(phi)
    p.t[:] = rng.uniform(0, t_max, n)
    p.p1[:] = rng.uniform(e1[0], e1[-1], n)
    p.p2[:] = rng.uniform(e2[0], e2[-1], n)
    p.ci[:] = np.clip(np.searchsorted(e1, p.p1, "right") - 1, 0, e1.size - 2)
    p.cj[:] = np.clip(np.searchsorted(e2, p.p2, "right") - 1, 0, e2.size - 2)
    return p


def _lam_tally(oracle, pool, e1, e2, sa, ss, dt, refl, lam, seed, step=1, scale_w=None):
    q = oracle.Pool2D(*[a.copy() for a in pool.arrays()])
    if scale_w is not None:
        q.w *= scale_w
    vol = oracle.volumes_2d(e1, e2, "xy")
    dep, si, bl, st = oracle.transport_2d(q, e1, e2, sa, ss, vol, dt, refl, 10**7, 0, 0.0, seed, step)
    return float(np.sum(lam * dep * vol))


def test_adjoint_zero_d_closed_form(oracle):
    e1 = np.array([0.0, 0.4]); e2 = np.array([0.0, 0.3])
    rng = np.random.default_rng(1)
    p = _pool(oracle, 2000, e1, e2, rng)
    sa = np.array([[7.0]]); ss = np.array([[25.0]]); dt = 0.01
    gA, gS, Lam = oracle.transport_2d_adjoint(p, e1, e2, "xy", sa, ss, dt, (True,) * 4, 10**7, 0.0, 3, 1,
                                              np.ones((1, 1)))
    L = C * (dt - p.t)
    exact = np.sum(p.w * L * np.exp(-sa[0, 0] * L))
    assert gA[0, 0] == pytest.approx(exact, rel=1e-10)
    np.testing.assert_allclose(Lam, p.w * (1 - np.exp(-sa[0, 0] * L)), rtol=1e-12)
    # the score term has zero mean (the tally does not depend on the path here)
    vals = []
    for seed in range(30):
        _, gs, _ = oracle.transport_2d_adjoint(p, e1, e2, "xy", sa, ss, dt, (True,) * 4, 10**7, 0.0, 100 + seed, 1,
                                               np.ones((1, 1)))
        vals.append(gs[0, 0])
    vals = np.array(vals)
    assert abs(vals.mean()) < 4 * vals.std(ddof=1) / np.sqrt(len(vals))


def _mesh():
    return np.linspace(0, 0.6, 5), np.linspace(0, 0.4, 4)


def test_adjoint_absorption_matches_crn_fd(oracle):
    """d/dsigma_a of sum lam E_abs: with the same random numbers the paths do
    not move, so central differences are exact to O(h^2)."""
    e1, e2 = _mesh()
    rng = np.random.default_rng(2)
    p = _pool(oracle, 3000, e1, e2, rng)
    sa = rng.uniform(2, 20, (4, 3)); ss = rng.uniform(1, 40, (4, 3))
    lam = rng.normal(size=(4, 3))
    dt, refl = 0.01, (False, True, False, True)
    gA, gS, _ = oracle.transport_2d_adjoint(p, e1, e2, "xy", sa, ss, dt, refl, 10**7, 0.0, 5, 1, lam)
    for (i, j) in [(0, 0), (1, 2), (3, 1), (2, 0)]:
        h = 1e-5 * sa[i, j]
        sp = sa.copy(); sp[i, j] += h
        sm = sa.copy(); sm[i, j] -= h
        fd = (_lam_tally(oracle, p, e1, e2, sp, ss, dt, refl, lam, 5)
              - _lam_tally(oracle, p, e1, e2, sm, ss, dt, refl, lam, 5)) / (2 * h)
        assert gA[i, j] == pytest.approx(fd, rel=1e-6, abs=1e-9 * np.abs(gA).max())


def test_adjoint_emission_weight_term(oracle):
    """Particles whose weight is proportional to sigma_a of their birth cell
    (material emission) add Lambda / sigma_a there."""
    e1, e2 = _mesh()
    rng = np.random.default_rng(3)
    p = _pool(oracle, 2000, e1, e2, rng)
    sa = rng.uniform(2, 20, (4, 3)); ss = rng.uniform(1, 40, (4, 3))
    lam = rng.normal(size=(4, 3))
    dt, refl = 0.01, (False,) * 4
    lo, hi = 500, 1700
    gA, _, _ = oracle.transport_2d_adjoint(p, e1, e2, "xy", sa, ss, dt, refl, 10**7, 0.0, 8, 1, lam, lo, hi)
    emit = np.zeros(len(p), bool); emit[lo:hi] = True
    for (i, j) in [(1, 1), (3, 2)]:
        h = 1e-5 * sa[i, j]
        born = emit & (p.ci == i) & (p.cj == j)
        vals = []
        for sgn in (1, -1):
            s2 = sa.copy(); s2[i, j] += sgn * h
            scale = np.where(born, s2[i, j] / sa[i, j], 1.0)
            vals.append(_lam_tally(oracle, p, e1, e2, s2, ss, dt, refl, lam, 8, scale_w=scale))
        fd = (vals[0] - vals[1]) / (2 * h)
        assert gA[i, j] == pytest.approx(fd, rel=1e-6)


def test_adjoint_scatter_score_matches_fd_in_mean(oracle):
    """d/dsigma_s: the score-function estimate against central differences of
    a global sigma_s scale (3%) with common random numbers.  The draws are
    indexed by event, so a scatter site moving across a face changes the
    rest of the path: the step must be large enough that such jumps are
    sampled, and the two estimates agree only in the mean."""
    e1, e2 = _mesh()
    rng = np.random.default_rng(4)
    sa = rng.uniform(1, 5, (4, 3)); ss = rng.uniform(5, 30, (4, 3))
    lam = np.ones((4, 3))
    dt, refl, h = 0.01, (False,) * 4, 0.03
    lr, fd = [], []
    for seed in range(60):
        p = _pool(oracle, 800, e1, e2, np.random.default_rng(100 + seed))
        _, gS, _ = oracle.transport_2d_adjoint(p, e1, e2, "xy", sa, ss, dt, refl, 10**7, 0.0, seed, 1, lam)
        lr.append(np.sum(gS * ss))
        fd.append((_lam_tally(oracle, p, e1, e2, sa, ss * (1 + h), dt, refl, lam, seed)
                   - _lam_tally(oracle, p, e1, e2, sa, ss * (1 - h), dt, refl, lam, seed)) / (2 * h))
    lr = np.array(lr); fd = np.array(fd)
    d = lr - fd
    z = abs(d.mean()) / (d.std(ddof=1) / np.sqrt(len(d)))
    assert z < 4.0, (z, lr.mean(), fd.mean())
    # more scattering keeps photons longer in the absorber: clearly positive
    assert lr.mean() > 3 * lr.std(ddof=1) / np.sqrt(len(lr))


# ---------------------------------------------------------------------------
# forward-mode likelihood-ratio reference (oracle.transport_2d_lr)
# ---------------------------------------------------------------------------
def test_lr_zero_d_closed_form(oracle):
    """One reflecting cell: E_abs = sum w (1 - e^{-sigma_a L}), L = c (dt - t),
    so the pathwise sigma_a derivative is sum w L e^{-sigma_a L} exactly; the
    tally does not depend on the path, so the score part has zero mean."""
    e1 = np.array([0.0, 0.4]); e2 = np.array([0.0, 0.3])
    p = _pool(oracle, 2000, e1, e2, np.random.default_rng(1))
    sa = np.array([[7.0]]); ss = np.array([[25.0]]); dt = 0.01
    one = np.ones((1, 1))
    path, score, X = oracle.transport_2d_lr(p, e1, e2, "xy", sa, ss, dt, (True,) * 4, 10**7, 0.0, 3, 1, one, one,
                                            0 * one)
    L = C * (dt - p.t)
    assert path == pytest.approx(np.sum(p.w * L * np.exp(-sa[0, 0] * L)), rel=1e-10)
    assert X == pytest.approx(np.sum(p.w * (1 - np.exp(-sa[0, 0] * L))), rel=1e-12)
    assert score == 0.0
    vals = [oracle.transport_2d_lr(p, e1, e2, "xy", sa, ss, dt, (True,) * 4, 10**7, 0.0, 100 + s, 1, one, 0 * one,
                                   one)[1] for s in range(30)]
    vals = np.array(vals)
    assert abs(vals.mean()) < 4 * vals.std(ddof=1) / np.sqrt(len(vals))


def test_lr_pathwise_equals_replay_adjoint(oracle):
    """The pathwise (sigma_a) and emission-weight parts of the forward-mode
    estimate and of the replay adjoint are the same estimator accumulated in
    opposite directions: on the same histories they agree to rounding for
    any direction vA."""
    e1, e2 = _mesh()
    rng = np.random.default_rng(6)
    p = _pool(oracle, 3000, e1, e2, rng)
    sa = rng.uniform(2, 20, (4, 3)); ss = rng.uniform(1, 40, (4, 3))
    lam = rng.normal(size=(4, 3))
    dt, refl, lo, hi = 0.01, (False, True, False, True), 700, 2400
    gA, gS, _ = oracle.transport_2d_adjoint(p, e1, e2, "xy", sa, ss, dt, refl, 10**7, 0.0, 5, 1, lam, lo, hi)
    for trial in range(3):
        vA = rng.normal(size=(4, 3))
        path, _, _ = oracle.transport_2d_lr(p, e1, e2, "xy", sa, ss, dt, refl, 10**7, 0.0, 5, 1, lam, vA,
                                            np.zeros((4, 3)), lo, hi)
        assert path == pytest.approx(np.sum(gA * vA), rel=1e-10, abs=1e-12 * np.sum(np.abs(gA * vA)))


def test_lr_score_matches_fd_and_adjoint_in_mean(oracle):
    """The score-function (sigma_s) part: its mean over seeds agrees with
    central differences of a global sigma_s scale (common random numbers,
    3%) and with the replay adjoint's causal score estimate, a different
    estimator of the same derivative."""
    e1, e2 = _mesh()
    rng = np.random.default_rng(4)
    sa = rng.uniform(1, 5, (4, 3)); ss = rng.uniform(5, 30, (4, 3))
    lam = np.ones((4, 3))
    dt, refl, h = 0.01, (False,) * 4, 0.03
    lr, adj, fd = [], [], []
    for seed in range(60):
        p = _pool(oracle, 800, e1, e2, np.random.default_rng(100 + seed))
        _, score, _ = oracle.transport_2d_lr(p, e1, e2, "xy", sa, ss, dt, refl, 10**7, 0.0, seed, 1, lam,
                                             np.zeros((4, 3)), ss)
        _, gS, _ = oracle.transport_2d_adjoint(p, e1, e2, "xy", sa, ss, dt, refl, 10**7, 0.0, seed, 1, lam)
        lr.append(score)
        adj.append(np.sum(gS * ss))
        fd.append((_lam_tally(oracle, p, e1, e2, sa, ss * (1 + h), dt, refl, lam, seed)
                   - _lam_tally(oracle, p, e1, e2, sa, ss * (1 - h), dt, refl, lam, seed)) / (2 * h))
    lr, adj, fd = np.array(lr), np.array(adj), np.array(fd)
    for other, name in ((fd, "fd"), (adj, "adjoint")):
        d = lr - other
        z = abs(d.mean()) / (d.std(ddof=1) / np.sqrt(len(d)))
        assert z < 4.0, (name, z, lr.mean(), other.mean())
    assert lr.mean() > 3 * lr.std(ddof=1) / np.sqrt(len(lr))
