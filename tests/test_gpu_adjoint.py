"""GPU replay adjoint of the gray 2-D transport and the torch.autograd step.

* FP64 device adjoint against the FP64 oracle adjoint on the same pool and
  streams (history by history, as the forward transport).
* Gradients of a loss through ``AbsorptionStep`` against central differences
  with common random numbers: exact (to O(h^2)) where nothing scatters, so
  paths do not depend on the opacity; in the mean over seeds otherwise.
* FP32-track adjoint against the FP64 adjoint."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

C = 29.98


@pytest.fixture(scope="module")
def L():
    from radtranbook_amd import _lib

    _lib.load()
    _lib.require_gpu()
    return _lib


def _announce(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _setup(L, oracle, geometry, precision, npart=12000, seed=7):
    e1 = np.linspace(0.0 if geometry == "rz" else 0.1, 0.6, 9)
    e2 = np.linspace(0.0, 0.8, 11)
    n, m = 8, 10
    rng = np.random.default_rng(5)
    T = np.exp(rng.uniform(np.log(0.08), np.log(0.5), (n, m)))
    sig = np.exp(rng.uniform(np.log(2.0), np.log(40.0), (n, m)))
    dt = 0.01
    gcode = 0 if geometry == "xy" else 1
    sa, ss, _ = oracle.fleck(T, sig, np.full((n, m), 0.3), 1.0, dt)
    sa = sa.reshape(n, m); ss = ss.reshape(n, m)
    vol = oracle.volumes_2d(e1, e2, geometry)
    Eb = oracle.boundary_energy(gcode, 2, 0.6, dt, e1, e2)
    census = oracle.sample_boundary_2d(gcode, 2, npart, Eb, dt, e1, e2, None, seed, 1)
    Ec = oracle.emission_energy(T, sa, vol, dt)
    cE, _ = oracle.counts(Ec, npart, False)
    emit = oracle.sample_volume_2d(e1, e2, Ec, cE, 0, 0 if geometry == "xy" else 2, T, dt, seed, 1,
                                   oracle.TAG_EMISSION)
    eng = L.Engine2D(e1, e2, geometry, precision, seed=seed)
    eng.set_step(1)
    eng.set_material(T, sig, np.full((n, m), 0.3), 1.0, dt)
    eng.put_census(*census.arrays())
    eng.sample_volume(L.SRC_EMISSION, npart, None, dt, True)
    pool = census.concat(emit)
    return dict(e1=e1, e2=e2, n=n, m=m, dt=dt, sa=sa, ss=ss, vol=vol, eng=eng, pool=pool, lo=len(census),
                hi=len(pool), T=T, sig=sig)


@pytest.mark.parametrize("geometry", ["xy", "rz"])
def test_adjoint_fp64_matches_oracle(L, oracle, geometry):
    _announce(f"test_adjoint_fp64_matches_oracle {geometry}")
    r = _setup(L, oracle, geometry, "fp64")
    eng, pool = r["eng"], r["pool"]
    got = eng.get_census()
    np.testing.assert_allclose(got[0], pool.w, rtol=1e-15)  # the same input pool
    reflect = (True, False, False, False) if geometry == "xy" else (False, False, False, False)
    floor = 1e-10 * oracle.canonical_sum(pool.w) / len(pool)
    eng.enable_adjoint(True)
    eng.transport(r["dt"], reflect, 10**6, floor)
    lam = np.random.default_rng(3).normal(size=(r["n"], r["m"]))
    gA_d, gS_d = eng.adjoint(r["dt"], reflect, 10**6, lam)
    gA, gS, Lam = oracle.transport_2d_adjoint(pool, r["e1"], r["e2"], geometry, r["sa"], r["ss"], r["dt"], reflect,
                                              10**6, floor, eng.seed, 1, lam, r["lo"], r["hi"], nthreads=4)
    sc = np.abs(gA).max()
    closeA = np.isclose(gA_d, gA, rtol=1e-6, atol=1e-9 * sc)
    closeS = np.isclose(gS_d, gS, rtol=1e-6, atol=1e-9 * np.abs(gS).max())
    _announce(f"  cells matching: gA {closeA.mean():.4f}, gS {closeS.mean():.4f}")
    assert closeA.mean() > 0.97 and closeS.mean() > 0.97
    assert np.sum(gA_d * r["sa"]) == pytest.approx(np.sum(gA * r["sa"]), rel=1e-4)
    assert np.sum(gS_d * r["ss"]) == pytest.approx(np.sum(gS * r["ss"]), rel=1e-3, abs=1e-6 * sc)


def test_adjoint_fp32_agrees_with_fp64(L, oracle):
    _announce("test_adjoint_fp32_agrees_with_fp64")
    out = {}
    for prec in ("fp64", "fp32"):
        r = _setup(L, oracle, "rz", prec, npart=40000)
        eng = r["eng"]
        eng.enable_adjoint(True)
        eng.transport(r["dt"], (False,) * 4, 10**6, -1.0)
        lam = np.ones((r["n"], r["m"]))
        out[prec] = eng.adjoint(r["dt"], (False,) * 4, 10**6, lam)
        sa, ss = r["sa"], r["ss"]
    a64, s64 = out["fp64"]; a32, s32 = out["fp32"]
    # total absorption derivative: sum_c sigma_a dL/dsigma_a (pathwise, low noise)
    assert np.sum(a32 * sa) == pytest.approx(np.sum(a64 * sa), rel=1e-2)
    np.testing.assert_allclose(a32, a64, rtol=0.05, atol=0.05 * np.abs(a64).max())


def _rz_step(L, precision, seed, cv_big=False, N=20000):
    from radtranbook_amd import IMC2D
    from radtranbook_amd.diff import AbsorptionStep

    IMC2D.VERBOSE = False
    e1 = np.linspace(0.0, 0.5, 9)
    e2 = np.linspace(0.0, 1.0, 13)
    n, m = 8, 12
    T0 = np.full((n, m), 0.05)
    cv0 = 1e6 if cv_big else 0.3
    sf = lambda T: 40.0 * (0.05 / np.maximum(T, 1e-3)) ** 3
    st = IMC2D.init_simulation(N // 2, T0, T0, e1, e2, lambda T: cv0 * T, lambda u: u / cv0, geometry="rz",
                               precision=precision, seed=seed)
    for _ in range(2):
        st, _ = IMC2D.step(st, N // 2, N // 2, 0, N, (0, 0, 0.3, 0), 0.01, e1, e2, sf, lambda u: u / cv0,
                           lambda T: cv0 + 0 * T, np.zeros((n, m)), geometry="rz")
    step = AbsorptionStep(st, N // 2, N // 2, (0, 0, 0.3, 0), 0.01, sf, lambda T: cv0 + 0 * T)
    return step, n, m


def _loss(E, w, Es):
    return np.sum(w * (E - Es) ** 2)


def test_autograd_matches_fd_without_scattering(L):
    """Huge c_v makes the Fleck factor 1: no effective scattering, so paths do
    not depend on the opacity scale and central differences with common
    random numbers (fixed emission counts) are exact to O(h^2)."""
    import torch

    _announce("test_autograd_matches_fd_without_scattering")
    step, n, m = _rz_step(L, "fp64", 3, cv_big=True)
    rng = np.random.default_rng(11)
    scale = torch.tensor(1.0 + 0.1 * rng.standard_normal((n, m)), dtype=torch.float32, requires_grad=True)
    E0 = step.forward_np(scale.detach().double().numpy(), 1.0)
    Es = E0 * rng.uniform(0.8, 1.2, (n, m))
    w = rng.uniform(0.5, 1.5, (n, m)) / np.sum(E0 ** 2)
    E = step(scale)
    loss = torch.sum(torch.from_numpy(w) * (E - torch.from_numpy(Es)) ** 2)
    loss.backward()
    g = scale.grad.double().numpy()
    step.fixed_emission_counts = step.last_counts
    s0 = scale.detach().double().numpy()
    v = rng.standard_normal((n, m))
    h = 1e-4
    lp = _loss(step.forward_np(s0 + h * v, 1.0), w, Es)
    lm = _loss(step.forward_np(s0 - h * v, 1.0), w, Es)
    fd = (lp - lm) / (2 * h)
    _announce(f"  directional derivative: autograd {np.sum(g * v):.6e}, fd {fd:.6e}")
    assert np.sum(g * v) == pytest.approx(fd, rel=2e-4)


def test_autograd_matches_fd_in_mean(L):
    """With scattering: the score-function gradient against 3% central
    differences with common random numbers, compared over seeds, along a
    random direction of the opacity scale and of the c_v scale (the chain rule
    dsigma_a/dcv_scale = sigma f (1 - f) / cv_scale = -dsigma_s/dcv_scale)."""
    import torch

    _announce("test_autograd_matches_fd_in_mean")
    ad, fd = [], []
    adc, fdc = [], []
    for seed in range(10):
        step, n, m = _rz_step(L, "fp64", 20 + seed, N=16000)
        rng = np.random.default_rng(99)
        s0 = 1.0 + 0.1 * rng.standard_normal((n, m))
        v = rng.standard_normal((n, m))
        E0 = step.forward_np(s0, 1.0)
        Es = E0 * rng.uniform(0.8, 1.2, (n, m))
        w = rng.uniform(0.5, 1.5, (n, m)) / np.sum(E0 ** 2)
        scale = torch.tensor(s0, dtype=torch.float64, requires_grad=True)
        cvs = torch.ones((n, m), dtype=torch.float64, requires_grad=True)
        E = step(scale, cvs)
        loss = torch.sum(torch.from_numpy(w) * (E - torch.from_numpy(Es)) ** 2)
        loss.backward()
        ad.append(float(np.sum(scale.grad.numpy() * v)))
        step.fixed_emission_counts = step.last_counts
        h = 0.03
        fd.append((_loss(step.forward_np(s0 + h * v, 1.0), w, Es) - _loss(step.forward_np(s0 - h * v, 1.0), w, Es))
                  / (2 * h))
        c = np.random.default_rng(77).standard_normal((n, m))
        adc.append(float(np.sum(cvs.grad.numpy() * c)))
        fdc.append((_loss(step.forward_np(s0, 1.0 + h * c), w, Es) - _loss(step.forward_np(s0, 1.0 - h * c), w, Es))
                   / (2 * h))
    for name, a_, f_ in (("opacity scale", ad, fd), ("c_v scale", adc, fdc)):
        a_ = np.array(a_); f_ = np.array(f_)
        d = a_ - f_
        z = abs(d.mean()) / (d.std(ddof=1) / np.sqrt(len(d)) + 1e-300)
        _announce(f"  {name}: autograd {a_.mean():.4e} +- {a_.std(ddof=1) / np.sqrt(len(a_)):.1e}, "
                  f"fd {f_.mean():.4e} +- {f_.std(ddof=1) / np.sqrt(len(f_)):.1e}, z {z:.2f}")
        assert z < 4.0
        assert np.sign(a_.mean()) == np.sign(f_.mean())


@pytest.mark.parametrize("start", ["cold", "warm"])
def test_replay_adjoint_matches_lr_reference_c4(L, oracle, start):
    """The device replay adjoint against the oracle's FP64 forward-mode
    likelihood-ratio gradient (oracle.transport_2d_lr, an independent
    estimator) on the differentiable rz configuration's mesh and material
    (128x256 cells, sigma_a = theta 100 T^-3, c_v = 0.3, bottom face 0.3 keV,
    dt = 0.01 ns; 4 steps of evolution, then the gradient step; 4e5
    particles per step so that the CPU reference stays short), along random
    directions of (sigma_a, sigma_s), over 8 seeds of the gradient step.
    The pathwise (absorption and emission-weight) parts are one estimator
    accumulated in opposite directions on the same histories: they agree
    seed by seed.  The score-function (sigma_s) parts are different
    estimators of the same derivative: they agree in the mean.  From the
    configuration's cold start (1e-3 keV) every scatter rate is still capped
    at the gradient step (sigma_s min(dr, dz) > 1e4), so its score parts are
    zero; a warm start (0.2 keV) leaves uncapped effective scattering."""
    from radtranbook_amd import IMC2D

    IMC2D.VERBOSE = False
    nr, nz, N = 128, 256, 400_000
    _announce(f"test_replay_adjoint_matches_lr_reference_c4 {start}: rz 128x256, 4e5 particles/step, 8 seeds")
    e1 = np.linspace(0.0, 0.5, nr + 1)
    e2 = np.linspace(0.0, 1.0, nz + 1)
    theta = 1.0 + 0.1 * np.random.default_rng(1).standard_normal((nr, nz))
    sf = lambda T: theta * 100.0 * np.maximum(T, 1e-3) ** -3  # noqa: E731
    cvf = lambda T: 0.3 + 0 * T  # noqa: E731
    T0 = np.full((nr, nz), 1e-3 if start == "cold" else 0.2)
    Tb, dt, refl = (0.0, 0.0, 0.3, 0.0), 0.01, (False,) * 4
    st = IMC2D.init_simulation(N // 2, T0, T0, e1, e2, lambda T: 0.3 * T, lambda u: u / 0.3, geometry="rz",
                               precision="fp64", seed=1)
    for _ in range(4):
        st, _ = IMC2D.step(st, N // 2, N // 2, 0, N, Tb, dt, e1, e2, sf, lambda u: u / 0.3, cvf,
                           np.zeros((nr, nz)), geometry="rz")
    eng = st._engine
    T = st.temperature
    rng = np.random.default_rng(7)
    lam = rng.normal(size=(nr, nz))
    sa, ss = eng.set_material(T, sf(T), cvf(T), 1.0, dt, want_opacity=True)
    capped = ss * np.minimum(np.diff(e1)[:, None], np.diff(e2)[None, :]) > 1e4
    vA = rng.normal(size=(nr, nz)) * sa
    vS = rng.normal(size=(nr, nz)) * ss
    Eb = IMC2D._face_energy("rz", 2, 0.3, dt, e1, e2)
    rows = []
    for seed in range(8):
        eng.set_seed(100 + seed)
        eng.snapshot(True)
        eng.set_step(st.count)
        eng.set_material(T, sf(T), cvf(T), 1.0, dt)
        eng.sample_boundary(2, N // 2, Eb, dt)
        lo = eng.n
        eng.sample_volume(L.SRC_EMISSION, N // 2, None, dt, True)
        hi = eng.n
        w, d1, d2, t, p1, p2, ci, cj = eng.get_census()
        pool = oracle.Pool2D(w, d1, d2, t, p1, p2, ci.astype(np.int32), cj.astype(np.int32))
        floor = 1e-10 * oracle.canonical_sum(w) / len(w)
        eng.enable_adjoint(True)
        eng.transport(dt, refl, 10**6, floor)
        gA, gS = eng.adjoint(dt, refl, 10**6, lam)
        eng.enable_adjoint(False)
        eng.snapshot(False)
        path, score, X = oracle.transport_2d_lr(pool, e1, e2, "rz", sa, ss, dt, refl, 10**6, floor, 100 + seed,
                                                st.count, lam, vA, vS, lo, hi)
        rows.append((float(np.sum(gA * vA)), path, float(np.sum(gS * vS)), score))
        _announce(f"  seed {seed}: pathwise adjoint {rows[-1][0]:.6e} LR {path:.6e}; "
                  f"score adjoint {rows[-1][2]:.5e} LR {score:.5e}")
    r = np.array(rows)
    # same histories (>99.9% identical between device and oracle): seed by seed
    np.testing.assert_allclose(r[:, 0], r[:, 1], rtol=5e-3, atol=5e-3 * np.abs(r[:, 1]).max())
    _announce(f"  capped cells {capped.mean():.3f}")
    if start == "cold":
        assert capped.all() and np.all(r[:, 2] == 0.0) and np.all(r[:, 3] == 0.0)
    else:
        assert not capped.all()
        d = r[:, 2] - r[:, 3]
        se = d.std(ddof=1) / np.sqrt(len(d))
        _announce(f"  score part: adjoint {r[:, 2].mean():.5e} +- {r[:, 2].std(ddof=1) / np.sqrt(8):.1e}, "
                  f"LR {r[:, 3].mean():.5e} +- {r[:, 3].std(ddof=1) / np.sqrt(8):.1e}, diff {d.mean():.3e} +- {se:.1e}")
        assert abs(d.mean()) <= 4 * se
        assert np.any(r[:, 2] != 0.0)
    st._engine.close()
