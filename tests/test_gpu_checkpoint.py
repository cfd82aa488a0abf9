"""Restart from a checkpoint on the GPU: a run interrupted after k steps,
saved, restored into a new device context and continued, matches the
uninterrupted run (same census, same random streams); the only difference
allowed is the summation order of the FP64 tally atomics."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EDGES = np.logspace(-4, 1, 11)
NU = np.sqrt(EDGES[:-1] * EDGES[1:])


@pytest.fixture(scope="module")
def L():
    from radtranbook_amd import _lib

    _lib.load()
    _lib.require_gpu()
    return _lib


def _announce(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _interrupted(run_steps, init, nsteps, k, path):
    """States after nsteps uninterrupted, and after k steps + save/load +
    nsteps - k steps."""
    from radtranbook_amd import checkpoint

    a = init()
    for _ in range(nsteps):
        a = run_steps(a)
    b = init()
    for _ in range(k):
        b = run_steps(b)
    checkpoint.save_checkpoint(b, path)
    b._engine.close()
    b, _ = checkpoint.load_checkpoint(path)
    for _ in range(nsteps - k):
        b = run_steps(b)
    return a, b


def _compare(a, b, fields):
    np.testing.assert_allclose(b.temperature, a.temperature, rtol=1e-9, atol=0)
    assert b.count == a.count and b.time == a.time
    for f in fields:
        x, y = getattr(a, f), getattr(b, f)
        assert x.shape == y.shape, f
        np.testing.assert_allclose(y, x, rtol=1e-9, atol=1e-300, err_msg=f)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_restart_imc2d(L, tmp_path, precision):
    _announce(f"test_restart_imc2d {precision}")
    from radtranbook_amd import IMC2D

    IMC2D.VERBOSE = False
    n = 12
    e1, e2 = np.linspace(0, 1, n + 1), np.linspace(0, 1, n + 1)
    T0 = np.full((n, n), 1e-3)
    sf = lambda T: 50.0 * np.maximum(T, 1e-3) ** -3
    inv = lambda u: u / 0.3

    def init():
        return IMC2D.init_simulation(3000, T0, T0, e1, e2, lambda T: 0.3 * T, inv, precision=precision, seed=21)

    def run(st):
        return IMC2D.step(st, 3000, 3000, 0, 6000, (0.5, 0, 0, 0), 0.005, e1, e2, sf, inv, lambda T: 0.3 + 0 * T,
                          np.zeros((n, n)))[0]

    a, b = _interrupted(run, init, 4, 2, tmp_path / "ck.npz")
    _compare(a, b, ("weights", "dir1", "pos1", "pos2", "cell_i", "cell_j"))


def test_restart_imc1d(L, tmp_path):
    _announce("test_restart_imc1d")
    from radtranbook_amd import IMC1D

    IMC1D.VERBOSE = False
    I = 40
    mesh = np.column_stack([np.linspace(0, 0.2, I + 1)[:-1], np.linspace(0, 0.2, I + 1)[1:]])
    T0 = np.full(I, 1e-3)
    inv = lambda u: u / 0.3

    def init():
        return IMC1D.init_simulation(2000, T0, T0, mesh, lambda T: 0.3 * T, inv, seed=5)

    def run(st):
        return IMC1D.step(st, 2000, 2000, 0, 4000, (1.0, 0.0), 0.01, mesh, lambda T: 300 * np.maximum(T, 1e-3) ** -3,
                          inv, lambda T: 0.3 + 0 * T, np.zeros(I))[0]

    a, b = _interrupted(run, init, 4, 1, tmp_path / "ck1.npz")
    _compare(a, b, ("weights", "mus", "positions", "cell_indices"))


def test_restart_mg2d(L, tmp_path):
    _announce("test_restart_mg2d")
    from radtranbook_amd import MG_IMC2D

    MG_IMC2D.VERBOSE = False
    e1, e2 = np.linspace(0, 1, 9), np.linspace(0, 0.8, 7)
    T0 = np.full((8, 6), 0.05)
    sf = [(lambda g: (lambda T: np.minimum(3.0 * (NU[g] / np.maximum(T, 1e-3)) ** -3, 2e3)))(g) for g in range(10)]
    inv = lambda u: u / 0.3

    def init():
        return MG_IMC2D.init_simulation(2000, T0, T0, e1, e2, EDGES, lambda T: 0.3 * T, inv, precision="fp32",
                                        seed=8)

    def run(st):
        return MG_IMC2D.step(st, 2000, 2000, 0, 4000, (0.5, 0, 0, 0), 0.004, e1, e2, EDGES, sf, inv,
                             lambda T: 0.3 + 0 * T, np.zeros((8, 6)))[0]

    a, b = _interrupted(run, init, 3, 1, tmp_path / "ck2.npz")
    _compare(a, b, ("weights", "dir1", "pos1", "cell_i", "groups", "radiation_energy_by_group"))


def test_restart_mg1d(L, tmp_path):
    _announce("test_restart_mg1d")
    from radtranbook_amd import MG_IMC1D

    MG_IMC1D.VERBOSE = False
    I = 30
    r = np.linspace(0, 1.0, I + 1)
    mesh = np.column_stack([r[:-1], r[1:]])
    T0 = np.full(I, 0.05)
    funcs = [(lambda g: (lambda T: 1.0 + 0.5 * g + 0.0 * T))(g) for g in range(10)]
    inv = lambda u: u / 0.2

    def init():
        return MG_IMC1D.init_simulation(2000, T0, T0, mesh, EDGES, lambda T: 0.2 * T, inv, seed=9)

    def run(st):
        return MG_IMC1D.step(st, 2000, 2000, 0, 4000, (0.0, 0.8), 0.005, mesh, EDGES, funcs, inv,
                             lambda T: 0.2 + 0 * T, np.zeros(I))[0]

    a, b = _interrupted(run, init, 3, 2, tmp_path / "ck3.npz")
    _compare(a, b, ("weights", "mus", "positions", "groups", "radiation_flux_by_group"))
