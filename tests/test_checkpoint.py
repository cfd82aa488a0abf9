"""Checkpoint file round trip on the CPU: a stand-in engine with the drop-in
engine interface (no device), saved and restored through the .npz format."""
import os

import numpy as np
import pytest


class FakeEngine2D:
    """The parts of _lib.Engine2D that checkpointing uses."""

    def __init__(self, e1, e2, geometry="xy", precision=0, seed=5, census=None):
        self.e1, self.e2, self.geometry, self.precision, self.seed = e1, e2, geometry, precision, seed
        self.census = census
        self.step = None

    def get_census(self):
        return self.census

    def put_census(self, *arrs):
        self.census = tuple(np.array(a) for a in arrs)

    def set_seed(self, seed):
        self.seed = seed

    def set_step(self, step):
        self.step = step


def _state(rng):
    from radtranbook_amd import IMC2D

    e1, e2 = np.linspace(0, 1, 6), np.linspace(0, 2, 4)
    n = 37
    census = (rng.uniform(0, 1, n), rng.uniform(-1, 1, n), rng.uniform(-1, 1, n), rng.uniform(0, 0.01, n),
              rng.uniform(0, 1, n), rng.uniform(0, 2, n), rng.integers(0, 5, n), rng.integers(0, 3, n))
    eng = FakeEngine2D(e1, e2, precision=1, seed=1234567890123, census=census)
    T = rng.uniform(0.1, 1, (5, 3))
    return IMC2D.SimulationState2D(eng, 0.3 * T, T, 0.9 * T, 0.125, 4.5, 7), census


def test_round_trip(tmp_path):
    from radtranbook_amd import checkpoint

    rng = np.random.default_rng(3)
    st, census = _state(rng)
    path = tmp_path / "ck.npz"
    checkpoint.save_checkpoint(st, path, extra={"dt": 0.005, "outputs": np.arange(3.0)})
    assert path.exists() and not os.path.exists(str(path) + ".tmp")
    eng2 = FakeEngine2D(st._engine.e1.copy(), st._engine.e2.copy(), precision=1, seed=1)
    st2, extra = checkpoint.load_checkpoint(path, engine=eng2)
    assert type(st2).__name__ == "SimulationState2D"
    for f in ("internal_energy", "temperature", "radiation_temperature"):
        assert np.array_equal(getattr(st2, f), getattr(st, f))
    assert (st2.time, st2.previous_total_energy, st2.count) == (0.125, 4.5, 7)
    assert eng2.seed == 1234567890123 and eng2.step == 7
    for a, b in zip(eng2.census, census):
        assert a.dtype == np.asarray(b).dtype and np.array_equal(a, b)
    assert extra["dt"] == 0.005 and np.array_equal(extra["outputs"], np.arange(3.0))


def test_rejects_other_mesh_and_files(tmp_path):
    from radtranbook_amd import checkpoint

    st, _ = _state(np.random.default_rng(4))
    path = tmp_path / "ck.npz"
    checkpoint.save_checkpoint(st, path)
    other = FakeEngine2D(np.linspace(0, 1, 7), st._engine.e2.copy(), precision=1)
    with pytest.raises(ValueError, match="edges1"):
        checkpoint.load_checkpoint(path, engine=other)
    fp64 = FakeEngine2D(st._engine.e1.copy(), st._engine.e2.copy(), precision=0)
    with pytest.raises(ValueError, match="precision"):
        checkpoint.load_checkpoint(path, engine=fp64)
    bad = tmp_path / "bad.npz"
    np.savez(bad, x=np.zeros(3))
    with pytest.raises(ValueError, match="not a"):
        checkpoint.load_checkpoint(bad)
    with pytest.raises(TypeError):
        checkpoint.save_checkpoint(object(), tmp_path / "x.npz")

