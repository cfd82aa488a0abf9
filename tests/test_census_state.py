"""Field-by-field census replacement on the drop-in state classes (CPU, stub
engine).  The reference's drivers replace a census one field at a time
(MG_IMC/problems/test_infinite_medium_multigroup_expband.py:187-195,
IMC/CrookedPipe2D.py:532); a larger or smaller replacement census must reach
the device whole, never cut to the shortest field."""
import numpy as np
import pytest

from radtranbook_amd.IMC1D import SimulationState
from radtranbook_amd.IMC2D import SimulationState2D
from radtranbook_amd.MG_IMC2D import SimulationState2DMG


class StubEngine:
    def __init__(self, fields, n, with_prob=False):
        self.with_prob = with_prob
        self.arrs = [np.ones(n) if f == "weights" else np.zeros(n) for f in fields]
        self.uploads = 0

    @property
    def n(self):
        return len(self.arrs[0])

    def get_census(self):
        out = [a.copy() for a in self.arrs]
        if self.with_prob:
            out.append(np.zeros(self.n, dtype=np.int32))
        return tuple(out)

    def put_census(self, *arrs):
        if self.with_prob:  # IMC1D: counts per problem first
            counts, arrs = arrs[0], arrs[1:]
            assert list(counts) == [len(arrs[0])]
        self.arrs = [np.asarray(a).copy() for a in arrs]
        self.uploads += 1

    def census_counts(self):
        return np.array([self.n]), self.n


def _new_state(cls, n):
    eng = StubEngine(cls._PARTICLE_FIELDS, n, with_prob=cls is SimulationState)
    if cls is SimulationState2DMG:
        st = cls(eng, None, None, None, None, 0.0, 0.0)
    else:
        st = cls(eng, None, None, None, 0.0, 0.0)
    return st, eng


@pytest.mark.parametrize("cls", [SimulationState2D, SimulationState, SimulationState2DMG])
def test_larger_census_assigned_field_by_field(cls):
    st, eng = _new_state(cls, 100)
    assert float(np.sum(st.weights)) == 100.0
    fields = cls._PARTICLE_FIELDS
    for k, f in enumerate(fields):
        val = np.full(150, 2.0) if f == "weights" else np.full(150, float(k))
        setattr(st, f, val)
        if k < len(fields) - 1:
            # lengths still disagree: nothing reaches the device, reads see the new field
            assert eng.n == 100
            assert len(getattr(st, f)) == 150
    assert eng.n == 150
    assert st.n_particles == 150
    assert float(np.sum(st.weights)) == 300.0
    for k, f in enumerate(fields):
        if f != "weights":
            assert np.all(getattr(st, f) == float(k))


@pytest.mark.parametrize("cls", [SimulationState2D, SimulationState, SimulationState2DMG])
def test_same_length_field_uploads_at_once(cls):
    st, eng = _new_state(cls, 64)
    st.weights = np.full(64, 3.0)
    assert eng.uploads == 1
    assert float(np.sum(st.weights)) == 192.0


@pytest.mark.parametrize("cls", [SimulationState2D, SimulationState, SimulationState2DMG])
def test_incomplete_replacement_raises_before_step(cls):
    st, eng = _new_state(cls, 100)
    st.weights = np.full(150, 2.0)
    with pytest.raises(ValueError, match="different lengths"):
        st._invalidate()  # the first thing every step() does
    with pytest.raises(ValueError, match="different lengths"):
        _ = st.n_particles
    assert eng.n == 100 and eng.uploads == 0
