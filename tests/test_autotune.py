"""Launch-shape autotuner: candidate sets, the pick (never slower than the
default on the sampled steps), the engine settings it applies, the pool it
puts back before each trial, and the JSON cache of picks."""
import json

import pytest

from radtranbook_amd import autotune as AT


@pytest.mark.parametrize("kind", ["gray2d", "mg2d", "gray1d"])
def test_candidates_unique_default_first(kind):
    c = AT.candidates(kind)
    assert c[0] == AT.default_shape(kind)
    assert len(set(c)) == len(c) > 1
    for s in c:
        assert s.block in (256, 512) and s.batch > 0 and s.tally_copies > 0 and 1 <= s.refill_min <= 64
        if kind != "gray2d":
            assert s.block == 256  # the MG and 1-D trackers take 256-thread workgroups only


def test_unknown_kind_rejected():
    with pytest.raises(ValueError):
        AT.candidates("gray3d")


class _FakeEngine:
    """Records the launch settings; a trial's time is a fixed function of the
    shape and the sample step, plus a trial counter that detects a pool that
    was not put back before a trial."""

    def __init__(self, cost):
        self.cost = cost
        self.launch = None
        self.tracker = None
        self.step = 0
        self.pool = "staged"
        self.saved = None
        self.trials = []

    def set_launch(self, grid, batch, copies):
        self.launch = (grid, batch, copies)

    def set_tracker(self, block, refill_min):
        self.tracker = (block, refill_min)

    def snapshot(self, save):
        if save:
            self.saved = self.pool
        else:
            assert self.saved is not None
            self.pool = self.saved


def _transport(eng):
    assert eng.pool == "staged", "a trial ran on a pool left over by the previous trial"
    eng.pool = "transported"
    s = AT.Shape(eng.launch[0], eng.launch[1], eng.launch[2], eng.tracker[0], eng.tracker[1])
    eng.trials.append(s)
    return eng.cost(s, eng.step)


def test_pick_is_the_minimum_of_summed_medians_and_default_restored():
    shapes = [AT.Shape(), AT.Shape(batch=128), AT.Shape(block=512, refill_min=8), AT.Shape(batch=512)]
    # batch=128 wins step 0 by far but loses steps 1 and 2: summed medians pick block 512
    table = {0: [10.0, 5.0, 9.0, 11.0], 1: [10.0, 12.0, 9.0, 11.0], 2: [10.0, 12.0, 9.0, 11.0]}

    def cost(s, step):
        return table[step][shapes.index(s)]

    eng = _FakeEngine(cost)
    t = AT.StepTuner("gray2d", _transport, lambda e, save: e.snapshot(save), repeats=3, shapes=shapes)
    for step in range(3):
        eng.step = step
        eng.pool = "staged"
        t.sample(eng)
        # every candidate timed `repeats` times, the pool put back afterwards,
        # the default shape set again for the step's own transport
        assert eng.pool == "staged"
        assert eng.launch == (0, 256, 8) and eng.tracker == (256, 1)
    assert len(eng.trials) == 3 * 3 * len(shapes)
    for s in shapes:
        assert eng.trials.count(s) == 9
    tot = t.totals()
    assert tot[shapes[1]] == pytest.approx(29.0) and tot[shapes[2]] == pytest.approx(27.0)
    assert t.best() == shapes[2]
    rep = t.report()
    assert rep["tuned_ms"] <= rep["default_ms"]
    assert rep["shape"]["block"] == 512 and rep["sample_steps"] == 3


def test_rotated_order_and_median_reject_a_slow_outlier():
    shapes = [AT.Shape(), AT.Shape(batch=128)]
    calls = []

    def cost(s, step):
        calls.append(s)
        # the first trial of the run is slow (a cold cache): the median drops it
        return 100.0 if len(calls) == 1 else (4.0 if s == shapes[1] else 5.0)

    eng = _FakeEngine(cost)
    t = AT.StepTuner("gray2d", _transport, lambda e, save: e.snapshot(save), repeats=3, shapes=shapes)
    t.sample(eng)
    # rotation: repeat r starts with candidate r mod n
    assert calls[0] == shapes[0] and calls[2] == shapes[1] and calls[4] == shapes[0]
    assert t.best() == shapes[1]
    assert t.totals()[shapes[0]] == pytest.approx(5.0)


def test_ties_keep_the_default():
    shapes = [AT.Shape(), AT.Shape(batch=128)]
    eng = _FakeEngine(lambda s, step: 3.0)
    t = AT.StepTuner("mg2d", _transport, lambda e, save: e.snapshot(save), repeats=1, shapes=shapes)
    t.sample(eng)
    assert t.best() == shapes[0]


def test_cache_round_trip(tmp_path, monkeypatch):
    path = tmp_path / "sub" / "shapes.json"
    monkeypatch.setenv("RADTRANBOOK_AMD_TUNE_CACHE", str(path))
    assert AT.cache_path() == str(path)
    assert AT.load_shape("gray2d", "marshak", "fp32") is None
    rep = {"shape": {"grid": 0, "batch": 128, "tally_copies": 8, "block": 512, "refill_min": 8}, "tuned_ms": 1.0}
    AT.store_shape("gray2d", "marshak", "fp32", rep)
    AT.store_shape("gray2d", "marshak", "fp64", {"shape": {"grid": 0, "batch": 256, "tally_copies": 8,
                                                           "block": 256, "refill_min": 1}})
    assert AT.load_shape("gray2d", "marshak", "fp32") == AT.Shape(0, 128, 8, 512, 8)
    assert AT.load_shape("gray2d", "marshak", "fp64") == AT.Shape()
    assert AT.load_shape("mg2d", "marshak", "fp32") is None
    d = json.loads(path.read_text())
    assert set(d) == {"gray2d|marshak|fp32", "gray2d|marshak|fp64"}
    # a damaged cache reads as empty and is replaced on the next store
    path.write_text("{not json")
    assert AT.load_shape("gray2d", "marshak", "fp32") is None
    AT.store_shape("gray1d", "slab", "fp64", rep)
    assert AT.load_shape("gray1d", "slab", "fp64") == AT.Shape(0, 128, 8, 512, 8)
