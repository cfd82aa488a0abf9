"""Launch-shape autotuner for the persistent transport kernels.

The 1-D engine rebuilds its transport pool from the census and the step's
source blocks on every ``transport`` call, so a staged step can be tracked
several times with different launch shapes without changing any result
(histories are keyed by their Philox index, not by the lane or workgroup
that tracks them).  The tuner times each candidate shape with the kernel's
own HIP events and keeps the fastest.

Shape = (grid, batch, tally_copies):
  grid          persistent workgroups (0 = occupancy-derived, one full wave
                of workgroups over every CU),
  batch         particles a wave claims per visit to the global queue,
  tally_copies  private copies of the per-cell tallies (fewer copies = less
                memory to clear and reduce, more contention on atomics).
"""
from __future__ import annotations

import itertools

DEFAULT_GRIDS = (0, 1024, 4096)
DEFAULT_BATCHES = (64, 256)
DEFAULT_COPIES = (4, 16)


def tune_engine1d(engine, reflect, weight_floor=-1.0, grids=DEFAULT_GRIDS, batches=DEFAULT_BATCHES,
                  copies=DEFAULT_COPIES, repeats=1):
    """Time every shape on the engine's staged step; set and return the best.

    Call after the step's sources are sampled and before its ``transport``.
    Returns (best_shape, {shape: best ms})."""
    times = {}
    for shape in itertools.product(grids, batches, copies):
        engine.set_launch(*shape)
        best = float("inf")
        for _ in range(repeats):
            engine.transport(reflect, weight_floor, want_tallies=False)
            best = min(best, engine.last_transport_ms())
        times[shape] = best
    best_shape = min(times, key=times.get)
    engine.set_launch(*best_shape)
    return best_shape, times
