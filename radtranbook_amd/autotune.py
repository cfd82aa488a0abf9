"""Launch-shape autotuner for the persistent transport kernels.

A shape is (grid, batch, tally_copies, block, refill_min):
  grid          persistent workgroups (0 = occupancy-derived: every CU full),
  batch         particles a wave claims per visit to the global queue (also
                the chunk of the longest-first pickup order),
  tally_copies  private copies of the per-cell tallies,
  block         threads per workgroup of the gray 2-D tracker (256 or 512;
                its waves share the LDS tally cache), 256 elsewhere,
  refill_min    idle lanes a wave with live histories waits for before it
                takes new particles (the events per refill pass).

Results do not depend on the shape: every history draws from its own Philox
stream, whichever lane or workgroup tracks it.  So a staged step (census
plus the step's sources) can be transported again and again: the tuner keeps
the pool (the engine's snapshot), and at each sample step of a run of the
workload itself times every candidate on that pool, ``repeats`` times with
the candidates in rotated order, the pool put back before each trial.  A
candidate's cost is the sum over the sample steps of its median kernel time;
the default shape is always a candidate, so the pick is never slower than it
on those inputs.  Picks are kept in a JSON cache (``cache_path()``) under a
key naming the engine, the workload and the precision.
"""
from __future__ import annotations

import json
import os
from dataclasses import asdict, dataclass

import numpy as np


@dataclass(frozen=True)
class Shape:
    grid: int = 0
    batch: int = 256
    tally_copies: int = 8
    block: int = 256
    refill_min: int = 1


def default_shape(kind):
    """The engines' built-in shape (what runs without tuning)."""
    return Shape(tally_copies=32) if kind == "mg2d" else Shape()


def candidates(kind):
    """Candidate shapes per engine kind ("gray2d", "mg2d", "gray1d"); the
    default first."""
    d = default_shape(kind)
    out = [d]
    if kind == "gray2d":
        for block in (256, 512):
            for batch in (128, 256, 512):
                for refill in (1, 8, 32):
                    out.append(Shape(0, batch, 8, block, refill))
    elif kind == "mg2d":
        for batch in (128, 256, 512):
            for copies in (16, 32):
                for refill in (1, 8, 32):
                    out.append(Shape(0, batch, copies, 256, refill))
    elif kind == "gray1d":
        for grid in (0, 4096):
            for batch in (64, 256):
                for copies in (4, 8, 16):
                    for refill in (1, 16):
                        out.append(Shape(grid, batch, copies, 256, refill))
    else:
        raise ValueError(f"unknown engine kind {kind!r}")
    seen, uniq = set(), []
    for s in out:
        if s not in seen:
            seen.add(s)
            uniq.append(s)
    return uniq


def apply(engine, shape):
    engine.set_launch(shape.grid, shape.batch, shape.tally_copies)
    engine.set_tracker(shape.block, shape.refill_min)


class StepTuner:
    """Times candidate shapes on staged steps of a run.

    ``transport(engine)`` runs the engine's transport once on its resident
    pool and returns the kernel time in ms; ``keep(engine, save)`` saves
    (True) or puts back (False) the pool; ``restage`` (optional) rebuilds the
    pool instead of ``keep`` (engines that stage on every transport).  Call
    ``sample(engine)`` at each sample step, after the sources are sampled
    and before the step's own transport; then ``best()``."""

    def __init__(self, kind, transport, keep=None, repeats=3, shapes=None):
        self.kind = kind
        self.shapes = list(shapes) if shapes is not None else candidates(kind)
        self.transport = transport
        self.keep = keep
        self.repeats = int(repeats)
        self.samples = []  # per sample step: {shape: median ms}

    def sample(self, engine):
        times = {s: [] for s in self.shapes}
        if self.keep is not None:
            self.keep(engine, True)
        n = len(self.shapes)
        for r in range(self.repeats):
            for q in range(n):
                s = self.shapes[(q + r) % n]
                if self.keep is not None:
                    self.keep(engine, False)
                apply(engine, s)
                times[s].append(float(self.transport(engine)))
        if self.keep is not None:
            self.keep(engine, False)
        apply(engine, self.shapes[0])
        self.samples.append({s: float(np.median(v)) for s, v in times.items()})

    def totals(self):
        return {s: float(sum(smp[s] for smp in self.samples)) for s in self.shapes}

    def best(self):
        tot = self.totals()
        return min(self.shapes, key=lambda s: (tot[s], self.shapes.index(s)))

    def report(self):
        tot = self.totals()
        b = self.best()
        d = self.shapes[0]
        return {"shape": asdict(b), "default": asdict(d), "tuned_ms": tot[b], "default_ms": tot[d],
                "sample_steps": len(self.samples), "candidates": len(self.shapes), "repeats": self.repeats}


# -- engine adapters ----------------------------------------------------------
def gray2d_tuner(dt, reflect, max_events, repeats=3):
    def transport(eng):
        eng.transport(dt, reflect, max_events, -1.0, want_tallies=False)
        return eng.last_transport_ms()

    return StepTuner("gray2d", transport, lambda eng, save: eng.snapshot(save), repeats)


def mg2d_tuner(dt, reflect, max_events, repeats=3):
    def transport(eng):
        eng.transport(dt, reflect, max_events, -1.0, want_tallies=False)
        return eng.last_transport_ms()

    return StepTuner("mg2d", transport, lambda eng, save: eng.snapshot(save), repeats)


def gray1d_tuner(reflect, repeats=3):
    # the 1-D engine rebuilds its transport pool from the census and the
    # step's source blocks on every transport: nothing to keep
    def transport(eng):
        eng.transport(reflect, -1.0, want_tallies=False)
        return eng.last_transport_ms()

    return StepTuner("gray1d", transport, None, repeats)


# -- cache --------------------------------------------------------------------
def cache_path():
    p = os.environ.get("RADTRANBOOK_AMD_TUNE_CACHE")
    if p:
        return p
    return os.path.join(os.path.expanduser("~"), ".cache", "radtranbook_amd", "launch_shapes.json")


def cache_key(kind, workload, precision):
    return f"{kind}|{workload}|{precision}"


def load_report(kind, workload, precision):
    """The cached tuner report (its "shape" checked), or None."""
    try:
        with open(cache_path()) as f:
            d = json.load(f)
        rep = d[cache_key(kind, workload, precision)]
        Shape(**rep["shape"])
        return rep
    except (OSError, KeyError, TypeError, ValueError):
        return None


def load_shape(kind, workload, precision):
    """The cached pick, or None."""
    rep = load_report(kind, workload, precision)
    return None if rep is None else Shape(**rep["shape"])


def store_shape(kind, workload, precision, report):
    """Keep a tuner report in the cache (written to a temporary file, then
    renamed)."""
    path = cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        try:
            with open(path) as f:
                d = json.load(f)
        except (OSError, ValueError):
            d = {}
        d[cache_key(kind, workload, precision)] = report
        tmp = path + f".tmp{os.getpid()}"
        with open(tmp, "w") as f:
            json.dump(d, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass


def tune_engine1d(engine, reflect, repeats=3):
    """Tune the batched 1-D engine on its staged step (call after the step's
    sources are sampled, before its transport); set and return the best
    shape with the tuner's report."""
    t = gray1d_tuner(reflect, repeats)
    t.sample(engine)
    b = t.best()
    apply(engine, b)
    return b, t.report()
