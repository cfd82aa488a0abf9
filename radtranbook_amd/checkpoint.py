"""Checkpoint and restart of a drop-in simulation state.

One ``.npz`` file (numeric and unicode arrays only, nothing pickled) holds
the material fields and scalars of the state, the whole census (every
particle field at the engine's precision, so FP32-track values round-trip
exactly through float64) and the random-stream position (seed, step count).
It is written to ``path + ".tmp"`` and renamed, so a run killed mid-save
leaves the previous checkpoint intact (the reference's
``MG_IMC/mg_utils.py:275-292`` does the same with a pickle).

The reference's ``IMC/CrookedPipe2D.py:459-570`` saves census attributes the
2-D state does not have (``mu``, ``eta``, ``r``, ``z``), so its restart loses
the census.  Here a state restored from step k continues with the census it
had and draws the streams of steps k, k+1, ... of the same seed, as the
uninterrupted run does; the restart is exact up to the summation order of
the FP64 tally atomics.

Usage::

    from radtranbook_amd import checkpoint
    checkpoint.save_checkpoint(state, "run.npz", extra={"dt": dt})
    state, extra = checkpoint.load_checkpoint("run.npz")
"""
from __future__ import annotations

import os

import numpy as np

from . import _lib

FORMAT = "radtranbook_amd checkpoint 1"

_PRECISION = {_lib.FP64: "fp64", _lib.FP32_TRACK: "fp32"}

# state class -> (drop-in module, material arrays that may be None, scalars)
_KINDS = {
    "SimulationState2D": ("IMC2D", ("internal_energy", "temperature", "radiation_temperature"),
                          ("time", "previous_total_energy", "count")),
    "SimulationState": ("IMC1D", ("internal_energy", "temperature", "radiation_temperature"),
                        ("time", "previous_total_energy", "count")),
    "SimulationState2DMG": ("MG_IMC2D", ("internal_energy", "temperature", "radiation_temperature",
                                         "radiation_energy_by_group", "radiation_energy_by_group_postcomb"),
                            ("time", "previous_total_energy", "count")),
    "SimulationState1DMG": ("MG_IMC1D", ("internal_energy", "temperature", "radiation_temperature",
                                         "radiation_energy_by_group", "radiation_flux_by_group",
                                         "radiation_energy_by_group_postcomb"),
                            ("time", "previous_total_energy", "count")),
}


def _engine_record(kind, eng):
    rec = {"precision": _PRECISION[eng.precision], "seed": np.int64(eng.seed)}
    if kind == "IMC2D":
        rec.update(geometry=eng.geometry, edges1=eng.e1, edges2=eng.e2)
    elif kind == "IMC1D":
        if eng.P != 1:
            raise ValueError("save_checkpoint(): an IMC1D state holds one problem")
        rec.update(geometry=eng.geometry, edges=eng.edges[0])
    elif kind == "MG_IMC2D":
        rec.update(geometry=eng.geometry, edges1=eng.e1, edges2=eng.e2, energy_edges=eng.edges)
    else:
        rec.update(edges=eng.e, energy_edges=eng.edges)
    return rec


def save_checkpoint(state, path, extra=None):
    """Write ``state`` (any drop-in module's simulation state) to ``path``.

    ``extra``: optional dict of numbers or arrays stored beside the state
    (time-step size, output bookkeeping, ...) and returned by
    load_checkpoint."""
    kind_name = type(state).__name__
    if kind_name not in _KINDS:
        raise TypeError(f"save_checkpoint(): not a drop-in simulation state: {kind_name}")
    kind, fields, scalars = _KINDS[kind_name]
    state._flush()  # pending census field assignments reach the device first
    eng = state._engine
    data = {"format": np.array(FORMAT), "kind": np.array(kind)}
    for k, v in _engine_record(kind, eng).items():
        data["engine_" + k] = np.asarray(v)
    for f in fields:
        v = getattr(state, f)
        data["has_" + f] = np.array(v is not None)
        if v is not None:
            data["state_" + f] = np.asarray(v, dtype=np.float64)
    for sname in scalars:
        data["scalar_" + sname] = np.asarray(getattr(state, sname))
    census = eng.get_census()
    if kind == "IMC1D":
        census = census[:5]
    data["census_fields"] = np.array(len(census))
    for a, arr in enumerate(census):
        data[f"census_{a}"] = np.asarray(arr)
    for k, v in (extra or {}).items():
        data["extra_" + k] = np.asarray(v)
    tmp = str(path) + ".tmp"
    with open(tmp, "wb") as fh:
        np.savez(fh, **data)
    os.replace(tmp, str(path))


def _new_engine(kind, rec, device):
    prec = str(rec["precision"])
    seed = int(rec["seed"])
    if kind == "IMC2D":
        return _lib.Engine2D(rec["edges1"], rec["edges2"], geometry=str(rec["geometry"]), precision=prec, seed=seed,
                             device=device)
    if kind == "IMC1D":
        return _lib.Engine1D([rec["edges"]], geometry=str(rec["geometry"]), precision=prec, seed=seed, device=device)
    if kind == "MG_IMC2D":
        return _lib.EngineMG(rec["edges1"], rec["edges2"], rec["energy_edges"], geometry=str(rec["geometry"]),
                             precision=prec, seed=seed, device=device)
    return _lib.EngineMG1D(rec["edges"], rec["energy_edges"], precision=prec, seed=seed, device=device)


def _check_engine(kind, rec, eng):
    mine = _engine_record(kind, eng)
    for k, v in rec.items():
        if k == "seed":
            continue
        a, b = np.asarray(mine[k]), np.asarray(v)
        if a.shape != b.shape or not np.array_equal(a, b):
            raise ValueError(f"load_checkpoint(): the engine differs from the checkpoint in {k}")


def load_checkpoint(path, device=0, engine=None):
    """Rebuild the simulation state saved in ``path``; returns (state, extra).

    ``engine``: an existing device context of the same mesh, geometry,
    groups and precision to reuse (its census is replaced); otherwise a new
    one is created on ``device``."""
    with np.load(str(path), allow_pickle=False) as z:
        if "format" not in z.files or str(z["format"]) != FORMAT:
            raise ValueError(f"load_checkpoint(): {path} is not a {FORMAT} file")
        kind = str(z["kind"])
        module_name, fields, scalars = next(v for v in _KINDS.values() if v[0] == kind)
        rec = {k[len("engine_"):]: z[k] for k in z.files if k.startswith("engine_")}
        vals = {f: (z["state_" + f].copy() if bool(z["has_" + f]) else None) for f in fields}
        sc = {s: z["scalar_" + s][()] for s in scalars}
        census = [z[f"census_{a}"].copy() for a in range(int(z["census_fields"]))]
        extra = {k[len("extra_"):]: z[k][()] if z[k].ndim == 0 else z[k].copy() for k in z.files
                 if k.startswith("extra_")}
    if engine is None:
        eng = _new_engine(kind, rec, device)
    else:
        _check_engine(kind, rec, engine)
        eng = engine
        eng.set_seed(int(rec["seed"]))
    if kind == "IMC1D":
        eng.put_census([census[0].size], *census)
    else:
        eng.put_census(*census)
    from importlib import import_module

    mod = import_module("." + module_name, __package__)
    cls = {"IMC2D": "SimulationState2D", "IMC1D": "SimulationState", "MG_IMC2D": "SimulationState2DMG",
           "MG_IMC1D": "SimulationState1DMG"}[kind]
    count = int(sc["count"])
    time = float(sc["time"])
    prev = float(sc["previous_total_energy"])
    if kind in ("IMC2D", "IMC1D"):
        state = getattr(mod, cls)(eng, vals["internal_energy"], vals["temperature"], vals["radiation_temperature"],
                                  time, prev, count)
    elif kind == "MG_IMC2D":
        state = getattr(mod, cls)(eng, vals["internal_energy"], vals["temperature"], vals["radiation_temperature"],
                                  vals["radiation_energy_by_group"], time, prev, count)
        state.radiation_energy_by_group_postcomb = vals["radiation_energy_by_group_postcomb"]
    else:
        state = getattr(mod, cls)(eng, vals["internal_energy"], vals["temperature"], vals["radiation_temperature"],
                                  vals["radiation_energy_by_group"], vals["radiation_flux_by_group"], time, prev,
                                  count)
        state.radiation_energy_by_group_postcomb = vals["radiation_energy_by_group_postcomb"]
    eng.set_step(count)
    return state, extra
