#!/usr/bin/env python3
"""Benchmark: gray 2-D heterogeneous Marshak IMC, 200x200 cells, 1e7 particles
per step, FP32-track/FP64-tally against plain FP64, on one MI355X.

One "step" is one IMC time step (Fleck factors, boundary + emission sourcing,
transport of every particle to census, material update, comb) of the
200x200 problem with the census resident in HBM.  W warm-up steps run on a
separate seed; the K timed steps then start from t = 0.

Prints one JSON line.  ``value`` = transported segments per second of the
FP32-track step (whole step, host orchestration included).  FOM = 1/(R^2 t)
with R^2 the median per-cell relative variance of the final-step absorbed
energy over ``--fom-seeds`` independent Philox seeds (32 by default) and t the
timed wall time (``fom``) or the summed transport-kernel time of the timed
steps (``fom_transport``).  The CPU baseline is the FP64 C oracle (OpenMP) on
a bounded sample of the same workload.  ``roofline.traffic`` is the HBM bytes
per transport launch that rocprofv3 measured for the same workload and timed
region (profiles/), null for other sizes.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "IMC FOM & segments/s, 2D 200² cells 1e7 particles: FP32-track vs FP64 GPU"
BYTES_PER_SEGMENT = {"fp32": 104, "fp64": 152}
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def profiled(profile, precision, sizes_match):
    """Figures rocprofv3 measured for this workload (the committed summary in
    profiles/): HBM bytes per segment of the transport kernel (FETCH_SIZE +
    WRITE_SIZE per dispatch, each scaled by the same counter's reading for a
    copy kernel of known size, over the timed launches) and, where recorded,
    its VALU busy fraction and active-lane fraction.  None unless this run
    has the profiled sizes."""
    if not sizes_match:
        return None
    try:
        with open(os.path.join(ROOT, "profiles", profile)) as fh:
            text = fh.read()
    except OSError:
        return None
    out = {}
    m = re.search(r"^bytes_per_segment %s (\S+)" % precision, text, re.M)
    if m:
        out["bytes_per_segment"] = float(m.group(1))
    m = re.search(r"^valu_busy %s (\S+) lanes_active (\S+)" % precision, text, re.M)
    if m:
        out["valu_busy"] = float(m.group(1))
        out["lanes_active"] = float(m.group(2))
    return out or None


# Algorithmic HBM bytes per history of each tracker: the history's state read
# once when a lane picks it up and written once at census (a persistent
# tracker keeps the history in registers in between).  The per-cell material
# and tallies are a few MB per launch and are left out.
BYTES_PER_HISTORY = {
    "k_transport<float,xy>": 36 + 32, "k_transport<double,xy>": 56 + 48,
    "k_mg_transport<float,xy>": 40 + 32, "k_mg_transport<double,xy>": 60 + 48,
    "k1_transport<float,0>": 28 + 20, "k1_transport<double,0>": 40 + 28,
    "k_transport_cf<xy>": 56 + 56, "k1_transport_cf<0>": 40 + 36,
}


def roofline(kernel, precision, seg_per_launch, hist_per_launch, avg_launch_s, model_bps, profile, sizes_match):
    """The roofline object of the bench line.  ``achieved``/``frac``: the
    algorithmic bytes per launch (BYTES_PER_HISTORY x histories) over the
    launch's average duration.  ``traffic``: HBM bytes per launch that
    rocprofv3 measured (bytes per segment x this run's segments per launch),
    ``traffic_frac`` the bandwidth that moves.  ``issue_frac``: the VALU busy
    fraction rocprofv3 measured (the kernels are issue/latency bound).  The
    per-segment event model (every segment loads and stores the state) is kept
    as ``model_*``."""
    alg = hist_per_launch * BYTES_PER_HISTORY[kernel]
    achieved = alg / avg_launch_s / 1e9
    meas = profiled(profile, precision, sizes_match)
    traffic = meas["bytes_per_segment"] * seg_per_launch if meas and "bytes_per_segment" in meas else None
    model = seg_per_launch * model_bps / avg_launch_s / 1e9
    return {
        "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
        "traffic": traffic,
        "traffic_frac": traffic / avg_launch_s / 1e9 / HBM_PEAK_GBS if traffic else None,
        "issue_frac": meas.get("valu_busy") if meas else None,
        "lanes_active_frac": meas.get("lanes_active") if meas else None,
        "kernel": kernel, "algorithmic_bytes_per_history": BYTES_PER_HISTORY[kernel],
        "histories_per_launch": hist_per_launch, "segments_per_launch": seg_per_launch,
        "model_bytes_per_segment": model_bps, "model_achieved": model, "model_frac": model / HBM_PEAK_GBS,
    }


def _sample_steps(steps):
    return sorted({max(steps // 4, 0), steps // 2, (3 * steps) // 4, steps - 1})


RETUNE = False  # --retune: tune even where a pick is cached


def cached_pick(kind, workload, precision):
    """(shape, report) kept by an earlier tuning run of this workload and
    precision, or None (none kept, or --retune)."""
    from radtranbook_amd import autotune as AT

    if RETUNE:
        return None
    rep = AT.load_report(kind, workload, precision)
    if rep is None:
        return None
    log(f"[{precision}] autotune: cached pick {rep['shape']} ({AT.cache_path()})")
    return AT.Shape(**rep["shape"]), dict(rep, cached=True)


def tune_key_2d(prob):
    return f"{prob.name}_N{prob.Ntarget}_Nb{prob.Nboundary}_Nmax{prob.Nmax}"


def tune_2d(prob, precision, steps):
    """Tuning pass: the workload's own run (timed seed) with the launch-shape
    candidates timed on the staged pool of four of its steps."""
    from radtranbook_amd import IMC2D
    from radtranbook_amd import autotune as AT

    IMC2D.VERBOSE = False
    t = AT.gray2d_tuner(prob.dt, prob.reflect, prob.max_events)
    e1, e2 = prob.edges1, prob.edges2
    T0 = np.full(prob.shape, prob.T0)
    src = np.zeros(prob.shape)
    samples = _sample_steps(steps)
    cur = [0]

    def hook(eng):
        if cur[0] in samples:
            t.sample(eng)

    st = IMC2D.init_simulation(prob.Ntarget, T0, T0, e1, e2, prob.eos.eos, prob.eos.inv_eos, geometry=prob.geometry,
                               precision=precision, seed=1)
    IMC2D._TRANSPORT_HOOK = hook
    try:
        for s in range(steps):
            cur[0] = s
            st, _ = IMC2D.step(st, prob.Ntarget, prob.Nboundary, 0, prob.Nmax, prob.T_boundary, prob.dt, e1, e2,
                               prob.opacity, prob.eos.inv_eos, prob.eos.cv_func, src, reflect=prob.reflect,
                               geometry=prob.geometry, max_events_per_particle=prob.max_events)
    finally:
        IMC2D._TRANSPORT_HOOK = None
        st._engine.close()
    rep = t.report()
    AT.store_shape("gray2d", tune_key_2d(prob), precision, rep)
    log(f"[{precision}] autotune: {rep['shape']} {rep['tuned_ms']:.2f} ms over {rep['sample_steps']} sample steps "
        f"(default {rep['default_ms']:.2f} ms, {rep['candidates']} candidates x {rep['repeats']})")
    return t.best(), rep


def run_gpu(prob, precision, steps, warmup, fom_seeds, launch=None):
    from radtranbook_amd import IMC2D
    from radtranbook_amd import autotune as AT

    IMC2D.VERBOSE = False
    e1, e2 = prob.edges1, prob.edges2
    shape = prob.shape
    T0 = np.full(shape, prob.T0)
    src = np.zeros(shape)
    opac, eos = prob.opacity, prob.eos

    engine = [None]

    def one_run(seed, nsteps, timed):
        # every run after the first reuses the first run's device context
        # (same buffers, new Philox seed), so no allocation is timed
        st = IMC2D.init_simulation(prob.Ntarget, T0, T0, e1, e2, eos.eos, eos.inv_eos, geometry=prob.geometry,
                                   precision=precision, seed=seed, engine=engine[0])
        if engine[0] is None and launch is not None:
            AT.apply(st._engine, launch)
        engine[0] = st._engine
        infos = []
        t0 = time.perf_counter()
        for s in range(nsteps):
            st, info = IMC2D.step(st, prob.Ntarget, prob.Nboundary, 0, prob.Nmax, prob.T_boundary, prob.dt, e1, e2,
                                  opac, eos.inv_eos, eos.cv_func, src, reflect=prob.reflect, geometry=prob.geometry,
                                  max_events_per_particle=prob.max_events)
            infos.append(info)
            if timed:
                te = info["profiling"]["transport_events"]
                log(f"  [{precision}] step {s + 1}/{nsteps}: {info['profiling']['phase_times_s']['total'] * 1e3:.1f} ms,"
                    f" transport kernel {info['profiling']['transport_kernel_ms']:.1f} ms, segments {te['total']:.3e},"
                    f" histories {te['n_particles_transported']:.3e}")
        wall = time.perf_counter() - t0
        return infos, wall

    if warmup > 0:
        log(f"[{precision}] warmup {warmup} steps")
        one_run(1000, warmup, False)
    log(f"[{precision}] timed {steps} steps")
    infos, wall = one_run(1, steps, True)
    segs = sum(i["profiling"]["transport_events"]["total"] for i in infos)
    hist = sum(i["profiling"]["transport_events"]["n_particles_transported"] for i in infos)
    kern_ms = np.array([i["profiling"]["transport_kernel_ms"] for i in infos])
    seg_per_launch = np.array([i["profiling"]["transport_events"]["total"] for i in infos])
    residual = max(abs(i["energy_residual"]) / i["total_energy"] for i in infos)
    deps = [infos[-1]["deposited"]]
    temps = [infos[-1]["temperature"]]
    for s in range(2, fom_seeds + 1):
        log(f"[{precision}] FOM seed {s}/{fom_seeds}")
        inf_s, _ = one_run(s, steps, False)
        deps.append(inf_s[-1]["deposited"])
        temps.append(inf_s[-1]["temperature"])
    engine[0].close()
    return dict(wall=wall, segments=int(segs), histories=int(hist), kernel_ms=kern_ms, seg_per_launch=seg_per_launch,
                residual=residual, deps=np.array(deps), temps=np.array(temps))


def fom_cells(results, T0):
    """Cells the radiation front has heated (ensemble-mean final T above 10 T0
    in every precision) with E_abs above 1e-3 of the maximum: the same set for
    every precision.  Cold cells that only re-absorb their own emission have a
    deterministic tally and carry no Monte Carlo noise."""
    mask = None
    for r in results.values():
        mean = r["deps"].mean(0)
        m = (r["temps"].mean(0) > 10.0 * T0) & (mean > 1e-3 * mean.max())
        mask = m if mask is None else (mask & m)
    return mask


def fom_of(r, mask):
    """R^2 and FOM = 1/(R^2 t) with t the wall time of the timed steps, and
    the same with t the summed transport-kernel time of those steps."""
    deps = r["deps"]
    if len(deps) < 2 or not mask.any():
        return float("nan"), None, None
    mean = deps.mean(0)[mask]
    var = deps.var(0, ddof=1)[mask]
    R2 = float(np.median(var / mean ** 2))
    if not R2 > 0:
        return R2, None, None
    return R2, 1.0 / (R2 * r["wall"]), 1.0 / (R2 * r["kernel_ms"].sum() / 1e3)


def run_cpu_baseline(prob, n_particles, nsteps):
    from oracle import oracle as O

    O.build()
    cores = len(os.sched_getaffinity(0))
    e1, e2 = prob.edges1, prob.edges2
    shape = prob.shape
    T0 = np.full(shape, prob.T0)
    orc = O.Oracle2D(e1, e2, prob.geometry, seed=1, nthreads=cores)
    N = int(n_particles)
    st = orc.init_simulation(N // 2, T0, T0, prob.eos.eos, prob.eos.inv_eos)
    segs = 0
    tt = 0.0
    for s in range(nsteps):
        t0 = time.perf_counter()
        st, info = orc.step(st, N // 2, N // 2, 0, N, prob.T_boundary, prob.dt, prob.opacity, prob.eos.inv_eos,
                            prob.eos.cv_func, np.zeros(shape), reflect=prob.reflect)
        dt = time.perf_counter() - t0
        if s >= 1:  # first step is the warm-up
            segs += int(info["stats"][0])
            tt += dt
        log(f"  [cpu] step {s + 1}/{nsteps}: {dt:.2f} s, segments {int(info['stats'][0]):.3e}")
    return dict(value=segs / tt if tt > 0 else None, cores=cores, seconds=tt, segments=segs, N=N, steps=nsteps)


# ---------------------------------------------------------------------------
# gray 1-D workloads: the reference's slab case and the batched slabs
# ---------------------------------------------------------------------------
# event-based model, as the 2-D figures: state (x, mu, d_census, cell, history
# id, event counter: 4 B each in FP32 / 8,8,8,4,4,4 in FP64; weight 8 B) loaded
# and stored once per segment, (sigma_a, sigma_s) gathered, two FP64 tallies
BYTES_PER_SEGMENT_1D = {"fp32": 2 * 32 + 8 + 16, "fp64": 2 * 44 + 16 + 16}


def _batch_inputs(probs):
    meshes = [p.mesh for p in probs]
    T0 = [np.full(p.I, p.T0) for p in probs]
    eos = [p.eos.eos for p in probs]
    inv = [p.eos.inv_eos for p in probs]
    cvf = [p.eos.cv_func for p in probs]
    sf = [p.opacity for p in probs]
    N = np.array([p.n_particles for p in probs], dtype=np.int64)
    Tb = np.array([[p.T_boundary[0], p.T_boundary[1]] for p in probs])
    return meshes, T0, eos, inv, cvf, sf, N, Tb


def run_gpu_batch1d(probs, precision, steps, warmup, tune):
    from radtranbook_amd.IMC1D import BatchIMC1D

    meshes, T0, eos, inv, cvf, sf, N, Tb = _batch_inputs(probs)
    Nt, Nb = N // 2, N - N // 2
    dt = probs[0].dt
    from radtranbook_amd import autotune as AT

    b = BatchIMC1D(meshes, T0, T0, eos, inv, N, precision=precision, seed=1)
    tuned = None
    key = f"batch_slabs_{len(probs)}_N{int(N.sum())}"
    cached = cached_pick("gray1d", key, precision) if tune else None
    if cached:
        AT.apply(b.eng, cached[0])
        tuned = cached[1]
    elif tune:
        # tuning pass: the timed run itself (seed 1), candidates timed on the
        # staged step at four of its steps
        tuner = AT.gray1d_tuner((False, False))
        samples = _sample_steps(steps)
        for s in range(steps):
            b.step(Nt, Nb, N, Tb, dt, sf, cvf, tune=(tuner.sample if s in samples else False))
        tuned = tuner.report()
        AT.apply(b.eng, tuner.best())
        AT.store_shape("gray1d", key, precision, tuned)
        log(f"[{precision}] autotune: {tuned['shape']} {tuned['tuned_ms']:.2f} ms over {tuned['sample_steps']} "
            f"sample steps (default {tuned['default_ms']:.2f} ms, {tuned['candidates']} candidates x "
            f"{tuned['repeats']})")
    b.reset(T0, T0, eos, inv, N, seed=1000)
    for s in range(warmup):
        b.step(Nt, Nb, N, Tb, dt, sf, cvf)
    b.reset(T0, T0, eos, inv, N, seed=1)
    infos = []
    t0 = time.perf_counter()
    for s in range(steps):
        ts = time.perf_counter()
        out = b.step(Nt, Nb, N, Tb, dt, sf, cvf)
        infos.append(out)
        log(f"  [{precision}] step {s + 1}/{steps}: {(time.perf_counter() - ts) * 1e3:.1f} ms, transport kernel "
            f"{out['transport_kernel_ms']:.2f} ms, segments {out['stats'][0]:.3e}, histories {out['stats'][7]:.3e}")
    wall = time.perf_counter() - t0
    b.eng.close()
    segs = sum(int(i["stats"][0]) for i in infos)
    hist = sum(int(i["stats"][7]) for i in infos)
    kern = np.array([i["transport_kernel_ms"] for i in infos])
    seg_launch = np.array([i["stats"][0] for i in infos], dtype=np.float64)
    resid = max(float(np.max(np.abs(i["energy_residual"]) / np.abs(i["total_energy"]))) for i in infos)
    return dict(wall=wall, segments=segs, histories=hist, kernel_ms=kern, seg_per_launch=seg_launch, residual=resid,
                tuned=tuned, final_T=[t.copy() for t in b.T])


def run_cpu_batch1d(probs, nsteps):
    from oracle import oracle as O

    O.build()
    cores = len(os.sched_getaffinity(0))
    segs, tt = 0, 0.0
    for p in probs:
        orc = O.Oracle1D(p.mesh, "slab", seed=1, nthreads=cores)
        T0 = np.full(p.I, p.T0)
        st = orc.init_simulation(p.n_particles, T0, T0, p.eos.eos, p.eos.inv_eos)
        for s in range(nsteps):
            t0 = time.perf_counter()
            st, info = orc.step(st, p.n_particles // 2, p.n_particles - p.n_particles // 2, 0, p.n_particles,
                                p.T_boundary, p.dt, p.opacity, p.eos.inv_eos, p.eos.cv_func, np.zeros(p.I))
            tt += time.perf_counter() - t0
            segs += int(info["stats"][0])
    return dict(value=segs / tt if tt > 0 else None, cores=cores, seconds=tt, segments=segs)


def main_1d(args, workload):
    from radtranbook_amd import _lib
    from radtranbook_amd.problems import batch_slabs, marshak_slab

    _lib.load()
    _lib.require_gpu()
    steps = args.steps if args.steps is not None else 10
    warmup = args.warmup if args.warmup is not None else 2
    if workload == "batch1d":
        probs = batch_slabs(P=args.problems, n_steps=steps)
        desc = (f"batched gray IMC 1D: {args.problems} independent slabs (cells log-U[50,2000], particles/step "
                f"log-U[1e4,1e6]), dt=0.01 ns x {steps} steps, one launch per stage over all problems")
        metric = f"IMC segments/s, batched 1D: {args.problems} slab problems in one persistent launch"
    else:
        probs = [marshak_slab(n_steps=steps)]
        desc = f"gray IMC 1D slab: 100 cells on [0,0.2] cm, 1e5 particles/step, dt=0.01 ns x {steps} steps"
        metric = "IMC segments/s, 1D slab 100 cells 1e5 particles/step"
    precs = ["fp32", "fp64"] if args.precision == "both" else [args.precision]
    res = {p: run_gpu_batch1d(probs, p, steps, warmup, args.tune and workload == "batch1d") for p in precs}
    head = res[precs[0]]
    value = head["segments"] / head["wall"]
    avg_launch_s = head["kernel_ms"].mean() / 1e3
    out = {
        "metric": metric,
        "value": value,
        "unit": "segments/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": head["wall"] / steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 geometry / f64 weights+tallies" if precs[0] == "fp32" else "f64",
        "data": "synthetic (seeded problem parameters, Philox particle streams)",
        "config": {"workload": desc, "problems": len(probs), "cells": int(sum(p.I for p in probs)),
                   "new_particles_per_step": int(sum(p.n_particles for p in probs))},
        "histories_per_s": head["histories"] / head["wall"],
        "transport_segments_per_s": head["segments"] / (head["kernel_ms"].sum() / 1e3),
        "roofline": roofline(f"k1_transport<{'float' if precs[0] == 'fp32' else 'double'},0>", precs[0],
                             float(head["seg_per_launch"].mean()), head["histories"] / steps, avg_launch_s,
                             BYTES_PER_SEGMENT_1D[precs[0]], "k1_transport_batch1d_256slabs.md",
                             workload == "batch1d" and args.problems == 256 and not args.tune),
        "autotune": {p: res[p]["tuned"] for p in precs} if args.tune and workload == "batch1d" else None,
    }
    for p in precs:
        r = res[p]
        out[p] = {"segments_per_s": r["segments"] / r["wall"], "ms_per_step": r["wall"] / steps * 1e3,
                  "transport_kernel_ms_per_step": float(r["kernel_ms"].mean()),
                  "segments_per_step": float(r["seg_per_launch"].mean()),
                  "histories_per_step": r["histories"] / steps, "max_energy_residual_rel": r["residual"]}
    if not args.no_cpu:
        if workload == "batch1d":
            sample = batch_slabs(P=args.problems, n_steps=steps, scale=args.cpu_scale)[:args.cpu_problems]
            nst = args.cpu_steps if args.cpu_steps is not None else 3
        else:
            sample, nst = probs, steps
        cb = run_cpu_batch1d(sample, nst)
        out["cpu_baseline"] = {
            "value": cb["value"], "unit": "segments/s", "cores": cb["cores"], "kind": "port",
            "sample": f"FP64 C oracle (OpenMP, history-based) on {len(sample)} of the problems"
                      + (f" at {args.cpu_scale:g} x their particle counts" if workload == "batch1d" else "")
                      + f", {nst} steps ({cb['segments']:.3e} segments in {cb['seconds']:.1f} s)"}
        if cb["value"]:
            out["gpu_over_cpu_segments_per_s"] = value / cb["value"]
    print(json.dumps(out), flush=True)


# ---------------------------------------------------------------------------
# multigroup 2-D workload: the crooked pipe
# ---------------------------------------------------------------------------
BYTES_PER_SEGMENT_MG = {"fp32": 2 * 44 + 8 + 16, "fp64": 2 * 64 + 16 + 16}


def run_gpu_mg(prob, precision, steps, warmup, n_particles, tune=False):
    from radtranbook_amd import MG_IMC2D
    from radtranbook_amd import autotune as AT

    MG_IMC2D.VERBOSE = False
    T0 = np.full(prob.shape, prob.T0)
    zeros = np.zeros(prob.shape)
    sf = prob.sigma_funcs()
    engine = [None]
    shape = [None]
    tuner = AT.mg2d_tuner(prob.dt, (False,) * 4, 10**6) if tune else None
    samples = _sample_steps(steps)
    cur = [-1]

    def hook(eng):
        if cur[0] in samples:
            tuner.sample(eng)

    def one_run(seed, nsteps, timed, tuning=False):
        st = MG_IMC2D.init_simulation(n_particles // 2, T0, T0, prob.edges1, prob.edges2, prob.energy_edges, prob.eos,
                                      prob.inv_eos, geometry=prob.geometry, precision=precision, seed=seed,
                                      engine=engine[0])
        if shape[0] is not None:
            AT.apply(st._engine, shape[0])
        engine[0] = st._engine
        infos = []
        t0 = time.perf_counter()
        for s in range(nsteps):
            ts = time.perf_counter()
            cur[0] = s if tuning else -1
            st, info = MG_IMC2D.step(st, 0, 0, 0, n_particles, (0.0, 0.0, 0.0, 0.0), prob.dt, prob.edges1, prob.edges2,
                                     prob.energy_edges, sf, prob.inv_eos, prob.cv_func, zeros,
                                     boundary_source_func=prob.boundary_source_func, Ntotal=n_particles,
                                     particle_budget_fmin=0.1)
            infos.append(info)
            if timed:
                te = info["profiling"]["transport_events"]
                log(f"  [{precision}] step {s + 1}/{nsteps}: {(time.perf_counter() - ts) * 1e3:.1f} ms, transport "
                    f"kernel {info['profiling']['transport_kernel_ms']:.2f} ms, segments {te['total']:.3e}, "
                    f"histories {te['n_particles_transported']:.3e}, split {info['N_boundary']}/{info['N_target']}")
        return infos, time.perf_counter() - t0

    rep = None
    key = f"{prob.name}_N{n_particles}"
    cached = cached_pick("mg2d", key, precision) if tune else None
    if cached:
        shape[0], rep = cached
    elif tune:
        MG_IMC2D._TRANSPORT_HOOK = hook
        try:
            one_run(1, steps, False, tuning=True)
        finally:
            MG_IMC2D._TRANSPORT_HOOK = None
        rep = tuner.report()
        shape[0] = tuner.best()
        AT.store_shape("mg2d", key, precision, rep)
        log(f"[{precision}] autotune: {rep['shape']} {rep['tuned_ms']:.2f} ms over {rep['sample_steps']} sample "
            f"steps (default {rep['default_ms']:.2f} ms, {rep['candidates']} candidates x {rep['repeats']})")
        AT.apply(engine[0], shape[0])
    if warmup > 0:
        one_run(1000, warmup, False)
    infos, wall = one_run(1, steps, True)
    engine[0].close()
    segs = sum(i["profiling"]["transport_events"]["total"] for i in infos)
    hist = sum(i["profiling"]["transport_events"]["n_particles_transported"] for i in infos)
    kern = np.array([i["profiling"]["transport_kernel_ms"] for i in infos])
    seg_launch = np.array([i["profiling"]["transport_events"]["total"] for i in infos], dtype=np.float64)
    resid = max(abs(i["balance_residual"]) / i["total_energy"] for i in infos)
    return dict(wall=wall, segments=segs, histories=hist, kernel_ms=kern, seg_per_launch=seg_launch, residual=resid,
                final_T=infos[-1]["temperature"], tuned=rep)


def run_cpu_mg(prob, n_particles, nsteps):
    from oracle import oracle_mg as OM

    OM.O.build()
    orc = OM.OracleMG2D(prob.edges1, prob.edges2, prob.energy_edges, prob.geometry, seed=1)
    T0 = np.full(prob.shape, prob.T0)
    st = orc.init_simulation(n_particles // 2, T0, T0, prob.eos, prob.inv_eos)
    sf = prob.sigma_funcs()
    segs, tt = 0, 0.0
    for s in range(nsteps):
        t0 = time.perf_counter()
        st, info = orc.step(st, 0, 0, 0, n_particles, (0.0, 0.0, 0.0, 0.0), prob.dt, sf, prob.inv_eos, prob.cv_func,
                            np.zeros(prob.shape), boundary_source_func=prob.boundary_source_func,
                            Ntotal=n_particles, particle_budget_fmin=0.1)
        tt += time.perf_counter() - t0
        segs += int(info["stats"][0])
        log(f"  [cpu] step {s + 1}/{nsteps}: {time.perf_counter() - t0:.2f} s, segments {int(info['stats'][0]):.3e}")
    return dict(value=segs / tt if tt > 0 else None, cores=1, seconds=tt, segments=segs)


def main_mg(args):
    from radtranbook_amd import _lib
    from radtranbook_amd.problems import crooked_pipe_mg

    _lib.load()
    _lib.require_gpu()
    steps = args.steps if args.steps is not None else 20
    warmup = args.warmup if args.warmup is not None else 2
    npart = int(args.particles) if args.particles is not None else 20_000_000
    prob = crooked_pipe_mg(n_particles=npart, n_steps=steps)
    precs = ["fp32", "fp64"] if args.precision == "both" else [args.precision]
    res = {p: run_gpu_mg(prob, p, steps, warmup, npart, args.tune) for p in precs}
    head = res[precs[0]]
    value = head["segments"] / head["wall"]
    out = {
        "metric": "IMC segments/s, multigroup 2D crooked pipe 240x240 x 10 groups, 2e7 particles/step",
        "value": value, "unit": "segments/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
        "ms_per_step": head["wall"] / steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 geometry / f64 weights+tallies" if precs[0] == "fp32" else "f64",
        "data": "synthetic (seeded channel polyline, Philox particle streams)",
        "config": {"workload": f"multigroup gray-body IMC 2D xy crooked pipe, 240x240 cells on [0,2]^2, 10 groups "
                               f"logspace(-4,1), {npart:.0e} particles/step (energy-proportional split, fmin 0.1), "
                               f"dt=0.002 ns x {steps} steps", "cells": 240 * 240, "groups": 10,
                   "particles_per_step": npart},
        "histories_per_s": head["histories"] / head["wall"],
        "transport_segments_per_s": head["segments"] / (head["kernel_ms"].sum() / 1e3),
        "roofline": roofline(f"k_mg_transport<{'float' if precs[0] == 'fp32' else 'double'},xy>", precs[0],
                             float(head["seg_per_launch"].mean()), head["histories"] / steps,
                             head["kernel_ms"].mean() / 1e3, BYTES_PER_SEGMENT_MG[precs[0]],
                             "k_mg_transport_crooked_pipe_240x240x10.md", npart == 20_000_000),
    }
    if args.tune:
        out["autotune"] = {p: res[p]["tuned"] for p in precs}
    for p in precs:
        r = res[p]
        out[p] = {"segments_per_s": r["segments"] / r["wall"], "ms_per_step": r["wall"] / steps * 1e3,
                  "transport_kernel_ms_per_step": float(r["kernel_ms"].mean()),
                  "segments_per_step": float(r["seg_per_launch"].mean()), "histories_per_step": r["histories"] / steps,
                  "max_balance_residual_rel": r["residual"]}
    if "fp32" in res and "fp64" in res:
        a, b = res["fp32"]["final_T"], res["fp64"]["final_T"]
        out["fp32_vs_fp64_final_T_max_rel"] = float(np.max(np.abs(a - b) / b))
    if not args.no_cpu:
        ncpu = int(args.cpu_particles) if args.cpu_particles is not None else 2_000_000
        nst = args.cpu_steps if args.cpu_steps is not None else 3
        cb = run_cpu_mg(prob, ncpu, nst)
        out["cpu_baseline"] = {"value": cb["value"], "unit": "segments/s", "cores": 1, "kind": "port",
                               "sample": f"FP64 C oracle (serial, like the reference's MG tracker) on the same "
                                         f"problem at {ncpu:.0e} particles/step, {nst} steps "
                                         f"({cb['segments']:.3e} segments in {cb['seconds']:.1f} s)"}
        if cb["value"]:
            out["gpu_over_cpu_segments_per_s"] = value / cb["value"]
    print(json.dumps(out), flush=True)


# ---------------------------------------------------------------------------
# differentiable gray rz step (opacity-scale gradients by replay adjoint)
# ---------------------------------------------------------------------------
def main_diff(args):
    import torch

    from radtranbook_amd import IMC2D, _lib
    from radtranbook_amd.diff import AbsorptionStep

    _lib.load()
    _lib.require_gpu()
    IMC2D.VERBOSE = False
    steps = args.steps if args.steps is not None else 5
    N = int(args.particles) if args.particles is not None else 4_000_000
    prec = "fp64" if args.precision == "fp64" else "fp32"
    nr, nz = 128, 256
    e1 = np.linspace(0.0, 0.5, nr + 1)
    e2 = np.linspace(0.0, 1.0, nz + 1)
    rng = np.random.default_rng(1)
    theta = (1.0 + 0.1 * rng.standard_normal((nr, nz))).astype(np.float32)
    sig = lambda T: 100.0 * np.maximum(T, 1e-3) ** -3
    sf = lambda T: theta.astype(np.float64) * sig(T)
    cv = 0.3
    T0 = np.full((nr, nz), 1e-3)
    Tb = (0.0, 0.0, 0.3, 0.0)
    st = IMC2D.init_simulation(N // 2, T0, T0, e1, e2, lambda T: cv * T, lambda u: u / cv, geometry="rz",
                               precision=prec, seed=1)
    t0 = time.perf_counter()
    for s in range(steps - 1):
        st, info = IMC2D.step(st, N // 2, N // 2, 0, N, Tb, 0.01, e1, e2, sf, lambda u: u / cv, lambda T: cv + 0 * T,
                              np.zeros((nr, nz)), geometry="rz")
        log(f"  step {s + 1}: {info['profiling']['phase_times_s']['total'] * 1e3:.1f} ms, "
            f"transport kernel {info['profiling']['transport_kernel_ms']:.2f} ms")
    t_pre = time.perf_counter() - t0
    step = AbsorptionStep(st, N // 2, N // 2, Tb, 0.01, sig, lambda T: cv + 0 * T)
    g2 = np.random.default_rng(2)
    E0 = step.forward_np(theta.astype(np.float64), 1.0)
    Estar = E0 * g2.uniform(0.8, 1.2, E0.shape)
    w = g2.uniform(0.5, 1.5, E0.shape) / max(np.sum(E0 ** 2), 1e-300)
    scale = torch.tensor(theta, requires_grad=True)
    cvs = torch.ones((nr, nz), dtype=torch.float32, requires_grad=True)
    t0 = time.perf_counter()
    E = step(scale, cvs)
    t_fwd = time.perf_counter() - t0
    fwd_kernel_ms = step.last_forward_ms
    loss = torch.sum(torch.from_numpy(w) * (E - torch.from_numpy(Estar)) ** 2)
    t0 = time.perf_counter()
    loss.backward()
    t_bwd = time.perf_counter() - t0
    g = scale.grad.double().numpy()
    # directional finite differences with common random numbers (fixed emission counts)
    step.fixed_emission_counts = step.last_counts
    v = np.random.default_rng(3).standard_normal((nr, nz))
    s0 = theta.astype(np.float64)
    fds = {}
    for h in (1e-3, 3e-2):
        lp = float(np.sum(w * (step.forward_np(s0 + h * v, 1.0) - Estar) ** 2))
        lm = float(np.sum(w * (step.forward_np(s0 - h * v, 1.0) - Estar) ** 2))
        fds[h] = (lp - lm) / (2 * h)
    ad = float(np.sum(g * v))
    out = {
        "metric": "differentiable IMC: replay-adjoint gradient of a tally loss, 2D rz 128x256, 4e6 particles/step",
        "value": float(np.asarray(step.last_stats)[0]) / (t_fwd + t_bwd), "unit": "forward segments/s (fwd+bwd)",
        "n_gpus": 1, "steps": steps, "warmup": 0, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 geometry / f64 weights+tallies" if prec == "fp32" else "f64",
        "data": "synthetic (seeded opacity scales, loss weights and targets; Philox particle streams)",
        "config": {"workload": f"gray IMC rz r in [0,0.5], z in [0,1], 128x256, sigma_a = theta 100 T^-3, "
                               f"c_v 0.3, zmin 0.3 keV, dt 0.01 x {steps} (gradient at the last step), "
                               f"{N:.0e} particles/step", "cells": nr * nz},
        "forward_step_s": t_fwd, "backward_s": t_bwd, "forward_transport_kernel_ms": fwd_kernel_ms,
        "segments_last_step": int(np.asarray(step.last_stats)[0]), "steps_before_s": t_pre,
        "loss": float(loss.detach()), "grad_dot_v": ad,
        "fd_dot_v": {f"{h:g}": fds[h] for h in fds},
        "grad_cv_scale_finite": bool(np.all(np.isfinite(cvs.grad.numpy()))),
    }
    print(json.dumps(out), flush=True)


# ---------------------------------------------------------------------------
# Carter-Forest variant on the C2 problem (FP64 tracker)
# ---------------------------------------------------------------------------
# event model: state (pos 2x8, dir 2x8, weight 8, d_census 8, time 8, cell 4,
# history id 4, event counter 4 = 68 B) loaded and stored once per segment,
# (sigma_a, beta) gathered, two FP64 tallies
BYTES_PER_SEGMENT_CF = 2 * 68 + 16 + 16


def main_cf(args):
    from radtranbook_amd import IMC2D_CarterForest as CFM, _lib
    from radtranbook_amd.problems import marshak_2d_heterogeneous

    _lib.load()
    _lib.require_gpu()
    CFM.VERBOSE = False
    steps = args.steps if args.steps is not None else 20
    warmup = args.warmup if args.warmup is not None else 2
    npart = int(args.particles) if args.particles is not None else 10_000_000
    prob = marshak_2d_heterogeneous(n=args.cells, n_particles=npart, n_steps=steps)
    e1, e2 = prob.edges1, prob.edges2
    T0 = np.full(prob.shape, prob.T0)
    src = np.zeros(prob.shape)
    eng = [None]

    def one_run(seed, nsteps, timed):
        st = CFM.init_simulation(prob.Ntarget, T0, T0, e1, e2, prob.eos.eos, prob.eos.inv_eos, geometry=prob.geometry,
                                 precision="fp64", seed=seed, engine=eng[0])
        eng[0] = st._engine
        infos = []
        t0 = time.perf_counter()
        for s in range(nsteps):
            st, info = CFM.step(st, prob.Ntarget, prob.Nboundary, 0, prob.Nmax, prob.T_boundary, prob.dt, e1, e2,
                                prob.opacity, prob.eos.inv_eos, prob.eos.cv_func, src, reflect=prob.reflect,
                                geometry=prob.geometry, max_events_per_particle=prob.max_events)
            infos.append(info)
            if timed:
                te = info["profiling"]["transport_events"]
                log(f"  [cf] step {s + 1}/{nsteps}: {info['profiling']['phase_times_s']['total'] * 1e3:.1f} ms, "
                    f"transport kernel {info['profiling']['transport_kernel_ms']:.2f} ms, segments {te['total']:.3e}, "
                    f"captures {te['absorption_capture_events']:.3e}, re-emissions "
                    f"{te['absorption_continue_events']:.3e}")
        return infos, time.perf_counter() - t0

    if warmup > 0:
        one_run(1000, warmup, False)
    infos, wall = one_run(1, steps, True)
    eng[0].close()
    segs = sum(i["profiling"]["transport_events"]["total"] for i in infos)
    hist = sum(i["profiling"]["transport_events"]["n_particles_transported"] for i in infos)
    kern = np.array([i["profiling"]["transport_kernel_ms"] for i in infos])
    seg_launch = np.array([i["profiling"]["transport_events"]["total"] for i in infos], dtype=np.float64)
    resid = max(abs(i["energy_residual"]) / i["total_energy"] for i in infos)
    value = segs / wall
    out = {
        "metric": "IMC segments/s, Carter-Forest 2D 200x200 heterogeneous Marshak, 1e7 particles/step",
        "value": value, "unit": "segments/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
        "ms_per_step": wall / steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (seeded sigma0 field, Philox particle streams)",
        "config": {"workload": f"Carter-Forest gray IMC 2D xy {args.cells}x{args.cells} heterogeneous Marshak, "
                               f"{npart:.0e} particles/step, dt=0.005 ns x {steps} steps",
                   "cells": args.cells * args.cells, "particles_per_step": npart},
        "histories_per_s": hist / wall,
        "transport_segments_per_s": segs / (kern.sum() / 1e3),
        "transport_kernel_ms_per_step": float(kern.mean()),
        "segments_per_step": float(seg_launch.mean()),
        "max_energy_residual_rel": resid,
        "roofline": roofline("k_transport_cf<xy>", "fp64", float(seg_launch.mean()), hist / steps, kern.mean() / 1e3,
                             BYTES_PER_SEGMENT_CF, "k_transport_cf_marshak2d_200x200_1e7.md",
                             args.cells == 200 and npart == 10_000_000),
    }
    if not args.no_cpu:
        from oracle import oracle as O, oracle_cf as OCF

        O.build()
        ncpu = int(args.cpu_particles) if args.cpu_particles is not None else 4_000_000
        nst = args.cpu_steps if args.cpu_steps is not None else 8
        cores = len(os.sched_getaffinity(0))
        orc = OCF.OracleCF2D(e1, e2, prob.geometry, seed=1, nthreads=cores)
        st = orc.init_simulation(ncpu // 2, T0, T0, prob.eos.eos, prob.eos.inv_eos)
        csegs, ct = 0, 0.0
        for s in range(nst):
            t0 = time.perf_counter()
            st, info = orc.step(st, ncpu // 2, ncpu // 2, 0, ncpu, prob.T_boundary, prob.dt, prob.opacity,
                                prob.eos.inv_eos, prob.eos.cv_func, src, reflect=prob.reflect)
            dtc = time.perf_counter() - t0
            if s >= 1:
                csegs += int(info["stats"][0])
                ct += dtc
            log(f"  [cpu cf] step {s + 1}/{nst}: {dtc:.2f} s, segments {int(info['stats'][0]):.3e}")
        out["cpu_baseline"] = {"value": csegs / ct if ct > 0 else None, "unit": "segments/s", "cores": cores,
                               "kind": "port",
                               "sample": f"FP64 Carter-Forest C oracle (OpenMP, history-based), same problem at "
                                         f"{ncpu:.0e} particles/step, steps 2..{nst} timed ({csegs:.3e} segments in "
                                         f"{ct:.1f} s)"}
        if ct > 0:
            out["gpu_over_cpu_segments_per_s"] = value / (csegs / ct)
    print(json.dumps(out), flush=True)


# ---------------------------------------------------------------------------
# Carter-Forest 1-D: the reference's own FC vs CF Marshak-wave comparison case
# ---------------------------------------------------------------------------
# event model: state (w, mu, x, t, d_census 5x8, cell 4, history id 4, event
# counter 4 = 52 B) loaded and stored once per segment, (sigma_a, beta)
# gathered, two FP64 tallies
BYTES_PER_SEGMENT_CF1D = 2 * 52 + 16 + 16
# the reference's published run of this case on its CPU path
# (IMC/marshak_wave_profile_compare_1ns.json: wall_s, phase_totals_s.transport,
# event_totals.total)
REFERENCE_CF1D = {"cf": {"wall_s": 7.253435118997004, "transport_s": 4.730065080017084, "segments": 67963139},
                  "fc": {"wall_s": 16.950405291994684, "transport_s": 15.079742463007278, "segments": 427251321}}


def main_cf1d(args):
    from radtranbook_amd import IMC1D, IMC1D_CarterForest as CFM, _lib
    from radtranbook_amd.problems import marshak_wave_compare

    _lib.load()
    _lib.require_gpu()
    IMC1D.VERBOSE = False
    CFM.VERBOSE = False
    case = marshak_wave_compare()
    warmup = args.warmup if args.warmup is not None else 2
    T0 = np.full(case.I, case.T0)

    def run(mod, seed, nsteps):
        st = mod.init_simulation(case.n_target, T0, T0, case.mesh, case.eos, case.inv_eos, geometry="slab", seed=seed)
        segs, hist, kern, k = 0, 0, 0.0, 0
        t0 = time.perf_counter()
        while st.time < case.t_final - 1e-12 and k < nsteps:
            st, info = mod.step(st, case.n_target, case.n_boundary, 0, case.n_max, case.T_boundary,
                                min(case.dt, case.t_final - st.time), case.mesh, case.sigma_a, case.inv_eos,
                                case.cv_func, np.zeros(case.I), reflect=case.reflect, geometry="slab")
            segs += info["profiling"]["transport_events"]["total"]
            hist += info["profiling"]["transport_events"]["n_particles_transported"]
            kern += info["profiling"]["transport_kernel_ms"]
            k += 1
        return dict(wall=time.perf_counter() - t0, segs=segs, hist=hist, kern_ms=kern, steps=k,
                    energy=st.previous_total_energy)

    full = int(round(case.t_final / case.dt))
    res = {}
    for name, mod in (("cf", CFM), ("fc", IMC1D)):
        if warmup > 0:
            run(mod, 1000, warmup)
        res[name] = run(mod, 1, full)
        log(f"  [{name}] {res[name]['steps']} steps: {res[name]['wall'] * 1e3:.1f} ms, {res[name]['segs']:.4e} "
            f"segments, kernels {res[name]['kern_ms']:.1f} ms, final energy {res[name]['energy']:.6e}")
    cf = res["cf"]
    value = cf["segs"] / cf["wall"]
    out = {
        "metric": "IMC segments/s, Carter-Forest 1-D slab Marshak wave (the reference's FC vs CF comparison case)",
        "value": value, "unit": "segments/s", "n_gpus": 1, "steps": cf["steps"], "warmup": warmup,
        "ms_per_step": cf["wall"] / cf["steps"] * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "the reference's deterministic case, Philox particle streams",
        "config": {"workload": "Carter-Forest gray IMC 1-D slab, 50 cells on [0, 0.2] cm, sigma_a = 300 T^-3, "
                               "c_v = 0.3, T0 = 1e-4 keV, 1 keV left face, reflecting right face, dt = 0.025 ns "
                               "to 1 ns, 1e4 + 1e4 new particles per step, census combed to 4e4",
                   "cells": case.I},
        "transport_segments_per_s": cf["segs"] / (cf["kern_ms"] / 1e3),
        "final_total_energy": cf["energy"],
        "fc": {"segments_per_s": res["fc"]["segs"] / res["fc"]["wall"], "wall_s": res["fc"]["wall"],
               "segments": res["fc"]["segs"], "final_total_energy": res["fc"]["energy"]},
        "reference_published_cpu": REFERENCE_CF1D,
        "wall_speedup_vs_reference_published": {m: REFERENCE_CF1D[m]["wall_s"] / res[m]["wall"] for m in res},
        "roofline": dict(roofline("k1_transport_cf<0>", "fp64", cf["segs"] / cf["steps"], cf["hist"] / cf["steps"],
                                  cf["kern_ms"] / cf["steps"] / 1e3, BYTES_PER_SEGMENT_CF1D, "", False),
                         note="5e4-8e4 histories per launch: latency-bound, far below the device's size"),
    }
    if not args.no_cpu:
        from oracle import oracle as O, oracle_cf1d as OCF

        O.build()
        cores = len(os.sched_getaffinity(0))
        orc = OCF.OracleCF1D(case.mesh, "slab", seed=1, nthreads=cores)
        st = orc.init_simulation(case.n_target, T0, T0, case.eos, case.inv_eos)
        csegs = 0
        t0 = time.perf_counter()
        while st.time < case.t_final - 1e-12:
            st, info = orc.step(st, case.n_target, case.n_boundary, 0, case.n_max, case.T_boundary,
                                min(case.dt, case.t_final - st.time), case.sigma_a, case.inv_eos, case.cv_func,
                                np.zeros(case.I), reflect=case.reflect)
            csegs += int(info["stats"][0])
        ct = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": csegs / ct, "unit": "segments/s", "cores": cores, "kind": "port",
                               "sample": f"FP64 C Carter-Forest oracle (OpenMP), the whole case: {csegs:.3e} "
                                         f"segments in {ct:.2f} s"}
        out["gpu_over_cpu_segments_per_s"] = value / (csegs / ct)
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", choices=["marshak2d", "slab1d", "batch1d", "mg2d", "diff_rz", "cf2d", "cf1d"],
                    default="marshak2d",
                    help="marshak2d: 2D 200x200 heterogeneous (default, headline); slab1d: the reference's 1D slab "
                         "case; batch1d: 256 independent slabs; mg2d: multigroup crooked pipe; diff_rz: "
                         "differentiable rz step; cf2d: Carter-Forest variant of marshak2d; cf1d: the reference's "
                         "FC vs Carter-Forest 1D Marshak-wave comparison case")
    ap.add_argument("--problems", type=int, default=256)
    ap.add_argument("--tune", action=argparse.BooleanOptionalAction, default=True,
                    help="before the warm-up, run the workload once (timed seed) and time the transport "
                         "launch-shape candidates on four of its staged steps; the timed runs use the fastest, "
                         "which is kept in the launch-shape cache and reused by later runs of the same workload "
                         "and precision (marshak2d, mg2d, batch1d; --no-tune keeps the default shape)")
    ap.add_argument("--retune", action="store_true", help="tune even where the cache holds a pick")
    ap.add_argument("--cpu-problems", type=int, default=32)
    ap.add_argument("--cpu-scale", type=float, default=0.1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--particles", type=float, default=None)
    ap.add_argument("--cells", type=int, default=200)
    ap.add_argument("--fom-seeds", type=int, default=32)
    ap.add_argument("--precision", choices=["both", "fp32", "fp64"], default="both")
    ap.add_argument("--cpu-particles", type=float, default=None)
    ap.add_argument("--cpu-steps", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    global RETUNE
    RETUNE = args.retune

    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU; --gpus {args.gpus} is not supported "
              "(the IMC step shards by particles with one tally all-reduce per step; that path is not built)",
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "mg2d":
        main_mg(args)
        return
    if args.workload == "diff_rz":
        main_diff(args)
        return
    if args.workload == "cf2d":
        main_cf(args)
        return
    if args.workload == "cf1d":
        main_cf1d(args)
        return
    if args.workload != "marshak2d":
        main_1d(args, args.workload)
        return
    if args.steps is None:
        args.steps = 20
    if args.warmup is None:
        args.warmup = 2
    if args.cpu_steps is None:
        args.cpu_steps = 8
    if args.particles is None:
        args.particles = 1e7
    if args.cpu_particles is None:
        args.cpu_particles = 4e6

    from radtranbook_amd import _lib
    from radtranbook_amd.problems import marshak_2d_heterogeneous

    _lib.load()
    _lib.require_gpu()
    prob = marshak_2d_heterogeneous(n=args.cells, n_particles=int(args.particles), n_steps=args.steps)
    precs = ["fp32", "fp64"] if args.precision == "both" else [args.precision]
    res = {}
    tuned = {}
    for p in precs:
        launch = None
        if args.tune:
            launch, tuned[p] = cached_pick("gray2d", tune_key_2d(prob), p) or tune_2d(prob, p, args.steps)
        res[p] = run_gpu(prob, p, args.steps, args.warmup, args.fom_seeds, launch)
    head = res[precs[0]]
    value = head["segments"] / head["wall"]
    kern_s = head["kernel_ms"].sum() / 1e3
    seg_per_launch = float(head["seg_per_launch"].mean())
    avg_launch_s = head["kernel_ms"].mean() / 1e3
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "segments/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["wall"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 geometry / f64 weights+tallies" if precs[0] == "fp32" else "f64",
        "data": "synthetic (seeded sigma0 field, Philox particle streams)",
        "config": {
            "workload": f"gray IMC 2D xy {args.cells}x{args.cells} heterogeneous Marshak, "
                        f"{args.particles:.0e} particles/step, dt=0.005 ns x {args.steps} steps",
            "cells": args.cells * args.cells,
            "particles_per_step": int(args.particles),
            "census_Nmax": int(args.particles),
        },
        "histories_per_s": head["histories"] / head["wall"],
        "transport_segments_per_s": head["segments"] / kern_s,
        "fom_seeds": args.fom_seeds,
        "roofline": roofline(f"k_transport<{'float' if precs[0] == 'fp32' else 'double'},xy>", precs[0],
                             seg_per_launch, head["histories"] / args.steps, avg_launch_s,
                             BYTES_PER_SEGMENT[precs[0]], "k_transport_marshak2d_200x200_1e7.md",
                             args.cells == 200 and int(args.particles) == 10_000_000),
    }
    if tuned:
        out["autotune"] = tuned
    mask = fom_cells(res, prob.T0)
    out["fom_cells"] = int(mask.sum())
    # the FOM-cell set of the FOM definition itself: every cell with E_abs
    # above 1e-3 of the maximum (cold cells that re-absorb their own emission
    # included)
    mask_all = None
    for r in res.values():
        mean = r["deps"].mean(0)
        m = mean > 1e-3 * mean.max()
        mask_all = m if mask_all is None else (mask_all & m)
    out["fom_cells_all"] = int(mask_all.sum())
    out["fom_cells_all_note"] = ("R2 median over every cell with E_abs > 1e-3 max: most of those cells are cold and "
                                 "re-absorb their own emission, so their FP64 tally has no Monte Carlo noise and "
                                 "the median is degenerate; fom uses the fom_cells the front has heated")
    for p in precs:
        r = res[p]
        r["R2"], r["fom"], r["fom_transport"] = fom_of(r, mask)
        r["R2_all"], r["fom_all"], _ = fom_of(r, mask_all)
        r["final_dep"] = r["deps"].mean(0)
        out[p] = {
            "segments_per_s": r["segments"] / r["wall"],
            "ms_per_step": r["wall"] / args.steps * 1e3,
            "transport_kernel_ms_per_step": float(r["kernel_ms"].mean()),
            "segments_per_step": float(r["seg_per_launch"].mean()),
            "R2_median": r["R2"],
            "fom": r["fom"],
            "fom_transport": r["fom_transport"],
            "R2_median_all_cells": r["R2_all"],
            "fom_all_cells": r["fom_all"],
            "max_energy_residual_rel": r["residual"],
        }
    if "fp32" in res and "fp64" in res:
        if res["fp32"]["fom"] and res["fp64"]["fom"]:
            out["fom_ratio_fp32_over_fp64"] = res["fp32"]["fom"] / res["fp64"]["fom"]
        if res["fp32"]["fom_transport"] and res["fp64"]["fom_transport"]:
            out["fom_ratio_fp32_over_fp64_transport"] = res["fp32"]["fom_transport"] / res["fp64"]["fom_transport"]
        a, b = res["fp32"]["final_dep"], res["fp64"]["final_dep"]
        m = b > 1e-3 * b.max()
        out["fp32_vs_fp64_final_dep_max_rel"] = float(np.max(np.abs(a[m] - b[m]) / b[m]))
        # FP32-track tallies against FP64 in combined standard errors of the
        # seed means, over the FOM cells
        d32, d64 = res["fp32"]["deps"], res["fp64"]["deps"]
        if len(d32) > 1 and len(d64) > 1:
            se = np.sqrt(d32.var(0, ddof=1) / len(d32) + d64.var(0, ddof=1) / len(d64))
            z = np.abs(d32.mean(0) - d64.mean(0))[mask] / np.maximum(se[mask], 1e-300)
            out["fp32_vs_fp64_final_dep_z"] = {"max": float(z.max()), "frac_above_3": float(np.mean(z > 3.0)),
                                               "cells": int(mask.sum()), "seeds": int(len(d32))}
    if not args.no_cpu:
        cb = run_cpu_baseline(prob, args.cpu_particles, args.cpu_steps)
        out["cpu_baseline"] = {
            "value": cb["value"],
            "unit": "segments/s",
            "cores": cb["cores"],
            "kind": "port",
            "sample": f"FP64 C oracle (OpenMP, history-based), same 200x200 problem at {cb['N']:.0e} particles/step, "
                      f"steps 2..{cb['steps']} timed ({cb['segments']:.3e} segments in {cb['seconds']:.1f} s)",
        }
        if cb["value"]:
            out["gpu_over_cpu_segments_per_s"] = value / cb["value"]
            # the FOM the CPU oracle would reach on the same timed steps: the
            # same FP64 histories (R^2 of the FP64 GPU runs) at its measured
            # segments/s
            r64 = res.get("fp64")
            if r64 is not None and r64["R2"] and r64["R2"] > 0:
                t_cpu = r64["segments"] / cb["value"]
                out["cpu_baseline"]["fom_estimate"] = 1.0 / (r64["R2"] * t_cpu)
                if head.get("fom"):
                    out["gpu_over_cpu_fom"] = head["fom"] / out["cpu_baseline"]["fom_estimate"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
