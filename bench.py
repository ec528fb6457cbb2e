#!/usr/bin/env python3
"""Throughput of the electron kinetic step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c1|c0|c2|c3|c4]

A step is one kinetic update of the whole particle population.  For the
discharges (c1, c3, c4) that is the fused collision + push + deposit kernel
(in place, k_push_inplace), the births append, the deposit reduction, the device
field step (moment update, fluid ions, Poisson) and, on the steps the
workload asks for it, binning + merge/split resampling or the periodic
compaction.  The discharges replay the pushes and field steps between two
resampling checks from captured hipGraphs (the run is the same step for
step; --no-graphs launches every step directly).  For the single swarm (c0)
the push compacts.  For the batched swarms (c2) it is one in-place launch
over all 256 instances, replayed from a captured hipGraph with population
control every 50 steps.  The state is synthetic (Maxwellian, uniform
density) and resident in HBM before timing.

For c4 the line also carries `closure`, the metric's second half: the
closure-accelerated c3 run against the plain one (simulated time per hour
and the closure's ionization-rate error; --no-closure leaves it out).

Prints one JSON line.  The roofline object is for the dominant kernel
(k_push_inplace; k_push for the single 0-D swarm), timed with HIP events on the
library's stream over directly launched steps right before the timed region
(graph replays record no events); its `traffic` is the HBM bytes per launch
measured with rocprofv3 (profiles/traffic.json).  The CPU baseline is the oracle port
timed on a bounded sample of the same workload on the host.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import sys
import time

HBM_PEAK_GBS = 8000.0


def _cores() -> dict:
    """Threads the CPU baseline uses: every core this process may run on, up
    to 16 (the host-CPU share of one GPU on the box: 16 cores per GPU), with
    the counts it was chosen from."""
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        aff = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS")
    share = 16
    if env and env.isdigit() and int(env) > 0:
        share = int(env)
    return {"threads": max(1, min(aff, share)), "nproc": os.cpu_count() or 1,
            "affinity": aff, "omp_num_threads": env}


def _time_oracle(setup, state, E, n, table, seed, seconds, capacity):
    """Oracle steps on `state` for about `seconds`: each step draws its Philox
    uniforms inside the timed loop (R=None), as the device push does."""
    import dataclasses

    from oracle import oracle as O
    done = 0
    psteps = 0
    cur = state
    t0 = time.perf_counter()
    while True:
        m = cur[0].size
        s = dataclasses.replace(setup, seed=seed, step=done)
        cur, _, _ = O.push_step(s, cur, E, None, table=table, capacity=capacity)
        psteps += m
        done += 1
        if time.perf_counter() - t0 > seconds or done >= 1000:
            break
    return psteps / (time.perf_counter() - t0), done


def _one_core(setup, state, E, table, seed, seconds):
    """The same oracle step on one host thread over the first 200k particles
    (the per-core figure beside the all-core one)."""
    import dataclasses
    m = min(state[0].size, 200_000)
    sub = tuple(a[:m].copy() for a in state)
    s1 = dataclasses.replace(setup, nthreads=1)
    rate, done = _time_oracle(s1, sub, E, m, table, seed, seconds, 2 * m)
    return {"value": rate, "cores": 1, "sample": f"{m} particles x {done} steps, 1 thread"}


def _full_step(cfg, seconds, threads, E):
    """The oracle's particle stages of a whole 1-D step on a bounded sample
    with the workload's particles per cell (up to 1M particles, the cells cut
    in proportion): in-place resampling (key histogram, trigger, splits and
    merges) and push + collisions + deposit + compaction, for about
    `seconds`.  The field step, O(cells), is left out."""
    import dataclasses

    import numpy as np

    from boltzmann_amd import _lib
    from boltzmann_amd import physics as ph
    from oracle import oracle as O

    n = min(cfg.n_particles, 1_000_000)
    cells = max(8, int(round(cfg.n_cells * n / cfg.n_particles)))
    sub = cfg.with_(n_particles=n, n_cells=cells, precision=_lib.BZ_F64)
    cur = ph.initial_state(sub)
    table = ph.build_table(cfg.xsec, cfg.gas_density_m3, sub.np_dtype) \
        if cfg.xs_mode != _lib.BZ_XS_FIT else None
    tab = None if table is None else (table.energy, table.sig, table.slope)
    setup = O.Setup(dim=1, dt=cfg.dt, xs_mode=cfg.xs_mode, n_cells=cells,
                    length_cm=cfg.length_cm, bnf=cfg.bnf, gas_density_m3=cfg.gas_density_m3,
                    nu_max=table.nu_max if table is not None else 0.0, nthreads=threads,
                    push_order=cfg.push_order)
    nb = max(cfg.nb, 1)
    fr = cfg.bin_fractions if len(cfg.bin_fractions) == nb else ph.bin_fractions(nb)
    thr = cfg.bin_thresholds if len(cfg.bin_thresholds) == nb - 1 else ph.bin_thresholds(nb)
    targets = np.array([int(f * cfg.target_ppc) for f in fr] * (cells + 1), dtype=np.int32)
    pol = O.policy(split_below=cfg.split_below, merge_at=cfg.merge_at,
                   merge_keep=cfg.merge_keep, trigger_at=cfg.trigger_at, min_wt=cfg.min_wt,
                   split_to=cfg.split_to)
    # the GPU run's field averaged onto the sample's cells (zero without one)
    Es = np.zeros(cells)
    if E is not None and np.any(E):
        xc = (np.arange(cells) + 0.5) / cells
        xg = (np.arange(len(E)) + 0.5) / len(E)
        Es = np.interp(xc, xg, np.asarray(E, dtype=np.float64))
    done = psteps = 0
    t0 = time.perf_counter()
    while True:
        s = dataclasses.replace(setup, seed=cfg.seed, step=done)
        if cfg.resample_every and done % cfg.resample_every == 0:
            cur, _ = O.resample_inplace(s, pol, nb, thr, targets, cfg.seed, done, cur,
                                        capacity=3 * cur[0].size + 1024)
        psteps += int(np.count_nonzero(cur[4] > 0))
        cur, _, _ = O.push_step(s, cur, Es, None, table=tab, capacity=2 * cur[0].size + 1024)
        done += 1
        if time.perf_counter() - t0 > seconds or done >= 1000:
            break
    rate = psteps / (time.perf_counter() - t0)
    return {"value": rate, "unit": "particle-steps/s", "cores": threads,
            "sample": f"{n} particles on {cells} cells ({cfg.n_particles // cfg.n_cells} per "
                      f"cell as the workload) x {done} steps of the oracle's in-place "
                      f"resampling (one thread) then push+MCC+deposit+compaction (OpenMP "
                      f"{threads} threads), FP64; the O(cells) field step left out"}


def cpu_baseline(cfg, seconds: float = 10.0, E=None):
    """Oracle (C++ restatement of the reference step, OpenMP over particles)
    on a bounded sample of the same workload: push + collisions + deposit +
    compaction of up to 1M particles in the field E of the GPU run (V/cm per
    cell), repeated for about `seconds`, each step drawing its uniforms."""
    import numpy as np

    from boltzmann_amd import physics as ph
    from oracle import oracle as O

    from boltzmann_amd import _lib

    n = min(cfg.n_particles, 1_000_000)
    # FP64, as the reference computes every stage
    sub = cfg.with_(n_particles=n, precision=_lib.BZ_F64)
    st = ph.initial_state(sub)
    table = ph.build_table(cfg.xsec, cfg.gas_density_m3, sub.np_dtype) \
        if cfg.xs_mode != _lib.BZ_XS_FIT else None
    cores = _cores()
    setup = O.Setup(dim=cfg.dim, dt=cfg.dt, xs_mode=cfg.xs_mode,
                    n_cells=cfg.n_cells, length_cm=cfg.length_cm, bnf=cfg.bnf,
                    gas_density_m3=cfg.gas_density_m3,
                    nu_max=table.nu_max if table is not None else 0.0,
                    e_field0=cfg.e_field0_vcm() if cfg.dim == 0 else 0.0,
                    nthreads=cores["threads"], push_order=cfg.push_order)
    tab = None if table is None else (table.energy, table.sig, table.slope)
    if cfg.dim == 1:
        E = np.zeros(cfg.n_cells) if E is None else np.asarray(E, dtype=np.float64)
        field = "the GPU run's final field" if np.any(E) else "zero field"
    else:
        E, field = None, "the uniform 0-D field"
    rate, done = _time_oracle(setup, st, E, n, tab, cfg.seed, seconds, 2 * n)
    out = {"value": rate, "unit": "particle-steps/s", "cores": cores["threads"], "kind": "port",
           "host": cores,
           "sample": f"{n} particles x {done} steps of the oracle push+MCC+deposit+compaction "
                     f"in {field} (FP64 like the reference, Philox uniforms drawn in the "
                     f"loop, OpenMP {cores['threads']} threads)",
           "one_core": _one_core(setup, st, E, tab, cfg.seed, seconds / 4)}
    if cfg.dim == 1 and cfg.target_ppc > 0:
        out["full_step"] = _full_step(cfg, seconds, cores["threads"], E)
    return out


def cpu_baseline_swarms(scfg, seconds: float = 10.0):
    """Oracle on one 1M-electron 0-D swarm at the median reduced field of
    the batch (the per-instance work of the batched launch)."""
    import numpy as np

    from boltzmann_amd import physics as ph
    from oracle import oracle as O

    n = 1_000_000
    en = float(np.median(scfg.fields_td()))
    table = ph.build_table(scfg.xsec, scfg.gas_density_m3, np.float64)
    rng = np.random.default_rng(scfg.seed)
    sv = math.sqrt(scfg.te_ev * ph.Q_E / ph.M_E)
    dt = np.float64
    st = (np.zeros(n, dt), *(rng.normal(0, sv, n).astype(dt) for _ in range(3)), np.ones(n, dt))
    cores = _cores()
    setup = O.Setup(dim=0, dt=scfg.dt, xs_mode=1, gas_density_m3=scfg.gas_density_m3,
                    nu_max=table.nu_max, e_field0=en * ph.TD * scfg.gas_density_m3 / 100.0,
                    nthreads=cores["threads"])
    tab = (table.energy, table.sig, table.slope)
    rate, done = _time_oracle(setup, st, None, n, tab, scfg.seed, seconds, 2 * n)
    return {"value": rate, "unit": "particle-steps/s", "cores": cores["threads"], "kind": "port",
            "host": cores,
            "sample": f"one 0-D swarm of {n} particles at {en:.1f} Td x {done} steps of the "
                      f"oracle push+MCC (FP64 like the reference, Philox uniforms drawn in the "
                      f"loop, OpenMP {cores['threads']} threads)",
            "one_core": _one_core(setup, st, None, tab, scfg.seed, seconds / 4)}


TRAFFIC_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                            "traffic.json")


def measured_traffic(workload: str, kernel: str, psteps_per_launch: float):
    """HBM bytes per launch of `kernel` from the rocprofv3 FETCH_SIZE /
    WRITE_SIZE runs recorded in profiles/traffic.json (per particle-step,
    already scaled by the copy-kernel calibration), or None."""
    try:
        with open(TRAFFIC_FILE) as f:
            rec = json.load(f)[workload]
    except (OSError, KeyError, ValueError):
        return None
    rec = rec.get(kernel) if "kernel" not in rec else (rec if rec["kernel"] == kernel else None)
    if not rec:
        return None
    per = rec["fetch_bytes_per_unit"] + rec["write_bytes_per_unit"]
    return per * psteps_per_launch


def measured_launch_traffic(workload: str, kernel: str):
    """HBM bytes per launch of a kernel whose work has no per-unit model
    (profiles/traffic.json, fetch and write per launch), or None."""
    try:
        with open(TRAFFIC_FILE) as f:
            rec = json.load(f)[workload][kernel]
        return rec["fetch_bytes_per_launch"] + rec["write_bytes_per_launch"]
    except (OSError, KeyError, ValueError, TypeError):
        return None


def copy_probe(mib: int) -> None:
    """Three device-to-device copies of `mib` MiB: a kernel of known traffic."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()


def _roofline(bytes_per_unit, units_per_launch, avg_s, kernel, workload=None, unit="particle"):
    achieved = bytes_per_unit * units_per_launch / avg_s / 1e9
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(workload, kernel, units_per_launch) if workload else None,
            "kernel": kernel, "kernel_avg_us": avg_s * 1e6,
            "bytes_per_unit": bytes_per_unit, "units_per_launch": units_per_launch,
            "work_unit": unit}


# timed kinds of the library's event timing (bz_timing_get_n)
T_PUSH, T_SORT, T_RESAMPLE, T_COPY, T_DOWN, T_SEG, T_HIST, T_FILL = range(8)
N_TIMED_KINDS = 8


def bench_engine(args):
    import torch

    from boltzmann_amd import _lib, configs
    from boltzmann_amd import physics as ph
    from boltzmann_amd.engine import KineticEngine

    cfg = configs.get(args.workload)
    eng = KineticEngine(cfg)
    _lib.check(eng.lib.bz_set_push_schedule(eng.ctx, args.sched))
    eng.load_state(ph.initial_state(cfg))
    eng.initialize_fields()
    index = 0
    tuned = None
    if args.shape >= 0:
        _lib.check(eng.lib.bz_set_push_shape(eng.ctx, args.shape))
    else:
        from boltzmann_amd.autotune import tune_push
        tuned = tune_push(eng, steps=10)
        index += tuned["steps_used"]
    graphs = not args.no_graphs
    index = eng.run(index, args.warmup, graphs=graphs)
    torch.cuda.synchronize()
    # kernel times from directly launched steps with the library's HIP events
    # on its stream (graph replays record none)
    n_timed = max(10, min(50, args.steps // 4))
    ct0 = eng.counters()
    _lib.check(eng.lib.bz_timing_enable(eng.ctx, 1))
    # step by step: the slots each step starts from (the key histogram's units)
    slots = 0
    for _ in range(n_timed):
        slots += eng.count()
        index = eng.run(index, 1, graphs=False)
    torch.cuda.synchronize()
    ms = (C.c_double * N_TIMED_KINDS)()
    nl = (C.c_int64 * N_TIMED_KINDS)()
    _lib.check(eng.lib.bz_timing_get_n(eng.ctx, N_TIMED_KINDS, ms, nl))
    _lib.check(eng.lib.bz_timing_enable(eng.ctx, 0))
    ct1 = eng.counters()
    # every block the timed run replays is captured before the clock starts
    n_graphs = eng.precapture() if graphs else 0
    c0 = eng.counters()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    index = eng.run(index, args.steps, graphs=graphs)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    c1 = eng.counters()
    psteps = c1["particle_steps"] - c0["particle_steps"]
    elapsed = t1 - t0
    word = 8 if cfg.precision == _lib.BZ_F64 else 4
    # push: read + write of the carried components (0-D runs carry no position)
    bytes_push = 2 * (5 if cfg.dim == 1 else 4) * word
    if eng.push_keys:
        # the binning key of the next sort (or in-place pass), written by the push
        bytes_push += 2 if eng.push_keys16 else 4
    dct = {k: ct1[k] - ct0[k] for k in ct1}

    def avg(kind):
        return ms[kind] / 1e3 / max(int(nl[kind]), 1)

    kernels = {}
    if nl[T_PUSH]:
        kernels["push"] = _roofline(bytes_push, dct["particle_steps"] / nl[T_PUSH], avg(T_PUSH),
                                    "k_push_inplace" if cfg.dim == 1 else "k_push", cfg.name,
                                    "particle-step")
    if nl[T_COPY]:
        # sorted key 4 + permutation 4 + the five words gathered 20, written 20
        kernels["res_copy"] = _roofline(8 + 10 * word, dct["copied"] / nl[T_COPY], avg(T_COPY),
                                        "k_res_copy", cfg.name, "copied particle")
    if nl[T_DOWN] and nl[T_SORT]:
        # the full sort's passes (in-place resampling sorts only its candidate
        # subsets: their time is in kernel_ms["downsweep"], with no slot count)
        passes = nl[T_DOWN] / nl[T_SORT]
        # per slot and pass: key read 4, key + value written 8, value read 4
        # on every pass but the first (the first pass's values are the indices)
        per = (12 + 16 * (passes - 1)) / passes
        kernels["downsweep"] = _roofline(per, dct["sorted"] / max(nl[T_SORT], 1), avg(T_DOWN),
                                         "k_downsweep", cfg.name, "sorted slot per pass")
    if nl[T_HIST]:
        # in-place resampling's key histogram: each slot's key read once
        kernels["key_histogram"] = _roofline(2 if eng.push_keys16 else 4,
                                             slots / n_timed, avg(T_HIST), "k_ip_hist",
                                             None, "slot")
        kernels["key_histogram"]["traffic"] = measured_launch_traffic(cfg.name, "k_ip_hist")
    # merges and splits (member gathers) and the hole filling's moves (scattered
    # writes) have no per-unit byte model: their time per launch and the HBM
    # bytes rocprofv3 counted for a launch of the same workload
    for key, kind, kname, bound in (("segments", T_SEG, "k_res_segments", "gather"),
                                    ("hole_fill_moves", T_FILL, "k_fill_move", "scatter")):
        if not nl[kind]:
            continue
        tr = measured_launch_traffic(cfg.name, kname)
        kernels[key] = {"bound": bound, "kernel": kname, "kernel_avg_us": avg(kind) * 1e6,
                        "launches": int(nl[kind]), "traffic": tr,
                        "achieved_measured": tr / avg(kind) / 1e9 if tr else None,
                        "frac_measured": tr / avg(kind) / 1e9 / HBM_PEAK_GBS if tr else None,
                        "unit": "GB/s", "peak": HBM_PEAK_GBS}
    # the roofline line names the kernel with the most time per step among
    # those with an algorithmic byte count
    share = {"push": ms[T_PUSH], "res_copy": ms[T_COPY], "downsweep": ms[T_DOWN],
             "key_histogram": ms[T_HIST]}
    dom = max((k for k in kernels if k in share), key=lambda k: share[k])
    out = {
        "metric": "particle-steps/s on 1 MI355X",
        "value": psteps / elapsed,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp64" if cfg.precision == _lib.BZ_F64 else "fp32",
        "data": "synthetic (Maxwellian electrons, uniform density, tabulated argon-like sigma)",
        "config": {"workload": cfg.name, "particles_start": cfg.n_particles,
                   "particles_end": eng.count(), "cells": cfg.n_cells if cfg.dim == 1 else 0,
                   "energy_bins": cfg.nb, "resample_every": cfg.resample_every,
                   "dt_s": cfg.dt, "parallelism": "single GPU", "push_schedule": args.sched,
                   "graphs": n_graphs},
        "roofline": kernels[dom],
        "roofline_kernels": kernels,
        "counters": {k: c1[k] - c0[k] for k in ("births", "absorbed", "ionization", "overflow",
                                                "resample", "split", "merge", "copied",
                                                "sorted")},
        "kernel_ms": {"push": ms[T_PUSH], "sort": ms[T_SORT], "resample": ms[T_RESAMPLE],
                      "res_copy": ms[T_COPY], "downsweep": ms[T_DOWN],
                      "segments": ms[T_SEG], "key_histogram": ms[T_HIST],
                      "hole_fill_moves": ms[T_FILL],
                      "launches": [int(v) for v in nl], "over_steps": n_timed},
        "autotune": tuned,
    }
    E_run = eng.E.cpu().numpy() if cfg.dim == 1 else None
    eng.close()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(cfg, args.cpu_seconds, E_run)
    return out


def bench_swarms(args):
    """256 batched 0-D swarms: graph replay of 100 steps + 2 controls."""
    import torch

    from boltzmann_amd import _lib
    from boltzmann_amd.autotune import tune_swarms
    from boltzmann_amd.swarms import SwarmBatch, c2

    scfg = c2()
    b = SwarmBatch(scfg)
    b.load_initial()
    tuned = None
    if args.shape >= 0:
        _lib.check(b.lib.bz_set_push_shape(b.ctx, args.shape))
    else:
        tuned = tune_swarms(b, steps=10)
    per_graph = 2 * scfg.control_every
    gid = b.capture(per_graph, accumulate=True, control_every=scfg.control_every)
    reps_w = max(1, args.warmup // per_graph)
    reps = max(1, args.steps // per_graph)
    b.replay(gid, reps_w)
    b.synchronize()
    # kernel time of the push, from eager steps with the library's events
    _lib.check(b.lib.bz_timing_enable(b.ctx, 1))
    n_before = b.count()
    for _ in range(20):
        b.step(accumulate=True)
    ms = (C.c_double * 3)()
    nl = (C.c_int64 * 3)()
    b.synchronize()
    _lib.check(b.lib.bz_timing_get(b.ctx, ms, nl))
    _lib.check(b.lib.bz_timing_enable(b.ctx, 0))
    n_mid = b.count()
    b.synchronize()
    c0 = b.counters()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b.replay(gid, reps)
    b.synchronize()
    t1 = time.perf_counter()
    c1 = b.counters()
    psteps = c1["particle_steps"] - c0["particle_steps"]
    steps = reps * per_graph
    elapsed = t1 - t0
    word = 8 if scfg.precision == _lib.BZ_F64 else 4
    bytes_per_pstep = 8 * word  # x (instance), vx, vy, vz, w read; vx, vy, vz written
    avg_push_s = ms[0] / 1e3 / max(int(nl[0]), 1)
    res = b.results()
    out = {
        "metric": "particle-steps/s on 1 MI355X",
        "value": psteps / elapsed,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": reps_w * per_graph,
        "ms_per_step": elapsed / steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp64" if scfg.precision == _lib.BZ_F64 else "fp32",
        "data": "synthetic (Maxwellian 1 eV swarms, tabulated argon-like sigma)",
        "config": {"workload": scfg.name, "instances": b.n_inst,
                   "particles_start": b.total0, "particles_end": b.count(),
                   "en_td": [scfg.en_td_lo, scfg.en_td_hi], "dt_s": scfg.dt,
                   "graph": f"{per_graph} steps + control every {scfg.control_every}",
                   "parallelism": "single GPU"},
        "roofline": _roofline(bytes_per_pstep, 0.5 * (n_before + n_mid), avg_push_s, "k_push_inplace",
                              scfg.name, "particle-step"),
        "counters": {k: c1[k] - c0[k] for k in ("births", "ionization", "split", "merge")},
        "sample_results": {"en_td": [float(res["en_td"][i]) for i in (0, 128, 255)],
                           "mean_energy_ev": [float(res["mean_energy"][i]) for i in (0, 128, 255)]},
        "autotune": tuned,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline_swarms(scfg, args.cpu_seconds)
    b.close()
    return out


def bench_closure(args):
    """Closure-accelerated 1-D run (1024 cells, 4M electrons) against the
    plain run: simulated time per wall-clock hour of each, and the closure's
    ionization-rate error checked against the kinetic step at every update.  The
    plain run is the workload as configured (direct collision probability,
    nu dt ~ 0.1).  The closure run's kinetic updates span each interval with
    a few longer free-flight pushes (exact collisions at any push length,
    omega_pe h < 2).  The closure network is trained first, on the GPU, on
    batched-swarm samples."""
    import numpy as np
    import torch

    from boltzmann_amd import _lib, configs
    from boltzmann_amd import physics as ph
    from boltzmann_amd.closure_run import ClosureDischarge, swarm_training_set, train_closure
    from boltzmann_amd.engine import KineticEngine

    def log(msg):
        print(f"[c3] {msg}", file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    x, t = swarm_training_set(n_samples=args.closure_samples, log=log)
    t_data = time.perf_counter() - t0
    net, rep = train_closure(x, t, epochs=args.closure_epochs, log=log)
    log(f"trained: {rep['holdout_trained']}")
    params = {"master": net.parameters(), "norm": dict(net.norm)}
    net.close()
    cfg = configs.c3()
    cfg_closure = cfg.with_(xs_mode=_lib.BZ_XS_FLIGHT)
    st = ph.initial_state(cfg)
    n_sub = 8
    steps = max(n_sub, (args.steps // n_sub) * n_sub)
    warm = max(n_sub, (args.warmup // n_sub) * n_sub)
    # plain: a kinetic step every dt
    eng = KineticEngine(cfg)
    eng.load_state(st)
    eng.initialize_fields()
    nn = cfg.n_nodes
    for i in range(warm):
        eng.step(i)
    acc_ne = torch.zeros(1, dtype=torch.float64, device=eng.dev)
    acc_nrg = torch.zeros(1, dtype=torch.float64, device=eng.dev)
    acc_ion = torch.zeros(1, dtype=torch.float64, device=eng.dev)
    torch.cuda.synchronize()
    ta = time.perf_counter()
    for i in range(warm, warm + steps):
        eng.step(i)
    torch.cuda.synchronize()
    tb = time.perf_counter()
    # moments of the plain run (diagnostic pass, outside the timed region)
    for i in range(warm + steps, warm + 2 * steps):
        eng.step(i)
        m = eng.moments
        acc_ne += m[:nn].sum()
        acc_ion += m[nn:2 * nn].sum()
        acc_nrg += m[2 * nn:].sum()
    plain_s = tb - ta
    plain = {"ms_per_step": plain_s / steps * 1e3,
             "sim_seconds_per_hour": cfg.dt * steps / plain_s * 3600,
             "mean_energy_ev": float(acc_nrg.item() / acc_ne.item()),
             "ionization_per_step": float(acc_ion.item() / steps),
             "particles_end": eng.count()}
    eng.close()
    # closure: a kinetic step every n_sub sub-steps
    run = ClosureDischarge(cfg_closure, params, n_sub=n_sub,
                           resample_every=max(1, cfg.resample_every // n_sub))
    run.load(st)
    for _ in range(warm // n_sub):
        run.macro_step()
    torch.cuda.synchronize()
    run.tally.zero_()
    tc = time.perf_counter()
    for _ in range(steps // n_sub):
        run.macro_step()
    torch.cuda.synchronize()
    td = time.perf_counter()
    for _ in range(steps // n_sub):
        run.macro_step()
    rep_c = run.report()
    run_n_kin = run.n_kin
    closure_s = td - tc
    closure = {"ms_per_sim_step": closure_s / steps * 1e3,
               "sim_seconds_per_hour": cfg.dt * steps / closure_s * 3600,
               "particles_end": run.eng.count(), **rep_c,
               "ionization_per_step": rep_c["ionization_kinetic"] / (2 * steps)}
    run.close()
    speedup = closure["sim_seconds_per_hour"] / plain["sim_seconds_per_hour"]
    e_mean = abs(closure["mean_energy_ev"] / plain["mean_energy_ev"] - 1)
    e_ion = abs(closure["ionization_per_step"] / plain["ionization_per_step"] - 1)
    return {
        "metric": "simulated time per hour, closure-accelerated vs plain 1-D run (1 MI355X)",
        "value": closure["sim_seconds_per_hour"],
        "unit": "simulated s / wall h",
        "n_gpus": 1, "steps": steps, "warmup": warm,
        "ms_per_step": closure["ms_per_sim_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic (uniform Maxwellian discharge; closure trained on synthetic swarms)",
        "config": {"workload": cfg.name, "cells": cfg.n_cells,
                   "kinetic_pushes_per_update": run_n_kin,
                   "particles_start": cfg.n_particles, "dt_s": cfg.dt, "sub_steps": n_sub,
                   "parallelism": "single GPU"},
        "speedup_vs_plain": speedup,
        "rate_error": {"closure_vs_kinetic_ionization": closure["ionization_rate_error"],
                       "closure_vs_kinetic_per_update": closure["ionization_rate_error_per_update"],
                       "run_vs_plain_ionization_per_step": e_ion,
                       "run_vs_plain_mean_energy": e_mean},
        "plain": plain, "closure": closure,
        "training": {"data_seconds": t_data, **rep},
    }


def closure_half(args) -> dict:
    """The second half of the metric next to the particle-steps/s line: the
    closure-accelerated 1-D run (c3) against the plain run, simulated time
    per wall-clock hour and the closure's ionization-rate error (target:
    >= 5x at <= 2 %).  A failure is reported in the object, not raised."""
    try:
        r = bench_closure(argparse.Namespace(**{**vars(args), "steps": 400, "warmup": 100}))
    except Exception as exc:  # noqa: BLE001 - reported, the particle line stands
        return {"error": f"{type(exc).__name__}: {exc}"}
    err = r["rate_error"]["closure_vs_kinetic_ionization"]
    return {"metric": "simulated time per hour, closure-accelerated vs plain 1-D run",
            "workload": r["config"]["workload"], "speedup_vs_plain": r["speedup_vs_plain"],
            "closure_sim_seconds_per_hour": r["value"],
            "plain_sim_seconds_per_hour": r["plain"]["sim_seconds_per_hour"],
            "ionization_rate_error": err, "rate_error": r["rate_error"],
            "target": {"speedup": 5.0, "rate_error": 0.02},
            "met": bool(r["speedup_vs_plain"] >= 5.0 and err is not None and err <= 0.02),
            "sub_steps": r["config"]["sub_steps"],
            "kinetic_pushes_per_update": r["config"]["kinetic_pushes_per_update"],
            "training_holdout": r["training"]["holdout_trained"]}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--workload", default="c4",
                    help="c4 (default: the largest single-GPU discharge), c0, c1, c2, c3")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-graphs", action="store_true",
                    help="launch every step directly instead of replaying captured blocks")
    ap.add_argument("--shape", type=int, default=-1, help="push tile shape (-1: autotune)")
    ap.add_argument("--sched", type=int, default=0,
                    help="push tile schedule: 0 static + cooperative launch, 1 ticket counter")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="after the run, copy MIB MiB device to device 3 times: the plain "
                         "copy kernel that scales rocprofv3 FETCH_SIZE / WRITE_SIZE counts")
    ap.add_argument("--closure-samples", type=int, default=1_000_000)
    ap.add_argument("--closure-epochs", type=int, default=30)
    ap.add_argument("--no-closure", action="store_true",
                    help="c4: leave out the closure half of the metric (the c3 run, ~15 s)")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the particle "
              "decomposition across GPUs is a design note, not a measured path.",
              file=sys.stderr)
        return 2
    if args.workload == "c2":
        out = bench_swarms(args)
    elif args.workload == "c3":
        out = bench_closure(args)
    else:
        out = bench_engine(args)
        if args.workload == "c4" and not args.no_closure:
            out["closure"] = closure_half(args)
    if args.copy_probe > 0:
        copy_probe(args.copy_probe)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
