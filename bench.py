#!/usr/bin/env python
"""Benchmark: FP64 CW-FDTD Yee cell-updates/s on one MI355X.

Workload (the single-GPU headline case): 400x200x120 grid at dx = 0.03 um,
12-cell CPML on every face, a 240x120x10 design slab of 40 seeded ellipses
(eps 3.48^2 in 1.444^2, boundary cells area-averaged), lambda = 1.31 um,
planar Ey source across the input face.  Synthetic inputs, built on the host
and uploaded once; the timed region starts with everything resident in HBM.

A step is one full Yee time step (H pass + E pass, CPML slabs, source
injection) with the on-device convergence check every Tn steps, as the solve
loop runs it.  value = Nx*Ny*Nz*K / (wall time of K steps).

Also reported: per-kernel HBM roofline (HIP events on the library's own
stream), a plain device-copy bandwidth, one full converged solve, and the
reference CPU kernel (EMopt's src/fdtd.cpp compiled from its sources,
single thread) timed on a bounded sample of the same workload on this host.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
import warnings

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
B_ALG_H = 72    # bytes per cell, H pass: read E(3) H(3), write H(3)
B_ALG_E = 96    # bytes per cell, E pass: read H(3) E(3) b_eps(3), write E(3)
B_ALG_STEP = B_ALG_H + B_ALG_E


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_sim(case):
    from emopt_amd.fdtd import FDTD
    from emopt_amd.misc import DomainCoordinates
    from tests.cases import _configure
    d = case["d"]
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        sim = FDTD(case["X"], case["Y"], case["Z"], d, d, d, case["wl"], rtol=case["rtol"])
        _configure(sim, case, DomainCoordinates)
    sim.set_material_arrays(case["mats"][:3], case["mats"][3:])
    sim._register(sim._sources)
    return sim


def time_steps(sim, n_start, nsteps):
    out = np.zeros(3)
    sim._lib.emf_time_steps(sim._h, n_start, nsteps, sim.Tn, out.ctypes.data_as(
        __import__("ctypes").c_void_p))
    return out


def cpu_baseline(case, seconds):
    """EMopt's own kernel (oracle/_ref, compiled from src/fdtd.cpp) driven by
    the oracle's restatement of the Python time loop, one thread."""
    from oracle import driver as D
    from tests.cases import _configure
    kind = "reference" if os.path.exists(D.REF_LIB) else "port"
    lib = D.load(D.REF_LIB if kind == "reference" else D.ORACLE_LIB)
    d = case["d"]
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        o = D.OracleFDTD(lib, case["X"], case["Y"], case["Z"], d, d, d, case["wl"], rtol=case["rtol"])
        _configure(o, case, D.Domain)
    o.set_material_arrays(case["mats"][:3], case["mats"][3:])
    o.lib.FDTD_clear_sources(o.h)
    for arrs, i0, j0, k0, sI, sJ, sK in o.sources:
        o.lib.FDTD_add_source(o.h, *arrs, i0, j0, k0, sI, sJ, sK, False)
    o.lib.FDTD_reset_pml(o.h)
    o._step(0)   # warm-up (page-in)
    n, t0 = 1, time.perf_counter()
    while True:
        o._step(n)
        n += 1
        el = time.perf_counter() - t0
        if el >= seconds and n >= 3:
            break
    steps = n - 1
    cells = case["Nx"] * case["Ny"] * case["Nz"]
    return dict(value=cells * steps / el, unit="cell-updates/s", cores=1, kind=kind,
                sample="%d time steps of the %dx%dx%d workload (%.1f s), EMopt src/fdtd.cpp "
                       "-O3 -march=x86-64-v3, 1 thread, Python time loop with one-rank ghost "
                       "exchange" % (steps, case["Nx"], case["Ny"], case["Nz"], el))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=40)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-solve", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on exactly one GPU (single-GPU solver, no "
              "multi-GPU path); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    from tests.cases import large_case
    from emopt_amd import _lib
    t = time.perf_counter()
    case = large_case()
    log("[bench] workload built in %.1fs" % (time.perf_counter() - t))
    sim = make_sim(case)
    cells = case["Nx"] * case["Ny"] * case["Nz"]
    L = sim._lib

    # warm-up, then the timed region (emf_time_steps synchronises on return)
    L.emf_reset_pml(sim._h)
    L.emf_zero_fields(sim._h)
    time_steps(sim, 0, args.warmup)
    log("[bench] warm-up done")
    t0 = time.perf_counter()
    dev = time_steps(sim, args.warmup, args.steps)
    wall = time.perf_counter() - t0
    value = cells * args.steps / wall
    ms_h = dev[1] / args.steps
    ms_e = dev[2] / args.steps
    log("[bench] %d steps: wall %.3fs, device %.3fs, H %.3f ms, E %.3f ms"
        % (args.steps, wall, dev[0] / 1e3, ms_h, ms_e))

    gbs = np.zeros(1)
    L.emf_copy_bandwidth(0, 1 << 31, 10, gbs.ctypes.data_as(_lib.c_void_p))

    ach_e = B_ALG_E * cells / (ms_e * 1e-3) / 1e9
    ach_h = B_ALG_H * cells / (ms_h * 1e-3) / 1e9
    res = {
        "metric": "Yee cell-updates/s (FP64 CW-FDTD, 400x200x120)",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "single_large_400x200x120_dx0.03_cpml12_ellipses40_ey_plane",
                   "grid": [case["Nx"], case["Ny"], case["Nz"]], "cells": cells,
                   "cpml_cells": 12, "wavelength_um": case["wl"], "Tn": sim.Tn,
                   "bytes_per_cell_update": B_ALG_STEP},
        "hbm_frac_168B": value * B_ALG_STEP / (PEAK_HBM_GBS * 1e9),
        "roofline": {"bound": "hbm", "kernel": "k_update_e (E pass)", "achieved": ach_e,
                     "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": ach_e / PEAK_HBM_GBS,
                     "traffic": None},
        "kernels": {"k_update_h": {"avg_ms": ms_h, "alg_bytes": B_ALG_H * cells, "GBps": ach_h},
                    "k_update_e": {"avg_ms": ms_e, "alg_bytes": B_ALG_E * cells, "GBps": ach_e}},
        "copy_GBps": float(gbs[0]),
    }

    if not args.no_solve:
        t = time.perf_counter()
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            sim._solve(0)
        ls = sim.last_solve
        el = time.perf_counter() - t
        res["solve"] = {"n0": ls["n0"], "steps": ls["steps"], "status": ls["status"],
                        "seconds": el, "solves_per_hr": 3600.0 / el,
                        "cell_updates_per_s": cells * ls["steps"] / el}
        log("[bench] full solve: %d steps in %.2fs (%s)" % (ls["steps"], el, ls["status"]))

    if not args.no_cpu:
        cb = cpu_baseline(case, args.cpu_seconds)
        res["cpu_baseline"] = cb
        log("[bench] cpu baseline %.3g cell-updates/s" % cb["value"])
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
