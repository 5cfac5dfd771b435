"""Launch-shape autotuning for the step kernels.

Single solves: the step kernel and its work-group shape -- kind (two passes,
two passes with the H pass's neighbour row through LDS, fused with row
recompute, fused through LDS), rows per work-group, z planes per
work-group and work-group order -- are timed with HIP events on the solver
itself (``emf_time_steps``), on its own grid, materials and CPML, and the
fastest is kept.  Every candidate computes bit-for-bit the same fields (the
parity tests check each one), so tuning never changes results.  Winners are
cached per (grid, complex eps, uniform mu) in memory and, when a path is given,
in a JSON file.

Batches: steps per graph replay, z planes per work-group and persistent
work-groups per CU are timed on a sample of the batch's members (a fixed
number of replays from the start of their solves), stepped as one group so
that every replay advances every member.
"""
from __future__ import annotations

import ctypes
import json
import os
import time

import numpy as np

__all__ = ["SOLVE_CANDIDATES", "tune_solver", "BATCH_CANDIDATES", "tune_batch"]

# (kind, rows per work-group, z planes per work-group, order)
SOLVE_CANDIDATES = [(1, 4, 2, 3), (1, 2, 2, 3), (1, 2, 4, 3), (1, 2, 2, 0), (1, 4, 2, 0),
                    (1, 8, 2, 3), (1, 2, 1, 3), (1, 4, 2, 1), (3, 2, 2, 3), (2, 8, 4, 0),
                    (2, 4, 4, 0), (0, 1, 8, 0), (4, 4, 2, 3), (4, 2, 2, 3)]
_cache = {}


def _key(sim):
    return "%dx%dx%d/ceps%d/mu%d" % (sim.Nx, sim.Ny, sim.Nz, int(sim._complex_eps),
                                     int(getattr(sim, "_mu_uniform", True)))


def _load(path):
    if path and os.path.exists(path):
        try:
            with open(path) as f:
                return json.load(f)
        except (OSError, ValueError):
            return {}
    return {}


def tune_solver(sim, candidates=None, steps=40, warmup=8, cache_path=None, verbose=False):
    """Time each candidate on `sim` and apply the fastest; returns
    (best candidate, {candidate: ms per step}).  The solver's fields and CPML
    are reset afterwards (a solve resets them anyway)."""
    key = _key(sim)
    disk = _load(cache_path)
    if key in _cache or key in disk:
        best = tuple(_cache.get(key) or disk[key])
        sim._lib.emf_set_tuning(sim._h, *best)
        return best, {}
    if not sim._mats_ready:
        sim.build()
    sim._register(sim._sources)
    L, h = sim._lib, sim._h
    times = {}
    out = np.zeros(3)
    for cand in candidates or SOLVE_CANDIDATES:
        try:
            L.emf_set_tuning(h, *cand)
        except Exception:
            continue
        L.emf_reset_pml(h)
        L.emf_zero_fields(h)
        L.emf_time_steps(h, 0, warmup, sim.Tn, 0, out.ctypes.data_as(ctypes.c_void_p))
        L.emf_time_steps(h, warmup, steps, sim.Tn, 0, out.ctypes.data_as(ctypes.c_void_p))
        times[cand] = out[0] / steps
        if verbose:
            print("tune %s: %.4f ms/step" % (cand, times[cand]))
    best = min(times, key=times.get)
    L.emf_set_tuning(h, *best)
    L.emf_reset_pml(h)
    L.emf_zero_fields(h)
    _cache[key] = list(best)
    if cache_path:
        disk[key] = list(best)
        with open(cache_path, "w") as f:
            json.dump(disk, f, indent=1)
    return best, times


# (steps per graph, z planes per work-group, work-groups per CU)
BATCH_CANDIDATES = [(64, 2, 16), (64, 2, 24), (64, 2, 8), (64, 4, 16), (32, 2, 16), (128, 2, 16), (64, 1, 16)]


def tune_batch(sims, candidates=None, graphs=3, device=0, verbose=False):
    """Time batch settings on `sims` (configured FDTD objects): each
    candidate runs `graphs` replays of all of them from the start of their
    solves; returns (best, {candidate: cell-updates/s})."""
    from .batch import Batch
    cells = sum(s.Nx * s.Ny * s.Nz for s in sims)
    rates = {}
    for cand in candidates or BATCH_CANDIDATES:
        spg, zc, bpc = cand
        b = Batch(device, spg, zc, bpc, group_bytes=0)
        ids = [b.add(s, postprocess=False) for s in sims]
        b.run(until="graphs", max_graphs=1)          # warm-up (graph capture)
        t = time.perf_counter()
        b.run(until="graphs", max_graphs=graphs)
        el = time.perf_counter() - t
        for mid in ids:
            if mid in b._members:
                b.remove(mid)
        b.close()
        rates[cand] = cells * spg * graphs / el
        if verbose:
            print("batch tune %s: %.3g cell-updates/s" % (cand, rates[cand]))
    best = max(rates, key=rates.get)
    return best, rates
