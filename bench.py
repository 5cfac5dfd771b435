#!/usr/bin/env python
"""Benchmark: FP64 CW-FDTD Yee cell-updates/s on one MI355X.

Workload (the single-GPU headline case): 400x200x120 grid at dx = 0.03 um,
12-cell CPML on every face, a 240x120x10 design slab of 40 seeded ellipses
(eps 3.48^2 in 1.444^2, boundary cells area-averaged), lambda = 1.31 um,
planar Ey source across the input face.  Synthetic inputs, built on the host
and uploaded once; the timed region starts with everything resident in HBM.

A step is one full Yee time step (H pass + E pass, CPML slabs, source
injection) with the on-device convergence check every Tn steps, as the solve
loop runs it.  value = Nx*Ny*Nz*K / (wall time of K steps); the library call
that runs them synchronises its stream before it returns.

Also reported:
  roofline      the dominant kernel's algorithmic bytes per launch divided by
                its average launch time (each launch's own start and stop HIP
                events, recorded by hipExtLaunchKernelGGL on the solver's
                stream, in a second run of min(K, 200) steps after the timed
                region), against 8 TB/s; `traffic` is the HBM bytes per launch
                measured by rocprofv3 counters (FETCH_SIZE, WRITE_SIZE, scaled
                by the same counters on a plain copy of known size), read from
                the committed summary profiles/hbm_traffic_step_400x200x120.md
  cpu_baseline  EMopt's own kernel (src/fdtd.cpp compiled from its sources),
                one thread, on a bounded sample of the same workload
  cpu_all_cores the C restatement built with OpenMP, all host threads
  solve         one full converged solve (solves/hour)
  batched       32 independent solves of uneven size and wavelength (half
                forward dipoles, half adjoint plane sources): one after the
                other through the single-solve path, then all together in one
                batch; solves/hour, latency p50/p99, result agreement
  serving       Poisson arrivals (2, 8 and 32 per second) of solve requests of
                mixed grid size served by the continuous-batching server under
                an HBM budget, and the same requests through the single-solve
                path as a FIFO queue; solves/hour and latency p50/p99 of both
  differentiation  the grating workload (360x60x120, 64 edge parameters):
                forward and backward (adjoint solve + gradient assembly)
                times, the gradient against central differences on 4 edges
  error_vs_reference  the 400x200x120 phasors at a fixed step count against
                the oracle, relative L2 per component

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--batch N] [--serve-requests R]
    python bench.py --profile-summary --stats S.csv [--fetch F.csv --write W.csv]
"""
from __future__ import annotations

import argparse
import csv
import ctypes
import json
import os
import re
import sys
import time
import warnings

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
B_ALG_H = 72    # bytes per cell, H pass: read E(3) H(3), write H(3); mu == 1
B_ALG_E = 96    # bytes per cell, E pass: read H(3) E(3) b = dt/eps (3), write E(3)
B_ALG_STEP = B_ALG_H + B_ALG_E
KERNELS = {"k_pass_h": B_ALG_H, "k_pass_e": B_ALG_E}
COPY_BYTES = 1 << 31            # k_copy reads and writes this many bytes
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_step_400x200x120.md")


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


def time_steps(sim, n_start, nsteps, per_kernel):
    out = np.zeros(3)
    sim._lib.emf_time_steps(sim._h, n_start, nsteps, sim.Tn, 1 if per_kernel else 0,
                            out.ctypes.data_as(ctypes.c_void_p))
    return out


def _oracle_sim(case, path):
    from oracle import driver as D
    from tests.cases import _configure
    lib = D.load(path)
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
    return o


def _time_cpu(o, seconds):
    o._step(0)   # warm-up (page-in)
    n, t0 = 1, time.perf_counter()
    while True:
        o._step(n)
        n += 1
        el = time.perf_counter() - t0
        if el >= seconds and n >= 3:
            return n - 1, el


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(case, seconds):
    """EMopt's own kernel (oracle/_ref, compiled from src/fdtd.cpp) driven by
    the oracle's restatement of the Python time loop, one thread."""
    from oracle import driver as D
    kind = "reference" if os.path.exists(D.REF_LIB) else "port"
    o = _oracle_sim(case, D.REF_LIB if kind == "reference" else D.ORACLE_LIB)
    steps, el = _time_cpu(o, seconds)
    cells = case["Nx"] * case["Ny"] * case["Nz"]
    src = "EMopt src/fdtd.cpp" if kind == "reference" else "C restatement of src/fdtd.cpp"
    return dict(value=cells * steps / el, unit="cell-updates/s", cores=1, kind=kind,
                sample="%d time steps of the %dx%dx%d workload (%.1f s), %s -O3 "
                       "-march=x86-64-v3, 1 thread, Python time loop with one-rank ghost "
                       "exchange; host CPU %s" % (steps, case["Nx"], case["Ny"], case["Nz"], el,
                                                  src, _cpu_model()))


def cpu_all_cores(case, seconds):
    """The C restatement built with -fopenmp (cell loops over all threads,
    bit-identical per cell to the serial build)."""
    from oracle import driver as D
    if not os.path.exists(D.ORACLE_OMP_LIB):
        return None
    threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    o = _oracle_sim(case, D.ORACLE_OMP_LIB)
    steps, el = _time_cpu(o, seconds)
    cells = case["Nx"] * case["Ny"] * case["Nz"]
    return dict(value=cells * steps / el, unit="cell-updates/s", cores=threads, kind="port",
                sample="%d time steps of the %dx%dx%d workload (%.1f s), oracle/cfdtd.c -O3 "
                       "-fopenmp, %d threads; host CPU %s"
                       % (steps, case["Nx"], case["Ny"], case["Nz"], el, threads, _cpu_model()))


def error_vs_reference(case, sim, n0):
    """Relative L2 error per recovered complex component (Ex..Hz) of the
    workload's solve at a fixed step count n0, against the oracle (the C
    restatement of src/fdtd.cpp, OpenMP build, bit-identical per cell to the
    serial build that the compiled reference kernel pins) on the host."""
    from tests.cases import all_fields, rel_l2, run_oracle
    sim.fixed_n0 = n0
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        sim._register(sim._sources)
        sim._solve(0)
    sim.fixed_n0 = None
    got = all_fields(sim, case)
    ora = run_oracle(case, fixed_n0=n0, lib="omp")
    errs = [rel_l2(a, b) for a, b in zip(got, ora["fields"])]
    threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    return {"n0": int(n0), "Tn": int(ora["Tn"]), "steps": int(n0 + 2 * ora["Tn"]),
            "rel_l2": dict(zip(("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"), errs)),
            "max_rel_l2": max(errs), "tolerance": 1e-6,
            "oracle": "oracle/cfdtd.c (OpenMP, %d threads) under the restated Python loop, "
                      "%.1f s" % (threads, ora["seconds"])}


def read_traffic(path=TRAFFIC_FILE):
    """Per-launch HBM bytes recorded by the counter runs (see --profile-summary)."""
    out = {}
    if os.path.exists(path):
        with open(path) as f:
            for line in f:
                m = re.match(r"^- (\w+): hbm_bytes_per_launch = ([0-9.eE+]+)", line.strip())
                if m:
                    out[m.group(1)] = float(m.group(2))
    return out


def _pcts(lat):
    lat = np.asarray(lat, dtype=np.double)
    return dict(p50_s=float(np.percentile(lat, 50)), p99_s=float(np.percentile(lat, 99)))


def _plain_solve(sim, adjoint):
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        if adjoint:
            sim._register(sim._adj_sources)
            sim._solve(1)
        else:
            sim._register(sim._sources)
            sim._solve(0)


def copy_curve():
    """Copy rate (GB/s, read + write) against the working set (both buffers),
    from 16 MB to 4 GB: working sets inside the 256 MB Infinity Cache are
    served from it when the copy is repeated."""
    from emopt_amd import _lib
    L = _lib.load()
    out = []
    for mb in (8, 16, 32, 48, 64, 96, 128, 192, 256, 512, 1024, 2048):
        b = mb << 20
        reps = int(min(2000, max(10, (40 << 30) // b)))
        g = np.zeros(1)
        L.emf_copy_bandwidth(0, b, reps, g.ctypes.data_as(_lib.c_void_p))
        out.append({"working_set_MB": 2 * mb, "GBps": float(g[0]), "reps": reps})
        log("[bench] copy, working set %5d MB: %.0f GB/s" % (2 * mb, g[0]))
    return out


def bench_batched(n, seed=3, direct=False, plain=True):
    """The batched workload: n solves, dims uniform in [64,256]x[64,160]x
    [48,96] at d = 0.04 um, lambda 1.50-1.60 um, 8-12 cell CPML, half with a
    forward Jz dipole, half with an adjoint plane source.  All submitted at
    t = 0: the single-solve path runs them one after the other, the batch
    runs them together."""
    from emopt_amd.batch import Batch
    from tests.cases import all_fields, batch_case, build_sim, rel_l2
    t = time.perf_counter()
    cases = [batch_case(q, seed=seed, n_fwd=n // 2) for q in range(n)]
    cells = [c["Nx"] * c["Ny"] * c["Nz"] for c in cases]
    log("[bench] batched workload built in %.1fs (%d solves, %.3g cells)"
        % (time.perf_counter() - t, n, sum(cells)))
    sims = [build_sim(c) for c in cases]
    _plain_solve(build_sim(cases[0]), cases[0]["adjoint"])        # warm-up
    plain_t, plain_info = [], []
    sample = list(range(0, n, max(1, n // 8)))
    if plain:
        t0 = time.perf_counter()
        for sim, c in zip(sims, cases):
            _plain_solve(sim, c["adjoint"])
            plain_t.append(time.perf_counter() - t0)
            plain_info.append(dict(sim.last_solve))
        tp = time.perf_counter() - t0
        ref = {q: all_fields(sims[q], cases[q], cases[q]["adjoint"]) for q in sample}
        log("[bench] batched workload, single-solve path: %.2fs" % tp)
    b = Batch(use_graph=not direct)
    t0 = time.perf_counter()
    for sim, c in zip(sims, cases):
        b.add(sim, adjoint=c["adjoint"], postprocess=False)
    if direct:
        # under a tracer: report progress every 50 replays
        while b.running:
            b.run(until="graphs", max_graphs=50)
            log("[bench] batched workload: %d solves running after %.1fs"
                % (b.running, time.perf_counter() - t0))
    else:
        b.run()
    tb = time.perf_counter() - t0
    log("[bench] batched workload, one batch%s: %.2fs" % (" (direct launches)" if direct else "", tb))
    if not plain:
        info = [dict(sims[q].last_solve) for q in range(n)]
        steps = sum(i["steps"] * c for i, c in zip(info, cells))
        b.close()
        return {"workload": "batch_%d_dims64-256x64-160x48-96_dx0.04_wl1.50-1.60_cpml8-12" % n,
                "solves": n, "cells_total": int(sum(cells)), "cell_updates_total": int(steps),
                "direct_launches": bool(direct),
                "batched": dict(seconds=tb, solves_per_hr=n * 3600 / tb,
                                cell_updates_per_s=steps / tb)}
    n0_same = all(sims[q].last_solve["n0"] == plain_info[q]["n0"] for q in range(n))
    err = max(max(rel_l2(x, y) for x, y in zip(all_fields(sims[q], cases[q], cases[q]["adjoint"]),
                                               ref[q])) for q in sample)
    steps = sum(i["steps"] * c for i, c in zip(plain_info, cells))
    lat_b = [v["seconds"] for v in b.finished.values()]
    b.close()
    return {"workload": "batch_%d_dims64-256x64-160x48-96_dx0.04_wl1.50-1.60_cpml8-12" % n,
            "solves": n, "cells_total": int(sum(cells)), "cell_updates_total": int(steps),
            "plain": dict(seconds=tp, solves_per_hr=n * 3600 / tp, cell_updates_per_s=steps / tp,
                          **_pcts(plain_t)),
            "batched": dict(seconds=tb, solves_per_hr=n * 3600 / tb, cell_updates_per_s=steps / tb,
                            **_pcts(lat_b)),
            "direct_launches": bool(direct),
            "speedup_solves_per_hr": tp / tb, "n0_identical": bool(n0_same),
            "max_rel_l2_vs_plain_sample": err, "sample": sample}


def bench_serving(rates, n_requests, seed=5, budget=200e9, max_running=8):
    """The serving workload (dims log-uniform 64^3..320x240x120, d 0.03-0.05
    um, lambda 1.26-1.62 um, random ellipses) at each Poisson arrival rate:
    the same n_requests requests arrive at that rate (the window is the last
    arrival time, stated per rate) and are served by the continuous-batching
    server under the HBM budget; then the single-solve path as a FIFO queue
    with the same arrival times, each request's service time being its
    problem build (host material sampling and upload) plus its solve,
    measured one by one once -- the work a plain request loop does per
    request."""
    from emopt_amd.serve import Server
    from tests.cases import build_sim, serve_case, serve_request
    # single-solve service times, once for all rates
    service, steps, build_s = [], 0, 0.0
    for q in range(n_requests):
        t0 = time.perf_counter()
        c = serve_case(q, seed=seed)
        sim = build_sim(c)
        t1 = time.perf_counter()
        _plain_solve(sim, False)
        service.append(time.perf_counter() - t0)
        build_s += t1 - t0
        steps += sim.last_solve["steps"] * c["Nx"] * c["Ny"] * c["Nz"]
        del sim
        if (q + 1) % 32 == 0:
            log("[bench] serving: single-solve service, %d of %d requests" % (q + 1, n_requests))
    log("[bench] serving: single-solve service of %d requests %.1fs" % (n_requests, sum(service)))
    out = {"requests": n_requests, "budget_bytes": budget, "max_running": max_running,
           "plain_busy_seconds": float(sum(service)), "plain_build_seconds": build_s,
           "rates": []}
    for k, rate in enumerate(rates):
        rng = np.random.default_rng([seed, k])
        arr = list(np.cumsum(rng.exponential(1.0 / rate, size=n_requests)))
        reqs = [serve_request(q, seed=seed) for q in range(n_requests)]
        ndone = [0]

        def progress(req, sim):
            ndone[0] += 1
            if ndone[0] % 32 == 0:
                log("[bench] serving %g/s: %d of %d requests done" % (rate, ndone[0], n_requests))

        srv = Server(hbm_budget=budget, prep_workers=8, max_running=max_running, on_done=progress)
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            st = srv.run_trace(list(zip(arr, reqs)))
        srv.close()
        fin, lat = 0.0, []
        for a, sv in zip(arr, service):
            fin = max(fin, a) + sv
            lat.append(fin - a)
        plain = dict(solves_per_hr=n_requests * 3600 / fin, cell_updates_per_s=steps / fin,
                     **_pcts(lat))
        log("[bench] serving %g/s over %.1fs: served p50 %.2fs p99 %.2fs %.0f solves/h; FIFO p50 "
            "%.2fs p99 %.2fs %.0f solves/h" % (rate, arr[-1], st.get("p50_s", 0),
                                               st.get("p99_s", 0), st.get("solves_per_hr", 0),
                                               plain["p50_s"], plain["p99_s"],
                                               plain["solves_per_hr"]))
        out["rates"].append({
            "workload": "serve_poisson_%gps_%dreq_window%.1fs_dims64^3-320x240x120"
                        % (rate, n_requests, arr[-1]),
            "rate_per_s": rate, "window_s": float(arr[-1]), "served": st, "plain_fifo": plain,
            "waits_on_free_hbm": srv.waits_on_free,
            "p99_ratio_plain_over_served": plain["p99_s"] / st["p99_s"] if st.get("p99_s") else None,
            "solves_per_hr_ratio": (st["solves_per_hr"] / plain["solves_per_hr"]
                                    if st.get("solves_per_hr") else None)})
    return out


def bench_materials(reps=3):
    """Material sampling of the 400x200x120 workload (40 ellipses in a
    240x120x10 design slab over cladding): the three eps components at their
    Yee offsets on the host sampler and on the device (emf_sample_structured,
    including the copy back to host memory), best of `reps`, under the
    reference's exact overlap rule (the default) and the supersampled one;
    the host and device results are compared for equality."""
    from emopt_amd.fdtd import FDTD
    from tests.cases import ellipse_design_eps
    Nx, Ny, Nz, d = 400, 200, 120, 0.03
    shifts = FDTD.YEE_SHIFTS[:3]
    res = {}
    for overlap in ("exact", "supersample"):
        eps = ellipse_design_eps(Nx, Ny, Nz, d, seed=2, overlap=overlap)
        eps.get_values_device(0, 8, 0, 8, 0, 8)    # warm-up
        th, td, same = [], [], True
        for _ in range(reps):
            t = time.perf_counter()
            h = [eps.get_values(0, Nx, 0, Ny, 0, Nz, sx=a, sy=b, sz=c) for a, b, c in shifts]
            th.append(time.perf_counter() - t)
            t = time.perf_counter()
            g = [eps.get_values_device(0, Nx, 0, Ny, 0, Nz, sx=a, sy=b, sz=c) for a, b, c in shifts]
            td.append(time.perf_counter() - t)
            same = same and all(np.array_equal(x, y) for x, y in zip(g, h))
        res[overlap] = {"host_s": min(th), "device_s": min(td), "speedup": min(th) / min(td),
                        "identical": bool(same)}
    out = {"workload": "eps_xyz_400x200x120_ellipses40", "rule": "exact"}
    out.update(res["exact"])
    out["supersample"] = res["supersample"]
    return out


def bench_diff(n_fd=4, rtol=1e-5):
    """The differentiation workload at full size (360x60x120 grating, 64
    edge parameters): gradient through the autograd function (forward +
    adjoint solve + device assembly), the reference gradient formula on the
    same fields (device assembly without the CPML / dispersion factors, and
    the reference's per-parameter loop), and central differences (h = 1e-3
    um) on n_fd edges.  After a first converged solve every solve runs a
    fixed step count, three times the converged one, so the differences are
    not limited by the convergence test's phase tolerance sqrt(rtol)."""
    import torch
    from emopt_amd.adjoint import fdtd_fom
    from tests.cases import grating_problem
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        sim, am, p = grating_problem(rtol=rtol)
        am.fom(p)
        n0 = sim.last_solve["n0"]
        sim.fixed_n0 = 3 * n0 - (3 * n0) % sim.Tn
        pt = torch.tensor(p, dtype=torch.float64, requires_grad=True)
        t = time.perf_counter()
        F = fdtd_fom(pt, am)
        t_fwd = time.perf_counter() - t
        t = time.perf_counter()
        F.backward()
        t_bwd = time.perf_counter() - t
        g = pt.grad.numpy().copy()
        pidx, comp, cell, deps = am.calc_deps(sim, p)
        g_formula = sim.grad_assemble(pidx, comp, cell, deps, len(p))
        t = time.perf_counter()
        g_ref = am.calc_gradient(sim, p)
        t_ref = time.perf_counter() - t
        log("[bench] diff: F %.4e, forward %.2fs, backward %.2fs, reference loop %.2fs"
            % (float(F.detach()), t_fwd, t_bwd, t_ref))
        idx = list(np.linspace(0, len(p) - 1, n_fd).astype(int))
        fd = []
        for i in idx:
            v = []
            for sgn in (1, -1):
                q = p.copy()
                q[i] += sgn * 1e-3
                v.append(am.fom(q))
            fd.append((v[0] - v[1]) / 2e-3)
        am.fom(p)
    fd = np.array(fd)
    F = float(F.detach())
    return {"workload": "grating_360x60x120_dx0.03_64edges_gauss_w0_10.4", "F": float(F),
            "params": len(p), "converged_n0": n0, "fixed_steps": int(sim.fixed_n0),
            "forward_s": t_fwd, "backward_s": t_bwd, "reference_loop_s": t_ref,
            "formula_device_vs_reference_loop": float(np.linalg.norm(g_formula - g_ref)
                                                      / np.linalg.norm(g_ref)),
            "exact_vs_reference_formula": float(np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref)),
            "fd_formula": g_ref[idx].tolist(),
            "fd_edges": [int(i) for i in idx], "fd": fd.tolist(), "adjoint": g[idx].tolist(),
            "fd_rel_err": (np.abs(g[idx] - fd) / np.abs(fd)).tolist()}


# --------------------------------------------------------- profile summaries

def _short(name):
    for k in list(KERNELS) + ["k_copy", "k_conv_check", "k_step_fused", "k_step_lds"]:
        if k in name:
            return k
    return name.split("(")[0][:48]


def _read_counters(path, counter, totals=False):
    """Average counter value per dispatch (totals: (sum, dispatches)), by
    kernel, from a rocprofv3 counter_collection.csv."""
    acc = {}
    with open(path) as f:
        for row in csv.DictReader(f):
            if row.get("Counter_Name") != counter:
                continue
            k = _short(row["Kernel_Name"])
            s, n = acc.get(k, (0.0, 0))
            acc[k] = (s + float(row["Counter_Value"]), n + 1)
    if totals:
        return acc
    return {k: s / n for k, (s, n) in acc.items()}


def profile_summary(args):
    lines = []
    cmd = args.command or "(command not given)"
    title = args.workload or "400x200x120 step workload"
    step_workload = args.workload is None
    if args.stats:
        lines += ["# Kernel durations, " + title, "",
                  "Command:", "", "    " + cmd, "",
                  "| kernel | calls | avg us | min us | max us | total % |",
                  "|---|---|---|---|---|---|"]
        with open(args.stats) as f:
            for row in csv.DictReader(f):
                lines.append("| %s | %s | %.2f | %.2f | %.2f | %.1f |" % (
                    _short(row["Name"]), row["Calls"], float(row["AverageNs"]) / 1e3,
                    float(row["MinNs"]) / 1e3, float(row["MaxNs"]) / 1e3,
                    float(row["Percentage"])))
        if step_workload:
            cells = 400 * 200 * 120
            lines += ["", "Algorithmic bytes per launch: k_pass_h %d B/cell, k_pass_e %d B/cell "
                      "x %d cells." % (B_ALG_H, B_ALG_E, cells)]
        if args.cell_updates:
            lines += ["", "Time per cell-update (all launches of the kernel / %.6g cell-updates):"
                      % args.cell_updates, ""]
            with open(args.stats) as f:
                for row in csv.DictReader(f):
                    lines.append("- %s: %.2f ps" % (_short(row["Name"]),
                                                   float(row["TotalDurationNs"]) * 1e3
                                                   / args.cell_updates))
    if args.fetch and args.write and args.cell_updates:
        # a workload whose launches differ in size: total bytes per kernel
        fetch = _read_counters(args.fetch, "FETCH_SIZE", totals=True)
        write = _read_counters(args.write, "WRITE_SIZE", totals=True)
        ff = COPY_BYTES / (fetch["k_copy"][0] / fetch["k_copy"][1] * 1024.0)
        fw = COPY_BYTES / (write["k_copy"][0] / write["k_copy"][1] * 1024.0)
        lines += ["# HBM traffic, " + title, "",
                  "Commands (one counter per run):", "", "    " + (args.fetch_command or "?"),
                  "    " + (args.write_command or "?"), "",
                  "Scale: k_copy moves %d B each way; FETCH_SIZE reads %.1f KiB and WRITE_SIZE "
                  "%.1f KiB for it, so the factors are %.3f (fetch) and %.3f (write)."
                  % (COPY_BYTES, fetch["k_copy"][0] / fetch["k_copy"][1],
                     write["k_copy"][0] / write["k_copy"][1], ff, fw), "",
                  "Totals over all launches; per cell-update = total / %.6g cell-updates." %
                  args.cell_updates, "",
                  "| kernel | launches | scaled read GB | scaled write GB | B per cell-update |",
                  "|---|---|---|---|---|"]
        for k in sorted(set(fetch) & set(write)):
            rd, wr = fetch[k][0] * 1024 * ff, write[k][0] * 1024 * fw
            lines.append("| %s | %d | %.4g | %.4g | %.2f |" % (
                k, fetch[k][1], rd / 1e9, wr / 1e9,
                0.0 if k == "k_copy" else (rd + wr) / args.cell_updates))
    elif args.fetch and args.write:
        fetch = _read_counters(args.fetch, "FETCH_SIZE")
        write = _read_counters(args.write, "WRITE_SIZE")
        ff = COPY_BYTES / (fetch["k_copy"] * 1024.0)
        fw = COPY_BYTES / (write["k_copy"] * 1024.0)
        cells = 400 * 200 * 120
        lines += ["# HBM traffic per launch, " + title, "",
                  "Commands (one counter per run):", "", "    " + (args.fetch_command or "?"),
                  "    " + (args.write_command or "?"), "",
                  "Scale: k_copy moves %d B each way; FETCH_SIZE reads %.1f KiB and WRITE_SIZE "
                  "%.1f KiB for it, so the factors are %.3f (fetch) and %.3f (write)."
                  % (COPY_BYTES, fetch["k_copy"], write["k_copy"], ff, fw), "",
                  "| kernel | FETCH_SIZE KiB | WRITE_SIZE KiB | scaled read B | scaled write B | "
                  "HBM B/launch | algorithmic B/launch |", "|---|---|---|---|---|---|---|"]
        res = []
        for k in sorted(set(fetch) & set(write)):
            rd, wr = fetch[k] * 1024 * ff, write[k] * 1024 * fw
            alg = (KERNELS.get(k, 0) * cells if (k in KERNELS and step_workload)
                   else (2 * COPY_BYTES if k == "k_copy" else 0))
            lines.append("| %s | %.0f | %.0f | %.4g | %.4g | %.4g | %s |" % (
                k, fetch[k], write[k], rd, wr, rd + wr, "%.4g" % alg if alg else "-"))
            res.append((k, rd + wr))
        lines += [""] + ["- %s: hbm_bytes_per_launch = %.6g" % (k, t) for k, t in res]
    print("\n".join(lines))


# ---------------------------------------------------------------------- main

def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=40)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--omp-seconds", type=float, default=6.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--error-periods", type=int, default=28,
                    help="compare the workload's phasors with the oracle's at n0 = this many "
                         "Tn (0: skip); 28 Tn is past the source ramp")
    ap.add_argument("--batch", type=int, default=32, help="solves in the batched section (0: skip)")
    ap.add_argument("--batch-direct", action="store_true",
                    help="batched section: launch each replay's kernels directly (tracing)")
    ap.add_argument("--batch-no-plain", action="store_true",
                    help="batched section: skip the single-solve leg")
    ap.add_argument("--copy-curve", action="store_true",
                    help="with --only batched: copy rate against working-set size")
    ap.add_argument("--serve-requests", type=int, default=48,
                    help="requests per arrival rate in the serving section (0: skip)")
    ap.add_argument("--serve-rates", default="2,8,32", help="Poisson arrival rates, per second")
    ap.add_argument("--serve-running", type=int, default=8, help="solves admitted to the batch at once")
    ap.add_argument("--no-diff", action="store_true", help="skip the differentiation workload")
    ap.add_argument("--only", choices=["batched", "serving", "differentiation"],
                    help="run only this section (for profiling it) and print its JSON")
    ap.add_argument("--profile-summary", action="store_true")
    ap.add_argument("--stats")
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--command")
    ap.add_argument("--workload", help="title of the profiled workload (default: the step one)")
    ap.add_argument("--cell-updates", type=float,
                    help="profile summary of a workload whose launches differ in size: report "
                         "totals per cell-update")
    ap.add_argument("--fetch-command")
    ap.add_argument("--write-command")
    args = ap.parse_args()
    if args.profile_summary:
        profile_summary(args)
        return
    if args.gpus != 1:
        print("bench.py: this build runs on exactly one GPU (single-GPU solver, no "
              "multi-GPU path); use --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.only:
        if args.only == "batched":
            out = bench_batched(args.batch, direct=args.batch_direct,
                                plain=not args.batch_no_plain)
            if args.copy_curve:
                out["copy_curve"] = copy_curve()
            from emopt_amd import _lib
            gbs = np.zeros(1)     # the copy kernel that scales the traffic counters
            _lib.load().emf_copy_bandwidth(0, COPY_BYTES, 3, gbs.ctypes.data_as(_lib.c_void_p))
            out["copy_GBps"] = float(gbs[0])
        elif args.only == "serving":
            out = bench_serving([float(r) for r in args.serve_rates.split(",")],
                                args.serve_requests, max_running=args.serve_running)
        else:
            out = bench_diff()
        print(json.dumps({args.only: out}), flush=True)
        return

    from tests.cases import large_case
    from emopt_amd import _lib
    t = time.perf_counter()
    case = large_case()
    log("[bench] workload built in %.1fs" % (time.perf_counter() - t))
    sim = make_sim(case)
    cells = case["Nx"] * case["Ny"] * case["Nz"]
    L = sim._lib

    # Set-up: step until the wave from the source has reached every cell
    # (Nx + Ny + Nz steps; the stencil spreads one cell per step).  Cells
    # still exactly zero make a step measurably cheaper (~5 % on an all-zero
    # grid), and the bulk of a real solve has none.  Then the W warm-up
    # steps, then the timed region: exactly K steps, no per-launch events.
    L.emf_reset_pml(sim._h)
    L.emf_zero_fields(sim._h)
    n_fill = case["Nx"] + case["Ny"] + case["Nz"]
    time_steps(sim, 0, n_fill, False)
    time_steps(sim, n_fill, args.warmup, False)
    log("[bench] set-up (%d steps) and warm-up (%d steps) done" % (n_fill, args.warmup))
    n_start = n_fill + args.warmup
    t0 = time.perf_counter()
    dev = time_steps(sim, n_start, args.steps, False)
    wall = time.perf_counter() - t0
    value = cells * args.steps / wall
    log("[bench] %d steps: wall %.3fs, device %.3fs" % (args.steps, wall, dev[0] / 1e3))

    # per-kernel launch times, measured after the timed region
    nk = min(args.steps, 200)
    pk = time_steps(sim, n_start + args.steps, nk, True)
    ms = {"k_pass_h": pk[1] / nk, "k_pass_e": pk[2] / nk}
    log("[bench] per launch: k_pass_h %.3f ms, k_pass_e %.3f ms" % (ms["k_pass_h"], ms["k_pass_e"]))
    gbs = np.zeros(1)
    L.emf_copy_bandwidth(0, COPY_BYTES, 10, gbs.ctypes.data_as(_lib.c_void_p))

    traffic = read_traffic()
    kernels = {}
    for k, b in KERNELS.items():
        ach = b * cells / (ms[k] * 1e-3) / 1e9
        kernels[k] = {"avg_ms": ms[k], "alg_bytes_per_launch": b * cells, "GBps": ach,
                      "hbm_bytes_per_launch": traffic.get(k)}
    dom = max(kernels, key=lambda k: kernels[k]["avg_ms"])
    ach = kernels[dom]["GBps"]
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
                   "eps": "40 seeded ellipses as 128-gons in a 240x120x10 slab, the "
                          "reference's overlap rule",
                   "bytes_per_cell_update": B_ALG_STEP,
                   "step_kernels": "k_pass_h + k_pass_e (two launches per step)"},
        "hbm_frac_168B": value * B_ALG_STEP / (PEAK_HBM_GBS * 1e9),
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": ach, "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": ach / PEAK_HBM_GBS,
                     "traffic": traffic.get(dom),
                     "alg_bytes_per_launch": kernels[dom]["alg_bytes_per_launch"],
                     "avg_launch_ms": kernels[dom]["avg_ms"]},
        "kernels": kernels,
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

    if args.error_periods > 0:
        t = time.perf_counter()
        res["error_vs_reference"] = error_vs_reference(case, sim, args.error_periods * sim.Tn)
        ev = res["error_vs_reference"]
        log("[bench] error vs oracle at n0 = %d: max rel L2 %.2e (%.1fs)"
            % (ev["n0"], ev["max_rel_l2"], time.perf_counter() - t))

    if not args.no_solve:
        res["materials"] = bench_materials()
        mt = res["materials"]
        log("[bench] materials, exact rule: host %.3fs, device %.3fs, identical %s; supersampled: "
            "host %.3fs, device %.3fs, identical %s" % (
                mt["host_s"], mt["device_s"], mt["identical"], mt["supersample"]["host_s"],
                mt["supersample"]["device_s"], mt["supersample"]["identical"]))
    if args.batch > 0:
        res["batched"] = bench_batched(args.batch)
    if args.serve_requests > 0:
        res["serving"] = bench_serving([float(r) for r in args.serve_rates.split(",")],
                                       args.serve_requests, max_running=args.serve_running)
    if not args.no_diff:
        res["differentiation"] = bench_diff()

    if not args.no_cpu:
        del sim
        cb = cpu_baseline(case, args.cpu_seconds)
        res["cpu_baseline"] = cb
        log("[bench] cpu baseline %.3g cell-updates/s" % cb["value"])
        if args.omp_seconds > 0:
            co = cpu_all_cores(case, args.omp_seconds)
            if co:
                res["cpu_all_cores"] = co
                log("[bench] cpu all cores %.3g cell-updates/s" % co["value"])
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
