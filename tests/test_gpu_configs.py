"""Parity at the workloads' full sizes: the HIP path against the CPU oracle.

Every workload the solver is measured on runs here at its real size, through
the C ABI, and its recovered complex fields are compared with the oracle's
(the C restatement of src/fdtd.cpp under the restated Python driver, built
with OpenMP: bit-identical per cell to the serial build, which the compiled
reference kernel pins).  The bar is a relative L2 error <= 1e-6 per complex
component at an identical step count n0; each test prints the six errors.

  strip 120x80x60, 10-cell CPML      converged on both sides: same n0, the
                                     same convergence history, fields
  single large 400x200x120           fixed n0 (the oracle needs ~1 s per
                                     100 steps here): fields
  batch of 32 solves                 every member equal to its own
                                     single-solve run; the smallest forward
                                     (dipole) and adjoint (plane) members
                                     converged against the oracle
  grating 360x60x120, 64 edges       forward fields at a fixed n0 against the
                                     oracle on the solver's own materials and
                                     sources; the gradient of the autograd
                                     function against central differences
                                     (h = 1e-3 um) on two edges, and the
                                     device assembly without the CPML factors
                                     against the reference's per-parameter
                                     loop
  16 served requests 64^3..320x240x120
                                     each equal to its single-solve run; the
                                     smallest against the oracle, converged
"""
import warnings

import numpy as np
import pytest

from tests.cases import (all_fields, batch_case, build_sim, case_from_sim, large_case, rel_l2,
                         run_gpu, run_oracle, serve_case, serve_request, strip_case)

pytestmark = pytest.mark.gpu

TOL = 1e-6
NAMES = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")


def _report(label, got, ref):
    errs = [rel_l2(a, b) for a, b in zip(got, ref)]
    print("%s: rel L2 %s" % (label, " ".join("%s %.2e" % (n, e) for n, e in zip(NAMES, errs))),
          flush=True)
    return errs


def _plain_solve(case):
    sim = build_sim(case)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        if case["adjoint"]:
            sim.solve_adjoint()
        else:
            sim._register(sim._sources)
            sim._solve(0)
    return sim


def test_strip_120x80x60_converged():
    case = strip_case(120, 80, 60, w=10)
    gpu = run_gpu(case)
    ora = run_oracle(case, lib="omp")
    print("strip 120x80x60: n0 gpu %d oracle %d, Tn %d, %d checks; oracle %.1f s"
          % (gpu["n0"], ora["n0"], ora["Tn"], len(ora["history"]), ora["seconds"]), flush=True)
    assert gpu["status"] == "converged" and ora["status"] == "converged"
    assert gpu["n0"] == ora["n0"] and gpu["Tn"] == ora["Tn"]
    np.testing.assert_allclose(gpu["history"], ora["history"], rtol=1e-8, atol=0)
    errs = _report("strip 120x80x60", gpu["fields"], ora["fields"])
    assert max(errs) <= TOL


def test_single_large_400x200x120_fixed_n0():
    case = large_case()
    gpu = run_gpu(case, fixed_n0=0)
    n0 = 12 * gpu["Tn"]
    gpu = run_gpu(case, fixed_n0=n0)
    ora = run_oracle(case, fixed_n0=n0, lib="omp")
    print("400x200x120: n0 %d (Tn %d); oracle %.1f s" % (n0, ora["Tn"], ora["seconds"]), flush=True)
    assert gpu["n0"] == ora["n0"] == n0
    errs = _report("400x200x120", gpu["fields"], ora["fields"])
    assert max(errs) <= TOL
    assert min(np.linalg.norm(f) for f in ora["fields"]) > 0


def test_batch_32_solves_full_size():
    from emopt_amd.batch import Batch
    cases = [batch_case(q) for q in range(32)]
    sims = [build_sim(c) for c in cases]
    b = Batch()
    for s, c in zip(sims, cases):
        b.add(s, adjoint=c["adjoint"], postprocess=False)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        done = b.run()
    b.close()
    assert len(done) == 32
    worst = 0.0
    for q, (s, c) in enumerate(zip(sims, cases)):
        got = all_fields(s, c, c["adjoint"])
        ref_sim = _plain_solve(c)
        for key in ("n0", "Tn", "steps", "status"):
            assert s.last_solve[key] == ref_sim.last_solve[key], (q, key)
        errs = [rel_l2(x, y) for x, y in zip(got, all_fields(ref_sim, c, c["adjoint"]))]
        worst = max(worst, max(errs))
        assert max(errs) <= 1e-12, (q, errs)
        del ref_sim
    print("batch of 32: every member equals its single solve (worst rel L2 %.2e)" % worst)
    cells = [c["Nx"] * c["Ny"] * c["Nz"] for c in cases]
    fwd = min((n, q) for q, n in enumerate(cells) if not cases[q]["adjoint"])[1]
    adj = min((n, q) for q, n in enumerate(cells) if cases[q]["adjoint"])[1]
    for q in (fwd, adj):
        c = cases[q]
        ora = run_oracle(c, lib="omp", adjoint=c["adjoint"])
        got = all_fields(sims[q], c, c["adjoint"])
        label = "batch member %d (%s, %dx%dx%d)" % (q, "adjoint plane" if c["adjoint"] else
                                                    "forward dipole", c["Nx"], c["Ny"], c["Nz"])
        print("%s: n0 batch %d oracle %d" % (label, sims[q].last_solve["n0"], ora["n0"]))
        assert sims[q].last_solve["n0"] == ora["n0"]
        np.testing.assert_allclose(sims[q].convergence_history, ora["history"], rtol=1e-8)
        errs = _report(label, got, ora["fields"])
        assert max(errs) <= TOL


def test_grating_360x60x120_gradient_and_fields():
    import torch
    from emopt_amd.adjoint import fdtd_fom
    from tests.cases import grating_problem
    rtol = 1e-5
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        sim, am, p = grating_problem(rtol=rtol)
        am.fom(p)
    assert (sim.Nx, sim.Ny, sim.Nz) == (360, 60, 120) and len(p) == 64
    n0 = sim.last_solve["n0"]
    # forward fields against the oracle at a fixed step count
    case = case_from_sim(sim, rtol)
    nf = 4 * sim.Tn
    sim.fixed_n0 = nf
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        sim._register(sim._sources)
        sim._solve(0)
    got = all_fields(sim, case)
    ora = run_oracle(case, fixed_n0=nf, lib="omp")
    errs = _report("grating 360x60x120 forward, n0 %d" % nf, got, ora["fields"])
    assert max(errs) <= TOL
    # gradient: every solve runs three times the converged step count
    sim.fixed_n0 = 3 * n0 - (3 * n0) % sim.Tn
    pt = torch.tensor(p, dtype=torch.float64, requires_grad=True)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        F = fdtd_fom(pt, am)
        F.backward()
        g = pt.grad.numpy().copy()
        pidx, comp, cell, deps = am.calc_deps(sim, p)
        g_formula = sim.grad_assemble(pidx, comp, cell, deps, len(p))
        g_ref = am.calc_gradient(sim, p)
        rel_formula = np.linalg.norm(g_formula - g_ref) / np.linalg.norm(g_ref)
        print("grating: F %.6e, converged n0 %d, solves at %d steps; device assembly without "
              "CPML factors vs reference loop %.2e" % (float(F.detach()), n0, sim.fixed_n0,
                                                      rel_formula), flush=True)
        assert rel_formula <= 1e-6
        h = 1e-3
        for i in (0, 42):
            v = []
            for sgn in (1, -1):
                q = p.copy()
                q[i] += sgn * h
                v.append(am.fom(q))
            fd = (v[0] - v[1]) / (2 * h)
            rel = abs(g[i] - fd) / abs(fd)
            print("grating edge %d: adjoint %.6e fd %.6e rel %.2e" % (i, g[i], fd, rel), flush=True)
            assert rel <= 1e-2, (i, g[i], fd)
        am.fom(p)


def test_serving_16_requests_full_range():
    from emopt_amd.serve import Server
    qs = list(range(16))
    svals = np.linspace(0.0, 1.0, len(qs))
    results = {}

    def keep(req, sim):
        c = serve_case(req.rid, materials=False, s=svals[req.rid])
        results[req.rid] = (dict(req.info), all_fields(sim, c))
    srv = Server(hbm_budget=200e9, on_done=keep)
    trace = [(0.05 * q, serve_request(q, s=svals[q])) for q in qs]
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        st = srv.run_trace(trace)
    srv.close()
    print("served 16: %s" % {k: v for k, v in st.items() if k != "span_s"}, flush=True)
    assert st["completed"] == 16 and st["failed"] == 0
    dims = []
    for q in qs:
        c = serve_case(q, s=svals[q])
        dims.append((c["Nx"], c["Ny"], c["Nz"]))
        sim = _plain_solve(c)
        info, got = results[q]
        assert info["n0"] == sim.last_solve["n0"], q
        errs = [rel_l2(x, y) for x, y in zip(got, all_fields(sim, c))]
        assert max(errs) <= 1e-12, (q, errs)
        if q == 0:
            ora = run_oracle(c, lib="omp")
            assert ora["n0"] == info["n0"]
            errs = _report("served request 0 (%dx%dx%d)" % dims[0], got, ora["fields"])
            assert max(errs) <= TOL
        del sim
    print("served dims:", dims)
    assert dims[0] == (64, 64, 64) and dims[-1] == (320, 240, 120)
