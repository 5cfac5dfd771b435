"""Continuous-batching server on the GPU: requests of different grid sizes
arrive over time, are admitted by the HBM budget and join the running batch;
each result is the same alone and in the mix (exactly), and agrees with the
single-solve path (same n0, phasors to 1e-12 relative L2)."""
import warnings

import numpy as np
import pytest

from tests.cases import all_fields, build_sim, rel_l2, serve_case, serve_request

pytestmark = pytest.mark.gpu

SMALL = dict(lo=(32, 30, 28), hi=(60, 50, 40), w=6, n_ell=4)
NREQ = 6


def _serve(qs, offsets, budget=200e9):
    from emopt_amd.serve import Server
    out = {}

    def keep(req, sim):
        c = serve_case(req.rid, materials=False, **SMALL)
        out[req.rid] = (dict(req.info), all_fields(sim, c))
    srv = Server(hbm_budget=budget, steps_per_graph=16, graphs_per_poll=2, on_done=keep)
    trace = [(t, serve_request(q, **SMALL)) for q, t in zip(qs, offsets)]
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        st = srv.run_trace(trace)
    srv.close()
    return out, st


def test_served_results_alone_and_in_mix():
    mix, st = _serve(range(NREQ), [0.0, 0.0, 0.02, 0.05, 0.1, 0.15])
    print("mix:", {k: v for k, v in st.items() if k != "span_s"})
    assert st["completed"] == NREQ and st["failed"] == 0
    for q in range(NREQ):
        alone, _ = _serve([q], [0.0])
        a_info, a_f = alone[q]
        m_info, m_f = mix[q]
        assert a_info["n0"] == m_info["n0"]
        for x, y in zip(a_f, m_f):
            assert np.array_equal(x, y), q
    # against the single-solve path
    for q in (0, 3):
        c = serve_case(q, **SMALL)
        sim = build_sim(c)
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            sim._register(sim._sources)
            sim._solve(0)
        ref = all_fields(sim, c)
        assert sim.last_solve["n0"] == mix[q][0]["n0"]
        errs = [rel_l2(x, y) for x, y in zip(mix[q][1], ref)]
        print("request %d vs single solve: max rel L2 %.2e" % (q, max(errs)))
        assert max(errs) <= 1e-12


def test_budget_limits_concurrency():
    """With a budget that holds about two requests, everything still completes."""
    from emopt_amd.serve import state_bytes
    sizes = []
    for q in range(4):
        c = serve_case(q, materials=False, **SMALL)
        sizes.append(state_bytes(c["Nx"], c["Ny"], c["Nz"], c["w"]))
    budget = 2.2 * max(sizes)
    out, st = _serve(range(4), [0.0] * 4, budget=budget)
    assert st["completed"] == 4


def test_state_estimate_covers_device_allocation():
    """state_bytes (the admission estimate) against the device memory a
    request of the largest serving size (320x240x120) actually takes once
    built and admitted to a batch: the estimate covers it, within 1.5x."""
    from emopt_amd.batch import Batch
    from emopt_amd.serve import state_bytes
    b = Batch()
    c = serve_case(0, s=1.0)
    assert (c["Nx"], c["Ny"], c["Nz"]) == (320, 240, 120)
    free0 = b.free_bytes()
    sim = build_sim(c)
    mid = b.add(sim, postprocess=False)
    used = free0 - b.free_bytes()
    est = state_bytes(c["Nx"], c["Ny"], c["Nz"], c["w"], wavelength=c["wl"], d=c["d"])
    print("320x240x120 request: device memory taken %.3f GB, estimate %.3f GB" % (used / 1e9,
                                                                                 est / 1e9))
    b.remove(mid)
    b.close()
    assert 0 < used <= est <= 1.5 * used
