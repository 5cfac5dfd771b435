"""The launch-shape autotuner (emopt_amd.autotune) on the GPU.

tune_solver times the candidate step kernels and shapes on the solver's own
problem and keeps the fastest: every candidate computes the same fields (the
step-kernel parity tests check each), so the tuned solve's n0 and phasors
equal the default solve's; a second call on the same grid is answered from
the cache, in memory and in the JSON file.  tune_batch times batch settings
on a set of members; a batch run with its choice gives each member's
single-solve result.
"""
import json
import warnings

import numpy as np
import pytest

from tests.cases import all_fields, batch_case, build_sim, rel_l2, run_gpu, strip_case

pytestmark = pytest.mark.gpu


def test_tune_solver_keeps_results_and_caches(tmp_path):
    from emopt_amd import autotune
    case = strip_case(40, 28, 22, w=5)
    ref = run_gpu(case)
    sim = build_sim(case)
    path = str(tmp_path / "tune.json")
    best, times = autotune.tune_solver(sim, steps=20, warmup=4, cache_path=path)
    print("timed %d candidates, best %s: %s" % (len(times), best,
          " ".join("%s %.4f" % (c, t) for c, t in sorted(times.items(), key=lambda x: x[1]))))
    assert best in autotune.SOLVE_CANDIDATES
    assert len(times) >= len(autotune.SOLVE_CANDIDATES) - 2
    assert times[best] == min(times.values())
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        sim._register(sim._sources)
        sim._solve(0)
    assert sim.last_solve["n0"] == ref["n0"]
    errs = [rel_l2(a, b) for a, b in zip(all_fields(sim, case), ref["fields"])]
    assert max(errs) <= 1e-12, errs
    again, t2 = autotune.tune_solver(build_sim(case), cache_path=path)
    assert again == best and t2 == {}
    with open(path) as f:
        assert list(json.load(f).values()) == [list(best)]


def test_tune_batch_then_batch_equals_single_solves():
    from emopt_amd import autotune
    from emopt_amd.batch import Batch
    small = dict(dims=((40, 64), (36, 52), (30, 40)), w_range=(4, 6))
    cases = [batch_case(q, n_fwd=3, **small) for q in range(3)]
    sims = [build_sim(c) for c in cases]
    cands = [(64, 2, 16), (32, 2, 8), (64, 1, 16)]
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        best, rates = autotune.tune_batch(sims, candidates=cands, graphs=2)
    print("batch candidates %s, best %s" % ({c: "%.3g" % r for c, r in rates.items()}, best))
    assert best in cands and set(rates) == set(cands)
    assert rates[best] == max(rates.values())
    b = Batch(steps_per_graph=best[0], zc=best[1], blocks_per_cu=best[2])
    for s in sims:
        b.add(s, postprocess=False)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        assert sorted(b.run()) == [0, 1, 2]
    b.close()
    for s, c in zip(sims, cases):
        ref = run_gpu(c)
        errs = [rel_l2(a, r) for a, r in zip(all_fields(s, c), ref["fields"])]
        assert s.last_solve["n0"] == ref["n0"] and max(errs) <= 1e-12, (s.last_solve, errs)
