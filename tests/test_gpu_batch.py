"""Batched solves on the GPU (emf_batch_* through emopt_amd.batch.Batch).

Every member of a batch must give what the single-solve path (emf_solve,
itself pinned to the oracle and the reference fixtures) gives on the same
problem: n0, Tn, step count, status and number of checks equal, the
convergence history to 1e-8 relative (its changes are ratios of nearly equal
sums, which amplify roundoff), the phasors to 1e-12 relative L2 per component
(the per-cell arithmetic is the same; the two kernels are compiled separately
and agree to roundoff, observed ~3e-14).  A member's result must not depend at
all on what else is in the batch, nor on when it was admitted: those
comparisons are exact.
"""
import warnings

import numpy as np
import pytest

from tests.cases import all_fields, batch_case, build_sim, rel_l2

pytestmark = pytest.mark.gpu

SMALL = dict(dims=((40, 64), (36, 52), (30, 40)), w_range=(4, 6))
TOL = 1e-12


def _plain(case):
    sim = build_sim(case)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        if case["adjoint"]:
            sim.solve_adjoint()
        else:
            sim._register(sim._sources)
            sim._solve(0)
    return dict(sim.last_solve), all_fields(sim, case, case["adjoint"]), sim.convergence_history


def _check(sim, case, ref, label):
    ls, f, h = ref
    got = all_fields(sim, case, case["adjoint"])
    errs = [rel_l2(x, y) for x, y in zip(got, f)]
    print(label, "n0", sim.last_solve["n0"], "vs", ls["n0"], "max rel L2 %.2e" % max(errs))
    for key in ("n0", "Tn", "steps", "status", "nchecks"):
        assert sim.last_solve[key] == ls[key], (label, key, sim.last_solve[key], ls[key])
    np.testing.assert_allclose(sim.convergence_history, h, rtol=1e-8)
    assert max(errs) <= TOL, (label, errs)


@pytest.fixture(scope="module")
def cases():
    return [batch_case(q, n_fwd=3, **SMALL) for q in range(6)]


@pytest.fixture(scope="module")
def refs(cases):
    return [_plain(c) for c in cases]


def test_batch_matches_single_solves(cases, refs):
    from emopt_amd.batch import Batch
    b = Batch(steps_per_graph=32)
    sims = [build_sim(c) for c in cases]
    for s, c in zip(sims, cases):
        b.add(s, adjoint=c["adjoint"], postprocess=False)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        done = b.run()
    assert sorted(done) == list(range(len(cases)))
    assert b.running == 0
    for q, (s, c, r) in enumerate(zip(sims, cases, refs)):
        _check(s, c, r, "member %d (%s)" % (q, "adj" if c["adjoint"] else "fwd"))
    b.close()


def test_batch_many_convergence_points():
    """A member with nconv=200 (more sample points than one wave's lanes)
    gives what its single solve gives."""
    from emopt_amd.batch import Batch
    case = batch_case(1, n_fwd=2, **SMALL)
    case["nconv"] = 200
    ref = _plain(case)
    sim = build_sim(case)
    b = Batch(steps_per_graph=32)
    b.add(sim, adjoint=case["adjoint"], postprocess=False)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        assert b.run() == [0]
    _check(sim, case, ref, "nconv=200 member")
    b.close()


def test_batch_members_with_symmetry_boundaries():
    """Members with E / H symmetry on min faces (no CPML there) and
    different CPML widths per face: each equals its single solve, and one
    forward and one adjoint member equal the oracle's solves."""
    from emopt_amd.batch import Batch
    from tests.cases import run_oracle
    bcs = ["0H0", "E00", "HE0", "EEH"]
    cases = []
    for q, bc in enumerate(bcs):
        c = batch_case(q, seed=11, n_fwd=2, **SMALL)      # 0, 1 forward; 2, 3 adjoint
        w = list(c["w"])
        for a, f in enumerate(bc):
            if f in "EH":
                w[2 * a] = 0
        w[1] = max(w[1] - 1, 2)
        c["bc"], c["w"] = bc, w
        cases.append(c)
    refs = [_plain(c) for c in cases]
    b = Batch(steps_per_graph=32)
    sims = [build_sim(c) for c in cases]
    for s_, c in zip(sims, cases):
        b.add(s_, adjoint=c["adjoint"], postprocess=False)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        assert sorted(b.run()) == list(range(len(cases)))
    for q, (s_, c, r) in enumerate(zip(sims, cases, refs)):
        _check(s_, c, r, "%s member (%s)" % (c["bc"], "adj" if c["adjoint"] else "fwd"))
    b.close()
    for q in (0, 2):                       # one forward, one adjoint member
        c = cases[q]
        o = run_oracle(c, lib="omp", adjoint=c["adjoint"])
        got = all_fields(sims[q], c, c["adjoint"])
        errs = [rel_l2(x, y) for x, y in zip(got, o["fields"])]
        print("%s vs oracle: n0 %d / %d, max rel L2 %.2e"
              % (c["bc"], sims[q].last_solve["n0"], o["n0"], max(errs)))
        assert sims[q].last_solve["n0"] == o["n0"]
        assert max(errs) <= 1e-6


def test_batch_alone_equals_in_mix(cases, refs):
    """A member alone in a batch and the same problem among others."""
    from emopt_amd.batch import Batch
    c = cases[4]
    b = Batch(steps_per_graph=16)
    alone = build_sim(c)
    b.add(alone, adjoint=c["adjoint"], postprocess=False)
    b.run()
    mix = [build_sim(x) for x in cases]
    for s, x in zip(mix, cases):
        b.add(s, adjoint=x["adjoint"], postprocess=False)
    b.run()
    a = all_fields(alone, c, c["adjoint"])
    m = all_fields(mix[4], c, c["adjoint"])
    for x, y in zip(a, m):
        assert np.array_equal(x, y)
    _check(mix[4], c, refs[4], "in mix")
    b.close()


def test_batch_continuous_admission(cases, refs):
    """Members admitted while others are half-way through their loops."""
    from emopt_amd.batch import Batch
    b = Batch(steps_per_graph=8)
    sims = {}
    for q in (0, 3):
        sims[q] = build_sim(cases[q])
        b.add(sims[q], adjoint=cases[q]["adjoint"], postprocess=False)
    b.run(until="graphs", max_graphs=20)          # 160 steps in
    for q in (1, 4):
        sims[q] = build_sim(cases[q])
        b.add(sims[q], adjoint=cases[q]["adjoint"], postprocess=False)
    first = b.run(until="any")
    assert len(first) >= 1
    for q in (2, 5):
        sims[q] = build_sim(cases[q])
        b.add(sims[q], adjoint=cases[q]["adjoint"], postprocess=False)
    b.run()
    assert b.running == 0
    for q, s in sims.items():
        _check(s, cases[q], refs[q], "admitted member %d" % q)
    b.close()


@pytest.mark.parametrize("rotate", [False, True])
def test_batch_grouped_stepping(cases, refs, rotate):
    """Cache-sized groups (here about two members per group), stepped first
    group first or in turns replay by replay, with members admitted
    part-way: every member still matches its single solve, and exactly
    matches the same member stepped without groups."""
    from emopt_amd.batch import Batch
    one = max(80 * c["Nx"] * c["Ny"] * c["Nz"] for c in cases)
    b = Batch(steps_per_graph=8, group_bytes=2 * one, rotate_groups=rotate)
    sims = {}
    for q in (0, 1, 2):
        sims[q] = build_sim(cases[q])
        b.add(sims[q], adjoint=cases[q]["adjoint"], postprocess=False)
    b.run(until="graphs", max_graphs=15)
    for q in (3, 4, 5):
        sims[q] = build_sim(cases[q])
        b.add(sims[q], adjoint=cases[q]["adjoint"], postprocess=False)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        b.run()
    assert b.running == 0
    for q, s in sims.items():
        _check(s, cases[q], refs[q], "grouped member %d" % q)
    b.close()
    flat = Batch(steps_per_graph=8, group_bytes=0)
    s4 = build_sim(cases[4])
    flat.add(s4, adjoint=cases[4]["adjoint"], postprocess=False)
    flat.run()
    for x, y in zip(all_fields(s4, cases[4], cases[4]["adjoint"]),
                    all_fields(sims[4], cases[4], cases[4]["adjoint"])):
        assert np.array_equal(x, y)
    flat.close()


def test_batch_lanes_give_identical_results(cases, refs):
    """One group per member, stepped by 1, 2 or 3 lanes (streams) at once,
    the same groups throughout or in turns: every member's result is exactly
    the same, and matches its single solve."""
    from emopt_amd.batch import Batch
    one = max(80 * c["Nx"] * c["Ny"] * c["Nz"] for c in cases)
    out = {}
    for lanes, rotate in ((1, False), (2, False), (3, True), (2, True)):
        b = Batch(steps_per_graph=8, group_bytes=one, rotate_groups=rotate, lanes=lanes)
        sims = [build_sim(c) for c in cases]
        for s, c in zip(sims, cases):
            b.add(s, adjoint=c["adjoint"], postprocess=False)
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            b.run()
        assert b.running == 0
        out[lanes, rotate] = [all_fields(s, c, c["adjoint"]) for s, c in zip(sims, cases)]
        if (lanes, rotate) == (2, False):
            for q, (s, c, r) in enumerate(zip(sims, cases, refs)):
                _check(s, c, r, "2 lanes, member %d" % q)
        b.close()
    base = out[1, False]
    for key, f in out.items():
        for q in range(len(cases)):
            for x, y in zip(f[q], base[q]):
                assert np.array_equal(x, y), (key, q)


def test_batch_direct_launches_equal_graph_replays(cases):
    """The replay's kernels launched one by one (for tracing) give exactly
    the graph's result."""
    from emopt_amd.batch import Batch
    out = []
    for graph in (True, False):
        b = Batch(steps_per_graph=16, use_graph=graph)
        sims = [build_sim(cases[q]) for q in (1, 4)]
        for s, q in zip(sims, (1, 4)):
            b.add(s, adjoint=cases[q]["adjoint"], postprocess=False)
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            b.run()
        out.append([(dict(s.last_solve), all_fields(s, cases[q], cases[q]["adjoint"]))
                    for s, q in zip(sims, (1, 4))])
        b.close()
    for (la, fa), (lb, fb) in zip(*out):
        assert la["n0"] == lb["n0"] and la["steps"] == lb["steps"]
        for x, y in zip(fa, fb):
            assert np.array_equal(x, y)


def test_batch_fixed_n0_and_nmax(cases):
    """fixed_n0 members (status 'fixed') and a member stopped by Nmax
    (status 'nmax') against the single-solve path."""
    from emopt_amd.batch import Batch
    c0, c1 = cases[0], cases[3]
    refs = []
    for c, fix, nmax in ((c0, 300, None), (c1, None, 170.5)):
        s = build_sim(c)
        s.fixed_n0 = fix
        if nmax is not None:
            s.Nmax = nmax
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            if c["adjoint"]:
                s.solve_adjoint()
            else:
                s._register(s._sources)
                s._solve(0)
        refs.append((dict(s.last_solve), all_fields(s, c, c["adjoint"]), s.convergence_history))
    assert refs[0][0]["status"] == "fixed" and refs[1][0]["status"] == "nmax"
    b = Batch(steps_per_graph=16)
    sims = []
    for c, fix, nmax in ((c0, 300, None), (c1, None, 170.5)):
        s = build_sim(c)
        s.fixed_n0 = fix
        if nmax is not None:
            s.Nmax = nmax
        b.add(s, adjoint=c["adjoint"], postprocess=False)
        sims.append(s)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        b.run()
    for s, c, r, lab in zip(sims, (c0, c1), refs, ("fixed", "nmax")):
        _check(s, c, r, lab)
    b.close()


def test_batch_member_handle_is_locked(cases):
    """While a handle is a batch member, other calls on it are refused."""
    from emopt_amd import _lib
    from emopt_amd.batch import Batch
    b = Batch()
    s = build_sim(cases[0])
    mid = b.add(s, postprocess=False)
    with pytest.raises(_lib.EmfError):
        s._lib.emf_reset_pml(s._h)
    b.remove(mid)
    s._lib.emf_reset_pml(s._h)
    b.close()
