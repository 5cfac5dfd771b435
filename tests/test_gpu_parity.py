"""GPU parity: the HIP path (through libemfdtd.so) against the CPU oracle.

Bar (from the workload definition): relative L2 error <= 1e-6 per recovered
complex component at an identical step count; in practice the two agree to
~1e-12, the residue of FMA contraction and libm last-bit differences.
"""
import numpy as np
import pytest

from tests.cases import random_case, rel_l2, rel_l2_set, run_gpu, run_oracle, strip_case

TOL = 1e-6          # per-component relative L2, the workload's bar
STEP_TOL = 1e-11    # time-domain fields after a few steps

pytestmark = pytest.mark.gpu


def _step_gpu(case, nsteps):
    """Run nsteps raw H/E half steps through the ABI; return Ex..Hz."""
    import warnings
    from emopt_amd.fdtd import FDTD
    from emopt_amd.misc import DomainCoordinates
    from tests.cases import _configure
    d = case["d"]
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        sim = FDTD(case["X"], case["Y"], case["Z"], d, d, d, case["wl"],
                   complex_eps=case.get("ceps", False))
        _configure(sim, case, DomainCoordinates)
    sim.set_material_arrays(case["mats"][:3], case["mats"][3:])
    sim._register(sim._sources)
    L, h = sim._lib, sim._h
    L.emf_reset_pml(h)
    L.emf_zero_fields(h)
    dt = sim.dt
    for n in range(nsteps):
        L.emf_update_H(h, n, n * dt)
        L.emf_update_E(h, n, (n + 0.5) * dt)
    return sim.time_fields()


def _step_oracle(case, nsteps):
    import warnings
    from oracle import driver as D
    from tests.cases import _configure
    d = case["d"]
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        o = D.OracleFDTD(D.load(D.ORACLE_LIB), case["X"], case["Y"], case["Z"], d, d, d,
                         case["wl"], complex_eps=case.get("ceps", False))
        _configure(o, case, D.Domain)
    o.set_material_arrays(case["mats"][:3], case["mats"][3:])
    o._bind(o.fwd)
    o.lib.FDTD_clear_sources(o.h)
    for arrs, i0, j0, k0, sI, sJ, sK in o.sources:
        o.lib.FDTD_add_source(o.h, *arrs, i0, j0, k0, sI, sJ, sK, False)
    o.lib.FDTD_reset_pml(o.h)
    for n in range(nsteps):
        o._step(n)
    return [o.padded(q)[1:-1, 1:-1, 1:-1] for q in range(6)]


def test_src_func_t_matches_oracle():
    """FDTD.src_func_t (emf_src_func_t) against the oracle's FDTD_src_func_t
    with the same ramp, through and past the ramp, at several phases;
    test_src_func returns the same samples."""
    import warnings
    from emopt_amd.fdtd import FDTD
    from oracle import driver as D
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        sim = FDTD(1.0, 0.8, 0.6, 0.05, 0.05, 0.05, 1.31)
        o = D.OracleFDTD(D.load(D.ORACLE_LIB), 1.0, 0.8, 0.6, 0.05, 0.05, 0.05, 1.31)
    for T in (None, 7.5):
        if T is not None:
            sim.src_ramp_time = T
            o.src_ramp_time = T
        for n in (0, 1, 37, 400, 2999):
            for ph in (0.0, 0.7, -2.1):
                t = n * sim.dt
                a, b = sim.src_func_t(n, t, ph), o.lib.FDTD_src_func_t(o.h, n, t, ph)
                assert abs(a - b) <= 1e-15 * max(1.0, abs(b)), (T, n, ph, a, b)
    import importlib.util
    if importlib.util.find_spec("matplotlib") is None:   # otherwise it would plot
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            ramp = sim.test_src_func()
        assert ramp.shape == (3000,) and ramp[2999] == sim.src_func_t(2999, 2999 * sim.dt, 0.0)


def test_two_sample_phasors_match_oracle():
    """emf_calc_complex_fields_2T (k_phasor2; the reference's
    FDTD::calc_complex_fields(t0, t1), src/fdtd.cpp:1346-1390) against the
    oracle on the same state: the t0 snapshot after 12 steps, the live fields
    after 25 steps."""
    import warnings
    from emopt_amd.fdtd import FDTD, FIELD_INDEX, ptr
    from emopt_amd.misc import DomainCoordinates
    from oracle import driver as D
    from tests.cases import _configure
    case = random_case(13, 11, 9, seed=21, nsrc=2)
    d, n_cap, n_end = case["d"], 12, 25
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        sim = FDTD(case["X"], case["Y"], case["Z"], d, d, d, case["wl"])
        _configure(sim, case, DomainCoordinates)
        o = D.OracleFDTD(D.load(D.ORACLE_LIB), case["X"], case["Y"], case["Z"], d, d, d,
                         case["wl"])
        _configure(o, case, D.Domain)
    sim.set_material_arrays(case["mats"][:3], case["mats"][3:])
    sim._register(sim._sources)
    L, h, dt = sim._lib, sim._h, sim.dt
    L.emf_reset_pml(h)
    L.emf_zero_fields(h)
    o.set_material_arrays(case["mats"][:3], case["mats"][3:])
    o._bind(o.fwd)
    o.lib.FDTD_clear_sources(o.h)
    for arrs, i0, j0, k0, sI, sJ, sK in o.sources:
        o.lib.FDTD_add_source(o.h, *arrs, i0, j0, k0, sI, sJ, sK, False)
    o.lib.FDTD_reset_pml(o.h)
    for n in range(n_end):
        if n == n_cap:
            L.emf_capture_t0_fields(h, 0)
            o.lib.FDTD_capture_t0_fields(o.h)
        L.emf_update_H(h, n, n * dt)
        L.emf_update_E(h, n, (n + 0.5) * dt)
        o._step(n)
    L.emf_calc_complex_fields_2T(h, 0, n_cap * dt, n_end * dt)
    o.lib.FDTD_calc_complex_fields_2T(o.h, n_cap * dt, n_end * dt)
    ref = o.phasors()
    errs = []
    for q, c in enumerate(("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")):
        g = np.zeros((case["Nz"], case["Ny"], case["Nx"]), dtype=np.complex128)
        L.emf_get_field(h, 0, FIELD_INDEX[c], 0, case["Nx"], 0, case["Ny"], 0, case["Nz"], ptr(g))
        errs.append(rel_l2(g, ref[q]))
    print("2T rel L2 Ex..Hz:", " ".join("%.2e" % e for e in errs))
    assert max(errs) < 1e-9, errs


def test_set_fields_captures_and_three_sample_phasors_match_oracle():
    """The step-level calls EMopt's own loop makes through its ctypes layer
    (emopt/fdtd.py:1049-1194), each against the oracle making the same calls:
    random time-domain fields set with emf_set_time_fields, 30 steps with the
    t0 and t1 snapshots (emf_capture_t0_fields / emf_capture_t1_fields) at
    steps 12 and 21, the live fields at step 30, and the 3T phasors of
    emf_calc_complex_fields_3T (src/fdtd.cpp:1346-1443); also the library's
    version string and device count."""
    import warnings
    from emopt_amd.fdtd import FDTD, FIELD_INDEX, ptr
    from emopt_amd.misc import DomainCoordinates
    from oracle import driver as D
    from tests.cases import _configure
    case = random_case(14, 12, 10, seed=23, nsrc=2, bc="0H0")
    d, n1, n2, n3 = case["d"], 12, 21, 30
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        sim = FDTD(case["X"], case["Y"], case["Z"], d, d, d, case["wl"])
        _configure(sim, case, DomainCoordinates)
        o = D.OracleFDTD(D.load(D.ORACLE_LIB), case["X"], case["Y"], case["Z"], d, d, d,
                         case["wl"])
        _configure(o, case, D.Domain)
    L, h, dt = sim._lib, sim._h, sim.dt
    assert L.emf_device_count() >= 1 and L.emf_version()
    sim.set_material_arrays(case["mats"][:3], case["mats"][3:])
    sim._register(sim._sources)
    L.emf_reset_pml(h)
    shp = (case["Nz"], case["Ny"], case["Nx"])
    rng = np.random.default_rng(5)
    init = [np.ascontiguousarray(rng.normal(size=shp)) for _ in range(6)]
    L.emf_set_time_fields(h, *[ptr(a) for a in init])
    o.set_material_arrays(case["mats"][:3], case["mats"][3:])
    o._bind(o.fwd)
    for q in range(6):
        o.padded(q)[1:-1, 1:-1, 1:-1] = init[q]
    o.lib.FDTD_clear_sources(o.h)
    for arrs, i0, j0, k0, sI, sJ, sK in o.sources:
        o.lib.FDTD_add_source(o.h, *arrs, i0, j0, k0, sI, sJ, sK, False)
    o.lib.FDTD_reset_pml(o.h)
    assert all(np.array_equal(a, b) for a, b in zip(sim.time_fields(), init))
    for n in range(n3):
        if n == n1:
            L.emf_capture_t0_fields(h, 0)
            o.lib.FDTD_capture_t0_fields(o.h)
        if n == n2:
            L.emf_capture_t1_fields(h, 0)
            o.lib.FDTD_capture_t1_fields(o.h)
        L.emf_update_H(h, n, n * dt)
        L.emf_update_E(h, n, (n + 0.5) * dt)
        o._step(n)
    live = [o.padded(q)[1:-1, 1:-1, 1:-1] for q in range(6)]
    errs = rel_l2_set(sim.time_fields(), live)
    assert max(errs) < STEP_TOL, errs
    L.emf_calc_complex_fields_3T(h, 0, n1 * dt, n2 * dt, n3 * dt)
    o.lib.FDTD_calc_complex_fields_3T(o.h, n1 * dt, n2 * dt, n3 * dt)
    ref = o.phasors()
    errs = []
    for q, c in enumerate(("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")):
        g = np.zeros(shp, dtype=np.complex128)
        L.emf_get_field(h, 0, FIELD_INDEX[c], 0, case["Nx"], 0, case["Ny"], 0, case["Nz"], ptr(g))
        errs.append(rel_l2(g, ref[q]))
    print("3T rel L2 Ex..Hz:", " ".join("%.2e" % e for e in errs))
    assert max(errs) < 1e-9, errs


@pytest.mark.parametrize("bc", ["000", "0H0", "E00", "HE0", "EEH"])
@pytest.mark.parametrize("ceps", [False, True])
def test_step_level(bc, ceps):
    case = random_case(13, 11, 9, seed=sum(map(ord, bc)) + int(ceps), ceps=ceps, bc=bc, nsrc=2)
    g = _step_gpu(case, 20)
    o = _step_oracle(case, 20)
    errs = rel_l2_set(g, o)
    assert max(errs) < STEP_TOL, (bc, ceps, errs)


_ALL_BCS = [a + b + c for a in "0EH" for b in "0EH" for c in "0EH"]


@pytest.mark.parametrize("bc", _ALL_BCS)
def test_step_level_every_boundary_combination(bc):
    """Every combination of PEC / E-symmetry / H-symmetry on the three min
    faces (src/fdtd.cpp:398-501 ghost rules; the max faces are PEC), each
    with its own CPML widths (zero on some faces), random materials with
    non-uniform mu, complex eps on every other case, and two source blocks,
    20 steps against the oracle."""
    q = _ALL_BCS.index(bc)
    rng = np.random.default_rng(100 + q)
    w = tuple(int(v) for v in rng.integers(0, 4, size=6))
    case = random_case(12, 11, 10, w=w, seed=200 + q, ceps=bool(q % 2), bc=bc, nsrc=2)
    g = _step_gpu(case, 20)
    o = _step_oracle(case, 20)
    errs = rel_l2_set(g, o)
    assert max(errs) < STEP_TOL, (bc, w, errs)


def test_step_level_many_sources_and_no_pml():
    # more source blocks than the kernel carries inline, and zero CPML widths
    case = random_case(12, 10, 8, w=(0, 0, 0, 0, 0, 0), seed=7, nsrc=6)
    g = _step_gpu(case, 15)
    o = _step_oracle(case, 15)
    errs = rel_l2_set(g, o)
    assert max(errs) < STEP_TOL, errs


def test_step_level_more_than_64_sources():
    """A wave tests the first 64 source blocks in one ballot and every later
    block one by one: 70 overlapping blocks (applied in order) against the
    oracle, and a grid wider than one wave."""
    case = random_case(150, 11, 9, w=(3, 4, 2, 3, 2, 3), seed=9, nsrc=70)
    g = _step_gpu(case, 12)
    o = _step_oracle(case, 12)
    errs = rel_l2_set(g, o)
    assert max(errs) < STEP_TOL, errs


def test_step_level_wide_pml_uniform_mu():
    case = random_case(40, 24, 20, w=(6, 5, 4, 7, 5, 6), seed=3, mu_rand=False, nsrc=1)
    g = _step_gpu(case, 40)
    o = _step_oracle(case, 40)
    errs = rel_l2_set(g, o)
    assert max(errs) < STEP_TOL, errs


def test_strip_solve_converged_matches_oracle():
    case = strip_case(24, 20, 16, w=4)
    g = run_gpu(case)
    o = run_oracle(case)
    assert g["status"] == "converged"
    assert g["n0"] == o["n0"], (g["n0"], o["n0"])
    np.testing.assert_allclose(g["history"][:, 0], o["history"][:, 0])
    np.testing.assert_allclose(g["history"][:, 1:], o["history"][:, 1:], rtol=1e-8, atol=1e-300)
    errs = [rel_l2(a, b) for a, b in zip(g["fields"], o["fields"])]
    assert max(errs) <= TOL, errs
    print("rel L2 Ex..Hz:", " ".join("%.2e" % e for e in errs))


def test_solve_with_200_convergence_points_matches_oracle():
    """nconv=200 (the reference's 3D MMI example, emopt/examples/
    MMI_splitter_3D/mmi_1x2_splitter_3D_fdtd.py) samples about 200 cells,
    more than one 64-lane wave holds: each lane takes several points.  The
    same n0, the same checks and the same phasors as the oracle's."""
    case = strip_case(24, 20, 16, w=4)
    case["nconv"] = 200
    g = run_gpu(case)
    o = run_oracle(case)
    assert len(o["sim"].conv_pts) > 128, len(o["sim"].conv_pts)
    assert g["status"] == "converged"
    assert g["n0"] == o["n0"], (g["n0"], o["n0"])
    np.testing.assert_allclose(g["history"][:, 0], o["history"][:, 0])
    np.testing.assert_allclose(g["history"][:, 1:], o["history"][:, 1:], rtol=1e-8, atol=1e-300)
    errs = [rel_l2(a, b) for a, b in zip(g["fields"], o["fields"])]
    assert max(errs) <= TOL, errs


def test_nan_fields_stop_at_nmax_with_status_nan():
    """A NaN in a source current makes the fields NaN; the reference's loop
    runs on past NaN/Inf (emopt/fdtd.py:1050-1052,1137-1138) and stops at
    Nmax.  The device loop stops at the same step as the oracle's, after the
    same number of checks, and reports status 'nan' (the oracle: 'nmax')."""
    import warnings
    from tests.cases import build_sim
    case = strip_case(24, 20, 16, w=4)
    case["sources"][0][0][2][0, 3, 0] = complex(np.nan, 0.0)
    sim = build_sim(case)
    nmax = 6.5 * sim.Tn
    sim.Nmax = nmax
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        sim._register(sim._sources)
        sim._solve(0)
    o = run_oracle(case, nmax=nmax)
    g = sim.last_solve
    print("device: n0 %d status %s checks %d; oracle: n0 %d status %s checks %d"
          % (g["n0"], g["status"], g["nchecks"], o["n0"], o["status"], len(o["history"])))
    assert g["status"] == "nan" and o["status"] == "nmax"
    assert g["n0"] == o["n0"] and g["nchecks"] == len(o["history"])
    assert any("not finite" in str(x.message) for x in w)
    full = sim.get_field("Ez")
    assert np.isnan(full).any()


def test_strip_solve_fixed_n0_bc_symmetry():
    case = strip_case(30, 22, 18, w=4, bc="0H0")
    g = run_gpu(case, fixed_n0=600)
    o = run_oracle(case, fixed_n0=600)
    errs = [rel_l2(a, b) for a, b in zip(g["fields"], o["fields"])]
    assert max(errs) <= TOL, errs


def _golden(name):
    import os
    p = os.path.join(os.path.dirname(__file__), "golden", name)
    return np.load(p, allow_pickle=False)


def test_gpu_matches_reference_golden_steps():
    """Time-domain fields against EMopt's own kernel (fixtures written by the
    compiled src/fdtd.cpp)."""
    g = _golden("steps.npz")
    for key in sorted({k.rsplit("_", 1)[0] for k in g.files if k.endswith("_meta")}):
        meta = g[key + "_meta"]
        Nx, Ny, Nz, seed, ceps, nsrc, w0, w1, w2, w3, w4, w5 = [int(v) for v in meta[:12]]
        bc = "".join(chr(int(c)) for c in meta[12:15])
        case = random_case(Nx, Ny, Nz, w=(w0, w1, w2, w3, w4, w5), seed=seed, ceps=bool(ceps),
                           bc=bc, nsrc=nsrc)
        for n_target, fld in zip(g[key + "_steps"], g[key + "_fields"]):
            errs = rel_l2_set(_step_gpu(case, int(n_target)), fld)
            assert max(errs) < STEP_TOL, (key, int(n_target), errs)


@pytest.mark.parametrize("dims,w", [((24, 20, 16), 4), ((40, 32, 24), 5)])
def test_gpu_matches_reference_golden_solve(dims, w):
    g = _golden("solve_strip_%dx%dx%d.npz" % dims)
    r = run_gpu(strip_case(*dims, w=w))
    assert r["n0"] == int(g["n0"]) and r["Tn"] == int(g["Tn"])
    np.testing.assert_allclose(r["history"], g["history"], rtol=1e-8)
    errs = [rel_l2(a, b) for a, b in zip(r["fields"], g["fields"])]
    print("vs reference: rel L2 Ex..Hz:", " ".join("%.2e" % e for e in errs))
    assert max(errs) <= TOL, errs


def _run_steps(case, nsteps, kind, tyg=4, zc=8, order=0, compress=1, stream=-1, wide=0):
    """nsteps time steps through emf_time_steps with the given step kernel."""
    import ctypes
    import warnings
    from emopt_amd.fdtd import FDTD
    from emopt_amd.misc import DomainCoordinates
    from tests.cases import _configure
    d = case["d"]
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        sim = FDTD(case["X"], case["Y"], case["Z"], d, d, d, case["wl"],
                   complex_eps=case.get("ceps", False))
        _configure(sim, case, DomainCoordinates)
    sim.set_material_arrays(case["mats"][:3], case["mats"][3:])
    sim._register(sim._sources)
    L, h = sim._lib, sim._h
    L.emf_set_tuning(h, kind, tyg, zc, order)
    L.emf_set_compression(h, compress)
    L.emf_set_streaming(h, stream)
    L.emf_set_wide_offsets(h, wide)
    L.emf_reset_pml(h)
    L.emf_zero_fields(h)
    out = np.zeros(3)
    L.emf_time_steps(h, 0, nsteps, 0, 1, out.ctypes.data_as(ctypes.c_void_p))
    return sim.time_fields()


@pytest.mark.parametrize("variant", ["plain", "ceps", "mu", "bcs", "manysrc"])
def test_step_kernels_agree(variant):
    """Every step kernel and launch shape against the default two-pass one:
    the fused H+E kernels (ping-pong buffers; kind 0 recomputes the halo row,
    kind 2 shares it through LDS) for every work-group shape and z-chunk, the
    two passes for other row counts, z-chunks and work-group orders, and the
    H pass with its y-neighbour row through LDS (kind 4)."""
    kw = dict(w=(5, 4, 3, 6, 4, 5), seed=31, nsrc=2, mu_rand=False)
    if variant == "ceps":
        kw["ceps"] = True
    if variant == "mu":
        kw["mu_rand"] = True
    if variant == "bcs":
        kw["bc"] = "EH0"
    if variant == "manysrc":
        kw["nsrc"] = 7
    case = random_case(70, 37, 23, **kw)
    ref = _run_steps(case, 33, kind=1)
    for kind, tygs in ((0, (1, 2, 4, 8)), (2, (4, 8, 16))):
        for tyg in tygs:
            for zc in (1, 5, 8, 64):
                got = _run_steps(case, 33, kind=kind, tyg=tyg, zc=zc)
                errs = rel_l2_set(got, ref)
                assert max(errs) < 1e-13, (variant, kind, tyg, zc, errs)
    for kind in (1, 3):
        for tyg in (1, 2, 4, 8):
            for zc in (1, 5):
                for order in (0, 1, 2, 3):
                    if kind == 1 and tyg == 4 and zc == 8 and order == 0:
                        continue
                    got = _run_steps(case, 33, kind=kind, tyg=tyg, zc=zc, order=order)
                    errs = rel_l2_set(got, ref)
                    assert max(errs) < 1e-13, (variant, kind, tyg, zc, order, errs)
    # kind 4 (the H pass's y-neighbour row through LDS) reads the values kind 1
    # loads: identical fields for the same shape, partial rows and lanes
    # included (37 rows, 70 columns)
    for tyg in (2, 4, 8):
        for zc in (1, 5):
            for order in (0, 3):
                a = _run_steps(case, 33, kind=1, tyg=tyg, zc=zc, order=order)
                b = _run_steps(case, 33, kind=4, tyg=tyg, zc=zc, order=order)
                assert all(np.array_equal(x, y) for x, y in zip(a, b)), (variant, tyg, zc, order)


@pytest.mark.parametrize("tuning", [(0, 2, 8, 0), (2, 8, 4, 0), (1, 8, 8, 2)])
def test_tuned_solve_matches_default_solve(tuning):
    """A whole solve (ping-pong set parity across the capture points and the
    gated steps after convergence) with another step kernel or launch shape."""
    case = strip_case(40, 28, 22, w=5)
    import warnings
    a = run_gpu(case, fixed_n0=700)
    sim = a["sim"]
    sim._lib.emf_set_tuning(sim._h, *tuning)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        sim._register(sim._sources)
        sim.fixed_n0 = 700
        sim._solve(0)
    from emopt_amd.misc import DomainCoordinates
    full = DomainCoordinates.from_indices(0, case["Nx"], 0, case["Ny"], 0, case["Nz"], *([case["d"]] * 3))
    b = [sim.get_field(c, full) for c in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")]
    errs = [rel_l2(x, y) for x, y in zip(a["fields"], b)]
    assert max(errs) < 1e-12, (tuning, errs)


@pytest.mark.parametrize("which", ["strip", "random", "random_ceps"])
def test_coefficient_compression_is_exact(which):
    """The E pass reading a per-cell index into the table of distinct eps
    triples gives exactly the fields of the full coefficient arrays: few
    distinct values (strip, 2-byte index), every cell distinct (random,
    4-byte index) and complex eps (compression not applicable); for the two
    passes and for the fused LDS step."""
    if which == "strip":
        case = strip_case(40, 28, 22, w=5)
    else:
        case = random_case(70, 37, 23, w=(5, 4, 3, 6, 4, 5), seed=31, nsrc=2, mu_rand=False,
                           ceps=(which == "random_ceps"))
    for kind, tyg, zc in ((1, 4, 2), (3, 4, 2), (2, 8, 5)):
        a = _run_steps(case, 41, kind=kind, tyg=tyg, zc=zc, compress=1)
        b = _run_steps(case, 41, kind=kind, tyg=tyg, zc=zc, compress=0)
        for x, y in zip(a, b):
            assert np.array_equal(x, y), (which, kind)


def test_psi_streaming_is_exact():
    """Streaming (non-temporal) psi access is a cache policy only: the fields
    after 41 steps are bit-identical with it forced on and off."""
    case = random_case(70, 37, 23, w=(5, 4, 3, 6, 4, 5), seed=31, nsrc=2, mu_rand=True)
    for tyg, zc in ((2, 2), (4, 5)):
        a = _run_steps(case, 41, kind=1, tyg=tyg, zc=zc, order=3, stream=1)
        b = _run_steps(case, 41, kind=1, tyg=tyg, zc=zc, order=3, stream=0)
        for x, y in zip(a, b):
            assert np.array_equal(x, y), (tyg, zc)


def test_h_pass_shape_is_launch_shape_only():
    """The H pass's own work-group shape (emf_set_h_shape) changes the
    launch, not the arithmetic: the fields after 25 steps are identical to
    the default shape's, for several rows x planes per work-group."""
    case = random_case(70, 23, 19, w=(3, 4, 2, 3, 4, 3), seed=12, nsrc=2)
    ref = _step_gpu(case, 25)
    import warnings
    from emopt_amd.fdtd import FDTD
    from emopt_amd.misc import DomainCoordinates
    from tests.cases import _configure
    for tyg, zc in ((1, 1), (2, 3), (8, 4), (4, 7), (1, 19)):
        d = case["d"]
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            sim = FDTD(case["X"], case["Y"], case["Z"], d, d, d, case["wl"])
            _configure(sim, case, DomainCoordinates)
        sim.set_material_arrays(case["mats"][:3], case["mats"][3:])
        sim._register(sim._sources)
        L, h = sim._lib, sim._h
        L.emf_set_h_shape(h, tyg, zc)
        L.emf_reset_pml(h)
        L.emf_zero_fields(h)
        for n in range(25):
            L.emf_update_H(h, n, n * sim.dt)
            L.emf_update_E(h, n, (n + 0.5) * sim.dt)
        got = sim.time_fields()
        for a, b in zip(got, ref):
            assert np.array_equal(a, b), (tyg, zc)


@pytest.mark.parametrize("variant", ["index4", "ceps", "mu", "bcs", "index2"])
def test_wide_offsets_are_exact(variant):
    """The passes with 64-bit cell offsets (which grids above 178 M cells
    use) give exactly the fields of the 32-bit ones: every coefficient mode
    (4- and 2-byte compressed index, complex eps), non-uniform mu, symmetric
    boundaries, streamed and cached psi, unrolled (2 planes) and looped (5)
    E-pass columns."""
    if variant == "index2":
        case = strip_case(40, 28, 22, w=5)
    else:
        kw = dict(w=(5, 4, 3, 6, 4, 5), seed=37, nsrc=2, mu_rand=variant == "mu",
                  ceps=variant == "ceps", bc="EH0" if variant == "bcs" else "000")
        case = random_case(70, 37, 23, **kw)
    for zc in (2, 5):
        for stream in (0, 1):
            a = _run_steps(case, 29, kind=1, tyg=4, zc=zc, stream=stream, wide=1)
            b = _run_steps(case, 29, kind=1, tyg=4, zc=zc, stream=stream, wide=0)
            for x, y in zip(a, b):
                assert np.array_equal(x, y), (variant, zc, stream)
