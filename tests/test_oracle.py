"""Pinning the oracle (CPU only).

The oracle's C restatement (oracle/cfdtd.c) and driver (oracle/driver.py) are
checked against
  * the committed golden fixtures, generated from EMopt's own src/fdtd.cpp
    compiled from its sources (tests/golden/make_golden.py),
  * that compiled reference itself when it is present (oracle/_ref/),
  * closed forms: exact recovery of A*sin(t+phi)+B by the 3-sample phasor,
    the ramp envelope's end values,
  * invariants: bounded discrete energy of a source-free lossless PEC cavity,
  * a planted-mistake check showing that a dropped CPML term is detected.
"""
import os

import numpy as np
import pytest

from oracle import driver as D
from tests.cases import random_case, rel_l2, rel_l2_set, run_oracle, strip_case

GOLD = os.path.join(os.path.dirname(__file__), "golden")
HAVE_REF = os.path.exists(D.REF_LIB)


@pytest.fixture(scope="module")
def olib():
    return D.load(D.ORACLE_LIB)


# ------------------------------------------------------------- closed forms

def test_phasor_3t_recovers_sinusoid(olib):
    rng = np.random.default_rng(1)
    for _ in range(200):
        A = rng.uniform(0.1, 5); phi = rng.uniform(-np.pi, np.pi); B = rng.uniform(-1, 1)
        t0 = rng.uniform(0, 100); t1 = t0 + rng.uniform(0.3, 2.5); t2 = t1 + rng.uniform(0.3, 2.5)
        f = [A * np.sin(t + phi) + B for t in (t0, t1, t2)]
        ph = olib.FDTD_calc_phase_3T(t0, t1, t2, *f)
        amp = olib.FDTD_calc_amplitude_3T(t0, t1, t2, *f, ph)
        if amp < 0:
            amp, ph = -amp, ph + np.pi
        assert abs(amp - A) < 1e-9 * A
        assert abs(np.angle(np.exp(1j * (ph - phi)))) < 1e-9


def test_phasor_zero_signal(olib):
    assert olib.FDTD_calc_phase_3T(0.1, 0.2, 0.3, 2.0, 2.0, 2.0) == 0.0
    assert olib.FDTD_calc_amplitude_3T(0.1, 0.2, 0.3, 2.0, 2.0, 2.0, 0.0) == 0.0
    assert olib.FDTD_calc_phase_2T(0.1, 0.2, 0.0, 0.0) == 0.0


def test_phasor_2t_recovers_pure_sinusoid(olib):
    rng = np.random.default_rng(2)
    for _ in range(100):
        A = rng.uniform(0.1, 5); phi = rng.uniform(-np.pi / 2 + 0.1, np.pi / 2 - 0.1)
        t0 = rng.uniform(0, 10); t1 = t0 + rng.uniform(0.3, 1.2)
        f0, f1 = A * np.sin(t0 + phi), A * np.sin(t1 + phi)
        ph = olib.FDTD_calc_phase_2T(t0, t1, f0, f1)
        amp = olib.FDTD_calc_amplitude_2T(t0, t1, f0, f1, ph)
        assert abs(abs(amp) - A) < 1e-8 * A


def test_ramp_envelope_end_values(olib):
    h = olib.FDTD_new()
    T, m = 30.0, 1e-4
    olib.FDTD_set_source_properties(h, T, m)
    # envelope(0) = 0 and envelope(T) = 1 analytically: sin(t + phase) factor
    # removed with phase = pi/2 - t
    assert abs(olib.FDTD_src_func_t(h, 0, 0.0, np.pi / 2)) < 1e-15
    assert abs(olib.FDTD_src_func_t(h, 0, T, np.pi / 2 - T) - 1.0) < 1e-14
    assert olib.FDTD_src_func_t(h, 0, T + 1.0, 0.3) == np.sin(T + 1.3)


def run_small_steps(lib, case, nsteps):
    from tests.cases import _configure
    import warnings
    d = case["d"]
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        o = D.OracleFDTD(lib, case["X"], case["Y"], case["Z"], d, d, d, case["wl"],
                         complex_eps=case.get("ceps", False))
        _configure(o, case, D.Domain)
    o.set_material_arrays(case["mats"][:3], case["mats"][3:])
    o._bind(o.fwd)
    o.lib.FDTD_clear_sources(o.h)
    for arrs, i0, j0, k0, sI, sJ, sK in o.sources:
        o.lib.FDTD_add_source(o.h, *arrs, i0, j0, k0, sI, sJ, sK, False)
    o.lib.FDTD_reset_pml(o.h)
    for n in range(nsteps):
        o._step(n)
    return o


# --------------------------------------------------------------- invariants

def _yee_energy(o):
    I, J, K = o.I, o.J, o.K
    E = [o.padded(q)[1:-1, 1:-1, 1:-1] for q in range(3)]
    H = [o.padded(q)[1:-1, 1:-1, 1:-1] for q in range(3, 6)]
    eps = [o.mats[q].real.reshape(I, J, K) for q in range(3)]
    return sum(np.sum(e * x * x) for e, x in zip(eps, E)) + sum(np.sum(h * h) for h in H)


def test_pec_cavity_energy_is_bounded(olib):
    """No sources, no CPML, PEC walls: the leapfrog scheme conserves a
    discrete energy, so the time-averaged E/H energy neither grows nor
    decays beyond O(dt) oscillation."""
    case = random_case(10, 9, 8, w=(0,) * 6, seed=11, nsrc=0, mu_rand=False)
    o = run_small_steps(olib, case, 0)
    rng = np.random.default_rng(3)
    for q in range(3):
        o.padded(q)[1:-1, 1:-1, 1:-1] = rng.standard_normal((o.I, o.J, o.K))
    e = []
    for n in range(400):
        o._step(n)
        e.append(_yee_energy(o))
    e = np.array(e)
    assert np.all(np.isfinite(e))
    a, b = e[:100].mean(), e[-100:].mean()
    assert abs(a - b) / a < 0.05


# ------------------------------------------------ against the compiled ref

@pytest.mark.skipif(not HAVE_REF, reason="reference sources not compiled here")
@pytest.mark.parametrize("bc", ["000", "0H0", "E00", "HE0"])
@pytest.mark.parametrize("ceps", [False, True])
def test_oracle_matches_compiled_reference_steps(olib, bc, ceps):
    case = random_case(12, 10, 8, seed=3, ceps=ceps, bc=bc, nsrc=2)
    a = run_small_steps(olib, case, 30)
    b = run_small_steps(D.load(D.REF_LIB), case, 30)
    errs = rel_l2_set(a.fields, b.fields)
    assert max(errs) < 1e-13, errs


@pytest.mark.skipif(not os.path.exists(D.ORACLE_OMP_LIB), reason="OpenMP oracle not built")
def test_threaded_oracle_is_bit_identical(olib):
    """The -fopenmp build (the all-cores CPU baseline) splits only the cell
    loops over threads, so every field must match the serial build exactly."""
    case = random_case(16, 12, 10, seed=5, ceps=True, bc="HE0", nsrc=2)
    a = run_small_steps(olib, case, 25)
    b = run_small_steps(D.load(D.ORACLE_OMP_LIB), case, 25)
    for x, y in zip(a.fields, b.fields):
        assert np.array_equal(x, y)


@pytest.mark.skipif(not HAVE_REF, reason="reference sources not compiled here")
def test_oracle_matches_compiled_reference_solve():
    case = strip_case(24, 20, 16, w=4)
    a = run_oracle(case, lib="oracle")
    b = run_oracle(case, lib="ref")
    assert a["n0"] == b["n0"]
    assert max(rel_l2(x, y) for x, y in zip(a["fields"], b["fields"])) < 1e-12


# ----------------------------------------------------------- golden fixtures

def _golden(name):
    p = os.path.join(GOLD, name)
    if not os.path.exists(p):
        pytest.skip("fixture %s not generated" % name)
    return np.load(p, allow_pickle=False)


def test_oracle_matches_golden_steps(olib):
    g = _golden("steps.npz")
    for key in sorted({k.rsplit("_", 1)[0] for k in g.files if k.endswith("_meta")}):
        meta = g[key + "_meta"]
        Nx, Ny, Nz, seed, ceps, nsrc, w0, w1, w2, w3, w4, w5 = [int(v) for v in meta[:12]]
        bc = "".join(chr(int(c)) for c in meta[12:15])
        case = random_case(Nx, Ny, Nz, w=(w0, w1, w2, w3, w4, w5), seed=seed, ceps=bool(ceps),
                           bc=bc, nsrc=nsrc)
        nmax = int(g[key + "_steps"].max())
        o = run_small_steps(olib, case, 0)
        done = 0
        for n_target, fld in zip(g[key + "_steps"], g[key + "_fields"]):
            while done < n_target:
                o._step(done)
                done += 1
            got = [o.padded(q)[1:-1, 1:-1, 1:-1] for q in range(6)]
            errs = rel_l2_set(got, fld)
            assert max(errs) < 1e-13, (key, n_target, errs)
        assert done == nmax


@pytest.mark.parametrize("dims,w", [((24, 20, 16), 4), ((40, 32, 24), 5)])
def test_oracle_matches_golden_solve(dims, w):
    g = _golden("solve_strip_%dx%dx%d.npz" % dims)
    case = strip_case(*dims, w=w)
    a = run_oracle(case, lib="oracle")
    assert a["n0"] == int(g["n0"])
    np.testing.assert_allclose(a["history"], g["history"], rtol=1e-9)
    for q in range(6):
        assert rel_l2(a["fields"][q], g["fields"][q]) < 1e-12


def test_oracle_matches_golden_kats(olib):
    g = _golden("kats.npz")
    for row, want in zip(g["phase3_in"], g["phase3_out"]):
        assert olib.FDTD_calc_phase_3T(*row) == pytest.approx(want, rel=1e-14, abs=1e-15)
    for row, want in zip(g["amp3_in"], g["amp3_out"]):
        assert olib.FDTD_calc_amplitude_3T(*row) == pytest.approx(want, rel=1e-14, abs=1e-15)
    h = olib.FDTD_new()
    olib.FDTD_set_source_properties(h, float(g["src_T"]), float(g["src_min"]))
    for row, want in zip(g["src_in"], g["src_out"]):
        assert olib.FDTD_src_func_t(h, 0, row[0], row[1]) == pytest.approx(want, rel=1e-14, abs=1e-16)


# ------------------------------------------------------- planted mistake

def test_planted_mistake_is_detected(olib, tmp_path):
    """A copy of the oracle with the CPML 'd/kappa' term dropped from the H
    update must fail the step comparison against the intact oracle."""
    import shutil, subprocess
    src = open(os.path.join(os.path.dirname(D.__file__), "cfdtd.c")).read()
    bad = src.replace("hz = hz", "hz = hz").replace(
        "Hz[p] = Hz[p] - mz * (p0 - dEydx + dEydx / kap);",
        "Hz[p] = Hz[p] - mz * (p0 - dEydx);")
    assert bad != src
    cpath = tmp_path / "bad.c"; cpath.write_text(bad)
    so = tmp_path / "libbad.so"
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    subprocess.check_call(["gcc", "-O2", "-fPIC", "-shared", "-o", str(so), str(cpath), "-lm"])
    case = random_case(12, 10, 8, w=(3, 3, 3, 3, 3, 3), seed=4, nsrc=1)
    # make kappa != 1 so the dropped term matters
    a = run_small_steps(olib, case, 0); a.set_pml_properties(3.0, 0.0, 2.0, 3.0)
    b = run_small_steps(D.load(str(so)), case, 0); b.set_pml_properties(3.0, 0.0, 2.0, 3.0)
    for n in range(20):
        a._step(n); b._step(n)
    assert rel_l2(b.fields[5], a.fields[5]) > 1e-6
