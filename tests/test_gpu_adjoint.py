"""Adjoint gradient of a grating-coupler FOM on the GPU.

The differentiable solve (``fdtd_fom``: forward solve + FOM; backward =
adjoint solve + the device gradient assembly over the boundary cells each
trench edge moves) is checked against
  * EMopt's own gradient formula, restated in AdjointMethod.calc_gradient
    (emopt/adjoint_method.py:448-547): per parameter, step it by 1e-8 um,
    re-sample eps in its update box and take -2 Re(y^T dA x)/step with the
    same forward and adjoint fields.  The device assembly without the
    exact-adjoint factors must agree with it to 1e-6 (they differ only in how
    dEps/dp is formed);
  * central finite differences of the FOM with h = 1e-3 um on several edges
    -- agreement to 1 % per parameter for the exact gradient (the trenches
    cross the y CPML slabs, where the reference formula alone is off by up to
    a factor of a few: its adjoint is that of the unscaled CPML operator).
    Every solve here runs the same fixed number of steps, four times the step
    count at which the first forward solve converged: the convergence test
    accepts phase changes up to sqrt(rtol), and that tolerance alone would put
    more noise on a difference quotient than the check allows;
  * a point detector with eps perturbed in one cell inside and one outside
    the CPML: the exact gradient agrees with central differences to 1e-4 of
    the largest of these derivatives (the reference formula is off by orders
    of magnitude inside the slab, by the dispersion factor 1 - dt^2/24 outside).
The grid is a reduced grating (16 edges) that runs in seconds.
"""
import warnings

import numpy as np
import pytest

from tests.cases import grating_problem

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def problem():
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        sim, am, p = grating_problem(Nx=120, Ny=24, Nz=50, d=0.04, ntr=8, w=8, t_box=0.6,
                                     w0=2.0, rtol=1e-6)
        am.fom(p)
    n0 = sim.last_solve["n0"]
    sim.fixed_n0 = 4 * n0 - (4 * n0) % sim.Tn
    print("converged at n0 = %d; all solves run %d steps before the snapshots" % (n0, sim.fixed_n0))
    return sim, am, p


@pytest.fixture(scope="module")
def grads(problem):
    import torch
    from emopt_amd.adjoint import fdtd_fom
    sim, am, p = problem
    pt = torch.tensor(p, dtype=torch.float64, requires_grad=True)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        F = fdtd_fom(pt, am)
        F.backward()
    g_dev = pt.grad.numpy().copy()
    pidx, comp, cell, deps = am.calc_deps(sim, p)
    g_formula = sim.grad_assemble(pidx, comp, cell, deps, len(p))   # same fields
    g_ref = am.calc_gradient(sim, p)
    return float(F.detach()), g_dev, g_formula, g_ref


def test_device_assembly_matches_reference_formula(grads):
    F, g_dev, g_formula, g_ref = grads
    print("F = %.6e" % F)
    print("exact   :", np.array2string(g_dev, precision=4))
    print("formula :", np.array2string(g_formula, precision=4))
    print("ref loop:", np.array2string(g_ref, precision=4))
    assert np.linalg.norm(g_formula - g_ref) <= 1e-6 * np.linalg.norm(g_ref)
    assert np.linalg.norm(g_ref) > 0


def test_gradient_matches_central_differences(problem, grads):
    sim, am, p = problem
    F0, g_dev, _, _ = grads
    h = 1e-3
    idx = [0, 5, 8, 13]
    for i in idx:
        vals = []
        for s in (+1, -1):
            q = p.copy()
            q[i] += s * h
            with warnings.catch_warnings():
                warnings.simplefilter("ignore")
                vals.append(am.fom(q))
        fd = (vals[0] - vals[1]) / (2 * h)
        print("edge %2d: adjoint %.5e  fd %.5e  rel %.2e" % (i, g_dev[i], fd,
                                                             abs(g_dev[i] - fd) / abs(fd)))
        assert abs(g_dev[i] - fd) <= 0.01 * abs(fd) + 1e-3 * np.abs(g_dev).max(), (i, g_dev[i], fd)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        am.fom(p)   # leave the problem at p


def test_point_detector_gradient_inside_cpml():
    """F = |Ez(detector)|^2, parameter eps_z (and eps_x) of one cell; the
    exact adjoint gradient against central differences for a cell inside an
    x CPML slab and one in the interior."""
    from emopt_amd.fdtd import FDTD
    from emopt_amd.misc import DomainCoordinates as DC
    from tests.cases import span
    N, d, w = (36, 30, 28), 0.05, 6
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        sim = FDTD(span(N[0], d), span(N[1], d), span(N[2], d), d, d, d, 1.2, rtol=1e-6)
        sim.w_pml = [w * d + 1e-9 * d] * 6
    shp = (N[2], N[1], N[0])
    rng = np.random.default_rng(1)
    eps = [np.full(shp, 2.0, complex) + 0.3 * rng.random(shp) for _ in range(3)]
    mu = [np.ones(shp, complex) for _ in range(3)]
    src = [np.zeros((1, 1, 1), complex) for _ in range(6)]
    src[2][0, 0, 0] = 1.0
    sim.set_sources(src, DC.from_indices(12, 13, 14, 15, 13, 14, d, d, d))
    det = DC.from_indices(20, 21, 16, 17, 15, 16, d, d, d)

    def solve(e):
        sim.set_material_arrays(e, mu)
        sim._register(sim._sources)
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            sim._solve(0)
        return sim.get_field("Ez", det)[0, 0, 0]
    solve(eps)
    n0 = sim.last_solve["n0"]
    sim.fixed_n0 = 4 * n0 - (4 * n0) % sim.Tn
    ez = solve(eps)
    a = [np.zeros((1, 1, 1), complex) for _ in range(6)]
    a[2][0, 0, 0] = np.conj(ez)
    sim.clear_adjoint_sources()
    sim.add_adjoint_sources(a, det)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        sim.solve_adjoint()
    fac = 2 * np.sin(0.5 * sim.dt) / sim.dt
    res = []
    for (k, j, i) in ((2, 15, 14), (17, 15, 14)):          # inside the xmin slab; interior
        for c in (2, 0):
            g = (i * N[1] + j) * N[0] + k
            P = sim.pml_stretch([c], [g])[0]
            assert (P != 1) == (k < w)
            exact = sim.grad_assemble([0], [c], [g], [P * fac], 1)[0]
            formula = sim.grad_assemble([0], [c], [g], [1.0], 1)[0]
            v = []
            for sgn in (1, -1):
                e2 = [x.copy() for x in eps]
                e2[c][i, j, k] += sgn * 1e-4
                v.append(abs(solve(e2)) ** 2)
            fd = (v[0] - v[1]) / 2e-4
            print("cell %s eps_%s: exact %.6e formula %.6e fd %.6e" % ((k, j, i), "xyz"[c], exact,
                                                                     formula, fd))
            res.append((exact, fd))
    solve(eps)
    scale = max(abs(f) for _, f in res)
    for exact, fd in res:
        assert abs(exact - fd) <= 1e-4 * max(abs(fd), scale), (exact, fd)


def test_backward_after_another_forward_uses_its_own_params(problem, grads):
    """fdtd_fom at p, then at a second point, then backward of the first:
    the backward must not reuse the second point's forward solve (the
    reference's prev_params guard, emopt/adjoint_method.py:569-574)."""
    import torch
    from emopt_amd.adjoint import fdtd_fom
    sim, am, p = problem
    _, g_dev, _, _ = grads
    p1 = torch.tensor(p, dtype=torch.float64, requires_grad=True)
    q = p.copy()
    q[3] += 0.02
    p2 = torch.tensor(q, dtype=torch.float64, requires_grad=True)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        F1 = fdtd_fom(p1, am)
        F2 = fdtd_fom(p2, am)
        F1.backward()
        F2.backward()
        g2 = am.gradient_device(q)
        am.fom(p)
    print("F1 %.6e F2 %.6e" % (float(F1.detach()), float(F2.detach())))
    assert float(F1.detach()) != float(F2.detach())
    np.testing.assert_allclose(p1.grad.numpy(), g_dev, rtol=1e-12, atol=0)
    np.testing.assert_allclose(p2.grad.numpy(), g2, rtol=1e-12, atol=0)


def test_analytic_deps_match_device_differences(problem):
    """The grating's d eps / d edge from the overlap rule's geometry
    (emopt_amd/shape.py) against central differences of the solver's device
    sampling, entry by entry over all parameters, components and cells."""
    sim, am, p = problem

    def dense(deps):
        pidx, comp, cell, dv = deps
        out = {}
        for a, b, c, v in zip(pidx, comp, cell, dv):
            out[(int(a), int(b), int(c))] = out.get((int(a), int(b), int(c)), 0) + v
        return out
    am.deps = "difference"
    cd = dense(am.calc_deps(sim, p))
    am.deps = "analytic"
    an = dense(am.calc_deps(sim, p))
    scale = max(abs(v) for v in cd.values())
    keys = set(cd) | set(an)
    err = max(abs(an.get(k, 0) - cd.get(k, 0)) for k in keys)
    print("%d entries (analytic %d, differenced %d), max |deps| %.3e, max error %.3e"
          % (len(keys), len(an), len(cd), scale, err))
    assert err <= 1e-7 * scale
