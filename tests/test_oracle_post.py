"""Pinning the oracle's post-processing and solve-loop restatements (CPU only).

oracle/driver.py restates, besides the time loop, the reference's field
post-processing: get_field_interp (emopt/fdtd.py:1401-1553),
get_source_power (:1555-1667), get_A_diag (:1669-1692) and calc_ydAx
(:1694-1717).  EMopt's Python package cannot be imported here (petsc4py and
mpi4py are absent), so each is pinned by what the mathematics fixes:

* interpolation onto the Ez grid: a field linear in position interpolates
  exactly to its value at the Ez points; 'E'/'H' mirrors keep odd/even
  fields exact at the min faces; no mirror means a zero neighbour; and the
  reference's domain growth, which skips the max faces, returns a domain that
  reaches a max face one cell short along that axis (its last cell is cropped
  with the growth that did not happen) -- a symmetric growth would fail it;
* source power: the box flux of a field with S = 0.5 (u - c) along one axis
  is the face count times the spacing, the min face dropped under 'E'/'H';
* calc_ydAx: the reference's gradient formula -2 Re(y^T dA x)/step, with
  forward and adjoint fields from the oracle's own solves, against central
  differences of a point-detector FOM |Ez|^2 for a change of eps in one cell
  and of mu in one cell (a sign or a dropped i fails; the residue is the
  leapfrog's dispersion factor 1 - dt^2/24 ~ 1e-3);
* the solve loop: the recovered phasor equals a least-squares fit of
  A sin(t + phi) + B to the fields over every step of the capture window
  (in a lossy medium, where the fields have settled), with E
  after n steps at t = n dt and H at (n - 1/2) dt; a capture time off by one
  step or a missing H shift fails it.

The same interpolation and source-power forms run on the product's host
restatement (emopt_amd/fields.py) beside the oracle's.
"""
import numpy as np
import pytest

from emopt_amd import fields
from emopt_amd.misc import DomainCoordinates
from oracle import driver as D
from tests.cases import span, strip_case, run_oracle
from tests.test_host import YEE, FakeSim

NAMES = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")


def _yee_fields(N, d, funcs):
    Nx, Ny, Nz = N
    out = []
    for c in NAMES:
        sx, sy, sz = YEE[c]
        Z, Y, X = np.meshgrid((np.arange(Nz) + sz) * d, (np.arange(Ny) + sy) * d,
                              (np.arange(Nx) + sx) * d, indexing="ij")
        f = funcs.get(c)
        out.append(np.zeros(Z.shape, complex) if f is None else f(X, Y, Z).astype(complex))
    return out


def _impl(kind, N, d, funcs, bc="000", w_pml=(0.0,) * 6):
    """(interp(comp, k1, k2, j1, j2, i1, i2), source_power()) of the oracle
    or of the product's host restatement, on analytic fields."""
    if kind == "oracle":
        o = D.holder(*N, d, d, d, _yee_fields(N, d, funcs), bc=bc, w_pml=w_pml)
        return (lambda c, *b: o.get_field_interp(c, D.Domain.from_indices(*b)),
                o.get_source_power)
    sim = FakeSim(N, d, funcs, bc=bc, w_pml=w_pml)
    return (lambda c, *b: sim.get_field_interp(c, DomainCoordinates.from_indices(*b, d, d, d)),
            lambda: fields.source_power(sim))


def _ez_points(k1, k2, j1, j2, i1, i2, d):
    Z, Y, X = np.meshgrid(np.arange(i1, i2), np.arange(j1, j2), np.arange(k1, k2), indexing="ij")
    return X * d, Y * d, Z * d


LIN = lambda x, y, z: 1.0 + 2.0 * x - 3.0 * y + 5.0 * z + 0.5j * x


@pytest.mark.parametrize("kind", ["oracle", "host"])
@pytest.mark.parametrize("comp", ["Ex", "Ey", "Hx", "Hy", "Hz"])
def test_interp_linear_exact(kind, comp):
    interp, _ = _impl(kind, (20, 16, 12), 0.1, {comp: LIN})
    b = (3, 15, 2, 12, 2, 9)
    got = interp(comp, *b)
    assert got.shape == (7, 10, 12)
    np.testing.assert_allclose(got, LIN(*_ez_points(*b, 0.1)), rtol=0, atol=1e-12)


@pytest.mark.parametrize("kind", ["oracle", "host"])
@pytest.mark.parametrize("comp", ["Ex", "Ey", "Hx", "Hy", "Hz"])
def test_interp_domain_reaching_max_faces(kind, comp):
    """The domain grows by one cell at a min face (k1 > 0) and at a max face
    only while k2 < Nx - 1, and the grown box goes through
    DomainCoordinates, whose upper index is int(x/dx) + 1.  A domain with
    k2 = Nx is therefore not grown, its box ends at Nx + 1, and the crop
    -1 - (k2_box - k2) drops its last cell: the result is one cell short along
    every axis that reaches its max face, and exact on the cells it keeps.
    A domain ending one cell before the max face keeps its full shape."""
    N, d = (20, 16, 12), 0.1
    interp, _ = _impl(kind, N, d, {comp: LIN})
    got = interp(comp, 3, 20, 2, 16, 2, 12)
    assert got.shape == (12 - 2 - 1, 16 - 2 - 1, 20 - 3 - 1)
    np.testing.assert_allclose(got, LIN(*_ez_points(3, 19, 2, 15, 2, 11, d)), rtol=0, atol=1e-12)
    got = interp(comp, 3, 19, 2, 15, 2, 11)
    assert got.shape == (9, 13, 16)
    np.testing.assert_allclose(got, LIN(*_ez_points(3, 19, 2, 15, 2, 11, d)), rtol=0, atol=1e-12)
    # only the y axis reaches its max face
    got = interp(comp, 3, 15, 4, 16, 2, 9)
    assert got.shape == (7, 11, 12)
    np.testing.assert_allclose(got, LIN(*_ez_points(3, 15, 4, 15, 2, 9, d)), rtol=0, atol=1e-12)


_ODD_X = lambda x, y, z: 3.0 * x
_ODD_Y = lambda x, y, z: 3.0 * y
_EVEN_X = lambda x, y, z: 2.0 + y - z
_EVEN_Y = lambda x, y, z: 2.0 + x - z


@pytest.mark.parametrize("kind", ["oracle", "host"])
@pytest.mark.parametrize("comp,bc,f", [
    ("Ex", "E00", _ODD_X), ("Ex", "H00", _EVEN_X), ("Hy", "E00", _ODD_X), ("Hy", "H00", _EVEN_X),
    ("Ey", "0E0", _ODD_Y), ("Ey", "0H0", _EVEN_Y), ("Hx", "0E0", _ODD_Y), ("Hx", "0H0", _EVEN_Y),
    ("Hz", "EE0", lambda x, y, z: 3.0 * x * y), ("Hz", "HH0", lambda x, y, z: 2.0 - z),
    ("Hz", "EH0", _ODD_X), ("Hz", "HE0", _ODD_Y)])
def test_interp_min_face_mirrors(kind, comp, bc, f):
    interp, _ = _impl(kind, (14, 12, 10), 0.1, {comp: f}, bc=bc)
    got = interp(comp, 0, 6, 0, 6, 2, 7)
    np.testing.assert_allclose(got, f(*_ez_points(0, 6, 0, 6, 2, 7, 0.1)), rtol=0, atol=1e-12)


@pytest.mark.parametrize("kind", ["oracle", "host"])
def test_interp_zero_neighbour_without_mirror(kind):
    interp, _ = _impl(kind, (10, 10, 10), 0.1, {"Ex": lambda x, y, z: 1.0 + 0 * x})
    got = interp("Ex", 0, 4, 3, 6, 3, 6)
    np.testing.assert_allclose(got[:, :, 0], 0.5)
    np.testing.assert_allclose(got[:, :, 1:], 1.0)


@pytest.mark.parametrize("kind", ["oracle", "host"])
@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("bc", ["000", "EH0", "HE0"])
def test_source_power_closed_form(kind, axis, bc):
    d, N = 0.05, (24, 20, 18)
    w = (0.2, 0.2, 0.15, 0.15, 0.1, 0.1)
    c = 0.37
    fs = [{"Ey": lambda x, y, z: 1.0 + 0 * x, "Hz": lambda x, y, z: x - c},
          {"Ez": lambda x, y, z: 1.0 + 0 * x, "Hx": lambda x, y, z: y - c},
          {"Ex": lambda x, y, z: 1.0 + 0 * x, "Hy": lambda x, y, z: z - c}][axis]
    _, power = _impl(kind, N, d, fs, bc=bc, w_pml=w)
    X, Y, Z = [(n - 1) * d for n in N]
    lims = [w[0] + d, X - w[1] - d, w[2] + d, Y - w[3] - d, w[4] + d, Z - w[5] - d]
    k1, k2, j1, j2, i1, i2 = [int(v / d) for v in lims]
    lo, hi = [(k1, k2), (j1, j2), (i1, i2)][axis]
    n = [(j2 - j1 + 1) * (i2 - i1 + 1), (k2 - k1 + 1) * (i2 - i1 + 1),
         (k2 - k1 + 1) * (j2 - j1 + 1)][axis]
    area = n * d * d
    expect = 0.5 * area * (hi * d - c) + (-0.5 * area * (lo * d - c) if bc[axis] not in "EH" else 0)
    assert power() == pytest.approx(expect, rel=1e-12, abs=1e-14)


@pytest.mark.parametrize("kind", ["oracle", "host"])
def test_default_domain_drops_a_plane_where_the_span_rounds_down(kind):
    """get_field without a domain takes DomainCoordinates(0, X, 0, Y, 0, Z)
    (emopt/fdtd.py:1325-1327) with X = (N - 1) d: int(X / d) + 1 is N - 1
    where (N - 1) d / d rounds below N - 1 (29 * 0.04 / 0.04 = 28.999...),
    so that axis comes back one plane short; elsewhere it is whole."""
    d, N = 0.04, (36, 30, 24)
    assert int((N[1] - 1) * d / d) == N[1] - 2          # y rounds down
    assert int((N[0] - 1) * d / d) == N[0] - 1          # x does not
    lin = {"Ez": LIN}
    if kind == "oracle":
        o = D.holder(*N, d, d, d, _yee_fields(N, d, lin))
        got = o.get_field("Ez")
    else:
        got = FakeSim(N, d, lin).get_field("Ez")
    assert got.shape == (N[2], N[1] - 1, N[0])


def test_get_A_diag_returns_copies():
    rng = np.random.default_rng(0)
    mats = [rng.random((4, 5, 6)) + 1j * rng.random((4, 5, 6)) for _ in range(6)]
    o = D.holder(6, 5, 4, 0.1, 0.1, 0.1, None, mats=mats)
    A = o.get_A_diag()
    for a, m in zip(A, mats):
        assert np.array_equal(a, m.ravel())
    A[0][:] = 0
    assert np.array_equal(o.mats[0], mats[0].ravel())


# ------------------------------------------------ calc_ydAx against FD

def _point_problem():
    N, d, wl, w = (24, 22, 20), 0.05, 1.2, 5
    rng = np.random.default_rng(11)
    shp = (N[2], N[1], N[0])
    eps = [np.full(shp, 2.0, complex) + 0.3 * rng.random(shp) for _ in range(3)]
    mu = [np.full(shp, 1.0, complex) + 0.2 * rng.random(shp) for _ in range(3)]
    o = D.OracleFDTD(D.load(D.ORACLE_LIB), span(N[0], d), span(N[1], d), span(N[2], d),
                     d, d, d, wl, rtol=1e-6)
    o.w_pml = [w * d + 1e-9 * d] * 6
    src = [np.zeros((1, 1, 1), complex) for _ in range(6)]
    src[2][0, 0, 0] = 1.0
    o.set_sources(src, D.Domain.from_indices(10, 11, 11, 12, 10, 11))
    return o, eps, mu


def test_calc_ydAx_gradient_matches_central_differences():
    """F = |Ez(detector)|^2.  Adjoint source conj(Ez) at the detector; the
    reference's gradient -2 Re(calc_ydAx)/step for a change `step` of eps_z
    in one cell and of mu_x in another, against central differences of F
    (every solve runs the same fixed step count, four times the converged
    one)."""
    import warnings
    o, eps, mu = _point_problem()
    det = D.Domain.from_indices(14, 15, 12, 13, 11, 12)

    def solve(e, m, n0=None):
        o.set_material_arrays(e, m)
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            o.solve_forward(fixed_n0=n0)
        return o.get_field("Ez", det)[0, 0, 0]
    solve(eps, mu)
    n0 = 4 * o.n0 - (4 * o.n0) % o.Tn
    ez = solve(eps, mu, n0)
    a = [np.zeros((1, 1, 1), complex) for _ in range(6)]
    a[2][0, 0, 0] = np.conj(ez)
    o.adj_sources = []
    o.add_adjoint_sources(a, det)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        o.solve_adjoint(fixed_n0=n0)
    A0 = o.get_A_diag()
    h = 1e-4
    where = ((2, (11, 12, 12)), (3, (10, 11, 13)), (4, (10, 11, 13)))   # eps_z, mu_x, mu_y
    grads = []
    for q, (i, j, k) in where:        # all formula gradients on the unperturbed solves
        g = (i * o.J + j) * o.K + k
        o.mats[q][g] = A0[q][g] + h
        grads.append(-2.0 * np.real(o.calc_ydAx(A0)) / h)
        o.mats[q][g] = A0[q][g]
    res = []
    for (q, (i, j, k)), grad in zip(where, grads):
        v = []
        for sgn in (1, -1):
            e2 = [x.copy() for x in eps]
            m2 = [x.copy() for x in mu]
            (e2 + m2)[q][i, j, k] += sgn * h
            v.append(abs(solve(e2, m2, n0)) ** 2)
        fd = (v[0] - v[1]) / (2 * h)
        print("%s at %s: calc_ydAx gradient %.6e, central difference %.6e"
              % (("eps_x", "eps_y", "eps_z", "mu_x", "mu_y", "mu_z")[q], (k, j, i), grad, fd))
        res.append((grad, fd))
    for grad, fd in res:
        assert abs(grad - fd) <= 5e-3 * abs(fd), (grad, fd)
        assert abs(fd) > 0


# --------------------------------------- the solve loop against a sinusoid fit

def _lossy_case():
    """A source in a uniformly lossy medium (eps = 2 + 1i): transients decay
    by ~e^-pi per period, so after a few tens of periods the fields are
    sinusoids plus a constant to ~1e-6 of their amplitude (a slow drift of
    the static H remains)."""
    from tests.cases import random_case
    c = random_case(20, 18, 16, w=(4,) * 6, d=0.05, wl=1.0, seed=3, ceps=True, mu_rand=False,
                    nsrc=1, ramp=5 * 2 * np.pi)
    shp = (16, 18, 20)
    c["mats"] = [np.full(shp, 2.0 + 1.0j) for _ in range(3)] + [np.ones(shp, complex)] * 3
    return c


def _window_fit_error(case, n0, cells, e_shift=0.0, h_shift=-0.5):
    """Solve `case` on the oracle with the captures at n0, n0 + Tn, n0 + 2 Tn,
    recording the fields at `cells` after every step of that window; fit
    a sin t + b cos t + B to each component with E after n steps at
    t = (n + e_shift) dt and H at (n + h_shift) dt, and return the largest
    |P_fit - P| / |P| against the recovered phasor P (P = A cos(phi) -
    i A sin(phi) of f(t) = A sin(t + phi) + B, i.e. P = a - i b)."""
    import warnings
    from tests.cases import _configure
    d = case["d"]
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        o = D.OracleFDTD(D.load(D.ORACLE_LIB), case["X"], case["Y"], case["Z"], d, d, d,
                         case["wl"], rtol=case["rtol"], complex_eps=True)
        _configure(o, case, D.Domain)
    o.set_material_arrays(case["mats"][:3], case["mats"][3:])
    rec = []
    step = o._step

    def recording_step(n):
        step(n)
        if n + 1 >= n0:
            rec.append((n + 1, [[o.padded(q)[i + 1, j + 1, k + 1] for (i, j, k) in cells]
                                for q in range(6)]))
    o._step = recording_step
    o.solve_forward(fixed_n0=n0)
    ph = o.phasors()
    Tn, dt = o.Tn, o.dt
    steps = np.array([n for n, _ in rec], dtype=float)
    vals = np.array([v for _, v in rec])              # [sample, component, cell]
    assert steps[0] == n0 and steps[-1] == n0 + 2 * Tn
    worst = 0.0
    for q in range(6):
        t = (steps + (e_shift if q < 3 else h_shift)) * dt
        M = np.stack([np.sin(t), np.cos(t), np.ones_like(t)], axis=1)
        coef, *_ = np.linalg.lstsq(M, vals[:, q], rcond=None)
        for c, (i, j, k) in enumerate(cells):
            P = ph[q][i, j, k]
            worst = max(worst, abs(coef[0, c] - 1j * coef[1, c] - P) / abs(P))
    return worst


def test_solve_phasor_matches_sinusoid_fit():
    """The phasor the solve loop recovers from its three captures (n0, n0 +
    Tn, n0 + 2 Tn) against a least-squares sinusoid over every step of that
    window, in a lossy medium where the fields have settled to rounding.
    Planted label mistakes -- every capture one step late (E and H), and H
    without its half-step shift -- are far outside the tolerance."""
    case = _lossy_case()
    cells = [(8, 9, 10), (6, 7, 12), (10, 12, 15), (7, 9, 6)]
    n0 = 1200
    err = _window_fit_error(case, n0, cells)
    print("phasor vs least-squares sinusoid over the capture window: %.2e" % err)
    assert err < 2e-3
    for e_shift, h_shift in ((1.0, 0.5), (0.0, 0.0)):
        bad = _window_fit_error(case, n0, cells, e_shift=e_shift, h_shift=h_shift)
        print("labels E %+g, H %+g steps: %.2e" % (e_shift, h_shift, bad))
        assert bad > 2e-2
