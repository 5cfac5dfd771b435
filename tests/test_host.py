"""Host-side logic pinned by closed forms (CPU only).

* Field interpolation onto the Ez grid (emopt/fdtd.py:1401-1553): the Yee
  components of a field that is LINEAR in position must interpolate exactly to
  its value at the Ez points (a wrong stencil offset or sign breaks this); at
  a min face with an 'E'/'H' mirror, odd/even fields stay exact.
* Source power (emopt/fdtd.py:1555-1667): for E and H fields whose Poynting
  vector is S = 0.5 (u - u_c) along one axis, the box flux is the face count
  times the face spacing, with the min face dropped under an 'E'/'H' bc.
* Material cell averaging (emopt_amd/grid.py): exact rectangle fill fractions
  (area invariant, 1-D overlap products), z-slab fractions, layer priority,
  supersampled ellipse / polygon areas against pi*a*b and the shoelace area.

The reference's own versions of these routines cannot be imported here
(emopt needs petsc4py/mpi4py), so these closed forms are the pins.
"""
import numpy as np
import pytest

from emopt_amd import fields, grid
from emopt_amd.misc import DomainCoordinates

# Positions of the six components in cell units (x, y, z) relative to the Ez
# point (k, j, i): the Yee offsets EMopt samples materials at
# (emopt/fdtd.py:842-880).
YEE = {"Ex": (0.5, 0.0, -0.5), "Ey": (0.0, 0.5, -0.5), "Ez": (0.0, 0.0, 0.0),
       "Hx": (0.0, 0.5, 0.0), "Hy": (0.5, 0.0, 0.0), "Hz": (0.5, 0.5, -0.5)}


class FakeSim:
    """Just enough of the FDTD surface for fields.py: analytic complex
    fields sampled at each component's Yee position."""

    def __init__(self, N, d, funcs, bc="000", w_pml=(0.0,) * 6):
        self.Nx, self.Ny, self.Nz = N
        self.dx = self.dy = self.dz = d
        self.X, self.Y, self.Z = [(n - 1) * d for n in N]
        self.bc = bc
        self.w_pml = list(w_pml)
        self.funcs = funcs

    def get_field(self, comp, domain=None):
        if domain is None:
            domain = DomainCoordinates(0, self.X, 0, self.Y, 0, self.Z, self.dx, self.dy, self.dz)
        k = np.arange(max(0, domain.k1), min(self.Nx, domain.k2))
        j = np.arange(max(0, domain.j1), min(self.Ny, domain.j2))
        i = np.arange(max(0, domain.i1), min(self.Nz, domain.i2))
        sx, sy, sz = YEE[comp]
        Z, Y, X = np.meshgrid((i + sz) * self.dz, (j + sy) * self.dy, (k + sx) * self.dx,
                              indexing="ij")
        f = self.funcs.get(comp)
        return np.zeros(Z.shape, complex) if f is None else f(X, Y, Z).astype(complex)

    get_adjoint_field = get_field

    def get_field_interp(self, comp, domain=None, squeeze=False, adjoint=False):
        return fields.interp_to_ez(self, comp, domain, squeeze, adjoint)


def _ez_points(sim, dom):
    Z, Y, X = np.meshgrid(dom.i1 + np.arange(dom.Nz), dom.j1 + np.arange(dom.Ny),
                          dom.k1 + np.arange(dom.Nx), indexing="ij")
    return X * sim.dx, Y * sim.dy, Z * sim.dz


@pytest.mark.parametrize("comp", ["Ex", "Ey", "Hx", "Hy", "Hz"])
def test_interp_of_linear_field_is_exact(comp):
    lin = lambda x, y, z: 1.0 + 2.0 * x - 3.0 * y + 5.0 * z + 0.5j * x
    sim = FakeSim((20, 16, 12), 0.1, {comp: lin})
    dom = DomainCoordinates.from_indices(3, 15, 2, 12, 2, 9, 0.1, 0.1, 0.1)
    got = sim.get_field_interp(comp, dom)
    assert got.shape == dom.shape
    np.testing.assert_allclose(got, lin(*_ez_points(sim, dom)), rtol=0, atol=1e-12)


_ODD_X = lambda x, y, z: 3.0 * x
_ODD_Y = lambda x, y, z: 3.0 * y
_EVEN_X = lambda x, y, z: 2.0 + y - z      # constant along x
_EVEN_Y = lambda x, y, z: 2.0 + x - z      # constant along y
_MIRROR_CASES = [
    ("Ex", "E00", _ODD_X), ("Ex", "H00", _EVEN_X), ("Hy", "E00", _ODD_X), ("Hy", "H00", _EVEN_X),
    ("Ey", "0E0", _ODD_Y), ("Ey", "0H0", _EVEN_Y), ("Hx", "0E0", _ODD_Y), ("Hx", "0H0", _EVEN_Y),
    ("Hz", "EE0", lambda x, y, z: 3.0 * x * y), ("Hz", "HH0", lambda x, y, z: 2.0 - z),
    ("Hz", "EH0", _ODD_X), ("Hz", "HE0", _ODD_Y),
]


@pytest.mark.parametrize("comp,bc,f", _MIRROR_CASES)
def test_interp_mirror_at_min_face(comp, bc, f):
    """At k=0 (or j=0) with bc 'E' the ghost is -f(first), with 'H' +f(first):
    a field odd ('E') or even ('H') about the face plane, and linear (or
    bilinear) elsewhere, stays exact there."""
    sim = FakeSim((14, 12, 10), 0.1, {comp: f}, bc=bc)
    dom = DomainCoordinates.from_indices(0, 6, 0, 6, 2, 7, 0.1, 0.1, 0.1)
    got = sim.get_field_interp(comp, dom)
    np.testing.assert_allclose(got, f(*_ez_points(sim, dom)), rtol=0, atol=1e-12)


def test_interp_zero_pads_without_mirror():
    """With bc '0' the neighbour past the min face is 0 (not mirrored)."""
    f = lambda x, y, z: 1.0 + 0 * x
    sim = FakeSim((10, 10, 10), 0.1, {"Ex": f}, bc="000")
    dom = DomainCoordinates.from_indices(0, 4, 3, 6, 3, 6, 0.1, 0.1, 0.1)
    got = sim.get_field_interp("Ex", dom)
    np.testing.assert_allclose(got[:, :, 0], 0.5)
    np.testing.assert_allclose(got[:, :, 1:], 1.0)


def _box_faces(sim):
    """Ez-grid index of the six faces of the source-power box."""
    d = sim.dx
    w = sim.w_pml
    xmin = w[0] + d if w[0] > 0 else 0.0
    xmax = sim.X - w[1] - d if w[1] > 0 else sim.X - d
    ymin = w[2] + d if w[2] > 0 else 0.0
    ymax = sim.Y - w[3] - d if w[3] > 0 else sim.Y - d
    zmin = w[4] + d if w[4] > 0 else 0.0
    zmax = sim.Z - w[5] - d if w[5] > 0 else sim.Z - d
    return [int(v / d) for v in (xmin, xmax, ymin, ymax, zmin, zmax)]


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("bc", ["000", "EH0", "HE0"])
def test_source_power_closed_form(axis, bc):
    """S = 0.5 * (u - uc) along one axis (E constant 1, H linear): the flux out
    of the box is 0.5 * (face area sum) * (u2 - u1) for the two faces, each
    face counted with its own sign; under an 'E'/'H' min face only the max
    face contributes."""
    d, N = 0.05, (24, 20, 18)
    w = (0.2, 0.2, 0.15, 0.15, 0.1, 0.1)
    c = 0.37
    if axis == 0:     # Sx = 0.5 (Ey Hz* - Ez Hy*)
        fs = {"Ey": lambda x, y, z: 1.0 + 0 * x, "Hz": lambda x, y, z: x - c}
    elif axis == 1:   # Sy = 0.5 (Ez Hx* - Ex Hz*)
        fs = {"Ez": lambda x, y, z: 1.0 + 0 * x, "Hx": lambda x, y, z: y - c}
    else:             # Sz = 0.5 (Ex Hy* - Ey Hx*)
        fs = {"Ex": lambda x, y, z: 1.0 + 0 * x, "Hy": lambda x, y, z: z - c}
    sim = FakeSim(N, d, fs, bc=bc, w_pml=w)
    P = fields.source_power(sim)
    k1, k2, j1, j2, i1, i2 = _box_faces(sim)
    lo, hi = [(k1, k2), (j1, j2), (i1, i2)][axis]
    n = [(j2 - j1 + 1) * (i2 - i1 + 1), (k2 - k1 + 1) * (i2 - i1 + 1),
         (k2 - k1 + 1) * (j2 - j1 + 1)][axis]
    area = n * d * d
    out_hi = 0.5 * area * (hi * d - c)
    out_lo = -0.5 * area * (lo * d - c)
    min_face_on = bc[axis] not in "EH"
    expect = out_hi + (out_lo if min_face_on else 0.0)
    assert P == pytest.approx(expect, rel=1e-12, abs=1e-14)


def test_source_power_zero_fields():
    sim = FakeSim((16, 14, 12), 0.05, {}, w_pml=(0.1,) * 6)
    assert fields.source_power(sim) == 0.0


# -------------------------------------------------------------- materials

def _struct(prims, zr=(0.0, 10.0), d=0.1, N=(30, 24, 8)):
    m = grid.StructuredMaterial3D((N[0] - 1) * d, (N[1] - 1) * d, (N[2] - 1) * d, d, d, d,
                                  supersample=16)
    for p in prims:
        m.add_primitive(p, *zr)
    return m


def test_rectangle_fill_fractions_exact():
    d = 0.1
    r = grid.Rectangle(1.234, 0.987, 0.731, 0.419, material_value=5.0)
    m = _struct([r], d=d)
    v = m.get_values(0, 30, 0, 24, 2, 3)[0]
    # area invariant: sum of (value - 1) * cell area = (5 - 1) * rectangle area
    assert np.sum(v.real - 1.0) * d * d == pytest.approx(4.0 * 0.731 * 0.419, rel=1e-12)
    # per cell: product of the 1-D overlaps of the cell with the spans
    x0, x1 = 1.234 - 0.731 / 2, 1.234 + 0.731 / 2
    y0, y1 = 0.987 - 0.419 / 2, 0.987 + 0.419 / 2
    xc, yc = np.arange(30) * d, np.arange(24) * d
    ox = np.clip(np.minimum(xc + d / 2, x1) - np.maximum(xc - d / 2, x0), 0, None) / d
    oy = np.clip(np.minimum(yc + d / 2, y1) - np.maximum(yc - d / 2, y0), 0, None) / d
    np.testing.assert_allclose(v.real, 1.0 + 4.0 * oy[:, None] * ox[None, :], atol=1e-12)


def test_z_slab_fraction():
    d = 0.1
    r = grid.Rectangle(1.5, 1.2, 10.0, 10.0, material_value=3.0)
    m = _struct([r], zr=(0.23, 0.47), d=d)
    v = m.get_values(5, 6, 5, 6, 0, 8)[:, 0, 0].real
    zc = np.arange(8) * d
    fz = np.clip(np.minimum(zc + d / 2, 0.47) - np.maximum(zc - d / 2, 0.23), 0, None) / d
    np.testing.assert_allclose(v, 1.0 + 2.0 * fz, atol=1e-12)


def test_layer_priority():
    a = grid.Rectangle(1.0, 1.0, 1.0, 1.0, material_value=4.0, layer=1)
    b = grid.Rectangle(1.0, 1.0, 2.0, 2.0, material_value=2.0, layer=2)
    m = _struct([b, a])
    v = m.get_values(10, 11, 10, 11, 2, 3).ravel()[0]
    assert v == pytest.approx(4.0)       # lower layer number wins the overlap
    v = m.get_values(4, 5, 10, 11, 2, 3).ravel()[0]
    assert v == pytest.approx(2.0)


def test_ellipse_and_polygon_areas():
    d = 0.05
    e = grid.Ellipse(0.7, 0.6, 0.33, 0.21, angle=0.4, material_value=2.0)
    m = _struct([e], d=d, N=(30, 26, 4))
    v = m.get_values(0, 30, 0, 26, 1, 2)[0].real
    assert np.sum(v - 1.0) * d * d == pytest.approx(np.pi * 0.33 * 0.21, rel=2e-3)
    xs = np.array([0.2, 1.1, 0.9, 0.35])
    ys = np.array([0.15, 0.3, 1.0, 0.8])
    p = grid.Polygon(xs, ys, material_value=2.0)
    m = _struct([p], d=d, N=(30, 26, 4))
    v = m.get_values(0, 30, 0, 26, 1, 2)[0].real
    shoelace = 0.5 * abs(np.dot(xs, np.roll(ys, -1)) - np.dot(ys, np.roll(xs, -1)))
    assert np.sum(v - 1.0) * d * d == pytest.approx(shoelace, rel=2e-3)


def test_gaussian_fields_closed_forms():
    """misc.gaussian_fields (emopt/misc.py:580-636): at the waist (theta = 0)
    Ez = exp(-(x^2 + z^2) / w0^2), Hx = n0 Ez, Hy = 0; along the axis one
    Rayleigh range from the waist (theta = pi/2, x = y_R) the amplitude is
    1/sqrt(2) and the phase k0 y_R - pi/4 (the Gouy shift)."""
    from emopt_amd.misc import gaussian_fields
    w0, wl, n0 = 1.3, 1.55, 1.444
    x = np.linspace(-3, 3, 13)
    z = np.linspace(-2, 2, 13)
    Ez, Hx, Hy = gaussian_fields(x, z, 0.0, w0, 0.0, wl, n0)
    np.testing.assert_allclose(Ez, np.exp(-(x ** 2 + z ** 2) / w0 ** 2), rtol=1e-14)
    np.testing.assert_allclose(Hx, n0 * Ez, rtol=1e-14)
    assert np.all(Hy == 0)
    yR = np.pi * w0 ** 2 / wl
    Ez, Hx, Hy = gaussian_fields(np.array([yR]), np.array([0.0]), 0.0, w0, np.pi / 2, wl, n0)
    assert abs(abs(Ez[0]) - 1 / np.sqrt(2)) < 1e-14
    k0 = 2 * np.pi * n0 / wl
    dphi = np.angle(Ez[0] * np.exp(-1j * (k0 * yR - np.pi / 4)))
    assert abs(dphi) < 1e-12
    assert abs(Hy[0] + n0 * Ez[0]) < 1e-14
