"""Analytic shape derivatives (emopt_amd/shape.py) against central
differences of the material sampling, cell by cell (CPU only).

The exact overlap rule (src/Grid.cpp:460-525, restated in grid.py) makes a
cell's value piecewise smooth in a boundary position: linear in a rectangle
edge, quadratic in a polygon vertex.  Central differences with a step well
inside one cell are then exact to rounding (rectangles) or to O(h^2)
(polygons), so the analytic derivative must match them on every cell the
move touches, with no cell missing and none extra:
  * a rectangle edge crossing layers of different priority, partly hidden
    by a higher-priority primitive and partly over a lower-priority one;
  * a trench edge of the grating workload (three eps components, Yee
    offsets, z slabs partly filled);
  * a vertex of a concave polygon overlapping a rectangle;
  * a tilted polygon edge over the background.
"""
import numpy as np
import pytest

from emopt_amd import grid
from emopt_amd.adjoint import rect_edge_deps
from emopt_amd.shape import boundary_deps, polygon_vertex_velocity, rect_edge_velocity


class GridSim:
    YEE_SHIFTS = ((0.5, 0.0, -0.5), (0.0, 0.5, -0.5), (0.0, 0.0, 0.0))

    def __init__(self, N):
        self.Nx, self.Ny, self.Nz = N


def _dense(N, deps):
    comp, cell, dv = deps
    out = np.zeros((3, N[0] * N[1] * N[2]), dtype=complex)
    np.add.at(out, (comp, cell), dv)
    return out


def _cd(sim, eps, setp, p0, h):
    """Central difference of the three eps components (host sampler)."""
    vals = []
    N = (sim.Nx, sim.Ny, sim.Nz)
    for s in (1, -1):
        setp(p0 + s * h)
        vals.append(np.array([eps.get_values(0, N[0], 0, N[1], 0, N[2], sx=a, sy=b, sz=c).ravel()
                              for a, b, c in sim.YEE_SHIFTS]))
    setp(p0)
    return (vals[0] - vals[1]) / (2 * h)


def _compare(an, cd, rtol):
    scale = np.abs(cd).max()
    assert scale > 0
    err = np.abs(an - cd).max()
    touched_an = np.abs(an) > 1e-9 * scale
    touched_cd = np.abs(cd) > 1e-6 * scale
    print("max |d eps/dp| %.4e, max error %.3e, cells %d analytic / %d differenced"
          % (scale, err, touched_an.sum(), touched_cd.sum()))
    assert err <= rtol * scale
    assert np.all(touched_an[touched_cd])


def _edge_rect(xl, xr, y0, h, v, layer):
    r = grid.Rectangle(0.5 * (xl + xr), y0, xr - xl, h, material_value=v, layer=layer)
    return r


def _set_edge(r, edge):
    def setp(v):
        x0, x1 = r.x0 - r.width / 2, r.x0 + r.width / 2
        y0, y1 = r.y0 - r.height / 2, r.y0 + r.height / 2
        if edge == "xl":
            x0 = v
        elif edge == "xr":
            x1 = v
        elif edge == "yb":
            y0 = v
        else:
            y1 = v
        r.x0, r.width = 0.5 * (x0 + x1), x1 - x0
        r.y0, r.height = 0.5 * (y0 + y1), y1 - y0
    return setp


def _edge_value(r, edge):
    return {"xl": r.x0 - r.width / 2, "xr": r.x0 + r.width / 2,
            "yb": r.y0 - r.height / 2, "yt": r.y0 + r.height / 2}[edge]


@pytest.mark.parametrize("edge", ["xl", "xr", "yb", "yt"])
def test_rectangle_edge_over_layers(edge):
    d, N = 0.05, (24, 22, 6)
    X, Y, Z = [(n - 1) * d for n in N]
    eps = grid.StructuredMaterial3D(X, Y, Z, d, d, d)
    eps.add_primitive(grid.Rectangle(X / 2, Y / 2, 2 * X, 2 * Y, material_value=2.0, layer=5),
                      -1, 0.17)
    lo = grid.Rectangle(0.61, 0.44, 0.37, 0.29, material_value=7.0, layer=3)     # below p
    hi = grid.Rectangle(0.37, 0.71, 0.21, 0.33, material_value=4.5, layer=1)     # above p
    p = _edge_rect(0.303, 0.717, 0.412, 0.513, 11.0, 2)
    for r in (lo, hi, p):
        eps.add_primitive(r, 0.031, 0.143)
    sim = GridSim(N)
    an = _dense(N, boundary_deps(sim, eps, p, rect_edge_velocity(edge)))
    cd = _cd(sim, eps, _set_edge(p, edge), _edge_value(p, edge), 1e-5 * d)
    _compare(an, cd, 1e-8)


def test_grating_trench_edge_matches_rect_edge_deps():
    """A trench edge of the grating workload's stack against rect_edge_deps
    (the central difference the gradient used before)."""
    from tests.cases import _edge_rect_cls, span, EPS_OX, EPS_SI
    EdgeRect = _edge_rect_cls()
    N, d = (40, 10, 30), 0.03
    X, Y, Z = (span(n, d) for n in N)
    eps = grid.StructuredMaterial3D(X, Y, Z, d, d, d)
    eps.add_primitive(grid.Rectangle(X / 2, Y / 2, 4 * X, 4 * Y, material_value=EPS_OX, layer=3),
                      -Z, 0.4)
    eps.add_primitive(grid.Rectangle(X / 2, Y / 2, 4 * X, 4 * Y, material_value=EPS_SI, layer=2),
                      0.4, 0.62)
    tr = EdgeRect(0.4137, 0.6851, Y / 2, 4 * Y, material_value=1.0, layer=1)
    eps.add_primitive(tr, 0.62 - 0.11, 0.62)
    sim = GridSim(N)
    sim.YEE_SHIFTS = GridSim.YEE_SHIFTS
    for edge, attr in (("xl", "xl"), ("xr", "xr")):
        # EdgeRect's ring is the Rectangle ring of its bbox: the x edges move
        an = _dense(N, boundary_deps(sim, eps, tr, rect_edge_velocity(edge)))
        cd = _dense(N, rect_edge_deps(sim, eps, tr, attr, 1e-6 * d, (0, N[0], 0, N[1], 0, N[2])))
        _compare(an, cd, 1e-7)


@pytest.mark.parametrize("vertex", [0, 2, 4])
def test_polygon_vertex(vertex):
    d, N = 0.05, (26, 24, 4)
    X, Y, Z = [(n - 1) * d for n in N]
    eps = grid.StructuredMaterial3D(X, Y, Z, d, d, d)
    # no vertex on a cell boundary (the value has a kink there)
    xs = np.array([0.3113, 0.9307, 0.8791, 0.6218, 0.5537, 0.2711])
    ys = np.array([0.2231, 0.2794, 0.8309, 0.5117, 0.8593, 0.7086])
    poly = grid.Polygon(xs.copy(), ys.copy(), material_value=9.0, layer=2)
    over = grid.Rectangle(0.71, 0.37, 0.33, 0.19, material_value=3.0, layer=1)   # hides a part
    under = grid.Rectangle(0.45, 0.66, 0.41, 0.37, material_value=5.0, layer=4)
    for r in (poly, over, under):
        eps.add_primitive(r, 0.0, 0.11)
    sim = GridSim(N)
    direction = np.array([0.6, -0.8])
    ring = eps._ring(poly)
    assert np.array_equal(ring.xs, xs)
    an = _dense(N, boundary_deps(sim, eps, poly, polygon_vertex_velocity(len(xs), vertex,
                                                                         direction)))

    def setp(t):
        poly.xs = xs.copy()
        poly.ys = ys.copy()
        poly.xs[vertex] += t * direction[0]
        poly.ys[vertex] += t * direction[1]
    cd = _cd(sim, eps, setp, 0.0, 1e-5 * d)
    _compare(an, cd, 1e-6)


def test_planted_missing_crossings_fails(monkeypatch):
    """Planted in the derivative: the breakpoints where the moving edge
    enters another primitive are dropped, so a partly hidden sub-segment is
    taken as wholly visible or hidden.  The differences catch it."""
    import emopt_amd.shape as S
    d, N = 0.05, (24, 22, 6)
    X, Y, Z = [(n - 1) * d for n in N]
    eps = grid.StructuredMaterial3D(X, Y, Z, d, d, d)
    hi = grid.Rectangle(0.37, 0.71, 0.21, 0.33, material_value=4.5, layer=1)
    p = _edge_rect(0.303, 0.717, 0.412, 0.513, 11.0, 2)
    for r in (hi, p):
        eps.add_primitive(r, 0.031, 0.143)
    sim = GridSim(N)
    cd = _cd(sim, eps, _set_edge(p, "yt"), _edge_value(p, "yt"), 1e-5 * d)
    good = _dense(N, S.boundary_deps(sim, eps, p, rect_edge_velocity("yt")))
    assert np.abs(good - cd).max() <= 1e-8 * np.abs(cd).max()
    monkeypatch.setattr(S, "_crossings", lambda *a: [])
    bad = _dense(N, S.boundary_deps(sim, eps, p, rect_edge_velocity("yt")))
    assert np.abs(bad - cd).max() > 0.1 * np.abs(cd).max()
