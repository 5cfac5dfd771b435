"""Analytic shape derivatives of a structured material (d eps / d p).

The reference differentiates eps with respect to a design parameter by
re-sampling the material in the parameter's update box after a small step
and differencing (emopt/adjoint_method.py:504-533, emopt/fdtd.py:882-992).
Here the derivative of every boundary-smoothed cell comes from the geometry
of the overlap rule itself (StructuredMaterial2D::get_value,
src/Grid.cpp:460-525; restated in grid.py):

    value = sum over counted primitives r of val_r * (U_r - U_{r-1}) / A
            + background * (A - U_last) / A,

U_r being the area of the cell covered by the union of the counted
primitives up to r in priority order (a primitive counts when a cell corner
is strictly inside it; a first counted primitive covering all four corners
gives the cell its value outright).  Moving the boundary of a counted
primitive p with normal velocity V_n changes U_r, for r >= p, by the
integral of V_n over the part of p's boundary inside the cell and outside the
other counted primitives up to r.  Summed, the derivative of the value is

    d value / d p = (1/A) * integral over the visible part of p's boundary
                    in the cell of (val_p - val_next(x)) * V_n(x) dl,

where a boundary point is visible when no counted primitive before p
contains it, and val_next(x) is the value of the first counted primitive
after p containing it (the background when none does).  The gating corners
and the full-cell rule do not change under an infinitesimal move (except on
a set of measure zero), so they only decide which primitives count.

A parameter moves vertices of one primitive's ring (a rectangle's edges or a
polygon's vertices) with given velocities; along an edge the velocity is
linear between its two vertices, so each visible sub-segment integrates
exactly.  The 3D value of a Yee cell is the z-overlap-weighted sum of the
in-plane values of the z segments of the stack (grid.py), so its derivative
is the same weighted sum of in-plane derivatives over the segments where the
primitive is present.

Checked cell by cell against central differences of the material sampling
(tests/test_shape.py).
"""
from __future__ import annotations

import numpy as np

from .grid import Ellipse, Polygon, Rectangle

__all__ = ["boundary_deps", "rect_edge_velocity", "polygon_vertex_velocity"]

_EDGES = {"xl": (0, 1, 0), "xr": (2, 3, 0), "yb": (0, 3, 1), "yt": (1, 2, 1)}


def rect_edge_velocity(edge):
    """Per-vertex velocity (4, 2) of a Rectangle's ring (x0,y0), (x0,y1),
    (x1,y1), (x1,y0) when one edge coordinate ('xl', 'xr', 'yb', 'yt')
    increases by one."""
    a, b, ax = _EDGES[edge]
    u = np.zeros((4, 2))
    u[a, ax] = u[b, ax] = 1.0
    return u


def polygon_vertex_velocity(nv, index, direction):
    """Per-vertex velocity (nv, 2) when vertex `index` moves along
    `direction` (a 2-vector) by one."""
    u = np.zeros((nv, 2))
    u[index] = direction
    return u


def _clip_t(p0, p1, xa, xb, ya, yb):
    """Parameter range [t0, t1] of the segment p0 + t (p1 - p0) inside the
    box (Liang-Barsky), or None."""
    t0, t1 = 0.0, 1.0
    d = (p1[0] - p0[0], p1[1] - p0[1])
    for dv, lo, hi, pv in ((d[0], xa, xb, p0[0]), (d[1], ya, yb, p0[1])):
        if dv == 0.0:
            if pv < lo or pv > hi:
                return None
            continue
        ta, tb = (lo - pv) / dv, (hi - pv) / dv
        if ta > tb:
            ta, tb = tb, ta
        t0, t1 = max(t0, ta), min(t1, tb)
        if t0 >= t1:
            return None
    return t0, t1


def _crossings(p0, p1, ring, t0, t1):
    """Parameters in (t0, t1) where the segment crosses an edge of `ring`."""
    out = []
    xs, ys = ring.xs, ring.ys
    n = len(xs)
    dx, dy = p1[0] - p0[0], p1[1] - p0[1]
    for a in range(n):
        c = a + 1 if a + 1 < n else 0
        ex, ey = xs[c] - xs[a], ys[c] - ys[a]
        den = dx * ey - dy * ex
        if den == 0.0:
            continue
        t = ((xs[a] - p0[0]) * ey - (ys[a] - p0[1]) * ex) / den
        s = ((xs[a] - p0[0]) * dy - (ys[a] - p0[1]) * dx) / den
        if t0 < t < t1 and 0.0 <= s <= 1.0:
            out.append(t)
    return out


def _inside(ring, x, y):
    """Strictly inside (the rule's `within`: boundary points are out)."""
    b = ring.bbox
    if not (b[0] < x < b[1] and b[2] < y < b[3]):
        return False
    if ring.kind == 0:
        return True
    return bool(ring.within(np.array([x]), np.array([y]))[0])


def _inplane_derivative(rings, p, edges, xa, xb, ya, yb, bg):
    """d(in-plane value)/dp of the cell [xa, xb] x [ya, yb]; rings in
    priority order, p the index of the moving ring, edges a list of
    (P0, P1, U0, U1, n_out) for its moving edges."""
    corners = ((xa, ya), (xb, ya), (xb, yb), (xa, yb))
    counted = []
    for q, r in enumerate(rings):
        b = r.bbox
        if b[1] <= xa or b[0] >= xb or b[3] <= ya or b[2] >= yb:
            continue                     # no corner can be inside
        c = [_inside(r, x, y) for x, y in corners]
        if any(c):
            if not counted and all(c):
                return 0.0              # the cell takes this primitive's value outright
            counted.append(q)
    if p not in counted:
        return 0.0
    before = [rings[q] for q in counted if q < p]
    after = [rings[q] for q in counted if q > p]
    vp = rings[p].value
    total = 0.0
    for P0, P1, U0, U1, nrm in edges:
        tr = _clip_t(P0, P1, xa, xb, ya, yb)
        if tr is None:
            continue
        t0, t1 = tr
        ts = [t0, t1]
        for r in before + after:
            ts += _crossings(P0, P1, r, t0, t1)
        ts.sort()
        L = np.hypot(P1[0] - P0[0], P1[1] - P0[1])
        w0 = U0[0] * nrm[0] + U0[1] * nrm[1]            # V_n at t = 0 and t = 1
        w1 = U1[0] * nrm[0] + U1[1] * nrm[1]
        for ta, tb in zip(ts[:-1], ts[1:]):
            if not tb > ta:
                continue
            tm = 0.5 * (ta + tb)
            xm = P0[0] + tm * (P1[0] - P0[0])
            ym = P0[1] + tm * (P1[1] - P0[1])
            if any(_inside(r, xm, ym) for r in before):
                continue
            nxt = next((r.value for r in after if _inside(r, xm, ym)), bg)
            # integral of V_n(t) L dt with V_n linear from w0 to w1
            integ = L * (w0 * (tb - ta) + (w1 - w0) * 0.5 * (tb * tb - ta * ta))
            total += (vp - nxt) * integ
    return total / ((xb - xa) * (yb - ya))


def boundary_deps(sim, eps, prim, velocity, tol=0.0):
    """Sparse d eps / d p for a parameter p that moves the ring of `prim`
    (a Rectangle or Polygon of the StructuredMaterial3D `eps`, exact overlap
    rule) with per-vertex velocity `velocity` (nv, 2), at the three eps Yee
    offsets of `sim`.  Returns (component, flat cell index, complex dEps)
    arrays, as rect_edge_deps does, entries with |dEps| <= tol dropped."""
    if eps.overlap != "exact":
        raise ValueError("analytic derivatives follow the exact overlap rule")
    if not isinstance(prim, (Rectangle, Polygon)) or isinstance(prim, Ellipse):
        raise TypeError("boundary_deps moves rectangles and polygons")
    prims = eps._prims
    entry = next(e for e in prims if e[0] is prim)
    ring_p = eps._ring(prim)
    xs, ys = ring_p.xs, ring_p.ys
    nv = len(xs)
    u = np.asarray(velocity, dtype=float).reshape(nv, 2)
    sgn_area = np.sum(xs * np.roll(ys, -1) - np.roll(xs, -1) * ys)
    orient = 1.0 if sgn_area > 0 else -1.0      # counter-clockwise: +1
    edges = []
    for a in range(nv):
        b = a + 1 if a + 1 < nv else 0
        if not (np.any(u[a]) or np.any(u[b])):
            continue
        ex, ey = xs[b] - xs[a], ys[b] - ys[a]
        L = np.hypot(ex, ey)
        if L == 0.0:
            continue
        nrm = (orient * ey / L, -orient * ex / L)     # outward normal
        edges.append(((xs[a], ys[a]), (xs[b], ys[b]), u[a], u[b], nrm))
    if not edges:
        return (np.zeros(0, np.int32), np.zeros(0, np.int64), np.zeros(0, np.complex128))
    ex0 = min(min(P0[0], P1[0]) for P0, P1, *_ in edges)
    ex1 = max(max(P0[0], P1[0]) for P0, P1, *_ in edges)
    ey0 = min(min(P0[1], P1[1]) for P0, P1, *_ in edges)
    ey1 = max(max(P0[1], P1[1]) for P0, P1, *_ in edges)
    dx, dy, dz = eps._dx, eps._dy, eps._dz
    Nx, Ny, Nz = sim.Nx, sim.Ny, sim.Nz
    bg = eps.background
    active_all = [e for e in prims if not isinstance(e[0], Ellipse) or eps._ring(e[0]) is not None]
    bounds = sorted({b for _, z1, z2, _ in active_all for b in (z1, z2)})
    comps, cells, vals = [], [], []
    for c, (sx, sy, sz) in enumerate(sim.YEE_SHIFTS[:3]):
        k1 = max(0, int(np.floor(ex0 / dx - sx - 0.5)) - 1)
        k2 = min(Nx, int(np.ceil(ex1 / dx - sx + 0.5)) + 2)
        j1 = max(0, int(np.floor(ey0 / dy - sy - 0.5)) - 1)
        j2 = min(Ny, int(np.ceil(ey1 / dy - sy + 0.5)) + 2)
        zc = (np.arange(Nz) + sz) * dz
        zlo, zhi = zc - dz / 2, zc + dz / 2
        acc = {}
        for za, zb in zip(bounds[:-1], bounds[1:]):
            if not (entry[1] <= za and entry[2] >= zb):
                continue                         # the primitive is not in this segment
            fz = np.clip(np.minimum(zhi, zb) - np.maximum(zlo, za), 0, None) / dz
            iz = np.nonzero(fz > 0)[0]
            if len(iz) == 0:
                continue
            active = sorted([e for e in active_all if e[1] <= za and e[2] >= zb],
                            key=lambda e: (e[0].layer, e[3]))
            rings, p = [], None
            for e in active:
                r = eps._ring(e[0])
                if r is None:
                    continue
                if e[0] is prim:
                    p = len(rings)
                rings.append(r)
            for j in range(j1, j2):
                yc = (j + sy) * dy
                ya, yb = yc - dy / 2, yc + dy / 2
                if yb < ey0 or ya > ey1:
                    continue
                for k in range(k1, k2):
                    xc = (k + sx) * dx
                    xa, xb = xc - dx / 2, xc + dx / 2
                    if xb < ex0 or xa > ex1:
                        continue
                    d2 = _inplane_derivative(rings, p, edges, xa, xb, ya, yb, bg)
                    if d2 == 0.0:
                        continue
                    for i in iz:
                        acc[(i, j, k)] = acc.get((i, j, k), 0.0) + fz[i] * d2
        keys = [key for key, v in acc.items() if abs(v) > tol]
        if not keys:
            continue
        ijk = np.array(keys, dtype=np.int64)
        comps.append(np.full(len(keys), c, np.int32))
        cells.append((ijk[:, 0] * Ny + ijk[:, 1]) * Nx + ijk[:, 2])
        vals.append(np.array([acc[key] for key in keys], dtype=np.complex128))
    if not comps:
        return (np.zeros(0, np.int32), np.zeros(0, np.int64), np.zeros(0, np.complex128))
    return np.concatenate(comps), np.concatenate(cells), np.concatenate(vals)
