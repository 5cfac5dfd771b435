"""The reference's in-plane overlap rule (StructuredMaterial2D::get_value,
src/Grid.cpp:460-525; GridCell::intersect, src/Grid.cpp:74-108) in
emopt_amd.grid, checked against geometry computed another way:

* cell overlaps from Sutherland-Hodgman clipping of the primitive to the
  cell (a convex window, so concave subjects clip correctly) and the shoelace
  area, instead of the slab sweep;
* two overlapping convex primitives by inclusion-exclusion, their
  intersection clipped the same way;
* corner containment by half-plane tests (convex primitives) or by closed
  form (rectangles);
* the rule's control flow on hand-built cells: a primitive with no corner
  inside a cell is skipped even where it crosses it, four corners inside
  give the full value even when the primitive has a notch inside the cell,
  later primitives take only what is left.

boost.geometry is not available here, so the reference itself is not run.
"""
import numpy as np
import pytest

from emopt_amd import grid
from emopt_amd.grid import Ellipse, Polygon, Rectangle, StructuredMaterial3D


def shoelace(xs, ys):
    return 0.5 * abs(np.dot(xs, np.roll(ys, -1)) - np.dot(ys, np.roll(xs, -1)))


def clip_convex(subject, window):
    """Sutherland-Hodgman: subject polygon clipped to a convex window (both
    lists of (x, y), window counter-clockwise)."""
    out = list(subject)
    n = len(window)
    for a in range(n):
        (x1, y1), (x2, y2) = window[a], window[(a + 1) % n]
        inp, out = out, []
        if not inp:
            break

        def side(p):
            return (x2 - x1) * (p[1] - y1) - (y2 - y1) * (p[0] - x1)

        for b in range(len(inp)):
            p, q = inp[b - 1], inp[b]
            sp, sq = side(p), side(q)
            if (sp >= 0) != (sq >= 0):
                t = sp / (sp - sq)
                out.append((p[0] + t * (q[0] - p[0]), p[1] + t * (q[1] - p[1])))
            if sq >= 0:
                out.append(q)
    return out


def area(poly):
    if len(poly) < 3:
        return 0.0
    xs, ys = np.array(poly).T
    return shoelace(xs, ys)


def ccw(poly):
    xs, ys = np.array(poly).T
    s = np.dot(xs, np.roll(ys, -1)) - np.dot(ys, np.roll(xs, -1))
    return list(poly) if s > 0 else list(poly)[::-1]


def inside_convex(poly, x, y):
    """Strictly inside a convex counter-clockwise polygon."""
    n = len(poly)
    for a in range(n):
        (x1, y1), (x2, y2) = poly[a], poly[(a + 1) % n]
        if (x2 - x1) * (y - y1) - (y2 - y1) * (x - x1) <= 0:
            return False
    return True


def cell_box(xc, yc, d):
    h = d / 2
    return [(xc - h, yc - h), (xc + h, yc - h), (xc + h, yc + h), (xc - h, yc + h)]


def plane(m, n, sx=0.0, sy=0.0):
    return m.get_values(0, n, 0, n, 0, 1, sx=sx, sy=sy, sz=0.5)[0]


def material(prims, d=0.1, L=2.0, **kw):
    m = StructuredMaterial3D(L, L, 1.0, d, d, d, **kw)
    for p in prims:
        m.add_primitive(p, 0.0, 1.0)
    return m


def rotated_square(cx, cy, h, ang):
    t = ang + np.pi / 4 + np.arange(4) * np.pi / 2
    r = h * np.sqrt(2)
    return list(zip(cx + r * np.cos(t), cy + r * np.sin(t)))


@pytest.mark.parametrize("sx,sy", [(0.0, 0.0), (0.5, 0.0), (0.0, 0.5), (0.5, 0.5)])
def test_rectangle_closed_form(sx, sy):
    d, n = 0.1, 20
    x0, x1, y0, y1 = 0.537, 1.361, 0.402, 1.219
    m = material([Rectangle((x0 + x1) / 2, (y0 + y1) / 2, x1 - x0, y1 - y0, material_value=5.0)], d)
    v = plane(m, n, sx, sy)
    for j in range(n):
        for k in range(n):
            xc, yc = (k + sx) * d, (j + sy) * d
            corners = [(xc - d / 2, yc - d / 2), (xc + d / 2, yc - d / 2),
                       (xc + d / 2, yc + d / 2), (xc - d / 2, yc + d / 2)]
            ins = [x0 < x < x1 and y0 < y < y1 for x, y in corners]
            fx = max(min(xc + d / 2, x1) - max(xc - d / 2, x0), 0) / d
            fy = max(min(yc + d / 2, y1) - max(yc - d / 2, y0), 0) / d
            want = 5.0 if all(ins) else (fx * fy * 5.0 + (1 - fx * fy) if any(ins) else 1.0)
            assert abs(v[j, k] - want) < 1e-12, (j, k, v[j, k], want)


@pytest.mark.parametrize("shape", ["square", "triangle", "hexagon"])
def test_convex_polygon_against_clipping(shape):
    d, n = 0.1, 20
    if shape == "square":
        poly = rotated_square(1.03, 0.97, 0.47, 0.37)
    elif shape == "triangle":
        poly = [(0.31, 0.27), (1.73, 0.52), (0.88, 1.69)]
    else:
        t = 0.2 + np.arange(6) * np.pi / 3
        poly = list(zip(1.01 + 0.63 * np.cos(t), 0.98 + 0.63 * np.sin(t)))
    poly = ccw(poly)
    xs, ys = np.array(poly).T
    m = material([Polygon(xs, ys, material_value=3.0 + 0.5j)], d)
    v = plane(m, n, 0.5, 0.5)
    checked = 0
    for j in range(n):
        for k in range(n):
            xc, yc = (k + 0.5) * d, (j + 0.5) * d
            box = cell_box(xc, yc, d)
            ins = [inside_convex(poly, x, y) for x, y in box]
            ov = area(clip_convex(poly, box)) / (d * d)
            if all(ins):
                want = 3.0 + 0.5j
            elif any(ins):
                want = ov * (3.0 + 0.5j) + (1 - ov)
                checked += 1
            else:
                want = 1.0
            assert abs(v[j, k] - want) < 1e-12, (j, k, v[j, k], want)
    assert checked > 10


def test_concave_polygon_against_clipping():
    """An L-shape and a five-pointed star (concave subjects, clipped to each
    convex cell)."""
    d, n = 0.1, 20
    t = np.pi / 2 + np.arange(10) * np.pi / 5
    r = np.where(np.arange(10) % 2 == 0, 0.83, 0.34)
    star = list(zip(1.02 + r * np.cos(t), 0.99 + r * np.sin(t)))
    ell = [(0.23, 0.21), (1.71, 0.21), (1.71, 0.77), (0.81, 0.77), (0.81, 1.83), (0.23, 1.83)]
    for poly in (star, ell):
        poly = ccw(poly)
        xs, ys = np.array(poly).T
        m = material([Polygon(xs, ys, material_value=2.0)], d)
        v = plane(m, n, 0.5, 0.5)
        inside = lambda x, y: bool(grid._Ring(2, xs, ys, (xs.min(), xs.max(), ys.min(), ys.max()),
                                              2.0).within(np.array([x]), np.array([y]))[0])
        for j in range(n):
            for k in range(n):
                xc, yc = (k + 0.5) * d, (j + 0.5) * d
                box = cell_box(xc, yc, d)
                ins = [inside(x, y) for x, y in box]
                ov = area(clip_convex(poly, box)) / (d * d)
                if all(ins):
                    want = 2.0
                elif any(ins):
                    want = 2.0 * ov + (1 - ov)
                else:
                    want = 1.0
                assert abs(v[j, k] - want) < 1e-12, (j, k, v[j, k], want)


def test_finely_sampled_curve_against_clipping():
    """A 1200-vertex ellipse polygon: near its x extremes a hundred edges
    span one cell's x range, few pass through the cell.  Both area methods
    sum many short pieces here, hence 1e-11."""
    d, n = 0.07, 28
    e = Ellipse(1.01, 0.98, 0.77, 0.52, angle=0.3)
    xs, ys = e.polygon(1200)
    poly = ccw(list(zip(xs, ys)))
    m = material([Polygon(xs, ys, material_value=2.0)], d)
    v = plane(m, n, 0.5, 0.5)
    r = grid._Ring(2, xs, ys, (xs.min(), xs.max(), ys.min(), ys.max()), 2.0)
    c = (np.arange(n) + 0.5) * d
    lo, hi = c - d / 2, c + d / 2
    ins = np.stack([r.within(*np.meshgrid(px, py)) for px, py in
                    ((lo, lo), (hi, lo), (hi, hi), (lo, hi))])
    for j in range(n):
        for k in range(n):
            if ins[:, j, k].all() or not ins[:, j, k].any():
                continue
            box = cell_box(c[k], c[j], d)
            ov = area(clip_convex(poly, box)) / (d * d)
            assert abs(v[j, k] - (2.0 * ov + (1 - ov))) < 1e-11, (j, k)


def test_within_against_half_planes():
    poly = ccw(rotated_square(1.0, 1.0, 0.5, 0.3))
    xs, ys = np.array(poly).T
    r = grid._Ring(2, xs, ys, (xs.min(), xs.max(), ys.min(), ys.max()), 1.0)
    rng = np.random.default_rng(0)
    px, py = rng.uniform(0.2, 1.8, 4000), rng.uniform(0.2, 1.8, 4000)
    got = r.within(px, py)
    want = np.array([inside_convex(poly, x, y) for x, y in zip(px, py)])
    assert np.array_equal(got, want)
    # vertices and edge midpoints of an axis-aligned square are on the boundary: out
    sq = grid._Ring(2, np.array([0.0, 1.0, 1.0, 0.0]), np.array([0.0, 0.0, 1.0, 1.0]),
                    (0.0, 1.0, 0.0, 1.0), 1.0)
    assert not sq.within(np.array([0.0, 0.5, 1.0, 0.5]), np.array([0.0, 0.0, 0.5, 1.0])).any()
    assert sq.within(np.array([0.5]), np.array([0.5]))[0]
    # a point on a slanted edge, strictly inside the bounding box, where the
    # even-odd count alone says inside
    dia = grid._Ring(2, np.array([1.0, 2.0, 1.0, 0.0]), np.array([0.0, 1.0, 2.0, 1.0]),
                     (0.0, 2.0, 0.0, 2.0), 1.0)
    assert not dia.within(np.array([0.5]), np.array([0.5]))[0]
    assert dia.within(np.array([0.5 + 1e-9]), np.array([0.5 + 1e-9]))[0]


def test_two_convex_primitives_inclusion_exclusion():
    """P1 (layer 1) over P2 (layer 2): where both count, P2 takes
    area(cell & P2) - area(cell & P1 & P2)."""
    d, n = 0.1, 20
    p1 = ccw(rotated_square(0.83, 0.91, 0.41, 0.21))
    p2 = ccw([(0.52, 0.38), (1.67, 0.61), (1.42, 1.58), (0.61, 1.37)])
    m1, m2 = 4.0, 2.5
    a1, b1 = np.array(p1).T
    a2, b2 = np.array(p2).T
    m = material([Polygon(a2, b2, material_value=m2, layer=2),
                  Polygon(a1, b1, material_value=m1, layer=1)], d)
    v = plane(m, n, 0.5, 0.5)
    both = 0
    for j in range(n):
        for k in range(n):
            xc, yc = (k + 0.5) * d, (j + 0.5) * d
            box = cell_box(xc, yc, d)
            A = d * d
            val, taken, done = 0.0, 0.0, False
            i1 = [inside_convex(p1, x, y) for x, y in box]
            i2 = [inside_convex(p2, x, y) for x, y in box]
            if all(i1):
                val, done = m1, True
            elif any(i1):
                o1 = area(clip_convex(p1, box)) / A
                val += o1 * m1
                taken += o1
            if not done:
                if all(i2) and taken == 0.0:
                    val, done = m2, True
                elif any(i2):
                    o12 = area(clip_convex(clip_convex(p2, p1), box)) / A if any(i1) else 0.0
                    o2 = area(clip_convex(p2, box)) / A - o12
                    val += o2 * m2
                    taken += o2
                    both += any(i1)
            if not done:
                val += max(1.0 - taken, 0.0)
            assert abs(v[j, k] - val) < 1e-12, (j, k, v[j, k], val)
    assert both > 3


def test_control_flow_cells():
    d = 0.1
    # a strip narrower than a cell through the middle of cell (10, 10)
    # (corners at 0.95 and 1.05): no corner inside, so the cell is background
    m = material([Rectangle(1.0, 1.0, 0.04, 0.04, material_value=9.0)], d)
    assert plane(m, 20, 0.0, 0.0)[10, 10] == 1.0
    # the same strip covering one corner: it counts, with its exact area
    m = material([Rectangle(0.95, 0.95, 0.04, 0.04, material_value=9.0)], d)
    v = plane(m, 20, 0.0, 0.0)[10, 10]
    assert abs(v - (0.02 * 0.02 / 0.01 * 9.0 + (1 - 0.04))) < 1e-12
    # a square with a notch cut inside cell (10, 10) but all four cell
    # corners inside: the cell takes the full value
    xs = [0.5, 1.5, 1.5, 1.02, 1.02, 0.98, 0.98, 0.5]
    ys = [0.5, 0.5, 1.5, 1.5, 1.0, 1.0, 1.5, 1.5]
    m = material([Polygon(xs, ys, material_value=6.0)], d)
    assert plane(m, 20, 0.0, 0.0)[10, 10] == 6.0
    # a cell that a first primitive covers in part, crossed by a strip with no
    # corner inside it: the strip is skipped, the rest is background
    m = material([Rectangle(0.9, 0.9, 0.2, 0.2, material_value=4.0, layer=1),
                  Rectangle(1.0, 1.0, 0.02, 2.0, material_value=9.0, layer=2)], d)
    v = plane(m, 20, 0.0, 0.0)[10, 10]      # cell [0.95, 1.05]^2, first covers 0.05^2
    assert abs(v - (0.25 * 4.0 + 0.75)) < 1e-12
    # a diamond whose left edge passes exactly through a cell corner: that
    # corner is on the boundary, so not inside
    m = material([Polygon([1.0, 2.0, 1.0, 0.0], [0.0, 1.0, 2.0, 1.0], material_value=5.0)], 0.5,
                 L=2.0)
    v = m.get_values(0, 4, 0, 4, 0, 1, sx=0.5, sy=0.5, sz=0.5)[0]
    # cell (k=0, j=0) = [0, 0.5]^2: only its corner (0.5, 0.5) touches the
    # diamond, on the edge, so the cell is background
    assert v[0, 0] == 1.0
    # priority: a full-cell low layer hides a later one
    m = material([Rectangle(1.0, 1.0, 1.0, 1.0, material_value=7.0, layer=2),
                  Rectangle(1.0, 1.0, 1.0, 1.0, material_value=3.0, layer=1)], d)
    assert plane(m, 20)[10, 10] == 3.0


def test_ellipse_polygon_closed_form():
    e = Ellipse(1.0, 1.0, 0.6, 0.35, angle=0.4)
    for n in (16, 64, 128):
        xs, ys = e.polygon(n)
        assert abs(shoelace(xs, ys) - 0.6 * 0.35 * n / 2 * np.sin(2 * np.pi / n)) < 1e-13
    ec = Ellipse(1.0, 1.0, 0.6, 0.6, clip=(0.0, 1.0, 0.0, 2.0))      # left half disc
    xs, ys = ec.polygon(128)
    assert xs.max() <= 1.0
    assert abs(shoelace(xs, ys) - 0.36 * 64 * np.sin(2 * np.pi / 128) / 2) < 1e-13
    assert len(Ellipse(1.0, 1.0, 0.2, 0.2, clip=(3.0, 4.0, 0.0, 1.0)).polygon(64)[0]) == 0


def test_exact_plane_totals():
    """Sum of (value - 1) x cell area over a plane = (eps - 1) x polygon area
    when every boundary cell of the polygon has a corner inside it (a large
    convex polygon with obtuse corners)."""
    d, n = 0.05, 40
    t = 0.1 + np.arange(8) * np.pi / 4
    poly = ccw(list(zip(1.0 + 0.71 * np.cos(t), 1.0 + 0.71 * np.sin(t))))
    xs, ys = np.array(poly).T
    m = material([Polygon(xs, ys, material_value=3.0)], d, L=2.0)
    v = plane(m, n, 0.5, 0.5)
    assert abs(((v.real - 1.0) * d * d).sum() - 2.0 * shoelace(xs, ys)) < 1e-12


def test_limits_raise():
    d = 0.1
    prims = [Rectangle(1.0, 1.0, 1.0 - 0.01 * q, 1.0 + 0.013 * q, material_value=2.0, layer=q)
             for q in range(10)]
    m = material(prims, d)
    with pytest.raises(ValueError, match="more than 8 primitives"):
        m.get_values(0, 20, 0, 20, 0, 1, sx=0.3, sy=0.3, sz=0.5)


def test_overlap_argument_and_device_records():
    """The exact rule hands the device kernel rectangles and polygons only:
    an ellipse goes as its inscribed polygon (kind 2, ellipse_vertices
    vertices) and one clipped to nothing is left out, on the host too."""
    with pytest.raises(ValueError):
        StructuredMaterial3D(1, 1, 1, 0.1, 0.1, 0.1, overlap="boost")
    m = StructuredMaterial3D(2.0, 2.0, 1.0, 0.1, 0.1, 0.1, ellipse_vertices=40)
    m.add_primitive(Rectangle(1.0, 1.0, 0.5, 0.4, material_value=3.0), 0.0, 1.0)
    m.add_primitive(Ellipse(1.0, 1.0, 0.3, 0.2, angle=0.2, material_value=5.0, layer=0), 0.2, 0.6)
    m.add_primitive(Ellipse(1.0, 1.0, 0.3, 0.2, clip=(5.0, 6.0, 0.0, 1.0), material_value=7.0),
                    0.33, 0.77)
    recs, vx, vy, nv = m.device_primitives()
    assert list(recs["kind"]) == [2, 0]             # priority order: layer 0 first
    assert recs["nv"][0] == 40 and nv == 40
    xs, ys = Ellipse(1.0, 1.0, 0.3, 0.2, angle=0.2).polygon(40)
    assert np.array_equal(vx[:40], xs) and np.array_equal(vy[:40], ys)
    assert tuple(recs["bbox"][0]) == (xs.min(), xs.max(), ys.min(), ys.max())
    # on the host the clipped-away ellipse is left out too: the same doubles as
    # the material built without it
    m2 = StructuredMaterial3D(2.0, 2.0, 1.0, 0.1, 0.1, 0.1, ellipse_vertices=40)
    m2.add_primitive(Rectangle(1.0, 1.0, 0.5, 0.4, material_value=3.0), 0.0, 1.0)
    m2.add_primitive(Ellipse(1.0, 1.0, 0.3, 0.2, angle=0.2, material_value=5.0, layer=0), 0.2, 0.6)
    for sz in (0.0, -0.5, 0.3):
        assert np.array_equal(m.get_values(0, 20, 0, 20, 0, 10, sz=sz),
                              m2.get_values(0, 20, 0, 20, 0, 10, sz=sz))
    sm = StructuredMaterial3D(2.0, 2.0, 1.0, 0.1, 0.1, 0.1, overlap="supersample")
    sm.add_primitive(Ellipse(1.0, 1.0, 0.3, 0.2, material_value=5.0), 0.0, 1.0)
    assert list(sm.device_primitives()[0]["kind"]) == [1]
