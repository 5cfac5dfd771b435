"""Host parts of the experimental path (emopt_amd.experimental): the torch
shapes, the three materials and the sigmoid parameter maps.  Each is pinned
by a closed form, an identity or brute force, not by restating its code:

* rect1d with the linear ramp and k = 1/dx gives each cell its exact fill
  fraction of the interval (overlap length / dx);
* an axis-aligned square as a clockwise polygon equals the rectangle;
  reversing the vertex order (counter-clockwise) leaves an empty shape;
* union_b is 1 - prod(1 - s); intersection of 0/1 masks is the logical and;
* autograd of a shape equals central differences of the same shape;
* TopologyMaterial3D / HybridMaterial3D values equal a cell-by-cell loop
  over the box (background outside the design domain, the grid or the
  function there, the function seeing the background);
* the sigmoid maps invert each other over the squished bounds.
"""
import numpy as np
import pytest
import torch

from emopt_amd import grid
from emopt_amd.misc import DomainCoordinates as DC
from emopt_amd.experimental import autodiff_geometry as adg
from emopt_amd.experimental.adjoint_method import (inverse_scaled_sigmoid, scaled_sigmoid,
                                                   sigmoid)
from emopt_amd.experimental.grid import (AutoDiffMaterial3D, HybridMaterial3D,
                                         TopologyMaterial3D)

F64 = torch.float64


def test_linear_ramp_rect_is_exact_fill_fraction():
    dx = 0.04
    x = torch.arange(0, 60, dtype=F64) * dx
    rng = np.random.default_rng(1)
    for _ in range(20):
        a = rng.uniform(0.1, 1.0)
        b = a + rng.uniform(2 * dx, 1.2)
        s = adg.rect1d(1.0 / dx, x, torch.tensor([a, b], dtype=F64), nl=adg.nl_lin).numpy()
        xc = x.numpy()
        exact = np.clip(np.minimum(xc + dx / 2, b) - np.maximum(xc - dx / 2, a), 0, None) / dx
        np.testing.assert_allclose(s, exact, rtol=0, atol=1e-12)


def test_square_polygon_equals_rectangle():
    d = 0.05
    x = torch.arange(0, 40, dtype=F64) * d
    y = torch.arange(0, 30, dtype=F64) * d
    x0, x1, y0, y1 = 0.33, 1.41, 0.27, 1.02
    cw = torch.tensor([[x0, y0], [x0, y1], [x1, y1], [x1, y0]], dtype=F64)   # clockwise
    r = adg.rect2d(1 / d, x, y, torch.tensor([x0, x1, y0, y1], dtype=F64), nl=adg.nl_lin)
    p = adg.poly2d(1 / d, x, y, cw, nl=adg.nl_lin)
    assert p.shape == (30, 40)
    np.testing.assert_allclose(p.numpy(), r.numpy(), rtol=0, atol=1e-12)
    ccw = torch.flip(cw, dims=[0])
    assert float(adg.poly2d(1 / d, x, y, ccw, nl=adg.nl_lin).max()) == 0.0


def test_set_operators():
    rng = np.random.default_rng(2)
    s = [torch.tensor(rng.uniform(size=(5, 7))) for _ in range(4)]
    ub = adg.union_b(s).numpy()
    np.testing.assert_allclose(ub, 1 - np.prod([1 - t.numpy() for t in s], axis=0), atol=1e-14)
    m = [torch.tensor((rng.uniform(size=(5, 7)) > 0.5).astype(np.float64)) for _ in range(3)]
    both = np.logical_and.reduce([t.numpy() > 0.5 for t in m]).astype(np.float64)
    either = np.logical_or.reduce([t.numpy() > 0.5 for t in m]).astype(np.float64)
    assert np.array_equal(adg.intersection(m).numpy(), both)
    assert np.array_equal(adg.intersection_b(m).numpy(), both)
    assert np.array_equal(adg.union(m).numpy(), either)
    assert np.array_equal(adg.union_b(m).numpy(), either)


@pytest.mark.parametrize("nl", [adg.nl_sig, adg.nl_erf, adg.nl_sin, adg.nl_quad, adg.nl_lin])
def test_nonlinearities_are_monotone_steps(nl):
    t = torch.linspace(-3, 3, 601, dtype=F64)
    v = nl(2.0, t).numpy()
    assert v.min() >= 0 and v.max() <= 1
    assert np.all(np.diff(v) >= 0)
    assert abs(nl(2.0, torch.zeros(1, dtype=F64)).item() - 0.5) < 1e-15
    assert v[0] < 1e-2 and v[-1] > 1 - 1e-2


def test_shape_gradients_match_central_differences():
    d = 0.05
    x = torch.arange(0, 30, dtype=F64) * d
    y = torch.arange(0, 26, dtype=F64) * d
    z = torch.arange(0, 12, dtype=F64) * d
    w = torch.tensor(np.random.default_rng(3).normal(size=(12, 26, 30)))

    def f(v):
        disc = adg.circ2d(1 / d, x, y, v[:3])
        box = adg.rect2d(1 / d, x, y, v[3:7], nl=adg.nl_sig)
        star = adg.polar2d(1 / d, x, y, v[7:12], order=3)
        shape = adg.depth(adg.union_b([disc, box, star]), 1 / d, z, v[12:14], nl=adg.nl_erf)
        return (w * shape).sum()

    # the star's centre (v[10], v[11]) lies off the grid points: at a grid
    # point atan2's derivative is singular
    v0 = np.array([0.4, 0.7, 0.6, 0.2, 0.9, 0.3, 1.0, 0.35, 0.2, 0.3, 0.813, 0.617, 0.15, 0.42])
    v = torch.tensor(v0, requires_grad=True)
    g = torch.autograd.grad(f(v), v)[0].numpy()
    h = 1e-6
    fd = np.array([(f(torch.tensor(v0 + h * e)).item() - f(torch.tensor(v0 - h * e)).item()) / (2 * h)
                   for e in np.eye(len(v0))])
    np.testing.assert_allclose(g, fd, rtol=1e-6, atol=1e-6 * np.abs(fd).max())


def _background(X=1.6, Y=1.2, Z=0.8, d=0.04):
    eps = grid.StructuredMaterial3D(X, Y, Z, d, d, d)
    eps.add_primitive(grid.Rectangle(X / 2, Y / 2, 2 * X, 2 * Y, material_value=2.0, layer=2),
                      -Z, 2 * Z)
    eps.add_primitive(grid.Rectangle(X / 2, Y / 2, 0.5, 0.3, material_value=12.0, layer=1),
                      Z / 2 - 0.1, Z / 2 + 0.1)
    return eps


def test_topology_material_values():
    d = 0.04
    eps = _background(d=d)
    fd = DC(0.4, 1.0, 0.3, 0.8, 0.2, 0.5, d, d, d)
    top = TopologyMaterial3D(eps, fd)
    assert top.grid.shape == (fd.Nz, fd.Ny, fd.Nx)
    np.testing.assert_array_equal(top.grid, eps.get_values_in(fd))
    g = np.random.default_rng(4).uniform(2, 12, size=top.grid.shape)
    top.grid = g
    box = (3, 36, 2, 28, 1, 18)          # (k1, k2, j1, j2, i1, i2), overlaps part of fd
    for sx, sy, sz in ((0.5, 0.0, -0.5), (0.0, 0.0, 0.0)):
        got = top.get_values(*box, sx=sx, sy=sy, sz=sz)
        bg = eps.get_values(*box, sx=sx, sy=sy, sz=sz)
        k1, k2, j1, j2, i1, i2 = box
        for i in range(i1, i2):
            for j in range(j1, j2):
                for k in range(k1, k2):
                    inside = fd.i1 <= i < fd.i2 and fd.j1 <= j < fd.j2 and fd.k1 <= k < fd.k2
                    want = g[i - fd.i1, j - fd.j1, k - fd.k1] if inside else bg[i - i1, j - j1, k - k1]
                    assert got[i - i1, j - j1, k - k1] == want
    # a flat array argument receives the values in place
    flat = np.zeros(np.prod(got.shape), dtype=np.complex128)
    top.get_values(*box, arr=flat)
    np.testing.assert_array_equal(flat.reshape(got.shape), top.get_values(*box))


def test_autodiff_and_hybrid_material_values():
    d = 0.04
    eps = _background(d=d)

    def func(v, coords, bg=None):
        z, y, x = coords
        core = adg.rect2d(1 / d, x, y, v[:4], nl=adg.nl_lin)
        shape = adg.depth(core, 1 / d, z, v[4:6], nl=adg.nl_lin)
        base = 2.0 if bg is None else torch.as_tensor(np.real(bg))
        return base + 10.0 * shape

    v = np.array([0.5, 1.1, 0.4, 0.7, 0.3, 0.5])
    ad = AutoDiffMaterial3D(d, d, d, func, v)
    k1, k2, j1, j2, i1, i2 = 0, 41, 0, 31, 0, 21
    for sx, sy, sz in ((0.5, 0.0, -0.5), (0.0, 0.5, 0.0)):
        got = ad.get_values(k1, k2, j1, j2, i1, i2, sx=sx, sy=sy, sz=sz)
        x = (np.arange(k1, k2) + sx) * d
        y = (np.arange(j1, j2) + sy) * d
        z = (np.arange(i1, i2) + sz) * d
        fx = np.clip(np.minimum(x + d / 2, v[1]) - np.maximum(x - d / 2, v[0]), 0, None) / d
        fy = np.clip(np.minimum(y + d / 2, v[3]) - np.maximum(y - d / 2, v[2]), 0, None) / d
        fz = np.clip(np.minimum(z + d / 2, v[5]) - np.maximum(z - d / 2, v[4]), 0, None) / d
        want = 2.0 + 10.0 * fz[:, None, None] * fy[None, :, None] * fx[None, None, :]
        np.testing.assert_allclose(got, want, rtol=0, atol=1e-12)
    fd = DC(0.4, 1.0, 0.3, 0.8, 0.2, 0.5, d, d, d)
    hy = HybridMaterial3D(eps, ad, fd)
    box = (3, 36, 2, 28, 1, 18)
    got = hy.get_values(*box, sx=0.5, sz=-0.5)
    bg = eps.get_values(*box, sx=0.5, sz=-0.5)
    fv = ad.get_values(*box, sx=0.5, sz=-0.5, bg=bg)     # the function over the whole box
    k1, k2, j1, j2, i1, i2 = box
    for i in range(i1, i2):
        for j in range(j1, j2):
            for k in range(k1, k2):
                inside = fd.i1 <= i < fd.i2 and fd.j1 <= j < fd.j2 and fd.k1 <= k < fd.k2
                src = fv if inside else bg
                assert got[i - i1, j - j1, k - k1] == src[i - i1, j - j1, k - k1]


def test_sigmoid_maps():
    lb, ub, sq = 2.0, 12.0, 0.05
    x = np.linspace(lb, ub, 11)
    p = inverse_scaled_sigmoid(x, lb, ub, squish=sq)
    lo, hi = lb - sq * (ub - lb), ub + sq * (ub - lb)
    np.testing.assert_allclose(lo + (hi - lo) * sigmoid(p), x, rtol=1e-14)
    np.testing.assert_allclose(scaled_sigmoid(p, lb, ub), lb + (ub - lb) * sigmoid(p))
    with pytest.raises(ValueError):
        inverse_scaled_sigmoid(np.array([ub + 0.6 * (ub - lb)]), lb, ub, squish=sq)


def test_module_surface():
    """The names scripts reach through the package, as in emopt/__init__.py."""
    import emopt_amd as emopt
    assert emopt.experimental.grid.TopologyMaterial3D is TopologyMaterial3D
    assert emopt.experimental.adjoint_method.sigmoid is sigmoid
    assert emopt.experimental.autodiff_geometry.rect2d is adg.rect2d
    assert emopt.optimizer.Optimizer.__name__ == "Optimizer"
    assert emopt.misc.NOT_PARALLEL and emopt.misc.run_on_master(lambda: 3)() == 3
    assert emopt.fomutils.ModeMatch is emopt.adjoint.ModeMatch
    assert emopt.adjoint_method.AdjointMethodMO is emopt.adjoint.AdjointMethodMO


def test_remaining_shapes_by_identities():
    """rectnd over [z, y, x] equals rect3d; step1d reversed is the
    complement for the odd-symmetric sigmoid; step2d with normal (1, 0) is
    step1d along x broadcast over y; union_c / intersection_c of 0/1 masks
    with a steep ramp are the logical or / and."""
    d = 0.05
    x = torch.arange(0, 20, dtype=F64) * d
    y = torch.arange(0, 14, dtype=F64) * d
    z = torch.arange(0, 9, dtype=F64) * d
    v = torch.tensor([0.21, 0.74, 0.12, 0.55, 0.08, 0.33], dtype=F64)
    r3 = adg.rect3d(1 / d, x, y, z, v, nl=adg.nl_lin)
    rn = adg.rectnd(1 / d, [z, y, x], torch.stack([v[4:6], v[2:4], v[0:2]]), nl=adg.nl_lin)
    assert rn.shape == r3.shape == (9, 14, 20)
    np.testing.assert_allclose(rn.numpy(), r3.numpy(), rtol=0, atol=1e-15)
    s = adg.step1d(7.0, x, 0.4)
    np.testing.assert_allclose(adg.step1d(7.0, x, 0.4, reverse=True).numpy(), 1 - s.numpy(),
                               atol=1e-15)
    s2 = adg.step2d(7.0, x, y, torch.tensor([1.0, 0.0, 0.4, 0.3], dtype=F64))
    np.testing.assert_allclose(s2.numpy(), np.broadcast_to(s.numpy(), (14, 20)), atol=1e-15)
    rng = np.random.default_rng(9)
    m = [torch.tensor((rng.uniform(size=(4, 6)) > 0.5).astype(np.float64)) for _ in range(3)]
    any_ = np.logical_or.reduce([t.numpy() > 0.5 for t in m])
    all_ = np.logical_and.reduce([t.numpy() > 0.5 for t in m])
    np.testing.assert_allclose(adg.union_c(200.0, m).numpy(), any_.astype(float), atol=1e-40)
    np.testing.assert_allclose(adg.intersection_c(200.0, m).numpy(), all_.astype(float), atol=1e-40)


def test_polygon_set_points():
    from emopt_amd.grid import Polygon
    p = Polygon()
    p.set_points([0, 1, 1, 0], [0, 0, 1, 1])
    assert p.Np == 4 and p.contains_point(0.5, 0.5) and p.bbox() == (0.0, 1.0, 0.0, 1.0)
