"""Differentiable shapes for AutoDiffMaterial3D functions
(emopt/experimental/autodiff_geometry.py).

A shape is a torch tensor of values in [0, 1] over a grid given by 1-D
coordinate tensors (no meshgrid: the axes broadcast), 1 inside and 0
outside, with a transition of width ~1/k about the boundary shaped by a
nonlinearity ``nl(k, s)`` of the signed distance s (positive inside).
Combine shapes with the union / intersection operators and extrude 2-D
shapes along z with :func:`depth`; a material is then e.g.
``eps_lo + (eps_hi - eps_lo) * shape``.  Everything is plain torch, so
gradients with respect to the parameters v flow through.

For rectangles and cuboids the piecewise-linear ``nl_lin`` with k = 1/dx
gives each boundary cell its exact fill fraction along that axis.
"""
from __future__ import annotations

import math

import numpy as np
import torch

NONZERO = 1e-16
SQRT2 = np.sqrt(2.0)
ISQRT2 = np.sqrt(2.0) / 2.0

# ------------------------------------------------------------- nonlinearities
# Each maps k * s onto [0, 1], 1/2 at s = 0, increasing.


def nl_sig(k, x):
    """Logistic sigmoid of k x."""
    return torch.sigmoid(k * x)


def nl_erf(k, x):
    """(1 + erf(k x)) / 2."""
    return 0.5 * (1.0 + torch.erf(k * x))


def nl_lin(k, x):
    """k x + 1/2 clamped to [0, 1]: a linear ramp one 1/k wide."""
    return torch.clamp(k * x + 0.5, min=0.0, max=1.0)


def nl_sin(k, x):
    """(1 + sin(k x)) / 2 on |k x| <= pi/2, 0 below, 1 above."""
    kx = k * x
    mid = 0.5 * (torch.sin(kx) + 1.0)
    return torch.where(kx > np.pi / 2.0, torch.ones_like(mid),
                       torch.where(kx < -np.pi / 2.0, torch.zeros_like(mid), mid))


def nl_quad(k, x):
    """Two parabolic arcs meeting at k x = 0, reaching 0 and 1 at
    k x = -/+ 1/sqrt(2) with zero slope."""
    kx = k * x
    lo = (ISQRT2 + kx) ** 2
    hi = 1.0 - (ISQRT2 - kx) ** 2
    return torch.where(kx < -ISQRT2, torch.zeros_like(lo),
                       torch.where(kx < 0, lo, torch.where(kx > ISQRT2, torch.ones_like(hi), hi)))


# ------------------------------------------------------------- set operators

def union(list_shapes):
    """Sum of the shapes clamped to [0, 1]."""
    return torch.clamp(sum(list_shapes), min=0.0, max=1.0)


def union_b(list_shapes):
    """Probabilistic union: a + b - a b, folded over the list."""
    out = list_shapes[0]
    for s in list_shapes[1:]:
        out = s + out - s * out
    return out


def union_c(k, list_shapes, nl=nl_sig):
    """nl(k, sum - 1/2)."""
    return nl(k, sum(list_shapes) - 0.5)


def intersection(list_shapes):
    """Sum of the N shapes clamped to [N - 1, N], shifted down by N - 1."""
    n = len(list_shapes)
    return torch.clamp(sum(list_shapes), min=n - 1.0, max=float(n)) - (n - 1.0)


def intersection_b(list_shapes):
    """Product of the shapes."""
    return math.prod(list_shapes)


def intersection_c(k, list_shapes, nl=nl_sig):
    """nl(k, sum - (N - 1/2))."""
    return nl(k, sum(list_shapes) - (len(list_shapes) - 0.5))


# ---------------------------------------------------------------- shapes

def step1d(k, x, v, reverse=False, nl=nl_sig):
    """Step at x = v: 0 -> 1 with increasing x (1 -> 0 with reverse)."""
    return nl(k, v - x) if reverse else nl(k, x - v)


def rect1d(k, x, v, nl=nl_sig):
    """1 on v[0] < x < v[1]."""
    return nl(k, x - v[0]) * nl(k, v[1] - x)


def rect2d(k, x, y, v, nl=nl_sig):
    """Rectangle [v0, v1] x [v2, v3]; shape (len(y), len(x))."""
    return rect1d(k, x, v[0:2], nl=nl).reshape(1, -1) * rect1d(k, y, v[2:4], nl=nl).reshape(-1, 1)


def rect3d(k, x, y, z, v, nl=nl_sig):
    """Cuboid [v0, v1] x [v2, v3] x [v4, v5]; shape (len(z), len(y), len(x))."""
    return (rect1d(k, x, v[0:2], nl=nl).reshape(1, 1, -1)
            * rect1d(k, y, v[2:4], nl=nl).reshape(1, -1, 1)
            * rect1d(k, z, v[4:6], nl=nl).reshape(-1, 1, 1))


def rectnd(k, list_coords, v, nl=nl_sig):
    """Box over n axes: axis a spans [v[a, 0], v[a, 1]] of list_coords[a];
    with list_coords = [z, y, x] it equals rect3d."""
    n = len(list_coords)
    out = torch.ones([1] * n, dtype=torch.as_tensor(list_coords[0]).dtype)
    for a, r in enumerate(list_coords):
        view = [1] * n
        view[a] = -1
        out = out * rect1d(k, r, v[a, :], nl=nl).reshape(view)
    return out


def depth(shape, k, r, v, nl=nl_sig):
    """Extrude an n-D shape along a new leading axis with coordinates r
    over [v[0], v[1]]."""
    view = [-1] + [1] * shape.dim()
    return rect1d(k, r, v, nl=nl).reshape(view) * shape.unsqueeze(0)


def step2d(k, x, y, v, nl=nl_sig):
    """Half plane through (v2, v3) with normal (v0, v1) (not normalised:
    its length scales the transition); shape (len(y), len(x))."""
    return nl(k, (v[1] * (y - v[3])).reshape(-1, 1) + (v[0] * (x - v[2])).reshape(1, -1))


def poly2d(k, x, y, v, nl=nl_sig):
    """Convex polygon with clockwise vertices v (n x 2): the product of
    the half planes of its edges, each with the unit normal pointing
    inwards and anchored at the edge's midpoint."""
    edge = torch.roll(v, -1, dims=0) - v
    mid = v + 0.5 * edge
    nrm = torch.stack([edge[:, 1], -edge[:, 0]], dim=1)
    nrm = nrm / nrm.norm(dim=-1, keepdim=True)
    s = nrm[:, 0] * (x.reshape(1, -1, 1) - mid[:, 0]) + nrm[:, 1] * (y.reshape(-1, 1, 1) - mid[:, 1])
    return nl(k, s).prod(-1)


def circ2d(k, x, y, v, nl=nl_sig):
    """Disc of radius v0 centred at (v1, v2)."""
    r = torch.sqrt(NONZERO + ((x - v[1]) ** 2).reshape(1, -1) + ((y - v[2]) ** 2).reshape(-1, 1))
    return nl(k, v[0] - r)


def polar2d(k, x, y, v, order=4, nl=nl_sig):
    """Star-shaped region r(theta) < v0 (1 + v1 cos(order theta)) about
    (v3, v4), theta = v2 + atan2(x - v3, y - v4).  (The reference's body
    multiplies by the builtin ``ord`` instead of ``order``, which raises; the
    documented boundary is what is computed here.)"""
    dx = (x - v[3]).reshape(1, -1)
    dy = (y - v[4]).reshape(-1, 1)
    theta = v[2] + torch.atan2(dx, dy)
    r = torch.sqrt(NONZERO + dx ** 2 + dy ** 2)
    return nl(k, v[0] * (1.0 + v[1] * torch.cos(order * theta)) - r)
