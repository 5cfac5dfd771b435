"""Material distributions for the solver (host side, numpy).

The solver only needs an object with ``get_values(k1, k2, j1, j2, i1, i2, sx,
sy, sz, arr)`` returning complex values on the (shifted) Yee grid, the
interface of ``emopt.grid.Material3D`` (emopt/grid.py:723-845).  This module
provides:

* :class:`ConstantMaterial3D` (emopt/grid.py:846-872)
* :class:`StructuredMaterial3D` -- z-stacked primitives with layer priority
  and cell averaging (emopt/grid.py:874-961, src/Grid.cpp:460-525,736-874)
* primitives :class:`Rectangle`, :class:`Polygon`, :class:`Ellipse`
* :class:`FunctionMaterial3D` for analytic / array-backed distributions.

StructuredMaterial3D also samples on the GPU (``get_values_device``, the
emf_sample_structured kernels), with the arithmetic below operation for
operation; FDTD.build/update use that path for it.

Cell averaging: a Yee cell value is the sum over z-slabs of the stack of
(slab fraction of the cell's z extent) x (in-plane average), background 1.0
outside the stack, exactly as src/Grid.cpp:736-824 combines layers.  In-plane,
primitives are visited in priority order (lower ``layer`` first); each takes
its overlap with the part of the cell not yet covered, the remainder is
background 1.0 (src/Grid.cpp:460-525).  Rectangle overlaps are exact; polygon
and ellipse overlaps are supersampled (``supersample`` points per axis), since
the reference's polygon clipping depends on boost.geometry, which this
project does not use.
"""
from __future__ import annotations

import numpy as np


class Material3D:
    def get_value(self, x, y, z):
        v = self.get_values(0, 1, 0, 1, 0, 1, sx=x, sy=y, sz=z)
        return v.ravel()[0]

    def get_values(self, k1, k2, j1, j2, i1, i2, sx=0.0, sy=0.0, sz=0.0, arr=None, reshape=True):
        raise NotImplementedError

    def get_values_in(self, domain, sx=0.0, sy=0.0, sz=0.0, arr=None, squeeze=False):
        v = self.get_values(domain.k1, domain.k2, domain.j1, domain.j2, domain.i1, domain.i2,
                            sx, sy, sz, arr)
        return np.squeeze(v) if squeeze else v

    @staticmethod
    def _out(vals, arr, reshape):
        if arr is not None:
            flat = np.ravel(arr)
            flat[:] = vals.ravel()
            vals = flat if not reshape else flat.reshape(vals.shape)
        return vals if reshape else vals.ravel()


class ConstantMaterial3D(Material3D):
    def __init__(self, value):
        self._value = complex(value)

    @property
    def material_value(self):
        return self._value

    @material_value.setter
    def material_value(self, v):
        self._value = complex(v)

    def get_values(self, k1, k2, j1, j2, i1, i2, sx=0.0, sy=0.0, sz=0.0, arr=None, reshape=True):
        vals = np.full((i2 - i1, j2 - j1, k2 - k1), self._value, dtype=np.complex128)
        return self._out(vals, arr, reshape)


class FunctionMaterial3D(Material3D):
    """Values from ``func(x, y, z)`` evaluated at the Yee-shifted cell
    positions (x = (k + sx) * dx etc.), broadcast over [Nz, Ny, Nx]."""

    def __init__(self, func, dx, dy, dz):
        self._f = func
        self._d = (dx, dy, dz)

    def get_values(self, k1, k2, j1, j2, i1, i2, sx=0.0, sy=0.0, sz=0.0, arr=None, reshape=True):
        dx, dy, dz = self._d
        x = (np.arange(k1, k2) + sx) * dx
        y = (np.arange(j1, j2) + sy) * dy
        z = (np.arange(i1, i2) + sz) * dz
        Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
        vals = np.asarray(self._f(X, Y, Z), dtype=np.complex128)
        vals = np.broadcast_to(vals, (i2 - i1, j2 - j1, k2 - k1)).copy()
        return self._out(vals, arr, reshape)


# ------------------------------------------------------------- primitives

class MaterialPrimitive:
    def __init__(self, material_value=1.0, layer=1):
        self.material_value = complex(material_value)
        self.layer = layer

    def set_material(self, mat):
        self.material_value = complex(mat)

    def coverage(self, xc, yc, dx, dy, ss):
        """Area fraction of each cell centred at (yc[:,None], xc[None,:]) of
        size dx*dy covered by the primitive, shape [len(yc), len(xc)]."""
        raise NotImplementedError

    def bbox(self):
        raise NotImplementedError


class Rectangle(MaterialPrimitive):
    """Axis-aligned rectangle centred at (x0, y0) (emopt/grid.py:367-501)."""

    def __init__(self, x0, y0, xspan, yspan, material_value=1.0, layer=1):
        super().__init__(material_value, layer)
        self.x0, self.y0, self.width, self.height = x0, y0, xspan, yspan

    def bbox(self):
        return (self.x0 - self.width / 2, self.x0 + self.width / 2,
                self.y0 - self.height / 2, self.y0 + self.height / 2)

    def coverage(self, xc, yc, dx, dy, ss):
        x0, x1, y0, y1 = self.bbox()
        fx = np.clip(np.minimum(xc + dx / 2, x1) - np.maximum(xc - dx / 2, x0), 0, None) / dx
        fy = np.clip(np.minimum(yc + dy / 2, y1) - np.maximum(yc - dy / 2, y0), 0, None) / dy
        return fy[:, None] * fx[None, :]


class _Sampled(MaterialPrimitive):
    def inside(self, x, y):
        raise NotImplementedError

    def coverage(self, xc, yc, dx, dy, ss):
        out = np.zeros((len(yc), len(xc)))
        xb0, xb1, yb0, yb1 = self.bbox()
        kx = np.nonzero((xc + dx / 2 > xb0) & (xc - dx / 2 < xb1))[0]
        ky = np.nonzero((yc + dy / 2 > yb0) & (yc - dy / 2 < yb1))[0]
        if len(kx) == 0 or len(ky) == 0:
            return out
        off = (np.arange(ss) + 0.5) / ss - 0.5
        xs = (xc[kx][:, None] + off[None, :] * dx).ravel()
        ys = (yc[ky][:, None] + off[None, :] * dy).ravel()
        Y, X = np.meshgrid(ys, xs, indexing="ij")
        ins = self.inside(X, Y).astype(np.float64)
        ins = ins.reshape(len(ky), ss, len(kx), ss).mean(axis=(1, 3))
        out[np.ix_(ky, kx)] = ins
        return out


class Ellipse(_Sampled):
    """Ellipse centred at (x0, y0) with semi-axes (a, b), rotated by angle,
    optionally clipped to the box clip = (xmin, xmax, ymin, ymax)."""

    def __init__(self, x0, y0, a, b, angle=0.0, material_value=1.0, layer=1, clip=None):
        super().__init__(material_value, layer)
        self.x0, self.y0, self.a, self.b, self.angle = x0, y0, a, b, angle
        self.clip = clip

    def bbox(self):
        r = max(self.a, self.b)
        x0, x1, y0, y1 = self.x0 - r, self.x0 + r, self.y0 - r, self.y0 + r
        if self.clip is not None:
            x0, x1 = max(x0, self.clip[0]), min(x1, self.clip[1])
            y0, y1 = max(y0, self.clip[2]), min(y1, self.clip[3])
        return x0, x1, y0, y1

    def inside(self, x, y):
        c, s = np.cos(self.angle), np.sin(self.angle)
        u = (x - self.x0) * c + (y - self.y0) * s
        v = -(x - self.x0) * s + (y - self.y0) * c
        ins = (u / self.a) ** 2 + (v / self.b) ** 2 <= 1.0
        if self.clip is not None:
            ins &= (x >= self.clip[0]) & (x < self.clip[1]) & (y >= self.clip[2]) & (y < self.clip[3])
        return ins


class Polygon(_Sampled):
    """Simple polygon from vertex lists (emopt/grid.py:503-651); even-odd rule."""

    def __init__(self, xs=None, ys=None, material_value=1.0, layer=1):
        super().__init__(material_value, layer)
        self.xs = np.asarray(xs if xs is not None else [], dtype=np.double)
        self.ys = np.asarray(ys if ys is not None else [], dtype=np.double)

    def set_points(self, xs, ys):
        self.xs = np.asarray(xs, dtype=np.double)
        self.ys = np.asarray(ys, dtype=np.double)

    def bbox(self):
        return self.xs.min(), self.xs.max(), self.ys.min(), self.ys.max()

    def inside(self, x, y):
        res = np.zeros(x.shape, dtype=bool)
        xs, ys = self.xs, self.ys
        n = len(xs)
        for a in range(n):
            b = (a + 1) % n
            xa, ya, xb, yb = xs[a], ys[a], xs[b], ys[b]
            if ya == yb:
                continue
            cond = (ya > y) != (yb > y)
            xint = xa + (y - ya) * (xb - xa) / (yb - ya)
            res ^= cond & (x < xint)
        return res


class StructuredMaterial3D(Material3D):
    """Primitives extruded over [z1, z2) ranges (emopt/grid.py:874-961)."""

    def __init__(self, X, Y, Z, dx, dy, dz, supersample=8, background=1.0):
        self._X, self._Y, self._Z = X, Y, Z
        self._dx, self._dy, self._dz = dx, dy, dz
        self.supersample = supersample
        self.background = complex(background)
        self._prims = []    # (prim, z1, z2, insertion order)

    @property
    def primitives(self):
        return [p for p, _, _, _ in self._prims]

    def add_primitive(self, prim, z1, z2):
        if z2 <= z1:
            return
        self._prims.append((prim, float(z1), float(z2), len(self._prims)))

    def add_primitives(self, prims, z1, z2):
        for p in prims:
            self.add_primitive(p, z1, z2)

    def _inplane(self, active, xc, yc):
        dx, dy = self._dx, self._dy
        remaining = np.ones((len(yc), len(xc)))
        val = np.zeros((len(yc), len(xc)), dtype=np.complex128)
        for prim in sorted(active, key=lambda p: (p[0].layer, p[3])):
            cov = np.minimum(prim[0].coverage(xc, yc, dx, dy, self.supersample), remaining)
            val += cov * prim[0].material_value
            remaining -= cov
        val += remaining * self.background
        return val

    # emf_prim of include/emfdtd.h
    PRIM_DTYPE = np.dtype([("kind", "<i4"), ("clip", "<i4"), ("v0", "<i4"), ("nv", "<i4"),
                           ("re", "<f8"), ("im", "<f8"), ("z1", "<f8"), ("z2", "<f8"),
                           ("bbox", "<f8", (4,)), ("q", "<f8", (10,))])

    def device_primitives(self):
        """The primitives as emf_prim records in priority order (layer, then
        insertion), with the polygon vertex arrays."""
        order = sorted(self._prims, key=lambda p: (p[0].layer, p[3]))
        recs = np.zeros(len(order), dtype=self.PRIM_DTYPE)
        vx, vy = [], []
        nv = 0
        for r, (prim, z1, z2, _) in zip(recs, order):
            r["re"], r["im"] = prim.material_value.real, prim.material_value.imag
            r["z1"], r["z2"] = z1, z2
            r["bbox"] = prim.bbox()
            if isinstance(prim, Rectangle):
                r["kind"] = 0
            elif isinstance(prim, Ellipse):
                r["kind"] = 1
                r["q"][:6] = (prim.x0, prim.y0, prim.a, prim.b, np.cos(prim.angle), np.sin(prim.angle))
                if prim.clip is not None:
                    r["clip"] = 1
                    r["q"][6:] = prim.clip
            elif isinstance(prim, Polygon):
                r["kind"] = 2
                r["v0"], r["nv"] = nv, len(prim.xs)
                vx.append(prim.xs)
                vy.append(prim.ys)
                nv += len(prim.xs)
            else:
                raise TypeError("no device sampler for %s" % type(prim).__name__)
        vx = np.ascontiguousarray(np.concatenate(vx) if vx else np.zeros(1), dtype=np.double)
        vy = np.ascontiguousarray(np.concatenate(vy) if vy else np.zeros(1), dtype=np.double)
        return recs, vx, vy, nv

    def get_values_device(self, k1, k2, j1, j2, i1, i2, sx=0.0, sy=0.0, sz=0.0, device=0,
                          arr=None, reshape=True):
        """get_values computed on GPU `device` (emf_sample_structured): the
        same arithmetic in the same order, so the same doubles."""
        from . import _lib
        lib = _lib.load()
        recs, vx, vy, nv = self.device_primitives()
        out = np.empty((i2 - i1, j2 - j1, k2 - k1), dtype=np.complex128)
        lib.emf_sample_structured(int(device), recs.ctypes.data, len(recs), vx.ctypes.data,
                                  vy.ctypes.data, nv, self._dx, self._dy, self._dz,
                                  int(self.supersample), self.background.real, self.background.imag,
                                  k1, k2, j1, j2, i1, i2, sx, sy, sz, out.ctypes.data)
        return self._out(out, arr, reshape)

    def get_values(self, k1, k2, j1, j2, i1, i2, sx=0.0, sy=0.0, sz=0.0, arr=None, reshape=True):
        dx, dy, dz = self._dx, self._dy, self._dz
        xc = (np.arange(k1, k2) + sx) * dx
        yc = (np.arange(j1, j2) + sy) * dy
        zc = (np.arange(i1, i2) + sz) * dz
        zlo, zhi = zc - dz / 2, zc + dz / 2
        out = np.zeros((i2 - i1, j2 - j1, k2 - k1), dtype=np.complex128)
        covered = np.zeros(i2 - i1)
        bounds = sorted({b for _, z1, z2, _ in self._prims for b in (z1, z2)})
        for za, zb in zip(bounds[:-1], bounds[1:]):
            fz = np.clip(np.minimum(zhi, zb) - np.maximum(zlo, za), 0, None) / dz
            if not np.any(fz > 0):
                continue
            active = [p for p in self._prims if p[1] <= za and p[2] >= zb]
            v2 = self._inplane(active, xc, yc)
            nz = np.nonzero(fz > 0)[0]
            out[nz] += fz[nz, None, None] * v2[None, :, :]
            covered += fz
        out += (1.0 - covered)[:, None, None] * self.background
        return self._out(out, arr, reshape)
