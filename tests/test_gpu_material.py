"""Structured material sampling on the GPU (emf_sample_structured) against
the host sampler of emopt_amd/grid.py, which restates the reference's
StructuredMaterial3D (src/Grid.cpp:736-874, emopt/grid.py:874-961), under
both in-plane rules: the reference's exact one (corner-gated primitives
taking exact areas, src/Grid.cpp:460-525) and the supersampled one.  The
device kernels repeat the host arithmetic operation for operation, so the
comparison is exact (np.array_equal) at all six Yee offsets, on whole grids
and boxes: rectangles with edges on and off cell boundaries, rotated and
clipped ellipses, convex and concave polygons, overlapping layers and z
ranges, complex values and background."""
import time

import numpy as np
import pytest

from emopt_amd.grid import Ellipse, Polygon, Rectangle, StructuredMaterial3D
from tests.cases import ellipse_design_eps, span

pytestmark = pytest.mark.gpu

SHIFTS = ((0.5, 0.0, -0.5), (0.0, 0.5, -0.5), (0.0, 0.0, 0.0),
          (0.0, 0.5, 0.0), (0.5, 0.0, 0.0), (0.5, 0.5, -0.5))


def _random_structure(seed, N=(40, 34, 26), d=0.05, overlap="supersample"):
    rng = np.random.default_rng(seed)
    X, Y, Z = (n * d for n in N)
    m = StructuredMaterial3D(X, Y, Z, d, d, d, supersample=int(rng.integers(3, 9)),
                             background=1.0 + 0.05j * rng.random(), overlap=overlap,
                             ellipse_vertices=int(rng.integers(12, 200)))
    for q in range(14):
        kind = q % 3
        val = complex(1 + 11 * rng.random(), 0.3 * rng.random() * (q % 2))
        layer = int(rng.integers(0, 4))
        z1 = float(rng.uniform(-0.2, Z * 0.8))
        z2 = z1 + float(rng.uniform(0.02, Z * 0.6))
        if q == 0:      # edges exactly on cell boundaries
            prim = Rectangle(10 * d, 8 * d, 6 * d, 4 * d, material_value=val, layer=layer)
            z1, z2 = 4 * d, 15 * d
        elif kind == 0:
            prim = Rectangle(rng.uniform(0, X), rng.uniform(0, Y), rng.uniform(0.05, X / 2),
                             rng.uniform(0.05, Y / 2), material_value=val, layer=layer)
        elif kind == 1:
            clip = None
            if rng.random() < 0.5:
                x0, y0 = rng.uniform(0, X / 2), rng.uniform(0, Y / 2)
                clip = (x0, x0 + rng.uniform(0.2, X / 2), y0, y0 + rng.uniform(0.2, Y / 2))
            prim = Ellipse(rng.uniform(0, X), rng.uniform(0, Y), rng.uniform(0.05, 0.6),
                           rng.uniform(0.05, 0.6), rng.uniform(0, np.pi), material_value=val,
                           layer=layer, clip=clip)
        else:
            nv = int(rng.integers(3, 9))
            ang = np.sort(rng.uniform(0, 2 * np.pi, nv))
            r = rng.uniform(0.1, 0.7, nv)           # star-shaped, often concave
            cx, cy = rng.uniform(0, X), rng.uniform(0, Y)
            prim = Polygon(cx + r * np.cos(ang), cy + r * np.sin(ang), material_value=val,
                           layer=layer)
        m.add_primitive(prim, z1, z2)
    return m, N


@pytest.mark.parametrize("overlap", ["exact", "supersample"])
@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_device_sampler_matches_host(seed, overlap):
    m, (Nx, Ny, Nz) = _random_structure(seed, overlap=overlap)
    boxes = [(0, Nx, 0, Ny, 0, Nz), (3, 29, 5, 21, 2, 17), (0, 1, 0, Ny, 7, 8)]
    for (k1, k2, j1, j2, i1, i2) in boxes:
        for (sx, sy, sz) in SHIFTS:
            h = m.get_values(k1, k2, j1, j2, i1, i2, sx=sx, sy=sy, sz=sz)
            g = m.get_values_device(k1, k2, j1, j2, i1, i2, sx=sx, sy=sy, sz=sz)
            assert g.shape == h.shape
            assert np.array_equal(g, h), (seed, (k1, k2, j1, j2, i1, i2), (sx, sy, sz),
                                          np.abs(g - h).max())
            if k2 - k1 == Nx:
                assert len(np.unique(h)) > 3     # the structure shows up in the box


def test_device_sampler_outside_all_primitives():
    m = StructuredMaterial3D(1.0, 1.0, 1.0, 0.1, 0.1, 0.1, background=2.25)
    m.add_primitive(Rectangle(0.5, 0.5, 0.2, 0.2, material_value=4.0), 5.0, 6.0)  # above the grid
    g = m.get_values_device(0, 10, 0, 10, 0, 10)
    assert np.array_equal(g, m.get_values(0, 10, 0, 10, 0, 10))
    assert np.all(g == 2.25)


def test_device_sampler_exact_rule_limits():
    """A cell where ten partial rectangles count: the device reports the
    limit as an error, as the host sampler does, and an ellipse is refused by
    the exact rule's entry point unless given as a polygon."""
    from emopt_amd import _lib
    m = StructuredMaterial3D(2.0, 2.0, 1.0, 0.1, 0.1, 0.1)
    for q in range(10):
        m.add_primitive(Rectangle(1.0, 1.0, 1.0 - 0.01 * q, 1.0 + 0.013 * q, material_value=2.0,
                                  layer=q), 0.0, 1.0)
    with pytest.raises(ValueError, match="more than 8 primitives"):
        m.get_values(0, 20, 0, 20, 0, 1, sx=0.3, sy=0.3, sz=0.5)
    with pytest.raises(_lib.EmfError, match="more than 8 primitives"):
        m.get_values_device(0, 20, 0, 20, 0, 1, sx=0.3, sy=0.3, sz=0.5)
    recs, vx, vy, nv = m.device_primitives()
    recs["kind"][0] = 1
    out = np.empty((1, 20, 20), dtype=np.complex128)
    with pytest.raises(_lib.EmfError, match="ellipses as polygons"):
        _lib.load().emf_sample_structured(0, recs.ctypes.data, len(recs), vx.ctypes.data,
                                          vy.ctypes.data, nv, 0.1, 0.1, 0.1, 0, 1.0, 0.0,
                                          0, 20, 0, 20, 0, 1, 0.0, 0.0, 0.0, out.ctypes.data)


def test_device_sampler_large_design_region():
    """The 400x200x120 workload's material (40 ellipses as 128-gons in a
    240x120x10 design slab over cladding, the reference's overlap rule):
    identical values at one offset; prints the host and device times."""
    Nx, Ny, Nz, d = 400, 200, 120, 0.03
    eps = ellipse_design_eps(Nx, Ny, Nz, d, seed=2, overlap="exact")
    t = time.perf_counter()
    h = eps.get_values(0, Nx, 0, Ny, 0, Nz, sx=0.5, sy=0.0, sz=-0.5)
    th = time.perf_counter() - t
    eps.get_values_device(0, 8, 0, 8, 0, 8)   # warm-up
    t = time.perf_counter()
    g = eps.get_values_device(0, Nx, 0, Ny, 0, Nz, sx=0.5, sy=0.0, sz=-0.5)
    tg = time.perf_counter() - t
    print("eps_x of 400x200x120: host %.3f s, device %.3f s (incl. copy to host)" % (th, tg))
    assert np.array_equal(g, h)


def test_fdtd_build_uses_device_sampler():
    """FDTD.build samples a structured eps on the GPU; the uploaded arrays
    equal the host sampler's."""
    import warnings
    from emopt_amd.fdtd import FDTD
    from emopt_amd.grid import ConstantMaterial3D
    Nx, Ny, Nz, d = 40, 30, 24, 0.05
    eps = ellipse_design_eps(Nx, Ny, Nz, d, design=(24, 16, 6), n_ell=6, seed=4, overlap="exact")
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        sim = FDTD(span(Nx, d), span(Ny, d), span(Nz, d), d, d, d, 1.55)
    sim.set_materials(eps, ConstantMaterial3D(1.0))
    sim.build()
    diag = sim.get_A_diag()
    for c, (sx, sy, sz) in enumerate(SHIFTS[:3]):
        host = eps.get_values(0, Nx, 0, Ny, 0, Nz, sx=sx, sy=sy, sz=sz).ravel()
        assert np.array_equal(diag[c], host), c
    calls = []
    orig = StructuredMaterial3D.get_values_device

    def spy(self, *a, **k):
        calls.append(1)
        return orig(self, *a, **k)
    StructuredMaterial3D.get_values_device = spy
    try:
        sim.build()
    finally:
        StructuredMaterial3D.get_values_device = orig
    assert len(calls) == 3
