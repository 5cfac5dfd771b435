"""Shared workload builders for the parity tests, smoke() and bench.py.

A case is a plain dict of inputs (grid, CPML, materials already sampled at
the Yee offsets, sources), so the GPU product path and the CPU oracle receive
identical arrays.
"""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

EPS_SI = 3.48 ** 2
EPS_OX = 1.444 ** 2
YEE = ((0.5, 0.0, -0.5), (0.0, 0.5, -0.5), (0.0, 0.0, 0.0),
       (0.0, 0.5, 0.0), (0.5, 0.0, 0.0), (0.5, 0.5, -0.5))


def rel_l2(a, b):
    a = np.asarray(a); b = np.asarray(b)
    nb = np.linalg.norm(b.ravel())
    return float(np.linalg.norm((a - b).ravel()) / (nb if nb > 0 else 1.0))


def rel_l2_set(got, ref):
    """Per-component L2 error of a field set, normalised by the largest
    component norm of `ref`.  Used for time-domain snapshots, where a
    component can legitimately be pure rounding noise (H after the first half
    step sees only the ramp envelope at t=0, an exact cancellation)."""
    scale = max(np.linalg.norm(np.asarray(r).ravel()) for r in ref)
    scale = scale if scale > 0 else 1.0
    return [float(np.linalg.norm((np.asarray(g) - np.asarray(r)).ravel()) / scale)
            for g, r in zip(got, ref)]


def span(N, d):
    """Physical size X with int(ceil(X/d)) + 1 == N (the reference's grid
    sizing, emopt/fdtd.py:252-254), nudged down when (N-1)*d rounds up."""
    X = (N - 1) * d
    while int(np.ceil(X / d) + 1) > N:
        X = np.nextafter(X, 0.0)
    assert int(np.ceil(X / d) + 1) == N
    return X


def sample(mat, Nx, Ny, Nz, q):
    sx, sy, sz = YEE[q]
    return mat.get_values(0, Nx, 0, Ny, 0, Nz, sx=sx, sy=sy, sz=sz)


def strip_case(Nx, Ny, Nz, w=10, d=0.04, wl=1.55, core=(0.5, 0.22), rtol=1e-4,
               ramp=5 * 2 * np.pi, bc="000", src_comp=2, sigma_y=0.5):
    """Straight strip waveguide along x (core eps 3.48^2 in cladding 1.444^2),
    centred in y and z, with a Gaussian line current along y at x index
    w + 5 and the core-centre z (the 3D strip-waveguide workload)."""
    from emopt_amd.grid import StructuredMaterial3D, Rectangle, ConstantMaterial3D
    X, Y, Z = span(Nx, d), span(Ny, d), span(Nz, d)
    eps = StructuredMaterial3D(X, Y, Z, d, d, d)
    clad = Rectangle(X / 2, Y / 2, 2 * X, 2 * Y, material_value=EPS_OX, layer=2)
    wg = Rectangle(X / 2, Y / 2, 2 * X, core[0], material_value=EPS_SI, layer=1)
    eps.add_primitive(clad, -Z, 2 * Z)
    eps.add_primitive(wg, Z / 2 - core[1] / 2, Z / 2 + core[1] / 2)
    mu = ConstantMaterial3D(1.0)
    mats = [sample(eps if q < 3 else mu, Nx, Ny, Nz, q) for q in range(6)]
    ks = min(w + 5, Nx - 1)
    iz = int(round((Z / 2) / d))
    y = np.arange(Ny) * d
    line = np.exp(-((y - Y / 2) / sigma_y) ** 2).astype(np.complex128)
    src = [np.zeros((1, Ny, 1), dtype=np.complex128) for _ in range(6)]
    src[src_comp][0, :, 0] = line
    return dict(Nx=Nx, Ny=Ny, Nz=Nz, d=d, wl=wl, X=X, Y=Y, Z=Z, w=[w] * 6, rtol=rtol,
                ramp=ramp, bc=bc, mats=mats, sources=[(src, (ks, ks + 1, 0, Ny, iz, iz + 1))])


def ellipse_design_eps(Nx, Ny, Nz, d, design=(240, 120, 10), n_ell=40, seed=2,
                       axes=(0.1, 0.6), overlap="supersample"):
    """Cladding (1.444^2) everywhere; a design slab of design[0] x design[1]
    x design[2] cells centred in the grid filled with n_ell seeded ellipses
    (semi-axes uniform in `axes` um, eps 3.48^2 inside), boundary cells
    area-averaged (by default supersampled, the rule these workloads were
    defined with; overlap="exact" gives the reference's rule)."""
    from emopt_amd.grid import StructuredMaterial3D, Rectangle, Ellipse
    X, Y, Z = span(Nx, d), span(Ny, d), span(Nz, d)
    rng = np.random.default_rng(seed)
    eps = StructuredMaterial3D(X, Y, Z, d, d, d, overlap=overlap)
    eps.add_primitive(Rectangle(X / 2, Y / 2, 2 * X, 2 * Y, material_value=EPS_OX, layer=10), -Z, 2 * Z)
    dxw, dyw, dzw = design[0] * d, design[1] * d, design[2] * d
    box = (X / 2 - dxw / 2, X / 2 + dxw / 2, Y / 2 - dyw / 2, Y / 2 + dyw / 2)
    z0 = Z / 2 - dzw / 2
    for _ in range(n_ell):
        cx = rng.uniform(box[0], box[1]); cy = rng.uniform(box[2], box[3])
        a, b = rng.uniform(*axes, size=2)
        ang = rng.uniform(0, np.pi)
        eps.add_primitive(Ellipse(cx, cy, a, b, ang, material_value=EPS_SI, layer=1, clip=box),
                          z0, z0 + dzw)
    return eps


def large_case(Nx=400, Ny=200, Nz=120, d=0.03, w=12, wl=1.31, seed=2, rtol=1e-4):
    """Single large FP64 solve: 400x200x120 at dx=0.03 um, 12-cell CPML,
    240x120x10 design slab of 40 seeded ellipses (128-vertex polygons under
    the reference's overlap rule), lambda 1.31 um, planar Ey source across
    the input face (x index w + 2, interior y-z)."""
    from emopt_amd.grid import ConstantMaterial3D
    X, Y, Z = span(Nx, d), span(Ny, d), span(Nz, d)
    eps = ellipse_design_eps(Nx, Ny, Nz, d, seed=seed, overlap="exact")
    mu = ConstantMaterial3D(1.0)
    mats = [sample(eps if q < 3 else mu, Nx, Ny, Nz, q) for q in range(6)]
    ks = w + 2
    J0, J1, I0, I1 = w, Ny - w, w, Nz - w
    src = [np.zeros((I1 - I0, J1 - J0, 1), dtype=np.complex128) for _ in range(6)]
    src[1][:] = 1.0
    return dict(Nx=Nx, Ny=Ny, Nz=Nz, d=d, wl=wl, X=X, Y=Y, Z=Z, w=[w] * 6, rtol=rtol,
                ramp=None, bc="000", mats=mats, sources=[(src, (ks, ks + 1, J0, J1, I0, I1))])


def random_case(Nx, Ny, Nz, w=(3, 3, 4, 3, 3, 4), d=0.05, wl=1.0, seed=0, ceps=False,
                mu_rand=True, bc="000", nsrc=1, ramp=5.0):
    """Random materials and random complex J/M blocks (step-level fixtures)."""
    rng = np.random.default_rng(seed)
    shp = (Nz, Ny, Nx)
    eps = [1 + 3 * rng.random(shp) + (0.2j * rng.random(shp) if ceps else 0) for _ in range(3)]
    mu = [1 + 0.5 * rng.random(shp) if mu_rand else np.ones(shp) for _ in range(3)]
    sources = []
    for q in range(nsrc):
        I = int(rng.integers(1, max(2, Nz // 3))); J = int(rng.integers(1, max(2, Ny // 3)))
        K = int(rng.integers(1, max(2, Nx // 3)))
        i0 = int(rng.integers(0, Nz - I + 1)); j0 = int(rng.integers(0, Ny - J + 1))
        k0 = int(rng.integers(0, Nx - K + 1))
        arrs = [rng.standard_normal((I, J, K)) + 1j * rng.standard_normal((I, J, K)) for _ in range(6)]
        sources.append((arrs, (k0, k0 + K, j0, j0 + J, i0, i0 + I)))
    X, Y, Z = span(Nx, d), span(Ny, d), span(Nz, d)
    return dict(Nx=Nx, Ny=Ny, Nz=Nz, d=d, wl=wl, X=X, Y=Y, Z=Z, w=list(w), rtol=1e-4,
                ramp=ramp, bc=bc, mats=[np.asarray(m, dtype=np.complex128) for m in eps + mu],
                sources=sources, ceps=ceps)


def _configure(sim, case, dom_cls):
    d = case["d"]
    sim.w_pml = [wc * d + 1e-9 * d for wc in case["w"]]
    sim.bc = case["bc"]
    if case.get("ramp") is not None:
        sim.src_ramp_time = case["ramp"]
    for arrs, (k1, k2, j1, j2, i1, i2) in case["sources"]:
        sim.set_sources(arrs, dom_cls.from_indices(k1, k2, j1, j2, i1, i2, *([d] * 3))
                        if dom_cls.__name__ == "DomainCoordinates" else
                        dom_cls.from_indices(k1, k2, j1, j2, i1, i2))


def run_gpu(case, fixed_n0=None, device=0):
    import warnings
    from emopt_amd.fdtd import FDTD
    from emopt_amd.misc import DomainCoordinates
    d = case["d"]
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        sim = FDTD(case["X"], case["Y"], case["Z"], d, d, d, case["wl"], rtol=case["rtol"],
                   nconv=case.get("nconv"), min_rindex=case.get("min_rindex", 1.0),
                   complex_eps=case.get("ceps", False), device=device)
        _configure(sim, case, DomainCoordinates)
    sim.set_material_arrays(case["mats"][:3], case["mats"][3:])
    sim.fixed_n0 = fixed_n0
    sim.verbose = 0
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        sim._register(sim._sources)
        sim._solve(0)
    full = DomainCoordinates.from_indices(0, case["Nx"], 0, case["Ny"], 0, case["Nz"], d, d, d)
    fields = [sim.get_field(c, full) for c in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")]
    out = dict(sim.last_solve)
    out["fields"] = fields
    out["history"] = sim.convergence_history
    out["sim"] = sim
    return out


def run_oracle(case, fixed_n0=None, lib="oracle", max_steps=None, adjoint=False, nmax=None):
    """The oracle's solve of `case`.  lib: "oracle" (serial C restatement),
    "omp" (the same built with OpenMP, bit-identical per cell) or "ref"
    (EMopt's own src/fdtd.cpp compiled from its sources).  adjoint=True
    solves with case["adj_sources"] through solve_adjoint."""
    import warnings
    from oracle import driver as D
    path = {"oracle": D.ORACLE_LIB, "omp": D.ORACLE_OMP_LIB, "ref": D.REF_LIB}[lib]
    d = case["d"]
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        o = D.OracleFDTD(D.load(path), case["X"], case["Y"], case["Z"], d, d, d, case["wl"],
                         rtol=case["rtol"], nconv=case.get("nconv"),
                         min_rindex=case.get("min_rindex", 1.0),
                         complex_eps=case.get("ceps", False))
        _configure(o, case, D.Domain)
        for arrs, (k1, k2, j1, j2, i1, i2) in case.get("adj_sources", []):
            o.add_adjoint_sources(arrs, D.Domain.from_indices(k1, k2, j1, j2, i1, i2))
        o.set_material_arrays(case["mats"][:3], case["mats"][3:])
        if nmax is not None:
            o.Nmax = nmax
        if adjoint:
            o.solve_adjoint(fixed_n0=fixed_n0, max_steps=max_steps)
        else:
            o.solve_forward(fixed_n0=fixed_n0, max_steps=max_steps)
    return dict(n0=o.n0, Tn=o.Tn, fields=o.phasors(adjoint), history=np.array(o.history),
                seconds=o.solve_seconds, status=o.status, sim=o)


def batch_case(q, seed=3, dims=((64, 256), (64, 160), (48, 96)), w_range=(8, 12), d=0.04,
               rtol=1e-4, n_fwd=16):
    """Solve q of the batched workload: grid dims uniform in `dims`, lambda
    uniform in 1.50-1.60 um, CPML width uniform in w_range cells, a random
    slab (eps in [1.444^2, 3.48^2]) holding 4-12 random ellipses; solves
    q < n_fwd carry a forward Jz dipole at a random interior point, the others
    an adjoint source: a complex Gaussian times a random phase on an x
    monitor plane."""
    from emopt_amd.grid import StructuredMaterial3D, Rectangle, Ellipse, ConstantMaterial3D
    rng = np.random.default_rng([seed, q])
    Nx, Ny, Nz = (int(rng.integers(lo, hi + 1)) for lo, hi in dims)
    wl = float(rng.uniform(1.50, 1.60))
    w = int(rng.integers(w_range[0], w_range[1] + 1))
    X, Y, Z = span(Nx, d), span(Ny, d), span(Nz, d)
    eps = StructuredMaterial3D(X, Y, Z, d, d, d, overlap="supersample")
    eps.add_primitive(Rectangle(X / 2, Y / 2, 2 * X, 2 * Y, material_value=EPS_OX, layer=10),
                      -Z, 2 * Z)
    t = float(rng.uniform(0.15, 0.35))
    zc = Z / 2
    slab_eps = float(rng.uniform(EPS_OX, EPS_SI))
    eps.add_primitive(Rectangle(X / 2, Y / 2, 2 * X, 2 * Y, material_value=slab_eps, layer=5),
                      zc - t / 2, zc + t / 2)
    for _ in range(int(rng.integers(4, 13))):
        cx, cy = rng.uniform(0.2 * X, 0.8 * X), rng.uniform(0.2 * Y, 0.8 * Y)
        a, b = rng.uniform(0.1, 0.5, size=2)
        eps.add_primitive(Ellipse(cx, cy, a, b, rng.uniform(0, np.pi),
                                  material_value=float(rng.uniform(EPS_OX, EPS_SI)), layer=1),
                          zc - t / 2, zc + t / 2)
    mu = ConstantMaterial3D(1.0)
    mats = [sample(eps if c < 3 else mu, Nx, Ny, Nz, c) for c in range(6)]
    adjoint = q >= n_fwd
    src = None
    if not adjoint:
        k = int(rng.integers(w + 2, Nx - w - 2)); j = int(rng.integers(w + 2, Ny - w - 2))
        i = int(round(zc / d))
        arrs = [np.zeros((1, 1, 1), dtype=np.complex128) for _ in range(6)]
        arrs[2][0, 0, 0] = 1.0
        src = (arrs, (k, k + 1, j, j + 1, i, i + 1))
    else:
        k = int(rng.integers(w + 2, Nx - w - 2))
        J0, J1, I0, I1 = w + 1, Ny - w - 1, w + 1, Nz - w - 1
        y = (np.arange(J0, J1) * d)[None, :]
        z = (np.arange(I0, I1) * d)[:, None]
        g = np.exp(-((y - Y / 2) ** 2 + (z - zc) ** 2) / (2 * 0.6 ** 2))
        ph = np.exp(1j * rng.uniform(0, 2 * np.pi))
        arrs = [np.zeros((I1 - I0, J1 - J0, 1), dtype=np.complex128) for _ in range(6)]
        arrs[2][:, :, 0] = g * ph
        arrs[4][:, :, 0] = 0.5 * g * ph * 1j
        src = (arrs, (k, k + 1, J0, J1, I0, I1))
    return dict(Nx=Nx, Ny=Ny, Nz=Nz, d=d, wl=wl, X=X, Y=Y, Z=Z, w=[w] * 6, rtol=rtol,
                ramp=None, bc="000", mats=mats, sources=[] if adjoint else [src],
                adj_sources=[src] if adjoint else [], adjoint=adjoint)


def build_sim(case, device=0):
    """FDTD object for `case` with materials uploaded and (adjoint) sources set."""
    import warnings
    from emopt_amd.fdtd import FDTD
    from emopt_amd.misc import DomainCoordinates
    d = case["d"]
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        sim = FDTD(case["X"], case["Y"], case["Z"], d, d, d, case["wl"], rtol=case["rtol"],
                   nconv=case.get("nconv"), min_rindex=case.get("min_rindex", 1.0),
                   complex_eps=case.get("ceps", False), device=device)
        _configure(sim, case, DomainCoordinates)
    for arrs, (k1, k2, j1, j2, i1, i2) in case.get("adj_sources", []):
        sim.add_adjoint_sources(arrs, DomainCoordinates.from_indices(k1, k2, j1, j2, i1, i2, d, d, d))
    sim.set_material_arrays(case["mats"][:3], case["mats"][3:])
    sim.verbose = 0
    return sim


def case_from_sim(sim, rtol):
    """The oracle's inputs for a configured solver: the material arrays it
    holds on the device, its sources, CPML, bc and ramp."""
    shp = (sim.Nz, sim.Ny, sim.Nx)
    A = sim.get_A_diag()
    srcs = [([a.reshape(s.I, s.J, s.K) for a in s.arrays],
             (s.k0, s.k0 + s.K, s.j0, s.j0 + s.J, s.i0, s.i0 + s.I)) for s in sim._sources]
    d = sim.dx
    return dict(Nx=sim.Nx, Ny=sim.Ny, Nz=sim.Nz, d=d, wl=sim.wavelength,
                X=span(sim.Nx, d), Y=span(sim.Ny, d), Z=span(sim.Nz, d), w=list(sim._w_cells),
                rtol=rtol, ramp=sim.src_ramp_time, bc="".join(sim.bc),
                mats=[a.reshape(shp) for a in A], sources=srcs, nconv=sim._nconv,
                min_rindex=sim._min_rindex)


def all_fields(sim, case, adjoint=False):
    from emopt_amd.misc import DomainCoordinates
    d = case["d"]
    full = DomainCoordinates.from_indices(0, case["Nx"], 0, case["Ny"], 0, case["Nz"], d, d, d)
    get = sim.get_adjoint_field if adjoint else sim.get_field
    return [get(c, full) for c in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")]


# ------------------------------------------------------- grating coupler

def _edge_rect_cls():
    from emopt_amd.grid import Rectangle

    class EdgeRect(Rectangle):
        """Axis-aligned rectangle given by its x edges (xl, xr) and y span."""

        def __init__(self, xl, xr, y0, yspan, material_value=1.0, layer=1):
            super().__init__(0.5 * (xl + xr), y0, xr - xl, yspan, material_value, layer)
            self.xl, self.xr = xl, xr

        def bbox(self):
            return (self.xl, self.xr, self.y0 - self.height / 2, self.y0 + self.height / 2)

    return EdgeRect


def grating_problem(Nx=360, Ny=60, Nz=120, d=0.03, ntr=32, w=10, wl=1.55, seed=4,
                    w0=10.4, theta_deg=8.0, rtol=1e-4, t_si=0.22, t_box=2.0, etch=0.11,
                    pitch=(0.3, 0.7)):
    """Grating-coupler-like slab: Si (t_si) on oxide (t_box above the bottom
    CPML), air above; ntr etched trenches (depth `etch`) whose 2*ntr edge
    positions are the design parameters, pitches drawn uniform in `pitch`
    (seed) and scaled together so the grating fits the window between 0.3 um
    past the x CPML on both sides, duty cycle 0.5.  A Jy line current across
    the slab at x index w + 3 launches the slab's TE wave; the FOM is the
    power-normalised overlap, on a z plane 0.4 um above the slab, with a
    Gaussian beam of waist w0 tilted by theta in the x-z plane, E along y.
    Returns (sim, am, params)."""
    import warnings
    from emopt_amd.adjoint import AdjointMethodPNF3D, ModeMatch, rect_edge_deps
    from emopt_amd.shape import boundary_deps, rect_edge_velocity
    from emopt_amd.fdtd import FDTD
    from emopt_amd.grid import StructuredMaterial3D, Rectangle, ConstantMaterial3D
    from emopt_amd.misc import DomainCoordinates, gaussian_mode
    EdgeRect = _edge_rect_cls()
    X, Y, Z = span(Nx, d), span(Ny, d), span(Nz, d)
    wd = w * d
    z_si0 = wd + t_box
    z_si1 = z_si0 + t_si
    z_mon = z_si1 + 0.4
    if z_mon > Z - wd - 2 * d:
        raise ValueError("grid too short for the monitor plane")
    eps = StructuredMaterial3D(X, Y, Z, d, d, d, supersample=4)
    eps.add_primitive(Rectangle(X / 2, Y / 2, 4 * X, 4 * Y, material_value=EPS_OX, layer=3), -Z, z_si0)
    eps.add_primitive(Rectangle(X / 2, Y / 2, 4 * X, 4 * Y, material_value=EPS_SI, layer=2), z_si0, z_si1)
    rng = np.random.default_rng(seed)
    pit = rng.uniform(pitch[0], pitch[1], size=ntr)
    x_start, x_end = wd + 0.3 + 4 * d, X - wd - 0.3
    pit *= (x_end - x_start) / pit.sum()
    edges = []
    x = x_start
    trenches = []
    for q in range(ntr):
        xl, xr = x + 0.25 * pit[q], x + 0.75 * pit[q]
        edges += [xl, xr]
        r = EdgeRect(xl, xr, Y / 2, 4 * Y, material_value=1.0, layer=1)
        eps.add_primitive(r, z_si1 - etch, z_si1)
        trenches.append(r)
        x += pit[q]
    params = np.array(edges)
    mu = ConstantMaterial3D(1.0)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        sim = FDTD(X, Y, Z, d, d, d, wl, rtol=rtol)
        sim.w_pml = [wd + 1e-9 * d] * 6
    sim.set_materials(eps, mu)
    # source: Jy across the slab (Gaussian in z), all y
    ks = w + 3
    i0, i1 = int(z_si0 / d) - 3, int(z_si1 / d) + 4
    z = np.arange(i0, i1) * d
    prof = np.exp(-((z - 0.5 * (z_si0 + z_si1)) / 0.12) ** 2)
    arrs = [np.zeros((i1 - i0, Ny, 1), dtype=np.complex128) for _ in range(6)]
    arrs[1][:, :, 0] = prof[:, None]
    sim.set_sources(arrs, DomainCoordinates.from_indices(ks, ks + 1, 0, Ny, i0, i1, d, d, d))
    # FOM plane and target beam
    mon = DomainCoordinates(wd + d, X - wd - d, wd + d, Y - wd - d, z_mon, z_mon, d, d, d)
    xs = mon.x - 0.5 * (x_start + x_end)
    ys = mon.y - Y / 2
    XX, YY = np.meshgrid(xs, ys)
    Eg, Hxg, Hyg = gaussian_mode(XX, YY, w0, theta_deg * np.pi / 180, wl, 1.0)
    mm = ModeMatch([0, 0, 1], d, d, Eym=Eg[None], Hxm=-Hxg[None], Hzm=-Hyg[None])
    e_i0, e_i1 = int((z_si1 - etch) / d) - 1, int(z_si1 / d) + 2

    class GratingAM(AdjointMethodPNF3D):
        def update_system(self, p):
            for q, r in enumerate(trenches):
                r.xl, r.xr = float(p[2 * q]), float(p[2 * q + 1])
                r.x0, r.width = 0.5 * (r.xl + r.xr), r.xr - r.xl

        def get_update_boxes(self, sim, p):
            return [(p[i] - 2 * d, p[i] + 2 * d, 0, Y, z_si1 - etch - 2 * d, z_si1 + 2 * d)
                    for i in range(len(p))]

        def get_fom_domains(self):
            return [mon]

        def calc_f(self, sim, p):
            mm.compute_device(sim, mon)          # overlap sums reduced on the GPU
            return mm.get_mode_match_forward(1.0)

        def calc_dfdx(self, sim, p):
            mm.compute_device(sim, mon)
            zero = np.zeros(mon.shape, dtype=np.complex128)
            return [(mm.get_dFdEx(), mm.get_dFdEy(), zero, mm.get_dFdHx(), mm.get_dFdHy(), zero)]

        def calc_dfdx_device(self, sim, p):
            # the same planes as calc_dfdx, left on the GPU (Ez and Hz zero)
            mm.compute_device(sim, mon)
            return [mm.dFdx_device(sim, 0, mask=0b011011)]

        def calc_deps(self, sim, p):
            """d eps / d edge for every edge: analytic from the overlap rule
            (deps="analytic", the default), or central differences of the
            device sampling (deps="difference")."""
            out = [[], [], [], []]
            for i in range(len(p)):
                r = trenches[i // 2]
                edge = "xl" if i % 2 == 0 else "xr"
                if self.deps == "analytic":
                    comp, cell, dv = boundary_deps(sim, eps, r, rect_edge_velocity(edge))
                else:
                    k = int(p[i] / d)
                    comp, cell, dv = rect_edge_deps(sim, eps, r, edge, 1e-6 * d,
                                                    (max(k - 2, 0), min(k + 3, sim.Nx), 0,
                                                     sim.Ny, e_i0, e_i1))
                out[0].append(np.full(len(comp), i, np.int32))
                out[1].append(comp); out[2].append(cell); out[3].append(dv)
            return tuple(np.concatenate(a) for a in out)

    am = GratingAM(sim, step=1e-8)
    am.deps = "analytic"
    am.update_system(params)
    return sim, am, params


def serve_case(q, seed=5, lo=(64, 64, 64), hi=(320, 240, 120), n_ell=12, w=10, rtol=1e-4,
               materials=True, s=None):
    """Request q of the serving workload: grid dims log-uniform between `lo`
    and `hi` (one shared log factor), d uniform in 0.03-0.05 um, lambda
    uniform in 1.26-1.62 um, a slab of random ellipses (eps 3.48^2 in
    1.444^2) across the middle, CPML `w` cells, planar Ey source across the
    input face as in the single large workload.  materials=False skips the
    (host-heavy) material sampling; s in [0, 1] overrides the drawn log
    factor (0 gives `lo`, 1 gives `hi`)."""
    from emopt_amd.grid import ConstantMaterial3D
    rng = np.random.default_rng([seed, q])
    s_drawn = rng.uniform()
    s = s_drawn if s is None else float(s)
    Nx, Ny, Nz = (int(round(a * (b / a) ** s)) for a, b in zip(lo, hi))
    d = float(rng.uniform(0.03, 0.05))
    wl = float(rng.uniform(1.26, 1.62))
    eseed = int(rng.integers(1 << 30))
    X, Y, Z = span(Nx, d), span(Ny, d), span(Nz, d)
    ks = w + 2
    J0, J1, I0, I1 = w, Ny - w, w, Nz - w
    src = [np.zeros((I1 - I0, J1 - J0, 1), dtype=np.complex128) for _ in range(6)]
    src[1][:] = 1.0
    mats = None
    if materials:
        design = (max(Nx - 2 * w - 8, 4), max(Ny - 2 * w - 8, 4), min(10, max(Nz // 6, 2)))
        eps = ellipse_design_eps(Nx, Ny, Nz, d, design=design, n_ell=n_ell, seed=eseed,
                                 axes=(0.1, 0.6))
        mu = ConstantMaterial3D(1.0)
        mats = [sample(eps if c < 3 else mu, Nx, Ny, Nz, c) for c in range(6)]
    return dict(Nx=Nx, Ny=Ny, Nz=Nz, d=d, wl=wl, X=X, Y=Y, Z=Z, w=[w] * 6, rtol=rtol,
                ramp=None, bc="000", mats=mats, sources=[(src, (ks, ks + 1, J0, J1, I0, I1))],
                adjoint=False)


def serve_request(q, **kw):
    """A serving Request for serve_case(q): dims (for the admission estimate)
    now, the problem itself built on the server's worker thread."""
    from emopt_amd.serve import Request, state_bytes
    c = serve_case(q, materials=False, **kw)

    def build():
        return build_sim(serve_case(q, **kw))
    req = Request(q, build, state_bytes(c["Nx"], c["Ny"], c["Nz"], c["w"], wavelength=c["wl"],
                                        d=c["d"]), adjoint=False,
                  cells=c["Nx"] * c["Ny"] * c["Nz"],
                  meta=dict(dims=(c["Nx"], c["Ny"], c["Nz"]), wl=c["wl"], d=c["d"]))
    return req
