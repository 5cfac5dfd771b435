"""Field post-processing on the host: interpolation of every component onto
the Ez grid and the Poynting source power.

Semantics of emopt/fdtd.py:1401-1553 (``get_field_interp``) and
emopt/fdtd.py:1555-1667 (``get_source_power``).  On a solver the averaging
runs on its GPU (emf_get_field_interp) for domains inside the grid, and only
the domain comes back; the source power is reduced there too
(emf_interp_flux), so only six scalars come back.  The numpy restatement
below serves other domains and solver stand-ins.
"""
from __future__ import annotations

import numpy as np

from .misc import DomainCoordinates

# neighbour offsets (di, dj, dk) averaged onto the Ez grid, per component;
# the stencils of emopt/fdtd.py:1481-1547 written as offset lists
_STENCIL = {
    "Ex": ((0, 0, 0), (0, 0, -1), (1, 0, 0), (1, 0, -1)),
    "Ey": ((0, 0, 0), (0, -1, 0), (1, 0, 0), (1, -1, 0)),
    "Hx": ((0, 0, 0), (0, -1, 0)),
    "Hy": ((0, 0, 0), (0, 0, -1)),
    "Hz": ((0, 0, 0), (0, 0, -1), (0, -1, 0), (0, -1, -1),
           (1, 0, 0), (1, 0, -1), (1, -1, 0), (1, -1, -1)),
}
# which min faces get the 'E'/'H' mirror before averaging (x: axis 2, y: axis 1)
_MIRROR = {"Ex": (2,), "Ey": (1,), "Hx": (1,), "Hy": (2,), "Hz": (2, 1)}
# region of the padded array that is updated (others keep the raw value):
# slices along (i, j, k) as (start, stop) offsets from (0, n)
_REGION = {
    "Ex": ((1, -1), (0, 0), (1, -1)),
    "Ey": ((1, -1), (1, -1), (0, 0)),
    "Hx": ((0, 0), (1, 0), (0, 0)),
    "Hy": ((0, 0), (0, 0), (1, 0)),
    "Hz": ((0, -1), (1, 0), (1, 0)),
}


def _avg(fr, comp):
    """Average `fr` (padded by one zero cell on every side) over the stencil
    of `comp`, inside the component's update region."""
    (i0, i1), (j0, j1), (k0, k1) = _REGION[comp]
    ni, nj, nk = fr.shape
    si, sj, sk = slice(i0, ni + i1), slice(j0, nj + j1), slice(k0, nk + k1)
    out = np.copy(fr)
    part = out[si, sj, sk]   # neighbours are added one by one, in stencil order
    for (di, dj, dk) in _STENCIL[comp][1:]:
        part += fr[slice(si.start + di, si.stop + di), slice(sj.start + dj, sj.stop + dj),
                   slice(sk.start + dk, sk.stop + dk)]
    return out / float(len(_STENCIL[comp]))


def grown_boxes(sim, domain):
    """The domain get_field_interp averages on and the box it reads
    (emopt/fdtd.py:1427-1457): (domain, dint, dom, ext, on_device) with dint
    the grown box as a DomainCoordinates, dom and ext as (k1, k2, j1, j2, i1,
    i2) -- ext is the domain grown by one cell at each min face above 0 and
    at each max face below N - 1, rounded through DomainCoordinates as the
    reference rounds it -- and on_device True when both lie in the grid (the
    device kernels' precondition)."""
    Nx, Ny, Nz = sim.Nx, sim.Ny, sim.Nz
    dx, dy, dz = sim.dx, sim.dy, sim.dz
    if domain is None:
        dint = DomainCoordinates(0, sim.X, 0, sim.Y, 0, sim.Z, dx, dy, dz)
        domain = dint
    else:
        k1, k2, j1, j2, i1, i2 = domain.k1, domain.k2, domain.j1, domain.j2, domain.i1, domain.i2
        if k1 > 0: k1 -= 1
        if k2 < Nx - 1: k2 += 1
        if j1 > 0: j1 -= 1
        if j2 < Ny - 1: j2 += 1
        if i1 > 0: i1 -= 1
        if i2 < Nz - 1: i2 += 1
        dint = DomainCoordinates(k1 * dx, k2 * dx, j1 * dy, j2 * dy, i1 * dz, i2 * dz, dx, dy, dz)
    ext = (dint.k1, dint.k2, dint.j1, dint.j2, dint.i1, dint.i2)
    dom = (domain.k1, domain.k2, domain.j1, domain.j2, domain.i1, domain.i2)
    k1, k2, j1, j2, i1, i2 = ext
    ok = (0 <= k1 <= dom[0] <= dom[1] <= k2 <= Nx and 0 <= j1 <= dom[2] <= dom[3] <= j2 <= Ny
          and 0 <= i1 <= dom[4] <= dom[5] <= i2 <= Nz)
    return domain, dint, dom, ext, ok


def interp_to_ez(sim, component, domain=None, squeeze=False, adjoint=False):
    get = sim.get_adjoint_field if adjoint else sim.get_field
    if component == "Ez":
        f = get(component, domain)
        return np.squeeze(f) if squeeze else f
    domain, dint, dom, ext, ok = grown_boxes(sim, domain)
    k1, k2, j1, j2, i1, i2 = ext
    dev = getattr(sim, "_interp_device", None)
    if dev is not None and ok:
        # on the solver's GPU: the same stencil sums on the same grown box
        f = dev(component, dom, ext, 1 if adjoint else 0)
        return np.squeeze(f) if squeeze else f
    fr = np.pad(get(component, dint), 1, "constant", constant_values=0)
    bc = sim.bc
    for ax in _MIRROR[component]:
        lo = k1 if ax == 2 else j1
        b = bc[0] if ax == 2 else bc[1]
        if lo == 0 and b in "EH":
            sgn = -1 if b == "E" else 1
            if ax == 2:
                fr[:, :, 0] = sgn * fr[:, :, 1]
            else:
                fr[:, 0, :] = sgn * fr[:, 1, :]
    f = _avg(fr, component)
    f = f[1 + domain.i1 - i1:-1 - (i2 - domain.i2),
          1 + domain.j1 - j1:-1 - (j2 - domain.j2),
          1 + domain.k1 - k1:-1 - (k2 - domain.k2)]
    return np.squeeze(f) if squeeze else f


# faces of the source-power box: (axis, min face?, sign, the reference's
# component pairs a conj(b) - c conj(d)) in the order of emopt/fdtd.py:1611-1665
_FACES = ((0, True, -1.0, ("Ey", "Hz", "Ez", "Hy")), (0, False, 1.0, ("Ey", "Hz", "Ez", "Hy")),
          (1, True, 1.0, ("Ex", "Hz", "Ez", "Hx")), (1, False, -1.0, ("Ex", "Hz", "Ez", "Hx")),
          (2, True, -1.0, ("Ex", "Hy", "Ey", "Hx")), (2, False, 1.0, ("Ex", "Hy", "Ey", "Hx")))


def source_power(sim):
    """Poynting flux out of the box one cell inside the CPML
    (emopt/fdtd.py:1555-1667), min faces with an 'E'/'H' symmetry skipped.
    On a solver each face's sum is reduced on its GPU (emf_interp_flux) and
    only a scalar comes back; otherwise the interpolated planes are summed
    here."""
    dx, dy, dz = sim.dx, sim.dy, sim.dz
    w = sim.w_pml
    X, Y, Z = sim.X, sim.Y, sim.Z
    xmin = w[0] + dx if w[0] > 0 else 0.0
    xmax = X - w[1] - dx if w[1] > 0 else X - dx
    ymin = w[2] + dy if w[2] > 0 else 0.0
    ymax = Y - w[3] - dy if w[3] > 0 else Y - dy
    zmin = w[4] + dz if w[4] > 0 else 0.0
    zmax = Z - w[5] - dz if w[5] > 0 else Z - dz
    D = lambda *a: DomainCoordinates(*a, dx, dy, dz)
    planes = {(0, True): D(xmin, xmin, ymin, ymax, zmin, zmax),
              (0, False): D(xmax, xmax, ymin, ymax, zmin, zmax),
              (1, True): D(xmin, xmax, ymin, ymin, zmin, zmax),
              (1, False): D(xmin, xmax, ymax, ymax, zmin, zmax),
              (2, True): D(xmin, xmax, ymin, ymax, zmin, zmin),
              (2, False): D(xmin, xmax, ymin, ymax, zmax, zmax)}
    area = (dy * dz, dx * dz, dx * dy)
    bc = sim.bc
    flux_dev = getattr(sim, "_flux_device", None)
    g = sim.get_field_interp
    re = np.real
    P = 0.0
    for axis, is_min, sign, (a, b, c, d) in _FACES:
        if is_min and bc[axis] in "EH":
            continue
        dom = planes[(axis, is_min)]
        if flux_dev is not None:
            _, _, db, ext, ok = grown_boxes(sim, dom)
            if ok:
                P += sign * 0.5 * area[axis] * flux_dev(axis, db, ext, 0)
                continue
        A, B, C, Dd = g(a, dom), g(b, dom), g(c, dom), g(d, dom)
        P += sign * 0.5 * area[axis] * np.sum(re(A * np.conj(B) - C * np.conj(Dd)))
    return P
