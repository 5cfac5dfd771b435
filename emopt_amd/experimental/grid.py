"""Materials of ``emopt.experimental.grid`` for the 3D solver
(emopt/experimental/grid.py:284-565).

* :class:`AutoDiffMaterial3D` -- values of a torch function of the design
  parameters at the (Yee-shifted) cell positions, so that the adjoint
  method can differentiate the material with torch.autograd.
* :class:`HybridMaterial3D` -- an ordinary material with a box in which an
  AutoDiffMaterial3D takes over.
* :class:`TopologyMaterial3D` -- an ordinary material with a box whose
  values are a free array (one value per cell, the same for the three
  staggered components), for topology optimisation.

The solver samples these through get_values like any other material
(FDTD.build/update run the function on the host and upload the arrays)."""
from __future__ import annotations

import numpy as np

from ..grid import Material3D

__all__ = ["AutoDiffMaterial3D", "HybridMaterial3D", "TopologyMaterial3D"]


def _torch():
    import torch
    return torch


def _overlap(fd, k1, k2, j1, j2, i1, i2):
    """The part of the index box (k1..i2) inside the domain fd: slices
    into the box and into the domain, or None when it is empty
    (emopt/experimental/grid.py:426-453, 520-563)."""
    lo = (max(fd.i1, i1), max(fd.j1, j1), max(fd.k1, k1))
    hi = (min(fd.i2, i2), min(fd.j2, j2), min(fd.k2, k2))
    if any(h < l for l, h in zip(lo, hi)):
        return None
    box = tuple(slice(l - o, h - o) for l, h, o in zip(lo, hi, (i1, j1, k1)))
    dom = tuple(slice(l - o, h - o) for l, h, o in zip(lo, hi, (fd.i1, fd.j1, fd.k1)))
    return box, dom, lo, hi


class AutoDiffMaterial3D(Material3D):
    """Material given by ``func(v, [z, y, x], bg=None)``: v the design
    parameters (a torch tensor), z, y, x 1-D torch tensors of coordinates;
    the result has shape (len(z), len(y), len(x))
    (emopt/experimental/grid.py:284-363).  Coordinates are float64 whatever
    torch's default dtype (the reference builds them in the default dtype,
    float32 unless a script changes it)."""

    def __init__(self, dx, dy, dz, func, v):
        self._dx, self._dy, self._dz = dx, dy, dz
        self._func = func
        self.v = v

    @property
    def v(self):
        return self._v

    @v.setter
    def v(self, new_v):
        self._v = _torch().as_tensor(new_v).squeeze()

    def _coords(self, k1, k2, j1, j2, i1, i2, sx, sy, sz):
        torch = _torch()
        f64 = torch.float64
        return [(torch.arange(i1, i2, dtype=f64) + sz) * self._dz,
                (torch.arange(j1, j2, dtype=f64) + sy) * self._dy,
                (torch.arange(k1, k2, dtype=f64) + sx) * self._dx]

    def get_value(self, x, y, z, bg=None):
        torch = _torch()
        c = [torch.tensor([z * self._dz], dtype=torch.float64),
             torch.tensor([y * self._dy], dtype=torch.float64),
             torch.tensor([x * self._dx], dtype=torch.float64)]
        return complex((self._func(self._v, c, bg=bg) + 0j).detach().reshape(-1)[0].item())

    def get_values(self, k1, k2, j1, j2, i1, i2, sx=0.0, sy=0.0, sz=0.0, arr=None, reshape=True,
                   bg=None):
        vals = self._func(self._v, self._coords(k1, k2, j1, j2, i1, i2, sx, sy, sz), bg=bg)
        vals = np.asarray((vals + 0j).detach().cpu().numpy(), dtype=np.complex128)
        vals = np.broadcast_to(vals, (i2 - i1, j2 - j1, k2 - k1))
        return self._out(np.array(vals), arr, reshape)


class HybridMaterial3D(Material3D):
    """``mats`` everywhere except inside the domain ``fdomain``, where the
    AutoDiffMaterial3D ``matf`` takes over; matf's function receives the
    background values there as ``bg`` (emopt/experimental/grid.py:365-459)."""

    def __init__(self, mats, matf, fdomain):
        self._mats = mats
        self._matf = matf
        self._fd = fdomain
        self._func = matf._func

    @property
    def v(self):
        return self._matf.v

    @v.setter
    def v(self, new_v):
        self._matf.v = new_v

    def get_values(self, k1, k2, j1, j2, i1, i2, sx=0.0, sy=0.0, sz=0.0, arr=None, reshape=True):
        vals = np.array(self._mats.get_values(k1, k2, j1, j2, i1, i2, sx=sx, sy=sy, sz=sz),
                        dtype=np.complex128).reshape(i2 - i1, j2 - j1, k2 - k1)
        ov = _overlap(self._fd, k1, k2, j1, j2, i1, i2)
        if ov is not None:
            box, _, lo, hi = ov
            if all(h > l for l, h in zip(lo, hi)):
                vals[box] = self._matf.get_values(lo[2], hi[2], lo[1], hi[1], lo[0], hi[0],
                                                  sx=sx, sy=sy, sz=sz, bg=vals[box].copy())
        return self._out(vals, arr, reshape)


class TopologyMaterial3D(Material3D):
    """``mats`` everywhere except inside ``fdomain``, whose cells take the
    values of the array ``grid`` (shape [nz, ny, nx] of the domain,
    initialised from mats); the three staggered eps components of a cell
    share its grid value (emopt/experimental/grid.py:461-565).  The grid
    keeps the domain's full 3-D shape (the reference squeezes it, which
    breaks its own indexing for a domain one cell thick)."""

    def __init__(self, mats, fdomain):
        self._mats = mats
        self._fd = fdomain
        self._grid = np.array(mats.get_values_in(fdomain), dtype=np.complex128).reshape(
            fdomain.Nz, fdomain.Ny, fdomain.Nx)

    @property
    def grid(self):
        return self._grid

    @grid.setter
    def grid(self, new_grid):
        self._grid = new_grid

    def get_values(self, k1, k2, j1, j2, i1, i2, sx=0.0, sy=0.0, sz=0.0, arr=None, reshape=True):
        vals = np.array(self._mats.get_values(k1, k2, j1, j2, i1, i2, sx=sx, sy=sy, sz=sz),
                        dtype=np.complex128).reshape(i2 - i1, j2 - j1, k2 - k1)
        ov = _overlap(self._fd, k1, k2, j1, j2, i1, i2)
        if ov is not None:
            box, dom, _, _ = ov
            vals[box] = np.reshape(self._grid, (self._fd.Nz, self._fd.Ny, self._fd.Nx))[dom]
        return self._out(vals, arr, reshape)
