"""Index-box helper with the semantics of ``emopt.misc.DomainCoordinates``
(emopt/misc.py:169-490): real-space bounds map to grid indices by
``i1 = int(zmin/dz)``, ``i2 = int(zmax/dz) + 1`` (likewise j, k), so a domain
whose min equals its max is a single plane."""
from __future__ import annotations

import numpy as np

# One process drives one GPU: it is the reference's master rank
# (emopt/misc.py:22-25 with a single MPI process).
RANK = 0
N_PROC = 1
NOT_PARALLEL = True


def run_on_master(func):
    """Decorator of emopt/misc.py:27-38: the function runs on the master
    rank only, which is this process."""
    import functools

    @functools.wraps(func)
    def wrapper(*args, **kwargs):
        return func(*args, **kwargs) if NOT_PARALLEL else None

    return wrapper


class DomainCoordinates:
    def __init__(self, xmin, xmax, ymin, ymax, zmin, zmax, dx, dy, dz):
        i1 = int(zmin / dz); i2 = int(zmax / dz) + 1
        j1 = int(ymin / dy); j2 = int(ymax / dy) + 1
        k1 = int(xmin / dx); k2 = int(xmax / dx) + 1
        self.set_region(k1, k2, j1, j2, i1, i2, dx, dy, dz)

    @classmethod
    def from_indices(cls, k1, k2, j1, j2, i1, i2, dx, dy, dz):
        d = cls.__new__(cls)
        d.set_region(k1, k2, j1, j2, i1, i2, dx, dy, dz)
        return d

    def set_region(self, k1, k2, j1, j2, i1, i2, dx, dy, dz):
        self.k1, self.k2, self.j1, self.j2, self.i1, self.i2 = k1, k2, j1, j2, i1, i2
        self._dx, self._dy, self._dz = dx, dy, dz
        self._x = dx * np.arange(k1, k2, dtype=np.int64).astype(np.double)
        self._y = dy * np.arange(j1, j2, dtype=np.int64).astype(np.double)
        self._z = dz * np.arange(i1, i2, dtype=np.int64).astype(np.double)

    x = property(lambda self: self._x)
    y = property(lambda self: self._y)
    z = property(lambda self: self._z)
    i = property(lambda self: slice(self.i1, self.i2))
    j = property(lambda self: slice(self.j1, self.j2))
    k = property(lambda self: slice(self.k1, self.k2))
    Nx = property(lambda self: self.k2 - self.k1)
    Ny = property(lambda self: self.j2 - self.j1)
    Nz = property(lambda self: self.i2 - self.i1)
    dx = property(lambda self: self._dx)
    dy = property(lambda self: self._dy)
    dz = property(lambda self: self._dz)
    xspan = property(lambda self: self._dx * (self.Nx - 1))
    yspan = property(lambda self: self._dy * (self.Ny - 1))
    zspan = property(lambda self: self._dz * (self.Nz - 1))

    @property
    def shape(self):
        return (self.Nz, self.Ny, self.Nx)

    def copy(self):
        return DomainCoordinates.from_indices(self.k1, self.k2, self.j1, self.j2, self.i1,
                                              self.i2, self._dx, self._dy, self._dz)

    def __repr__(self):
        return "DomainCoordinates(k=[%d,%d) j=[%d,%d) i=[%d,%d))" % (
            self.k1, self.k2, self.j1, self.j2, self.i1, self.i2)


def gaussian_fields(x, z, prop_dist, w0, theta, wavelength, n0):
    """Fields of a Gaussian beam in the x-z plane, beam axis at angle theta in
    the x-y plane (emopt/misc.py:580-636).  Returns (Ez, Hx, Hy),
    non-dimensionalised like the solver (H scaled by n0)."""
    xg = np.asarray(x * np.cos(theta))
    yg = np.asarray(x * np.sin(theta))
    zg = np.asarray(z)
    r2 = xg ** 2 + zg ** 2
    k0 = 2 * np.pi * n0 / wavelength
    y_r = np.pi * w0 ** 2 / wavelength
    w_y = w0 * np.sqrt(1 + (yg / y_r) ** 2)
    inv_r = yg / (yg ** 2 + y_r ** 2)
    gouy = np.nan_to_num(np.arctan(yg / y_r))
    Ez = w0 / w_y * np.exp(-r2 / w_y ** 2) * np.exp(1j * (k0 * yg + k0 * inv_r * r2 / 2.0 - gouy))
    return Ez, Ez * np.cos(theta) * n0, -1 * Ez * np.sin(theta) * n0


def gaussian_mode(x, z, w0, theta, wavelength, n0):
    """Single-wavevector Gaussian 'fibre' mode, E along z, unit amplitude
    (emopt/misc.py:638-687).  Returns (Ez, Hx, Hy)."""
    xg = np.asarray(x * np.cos(theta))
    yg = np.asarray(x * np.sin(theta))
    zg = np.asarray(z)
    k0 = 2 * np.pi * n0 / wavelength
    Ez = np.exp(-(xg ** 2 + zg ** 2) / w0 ** 2) * np.exp(1j * k0 * yg)
    Hx = (n0 * np.cos(theta) + 2 * xg * np.sin(theta) / (1j * w0 ** 2)) * Ez
    Hy = (2 * xg * np.cos(theta) / (1j * w0 ** 2) - n0 * np.sin(theta)) * Ez
    return Ez, Hx, Hy
