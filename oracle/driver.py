"""ORACLE — TEST INFRASTRUCTURE ONLY.

Single-rank restatement of EMopt's Python CW-FDTD driver (``emopt/fdtd.py``)
that drives any shared library exporting the reference's ``FDTD_*`` C ABI
(``src/fdtd.hpp:571-643``): either this directory's C restatement
(``oracle/libcfdtd.so``) or the reference kernel compiled from its own sources
(``oracle/_ref/libfdtd_ref.so``).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg
may import this module; the product package never does.

Restated pieces and the lines they follow:

* grid sizing, dt, CPML defaults, source ramp, Nmax, convergence sample
  indices ........................................ emopt/fdtd.py:236-437
* pml width / bc / ramp-time setters ............. emopt/fdtd.py:505-616
* ``build`` with the six Yee-shifted samplings .... emopt/fdtd.py:842-880
* source cropping and registration ............... emopt/fdtd.py:658-747
* ``__solve`` convergence loop, capture, phasor .. emopt/fdtd.py:994-1194
* ``solve_forward`` / ``solve_adjoint`` .......... emopt/fdtd.py:1196-1300
* field gather / interpolation / source power .... emopt/fdtd.py:1322-1667
* ``get_A_diag`` / ``calc_ydAx`` .................. emopt/fdtd.py:1669-1717
* GhostComm on one rank = periodic copy of the six faces
  ................................................. emopt/fdtd.py:1733-1873,
                                                    src/fdtd.cpp:1789-1929
"""
from __future__ import annotations

import ctypes
import os
import time
import warnings
from math import pi

import numpy as np
from numpy.ctypeslib import ndpointer

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "libcfdtd.so")
ORACLE_OMP_LIB = os.path.join(HERE, "libcfdtd_omp.so")
REF_LIB = os.path.join(HERE, "_ref", "libfdtd_ref.so")

_c_complex_p = ndpointer(np.complex128, ndim=1, flags="C")
_c_double_p = ndpointer(np.double, ndim=1, flags="C")


def load(path: str):
    """Load a library exporting the FDTD_* ABI and declare its argtypes
    (mirrors emopt/fdtd_ctypes.py:26-149)."""
    lib = ctypes.CDLL(path)
    vp, d, i, b = ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_bool
    sig = {
        "FDTD_new": ([], vp),
        "FDTD_set_wavelength": ([vp, d], None),
        "FDTD_set_physical_dims": ([vp, d, d, d, d, d, d], None),
        "FDTD_set_grid_dims": ([vp, i, i, i], None),
        "FDTD_set_local_grid": ([vp, i, i, i, i, i, i], None),
        "FDTD_set_dt": ([vp, d], None),
        "FDTD_set_complex_eps": ([vp, b], None),
        "FDTD_set_field_arrays": ([vp] + [_c_double_p] * 6, None),
        "FDTD_set_mat_arrays": ([vp] + [_c_complex_p] * 6, None),
        "FDTD_update_H": ([vp, i, d], None),
        "FDTD_update_E": ([vp, i, d], None),
        "FDTD_set_pml_widths": ([vp, i, i, i, i, i, i], None),
        "FDTD_set_pml_properties": ([vp, d, d, d, d], None),
        "FDTD_build_pml": ([vp], None),
        "FDTD_reset_pml": ([vp], None),
        "FDTD_set_t0_arrays": ([vp] + [_c_complex_p] * 6, None),
        "FDTD_set_t1_arrays": ([vp] + [_c_complex_p] * 6, None),
        "FDTD_calc_phase_2T": ([d, d, d, d], d),
        "FDTD_calc_amplitude_2T": ([d, d, d, d, d], d),
        "FDTD_calc_phase_3T": ([d, d, d, d, d, d], d),
        "FDTD_calc_amplitude_3T": ([d, d, d, d, d, d, d], d),
        "FDTD_capture_t0_fields": ([vp], None),
        "FDTD_capture_t1_fields": ([vp], None),
        "FDTD_calc_complex_fields_2T": ([vp, d, d], None),
        "FDTD_calc_complex_fields_3T": ([vp, d, d, d], None),
        "FDTD_add_source": ([vp] + [_c_complex_p] * 6 + [i] * 6 + [b], None),
        "FDTD_clear_sources": ([vp], None),
        "FDTD_set_source_properties": ([vp, d, d], None),
        "FDTD_src_func_t": ([vp, i, d, d], d),
        "FDTD_set_bc": ([vp, ctypes.c_char_p], None),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    return lib


class Domain:
    """Index box restating emopt.misc.DomainCoordinates (emopt/misc.py:212-279)."""

    def __init__(self, xmin, xmax, ymin, ymax, zmin, zmax, dx, dy, dz):
        self.i1 = int(zmin / dz); self.i2 = int(zmax / dz) + 1
        self.j1 = int(ymin / dy); self.j2 = int(ymax / dy) + 1
        self.k1 = int(xmin / dx); self.k2 = int(xmax / dx) + 1

    @classmethod
    def from_indices(cls, k1, k2, j1, j2, i1, i2):
        d = cls.__new__(cls)
        d.k1, d.k2, d.j1, d.j2, d.i1, d.i2 = k1, k2, j1, j2, i1, i2
        return d

    @property
    def i(self):
        return slice(self.i1, self.i2)

    @property
    def j(self):
        return slice(self.j1, self.j2)

    @property
    def k(self):
        return slice(self.k1, self.k2)


class OracleFDTD:
    """One-rank restatement of ``emopt.fdtd.FDTD`` over an FDTD_* library."""

    def __init__(self, lib, X, Y, Z, dx, dy, dz, wavelength, rtol=1e-6, nconv=None,
                 min_rindex=1.0, complex_eps=False):
        self.lib = lib
        nproc = 1
        if nconv is None:
            nconv = nproc * 10
        self._dx, self._dy, self._dz = dx, dy, dz
        Nx = int(np.ceil(X / dx) + 1); Ny = int(np.ceil(Y / dy) + 1); Nz = int(np.ceil(Z / dz) + 1)
        self.Nx, self.Ny, self.Nz = Nx, Ny, Nz
        self._X = dx * (Nx - 1); self._Y = dy * (Ny - 1); self._Z = dz * (Nz - 1)
        self._wavelength = wavelength
        self._R = wavelength / (2 * pi)
        self._Sc = 0.95
        self._min_rindex = min_rindex
        self.dt = self._Sc * np.min([dx, dy, dz]) / self._R / np.sqrt(3) * min_rindex
        I, J, K = Nz, Ny, Nx
        self.I, self.J, self.K = I, J, K
        npad = (K + 2) * (J + 2) * (I + 2)
        self.fields = [np.zeros(npad) for _ in range(6)]   # Ex Ey Ez Hx Hy Hz
        self.mats = [np.zeros(I * J * K, dtype=np.complex128) for _ in range(6)]
        self.fwd = [[np.zeros(I * J * K, dtype=np.complex128) for _ in range(6)] for _ in range(2)]
        self.adj = [[np.zeros(I * J * K, dtype=np.complex128) for _ in range(6)] for _ in range(2)]

        h = lib.FDTD_new()
        self.h = h
        lib.FDTD_set_wavelength(h, wavelength)
        lib.FDTD_set_physical_dims(h, X, Y, Z, dx, dy, dz)
        lib.FDTD_set_grid_dims(h, Nx, Ny, Nz)
        lib.FDTD_set_local_grid(h, 0, 0, 0, K, J, I)
        lib.FDTD_set_dt(h, self.dt)
        lib.FDTD_set_field_arrays(h, *self.fields)
        lib.FDTD_set_mat_arrays(h, *self.mats)
        lib.FDTD_set_complex_eps(h, complex_eps)

        w = 15
        self._w_pml = [w * dx, w * dx, w * dy, w * dy, w * dz, w * dz]
        self.w_cells = [w] * 6
        self._pml = (3.0, 0.0, 1.0, 3.0)
        lib.FDTD_set_pml_widths(h, *self.w_cells)
        lib.FDTD_set_pml_properties(h, *self._pml)
        lib.FDTD_build_pml(h)

        Nlambda = wavelength / np.min([dx, dy, dz]) / min_rindex
        Ncycle = Nlambda * np.sqrt(3)
        self.Nlambda, self.Ncycle = Nlambda, Ncycle
        self._src_T = Ncycle * 20.0 * self.dt
        self._src_min = 1e-4
        self.Nmax = Nlambda * 500
        lib.FDTD_set_source_properties(h, self._src_T, self._src_min)
        self.rtol = rtol
        n_pts = int(nconv / nproc)
        self.conv_pts = np.arange(0, I * J * K, int(I * J * K / n_pts))
        self.nconv = len(self.conv_pts)
        self.sources = []
        self.adj_sources = []
        self._bc = "000"
        lib.FDTD_set_bc(h, b"000")
        self.history = []
        self.n0 = None

    # ---------------------------------------------------------- properties
    @property
    def bc(self):
        return self._bc

    @bc.setter
    def bc(self, newbc):
        if len(newbc) != 3:
            raise ValueError("Incorrect number of boundary conditions specified!")
        for c in newbc:
            if c not in "0EHP":
                raise ValueError("Unrecognized boundary condition: %s" % c)
            if c == "P":
                raise NotImplementedError("Periodic boundary conditions not yet implemented")
        self._bc = newbc
        self.lib.FDTD_set_bc(self.h, newbc.encode("ascii"))

    @property
    def w_pml(self):
        return self._w_pml

    @w_pml.setter
    def w_pml(self, w):
        if len(w) != 6:
            raise ValueError("Incorrect number of pml widths specified.")
        d = [self._dx, self._dx, self._dy, self._dy, self._dz, self._dz]
        self.w_cells = [int(w[q] / d[q]) for q in range(6)]
        self._w_pml = [self.w_cells[q] * d[q] for q in range(6)]
        self.lib.FDTD_set_pml_widths(self.h, *self.w_cells)
        self.lib.FDTD_build_pml(self.h)

    def set_pml_properties(self, sigma, alpha, kappa, pw):
        self._pml = (sigma, alpha, kappa, pw)
        self.lib.FDTD_set_pml_properties(self.h, sigma, alpha, kappa, pw)

    @property
    def src_ramp_time(self):
        return self._src_T

    @src_ramp_time.setter
    def src_ramp_time(self, T):
        if T <= 0:
            raise ValueError("Source ramp time must be greater than zero.")
        self._src_T = T
        self.lib.FDTD_set_source_properties(self.h, self._src_T, self._src_min)

    @property
    def src_min_value(self):
        return self._src_min

    @src_min_value.setter
    def src_min_value(self, m):
        if m <= 0:
            raise ValueError("Minimum source value must be greater than zero.")
        self._src_min = m
        self.lib.FDTD_set_source_properties(self.h, self._src_T, self._src_min)

    # ------------------------------------------------------------ materials
    def set_material_arrays(self, eps_xyz, mu_xyz):
        """Install the six Yee-sampled material arrays ([Nz,Ny,Nx] complex)."""
        for q, a in enumerate(list(eps_xyz) + list(mu_xyz)):
            self.mats[q][:] = np.asarray(a, dtype=np.complex128).ravel()

    def build(self, eps, mu):
        """Sample two Material3D-like objects at the Yee offsets of
        emopt/fdtd.py:858-880 (sx, sy, sz per component)."""
        K, J, I = self.K, self.J, self.I
        shifts = [(0.5, 0.0, -0.5), (0.0, 0.5, -0.5), (0.0, 0.0, 0.0),
                  (0.0, 0.5, 0.0), (0.5, 0.0, 0.0), (0.5, 0.5, -0.5)]
        for q, (sx, sy, sz) in enumerate(shifts):
            m = eps if q < 3 else mu
            self.mats[q][:] = np.ravel(m.get_values(0, K, 0, J, 0, I, sx=sx, sy=sy, sz=sz))

    # -------------------------------------------------------------- sources
    def _overlap(self, dom):
        I, J, K = self.I, self.J, self.K
        if 0 >= dom.i2 or I <= dom.i1 or 0 >= dom.j2 or J <= dom.j1 or 0 >= dom.k2 or K <= dom.k1:
            return None
        imax = min(I, dom.i2); jmax = min(J, dom.j2); kmax = min(K, dom.k2)
        imin = max(0, dom.i1); jmin = max(0, dom.j1); kmin = max(0, dom.k1)
        return (imin, jmin, kmin), (imin - dom.i1, jmin - dom.j1, kmin - dom.k1), \
               (imax - imin, jmax - jmin, kmax - kmin)

    def _set_sources(self, src, dom, adjoint=False):
        ov = self._overlap(dom)
        if ov is None:
            return
        (i0, j0, k0), (id1, jd1, kd1), (sI, sJ, sK) = ov
        sl = (slice(id1, id1 + sI), slice(jd1, jd1 + sJ), slice(kd1, kd1 + sK))
        arrs = [np.ascontiguousarray(np.asarray(a, dtype=np.complex128)[sl]).ravel().copy() for a in src]
        rec = (arrs, i0, j0, k0, sI, sJ, sK)
        (self.adj_sources if adjoint else self.sources).append(rec)
        self.lib.FDTD_add_source(self.h, *arrs, i0, j0, k0, sI, sJ, sK, True)

    def set_sources(self, src, dom):
        self._set_sources(src, dom, adjoint=False)

    def add_adjoint_sources(self, src, dom):
        self._set_sources(src, dom, adjoint=True)

    def set_adjoint_sources(self, src):
        self.adj_sources = []
        for dFdx, dom in zip(src[0], src[1]):
            self.add_adjoint_sources(dFdx, dom)

    def clear_sources(self):
        self.sources = []

    # ---------------------------------------------------------- time loop
    def ghost_exchange(self, arr):
        """One-rank GhostComm: every ghost face takes the opposite face's
        outermost interior layer (emopt/fdtd.py:1795-1842)."""
        I, J, K = self.I, self.J, self.K
        a = arr.reshape(I + 2, J + 2, K + 2)
        it, jt, kt = slice(1, I + 1), slice(1, J + 1), slice(1, K + 1)
        xlo = a[it, jt, K].copy(); xhi = a[it, jt, 1].copy()
        ylo = a[it, J, kt].copy(); yhi = a[it, 1, kt].copy()
        zlo = a[I, jt, kt].copy(); zhi = a[1, jt, kt].copy()
        a[it, jt, 0] = xlo; a[it, jt, K + 1] = xhi
        a[it, 0, kt] = ylo; a[it, J + 1, kt] = yhi
        a[0, jt, kt] = zlo; a[I + 1, jt, kt] = zhi

    def _step(self, n):
        lib, h, dt = self.lib, self.h, self.dt
        lib.FDTD_update_H(h, n, n * dt)
        for q in (3, 4, 5):
            self.ghost_exchange(self.fields[q])
        lib.FDTD_update_E(h, n, (n + 0.5) * dt)
        for q in (0, 1, 2):
            self.ghost_exchange(self.fields[q])

    def _solve(self, fixed_n0=None, max_steps=None):
        """Convergence loop of emopt/fdtd.py:994-1194.  ``fixed_n0`` replaces
        the loop by exactly that many steps (used to pin the capture step)."""
        lib, h, dt = self.lib, self.h, self.dt
        lib.FDTD_reset_pml(h)
        for f in self.fields:
            f.fill(0)
        Tn = int(np.int64(self.Ncycle * 3 / 4))
        self.Tn = Tn
        nc = self.nconv
        Eh = np.zeros((3, 3, nc))               # [sample slot 0..2][component][point]
        phi0 = np.zeros(nc); phi1 = np.zeros(nc)
        A0 = np.zeros(nc); A1 = np.zeros(nc)
        t0 = t1 = t2 = 0.0
        A_change = 1.0; phi_change = 1.0
        amp_rtol = self.rtol; phi_rtol = np.sqrt(self.rtol)
        self.history = []
        self.status = "converged"
        n = 0
        p = 0
        if fixed_n0 is not None:
            for n in range(fixed_n0):
                self._step(n)
            n = fixed_n0
        else:
            with np.errstate(divide="ignore", invalid="ignore"):
                while (A_change > amp_rtol or phi_change > phi_rtol or np.isnan(A_change)
                       or np.isinf(A_change) or np.isnan(phi_change) or np.isinf(phi_change)):
                    if n > self.Nmax or (max_steps is not None and n >= max_steps):
                        warnings.warn("Maximum number of time steps exceeded.")
                        self.status = "nmax"
                        break
                    self._step(n)
                    if p == Tn - 1:
                        t0, t1, t2 = t1, t2, n * dt
                        phi0[:] = phi1
                        A0[:] = A1
                        Eh[0] = Eh[1]; Eh[1] = Eh[2]
                        for c in range(3):
                            Eh[2, c] = np.real(self.fields[c][self.conv_pts])
                        for q in range(nc):
                            ph = [0.0] * 3; am = [0.0] * 3
                            for c in range(3):
                                ph[c] = lib.FDTD_calc_phase_3T(t0, t1, t2, Eh[0, c, q], Eh[1, c, q], Eh[2, c, q])
                                am[c] = lib.FDTD_calc_amplitude_3T(t0, t1, t2, Eh[0, c, q], Eh[1, c, q],
                                                                   Eh[2, c, q], ph[c])
                            for c in range(3):
                                if am[c] < 0:
                                    am[c] *= -1
                                    ph[c] += pi
                            phi1[q] = ph[0] + ph[1] + ph[2]
                            A1[q] = am[0] + am[1] + am[2]
                        A_change = np.linalg.norm(A1 - A0) / np.linalg.norm(A0)
                        phi_change = np.linalg.norm(phi1 - phi0) / np.linalg.norm(phi0)
                        self.history.append((n, float(A_change), float(phi_change)))
                        p = 0
                    else:
                        p += 1
                    n += 1
        lib.FDTD_capture_t0_fields(h)
        n0 = n
        self.n0 = n0
        for q in range(Tn):
            self._step(q + n0)
        lib.FDTD_capture_t1_fields(h)
        for q in range(Tn):
            self._step(q + n0 + Tn)
        lib.FDTD_calc_complex_fields_3T(h, n0 * dt, (n0 + Tn) * dt, (n0 + 2 * Tn) * dt)

    def _bind(self, sets):
        for a in sets[0] + sets[1]:
            a.fill(0)
        self.lib.FDTD_set_t0_arrays(self.h, *sets[0])
        self.lib.FDTD_set_t1_arrays(self.h, *sets[1])

    def solve_forward(self, fixed_n0=None, max_steps=None):
        self._bind(self.fwd)
        self.lib.FDTD_clear_sources(self.h)
        for arrs, i0, j0, k0, sI, sJ, sK in self.sources:
            self.lib.FDTD_add_source(self.h, *arrs, i0, j0, k0, sI, sJ, sK, False)
        t = time.perf_counter()
        self._solve(fixed_n0, max_steps)
        self.solve_seconds = time.perf_counter() - t

    def solve_adjoint(self, fixed_n0=None, max_steps=None):
        self._bind(self.adj)
        self.lib.FDTD_clear_sources(self.h)
        for arrs, i0, j0, k0, sI, sJ, sK in self.adj_sources:
            self.lib.FDTD_add_source(self.h, *arrs, i0, j0, k0, sI, sJ, sK, False)
        t = time.perf_counter()
        self._solve(fixed_n0, max_steps)
        self.solve_seconds = time.perf_counter() - t

    # ---------------------------------------------------------- field access
    _NAMES = ["Ex", "Ey", "Ez", "Hx", "Hy", "Hz"]

    def get_field(self, comp, domain=None, adjoint=False):
        """Raw phasor of one component on a domain (emopt/fdtd.py:1322-1359);
        no domain means DomainCoordinates(0, X, 0, Y, 0, Z), which drops the
        last plane along an axis where (N - 1) * d / d rounds below N - 1."""
        q = self._NAMES.index(comp)
        f = (self.adj if adjoint else self.fwd)[0][q].reshape(self.Nz, self.Ny, self.Nx)
        if domain is None:
            domain = Domain(0, self._X, 0, self._Y, 0, self._Z, self._dx, self._dy, self._dz)
        return f[domain.i, domain.j, domain.k].copy()

    def get_adjoint_field(self, comp, domain=None):
        return self.get_field(comp, domain, adjoint=True)

    def padded(self, q):
        return self.fields[q].reshape(self.I + 2, self.J + 2, self.K + 2)

    def get_field_interp(self, comp, domain=None, adjoint=False):
        """Interpolation of every component onto the Ez grid
        (emopt/fdtd.py:1401-1553)."""
        if comp == "Ez":
            return self.get_field(comp, domain, adjoint)
        Nx, Ny, Nz = self.Nx, self.Ny, self.Nz
        if domain is None:
            domain = Domain(0, self._X, 0, self._Y, 0, self._Z, self._dx, self._dy, self._dz)
            k1, k2, j1, j2, i1, i2 = domain.k1, domain.k2, domain.j1, domain.j2, domain.i1, domain.i2
            dint = domain
        else:
            k1, k2, j1, j2, i1, i2 = domain.k1, domain.k2, domain.j1, domain.j2, domain.i1, domain.i2
            if k1 > 0: k1 -= 1
            if k2 < Nx - 1: k2 += 1
            if j1 > 0: j1 -= 1
            if j2 < Ny - 1: j2 += 1
            if i1 > 0: i1 -= 1
            if i2 < Nz - 1: i2 += 1
            dint = Domain(k1 * self._dx, k2 * self._dx, j1 * self._dy, j2 * self._dy,
                          i1 * self._dz, i2 * self._dz, self._dx, self._dy, self._dz)
            k1, k2, j1, j2, i1, i2 = dint.k1, dint.k2, dint.j1, dint.j2, dint.i1, dint.i2
        fr = np.pad(self.get_field(comp, dint, adjoint), 1, "constant", constant_values=0)

        def crop(f):
            return f[1 + domain.i1 - i1:-1 - (i2 - domain.i2),
                     1 + domain.j1 - j1:-1 - (j2 - domain.j2),
                     1 + domain.k1 - k1:-1 - (k2 - domain.k2)]

        bc = self._bc

        def mirror_x():
            if k1 == 0 and bc[0] == "E": fr[:, :, 0] = -1 * fr[:, :, 1]
            elif k1 == 0 and bc[0] == "H": fr[:, :, 0] = fr[:, :, 1]

        def mirror_y():
            if j1 == 0 and bc[1] == "E": fr[:, 0, :] = -1 * fr[:, 1, :]
            elif j1 == 0 and bc[1] == "H": fr[:, 0, :] = fr[:, 1, :]

        if comp == "Ex":
            mirror_x()
            f = np.copy(fr)
            f[1:-1, :, 1:-1] += fr[1:-1, :, 0:-2]
            f[1:-1, :, 1:-1] += fr[2:, :, 1:-1]
            f[1:-1, :, 1:-1] += fr[2:, :, 0:-2]
            return crop(f / 4.0)
        if comp == "Ey":
            mirror_y()
            f = np.copy(fr)
            f[1:-1, 1:-1, :] += fr[1:-1, 0:-2, :]
            f[1:-1, 1:-1, :] += fr[2:, 1:-1, :]
            f[1:-1, 1:-1, :] += fr[2:, 0:-2, :]
            return crop(f / 4.0)
        if comp == "Hx":
            mirror_y()
            f = np.copy(fr)
            f[:, 1:, :] += fr[:, 0:-1, :]
            return crop(f / 2.0)
        if comp == "Hy":
            mirror_x()
            f = np.copy(fr)
            f[:, :, 1:] += fr[:, :, 0:-1]
            return crop(f / 2.0)
        # Hz
        mirror_x()
        mirror_y()
        f = np.copy(fr)
        f[0:-1, 1:, 1:] += fr[0:-1, 1:, 0:-1]
        f[0:-1, 1:, 1:] += fr[0:-1, 0:-1, 1:]
        f[0:-1, 1:, 1:] += fr[0:-1, 0:-1, 0:-1]
        f[0:-1, 1:, 1:] += fr[1:, 1:, 1:]
        f[0:-1, 1:, 1:] += fr[1:, 1:, 0:-1]
        f[0:-1, 1:, 1:] += fr[1:, 0:-1, 1:]
        f[0:-1, 1:, 1:] += fr[1:, 0:-1, 0:-1]
        return crop(f / 8.0)

    def get_source_power(self):
        """Poynting flux through a box one cell inside the CPML
        (emopt/fdtd.py:1555-1667)."""
        dx, dy, dz = self._dx, self._dy, self._dz
        w = self._w_pml
        xmin = w[0] + dx if w[0] > 0 else 0.0
        xmax = self._X - w[1] - dx if w[1] > 0 else self._X - dx
        ymin = w[2] + dy if w[2] > 0 else 0.0
        ymax = self._Y - w[3] - dy if w[3] > 0 else self._Y - dy
        zmin = w[4] + dz if w[4] > 0 else 0.0
        zmax = self._Z - w[5] - dz if w[5] > 0 else self._Z - dz
        D = lambda *a: Domain(*a, dx, dy, dz)
        x1 = D(xmin, xmin, ymin, ymax, zmin, zmax); x2 = D(xmax, xmax, ymin, ymax, zmin, zmax)
        y1 = D(xmin, xmax, ymin, ymin, zmin, zmax); y2 = D(xmin, xmax, ymax, ymax, zmin, zmax)
        z1 = D(xmin, xmax, ymin, ymax, zmin, zmin); z2 = D(xmin, xmax, ymin, ymax, zmax, zmax)
        g = self.get_field_interp
        P = 0.0
        Ey, Ez, Hy, Hz = g("Ey", x1), g("Ez", x1), g("Hy", x1), g("Hz", x1)
        if self._bc[0] not in "EH":
            P += -0.5 * dy * dz * np.sum(np.real(Ey * np.conj(Hz) - Ez * np.conj(Hy)))
        Ey, Ez, Hy, Hz = g("Ey", x2), g("Ez", x2), g("Hy", x2), g("Hz", x2)
        P += 0.5 * dy * dz * np.sum(np.real(Ey * np.conj(Hz) - Ez * np.conj(Hy)))
        Ex, Ez, Hx, Hz = g("Ex", y1), g("Ez", y1), g("Hx", y1), g("Hz", y1)
        if self._bc[1] not in "EH":
            P += 0.5 * dx * dz * np.sum(np.real(Ex * np.conj(Hz) - Ez * np.conj(Hx)))
        Ex, Ez, Hx, Hz = g("Ex", y2), g("Ez", y2), g("Hx", y2), g("Hz", y2)
        P += -0.5 * dx * dz * np.sum(np.real(Ex * np.conj(Hz) - Ez * np.conj(Hx)))
        Ex, Ey, Hx, Hy = g("Ex", z1), g("Ey", z1), g("Hx", z1), g("Hy", z1)
        if self._bc[2] not in "EH":
            P += -0.5 * dx * dy * np.sum(np.real(Ex * np.conj(Hy) - Ey * np.conj(Hx)))
        Ex, Ey, Hx, Hy = g("Ex", z2), g("Ey", z2), g("Hx", z2), g("Hy", z2)
        P += 0.5 * dx * dy * np.sum(np.real(Ex * np.conj(Hy) - Ey * np.conj(Hx)))
        return P

    def get_A_diag(self):
        return tuple(np.copy(m) for m in self.mats)

    def calc_ydAx(self, Adiag0):
        """sum(E_adj*1j*(eps-eps0)*E) - sum(H_adj*1j*(mu-mu0)*H)
        (emopt/fdtd.py:1694-1717)."""
        out = np.zeros(self.mats[0].shape, dtype=np.complex128)
        for q in range(3):
            out = out + self.adj[0][q] * 1j * (self.mats[q] - Adiag0[q]) * self.fwd[0][q]
        for q in range(3, 6):
            out = out + self.adj[0][q] * -1j * (self.mats[q] - Adiag0[q]) * self.fwd[0][q]
        return np.sum(out)

    def phasors(self, adjoint=False):
        """The six recovered complex components, each shaped [Nz,Ny,Nx]."""
        sets = self.adj if adjoint else self.fwd
        return [a.reshape(self.Nz, self.Ny, self.Nx).copy() for a in sets[0]]


def holder(Nx, Ny, Nz, dx, dy, dz, fwd, adj=None, mats=None, bc="000", w_pml=(0.0,) * 6):
    """An OracleFDTD with no library behind it, holding given complex phasors
    (six [Nz,Ny,Nx] arrays, forward and optionally adjoint) and material
    arrays, so that the post-processing restatements above
    (get_field_interp, get_source_power, get_A_diag, calc_ydAx) can check
    fields computed elsewhere."""
    o = OracleFDTD.__new__(OracleFDTD)
    o.lib = None
    o.Nx, o.Ny, o.Nz = Nx, Ny, Nz
    o.I, o.J, o.K = Nz, Ny, Nx
    o._dx, o._dy, o._dz = dx, dy, dz
    o._X, o._Y, o._Z = dx * (Nx - 1), dy * (Ny - 1), dz * (Nz - 1)
    o._bc = bc
    o._w_pml = list(w_pml)
    n = Nx * Ny * Nz

    def flat(arrs):
        if arrs is None:
            return [np.zeros(n, dtype=np.complex128) for _ in range(6)]
        return [np.ascontiguousarray(a, dtype=np.complex128).ravel().copy() for a in arrs]
    o.fwd = [flat(fwd), flat(None)]
    o.adj = [flat(adj), flat(None)]
    o.mats = flat(mats) if mats is not None else [np.ones(n, dtype=np.complex128)
                                                  for _ in range(6)]
    return o
