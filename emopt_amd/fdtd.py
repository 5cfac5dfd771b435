"""CW-FDTD solver with the surface of ``emopt.fdtd.FDTD`` (emopt/fdtd.py:93-1730),
computed on an MI355X through ``libemfdtd.so``.

What moved to the device compared with the reference:

* the whole time loop of ``__solve`` (emopt/fdtd.py:994-1194): H/E updates,
  CPML, sources, the every-Tn convergence check, the two snapshots and the
  3-sample phasor recovery run as HIP kernels; Python makes one call per solve
  instead of two ctypes calls per step plus six ghost exchanges;
* field state, materials and results stay resident in HBM; ``get_field``
  copies back only the requested domain.

Everything else -- grid sizing, dt, CPML defaults, ramp, Nmax, property
validation, source cropping, field interpolation, source power, ``get_A_diag``
and ``calc_ydAx`` -- keeps the reference's semantics; each method cites the
lines it follows.  There is one process and one GPU: the reference's MPI
domain decomposition (PETSc DMDA, GhostComm) has no counterpart.
"""
from __future__ import annotations

import warnings
from math import pi

import numpy as np

from . import _lib
from ._lib import SolveInfo, c128, ptr
from .defs import FIELD_INDEX, FieldComponent
from .grid import StructuredMaterial3D
from .misc import DomainCoordinates
from .simulation import MaxwellSolver

__all__ = ["FDTD", "SourceArray"]

STATUS = {0: "converged", 1: "nmax", 2: "fixed", 3: "nan"}


class SourceArray:
    """Source block container (emopt/fdtd.py:49-91)."""

    def __init__(self, Jx, Jy, Jz, Mx, My, Mz, i0, j0, k0, I, J, K):
        self.Jx, self.Jy, self.Jz, self.Mx, self.My, self.Mz = Jx, Jy, Jz, Mx, My, Mz
        self.i0, self.j0, self.k0, self.I, self.J, self.K = i0, j0, k0, I, J, K

    @property
    def arrays(self):
        return (self.Jx, self.Jy, self.Jz, self.Mx, self.My, self.Mz)


class FDTD(MaxwellSolver):
    """A 3D CW-FDTD solver on one MI355X (constructor: emopt/fdtd.py:236-437).

    Extra keyword ``device`` selects the HIP device.  ``dist_method`` is
    accepted for signature compatibility; one GPU owns the whole grid.
    """

    def __init__(self, X, Y, Z, dx, dy, dz, wavelength, rtol=1e-6, nconv=None,
                 min_rindex=1.0, complex_eps=False, dist_method="auto", device=0):
        super().__init__(3)
        nproc = 1
        if nconv is None:
            nconv = nproc * 10
        elif nconv < nproc * 2:
            warnings.warn("Number of convergence test points (nconv) is likely too low.")
        self._dx, self._dy, self._dz = dx, dy, dz
        Nx = int(np.ceil(X / dx) + 1); self._Nx = Nx
        Ny = int(np.ceil(Y / dy) + 1); self._Ny = Ny
        Nz = int(np.ceil(Z / dz) + 1); self._Nz = Nz
        self._X = dx * (Nx - 1); self._Y = dy * (Ny - 1); self._Z = dz * (Nz - 1)
        self._wavelength = wavelength
        self._R = wavelength / (2 * pi)
        self._Sc = 0.95
        self._min_rindex = min_rindex
        self._dt = self._Sc * np.min([dx, dy, dz]) / self._R / np.sqrt(3) * min_rindex
        self._complex_eps = bool(complex_eps)

        self._lib = _lib.load()
        L = self._lib
        h = _lib.c_void_p()
        L.emf_new(int(device), _lib.ctypes.byref(h))
        self._h = h
        self._device = device
        L.emf_set_wavelength(h, wavelength)
        L.emf_set_physical_dims(h, X, Y, Z, dx, dy, dz)
        L.emf_set_grid_dims(h, Nx, Ny, Nz)
        L.emf_set_local_grid(h, 0, 0, 0, Nx, Ny, Nz)
        L.emf_set_dt(h, self._dt)
        L.emf_set_complex_eps(h, int(self._complex_eps))

        w_pml = 15   # emopt/fdtd.py:368-390
        self._w_pml = [w_pml * dx, w_pml * dx, w_pml * dy, w_pml * dy, w_pml * dz, w_pml * dz]
        self._w_cells = [w_pml] * 6
        self._pml_sigma, self._pml_alpha, self._pml_kappa, self._pml_pow = 3.0, 0.0, 1.0, 3.0
        L.emf_set_pml_widths(h, *self._w_cells)
        L.emf_set_pml_properties(h, self._pml_sigma, self._pml_alpha, self._pml_kappa,
                                 self._pml_pow)
        L.emf_build_pml(h)

        Nlambda = wavelength / np.min([dx, dy, dz]) / self._min_rindex   # :393-402
        Ncycle = Nlambda * np.sqrt(3)
        self._Nlambda, self._Ncycle = Nlambda, Ncycle
        self._src_T = Ncycle * 20.0 * self._dt
        self._src_min = 1e-4
        self.Nmax = Nlambda * 500
        L.emf_set_source_properties(h, self._src_T, self._src_min)
        self._rtol = rtol
        self.verbose = 2
        self._nconv = int(nconv / nproc)
        L.emf_set_nconv(h, self._nconv)

        self._sources = []
        self._adj_sources = []
        self._adj_staged = False       # adjoint sources include the device-staged blocks
        self._bc = ["0", "0", "0"]
        L.emf_set_bc(h, b"000")
        self._eps = None
        self._mu = None
        self._mats_ready = False
        self._have = [False, False]    # result slots filled (fwd, adj)
        self.forward_count = 0
        self.adjoint_count = 0
        self.fixed_n0 = None           # pin the capture step (parity runs)
        self.last_solve = None         # dict of the last emf_solve outcome
        self._history = np.zeros((0, 3))

    def __del__(self):
        try:
            if getattr(self, "_h", None) is not None and self._h.value:
                self._lib.emf_delete(self._h)
                self._h = None
        except Exception:
            pass

    # ------------------------------------------------------------ properties
    @property
    def wavelength(self):
        return self._wavelength

    @wavelength.setter
    def wavelength(self, wlen):   # emopt/fdtd.py:443-455
        if wlen <= 0:
            raise ValueError("Wavelength must be >= 0.")
        self._wavelength = wlen
        self._R = wlen / 2 / pi
        self._lib.emf_set_wavelength(self._h, wlen)
        ds = np.min([self._dx, self._dy, self._dz])
        self._dt = self._Sc * ds / self._R / np.sqrt(3) * self._min_rindex
        self._lib.emf_set_dt(self._h, self._dt)

    X = property(lambda self: self._X)
    Y = property(lambda self: self._Y)
    Z = property(lambda self: self._Z)
    dx = property(lambda self: self._dx)
    dy = property(lambda self: self._dy)
    dz = property(lambda self: self._dz)
    Nx = property(lambda self: self._Nx)
    Ny = property(lambda self: self._Ny)
    Nz = property(lambda self: self._Nz)
    eps = property(lambda self: self._eps)
    mu = property(lambda self: self._mu)
    dt = property(lambda self: self._dt)
    Nlambda = property(lambda self: self._Nlambda)
    Ncycle = property(lambda self: self._Ncycle)

    @property
    def Tn(self):
        return int(np.int64(self._Ncycle * 3 / 4))

    @property
    def courant_num(self):
        return self._Sc

    @courant_num.setter
    def courant_num(self, Sc):   # emopt/fdtd.py:505-514
        if Sc > 1.0:
            warnings.warn("A courant number greater than 1 will lead to instability.")
        self._Sc = Sc
        ds = np.min([self._dx, self._dy, self._dz])
        self._dt = self._Sc * ds / self._R / np.sqrt(3) * self._min_rindex
        self._lib.emf_set_dt(self._h, self._dt)

    @property
    def rtol(self):
        return self._rtol

    @rtol.setter
    def rtol(self, rtol):
        if rtol <= 0:
            raise ValueError("Relative tolerance must be greater than zero.")
        self._rtol = rtol

    @property
    def src_min_value(self):
        return self._src_min

    @src_min_value.setter
    def src_min_value(self, new_min):
        if new_min <= 0:
            raise ValueError("Minimum source value must be greater than zero.")
        self._src_min = new_min
        self._lib.emf_set_source_properties(self._h, self._src_T, self._src_min)

    @property
    def src_ramp_time(self):
        return self._src_T

    @src_ramp_time.setter
    def src_ramp_time(self, new_T):   # emopt/fdtd.py:551-565
        if new_T <= 0:
            raise ValueError("Source ramp time must be greater than zero.")
        if new_T <= self._Nlambda * 5:
            warnings.warn("Source ramp time is likely too short. Simulations may not converge.")
        self._src_T = new_T
        self._lib.emf_set_source_properties(self._h, self._src_T, self._src_min)

    @property
    def bc(self):
        return "".join(self._bc)

    @bc.setter
    def bc(self, newbc):   # emopt/fdtd.py:570-583
        if len(newbc) != 3:
            raise ValueError("Incorrect number of boundary conditions specified!")
        for b in newbc:
            if b not in "0EHP":
                raise ValueError("Unrecognized boundary condition: %s" % (b))
            if b == "P":
                raise NotImplementedError("Periodic boundary conditions not yet implemented")
        self._bc = list(newbc)
        self._lib.emf_set_bc(self._h, "".join(self._bc).encode("ascii"))

    @property
    def w_pml(self):
        return self._w_pml

    @w_pml.setter
    def w_pml(self, w_pml):   # emopt/fdtd.py:589-616
        if len(w_pml) != 6:
            raise ValueError("Incorrect number of pml widths specified. You must specify 6 "
                             "widths (xmin, xmax, ymin, ymax, zmin, zmax).")
        d = [self._dx, self._dx, self._dy, self._dy, self._dz, self._dz]
        self._w_cells = [int(w_pml[q] / d[q]) for q in range(6)]
        self._w_pml = [self._w_cells[q] * d[q] for q in range(6)]
        self._lib.emf_set_pml_widths(self._h, *self._w_cells)
        self._lib.emf_build_pml(self._h)

    w_pml_xmin = property(lambda self: self._w_cells[0])
    w_pml_xmax = property(lambda self: self._w_cells[1])
    w_pml_ymin = property(lambda self: self._w_cells[2])
    w_pml_ymax = property(lambda self: self._w_cells[3])
    w_pml_zmin = property(lambda self: self._w_cells[4])
    w_pml_zmax = property(lambda self: self._w_cells[5])
    X_real = property(lambda self: self._X - self._w_pml[0] - self._w_pml[1])
    Y_real = property(lambda self: self._Y - self._w_pml[2] - self._w_pml[3])
    Z_real = property(lambda self: self._Z - self._w_pml[4] - self._w_pml[5])

    def set_pml_properties(self, sigma=3.0, alpha=0.0, kappa=1.0, pow=3.0):
        self._pml_sigma, self._pml_alpha, self._pml_kappa, self._pml_pow = sigma, alpha, kappa, pow
        self._lib.emf_set_pml_properties(self._h, sigma, alpha, kappa, pow)

    # ------------------------------------------------------------- materials
    def set_materials(self, eps, mu):
        self._eps = eps
        self._mu = mu

    # Yee offsets (sx, sy, sz) of eps_x, eps_y, eps_z, mu_x, mu_y, mu_z
    # (emopt/fdtd.py:858-880)
    YEE_SHIFTS = ((0.5, 0.0, -0.5), (0.0, 0.5, -0.5), (0.0, 0.0, 0.0),
                  (0.0, 0.5, 0.0), (0.5, 0.0, 0.0), (0.5, 0.5, -0.5))

    def build(self):
        """Sample eps and mu at the six Yee offsets and upload them
        (emopt/fdtd.py:842-880)."""
        if self._eps is None or self._mu is None:
            raise RuntimeError("set_materials(eps, mu) must be called before build()")
        Nx, Ny, Nz = self._Nx, self._Ny, self._Nz
        arrs = []
        for q, (sx, sy, sz) in enumerate(self.YEE_SHIFTS):
            m = self._eps if q < 3 else self._mu
            a = np.empty(Nx * Ny * Nz, dtype=np.complex128)
            self._sample(m, 0, Nx, 0, Ny, 0, Nz, sx, sy, sz, arr=a)
            arrs.append(c128(a))
        self.set_material_arrays(arrs[:3], arrs[3:])

    def _sample(self, m, k1, k2, j1, j2, i1, i2, sx, sy, sz, arr=None):
        """Material values of a box at one Yee offset: structured materials
        are sampled on the GPU (the same doubles as their host sampler),
        anything else by its own get_values."""
        if isinstance(m, StructuredMaterial3D):
            return m.get_values_device(k1, k2, j1, j2, i1, i2, sx=sx, sy=sy, sz=sz,
                                       device=self._device, arr=arr)
        return m.get_values(k1, k2, j1, j2, i1, i2, sx=sx, sy=sy, sz=sz, arr=arr)

    def set_material_arrays(self, eps_xyz, mu_xyz):
        """Upload already-sampled material arrays ([Nz,Ny,Nx] complex each)."""
        arrs = [c128(a).ravel() for a in list(eps_xyz) + list(mu_xyz)]
        n = self._Nx * self._Ny * self._Nz
        for a in arrs:
            if a.size != n:
                raise ValueError("material arrays must have Nz*Ny*Nx = %d entries" % n)
        self._lib.emf_set_mat_arrays(self._h, *[ptr(a) for a in arrs])
        self._mu_uniform = all(bool(np.all(np.real(a) == np.real(a[0]))) for a in arrs[3:])
        self._mats_ready = True

    def update(self, bbox=None):
        """Re-sample eps and mu inside a real-space bounding box [xmin, xmax,
        ymin, ymax, zmin, zmax] at their six Yee offsets (emopt/fdtd.py:
        882-992); None rebuilds everything.  mu values equal to the grid's
        uniform mu change nothing on the device."""
        if bbox is None:
            v = self.verbose
            self.verbose = 0
            self.build()
            self.verbose = v
            return
        k1, k2, j1, j2, i1, i2 = self.index_box(bbox)
        if k1 >= k2 or j1 >= j2 or i1 >= i2:
            return
        for c in range(3):
            sx, sy, sz = self.YEE_SHIFTS[c]
            tmp = c128(self._sample(self._eps, k1, k2, j1, j2, i1, i2, sx, sy, sz))
            self._lib.emf_update_eps_box(self._h, c, ptr(tmp), k1, k2, j1, j2, i1, i2)
        for c in range(3):
            sx, sy, sz = self.YEE_SHIFTS[3 + c]
            tmp = c128(self._sample(self._mu, k1, k2, j1, j2, i1, i2, sx, sy, sz))
            self._lib.emf_update_mu_box(self._h, c, ptr(tmp), k1, k2, j1, j2, i1, i2)
            if self._mu_uniform and not np.all(tmp == tmp.flat[0]):
                self._mu_uniform = False

    def index_box(self, bbox):
        """(k1, k2, j1, j2, i1, i2) of the cells a real-space box [xmin,
        xmax, ymin, ymax, zmin, zmax] covers (DomainCoordinates rounding,
        emopt/misc.py:212-279), clipped to the grid."""
        dom = DomainCoordinates(bbox[0], bbox[1], bbox[2], bbox[3], bbox[4], bbox[5],
                                self._dx, self._dy, self._dz)
        return (max(0, dom.k1), min(self._Nx, dom.k2), max(0, dom.j1), min(self._Ny, dom.j2),
                max(0, dom.i1), min(self._Nz, dom.i2))

    # --------------------------------------------------------------- sources
    def _local_overlap(self, domain):
        # emopt/fdtd.py:658-708 with one chunk covering the whole grid
        I, J, K = self._Nz, self._Ny, self._Nx
        if 0 >= domain.i2 or I <= domain.i1 or 0 >= domain.j2 or J <= domain.j1 \
                or 0 >= domain.k2 or K <= domain.k1:
            return None
        imin, jmin, kmin = max(0, domain.i1), max(0, domain.j1), max(0, domain.k1)
        imax, jmax, kmax = min(I, domain.i2), min(J, domain.j2), min(K, domain.k2)
        return ((imin, jmin, kmin), (imin - domain.i1, jmin - domain.j1, kmin - domain.k1),
                (imax - imin, jmax - jmin, kmax - kmin))

    def _set_sources(self, src, domain, adjoint=False):
        """Crop the six arrays to the grid and keep them (emopt/fdtd.py:710-747).
        The complex -> (amplitude, phase) conversion of src/fdtd.cpp:1517-1576
        happens in the library when the solve registers them."""
        ov = self._local_overlap(domain)
        if ov is None:
            return
        (i0, j0, k0), (id1, jd1, kd1), (sI, sJ, sK) = ov
        sl = (slice(id1, id1 + sI), slice(jd1, jd1 + sJ), slice(kd1, kd1 + sK))
        arrs = [c128(np.copy(np.asarray(a)[sl])).ravel() for a in src]
        sa = SourceArray(*arrs, i0, j0, k0, sI, sJ, sK)
        (self._adj_sources if adjoint else self._sources).append(sa)

    def set_sources(self, src, domain, mindex=0):
        """Add a source given as (Jx, Jy, Jz, Mx, My, Mz) arrays, or as a
        solved :class:`emopt_amd.modes.ModeFullVector` whose mode `mindex`
        is launched, on `domain` (emopt/fdtd.py:749-788)."""
        from .modes import ModeFullVector
        if isinstance(src, ModeFullVector):
            src = src.get_source(mindex, self._dx, self._dy, self._dz)
        if not isinstance(src, (tuple, list)) or len(src) != 6:
            raise ValueError("sources must be six current-density arrays (Jx, Jy, Jz, Mx, My, "
                             "Mz) or a ModeFullVector")
        self._set_sources(src, domain, adjoint=False)

    def set_adjoint_sources(self, src):
        """src = [list of 6-array tuples, list of domains] (emopt/fdtd.py:790-813)."""
        self.clear_adjoint_sources()
        for dFdx, domain in zip(src[0], src[1]):
            self.add_adjoint_sources(dFdx, domain)

    def add_adjoint_sources(self, src, domain):
        self._set_sources(src, domain, adjoint=True)

    def clear_sources(self):
        self._sources = []

    def clear_adjoint_sources(self):
        self._adj_sources = []
        self._adj_staged = False

    def set_adjoint_sources_staged(self):
        """Make the blocks staged on the device (stage_planes, stage_power_face)
        the adjoint sources; sources added with add_adjoint_sources join them."""
        self._adj_sources = []
        self._adj_staged = True

    def _register(self, sources):
        L, h = self._lib, self._h
        if sources is self._adj_sources and self._adj_staged:
            L.emf_stage_register(h)          # replaces the current sources
        else:
            L.emf_clear_sources(h)
        for s in sources:
            L.emf_add_source(h, *[ptr(a) for a in s.arrays], s.i0, s.j0, s.k0, s.I, s.J, s.K, 1)

    # ------------------------------------- adjoint sources built on the device
    def _grown_place(self, domain):
        """(i0, j0, k0, id1, jd1, kd1, I, J, K) of interpolated_dFdx_3D's grown
        domain (one cell more at -x, -y and +z) clipped to the grid, as
        add_adjoint_sources would place it; None when it misses the grid."""
        g = domain.copy()
        g.set_region(domain.k1 - 1, domain.k2, domain.j1 - 1, domain.j2, domain.i1,
                     domain.i2 + 1, self._dx, self._dy, self._dz)
        ov = self._local_overlap(g)
        if ov is None:
            return None
        return np.array([v for t in ov for v in t], dtype=np.int32)

    def stage_clear(self):
        self._lib.emf_stage_clear(self._h)

    def stage_count(self):
        n = _lib.ctypes.c_int()
        self._lib.emf_stage_count(self._h, _lib.ctypes.byref(n))
        return n.value

    def stage_planes(self, weight_set, scale, div, domain, mask=0b111111):
        """Stage interpolated_dFdx_3D of the planes (scale_q * W_q) / div on
        `domain`, W the weight set (set_plane_weights) and scale six complex
        factors (Ex..Hz); components whose mask bit is clear are zero."""
        pl = self._grown_place(domain)
        if pl is None:
            return
        sc = np.zeros(12)
        sc[0::2] = np.real(scale)
        sc[1::2] = np.imag(scale)
        self._lib.emf_stage_interp_T(self._h, int(weight_set), ptr(sc), int(mask), 1.0 / div,
                                     domain.k1, domain.k2, domain.j1, domain.j2, domain.i1,
                                     domain.i2, ptr(pl))

    def stage_power_face(self, comps, s1, mf, rp2, domain, adjoint=False):
        """Stage one face term of power_norm_dFdx_3D: comps = indices of
        (E1, E2, H1, H2), s1 = sign * area / 4, mf = -f, rp2 = 1 / Psrc^2.
        False (nothing staged) when the face's grown box leaves the grid."""
        from .fields import grown_boxes
        _, _, dom, ext, ok = grown_boxes(self, domain)
        pl = self._grown_place(domain)
        if not ok:
            return False
        if pl is None:
            return True
        cs = np.array(comps, dtype=np.int32)
        self._lib.emf_stage_power_face(self._h, 1 if adjoint else 0, ptr(cs), float(s1), float(mf),
                                       float(rp2), *dom, *ext, ptr(pl))
        return True

    def staged_block(self, idx):
        """(i0, j0, k0, I, J, K) and the six stored arrays of staged block idx."""
        geom = np.zeros(6, dtype=np.int32)
        self._lib.emf_stage_get(self._h, int(idx), ptr(geom), None)
        I, J, K = int(geom[3]), int(geom[4]), int(geom[5])
        out = np.zeros((6, I, J, K), dtype=np.complex128)
        self._lib.emf_stage_get(self._h, int(idx), ptr(geom), ptr(out))
        return tuple(int(v) for v in geom), out

    # ----------------------------------------------------------------- solve
    def _solve(self, slot):
        if not self._mats_ready:
            self.build()
        info = SolveInfo()
        n0 = -1 if self.fixed_n0 is None else int(self.fixed_n0)
        self._lib.emf_solve(self._h, slot, float(self._rtol), float(self.Nmax), self.Tn, n0,
                            _lib.ctypes.byref(info))
        self._solved(slot, info)

    def _solved(self, slot, info):
        """Bookkeeping after a solve into `slot` (by emf_solve or a batch)."""
        nrec = self._lib.emf_get_history(self._h, None, 0)
        hist = np.zeros((max(nrec, 0), 3))
        if nrec > 0:
            self._lib.emf_get_history(self._h, ptr(hist), nrec)
        self._history = hist
        self.last_solve = dict(n0=int(info.n0), Tn=int(info.Tn), steps=int(info.steps),
                               status=STATUS.get(info.status, str(info.status)),
                               nchecks=int(info.nchecks), a_change=float(info.a_change),
                               phi_change=float(info.phi_change), seconds=float(info.seconds))
        if info.status in (1, 3):
            warnings.warn("Maximum number of time steps exceeded.")
        if info.status == 3:
            warnings.warn("The fields are not finite (NaN or Inf) at the last convergence check.")
        self._have[slot] = True

    @property
    def convergence_history(self):
        """(n, amplitude change, phase change) of every convergence check."""
        return self._history

    def solve_forward(self):
        """Forward solve (emopt/fdtd.py:1196-1250)."""
        self._register(self._sources)
        self._solve(0)
        self._forward_done()

    def _forward_done(self):
        self.update_saved_fields()
        self._source_power = self.get_source_power()
        self.forward_count += 1

    def solve_adjoint(self):
        """Adjoint solve: the same loop with the adjoint sources
        (emopt/fdtd.py:1252-1300)."""
        self._register(self._adj_sources)
        self._solve(1)
        self.adjoint_count += 1

    # ----------------------------------------------------------- field access
    def update_saved_fields(self):
        self._saved_fields = []
        for domain in self._field_domains:
            self._saved_fields.append(tuple(self.get_field_interp(c, domain) for c in
                                            ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")))

    def _get(self, component, domain, slot):
        if not self._have[slot]:
            raise RuntimeError("no %s solution available" % ("forward", "adjoint")[slot])
        if domain is None:
            domain = DomainCoordinates(0, self._X, 0, self._Y, 0, self._Z,
                                       self._dx, self._dy, self._dz)
        k1, k2 = max(0, domain.k1), min(self._Nx, domain.k2)
        j1, j2 = max(0, domain.j1), min(self._Ny, domain.j2)
        i1, i2 = max(0, domain.i1), min(self._Nz, domain.i2)
        out = np.zeros((max(i2 - i1, 0), max(j2 - j1, 0), max(k2 - k1, 0)), dtype=np.complex128)
        if out.size:
            self._lib.emf_get_field(self._h, slot, FIELD_INDEX[component], k1, k2, j1, j2,
                                    i1, i2, ptr(out))
        return out

    def _interp_device(self, component, dom, ext, slot):
        """get_field_interp's averaging on the GPU (emf_get_field_interp):
        dom, ext = (k1, k2, j1, j2, i1, i2) of the domain and of the grown box."""
        if not self._have[slot]:
            raise RuntimeError("no %s solution available" % ("forward", "adjoint")[slot])
        k1, k2, j1, j2, i1, i2 = dom
        out = np.zeros((i2 - i1, j2 - j1, k2 - k1), dtype=np.complex128)
        if out.size:
            self._lib.emf_get_field_interp(self._h, slot, FIELD_INDEX[component], *dom, *ext,
                                           ptr(out))
        return out

    def _flux_device(self, axis, dom, ext, slot):
        """One face of get_source_power reduced on the GPU
        (emf_interp_flux): sum of Re(a conj(b) - c conj(d)) of the fields
        interpolated on dom (grown box ext)."""
        if not self._have[slot]:
            raise RuntimeError("no %s solution available" % ("forward", "adjoint")[slot])
        out = np.zeros(1)
        self._lib.emf_interp_flux(self._h, slot, int(axis), *dom, *ext, ptr(out))
        return float(out[0])

    def set_plane_weights(self, wid, weights):
        """Upload six weight planes (Ex..Hz, None: zero) of one domain's shape
        as set `wid` (0..7) for interp_dot."""
        arrs = [None if w is None else c128(w).ravel() for w in weights]
        count = next((a.size for a in arrs if a is not None), 0)
        if any(a is not None and a.size != count for a in arrs):
            raise ValueError("weight planes must have one shape")
        self._lib.emf_set_plane_weights(self._h, int(wid), int(count), *[ptr(a) for a in arrs])
        self._weights_shape = getattr(self, "_weights_shape", {})
        self._weights_shape[wid] = count

    def interp_dot(self, wid, domain, adjoint=False):
        """(sum_E, sum_H) = sums over `domain` of weight set `wid` times the
        fields interpolated onto the Ez grid (E components, H components),
        reduced on the GPU (emf_interp_dot)."""
        from .fields import grown_boxes
        slot = 1 if adjoint else 0
        if not self._have[slot]:
            raise RuntimeError("no %s solution available" % ("forward", "adjoint")[slot])
        _, _, dom, ext, ok = grown_boxes(self, domain)
        if not ok:
            raise ValueError("interp_dot needs a domain inside the grid")
        out = np.zeros(4)
        self._lib.emf_interp_dot(self._h, slot, int(wid), *dom, *ext, ptr(out))
        return complex(out[0], out[1]), complex(out[2], out[3])

    def get_field(self, component, domain=None):
        """Raw (uninterpolated) forward field (emopt/fdtd.py:1322-1380)."""
        return self._get(component, domain, 0)

    def get_adjoint_field(self, component, domain=None):
        return self._get(component, domain, 1)

    def get_field_interp(self, component, domain=None, squeeze=False, adjoint=False):
        """Field interpolated onto the Ez grid (emopt/fdtd.py:1401-1553)."""
        from .fields import interp_to_ez
        return interp_to_ez(self, component, domain, squeeze, adjoint)

    def get_source_power(self):
        """Power out of a box one cell inside the CPML (emopt/fdtd.py:1555-1667)."""
        from .fields import source_power
        return source_power(self)

    def src_func_t(self, n, t, phase=0.0):
        """The source time function at step n, time t (FDTD_src_func_t,
        src/fdtd.cpp:1595-1602): the ramp envelope times sin(t + phase)."""
        return self._lib.emf_src_func_t(self._h, int(n), float(t), float(phase))

    def test_src_func(self):
        """Plot the source time function over 3000 steps (emopt/fdtd.py:1720-1730)."""
        time = np.arange(0, 3000, 1) * self._dt
        ramp = np.array([self.src_func_t(i, time[i], 0.0) for i in range(len(time))])
        try:
            import matplotlib.pyplot as plt
        except ImportError:
            warnings.warn("matplotlib is not available: no plot")
            return ramp
        plt.plot(time, ramp)
        plt.show()
        return ramp

    def get_A_diag(self):
        """Copies of (eps_x, eps_y, eps_z, mu_x, mu_y, mu_z) (emopt/fdtd.py:1669-1692)."""
        n = self._Nx * self._Ny * self._Nz
        arrs = [np.empty(n, dtype=np.complex128) for _ in range(6)]
        self._lib.emf_get_mat_arrays(self._h, *[ptr(a) for a in arrs])
        return tuple(arrs)

    def field_products(self, domain, which="E"):
        """(F_x F_adj_x, F_y F_adj_y, F_z F_adj_z) on `domain` for F = E or H, as a
        numpy array [3, nz, ny, nx]: the elementwise products of the forward
        and adjoint phasors, formed on the GPU (emf_field_products)."""
        if domain is None:
            domain = DomainCoordinates(0, self._X, 0, self._Y, 0, self._Z,
                                       self._dx, self._dy, self._dz)
        k1, k2 = max(0, domain.k1), min(self._Nx, domain.k2)
        j1, j2 = max(0, domain.j1), min(self._Ny, domain.j2)
        i1, i2 = max(0, domain.i1), min(self._Nz, domain.i2)
        out = np.zeros((3, max(i2 - i1, 0), max(j2 - j1, 0), max(k2 - k1, 0)),
                       dtype=np.complex128)
        if out.size:
            self._lib.emf_field_products(self._h, {"E": 0, "H": 1}[which], k1, k2, j1, j2, i1, i2,
                                         ptr(out))
        return out

    def calc_ydAx_topology(self, domain, update_mu, sig_eps=None, sig_mu=None, del_eps=1.,
                           del_mu=1., planar=False, lam=0.):
        """Topology gradient 2 sig (lam / size + del Im(sum_c F_c F_adj_c)) per
        design cell, eps terms minus mu terms (emopt/experimental/fdtd.py:
        37-158), with the field products formed on the GPU.  planar sums the
        products over z first.  sig_eps / sig_mu = None weigh every cell by 1
        (an unbounded problem, as the reference documents; its code would
        multiply by None)."""
        P = self.field_products(domain, "E")
        s_e = 1.0 if sig_eps is None else sig_eps
        if planar:
            grad_eps = s_e * (lam / np.size(s_e) + 2 * del_eps * np.imag(
                np.sum(P[0], axis=0) + np.sum(P[1], axis=0) + np.sum(P[2], axis=0)))
        else:
            grad_eps = s_e * (lam / np.size(s_e) + 2 * del_eps * np.imag(P[0] + P[1] + P[2]))
        if not update_mu:
            return np.ravel(grad_eps)
        Q = self.field_products(domain, "H")
        s_m = 1.0 if sig_mu is None else sig_mu
        if planar:
            grad_mu = s_m * (lam / np.size(s_m) - 2 * del_mu * np.imag(
                np.sum(Q[0], axis=0) + np.sum(Q[1], axis=0) + np.sum(Q[2], axis=0)))
        else:
            grad_mu = s_m * (lam / np.size(s_m) - 2 * del_mu * np.imag(Q[0] + Q[1] + Q[2]))
        return np.concatenate([np.ravel(grad_eps), np.ravel(grad_mu)], axis=0)

    def calc_ydAx_autograd(self, domain, update_mu, eps_autograd, mu_autograd):
        """Pseudo-loss 2 Im sum(eps o E o E_adj - mu o H o H_adj) over `domain`
        for reverse-mode differentiation (emopt/experimental/fdtd.py:160-223):
        its gradient with respect to the torch tensors eps_autograd (and
        mu_autograd with update_mu) is the adjoint gradient per cell.  The
        field products are formed on the GPU; torch moves them to the tensors'
        device.  Tensors on the GPU need torch imported before this library is
        first used: a HIP runtime that this library initialised first leaves
        torch without a device."""
        import torch
        dev = eps_autograd[0].device
        P = torch.as_tensor(self.field_products(domain, "E")).to(dev)
        loss = eps_autograd[0] * P[0] + eps_autograd[1] * P[1] + eps_autograd[2] * P[2]
        if update_mu:
            Q = torch.as_tensor(self.field_products(domain, "H")).to(dev)
            loss = loss - mu_autograd[0] * Q[0] - mu_autograd[1] * Q[1] - mu_autograd[2] * Q[2]
        return 2 * loss.imag.sum()

    def calc_ydAx(self, Adiag0, box=None):
        """y^T (A - A0) x on the device (emopt/fdtd.py:1694-1717), the eps
        and mu terms.  Adiag0 holds full-grid arrays (as get_A_diag returns
        them: eps_x..eps_z, mu_x..mu_z; with only three, mu is taken as
        unchanged); with box = (k1, k2, j1, j2, i1, i2) only that index box is
        summed (the rest of A - A0 is zero)."""
        out = np.zeros(2)
        k1, k2, j1, j2, i1, i2 = box if box is not None else (0, self._Nx, 0, self._Ny, 0, self._Nz)
        shp = (self._Nz, self._Ny, self._Nx)
        z = [c128(np.asarray(a).reshape(shp)[i1:i2, j1:j2, k1:k2]).ravel() for a in Adiag0[:6]]
        if len(z) < 6:
            z = z[:3] + [None] * 3
        self._lib.emf_calc_ydAx(self._h, *[ptr(a) for a in z], k1, k2, j1, j2, i1, i2, ptr(out))
        return complex(out[0], out[1])

    def pml_stretch(self, comp, cell):
        """CPML stretch product s_x s_y s_z at the positions of E component
        `comp` of the given cells (1 outside the CPML slabs); see
        emf_pml_stretch."""
        comp = np.ascontiguousarray(comp, dtype=np.int32)
        cell = np.ascontiguousarray(cell, dtype=np.int64)
        out = np.zeros(2 * len(comp))
        self._lib.emf_pml_stretch(self._h, len(comp), ptr(comp), ptr(cell), ptr(out))
        return out[0::2] + 1j * out[1::2]

    def grad_assemble(self, pidx, comp, cell, deps, nparam, step=1.0, noalloc=False):
        """grad[p] = -2 Re(sum E_adj * 1j * deps * E) / step over sparse
        (parameter, component, cell, dEps) entries, on the device
        (k_grad): the reference's per-parameter loop
        (emopt/adjoint_method.py:504-533) in one pass.  Components 3..5 are
        mu_x..mu_z entries (term -1j * dmu * H_adj * H).  noalloc=True uses
        the entry point that allocates no device memory (up to ~1 M
        entries)."""
        pidx = np.asarray(pidx, dtype=np.int32)
        order = np.argsort(pidx, kind="stable")
        pidx = np.ascontiguousarray(pidx[order])
        comp = np.ascontiguousarray(np.asarray(comp, dtype=np.int32)[order])
        cell = np.ascontiguousarray(np.asarray(cell, dtype=np.int64)[order])
        deps = c128(np.asarray(deps, dtype=np.complex128)[order])
        grad = np.zeros(int(nparam))
        fn = self._lib.emf_grad_assemble_noalloc if noalloc else self._lib.emf_grad_assemble
        fn(self._h, len(pidx), ptr(pidx), ptr(comp), ptr(cell), ptr(deps), int(nparam),
           float(step), ptr(grad))
        return grad

    # ---------------------------------------------------------- step-level
    def time_fields(self):
        """Current real time-domain fields (Ex..Hz), each [Nz, Ny, Nx]."""
        n = self._Nx * self._Ny * self._Nz
        arrs = [np.empty(n) for _ in range(6)]
        self._lib.emf_get_time_fields(self._h, *[ptr(a) for a in arrs])
        return [a.reshape(self._Nz, self._Ny, self._Nx) for a in arrs]
