"""Full-vector eigenmodes of a 2D slice of a 3D structure, and the current
sources that launch them into the FDTD solver.

The interface of emopt.modes.ModeFullVector (emopt/modes.py:1099-2300):
construct with the wavelength, the eps and mu materials, a planar domain
(one cell thick along the propagation direction) and a guess n0 for the
effective index; build(), solve(), then neff[i], get_field(i, c),
get_field_interp(i, c) and get_source(i, dx, dy, dz), whose six current
arrays FDTD.set_sources accepts directly (emopt/fdtd.py:771-783).

The discretisation is the reference's: the slice's six field components on
a compressed 2D Yee cell, non-dimensional lengths (R = wavelength / 2 pi),
Maxwell's curl equations for fields ~ exp(i n z) written as the generalised
eigenproblem A x = n B x (B non-zero only in the four transverse-field rows),
the same finite-difference stencils, material offsets and boundary rows
('0' PEC, 'E' / 'H' symmetry at the min faces; the max faces are PEC) of
emopt/modes.py:1278-1588.  The reference hands A and B to SLEPc with a
shift-and-invert transform about n0 and MUMPS; here the same transform is
applied with scipy: the sparse LU of A - n0 B and ARPACK on
(A - n0 B)^-1 B, whose largest eigenvalues mu give the n = n0 + 1/mu
closest to n0.  The modes come out ordered by distance from n0, as SLEPc's
target ordering gives them.

SLEPc and PETSc are not available here, so the eigenvalues are pinned by
mathematics instead of by the reference (tests/test_modes.py): the TE mode
of a dielectric slab converges at second order in the cell size to the
root of the slab's dispersion relation, a homogeneous guide between PEC
walls to its closed form, the source of get_source carries unit power,
and a launched mode propagates one way (tests/test_gpu_modes.py).

Propagation along x and along z is supported; the reference's y-normal
branch maps the slice's first axis to the grid's z index while laying the
slice out with x first, which is only consistent for square slices, so a
y-normal domain is rejected.
"""
from __future__ import annotations

import warnings

import numpy as np

__all__ = ["ModeFullVector"]

_COMPS = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")
# component as stored for a slice with normal x / z (the solver's frame has
# the mode propagating along its own z): emopt/modes.py:1629-1649
_PERMUTE = {"x": {"Ex": "Ez", "Ey": "Ex", "Ez": "Ey", "Hx": "Hz", "Hy": "Hx", "Hz": "Hy"},
            "z": {c: c for c in _COMPS}}


class ModeFullVector:
    def __init__(self, wavelength, eps, mu, domain, n0=1.0, neigs=1, backwards=False,
                 verbose=True):
        self.wavelength = wavelength
        self.n0 = n0
        self.neigs = neigs
        self.eps = eps
        self.mu = mu
        self.domain = domain
        self._fshape = domain.shape
        if domain.Nx == 1:
            self.ndir = "x"
            self._M, self._N = domain.Nz, domain.Ny
            self.dx, self.dy = domain.dy, domain.dz
        elif domain.Nz == 1:
            self.ndir = "z"
            self._M, self._N = domain.Ny, domain.Nx
            self.dx, self.dy = domain.dx, domain.dy
        elif domain.Ny == 1:
            raise NotImplementedError("y-normal mode slices are not supported (see the module "
                                      "notes); use an x- or z-normal slice")
        else:
            raise AttributeError("3D domains are not supported for 2D mode calculations!")
        self.verbose = verbose
        self._dir = -1.0 if backwards else 1.0
        self.R = wavelength / (2 * np.pi)
        self._bc = ["0", "0"]
        self._neff = np.zeros(neigs, dtype=np.complex128)
        self._x = []
        self._A = self._B = None

    # --------------------------------------------------------- properties
    @property
    def neff(self):
        return self._neff

    @neff.setter
    def neff(self, value):
        warnings.warn("neff cannot be set by the user.")

    @property
    def bc(self):
        return self._bc

    @bc.setter
    def bc(self, bc):
        for b in bc[:2]:
            if b not in ("0", "E", "H"):
                raise ValueError("Boundary condition type '%s' is not supported here: use 0, E "
                                 "or H." % b)
        self._bc = list(bc[:2])

    # ----------------------------------------------------------- materials
    def _materials(self):
        """eps_x, eps_y, eps_z, mu_x, mu_y, mu_z of the slice as [M, N]
        arrays at the offsets of emopt/modes.py:1312-1352 (index
        coordinates; x-normal: the solver's x, y are the grid's y, z)."""
        d = self.domain
        M, N = self._M, self._N
        if self.ndir == "x":
            box = (d.k1, d.k1 + 1, d.j1, d.j1 + N, d.i1, d.i1 + M)
            shifts = ((0, 0.5, -0.5), (0, 0, 0), (0, 0, -0.5),     # eps_x, eps_y, eps_z
                      (0, 0, 0), (0, 0.5, -0.5), (0, 0.5, 0))      # mu_x, mu_y, mu_z
        else:
            box = (d.k1, d.k1 + N, d.j1, d.j1 + M, d.i1, d.i1 + 1)
            shifts = ((0, 0, 0),) * 6
        out = []
        for q, (sx, sy, sz) in enumerate(shifts):
            m = self.eps if q < 3 else self.mu
            v = m.get_values(*box, sx=sx, sy=sy, sz=sz)
            out.append(np.asarray(v, dtype=np.complex128).reshape(M, N))
        return out

    # --------------------------------------------------------------- build
    def build(self):
        """Assemble A and B of A x = n B x (emopt/modes.py:1278-1588)."""
        import scipy.sparse as sp
        if self.verbose:
            print("Building mode solver system matrix...")
        M, N = self._M, self._N
        odx = self.R / self.dx
        ody = self.R / self.dy
        ex, ey, ez, mx, my, mz = self._materials()
        bc = self._bc
        d = self._dir
        MN = M * N
        rows, cols, vals = [], [], []
        brow, bcol, bval = [], [], []

        def emit(I, ent):
            for J, v in ent.items():
                rows.append(I); cols.append(J); vals.append(v)

        for y in range(M):
            for x in range(N):
                c = y * N + x
                # (stuff) = n B Hy
                ent = {5 * MN + c: ody, c: 1j * ex[y, x]}
                if y > 0:
                    ent[5 * MN + c - N] = -ody
                if y == 0:
                    if bc[1] == "0":
                        ent[c] = 0
                    elif bc[1] == "E":
                        ent[5 * MN + c] = 2 * ody
                    elif bc[1] == "H":
                        ent[5 * MN + c] = 0
                        ent[c] = 0
                emit(c, ent)
                brow.append(c); bcol.append(4 * MN + c); bval.append(1j * d)
                # (stuff) = n B Hx
                I = MN + c
                ent = {5 * MN + c: -odx, MN + c: 1j * ey[y, x]}
                if x > 0:
                    ent[5 * MN + c - 1] = odx
                if x == 0:
                    if bc[0] == "0":
                        ent[MN + c] = 0
                    elif bc[0] == "E":
                        ent[5 * MN + c] = -2 * odx
                    elif bc[0] == "H":
                        ent[5 * MN + c] = 0
                emit(I, ent)
                brow.append(I); bcol.append(3 * MN + c); bval.append(-1j * d)
                # (stuff) = 0: the Ez row
                I = 2 * MN + c
                ent = {MN + c: -odx, c: ody, 5 * MN + c: -1j * mz[y, x]}
                if x < N - 1:
                    ent[MN + c + 1] = odx
                if y < M - 1:
                    ent[c + N] = -ody
                if x == 0 and bc[0] == "0":
                    ent[MN + c] = 0
                if y == 0 and bc[1] in ("0", "H"):
                    ent[c] = 0
                emit(I, ent)
                # (stuff) = n B Ey
                I = 3 * MN + c
                ent = {2 * MN + c: -ody, 3 * MN + c: -1j * mx[y, x]}
                if y < M - 1:
                    ent[2 * MN + c + N] = ody
                if x == 0 and bc[0] == "0":
                    ent[3 * MN + c] = 0
                if y == 0 and bc[1] in ("0", "H"):
                    ent[2 * MN + c] = 0
                emit(I, ent)
                brow.append(I); bcol.append(MN + c); bval.append(1j * d)
                # (stuff) = n B Ex
                I = 4 * MN + c
                ent = {2 * MN + c: odx, 4 * MN + c: -1j * my[y, x]}
                if x < N - 1:
                    ent[2 * MN + c + 1] = -odx
                if x == 0 and bc[0] == "0":
                    ent[2 * MN + c] = 0
                if y == 0 and bc[1] in ("0", "H"):
                    ent[4 * MN + c] = 0
                emit(I, ent)
                brow.append(I); bcol.append(c); bval.append(-1j * d)
                # (stuff) = 0: the Hz row
                I = 5 * MN + c
                ent = {4 * MN + c: odx, 3 * MN + c: -ody, 2 * MN + c: 1j * ez[y, x]}
                if x > 0:
                    ent[4 * MN + c - 1] = -odx
                if y > 0:
                    ent[3 * MN + c - N] = ody
                if x == 0:
                    if bc[0] == "0":
                        ent[3 * MN + c] = 0
                        if y > 0:
                            ent[3 * MN + c - N] = 0
                    elif bc[0] == "E":
                        ent[4 * MN + c] = 2 * odx
                    elif bc[0] == "H":
                        ent[4 * MN + c] = 0
                if y == 0:
                    if bc[1] == "0":
                        ent[4 * MN + c] = 0
                        if x > 0:
                            ent[4 * MN + c - 1] = 0
                    elif bc[1] == "E":
                        ent[3 * MN + c] = -2 * ody
                    elif bc[1] == "H":
                        ent[3 * MN + c] = 0
                        ent[4 * MN + c] = 0
                        if x > 0:
                            ent[4 * MN + c - 1] = 0
                emit(I, ent)
        n = 6 * MN
        self._A = sp.csc_matrix((np.array(vals, dtype=np.complex128), (rows, cols)), shape=(n, n))
        self._B = sp.csc_matrix((np.array(bval, dtype=np.complex128), (brow, bcol)), shape=(n, n))

    # --------------------------------------------------------------- solve
    def solve(self):
        """The neigs modes whose effective index is closest to n0."""
        import scipy.sparse.linalg as sla
        if self._A is None:
            self.build()
        if self.verbose:
            print("Solving...")
        sigma = complex(self.n0)
        lu = sla.splu((self._A - sigma * self._B).tocsc())
        B = self._B.tocsr()
        n = B.shape[0]
        op = sla.LinearOperator((n, n), matvec=lambda v: lu.solve(B @ v), dtype=np.complex128)
        k = min(self.neigs, n - 2)
        rng = np.random.default_rng(0)
        v0 = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        mu, vecs = sla.eigs(op, k=k, which="LM", v0=v0, ncv=min(n - 1, max(2 * k + 1, 20)))
        nvals = sigma + 1.0 / mu
        order = np.argsort(np.abs(nvals - sigma))
        if k < self.neigs:
            warnings.warn("%d eigenmodes were requested, however only %d eigenmodes were found."
                          % (self.neigs, k))
        self._x = [vecs[:, q] for q in order]
        for i, q in enumerate(order):
            self._neff[i] = nvals[q]

    # -------------------------------------------------------------- fields
    def _component(self, component, permute):
        return _PERMUTE[self.ndir][component] if permute else component

    def get_field(self, i, component, permute=True, squeeze=False):
        """Raw field component of mode i on the slice (emopt/modes.py:1651-1787)."""
        M, N = self._M, self._N
        c = _COMPS.index(self._component(component, permute))
        field = np.reshape(self._x[i][c * M * N:(c + 1) * M * N], (M, N))
        return field if squeeze else np.reshape(field, self._fshape)

    def get_field_interp(self, i, component, squeeze=False):
        """Field of mode i interpolated as the reference interpolates it
        (emopt/modes.py:1789-2052): onto the FDTD's Ez grid for a slice
        normal to x, onto the Ez points for a slice normal to z."""
        f_raw = np.pad(self.get_field(i, component, squeeze=True), 1, "constant",
                       constant_values=0)
        comp = self._component(component, True)
        bc = self._bc

        def mirror_x(f):
            if bc[0] == "E":
                f[:, 0] = -1 * f[:, 1]
            elif bc[0] == "H":
                f[:, 0] = f[:, 1]

        def mirror_y(f):
            if bc[1] == "E":
                f[0, :] = -1 * f[1, :]
            elif bc[1] == "H":
                f[0, :] = f[1, :]
        if self.ndir == "x":
            if comp in ("Ex", "Hy"):          # average along x and y
                out = np.copy(f_raw)
                mirror_x(f_raw)
                out[:, 1:] += f_raw[:, 0:-1]
                out[0:-1, 1:] += f_raw[1:, 0:-1]
                out[0:-1, :] += f_raw[1:, :]
                field = out[1:-1, 1:-1] / 4.0
            elif comp in ("Ey", "Hx"):        # no averaging
                field = f_raw[1:-1, 1:-1]
            elif comp == "Ez":                # average along y
                out = np.copy(f_raw)
                out[0:-1, :] += out[1:, :]
                field = out[1:-1, 1:-1] / 2.0
            else:                             # Hz: average along x
                out = np.copy(f_raw)
                mirror_x(f_raw)
                out[:, 1:] += f_raw[:, 0:-1]
                field = out[1:-1, 1:-1] / 2.0
        else:
            if comp == "Ex":
                out = np.copy(f_raw)
                mirror_x(f_raw)
                out[:, 1:] += f_raw[:, 0:-1]
                field = out[1:-1, 1:-1] / 2.0
            elif comp == "Ey":
                out = np.copy(f_raw)
                mirror_y(f_raw)
                out[1:, :] += f_raw[0:-1, :]
                field = out[1:-1, 1:-1] / 2.0
            elif comp == "Ez":
                field = f_raw[1:-1, 1:-1]
            elif comp == "Hx":
                out = np.copy(f_raw)
                mirror_y(f_raw)
                out[1:, :] += f_raw[0:-1, :]
                field = out[1:-1, 1:-1] / 2.0
            elif comp == "Hy":
                out = np.copy(f_raw)
                mirror_x(f_raw)
                out[:, 1:] += f_raw[:, 0:-1]
                field = out[1:-1, 1:-1] / 2.0
            else:
                out = np.copy(f_raw)
                mirror_y(f_raw)
                mirror_x(f_raw)
                out[:, 1:] += f_raw[:, 0:-1]
                out[1:, :] += f_raw[0:-1, :]
                out[1:, 1:] += f_raw[0:-1, 0:-1]
                field = out[1:-1, 1:-1] / 4.0
        return field if squeeze else np.reshape(field, self._fshape)

    def component_energy(self, i):
        """Fraction of the mode's energy in each component, Ex..Hz
        (emopt/modes.py:2054-2097)."""
        eps = np.squeeze(self.eps.get_values_in(self.domain))
        mu = np.squeeze(self.mu.get_values_in(self.domain))
        W = []
        for c in _COMPS:
            f = np.reshape(self.get_field(i, c, squeeze=True), eps.shape)
            W.append(np.sum((eps if c[0] == "E" else mu).real * np.abs(f) ** 2))
        Wtot = sum(W)
        return [w / Wtot for w in W]

    def get_mode_number(self, i):
        """(P, Q): the numbers of sign changes of mode i's stronger
        transverse electric component along the slice's first and second
        axes, through the component's largest value, after a global phase
        makes that value real.  Samples under 1e-3 of it (the tails, and the
        node itself) are skipped.  A slice with an E or H symmetry face holds
        half the mode, and the count is that half's.  The reference declares
        this method but leaves it empty (emopt/modes.py:2099-2118), so its
        find_mode_index cannot run."""
        M, N = self._M, self._N
        comps = [np.reshape(self._x[i][c * M * N:(c + 1) * M * N], (M, N)) for c in (0, 1)]
        f = max(comps, key=lambda a: float(np.sum(np.abs(a) ** 2)))
        peak = int(np.argmax(np.abs(f)))
        m0, n0 = divmod(peak, N)
        f = f * np.exp(-1j * np.angle(f.flat[peak]))
        tol = 1e-3 * abs(f.flat[peak])

        def changes(line):
            v = np.real(line)[np.abs(line) > tol]
            return int(np.count_nonzero(np.diff(np.sign(v)) != 0))

        return changes(f[m0, :]), changes(f[:, n0])

    def find_mode_index(self, P, Q):
        """Index of the first mode whose get_mode_number is (P, Q); 0 with a
        warning when there is none (emopt/modes.py:2120-2141)."""
        for i in range(self.neigs):
            if self.get_mode_number(i) == (P, Q):
                return i
        warnings.warn("Desired mode number was not found.")
        return 0

    # -------------------------------------------------------------- source
    def get_source(self, i, dx, dy, dz):
        """The current sources (Jx, Jy, Jz, Mx, My, Mz) on the slice that
        launch mode i (emopt/modes.py:2144-2298): the fields normalised to
        unit power, then the electric and magnetic currents of the
        discontinuity, with the half-cell phase exp(i n dz/2) along the
        propagation direction."""
        neff = self._neff[i]
        bc = self._bc
        M, N = self._M, self._N
        if self.ndir == "x":
            get_mu = lambda sx, sy: self.mu.get_values_in(self.domain, sy=sx, sz=sy,
                                                          squeeze=True)
            mydx, mydy, mydz = dy, dz, dx
        else:
            get_mu = lambda sx, sy: self.mu.get_values_in(self.domain, sx=sx, sy=sy,
                                                          squeeze=True)
            mydx, mydy, mydz = dx, dy, dz
        dx_, dy_, dz_ = mydx / self.R, mydy / self.R, mydz / self.R
        ekz = np.exp(1j * self._dir * neff * dz_ / 2)
        g = lambda c: self.get_field(i, c, permute=False, squeeze=True)
        Ex, Ey, Ez, Hx, Hy, Hz = (g(c) for c in _COMPS)
        S = 0.5 * mydx * mydy * np.sum(Ex * np.conj(Hy) - Ey * np.conj(Hx))
        rs = np.sqrt(S)
        Ex, Ey, Ez, Hx, Hy = Ex / rs, Ey / rs, Ez / rs, Hx / rs, Hy / rs
        z = lambda: np.zeros((M, N), dtype=np.complex128)
        Jx, Jy, Jz, Mx, My, Mz = z(), z(), z(), z(), z(), z()
        pad = lambda f: np.pad(f, 1, "constant", constant_values=0)
        Ex = pad(Ex)
        My += Ex[1:-1, 1:-1] * ekz / dz_
        Ey = pad(Ey)
        Mx += -Ey[1:-1, 1:-1] * ekz / dz_
        Ez = pad(Ez)
        eps = np.reshape(self.eps.get_values_in(self.domain, squeeze=True), (M, N))
        Jz += 1j * eps * Ez[1:-1, 1:-1]
        Mx += (Ez[2:, 1:-1] - Ez[1:-1, 1:-1]) / dy_
        My += -1 * (Ez[1:-1, 2:] - Ez[1:-1, 1:-1]) / dx_
        Hx = pad(Hx)
        if bc[1] == "E":
            Hx[0, :] = -1 * Hx[1, :]
        elif bc[1] == "H":
            Hx[0, :] = Hx[1, :]
        mu = np.reshape(get_mu(0.0, 0.5), (M, N))
        Jy += Hx[1:-1, 1:-1] / dz_
        Jz += -1 * (Hx[1:-1, 1:-1] - Hx[0:-2, 1:-1]) / dy_
        Mx += -1j * mu * Hx[1:-1, 1:-1]
        Hy = pad(Hy)
        if bc[0] == "E":
            Hy[:, 0] = -1 * Hy[:, 1]
        elif bc[0] == "H":
            Hy[:, 0] = Hy[:, 1]
        mu = np.reshape(get_mu(0.5, 0.0), (M, N))
        Jx += -Hy[1:-1, 1:-1] / dz_
        Jz += (Hy[1:-1, 1:-1] - Hy[1:-1, 0:-2]) / dx_
        My += -1j * mu * Hy[1:-1, 1:-1]
        out = [np.reshape(a, self.domain.shape) for a in (Jx, Jy, Jz, Mx, My, Mz)]
        Jx, Jy, Jz, Mx, My, Mz = out
        if self.ndir == "x":
            return Jz, Jx, Jy, Mz, Mx, My
        return Jx, Jy, Jz, Mx, My, Mz
