"""Adjoint gradients of CW-FDTD figures of merit, and a differentiable solve.

Restated from EMopt (semantics and conventions kept, so a user's adjoint
method ports unchanged):

* :class:`ModeMatch`             emopt/fomutils.py:925-1196
* :func:`interpolated_dFdx_3D`   emopt/fomutils.py:1282-1429
* :func:`power_norm_dFdx_3D`     emopt/fomutils.py:1654-1920
* :class:`AdjointMethod`         emopt/adjoint_method.py:97-547 (the FDTD
  subset: fom, dFdx, adjoint solve, the per-parameter gradient loop; the
  device assembly adds the CPML-stretch and discrete-frequency factors that
  make the gradient exact for the time-stepped system, see gradient_device)
* :class:`AdjointMethodPNF3D`    emopt/adjoint_method.py:1329-1486

Conventions (those of the reference): dFdx are the derivatives of F with
respect to the complex fields (z, not conj(z)); they are injected as the
adjoint J (from dF/dE) and M (from dF/dH) currents; the gradient is
``-2 Re(sum E_adj * 1j * dA/dp * E)``, A's diagonal being eps (mu is never a
design variable here).

On top of that, :class:`FDTDAdjointFunction` is a ``torch.autograd.Function``:
forward = material update + forward solve + FOM; backward = adjoint solve +
the hand-written gradient-assembly kernel (``emf_grad_assemble``) over the
boundary cells whose eps moves with each parameter, fed by
``AdjointMethod.calc_deps`` (sparse dEps/dp, exact for the piecewise-linear
fill fractions of moving rectangle edges).
"""
from __future__ import annotations

import warnings

import numpy as np

from .misc import DomainCoordinates

__all__ = ["ModeMatch", "interpolated_dFdx_3D", "power_norm_dFdx_3D",
           "power_norm_dFdx_3D_device", "AdjointMethod",
           "AdjointMethodPNF3D", "FDTDAdjointFunction", "fdtd_fom", "rect_edge_deps"]


class ModeMatch:
    """Projection of fields onto reference fields through a plane
    (emopt/fomutils.py:925-1196): ``mode_match_fwd`` is the power of the
    fields carried by the reference mode."""

    def __init__(self, normal, ds1, ds2=1.0, Exm=None, Eym=None, Ezm=None, Hxm=None, Hym=None,
                 Hzm=None):
        given = [Exm, Eym, Ezm, Hxm, Hym, Hzm]
        shp = next((np.shape(f) for f in given if f is not None), None)
        if shp is None:
            raise ValueError("ModeMatch needs at least one reference field")
        z = lambda f: np.zeros(shp, dtype=np.complex128) if f is None else np.asarray(f)
        self.fshape = shp
        self.Exm, self.Eym, self.Ezm = z(Exm), z(Eym), z(Ezm)
        self.Hxm, self.Hym, self.Hzm = z(Hxm), z(Hym), z(Hzm)
        self.normal = np.array(normal)
        self.ds1, self.ds2 = ds1, ds2
        self.sx, self.sy, self.sz = (float(v) for v in self.normal)
        ds = ds1 * ds2
        Pxm = self.Eym * np.conj(self.Hzm) - self.Ezm * np.conj(self.Hym)
        Pym = -self.Exm * np.conj(self.Hzm) + self.Ezm * np.conj(self.Hxm)
        Pzm = self.Exm * np.conj(self.Hym) - self.Eym * np.conj(self.Hxm)
        self.Pm = ds * np.sum(self.sx * Pxm + self.sy * Pym + self.sz * Pzm)
        self.am = self.bm = 0.0
        self.mode_match_fwd = self.mode_match_back = 0.0

    def compute(self, Ex=None, Ey=None, Ez=None, Hx=None, Hy=None, Hz=None):
        z = lambda f: np.zeros(self.fshape, dtype=np.complex128) if f is None else np.asarray(f)
        Ex, Ey, Ez, Hx, Hy, Hz = z(Ex), z(Ey), z(Ez), z(Hx), z(Hy), z(Hz)
        ds = self.ds1 * self.ds2
        sx, sy, sz = self.sx, self.sy, self.sz
        c = np.conj
        self.am1 = 0.5 * ds * np.sum((Ey * c(self.Hzm) - Ez * c(self.Hym)) * sx
                                     - (Ex * c(self.Hzm) - Ez * c(self.Hxm)) * sy
                                     + (Ex * c(self.Hym) - Ey * c(self.Hxm)) * sz) / self.Pm
        self.am2 = 0.5 * ds * np.sum((c(self.Eym) * Hz - c(self.Ezm) * Hy) * sx
                                     - (c(self.Exm) * Hz - c(self.Ezm) * Hx) * sy
                                     + (c(self.Exm) * Hy - c(self.Eym) * Hx) * sz) / c(self.Pm)
        self.am = self.am1 + self.am2
        self.bm = self.am1 - self.am2
        self.mode_match_fwd = 0.5 * self.am * np.conj(self.am) * np.real(self.Pm)
        self.mode_match_back = 0.5 * self.bm * np.conj(self.bm) * np.real(self.Pm)

    def _weights(self):
        """Weight planes W (Ex..Hz) with am1 = 0.5 ds sum(W_E . E) / Pm and
        am2 = 0.5 ds sum(W_H . H) / conj(Pm): compute()'s sums regrouped per
        field component."""
        c = np.conj
        sx, sy, sz = self.sx, self.sy, self.sz
        return (-sy * c(self.Hzm) + sz * c(self.Hym), sx * c(self.Hzm) - sz * c(self.Hxm),
                -sx * c(self.Hym) + sy * c(self.Hxm), sy * c(self.Ezm) - sz * c(self.Eym),
                -sx * c(self.Ezm) + sz * c(self.Exm), sx * c(self.Eym) - sy * c(self.Exm))

    def compute_device(self, sim, domain, adjoint=False, weight_set=0):
        """compute() on the solver's GPU: the overlap sums are reduced there
        over the fields interpolated on `domain` (emf_interp_dot), with the
        weight planes uploaded once per solver, so no field plane crosses to
        the host.  Same results as compute() on get_field_interp's planes up
        to the summation order."""
        key = (id(sim), weight_set)
        if getattr(self, "_dev_key", None) != key:
            sim.set_plane_weights(weight_set, self._weights())
            self._dev_key = key
        se, sh = sim.interp_dot(weight_set, domain, adjoint=adjoint)
        ds = self.ds1 * self.ds2
        self.am1 = 0.5 * ds * se / self.Pm
        self.am2 = 0.5 * ds * sh / np.conj(self.Pm)
        self.am = self.am1 + self.am2
        self.bm = self.am1 - self.am2
        self.mode_match_fwd = 0.5 * self.am * np.conj(self.am) * np.real(self.Pm)
        self.mode_match_back = 0.5 * self.bm * np.conj(self.bm) * np.real(self.Pm)

    def dFdx_device(self, sim, weight_set=0, mask=0b111111):
        """The get_dFd* planes in the form the device staging takes:
        (weight_set, scale, mask) with dF/dfield_q = scale_q * W_q for the
        weight planes W already on `sim` (compute_device uploads them).  Call
        after compute_device; a component whose mask bit is clear is zero."""
        key = (id(sim), weight_set)
        if getattr(self, "_dev_key", None) != key:
            sim.set_plane_weights(weight_set, self._weights())
            self._dev_key = key
        kE, kH = self._k(False), self._k(True)
        return weight_set, (kE, kE, kE, kH, kH, kH), mask

    def get_mode_match_forward(self, P_in):
        return np.real(self.mode_match_fwd) / P_in

    def get_mode_match_back(self, P_in):
        return np.real(self.mode_match_back) / P_in

    # dF/dfield of the unnormalised forward mode match (emopt/fomutils.py:1118-1196)
    def _k(self, conj_pm):
        ds = self.ds1 * self.ds2
        Pm = np.conj(self.Pm) if conj_pm else self.Pm
        return 0.25 * ds * np.real(self.Pm) * np.conj(self.am) / Pm

    def get_dFdEx(self):
        return self._k(False) * (-np.conj(self.Hzm) * self.sy + np.conj(self.Hym) * self.sz)

    def get_dFdEy(self):
        return self._k(False) * (np.conj(self.Hzm) * self.sx - np.conj(self.Hxm) * self.sz)

    def get_dFdEz(self):
        return self._k(False) * (-np.conj(self.Hym) * self.sx + np.conj(self.Hxm) * self.sy)

    def get_dFdHx(self):
        return self._k(True) * (np.conj(self.Ezm) * self.sy - np.conj(self.Eym) * self.sz)

    def get_dFdHy(self):
        return self._k(True) * (-np.conj(self.Ezm) * self.sx + np.conj(self.Exm) * self.sz)

    def get_dFdHz(self):
        return self._k(True) * (np.conj(self.Eym) * self.sx - np.conj(self.Exm) * self.sy)


def interpolated_dFdx_3D(sim, domain, dFdExi, dFdEyi, dFdEzi, dFdHxi, dFdHyi, dFdHzi):
    """Map derivatives with respect to fields interpolated onto the Ez grid
    (get_field_interp) back onto the Yee fields: the transpose of the
    interpolation stencils (emopt/fomutils.py:1282-1429).  Returns the grown
    domain (one cell more at -x, -y and +z) and six arrays on it."""
    Nz, Ny, Nx = domain.shape
    shp = (Nz + 2, Ny + 2, Nx + 2)
    I, O, M = slice(1, -1), slice(0, -2), slice(2, None)   # i/j/k, minus one, plus one
    bc = sim.bc

    def zeros():
        return np.zeros(shp, dtype=np.complex128)

    dEx = zeros()
    for s in ((I, I, I), (I, I, O), (M, I, I), (M, I, O)):
        dEx[s] += dFdExi / 4.0
    if domain.k1 == 0 and bc[0] == "H":
        dEx[I, I, 1] += dFdExi[:, :, 0] / 4.0
        dEx[M, I, 1] += dFdExi[:, :, 0] / 4.0
    elif domain.k1 == 0 and bc[0] == "E":
        # as in the reference, this branch indexes the y axis
        dEx[I, 1, I] -= dFdExi[:, 0, :] / 4.0
        dEx[M, 1, I] -= dFdExi[:, 0, :] / 4.0

    dEy = zeros()
    for s in ((I, I, I), (I, O, I), (M, I, I), (M, O, I)):
        dEy[s] += dFdEyi / 4.0
    if domain.j1 == 0 and bc[1] in "HE":
        sgn = 1.0 if bc[1] == "H" else -1.0
        dEy[I, 1, I] += sgn * dFdEyi[:, 0, :] / 4.0
        dEy[M, 1, I] += sgn * dFdEyi[:, 0, :] / 4.0

    dEz = zeros()
    dEz[I, I, I] = dFdEzi

    dHx = zeros()
    for s in ((I, I, I), (I, O, I)):
        dHx[s] += dFdHxi / 2.0
    if domain.j1 == 0 and bc[1] in "HE":
        sgn = 1.0 if bc[1] == "H" else -1.0
        dHx[I, 1, I] += sgn * dFdHxi[:, 0, :] / 2.0

    dHy = zeros()
    for s in ((I, I, I), (I, I, O)):
        dHy[s] += dFdHyi / 2.0

    dHz = zeros()
    for s in ((I, I, I), (I, O, I), (I, I, O), (I, O, O), (M, I, I), (M, O, I), (M, I, O),
              (M, O, O)):
        dHz[s] += dFdHzi / 8.0
    if domain.j1 == 0 and bc[1] in "HE":
        sgn = 1.0 if bc[1] == "H" else -1.0
        for s in ((I, 1, I), (I, 1, O), (M, 1, I), (M, 1, O)):
            dHz[s] += sgn * dFdHzi[:, 0, :] / 8.0

    crop = (slice(1, None), slice(0, -1), slice(0, -1))
    out = [a[crop] for a in (dEx, dEy, dEz, dHx, dHy, dHz)]
    dom = domain.copy()
    dom.set_region(dom.k1 - 1, dom.k2, dom.j1 - 1, dom.j2, dom.i1, dom.i2 + 1, dom.dx, dom.dy,
                   dom.dz)
    return (dom, *out)


def _box_planes(sim):
    """The six planes of the source-power box (emopt/fomutils.py:1710-1745)."""
    w = sim.w_pml
    dx, dy, dz = sim.dx, sim.dy, sim.dz
    X, Y, Z = sim.X, sim.Y, sim.Z
    xmin = w[0] + dx if w[0] > 0 else 0.0
    xmax = X - w[1] - dx if w[1] > 0 else X
    ymin = w[2] + dy if w[2] > 0 else 0.0
    ymax = Y - w[3] - dy if w[3] > 0 else Y
    zmin = w[4] + dz if w[4] > 0 else 0.0
    zmax = Z - w[5] - dz if w[5] > 0 else Z
    D = lambda *a: DomainCoordinates(*a, dx, dy, dz)
    return (D(xmin, xmin, ymin, ymax, zmin, zmax), D(xmax, xmax, ymin, ymax, zmin, zmax),
            D(xmin, xmax, ymin, ymin, zmin, zmax), D(xmin, xmax, ymax, ymax, zmin, zmax),
            D(xmin, xmax, ymin, ymax, zmin, zmin), D(xmin, xmax, ymin, ymax, zmax, zmax))


def power_norm_dFdx_3D(sim, f, domain, dfdEx, dfdEy, dfdEz, dfdHx, dfdHy, dfdHz):
    """Adjoint sources of F = f / Psrc (emopt/fomutils.py:1654-1920): the
    interpolation-compensated df/dx / Psrc on `domain`, plus -f dPsrc/dx /
    Psrc^2 on the six planes of the source-power box (min planes skipped under
    an 'E'/'H' boundary).  Returns [sources, domains] for
    ``set_adjoint_sources``."""
    Psrc = sim.source_power
    srcs, doms = [], []
    d = interpolated_dFdx_3D(sim, domain, dfdEx / Psrc, dfdEy / Psrc, dfdEz / Psrc,
                             dfdHx / Psrc, dfdHy / Psrc, dfdHz / Psrc)
    srcs.append(d[1:]); doms.append(d[0])
    dx, dy, dz = sim.dx, sim.dy, sim.dz
    c = np.conj
    g = sim.get_field_interp
    x1, x2, y1, y2, z1, z2 = _box_planes(sim)
    # (plane, sign, skip-if-symmetric axis, components (E1, E2, H1, H2), area)
    # dS/dE and dS/dH of +/- 0.5 Re(E1 H2* - E2 H1*) for the plane's normal
    faces = ((x1, -1, 0, ("Ey", "Ez", "Hy", "Hz"), dy * dz),
             (x2, +1, None, ("Ey", "Ez", "Hy", "Hz"), dy * dz),
             (y1, -1, 1, ("Ez", "Ex", "Hz", "Hx"), dx * dz),
             (y2, +1, None, ("Ez", "Ex", "Hz", "Hx"), dx * dz),
             (z1, -1, 2, ("Ex", "Ey", "Hx", "Hy"), dx * dy),
             (z2, +1, None, ("Ex", "Ey", "Hx", "Hy"), dx * dy))
    order = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")
    for dom, sgn, sym_axis, (e1, e2, h1, h2), area in faces:
        if sym_axis is not None and sim.bc[sym_axis] in "EH":
            continue
        F = {k: g(k, dom) for k in (e1, e2, h1, h2)}
        dS = {k: np.zeros(dom.shape, dtype=np.complex128) for k in order}
        # S_n = 0.5 Re(E1 H2* - E2 H1*): dS/dE1 = H2*/4, dS/dE2 = -H1*/4,
        # dS/dH1 = -E2*/4, dS/dH2 = E1*/4 (times the plane's sign and area)
        dS[e1] = sgn * 0.25 * area * c(F[h2])
        dS[e2] = -sgn * 0.25 * area * c(F[h1])
        dS[h1] = -sgn * 0.25 * area * c(F[e2])
        dS[h2] = sgn * 0.25 * area * c(F[e1])
        dP = [-1 * f * dS[k] / Psrc ** 2 for k in order]
        d = interpolated_dFdx_3D(sim, dom, *dP)
        srcs.append(d[1:]); doms.append(d[0])
    return [srcs, doms]


_FACE_COMPS = (("Ey", "Ez", "Hy", "Hz"), ("Ez", "Ex", "Hz", "Hx"), ("Ex", "Ey", "Hx", "Hy"))


def power_norm_dFdx_3D_device(sim, f, domain, weight_set, scale, mask=0b111111):
    """power_norm_dFdx_3D built on the GPU into sim's staged sources: the FOM
    term from the derivative planes scale_q * W_q (W the weight set, as
    ModeMatch.dFdx_device gives them) and the six face terms from the solved
    fields, each mapped back onto the Yee grid there.  The same blocks as
    power_norm_dFdx_3D's host arrays (the same arithmetic, in the same order
    as numpy evaluates it on an FMA-capable host).  Returns False, with
    nothing staged for the faces from the first face whose grown box leaves
    the grid on, when a face needs the host path."""
    from .defs import FIELD_INDEX
    Psrc = sim.source_power
    sim.stage_planes(weight_set, scale, Psrc, domain, mask)
    dx, dy, dz = sim.dx, sim.dy, sim.dz
    x1, x2, y1, y2, z1, z2 = _box_planes(sim)
    faces = ((x1, -1, 0, 0, dy * dz), (x2, +1, None, 0, dy * dz),
             (y1, -1, 1, 1, dx * dz), (y2, +1, None, 1, dx * dz),
             (z1, -1, 2, 2, dx * dy), (z2, +1, None, 2, dx * dy))
    mf = -1 * f
    rp2 = 1.0 / Psrc ** 2
    for dom, sgn, sym_axis, axis, area in faces:
        if sym_axis is not None and sim.bc[sym_axis] in "EH":
            continue
        comps = [FIELD_INDEX[c] for c in _FACE_COMPS[axis]]
        if not sim.stage_power_face(comps, sgn * 0.25 * area, mf, rp2, dom):
            return False
    return True


def rect_edge_deps(sim, eps, prim, attr, delta, bbox):
    """Sparse dEps/dp of moving one coordinate of a primitive: the cells of
    `bbox` (index box k1, k2, j1, j2, i1, i2) whose sampled eps changes,
    from a central difference of the material sampling with a step `delta`
    that stays inside one cell.  Rectangle fill fractions are linear in an
    edge position within a cell, so this is the exact derivative.  Returns
    (component, flat cell index, complex dEps) arrays."""
    k1, k2, j1, j2, i1, i2 = bbox
    comps, cells, vals = [], [], []
    v0 = getattr(prim, attr)
    # the solver's own sampling (structured materials on its GPU, the same
    # doubles as the host sampler), else the material's get_values
    sample = getattr(sim, "_sample", None)
    if sample is None:
        sample = lambda m, *box: m.get_values(*box[:6], sx=box[6], sy=box[7], sz=box[8])
    samples = []  # eps at p + delta and p - delta, per component
    for sgn in (+1, -1):
        setattr(prim, attr, v0 + sgn * delta)
        samples.append([sample(eps, k1, k2, j1, j2, i1, i2, sx, sy, sz)
                        for (sx, sy, sz) in sim.YEE_SHIFTS[:3]])
    setattr(prim, attr, v0)
    for c in range(3):
        d = (samples[0][c] - samples[1][c]) / (2 * delta)
        nz = np.nonzero(np.abs(d) > 0)
        if len(nz[0]) == 0:
            continue
        i = nz[0] + i1; j = nz[1] + j1; k = nz[2] + k1
        comps.append(np.full(len(i), c, dtype=np.int32))
        cells.append(((i.astype(np.int64) * sim.Ny + j) * sim.Nx + k))
        vals.append(d[nz])
    if not comps:
        return (np.zeros(0, np.int32), np.zeros(0, np.int64), np.zeros(0, np.complex128))
    return np.concatenate(comps), np.concatenate(cells), np.concatenate(vals)


class AdjointMethod:
    """The FDTD subset of emopt.adjoint_method.AdjointMethod
    (emopt/adjoint_method.py:97-547).  Subclasses define update_system,
    calc_fom, calc_dFdx, calc_grad_p and get_update_boxes as in the
    reference; optionally calc_deps (sparse dEps/dp) for the device gradient
    assembly used by :class:`FDTDAdjointFunction`.

    exact_adjoint (default True) is a choice of this package, not of the
    reference: gradient_device and fdtd_fom then scale dEps/dp by the CPML
    stretch product and the leapfrog's discrete frequency factor, giving the
    gradient of the time-stepped system (what finite differences of the FOM
    measure).  The reference's formula, -2 Re(E_adj i dEps E), is what
    gradient() and calc_gradient() return, and what gradient_device returns
    with exact=False; the two differ wherever the design crosses a CPML slab
    (by 16.5 % in norm on the 360x60x120 grating workload)."""

    exact_adjoint = True

    # build the adjoint sources on the GPU when the method offers it
    # (stage_dFdx); False: always the host path of calc_dFdx
    device_sources = True

    def __init__(self, sim, step=1e-8):
        self.sim = sim
        self._step = step
        self.prev_params = []          # params of the last forward solve (fom)

    @property
    def step(self):
        """Parameter step of the gradient loop (emopt/adjoint_method.py:
        122-128)."""
        return self._step

    @step.setter
    def step(self, val):
        self._step = val

    # --- user hooks (emopt/adjoint_method.py:130-260)
    def update_system(self, params):
        raise NotImplementedError

    def calc_fom(self, sim, params):
        raise NotImplementedError

    def calc_dFdx(self, sim, params):
        raise NotImplementedError

    def calc_grad_p(self, sim, params):
        return np.zeros(len(params))

    def stage_dFdx(self, sim, params):
        """Build the adjoint sources on the GPU (sim's staged sources) and
        return True, or return False for the host path (calc_dFdx).  The
        base class has no device form."""
        return False

    def get_update_boxes(self, sim, params):
        N = len(params)
        return [(0, sim.X, 0, sim.Y, 0, sim.Z) for _ in range(N)]

    def calc_deps(self, sim, params):
        """(pidx, comp, cell, deps) of every parameter; needed only by the
        device gradient assembly."""
        raise NotImplementedError

    # --- the solves
    def fom(self, params):
        """Forward solve and FOM (emopt/adjoint_method.py:418-447).  The
        params are recorded (as a copy: the reference keeps the caller's
        array itself, so an in-place change of it went unnoticed) for the
        forward-solve guard of gradient."""
        self.update_system(params)
        self.prev_params = np.array(params, dtype=np.double, copy=True)
        self.sim.update()
        self.sim.solve_forward()
        return self.calc_fom(self.sim, params)

    def _forward_current(self, params):
        """True when the last forward solve was at exactly these params."""
        return np.array_equal(self.prev_params, np.asarray(params, dtype=np.double))

    def prepare_adjoint(self, params):
        """Set the solver's adjoint sources from the last forward solve:
        built on the device when stage_dFdx can, else calc_dFdx's planes."""
        if self.device_sources and self.stage_dFdx(self.sim, params):
            self.sim.set_adjoint_sources_staged()
        else:
            src = self.calc_dFdx(self.sim, params)
            self.sim.set_adjoint_sources(src)

    def solve_adjoint_for(self, params):
        self.prepare_adjoint(params)
        self.sim.solve_adjoint()

    def calc_gradient(self, sim, params):
        """The reference's gradient loop (emopt/adjoint_method.py:448-547):
        per parameter, step it, re-sample eps in its update box(es), and
        take -2 Re(y^T (A - A0) x) / step."""
        Ai = sim.get_A_diag()
        step = self._step
        boxes = self.get_update_boxes(sim, params)
        params = np.array(params, dtype=np.double)
        grad = np.zeros(len(params))
        for i in range(len(params)):
            p0 = params[i]
            ub = boxes[i]
            ubs = ub if isinstance(ub[0], (list, tuple, np.ndarray)) else [ub]
            params[i] = p0 + step
            self.update_system(params)
            for box in ubs:
                sim.update(box)
            # A - A0 vanishes outside the update boxes: sum over their hull
            idx = [sim.index_box(box) for box in ubs]
            hull = (min(b[0] for b in idx), max(b[1] for b in idx), min(b[2] for b in idx),
                    max(b[3] for b in idx), min(b[4] for b in idx), max(b[5] for b in idx))
            if hull[0] >= hull[1] or hull[2] >= hull[3] or hull[4] >= hull[5]:
                grad[i] = 0.0
            else:
                grad[i] = -2 * np.real(sim.calc_ydAx(Ai, box=hull) / step)
            params[i] = p0
            self.update_system(params)
            for box in ubs:
                sim.update(box)
        return grad

    def gradient(self, params):
        """Forward solve (skipped when the last one was at these params),
        adjoint solve, reference gradient loop (emopt/adjoint_method.py:
        548-606)."""
        self.update_system(params)
        self.sim.update()
        if not self._forward_current(params):
            self.prev_params = np.array(params, dtype=np.double, copy=True)
            self.sim.solve_forward()
        self.solve_adjoint_for(params)
        return self.calc_gradient(self.sim, params) + self.calc_grad_p(self.sim, params)

    def check_gradient(self, params, indices=[], plot=True, verbose=True,
                       return_gradients=False, fd_step=1e-10):
        """Adjoint gradient against one-sided finite differences
        (emopt/adjoint_method.py:608-721): F(p_i + fd_step) - F(p) over
        fd_step for each index in `indices` (all when empty).  Returns the
        total relative error ||g_fd - g_am|| / ||g_fd|| (and, with
        return_gradients, also g_fd and the full adjoint gradient); warns
        when it is over 1 %.  plot draws the three bar charts when
        matplotlib is importable."""
        params = np.array(params, dtype=np.double)
        if len(indices) == 0:
            indices = np.arange(0, len(params), 1)
        indices = np.asarray(indices, dtype=int)
        self.update_system(params)
        self.sim.update()
        grad_am = self.gradient(params)
        grad_fd = np.zeros(len(indices))
        fom0 = self.fom(params)
        if verbose:
            print("Checking gradient...")
        for i, j in enumerate(indices):
            if verbose:
                print("\tDerivative %d of %d" % (i + 1, len(indices)))
            p0 = params[j]
            params[j] += fd_step
            fom1 = self.fom(params)
            grad_fd[i] = (fom1 - fom0) / fd_step
            params[j] = p0
        errors = np.abs(grad_fd - grad_am[indices]) / np.abs(grad_fd)
        error_tot = np.linalg.norm(grad_fd - grad_am[indices]) / np.linalg.norm(grad_fd)
        if error_tot < 0.01:
            if verbose:
                print("The total error in the gradient is %0.4E" % error_tot)
        else:
            warnings.warn("The total error in the gradient is %0.4E which is over 1%%" % error_tot)
        if plot:
            try:
                import matplotlib.pyplot as plt
            except ImportError:
                warnings.warn("matplotlib is not available: no gradient plot")
            else:
                f = plt.figure()
                axs = [f.add_subplot(311), f.add_subplot(312), f.add_subplot(313)]
                xs = np.arange(len(indices))
                for ax, v, title in zip(axs, (grad_fd, grad_am[indices], errors),
                                        ("Finite Differences", "Adjoint Method",
                                         "Error in Adjoint Method")):
                    ax.bar(xs, v)
                    ax.set_title(title)
                axs[2].set_yscale("log")
                plt.show()
        if return_gradients:
            return error_tot, grad_fd, grad_am
        return error_tot

    def gradient_device(self, params, fom_done=False, exact=None):
        """Forward (unless fom_done) and adjoint solves, then the device
        gradient assembly from calc_deps.

        exact (default: self.exact_adjoint) applies the two factors by which
        the time-stepped system's adjoint differs from the reference formula
        -2 Re(E_adj i dEps E): the CPML stretch product s_x s_y s_z at each
        perturbed cell (the CPML operator is complex symmetric only after
        that scaling, so perturbations inside the slabs need it) and the
        discrete frequency 2 sin(dt/2)/dt of the leapfrog (the CW steady
        state of the Yee scheme solves Maxwell's equations at that
        frequency).  With exact=False the result is the reference formula."""
        if not fom_done or not self._forward_current(params):
            self.fom(params)      # another forward solve ran since: redo this one
        self.solve_adjoint_for(params)
        pidx, comp, cell, deps = self.calc_deps(self.sim, params)
        if self.exact_adjoint if exact is None else exact:
            dt = self.sim.dt
            deps = deps * self.sim.pml_stretch(comp, cell) * (2.0 * np.sin(0.5 * dt) / dt)
        g = self.sim.grad_assemble(pidx, comp, cell, deps, len(params))
        return g + self.calc_grad_p(self.sim, params)


class AdjointMethodPNF3D(AdjointMethod):
    """Power-normalised FOM F = f / Psrc (emopt/adjoint_method.py:1329-1486):
    subclasses define calc_f, calc_dfdx (derivatives of f with respect to
    the interpolated fields on the FOM domain) and get_fom_domains."""

    def calc_f(self, sim, params):
        raise NotImplementedError

    def calc_dfdx(self, sim, params):
        raise NotImplementedError

    def calc_dfdx_device(self, sim, params):
        """Optional device form of calc_dfdx: per FOM domain, (weight_set,
        scale, mask) with df/dfield_q = scale_q * W_q for weight planes W on
        the solver (ModeMatch.dFdx_device).  None: calc_dfdx's planes are
        uploaded instead."""
        return None

    def get_fom_domains(self):
        raise NotImplementedError

    # weight set that carries calc_dfdx's host planes to the device
    HOST_PLANES_SET = 7

    def stage_dFdx(self, sim, params):
        """calc_dFdx's sources built on the GPU (power_norm_dFdx_3D_device).
        The derivative planes come from calc_dfdx_device when the subclass
        has it; otherwise calc_dfdx's host planes are uploaded (weight set
        HOST_PLANES_SET) with unit factors, which gives the same values, and
        the face terms and all transpositions still run on the device.
        False (the host path) when a plane does not have its domain's shape."""
        specs = self.calc_dfdx_device(sim, params)
        f = self.calc_f(sim, params)
        doms = self.get_fom_domains()
        host = None
        if specs is None:
            host = self.calc_dfdx(sim, params)
            for dfdx, dom in zip(host, doms):
                n = int(np.prod(dom.shape))
                if any(np.size(a) != n for a in dfdx):
                    return False
        sim.stage_clear()
        for q, dom in enumerate(doms):
            if host is not None:
                sim.set_plane_weights(self.HOST_PLANES_SET, host[q])
                spec = (self.HOST_PLANES_SET, (1.0,) * 6, 0b111111)
            else:
                spec = specs[q]
            if not power_norm_dFdx_3D_device(sim, f, dom, *spec):
                sim.stage_clear()
                return False
        return True

    def calc_penalty(self, sim, params):
        return 0.0

    def calc_fom(self, sim, params):
        return self.calc_f(sim, params) / sim.source_power + self.calc_penalty(sim, params)

    def calc_dFdx(self, sim, params):
        f = self.calc_f(sim, params)
        dfdxs = self.calc_dfdx(sim, params)
        srcs = [[], []]
        for dfdx, dom in zip(dfdxs, self.get_fom_domains()):
            a = power_norm_dFdx_3D(sim, f, dom, *dfdx)
            srcs = [srcs[0] + a[0], srcs[1] + a[1]]
        return srcs


class AdjointMethodMO(AdjointMethod):
    """A figure of merit made of several adjoint methods' figures of merit,
    e.g. one per wavelength of a broadband design
    (emopt/adjoint_method.py:723-935).  Subclasses define
    calc_total_fom(foms) and calc_total_gradient(foms, grads).

    batched=True (not in the reference) runs the members' forward solves
    together as one batch on the GPU (emopt_amd.batch.Batch), and then their
    adjoint solves together; each member's result equals its own solve to
    roundoff.  The members must then drive distinct solver objects.  A member
    the batch does not take (complex eps, non-uniform mu, a grid above 32-bit
    offsets) is solved on its own after the batch."""

    def __init__(self, ams, step=1e-6, batched=False):
        self._ams = list(ams)
        self._foms_current = np.zeros(len(self._ams))
        self._step = step
        self.prev_params = []
        self.sim = self._ams[0].sim if self._ams else None
        self.batched = batched
        self._batch = None

    @property
    def adjoint_methods(self):
        return self._ams

    @adjoint_methods.setter
    def adjoint_methods(self, new_ams):
        self._ams = list(new_ams)

    def update_system(self, params):
        for am in self._ams:
            am.update_system(params)

    def calc_fom(self, sim, params):
        return self.calc_total_fom(self._foms_current)

    def calc_dFdx(self, sim, params):
        return None        # each member forms its own adjoint sources

    def calc_grad_p(self, sim, params):
        return None        # part of each member's gradient

    def calc_total_fom(self, foms):
        raise NotImplementedError

    def calc_total_gradient(self, foms, grads):
        raise NotImplementedError

    # --- batched member solves
    def _batch_of(self):
        from .batch import Batch
        sims = [am.sim for am in self._ams]
        if len({id(s) for s in sims}) != len(sims):
            raise ValueError("batched=True needs a distinct solver per member")
        if self._batch is None:
            self._batch = Batch(device=sims[0]._device)
        return self._batch

    def _solve_together(self, sims, adjoint):
        """The solves of `sims` as one batch; those the batch refuses run on
        their own afterwards."""
        from ._lib import EmfError
        b = self._batch_of()
        alone = []
        for sim in sims:
            try:
                b.add(sim, adjoint=adjoint)
            except EmfError:
                alone.append(sim)
        b.run()
        for sim in alone:
            sim.solve_adjoint() if adjoint else sim.solve_forward()

    def _forward_all(self, params, only_stale):
        """Update every member's structure and run the forward solves (all
        of them, or those not already at params) as one batch."""
        todo = []
        for am in self._ams:
            am.update_system(params)
            am.sim.update()
            if only_stale and am._forward_current(params):
                continue
            am.prev_params = np.array(params, dtype=np.double, copy=True)
            todo.append(am.sim)
        self._solve_together(todo, adjoint=False)

    def fom(self, params):
        """The members' figures of merit (each its own forward solve) and
        their total (emopt/adjoint_method.py:841-868)."""
        if self.batched:
            self._forward_all(params, only_stale=False)
            foms = [am.calc_fom(am.sim, params) for am in self._ams]
        else:
            foms = [am.fom(params) for am in self._ams]
        self._foms_current = foms
        self.prev_params = np.array(params, dtype=np.double, copy=True)
        return self.calc_total_fom(foms)

    def gradient(self, params):
        """The members' gradients and figures of merit combined by
        calc_total_gradient (emopt/adjoint_method.py:870-899)."""
        if not self.batched:
            grads, foms = [], []
            for am in self._ams:
                grads.append(am.gradient(params))
                foms.append(am.calc_fom(am.sim, params))
            return self.calc_total_gradient(foms, grads)
        self._forward_all(params, only_stale=True)
        for am in self._ams:
            am.prepare_adjoint(params)
        self._solve_together([am.sim for am in self._ams], adjoint=True)
        grads = [am.calc_gradient(am.sim, params) + am.calc_grad_p(am.sim, params)
                 for am in self._ams]
        foms = [am.calc_fom(am.sim, params) for am in self._ams]
        return self.calc_total_gradient(foms, grads)

    def check_gradient(self, params, indices=[], plot=True, verbose=True,
                       return_gradients=False, fd_step=1e-10):
        """check_gradient of the total figure of merit
        (emopt/adjoint_method.py:901-935)."""
        self.sim = self._ams[0].sim
        for am in self._ams:
            am.update_system(params)
            am.sim.update()
        return super().check_gradient(params, indices, plot, verbose, return_gradients, fd_step)

    def close(self):
        if self._batch is not None:
            self._batch.close()
            self._batch = None


def _torch():
    import torch
    return torch


class FDTDAdjointFunction:
    """``torch.autograd.Function`` whose forward is the FDTD forward solve
    and FOM, and whose backward is the adjoint solve plus the device gradient
    assembly.  Use through :func:`fdtd_fom`."""
    _cls = None

    @classmethod
    def get(cls):
        if cls._cls is None:
            torch = _torch()

            class _Fn(torch.autograd.Function):
                @staticmethod
                def forward(ctx, params, am):
                    p = params.detach().cpu().numpy().astype(np.double).copy()
                    F = am.fom(p)
                    ctx.am = am
                    ctx.p = p
                    return torch.tensor(float(np.real(F)), dtype=torch.float64)

                @staticmethod
                def backward(ctx, gF):
                    g = ctx.am.gradient_device(ctx.p, fom_done=True)
                    return gF * torch.from_numpy(np.asarray(g, dtype=np.double)), None

            cls._cls = _Fn
        return cls._cls


def fdtd_fom(params, am):
    """Differentiable FOM: ``F = fdtd_fom(p, am); F.backward()`` gives
    p.grad from one forward and one adjoint solve.  The gradient is
    gradient_device's: with am.exact_adjoint (the default) the gradient of
    the time-stepped system, which differs from EMopt's formula where the
    design crosses a CPML slab; am.exact_adjoint = False gives the
    formula."""
    return FDTDAdjointFunction.get().apply(params, am)
