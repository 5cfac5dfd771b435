"""The full-vector mode solver (emopt_amd/modes.py) pinned by mathematics
(CPU only).

EMopt's ModeFullVector needs PETSc and SLEPc, which are not importable here,
so its eigenvalues are checked against closed forms instead:
  * a dielectric slab (Si 0.22 um in SiO2 at 1.55 um) uniform along the
    slice's other axis, in slices normal to x and to z: the first mode's
    effective index converges at second order in the cell size to the root
    of the slab's TE dispersion relation, and in the x-normal slice the
    second lies within 2e-2 of the TM root (the area-averaged eps across the
    interfaces limits it); a wrong stencil sign or material offset breaks
    either;
  * each returned pair satisfies A x = n B x;
  * the source of get_source carries unit power (|S| = 1 after the
    normalisation), and a backward mode is launched with the opposite phase
    factor.
"""
import numpy as np
import pytest
from scipy.optimize import brentq

from emopt_amd import grid
from emopt_amd.misc import DomainCoordinates
from emopt_amd.modes import ModeFullVector

WL, N1, N2, T = 1.55, 3.48, 1.444, 0.22
R = WL / (2 * np.pi)


def _slab_neff(tm):
    r = (N2 / N1) ** 2 if tm else 1.0

    def f(n):
        kx, g = np.sqrt(N1 ** 2 - n ** 2), np.sqrt(n ** 2 - N2 ** 2)
        return r * kx * np.tan(kx * T / R / 2) - g
    # the fundamental root lies where tan() is on its first branch
    lo = np.sqrt(max(N1 ** 2 - (np.pi * R / T) ** 2, N2 ** 2)) + 1e-12
    return brentq(f, lo, N1 - 1e-12)


def _slab_solver(d, normal="x", bc=("H", "0"), H=2.0, neigs=2, n0=None):
    Nt = int(round(H / d)) + 1
    if normal == "x":
        N = (3, 6, Nt)
        dom = DomainCoordinates.from_indices(1, 2, 0, 6, 0, Nt, d, d, d)
    else:
        N = (6, Nt, 3)
        dom = DomainCoordinates.from_indices(0, 6, 0, Nt, 1, 2, d, d, d)
    X, Y, Z = [(n - 1) * d for n in N]
    eps = grid.StructuredMaterial3D(X, Y, Z, d, d, d)
    big = grid.Rectangle(X / 2, Y / 2, 10 * X + 10, 10 * Y + 10, material_value=N2 ** 2, layer=2)
    eps.add_primitive(big, -1, 10)
    if normal == "x":           # slab normal to the grid's z axis
        eps.add_primitive(grid.Rectangle(X / 2, Y / 2, 10 * X + 10, 10 * Y + 10,
                                         material_value=N1 ** 2, layer=1), Z / 2 - T / 2,
                          Z / 2 + T / 2)
    else:                       # slab normal to the grid's y axis, constant in z
        eps.add_primitive(grid.Rectangle(X / 2, Y / 2, 10 * X + 10, T, material_value=N1 ** 2,
                                         layer=1), -1, 10)
    m = ModeFullVector(WL, eps, grid.ConstantMaterial3D(1.0), dom,
                       n0=(n0 if n0 is not None else _slab_neff(False) + 0.2), neigs=neigs,
                       verbose=False)
    m.bc = list(bc)
    m.build()
    m.solve()
    return m


@pytest.mark.parametrize("normal", ["x", "z"])
def test_slab_modes_converge_to_dispersion_roots(normal):
    te, tm = _slab_neff(False), _slab_neff(True)
    bc = ("H", "0") if normal == "x" else ("0", "H")
    errs = []
    for d in (0.02, 0.01, 0.005):
        m = _slab_solver(d, normal, bc)
        n = np.sort(m.neff.real)[::-1]
        assert np.all(np.abs(m.neff.imag) < 1e-10)
        errs.append((abs(n[0] - te), abs(n[1] - tm)))
        print("%s-normal d=%.3f: TE0 %.6f (%.6f) TM0 %.6f (%.6f)" % (normal, d, n[0], te, n[1], tm))
    errs = np.array(errs)
    # TE (E along the interfaces): second order, halving the cell divides the
    # error by about four
    assert errs[-1, 0] < 5e-4
    assert errs[1, 0] / errs[2, 0] > 3.0
    # TM (E across them; the x-normal slice, whose symmetric min face admits
    # it): the area-averaged eps of the interface cells (the reference's
    # sampling) leaves an error that depends on where the interfaces fall in
    # the cells, not on their size
    if normal == "x":
        assert np.all(errs[:, 1] < 2e-2)


def test_eigenpairs_and_source_power():
    m = _slab_solver(0.01)
    for i in range(2):
        x = m._x[i]
        r = m._A @ x - m.neff[i] * (m._B @ x)
        assert np.linalg.norm(r) <= 1e-8 * np.linalg.norm(m._A @ x)
    d = 0.01
    J = m.get_source(0, d, d, d)
    assert all(a.shape == m.domain.shape for a in J)
    # unit power of the normalised fields, as get_source normalises them
    g = lambda c: m.get_field(0, c, permute=False, squeeze=True)
    Ex, Ey, Hx, Hy = g("Ex"), g("Ey"), g("Hx"), g("Hy")
    S = 0.5 * m.dx * m.dy * np.sum(Ex * np.conj(Hy) - Ey * np.conj(Hx))
    rs = np.sqrt(S)
    S1 = 0.5 * m.dx * m.dy * np.sum(Ex / rs * np.conj(Hy / rs) - Ey / rs * np.conj(Hx / rs))
    assert abs(abs(S1) - 1.0) < 1e-12
    # x-normal: the current along the propagation axis (grid x) comes from
    # the slice's longitudinal terms; the transverse magnetic currents carry
    # the half-cell phase exp(i n dz / 2)
    Jx, Jy, Jz, Mx, My, Mz = J
    assert np.abs(Jy).max() > 0 and np.abs(Mz).max() > 0


def test_backwards_mode_has_opposite_direction():
    """A backward mode (dir = -1) of the same slab has the same index and
    its source has the conjugate half-cell phase on the E-derived magnetic
    currents."""
    from emopt_amd.misc import DomainCoordinates as DC
    d, H = 0.01, 2.0
    f = _slab_solver(d)
    Nt = int(round(H / d)) + 1
    b = ModeFullVector(WL, f.eps, f.mu, DC.from_indices(1, 2, 0, 6, 0, Nt, d, d, d),
                       n0=f.n0, neigs=1, backwards=True, verbose=False)
    b.bc = ["H", "0"]
    b.build()
    b.solve()
    assert abs(b.neff[0] - f.neff[0]) < 1e-9


def test_y_normal_slice_rejected():
    d = 0.05
    eps = grid.ConstantMaterial3D(2.0)
    with pytest.raises(NotImplementedError):
        ModeFullVector(WL, eps, eps, DomainCoordinates.from_indices(0, 5, 2, 3, 0, 5, d, d, d))


def test_mode_numbers_of_a_wide_guide():
    """A 2 um x 0.22 um Si strip in SiO2 (an MMI section), x-normal slice
    with PEC walls: its quasi-TE modes in order of decreasing index have 0, 1,
    2, ... sign changes across the width and none across the height.  The
    reference leaves get_mode_number empty, so the count is pinned by the
    guide's mode order instead."""
    d = 0.04
    N = (3, 81, 41)
    X, Y, Z = [(n - 1) * d for n in N]
    eps = grid.StructuredMaterial3D(X, Y, Z, d, d, d)
    eps.add_primitive(grid.Rectangle(X / 2, Y / 2, 10 * X, 10 * Y, material_value=N2 ** 2,
                                     layer=2), -1, 10)
    eps.add_primitive(grid.Rectangle(X / 2, Y / 2, 10 * X, 2.0, material_value=N1 ** 2, layer=1),
                      Z / 2 - T / 2, Z / 2 + T / 2)
    dom = DomainCoordinates.from_indices(1, 2, 0, N[1], 0, N[2], d, d, d)
    m = ModeFullVector(WL, eps, grid.ConstantMaterial3D(1.0), dom, n0=3.2, neigs=6,
                       verbose=False)
    m.build()
    m.solve()
    te = []
    for i in range(m.neigs):
        W = m.component_energy(i)
        # quasi-TE in this slice: the field lies along the width (grid y)
        if W[1] > W[2]:
            te.append((m.neff[i].real, m.get_mode_number(i)))
    te.sort(reverse=True)
    print("quasi-TE neff and mode numbers:", te)
    assert [pq for _, pq in te[:3]] == [(0, 0), (1, 0), (2, 0)]
    i1 = m.find_mode_index(1, 0)
    assert m.get_mode_number(i1) == (1, 0)
    with pytest.warns(UserWarning):
        assert m.find_mode_index(9, 9) == 0
