"""get_field_interp on the GPU (emf_get_field_interp) against the numpy
restatement of emopt/fdtd.py:1401-1553 in emopt_amd/fields.py, on the same
solved fields: every component, forward and adjoint slots, domains inside
the grid and touching each min and max face, under bc '000', 'EH0' and 'HE0'
(the 'E'/'H' mirrors at the x and y min faces).  The kernel sums the stencil
in the same order and divides once, so the values are identical
(np.array_equal); the numpy restatement itself is pinned by the closed-form
checks of tests/test_host.py."""
import warnings

import numpy as np
import pytest

from emopt_amd import fields
from emopt_amd.misc import DomainCoordinates
from tests.cases import build_sim, strip_case

pytestmark = pytest.mark.gpu

COMPS = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")


class _HostOnly:
    """The solver without its device interpolation: fields.interp_to_ez then
    takes the numpy path on the same downloaded fields."""

    def __init__(self, sim):
        self._sim = sim

    def __getattr__(self, name):
        if name == "_interp_device":
            raise AttributeError(name)
        return getattr(self._sim, name)

    def get_field_interp(self, component, domain=None, squeeze=False, adjoint=False):
        return fields.interp_to_ez(self, component, domain, squeeze, adjoint)


def _domains(sim):
    Nx, Ny, Nz, d = sim.Nx, sim.Ny, sim.Nz, sim.dx
    box = lambda *b: DomainCoordinates.from_indices(*b, d, d, d)
    return [None,
            box(5, 20, 4, 17, 3, 15),              # interior
            box(0, 1, 0, Ny, 0, Nz),               # x = 0 plane (x mirror)
            box(0, Nx, 0, 1, 0, Nz),               # y = 0 plane (y mirror)
            box(0, 7, 0, 6, 0, 5),                 # the min corner (both mirrors)
            box(Nx - 1, Nx, 0, Ny, 0, Nz),         # the last x plane
            box(Nx - 2, Nx - 1, Ny - 3, Ny, 0, Nz),  # ends one short of the max face
            box(0, Nx, 0, Ny, Nz // 2, Nz // 2 + 1),  # a z plane
            DomainCoordinates(0.31, 0.93, 0.17, 0.81, 0.42, 0.42, d, d, d)]


@pytest.mark.parametrize("bc", ["000", "EH0", "HE0"])
def test_device_interp_matches_numpy(bc):
    case = strip_case(36, 30, 24, w=6, bc=bc)
    sim = build_sim(case)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        sim.solve_forward()
        src, (k1, k2, j1, j2, i1, i2) = case["sources"][0]
        sim.add_adjoint_sources(src, DomainCoordinates.from_indices(k1, k2, j1, j2, i1, i2,
                                                                    sim.dx, sim.dy, sim.dz))
        sim.solve_adjoint()
    host = _HostOnly(sim)
    checked = 0
    for adjoint in (False, True):
        for dom in _domains(sim):
            for c in COMPS:
                g = fields.interp_to_ez(sim, c, dom, adjoint=adjoint)
                h = fields.interp_to_ez(host, c, dom, adjoint=adjoint)
                assert g.shape == h.shape, (bc, c, dom and (dom.k1, dom.k2, dom.j1, dom.j2))
                assert np.array_equal(g, h), (bc, adjoint, c, np.abs(g - h).max())
                checked += g.size > 0 and np.any(g != 0)
    assert checked > 60
    # the public entry takes the device path and agrees too
    d = _domains(sim)[1]
    assert np.array_equal(sim.get_field_interp("Hz", d), fields.interp_to_ez(host, "Hz", d))


def test_device_interp_source_power():
    """get_source_power (24 interpolations on the box one cell inside the
    CPML) is the same number on the device and numpy paths."""
    sim = build_sim(strip_case(36, 30, 24, w=6))
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        sim.solve_forward()
    p_dev = fields.source_power(sim)
    p_host = fields.source_power(_HostOnly(sim))
    assert p_dev == p_host and p_dev > 0
