"""Adjoint sources built on the GPU (emf_stage_*) against the host path.

The host path forms the planes of power_norm_dFdx_3D and maps them back onto
the Yee grid with interpolated_dFdx_3D (emopt_amd/adjoint.py, a restatement
of emopt/fomutils.py:1282-1429,1654-1920 whose transposition is pinned by
tests/test_adjoint_host.py), then crops each block to the grid as
add_adjoint_sources does; emf_add_source stores a complex value v as
(Re v, -Im v).  The staged blocks must hold the same values: the device
repeats the host arithmetic in numpy's order (the transposition's terms in
the same sequence, scaled by powers of two; complex products as numpy forms
them with FMA; quotients by a real scalar as products with its reciprocal),
so on an FMA-capable host they agree bit for bit; the bar is 1e-14 of the
block's largest value.  Domains touch the x and y min faces under every
'E'/'H' boundary, which exercises the mirror terms.  End to end, the
gradient of the reduced grating with device-built adjoint sources equals
the one with host-built sources.
"""
import warnings

import numpy as np
import pytest

from emopt_amd.adjoint import interpolated_dFdx_3D, power_norm_dFdx_3D, power_norm_dFdx_3D_device
from emopt_amd.misc import DomainCoordinates
from tests.cases import build_sim, strip_case

pytestmark = pytest.mark.gpu

TOL = 1e-14


def _solved(bc):
    case = strip_case(30, 26, 20, w=5, bc=bc)
    sim = build_sim(case)
    sim.fixed_n0 = 6 * sim.Tn
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        sim.solve_forward()
    return sim


def _host_blocks(sim, src):
    """The host path's blocks: set_adjoint_sources crops them to the grid."""
    sim.set_adjoint_sources(src)
    out = [((b.i0, b.j0, b.k0, b.I, b.J, b.K),
            np.stack([a.reshape(b.I, b.J, b.K) for a in b.arrays])) for b in sim._adj_sources]
    sim.clear_adjoint_sources()
    return out


def _compare(sim, host):
    assert sim.stage_count() == len(host), (sim.stage_count(), len(host))
    exact = 0
    for q, (geom, arrs) in enumerate(host):
        g, dev = sim.staged_block(q)
        assert g == geom, (q, g, geom)
        want = arrs.real - 1j * arrs.imag          # what emf_add_source stores
        scale = max(np.abs(want).max(), 1e-300)
        err = np.abs(dev - want).max() / scale
        assert err <= TOL, (q, err)
        exact += np.array_equal(dev, want)
    return exact


def _box(sim, *b):
    return DomainCoordinates.from_indices(*b, sim.dx, sim.dy, sim.dz)


@pytest.mark.parametrize("bc", ["000", "H00", "E00", "0H0", "0E0", "HE0", "EH0"])
def test_staged_planes_equal_host_transposition(bc):
    sim = _solved(bc)
    rng = np.random.default_rng(len(bc) + sum(map(ord, bc)))
    doms = [_box(sim, 0, 12, 0, 10, 8, 9),         # z plane at the min corner (both mirrors)
            _box(sim, 0, 1, 3, 20, 2, 17),         # x = 0 plane
            _box(sim, 4, 22, 0, 1, 2, 17),         # y = 0 plane
            _box(sim, 5, 20, 4, 17, 10, 13),       # interior block
            _box(sim, sim.Nx - 6, sim.Nx, sim.Ny - 5, sim.Ny, sim.Nz - 3, sim.Nz)]   # max corner
    total_exact = 0
    for dom in doms:
        n = dom.shape
        W = [rng.standard_normal(n) + 1j * rng.standard_normal(n) for _ in range(6)]
        scale = rng.standard_normal(6) + 1j * rng.standard_normal(6)
        div = 1.37
        for mask in (0b111111, 0b011011):
            sim.set_plane_weights(0, W)
            sim.stage_clear()
            sim.stage_planes(0, scale, div, dom, mask)
            planes = [scale[q] * W[q] / div if mask >> q & 1 else np.zeros(n, np.complex128)
                      for q in range(6)]
            d = interpolated_dFdx_3D(sim, dom, *planes)
            total_exact += _compare(sim, _host_blocks(sim, [[d[1:]], [d[0]]]))
    print("bc %s: %d blocks bit-identical of %d" % (bc, total_exact, 2 * len(doms)))


@pytest.mark.parametrize("bc", ["000", "HE0", "0H0"])
def test_staged_power_norm_equals_host(bc):
    """power_norm_dFdx_3D on the device (FOM term from weight planes, six
    face terms from the solved fields) against the host function."""
    sim = _solved(bc)
    rng = np.random.default_rng(11)
    dom = _box(sim, 6, 24, 5, 21, 12, 13)
    n = dom.shape
    W = [rng.standard_normal(n) + 1j * rng.standard_normal(n) for _ in range(6)]
    kE, kH = complex(0.7, -0.2), complex(-0.1, 0.45)
    scale = (kE, kE, kE, kH, kH, kH)
    sim.set_plane_weights(0, W)
    f = 0.371
    sim.stage_clear()
    assert power_norm_dFdx_3D_device(sim, f, dom, 0, scale, 0b011011)
    dfdx = [scale[q] * W[q] if 0b011011 >> q & 1 else np.zeros(n, np.complex128)
            for q in range(6)]
    host = _host_blocks(sim, power_norm_dFdx_3D(sim, f, dom, *dfdx))
    nsym = sum(c in "EH" for c in bc[:3])
    assert len(host) == 7 - nsym
    exact = _compare(sim, host)
    print("bc %s: %d of %d blocks bit-identical" % (bc, exact, len(host)))


def test_staged_sources_drive_the_adjoint_solve():
    """The grating's gradient with adjoint sources staged on the GPU equals
    the one with host-built sources (same adjoint fields): with the FOM's
    planes formed on the device (calc_dfdx_device) and with calc_dfdx's host
    planes uploaded (a FOM without the device hook)."""
    from tests.cases import grating_problem
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        sim, am, p = grating_problem(Nx=120, Ny=24, Nz=50, d=0.04, ntr=8, w=8, t_box=0.6,
                                     w0=2.0, rtol=1e-6)
        am.fom(p)
        sim.fixed_n0 = 4 * sim.Tn * (1 + sim.last_solve["n0"] // (4 * sim.Tn))
        am.fom(p)
        gs = {}
        for dev in (True, False):
            am.device_sources = dev
            gs[dev] = am.gradient_device(p, fom_done=True)
            assert sim.last_solve["status"] == "fixed"
        assert sim.stage_count() == 7          # FOM plane + six faces
        am.device_sources = True
        am.calc_dfdx_device = lambda sim, params: None
        gs["upload"] = am.gradient_device(p, fom_done=True)
        assert sim.stage_count() == 7
    for key in (True, "upload"):
        err = np.abs(gs[key] - gs[False]).max() / np.abs(gs[False]).max()
        print("device (%s) vs host adjoint sources: max gradient difference %.2e "
              "(bit-identical: %s)" % (key, err, np.array_equal(gs[key], gs[False])))
        assert err <= 1e-12
