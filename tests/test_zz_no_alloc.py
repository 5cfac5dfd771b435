"""The allocation-free gradient assembly (emf_grad_assemble_noalloc) against
emf_grad_assemble, through the C ABI: two calls in a row on the same solved
problem, the second with fewer entries and parameters, each equal bit for bit
to the allocating entry point's result (same kernel, same entry order); a
request larger than the static scratch array is rejected with an error and
leaves the next call's result intact."""
import ctypes
import warnings

import numpy as np
import pytest

from tests.cases import build_sim, strip_case

pytestmark = pytest.mark.gpu


def _entries(rng, n, nparam, ncell):
    pidx = np.sort(rng.integers(0, nparam, size=n)).astype(np.int32)
    comp = rng.integers(0, 6, size=n).astype(np.int32)
    cell = rng.integers(0, ncell, size=n).astype(np.int64)
    deps = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex128)
    return pidx, comp, cell, deps


def _call(sim, name, pidx, comp, cell, deps, nparam, step):
    from emopt_amd._lib import ptr
    grad = np.full(nparam, np.nan)
    getattr(sim._lib, name)(sim._h, len(pidx), ptr(pidx), ptr(comp), ptr(cell), ptr(deps),
                            int(nparam), float(step), ptr(grad))
    return grad


def test_noalloc_assembly_matches_allocating_entry():
    from emopt_amd._lib import EmfError
    case = strip_case(32, 26, 20, w=5)
    sim = build_sim(case)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        sim.solve_forward()
        src, (k1, k2, j1, j2, i1, i2) = case["sources"][0]
        from emopt_amd.misc import DomainCoordinates
        sim.add_adjoint_sources(src, DomainCoordinates.from_indices(k1 + 5, k2 + 5, j1, j2, i1,
                                                                    i2, sim.dx, sim.dy, sim.dz))
        sim.solve_adjoint()
    ncell = sim.Nx * sim.Ny * sim.Nz
    rng = np.random.default_rng(7)
    for n, nparam in ((200000, 96), (3000, 17)):      # second call smaller
        e = _entries(rng, n, nparam, ncell)
        ref = _call(sim, "emf_grad_assemble", *e, nparam, 1e-3)
        got = _call(sim, "emf_grad_assemble_noalloc", *e, nparam, 1e-3)
        print("n %d, %d parameters: max |grad| %.3e, max |diff| %.3e"
              % (n, nparam, np.abs(ref).max(), np.abs(got - ref).max()))
        assert np.all(np.isfinite(ref)) and np.abs(ref).max() > 0
        assert np.array_equal(got, ref)
    # a parameter without entries gets 0
    e = _entries(rng, 50, 3, ncell)
    e = tuple(a[e[0] != 1] for a in e)
    got = _call(sim, "emf_grad_assemble_noalloc", *e, 3, 1.0)
    assert got[1] == 0.0 and np.array_equal(got, _call(sim, "emf_grad_assemble", *e, 3, 1.0))
    # beyond the scratch array: rejected on the host, nothing launched
    big = _entries(rng, 1_300_000, 4, ncell)
    with pytest.raises(EmfError, match="scratch"):
        _call(sim, "emf_grad_assemble_noalloc", *big, 4, 1.0)
    e = _entries(rng, 1000, 5, ncell)
    assert np.array_equal(_call(sim, "emf_grad_assemble_noalloc", *e, 5, 1.0),
                          _call(sim, "emf_grad_assemble", *e, 5, 1.0))
    # and through the solver's Python method
    assert np.array_equal(sim.grad_assemble(*e, 5, noalloc=True), sim.grad_assemble(*e, 5))
