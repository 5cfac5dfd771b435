"""Adjoint methods of ``emopt.experimental.adjoint_method`` for the 3D FDTD
(emopt/experimental/adjoint_method.py:22-238 AutoDiff, :240-568 Topology,
:570-591 the sigmoid helpers).

Both are generated from a base adjoint method (AdjointMethod or
AdjointMethodPNF3D here), keeping its figure of merit and adjoint sources
and replacing how the structure is updated and how the gradient is formed:

* **AutoDiff** -- the structure is a torch function of the parameters
  (AutoDiffMaterial3D / HybridMaterial3D).  update_system evaluates it at
  the six Yee offsets over the design domain with the parameters recording
  a graph; the gradient is torch.autograd.grad of the solver's pseudo-loss
  2 Im sum(eps o E o E_adj [- mu o H o H_adj]) (FDTD.calc_ydAx_autograd,
  whose field products come from the GPU).  One forward and one adjoint
  solve give the whole gradient, however many parameters.
* **Topology** -- every cell of the design domain is a parameter
  (TopologyMaterial3D), optionally through a sigmoid onto [lb, ub] and
  optionally extruded along z (planar).  The gradient is
  FDTD.calc_ydAx_topology.
"""
from __future__ import annotations

import numpy as np

from .. import adjoint as _am
from ..misc import DomainCoordinates, NOT_PARALLEL
from . import fdtd as _fdtd
from .grid import AutoDiffMaterial3D, HybridMaterial3D, TopologyMaterial3D

__all__ = ["AutoDiff", "AutoDiffPNF3D", "Topology", "TopologyPNF3D", "sigmoid",
           "scaled_sigmoid", "inverse_scaled_sigmoid"]

# Yee offsets [sz, sy, sx] of Ex, Ey, Ez and of Hx, Hy, Hz, in cells (the
# offsets at which the FDTD samples eps_x..mu_z)
_SHIFTS_E = ((-0.5, 0.0, 0.5), (-0.5, 0.5, 0.0), (0.0, 0.0, 0.0))
_SHIFTS_M = ((0.0, 0.5, 0.0), (0.0, 0.0, 0.5), (-0.5, 0.5, 0.5))


def _check_solver(sim):
    from ..fdtd import FDTD
    if not isinstance(sim, (_fdtd.FDTD, FDTD)):
        raise TypeError("the experimental adjoint methods drive the FDTD solver")


def _AutoDiffGenerator(BaseAM):
    """An adjoint method class derived from BaseAM whose structure is a
    torch function of the parameters (emopt/experimental/adjoint_method.py:
    22-231).  Users define the figure of merit hooks as for BaseAM, but not
    update_system."""
    if not issubclass(BaseAM, _am.AdjointMethod):
        raise TypeError("BaseAM must derive from AdjointMethod")

    class _AutoDiff(BaseAM):
        def __init__(self, sim, domain=None, update_mu=False):
            """domain: where the material function acts (default: the
            HybridMaterial3D's domain, else the whole grid); update_mu:
            differentiate mu too (sim.mu must then be a torch material)."""
            super().__init__(sim)
            _check_solver(sim)
            if not isinstance(sim.eps, (AutoDiffMaterial3D, HybridMaterial3D)):
                raise TypeError("sim.eps must be an AutoDiffMaterial3D or HybridMaterial3D")
            if update_mu and not isinstance(sim.mu, (AutoDiffMaterial3D, HybridMaterial3D)):
                raise TypeError("update_mu needs sim.mu to be an AutoDiffMaterial3D or "
                                "HybridMaterial3D")
            if domain is None:
                domain = sim.eps._fd if isinstance(sim.eps, HybridMaterial3D) else \
                    DomainCoordinates(0, sim.X, 0, sim.Y, 0, sim.Z, sim.dx, sim.dy, sim.dz)
            if not isinstance(domain, DomainCoordinates):
                raise TypeError("domain must be a DomainCoordinates")
            self._domain = domain
            self._update_mu = update_mu
            self._ds = (sim.dz, sim.dy, sim.dx)

        def _coords(self, shift):
            torch = _am._torch()
            d = self._domain
            base = (d.z, d.y, d.x)
            return [torch.as_tensor(np.asarray(base[a], dtype=np.double) + shift[a] * self._ds[a])
                    for a in range(3)]

        def update_system(self, params):
            """Evaluate the material function at the Yee offsets over the
            design domain with params recording a graph (kept for
            calc_gradient), and hand the parameters to the material for the
            solver's next build (emopt/experimental/adjoint_method.py:
            149-199)."""
            torch = _am._torch()
            if NOT_PARALLEL:
                self._v = torch.as_tensor(np.asarray(params, dtype=np.double)).requires_grad_()
                self._eps_autograd = [self.sim.eps._func(self._v, self._coords(s)) + 0j
                                      for s in _SHIFTS_E]
                self._mu_autograd = [self.sim.mu._func(self._v, self._coords(s)) + 0j
                                     for s in _SHIFTS_M] if self._update_mu else None
            v = torch.as_tensor(np.asarray(params, dtype=np.double))
            self.sim.eps.v = v
            if self._update_mu:
                self.sim.mu.v = v

        def calc_gradient(self, sim, params):
            """d(pseudo-loss)/d(params) by reverse-mode differentiation
            (emopt/experimental/adjoint_method.py:201-229)."""
            torch = _am._torch()
            loss = sim.calc_ydAx_autograd(self._domain, self._update_mu, self._eps_autograd,
                                          self._mu_autograd)
            return torch.autograd.grad(loss, self._v)[0].detach().cpu().numpy()

    _AutoDiff.__name__ = "AutoDiff" + BaseAM.__name__.replace("AdjointMethod", "")
    return _AutoDiff


def _TopologyGenerator(BaseAM):
    """An adjoint method class derived from BaseAM for topology
    optimisation over a TopologyMaterial3D (emopt/experimental/
    adjoint_method.py:240-556).  Call get_params() after construction for
    the starting parameter vector."""
    if not issubclass(BaseAM, _am.AdjointMethod):
        raise TypeError("BaseAM must derive from AdjointMethod")

    class _Topology(BaseAM):
        def __init__(self, sim, domain=None, update_mu=False, eps_bounds=None, mu_bounds=None,
                     planar=False, vol_penalty=0.):
            """eps_bounds / mu_bounds: (lb, ub) reached through a sigmoid
            (None: the parameters are the values); planar: one parameter
            per (y, x) column of the domain, extruded along z; vol_penalty:
            lam * mean(sigmoid(params)) added to the figure of merit."""
            super().__init__(sim)
            _check_solver(sim)
            if not isinstance(sim.eps, TopologyMaterial3D):
                raise TypeError("sim.eps must be a TopologyMaterial3D")
            if update_mu and not isinstance(sim.mu, TopologyMaterial3D):
                raise TypeError("update_mu needs sim.mu to be a TopologyMaterial3D")
            self._domain = domain if domain is not None else sim.eps._fd
            if not isinstance(self._domain, DomainCoordinates):
                raise TypeError("domain must be a DomainCoordinates")
            self._update_mu = update_mu
            self._epsb = eps_bounds
            self._mub = mu_bounds
            self._planar = planar
            self._lam = vol_penalty

        def _to_params(self, grid, bounds, squish):
            p = inverse_scaled_sigmoid(grid.real, bounds[0], bounds[1], squish=squish) \
                if bounds else np.copy(grid.real)
            return (np.mean(p, axis=0) if self._planar else p).ravel()

        def get_params(self, squish=0.05):
            """The parameter vector of the current grid(s): inverse sigmoid
            of the values (pulled squish x (ub - lb) inside the bounds),
            z-averaged when planar (emopt/experimental/adjoint_method.py:
            362-410)."""
            self._grid_eps = np.copy(self.sim.eps.grid)
            self._gse = self._grid_eps.shape
            params = self._to_params(self._grid_eps, self._epsb, squish)
            self._pse = params.size
            if self._update_mu:
                self._grid_mu = np.copy(self.sim.mu.grid)
                self._gsm = self._grid_mu.shape
                pm = self._to_params(self._grid_mu, self._mub, squish)
                self._psm = pm.size
                params = np.concatenate([params, pm], axis=0)
            return params

        def _set_grid(self, mat, p, bounds, shape):
            g = scaled_sigmoid(p, bounds[0], bounds[1]) if bounds else p
            if self._planar:
                mat.grid = np.broadcast_to(g.reshape(shape[1:])[np.newaxis, ...], shape)
            else:
                mat.grid = g.reshape(shape)

        def update_system(self, params):
            """Write the parameters into the material grid(s)
            (emopt/experimental/adjoint_method.py:412-455)."""
            params = np.asarray(params, dtype=np.double)
            if self._update_mu:
                self._set_grid(self.sim.eps, params[:self._pse], self._epsb, self._gse)
                self._set_grid(self.sim.mu, params[self._pse:], self._mub, self._gsm)
            else:
                self._set_grid(self.sim.eps, params, self._epsb, self._gse)

        def _chain(self, p, bounds, shape):
            """(d value / d param, value range) of the sigmoid map."""
            if not bounds:
                return np.array([1.0]), 1.0
            s = sigmoid(p.reshape(shape[1:] if self._planar else shape))
            return s * (1.0 - s), bounds[1] - bounds[0]

        def calc_gradient(self, sim, params):
            """FDTD.calc_ydAx_topology with the sigmoid's chain factor
            (emopt/experimental/adjoint_method.py:457-527)."""
            params = np.asarray(params, dtype=np.double)
            pe = params[:self._pse] if self._update_mu else params
            sig_eps, del_eps = self._chain(pe, self._epsb, self._gse)
            if self._update_mu:
                sig_mu, del_mu = self._chain(params[self._pse:], self._mub, self._gsm)
            else:
                sig_mu, del_mu = 1, 1
            return sim.calc_ydAx_topology(self._domain, self._update_mu, sig_eps=sig_eps,
                                          sig_mu=sig_mu, del_eps=del_eps, del_mu=del_mu,
                                          planar=self._planar, lam=self._lam)

        def calc_penalty(self, sim, params):
            """lam * mean(sigmoid(params)), kept as current_vol_penalty
            (emopt/experimental/adjoint_method.py:529-550)."""
            self.current_vol_penalty = 0 if self._lam == 0 else \
                self._lam * np.mean(sigmoid(np.asarray(params, dtype=np.double)))
            return self.current_vol_penalty

        def calc_grad_p(self, sim, params):
            # the penalty's gradient is part of calc_gradient
            return np.zeros_like(params)

    _Topology.__name__ = "Topology" + BaseAM.__name__.replace("AdjointMethod", "")
    return _Topology


AutoDiff = _AutoDiffGenerator(_am.AdjointMethod)
AutoDiffPNF3D = _AutoDiffGenerator(_am.AdjointMethodPNF3D)
Topology = _TopologyGenerator(_am.AdjointMethod)
TopologyPNF3D = _TopologyGenerator(_am.AdjointMethodPNF3D)


def sigmoid(x):
    return 1.0 / (1.0 + np.exp(-x))


def scaled_sigmoid(x, lb, ub):
    return lb + (ub - lb) * sigmoid(x)


def inverse_scaled_sigmoid(x, lb, ub, squish=0.05):
    """The parameter whose scaled_sigmoid over the bounds widened by
    squish x (ub - lb) on each side is x (emopt/experimental/
    adjoint_method.py:580-591)."""
    span = ub - lb
    lo, hi = lb - squish * span, ub + squish * span
    a = (np.asarray(x) - lo) / (hi - lo)
    if not (np.all(a >= 0.0) and np.all(a <= 1.0)):
        raise ValueError("values outside the widened bounds")
    return np.log(a / (1.0 - a))
