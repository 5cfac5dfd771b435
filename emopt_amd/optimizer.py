"""``emopt.optimizer.Optimizer`` (emopt/optimizer.py:67-265): minimise an
adjoint method's figure of merit over its design parameters with
scipy.optimize.minimize, the gradient coming from the adjoint method.

The reference runs scipy on MPI rank 0 and broadcasts fom / gradient
commands to the other ranks; one process drives the GPU here, so the
calls are made directly."""
from __future__ import annotations

import numpy as np
from scipy.optimize import minimize

__all__ = ["Optimizer"]


class Optimizer:
    def __init__(self, am, p0, callback_func=None, opt_method='BFGS', Nmax=1000, tol=1e-5,
                 bounds=None, scipy_verbose=True):
        """am: the adjoint method; p0: starting parameters; callback_func:
        called with the parameters before the first iteration and after each
        one; opt_method, tol, bounds: scipy.optimize.minimize's method, tol
        and bounds; Nmax: its maxiter."""
        self.am = am
        self.p0 = p0
        self.callback = callback_func if callback_func is not None else (lambda p: None)
        self.opt_method = opt_method
        self.Nmax = Nmax
        self.tol = tol
        self.bounds = bounds
        self.scipy_verbose = scipy_verbose

    def run(self):
        """Run the optimisation; returns (final figure of merit, final
        parameters)."""
        return self.run_sequence(self.am)

    def run_sequence(self, am):
        """The optimisation itself (emopt/optimizer.py:237-264): one figure
        of merit at p0 and the callback on it, then minimize with the adjoint
        gradient.  Override for custom sequences."""
        p0 = np.asarray(self.p0, dtype=np.double)
        am.fom(p0)
        self.callback(p0)
        result = minimize(am.fom, p0, method=self.opt_method, jac=am.gradient,
                          callback=self.callback, tol=self.tol, bounds=self.bounds,
                          options={'maxiter': self.Nmax, 'disp': self.scipy_verbose})
        return result.fun, result.x
