"""Optimizer (emopt/optimizer.py:67-265) on a figure of merit with a known
minimum: the order of the calls the reference makes (the figure of merit at
p0 and the callback on p0 before scipy starts), the minimum reached with
the supplied gradient, bounds honoured, and the iteration cap."""
import numpy as np

from emopt_amd.optimizer import Optimizer


class _Quadratic:
    """F(p) = sum(w (p - c)^2) with its exact gradient, logging the calls."""

    def __init__(self, c, w):
        self.c, self.w = np.asarray(c, float), np.asarray(w, float)
        self.calls = []

    def fom(self, p):
        self.calls.append(("fom", np.array(p)))
        return float(np.sum(self.w * (p - self.c) ** 2))

    def gradient(self, p):
        self.calls.append(("grad", np.array(p)))
        return 2 * self.w * (p - self.c)


def test_minimises_with_the_adjoint_gradient():
    am = _Quadratic([1.0, -2.0, 0.5], [1.0, 3.0, 10.0])
    seen = []
    opt = Optimizer(am, np.zeros(3), callback_func=lambda p: seen.append(np.array(p)),
                    opt_method="L-BFGS-B", Nmax=100, tol=1e-12, scipy_verbose=False)
    fom, p = opt.run()
    assert am.calls[0][0] == "fom" and np.array_equal(am.calls[0][1], np.zeros(3))
    assert np.array_equal(seen[0], np.zeros(3)) and len(seen) >= 2
    assert any(kind == "grad" for kind, _ in am.calls)
    np.testing.assert_allclose(p, [1.0, -2.0, 0.5], atol=1e-5)
    assert fom < 1e-9


def test_bounds_and_iteration_cap():
    am = _Quadratic([1.0, -2.0], [1.0, 1.0])
    fom, p = Optimizer(am, np.zeros(2), opt_method="L-BFGS-B", bounds=[(0, 0.5), (-1, 1)],
                       scipy_verbose=False).run()
    np.testing.assert_allclose(p, [0.5, -1.0], atol=1e-8)
    am = _Quadratic([100.0, -200.0], [1.0, 50.0])
    Optimizer(am, np.zeros(2), opt_method="BFGS", Nmax=1, scipy_verbose=False).run()
    assert sum(1 for kind, _ in am.calls if kind == "grad") <= 6
