"""The names of ``emopt.fomutils`` that 3D FDTD figures of merit use
(emopt/fomutils.py:925-1196 ModeMatch, :1282-1429 interpolated_dFdx_3D,
:1654-1920 power_norm_dFdx_3D, whose implementations live in
:mod:`emopt_amd.adjoint`), and its smooth step and rect functions
(emopt/fomutils.py:38-140) for shaping figures of merit and penalties."""
import numpy as np

from .adjoint import ModeMatch, interpolated_dFdx_3D, power_norm_dFdx_3D

__all__ = ["ModeMatch", "interpolated_dFdx_3D", "power_norm_dFdx_3D", "step", "step_derivative",
           "rect", "rect_derivative"]


def step(x, k, y0=0, A=1.0):
    """Logistic step A / (1 + exp(-k x)) + y0: y0 for x -> -inf, y0 + A
    for x -> inf, steepness k."""
    return A / (1 + np.exp(-k * x)) + y0


def step_derivative(x, k, y0=0, A=1.0):
    """d step / dx."""
    e = np.exp(-k * x)
    return A * k * e / (1 + e) ** 2


def _rect_edges(w1p, ws):
    # slope k from the side width ws, where a logistic edge rises 1 % -> 99 %;
    # edges placed so that rect(+-w1p / 2) = 0.01 for w1p >> ws
    k = 2 * np.log(99.0) / ws
    return k, -w1p / 2.0 + ws / 2.0, w1p / 2.0 - ws / 2.0


def rect(x, w1p, ws):
    """Smooth rect centred at 0: the difference of two logistic steps, ~1
    inside, 0.01 at x = +-w1p / 2 (the "one percent" width), sides ws wide."""
    k, x1, x2 = _rect_edges(w1p, ws)
    return 1 / (1 + np.exp(-k * (x - x1))) - 1 / (1 + np.exp(-k * (x - x2)))


def rect_derivative(x, w1p, ws):
    """d rect / dx."""
    k, x1, x2 = _rect_edges(w1p, ws)
    e1, e2 = np.exp(-k * (x - x1)), np.exp(-k * (x - x2))
    return k * e1 / (1 + e1) ** 2 - k * e2 / (1 + e2) ** 2
