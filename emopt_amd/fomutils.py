"""The names of ``emopt.fomutils`` that the 3D FDTD figures of merit use
(emopt/fomutils.py:925-1196 ModeMatch, :1282-1429 interpolated_dFdx_3D,
:1654-1920 power_norm_dFdx_3D), so that ``emopt.fomutils.X`` resolves here.
The implementations live in :mod:`emopt_amd.adjoint`."""
from .adjoint import ModeMatch, interpolated_dFdx_3D, power_norm_dFdx_3D

__all__ = ["ModeMatch", "interpolated_dFdx_3D", "power_norm_dFdx_3D"]
