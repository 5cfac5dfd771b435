"""The names of ``emopt.adjoint_method`` for the 3D FDTD
(emopt/adjoint_method.py:58-721 AdjointMethod, :723-935 AdjointMethodMO,
:1329-1486 AdjointMethodPNF3D), so that ``emopt.adjoint_method.X`` resolves here.  The
implementations live in :mod:`emopt_amd.adjoint`."""
from .adjoint import AdjointMethod, AdjointMethodMO, AdjointMethodPNF3D

__all__ = ["AdjointMethod", "AdjointMethodMO", "AdjointMethodPNF3D"]
