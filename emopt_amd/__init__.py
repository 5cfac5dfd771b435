"""MI355X-native CW-FDTD behind the surface of ``emopt.fdtd.FDTD``.

The compute path is ``libemfdtd.so`` (hand-written HIP for gfx950, built by
``python -m emopt_amd.build``), reached through ctypes.
"""
from .defs import FieldComponent, SourceComponent
from .misc import DomainCoordinates
from .grid import (ConstantMaterial3D, StructuredMaterial3D, FunctionMaterial3D, Rectangle,
                   Polygon, Ellipse)

__version__ = "0.1.0"


# submodules reachable as attributes (``emopt.fdtd.FDTD``, ``emopt.modes``,
# ...) as in emopt/__init__.py:2; imported on first use, since the solver
# modules load the native library
_SUBMODULES = ("fdtd", "modes", "adjoint", "adjoint_method", "fomutils", "batch", "serve",
               "shape", "fields", "simulation", "autotune", "optimizer", "experimental")


def __getattr__(name):
    import importlib
    if name in ("FDTD", "SourceArray"):
        from . import fdtd
        return getattr(fdtd, name)
    if name in _SUBMODULES:
        return importlib.import_module("." + name, __name__)
    raise AttributeError(name)
