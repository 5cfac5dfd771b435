"""MI355X-native CW-FDTD behind the surface of ``emopt.fdtd.FDTD``.

The compute path is ``libemfdtd.so`` (hand-written HIP for gfx950, built by
``python -m emopt_amd.build``), reached through ctypes.
"""
from .defs import FieldComponent, SourceComponent
from .misc import DomainCoordinates
from .grid import (ConstantMaterial3D, StructuredMaterial3D, FunctionMaterial3D, Rectangle,
                   Polygon, Ellipse)

__version__ = "0.1.0"


def __getattr__(name):
    # the solver pulls in the native library; import it on first use
    if name in ("FDTD", "SourceArray"):
        from . import fdtd
        return getattr(fdtd, name)
    raise AttributeError(name)
