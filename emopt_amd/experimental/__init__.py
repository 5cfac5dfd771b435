"""The 3D FDTD parts of ``emopt.experimental``: differentiable (torch)
material functions, topology-optimisation materials, and the adjoint
methods that differentiate through them (emopt/experimental/fdtd.py,
grid.py, adjoint_method.py, autodiff_geometry.py).  The 2D/FDFD classes of
that package are not part of this solver."""

_SUBMODULES = ("fdtd", "grid", "adjoint_method", "autodiff_geometry")


def __getattr__(name):
    # emopt.experimental.fdtd etc. as attributes, imported on first use
    import importlib
    if name in _SUBMODULES:
        return importlib.import_module("." + name, __name__)
    raise AttributeError(name)
