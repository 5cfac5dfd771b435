"""The 3D FDTD parts of ``emopt.experimental``: differentiable (torch)
material functions, topology-optimisation materials, and the adjoint
methods that differentiate through them (emopt/experimental/fdtd.py,
grid.py, adjoint_method.py, autodiff_geometry.py).  The 2D/FDFD classes of
that package are not part of this solver."""
