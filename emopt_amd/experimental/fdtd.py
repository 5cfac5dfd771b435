"""``emopt.experimental.fdtd.FDTD`` (emopt/experimental/fdtd.py:25-223).

The experimental solver is the FDTD plus two gradient reductions,
calc_ydAx_topology (:37-158) and calc_ydAx_autograd (:160-223).
emopt_amd.fdtd.FDTD carries both, with the forward x adjoint field products
formed on the GPU, so this class adds nothing; it exists so that scripts
written against emopt.experimental find the name, and so that the
experimental adjoint methods can recognise the solver they drive."""
from ..fdtd import FDTD as _FDTD

__all__ = ["FDTD"]


class FDTD(_FDTD):
    """The FDTD solver with the experimental gradient reductions."""
