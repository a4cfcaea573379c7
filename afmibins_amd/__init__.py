"""afmibins_amd -- the immersed-boundary hot path of the reference IB /
Navier-Stokes library on MI355X (gfx950): spreading, interpolation, the
multigrid pressure projection and the CG Helmholtz solve, as hand-written HIP
kernels behind the C ABI of include/afm.h.

The compute path is libafm.so (afmibins_amd/lib); this package is its Python
mirror (ctypes + torch device memory).
"""
from ._lib import AfmError, load  # noqa: F401
from .api import Context, MGSolver, NSSolver, NS3Solver, IBSolver, delta_host  # noqa: F401
