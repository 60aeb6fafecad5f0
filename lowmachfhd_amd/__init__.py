"""MI355X-native implicit staggered-grid Stokes solver (LowMachFHD's gmres / stag_mg_solver path).

The compute path is the HIP library ``liblmfhd.so`` (sources in ``csrc/``, C ABI in
``include/lmfhd.h``); ``solver.StokesSolver`` mirrors the reference's Fortran interface over it.
"""
from . import grid, problems, solver  # noqa: F401
from .grid import Grid  # noqa: F401
from .solver import FGMRES, FP64, GMRES_LEFT, MIXED, StokesSolver  # noqa: F401
