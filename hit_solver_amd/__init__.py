"""hit_solver_amd: MI355X-native (gfx950) hot path of HIT_Solver.

The right-hand side (shock sensor, hybrid compact/WENO7 fluxes, compact line
solves, viscous/forcing/drag/cooling sources), the SSP-RK3 stage update and
the compact-node hyperviscosity filter run as hand-written HIP kernels behind
the C ABI in include/hitx.h, exposed here with the reference's Python API as
the ``hip`` backend.
"""

from .backend import HIP, array_module, select_array_module, to_numpy
from .domain import Domain2D

__all__ = ["HIP", "Domain2D", "array_module", "select_array_module", "to_numpy"]
