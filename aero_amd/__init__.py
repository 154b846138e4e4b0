"""MI355X-native lattice-Boltzmann core: the step path of the Aero wind tunnel.

``Solver`` (D2Q9) and ``Solver3D`` (D3Q19) keep the reference's solver API;
every lattice update runs in hand-written HIP kernels for gfx950 reached
through the C ABI in include/aero_hip.h (libaero_hip.so, built in-tree by
``python -m aero_amd.build``).
"""

from .solver import Solver
from .solver3d import Solver3D

__all__ = ["Solver", "Solver3D"]
