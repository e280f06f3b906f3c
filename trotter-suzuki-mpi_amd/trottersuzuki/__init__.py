"""trottersuzuki -- the reference's Python API (src/Python/trottersuzuki/__init__.py)
on the MI355X HIP kernels: Lattice1D, Lattice2D, State and the built-in states,
Potential, HarmonicPotential, Hamiltonian, Hamiltonian2Component, Solver
(kernel_type "hip" or "hip-fp32"), map_lattice_to_coordinate_space and
get_vortex_position.

The compiled module links the HIP runtime.  PyTorch ships its own copy under the
same soname and the first one loaded serves the process, so PyTorch, when
installed, is imported first (its build of the runtime is the one it can use).
"""
try:  # noqa: SIM105
    import torch  # noqa: F401
except ImportError:
    pass

from ._trottersuzuki import HarmonicPotential, Hamiltonian, Hamiltonian2Component
from .classes import (BesselState, ExponentialState, GaussianState, Lattice1D, Lattice2D, Potential,
                      SinusoidState, Solver, State)
from .tools import get_vortex_position, map_lattice_to_coordinate_space

__all__ = ["Lattice1D", "Lattice2D", "State", "ExponentialState", "GaussianState", "SinusoidState", "BesselState",
           "Potential", "HarmonicPotential", "Hamiltonian", "Hamiltonian2Component", "Solver",
           "map_lattice_to_coordinate_space", "get_vortex_position"]
