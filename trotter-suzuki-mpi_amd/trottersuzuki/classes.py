"""Python-level parts of the trottersuzuki classes
(src/Python/trottersuzuki/classes_extension.py:1-403): lattice axes, states
and potentials from Python functions, time-dependent potentials, and the
Solver whose evolve loop applies them step by step."""
from __future__ import annotations

import numpy as np

from . import _trottersuzuki as _ts
from .tools import _arity_one, evaluate_on_tile, imprint, tile_coordinates


def _x_axis(grid):
    return (np.arange(grid.global_no_halo_dim_x) - grid.global_no_halo_dim_x * 0.5 + 0.5) * grid.delta_x


class Lattice1D(_ts.Lattice1D):
    """A chain of ``dim`` points over ``length`` (closed or periodic)."""

    def get_x_axis(self):
        """The x coordinates of the lattice points (no halo)."""
        return _x_axis(self)


class Lattice2D(_ts.Lattice2D):
    """A dim_x x dim_y lattice over length_x x length_y (square when the y
    values are omitted)."""

    def __init__(self, dim_x, length_x, dim_y=None, length_y=None, periodic_x_axis=False, periodic_y_axis=False,
                 angular_velocity=0., coordinate_system="cartesian"):
        super().__init__(dim_x, length_x, dim_x if dim_y is None else dim_y,
                         length_x if length_y is None else length_y, periodic_x_axis, periodic_y_axis,
                         angular_velocity, coordinate_system)

    def get_x_axis(self):
        """The x coordinates of the lattice points (no halo)."""
        return _x_axis(self)

    def get_y_axis(self):
        """The y coordinates of the lattice points (no halo)."""
        return (np.arange(self.global_no_halo_dim_y) - self.global_no_halo_dim_y * 0.5 + 0.5) * self.delta_y


class _Imprint:
    def imprint(self, function):
        """psi(x, y) <- function(x, y) psi(x, y) (function of x alone on a chain)."""
        imprint(self, function)


class State(_Imprint, _ts.State):
    """A wave function on the whole tile of ``grid``."""

    def init_state(self, state_function):
        """Sets psi(x, y) = state_function(x, y) (or state_function(x))."""
        m = evaluate_on_tile(self.grid, state_function)
        self.init_state_matrix(np.ascontiguousarray(m.real), np.ascontiguousarray(m.imag))


class GaussianState(_Imprint, _ts.GaussianState):
    """sqrt(norm sqrt(wx wy) / pi) exp(-(wx (x-mx)^2 + wy (y-my)^2) / 2) e^{i phase}."""


class SinusoidState(_Imprint, _ts.SinusoidState):
    """2 sqrt(norm / (Lx Ly)) sin(2 pi nx x / Lx) sin(2 pi ny y / Ly) e^{i phase}."""


class ExponentialState(_Imprint, _ts.ExponentialState):
    """sqrt(norm / (Lx Ly)) exp(i 2 pi (nx x / Lx + ny y / Ly)) e^{i phase}."""


class BesselState:
    """Bessel-function states belong to cylindrical coordinates, which are
    outside this path."""

    def __init__(self, *args, **kwargs):
        raise NotImplementedError("BesselState needs cylindrical coordinates, which the HIP path does not run")


class Potential(_ts.Potential):
    """An external potential on the tile: static (a function of (x, y) or of
    x) or time-dependent (a function of (x, y, t), applied step by step by
    Solver.evolve)."""

    def init_potential(self, pot_function):
        """Tabulates V from ``pot_function``; a function of (x, y, t) marks the
        potential as time-dependent and is tabulated at t = 0."""
        self.pot_function = None
        if _arity_one(pot_function):
            static = pot_function
        else:
            try:
                pot_function(0, 0, 0)
                timed = True
            except TypeError:
                timed = False
            if timed:
                self.updated_potential_matrix = True
                self.pot_function = pot_function
                self.exp_potential_matrix = np.zeros((self.grid.dim_y, self.grid.dim_x), dtype=np.complex128)

                def static(x, y):
                    return pot_function(x, y, 0)
            else:
                static = pot_function
        self.potential_matrix = evaluate_on_tile(self.grid, static, dtype=np.float64)
        self.init_potential_matrix(self.potential_matrix)

    def exponential_update(self, delta_t, t):
        """exp(-i delta_t V(x, y, t)) over the tile (time-dependent potentials)."""
        v = evaluate_on_tile(self.grid, self.pot_function, dtype=np.float64, extra=(t,))
        self.exp_potential_matrix = np.exp(-1j * delta_t * v)
        return self.exp_potential_matrix


class Solver(_ts.Solver):
    """Evolves one state (or two, with a Hamiltonian2Component) with the HIP
    kernel: kernel_type "hip" (FP64) or "hip-fp32" (FP32 bulk, FP64
    reductions)."""

    def __init__(self, Lattice, State, Hamiltonian, delta_t, Potential=None, State2=None, Potential2=None,
                 kernel_type="hip"):
        if kernel_type not in ("hip", "hip-fp32"):
            raise ValueError(f"kernel_type {kernel_type!r}: this build runs the HIP kernel "
                             "(\"hip\" for FP64, \"hip-fp32\" for FP32 bulk)")
        if State2 is None:
            super().__init__(Lattice, State, Hamiltonian, delta_t, kernel_type)
        else:
            super().__init__(Lattice, State, State2, Hamiltonian, delta_t, kernel_type)
        self.delta_t = delta_t
        self.potential = Potential
        if State2 is not None:
            self.potential2 = Potential if Potential2 is None else Potential2
        else:
            self.potential2 = None

    def evolve(self, iterations, imag_time=False):
        """``iterations`` steps.  With a time-dependent Python potential (real
        time), each step gets exp(-i dt V(t)) at the current time and runs as
        its own soft update; the last one copies the state back."""
        pot = self.potential
        if imag_time or pot is None or getattr(pot, "pot_function", None) is None or \
                not self.hamiltonian.potential.updated_potential_matrix:
            super().evolve(iterations, imag_time)
            return
        for i in range(iterations):
            e = pot.exponential_update(self.delta_t, self.current_evolution_time)
            self.set_exp_potential(np.ravel(e.real), np.ravel(e.imag), 0)
            super().evolve(-1 if i < iterations - 1 else 1, imag_time)


__all__ = ["Lattice1D", "Lattice2D", "State", "GaussianState", "SinusoidState", "ExponentialState", "BesselState",
           "Potential", "Solver", "tile_coordinates"]
