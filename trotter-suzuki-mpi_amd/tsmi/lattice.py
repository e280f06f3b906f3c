"""Tile geometry of the single-process 2D Cartesian lattice.

Mirrors ``Lattice2D::init`` (src/model.cpp:103-158) and ``calculate_borders``
(src/common.cpp:83-102) for one process: every axis keeps a lattice halo width of
4 (no rotation), but only a periodic axis stores it, so the tile is
``W = Nx + 8*periodic_x`` by ``H = Ny + 8*periodic_y`` and the inner cells are
``[h, h+N)`` on a periodic axis and ``[0, N)`` on a closed one.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

HALO = 4  # src/model.cpp:140-141 with angular velocity 0


@dataclass(frozen=True)
class Lattice2D:
    """Single-process lattice (src/trottersuzuki.h:83-117)."""

    dim_x: int                  # global inner points along x
    length_x: float
    dim_y: int = -1
    length_y: float = -1.0
    periodic_x: bool = False
    periodic_y: bool = False

    def __post_init__(self):
        if self.dim_y < 0:
            object.__setattr__(self, "dim_y", self.dim_x)
        if self.length_y < 0:
            object.__setattr__(self, "length_y", self.length_x)
        if self.dim_x < 1 or self.dim_y < 1:
            raise ValueError("lattice dimensions must be positive")

    # -- geometry -----------------------------------------------------------
    @property
    def halo_x(self) -> int:
        return HALO

    @property
    def halo_y(self) -> int:
        return HALO

    @property
    def delta_x(self) -> float:
        return self.length_x / float(self.dim_x)

    @property
    def delta_y(self) -> float:
        return self.length_y / float(self.dim_y)

    @property
    def W(self) -> int:
        """Tile width incl. halo (``grid->dim_x``)."""
        return self.dim_x + (2 * HALO if self.periodic_x else 0)

    @property
    def H(self) -> int:
        """Tile height incl. halo (``grid->dim_y``)."""
        return self.dim_y + (2 * HALO if self.periodic_y else 0)

    @property
    def start_x(self) -> int:
        return -HALO if self.periodic_x else 0

    @property
    def start_y(self) -> int:
        return -HALO if self.periodic_y else 0

    @property
    def inner_x(self) -> slice:
        o = HALO if self.periodic_x else 0
        return slice(o, o + self.dim_x)

    @property
    def inner_y(self) -> slice:
        o = HALO if self.periodic_y else 0
        return slice(o, o + self.dim_y)

    # -- coordinates ---------------------------------------------------------
    def coords_x(self) -> np.ndarray:
        """Physical x of every tile column (src/common.cpp:45-65, Cartesian)."""
        return _map_axis(self.start_x, self.delta_x, self.dim_x, self.length_x, self.W)

    def coords_y(self) -> np.ndarray:
        return _map_axis(self.start_y, self.delta_y, self.dim_y, self.length_y, self.H)

    def mesh(self):
        """(X, Y) arrays of shape (H, W)."""
        x = self.coords_x()
        y = self.coords_y()
        return np.broadcast_to(x[None, :], (self.H, self.W)), np.broadcast_to(y[:, None], (self.H, self.W))


def _map_axis(start: int, delta: float, n: int, length: float, count: int) -> np.ndarray:
    # Same operation order as map_lattice_to_coordinate_space, element by element.
    out = np.empty(count, dtype=np.float64)
    c = n * delta * 0.5
    for i in range(count):
        idx = start * delta + 0.5 * delta + i * delta
        if idx - c < -length * 0.5:
            idx += length
        if idx - c > length * 0.5:
            idx -= length
        out[i] = idx - c
    return out
