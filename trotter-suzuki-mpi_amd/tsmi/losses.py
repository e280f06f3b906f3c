"""Losses of a final state and their adjoint seeds.

The adjoint sweep (``HipKernel.potential_gradient``) starts from
G = dL/dRe psi_N + i dL/dIm psi_N, the same convention torch uses for the
gradient of a real loss with respect to a complex tensor.
"""
from __future__ import annotations

import numpy as np

from .lattice import Lattice2D


def overlap(lat: Lattice2D, target: np.ndarray, psi: np.ndarray) -> complex:
    """<target|psi> = sum conj(target) psi dx dy over the inner region."""
    t = target[lat.inner_y, lat.inner_x]
    p = psi[lat.inner_y, lat.inner_x]
    return complex(np.sum(np.conj(t) * p) * lat.delta_x * lat.delta_y)


def infidelity(lat: Lattice2D, target: np.ndarray, psi: np.ndarray):
    """L = 1 - |<target|psi>|^2 and its seed G = -2 <target|psi> target dx dy
    (zero outside the inner region)."""
    c = overlap(lat, target, psi)
    g = np.zeros_like(psi, dtype=np.complex128)
    g[lat.inner_y, lat.inner_x] = -2.0 * c * target[lat.inner_y, lat.inner_x] * lat.delta_x * lat.delta_y
    return 1.0 - abs(c) ** 2, g
