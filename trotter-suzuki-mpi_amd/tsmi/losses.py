"""Losses of a final state and their adjoint seeds.

The adjoint sweep (``HipKernel.potential_gradient``) starts from
G = dL/dRe psi_N + i dL/dIm psi_N, the same convention torch uses for the
gradient of a real loss with respect to a complex tensor.  ``infidelity``
gives the loss and its seed in numpy; ``energy_torch`` is the energy
functional as a torch function, whose seed torch's autograd forms.
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


def energy_torch(lat: Lattice2D, re, im, v, mass: float = 1.0, g: float = 0.0, g_lhy: float = 0.0):
    """The total energy of Solver::calculate_energy_expected_values
    (src/solver.cpp:244-564, one component, no rotation) as a differentiable
    torch function of the tile (re, im) and the raw potential v, all (H, W):
    kinetic -1/(2m) sum Re(conj(psi) D psi) / (sum |psi|^2 over the stencil
    margin) with the 4th-order 5-point D along each axis (2-cell margin next to
    closed edges), potential sum V |psi|^2 / N, interaction g/2 sum |psi|^4 / N,
    LHY 2 g_LHY / 5 sum |psi|^5 / N, N = sum |psi|^2 over the inner region.
    As a loss of the final state of tsmi.autograd.evolve, its gradient in v
    adds the direct potential term to the reverse sweep's."""
    import torch
    iy, ix = lat.inner_y, lat.inner_x
    y0, y1, x0, x1 = iy.start, iy.stop, ix.start, ix.stop
    H = re.shape[0]
    ky0 = y0 + (0 if lat.periodic_y or H == 1 else 2)
    ky1 = y1 - (0 if lat.periodic_y or H == 1 else 2)
    kx0 = x0 + (0 if lat.periodic_x else 2)
    kx1 = x1 - (0 if lat.periodic_x else 2)
    n2 = re * re + im * im
    n_in = n2[y0:y1, x0:x1].sum()
    c1, c2, c3 = -1. / 12., 4. / 3., -2.5

    def lap(a, axis):
        def sh(k):
            if axis == 1:
                return a[ky0:ky1, kx0 + k:kx1 + k]
            return a[ky0 + k:ky1 + k, kx0:kx1]
        return c1 * sh(2) + c2 * sh(1) + c2 * sh(-1) + c1 * sh(-2) + c3 * sh(0)

    pr, pi = re[ky0:ky1, kx0:kx1], im[ky0:ky1, kx0:kx1]
    nk = n2[ky0:ky1, kx0:kx1].sum()
    lx = (pr * lap(re, 1) + pi * lap(im, 1)).sum() / (lat.delta_x * lat.delta_x)
    ly = (pr * lap(re, 0) + pi * lap(im, 0)).sum() / (lat.delta_y * lat.delta_y) if H > 1 else 0.0
    kin = -1. / (2. * mass) * (lx + ly) / nk
    inner = n2[y0:y1, x0:x1]
    pot = (v[y0:y1, x0:x1] * inner).sum() / n_in
    itr = 0.5 * g * (inner * inner).sum() / n_in
    lhy = 0.4 * g_lhy * (inner * inner * torch.sqrt(inner)).sum() / n_in if g_lhy != 0.0 else 0.0
    return kin + pot + itr + lhy
