"""Device-side builders of the large workloads (same formulas as
tsmi.workloads, evaluated in FP64 with torch on any device), so that an
8192 x 8192 tile does not have to be built on the host."""
from __future__ import annotations

import math

import numpy as np
import torch

from . import workloads as W
from .lattice import Lattice2D


def mesh(lat: Lattice2D, device) -> tuple[torch.Tensor, torch.Tensor]:
    x = torch.as_tensor(lat.coords_x(), dtype=torch.float64, device=device)
    y = torch.as_tensor(lat.coords_y(), dtype=torch.float64, device=device)
    return x[None, :].expand(lat.H, lat.W), y[:, None].expand(lat.H, lat.W)


def normalise(lat: Lattice2D, psi: torch.Tensor, norm: float = 1.0) -> torch.Tensor:
    inner = psi[lat.inner_y, lat.inner_x]
    n2 = float(torch.sum(inner.real ** 2 + inner.imag ** 2)) * lat.delta_x * lat.delta_y
    return psi * math.sqrt(norm / n2)


def config2(n: int = 8192, length: float = 64.0, n_vortices: int = 16, g: float = 50.0, trap: float = 0.01,
            device="cuda"):
    """Real-time GPE vortex state of tsmi.workloads.config2; returns
    (lattice, psi complex128 (H, W), v float64 (H, W), params dict)."""
    seed = W.SEED_BASE + 2
    lat = Lattice2D(n, length, periodic_x=True, periodic_y=True)
    X, Y = mesh(lat, device)
    v = trap * (X * X + Y * Y)
    mu = math.sqrt(2.0 * trap * g / math.pi)
    r_tf = math.sqrt(mu / trap)
    amp = torch.sqrt(torch.clamp(mu - v, min=0.0) / g)
    u = W.uniform(seed, 3 * n_vortices)
    psi = amp.to(torch.complex128)
    for j in range(n_vortices):
        rad = 0.7 * r_tf * np.sqrt(u[3 * j])
        ang = 2.0 * np.pi * u[3 * j + 1]
        xj, yj = float(rad * np.cos(ang)), float(rad * np.sin(ang))
        s = 1.0 if u[3 * j + 2] < 0.5 else -1.0
        dx, dy = X - xj, Y - yj
        psi = psi * torch.tanh(torch.sqrt(dx * dx + dy * dy)) * torch.exp(1j * s * torch.atan2(dy, dx))
    psi = normalise(lat, psi)
    return lat, psi, v.contiguous(), dict(dt=5e-4, g=g, mass=1.0, mu=mu, r_tf=r_tf)


def exp_potential(v: torch.Tensor, dt: float, imag_time: bool):
    """E = exp(-i dt V) (or exp(-dt V)) as two contiguous FP64 planes."""
    if imag_time:
        return torch.exp(-dt * v).contiguous(), torch.zeros_like(v)
    e = torch.exp(-1j * dt * v.to(torch.complex128))
    return e.real.contiguous(), e.imag.contiguous()
