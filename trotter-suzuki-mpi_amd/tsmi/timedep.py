"""Time-dependent potentials with the E table rebuilt on the device.

The reference re-tabulates E = exp(-i dt V(t)) on the host before every step
but the first of an evolve call when the potential changes
(src/solver.cpp:186-192: Potential::update, initialize_exp_potential,
ITrotterKernel::update_potential).  Here V(t) comes from a callable that
returns a device tensor (H, W) FP64 and the table is built by the library
(tsmi_set_potential_exp_device), so nothing crosses PCIe per step.
"""
from __future__ import annotations

from typing import Callable

from .hip import HipKernel


def evolve(kernel: HipKernel, potential_at: Callable[[float], "object"], steps: int, t_start: float = 0.0) -> float:
    """Solver::evolve(steps) of the reference with a time-dependent potential:
    step i of the call uses V(t_start + i dt) for i > 0 and the table already
    set for i = 0; the halo exchange follows every step but the last.  Returns
    the evolution time after the call."""
    import torch
    dt = kernel.delta_t
    kernel.set_stream(torch.cuda.current_stream().cuda_stream)
    t = t_start
    for i in range(steps):
        if i > 0:
            v = potential_at(t).to(torch.float64).contiguous()
            kernel.set_potential_exp_device(v.data_ptr())
        kernel.step(exchange=i != steps - 1)
        t += dt
    return t
