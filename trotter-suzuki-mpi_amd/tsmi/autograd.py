"""torch.autograd bridge: the real-time evolution as a differentiable function
of the potential (and of the initial state).

``evolve(kernel, v, psi_re, psi_im, steps, every)`` runs ``steps`` real-time
steps on the device with E = exp(-i dt V) built from the FP64 tensor ``v`` and
returns the final state as two FP64 tensors.  Its backward pass is the
library's reverse sweep (tsmi_adjoint_device): the step at -dt applied to the
adjoint and to the state together, re-anchored on checkpoints kept every
``every`` steps, with dL/dV accumulated in FP64.  Real time, any g and g_LHY,
closed or periodic axes as tsmi_adjoint allows (on periodic axes the gradients
are those of the periodic fields: each inner cell's value shared with its halo
copies, zero on the halo); the kernel's precision (32 or 64) is the precision
of both sweeps.  Everything runs on torch's current stream.
"""
from __future__ import annotations

import torch

from .hip import HipKernel


def _stream(device) -> int:
    return torch.cuda.current_stream(device).cuda_stream


class _Evolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, v, psi_re, psi_im, kernel: HipKernel, steps: int, every: int):
        v = v.detach().to(torch.float64).contiguous()
        kernel.set_stream(_stream(v.device))
        phase = -kernel.delta_t * v
        e_re = torch.cos(phase).contiguous()
        e_im = torch.sin(phase).contiguous()
        kernel.set_exp_potential_device(e_re.data_ptr(), e_im.data_ptr())
        p_re = psi_re.detach().to(torch.float64).contiguous()
        p_im = psi_im.detach().to(torch.float64).contiguous()
        kernel.set_state_device(p_re.data_ptr(), p_im.data_ptr())
        kernel.evolve_checkpointed(steps, every)
        out_re = torch.empty_like(v)
        out_im = torch.empty_like(v)
        kernel.get_state_device(out_re.data_ptr(), out_im.data_ptr())
        ctx.kernel = kernel
        return out_re, out_im

    @staticmethod
    def backward(ctx, g_re, g_im):
        kernel: HipKernel = ctx.kernel
        ref = g_re if g_re is not None else g_im
        g_re = torch.zeros_like(ref) if g_re is None else g_re.to(torch.float64).contiguous()
        g_im = torch.zeros_like(ref) if g_im is None else g_im.to(torch.float64).contiguous()
        kernel.set_stream(_stream(g_re.device))
        grad_v = torch.empty_like(g_re)
        g0_re = torch.empty_like(g_re)
        g0_im = torch.empty_like(g_re)
        kernel.adjoint_device(g_re.data_ptr(), g_im.data_ptr(), grad_v.data_ptr(), g0_re.data_ptr(),
                              g0_im.data_ptr())
        return grad_v, g0_re, g0_im, None, None, None


def evolve(kernel: HipKernel, v: torch.Tensor, psi_re: torch.Tensor, psi_im: torch.Tensor, steps: int,
           every: int = 100):
    """Final state (re, im) after ``steps`` real-time steps; differentiable in
    v, psi_re and psi_im.  All tensors are (H, W) FP64 on the kernel's device."""
    return _Evolve.apply(v, psi_re, psi_im, kernel, int(steps), int(every))


def infidelity(target_re: torch.Tensor, target_im: torch.Tensor, re: torch.Tensor, im: torch.Tensor,
               dxdy: float, inner=(slice(None), slice(None))):
    """1 - |<target|psi>|^2 with <target|psi> = sum conj(target) psi dx dy over ``inner``."""
    tr, ti, pr, pi = target_re[inner], target_im[inner], re[inner], im[inner]
    c_re = (tr * pr + ti * pi).sum() * dxdy
    c_im = (tr * pi - ti * pr).sum() * dxdy
    return 1.0 - (c_re * c_re + c_im * c_im)
