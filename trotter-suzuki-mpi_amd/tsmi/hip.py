"""Python view of one HIP kernel context (libtsmi_hip.so).

``HipKernel`` holds what the reference's ``ITrotterKernel`` implementation
holds for one component (src/kernel.h:66-151): the lattice geometry, the
kinetic/nonlinear coefficients, the state and the potential factor E on the
device.  Arrays are complex128/float64 numpy arrays of the tile shape (H, W).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .lattice import Lattice2D


def exp_potential(v: np.ndarray, dt: float, imag_time: bool):
    """E = exp(-i dt V) or exp(-dt V) (src/solver.cpp:98-105) as two float64 planes."""
    v = np.asarray(v, dtype=np.float64)
    if imag_time:
        return np.exp(-dt * v), np.zeros_like(v)
    e = np.exp(-1j * dt * v)
    return np.ascontiguousarray(e.real), np.ascontiguousarray(e.imag)


class HipKernel:
    """One device context; ``evolve(n)`` advances n full Trotter-Suzuki steps."""

    def __init__(self, lattice: Lattice2D, *, delta_t: float, mass: float = 1.0, coupling: float = 0.0,
                 coupling_lhy: float = 0.0, imag_time: bool = False, precision: int = 64,
                 norm_target: float = 0.0, tile_rows: int = 0, sweeps: int = 0):
        L = _lib.load()
        self.lattice = lattice
        self.delta_t = float(delta_t)
        self.imag_time = bool(imag_time)
        self.precision = int(precision)
        d = _lib.Desc(lattice.dim_x, lattice.dim_y, int(lattice.periodic_x), int(lattice.periodic_y),
                      lattice.halo_x, lattice.halo_y, int(imag_time), int(precision), int(tile_rows), int(sweeps),
                      float(delta_t), float(mass), float(coupling), float(coupling_lhy),
                      lattice.delta_x, lattice.delta_y, lattice.length_x, lattice.length_y, float(norm_target))
        h = C.c_void_p()
        _lib.check(L.tsmi_create(C.byref(d), C.byref(h)), "tsmi_create")
        self._h = h
        self._L = L

    # -- lifetime ----------------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            self._L.tsmi_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- data --------------------------------------------------------------------
    def _plane(self, a) -> np.ndarray:
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.shape != (self.lattice.H, self.lattice.W):
            raise ValueError(f"expected tile shape {(self.lattice.H, self.lattice.W)}, got {a.shape}")
        return a

    def set_state(self, psi: np.ndarray) -> None:
        re, im = self._plane(psi.real), self._plane(psi.imag)
        _lib.check(self._L.tsmi_set_state(self._h, _lib.dptr(re), _lib.dptr(im)), "tsmi_set_state")

    def set_exp_potential(self, e_re: np.ndarray, e_im: np.ndarray | None) -> None:
        re = self._plane(e_re)
        im = self._plane(e_im) if e_im is not None else None
        _lib.check(self._L.tsmi_set_exp_potential(self._h, _lib.dptr(re), _lib.dptr(im) if im is not None else None),
                   "tsmi_set_exp_potential")

    def set_potential(self, v: np.ndarray) -> None:
        v = self._plane(v)
        _lib.check(self._L.tsmi_set_potential(self._h, _lib.dptr(v)), "tsmi_set_potential")

    # -- device-resident data (pointers to W*H FP64 arrays, e.g. torch tensors) ----
    def set_state_device(self, re_ptr: int, im_ptr: int) -> None:
        _lib.check(self._L.tsmi_set_state_device(self._h, C.c_void_p(re_ptr), C.c_void_p(im_ptr)),
                   "tsmi_set_state_device")

    def set_exp_potential_device(self, e_re_ptr: int, e_im_ptr: int | None) -> None:
        _lib.check(self._L.tsmi_set_exp_potential_device(self._h, C.c_void_p(e_re_ptr), C.c_void_p(e_im_ptr or 0)),
                   "tsmi_set_exp_potential_device")

    def set_potential_device(self, v_ptr: int) -> None:
        _lib.check(self._L.tsmi_set_potential_device(self._h, C.c_void_p(v_ptr)), "tsmi_set_potential_device")

    def set_potential_exp_device(self, v_ptr: int) -> None:
        """E = exp(-i dt V) (or exp(-dt V)) tabulated on the device from a W*H
        FP64 device array, which is also kept as the raw V for energy()."""
        _lib.check(self._L.tsmi_set_potential_exp_device(self._h, C.c_void_p(v_ptr)), "tsmi_set_potential_exp_device")

    def step(self, exchange: bool = True) -> None:
        """One step; the periodic halo exchange after it when ``exchange``."""
        _lib.check(self._L.tsmi_step(self._h, int(bool(exchange))), "tsmi_step")

    def get_state_device(self, re_ptr: int, im_ptr: int) -> None:
        _lib.check(self._L.tsmi_get_state_device(self._h, C.c_void_p(re_ptr), C.c_void_p(im_ptr)),
                   "tsmi_get_state_device")

    def set_stream(self, stream_handle: int | None) -> None:
        _lib.check(self._L.tsmi_set_stream(self._h, C.c_void_p(stream_handle or 0)), "tsmi_set_stream")

    def evolve(self, n_steps: int) -> None:
        _lib.check(self._L.tsmi_evolve(self._h, int(n_steps)), "tsmi_evolve")

    # -- gradient with respect to the potential (real time, closed, g = 0) -------
    def evolve_checkpointed(self, n_steps: int, every: int = 100) -> None:
        """evolve(n_steps), keeping the state every ``every`` steps for adjoint()."""
        _lib.check(self._L.tsmi_evolve_checkpointed(self._h, int(n_steps), int(every)), "tsmi_evolve_checkpointed")

    def adjoint(self, g_final: np.ndarray):
        """Reverse sweep from G = dL/dRe psi_N + i dL/dIm psi_N; returns
        (dL/dV, dL/dRe psi_0 + i dL/dIm psi_0) on the tile."""
        gr, gi = self._plane(g_final.real), self._plane(g_final.imag)
        gv = np.zeros_like(gr)
        g0r = np.zeros_like(gr)
        g0i = np.zeros_like(gr)
        _lib.check(self._L.tsmi_adjoint(self._h, _lib.dptr(gr), _lib.dptr(gi), _lib.dptr(gv), _lib.dptr(g0r),
                                        _lib.dptr(g0i)), "tsmi_adjoint")
        return gv, g0r + 1j * g0i

    def adjoint_device(self, g_re_ptr: int, g_im_ptr: int, grad_ptr: int, g0_re_ptr: int | None = None,
                       g0_im_ptr: int | None = None) -> None:
        """adjoint() on device W*H FP64 arrays; asynchronous on the context stream."""
        _lib.check(self._L.tsmi_adjoint_device(self._h, C.c_void_p(g_re_ptr), C.c_void_p(g_im_ptr),
                                               C.c_void_p(grad_ptr), C.c_void_p(g0_re_ptr or 0),
                                               C.c_void_p(g0_im_ptr or 0)), "tsmi_adjoint_device")

    def kernel_timing(self, enable: bool = True) -> None:
        """Brackets every step-kernel launch with HIP events from now on."""
        _lib.check(self._L.tsmi_kernel_timing(self._h, int(bool(enable))), "tsmi_kernel_timing")

    def kernel_time(self) -> dict:
        """Syncs; ms and launch counts of the fused and single-step kernels."""
        out = np.zeros(4)
        _lib.check(self._L.tsmi_kernel_time(self._h, _lib.dptr(out)), "tsmi_kernel_time")
        return {"fused_ms": out[0], "fused_launches": int(out[1]), "single_ms": out[2], "single_launches": int(out[3])}

    def synchronize(self) -> None:
        _lib.check(self._L.tsmi_synchronize(self._h), "tsmi_synchronize")

    def sample(self, x: int = 0, y: int = 0, width: int | None = None, height: int | None = None) -> np.ndarray:
        lat = self.lattice
        w = lat.W - x if width is None else width
        h = lat.H - y if height is None else height
        re = np.zeros((h, w))
        im = np.zeros((h, w))
        _lib.check(self._L.tsmi_get_sample(self._h, w, x, y, w, h, _lib.dptr(re), _lib.dptr(im)), "tsmi_get_sample")
        return re + 1j * im

    def squared_norm(self) -> float:
        out = C.c_double()
        _lib.check(self._L.tsmi_squared_norm(self._h, C.byref(out)), "tsmi_squared_norm")
        return out.value

    def energy(self) -> np.ndarray:
        out = np.zeros(6)
        _lib.check(self._L.tsmi_energy(self._h, _lib.dptr(out)), "tsmi_energy")
        return out

    def expected_values(self) -> np.ndarray:
        out = np.zeros(10)
        _lib.check(self._L.tsmi_expected_values(self._h, _lib.dptr(out)), "tsmi_expected_values")
        return out

    def geometry(self) -> dict:
        out = (C.c_int32 * 6)()
        _lib.check(self._L.tsmi_geometry(self._h, out), "tsmi_geometry")
        return dict(zip(("W", "H", "pitch", "strips", "segments", "rows_per_segment"), list(out)))


def kernel_for(wl, precision: int = 64, norm_target: float | None = None, tile_rows: int = 0,
               e_planes=None, sweeps: int = 0) -> HipKernel:
    """A context loaded with a workload's state, E table and raw potential."""
    lat = wl.lattice
    if norm_target is None:
        norm_target = 0.0
        if wl.imag_time:
            inner = wl.psi[lat.inner_y, lat.inner_x]
            norm_target = float(np.sum(inner.real ** 2 + inner.imag ** 2) * (lat.delta_x * lat.delta_y))
    k = HipKernel(lat, delta_t=wl.dt, mass=wl.mass, coupling=wl.g, coupling_lhy=wl.g_lhy,
                  imag_time=wl.imag_time, precision=precision, norm_target=norm_target, tile_rows=tile_rows,
                  sweeps=sweeps)
    er, ei = e_planes if e_planes is not None else exp_potential(wl.v, wl.dt, wl.imag_time)
    k.set_exp_potential(er, ei)
    k.set_potential(wl.v)
    k.set_state(wl.psi)
    return k
