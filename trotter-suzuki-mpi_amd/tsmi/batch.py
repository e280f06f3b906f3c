"""Python view of a batch of independent lattices (libtsmi_hip.so, tsmi_batch_*).

``HipBatch(workloads, precision)`` puts every workload's lattice, state and E
table on the device; ``evolve(n)`` advances all of them by n steps with one
persistent launch per step (HIP-graph captured); ``state(i)`` reads lattice i
back; ``tune()`` runs the segment-height autotuner.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .hip import exp_potential


def desc_of(wl, precision: int, norm_target: float = 0.0, tile_rows: int = 0) -> _lib.Desc:
    lat = wl.lattice
    return _lib.Desc(lat.dim_x, lat.dim_y, int(lat.periodic_x), int(lat.periodic_y), lat.halo_x, lat.halo_y,
                     int(wl.imag_time), int(precision), int(tile_rows), 0, float(wl.dt), float(wl.mass),
                     float(wl.g), float(wl.g_lhy), lat.delta_x, lat.delta_y, lat.length_x, lat.length_y,
                     float(norm_target))


class HipBatch:
    def __init__(self, workloads, precision: int = 32, tile_rows: int = 0, norm_targets=None, e_planes=None):
        L = _lib.load()
        self._L = L
        self.workloads = list(workloads)
        n = len(self.workloads)
        nts = norm_targets if norm_targets is not None else [0.0] * n
        descs = (_lib.Desc * n)(*[desc_of(w, precision, nts[i], tile_rows) for i, w in enumerate(self.workloads)])
        h = C.c_void_p()
        _lib.check(L.tsmi_batch_create(descs, n, C.byref(h)), "tsmi_batch_create")
        self._h = h
        for i, w in enumerate(self.workloads):
            er, ei = e_planes[i] if e_planes is not None else exp_potential(w.v, w.dt, w.imag_time)
            self.set_exp_potential(i, er, ei)
            self.set_state(i, w.psi)

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            self._L.tsmi_batch_destroy(self._h)
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

    def _shape(self, i):
        lat = self.workloads[i].lattice
        return lat.H, lat.W

    def _plane(self, i, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.shape != self._shape(i):
            raise ValueError(f"lattice {i}: expected {self._shape(i)}, got {a.shape}")
        return a

    def set_state(self, i: int, psi: np.ndarray) -> None:
        re, im = self._plane(i, psi.real), self._plane(i, psi.imag)
        _lib.check(self._L.tsmi_batch_set_state(self._h, i, _lib.dptr(re), _lib.dptr(im)), "tsmi_batch_set_state")

    def set_exp_potential(self, i: int, e_re, e_im) -> None:
        re = self._plane(i, e_re)
        im = self._plane(i, e_im if e_im is not None else np.zeros_like(re))
        _lib.check(self._L.tsmi_batch_set_exp_potential(self._h, i, _lib.dptr(re), _lib.dptr(im)),
                   "tsmi_batch_set_exp_potential")

    def evolve(self, n_steps: int) -> None:
        _lib.check(self._L.tsmi_batch_evolve(self._h, int(n_steps)), "tsmi_batch_evolve")

    def synchronize(self) -> None:
        _lib.check(self._L.tsmi_batch_synchronize(self._h), "tsmi_batch_synchronize")

    def state(self, i: int) -> np.ndarray:
        H, W = self._shape(i)
        re = np.zeros((H, W))
        im = np.zeros((H, W))
        _lib.check(self._L.tsmi_batch_get_state(self._h, i, _lib.dptr(re), _lib.dptr(im)), "tsmi_batch_get_state")
        return re + 1j * im

    def tune(self, steps: int = 3) -> float:
        """Autotunes the segment heights; returns the tuned ms per batch step."""
        out = C.c_double()
        _lib.check(self._L.tsmi_batch_tune(self._h, int(steps), C.byref(out)), "tsmi_batch_tune")
        return out.value

    def info(self, i: int) -> dict:
        out = (C.c_int32 * 7)()
        _lib.check(self._L.tsmi_batch_info(self._h, i, out), "tsmi_batch_info")
        return dict(zip(("W", "H", "pitch", "strips", "segments", "rows_per_segment", "steps_per_pass"), list(out)))
