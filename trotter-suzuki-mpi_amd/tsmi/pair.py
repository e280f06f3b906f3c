"""Python view of a two-component system on the GPU (libtsmi_hip.so, tsmi_pair_*).

``HipPair(pw, precision)`` builds the two-state kernel for a PairWorkload
(Hamiltonian2Component with two states, src/trottersuzuki.h:459-497): both
states, both E tables and both raw potentials on the device.  ``evolve(n)`` is
the two-component loop of Solver::evolve (src/solver.cpp:172-230);
``state(which)`` reads one component back; ``energy()`` gives the Solver's
two-component energies (src/solver.cpp:244-564).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .hip import exp_potential

ENERGY_KEYS = ("total", "kinetic_a", "kinetic_b", "potential_a", "potential_b", "intra_a", "intra_b", "inter",
               "rabi", "norm_a", "norm_b", "norm")


def _desc(pw, which: int, precision: int, norm_target: float, tile_rows: int, sweeps: int = 0) -> _lib.Desc:
    lat = pw.lattice
    mass, g = (pw.mass_a, pw.g_a) if which == 0 else (pw.mass_b, pw.g_b)
    return _lib.Desc(lat.dim_x, lat.dim_y, int(lat.periodic_x), int(lat.periodic_y), lat.halo_x, lat.halo_y,
                     int(pw.imag_time), int(precision), int(tile_rows), int(sweeps), float(pw.dt), float(mass), float(g), 0.0,
                     lat.delta_x, lat.delta_y, lat.length_x, lat.length_y, float(norm_target))


def squared_norm(lat, psi: np.ndarray) -> float:
    """State::get_squared_norm (src/model.cpp:418,534): inner sum times dx*dy."""
    inner = psi[lat.inner_y, lat.inner_x]
    return float(np.sum(inner.real ** 2 + inner.imag ** 2) * (lat.delta_x * lat.delta_y))


class HipPair:
    def __init__(self, pw, precision: int = 64, tile_rows: int = 0, e_planes=None, norm_targets=None,
                 sweeps: int = 0):
        L = _lib.load()
        self._L = L
        self.pw = pw
        if norm_targets is None:
            norm_targets = (squared_norm(pw.lattice, pw.psi_a), squared_norm(pw.lattice, pw.psi_b))
        d = _lib.PairDesc(_desc(pw, 0, precision, norm_targets[0], tile_rows, sweeps),
                          _desc(pw, 1, precision, norm_targets[1], tile_rows, sweeps), float(pw.g_ab), float(pw.omega_r),
                          float(pw.omega_i))
        h = C.c_void_p()
        _lib.check(L.tsmi_pair_create(C.byref(d), C.byref(h)), "tsmi_pair_create")
        self._h = h
        for which, (psi, v) in enumerate(((pw.psi_a, pw.v_a), (pw.psi_b, pw.v_b))):
            er, ei = e_planes[which] if e_planes is not None else exp_potential(v, pw.dt, pw.imag_time)
            self._call("tsmi_pair_set_exp_potential", which, self._plane(er), self._plane(ei))
            self._call("tsmi_pair_set_potential", which, self._plane(v))
            self.set_state(which, psi)

    def _plane(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        lat = self.pw.lattice
        if a.shape != (lat.H, lat.W):
            raise ValueError(f"expected {(lat.H, lat.W)}, got {a.shape}")
        return a

    def _call(self, name, which, *planes):
        _lib.check(getattr(self._L, name)(self._h, int(which), *[_lib.dptr(p) for p in planes]), name)

    def set_state(self, which: int, psi: np.ndarray) -> None:
        self._call("tsmi_pair_set_state", which, self._plane(psi.real), self._plane(psi.imag))

    def evolve(self, n_steps: int, initial_half: bool = True) -> None:
        _lib.check(self._L.tsmi_pair_evolve(self._h, int(n_steps), int(initial_half)), "tsmi_pair_evolve")

    def synchronize(self) -> None:
        _lib.check(self._L.tsmi_pair_synchronize(self._h), "tsmi_pair_synchronize")

    def state(self, which: int) -> np.ndarray:
        lat = self.pw.lattice
        re = np.zeros((lat.H, lat.W))
        im = np.zeros((lat.H, lat.W))
        _lib.check(self._L.tsmi_pair_get_state(self._h, int(which), _lib.dptr(re), _lib.dptr(im)),
                   "tsmi_pair_get_state")
        return re + 1j * im

    def energy(self) -> np.ndarray:
        out = np.zeros(12)
        _lib.check(self._L.tsmi_pair_energy(self._h, _lib.dptr(out)), "tsmi_pair_energy")
        return out

    def geometry(self) -> dict:
        out = (C.c_int32 * 6)()
        _lib.check(self._L.tsmi_pair_geometry(self._h, out), "tsmi_pair_geometry")
        return dict(zip(("W", "H", "pitch", "strips", "segments", "rows_per_segment"), list(out)))

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            self._L.tsmi_pair_destroy(self._h)
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
