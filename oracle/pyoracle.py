"""ctypes view of the CPU oracle (oracle/libts_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker; the product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libts_oracle.so")
REF_DIR = os.path.join(HERE, "_ref")

_dp = C.POINTER(C.c_double)


class Desc(C.Structure):
    _fields_ = [("W", C.c_int), ("H", C.c_int), ("halo_x", C.c_int), ("halo_y", C.c_int),
                ("periodic_x", C.c_int), ("periodic_y", C.c_int), ("imag_time", C.c_int),
                ("aH", C.c_double), ("bH", C.c_double), ("aV", C.c_double), ("bV", C.c_double),
                ("c_nl", C.c_double), ("c_lhy", C.c_double), ("dx", C.c_double), ("dy", C.c_double),
                ("norm_target", C.c_double)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        L = C.CDLL(LIB_PATH)
        L.ts_oracle_coefficients.argtypes = [C.c_double] * 6 + [C.c_int, _dp]
        L.ts_oracle_exp_potential.argtypes = [C.c_longlong, _dp, C.c_double, C.c_int, _dp, _dp]
        L.ts_oracle_full_step.argtypes = [C.POINTER(Desc), _dp, _dp, _dp, _dp]
        L.ts_oracle_pass_vertical.argtypes = [C.POINTER(Desc), C.c_int, C.c_double, C.c_double, _dp, _dp]
        L.ts_oracle_pass_horizontal.argtypes = [C.POINTER(Desc), C.c_int, C.c_double, C.c_double, _dp, _dp]
        L.ts_oracle_pass_potential.argtypes = [C.POINTER(Desc), _dp, _dp, _dp, _dp]
        L.ts_oracle_halo_exchange.argtypes = [C.POINTER(Desc), _dp, _dp]
        L.ts_oracle_squared_norm.argtypes = [C.POINTER(Desc), _dp, _dp]
        L.ts_oracle_squared_norm.restype = C.c_double
        L.ts_oracle_state_norm.argtypes = [C.POINTER(Desc), _dp, _dp]
        L.ts_oracle_state_norm.restype = C.c_double
        L.ts_oracle_evolve.argtypes = [C.POINTER(Desc), C.c_int, _dp, _dp, _dp, _dp]
        L.ts_oracle_energy.argtypes = [C.POINTER(Desc), _dp, _dp, _dp, C.c_double, C.c_double, C.c_double,
                                       C.c_double, C.c_int, _dp]
        L.ts_oracle_expected_values.argtypes = [C.POINTER(Desc), _dp, _dp, _dp, _dp, _dp]
        L.ts_oracle_adjoint.argtypes = [C.POINTER(Desc), C.c_int, C.c_int, _dp, _dp, _dp, _dp, C.c_double,
                                        _dp, _dp, _dp, _dp, _dp]
        L.ts_oracle_adjoint.restype = C.c_int
        L.ts_oracle_halo_exchange_adjoint.argtypes = [C.POINTER(Desc), _dp, _dp]
        _lib = L
    return _lib


def _p(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def coefficients(dt, mass, dx, dy, g, g_lhy, imag_time):
    out = np.zeros(6)
    lib().ts_oracle_coefficients(dt, mass, dx, dy, g, g_lhy, int(imag_time), _p(out))
    return out


def exp_potential(v: np.ndarray, dt: float, imag_time: bool):
    v = np.ascontiguousarray(v, dtype=np.float64)
    er = np.empty_like(v)
    ei = np.empty_like(v)
    lib().ts_oracle_exp_potential(v.size, _p(v), dt, int(imag_time), _p(er), _p(ei))
    return er, ei


def make_desc(wl, norm_target: float | None = None) -> Desc:
    lat = wl.lattice
    co = coefficients(wl.dt, wl.mass, lat.delta_x, lat.delta_y, wl.g, wl.g_lhy, wl.imag_time)
    d = Desc(lat.W, lat.H, lat.halo_x, lat.halo_y, int(lat.periodic_x), int(lat.periodic_y),
             int(wl.imag_time), *co, lat.delta_x, lat.delta_y, 0.0)
    if norm_target is None:
        re = np.ascontiguousarray(wl.psi.real, dtype=np.float64)
        im = np.ascontiguousarray(wl.psi.imag, dtype=np.float64)
        norm_target = lib().ts_oracle_state_norm(C.byref(d), _p(re), _p(im))
    d.norm_target = norm_target
    return d


def evolve(wl, steps: int | None = None, psi: np.ndarray | None = None):
    """Oracle Solver::evolve(steps) on the workload; returns the final tile."""
    steps = wl.steps if steps is None else steps
    psi = wl.psi if psi is None else psi
    d = make_desc(wl)
    er, ei = exp_potential(wl.v, wl.dt, wl.imag_time)
    re = np.ascontiguousarray(psi.real, dtype=np.float64).copy()
    im = np.ascontiguousarray(psi.imag, dtype=np.float64).copy()
    lib().ts_oracle_evolve(C.byref(d), steps, _p(re), _p(im), _p(er), _p(ei))
    return re + 1j * im


def full_step(wl, psi):
    d = make_desc(wl)
    er, ei = exp_potential(wl.v, wl.dt, wl.imag_time)
    re = np.ascontiguousarray(psi.real).copy()
    im = np.ascontiguousarray(psi.imag).copy()
    lib().ts_oracle_full_step(C.byref(d), _p(re), _p(im), _p(er), _p(ei))
    return re + 1j * im


def energy(wl, psi):
    lat = wl.lattice
    d = make_desc(wl, norm_target=1.0)
    re = np.ascontiguousarray(psi.real).copy()
    im = np.ascontiguousarray(psi.imag).copy()
    v = np.ascontiguousarray(wl.v, dtype=np.float64)
    out = np.zeros(6)
    lib().ts_oracle_energy(C.byref(d), _p(re), _p(im), _p(v), wl.mass, wl.g, wl.g_lhy,
                           lat.length_x, lat.dim_x, _p(out))
    return out


def expected_values(wl, psi):
    lat = wl.lattice
    d = make_desc(wl, norm_target=1.0)
    re = np.ascontiguousarray(psi.real).copy()
    im = np.ascontiguousarray(psi.imag).copy()
    xc = np.ascontiguousarray(lat.coords_x())
    yc = np.ascontiguousarray(lat.coords_y())
    out = np.zeros(10)
    lib().ts_oracle_expected_values(C.byref(d), _p(re), _p(im), _p(xc), _p(yc), _p(out))
    return out


def adjoint(wl, g_final: np.ndarray, steps: int | None = None, every: int = 100):
    """Oracle gradient of L(psi_N) with respect to V (FP64 discrete adjoint of
    the real-time evolve, every tile entry its own variable).

    g_final = dL/dRe psi_N + i dL/dIm psi_N on the tile.  Returns
    (dL/dV, dL/dRe psi_0 + i dL/dIm psi_0)."""
    steps = wl.steps if steps is None else steps
    d = make_desc(wl, norm_target=1.0)
    er, ei = exp_potential(wl.v, wl.dt, wl.imag_time)
    re = np.ascontiguousarray(wl.psi.real, dtype=np.float64)
    im = np.ascontiguousarray(wl.psi.imag, dtype=np.float64)
    gr = np.ascontiguousarray(g_final.real, dtype=np.float64)
    gi = np.ascontiguousarray(g_final.imag, dtype=np.float64)
    gv = np.zeros_like(re)
    g0r = np.zeros_like(re)
    g0i = np.zeros_like(re)
    rc = lib().ts_oracle_adjoint(C.byref(d), steps, every, _p(re), _p(im), _p(er), _p(ei), wl.dt,
                                 _p(gr), _p(gi), _p(gv), _p(g0r), _p(g0i))
    if rc != 0:
        raise ValueError(f"ts_oracle_adjoint: unsupported case (code {rc})")
    return gv, g0r + 1j * g0i


def halo_exchange(wl, field: np.ndarray) -> np.ndarray:
    """The periodic halo refresh of a real or complex tile field (a copy)."""
    d = make_desc(wl, norm_target=1.0)
    re = np.ascontiguousarray(np.real(field), dtype=np.float64).copy()
    im = np.ascontiguousarray(np.imag(field), dtype=np.float64).copy()
    lib().ts_oracle_halo_exchange(C.byref(d), _p(re), _p(im))
    return re + 1j * im if np.iscomplexobj(field) else re


def fold_halo(wl, field: np.ndarray) -> np.ndarray:
    """The transpose of the halo refresh (a copy): every halo entry added onto
    the cell it mirrors and cleared -- derivatives with respect to tile entries
    become derivatives with respect to a periodic field."""
    d = make_desc(wl, norm_target=1.0)
    re = np.ascontiguousarray(np.real(field), dtype=np.float64).copy()
    im = np.ascontiguousarray(np.imag(field), dtype=np.float64).copy()
    lib().ts_oracle_halo_exchange_adjoint(C.byref(d), _p(re), _p(im))
    return re + 1j * im if np.iscomplexobj(field) else re


# ---------------------------------------------------------------------------
# The reference library itself (oracle/_ref, built from the reference sources).
# ---------------------------------------------------------------------------

def ref_driver(kind: str = "strict") -> str:
    return os.path.join(REF_DIR, f"ref_driver_{kind}")


def ref_available(kind: str = "strict") -> bool:
    return os.access(ref_driver(kind), os.X_OK)


def write_case(path: str, wl, steps: int, n_warmup: int = 0, time_mode: bool = False) -> None:
    lat = wl.lattice
    hdr = np.array([0x54534D49, 1, lat.dim_x, lat.dim_y, int(lat.periodic_x), int(lat.periodic_y),
                    steps, int(wl.imag_time), n_warmup, int(time_mode)], dtype=np.int32)
    par = np.array([lat.length_x, lat.length_y, wl.dt, wl.mass, wl.g, wl.g_lhy], dtype=np.float64)
    with open(path, "wb") as f:
        f.write(hdr.tobytes())
        f.write(par.tobytes())
        f.write(np.ascontiguousarray(wl.psi.real, dtype=np.float64).tobytes())
        f.write(np.ascontiguousarray(wl.psi.imag, dtype=np.float64).tobytes())
        f.write(np.ascontiguousarray(wl.v, dtype=np.float64).tobytes())


def run_reference(wl, steps: int | None = None, kind: str = "strict", threads: int | None = None):
    """Runs the reference Solver on the workload; returns (psi, energies[6], expect[10])."""
    steps = wl.steps if steps is None else steps
    lat = wl.lattice
    n = lat.W * lat.H
    env = dict(os.environ)
    if threads:
        env["OMP_NUM_THREADS"] = str(threads)
    with tempfile.TemporaryDirectory() as td:
        cpath, opath = os.path.join(td, "case.bin"), os.path.join(td, "out.bin")
        write_case(cpath, wl, steps)
        subprocess.run([ref_driver(kind), cpath, opath], check=True, env=env)
        raw = np.fromfile(opath, dtype=np.float64)
    re, im = raw[:n].reshape(lat.H, lat.W), raw[n:2 * n].reshape(lat.H, lat.W)
    return re + 1j * im, raw[2 * n:2 * n + 6].copy(), raw[2 * n + 6:2 * n + 16].copy()


def time_reference(wl, steps: int, n_warmup: int = 1, kind: str = "fast", threads: int | None = None) -> dict:
    import json
    env = dict(os.environ)
    if threads:
        env["OMP_NUM_THREADS"] = str(threads)
    with tempfile.TemporaryDirectory() as td:
        cpath = os.path.join(td, "case.bin")
        write_case(cpath, wl, steps, n_warmup=n_warmup, time_mode=True)
        out = subprocess.run([ref_driver(kind), cpath], check=True, env=env, capture_output=True, text=True)
    return json.loads(out.stdout.strip().splitlines()[-1])


# ---------------------------------------------------------------------------
# Two-component systems (PairWorkload)
# ---------------------------------------------------------------------------

class Desc2(C.Structure):
    _fields_ = [("a", Desc), ("b", Desc), ("c_ab", C.c_double), ("omega_r", C.c_double), ("omega_i", C.c_double)]


def _lib2():
    L = lib()
    if not getattr(L, "_pair_ready", False):
        L.ts_oracle_rabi.argtypes = [C.POINTER(Desc2), C.c_double, C.c_double, _dp, _dp, _dp, _dp]
        L.ts_oracle_evolve2.argtypes = [C.POINTER(Desc2), C.c_int, C.c_int, C.c_double, _dp, _dp, _dp, _dp,
                                        _dp, _dp, _dp, _dp, _dp, C.c_double]
        L.ts_oracle_energy2.argtypes = [C.POINTER(Desc2), _dp, _dp, _dp, _dp, _dp, _dp] + [C.c_double] * 8 + \
            [C.c_int, _dp]
        L._pair_ready = True
    return L


def make_desc2(pw) -> Desc2:
    return Desc2(make_desc(pw.component(0), norm_target=0.0), make_desc(pw.component(1), norm_target=0.0),
                 pw.g_ab * pw.dt, pw.omega_r, pw.omega_i)


def _planes(psi):
    return (np.ascontiguousarray(psi.real, dtype=np.float64).copy(),
            np.ascontiguousarray(psi.imag, dtype=np.float64).copy())


def pair_exp_potentials(pw):
    return exp_potential(pw.v_a, pw.dt, pw.imag_time), exp_potential(pw.v_b, pw.dt, pw.imag_time)


def evolve2(pw, calls=None, psi_a=None, psi_b=None, initial_half: bool = True):
    """Oracle Solver::evolve for each entry of ``calls`` on one kernel (the
    per-component norm targets persist between calls).  Returns (psi_a, psi_b)."""
    calls = pw.calls if calls is None else calls
    d = make_desc2(pw)
    ar, ai = _planes(pw.psi_a if psi_a is None else psi_a)
    br, bi = _planes(pw.psi_b if psi_b is None else psi_b)
    L = _lib2()
    norm = np.array([L.ts_oracle_state_norm(C.byref(d.a), _p(ar), _p(ai)),
                     L.ts_oracle_state_norm(C.byref(d.b), _p(br), _p(bi))])
    tot = float(norm[0] + norm[1])
    (ear, eai), (ebr, ebi) = pair_exp_potentials(pw)
    for n in calls:
        L.ts_oracle_evolve2(C.byref(d), int(n), int(initial_half), pw.dt, _p(ar), _p(ai), _p(br), _p(bi),
                            _p(ear), _p(eai), _p(ebr), _p(ebi), _p(norm), tot)
    return ar + 1j * ai, br + 1j * bi


def rabi(pw, var: float, psi_a, psi_b):
    d = make_desc2(pw)
    ar, ai = _planes(psi_a)
    br, bi = _planes(psi_b)
    _lib2().ts_oracle_rabi(C.byref(d), var, pw.dt, _p(ar), _p(ai), _p(br), _p(bi))
    return ar + 1j * ai, br + 1j * bi


def energy2(pw, psi_a, psi_b):
    lat = pw.lattice
    d = make_desc2(pw)
    ar, ai = _planes(psi_a)
    br, bi = _planes(psi_b)
    va = np.ascontiguousarray(pw.v_a, dtype=np.float64)
    vb = np.ascontiguousarray(pw.v_b, dtype=np.float64)
    out = np.zeros(12)
    _lib2().ts_oracle_energy2(C.byref(d), _p(ar), _p(ai), _p(br), _p(bi), _p(va), _p(vb), pw.mass_a, pw.mass_b,
                              pw.g_a, pw.g_b, pw.g_ab, pw.omega_r, pw.omega_i, lat.length_x, lat.dim_x, _p(out))
    return out


def write_case_pair(path: str, pw) -> None:
    lat = pw.lattice
    calls = [int(c) for c in pw.calls]
    hdr = np.array([0x54534D49, 2, lat.dim_x, lat.dim_y, int(lat.periodic_x), int(lat.periodic_y),
                    len(calls), int(pw.imag_time), 0, 0], dtype=np.int32)
    par = np.array([lat.length_x, lat.length_y, pw.dt, pw.mass_a, pw.mass_b, pw.g_a, pw.g_ab, pw.g_b,
                    pw.omega_r, pw.omega_i], dtype=np.float64)
    with open(path, "wb") as f:
        f.write(hdr.tobytes())
        f.write(np.array(calls, dtype=np.int32).tobytes())
        f.write(par.tobytes())
        for a in (pw.psi_a.real, pw.psi_a.imag, pw.psi_b.real, pw.psi_b.imag, pw.v_a, pw.v_b):
            f.write(np.ascontiguousarray(a, dtype=np.float64).tobytes())


def run_reference_pair(pw, kind: str = "strict", threads: int | None = None):
    """Runs the reference two-state Solver; returns (psi_a, psi_b, getters[14])."""
    lat = pw.lattice
    n = lat.W * lat.H
    env = dict(os.environ)
    if threads:
        env["OMP_NUM_THREADS"] = str(threads)
    with tempfile.TemporaryDirectory() as td:
        cpath, opath = os.path.join(td, "case.bin"), os.path.join(td, "out.bin")
        write_case_pair(cpath, pw)
        subprocess.run([ref_driver(kind), cpath, opath], check=True, env=env)
        raw = np.fromfile(opath, dtype=np.float64)
    sh = (lat.H, lat.W)
    a = raw[:n].reshape(sh) + 1j * raw[n:2 * n].reshape(sh)
    b = raw[2 * n:3 * n].reshape(sh) + 1j * raw[3 * n:4 * n].reshape(sh)
    return a, b, raw[4 * n:4 * n + 14].copy()
