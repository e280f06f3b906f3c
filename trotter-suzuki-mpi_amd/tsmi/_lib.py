"""ctypes binding of libtsmi_hip.so (include/tsmi_hip.h).

The library is loaded from the package's lib/ directory only; there is no
fallback: if it is missing or fails to load, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "lib", "libtsmi_hip.so")

TSMI_OK = 0

# Every symbol include/tsmi_hip.h declares (checked by the test suite).
EXPORTS = (
    "tsmi_create", "tsmi_destroy", "tsmi_set_stream", "tsmi_set_state", "tsmi_set_state_device",
    "tsmi_set_exp_potential", "tsmi_set_exp_potential_device", "tsmi_set_potential",
    "tsmi_set_potential_device", "tsmi_set_potential_exp_device", "tsmi_evolve", "tsmi_run_steps", "tsmi_step", "tsmi_evolve_checkpointed", "tsmi_adjoint",
    "tsmi_adjoint_device", "tsmi_batch_create", "tsmi_batch_destroy", "tsmi_batch_set_state",
    "tsmi_batch_set_exp_potential", "tsmi_batch_evolve", "tsmi_batch_get_state", "tsmi_batch_tune",
    "tsmi_batch_info", "tsmi_batch_synchronize", "tsmi_kernel_timing", "tsmi_kernel_time", "tsmi_squared_norm", "tsmi_energy", "tsmi_energy_device_noalloc", "tsmi_expected_values",
    "tsmi_get_sample", "tsmi_get_state_device", "tsmi_synchronize", "tsmi_geometry", "tsmi_plan_geometry",
    "tsmi_last_error",
    "tsmi_pair_create", "tsmi_pair_destroy", "tsmi_pair_set_state", "tsmi_pair_set_exp_potential",
    "tsmi_pair_set_potential", "tsmi_pair_evolve", "tsmi_pair_step", "tsmi_pair_rabi", "tsmi_pair_get_state",
    "tsmi_pair_energy",
    "tsmi_pair_geometry", "tsmi_pair_synchronize",
)


class Desc(C.Structure):
    _fields_ = [("dim_x", C.c_int32), ("dim_y", C.c_int32),
                ("periodic_x", C.c_int32), ("periodic_y", C.c_int32),
                ("halo_x", C.c_int32), ("halo_y", C.c_int32),
                ("imag_time", C.c_int32), ("precision", C.c_int32),
                ("tile_rows", C.c_int32), ("sweeps", C.c_int32),
                ("delta_t", C.c_double), ("mass", C.c_double),
                ("coupling", C.c_double), ("coupling_lhy", C.c_double),
                ("delta_x", C.c_double), ("delta_y", C.c_double),
                ("length_x", C.c_double), ("length_y", C.c_double),
                ("norm_target", C.c_double)]


class PairDesc(C.Structure):
    _fields_ = [("a", Desc), ("b", Desc), ("coupling_ab", C.c_double), ("omega_r", C.c_double),
                ("omega_i", C.c_double)]


class TsmiError(RuntimeError):
    pass


_lib = None
_dp = C.POINTER(C.c_double)


def load():
    """Loads the HIP library (raises TsmiError if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise TsmiError(f"HIP library not built: {LIB_PATH} (run __graft_entry__.build())")
    # One HIP runtime per process.  PyTorch ships its own libamdhip64.so.7 and
    # this library links the system one under the same soname, so whichever is
    # loaded first serves both.  Let PyTorch load its runtime first when it is
    # installed: the reverse order hands PyTorch a runtime it was not built with
    # and it then sees no GPU.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    L.tsmi_create.argtypes = [C.POINTER(Desc), C.POINTER(C.c_void_p)]
    L.tsmi_destroy.argtypes = [C.c_void_p]
    L.tsmi_destroy.restype = None
    L.tsmi_set_stream.argtypes = [C.c_void_p, C.c_void_p]
    L.tsmi_set_state.argtypes = [C.c_void_p, _dp, _dp]
    L.tsmi_set_exp_potential.argtypes = [C.c_void_p, _dp, _dp]
    L.tsmi_set_potential.argtypes = [C.c_void_p, _dp]
    L.tsmi_set_state_device.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    L.tsmi_set_exp_potential_device.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    L.tsmi_set_potential_device.argtypes = [C.c_void_p, C.c_void_p]
    L.tsmi_set_potential_exp_device.argtypes = [C.c_void_p, C.c_void_p]
    L.tsmi_get_state_device.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    L.tsmi_evolve.argtypes = [C.c_void_p, C.c_int32]
    L.tsmi_step.argtypes = [C.c_void_p, C.c_int32]
    L.tsmi_run_steps.argtypes = [C.c_void_p, C.c_int32, C.c_int32]
    L.tsmi_evolve_checkpointed.argtypes = [C.c_void_p, C.c_int32, C.c_int32]
    L.tsmi_adjoint.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp]
    L.tsmi_adjoint_device.argtypes = [C.c_void_p] + [C.c_void_p] * 5
    L.tsmi_batch_create.argtypes = [C.POINTER(Desc), C.c_int32, C.POINTER(C.c_void_p)]
    L.tsmi_batch_destroy.argtypes = [C.c_void_p]
    L.tsmi_batch_destroy.restype = None
    L.tsmi_batch_set_state.argtypes = [C.c_void_p, C.c_int32, _dp, _dp]
    L.tsmi_batch_set_exp_potential.argtypes = [C.c_void_p, C.c_int32, _dp, _dp]
    L.tsmi_batch_evolve.argtypes = [C.c_void_p, C.c_int32]
    L.tsmi_batch_get_state.argtypes = [C.c_void_p, C.c_int32, _dp, _dp]
    L.tsmi_batch_tune.argtypes = [C.c_void_p, C.c_int32, _dp]
    L.tsmi_batch_info.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_int32)]
    L.tsmi_batch_synchronize.argtypes = [C.c_void_p]
    L.tsmi_kernel_timing.argtypes = [C.c_void_p, C.c_int32]
    L.tsmi_kernel_time.argtypes = [C.c_void_p, _dp]
    L.tsmi_squared_norm.argtypes = [C.c_void_p, _dp]
    L.tsmi_energy.argtypes = [C.c_void_p, _dp]
    L.tsmi_energy_device_noalloc.argtypes = [C.POINTER(Desc), C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p,
                                             C.c_void_p, C.c_void_p]
    L.tsmi_expected_values.argtypes = [C.c_void_p, _dp]
    L.tsmi_get_sample.argtypes = [C.c_void_p, C.c_size_t, C.c_size_t, C.c_size_t, C.c_size_t, C.c_size_t, _dp, _dp]
    L.tsmi_synchronize.argtypes = [C.c_void_p]
    L.tsmi_geometry.argtypes = [C.c_void_p, C.POINTER(C.c_int32)]
    L.tsmi_plan_geometry.argtypes = [C.POINTER(Desc), C.POINTER(C.c_int32)]
    L.tsmi_pair_create.argtypes = [C.POINTER(PairDesc), C.POINTER(C.c_void_p)]
    L.tsmi_pair_destroy.argtypes = [C.c_void_p]
    L.tsmi_pair_destroy.restype = None
    L.tsmi_pair_set_state.argtypes = [C.c_void_p, C.c_int32, _dp, _dp]
    L.tsmi_pair_set_exp_potential.argtypes = [C.c_void_p, C.c_int32, _dp, _dp]
    L.tsmi_pair_set_potential.argtypes = [C.c_void_p, C.c_int32, _dp]
    L.tsmi_pair_evolve.argtypes = [C.c_void_p, C.c_int32, C.c_int32]
    L.tsmi_pair_step.argtypes = [C.c_void_p, C.c_int32, C.c_double]
    L.tsmi_pair_rabi.argtypes = [C.c_void_p, C.c_double, C.c_int32]
    L.tsmi_pair_get_state.argtypes = [C.c_void_p, C.c_int32, _dp, _dp]
    L.tsmi_pair_energy.argtypes = [C.c_void_p, _dp]
    L.tsmi_pair_geometry.argtypes = [C.c_void_p, C.POINTER(C.c_int32)]
    L.tsmi_pair_synchronize.argtypes = [C.c_void_p]
    L.tsmi_last_error.argtypes = []
    L.tsmi_last_error.restype = C.c_char_p
    for name in EXPORTS:
        if name not in ("tsmi_destroy", "tsmi_batch_destroy", "tsmi_pair_destroy", "tsmi_last_error"):
            getattr(L, name).restype = C.c_int
    _lib = L
    return L


def check(status: int, what: str = "") -> None:
    if status != TSMI_OK:
        msg = load().tsmi_last_error().decode(errors="replace")
        raise TsmiError(f"{what}: {msg} (status {status})" if what else msg)


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)
