"""Helpers shared by the golden-vector generator and the tests."""
import hashlib
import os
import platform

import numpy as np


def input_digest(wl) -> str:
    """SHA-256 of the workload's input bytes (initial tile, then potential)."""
    h = hashlib.sha256()
    for a in (wl.psi.real, wl.psi.imag, wl.v):
        h.update(np.ascontiguousarray(a, dtype=np.float64).tobytes())
    return h.hexdigest()


def host_fingerprint() -> str:
    """Machine, C library and CPU model: where libm results are reproducible."""
    cpu = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    libc = os.confstr("CS_GNU_LIBC_VERSION") if hasattr(os, "confstr") else ""
    return f"{platform.machine()}|{libc}|{cpu}"


def pair_digest(pw) -> str:
    """SHA-256 of a two-component workload's input bytes (both initial tiles,
    then both potentials)."""
    h = hashlib.sha256()
    for a in (pw.psi_a.real, pw.psi_a.imag, pw.psi_b.real, pw.psi_b.imag, pw.v_a, pw.v_b):
        h.update(np.ascontiguousarray(a, dtype=np.float64).tobytes())
    return h.hexdigest()
