#!/usr/bin/env python3
"""Writes tests/golden/adjoint_config4.npz: central finite differences of the
reference solver for the differentiation workload.

Workload: 1024 x 1024 closed lattice of side 20, linear Schroedinger, V a
seeded sum of 16 Gaussians, a seeded Gaussian packet, 2000 real-time steps at
dt = 1e-3 (tsmi.workloads.config4).  Loss: L = 1 - |<target|psi(T)>|^2 with the
displaced Gaussian target of the same builder.  For 32 grid cells, drawn from
a SplitMix64 stream with probability proportional to |psi_0|^2, the reference
solver (oracle/_ref/ref_driver_strict: the reference's own sources, whose
Solver computes E = exp(-i dt V) itself) runs with V -/+ h at that cell,
h = 1e-4, and dL/dV = (L(V+h) - L(V-h)) / 2h.

Runs in the build container only (it needs the reference build); about 40
minutes on 8 cores.  The fixture holds data only: the cell indices, h, the loss
of the unperturbed run and the 32 differences.

    python tests/golden/make_adjoint_golden.py
"""
from __future__ import annotations

import dataclasses
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.abspath(os.path.join(HERE, "..", ".."))
sys.path.insert(0, os.path.join(ROOT, "trotter-suzuki-mpi_amd"))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from tsmi import workloads as W  # noqa: E402
from tsmi.losses import infidelity  # noqa: E402
from oracle import pyoracle as O  # noqa: E402
from golden_util import host_fingerprint, input_digest  # noqa: E402

N_POINTS = 32
H_FD = 1e-4


def fd_points(wl, count: int) -> np.ndarray:
    """count distinct cells, each drawn with probability ~ |psi_0|^2 by inverse
    CDF on SplitMix64 uniforms (seed 0x5EED0004, draws from offset 1000)."""
    w = np.abs(wl.psi.ravel()) ** 2
    cdf = np.cumsum(w) / w.sum()
    out: list[int] = []
    off = 1000
    while len(out) < count:
        u = W.uniform(W.SEED_BASE + 4, 1, offset=off)[0]
        off += 1
        j = int(min(np.searchsorted(cdf, u, side="right"), w.size - 1))
        if j not in out:
            out.append(j)
    return np.array(out, dtype=np.int64)


def main() -> None:
    if not O.ref_available("strict"):
        sys.exit("oracle/_ref/ref_driver_strict is not built (make -C oracle)")
    threads = min(8, os.cpu_count() or 1)
    wl, target = W.config4()
    lat = wl.lattice

    def loss(v):
        psi, _, _ = O.run_reference(dataclasses.replace(wl, v=v), kind="strict", threads=threads)
        return infidelity(lat, target, psi)[0]

    pts = fd_points(wl, N_POINTS)
    t0 = time.time()
    L0 = loss(wl.v)
    fd = np.zeros(N_POINTS)
    for i, j in enumerate(pts):
        vp = wl.v.copy(); vp.flat[j] += H_FD
        vm = wl.v.copy(); vm.flat[j] -= H_FD
        fd[i] = (loss(vp) - loss(vm)) / (2 * H_FD)
        print(f"[{i + 1}/{N_POINTS}] cell {j}: dL/dV = {fd[i]:.12e}  ({time.time() - t0:.0f} s)", flush=True)
    spec = dict(builder="config4", kwargs={}, steps=wl.steps, dt=wl.dt, h=H_FD, loss="1-|<target|psi(T)>|^2",
                driver="ref_driver_strict")
    np.savez(os.path.join(HERE, "adjoint_config4.npz"), spec=json.dumps(spec), points=pts, h=H_FD,
             loss0=L0, fd=fd, input_sha256=input_digest(wl), host=host_fingerprint())
    print("wrote adjoint_config4.npz", flush=True)


if __name__ == "__main__":
    main()
