#!/usr/bin/env python3
"""Writes the golden vectors under tests/golden/ from the reference solver.

Runs in the build container only: it needs oracle/_ref/ref_driver_strict, which
oracle/Makefile compiles from the reference sources (-O2 -ffp-contract=off).  The
driver calls the reference's public API (Lattice2D, State, Potential,
Hamiltonian, Solver::evolve and the energy / expected-value getters) on inputs
built by tsmi.workloads, single-threaded so that the reductions run in row order.

Each fixture is an .npz (no pickles) holding the case spec (JSON), the SHA-256 of
the input bytes, the host fingerprint, and the reference outputs: the final tile
(full, or strided for large tiles), the six Solver energies and the ten State
expected values.

    python tests/golden/make_golden.py [name ...]
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.abspath(os.path.join(HERE, "..", ".."))
sys.path.insert(0, os.path.join(ROOT, "trotter-suzuki-mpi_amd"))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from tsmi import workloads as W  # noqa: E402
from oracle import pyoracle as O  # noqa: E402
from golden_util import host_fingerprint, input_digest  # noqa: E402

# name -> (builder, kwargs, full_tile, stride)
CASES = {
    "small64_closed_real_1": ("small_case", dict(n_x=64, n_y=64, periodic_x=False, periodic_y=False, steps=1), True),
    "small64_closed_gpe_lhy_100": ("small_case", dict(n_x=64, n_y=64, periodic_x=False, periodic_y=False,
                                                      steps=100, g=5.0, g_lhy=1.0), True),
    "small64x48_periodic_x_30": ("small_case", dict(n_x=64, n_y=48, periodic_x=True, periodic_y=False,
                                                    steps=30, g=3.0), True),
    "small48x64_periodic_y_30": ("small_case", dict(n_x=48, n_y=64, periodic_x=False, periodic_y=True,
                                                    steps=30, g=3.0), True),
    "small64_periodic_50": ("small_case", dict(n_x=64, n_y=64, periodic_x=True, periodic_y=True,
                                               steps=50, g=2.0), True),
    "small64_closed_imag_100": ("small_case", dict(n_x=64, n_y=64, periodic_x=False, periodic_y=False,
                                                   imag_time=True, steps=100, g=10.0, g_lhy=2.0), True),
    "small64_periodic_imag_60": ("small_case", dict(n_x=64, n_y=64, periodic_x=True, periodic_y=True,
                                                    imag_time=True, steps=60, g=10.0), True),
    "blocks260x136_closed_20": ("small_case", dict(n_x=260, n_y=136, periodic_x=False, periodic_y=False,
                                                   steps=20, g=2.0), True),
    "config0": ("config0", dict(), True),
}
for _s in W.REFERENCE_SCENARIOS:
    CASES["scenario_" + _s] = ("reference_scenario", dict(name=_s), False)

STRIDE = 5


def build(builder: str, kwargs: dict):
    return getattr(W, builder)(**kwargs)


def make(name: str) -> str:
    builder, kwargs, full = CASES[name]
    wl = build(builder, kwargs)
    psi, e, x = O.run_reference(wl, threads=1)
    out = dict(spec=np.array(json.dumps(dict(builder=builder, kwargs=kwargs))),
               input_sha256=np.array(input_digest(wl)), host=np.array(host_fingerprint()),
               energies=e, expect=x, steps=np.array(wl.steps))
    if full:
        out["psi_re"], out["psi_im"] = psi.real.copy(), psi.imag.copy()
    else:
        out["stride"] = np.array(STRIDE)
        out["psi_re"], out["psi_im"] = psi.real[::STRIDE, ::STRIDE].copy(), psi.imag[::STRIDE, ::STRIDE].copy()
        # energies and expected values of the initial state (Solver built, no step)
        _, e0, x0 = O.run_reference(wl, steps=0, threads=1)
        out["energies0"], out["expect0"] = e0, x0
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **out)
    return path


def main(argv):
    if not O.ref_available("strict"):
        sys.exit("oracle/_ref/ref_driver_strict missing: run make -C oracle ref in the build container")
    names = argv or list(CASES)
    for n in names:
        print(make(n), flush=True)


if __name__ == "__main__":
    main(sys.argv[1:])
