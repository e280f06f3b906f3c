"""The C++ host API (include/trottersuzuki.h + HipKernel) and the drop-in.

* test_solver (tests/cpp/test_solver.cpp) runs the reference's single-component
  kernel-test scenarios through Solver with the "hip" / "hip-fp32" kernel types;
  the results are checked against the reference test suite's own known answers
  (test/kerneltest.cpp, TOLERANCE 1e-3, NORM_TOLERANCE 1e-5) and, for FP64,
  against the golden vectors the reference wrote for the same scenarios.
* dropin_check (oracle/dropin_check.cpp) builds HipKernel against the
  reference's own header and library, drives it with the reference's per-step
  protocol and compares it with the reference Solver's CPU run.

The programs are started once when the session starts, before this process
touches the GPU (tests/spawned.py); the tests read their recorded output.
"""
import os

import numpy as np
import pytest

import spawned
from spawned import DROPIN_CASES, IO_CASES, PAIR_DROPIN
from tsmi import workloads as W

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _need(key: str, what: str):
    if not spawned.has(key):
        pytest.skip(f"{what} not run at session start (not built, or no GPU session)")


def test_error_paths_without_gpu():
    _need("errors", "test_solver")
    r = spawned.rows("errors")[0]
    assert r["caught"] == r["expected"]


@pytest.mark.gpu
@pytest.mark.parametrize("kernel", ["hip", "hip-fp32"])
def test_reference_scenarios(kernel):
    key = f"scenarios-{kernel}"
    _need(key, "test_solver")
    rows = {r["scenario"]: r for r in spawned.rows(key) if "E" in r}
    assert len(rows) == 5
    tol, ntol = 1e-3, 1e-5
    for name, r in rows.items():
        assert abs(r["N0"] - r["N"]) < ntol, name
        if name in ("free_particle", "harmonic_oscillator", "intra_particle_interaction"):
            assert abs(r["E0"] - r["E"]) < tol, name
    assert abs(1.00001 - rows["imaginary_harmonic_oscillator"]["E"]) < tol
    assert abs(1.02321 - rows["intra_particle_interaction"]["XX"]) < tol
    assert abs(1.59273 - rows["imaginary_intra_particle_interaction"]["E"]) < tol
    assert abs(0.768148 - rows["imaginary_intra_particle_interaction"]["XX"]) < tol
    if kernel == "hip":
        # the reference's values for the same scenarios.  The reference sums its
        # 62500 cells in plain sequence (error up to ~N eps), this library with
        # compensated FP64 sums: norms and energies agree to ~1e-11.
        for name, r in rows.items():
            z = np.load(os.path.join(GOLDEN, f"scenario_{name}.npz"))
            np.testing.assert_allclose(r["E"], z["energies"][0], rtol=1e-11)
            np.testing.assert_allclose(r["N"], z["energies"][5], rtol=1e-11)
            np.testing.assert_allclose(r["XX"], z["expect"][2], rtol=1e-11)
            if r["E0"] != 0.0:
                np.testing.assert_allclose(r["E0"], z["energies0"][0], rtol=1e-11)


@pytest.mark.gpu
@pytest.mark.parametrize("kernel", ["hip", "hip-fp32"])
def test_reference_pair_scenarios(kernel):
    """mixed_BEC_test and imaginary_mixed_BEC_test (test/kerneltest.cpp:178-237)
    through the two-state Solver and HipPairKernel: the total norm is kept, the
    Rabi coupling moves the whole norm to state 2 in 5000 real-time steps, and the
    imaginary-time run ends at the reference's norms 0.915292 / 0.084708
    (NORM_TOLERANCE 1e-5)."""
    rows = {}
    for name in ("mixed_BEC", "imaginary_mixed_BEC"):
        key = f"pair-scenario-{kernel}-{name}"
        _need(key, "test_solver")
        rows.update({r["scenario"]: r for r in spawned.rows(key)})
    # FP64: the reference's NORM_TOLERANCE.  FP32 bulk: the real-time run is 5000
    # steps; each step's FP32 roundings (the E table's above all, the same at a
    # site every step) move the norm by ~3e-9, so it is held to 1e-4, the
    # FP32-vs-FP64 bar of this library.
    ntol = 1e-5
    ntol_real = 1e-5 if kernel == "hip" else 1e-4
    r = rows["mixed_BEC"]
    assert abs(r["N0"] - r["N"]) < ntol_real
    assert abs(r["N1_0"] - r["N2"]) < ntol_real
    assert abs(r["N2_0"] - r["N1"]) < ntol_real
    r = rows["imaginary_mixed_BEC"]
    assert abs(r["N0"] - r["N"]) < ntol
    assert abs(0.915292 - r["N1"]) < ntol
    assert abs(0.084708 - r["N2"]) < ntol
    if kernel == "hip":
        z = np.load(os.path.join(GOLDEN, "pair_mixed_bec_imag.npz"))
        np.testing.assert_allclose([r["N1"], r["N2"]], z["getters"][12:14], rtol=1e-11)


@pytest.mark.gpu
@pytest.mark.parametrize("case", PAIR_DROPIN, ids=lambda c: "_".join(f"{k}{v}" for k, v in c.items() if k != "calls"))
def test_pair_dropin_inside_reference(case):
    """HipPairKernel built against the reference's header and library, driven
    with the two-component protocol of the reference Solver, against the
    reference's two-state "cpu" Solver."""
    key = spawned.pair_dropin_key(case)
    _need(key, "oracle/_ref/dropin_check")
    r = spawned.rows(key)[0]
    assert r["kernel"] == "HIP"
    assert r["rel_l2_a"] < 1e-12 and r["rel_l2_b"] < 1e-12


def _two_step_lattice(case) -> bool:
    """Real time with both axes closed or both periodic: the library runs two
    steps per pass there."""
    return not case.get("imag_time", False) and case["periodic_x"] == case["periodic_y"]


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("case", DROPIN_CASES, ids=spawned.case_id)
def test_dropin_inside_reference(case, precision):
    """HipKernel inside the reference library, driven call by call as the
    reference Solver::evolve drives its kernel (one run_kernel per step, the halo
    exchange skipped after the last step of each evolve call, update_potential
    where the potential changes), against the reference Solver's "cpu" run.
    HipKernel queues the per-step calls, so on the lattices where two steps per
    pass are exact the run must reach the two-step pass."""
    key = spawned.dropin_key(case, precision)
    _need(key, "oracle/_ref/dropin_check")
    r = spawned.rows(key)[0]
    assert r["kernel"] == ("HIP" if precision == 64 else "HIP-FP32")
    if precision == 64:
        assert r["rel_l2"] < 1e-12
        assert abs(r["kernel_norm"] - r["solver_norm"]) < 1e-12
    else:
        # FP32 bulk against the reference's FP64: the FP32 bar of this library
        assert r["rel_l2"] < 2e-5
        assert abs(r["kernel_norm"] - r["solver_norm"]) < 1e-5
    steps = sum(case.get("calls", (case["steps"],)))
    if _two_step_lattice(case) and steps >= 4:
        assert r["fused_launches"] > 0
        # every step is accounted for: two per fused pass, one per single pass
        assert 2 * r["fused_launches"] + r["single_launches"] == steps
    else:
        assert r["fused_launches"] == 0 and r["single_launches"] == steps


@pytest.mark.parametrize("case", IO_CASES, ids=lambda c: f"{c['n_x']}x{c['n_y']}_p{int(c['periodic_x'])}{int(c['periodic_y'])}")
def test_text_io_matches_reference(case):
    """State::write_to_file / write_particle_density / write_phase, State::loadtxt
    (with the halo refill of periodic axes) and Potential(grid, file), against the
    reference library driven on the same tile: identical text, identical loaded
    tiles.  No device work."""
    key = spawned.io_key(case)
    _need(key + "-ours", "test_solver and the reference build (oracle/_ref)")
    spawned.rows(key + "-ours")
    spawned.rows(key + "-ref")
    wl = W.small_case(steps=0, g=1.0, **case)
    lat = wl.lattice
    ours, ref = spawned.path(key, "ours"), spawned.path(key, "ref")
    for suffix in ("-state", "-density", "-phase", "-potential"):
        with open(ours + suffix) as a, open(ref + suffix) as b:
            assert a.read() == b.read(), suffix
    a = np.fromfile(ours + "-loaded.bin")
    b = np.fromfile(ref + "-loaded.bin")
    assert np.array_equal(a, b)
    n = lat.W * lat.H
    loaded = (a[:n] + 1j * a[n:2 * n]).reshape(lat.H, lat.W)
    inner = (lat.inner_y, lat.inner_x)
    np.testing.assert_allclose(loaded[inner], wl.psi[inner], rtol=1e-5, atol=1e-7)
    assert np.array_equal(a[2 * n:].reshape(lat.H, lat.W), wl.v)
