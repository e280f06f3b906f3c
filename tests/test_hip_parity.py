"""Parity of the HIP step (libtsmi_hip.so, through its C ABI) with the oracle.

FP64: the whole tile, halo cells included, within 1e-12 relative L2 of the
oracle after the stated number of steps, energies within 1e-12 and expected
values within 1e-11.  FP32 bulk: within the stated FP32 tolerance.  Both sides
get the same inputs: the same tile, the same E table (the oracle's, built with the
C library's complex exponential) and the same imaginary-time norm target.
"""
import json
import os

import numpy as np
import pytest

from tsmi import workloads as W

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def run_hip(wl, oracle, precision=64, steps=None, tile_rows=0, calls=None, sweeps=0):
    from tsmi.hip import kernel_for
    d = oracle.make_desc(wl)
    e = oracle.exp_potential(wl.v, wl.dt, wl.imag_time)
    k = kernel_for(wl, precision=precision, norm_target=d.norm_target if wl.imag_time else 0.0,
                   tile_rows=tile_rows, e_planes=e, sweeps=sweeps)
    try:
        for n in (calls or [wl.steps if steps is None else steps]):
            k.evolve(n)
        psi = k.sample()
        return psi, k.energy(), k.expected_values(), k.squared_norm()
    finally:
        k.close()


CASES = [
    dict(n_x=64, n_y=64, periodic_x=False, periodic_y=False, steps=1),
    dict(n_x=64, n_y=64, periodic_x=False, periodic_y=False, steps=100, g=5.0, g_lhy=1.0),
    dict(n_x=64, n_y=48, periodic_x=True, periodic_y=False, steps=30, g=3.0),
    dict(n_x=48, n_y=64, periodic_x=False, periodic_y=True, steps=30, g=3.0),
    dict(n_x=64, n_y=64, periodic_x=True, periodic_y=True, steps=50, g=2.0),
    dict(n_x=64, n_y=64, periodic_x=False, periodic_y=False, imag_time=True, steps=100, g=10.0, g_lhy=2.0),
    dict(n_x=64, n_y=64, periodic_x=True, periodic_y=True, imag_time=True, steps=60, g=10.0),
    dict(n_x=601, n_y=75, periodic_x=False, periodic_y=False, steps=12, g=1.0),      # several strips, ragged
    dict(n_x=500, n_y=90, periodic_x=True, periodic_y=True, steps=12, g=1.0),        # strips across the seam
    dict(n_x=248, n_y=40, periodic_x=False, periodic_y=False, steps=5),              # reference column-gap width
    dict(n_x=13, n_y=7, periodic_x=False, periodic_y=False, steps=9, g=0.5),         # tiny, odd
    dict(n_x=8, n_y=8, periodic_x=True, periodic_y=True, steps=9, g=0.5),            # smallest periodic
    dict(n_x=37, n_y=1, periodic_x=False, periodic_y=False, steps=9),                # single row
    dict(n_x=40, n_y=33, periodic_x=True, periodic_y=False, imag_time=True, steps=25, g=4.0),
    # interior strips with segments long enough for the steady-state loop, where each
    # row is renormalised as it enters the pipeline
    dict(n_x=520, n_y=200, periodic_x=False, periodic_y=False, imag_time=True, steps=20, g=10.0),
    dict(n_x=500, n_y=96, periodic_x=True, periodic_y=True, imag_time=True, steps=15, g=5.0, g_lhy=1.0),
]


def case_id(c):
    return "_".join(f"{k}{int(v) if isinstance(v, bool) else v}" for k, v in c.items())


@pytest.mark.parametrize("case", CASES, ids=[case_id(c) for c in CASES])
def test_fp64_matches_oracle(case, oracle):
    wl = W.small_case(**case)
    psi, e, x, n2 = run_hip(wl, oracle)
    ref = oracle.evolve(wl)
    assert rel_l2(psi, ref) < 1e-12
    np.testing.assert_allclose(e, oracle.energy(wl, ref), rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(x, oracle.expected_values(wl, ref), rtol=1e-11, atol=1e-13)
    d = oracle.make_desc(wl)
    re, im = np.ascontiguousarray(ref.real), np.ascontiguousarray(ref.imag)
    assert abs(n2 - oracle.lib().ts_oracle_squared_norm(d, oracle._p(re), oracle._p(im))) < 1e-13


@pytest.mark.parametrize("tile_rows", [2, 8, 30])
def test_segment_heights(tile_rows, oracle):
    """Many row segments per strip: the 4-row aprons between segments."""
    wl = W.small_case(300, 70, True, False, steps=7, g=2.0)
    psi, *_ = run_hip(wl, oracle, tile_rows=tile_rows)
    assert rel_l2(psi, oracle.evolve(wl)) < 1e-12


def test_split_evolve_calls(oracle):
    """evolve(7) then evolve(5): the second call starts from the halo cells the
    first left unrefreshed, as the reference kernel does across calls."""
    wl = W.small_case(64, 64, True, True, steps=12, g=2.0)
    psi, *_ = run_hip(wl, oracle, calls=[7, 5])
    ref = oracle.evolve(wl, steps=7)
    ref = oracle.evolve(wl, steps=5, psi=ref)
    assert rel_l2(psi, ref) < 1e-12


def test_zero_steps_is_identity(oracle):
    wl = W.small_case(40, 40, True, True, steps=0)
    psi, *_ = run_hip(wl, oracle, steps=0)
    assert np.array_equal(psi, wl.psi)


@pytest.mark.parametrize("name", ["config0", "small64_periodic_50", "small64_periodic_imag_60",
                                  "blocks260x136_closed_20", "small64x48_periodic_x_30"])
def test_fp64_matches_reference_golden(name, oracle):
    """Directly against the reference's own outputs (tests/golden)."""
    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    spec = json.loads(str(z["spec"]))
    wl = getattr(W, spec["builder"])(**spec["kwargs"])
    psi, e, x, _ = run_hip(wl, oracle, steps=int(z["steps"]))
    ref = z["psi_re"] + 1j * z["psi_im"]
    assert rel_l2(psi, ref) < 1e-12
    # the reference sums its cells in plain sequence (relative error up to
    # ~N eps = 1.5e-11 for 65536 cells); this library sums with compensation
    np.testing.assert_allclose(e, z["energies"], rtol=1e-11, atol=1e-14)
    np.testing.assert_allclose(x, z["expect"], rtol=1e-11, atol=1e-13)


@pytest.mark.parametrize("case", [CASES[1], CASES[4], CASES[5], CASES[8], CASES[14], CASES[15]],
                         ids=lambda c: case_id(c))
def test_fp32_bulk_within_tolerance(case, oracle):
    """FP32 bulk with FP64 coefficients rounded once and FP64 reductions:
    relative L2 below 2e-5 after up to 100 steps on these cases."""
    wl = W.small_case(**case)
    psi, e, x, n2 = run_hip(wl, oracle, precision=32)
    ref = oracle.evolve(wl)
    assert rel_l2(psi, ref) < 2e-5
    np.testing.assert_allclose(e, oracle.energy(wl, ref), rtol=1e-5, atol=1e-7)
    if not wl.imag_time:
        assert abs(n2 - 1.0) < 1e-5


def test_real_time_norm_conserved_fp64(oracle):
    wl = W.config0(n=128, steps=200)
    psi, e, x, n2 = run_hip(wl, oracle)
    assert abs(n2 - 1.0) < 1e-12


def test_fp32_high_beta_accuracy_1000_steps(oracle):
    """FP32 bulk against FP64 on the same GPU after 1000 steps at the kinetic
    angle beta = dt/(4 dx^2) = 2.048 of the 8192^2 benchmark lattice (512^2,
    side 4): relative L2 below 1e-4 and norm drift below 1e-6."""
    import torch
    from tsmi.hip import HipKernel, exp_potential
    wl = W.small_case(512, 512, True, True, steps=1000, dt=5e-4, g=50.0, length=4.0)
    outs = {}
    for prec in (32, 64):
        k = HipKernel(wl.lattice, delta_t=wl.dt, coupling=wl.g, precision=prec)
        k.set_exp_potential(*exp_potential(wl.v, wl.dt, False))
        k.set_state(wl.psi)
        k.evolve(wl.steps)
        outs[prec] = (k.sample(), k.squared_norm())
        k.close()
    lat = wl.lattice
    a, b = outs[32][0][lat.inner_y, lat.inner_x], outs[64][0][lat.inner_y, lat.inner_x]
    assert rel_l2(a, b) < 1e-4
    assert abs(outs[32][1] - outs[64][1]) < 1e-6, (rel_l2(a, b), outs[32][1] - outs[64][1])


@pytest.mark.parametrize("sweeps", [1, 2])
@pytest.mark.parametrize("form,beta", [("B", 2.048), ("A", 0.512)])
def test_fp32_norm_drift_uniform_phase(form, beta, sweeps):
    """FP32 norm drift over 1000 steps, 512^2 periodic, V = 0 and g = 0, at a
    kinetic angle of each scaled form (B: beta = 2.048, A: beta = 0.512), one
    and two steps per pass.  The phase applied at the potential stage is then
    the same at every site, so a rounding of the magnitude correction that
    repeats at every site adds up over the lattice and the steps (such a
    rounding drifted 8.0e-6 in form B and -2.2e-5 in form A).  Bar: 1e-7."""
    from tsmi.lattice import Lattice2D
    from tsmi.hip import kernel_for
    n, side = 512, 8.0
    lat = Lattice2D(n, side, periodic_x=True, periodic_y=True)
    X, Y = lat.mesh()
    dt = beta * 4 * (side / n) ** 2
    psi = W.normalise(lat, np.exp(-(X * X + Y * Y) / 4.0) * np.exp(1j * 0.7 * X))
    wl = W.Workload("drift", lat, psi, np.zeros_like(X), dt, 1000)
    k = kernel_for(wl, precision=32, sweeps=sweeps)
    try:
        n0 = k.squared_norm()
        k.evolve(wl.steps)
        drift = k.squared_norm() / n0 - 1
    finally:
        k.close()
    assert abs(drift) < 1e-7, drift


FUSED = [
    dict(n_x=64, n_y=64, periodic_x=True, periodic_y=True, g=2.0),
    dict(n_x=500, n_y=90, periodic_x=True, periodic_y=True, g=1.0, g_lhy=0.5),   # strips across the seam
    dict(n_x=601, n_y=75, periodic_x=False, periodic_y=False, g=1.0),            # ragged, closed
    dict(n_x=8, n_y=8, periodic_x=True, periodic_y=True),
    dict(n_x=13, n_y=7, periodic_x=False, periodic_y=False, g=0.5),
    dict(n_x=45, n_y=37, periodic_x=True, periodic_y=True, g=1.0),               # odd: the per-cell exchange
]


@pytest.mark.parametrize("steps", [1, 2, 3, 4, 5, 8, 13])
@pytest.mark.parametrize("case", FUSED, ids=lambda c: case_id(c))
def test_two_steps_per_pass_match_oracle(case, steps, oracle):
    """Fused passes (two steps per HBM pass) against the oracle, and single
    steps against the same oracle, for every step-count parity."""
    wl = W.small_case(steps=steps, **case)
    ref = oracle.evolve(wl)
    for sweeps in (2, 1):
        psi, *_ = run_hip(wl, oracle, sweeps=sweeps)
        assert rel_l2(psi, ref) < 1e-12, sweeps


def test_two_steps_per_pass_split_calls(oracle):
    """evolve(6) then evolve(7) with fused passes: each call starts from the halo
    the previous one left, as the reference does."""
    wl = W.small_case(96, 80, True, True, steps=13, g=2.0)
    psi, *_ = run_hip(wl, oracle, calls=[6, 7], sweeps=2, tile_rows=16)
    ref = oracle.evolve(wl, steps=6)
    ref = oracle.evolve(wl, steps=7, psi=ref)
    assert rel_l2(psi, ref) < 1e-12


def test_two_steps_per_pass_fp32(oracle):
    wl = W.small_case(256, 256, True, True, steps=40, g=3.0)
    ref = oracle.evolve(wl)
    a, *_ = run_hip(wl, oracle, precision=32, sweeps=2)
    b, *_ = run_hip(wl, oracle, precision=32, sweeps=1)
    assert rel_l2(a, ref) < 2e-5 and rel_l2(b, ref) < 2e-5
    assert rel_l2(a, b) < 1e-5


@pytest.mark.parametrize("precision", [64, 32])
def test_two_steps_per_pass_short_segments_periodic(precision, oracle):
    """Doubly periodic, two steps per pass with 12-row segments: interior
    segments (the FP32 stream) then have output rows within 8 rows of the edge,
    and the 8-wide halo they read is refreshed between passes.  Against the
    oracle (FP64 1e-12, FP32 2e-5) and single steps (FP32 1e-6)."""
    wl = W.small_case(480, 120, True, True, steps=9, g=2.0)
    ref = oracle.evolve(wl)
    a, *_ = run_hip(wl, oracle, precision=precision, sweeps=2, tile_rows=12)
    b, *_ = run_hip(wl, oracle, precision=precision, sweeps=1)
    tol = 1e-12 if precision == 64 else 2e-5
    assert rel_l2(a, ref) < tol and rel_l2(b, ref) < tol
    if precision == 32:
        assert rel_l2(a, b) < 1e-6


def test_two_steps_per_pass_rejected_where_not_exact():
    from tsmi import _lib
    from tsmi.hip import kernel_for
    for wl in (W.small_case(32, 32, True, False, steps=2), W.small_case(32, 32, True, True, steps=2, imag_time=True)):
        with pytest.raises(_lib.TsmiError, match="two steps per pass"):
            kernel_for(wl, sweeps=2)


def run_hip_with_e(wl, e, precision=64, sweeps=0, norm_target=0.0):
    from tsmi.hip import kernel_for
    k = kernel_for(wl, precision=precision, norm_target=norm_target, e_planes=e, sweeps=sweeps)
    try:
        k.evolve(wl.steps)
        return k.sample()
    finally:
        k.close()


@pytest.mark.parametrize("sweeps", [1, 2])
@pytest.mark.parametrize("periodic", [False, True])
def test_non_unit_e_table(periodic, sweeps, oracle):
    """An E table that is not a pure phase (a complex potential: |E| < 1, varying
    across the tile) is applied as given: the steps then read both E planes
    instead of the one-plane phase form, which would keep the norm."""
    wl = W.small_case(96, 72, periodic, periodic, steps=9, g=2.0)
    er, ei = oracle.exp_potential(wl.v, wl.dt, False)
    damp = 0.999 - 0.01 * np.linspace(0.0, 1.0, er.shape[1])[None, :]
    er, ei = er * damp, ei * damp
    d = oracle.make_desc(wl)
    re = np.ascontiguousarray(wl.psi.real).copy()
    im = np.ascontiguousarray(wl.psi.imag).copy()
    import ctypes as C
    oracle.lib().ts_oracle_evolve(C.byref(d), wl.steps, oracle._p(re), oracle._p(im), oracle._p(er), oracle._p(ei))
    ref = re + 1j * im
    psi = run_hip_with_e(wl, (er, ei), sweeps=sweeps)
    assert rel_l2(psi, ref) < 1e-12
    # the same table read as a pure phase would keep the norm: it must drop
    lat = wl.lattice
    assert W.squared_norm(lat, psi) < W.squared_norm(lat, wl.psi) * (1 - 1e-4)


def test_imaginary_time_e_with_zero(oracle):
    """Imaginary time with an E table that has zeros (a hard wall) has no
    exponent form: the steps read E itself."""
    wl = W.small_case(64, 64, False, False, steps=7, imag_time=True, g=3.0)
    er, ei = oracle.exp_potential(wl.v, wl.dt, True)
    er = er.copy()
    er[20:24, 30:34] = 0.0
    d = oracle.make_desc(wl)
    re = np.ascontiguousarray(wl.psi.real).copy()
    im = np.ascontiguousarray(wl.psi.imag).copy()
    import ctypes as C
    oracle.lib().ts_oracle_evolve(C.byref(d), wl.steps, oracle._p(re), oracle._p(im), oracle._p(er), oracle._p(ei))
    psi = run_hip_with_e(wl, (er, ei), norm_target=d.norm_target)
    assert rel_l2(psi, re + 1j * im) < 1e-12
