"""The two-component step (tsmi_pair_*, pair_kernel) against the two-component
oracle: FP64 within 1e-12 relative L2 on both whole tiles, FP32 within 2e-5;
energies; the reference test suite's mixed-BEC known answers
(test/kerneltest.cpp:178-237); the closed-form Rabi transfer."""
import numpy as np
import pytest

from tsmi import workloads as W

pytestmark = pytest.mark.gpu

COUPLED = dict(g_a=3.0, g_b=2.0, g_ab=1.5)
CASES = [
    dict(n_x=40, n_y=36, periodic_x=False, periodic_y=False, omega_r=2.0, omega_i=0.7, mass_b=1.7, calls=[7, 5]),
    dict(n_x=40, n_y=36, periodic_x=True, periodic_y=True, omega_r=2.0, omega_i=-0.5, calls=[6]),
    dict(n_x=40, n_y=36, periodic_x=True, periodic_y=False, omega_r=1.0, calls=[4, 3]),
    dict(n_x=36, n_y=40, periodic_x=False, periodic_y=True, calls=[5]),
    dict(n_x=300, n_y=90, periodic_x=True, periodic_y=True, omega_r=1.5, mass_b=0.6, calls=[9]),
    dict(n_x=260, n_y=136, periodic_x=False, periodic_y=False, omega_r=1.5, calls=[5, 1]),
    dict(n_x=37, n_y=1, periodic_x=False, periodic_y=False, omega_r=1.0, calls=[6]),
    dict(n_x=40, n_y=36, periodic_x=False, periodic_y=False, imag_time=True, omega_r=2.0, omega_i=0.7,
         mass_b=1.7, calls=[7, 5]),
    dict(n_x=40, n_y=36, periodic_x=True, periodic_y=True, imag_time=True, calls=[6, 2]),
    dict(n_x=130, n_y=70, periodic_x=True, periodic_y=False, imag_time=True, omega_r=1.0, calls=[8]),
]


def case_id(c):
    return (f"{c['n_x']}x{c['n_y']}_p{int(c['periodic_x'])}{int(c['periodic_y'])}_"
            f"{'imag' if c.get('imag_time') else 'real'}_calls{'_'.join(map(str, c['calls']))}")


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def make(case):
    return W.pair_case(**case, **COUPLED)


def run_gpu(oracle, pw, precision, initial_half=True, tile_rows=0, sweeps=0):
    from tsmi.pair import HipPair
    e = oracle.pair_exp_potentials(pw)
    with HipPair(pw, precision=precision, e_planes=e, tile_rows=tile_rows, sweeps=sweeps) as h:
        for n in pw.calls:
            h.evolve(n, initial_half=initial_half)
        return h.state(0), h.state(1), h.energy()


@pytest.mark.parametrize("case", CASES, ids=[case_id(c) for c in CASES])
def test_pair_fp64_matches_oracle(oracle, case):
    pw = make(case)
    print(f"[pair] fp64 {case_id(case)}", flush=True)
    a, b, e = run_gpu(oracle, pw, 64)
    ra, rb = oracle.evolve2(pw)
    assert rel_l2(a, ra) < 1e-12
    assert rel_l2(b, rb) < 1e-12
    np.testing.assert_allclose(e, oracle.energy2(pw, ra, rb), rtol=1e-10, atol=1e-13)


@pytest.mark.parametrize("case", CASES[:2] + CASES[7:9], ids=[case_id(c) for c in CASES[:2] + CASES[7:9]])
def test_pair_fp32_matches_oracle(oracle, case):
    pw = make(case)
    print(f"[pair] fp32 {case_id(case)}", flush=True)
    a, b, _ = run_gpu(oracle, pw, 32)
    ra, rb = oracle.evolve2(pw)
    assert rel_l2(a, ra) < 2e-5
    assert rel_l2(b, rb) < 2e-5


def test_pair_without_initial_half_rotation(oracle):
    """The Python API's later calls skip the opening half rotation
    (src/solver.cpp:174-175)."""
    pw = make(CASES[0])
    a, b, _ = run_gpu(oracle, pw, 64, initial_half=False)
    ra, rb = oracle.evolve2(pw, initial_half=False)
    assert rel_l2(a, ra) < 1e-12 and rel_l2(b, rb) < 1e-12


def test_pair_segments_do_not_change_the_result(oracle):
    pw = make(CASES[4])
    a1, b1, _ = run_gpu(oracle, pw, 64, tile_rows=8)
    a2, b2, _ = run_gpu(oracle, pw, 64, tile_rows=0)
    assert rel_l2(a1, a2) < 1e-14 and rel_l2(b1, b2) < 1e-14


FUSED = [CASES[0], CASES[1], CASES[4], CASES[5], CASES[6]]


@pytest.mark.parametrize("tile_rows", [0, 16])
@pytest.mark.parametrize("case", FUSED, ids=[case_id(c) for c in FUSED])
def test_pair_two_steps_per_pass_fp32(oracle, case, tile_rows):
    """FP32 real time on closed or periodic lattices, two steps per pass
    (pair2_kernel; on periodic axes the first and last step of each call run
    alone): the same state as one step per pass, and the oracle's."""
    pw = make(case)
    print(f"[pair] fp32 two steps per pass {case_id(case)} tile_rows={tile_rows}", flush=True)
    a1, b1, _ = run_gpu(oracle, pw, 32, tile_rows=tile_rows, sweeps=1)
    a2, b2, _ = run_gpu(oracle, pw, 32, tile_rows=tile_rows, sweeps=2)
    ra, rb = oracle.evolve2(pw)
    assert rel_l2(a2, a1) < 1e-6 and rel_l2(b2, b1) < 1e-6
    assert rel_l2(a2, ra) < 2e-5 and rel_l2(b2, rb) < 2e-5


def test_pair_single_steps_on_two_step_layout(oracle):
    """tsmi_pair_rabi / tsmi_pair_step driven one step at a time (the reference
    solver's protocol) on a pair created for two steps per pass -- periodic
    axes, so the 8-wide halo layout and its full-wrap exchange -- give the state
    of one-step-per-pass evolve."""
    from tsmi import _lib
    from tsmi.pair import HipPair
    pw = make(CASES[4])
    n = pw.calls[0]
    e = oracle.pair_exp_potentials(pw)
    with HipPair(pw, precision=32, e_planes=e, sweeps=1) as h:
        h.evolve(n)
        a1, b1 = h.state(0), h.state(1)
    with HipPair(pw, precision=32, e_planes=e, sweeps=2) as h:
        _lib.check(h._L.tsmi_pair_rabi(h._h, 0.5, 0), "tsmi_pair_rabi")
        for i in range(n):
            last = i == n - 1
            _lib.check(h._L.tsmi_pair_step(h._h, 0 if last else 1, 0.5 if last else 1.0), "tsmi_pair_step")
        a2, b2 = h.state(0), h.state(1)
    assert rel_l2(a2, a1) < 1e-6 and rel_l2(b2, b1) < 1e-6


@pytest.mark.parametrize("precision,case", [(64, CASES[1]), (32, CASES[2]), (32, CASES[8])],
                         ids=["fp64", "fp32_mixed_axes", "fp32_imag"])
def test_pair_two_steps_per_pass_refused(precision, case):
    """sweeps 2 where two steps per pass are not exact or not built is an error."""
    from tsmi import _lib
    from tsmi.pair import HipPair
    with pytest.raises(_lib.TsmiError, match="two steps per pass"):
        HipPair(make(case), precision=precision, sweeps=2)


def test_mixed_bec_imaginary_known_answer(oracle):
    """imaginary_mixed_BEC_test (test/kerneltest.cpp:208-237) on the GPU in FP64:
    1000 imaginary-time steps, norms 0.915292 and 0.084708 within 1e-5."""
    from tsmi.pair import HipPair
    pw = W.mixed_bec(True)
    with HipPair(pw, precision=64, e_planes=oracle.pair_exp_potentials(pw)) as h:
        h.evolve(pw.steps)
        e = h.energy()
        a, b = h.state(0), h.state(1)
    assert abs(e[9] - 0.915292) < 1e-5
    assert abs(e[10] - 0.084708) < 1e-5
    assert abs(e[11] - 1.0) < 1e-5
    import os
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "pair_mixed_bec_imag.npz"))
    s = int(z["stride"])
    assert rel_l2(a[::s, ::s], z["a_re"] + 1j * z["a_im"]) < 1e-11
    assert rel_l2(b[::s, ::s], z["b_re"] + 1j * z["b_im"]) < 1e-11


@pytest.mark.parametrize("precision,tol", [(64, 1e-11), (32, 2e-5)])
def test_rabi_transfer_closed_form(precision, tol):
    """Identical non-interacting components: populations N0 cos^2(|omega| t / 2)
    and N0 sin^2(|omega| t / 2), over two calls."""
    from tsmi.pair import HipPair
    pw = W.pair_case(64, 64, True, True, omega_r=1.2, omega_i=-0.9, calls=[150, 101], dt=5e-3)
    pw.v_b = pw.v_a.copy()
    pw.psi_b = np.zeros_like(pw.psi_a)
    with HipPair(pw, precision=precision) as h:
        for n in pw.calls:
            h.evolve(n)
        e = h.energy()
    n0 = W.squared_norm(pw.lattice, pw.psi_a)
    th = np.hypot(1.2, -0.9) / 2 * pw.dt * pw.steps
    assert e[9] == pytest.approx(n0 * np.cos(th) ** 2, rel=tol, abs=tol)
    assert e[10] == pytest.approx(n0 * np.sin(th) ** 2, rel=tol, abs=tol)


@pytest.mark.parametrize("sweeps", [1, 2])
def test_pair_fp32_norm_drift_with_rabi(sweeps):
    """FP32 total norm over 400 steps of two overlapping interacting clouds
    with Rabi coupling (512^2 periodic, g_a = g_b = 50, g_ab = 40, omega = 1,
    dt = 5e-4), one and two steps per pass: within 1e-7 of the initial norm,
    relative.  The rotation's cosine 1 - d (d = 3e-8) differs from 1 by less
    than half an FP32 ulp of the state: -d x applied in a rounding of its own
    is lost at most sites while the sine terms still add |s|^2, a drift of 2d
    per step (2.4e-5 over these steps)."""
    from tsmi.lattice import Lattice2D
    from tsmi.pair import HipPair
    lat = Lattice2D(512, 8.0, periodic_x=True, periodic_y=True)
    X, Y = lat.mesh()
    v_a = 0.5 * 0.64 * (X * X + Y * Y)
    v_b = 0.5 * 0.768 * (X * X + Y * Y)
    amp = np.sqrt(np.maximum(10.0 - v_a, 0.0) / 50.0)
    psi_a = (amp * np.exp(1j * 2.4 * X)).astype(np.complex128)
    psi_b = (0.5 * amp * np.exp(-1j * 1.6 * Y)).astype(np.complex128)
    pw = W.PairWorkload("pair_drift_512", lat, psi_a, psi_b, v_a, v_b, 5e-4, (400,), g_a=50.0, g_b=50.0,
                        g_ab=40.0, omega_r=1.0)
    n0 = W.squared_norm(lat, psi_a) + W.squared_norm(lat, psi_b)
    with HipPair(pw, precision=32, sweeps=sweeps) as h:
        h.evolve(400)
        e = h.energy()
    drift = e[11] / n0 - 1
    assert abs(drift) < 1e-7, drift
