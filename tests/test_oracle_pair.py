"""The two-component oracle (ts_oracle_evolve2 / _rabi / _energy2) pinned by the
reference solver's golden vectors, by the closed-form Rabi population transfer
and by the reference test suite's own known answers for two components
(test/kerneltest.cpp:178-237)."""
import glob
import json
import os

import numpy as np
import pytest

from golden_util import host_fingerprint, pair_digest
from tsmi import workloads as W

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
FIXTURES = sorted(glob.glob(os.path.join(GOLDEN, "pair*.npz")))


def load(path):
    z = np.load(path, allow_pickle=False)
    spec = json.loads(str(z["spec"]))
    return z, getattr(W, spec["builder"])(**spec["kwargs"])


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p)[:-4] for p in FIXTURES])
def test_pair_oracle_matches_reference(path, oracle):
    z, pw = load(path)
    a, b = oracle.evolve2(pw)
    e = oracle.energy2(pw, a, b)
    lat = pw.lattice
    norms = [W.squared_norm(lat, a), W.squared_norm(lat, b)]
    if "stride" in z:
        s = int(z["stride"])
        a, b = a[::s, ::s], b[::s, ::s]
    ra, rb = z["a_re"] + 1j * z["a_im"], z["b_re"] + 1j * z["b_im"]
    assert rel_l2(a, ra) < 1e-13
    assert rel_l2(b, rb) < 1e-13
    g = z["getters"]
    # the getters are NaN where the reference divides by an empty component's norm
    np.testing.assert_allclose(e, g[:12], rtol=1e-11, atol=1e-14, equal_nan=True)
    np.testing.assert_allclose(norms, g[12:14], rtol=1e-12)
    if pair_digest(pw) == str(z["input_sha256"]) and str(z["host"]) == host_fingerprint():
        assert np.array_equal(a, ra) and np.array_equal(b, rb), "oracle no longer bit-identical on its host"


def test_mixed_bec_imaginary_known_answer():
    """imaginary_mixed_BEC_test (test/kerneltest.cpp:208-237): after 1000
    imaginary-time steps the norms are 0.915292 and 0.084708 (NORM_TOLERANCE
    1e-5), and the total norm is kept."""
    z = np.load(os.path.join(GOLDEN, "pair_mixed_bec_imag.npz"))
    g = z["getters"]
    assert abs(g[12] - 0.915292) < 1e-5
    assert abs(g[13] - 0.084708) < 1e-5
    assert abs(g[11] - 1.0) < 1e-5


@pytest.mark.parametrize("omega", [(2.0, 0.0), (1.2, -0.9)])
def test_rabi_transfer_closed_form(oracle, omega):
    """Two identical, non-interacting components with the Rabi coupling: the
    kinetic and potential factors commute with the Rabi rotation, so starting
    from an empty component b the populations are N0 cos^2(|Omega| t) and
    N0 sin^2(|Omega| t), |Omega| = |omega|/2 (the kernel's half constants,
    src/cpukernel.cpp:330-331), t = n dt over any split into calls (the half
    rotations at the ends of consecutive calls join into full ones)."""
    pw = W.pair_case(32, 32, False, False, omega_r=omega[0], omega_i=omega[1], calls=[60, 41], dt=5e-3)
    pw.v_b = pw.v_a.copy()
    pw.psi_b = np.zeros_like(pw.psi_a)
    a, b = oracle.evolve2(pw)
    lat = pw.lattice
    n0 = W.squared_norm(lat, pw.psi_a)
    th = np.hypot(*omega) / 2 * pw.dt * pw.steps
    assert W.squared_norm(lat, a) == pytest.approx(n0 * np.cos(th) ** 2, rel=1e-12, abs=1e-14)
    assert W.squared_norm(lat, b) == pytest.approx(n0 * np.sin(th) ** 2, rel=1e-12, abs=1e-14)
    # the single-component oracle gives the shared spatial factor
    single = oracle.evolve(pw.component(0), steps=pw.steps)
    assert rel_l2(a, np.cos(th) * single) < 1e-12


def test_pair_without_coupling_is_two_single_systems(oracle):
    """g_ab = 0 and no Rabi coupling: each component evolves as the
    single-component system with its own mass, coupling and potential."""
    pw = W.pair_case(40, 36, True, False, g_a=3.0, g_b=2.0, mass_b=1.7, calls=[6, 5])
    a, b = oracle.evolve2(pw)
    for psi, which in ((a, 0), (b, 1)):
        wl = pw.component(which)
        ref = oracle.evolve(wl, steps=6)
        ref = oracle.evolve(wl, steps=5, psi=ref)
        assert rel_l2(psi, ref) < 1e-14


def test_inter_species_term_matters(oracle):
    """A wrong start-of-step density (the other component after its step, or
    no cross term) changes the result well beyond rounding."""
    pw = W.pair_case(40, 36, False, False, g_a=3.0, g_b=2.0, g_ab=4.0, calls=[8])
    a, _ = oracle.evolve2(pw)
    pw0 = W.pair_case(40, 36, False, False, g_a=3.0, g_b=2.0, g_ab=0.0, calls=[8])
    a0, _ = oracle.evolve2(pw0)
    assert rel_l2(a, a0) > 1e-6
