"""The CPU oracle against the golden vectors written by the reference solver.

The fixtures under tests/golden/ hold the outputs of the reference library
(compiled from its own sources, single-threaded, no FMA contraction) for inputs
that tsmi.workloads rebuilds here.  On the host that wrote them the oracle must
reproduce the reference bit for bit; elsewhere (another libm code path) to 1e-13.
"""
import glob
import json
import os

import numpy as np
import pytest

from golden_util import host_fingerprint, input_digest
from tsmi import workloads as W

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
# single-component evolution fixtures (adjoint_*.npz hold finite differences:
# tests/test_adjoint_oracle.py; pair*.npz two-component systems: tests/test_oracle_pair.py)
FIXTURES = sorted(f for f in glob.glob(os.path.join(GOLDEN, "*.npz"))
                  if not os.path.basename(f).startswith(("adjoint_", "pair")))


def load(path):
    z = np.load(path, allow_pickle=False)
    spec = json.loads(str(z["spec"]))
    wl = getattr(W, spec["builder"])(**spec["kwargs"])
    return z, wl


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p)[:-4] for p in FIXTURES])
def test_oracle_matches_reference(path, oracle):
    z, wl = load(path)
    inputs_identical = input_digest(wl) == str(z["input_sha256"])
    psi = oracle.evolve(wl, steps=int(z["steps"]))
    ref = z["psi_re"] + 1j * z["psi_im"]
    if "stride" in z:
        s = int(z["stride"])
        psi = psi[::s, ::s]
    # the whole tile, halo included: the halo semantics are part of the result
    assert rel_l2(psi, ref) < 1e-13
    e = oracle.energy(wl, oracle.evolve(wl, steps=int(z["steps"])))
    np.testing.assert_allclose(e, z["energies"], rtol=1e-12, atol=1e-15)
    x = oracle.expected_values(wl, oracle.evolve(wl, steps=int(z["steps"])))
    np.testing.assert_allclose(x, z["expect"], rtol=1e-11, atol=1e-13)
    if inputs_identical and str(z["host"]) == host_fingerprint():
        assert np.array_equal(psi, ref), "oracle no longer bit-identical to the reference on its host"


@pytest.mark.parametrize("name", W.REFERENCE_SCENARIOS)
def test_reference_known_answers(name):
    """The reference test suite's own known answers (test/kerneltest.cpp:8-125,
    TOLERANCE 1e-3, NORM_TOLERANCE 1e-5) hold for the golden vectors."""
    z = np.load(os.path.join(GOLDEN, f"scenario_{name}.npz"))
    e0, e1 = z["energies0"], z["energies"]
    tol, ntol = 1e-3, 1e-5
    assert abs(e0[5] - e1[5]) < ntol
    if name in ("free_particle", "harmonic_oscillator", "intra_particle_interaction"):
        assert abs(e0[0] - e1[0]) < tol
    if name == "imaginary_harmonic_oscillator":
        assert abs(1.00001 - e1[0]) < tol
    if name == "intra_particle_interaction":
        assert abs(1.02321 - z["expect"][2]) < tol
    if name == "imaginary_intra_particle_interaction":
        assert abs(1.59273 - e1[0]) < tol
        assert abs(0.768148 - z["expect"][2]) < tol
