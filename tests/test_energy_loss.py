"""The energy functional as a torch loss (tsmi.losses.energy_torch) against
the oracle's energy (pinned to the reference's golden energies), on the CPU:
closed, periodic and one-row lattices, with interaction and LHY terms, plus its
gradient against central differences of the oracle's energy in the state."""
import numpy as np
import pytest

from tsmi import workloads as W
from tsmi.losses import energy_torch

torch = pytest.importorskip("torch")


@pytest.mark.parametrize("nx,ny,px,py,g,g_lhy", [(48, 40, False, False, 2.0, 0.5), (40, 36, True, True, 1.0, 0.0),
                                                 (36, 30, True, False, 0.0, 0.0), (37, 1, False, False, 3.0, 0.0)])
def test_energy_torch_matches_oracle(oracle, nx, ny, px, py, g, g_lhy):
    wl = W.small_case(nx, ny, px, py, g=g, g_lhy=g_lhy, steps=5)
    psi = oracle.evolve(wl)      # a state with structure, halo as the reference leaves it
    t = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=torch.float64)
    e = energy_torch(wl.lattice, t(psi.real), t(psi.imag), t(wl.v), wl.mass, g, g_lhy)
    np.testing.assert_allclose(float(e), oracle.energy(wl, psi)[0], rtol=1e-12)


def test_energy_torch_gradient_in_state(oracle):
    wl = W.small_case(24, 20, False, False, g=2.0, steps=3)
    psi = oracle.evolve(wl)
    re = torch.tensor(np.ascontiguousarray(psi.real), dtype=torch.float64, requires_grad=True)
    im = torch.tensor(np.ascontiguousarray(psi.imag), dtype=torch.float64, requires_grad=True)
    e = energy_torch(wl.lattice, re, im, torch.tensor(wl.v), wl.mass, wl.g)
    e.backward()
    h = 1e-6
    rng = np.random.default_rng(7)
    for _ in range(8):
        y, x = int(rng.integers(2, 18)), int(rng.integers(2, 22))
        for part, grad in ((1.0, re.grad), (1j, im.grad)):
            p, m = psi.copy(), psi.copy()
            p[y, x] += part * h
            m[y, x] -= part * h
            fd = (oracle.energy(wl, p)[0] - oracle.energy(wl, m)[0]) / (2 * h)
            assert abs(float(grad[y, x]) - fd) <= 1e-6 * max(1.0, abs(fd))


def test_energy_gradient_in_potential_through_oracle_adjoint(oracle):
    """The pipeline the autograd bridge runs on the device, in FP64 on the CPU:
    the seed dE/dpsi_N from torch, the oracle's reverse sweep for the
    evolution's share of dE/dV, plus the direct potential term -- against
    central differences of the oracle's evolution and energy (h = 1e-4)."""
    import dataclasses
    wl = W.small_case(40, 32, False, False, steps=25, dt=2e-3, length=12.0)
    psi = oracle.evolve(wl)
    re = torch.tensor(np.ascontiguousarray(psi.real), dtype=torch.float64, requires_grad=True)
    im = torch.tensor(np.ascontiguousarray(psi.imag), dtype=torch.float64, requires_grad=True)
    v = torch.tensor(wl.v, dtype=torch.float64, requires_grad=True)
    e = energy_torch(wl.lattice, re, im, v, wl.mass)
    e.backward()
    seed = re.grad.numpy() + 1j * im.grad.numpy()
    gv, _ = oracle.adjoint(wl, seed, every=5)
    grad = gv + v.grad.numpy()
    rng = np.random.default_rng(3)
    h = 1e-4
    worst, scale = 0.0, 0.0
    for _ in range(10):
        y, x = int(rng.integers(3, 29)), int(rng.integers(3, 37))
        vals = []
        for sgn in (1.0, -1.0):
            v2 = wl.v.copy()
            v2[y, x] += sgn * h
            w2 = dataclasses.replace(wl, v=v2)
            vals.append(oracle.energy(w2, oracle.evolve(w2))[0])
        fd = (vals[0] - vals[1]) / (2 * h)
        worst = max(worst, abs(grad[y, x] - fd))
        scale = max(scale, abs(fd))
    assert worst <= 1e-6 * scale
