"""The oracle's discrete adjoint (oracle/ts_oracle.c: ts_oracle_adjoint) against
central finite differences of the oracle forward, which matches the reference
solver bit for bit (tests/test_oracle_golden.py).  CPU only."""
import numpy as np
import pytest

from tsmi import workloads as W
from tsmi.losses import infidelity


def case(n_x=32, n_y=24, steps=40, dt=4e-3, length=8.0):
    wl = W.small_case(n_x, n_y, False, False, steps=steps, dt=dt, length=length)
    lat = wl.lattice
    tx, ty = W.uniform(W.SEED_BASE + 77, 2, lo=-1.0, hi=1.0)
    target = W.normalise(lat, W.gaussian_packet(lat, tx, ty, 0.0, 0.0, width=1.5))
    return wl, target


def loss_of(oracle, wl, target, v=None, psi0=None):
    import dataclasses
    w2 = dataclasses.replace(wl, v=wl.v if v is None else v, psi=wl.psi if psi0 is None else psi0)
    return infidelity(wl.lattice, target, oracle.evolve(w2))[0]


def points(wl, k, seed):
    """k cells drawn with probability ~ |psi_0|^2 (the gradient is negligible
    where the state never goes)."""
    w = np.abs(wl.psi.ravel()) ** 2
    rng = np.random.default_rng(seed)
    return rng.choice(w.size, size=k, replace=False, p=w / w.sum())


@pytest.mark.parametrize("every", [7, 100])
def test_oracle_adjoint_matches_finite_differences(oracle, every):
    wl, target = case()
    _, g = infidelity(wl.lattice, target, oracle.evolve(wl))
    gv, _ = oracle.adjoint(wl, g, every=every)
    h = 1e-4
    for j in points(wl, 8, 1):
        vp = wl.v.copy(); vp.flat[j] += h
        vm = wl.v.copy(); vm.flat[j] -= h
        fd = (loss_of(oracle, wl, target, v=vp) - loss_of(oracle, wl, target, v=vm)) / (2 * h)
        assert abs(gv.flat[j] - fd) <= 1e-6 * np.abs(gv).max() + 1e-6 * abs(fd), (j, gv.flat[j], fd)


def test_oracle_adjoint_initial_state_gradient(oracle):
    wl, target = case(steps=20)
    _, g = infidelity(wl.lattice, target, oracle.evolve(wl))
    _, g0 = oracle.adjoint(wl, g, every=6)
    h = 1e-6
    for j in points(wl, 4, 2):
        for unit in (1.0, 1j):
            pp = wl.psi.copy(); pp.flat[j] += h * unit
            pm = wl.psi.copy(); pm.flat[j] -= h * unit
            fd = (loss_of(oracle, wl, target, psi0=pp) - loss_of(oracle, wl, target, psi0=pm)) / (2 * h)
            got = g0.flat[j].real if unit == 1.0 else g0.flat[j].imag
            assert abs(got - fd) <= 1e-6 * np.abs(g0).max(), (j, unit, got, fd)


def test_oracle_adjoint_is_checkpoint_independent(oracle):
    wl, target = case(steps=30)
    _, g = infidelity(wl.lattice, target, oracle.evolve(wl))
    a, a0 = oracle.adjoint(wl, g, every=1)
    b, b0 = oracle.adjoint(wl, g, every=30)
    c, c0 = oracle.adjoint(wl, g, every=11)
    assert np.array_equal(a, b) and np.array_equal(a, c)
    assert np.array_equal(a0, b0) and np.array_equal(a0, c0)


def test_oracle_adjoint_rejects_unsupported(oracle):
    wl = W.small_case(16, 16, True, False, steps=3)
    with pytest.raises(ValueError):
        oracle.adjoint(wl, np.zeros_like(wl.psi))
    wl = W.small_case(16, 16, False, False, steps=3, g=1.0)
    with pytest.raises(ValueError):
        oracle.adjoint(wl, np.zeros_like(wl.psi))


def test_config4_fixture_matches_builder():
    """The reference finite differences were taken on exactly the inputs the
    builder makes today."""
    import os
    from golden_util import input_digest
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "adjoint_config4.npz"))
    wl, _ = W.config4()
    assert str(z["input_sha256"]) == input_digest(wl)
    assert len(set(z["points"].tolist())) == 32 and float(z["h"]) == 1e-4
    assert np.all(np.isfinite(z["fd"])) and np.max(np.abs(z["fd"])) > 1e-6
