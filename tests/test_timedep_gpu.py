"""Time-dependent potential with the table rebuilt on the device (tsmi.timedep)
against the oracle running the reference's loop: step i of an evolve call uses
E(V(t_i)) for i > 0 and the initial table for i = 0 (src/solver.cpp:186-192),
the exchange after every step but the last."""
import numpy as np
import pytest

from tsmi import workloads as W

pytestmark = pytest.mark.gpu


def v_of(wl, t):
    return (1.0 + 0.3 * np.sin(5.0 * t)) * wl.v


def oracle_run(oracle, wl, steps, t0=0.0):
    d = oracle.make_desc(wl)
    re = np.ascontiguousarray(wl.psi.real).copy()
    im = np.ascontiguousarray(wl.psi.imag).copy()
    er, ei = oracle.exp_potential(v_of(wl, t0), wl.dt, False)
    t = t0
    for i in range(steps):
        if i > 0:
            er, ei = oracle.exp_potential(v_of(wl, t), wl.dt, False)
        oracle.lib().ts_oracle_full_step(d, oracle._p(re), oracle._p(im), oracle._p(er), oracle._p(ei))
        if i != steps - 1:
            oracle.lib().ts_oracle_halo_exchange(d, oracle._p(re), oracle._p(im))
        t += wl.dt
    return re + 1j * im


@pytest.mark.parametrize("periodic", [True, False])
def test_time_dependent_potential_matches_oracle(oracle, periodic):
    import torch
    from tsmi import timedep
    from tsmi.hip import kernel_for
    wl = W.small_case(64, 48, periodic, periodic, steps=15, g=1.0, dt=2e-3)
    e0 = oracle.exp_potential(v_of(wl, 0.0), wl.dt, False)
    k = kernel_for(wl, e_planes=e0)
    try:
        v0 = torch.tensor(wl.v, device="cuda")
        t = timedep.evolve(k, lambda s: (1.0 + 0.3 * np.sin(5.0 * s)) * v0, wl.steps)
        psi = k.sample()
        energy = k.energy()
    finally:
        k.close()
    assert abs(t - wl.steps * wl.dt) < 1e-15
    ref = oracle_run(oracle, wl, wl.steps)
    assert float(np.linalg.norm(psi - ref) / np.linalg.norm(ref)) < 1e-12
    # the energies use the last V(t) tabulated
    import dataclasses
    wl_last = dataclasses.replace(wl, v=v_of(wl, (wl.steps - 1) * wl.dt))
    np.testing.assert_allclose(energy, oracle.energy(wl_last, ref), rtol=1e-11, atol=1e-14)
