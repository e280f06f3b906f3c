"""Synthetic inputs: the seeded generator and the device-side builder."""
import numpy as np

from tsmi import workloads as W


def test_splitmix64_known_values():
    # SplitMix64 reference stream for seed 0 (first outputs of the published generator)
    z = W.splitmix64(0, 3)
    assert [int(v) for v in z] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_uniform_range_and_offset():
    u = W.uniform(12345, 1000)
    assert u.min() >= 0.0 and u.max() < 1.0
    assert np.array_equal(W.uniform(12345, 10, offset=5), u[5:15])


def test_torch_builder_matches_numpy():
    import torch
    from tsmi import workloads_torch as WT
    wl = W.config2(n=128)
    lat, psi, v, params = WT.config2(n=128, device="cpu")
    assert lat == wl.lattice
    np.testing.assert_allclose(v.numpy(), wl.v, rtol=0, atol=1e-15)
    rel = np.linalg.norm(psi.numpy() - wl.psi) / np.linalg.norm(wl.psi)
    assert rel < 1e-13
    er, ei = WT.exp_potential(v, params["dt"], False)
    e = np.exp(-1j * params["dt"] * wl.v)
    np.testing.assert_allclose(er.numpy() + 1j * ei.numpy(), e, rtol=0, atol=1e-15)


def test_config_shapes():
    wl = W.config0(n=32, steps=3)
    assert wl.psi.shape == (32, 32) and abs(W.squared_norm(wl.lattice, wl.psi) - 1.0) < 1e-14
    wl = W.config2(n=64)
    assert wl.psi.shape == (72, 72) and wl.lattice.periodic_x
