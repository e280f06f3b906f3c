"""The HIP reverse sweep (tsmi_evolve_checkpointed + tsmi_adjoint) against the
oracle's FP64 discrete adjoint, which is pinned by central finite differences
(tests/test_adjoint_oracle.py), and the torch.autograd bridge: closed axes with
and without the nonlinear and LHY terms, periodic axes (the gradient of the
periodic fields: the oracle's tile gradient folded onto the inner cells).

Bars: FP64 within 1e-10 of the oracle (max-norm, relative to the largest
gradient entry); FP32 within 1e-4 (the FP32 adjoint target)."""
import numpy as np
import pytest

from tsmi import workloads as W
from tsmi.losses import infidelity

pytestmark = pytest.mark.gpu


def case(n_x, n_y, steps, dt=2e-3, length=12.0, seed=W.SEED_BASE + 91, g=0.0, g_lhy=0.0, px=False, py=False):
    wl = W.small_case(n_x, n_y, px, py, steps=steps, dt=dt, length=length, seed=seed, g=g, g_lhy=g_lhy)
    lat = wl.lattice
    tx, ty = W.uniform(seed + 1, 2, lo=-1.5, hi=1.5)
    target = W.normalise(lat, W.gaussian_packet(lat, tx, ty, 0.0, 0.0, width=1.5))
    return wl, target


def rel_max(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def hip_adjoint(wl, target, oracle, precision, every, tile_rows=0, sweeps=0):
    from tsmi.hip import kernel_for
    e = oracle.exp_potential(wl.v, wl.dt, False)
    k = kernel_for(wl, precision=precision, tile_rows=tile_rows, e_planes=e, sweeps=sweeps)
    try:
        k.evolve_checkpointed(wl.steps, every)
        psi = k.sample()
        L, g = infidelity(wl.lattice, target, psi)
        gv, g0 = k.adjoint(g)
        # the context still holds psi_N
        assert np.array_equal(k.sample(), psi)
        return L, gv, g0
    finally:
        k.close()


CASES = [
    dict(n_x=96, n_y=72, steps=120, every=25),
    dict(n_x=601, n_y=75, steps=40, every=7),       # several strips, ragged edges
    dict(n_x=37, n_y=1, steps=30, every=30),        # single row: no vertical passes
    dict(n_x=64, n_y=64, steps=50, every=1),        # a checkpoint at every step
]


@pytest.mark.parametrize("c", CASES, ids=lambda c: f"{c['n_x']}x{c['n_y']}_s{c['steps']}_k{c['every']}")
def test_adjoint_fp64_matches_oracle(c, oracle):
    wl, target = case(c["n_x"], c["n_y"], c["steps"])
    _, g = infidelity(wl.lattice, target, oracle.evolve(wl))
    ref_v, ref_0 = oracle.adjoint(wl, g, every=10)
    L, gv, g0 = hip_adjoint(wl, target, oracle, 64, c["every"])
    assert rel_max(gv, ref_v) < 1e-10
    assert rel_max(g0, ref_0) < 1e-10


@pytest.mark.parametrize("c", CASES[:2], ids=lambda c: f"{c['n_x']}x{c['n_y']}")
def test_adjoint_fp32_within_tolerance(c, oracle):
    wl, target = case(c["n_x"], c["n_y"], c["steps"])
    _, g = infidelity(wl.lattice, target, oracle.evolve(wl))
    ref_v, ref_0 = oracle.adjoint(wl, g, every=10)
    _, gv, g0 = hip_adjoint(wl, target, oracle, 32, c["every"])
    assert rel_max(gv, ref_v) < 1e-4
    assert rel_max(g0, ref_0) < 1e-4


NL_CASES = [
    dict(n_x=96, n_y=72, steps=80, every=20, g=40.0, g_lhy=0.0),
    dict(n_x=601, n_y=75, steps=40, every=7, g=25.0, g_lhy=12.0),
]


@pytest.mark.parametrize("c", NL_CASES, ids=lambda c: f"{c['n_x']}x{c['n_y']}_g{c['g']:g}_lhy{c['g_lhy']:g}")
@pytest.mark.parametrize("precision", [64, 32])
def test_adjoint_nonlinear_matches_oracle(c, precision, oracle):
    """g and g_LHY: the reverse step undoes the density-dependent phase and
    pulls the adjoint through it (the 2 phi'(s) Im(conj(mu) y) x term)."""
    wl, target = case(c["n_x"], c["n_y"], c["steps"], g=c["g"], g_lhy=c["g_lhy"])
    _, g = infidelity(wl.lattice, target, oracle.evolve(wl))
    ref_v, ref_0 = oracle.adjoint(wl, g, every=10)
    _, gv, g0 = hip_adjoint(wl, target, oracle, precision, c["every"])
    bar = 1e-10 if precision == 64 else 1e-4
    assert rel_max(gv, ref_v) < bar
    assert rel_max(g0, ref_0) < bar


def periodic_case(oracle, px, py, g=30.0, steps=60, n_x=72, n_y=56):
    """A packet crossing the periodic edges; V and psi_0 with refreshed halos
    (periodic fields, as the reference's coordinate mapping makes them)."""
    import dataclasses
    wl, target = case(n_x, n_y, steps, dt=2e-3, length=8.0, g=g, px=px, py=py)
    wl = dataclasses.replace(wl, v=oracle.halo_exchange(wl, wl.v), psi=oracle.halo_exchange(wl, wl.psi))
    return wl, target


@pytest.mark.parametrize("px,py,sweeps", [(True, True, 0), (True, True, 1), (False, True, 0)],
                         ids=["xy_halo8", "xy_single", "y"])
@pytest.mark.parametrize("precision", [64, 32])
def test_adjoint_periodic_matches_folded_oracle(px, py, sweeps, precision, oracle):
    """Periodic axes: the library's gradient of the periodic fields against the
    oracle's exact tile gradient folded onto the inner cells (fold_halo, pinned
    against differences along periodic perturbations); zero on the halo."""
    wl, target = periodic_case(oracle, px, py)
    lat = wl.lattice
    _, g = infidelity(lat, target, oracle.evolve(wl))
    ref_v, ref_0 = oracle.adjoint(wl, g, every=10)
    ref_v, ref_0 = oracle.fold_halo(wl, ref_v), oracle.fold_halo(wl, ref_0)
    _, gv, g0 = hip_adjoint(wl, target, oracle, precision, 13, sweeps=sweeps)
    halo = np.ones((lat.H, lat.W), bool)
    halo[lat.inner_y, lat.inner_x] = False
    assert not gv[halo].any() and not g0[halo].any()
    bar = 1e-10 if precision == 64 else 1e-4
    assert rel_max(gv, ref_v) < bar
    assert rel_max(g0, ref_0) < bar


def test_adjoint_segment_heights_agree(oracle):
    wl, target = case(120, 90, 30)
    _, a, _ = hip_adjoint(wl, target, oracle, 64, 10, tile_rows=8)
    _, b, _ = hip_adjoint(wl, target, oracle, 64, 10, tile_rows=30)
    assert rel_max(a, b) < 1e-12


def test_adjoint_errors():
    from tsmi import _lib
    from tsmi.hip import kernel_for
    for wl in (W.small_case(32, 32, False, False, steps=3, imag_time=True),
               W.small_case(32, 32, True, False, steps=3)):      # x periodic, closed y: rows left unrefreshed
        k = kernel_for(wl)
        try:
            with pytest.raises(_lib.TsmiError, match="real time"):
                k.evolve_checkpointed(3, 1)
        finally:
            k.close()
    wl = W.small_case(32, 32, False, False, steps=3)
    k = kernel_for(wl)
    try:
        with pytest.raises(_lib.TsmiError, match="checkpointed"):
            k.adjoint(np.zeros_like(wl.psi))
        k.evolve_checkpointed(3, 2)
        k.evolve(1)                      # drops the checkpoints
        with pytest.raises(_lib.TsmiError, match="checkpointed"):
            k.adjoint(np.zeros_like(wl.psi))
    finally:
        k.close()


def test_torch_autograd_matches_oracle(oracle):
    import torch
    from tsmi import autograd as TA
    from tsmi.hip import HipKernel
    wl, target = case(80, 64, 60)
    lat = wl.lattice
    dev = torch.device("cuda:0")
    k = HipKernel(lat, delta_t=wl.dt, precision=64)
    try:
        v = torch.tensor(wl.v, device=dev, requires_grad=True)
        p_re = torch.tensor(np.ascontiguousarray(wl.psi.real), device=dev, requires_grad=True)
        p_im = torch.tensor(np.ascontiguousarray(wl.psi.imag), device=dev, requires_grad=True)
        t_re = torch.tensor(np.ascontiguousarray(target.real), device=dev)
        t_im = torch.tensor(np.ascontiguousarray(target.imag), device=dev)
        re, im = TA.evolve(k, v, p_re, p_im, wl.steps, every=16)
        loss = TA.infidelity(t_re, t_im, re, im, lat.delta_x * lat.delta_y)
        loss.backward()
        torch.cuda.synchronize()
        L_ref, g = infidelity(lat, target, oracle.evolve(wl))
        ref_v, ref_0 = oracle.adjoint(wl, g, every=10)
        assert abs(float(loss) - L_ref) < 1e-12
        assert rel_max(v.grad.cpu().numpy(), ref_v) < 1e-10
        assert rel_max(p_re.grad.cpu().numpy() + 1j * p_im.grad.cpu().numpy(), ref_0) < 1e-10
    finally:
        k.close()


@pytest.mark.parametrize("g", [0.0, 20.0])
def test_energy_loss_gradient_against_oracle_fd(oracle, g):
    """dE/dV of the final state's energy (tsmi.losses.energy_torch) through the
    autograd bridge -- the reverse sweep plus the direct potential term --
    against central differences of the oracle's evolution and energy at 12
    seeded cells (h = 1e-4; closed axes; g = 0 and g = 20, the energy then
    with its interaction term).  Bar: 1e-6 of the largest |FD|."""
    import torch
    from tsmi import autograd as TA
    from tsmi.hip import HipKernel
    from tsmi.losses import energy_torch
    wl, _ = case(56, 48, 40, g=g)
    lat = wl.lattice
    dev = torch.device("cuda:0")
    k = HipKernel(lat, delta_t=wl.dt, precision=64, coupling=g)
    try:
        v = torch.tensor(wl.v, device=dev, requires_grad=True)
        p_re = torch.tensor(np.ascontiguousarray(wl.psi.real), device=dev)
        p_im = torch.tensor(np.ascontiguousarray(wl.psi.imag), device=dev)
        re, im = TA.evolve(k, v, p_re, p_im, wl.steps, every=8)
        loss = energy_torch(lat, re, im, v, wl.mass, g=g)
        loss.backward()
        torch.cuda.synchronize()
        grad = v.grad.cpu().numpy()
    finally:
        k.close()
    import dataclasses
    e_ref = oracle.energy(wl, oracle.evolve(wl))[0]
    assert abs(float(loss) - e_ref) < 1e-12 * abs(e_ref)
    rng = np.random.default_rng(11)
    h = 1e-4
    fd, got = [], []
    for _ in range(12):
        y, x = int(rng.integers(4, lat.H - 4)), int(rng.integers(4, lat.W - 4))
        vals = []
        for sgn in (1.0, -1.0):
            v2 = wl.v.copy()
            v2[y, x] += sgn * h
            w2 = dataclasses.replace(wl, v=v2)
            vals.append(oracle.energy(w2, oracle.evolve(w2))[0])
        fd.append((vals[0] - vals[1]) / (2 * h))
        got.append(grad[y, x])
    fd, got = np.array(fd), np.array(got)
    assert np.abs(got - fd).max() <= 1e-6 * np.abs(fd).max()


@pytest.mark.parametrize("precision", [64, 32])
def test_config4_gradient_against_reference_fd(precision):
    """The 1024^2 differentiation workload (2000 steps, checkpoints every 100)
    against central differences of the reference solver at 32 cells
    (tests/golden/adjoint_config4.npz, written by make_adjoint_golden.py).
    Bar: max |dL/dV - FD| over the cells <= 1e-4 of the largest |FD|; the
    oracle's FP64 adjoint meets it at 2.9e-8 (largest gap from the FD noise of
    the smallest entries)."""
    import os
    from golden_util import input_digest
    from tsmi.hip import kernel_for
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "adjoint_config4.npz"))
    wl, target = W.config4()
    assert str(z["input_sha256"]) == input_digest(wl)
    k = kernel_for(wl, precision=precision)
    try:
        k.evolve_checkpointed(wl.steps, 100)
        L, g = infidelity(wl.lattice, target, k.sample())
        gv, _ = k.adjoint(g)
    finally:
        k.close()
    assert abs(L - float(z["loss0"])) < (1e-12 if precision == 64 else 1e-5)
    fd = z["fd"]
    err = float(np.max(np.abs(gv.ravel()[z["points"]] - fd)) / np.max(np.abs(fd)))
    assert err < 1e-4, err
