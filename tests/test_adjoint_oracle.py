"""The oracle's discrete adjoint (oracle/ts_oracle.c: ts_oracle_adjoint) against
central finite differences of the oracle forward, which matches the reference
solver bit for bit (tests/test_oracle_golden.py): closed axes without and with
the nonlinear and LHY terms, periodic axes (every tile entry of V and psi_0 its
own variable, halo entries included), and the fold of tile derivatives onto a
periodic field (fold_halo) against differences along a periodic perturbation.
CPU only."""
import numpy as np
import pytest

from tsmi import workloads as W
from tsmi.losses import infidelity


def case(n_x=32, n_y=24, steps=40, dt=4e-3, length=8.0, px=False, py=False, g=0.0, g_lhy=0.0):
    wl = W.small_case(n_x, n_y, px, py, steps=steps, dt=dt, length=length, g=g, g_lhy=g_lhy)
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


def test_oracle_adjoint_rejects_imaginary_time(oracle):
    wl = W.small_case(16, 16, False, False, steps=3, imag_time=True)
    with pytest.raises(ValueError):
        oracle.adjoint(wl, np.zeros_like(wl.psi))


def fd_check(oracle, wl, target, cells, h=1e-4, every=9, rtol=1e-6):
    """dL/dV at tile cells against central differences of single entries."""
    _, g = infidelity(wl.lattice, target, oracle.evolve(wl))
    gv, _ = oracle.adjoint(wl, g, every=every)
    worst = 0.0
    for j in cells:
        vp = wl.v.copy(); vp.flat[j] += h
        vm = wl.v.copy(); vm.flat[j] -= h
        fd = (loss_of(oracle, wl, target, v=vp) - loss_of(oracle, wl, target, v=vm)) / (2 * h)
        worst = max(worst, abs(gv.flat[j] - fd) / np.abs(gv).max())
        assert abs(gv.flat[j] - fd) <= rtol * np.abs(gv).max() + rtol * abs(fd), (j, gv.flat[j], fd)
    return gv, worst


@pytest.mark.parametrize("g,g_lhy", [(40.0, 0.0), (25.0, 12.0)])
def test_oracle_adjoint_nonlinear_matches_finite_differences(oracle, g, g_lhy):
    """The potential stage's pullback with the density-dependent phase: a
    dropped 2 phi'(s) Im(conj(mu) y) chi term, or a wrong LHY factor, moves the
    gradient by far more than the bar."""
    wl, target = case(steps=30, g=g, g_lhy=g_lhy)
    assert np.abs(wl.psi).max() ** 2 * g * wl.dt > 1e-2     # the nonlinear phase is not negligible
    fd_check(oracle, wl, target, points(wl, 8, 3))


def test_oracle_adjoint_nonlinear_initial_state_gradient(oracle):
    wl, target = case(steps=20, g=40.0, g_lhy=10.0)
    _, g = infidelity(wl.lattice, target, oracle.evolve(wl))
    _, g0 = oracle.adjoint(wl, g, every=6)
    h = 1e-6
    for j in points(wl, 4, 4):
        for unit in (1.0, 1j):
            pp = wl.psi.copy(); pp.flat[j] += h * unit
            pm = wl.psi.copy(); pm.flat[j] -= h * unit
            fd = (loss_of(oracle, wl, target, psi0=pp) - loss_of(oracle, wl, target, psi0=pm)) / (2 * h)
            got = g0.flat[j].real if unit == 1.0 else g0.flat[j].imag
            assert abs(got - fd) <= 1e-6 * np.abs(g0).max(), (j, unit, got, fd)


def halo_cells(wl, k, seed):
    """k halo cells of the tile next to the inner region (where the forward's
    inner cells read them)."""
    lat = wl.lattice
    mask = np.ones((lat.H, lat.W), bool)
    mask[lat.inner_y, lat.inner_x] = False
    near = np.zeros_like(mask)
    near[max(0, lat.inner_y.start - 2):lat.inner_y.stop + 2, max(0, lat.inner_x.start - 2):lat.inner_x.stop + 2] = True
    cand = np.flatnonzero((mask & near).ravel())
    w = np.abs(oracle_free_density(wl)).ravel()[cand] + 1e-3
    rng = np.random.default_rng(seed)
    return rng.choice(cand, size=k, replace=False, p=w / w.sum())


def oracle_free_density(wl):
    return np.abs(wl.psi) ** 2


def periodic_case(px, py, g=30.0, steps=24):
    """A packet that crosses the periodic edges during the run."""
    wl, target = case(n_x=28, n_y=20, steps=steps, dt=4e-3, length=6.0, px=px, py=py, g=g)
    return wl, target


@pytest.mark.parametrize("px,py", [(True, True), (False, True), (True, False)])
def test_oracle_adjoint_periodic_matches_finite_differences(oracle, px, py):
    """Periodic axes: the halo exchange's transpose between the reverse steps;
    each tile entry of V is its own variable, halo entries included (the
    forward's inner cells read the halo's potential in the first half of a step
    and its values through the exchange)."""
    wl, target = periodic_case(px, py)
    gv, _ = fd_check(oracle, wl, target, points(wl, 6, 5))
    fd_check(oracle, wl, target, halo_cells(wl, 4, 6))
    lat = wl.lattice
    halo = np.ones((lat.H, lat.W), bool)
    halo[lat.inner_y, lat.inner_x] = False
    assert np.abs(gv[halo]).max() > 1e-3 * np.abs(gv).max()   # the halo's potential matters


def test_oracle_adjoint_periodic_initial_state_gradient(oracle):
    wl, target = periodic_case(True, True, steps=16)
    _, g = infidelity(wl.lattice, target, oracle.evolve(wl))
    _, g0 = oracle.adjoint(wl, g, every=5)
    h = 1e-6
    for j in list(points(wl, 3, 7)) + list(halo_cells(wl, 3, 8)):
        for unit in (1.0, 1j):
            pp = wl.psi.copy(); pp.flat[j] += h * unit
            pm = wl.psi.copy(); pm.flat[j] -= h * unit
            fd = (loss_of(oracle, wl, target, psi0=pp) - loss_of(oracle, wl, target, psi0=pm)) / (2 * h)
            got = g0.flat[j].real if unit == 1.0 else g0.flat[j].imag
            assert abs(got - fd) <= 1e-6 * np.abs(g0).max(), (j, unit, got, fd)


def test_oracle_fold_is_the_periodic_field_derivative(oracle):
    """fold_halo of the tile gradient = the derivative along a periodic
    perturbation (a cell and all its halo copies moved together), the gradient
    the periodic reverse sweep of the library reports."""
    wl, target = periodic_case(True, True)
    lat = wl.lattice
    wl = __import__("dataclasses").replace(wl, v=oracle.halo_exchange(wl, wl.v),
                                          psi=oracle.halo_exchange(wl, wl.psi))
    _, g = infidelity(lat, target, oracle.evolve(wl))
    gv, _ = oracle.adjoint(wl, g, every=8)
    folded = oracle.fold_halo(wl, gv)
    halo = np.ones((lat.H, lat.W), bool)
    halo[lat.inner_y, lat.inner_x] = False
    assert not folded[halo].any()
    h = 1e-4
    # corner, edge and interior cells of the inner region
    ys = [lat.inner_y.start, lat.inner_y.start + 1, lat.H // 2, lat.inner_y.stop - 1]
    xs = [lat.inner_x.start, lat.inner_x.stop - 2, lat.W // 2, lat.inner_x.stop - 1]
    for y, x in zip(ys, xs):
        e = np.zeros((lat.H, lat.W)); e[y, x] = 1.0
        e = oracle.halo_exchange(wl, e)
        assert e.sum() > 1.0 or (y, x) == (lat.H // 2, lat.W // 2)
        fd = (loss_of(oracle, wl, target, v=wl.v + h * e) - loss_of(oracle, wl, target, v=wl.v - h * e)) / (2 * h)
        assert abs(folded[y, x] - fd) <= 1e-6 * np.abs(folded).max(), (y, x, folded[y, x], fd)
    # the inner entries alone are not the periodic derivative: the halo copies add to them
    inner = ~halo
    assert np.abs(gv - folded)[inner].max() > 1e-3 * np.abs(folded).max()


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
