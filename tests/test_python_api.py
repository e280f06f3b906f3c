"""The trottersuzuki Python package (the reference's Python API over this
project's host library), on the CPU: lattice geometry, the built-in states
against their closed forms (src/model.cpp:640-748), states and potentials
from Python functions (array and point-by-point evaluation agree), the
time-dependent potential's table, coordinate mapping against the lattice
module, and get_vortex_position on an imprinted vortex.  No GPU work."""
import math

import numpy as np
import pytest

ts = pytest.importorskip("trottersuzuki")

from tsmi import lattice as L  # noqa: E402


@pytest.mark.parametrize("px,py", [(False, False), (True, False), (True, True)])
def test_lattice2d_geometry_matches_lattice_module(px, py):
    g = ts.Lattice2D(40, 12.0, 30, 9.0, px, py)
    ref = L.Lattice2D(40, 12.0, 30, 9.0, px, py)
    assert (g.dim_x, g.dim_y) == (ref.W, ref.H)
    assert (g.start_x, g.start_y) == (ref.start_x, ref.start_y)
    assert g.delta_x == ref.delta_x and g.delta_y == ref.delta_y
    assert g.periodic_x_axis == px and g.periodic_y_axis == py
    np.testing.assert_array_equal(g.get_x_axis(), (np.arange(40) - 20 + 0.5) * 0.3)
    # square lattice when the y values are omitted
    s = ts.Lattice2D(24, 5.0)
    assert (s.dim_x, s.dim_y, s.length_y) == (24, 24, 5.0)


def test_lattice1d_geometry():
    g = ts.Lattice1D(100, 10.0)
    assert (g.dim_x, g.dim_y, g.halo_y, g.delta_y, g.length_y) == (100, 1, 0, 1.0, 0.0)
    gp = ts.Lattice1D(100, 10.0, True)
    assert gp.dim_x == 108 and gp.start_x == -4
    with pytest.raises(RuntimeError):
        ts.Lattice1D(10, 1.0, False, "cylindrical")


def test_coordinates_match_lattice_module():
    g = ts.Lattice2D(33, 7.0, 21, 5.0, True, False)
    ref = L.Lattice2D(33, 7.0, 21, 5.0, True, False)
    X, Y = ref.mesh()
    for (y, x) in [(0, 0), (3, 17), (20, 40), (10, 1)]:
        cx, cy = ts.map_lattice_to_coordinate_space(g, x, y)
        assert cx == X[y, x] and cy == Y[y, x]


def test_builtin_states_closed_forms():
    g = ts.Lattice2D(48, 10.0, 40, 8.0)
    X, Y = ts.classes.tile_coordinates(g)
    lx, ly = 10.0, 8.0
    # Gaussian: sqrt(norm sqrt(wx wy)/pi) exp(-(wx (x-mx)^2 + wy (y-my)^2)/2) e^{i phase}
    s = ts.GaussianState(g, 1.5, 0.8, 0.3, -0.2, 2.0, 0.7)
    re, im = s._tile()
    want = np.sqrt(2.0 * np.sqrt(1.5 * 0.8) / np.pi) * np.exp(-(1.5 * (X - 0.3) ** 2 + 0.8 * (Y + 0.2) ** 2) / 2)
    np.testing.assert_allclose(re + 1j * im, want * np.exp(0.7j), rtol=1e-13, atol=1e-15)
    # omega_y = -1 takes omega_x
    re2, _ = ts.GaussianState(g, 1.5)._tile()
    np.testing.assert_allclose(re2, np.sqrt(1.5 / np.pi) * np.exp(-1.5 * (X ** 2 + Y ** 2) / 2), rtol=1e-13)
    # exponential: sqrt(norm/(Lx Ly)) e^{i phase} e^{i 2 pi (nx x/Lx + ny y/Ly)}
    re, im = ts.ExponentialState(g, 2, 3, 1.0, 0.4)._tile()
    want = np.sqrt(1.0 / (lx * ly)) * np.exp(0.4j) * np.exp(2j * np.pi * (2 * X / lx + 3 * Y / ly))
    np.testing.assert_allclose(re + 1j * im, want, rtol=1e-12, atol=1e-14)
    # sinusoid: 2 sqrt(norm/(Lx Ly)) e^{i phase} sin(2 pi nx x/Lx) sin(2 pi ny y/Ly)
    re, im = ts.SinusoidState(g, 1, 2)._tile()
    want = 2 * np.sqrt(1.0 / (lx * ly)) * np.sin(2 * np.pi * X / lx) * np.sin(4 * np.pi * Y / ly)
    np.testing.assert_allclose(re, want, rtol=1e-12, atol=1e-14)
    assert not im.any()


def test_chain_states():
    g = ts.Lattice1D(64, 8.0)
    x = ts.map_lattice_to_coordinate_space(g, np.arange(64))
    re, _ = ts.GaussianState(g, 2.0, 0.5)._tile()
    # on a chain omega_y is 1 and y = 0
    np.testing.assert_allclose(re[0], np.sqrt(np.sqrt(2.0) / np.pi) * np.exp(-2.0 * (x - 0.5) ** 2 / 2), rtol=1e-13)
    # the chain's sinusoid has n_y = 0, so it is zero (as in the reference)
    re, im = ts.SinusoidState(g, 3)._tile()
    assert not re.any() and not im.any()
    assert ts.GaussianState(g, 1.0).get_particle_density().shape == (1, 64)


def test_init_state_array_and_pointwise_agree():
    g = ts.Lattice2D(20, 4.0, 16, 3.0, True, False)
    a, b = ts.State(g), ts.State(g)
    a.init_state(lambda x, y: np.exp(-0.5 * (x * x + y * y)) * np.exp(1j * 0.3 * x))
    b.init_state(lambda x, y: math.exp(-0.5 * (x * x + y * y)) * complex(math.cos(0.3 * x), math.sin(0.3 * x)))
    ra, ia = a._tile()
    rb, ib = b._tile()
    np.testing.assert_allclose(ra, rb, rtol=1e-15, atol=1e-16)
    np.testing.assert_allclose(ia, ib, rtol=1e-15, atol=1e-16)
    # a function of x alone, and a constant
    c = ts.State(g)
    c.init_state(lambda x: 2.0)
    assert np.all(c._tile()[0] == 2.0)


def test_imprint_multiplies():
    g = ts.Lattice2D(24, 6.0)
    s = ts.GaussianState(g, 1.0)
    r0, _ = s._tile()
    s.imprint(lambda x, y: np.exp(1j * np.angle(x + 1j * y)))
    re, im = s._tile()
    X, Y = ts.classes.tile_coordinates(g)
    np.testing.assert_allclose(re + 1j * im, r0 * np.exp(1j * np.angle(X + 1j * Y)), rtol=1e-14, atol=1e-16)


def test_potential_from_functions():
    g = ts.Lattice2D(16, 4.0)
    X, Y = ts.classes.tile_coordinates(g)
    p = ts.Potential(g)
    p.init_potential(lambda x, y: 0.5 * (x * x + y * y))
    assert not p.updated_potential_matrix and p.pot_function is None
    assert p.get_value(3, 5) == 0.5 * (X[5, 3] ** 2 + Y[5, 3] ** 2)
    # time-dependent: tabulated at t = 0, flagged, table exp(-i dt V(t))
    q = ts.Potential(g)
    q.init_potential(lambda x, y, t: (1 + t) * x * x)
    assert q.updated_potential_matrix
    assert q.get_value(2, 7) == X[7, 2] ** 2
    e = q.exponential_update(1e-3, 0.5)
    np.testing.assert_allclose(e, np.exp(-1e-3j * 1.5 * X * X), rtol=1e-15)
    h = ts.Hamiltonian(g, q, 1.0, 2.0)
    assert h.potential is q and h.coupling_a == 2.0


def test_harmonic_potential():
    g = ts.Lattice2D(20, 5.0)
    X, Y = ts.classes.tile_coordinates(g)
    p = ts.HarmonicPotential(g, 2.0, 0.5, 1.5, 0.2, -0.1)
    # 0.5 m (wx^2 (x-mx)^2 + wy^2 (y-my)^2)
    want = 0.5 * 1.5 * (4.0 * (X[4, 9] - 0.2) ** 2 + 0.25 * (Y[4, 9] + 0.1) ** 2)
    assert abs(p.get_value(9, 4) - want) < 1e-13


def test_vortex_position():
    g = ts.Lattice2D(64, 16.0)
    s = ts.GaussianState(g, 0.2)
    s.imprint(lambda x, y: np.exp(1j * np.angle((x - 1.25) + 1j * (y + 0.75))))
    pos = ts.get_vortex_position(g, s, 6.0)
    assert abs(pos[0] - 1.25) <= g.delta_x and abs(pos[1] + 0.75) <= g.delta_y
    # no vortex: the phase has no jump around the candidate point
    assert np.isnan(ts.get_vortex_position(g, ts.GaussianState(g, 0.2), 6.0)).all()


def test_solver_kernel_types():
    g = ts.Lattice2D(16, 4.0)
    s = ts.GaussianState(g, 1.0)
    h = ts.Hamiltonian(g)
    with pytest.raises(ValueError):
        ts.Solver(g, s, h, 1e-3, kernel_type="cpu")
    with pytest.raises(NotImplementedError):
        ts.BesselState(g)
