"""The trottersuzuki Python package driving the HIP kernels, against the oracle
on the same inputs: the quick-start set-up of the reference's documentation
(src/Python/doc/source/quickstart.rst: a Python-function potential and state,
Hamiltonian, Solver, evolve, energies), periodic and FP32 runs, imaginary time,
a time-dependent Python potential (the Solver's own step loop,
src/Python/trottersuzuki/classes_extension.py:388-403), a 1D chain, and two
components with Rabi coupling."""
import numpy as np
import pytest

from tsmi import lattice as L
from tsmi import workloads as W

pytestmark = pytest.mark.gpu

ts = pytest.importorskip("trottersuzuki")


def workload_of(grid, state, v, dt, steps, g=0.0, imag=False, mass=1.0):
    """The oracle's view of a trottersuzuki set-up (same tile, same values)."""
    lat = L.Lattice2D(grid.global_no_halo_dim_x, grid.length_x, grid.global_no_halo_dim_y,
                      grid.length_y if grid.length_y > 0 else 1.0, grid.periodic_x_axis, grid.periodic_y_axis)
    re, im = state._tile()
    return W.Workload("pyapi", lat, re + 1j * im, np.asarray(v, dtype=np.float64), dt, steps, imag_time=imag,
                      mass=mass, g=g)


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def tile(state):
    re, im = state._tile()
    return re + 1j * im


@pytest.mark.parametrize("kernel,periodic,tol", [("hip", False, 1e-12), ("hip", True, 1e-12),
                                                 ("hip-fp32", True, 2e-5)])
def test_quickstart_matches_oracle(oracle, kernel, periodic, tol):
    grid = ts.Lattice2D(96, 20., periodic_x_axis=periodic, periodic_y_axis=periodic)

    def harmonic_potential(x, y):
        return 0.5 * (x ** 2 + y ** 2)

    potential = ts.Potential(grid)
    potential.init_potential(harmonic_potential)
    hamiltonian = ts.Hamiltonian(grid, potential, 1., 3.0)

    def state_wave_function(x, y):
        return np.exp(-0.5 * ((x - 1.0) ** 2 + y ** 2)) / np.sqrt(np.pi) * np.exp(1j * 0.8 * y)

    state = ts.State(grid)
    state.init_state(state_wave_function)
    wl = workload_of(grid, state, potential.potential_matrix, 1e-3, 40, g=3.0)
    solver = ts.Solver(grid, state, hamiltonian, 1e-3, kernel_type=kernel)
    solver.evolve(25)
    solver.evolve(15)
    ref = oracle.evolve(wl, steps=25)
    ref = oracle.evolve(wl, steps=15, psi=ref)
    r = rel(tile(state), ref)
    assert r < tol, r
    e = oracle.energy(wl, ref)
    etol = 1e-11 if kernel == "hip" else 1e-5
    np.testing.assert_allclose(solver.get_total_energy(), e[0], rtol=etol)
    np.testing.assert_allclose(solver.get_kinetic_energy(), e[1], rtol=etol)
    np.testing.assert_allclose(solver.get_potential_energy(), e[2], rtol=etol)
    np.testing.assert_allclose(solver.get_squared_norm(), e[5], rtol=etol)
    # the State's own expected values see the state copied back by evolve
    np.testing.assert_allclose(state.get_squared_norm(), e[5], rtol=1e-10 if kernel == "hip" else 1e-5)
    assert abs(solver.current_evolution_time - 40e-3) < 1e-15


def test_builtin_state_and_harmonic_potential_imaginary_time(oracle):
    grid = ts.Lattice2D(80, 16.)
    pot = ts.HarmonicPotential(grid, 1., 1.)
    h = ts.Hamiltonian(grid, pot, 1., 10.0)
    state = ts.GaussianState(grid, 0.6, 0.6, 0.5, -0.3)
    X, Y = ts.classes.tile_coordinates(grid)
    wl = workload_of(grid, state, 0.5 * (X ** 2 + Y ** 2), 1e-3, 50, g=10.0, imag=True)
    solver = ts.Solver(grid, state, h, 1e-3, kernel_type="hip")
    solver.evolve(50, True)
    ref = oracle.evolve(wl, steps=50)
    assert rel(tile(state), ref) < 1e-12
    np.testing.assert_allclose(solver.get_total_energy(), oracle.energy(wl, ref)[0], rtol=1e-11)


def test_time_dependent_python_potential(oracle):
    """Each step uses exp(-i dt V(x, y, t)) at the step's start time and runs as
    a one-step evolve call, so no halo exchange falls between the steps."""
    grid = ts.Lattice2D(64, 12., periodic_x_axis=True, periodic_y_axis=True)

    def v_t(x, y, t):
        return 0.5 * (1.0 + 0.3 * np.sin(5.0 * t)) * (x ** 2 + y ** 2)

    pot = ts.Potential(grid)
    pot.init_potential(v_t)
    h = ts.Hamiltonian(grid, pot, 1., 1.0)
    state = ts.GaussianState(grid, 1.0, 1.0, 0.7, 0.2)
    dt, steps = 2e-3, 12
    wl = workload_of(grid, state, pot.potential_matrix, dt, steps, g=1.0)
    solver = ts.Solver(grid, state, h, dt, Potential=pot, kernel_type="hip")
    solver.evolve(steps)
    d = oracle.make_desc(wl)
    re = np.ascontiguousarray(wl.psi.real).copy()
    im = np.ascontiguousarray(wl.psi.imag).copy()
    X, Y = ts.classes.tile_coordinates(grid)
    for i in range(steps):
        er, ei = oracle.exp_potential(v_t(X, Y, i * dt), dt, False)
        oracle.lib().ts_oracle_full_step(d, oracle._p(re), oracle._p(im), oracle._p(er), oracle._p(ei))
    assert rel(tile(state), re + 1j * im) < 1e-12
    assert abs(solver.current_evolution_time - steps * dt) < 1e-15


def test_chain_matches_oracle(oracle):
    grid = ts.Lattice1D(300, 30.)
    pot = ts.Potential(grid)
    pot.init_potential(lambda x: 0.5 * x * x)
    h = ts.Hamiltonian(grid, pot, 1., 2.0)
    state = ts.GaussianState(grid, 1.0, 1.5)
    state.imprint(lambda x: np.exp(1j * 0.5 * x))
    wl = workload_of(grid, state, pot.potential_matrix, 1e-3, 60, g=2.0)
    solver = ts.Solver(grid, state, h, 1e-3)
    solver.evolve(60)
    ref = oracle.evolve(wl, steps=60)
    assert rel(tile(state), ref) < 1e-12
    e = oracle.energy(wl, ref)
    np.testing.assert_allclose(solver.get_total_energy(), e[0], rtol=1e-11)


def test_two_components_match_oracle(oracle):
    pw = W.pair_case(48, 40, False, False, g_a=3.0, g_b=2.0, g_ab=1.5, omega_r=2.0, omega_i=0.7, mass_b=1.7,
                     calls=(9,))
    lat = pw.lattice
    grid = ts.Lattice2D(lat.dim_x, lat.length_x, lat.dim_y, lat.length_y)
    pa, pb = ts.Potential(grid), ts.Potential(grid)
    pa.init_potential_matrix(np.ascontiguousarray(pw.v_a))
    pb.init_potential_matrix(np.ascontiguousarray(pw.v_b))
    h = ts.Hamiltonian2Component(grid, pa, pb, 1.0, 1.7, 3.0, 1.5, 2.0, 2.0, 0.7)
    sa, sb = ts.State(grid), ts.State(grid)
    sa.init_state_matrix(np.ascontiguousarray(pw.psi_a.real), np.ascontiguousarray(pw.psi_a.imag))
    sb.init_state_matrix(np.ascontiguousarray(pw.psi_b.real), np.ascontiguousarray(pw.psi_b.imag))
    solver = ts.Solver(grid, sa, h, pw.dt, State2=sb, kernel_type="hip")
    solver.evolve(9)
    ra, rb = oracle.evolve2(pw)
    assert rel(tile(sa), ra) < 1e-12
    assert rel(tile(sb), rb) < 1e-12
    e = oracle.energy2(pw, ra, rb)
    np.testing.assert_allclose(solver.get_total_energy(), e[0], rtol=1e-11)
