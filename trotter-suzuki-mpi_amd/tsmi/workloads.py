"""Synthetic workloads for the Trotter-Suzuki step (seeded, reproducible).

Random numbers come from SplitMix64; draw k of a stream with seed s is
``mix(s + (k+1) * 0x9E3779B97F4A7C15)`` and a uniform in [0,1) is
``(draw >> 11) * 2**-53``.  The stream is evaluated vectorised with numpy's
wrapping uint64 arithmetic, so any host builds identical inputs.

Every builder returns a :class:`Workload`: a lattice, the initial state and the
raw potential on the whole tile (halo included, as ``State`` and ``Potential``
hold them in the reference), and the Hamiltonian/run parameters.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .lattice import Lattice2D

GAMMA = np.uint64(0x9E3779B97F4A7C15)
SEED_BASE = 0x5EED0000


def splitmix64(seed: int, count: int, offset: int = 0) -> np.ndarray:
    """Draws ``offset .. offset+count-1`` of the SplitMix64 stream ``seed``."""
    with np.errstate(over="ignore"):
        k = np.arange(offset + 1, offset + count + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * GAMMA
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed: int, count: int, offset: int = 0, lo: float = 0.0, hi: float = 1.0) -> np.ndarray:
    u = (splitmix64(seed, count, offset) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)
    return lo + (hi - lo) * u


@dataclass
class Workload:
    name: str
    lattice: Lattice2D
    psi: np.ndarray            # complex128, (H, W)
    v: np.ndarray              # float64, (H, W): raw external potential
    dt: float
    steps: int
    imag_time: bool = False
    mass: float = 1.0
    g: float = 0.0
    g_lhy: float = 0.0
    extra: dict = field(default_factory=dict)


def squared_norm(lat: Lattice2D, psi: np.ndarray) -> float:
    inner = psi[lat.inner_y, lat.inner_x]
    return float(np.sum(inner.real ** 2 + inner.imag ** 2) * lat.delta_x * lat.delta_y)


def normalise(lat: Lattice2D, psi: np.ndarray, norm: float = 1.0) -> np.ndarray:
    return psi * np.sqrt(norm / squared_norm(lat, psi))


def harmonic(lat: Lattice2D, omega_x: float = 1.0, omega_y: float = 1.0, mass: float = 1.0) -> np.ndarray:
    """``HarmonicPotential::get_value`` on the tile (src/model.cpp:920-926)."""
    X, Y = lat.mesh()
    return 0.5 * mass * (omega_x * omega_x * X * X + omega_y * omega_y * Y * Y)


def gaussian_packet(lat: Lattice2D, x0: float, y0: float, kx: float, ky: float, width: float = 1.0) -> np.ndarray:
    X, Y = lat.mesh()
    r2 = (X - x0) ** 2 + (Y - y0) ** 2
    return np.exp(-r2 / (2.0 * width * width)) * np.exp(1j * (kx * X + ky * Y))


def gaussian_sum(lat: Lattice2D, seed: int, count: int, offset: int = 0) -> np.ndarray:
    """Sum of ``count`` Gaussians d*exp(-|r-c|^2/(2 w^2)) with seeded centres in
    the box, widths in [0.5, 2] and depths in [-5, 5]."""
    u = uniform(seed, 5 * count, offset)
    X, Y = lat.mesh()
    v = np.zeros((lat.H, lat.W))
    for k in range(count):
        cx = (u[5 * k] - 0.5) * lat.length_x
        cy = (u[5 * k + 1] - 0.5) * lat.length_y
        w = 0.5 + 1.5 * u[5 * k + 2]
        d = -5.0 + 10.0 * u[5 * k + 3]
        v += d * np.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / (2.0 * w * w))
    return v


# ---------------------------------------------------------------------------
# The five configurations of the benchmark suite (sizes may be overridden for
# parity tests at reduced scale).
# ---------------------------------------------------------------------------

def config0(n: int = 256, steps: int = 1000, length: float = 20.0) -> Workload:
    """Plain method: closed 2D Schroedinger, harmonic V, seeded Gaussian packet,
    real time, dt=1e-3, FP64."""
    seed = SEED_BASE + 0
    lat = Lattice2D(n, length)
    x0, y0, kx, ky = uniform(seed, 4, lo=-2.0, hi=2.0)
    psi = normalise(lat, gaussian_packet(lat, x0, y0, kx, ky))
    return Workload("config0", lat, psi, harmonic(lat), 1e-3, steps,
                    extra=dict(x0=x0, y0=y0, kx=kx, ky=ky))


def config1(n: int = 2048, steps: int = 10000, length: float = 24.0) -> Workload:
    """Imaginary-time GPE ground state: closed, harmonic V, g=100, unit Gaussian
    plus 1% seeded complex uniform noise, dt=1e-4, renormalised every step."""
    seed = SEED_BASE + 1
    lat = Lattice2D(n, length)
    X, Y = lat.mesh()
    m = lat.H * lat.W
    noise = uniform(seed, 2 * m, lo=-1.0, hi=1.0)
    psi = np.exp(-(X * X + Y * Y) / 2.0) + 0.01 * (noise[:m] + 1j * noise[m:]).reshape(lat.H, lat.W)
    psi = normalise(lat, psi)
    return Workload("config1", lat, psi, harmonic(lat), 1e-4, steps, imag_time=True, g=100.0)


def config2(n: int = 8192, steps: int = 1000, length: float = 64.0, n_vortices: int = 16,
            g: float = 50.0, trap: float = 0.01) -> Workload:
    """Real-time GPE vortex dynamics: periodic, V = trap*(x^2+y^2), g=50,
    Thomas-Fermi amplitude with 16 seeded vortices (tanh(r) cores, +-1 charges),
    dt=5e-4."""
    seed = SEED_BASE + 2
    lat = Lattice2D(n, length, periodic_x=True, periodic_y=True)
    X, Y = lat.mesh()
    v = trap * (X * X + Y * Y)
    mu = np.sqrt(2.0 * trap * g / np.pi)          # unit norm for the 2D TF profile
    r_tf = np.sqrt(mu / trap)
    amp = np.sqrt(np.maximum(mu - v, 0.0) / g)
    u = uniform(seed, 3 * n_vortices)
    psi = amp.astype(np.complex128)
    for j in range(n_vortices):
        rad = 0.7 * r_tf * np.sqrt(u[3 * j])
        ang = 2.0 * np.pi * u[3 * j + 1]
        xj, yj = rad * np.cos(ang), rad * np.sin(ang)
        s = 1.0 if u[3 * j + 2] < 0.5 else -1.0
        dx, dy = X - xj, Y - yj
        psi = psi * np.tanh(np.sqrt(dx * dx + dy * dy)) * np.exp(1j * s * np.arctan2(dy, dx))
    psi = normalise(lat, psi)
    return Workload("config2", lat, psi, v, 5e-4, steps, g=g, extra=dict(mu=mu, r_tf=r_tf))


CONFIG3_SIDES = (64, 128, 256, 512, 1024)


def config3_specs(count: int = 512, sides=CONFIG3_SIDES):
    """The seeded shape of the batching workload: per lattice (side, periodic),
    side uniform over ``sides``, both axes periodic with probability 1/2."""
    u = uniform(SEED_BASE + 3, 2 * count)
    out = []
    for i in range(count):
        side = sides[min(int(u[2 * i] * len(sides)), len(sides) - 1)]
        out.append((side, bool(u[2 * i + 1] < 0.5)))
    return out


def config3_state(i: int, side: int, periodic: bool, steps: int = 500, length: float = 20.0) -> Workload:
    """Lattice i of the batching workload: V a seeded sum of 8 Gaussians, a
    seeded Gaussian packet as in config0, dt=1e-3, real time."""
    seed = SEED_BASE + 3000 + i
    lat = Lattice2D(side, length, periodic_x=periodic, periodic_y=periodic)
    x0, y0, kx, ky = uniform(seed, 4, lo=-2.0, hi=2.0)
    psi = normalise(lat, gaussian_packet(lat, x0, y0, kx, ky))
    v = gaussian_sum(lat, seed, 8, offset=4)
    return Workload(f"config3_{i}", lat, psi, v, 1e-3, steps)


def config3(count: int = 512, steps: int = 500, sides=CONFIG3_SIDES, length: float = 20.0):
    """Batching: ``count`` independent real-time Schroedinger lattices of seeded
    sides and boundaries (list of Workloads)."""
    return [config3_state(i, s, p, steps, length) for i, (s, p) in enumerate(config3_specs(count, sides))]


def config4(n: int = 1024, steps: int = 2000, length: float = 20.0):
    """Differentiation: closed linear Schroedinger with V a seeded sum of 16
    Gaussians, seeded Gaussian packet, target a displaced Gaussian at a seeded
    point.  Returns (workload, target)."""
    seed = SEED_BASE + 4
    lat = Lattice2D(n, length)
    x0, y0, kx, ky, tx, ty = uniform(seed, 6, lo=-2.0, hi=2.0)
    psi = normalise(lat, gaussian_packet(lat, x0, y0, kx, ky))
    v = gaussian_sum(lat, seed, 16, offset=6)
    target = normalise(lat, gaussian_packet(lat, tx, ty, 0.0, 0.0))
    return Workload("config4", lat, psi, v, 1e-3, steps, extra=dict(checkpoint=100)), target


def exponential_state(lat: Lattice2D, n_x: int = 1, n_y: int = 1, norm: float = 1.0, phase: float = 0.0) -> np.ndarray:
    """Plane wave of ``ExponentialState`` (src/model.cpp:668-672)."""
    X, Y = lat.mesh()
    lx, ly = lat.dim_x * lat.delta_x, lat.dim_y * lat.delta_y
    return (np.sqrt(norm / (lx * ly)) * np.exp(1j * phase)
            * np.exp(1j * (2 * np.pi * n_x / lx * X + 2 * np.pi * n_y / ly * Y)))


def gaussian_state(lat: Lattice2D, omega_x: float, omega_y: float = -1.0, mean_x: float = 0.0,
                   mean_y: float = 0.0, norm: float = 1.0, phase: float = 0.0) -> np.ndarray:
    """``GaussianState`` (src/model.cpp:692-714)."""
    if omega_y == -1.0:
        omega_y = omega_x
    X, Y = lat.mesh()
    amp = np.sqrt(norm * np.sqrt(omega_x * omega_y) / np.pi)
    return (amp * np.exp(-(omega_x * (X - mean_x) ** 2 + omega_y * (Y - mean_y) ** 2) * 0.5)
            * np.exp(1j * phase)).astype(np.complex128)


def sinusoid_state(lat: Lattice2D, n_x: int = 1, n_y: int = 1, norm: float = 1.0, phase: float = 0.0) -> np.ndarray:
    """``SinusoidState`` (src/model.cpp:744-748)."""
    X, Y = lat.mesh()
    lx, ly = lat.dim_x * lat.delta_x, lat.dim_y * lat.delta_y
    return (np.sqrt(norm / (lx * ly)) * 2.0 * np.exp(1j * phase)
            * np.sin(2 * np.pi * n_x / lx * X) * np.sin(2 * np.pi * n_y / ly * Y))


def reference_scenario(name: str) -> Workload:
    """The single-component scenarios of the reference's kernel tests
    (test/kerneltest.cpp:8-125; DIM 250, LENGTH 100)."""
    dim, length = 250, 100.0
    if name == "free_particle":
        lat = Lattice2D(dim, length, periodic_x=True, periodic_y=True)
        return Workload(name, lat, exponential_state(lat), np.zeros((lat.H, lat.W)), 5e-3, 100)
    lat = Lattice2D(dim, length)
    v = harmonic(lat)
    if name == "harmonic_oscillator":
        return Workload(name, lat, gaussian_state(lat, 1.0), v, 5e-3, 100)
    if name == "imaginary_harmonic_oscillator":
        return Workload(name, lat, gaussian_state(lat, 0.5), v, 5e-3, 1000, imag_time=True)
    if name == "intra_particle_interaction":
        return Workload(name, lat, gaussian_state(lat, 1.0), v, 1e-3, 1000, g=10.0)
    if name == "imaginary_intra_particle_interaction":
        return Workload(name, lat, gaussian_state(lat, 1.0), v, 1e-3, 1000, imag_time=True, g=10.0)
    raise ValueError(name)


REFERENCE_SCENARIOS = ("free_particle", "harmonic_oscillator", "imaginary_harmonic_oscillator",
                       "intra_particle_interaction", "imaginary_intra_particle_interaction")


def small_case(n_x: int, n_y: int, periodic_x: bool, periodic_y: bool, *, imag_time: bool = False,
               g: float = 0.0, g_lhy: float = 0.0, steps: int = 10, dt: float = 1e-3,
               length: float = 20.0, seed: int = SEED_BASE + 100) -> Workload:
    """Small parity case: seeded packet in a harmonic plus Gaussian-bump potential."""
    lat = Lattice2D(n_x, length, n_y, length * n_y / n_x, periodic_x, periodic_y)
    x0, y0, kx, ky = uniform(seed, 4, lo=-2.0, hi=2.0)
    psi = normalise(lat, gaussian_packet(lat, x0, y0, kx, ky, width=1.5))
    v = 0.5 * harmonic(lat) + gaussian_sum(lat, seed, 2, offset=4)
    return Workload(f"small_{n_x}x{n_y}_p{int(periodic_x)}{int(periodic_y)}_{'imag' if imag_time else 'real'}",
                    lat, psi, v, dt, steps, imag_time=imag_time, g=g, g_lhy=g_lhy)


@dataclass
class PairWorkload:
    """A two-component system on one lattice: Hamiltonian2Component
    (src/trottersuzuki.h:459-497) with the two states of the two-state Solver
    (src/solver.cpp:43-60).  ``calls`` are the step counts of consecutive
    Solver::evolve calls."""
    name: str
    lattice: Lattice2D
    psi_a: np.ndarray
    psi_b: np.ndarray
    v_a: np.ndarray
    v_b: np.ndarray
    dt: float
    calls: tuple
    imag_time: bool = False
    mass_a: float = 1.0
    mass_b: float = 1.0
    g_a: float = 0.0
    g_b: float = 0.0
    g_ab: float = 0.0
    omega_r: float = 0.0
    omega_i: float = 0.0

    @property
    def steps(self) -> int:
        return int(sum(self.calls))

    def component(self, which: int) -> Workload:
        """Single-component view of component 0 (a) or 1 (b): its state,
        potential, mass and self coupling."""
        if which == 0:
            return Workload(self.name + "_a", self.lattice, self.psi_a, self.v_a, self.dt, self.steps,
                            imag_time=self.imag_time, mass=self.mass_a, g=self.g_a)
        return Workload(self.name + "_b", self.lattice, self.psi_b, self.v_b, self.dt, self.steps,
                        imag_time=self.imag_time, mass=self.mass_b, g=self.g_b)


def pair_case(n_x: int, n_y: int, periodic_x: bool, periodic_y: bool, *, imag_time: bool = False,
              g_a: float = 0.0, g_b: float = 0.0, g_ab: float = 0.0, omega_r: float = 0.0,
              omega_i: float = 0.0, mass_b: float = 1.0, calls=(10,), dt: float = 1e-3,
              length: float = 20.0, seed: int = SEED_BASE + 500) -> PairWorkload:
    """Small two-component parity case: two seeded packets (the second carrying a
    third of the norm) in two different seeded potentials."""
    lat = Lattice2D(n_x, length, n_y, length * n_y / n_x, periodic_x, periodic_y)
    u = uniform(seed, 8, lo=-2.0, hi=2.0)
    psi_a = normalise(lat, gaussian_packet(lat, u[0], u[1], u[2], u[3], width=1.5), 1.0)
    psi_b = normalise(lat, gaussian_packet(lat, u[4], u[5], u[6], u[7], width=1.2), 1.0 / 3.0)
    v_a = 0.5 * harmonic(lat) + gaussian_sum(lat, seed, 2, offset=8)
    v_b = 0.4 * harmonic(lat) + gaussian_sum(lat, seed + 1, 2, offset=8)
    tag = f"pair_{n_x}x{n_y}_p{int(periodic_x)}{int(periodic_y)}_{'imag' if imag_time else 'real'}"
    return PairWorkload(tag, lat, psi_a, psi_b, v_a, v_b, dt, tuple(calls), imag_time=imag_time,
                        mass_b=mass_b, g_a=g_a, g_b=g_b, g_ab=g_ab, omega_r=omega_r, omega_i=omega_i)


def mixed_bec(imag_time: bool) -> PairWorkload:
    """The two-component scenarios of the reference's kernel tests
    (test/kerneltest.cpp:178-237): DIM 250, LENGTH 100, closed; component a a
    Gaussian of omega 1, component b empty; both in the harmonic potential of
    omega 1; no interactions; Rabi coupling omega_r = 2 pi / 10; dt = 1e-3;
    5000 real-time or 1000 imaginary-time steps in one call."""
    lat = Lattice2D(250, 100.0)
    v = harmonic(lat)
    psi_a = gaussian_state(lat, 1.0)
    psi_b = np.zeros_like(psi_a)
    steps = 1000 if imag_time else 5000
    return PairWorkload("mixed_bec_" + ("imag" if imag_time else "real"), lat, psi_a, psi_b, v, v.copy(), 1e-3,
                        (steps,), imag_time=imag_time, omega_r=2.0 * np.pi / 10.0)
