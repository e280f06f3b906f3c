"""Coordinate mapping and state analysis of the Python API
(src/Python/trottersuzuki/tools.py: map_lattice_to_coordinate_space,
imprint, get_vortex_position), written over numpy arrays."""
from __future__ import annotations

import numpy as np


def _axis(start, delta, dim_no_halo, length, idx):
    """Physical coordinate of lattice index ``idx`` (scalar or array) along
    one Cartesian axis, wrapped into [-L/2, L/2] as src/common.cpp:45-65."""
    pos = start * delta + 0.5 * delta + np.asarray(idx, dtype=np.float64) * delta
    centre = dim_no_halo * delta * 0.5
    pos = np.where(pos - centre < -length * 0.5, pos + length, pos)
    pos = np.where(pos - centre > length * 0.5, pos - length, pos)
    return pos - centre


def map_lattice_to_coordinate_space(grid, x, y=None):
    """Physical coordinates of the tile point (x, y) of ``grid`` (Cartesian):
    x alone for a chain (``y`` omitted), else (x, y).  x and y may be arrays."""
    if grid.coordinate_system != "cartesian":
        raise NotImplementedError("only Cartesian coordinates are on this path")
    xs = _axis(grid.start_x, grid.delta_x, grid.global_no_halo_dim_x, grid.length_x, x)
    if np.ndim(xs) == 0:
        xs = float(xs)
    if y is None:
        return xs
    ys = _axis(grid.start_y, grid.delta_y, grid.global_no_halo_dim_y, grid.length_y, y)
    if np.ndim(ys) == 0:
        ys = float(ys)
    return xs, ys


def tile_coordinates(grid):
    """(X, Y): the physical coordinates of every tile point, each (dim_y, dim_x)."""
    xs = map_lattice_to_coordinate_space(grid, np.arange(grid.dim_x), np.arange(grid.dim_y))
    return np.meshgrid(np.atleast_1d(xs[0]), np.atleast_1d(xs[1]))


def _arity_one(function):
    """True when ``function`` takes a single coordinate (the reference probes
    ``function(0)``)."""
    try:
        function(0)
        return True
    except TypeError:
        return False


def evaluate_on_tile(grid, function, dtype=np.complex128, extra=()):
    """``function`` over every tile point as a (dim_y, dim_x) array.  A function
    of one argument is taken as f(x).  The function is first called on the
    whole coordinate arrays; if it does not accept arrays (or returns the
    wrong shape) it is called point by point, as the reference does."""
    X, Y = tile_coordinates(grid)
    args = (X,) if _arity_one(function) else (X, Y)
    try:
        with np.errstate(all="ignore"):
            v = np.asarray(function(*args, *extra), dtype=dtype)
        v = np.broadcast_to(v, X.shape).astype(dtype, copy=True)
        return v
    except (TypeError, ValueError):
        pass
    out = np.empty(X.shape, dtype=dtype)
    for j in range(X.shape[0]):
        for i in range(X.shape[1]):
            a = (X[j, i],) if len(args) == 1 else (X[j, i], Y[j, i])
            out[j, i] = function(*a, *extra)
    return out


def imprint(state, function):
    """Multiplies the wave function of ``state`` by ``function`` (of (x, y), or
    of x alone): psi(x, y) <- f(x, y) psi(x, y)."""
    m = evaluate_on_tile(state.grid, function)
    state.imprint_matrix(np.ascontiguousarray(m.real), np.ascontiguousarray(m.imag))


def get_vortex_position(grid, state, approx_cloud_radius=0.):
    """Position (x, y) of the single vortex of ``state`` (Cartesian): the mean
    position of the points whose phase-gradient norm is within 10 % of the
    largest one inside the given radius, provided a loop of 32 points around
    the largest one crosses exactly one 2-pi jump of the phase; [nan, nan]
    otherwise."""
    if approx_cloud_radius == 0.:
        approx_cloud_radius = np.sqrt(2) * grid.length_x
    dx = grid.length_x / float(grid.dim_x)
    dy = grid.length_y / float(grid.dim_y)
    phase = state.get_phase()
    ny, nx = grid.dim_y, grid.dim_x
    grad = np.zeros((nx, ny))
    iy, ix = np.meshgrid(np.arange(1, ny - 1), np.arange(1, nx - 1), indexing="ij")
    inside = (ix - nx * 0.5) ** 2 + (iy - ny * 0.5) ** 2 < approx_cloud_radius ** 2 / dx ** 2
    up, dw = phase[iy + 1, ix], phase[iy - 1, ix]
    rg, lf = phase[iy, ix + 1], phase[iy, ix - 1]
    up = np.where(np.abs(up - dw) > np.pi, up - np.sign(up) * 2. * np.pi, up)
    rg = np.where(np.abs(rg - lf) > np.pi, rg - np.sign(rg) * 2. * np.pi, rg)
    g = np.sqrt(((rg - lf) / (2. * dx)) ** 2 + ((up - dw) / (2. * dy)) ** 2)
    grad[iy[inside], ix[inside]] = g[inside]
    gmax = np.nanmax(grad)
    centre = np.argwhere(grad == gmax)[0]

    side = 8

    def ring(i):
        i = int(i) % (4 * side)
        quad, rest = divmod(i, side)
        h = side // 2
        if quad == 0:
            return np.array([rest - h, -(h + 1)])
        if quad == 1:
            return np.array([h + 1, rest - h])
        if quad == 2:
            return np.array([-(rest - h), h + 1])
        return np.array([-(h + 1), -(rest - h)])

    def outside(p):
        return p[0] < 0 or p[0] >= nx or p[1] < 0 or p[1] >= nx

    jumps = 0
    for i in range(4 * side):
        p1, p2 = centre + ring(i), centre + ring(i + 1)
        if outside(p1) and outside(p2):
            break
        if abs(phase[p1[0], p1[1]] - phase[p2[0], p2[1]]) >= 1.7 * np.pi:
            jumps += 1
    if jumps != 1:
        return np.array([np.nan, np.nan])
    sel = np.argwhere(grad[1:ny - 1, 1:nx - 1] >= 0.9 * gmax) + 1
    coords = np.zeros(2)
    coords[0] = np.mean((sel[:, 1] + 0.5) * dx - 0.5 * grid.length_x)
    coords[1] = np.mean((sel[:, 0] + 0.5) * dy - 0.5 * grid.length_y)
    return coords
