"""Domain decomposition on one device: the way the path would shard across GPUs.

A doubly periodic N x N lattice is split into 2 x 2 tiles of (N/2 + 8)^2 cells:
each tile is the tile's N/2 x N/2 inner block plus a 4-wide halo, advanced as a
closed lattice, and after every step each halo is refilled from the neighbours'
inner cells (the reference's MPI exchange, src/cpukernel.cpp:691-747, with the
2D decomposition of src/model.cpp:127-157 and src/common.cpp:83-102).  The
assembled inner region must equal the single-lattice periodic evolution: the
halo (4 = the dependency cone of one step) and even tile origins (the
checkerboard parity) make every inner cell see exactly the operations it sees
on the whole lattice.

The CPU test pins the decomposition with the oracle; the GPU test runs the tiles
as separate HIP contexts (one per would-be GPU) against one HIP context.
"""
import dataclasses

import numpy as np
import pytest

from tsmi import workloads as W
from tsmi.lattice import Lattice2D

HALO = 4


def tile_views(n):
    """Origins (in the global tile, halo included) of the 2 x 2 tiles."""
    half = n // 2
    return [(ty * half, tx * half) for ty in (0, 1) for tx in (0, 1)], half + 2 * HALO


def tile_workload(wl, oy, ox, size, psi):
    lat = wl.lattice
    tl = Lattice2D(size, size * lat.delta_x, size, size * lat.delta_y, False, False)
    win = (slice(oy, oy + size), slice(ox, ox + size))
    return dataclasses.replace(wl, name=wl.name + f"_tile{oy}_{ox}", lattice=tl, psi=psi[win].copy(),
                               v=wl.v[win].copy(), steps=1)


def exchange(wl, tiles, origins, size):
    """Inner blocks of the tiles -> global tile, periodic halo wrap -> new tiles."""
    lat = wl.lattice
    n = lat.dim_x
    half = n // 2
    g = np.zeros((lat.H, lat.W), dtype=np.complex128)
    for (oy, ox), t in zip(origins, tiles):
        g[oy + HALO:oy + HALO + half, ox + HALO:ox + HALO + half] = t[HALO:HALO + half, HALO:HALO + half]
    g[:, :HALO] = g[:, n:n + HALO]
    g[:, n + HALO:] = g[:, HALO:2 * HALO]
    g[:HALO, :] = g[n:n + HALO, :]
    g[n + HALO:, :] = g[HALO:2 * HALO, :]
    return g, [g[oy:oy + size, ox:ox + size].copy() for oy, ox in origins]


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


CASES = [dict(n=64, g=2.0, steps=6), dict(n=48, g=0.0, steps=5)]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"n{c['n']}_g{c['g']}_s{c['steps']}")
def test_tiles_match_single_lattice_oracle(case, oracle):
    wl = W.small_case(case["n"], case["n"], True, True, g=case["g"], steps=case["steps"])
    lat = wl.lattice
    ref = oracle.evolve(wl)
    origins, size = tile_views(case["n"])
    tiles = [wl.psi[oy:oy + size, ox:ox + size].copy() for oy, ox in origins]
    for _ in range(case["steps"]):
        tiles = [oracle.evolve(tile_workload(wl, oy, ox, size, wl.psi), steps=1, psi=t)
                 for (oy, ox), t in zip(origins, tiles)]
        g, tiles = exchange(wl, tiles, origins, size)
    inner = (lat.inner_y, lat.inner_x)
    assert rel_l2(g[inner], ref[inner]) < 1e-13


@pytest.mark.gpu
@pytest.mark.parametrize("precision,tol", [(64, 1e-12), (32, 1e-5)])
def test_tiles_match_single_lattice_hip(precision, tol, oracle):
    from tsmi.hip import kernel_for
    wl = W.small_case(64, 64, True, True, g=2.0, steps=8)
    lat = wl.lattice
    e = oracle.exp_potential(wl.v, wl.dt, False)
    k = kernel_for(wl, precision=precision, e_planes=e)
    try:
        k.evolve(wl.steps)
        ref = k.sample()
    finally:
        k.close()
    origins, size = tile_views(64)
    ctxs = []
    try:
        for oy, ox in origins:
            tw = tile_workload(wl, oy, ox, size, wl.psi)
            ctxs.append(kernel_for(tw, precision=precision, e_planes=oracle.exp_potential(tw.v, tw.dt, False)))
        for _ in range(wl.steps):
            for c in ctxs:
                c.evolve(1)
            g, tiles = exchange(wl, [c.sample() for c in ctxs], origins, size)
            for c, t in zip(ctxs, tiles):
                c.set_state(t)
    finally:
        for c in ctxs:
            c.close()
    inner = (lat.inner_y, lat.inner_x)
    assert rel_l2(g[inner], ref[inner]) < tol
