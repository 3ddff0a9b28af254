"""Seeded initial conditions for the benchmark and parity workloads.

Every random number comes from a Philox-4x32-10 counter generator keyed by
(seed, stream) and counted by the global cell index, so the same physical
cell gets the same value whatever the block order or block size, and the CPU
oracle and the GPU receive bit-identical FP64 inputs.

States are returned in the device layout: SoA, interior cells only,
``U[var, block, k, j, i]`` with the reference's variable order
rho, rho*ux, rho*uy, rho*uz, Bx, By, Bz, p (srcEquation/ModEquationMhd.f90:20-35;
the state holds pressure, not energy).
"""
from __future__ import annotations

import math

import numpy as np

from .grid import BlockGrid

__all__ = ["philox4x32_10", "uniform01", "brio_wu", "orszag_tang_z",
           "blast_wave", "kelvin_helmholtz", "uniform_state"]

_M0 = np.uint64(0xD2511F53)
_M1 = np.uint64(0xCD9E8D57)
_W0 = np.uint64(0x9E3779B9)
_W1 = np.uint64(0xBB67AE85)
_MASK = np.uint64(0xFFFFFFFF)


def philox4x32_10(ctr, key):
    """Philox-4x32 with 10 rounds (Salmon et al., SC'11).

    ``ctr``: uint32-valued array (..., 4); ``key``: two uint32 values.
    Returns uint32-valued uint64 array (..., 4)."""
    c = [np.asarray(ctr[..., i], dtype=np.uint64) & _MASK for i in range(4)]
    k0 = np.uint64(int(key[0]) & 0xFFFFFFFF)
    k1 = np.uint64(int(key[1]) & 0xFFFFFFFF)
    for r in range(10):
        if r > 0:
            k0 = (k0 + _W0) & _MASK
            k1 = (k1 + _W1) & _MASK
        p0 = _M0 * c[0]
        p1 = _M1 * c[2]
        hi0, lo0 = p0 >> np.uint64(32), p0 & _MASK
        hi1, lo1 = p1 >> np.uint64(32), p1 & _MASK
        c = [hi1 ^ c[1] ^ k0, lo1, hi0 ^ c[3] ^ k1, lo0]
    return np.stack(c, axis=-1)


def uniform01(index, seed, stream=0):
    """Uniform doubles in [0, 1) for each uint64 ``index`` (53-bit mantissa
    from two 32-bit Philox outputs)."""
    index = np.asarray(index, dtype=np.uint64)
    ctr = np.zeros(index.shape + (4,), dtype=np.uint64)
    ctr[..., 0] = index & _MASK
    ctr[..., 1] = index >> np.uint64(32)
    ctr[..., 2] = np.uint64(stream)
    out = philox4x32_10(ctr, (seed & 0xFFFFFFFF, seed >> 32))
    a = out[..., 0] >> np.uint64(5)
    b = out[..., 1] >> np.uint64(6)
    return (a.astype(np.float64) * 67108864.0 + b.astype(np.float64)) * (
        1.0 / 9007199254740992.0)


def _empty(grid):
    return np.zeros((8, grid.nBlock, grid.nK, grid.nJ, grid.nI),
                    dtype=np.float64)


def uniform_state(grid, rho=1.0, u=(0.0, 0.0, 0.0), b=(0.0, 0.0, 0.0), p=1.0):
    U = _empty(grid)
    U[0] = rho
    for d in range(3):
        U[1 + d] = rho * u[d]
        U[4 + d] = b[d]
    U[7] = p
    return U


def brio_wu(grid: BlockGrid, seed=20261202, noise=1e-3, x0=0.5,
            left=(1.0, 1.0, 1.0), right=(0.125, 0.1, -1.0), bx=0.75,
            axis=0):
    """Brio-Wu shock tube along ``axis`` (configs: left rho=1, p=1, By=1;
    right rho=0.125, p=0.1, By=-1; Bx=0.75) with seeded uniform relative
    density noise rho*(1 + noise*(2r-1))."""
    xyz = grid.cell_centers()
    x = xyz[axis]
    U = _empty(grid)
    isleft = x < x0
    rho = np.where(isleft, left[0], right[0])
    if noise:
        r = uniform01(grid.global_cell_index(), seed, stream=0)
        rho = rho * (1.0 + noise * (2.0 * r - 1.0))
    U[0] = rho
    # normal field along the tube axis, transverse field on the next axis
    U[4 + axis] = bx
    U[4 + (axis + 1) % 3] = np.where(isleft, left[2], right[2])
    U[7] = np.where(isleft, left[1], right[1])
    return U


def orszag_tang_z(grid: BlockGrid, seed=20261203):
    """Orszag-Tang vortex extended in z with seeded per-mode phase shifts:
    rho=25/(36 pi), p=5/(12 pi), u=(-sin(2 pi y+f1), sin(2 pi x+f2),
    0.2 sin(2 pi z+f3)), B=(-sin(2 pi y+f4), sin(4 pi x+f5), 0)/sqrt(4 pi)."""
    x, y, z = grid.cell_centers()
    ph = 2.0 * math.pi * uniform01(np.arange(5, dtype=np.uint64), seed,
                                   stream=7)
    tp = 2.0 * math.pi
    rho = 25.0 / (36.0 * math.pi)
    p = 5.0 / (12.0 * math.pi)
    b0 = 1.0 / math.sqrt(4.0 * math.pi)
    U = _empty(grid)
    U[0] = rho
    U[1] = rho * -np.sin(tp * y + ph[0])
    U[2] = rho * np.sin(tp * x + ph[1])
    U[3] = rho * 0.2 * np.sin(tp * z + ph[2])
    U[4] = -np.sin(tp * y + ph[3]) * b0
    U[5] = np.sin(2.0 * tp * x + ph[4]) * b0
    U[6] = 0.0
    U[7] = p
    return U


def blast_wave(grid: BlockGrid, seed=20261204, run=0, radius=0.1,
               p_in=10.0, p_out=0.1, rho=1.0, center=(0.5, 0.5, 0.5)):
    """Spherical blast wave: rho=1, p=0.1 outside and p=10 inside r<0.1, a
    uniform field with |B| ~ U[0.1, 1] and an isotropic random direction drawn
    from (seed, run)."""
    x, y, z = grid.cell_centers()
    r3 = uniform01(np.arange(3, dtype=np.uint64) + np.uint64(3 * run), seed,
                   stream=11)
    bmag = 0.1 + 0.9 * r3[0]
    cost = 2.0 * r3[1] - 1.0
    sint = math.sqrt(max(0.0, 1.0 - cost * cost))
    phi = 2.0 * math.pi * r3[2]
    bvec = bmag * np.array([sint * math.cos(phi), sint * math.sin(phi), cost])
    rr = np.sqrt((x - center[0]) ** 2 + (y - center[1]) ** 2 +
                 (z - center[2]) ** 2)
    U = _empty(grid)
    U[0] = rho
    for d in range(3):
        U[4 + d] = bvec[d]
    U[7] = np.where(rr < radius, p_in, p_out)
    return U


def kelvin_helmholtz(grid: BlockGrid, seed=20261206, noise=1e-2):
    """Kelvin-Helmholtz shear layer: u_x = 0.5*tanh((|y-0.5|-0.25)/0.02)
    (sign flipped across the layer), rho = 1 inside |y-0.5|<0.25 and 2
    outside, p = 2.5, Bx = 0.1, seeded 1e-2 velocity noise."""
    x, y, z = grid.cell_centers()
    d = np.abs(y - 0.5) - 0.25
    ux = 0.5 * np.tanh(d / 0.02)
    rho = np.where(d < 0.0, 1.0, 2.0)
    gidx = grid.global_cell_index()
    U = _empty(grid)
    U[0] = rho
    vel = [ux, np.zeros_like(ux), np.zeros_like(ux)]
    for c in range(3):
        r = uniform01(gidx, seed, stream=20 + c)
        vel[c] = vel[c] + noise * (2.0 * r - 1.0)
        U[1 + c] = rho * vel[c]
    U[4] = 0.1
    U[7] = 2.5
    return U
