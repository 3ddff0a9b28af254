"""Shared helpers for the parity tests."""
import numpy as np

from batsrus_amd.grid import BlockGrid


def random_mhd_soa(grid: BlockGrid, seed=1, rho=(0.5, 1.5), u=1.0, b=1.0,
                   p=(0.3, 1.5)):
    """Cell-wise random but physical MHD state (every HLLD branch and every
    limiter case occurs), SoA layout U[var, block, k, j, i]."""
    rng = np.random.default_rng(seed)
    shp = (grid.nBlock, grid.nK, grid.nJ, grid.nI)
    U = np.zeros((8,) + shp)
    r = rng.uniform(rho[0], rho[1], shp)
    U[0] = r
    for d in range(3):
        U[1 + d] = r * rng.uniform(-u, u, shp)
        U[4 + d] = rng.uniform(-b, b, shp)
    U[7] = rng.uniform(p[0], p[1], shp)
    return U


def smooth_mhd_soa(grid: BlockGrid, seed=2, amp=0.3):
    """Smooth periodic state with random Fourier phases plus small noise."""
    rng = np.random.default_rng(seed)
    x, y, z = grid.cell_centers()
    tp = 2 * np.pi
    ph = rng.uniform(0, tp, 16)
    U = np.zeros((8,) + x.shape)
    rho = 1.0 + amp * np.sin(tp * x + ph[0]) * np.cos(tp * y + ph[1])
    U[0] = rho
    U[1] = rho * amp * np.sin(tp * y + ph[2])
    U[2] = rho * amp * np.sin(tp * z + ph[3])
    U[3] = rho * amp * np.sin(tp * x + ph[4])
    U[4] = 0.5 + amp * np.cos(tp * z + ph[5])
    U[5] = amp * np.sin(tp * x + ph[6])
    U[6] = amp * np.cos(tp * y + ph[7])
    U[7] = 1.0 + amp * np.cos(tp * (x + y + z) + ph[8])
    U += 1e-4 * rng.standard_normal(U.shape) * np.array(
        [1, 1, 1, 1, 1, 1, 1, 1.0])[:, None, None, None, None]
    return U


def rel_err(a, b, axis_var=0):
    """max |a-b| per variable divided by max |b| of that variable."""
    a = np.asarray(a)
    b = np.asarray(b)
    a2 = np.moveaxis(a, axis_var, 0).reshape(a.shape[axis_var], -1)
    b2 = np.moveaxis(b, axis_var, 0).reshape(b.shape[axis_var], -1)
    scale = np.maximum(np.abs(b2).max(axis=1), 1e-300)
    return (np.abs(a2 - b2).max(axis=1) / scale).max()


def rel_err_point(a, b, floor=1e-3, axis_var=0):
    """Elementwise relative error: max over every element of |a-b| divided
    by |b| itself, or by floor times the variable's max |b| where |b| is
    smaller than that (elements that cross zero)."""
    a = np.asarray(a)
    b = np.asarray(b)
    a2 = np.moveaxis(a, axis_var, 0).reshape(a.shape[axis_var], -1)
    b2 = np.moveaxis(b, axis_var, 0).reshape(b.shape[axis_var], -1)
    big = np.abs(b2).max(axis=1, keepdims=True)
    den = np.maximum(np.abs(b2), np.maximum(floor * big, 1e-300))
    return (np.abs(a2 - b2) / den).max()


def rel_err_vec(a, b):
    """max |a-b| per variable divided by the largest |b| of the physical
    quantity the variable belongs to: rho, the momentum vector (rho u_x,
    rho u_y, rho u_z), the field vector (B_x, B_y, B_z), p.  A vector
    component that stays at rounding level (B_z of the Orszag-Tang vortex,
    whose flux divergence cancels exactly) is measured against the
    vector's size, not against its own rounding noise.  SoA arrays
    U[var, ...]."""
    a = np.asarray(a).reshape(8, -1)
    b = np.asarray(b).reshape(8, -1)
    mx = np.abs(b).max(axis=1)
    group = [[0], [1, 2, 3], [1, 2, 3], [1, 2, 3], [4, 5, 6], [4, 5, 6],
             [4, 5, 6], [7]]
    scale = np.array([max(mx[g].max(), 1e-300) for g in group])
    return (np.abs(a - b).max(axis=1) / scale).max()


def block_geometry(grid):
    """Lower corner and cell size of every block, arrays (nBlock, 3), for a
    same-level grid (root positions) or a two-level grid (positions in
    fine-block units)."""
    n = np.array([grid.nI, grid.nJ, grid.nK])
    dx = np.asarray(grid.CellSize_DB, dtype=np.float64).reshape(-1, 3)
    if hasattr(grid, "pos"):
        dxf = (grid.xyzMax - grid.xyzMin) / (np.array(grid.nRoot) * 2 * n)
        org = grid.xyzMin[None, :] + grid.pos * n * dxf[None, :]
    else:
        org = grid.xyzMin[None, :] + grid.iRoot_DB * n[None, :] * dx
    return org, dx


def ghost_axes(grid):
    """Cell-centre coordinates of every block's ghosted cells along x, y and
    z, arrays (nBlock, n + 2 nG)."""
    g = grid.nG
    n = np.array([grid.nI, grid.nJ, grid.nK])
    org, dx = block_geometry(grid)
    return [org[:, d, None] + (np.arange(-g, n[d] + g) + 0.5)[None, :] *
            dx[:, d, None] for d in range(3)]


def b0_dgb(grid, field):
    """B0_DGB (nBlock, nKg, nJg, nIg, 3) of field(x, y, z) at the cell
    centres and ghost cells of every block."""
    ax = ghost_axes(grid)
    x = ax[0][:, None, None, :]
    y = ax[1][:, None, :, None]
    z = ax[2][:, :, None, None]
    shp = (grid.nBlock, ax[2].shape[1], ax[1].shape[1], ax[0].shape[1])
    x, y, z = (np.broadcast_to(a, shp) for a in (x, y, z))
    return np.ascontiguousarray(np.stack(field(x, y, z), axis=-1))


def body_mask(grid, rBody, center=(0.5, 0.5, 0.5)):
    """True cells outside a sphere of radius rBody at `center`, over every
    block and its ghost cells (cell centres of the ghosted blocks), shape
    (nBlock, nKg, nJg, nIg), and each block's first cell centre relative
    to the sphere's centre, shape (nBlock, 3)."""
    c = np.asarray(center, dtype=np.float64)
    ax = [a - c[d] for d, a in enumerate(ghost_axes(grid))]
    x = ax[0][:, None, None, :]
    y = ax[1][:, None, :, None]
    z = ax[2][:, :, None, None]
    used = (x * x + y * y + z * z) > rBody ** 2
    org, dx = block_geometry(grid)
    xyz0 = org + 0.5 * dx - c[None, :]
    return np.ascontiguousarray(used), np.ascontiguousarray(xyz0)

