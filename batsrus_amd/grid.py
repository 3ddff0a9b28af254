"""Same-level block grid: block list, neighbour table, cell sizes.

The reference gets its grid from BATL (``srcBATL``, not present): a tree of
equal-shape blocks of nI x nJ x nK cells, each with nG ghost layers, tiling a
box whose sides are periodic or carry outer boundary conditions.  This module
builds the same description for a uniform (single refinement level) grid:

* ``iNeighbor_EB[b, side]`` -- 0-based index of the block across face ``side``
  (x-, x+, y-, y+, z-, z+), or ``-1`` where the face is an outer boundary
  with the "float" (zero-gradient) condition of
  src/ModCellBoundary.f90:518-559.  Periodic sides wrap.
* ``CellSize_DB[b, :]`` -- dx, dy, dz of block ``b`` (BATL ``CellSize_DB``).
* ``iRoot_DB[b, :]`` -- integer root-grid position of block ``b``.

Blocks are numbered either in x-fastest row-major order or along a Morton
(Z-order) curve, the space-filling-curve order BATL uses for load balance
(src/ModLoadBalance.f90:250-619).  Morton order keeps face neighbours close
in the block list, so halo reads of neighbouring blocks hit in cache.
"""
from __future__ import annotations

import numpy as np

__all__ = ["BlockGrid"]


def _morton_key(ix, iy, iz, nbits=11):
    key = np.zeros_like(ix, dtype=np.int64)
    for b in range(nbits):
        key |= ((ix >> b) & 1).astype(np.int64) << (3 * b)
        key |= ((iy >> b) & 1).astype(np.int64) << (3 * b + 1)
        key |= ((iz >> b) & 1).astype(np.int64) << (3 * b + 2)
    return key


class BlockGrid:
    """Uniform block grid on ``[xyzMin, xyzMax]``.

    Parameters mirror the reference's ``#GRID`` command
    (src/ModSetParameters.f90): root block counts, cells per block
    (``Config.pl -g``), ghost layers (``-ng``) and the domain box.
    """

    def __init__(self, nRoot=(4, 4, 4), nCell=(8, 8, 8), xyzMin=(0.0, 0.0, 0.0),
                 xyzMax=(1.0, 1.0, 1.0), periodic=(True, True, True), nDim=3,
                 nG=2, order="linear"):
        self.nRoot = tuple(int(n) for n in nRoot)
        self.nI, self.nJ, self.nK = (int(n) for n in nCell)
        self.nG = int(nG)
        self.nDim = int(nDim)
        if self.nDim < 3 and self.nK != 1:
            raise ValueError("nDim < 3 requires nK == 1")
        if self.nDim < 2 and self.nJ != 1:
            raise ValueError("nDim < 2 requires nJ == 1")
        self.xyzMin = np.asarray(xyzMin, dtype=np.float64)
        self.xyzMax = np.asarray(xyzMax, dtype=np.float64)
        self.periodic = tuple(bool(p) for p in periodic)
        nbx, nby, nbz = self.nRoot
        self.nBlock = nbx * nby * nbz
        ib, jb, kb = np.meshgrid(np.arange(nbx), np.arange(nby), np.arange(nbz),
                                 indexing="ij")
        ib = ib.transpose(2, 1, 0).ravel()
        jb = jb.transpose(2, 1, 0).ravel()
        kb = kb.transpose(2, 1, 0).ravel()
        if order == "morton":
            perm = np.argsort(_morton_key(ib, jb, kb), kind="stable")
            ib, jb, kb = ib[perm], jb[perm], kb[perm]
        elif order != "linear":
            raise ValueError("order must be 'linear' or 'morton'")
        self.order = order
        self.iRoot_DB = np.stack([ib, jb, kb], axis=1).astype(np.int32)
        lookup = -np.ones((nbx, nby, nbz), dtype=np.int64)
        lookup[ib, jb, kb] = np.arange(self.nBlock)
        nei = -np.ones((self.nBlock, 6), dtype=np.int32)
        pos = [ib, jb, kb]
        for d in range(3):
            for hi in (0, 1):
                side = 2 * d + hi
                if d >= self.nDim:
                    continue
                q = [p.copy() for p in pos]
                q[d] = q[d] + (1 if hi else -1)
                n = self.nRoot[d]
                if self.periodic[d]:
                    q[d] %= n
                    nei[:, side] = lookup[q[0], q[1], q[2]]
                else:
                    ok = (q[d] >= 0) & (q[d] < n)
                    nb = -np.ones(self.nBlock, dtype=np.int64)
                    nb[ok] = lookup[q[0][ok], q[1][ok], q[2][ok]]
                    nei[:, side] = nb
        self.iNeighbor_EB = np.ascontiguousarray(nei)
        size = (self.xyzMax - self.xyzMin) / (
            np.asarray(self.nRoot) * np.asarray([self.nI, self.nJ, self.nK]))
        self.CellSize_DB = np.ascontiguousarray(
            np.tile(size, (self.nBlock, 1)).astype(np.float64))

    # ------------------------------------------------------------------
    @property
    def nCellTotal(self):
        return self.nBlock * self.nI * self.nJ * self.nK

    @property
    def ghost(self):
        """Ghost layers per dimension (BATL sets nG only in active dims)."""
        return (self.nG, self.nG if self.nDim >= 2 else 0,
                self.nG if self.nDim >= 3 else 0)

    def cell_centers(self):
        """Cell-centre coordinates, arrays of shape (nBlock, nK, nJ, nI)."""
        dx = self.CellSize_DB[0]
        i = np.arange(self.nI) + 0.5
        j = np.arange(self.nJ) + 0.5
        k = np.arange(self.nK) + 0.5
        org = self.xyzMin[None, :] + self.iRoot_DB * np.asarray(
            [self.nI, self.nJ, self.nK]) * dx[None, :]
        x = org[:, 0, None, None, None] + i[None, None, None, :] * dx[0]
        y = org[:, 1, None, None, None] + j[None, None, :, None] * dx[1]
        z = org[:, 2, None, None, None] + k[None, :, None, None] * dx[2]
        shp = (self.nBlock, self.nK, self.nJ, self.nI)
        return (np.broadcast_to(x, shp), np.broadcast_to(y, shp),
                np.broadcast_to(z, shp))

    def global_cell_index(self):
        """Global (i, j, k) of each cell flattened x-fastest over the whole
        box; used as the counter of the seeded random generator so that the
        same physical cell gets the same random number whatever the block
        order."""
        gi = (self.iRoot_DB[:, 0, None, None, None] * self.nI +
              np.arange(self.nI)[None, None, None, :])
        gj = (self.iRoot_DB[:, 1, None, None, None] * self.nJ +
              np.arange(self.nJ)[None, None, :, None])
        gk = (self.iRoot_DB[:, 2, None, None, None] * self.nK +
              np.arange(self.nK)[None, :, None, None])
        NX = self.nRoot[0] * self.nI
        NY = self.nRoot[1] * self.nJ
        shp = (self.nBlock, self.nK, self.nJ, self.nI)
        return np.broadcast_to(gi + NX * (gj + NY * gk), shp).astype(np.uint64)

    # ------------------------------------------------------------------
    # layout conversions between the device layout (SoA, interior only,
    # U[var, block, k, j, i]) and the reference's State_VGB exchange format
    # (Fortran order, variable fastest, ghosts included).
    def soa_to_state_vgb(self, U):
        """Interior SoA (8, nBlock, nK, nJ, nI) -> ghosted State_VGB
        (nBlock, nKg, nJg, nIg, 8) with ghosts left at zero."""
        gI, gJ, gK = self.ghost
        nv = U.shape[0]
        S = np.zeros((self.nBlock, self.nK + 2 * gK, self.nJ + 2 * gJ,
                      self.nI + 2 * gI, nv), dtype=U.dtype)
        S[:, gK:gK + self.nK, gJ:gJ + self.nJ, gI:gI + self.nI, :] = \
            np.moveaxis(U, 0, -1)
        return S

    def state_vgb_to_soa(self, S):
        gI, gJ, gK = self.ghost
        inner = S[:, gK:gK + self.nK, gJ:gJ + self.nJ, gI:gI + self.nI, :]
        return np.ascontiguousarray(np.moveaxis(inner, -1, 0))

    def fill_ghosts(self, S):
        """Same-level ghost fill (periodic neighbours, float outer sides) of a
        ghosted State_VGB array in place; face slabs only, as the dimension
        split stencil uses no edge or corner cells."""
        gI, gJ, gK = self.ghost
        n = (self.nI, self.nJ, self.nK)
        g = (gI, gJ, gK)
        for d in range(self.nDim):
            ax = 3 - d  # array axis of dimension d in (b, k, j, i, v)
            for hi in (0, 1):
                nb = self.iNeighbor_EB[:, 2 * d + hi]
                for layer in range(1, g[d] + 1):
                    dst = (g[d] + n[d] - 1 + layer) if hi else (g[d] - layer)
                    src_nb = (g[d] + layer - 1) if hi else (g[d] + n[d] - layer)
                    src_own = (g[d] + n[d] - 1) if hi else g[d]
                    sl_dst = [slice(None)] * 5
                    sl_dst[ax] = dst
                    for o in range(3):
                        if o != d:
                            sl_dst[3 - o] = slice(g[o], g[o] + n[o])
                    sl_nb = list(sl_dst)
                    sl_nb[ax] = src_nb
                    sl_own = list(sl_dst)
                    sl_own[ax] = src_own
                    has = nb >= 0
                    tmp_nb = S[tuple(sl_nb)][np.where(has, nb, 0)]
                    tmp_own = S[tuple(sl_own)]
                    val = np.where(has.reshape((-1,) + (1,) * (tmp_nb.ndim - 1)),
                                   tmp_nb, tmp_own)
                    S[tuple(sl_dst)] = val
        return S
