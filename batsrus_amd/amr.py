"""Two-level block-adaptive grids.

A periodic root grid of nRoot blocks (nCell^3 cells each, level 0) in which
chosen root blocks are split into 8 children of half the cell size (level
1).  With two levels every face of a block meets either one block of the
same level, one coarser block (the block is a child whose parent's neighbour
is unrefined), or four finer blocks (the block is unrefined and its
neighbour is split); the coarse/fine balance holds by construction.

Sides of a block that meet a different level are "resolution-change sides".
Each gets a slab: the two ghost layers of that side, filled before every
stage from the other level, plus, on the coarse side, the fine face fluxes
that replace the coarse face fluxes (the reference's ModConserveFlux).  The
neighbour table encodes a slab as iNeighbor_EB = -2 - slab; -1 stays a
float (zero-gradient) outer side and >= 0 a same-level neighbour.

Slab records (iSlab_IS, 8 ints each):
  [block, side, kind, n0, n1, n2, n3, target]
  kind 0 (coarser neighbour): n0 = the coarse block, (n1, n2) = offset of
         this block's face within the coarse block's face, in coarse cells;
         target = the coarse block's slab on the opposite side
  kind 1 (finer neighbours): n0..n3 = the fine blocks covering the face's
         quadrants (qp + 2*qq, in-plane coordinates halved), target = -1
In-plane coordinates (p, q) of a side: (j, k) on x sides, (i, k) on y sides,
(i, j) on z sides.
"""
from __future__ import annotations

import numpy as np

from .grid import BlockGrid, _morton_key

SLAB_FIELDS = 8


class AmrGrid:
    """Two-level grid with the attributes the sessions read (nI..nBlock,
    iNeighbor_EB, CellSize_DB) plus the slab table iSlab_IS and the block
    levels/positions."""

    def __init__(self, nRoot, refine, nCell=8, xyzMin=(0.0, 0.0, 0.0),
                 xyzMax=(1.0, 1.0, 1.0)):
        self.nRoot = tuple(int(n) for n in nRoot)
        self.nI = self.nJ = self.nK = int(nCell)
        self.nG, self.nDim = 2, 3
        self.xyzMin = np.asarray(xyzMin, float)
        self.xyzMax = np.asarray(xyzMax, float)
        nr = np.array(self.nRoot)
        refine = np.asarray(refine, bool).reshape(self.nRoot[::-1])  # (k,j,i)
        # blocks in Morton order of the root position, children in place
        roots = [(i, j, k) for k in range(nr[2]) for j in range(nr[1])
                 for i in range(nr[0])]
        roots.sort(key=lambda r: int(_morton_key(np.array([r[0]]),
                                                 np.array([r[1]]),
                                                 np.array([r[2]]))[0]))
        lev, pos = [], []           # pos in fine-block units
        for (i, j, k) in roots:
            if refine[k, j, i]:
                for c in range(8):
                    ox, oy, oz = c & 1, (c >> 1) & 1, c >> 2
                    lev.append(1)
                    pos.append((2 * i + ox, 2 * j + oy, 2 * k + oz))
            else:
                lev.append(0)
                pos.append((2 * i, 2 * j, 2 * k))
        self.level = np.array(lev, np.int32)
        self.pos = np.array(pos, np.int64)
        self.nBlock = len(lev)
        self.refine = refine
        fine_index = {}
        coarse_index = {}
        for b in range(self.nBlock):
            if lev[b]:
                fine_index[tuple(pos[b])] = b
            else:
                coarse_index[tuple(np.array(pos[b]) // 2)] = b
        nf = 2 * nr
        nei = np.full((self.nBlock, 6), -1, np.int64)
        slabs = []
        slab_of = {}
        for b in range(self.nBlock):
            for s in range(6):
                d, hi = s // 2, s % 2
                step = np.zeros(3, np.int64)
                step[d] = 1 if hi else -1
                if lev[b] == 1:
                    q = (np.array(pos[b]) + step) % nf
                    if tuple(q) in fine_index:
                        nei[b, s] = fine_index[tuple(q)]
                        continue
                    c = tuple(q // 2)
                    C = coarse_index[c]
                    # offset of this face in C's face, in coarse cells
                    o = np.array(pos[b]) % 2
                    inplane = [a for a in range(3) if a != d]
                    p0 = int(o[inplane[0]]) * (self.nI // 2)
                    q0 = int(o[inplane[1]]) * (self.nI // 2)
                    slab_of[(b, s)] = len(slabs)
                    slabs.append([b, s, 0, C, p0, q0, 0, -1])
                else:
                    rc = (np.array(pos[b]) // 2 + step) % nr
                    if tuple(rc) in coarse_index:
                        nei[b, s] = coarse_index[tuple(rc)]
                        continue
                    # four finer neighbours: the children of root rc facing b
                    inplane = [a for a in range(3) if a != d]
                    kids = [0, 0, 0, 0]
                    for qq in range(2):
                        for qp in range(2):
                            f = 2 * rc
                            f[d] += 0 if hi else 1
                            f[inplane[0]] += qp
                            f[inplane[1]] += qq
                            kids[qp + 2 * qq] = fine_index[tuple(f)]
                    slab_of[(b, s)] = len(slabs)
                    slabs.append([b, s, 1] + kids + [-1])
        # targets: the coarse slab on the side facing a fine block
        for sl in slabs:
            if sl[2] == 0:
                b, s, C = sl[0], sl[1], sl[3]
                sl[7] = slab_of[(C, s ^ 1)]
        for t, sl in enumerate(slabs):
            nei[sl[0], sl[1]] = -2 - t
        self.iNeighbor_EB = np.ascontiguousarray(nei.astype(np.int32))
        self.iSlab_IS = np.ascontiguousarray(
            np.array(slabs, np.int32).reshape(-1, SLAB_FIELDS))
        self.nSlab = len(slabs)
        dxc = (self.xyzMax - self.xyzMin) / (nr * self.nI)
        cs = np.where(self.level[:, None] == 1, dxc[None, :] / 2,
                      dxc[None, :])
        self.CellSize_DB = np.ascontiguousarray(cs)
        self.periodic = (True, True, True)

    @property
    def nCellTotal(self):
        return self.nBlock * self.nI * self.nJ * self.nK

    ghost = BlockGrid.ghost
    soa_to_state_vgb = BlockGrid.soa_to_state_vgb
    state_vgb_to_soa = BlockGrid.state_vgb_to_soa

    def cell_centers(self):
        """Cell-centre coordinates, arrays of shape (nBlock, nK, nJ, nI)."""
        n = self.nI
        dxf = (self.xyzMax - self.xyzMin) / (np.array(self.nRoot) * 2 * n)
        i = np.arange(n) + 0.5
        org = self.xyzMin[None, :] + self.pos * n * dxf[None, :]
        dx = self.CellSize_DB
        x = org[:, 0, None, None, None] + i[None, None, None, :] * dx[:, 0, None, None, None]
        y = org[:, 1, None, None, None] + i[None, None, :, None] * dx[:, 1, None, None, None]
        z = org[:, 2, None, None, None] + i[None, :, None, None] * dx[:, 2, None, None, None]
        shp = (self.nBlock, n, n, n)
        return (np.broadcast_to(x, shp), np.broadcast_to(y, shp),
                np.broadcast_to(z, shp))

    def global_cell_index(self):
        """Index of each cell in the fine-level cell lattice of the box (a
        coarse cell takes the index of its lowest fine sub-cell); seeds the
        random initial perturbations independently of the block order."""
        n = self.nI
        NX, NY = self.nRoot[0] * 2 * n, self.nRoot[1] * 2 * n
        r = np.arange(n)
        sc = np.where(self.level == 1, 1, 2)[:, None, None, None]
        gi = self.pos[:, 0, None, None, None] * n + r[None, None, None, :] * sc
        gj = self.pos[:, 1, None, None, None] * n + r[None, None, :, None] * sc
        gk = self.pos[:, 2, None, None, None] * n + r[None, :, None, None] * sc
        shp = (self.nBlock, n, n, n)
        return np.broadcast_to(gi + NX * (gj + NY * gk), shp).astype(np.uint64)


def sphere_refinement(nRoot, radius=0.25, center=(0.5, 0.5, 0.5)):
    """Root blocks whose centre lies within radius of center (unit box)."""
    nr = np.array(nRoot)
    k, j, i = np.meshgrid(*[np.arange(n) for n in nr[::-1]], indexing="ij")
    c = [(i + 0.5) / nr[0], (j + 0.5) / nr[1], (k + 0.5) / nr[2]]
    r2 = sum((c[d] - center[d]) ** 2 for d in range(3))
    return r2 < radius ** 2
