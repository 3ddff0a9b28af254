"""Ensembles of independent runs on one flat block list.

Each run is its own periodic box of 8^3-cell blocks; the runs' blocks are
concatenated run after run, so one stage launch advances all of them.  The
neighbour table of the concatenation only links blocks of the same run, and
the device keeps a time step, simulation time and step counter per run
(bats_set_runs).
"""
from __future__ import annotations

import math

import numpy as np

from .grid import BlockGrid
from . import ics


def root_shape(nb):
    """Near-cubic root-block grid (nx, ny, nz) with nx*ny*nz close to nb."""
    nx = max(1, int(round(nb ** (1.0 / 3.0))))
    ny = max(1, int(round(math.sqrt(nb / nx))))
    nz = max(1, int(round(nb / (nx * ny))))
    return nx, ny, nz


class EnsembleGrid:
    """Concatenation of independent BlockGrids (one per run) exposing the
    attributes the solver sessions read (nI..nBlock, iNeighbor_EB,
    CellSize_DB) plus the run offsets iBlockFirst_R."""

    def __init__(self, grids):
        g0 = grids[0]
        for g in grids:
            if (g.nI, g.nJ, g.nK, g.nG, g.nDim) != (g0.nI, g0.nJ, g0.nK,
                                                    g0.nG, g0.nDim):
                raise ValueError("runs must share the block shape")
        self.runs = list(grids)
        self.nI, self.nJ, self.nK = g0.nI, g0.nJ, g0.nK
        self.nG, self.nDim = g0.nG, g0.nDim
        counts = [g.nBlock for g in grids]
        self.iBlockFirst_R = np.concatenate(
            [[0], np.cumsum(counts)]).astype(np.int32)
        self.nBlock = int(self.iBlockFirst_R[-1])
        self.nRun = len(grids)
        nei = []
        for r, g in enumerate(grids):
            off = int(self.iBlockFirst_R[r])
            n = g.iNeighbor_EB.astype(np.int64)
            nei.append(np.where(n >= 0, n + off, -1))
        self.iNeighbor_EB = np.ascontiguousarray(
            np.concatenate(nei).astype(np.int32))
        self.CellSize_DB = np.ascontiguousarray(
            np.concatenate([g.CellSize_DB for g in grids]))

    @property
    def nCellTotal(self):
        return self.nBlock * self.nI * self.nJ * self.nK

    ghost = BlockGrid.ghost
    soa_to_state_vgb = BlockGrid.soa_to_state_vgb
    state_vgb_to_soa = BlockGrid.state_vgb_to_soa
    fill_ghosts = BlockGrid.fill_ghosts

    def run_slice(self, r):
        return slice(int(self.iBlockFirst_R[r]), int(self.iBlockFirst_R[r + 1]))

    def split(self, U):
        """Per-run views of an SoA state U[var, block, k, j, i]."""
        return [U[:, self.run_slice(r)] for r in range(self.nRun)]


def blast_ensemble(nRun=48, nb_min=8, nb_max=4096, seed=20261204,
                   nCell=8, order="morton"):
    """The batched blast-wave ensemble: nRun runs whose block counts are
    drawn log-uniformly (seeded Philox) from [nb_min, nb_max]; run r is a
    periodic box of near-cubic root shape with cubic cells, a blast wave of
    radius 0.1 at p=10 in a p=0.1, rho=1 background and a uniform field of
    random magnitude in [0.1, 1] and isotropic random direction.  Returns
    (EnsembleGrid, SoA state)."""
    u = ics.uniform01(np.arange(nRun, dtype=np.uint64), seed, stream=13)
    grids, states = [], []
    for r in range(nRun):
        nb = int(round(math.exp(math.log(nb_min) +
                                u[r] * (math.log(nb_max) - math.log(nb_min)))))
        nx, ny, nz = root_shape(nb)
        m = max(nx, ny, nz)
        ext = (nx / m, ny / m, nz / m)
        g = BlockGrid(nRoot=(nx, ny, nz), nCell=(nCell,) * 3,
                      xyzMax=ext, order=order)
        grids.append(g)
        states.append(ics.blast_wave(g, seed=seed, run=r,
                                     center=tuple(e / 2 for e in ext)))
    eg = EnsembleGrid(grids)
    return eg, np.ascontiguousarray(np.concatenate(states, axis=1))
