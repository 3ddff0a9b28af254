"""Ensemble construction: run offsets, neighbour links confined to runs,
seeded block counts in range."""
import numpy as np

from batsrus_amd.ensemble import blast_ensemble, EnsembleGrid, root_shape
from batsrus_amd.grid import BlockGrid


def test_blast_ensemble_layout():
    eg, U = blast_ensemble(nRun=6, nb_min=8, nb_max=64, seed=7)
    assert eg.nRun == 6 and U.shape == (8, eg.nBlock, 8, 8, 8)
    f = eg.iBlockFirst_R
    assert f[0] == 0 and f[-1] == eg.nBlock and np.all(np.diff(f) > 0)
    for r in range(eg.nRun):
        sl = eg.run_slice(r)
        n = eg.iNeighbor_EB[sl]
        assert np.all((n >= sl.start) & (n < sl.stop))  # periodic, in-run
        assert 1 <= sl.stop - sl.start <= 2 * 64
    # deterministic in the seed, different across seeds
    eg2, U2 = blast_ensemble(nRun=6, nb_min=8, nb_max=64, seed=7)
    assert np.array_equal(U, U2)
    eg3, _ = blast_ensemble(nRun=6, nb_min=8, nb_max=64, seed=8)
    assert not np.array_equal(eg3.iBlockFirst_R, eg.iBlockFirst_R)


def test_root_shape_near_cubic():
    for nb in (8, 27, 64, 100, 655, 4096):
        nx, ny, nz = root_shape(nb)
        assert abs(nx * ny * nz - nb) / nb < 0.5
        assert max(nx, ny, nz) <= 2 * min(nx, ny, nz) + 1


def test_ensemble_ghost_fill_matches_runs():
    """Ghost fill on the concatenation equals the fill of each run alone."""
    g1 = BlockGrid(nRoot=(2, 1, 1), nCell=(8, 8, 8))
    g2 = BlockGrid(nRoot=(1, 2, 2), nCell=(8, 8, 8))
    eg = EnsembleGrid([g1, g2])
    rng = np.random.default_rng(0)
    U = rng.random((8, eg.nBlock, 8, 8, 8))
    S = eg.soa_to_state_vgb(U)
    eg.fill_ghosts(S)
    for r, g in enumerate(eg.runs):
        Sr = g.soa_to_state_vgb(U[:, eg.run_slice(r)])
        g.fill_ghosts(Sr)
        assert np.array_equal(S[eg.run_slice(r)], Sr)
