"""GPU runs checked against the reference's committed golden output and
against invariants that hold at any size."""
import os

import numpy as np
import pytest

import pyoracle as po
from batsrus_amd.core import Session
from batsrus_amd.grid import BlockGrid
from batsrus_amd import ics
from helpers import rel_err

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def test_gpu_shocktube_golden():
    """Param/SHOCKTUBE/TestOutput/mhdhyp_1d_ref.out (first-order Rusanov, one
    stage, gamma=2, dt=0.05, 512 steps, float ends) reproduced by the fused
    3-D kernel on 128 blocks of 8^3 along x.  The run is uniform in y and z
    (one periodic block across), so the y and z face fluxes of every cell
    cancel exactly and the result is the 1-D solution.  Reference tolerance:
    DiffNum.pl -r=1e-5 -a=1e-11 (Makefile.test:319-322)."""
    print("case gpu_shocktube_golden", flush=True)
    ref = np.load(os.path.join(HERE, "golden", "mhdhyp_1d_ref.npz"))
    grid = BlockGrid(nRoot=(128, 1, 1), nCell=(8, 8, 8),
                     xyzMin=(-128.0, -1.0, -1.0), xyzMax=(128.0, 1.0, 1.0),
                     periodic=(False, True, True))
    U0 = ics.brio_wu(grid, noise=0.0, x0=0.0, left=(1.0, 1.0, 1.0),
                     right=(0.125, 0.1, -1.0), bx=0.75)
    with Session(grid, nOrder=1, flux="rusanov", nStage=1, cfl=1.0,
                 gamma=2.0, UseDtFixed=True, DtFixed=0.05) as s:
        s.upload_soa(U0)
        s.advance(512, 25.6)
        dt, t, n = s.time()
        U = s.download_soa()
    assert n == 512 and abs(t - 25.6) < 1e-12
    # every y,z column must hold the same 1-D solution
    col = U[:, :, 0, 0, :]
    assert np.array_equal(U, np.broadcast_to(col[:, :, None, None, :], U.shape))
    rho = col[0].reshape(-1)
    prim = {"Rho": rho, "Ux": col[1].reshape(-1) / rho,
            "Uy": col[2].reshape(-1) / rho, "Uz": col[3].reshape(-1) / rho,
            "Bx": col[4].reshape(-1), "By": col[5].reshape(-1),
            "Bz": col[6].reshape(-1), "P": col[7].reshape(-1)}
    names = list(ref["names"])
    for nm, arr in prim.items():
        r = ref["data"][:, names.index(nm)]
        err = np.abs(arr - r)
        assert np.all((err <= 1e-11) | (err <= 1e-5 * np.abs(r))), nm
        # the fixture prints 11 significant digits
        assert err.max() <= 1e-10, (nm, err.max())


@pytest.mark.parametrize("flux,limiter,NB", [("rusanov", "mc", 8),
                                             ("hlld", "mc3", 16),
                                             ("linde", "minmod", 8)])
def test_gpu_conservation_periodic(flux, limiter, NB):
    """Periodic box: total mass, momentum, magnetic field and energy are
    conserved to rounding over 20 steps (no Powell source, no DbTrick)."""
    print(f"case conservation {flux} {limiter} NB={NB}", flush=True)
    nr = (4, 4, 4) if NB == 8 else (2, 2, 2)
    grid = BlockGrid(nRoot=nr, nCell=(NB, NB, NB), order="morton")
    U0 = ics.brio_wu(grid)
    with Session(grid, flux=flux, limiter=limiter, beta=1.5, cfl=0.8) as s:
        s.upload_soa(U0)
        s.init_timestep()
        s0 = s.integrate_domain()
        s.advance(20)
        s1 = s.integrate_domain()
        U = s.download_soa()
    V = np.prod(grid.CellSize_DB[0])
    # natural magnitudes: mass, mass x unit speed, |B| volume, energy
    mass = s0[0]
    bmag = np.sqrt((U0[4:7] ** 2).sum(axis=0)).sum() * V
    scale = np.array([mass, mass, mass, mass, bmag, bmag, bmag, s0[7]])
    drift = np.abs(s1 - s0) / scale
    assert drift.max() <= 5e-14, drift
    assert np.all(np.isfinite(U))


def test_gpu_uniform_state_unchanged():
    """A uniform state has zero flux divergence: the state must stay
    bit-identical over several steps (fluxes on both faces of a cell are
    computed from identical inputs)."""
    grid = BlockGrid(nRoot=(2, 2, 2), nCell=(16, 16, 16))
    U0 = ics.uniform_state(grid, rho=1.3, u=(0.2, -0.1, 0.05),
                           b=(0.3, 0.4, -0.2), p=0.7)
    with Session(grid, flux="hlld", limiter="mc3") as s:
        s.upload_soa(U0)
        s.init_timestep()
        s.advance(3)
        U = s.download_soa()
    # p goes through p -> E -> p each stage: equal to rounding only
    assert np.array_equal(U[:7], U0[:7])
    assert np.max(np.abs(U[7] - U0[7])) <= 4e-16 * 0.7 * 8


def test_gpu_axis_permutation_symmetry():
    """The same 1-D Brio-Wu problem along x and along y: the y run equals
    the x run under the cyclic relabelling x->y->z->x of coordinates and
    vector components (rotate_state_vectors, src/ModFaceFlux.f90:4493-4552)."""
    print("case axis_permutation", flush=True)
    scheme = dict(flux="hlld", limiter="mc", beta=1.5)
    prof = []
    for axis in (0, 1):
        # cubic cells: the relabelling is a symmetry of the grid
        grid = BlockGrid(nRoot=(4, 4, 1), nCell=(8, 8, 8),
                         xyzMax=(1.0, 1.0, 0.25))
        U0 = ics.brio_wu(grid, noise=0.0, axis=axis)
        with Session(grid, **scheme) as s:
            s.upload_soa(U0)
            s.init_timestep()
            s.advance(5)
            U = s.download_soa()
        G = U.reshape(8, 1, 4, 4, 8, 8, 8).transpose(0, 1, 4, 2, 5, 3, 6)\
            .reshape(8, 8, 32, 32)            # (v, K, J, I)
        if axis == 0:
            assert np.array_equal(G, np.broadcast_to(G[:, :1, :1, :], G.shape))
            prof.append(G[:, 0, 0, :])
        else:
            assert np.array_equal(G, np.broadcast_to(G[:, :1, :, :1], G.shape))
            prof.append(G[:, 0, :, 0])
    px, py = prof
    comp = [0, 2, 3, 1, 5, 6, 4, 7]
    for v in range(8):
        a, b = py[comp[v]], px[v]
        assert np.max(np.abs(a - b)) <= 1e-13 * max(np.abs(b).max(), 1e-300), v
