"""Edge cases on the GPU against the oracle: one-block grids with periodic
or float sides, a supersonic flow (the HLLD upwind branch everywhere), the
reference's stop conditions (negative fast speed squared, NaN in the state)
reported as status codes, and a large grid that conserves to rounding."""
import numpy as np
import pytest

import pyoracle as po
from batsrus_amd.core import Session, StopMpi
from batsrus_amd.grid import BlockGrid
from helpers import random_mhd_soa, smooth_mhd_soa, rel_err

pytestmark = pytest.mark.gpu

GAMMA = 5.0 / 3.0


def _stages(grid, U0, scheme, nstep=2, tol=1e-12):
    cfg = po.config_from_dict(scheme, grid)
    o = po.Session(grid, cfg, grid.soa_to_state_vgb(U0))
    o.init_timestep()
    with Session(grid, **scheme) as s:
        s.upload_soa(U0)
        s.init_timestep()
        for step in range(nstep):
            o.s.nStep += 1
            dto = o.set_global_timestep()
            dtg = s.set_global_timestep()
            assert abs(dtg - dto) <= tol * dto
            for stage in (1, 2):
                assert o.advance_stage(stage) == 0
                if stage == 2:
                    o.exchange()
                s.advance_stage(stage)
                e = rel_err(s.download_soa(), o.soa())
                assert e <= tol, (step, stage, e)


@pytest.mark.parametrize("NB", [8, 16])
@pytest.mark.parametrize("periodic", [True, False])
def test_single_block(NB, periodic):
    print(f"case single block NB={NB} periodic={periodic}", flush=True)
    g = BlockGrid(nRoot=(1, 1, 1), nCell=(NB, NB, NB),
                  periodic=(periodic,) * 3)
    _stages(g, random_mhd_soa(g, seed=3 + NB),
            dict(flux="hlld", limiter="mc3", beta=1.5, cfl=0.8))


def test_supersonic_flow():
    """|u_x| several times the fast speed: every x face takes the HLLD
    upwind branch (and the Linde/Rusanov paths the same state)."""
    print("case supersonic", flush=True)
    g = BlockGrid(nRoot=(2, 2, 2), nCell=(8, 8, 8), order="morton")
    U0 = smooth_mhd_soa(g, seed=9, amp=0.05)
    U0[1] = U0[0] * 8.0
    for flux in ("hlld", "linde", "rusanov"):
        _stages(g, U0, dict(flux=flux, limiter="mc", beta=1.5, cfl=0.8))


def test_negative_density_stops_like_the_reference():
    """A negative density gives a negative fast speed squared: the oracle
    stops (ModFaceFlux.f90:4119-4128) and the GPU returns the same code."""
    print("case negative fast speed", flush=True)
    g = BlockGrid(nRoot=(2, 1, 1), nCell=(8, 8, 8))
    U0 = smooth_mhd_soa(g, seed=4)
    U0[0, 1, 3, 4, 5] = -1.0
    U0[4, 1, 3, 4, 5] = 1.0
    scheme = dict(flux="hlld", limiter="mc", beta=1.5)
    cfg = po.config_from_dict(scheme, g)
    o = po.Session(g, cfg, g.soa_to_state_vgb(U0))
    assert o.init_timestep() == 2
    with Session(g, **scheme) as s:
        s.upload_soa(U0)
        with pytest.raises(StopMpi) as ei:
            s.init_timestep()
            s.advance(1)
        assert ei.value.code == 2


def test_nan_in_state_is_reported():
    print("case nan", flush=True)
    g = BlockGrid(nRoot=(2, 1, 1), nCell=(8, 8, 8))
    U0 = smooth_mhd_soa(g, seed=5)
    with Session(g, flux="rusanov") as s:
        s.upload_soa(U0)
        s.check_nan()                      # clean state: no error
        U0[7, 0, 1, 2, 3] = np.nan
        s.upload_soa(U0)
        with pytest.raises(StopMpi) as ei:
            s.check_nan()
        assert ei.value.code == 4


def test_large_grid_conserves():
    """32768 blocks of 8^3 (16.8M cells), HLLD FP64, 3 steps on a periodic
    box: mass, momentum, field and energy totals to rounding."""
    print("case large grid", flush=True)
    g = BlockGrid(nRoot=(32, 32, 32), nCell=(8, 8, 8), order="morton")
    U0 = smooth_mhd_soa(g, seed=6, amp=0.2)
    with Session(g, flux="hlld", limiter="mc3", beta=1.5, cfl=0.8) as s:
        s.upload_soa(U0)
        s.init_timestep()
        s.advance(3)
        U = s.download_soa()

    def totals(U):
        E = U[7] / (GAMMA - 1) + 0.5 * (U[1] ** 2 + U[2] ** 2 + U[3] ** 2) / \
            U[0] + 0.5 * (U[4] ** 2 + U[5] ** 2 + U[6] ** 2)
        t = U.reshape(8, -1).sum(axis=1)
        t[7] = E.sum()
        return t
    t0, t1 = totals(U0), totals(U)
    scale = np.abs(U0).reshape(8, -1).sum(axis=1)
    scale[7] = t0[7]
    err = (np.abs(t1 - t0) / scale).max()
    print(f"  max relative change of the totals {err:.2e}")
    assert err < 1e-12
