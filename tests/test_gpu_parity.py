"""GPU parity: the gfx950 path through the C ABI against the CPU oracle on
identical FP64 inputs.

Tolerance (FP64): 1e-12 relative per stage, measured per variable as
max|gpu - oracle| / max|oracle|.  The GPU build contracts multiply-adds
into FMAs while the oracle does not, so results differ at the rounding
level, not bit for bit.
"""
import numpy as np
import pytest

import pyoracle as po
from batsrus_amd.core import Session
from batsrus_amd.grid import BlockGrid
from batsrus_amd import ics
from helpers import random_mhd_soa, smooth_mhd_soa, rel_err, rel_err_point

pytestmark = pytest.mark.gpu

TOL_STAGE = 1e-12
TOL_POINT = 1e-10   # elementwise, floor 1e-3 of the variable's max
TOL_FACE = 1e-13

SCHEMES = [
    ("rusanov", "mc", 2), ("rusanov", "mc3", 2), ("rusanov", "minmod", 2),
    ("rusanov", "beta", 2), ("rusanov", "no", 2), ("rusanov", "mc", 1),
    ("linde", "mc", 2), ("linde", "mc3", 2),
    ("hlld", "mc", 2), ("hlld", "mc3", 2), ("hlld", "minmod", 2),
    ("hlld", "mc", 1),
]


def _ghosted(grid, U):
    S = grid.soa_to_state_vgb(U)
    grid.fill_ghosts(S)
    return S


@pytest.mark.parametrize("NB", [8, 16])
@pytest.mark.parametrize("flux,limiter,nOrder", SCHEMES)
def test_block_routines(flux, limiter, nOrder, NB):
    """calc_face_value / calc_face_flux / update_state / calc_timestep per
    block against the oracle's per-block routines."""
    print(f"case block_routines flux={flux} lim={limiter} order={nOrder} "
          f"NB={NB}", flush=True)
    grid = BlockGrid(nRoot=(2, 2, 1) if NB == 16 else (2, 2, 2),
                     nCell=(NB, NB, NB), periodic=(True, False, True))
    U0 = random_mhd_soa(grid, seed=NB + nOrder)
    scheme = dict(flux=flux, limiter=limiter, nOrder=nOrder, beta=1.5)
    cfg = po.config_from_dict(scheme, grid)
    S0 = _ghosted(grid, U0)
    o = po.Session(grid, cfg, S0)
    o.init_timestep()
    o.s.nStep = 1
    dto = o.set_global_timestep()
    dt_unscaled = o.DtMax_CB[0]
    with Session(grid, **scheme) as s:
        s.upload_soa(U0)
        s.init_timestep()
        dtg = s.set_global_timestep()
        assert abs(dtg - dto) <= 1e-13 * abs(dto)
        s.exchange_messages()
        s.set_stage(1)
        S_or = S0.copy()
        Old_or = np.zeros_like(S0)
        for b in range(grid.nBlock):
            s.calc_face_value(b + 1)
            fv = s.get_face_values()
            ofv = po.calc_face_value(cfg, S0[b])
            for a, r in zip(fv, ofv):
                assert rel_err(a, r, axis_var=-1) <= TOL_FACE
            s.calc_face_flux(False, b + 1)
            ff = s.get_face_fluxes()
            err, off = po.calc_face_flux(cfg, grid.CellSize_DB[b], ofv)
            assert err == 0
            for a, r in zip(ff, off):
                sl = [0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 11]
                assert rel_err(a[..., sl], r[..., sl], axis_var=-1) <= TOL_STAGE
            s.update_state(b + 1)
            s.calc_timestep(b + 1)
            dtcb = np.full(grid.nI * grid.nJ * grid.nK, dt_unscaled)
            blk = np.ascontiguousarray(S_or[b])
            old = np.ascontiguousarray(Old_or[b])
            po.update_state(cfg, 1, dto, grid.CellSize_DB[b], dtcb, off, blk,
                            old)
            S_or[b] = blk
        Sg = s.download_soa()
        cb, bb = s.get_dtmax()
    assert rel_err(Sg, grid.state_vgb_to_soa(S_or)) <= TOL_STAGE
    # calc_timestep on the stage-input fluxes
    for b in range(grid.nBlock):
        ofv = po.calc_face_value(cfg, S0[b])
        _, off = po.calc_face_flux(cfg, grid.CellSize_DB[b], ofv)
        dcb, db = po.calc_timestep(cfg, grid.CellSize_DB[b], off)
        assert abs(bb[b] - db) <= 1e-13 * db
        assert np.max(np.abs(cb[b] - dcb) / dcb) <= 1e-13


FUSED = [("rusanov", "mc", 2), ("rusanov", "mc3", 2), ("rusanov", "minmod", 2),
         ("rusanov", "mc", 1), ("linde", "mc3", 2), ("hlld", "mc", 2),
         ("hlld", "mc3", 2), ("hlld", "beta", 2), ("hlld", "no", 2)]


@pytest.mark.parametrize("NB", [8, 16])
@pytest.mark.parametrize("flux,limiter,nOrder", FUSED)
def test_fused_stages(flux, limiter, nOrder, NB):
    """Fused stage kernel, stage by stage over two time steps (midpoint
    scheme), against the oracle's advance_explicit stages."""
    print(f"case fused flux={flux} lim={limiter} order={nOrder} NB={NB}",
          flush=True)
    grid = BlockGrid(nRoot=(3, 2, 2) if NB == 8 else (2, 2, 2),
                     nCell=(NB, NB, NB), periodic=(False, True, True),
                     order="morton")
    U0 = smooth_mhd_soa(grid, seed=NB)
    scheme = dict(flux=flux, limiter=limiter, nOrder=nOrder, beta=1.5,
                  cfl=0.8)
    cfg = po.config_from_dict(scheme, grid)
    o = po.Session(grid, cfg, _ghosted(grid, U0))
    o.init_timestep()
    with Session(grid, **scheme) as s:
        s.upload_soa(U0)
        s.init_timestep()
        for step in range(2):
            o.s.nStep += 1
            dto = o.set_global_timestep()
            dtg = s.set_global_timestep()
            assert abs(dtg - dto) <= TOL_STAGE * dto, (step, dtg, dto)
            for stage in (1, 2):
                assert o.advance_stage(stage) == 0
                if stage == 2:
                    o.exchange()   # BATS_advance's exchange after the step
                s.advance_stage(stage)
                Ug, Uo = s.download_soa(), o.soa()
                e = rel_err(Ug, Uo)
                # and element by element (elements below 1e-3 of their
                # variable's largest value against that 1e-3)
                ep = rel_err_point(Ug, Uo)
                print(f"  step {step} stage {stage} rel err {e:.2e}, "
                      f"pointwise {ep:.2e}")
                assert e <= TOL_STAGE, (step, stage, e)
                assert ep <= TOL_POINT, (step, stage, ep)
        s.synchronize()
        _, t, n = s.time()
        assert n == 2
        assert abs(t - o.tSimulation) <= 1e-14 * o.tSimulation


@pytest.mark.parametrize("scheme", [
    dict(flux="rusanov", limiter="mc", nStage=2, UseHalfStep=False),
    dict(flux="hlld", limiter="mc3", nStage=1),
    dict(flux="linde", limiter="minmod", nStage=2, UseDbTrick=True),
    dict(flux="hlld", limiter="mc", nStage=2, pMin=0.9),
])
def test_fused_scheme_options(scheme):
    """SSP-RK2 interpolation, one-stage scheme, the DbTrick energy fix and
    the pressure floor through the fused kernel."""
    print(f"case options {scheme}", flush=True)
    grid = BlockGrid(nRoot=(2, 2, 2), nCell=(8, 8, 8))
    U0 = smooth_mhd_soa(grid, seed=5)
    full = dict(beta=1.5, cfl=0.8)
    full.update(scheme)
    cfg = po.config_from_dict(full, grid)
    o = po.Session(grid, cfg, _ghosted(grid, U0))
    o.init_timestep()
    with Session(grid, **full) as s:
        s.upload_soa(U0)
        s.init_timestep()
        for step in range(3):
            assert o.advance_step() == 0
            s.advance(1)
            e = rel_err(s.download_soa(), o.soa())
            assert e <= TOL_STAGE, (step, e)


def test_first_step_dt_zero():
    """With a stop time the reference's first step has Dt = 0 and leaves the
    state untouched (src/ModTimeStepControl.f90:348-351)."""
    grid = BlockGrid(nRoot=(2, 2, 2), nCell=(8, 8, 8))
    U0 = smooth_mhd_soa(grid, seed=9)
    with Session(grid, flux="hlld", limiter="mc3") as s:
        s.upload_soa(U0)
        s.advance(1, 1.0)
        dt, t, n = s.time()
        assert dt == 0.0 and t == 0.0 and n == 1
        assert np.array_equal(s.download_soa(), U0)
        s.advance(1, 1.0)
        dt, t, n = s.time()
        assert dt > 0.0 and n == 2


def test_download_state_vgb_ghosts():
    """download_state fills the interior and the face ghost slabs the way
    exchange_messages does."""
    grid = BlockGrid(nRoot=(2, 3, 2), nCell=(8, 8, 8),
                     periodic=(True, False, True))
    U0 = random_mhd_soa(grid, seed=3)
    with Session(grid) as s:
        S = grid.soa_to_state_vgb(U0)
        s.upload_state(S)
        Sg = s.download_state()
    ref = _ghosted(grid, U0)
    g = grid.nG
    n = grid.nI
    # interior + face slabs (edges and corners are not defined)
    mask = np.zeros(ref.shape[1:4], bool)
    mask[g:g + n, g:g + n, :] = True
    mask[g:g + n, :, g:g + n] = True
    mask[:, g:g + n, g:g + n] = True
    assert np.array_equal(Sg[:, mask], ref[:, mask])


@pytest.mark.parametrize("NB", [8, 16])
@pytest.mark.parametrize("flux,limiter", [("rusanov", "mc"),
                                          ("hlld", "mc3"),
                                          ("linde", "minmod")])
def test_block_routines_powell(flux, limiter, NB):
    """calc_face_value -> calc_face_flux -> calc_source -> update_state per
    block with the 8-wave div B source, against the oracle's routines."""
    print(f"case block_powell flux={flux} lim={limiter} NB={NB}", flush=True)
    grid = BlockGrid(nRoot=(2, 2, 1) if NB == 16 else (2, 2, 2),
                     nCell=(NB, NB, NB), periodic=(True, False, True))
    U0 = random_mhd_soa(grid, seed=30 + NB)
    scheme = dict(flux=flux, limiter=limiter, beta=1.5, UseDivbSource=True)
    cfg = po.config_from_dict(scheme, grid)
    S0 = _ghosted(grid, U0)
    o = po.Session(grid, cfg, S0)
    o.init_timestep()
    o.s.nStep = 1
    dto = o.set_global_timestep()
    dt_unscaled = o.DtMax_CB[0]
    with Session(grid, **scheme) as s:
        s.upload_soa(U0)
        s.init_timestep()
        s.set_global_timestep()
        s.exchange_messages()
        s.set_stage(1)
        S_or = S0.copy()
        for b in range(grid.nBlock):
            s.calc_face_value(b + 1)
            s.calc_face_flux(False, b + 1)
            with pytest.raises(Exception):
                s.update_state(b + 1)      # calc_source not called yet
            s.calc_source(b + 1)
            s.update_state(b + 1)
            ofv = po.calc_face_value(cfg, S0[b])
            src = po.calc_source(cfg, grid.CellSize_DB[b], S0[b], ofv)
            _, off = po.calc_face_flux(cfg, grid.CellSize_DB[b],
                                       [f.copy() for f in ofv])
            dtcb = np.full(grid.nI * grid.nJ * grid.nK, dt_unscaled)
            blk = np.ascontiguousarray(S_or[b])
            old = np.zeros_like(blk)
            po.update_state(cfg, 1, dto, grid.CellSize_DB[b], dtcb, off, blk,
                            old, source=src)
            S_or[b] = blk
        Sg = s.download_soa()
    assert rel_err(Sg, grid.state_vgb_to_soa(S_or)) <= TOL_STAGE


@pytest.mark.parametrize("NB", [8, 16])
@pytest.mark.parametrize("flux,limiter,nOrder", [
    ("rusanov", "mc", 2), ("linde", "mc3", 2), ("hlld", "mc3", 2),
    ("hlld", "minmod", 1)])
def test_fused_stages_powell(flux, limiter, nOrder, NB):
    """Fused stage kernel with the 8-wave source, stage by stage over two
    steps, against the oracle."""
    print(f"case fused_powell flux={flux} lim={limiter} order={nOrder} "
          f"NB={NB}", flush=True)
    grid = BlockGrid(nRoot=(3, 2, 2) if NB == 8 else (2, 2, 2),
                     nCell=(NB, NB, NB), periodic=(False, True, True),
                     order="morton")
    U0 = smooth_mhd_soa(grid, seed=40 + NB)   # div B != 0 (random phases)
    scheme = dict(flux=flux, limiter=limiter, nOrder=nOrder, beta=1.5,
                  cfl=0.8, UseDivbSource=True)
    cfg = po.config_from_dict(scheme, grid)
    o = po.Session(grid, cfg, _ghosted(grid, U0))
    o.init_timestep()
    with Session(grid, **scheme) as s:
        s.upload_soa(U0)
        s.init_timestep()
        for step in range(2):
            o.s.nStep += 1
            dto = o.set_global_timestep()
            dtg = s.set_global_timestep()
            assert abs(dtg - dto) <= TOL_STAGE * dto, (step, dtg, dto)
            for stage in (1, 2):
                assert o.advance_stage(stage) == 0
                if stage == 2:
                    o.exchange()
                s.advance_stage(stage)
                e = rel_err(s.download_soa(), o.soa())
                print(f"  step {step} stage {stage} rel err {e:.2e}")
                assert e <= TOL_STAGE, (step, stage, e)
