"""Adaptive low order faces (#ADAPTIVELOWORDER, nLowOrder = 1) on the GPU
against the oracle: the face weights per block, the weighted per-block face
values, the fused stage kernel stage by stage (with and without the 8-wave
source, 8^3 and 16^3 blocks), a two-level grid, the low order everywhere
limit (the first order run) and the refused surrogate combination.

Tolerance (FP64): weights 1e-14 absolute (they are ratios in [0, 1]);
faces 1e-13 and stages 1e-12 relative per variable, as the other parity
tests.
"""
import numpy as np
import pytest

import pyoracle as po
from batsrus_amd.amr import AmrGrid, sphere_refinement
from batsrus_amd.core import Session, StopMpi
from batsrus_amd.grid import BlockGrid
from helpers import random_mhd_soa, smooth_mhd_soa, rel_err

pytestmark = pytest.mark.gpu

LO = dict(UseAdaptiveLowOrder=True, nLowOrder=1, pCritLow=2.0,
          pCritHigh=1.5, VelCrit=1.0)


def _shocked(grid, seed):
    """Smooth state with a converging flow and a pressure step across
    x = 0.5: faces near the step get weights 0 < w <= 1."""
    U = smooth_mhd_soa(grid, seed=seed, amp=0.2)
    x = grid.cell_centers()[0]
    U[1] = U[1] + U[0] * 2.0 * np.sign(0.5 - x) * (np.abs(x - 0.5) < 0.3)
    U[7] = U[7] * np.where(np.abs(x - 0.5) < 0.15, 2.5, 1.0)
    return U


def _ghosted(grid, U):
    S = grid.soa_to_state_vgb(U)
    grid.fill_ghosts(S)
    return S


@pytest.mark.parametrize("NB", [8, 16])
def test_weights_and_face_values_per_block(NB):
    print(f"case low_order block NB={NB}", flush=True)
    grid = BlockGrid(nRoot=(2, 2, 1) if NB == 16 else (2, 2, 2),
                     nCell=(NB, NB, NB), periodic=(True, False, True))
    U0 = random_mhd_soa(grid, seed=30 + NB)
    scheme = dict(flux="hlld", limiter="mc3", beta=1.5, **LO)
    cfg = po.config_from_dict(scheme, grid)
    S0 = _ghosted(grid, U0)
    nz = nfrac = 0
    with Session(grid, **scheme) as s:
        s.upload_soa(U0)
        s.exchange_messages()
        s.set_low_order_face()
        s.set_stage(1)
        for b in range(grid.nBlock):
            crit = po.set_low_order_face(cfg, S0[b])
            got = s.get_low_order_crit(b + 1)
            for a, r in zip(got, crit):
                assert np.abs(a - r).max() <= 1e-14
            nz += sum(int((r > 0).sum()) for r in crit)
            nfrac += sum(int(((r > 0) & (r < 1)).sum()) for r in crit)
            s.calc_face_value(b + 1)
            fv = s.get_face_values()
            ofv = po.calc_face_value(cfg, S0[b], crit)
            for a, r in zip(fv, ofv):
                assert rel_err(a, r, axis_var=-1) <= 1e-13
    print(f"  faces with weight > 0: {nz}, fractional: {nfrac}")
    assert nz > 0 and nfrac > 0


def _fused(grid, U0, scheme, tol=1e-12):
    cfg = po.config_from_dict(scheme, grid)
    o = po.Session(grid, cfg, grid.soa_to_state_vgb(U0))
    o.exchange()
    o.init_timestep()
    with Session(grid, **scheme) as s:
        s.upload_soa(U0)
        s.init_timestep()
        for step in range(2):
            o.s.nStep += 1
            dto = o.set_global_timestep()
            dtg = s.set_global_timestep()
            assert abs(dtg - dto) <= tol * dto, (step, dtg, dto)
            for stage in (1, 2):
                assert o.advance_stage(stage) == 0
                if stage == 2:
                    o.exchange()
                s.advance_stage(stage)
                e = rel_err(s.download_soa(), o.soa())
                print(f"  step {step} stage {stage} rel err {e:.2e}")
                assert e <= tol, (step, stage, e)
        crit = s.get_low_order_crit(1)
    return crit


@pytest.mark.parametrize("NB", [8, 16])
@pytest.mark.parametrize("flux,limiter,powell", [
    ("hlld", "mc3", False), ("rusanov", "mc", False), ("linde", "minmod", False),
    ("hlld", "mc", True)])
def test_fused_stages(flux, limiter, powell, NB):
    print(f"case low_order fused flux={flux} lim={limiter} powell={powell} "
          f"NB={NB}", flush=True)
    grid = BlockGrid(nRoot=(3, 2, 2) if NB == 8 else (2, 2, 2),
                     nCell=(NB, NB, NB), periodic=(False, True, True),
                     order="morton")
    U0 = _shocked(grid, seed=50 + NB)
    scheme = dict(flux=flux, limiter=limiter, beta=1.5, cfl=0.8,
                  UseDivbSource=powell, **LO)
    _fused(grid, U0, scheme)


def test_two_level_grid():
    print("case low_order amr", flush=True)
    g = AmrGrid((4, 4, 4), sphere_refinement((4, 4, 4), 0.3))
    U0 = _shocked(g, seed=61)
    _fused(g, U0, dict(flux="hlld", limiter="mc", beta=1.5, cfl=0.8, **LO))


def test_low_order_everywhere_is_the_first_order_run():
    """pCritLow = pCritHigh = 1, VelCrit = 0: every face is low order and
    the fused kernel reproduces its own first order run (to rounding: the
    two kernel variants are compiled separately and may contract
    multiply-adds differently)."""
    print("case low_order everywhere", flush=True)
    grid = BlockGrid(nRoot=(2, 2, 2), nCell=(8, 8, 8), order="morton")
    U0 = random_mhd_soa(grid, seed=77)
    base = dict(flux="hlld", limiter="mc", beta=1.5, cfl=0.8)
    out = []
    for sch in (dict(base, **dict(LO, pCritLow=1.0, pCritHigh=1.0,
                                  VelCrit=0.0)),
                dict(base, nOrder=1)):
        with Session(grid, **sch) as s:
            s.upload_soa(U0)
            s.init_timestep()
            s.advance(3)
            out.append(s.download_soa())
    e = rel_err(out[0], out[1])
    print(f"  rel err {e:.2e}")
    assert e <= 1e-12


def test_fp32_state_close_to_fp64_oracle():
    print("case low_order fp32", flush=True)
    grid = BlockGrid(nRoot=(2, 2, 2), nCell=(8, 8, 8), order="morton")
    U0 = _shocked(grid, seed=83)
    scheme = dict(flux="hlld", limiter="mc", beta=1.5, cfl=0.8, **LO)
    cfg = po.config_from_dict(scheme, grid)
    o = po.Session(grid, cfg, grid.soa_to_state_vgb(U0))
    o.exchange()
    o.init_timestep()
    for _ in range(3):
        assert o.advance_step() == 0
    with Session(grid, precision="f32", **scheme) as s:
        s.upload_soa(U0)
        s.init_timestep()
        s.advance(3)
        U = s.download_soa()
    ref = o.soa()
    l1 = (np.abs(U - ref).reshape(8, -1).sum(1) /
          np.abs(ref).reshape(8, -1).sum(1))
    print(f"  rel L1 {np.array2string(l1, precision=2)}")
    assert l1.max() < 1e-5


def test_surrogate_with_low_order_is_refused():
    from batsrus_amd.core import Mlp
    from batsrus_amd.workloads import init_weights
    grid = BlockGrid(nRoot=(1, 1, 1), nCell=(8, 8, 8))
    U0 = smooth_mhd_soa(grid, seed=1)
    with Session(grid, precision="f32", flux="hlld", **LO) as s, \
            Mlp(init_weights(0)) as m:
        s.upload_soa(U0)
        with pytest.raises(StopMpi):
            s.set_surrogate(m, "rusanov")
        s.init_timestep()
        s.advance(1)
