"""The sharper limiter next to sides whose DiLevel_EB is not zero
(#RESOLUTIONCHANGE BetaLimiterResChange / nFaceLimiterResChange,
src/ModFaceValue.f90:1982-2015) on the GPU against the oracle: two-level
grids (every limiter that reads beta, 8^3 and 16^3 blocks), outer float and
fixed sides with the 8-wave source, the per-block routines, the adaptive
low order faces, FP32 state, the switches that turn it off, and the
surrogate's refusal."""
import numpy as np
import pytest

import pyoracle as po
from batsrus_amd.amr import AmrGrid, sphere_refinement
from batsrus_amd.core import Mlp, Session, StopMpi
from batsrus_amd.grid import BlockGrid
from batsrus_amd.workloads import init_weights
from helpers import random_mhd_soa, smooth_mhd_soa, rel_err

pytestmark = pytest.mark.gpu

RC = dict(BetaLimiterResChange=1.0, nFaceLimiterResChange=2)
STATES = np.array([[1.0 + 0.1 * s, 0.2, -0.1 * s, 0.05, 0.4, -0.3, 0.1,
                    0.8 + 0.05 * s] for s in range(6)])


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
@pytest.mark.parametrize("limiter", ["mc", "mc3", "beta"])
def test_two_level_grid(limiter, NB):
    print(f"case reschange amr lim={limiter} NB={NB}", flush=True)
    if NB == 8:
        g = AmrGrid((4, 4, 4), sphere_refinement((4, 4, 4), 0.3))
    else:
        ref = np.zeros((2, 2, 2), bool)
        ref[0, 0, 0] = True                 # one refined root of eight
        g = AmrGrid((2, 2, 2), ref, nCell=16)
    assert g.nSlab > 0
    _stages(g, smooth_mhd_soa(g, seed=60 + NB, amp=0.3),
            dict(flux="hlld", limiter=limiter, beta=1.8, cfl=0.8, **RC))


def test_outer_sides_with_divb_source():
    print("case reschange outer powell", flush=True)
    g = BlockGrid(nRoot=(2, 2, 2), nCell=(8, 8, 8),
                  periodic=(False, False, True), order="morton")
    bc = dict(TypeCellBc=["fixed", "float", "float", "inflow", "periodic",
                          "periodic"], CellState_VI=STATES)
    _stages(g, random_mhd_soa(g, seed=61),
            dict(flux="hlld", limiter="mc3", beta=1.9, cfl=0.8,
                 UseDivbSource=True, **RC, **bc))


def test_block_routines_on_a_bounded_grid():
    print("case reschange block routines", flush=True)
    g = BlockGrid(nRoot=(2, 2, 1), nCell=(8, 8, 8),
                  periodic=(False, True, False))
    U0 = random_mhd_soa(g, seed=62)
    scheme = dict(flux="linde", limiter="mc", beta=1.8, **RC)
    cfg = po.config_from_dict(scheme, g)
    S0 = g.soa_to_state_vgb(U0)
    o = po.Session(g, cfg, S0)
    o.exchange()
    with Session(g, **scheme) as s:
        s.upload_soa(U0)
        s.exchange_messages()
        s.set_stage(1)
        for b in range(g.nBlock):
            s.calc_face_value(b + 1)
            fv = s.get_face_values()
            rc = [int(n < 0) for n in g.iNeighbor_EB[b]]
            ofv = po.calc_face_value(cfg, o.State[b], res_change=rc)
            for a, r in zip(fv, ofv):
                assert rel_err(a, r, axis_var=-1) <= 1e-13
            plain = po.calc_face_value(cfg, o.State[b])
            if any(rc):
                assert any(not np.array_equal(a, r) for a, r in
                           zip(ofv, plain))


def test_with_low_order_faces_on_two_levels():
    print("case reschange low order", flush=True)
    g = AmrGrid((4, 4, 4), sphere_refinement((4, 4, 4), 0.3))
    lo = dict(UseAdaptiveLowOrder=True, nLowOrder=1, pCritLow=1.3,
              pCritHigh=1.0, VelCrit=0.0)
    _stages(g, smooth_mhd_soa(g, seed=63, amp=0.3),
            dict(flux="hlld", limiter="mc", beta=1.8, cfl=0.8, **RC, **lo))


def test_fp32_close_to_fp64_oracle():
    print("case reschange fp32", flush=True)
    g = AmrGrid((4, 4, 4), sphere_refinement((4, 4, 4), 0.3))
    U0 = smooth_mhd_soa(g, seed=64, amp=0.3)
    scheme = dict(flux="hlld", limiter="mc3", beta=1.8, cfl=0.8, **RC)
    cfg = po.config_from_dict(scheme, g)
    o = po.Session(g, cfg, g.soa_to_state_vgb(U0))
    o.init_timestep()
    for _ in range(3):
        assert o.advance_step() == 0
    with Session(g, precision="f32", **scheme) as s:
        s.upload_soa(U0)
        s.init_timestep()
        s.advance(3)
        U = s.download_soa()
    ref = o.soa()
    l1 = (np.abs(U - ref).reshape(8, -1).sum(1) /
          np.abs(ref).reshape(8, -1).sum(1))
    assert l1.max() < 1e-5


def test_switches_off_give_the_plain_run():
    """nFace = 0, or LimiterBeta not above BetaLimiterResChange, run the
    plain kernel bit for bit; the active setting differs from it."""
    print("case reschange off", flush=True)
    g = AmrGrid((4, 4, 4), sphere_refinement((4, 4, 4), 0.3))
    U0 = smooth_mhd_soa(g, seed=65, amp=0.3)
    base = dict(flux="hlld", limiter="mc", beta=1.8, cfl=0.8)
    out = []
    for kw in (dict(nFaceLimiterResChange=0, BetaLimiterResChange=1.0),
               dict(BetaLimiterResChange=1.8),
               dict(BetaLimiterResChange=2.0), RC):
        with Session(g, **base, **kw) as s:
            s.upload_soa(U0)
            s.init_timestep()
            s.advance(2)
            out.append(s.download_soa())
    assert np.array_equal(out[0], out[1])
    assert np.array_equal(out[0], out[2])
    assert not np.array_equal(out[0], out[3])


def test_surrogate_refuses_the_sharper_limiter():
    print("case reschange surrogate refused", flush=True)
    g = AmrGrid((4, 4, 4), sphere_refinement((4, 4, 4), 0.3))
    U0 = smooth_mhd_soa(g, seed=66, amp=0.3)
    with Mlp(init_weights(0)) as m:
        with Session(g, precision="f32", flux="hlld", limiter="mc",
                     beta=1.8, **RC) as s:
            s.set_surrogate(m, "rusanov")
            s.upload_soa(U0)
            with pytest.raises(StopMpi) as ei:
                s.init_timestep()
                s.advance(1)
            assert ei.value.code == 1
