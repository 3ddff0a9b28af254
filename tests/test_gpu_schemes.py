"""GPU parity of the time-stepping options against the oracle: SSP-RK3
(three stages, a third state buffer) through the fused stages, two-level
grids, the per-block routines, FP32 state, graphs and persistent grids; the
dB trick's activation rule; the fixed time step's CFL of one; outer
boundary types on bats_init's default grid; the surrogate's refusal of the
8-wave source."""
import numpy as np
import pytest

import pyoracle as po
from batsrus_amd.amr import AmrGrid, sphere_refinement
from batsrus_amd.core import Mlp, Session, StopMpi
from batsrus_amd.grid import BlockGrid
from batsrus_amd.workloads import init_weights
from helpers import random_mhd_soa, smooth_mhd_soa, rel_err

pytestmark = pytest.mark.gpu

RK3 = dict(nStage=3, UseHalfStep=False)


def _stage_parity(grid, U0, scheme, nstep=2, tol=1e-12):
    cfg = po.config_from_dict(scheme, grid)
    o = po.Session(grid, cfg, grid.soa_to_state_vgb(U0))
    o.init_timestep()
    worst = 0.0
    with Session(grid, **scheme) as s:
        s.upload_soa(U0)
        s.init_timestep()
        for step in range(nstep):
            o.s.nStep += 1
            dto = o.set_global_timestep()
            dtg = s.set_global_timestep()
            assert abs(dtg - dto) <= tol * abs(dto), (step, dtg, dto)
            for stage in range(1, cfg.nStage + 1):
                assert o.advance_stage(stage) == 0
                if stage == cfg.nStage:
                    o.exchange()
                s.advance_stage(stage)
                e = rel_err(s.download_soa(), o.soa())
                worst = max(worst, e)
                assert e <= tol, (step, stage, e)
        s.synchronize()
        _, t, _ = s.time()
        assert abs(t - o.tSimulation) <= 1e-14 * max(o.tSimulation, 1e-300)
    print(f"  worst stage difference {worst:.2e}")


@pytest.mark.parametrize("NB,flux,limiter,powell", [
    (8, "rusanov", "mc", False), (16, "hlld", "mc3", False),
    (8, "linde", "minmod", True), (8, "hlld", "mc", True)])
def test_rk3_fused_stages(NB, flux, limiter, powell):
    print(f"case rk3 fused NB={NB} flux={flux} lim={limiter} powell={powell}",
          flush=True)
    g = BlockGrid(nRoot=(2, 2, 2) if NB == 8 else (2, 2, 1),
                  nCell=(NB, NB, NB), order="morton")
    _stage_parity(g, random_mhd_soa(g, seed=100 + NB),
                  dict(flux=flux, limiter=limiter, beta=1.5, cfl=0.8,
                       UseDivbSource=powell, **RK3))


def test_rk3_two_level_and_bounded():
    print("case rk3 two-level bounded", flush=True)
    g = AmrGrid((4, 4, 4), sphere_refinement((4, 4, 4), 0.3))
    _stage_parity(g, smooth_mhd_soa(g, seed=101, amp=0.3),
                  dict(flux="hlld", limiter="mc", beta=1.5, cfl=0.8, **RK3))
    b = BlockGrid(nRoot=(2, 2, 2), nCell=(8, 8, 8),
                  periodic=(False, True, False))
    _stage_parity(b, random_mhd_soa(b, seed=102),
                  dict(flux="hlld", limiter="mc3", beta=1.5, cfl=0.8,
                       TypeCellBc=["reflect", "outflow", "periodic",
                                   "periodic", "float", "float"], **RK3))


def test_rk3_block_routines():
    """The per-block routines in advance_explicit's order with three
    stages (State updated in place, StateOld saved at stage 1)."""
    print("case rk3 block routines", flush=True)
    g = BlockGrid(nRoot=(2, 2, 1), nCell=(8, 8, 8))
    U0 = random_mhd_soa(g, seed=103)
    scheme = dict(flux="hlld", limiter="mc3", beta=1.5, cfl=0.8, **RK3)
    cfg = po.config_from_dict(scheme, g)
    o = po.Session(g, cfg, g.soa_to_state_vgb(U0))
    o.init_timestep()
    with Session(g, **scheme) as s:
        s.upload_soa(U0)
        s.init_timestep()
        for step in range(2):
            o.s.nStep += 1
            o.set_global_timestep()
            s.set_global_timestep()
            for stage in (1, 2, 3):
                assert o.advance_stage(stage) == 0
                s.exchange_messages()
                s.set_stage(stage)
                for b in range(g.nBlock):
                    s.calc_face_value(b + 1)
                    s.calc_face_flux(False, b + 1)
                    s.update_state(b + 1)
                    if stage == 3:
                        s.calc_timestep(b + 1)
                e = rel_err(s.download_soa(), o.soa())
                assert e <= 1e-12, (step, stage, e)
            o.exchange()


def test_rk3_fp32_graph_persistent():
    """FP32 state stays close to the FP64 oracle; graph replays and a
    persistent grid give the eager bits."""
    print("case rk3 fp32 graph persistent", flush=True)
    g = BlockGrid(nRoot=(2, 2, 2), nCell=(16, 16, 16), order="morton")
    U0 = smooth_mhd_soa(g, seed=104, amp=0.3)
    scheme = dict(flux="hlld", limiter="mc3", beta=1.5, cfl=0.8, **RK3)
    cfg = po.config_from_dict(scheme, g)
    o = po.Session(g, cfg, g.soa_to_state_vgb(U0))
    o.init_timestep()
    for _ in range(3):
        assert o.advance_step() == 0
    out = []
    for prec, nwg, graph in (("f64", 0, False), ("f64", 3, False),
                             ("f64", 0, True), ("f32", 0, False)):
        with Session(g, precision=prec, **scheme) as s:
            s.upload_soa(U0)
            s.init_timestep()
            s.set_persistent_grid(nwg)
            s.set_step_graph(graph)
            s.advance(3)
            out.append(s.download_soa())
    assert rel_err(out[0], o.soa()) <= 1e-12
    assert np.array_equal(out[1], out[0]) and np.array_equal(out[2], out[0])
    ref = o.soa()
    l1 = (np.abs(out[3] - ref).reshape(8, -1).sum(1) /
          np.abs(ref).reshape(8, -1).sum(1))
    assert l1.max() < 1e-5


# ----------------------------------------------------------------------
# dB trick activation (src/ModSetParameters.f90:3903-3906,
# src/ModUpdateState.f90:916-918)
@pytest.mark.parametrize("scheme,applies", [
    (dict(nStage=2, UseHalfStep=False), False),          # SSP-RK2
    (dict(nStage=1), False),                             # one stage
    (dict(nStage=1, IsSteadyState=True), True),          # steady, one stage
    (dict(nStage=2, IsSteadyState=True), True),          # steady, stage 1
    (dict(nStage=2), True)])                             # half step
def test_dbtrick_activation(scheme, applies):
    print(f"case dbtrick {scheme}", flush=True)
    g = BlockGrid(nRoot=(2, 2, 2), nCell=(8, 8, 8))
    U0 = random_mhd_soa(g, seed=105)
    full = dict(flux="hlld", limiter="mc", beta=1.5, cfl=0.8, **scheme)
    runs = []
    for trick in (True, False):
        sch = dict(full, UseDbTrick=trick)
        cfg = po.config_from_dict(sch, g)
        o = po.Session(g, cfg, g.soa_to_state_vgb(U0))
        o.init_timestep()
        with Session(g, **sch) as s:
            s.upload_soa(U0)
            s.init_timestep()
            for step in range(2):
                assert o.advance_step() == 0
                s.advance(1)
                e = rel_err(s.download_soa(), o.soa())
                assert e <= 1e-12, (trick, step, e)
            runs.append(s.download_soa())
    assert np.array_equal(runs[0], runs[1]) != applies


def test_dt_fixed_runs_with_cfl_one():
    """#FIXEDTIMESTEP forces Cfl = 1 (src/ModSetParameters.f90:4314-4321):
    the step is DtFixed whatever Cfl the caller passed."""
    print("case dt fixed", flush=True)
    g = BlockGrid(nRoot=(2, 2, 2), nCell=(8, 8, 8))
    U0 = smooth_mhd_soa(g, seed=106, amp=0.2)
    with Session(g, flux="rusanov", cfl=0.8, UseDtFixed=True,
                 DtFixed=1e-3) as s:
        s.upload_soa(U0)
        s.init_timestep()
        s.advance(3)
        dt, t, n = s.time()
    assert n == 3 and dt == 1e-3 and abs(t - 3e-3) < 1e-17


def test_boundary_types_on_default_grid():
    """bats_set_cell_boundary applies to bats_init's default grid (each
    block a domain of unit cells with six outer sides): a run there equals
    the run on the same grid given explicitly; periodic types need a grid
    with neighbours."""
    print("case default grid boundary", flush=True)
    g = BlockGrid(nRoot=(2, 1, 1), nCell=(8, 8, 8))
    g.iNeighbor_EB[:] = -1
    g.CellSize_DB[:] = 1.0
    U0 = smooth_mhd_soa(g, seed=107, amp=0.2)
    types = ["reflect", "outflow", "float", "reflect", "float", "float"]
    scheme = dict(flux="hlld", limiter="mc", beta=1.5, cfl=0.8,
                  UseOutflowPressure=True, pOutflow=0.9)
    out = []
    for default in (True, False):
        with Session(g, default_grid=default, **scheme) as s:
            s.set_cell_boundary(types)
            s.upload_soa(U0)
            s.init_timestep()
            s.advance(2)
            out.append(s.download_soa())
    assert np.array_equal(out[0], out[1])
    with Session(g, default_grid=True, **scheme) as s:
        with pytest.raises(StopMpi) as ei:
            s.set_cell_boundary(["periodic"] * 6)
        assert ei.value.code == 1


def test_surrogate_refuses_divb_source():
    print("case surrogate divb refused", flush=True)
    g = BlockGrid(nRoot=(2, 2, 2), nCell=(8, 8, 8))
    with Mlp(init_weights(0)) as m:
        with Session(g, precision="f32", flux="hlld",
                     UseDivbSource=True) as s:
            for mode in ("network", "rusanov"):
                with pytest.raises(StopMpi) as ei:
                    s.set_surrogate(m, mode)
                assert ei.value.code == 1
