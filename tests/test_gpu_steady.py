"""Steady-state stepping (#TIMEACCURATE F) on the GPU against the oracle:
each cell advances with its own CFL step, set from the first stage's
fluxes; fused stages (two-stage HLLD, one-stage Rusanov, two-level grid,
bounded grid), the per-block routines in the reference's order, FP32
state, graph and persistent launches, and the surrogate's refusal."""
import numpy as np
import pytest

import pyoracle as po
from batsrus_amd.amr import AmrGrid, sphere_refinement
from batsrus_amd.core import Mlp, Session, StopMpi
from batsrus_amd.grid import BlockGrid
from batsrus_amd.workloads import init_weights
from helpers import random_mhd_soa, smooth_mhd_soa, rel_err

pytestmark = pytest.mark.gpu

ST = dict(IsSteadyState=True)


def _steps(grid, U0, scheme, nstep=3, tol=1e-12):
    cfg = po.config_from_dict(scheme, grid)
    o = po.Session(grid, cfg, grid.soa_to_state_vgb(U0))
    o.init_timestep()
    with Session(grid, **scheme) as s:
        s.upload_soa(U0)
        s.init_timestep()
        for step in range(nstep):
            o.s.nStep += 1
            s.set_global_timestep()     # counts the step; Dt is not used
            for stage in range(1, cfg.nStage + 1):
                assert o.advance_stage(stage) == 0
                if stage == cfg.nStage:
                    o.exchange()
                s.advance_stage(stage)
                e = rel_err(s.download_soa(), o.soa())
                assert e <= tol, (step, stage, e)
        _, t, _ = s.time()
        assert t == 0.0


@pytest.mark.parametrize("flux,limiter,nStage", [("hlld", "mc3", 2),
                                                 ("rusanov", "mc", 1),
                                                 ("linde", "minmod", 2)])
def test_fused_stages(flux, limiter, nStage):
    print(f"case steady flux={flux} nStage={nStage}", flush=True)
    g = BlockGrid(nRoot=(2, 2, 2), nCell=(8, 8, 8), order="morton")
    _steps(g, smooth_mhd_soa(g, seed=80, amp=0.3),
           dict(flux=flux, limiter=limiter, beta=1.5, cfl=0.8,
                nStage=nStage, **ST))


def test_two_level_grid_and_bounded_grid():
    print("case steady amr bounded", flush=True)
    g = AmrGrid((4, 4, 4), sphere_refinement((4, 4, 4), 0.3))
    _steps(g, smooth_mhd_soa(g, seed=81, amp=0.3),
           dict(flux="hlld", limiter="mc", beta=1.5, cfl=0.8, **ST))
    b = BlockGrid(nRoot=(2, 2, 2), nCell=(16, 16, 16),
                  periodic=(False, True, False))
    _steps(b, random_mhd_soa(b, seed=82),
           dict(flux="hlld", limiter="mc3", beta=1.5, cfl=0.8,
                TypeCellBc=["reflect", "outflow", "periodic", "periodic",
                            "float", "float"], **ST), nstep=2)


def test_block_routines_in_the_reference_order():
    """calc_timestep before update_state at the first stage."""
    print("case steady block routines", flush=True)
    g = BlockGrid(nRoot=(2, 2, 1), nCell=(8, 8, 8))
    U0 = random_mhd_soa(g, seed=83)
    scheme = dict(flux="hlld", limiter="mc3", beta=1.5, cfl=0.8, **ST)
    cfg = po.config_from_dict(scheme, g)
    o = po.Session(g, cfg, g.soa_to_state_vgb(U0))
    o.init_timestep()
    with Session(g, **scheme) as s:
        s.upload_soa(U0)
        for step in range(2):
            o.s.nStep += 1
            s.set_global_timestep()
            for stage in (1, 2):
                assert o.advance_stage(stage) == 0
                s.exchange_messages()
                s.set_stage(stage)
                for b in range(g.nBlock):
                    s.calc_face_value(b + 1)
                    s.calc_face_flux(False, b + 1)
                    if stage == 1:
                        s.calc_timestep(b + 1)
                    s.update_state(b + 1)
                e = rel_err(s.download_soa(), o.soa())
                assert e <= 1e-12, (step, stage, e)
            o.exchange()


def test_fp32_close_to_fp64_oracle():
    print("case steady fp32", flush=True)
    g = BlockGrid(nRoot=(2, 2, 2), nCell=(8, 8, 8), order="morton")
    U0 = smooth_mhd_soa(g, seed=84, amp=0.3)
    scheme = dict(flux="hlld", limiter="mc3", beta=1.5, cfl=0.8, **ST)
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


def test_graph_and_persistent_are_bitwise_equal():
    print("case steady graph persistent", flush=True)
    g = BlockGrid(nRoot=(2, 2, 2), nCell=(16, 16, 16), order="morton")
    U0 = smooth_mhd_soa(g, seed=85, amp=0.3)
    scheme = dict(flux="hlld", limiter="mc3", beta=1.5, cfl=0.8, **ST)
    out = []
    for nwg, graph in ((0, False), (3, False), (0, True)):
        with Session(g, **scheme) as s:
            s.upload_soa(U0)
            s.init_timestep()
            s.set_persistent_grid(nwg)
            s.set_step_graph(graph)
            s.advance(3)
            out.append(s.download_soa())
    for U in out[1:]:
        assert np.array_equal(U, out[0])


def test_surrogate_refuses_steady_state():
    print("case steady surrogate refused", flush=True)
    g = BlockGrid(nRoot=(2, 2, 2), nCell=(8, 8, 8))
    U0 = smooth_mhd_soa(g, seed=86, amp=0.3)
    with Mlp(init_weights(0)) as m:
        with Session(g, precision="f32", flux="hlld", **ST) as s:
            s.set_surrogate(m, "rusanov")
            s.upload_soa(U0)
            with pytest.raises(StopMpi) as ei:
                s.init_timestep()
                s.advance(1)
            assert ei.value.code == 1
