"""The inner body on the GPU (bats_set_body): the fused stage with the true
cell mask and the 'ionosphere' face boundary condition of the fast update
(src/ModUpdateStateFast.f90:230-265,343-345,924-937,1277-1327) against the
oracle's stage update with the same body, stage by stage over two steps
within 1e-12, with and without the 8-wave source and B0, first and second
order; body cells keep their state bit for bit and stay out of the block
minimum of the local time step; the captured step graph gives the same
state; the per-block routines carry the body the same way; and the
combinations the body variant does not carry are refused."""
import numpy as np
import pytest

import pyoracle as po
from batsrus_amd.amr import AmrGrid, sphere_refinement
from batsrus_amd.core import Session, StopMpi
from batsrus_amd.grid import BlockGrid
from helpers import body_mask, smooth_mhd_soa, rel_err
from test_gpu_b0 import b0_dgb

pytestmark = pytest.mark.gpu

BODY = dict(BodyRho=3.0, BodyP=2.0)


def dipole(x, y, z, m=(0.0, 0.3, 1.0), c=(0.5, 0.5, 0.5), scale=2e-3):
    """A dipole at the body centre (singular only there, off every cell
    centre of these grids)."""
    rx, ry, rz = x - c[0], y - c[1], z - c[2]
    r2 = rx * rx + ry * ry + rz * rz
    r5 = r2 ** 2.5
    mr = m[0] * rx + m[1] * ry + m[2] * rz
    return ((3 * mr * rx - m[0] * r2) / r5 * scale,
            (3 * mr * ry - m[1] * r2) / r5 * scale,
            (3 * mr * rz - m[2] * r2) / r5 * scale)


def _grid(NB):
    # 8^3 blocks: the body sits inside the middle block of 3^3, and its
    # neighbours reach it only through their ghost cells; 16^3 blocks: the
    # body sits on the corner that all 2^3 blocks share
    if NB == 8:
        g = BlockGrid(nRoot=(3, 3, 3), nCell=(8, 8, 8),
                      periodic=(False, False, False), order="morton")
        r = 0.13
    else:
        g = BlockGrid(nRoot=(2, 2, 2), nCell=(16, 16, 16),
                      periodic=(False, False, False), order="morton")
        r = 0.16
    used, xyz0 = body_mask(g, r)
    return g, used, xyz0


CASES = [
    # flux, limiter, nOrder, powell, NB, B0, B1rCoef
    ("rusanov", "mc", 2, False, 8, False, -1.0),
    ("hlld", "mc3", 2, True, 16, False, -1.0),
    ("linde", "minmod", 2, True, 8, True, 0.0),
    ("hlld", "mc", 2, False, 8, True, -1.0),
    ("hlld", "mc", 1, True, 8, False, 0.5),
]


@pytest.mark.parametrize("flux,limiter,order,powell,NB,b0,b1r", CASES)
def test_fused_stages_with_body_match_oracle(flux, limiter, order, powell,
                                             NB, b0, b1r):
    print(f"case body fused flux={flux} lim={limiter} order={order} "
          f"powell={powell} NB={NB} b0={b0} B1rCoef={b1r}", flush=True)
    grid, used, xyz0 = _grid(NB)
    g = grid.nG
    inner = used[:, g:-g, g:-g, g:-g]
    assert (~inner).any() and inner.any()
    U0 = smooth_mhd_soa(grid, seed=70 + NB, amp=0.2)
    scheme = dict(flux=flux, limiter=limiter, nOrder=order, beta=1.5,
                  cfl=0.8, UseDivbSource=powell)
    cfg = po.config_from_dict(scheme, grid)
    S = grid.soa_to_state_vgb(U0)
    grid.fill_ghosts(S)
    B0 = b0_dgb(grid, dipole) if b0 else None
    body = dict(Used_GB=used, XyzStart_DB=xyz0, B1rCoef=b1r, **BODY)
    o = po.Session(grid, cfg, S, B0_DGB=B0, UseB0Source=b0, **body)
    o.init_timestep()
    with Session(grid, **scheme) as s:
        if b0:
            s.set_b0(B0, UseB0Source=True)
        s.set_body(used, xyz0, B1rCoef=b1r, **BODY)
        s.upload_soa(U0)
        s.init_timestep()
        for step in range(2):
            o.s.nStep += 1
            dto = o.set_global_timestep()
            dtg = s.set_global_timestep()
            assert abs(dtg - dto) <= 1e-12 * dto, (step, dtg, dto)
            for stage in (1, 2):
                assert o.advance_stage(stage) == 0
                if stage == 2:
                    o.exchange()
                s.advance_stage(stage)
                Ug = s.download_soa()
                e = rel_err(Ug, o.soa())
                print(f"  step {step} stage {stage} rel err {e:.2e}")
                assert e <= 1e-12, (step, stage, e)
                # body cells keep their state bit for bit
                np.testing.assert_array_equal(Ug[:, ~inner], U0[:, ~inner])
        # the block minimum of the local time step leaves the body out
        np.testing.assert_allclose(s.get_dtmax_b(), o.DtMax_B, rtol=1e-12)


@pytest.mark.parametrize("flux,limiter,order,powell,b0", [
    ("hlld", "mc3", 2, True, True), ("rusanov", "minmod", 2, False, False),
    ("linde", "mc", 1, True, False)])
def test_per_block_routines_with_body_match_oracle(flux, limiter, order,
                                                   powell, b0):
    """The per-block routines in the reference's loop order carry the body
    as the fused stage does: the face treatment in calc_face_value, no flux
    between two body cells in calc_face_flux, body cells kept by
    update_state and masked by calc_timestep."""
    print(f"case body per-block flux={flux} lim={limiter} order={order} "
          f"powell={powell} b0={b0}", flush=True)
    grid, used, xyz0 = _grid(8)
    g = grid.nG
    inner = used[:, g:-g, g:-g, g:-g]
    U0 = smooth_mhd_soa(grid, seed=77, amp=0.2)
    scheme = dict(flux=flux, limiter=limiter, nOrder=order, beta=1.5,
                  cfl=0.8, UseDivbSource=powell)
    cfg = po.config_from_dict(scheme, grid)
    S = grid.soa_to_state_vgb(U0)
    grid.fill_ghosts(S)
    B0 = b0_dgb(grid, dipole) if b0 else None
    o = po.Session(grid, cfg, S, B0_DGB=B0, UseB0Source=b0, Used_GB=used,
                   XyzStart_DB=xyz0, **BODY)
    o.init_timestep()
    with Session(grid, **scheme) as s:
        if b0:
            s.set_b0(B0, UseB0Source=True)
        s.set_body(used, xyz0, **BODY)
        s.upload_soa(U0)
        s.init_timestep()
        for step in range(2):
            o.s.nStep += 1
            dto = o.set_global_timestep()
            dtg = s.set_global_timestep()
            assert abs(dtg - dto) <= 1e-12 * dto, (step, dtg, dto)
            for stage in (1, 2):
                assert o.advance_stage(stage) == 0
                if stage == 2:
                    o.exchange()
                s.exchange_messages()
                s.set_stage(stage)
                for b in range(1, grid.nBlock + 1):
                    s.calc_face_value(b)
                    s.calc_face_flux(False, b)
                    s.calc_source(b)
                    s.update_state(b)
                    if stage == 2:
                        s.calc_timestep(b)
                Ug = s.download_soa()
                e = rel_err(Ug, o.soa())
                print(f"  step {step} stage {stage} rel err {e:.2e}")
                assert e <= 1e-12, (step, stage, e)
                np.testing.assert_array_equal(Ug[:, ~inner], U0[:, ~inner])
        np.testing.assert_allclose(s.get_dtmax_b(), o.DtMax_B, rtol=1e-12)


def test_steady_state_with_body_and_b0_matches_oracle():
    """Local time stepping (#TIMEACCURATE F) with the body and a dipole B0:
    every true cell with its own step, body cells at rest."""
    print("case body steady", flush=True)
    grid, used, xyz0 = _grid(8)
    g = grid.nG
    inner = used[:, g:-g, g:-g, g:-g]
    U0 = smooth_mhd_soa(grid, seed=79, amp=0.2)
    scheme = dict(flux="hlld", limiter="mc", beta=1.5, cfl=0.8,
                  UseDivbSource=True, IsSteadyState=True)
    cfg = po.config_from_dict(scheme, grid)
    B0 = b0_dgb(grid, dipole)
    body = dict(Used_GB=used, XyzStart_DB=xyz0, **BODY)
    o = po.Session(grid, cfg, grid.soa_to_state_vgb(U0), B0_DGB=B0,
                   UseB0Source=True, **body)
    o.init_timestep()
    with Session(grid, **scheme) as s:
        s.set_b0(B0, UseB0Source=True)
        s.set_body(used, xyz0, **BODY)
        s.upload_soa(U0)
        s.init_timestep()
        for step in range(2):
            o.s.nStep += 1
            s.set_global_timestep()
            for stage in (1, 2):
                assert o.advance_stage(stage) == 0
                if stage == 2:
                    o.exchange()
                s.advance_stage(stage)
                Ug = s.download_soa()
                e = rel_err(Ug, o.soa())
                print(f"  step {step} stage {stage} rel err {e:.2e}")
                assert e <= 1e-12, (step, stage, e)
                np.testing.assert_array_equal(Ug[:, ~inner], U0[:, ~inner])
        cb, _ = s.get_dtmax()
        assert np.all(cb[~inner] == 0) and np.all(cb[inner] > 0)
        np.testing.assert_allclose(cb[inner],
                                   o.DtMax_CB.reshape(inner.shape)[inner],
                                   rtol=1e-12)


def test_body_changes_the_result_and_graph_agrees():
    """Without the body the run differs near it; the captured step graph
    (both launches of each stage) gives the stage-by-stage state."""
    print("case body graph", flush=True)
    grid, used, xyz0 = _grid(8)
    U0 = smooth_mhd_soa(grid, seed=77, amp=0.2)
    scheme = dict(flux="hlld", limiter="mc", beta=1.5, UseDivbSource=True)
    out = []
    for mode in ("stages", "graph", "nobody"):
        with Session(grid, **scheme) as s:
            if mode != "nobody":
                s.set_body(used, xyz0, **BODY)
            s.set_step_graph(mode == "graph")
            s.upload_soa(U0)
            s.init_timestep()
            if mode == "stages":
                for _ in range(3):
                    s.set_global_timestep()
                    s.advance_stage(1)
                    s.advance_stage(2)
            else:
                s.advance(3)
            out.append(s.download_soa())
    assert rel_err(out[1], out[0]) <= 1e-14
    assert rel_err(out[2], out[0]) > 1e-4


def test_body_refusals():
    print("case body refusals", flush=True)
    grid, used, xyz0 = _grid(8)
    with Session(grid, flux="hlld", limiter="mc", precision="f32") as s:
        with pytest.raises(StopMpi):
            s.set_body(used, xyz0, **BODY)
    with Session(grid, flux="hlld", limiter="mc",
                 UseAdaptiveLowOrder=True) as s:
        with pytest.raises(StopMpi):
            s.set_body(used, xyz0, **BODY)
    # the sharper resolution-change limiter on a two-level grid
    ag = AmrGrid((4, 4, 4), sphere_refinement((4, 4, 4), 0.3))
    n = ag.nI + 2 * ag.nG
    au = np.ones((ag.nBlock, n, n, n), bool)
    au[0, 4, 4, 4] = False
    ax = np.zeros((ag.nBlock, 3))
    with Session(ag, flux="hlld", limiter="mc", beta=1.5,
                 BetaLimiterResChange=1.2,
                 nFaceLimiterResChange=2) as s:
        with pytest.raises(StopMpi):
            s.set_body(au, ax, **BODY)
