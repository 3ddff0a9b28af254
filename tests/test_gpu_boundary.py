"""Fixed and inflow outer boundaries on the GPU against the oracle: stage
by stage on a same-level grid (HLLD and Rusanov, with the 8-wave source),
the reference-shaped per-block routines, a two-level grid with fixed x
sides, FP32 state, the ghosted download, graph and persistent launches,
and the refused boundary types."""
import numpy as np
import pytest

import pyoracle as po
from batsrus_amd.amr import AmrGrid, sphere_refinement
from batsrus_amd.core import Session, StopMpi
from batsrus_amd.grid import BlockGrid
from helpers import random_mhd_soa, smooth_mhd_soa, rel_err

pytestmark = pytest.mark.gpu


def _cons(rho, u, b, p):
    return [rho, rho * u[0], rho * u[1], rho * u[2], *b, p]


STATES = np.array([_cons(1.0 + 0.1 * s, (0.3 - 0.1 * s, 0.05 * s, -0.2),
                         (0.4, 0.2 * s - 0.3, 0.1), 0.8 + 0.05 * s)
                   for s in range(6)])
BC = dict(TypeCellBc=["fixed", "inflow", "float", "fixed", "periodic",
                      "periodic"], CellState_VI=STATES)


def _grid(NB=8):
    return BlockGrid(nRoot=(2, 2, 2), nCell=(NB,) * 3,
                     periodic=(False, False, True), order="morton")


def _stages(grid, U0, scheme, nstep=2, tol=1e-12, **sess):
    cfg = po.config_from_dict(scheme, grid)
    o = po.Session(grid, cfg, grid.soa_to_state_vgb(U0))
    o.init_timestep()
    with Session(grid, **scheme, **sess) as s:
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
@pytest.mark.parametrize("flux,limiter", [("hlld", "mc3"), ("rusanov", "mc"),
                                          ("linde", "minmod")])
def test_fixed_sides_stage_parity(flux, limiter, NB):
    print(f"case boundary stages flux={flux} lim={limiter} NB={NB}",
          flush=True)
    g = _grid(NB)
    _stages(g, random_mhd_soa(g, seed=40 + NB),
            dict(flux=flux, limiter=limiter, beta=1.5, cfl=0.8, **BC))


def test_fixed_sides_with_divb_source():
    print("case boundary powell", flush=True)
    g = _grid()
    _stages(g, smooth_mhd_soa(g, seed=41, amp=0.2),
            dict(flux="hlld", limiter="mc", beta=1.5, cfl=0.8,
                 UseDivbSource=True, **BC))


def test_fixed_sides_block_routines():
    """The reference-shaped loop (exchange_messages, then per block
    calc_face_value, calc_face_flux, update_state, calc_timestep) on a grid
    with fixed sides against the oracle's whole step."""
    print("case boundary block routines", flush=True)
    g = _grid()
    U0 = random_mhd_soa(g, seed=42)
    scheme = dict(flux="hlld", limiter="mc3", beta=1.5, cfl=0.8, **BC)
    cfg = po.config_from_dict(scheme, g)
    o = po.Session(g, cfg, g.soa_to_state_vgb(U0))
    o.init_timestep()
    with Session(g, **scheme) as s:
        s.upload_soa(U0)
        # the Dt = 0 pass of the first step, per block
        s.exchange_messages()
        s.set_stage(2)
        for b in range(g.nBlock):
            s.calc_face_value(b + 1)
            s.calc_face_flux(False, b + 1)
            s.calc_timestep(b + 1)
        for step in range(2):
            o.s.nStep += 1
            dto = o.set_global_timestep()
            dtg = s.set_global_timestep()
            assert abs(dtg - dto) <= 1e-12 * dto
            for stage in (1, 2):
                assert o.advance_stage(stage) == 0
                s.set_stage(stage)
                for b in range(g.nBlock):
                    s.calc_face_value(b + 1)
                    s.calc_face_flux(False, b + 1)
                    s.update_state(b + 1)
                    if stage == 2:
                        s.calc_timestep(b + 1)
                s.exchange_messages()
                e = rel_err(s.download_soa(), o.soa())
                assert e <= 1e-12, (step, stage, e)
            o.exchange()


def _amr_fixed_x():
    g = AmrGrid((4, 4, 4), sphere_refinement((4, 4, 4), 0.3))
    nei = g.iNeighbor_EB.copy()
    nf = 2 * 4
    for b in range(g.nBlock):
        x = g.pos[b, 0]
        top = nf - 1 if g.level[b] else nf - 2
        if x == 0:
            assert nei[b, 0] >= 0
            nei[b, 0] = -1
        if x == top:
            assert nei[b, 1] >= 0
            nei[b, 1] = -1
    g.iNeighbor_EB = np.ascontiguousarray(nei)
    g.periodic = (False, True, True)
    return g


def test_two_level_grid_with_fixed_sides():
    print("case boundary amr", flush=True)
    g = _amr_fixed_x()
    U0 = smooth_mhd_soa(g, seed=43, amp=0.2)
    bc = dict(TypeCellBc=["fixed", "inflow"] + ["periodic"] * 4,
              CellState_VI=STATES)
    _stages(g, U0, dict(flux="hlld", limiter="mc", beta=1.5, cfl=0.8, **bc))


def test_fp32_fixed_sides_close_to_fp64_oracle():
    print("case boundary fp32", flush=True)
    g = _grid()
    U0 = smooth_mhd_soa(g, seed=44, amp=0.2)
    scheme = dict(flux="hlld", limiter="mc3", beta=1.5, cfl=0.8, **BC)
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
    print(f"  rel L1 {np.array2string(l1, precision=2)}")
    assert l1.max() < 1e-5


def test_download_state_has_the_fixed_ghosts():
    print("case boundary download", flush=True)
    g = _grid()
    U0 = random_mhd_soa(g, seed=45)
    with Session(g, **BC) as s:
        s.upload_soa(U0)
        S = s.download_state()
    cfg = po.config_from_dict(BC, g)
    o = po.Session(g, cfg, g.soa_to_state_vgb(U0))
    o.exchange()
    n, gg = g.nI, g.nG
    mask = np.zeros(S.shape[1:4], bool)
    mask[gg:gg + n, gg:gg + n, :] = True
    mask[gg:gg + n, :, gg:gg + n] = True
    mask[:, gg:gg + n, gg:gg + n] = True
    assert np.array_equal(S[:, mask], o.State[:, mask])


def test_graph_and_persistent_launches_are_bitwise_equal():
    print("case boundary graph persistent", flush=True)
    g = _grid()
    U0 = smooth_mhd_soa(g, seed=46, amp=0.2)
    scheme = dict(flux="hlld", limiter="mc3", beta=1.5, cfl=0.8, **BC)
    out = []
    for nwg, graph in ((0, False), (5, False), (0, True)):
        with Session(g, **scheme) as s:
            s.upload_soa(U0)
            s.init_timestep()
            s.set_persistent_grid(nwg)
            s.set_step_graph(graph)
            s.advance(3)
            out.append(s.download_soa())
    for U in out[1:]:
        assert np.array_equal(U, out[0])


def test_boundary_changes_after_the_grid():
    """Types set after the grid (state reallocated, then uploaded) give the
    same run as types given with the grid; switching back to float gives
    the float run."""
    print("case boundary reapply", flush=True)
    g = _grid()
    U0 = smooth_mhd_soa(g, seed=47, amp=0.2)
    scheme = dict(flux="rusanov", limiter="mc", beta=1.5, cfl=0.8)
    res = []
    for mode in ("with", "after", "float", "float-after"):
        kw = BC if mode == "with" else {}
        with Session(g, **scheme, **kw) as s:
            if mode == "after":
                s.set_cell_boundary(BC["TypeCellBc"], BC["CellState_VI"])
            if mode == "float-after":
                s.set_cell_boundary(BC["TypeCellBc"], BC["CellState_VI"])
                s.set_cell_boundary(["float"] * 6)
            s.upload_soa(U0)
            s.init_timestep()
            s.advance(2)
            res.append(s.download_soa())
    assert np.array_equal(res[0], res[1])
    assert np.array_equal(res[2], res[3])
    assert not np.array_equal(res[0], res[2])


def test_refused_boundary_types():
    print("case boundary refused", flush=True)
    g = _grid()
    with Session(g) as s:
        with pytest.raises(StopMpi) as ei:
            s.set_cell_boundary(["float"] * 5 + [14])     # ShearBC_
        assert ei.value.code == 5
        with pytest.raises(StopMpi) as ei:                # vary: a state
            s.set_cell_boundary(["float"] * 5 + [11])     # is needed
        assert ei.value.code == 1
        with pytest.raises(StopMpi) as ei:
            s.set_cell_boundary(["periodic"] * 6)         # x, y sides are -1
        assert ei.value.code == 1
        bad = STATES.copy()
        bad[0, 0] = np.nan
        with pytest.raises(StopMpi) as ei:
            s.set_cell_boundary(["fixed"] * 6, bad)
        assert ei.value.code == 1
        with pytest.raises(StopMpi) as ei:
            s.set_cell_boundary(["fixed"] * 6)            # no state
        assert ei.value.code == 1


def test_block_routines_across_a_boundary_change():
    """Per-block routines, then a change of the outer boundaries (the state
    and the exchange snapshot are reallocated), a new upload and the
    per-block routines again: each run matches a fresh oracle run."""
    print("case boundary block routines regrid", flush=True)
    g = _grid()
    U0 = random_mhd_soa(g, seed=48)
    float_bc = dict(TypeCellBc=["float"] * 4 + ["periodic"] * 2)
    with Session(g, flux="rusanov", limiter="mc", beta=1.5, cfl=0.8) as s:
        for bc in (float_bc, BC, float_bc):
            s.set_cell_boundary(bc["TypeCellBc"], bc.get("CellState_VI"))
            s.upload_soa(U0)
            s.exchange_messages()
            s.set_stage(2)
            for b in range(g.nBlock):
                s.calc_face_value(b + 1)
                s.calc_face_flux(False, b + 1)
                s.calc_timestep(b + 1)
            dtg = s.set_global_timestep()
            for stage in (1, 2):
                s.set_stage(stage)
                for b in range(g.nBlock):
                    s.calc_face_value(b + 1)
                    s.calc_face_flux(False, b + 1)
                    s.update_state(b + 1)
                    if stage == 2:
                        s.calc_timestep(b + 1)
                s.exchange_messages()
            cfg = po.config_from_dict(dict(flux="rusanov", limiter="mc",
                                           beta=1.5, cfl=0.8, **bc), g)
            o = po.Session(g, cfg, g.soa_to_state_vgb(U0))
            o.init_timestep()
            o.s.nStep += 1
            dto = o.set_global_timestep()
            assert abs(dtg - dto) <= 1e-12 * dto
            assert o.advance_stage(1) == 0
            assert o.advance_stage(2) == 0
            assert rel_err(s.download_soa(), o.soa()) <= 1e-12


def test_surrogate_stage_keeps_a_uniform_flow_at_fixed_sides():
    """FP32 state, the gated surrogate stage (Rusanov and network modes) on
    a grid whose six sides are fixed at the uniform state: the flow stays
    uniform to FP32 rounding, so the surrogate's face pass reads the fixed
    ghost layers."""
    print("case boundary surrogate uniform", flush=True)
    from batsrus_amd.core import Mlp
    from batsrus_amd.workloads import init_weights
    g = BlockGrid(nRoot=(2, 2, 2), nCell=(8, 8, 8), periodic=(False,) * 3)
    U = STATES[2]
    U0 = np.broadcast_to(U[:, None, None, None, None],
                         (8, g.nBlock, 8, 8, 8)).copy()
    bc = dict(TypeCellBc=["fixed"] * 6, CellState_VI=np.tile(U, (6, 1)))
    with Mlp(init_weights(0)) as m:
        for mode in ("rusanov", "network"):
            with Session(g, precision="f32", flux="hlld", limiter="mc",
                         beta=1.5, cfl=0.8, **bc) as s:
                s.set_surrogate(m, mode)
                s.upload_soa(U0)
                s.init_timestep()
                s.advance(3)
                got = s.download_soa()
            err = np.abs(got - U0).max() / np.abs(U0).max()
            assert err < 1e-6, (mode, err)


REFLECT = dict(TypeCellBc=["reflect", "reflect", "fixed", "float",
                           "periodic", "periodic"], CellState_VI=STATES)


@pytest.mark.parametrize("flux,limiter,powell", [("hlld", "mc3", False),
                                                 ("rusanov", "mc", True)])
def test_reflecting_sides_stage_parity(flux, limiter, powell):
    print(f"case boundary reflect flux={flux} powell={powell}", flush=True)
    g = _grid()
    _stages(g, random_mhd_soa(g, seed=49),
            dict(flux=flux, limiter=limiter, beta=1.5, cfl=0.8,
                 UseDivbSource=powell, **REFLECT))


def test_reflecting_walls_equal_a_mirrored_periodic_domain():
    """GPU: [0,1] with reflecting x sides against [0,2] periodic whose
    second half is the mirror image: equal bit for bit after the first
    stage, to rounding after three steps."""
    print("case boundary reflect mirror", flush=True)
    nb = 2
    half = BlockGrid(nRoot=(nb, 1, 1), nCell=(8, 8, 8),
                     xyzMax=(1.0, 1.0 / nb, 1.0 / nb),
                     periodic=(False, True, True))
    full = BlockGrid(nRoot=(2 * nb, 1, 1), nCell=(8, 8, 8),
                     xyzMax=(2.0, 1.0 / nb, 1.0 / nb))
    U = random_mhd_soa(half, seed=50)
    M = U[:, ::-1, :, :, ::-1].copy()
    M[1] *= -1
    M[4] *= -1
    Uf = np.concatenate([U, M], axis=1)
    scheme = dict(flux="hlld", limiter="mc3", beta=1.5, cfl=0.8)
    out = []
    for grid, U0, kw in ((half, U, dict(TypeCellBc=["reflect"] * 2 +
                                        ["periodic"] * 4)),
                         (full, Uf, {})):
        with Session(grid, **scheme, **kw) as s:
            s.upload_soa(U0)
            s.init_timestep()
            s.set_global_timestep()
            s.advance_stage(1)
            first = s.download_soa()
            s.advance_stage(2)
            s.advance(3)
            out.append((first, s.download_soa()))
    assert np.array_equal(out[0][0], out[1][0][:, :nb])
    a, b = out[0][1], out[1][1][:, :nb]
    assert (np.abs(a - b).max(axis=(1, 2, 3, 4)) /
            np.abs(b).max(axis=(1, 2, 3, 4))).max() < 1e-13


def test_reflecting_sides_block_routines():
    print("case boundary reflect block routines", flush=True)
    g = _grid()
    U0 = random_mhd_soa(g, seed=51)
    scheme = dict(flux="linde", limiter="mc", beta=1.5, cfl=0.8, **REFLECT)
    cfg = po.config_from_dict(scheme, g)
    o = po.Session(g, cfg, g.soa_to_state_vgb(U0))
    o.init_timestep()
    with Session(g, **scheme) as s:
        s.upload_soa(U0)
        s.exchange_messages()
        s.set_stage(2)
        for b in range(g.nBlock):
            s.calc_face_value(b + 1)
            s.calc_face_flux(False, b + 1)
            s.calc_timestep(b + 1)
        for step in range(2):
            o.s.nStep += 1
            dto = o.set_global_timestep()
            dtg = s.set_global_timestep()
            assert abs(dtg - dto) <= 1e-12 * dto
            for stage in (1, 2):
                assert o.advance_stage(stage) == 0
                s.set_stage(stage)
                for b in range(g.nBlock):
                    s.calc_face_value(b + 1)
                    s.calc_face_flux(False, b + 1)
                    s.update_state(b + 1)
                    if stage == 2:
                        s.calc_timestep(b + 1)
                s.exchange_messages()
                e = rel_err(s.download_soa(), o.soa())
                assert e <= 1e-12, (step, stage, e)
            o.exchange()
        # the ghosted download holds the mirrored cells
        S = s.download_state()
    o.exchange()
    n, gg = g.nI, g.nG
    mask = np.zeros(S.shape[1:4], bool)
    mask[gg:gg + n, gg:gg + n, :] = True
    mask[gg:gg + n, :, gg:gg + n] = True
    mask[:, gg:gg + n, gg:gg + n] = True
    assert rel_err(S[:, mask], o.State[:, mask], axis_var=-1) <= 1e-12


def test_outflow_pressure_sides_stage_parity():
    print("case boundary outflow pressure", flush=True)
    g = _grid()
    bc = dict(TypeCellBc=["outflow", "outflow", "reflect", "outflow",
                          "periodic", "periodic"], UseOutflowPressure=True,
              pOutflow=0.9)
    _stages(g, random_mhd_soa(g, seed=52),
            dict(flux="hlld", limiter="mc3", beta=1.5, cfl=0.8, **bc))
