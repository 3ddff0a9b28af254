"""The TVD treatment at resolution changes on the GPU (UseTvdResChange with
first-order prolongation, the reference's default:
src/ModFaceValue.f90:233-408,680-852,2269-2343): on a two-level grid the
fused stages match the oracle's stage update with the same treatment
within 1e-12 per stage over two steps (Rusanov, Linde, HLLD; MC, Koren,
minmod, beta, with the sharper resolution-change beta, and first order;
with the 8-wave source, on 8^3 and 16^3 blocks; steady state, fused and
per-block),
FP32 state within its rounding, the grid conserves to rounding, the step
graph replays the eager stages, and the combinations the TVD variant does
not carry are refused.  The per-block routines with the TVD faces are
checked in test_gpu_reschange_blocks."""
import numpy as np
import pytest

import pyoracle as po
from batsrus_amd.amr import AmrGrid, sphere_refinement
from batsrus_amd.core import Session, StopMpi
from helpers import smooth_mhd_soa, rel_err, rel_err_vec

pytestmark = pytest.mark.gpu


def _grid():
    return AmrGrid((4, 4, 4), sphere_refinement((4, 4, 4), 0.3))


CASES = [
    dict(flux="hlld", limiter="mc"),
    dict(flux="rusanov", limiter="mc3"),
    dict(flux="linde", limiter="minmod"),
    dict(flux="hlld", limiter="beta", beta=1.8, BetaLimiterResChange=1.2,
         nFaceLimiterResChange=2),
    dict(flux="hlld", limiter="mc", nOrder=1),
    # with the 8-wave source (the reference's default for MHD)
    dict(flux="hlld", limiter="mc3", UseDivbSource=True),
    dict(flux="linde", limiter="beta", beta=1.8, BetaLimiterResChange=1.2,
         nFaceLimiterResChange=2, UseDivbSource=True),
]


def _run_pair(scheme, precision="f64", steps=2, seed=51, check=None,
              grid=None):
    g = grid or _grid()
    U0 = smooth_mhd_soa(g, seed=seed, amp=0.2)
    sch = dict(dict(beta=1.5, cfl=0.8), **scheme, UseTvdResChange=True)
    cfg = po.config_from_dict(sch, g)
    o = po.Session(g, cfg, g.soa_to_state_vgb(U0))
    o.exchange()
    o.init_timestep()
    with Session(g, precision=precision, **sch) as s:
        s.upload_soa(U0)
        s.init_timestep()
        for step in range(steps):
            o.s.nStep += 1
            dto = o.set_global_timestep()
            dtg = s.set_global_timestep()
            for stage in (1, 2):
                assert o.advance_stage(stage) == 0
                if stage == 2:
                    o.exchange()
                s.advance_stage(stage)
                if check:
                    check(step, stage, dtg, dto, s.download_soa(), o.soa())
    return g, U0


@pytest.mark.parametrize("scheme", CASES,
                         ids=lambda d: "-".join(str(v) for v in d.values()))
def test_tvd_stages_match_oracle(scheme):
    print(f"case tvd {scheme}", flush=True)

    def check(step, stage, dtg, dto, Ug, Uo):
        assert abs(dtg - dto) <= 1e-12 * dto, (step, dtg, dto)
        e = rel_err(Ug, Uo)
        print(f"  step {step} stage {stage} rel err {e:.2e}")
        assert e <= 1e-12, (step, stage, e)
    _run_pair(scheme, check=check)


def test_tvd_eight_wave_16cube_matches_oracle():
    print("case tvd 8-wave 16^3", flush=True)
    refine = np.zeros((2, 2, 2), bool)
    refine[0, 0, 0] = True
    refine[1, 1, 0] = True

    def check(step, stage, dtg, dto, Ug, Uo):
        assert abs(dtg - dto) <= 1e-12 * dto, (step, dtg, dto)
        e = rel_err(Ug, Uo)
        print(f"  step {step} stage {stage} rel err {e:.2e}")
        assert e <= 1e-12, (step, stage, e)
    _run_pair(dict(flux="hlld", limiter="mc", UseDivbSource=True),
              check=check, grid=AmrGrid((2, 2, 2), refine, nCell=16))


@pytest.mark.parametrize("per_block", [False, True])
def test_tvd_steady_state_matches_oracle(per_block):
    """Steady state (per-cell steps) with the TVD faces and the 8-wave
    source, fused and through the per-block routines in the reference's
    order (calc_timestep before update_state at the first stage)."""
    print(f"case tvd steady per_block={per_block}", flush=True)
    g = _grid()
    U0 = smooth_mhd_soa(g, seed=57, amp=0.2)
    sch = dict(flux="hlld", limiter="mc", beta=1.5, cfl=0.8,
               UseDivbSource=True, UseTvdResChange=True, IsSteadyState=True)
    o = po.Session(g, po.config_from_dict(sch, g), g.soa_to_state_vgb(U0))
    o.exchange()
    o.init_timestep()
    from test_gpu_reschange_blocks import _fine_blocks
    with Session(g, **sch) as s:
        s.upload_soa(U0)
        s.init_timestep()
        for step in range(2):
            o.s.nStep += 1
            s.set_global_timestep()
            for stage in (1, 2):
                assert o.advance_stage(stage) == 0
                if stage == 2:
                    o.exchange()
                if per_block:
                    s.exchange_messages()
                    s.set_stage(stage)
                    for b in _fine_blocks(g):
                        s.calc_face_value(b, True, True)
                        s.calc_face_flux(True, b)
                        s.save_cons_flux(b)
                    s.message_pass_face()
                    for b in range(1, g.nBlock + 1):
                        s.calc_face_value(b, False, True)
                        s.calc_face_flux(False, b)
                        s.apply_cons_flux(b)
                        s.calc_source(b)
                        if stage == 1:
                            s.calc_timestep(b)
                        s.update_state(b)
                else:
                    s.advance_stage(stage)
                e = rel_err(s.download_soa(), o.soa())
                print(f"  step {step} stage {stage} rel err {e:.2e}")
                assert e <= 1e-12, (step, stage, e)


def test_tvd_fp32_state_matches_oracle_within_rounding():
    print("case tvd fp32", flush=True)

    def check(step, stage, dtg, dto, Ug, Uo):
        # scaled by the largest magnitude of each variable group
        e = rel_err_vec(Ug, Uo)
        print(f"  step {step} stage {stage} rel err {e:.2e}")
        assert e <= 1e-5, (step, stage, e)
    _run_pair(dict(flux="hlld", limiter="mc"), precision="f32", check=check)


def _totals(g, U):
    V = np.prod(g.CellSize_DB, axis=1)[None, :, None, None, None]
    W = U.copy()
    W[7] = U[7] / (5 / 3 - 1) + 0.5 * (U[1] ** 2 + U[2] ** 2 + U[3] ** 2) / \
        U[0] + 0.5 * (U[4] ** 2 + U[5] ** 2 + U[6] ** 2)
    return (W * V).reshape(8, -1).sum(1), np.abs(W * V).reshape(8, -1).sum(1)


def test_tvd_conserves_and_graph_replays():
    print("case tvd conservation and graph", flush=True)
    g = _grid()
    U0 = smooth_mhd_soa(g, seed=53, amp=0.2)
    sch = dict(flux="hlld", limiter="mc3", beta=1.5, UseTvdResChange=True)
    out = []
    for graph in (False, True):
        with Session(g, **sch) as s:
            s.set_step_graph(graph)
            s.upload_soa(U0)
            s.init_timestep()
            s.advance(6)
            out.append(s.download_soa())
    assert np.array_equal(out[0], out[1])
    t0, a0 = _totals(g, U0)
    t1, _ = _totals(g, out[0])
    assert np.all(np.abs(t1 - t0) <= 1e-13 * a0), (t1 - t0) / a0
    # and the default prolongation gives another result
    with Session(g, flux="hlld", limiter="mc3", beta=1.5) as s:
        s.upload_soa(U0)
        s.init_timestep()
        s.advance(6)
        assert rel_err(s.download_soa(), out[0]) > 1e-8


def test_tvd_refusals():
    print("case tvd refusals", flush=True)
    g = _grid()
    U0 = smooth_mhd_soa(g, seed=54)
    for bad in (dict(UseAdaptiveLowOrder=True),):
        with Session(g, flux="hlld", limiter="mc", UseTvdResChange=True,
                     **bad) as s:
            s.upload_soa(U0)
            with pytest.raises(StopMpi):
                s.init_timestep()
    # nor do the per-block routines with low order faces
    with Session(g, flux="hlld", limiter="mc", UseTvdResChange=True,
                 UseAdaptiveLowOrder=True) as s:
        s.upload_soa(U0)
        s.exchange_messages()
        s.set_stage(1)
        with pytest.raises(StopMpi):
            s.calc_face_value(1)
