"""Two-level grids on the GPU: ghost slabs at resolution changes and the
coarse-face flux correction against the CPU oracle, stage by stage, and
conservation with the network in the loop."""
import numpy as np
import pytest

import mlp_oracle as mo
import pyoracle as po
from batsrus_amd.amr import AmrGrid, sphere_refinement
from batsrus_amd.core import Session, Mlp
from helpers import smooth_mhd_soa, rel_err

pytestmark = pytest.mark.gpu


def _grid():
    return AmrGrid((4, 4, 4), sphere_refinement((4, 4, 4), 0.3))


@pytest.mark.parametrize("flux,limiter,powell", [
    ("rusanov", "mc", False), ("hlld", "mc3", False),
    ("linde", "minmod", False), ("hlld", "mc", True), ("rusanov", "mc3", True)])
def test_two_level_stages_match_oracle(flux, limiter, powell):
    """With the 8-wave source the coarse cells next to a resolution change
    take the fine faces' normal field at that face."""
    print(f"case amr flux={flux} lim={limiter} powell={powell}", flush=True)
    g = _grid()
    U0 = smooth_mhd_soa(g, seed=71, amp=0.2)
    scheme = dict(flux=flux, limiter=limiter, beta=1.5, cfl=0.8,
                  UseDivbSource=powell)
    cfg = po.config_from_dict(scheme, g)
    S = g.soa_to_state_vgb(U0)
    o = po.Session(g, cfg, S)
    o.exchange()
    o.init_timestep()
    with Session(g, **scheme) as s:
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
                e = rel_err(s.download_soa(), o.soa())
                print(f"  step {step} stage {stage} rel err {e:.2e}")
                assert e <= 1e-12, (step, stage, e)


def _totals(g, U):
    V = np.prod(g.CellSize_DB, axis=1)[None, :, None, None, None]
    W = U.copy()
    W[7] = U[7] / (5 / 3 - 1) + 0.5 * (U[1] ** 2 + U[2] ** 2 + U[3] ** 2) / \
        U[0] + 0.5 * (U[4] ** 2 + U[5] ** 2 + U[6] ** 2)
    return (W * V).reshape(8, -1).sum(1), np.abs(W * V).reshape(8, -1).sum(1)


@pytest.mark.parametrize("precision,mode", [("f64", None), ("f32", None),
                                            ("f32", "network")])
def test_two_level_conservation(precision, mode):
    print(f"case amr conservation {precision} {mode}", flush=True)
    g = _grid()
    U0 = smooth_mhd_soa(g, seed=72, amp=0.1)
    with Session(g, flux="hlld", limiter="mc", beta=1.5,
                 cfl=0.8 if mode is None else 0.05,
                 precision=precision) as s, Mlp(mo.init_params(seed=5)) as m:
        if mode:
            s.set_surrogate(m, mode, BetaMin=0.0)
        s.upload_soa(U0)
        s.init_timestep()
        s.advance(1 if mode else 4)
        U = s.download_soa()
    t0, sc = _totals(g, U0)
    t1, _ = _totals(g, U)
    tol = 1e-13 if precision == "f64" else 2e-6
    assert (np.abs(t1 - t0) / sc)[:8].max() < tol
