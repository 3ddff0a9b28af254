"""Pins the oracle's steady-state stepping (#TIMEACCURATE F): every cell
advances with its own CFL step DtMax_CB taken from the first stage's
fluxes (src/ModAdvanceExplicit.f90:184-195, src/ModUpdateState.f90:332).
Against the time-accurate update from the same state, each cell's change
is scaled by exactly DtMax_CB / min(DtMax_CB): a one-stage step pins the
per-cell step, a two-stage step pins that stage 2 reuses stage 1's steps
(the stage-2 fluxes come from a different state, so only stage 1 is
compared there)."""
import numpy as np

import pyoracle as po
from batsrus_amd.grid import BlockGrid
from helpers import smooth_mhd_soa

NC = 8


def _pair(nStage, flux="rusanov", limiter="mc"):
    g = BlockGrid(nRoot=(2, 2, 1), nCell=(NC,) * 3)
    U0 = smooth_mhd_soa(g, seed=71, amp=0.4)
    kw = dict(nI=NC, nJ=NC, nK=NC, nStage=nStage, flux=flux,
              limiter=limiter, beta=1.5, cfl=0.8, pMin=-1.0)
    out = []
    for steady in (False, True):
        cfg = po.make_config(IsSteadyState=steady, **kw)
        o = po.Session(g, cfg, g.soa_to_state_vgb(U0))
        o.init_timestep()
        dtcb = o.DtMax_CB.copy()            # local steps of U0
        o.s.nStep += 1
        if not steady:
            o.set_global_timestep()
        assert o.advance_stage(1) == 0
        out.append((o, dtcb))
    return g, U0, out


def test_one_stage_steady_update_scales_per_cell():
    g, U0, ((ta, dtcb), (st, _)) = _pair(nStage=1)
    dmin = dtcb.min()
    ratio = (dtcb / dmin).reshape(g.nBlock, NC, NC, NC)
    dta = ta.soa() - U0
    dst = st.soa() - U0
    # momentum and field change linearly with the step
    for v in range(1, 7):
        want = dta[v] * ratio
        scale = np.abs(want).max()
        assert np.abs(dst[v] - want).max() <= 1e-12 * scale, v
    # the steps really differ from cell to cell
    assert ratio.max() > 1.05
    # no simulation time passes in steady state
    assert st.tSimulation == 0.0 and ta.tSimulation > 0.0


def test_two_stage_steady_first_stage_and_reuse():
    g, U0, ((ta, dtcb), (st, _)) = _pair(nStage=2, flux="hlld",
                                        limiter="mc3")
    ratio = (dtcb / dtcb.min()).reshape(g.nBlock, NC, NC, NC)
    dta = ta.soa() - U0
    dst = st.soa() - U0
    for v in range(1, 7):
        want = dta[v] * ratio
        assert np.abs(dst[v] - want).max() <= 1e-12 * np.abs(want).max(), v
    # stage 2 keeps stage 1's local steps (DtMax_CB is not recomputed)
    before = st.DtMax_CB.copy()
    assert np.array_equal(before, dtcb)
    assert st.advance_stage(2) == 0
    assert np.array_equal(st.DtMax_CB, before)


def test_uniform_state_is_steady():
    g = BlockGrid(nRoot=(2, 1, 1), nCell=(NC,) * 3)
    U = np.array([1.2, 0.3, -0.1, 0.2, 0.4, 0.1, -0.3, 0.9])
    U0 = np.broadcast_to(U[:, None, None, None, None],
                         (8, g.nBlock, NC, NC, NC)).copy()
    cfg = po.make_config(nI=NC, nJ=NC, nK=NC, flux="hlld", limiter="mc3",
                         IsSteadyState=True)
    o = po.Session(g, cfg, g.soa_to_state_vgb(U0))
    o.init_timestep()
    for _ in range(3):
        assert o.advance_step() == 0
    assert np.abs(o.soa() - U0).max() <= 1e-14 * np.abs(U0).max()
