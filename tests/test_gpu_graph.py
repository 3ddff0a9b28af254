"""Step graphs: bats_advance replaying one captured hipGraph per time step
gives the same bits as the eager launches, replay after replay with eager
work (downloads, time queries, per-stage launches) in between, for the two
buffer parities of a one-stage scheme, ensembles with per-run time steps,
two-level grids, the adaptive low order faces, a time limit that changes
and a change of the launch shape that drops the graph."""
import time

import numpy as np
import pytest

from batsrus_amd.amr import AmrGrid, sphere_refinement
from batsrus_amd.core import Session
from batsrus_amd.ensemble import blast_ensemble
from batsrus_amd.grid import BlockGrid
from batsrus_amd import ics
from helpers import smooth_mhd_soa

pytestmark = pytest.mark.gpu


def _pair(grid, U0, runs=None, **scheme):
    ss = []
    for graph in (False, True):
        s = Session(grid, **scheme)
        if runs is not None:
            s.set_runs(runs)
        s.upload_soa(U0)
        s.init_timestep()
        if graph:
            s.set_step_graph(True)
        ss.append(s)
    return ss


def _replays(eager, graph, rounds=4, nstep=2, tlimit=-1.0):
    for r in range(rounds):
        eager.advance(nstep, tlimit)
        graph.advance(nstep, tlimit)
        a, b = eager.download_soa(), graph.download_soa()
        assert np.array_equal(a, b), r
        assert eager.time() == graph.time(), r


@pytest.mark.parametrize("nStage", [1, 2])
@pytest.mark.parametrize("NB", [8, 16])
def test_graph_equals_eager(nStage, NB):
    print(f"case graph nStage={nStage} NB={NB}", flush=True)
    g = BlockGrid(nRoot=(4, 2, 2) if NB == 8 else (2, 2, 1),
                  nCell=(NB, NB, NB), order="morton",
                  periodic=(False, True, True))
    U0 = smooth_mhd_soa(g, seed=5)
    e, gr = _pair(g, U0, flux="hlld", limiter="mc3", beta=1.5,
                  nStage=nStage, cfl=0.8 if nStage == 2 else 0.4)
    try:
        _replays(e, gr, nstep=3)
        # eager per-stage launches on the graph session between replays
        for s in (e, gr):
            s.set_global_timestep()
            for st in range(1, nStage + 1):
                s.advance_stage(st)
        assert np.array_equal(e.download_soa(), gr.download_soa())
        _replays(e, gr, rounds=2)
    finally:
        e.close()
        gr.close()


def test_graph_ensemble_amr_and_low_order():
    print("case graph ensemble/amr/low order", flush=True)
    eg, U0 = blast_ensemble(nRun=4, nb_min=8, nb_max=40, seed=2)
    e, gr = _pair(eg, U0, runs=eg.iBlockFirst_R, flux="hlld", limiter="mc",
                  beta=1.5, cfl=0.8)
    try:
        _replays(e, gr, rounds=3)
        for s in (e, gr):
            s.set_persistent_grid(7)   # drops the graph: recaptured
        _replays(e, gr, rounds=2)
    finally:
        e.close()
        gr.close()
    ag = AmrGrid((4, 4, 4), sphere_refinement((4, 4, 4), 0.3))
    Ua = smooth_mhd_soa(ag, seed=9, amp=0.2)
    e, gr = _pair(ag, Ua, flux="hlld", limiter="mc", beta=1.5, cfl=0.8,
                  UseAdaptiveLowOrder=True, nLowOrder=1, pCritLow=1.3,
                  pCritHigh=1.0, VelCrit=0.0)
    try:
        _replays(e, gr, rounds=3)
    finally:
        e.close()
        gr.close()


def test_graph_time_limit_recapture():
    print("case graph time limit", flush=True)
    g = BlockGrid(nRoot=(4, 4, 4), nCell=(8, 8, 8), order="morton",
                  periodic=(False, True, True))
    U0 = ics.brio_wu(g)
    e, gr = _pair(g, U0, flux="rusanov", limiter="mc", beta=1.5, cfl=0.8)
    try:
        _replays(e, gr, rounds=2, nstep=3)
        t = e.time()[1]
        # a stop time inside the next steps: the last step is clipped
        _replays(e, gr, rounds=2, nstep=3, tlimit=t + 0.01)
        assert abs(e.time()[1] - (t + 0.01)) < 1e-14
    finally:
        e.close()
        gr.close()


def test_graph_launch_rate_small_grid():
    """64 blocks of 8^3 (Brio-Wu, Rusanov): the step time with and without
    the step graph (printed; the graph must not be slower by more than
    noise)."""
    g = BlockGrid(nRoot=(4, 4, 4), nCell=(8, 8, 8), order="morton",
                  periodic=(False, True, True))
    U0 = ics.brio_wu(g)
    res = {}
    for graph in (False, True):
        with Session(g, flux="rusanov", limiter="mc", beta=1.5,
                     cfl=0.8) as s:
            s.upload_soa(U0)
            s.init_timestep()
            s.set_step_graph(graph)
            s.advance(20)
            s.synchronize()
            t0 = time.perf_counter()
            s.advance(400)
            s.synchronize()
            res[graph] = (time.perf_counter() - t0) / 400 * 1e6
    print(f"  step time eager {res[False]:.1f} us, graph {res[True]:.1f} us")
    assert res[True] < 1.5 * res[False]
