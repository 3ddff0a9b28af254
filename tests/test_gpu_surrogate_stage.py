"""The gated surrogate stage on the GPU against its CPU oracle
(oracle/surrogate_oracle.py): the smoothness gate, the Rusanov base flux,
the network correction in the rotated frame and its scaling, the sum of the
fine fluxes at resolution changes and the FP64 stage update, compared one
stage at a time from the same FP32 input state, on a same-level grid and a
two-level grid, with the network and in the Rusanov mode.

The GPU forms face states and fluxes in FP32, the oracle in FP64, and the
network's BF16 inputs then round to neighbouring BF16 values on a few
faces; the bounds below follow from that: per variable the 99.9th
percentile of |GPU - oracle| / max|oracle| stays below 1e-6 and the
largest below 1e-4.  Cells next to a face whose gate quantities lie within
1e-4 (relative) of a threshold, where FP32 and FP64 may decide the gate
differently, are left out and counted.  Plausible mistakes in the
correction (the |j| factor or the face area dropped, the sign or the
rotated slots wrong) break the 99.9th-percentile bound by more than a
factor of ten (a random-init network's corrections are small, so the sign
and slot mistakes move the state by about 1e-4 of its size)."""
import numpy as np
import pytest

import mlp_oracle as mo
import surrogate_oracle as so
from batsrus_amd.amr import AmrGrid, sphere_refinement
from batsrus_amd.core import Mlp, Session
from batsrus_amd.grid import BlockGrid
from helpers import smooth_mhd_soa

pytestmark = pytest.mark.gpu

SCHEME = dict(flux="hlld", limiter="mc", beta=1.5, cfl=0.8, nStage=2)
JUMP_TOL = 0.05
Q_BOUND, MAX_BOUND = 1e-6, 1e-4   # 99.9th percentile and largest


def _errs(Ug, Uo, mask):
    d = np.abs(Ug - Uo).reshape(8, -1)[:, mask.reshape(-1)]
    sc = np.abs(Uo).reshape(8, -1).max(1)[:, None]
    r = d / sc
    return np.quantile(r, 0.999, axis=1).max(), r.max()


def _run(grid, U0, mode, seed, nstep=2):
    params = mo.init_params(seed=seed)
    worst_q, worst_m, left_out = 0.0, 0.0, 0
    with Session(grid, precision="f32", **SCHEME) as s, Mlp(params) as m:
        s.set_surrogate(m, mode, BetaMin=0.0, JumpTol=JUMP_TOL)
        s.upload_soa(U0)
        s.init_timestep()
        for step in range(nstep):
            s.set_global_timestep()
            Dt = s.time()[0]
            Un = s.download_soa()
            for stage in (1, 2):
                Uin = s.download_soa()
                nf0, nn0 = s.surrogate_counts()
                Uo, dtb, nNet, cm = so.surrogate_stage(
                    grid, SCHEME, Uin, Un, stage, Dt, params, mode,
                    0.0, JUMP_TOL)
                s.advance_stage(stage)
                Ug = s.download_soa()
                nf1, nn1 = s.surrogate_counts()
                mask = np.broadcast_to(cm >= 1e-4, Ug.shape[1:])
                q, mx = _errs(Ug, Uo, mask)
                left_out += int((~mask).sum())
                worst_q, worst_m = max(worst_q, q), max(worst_m, mx)
                print(f"  step {step} stage {stage}: network faces "
                      f"GPU {nn1 - nn0} oracle {nNet}; 99.9% {q:.1e} "
                      f"max {mx:.1e}, cells left out {(~mask).sum()}")
                if mode == "network":
                    assert nNet > 0
                    assert abs((nn1 - nn0) - nNet) <= max(
                        2, 6 * int((cm < 1e-4).sum()))
                assert q < Q_BOUND and mx < MAX_BOUND, (step, stage, q, mx)
                if stage == 2:
                    dg = s.get_dtmax_b()
                    assert np.abs(dg - dtb).max() <= 1e-5 * dtb.max()
    assert left_out <= 0.01 * U0[0].size * 2 * nstep
    return worst_q, worst_m


@pytest.mark.parametrize("mode", ["network", "rusanov"])
def test_same_level_grid(mode):
    print(f"case surrogate stage same-level {mode}", flush=True)
    g = BlockGrid(nRoot=(2, 2, 2), nCell=(8, 8, 8), order="morton")
    _run(g, smooth_mhd_soa(g, seed=71, amp=0.2), mode, seed=5)


def test_two_level_grid():
    print("case surrogate stage two-level network", flush=True)
    g = AmrGrid((4, 4, 4), sphere_refinement((4, 4, 4), 0.25))
    _run(g, smooth_mhd_soa(g, seed=72, amp=0.2), "network", seed=6)


def test_assembly_mistakes_are_detected():
    """The GPU stage differs from oracles with a planted mistake in the
    correction by far more than from the correct oracle."""
    print("case surrogate stage planted mistakes", flush=True)
    g = BlockGrid(nRoot=(2, 2, 2), nCell=(8, 8, 8), order="morton")
    U0 = smooth_mhd_soa(g, seed=73, amp=0.2)
    params = mo.init_params(seed=7)
    with Session(g, precision="f32", **SCHEME) as s, Mlp(params) as m:
        s.set_surrogate(m, "network", BetaMin=0.0, JumpTol=JUMP_TOL)
        s.upload_soa(U0)
        s.init_timestep()
        s.set_global_timestep()
        Dt = s.time()[0]
        Uin = s.download_soa()
        s.advance_stage(1)
        Ug = s.download_soa()
    good, _, _, cm = so.surrogate_stage(g, SCHEME, Uin, Uin, 1, Dt, params,
                                        "network", 0.0, JUMP_TOL)
    mask = np.broadcast_to(cm >= 1e-4, Ug.shape[1:])
    q_good, e_good = _errs(Ug, good, mask)
    assert q_good < Q_BOUND and e_good < MAX_BOUND
    for mistake in ("no_jump", "no_area", "sign", "slot"):
        bad, _, _, _ = so.surrogate_stage(g, SCHEME, Uin, Uin, 1, Dt, params,
                                          "network", 0.0, JUMP_TOL,
                                          mistake=mistake)
        q_bad, e_bad = _errs(Ug, bad, mask)
        print(f"  {mistake}: 99.9% {q_bad:.1e} max {e_bad:.1e} (correct "
              f"oracle {q_good:.1e}, {e_good:.1e})")
        assert q_bad > 10 * Q_BOUND and e_bad > 100 * e_good, mistake
