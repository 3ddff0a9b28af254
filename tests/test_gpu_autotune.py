"""The launch-shape autotuner: it measures every candidate on a scratch
session, stores the fastest under the workload's key, reloads it without
timing again, and the chosen launch shape gives the same bits as the
default one (within rounding where it splits the blocks' planes).  On
same-level, two-level, body, ensemble, FP32, steady-state and low-order
workloads the XCD-contiguous order and a persistent grid give the
default's state, block time steps and time bit for bit; fewer cells per
thread (NB/2, NB/4) start a z window at each part boundary, whose face
rounds like the block's first face (the compiler contracts the prologue's
arithmetic differently), so the state and time steps agree within a few
ulps of the storage type (1e-14 relative for FP64, 1e-6 for FP32)."""
import os
import tempfile

import numpy as np
import pytest

from batsrus_amd.amr import AmrGrid, sphere_refinement
from batsrus_amd.autotune import (Autotuner, apply, compare_block_size,
                                  markdown)
from batsrus_amd.core import Session
from batsrus_amd.ensemble import blast_ensemble
from batsrus_amd.grid import BlockGrid
from helpers import body_mask, rel_err, smooth_mhd_soa

pytestmark = pytest.mark.gpu

SCHEME = dict(flux="hlld", limiter="mc", beta=1.5, cfl=0.8)


def test_autotune_stores_and_applies():
    print("case autotune ensemble", flush=True)
    eg, U0 = blast_ensemble(nRun=6, nb_min=64, nb_max=400, seed=4)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "tune.json")
        at = Autotuner(path)
        e = at.tune(eg, U0, SCHEME, runs=eg.iBlockFirst_R)
        times = [m["step_us"] for m in e["measured"]]
        print(markdown("blast ensemble", eg, SCHEME, e))
        assert e["step_us"] == min(times)
        assert os.path.exists(path)
        again = Autotuner(path).tune(eg, U0, SCHEME, runs=eg.iBlockFirst_R)
        assert again == e        # stored, not re-measured
        out = []
        for choice in (None, e["choice"]):
            with Session(eg, **SCHEME) as s:
                s.set_runs(eg.iBlockFirst_R)
                s.upload_soa(U0)
                s.init_timestep()
                if choice:
                    apply(s, choice)
                s.advance(4)
                out.append(s.download_soa())
        if e["choice"]["cells_per_thread"] == eg.nI:
            assert np.array_equal(out[0], out[1])
        else:
            assert rel_err(out[1], out[0]) <= 1e-14


def test_compare_block_size():
    print("case autotune block size", flush=True)

    def make_grid(nc):
        n = 64 // nc
        return BlockGrid(nRoot=(n, n, n), nCell=(nc, nc, nc), order="morton")
    r = compare_block_size(make_grid, lambda g: smooth_mhd_soa(g, seed=3),
                           SCHEME)
    print(f"  ns per cell update: 8^3 {r[8]:.3f}, 16^3 {r[16]:.3f}")
    assert r[8] > 0 and r[16] > 0


def test_tune_block_size_stores_the_domain_choice():
    print("case autotune block size per domain", flush=True)

    def make_grid(nc):
        n = 32 // nc
        return BlockGrid(nRoot=(n, n, n), nCell=(nc, nc, nc), order="morton")
    with tempfile.TemporaryDirectory() as d:
        at = Autotuner(os.path.join(d, "t.json"))
        e = at.tune_block_size(make_grid, lambda g: smooth_mhd_soa(g, seed=5),
                               SCHEME, nstep=3, warm=1)
        print("  ", e["choice"], e["ns_per_cell_update"])
        ns = {int(k): v for k, v in e["ns_per_cell_update"].items()}
        assert e["choice"]["NB"] == min(ns, key=ns.get)
        assert Autotuner(at.path).tune_block_size(
            make_grid, None, SCHEME) == e        # stored, not re-measured


def _shape_cases():
    g8 = BlockGrid(nRoot=(4, 4, 4), nCell=(8, 8, 8), order="morton")
    g16 = BlockGrid(nRoot=(2, 2, 2), nCell=(16, 16, 16), order="morton")
    amr = AmrGrid((4, 4, 4), sphere_refinement((4, 4, 4), 0.3))
    gb = BlockGrid(nRoot=(3, 3, 3), nCell=(8, 8, 8),
                   periodic=(False, False, False), order="morton")
    eg, Ue = blast_ensemble(nRun=5, nb_min=8, nb_max=120, seed=6)
    hl = dict(flux="hlld", limiter="mc", beta=1.5)
    return {
        "8^3 powell": (g8, smooth_mhd_soa(g8, seed=11), dict(hl,
                       UseDivbSource=True), {}),
        "16^3 mc3": (g16, smooth_mhd_soa(g16, seed=12),
                     dict(hl, limiter="mc3"), {}),
        "two-level powell": (amr, smooth_mhd_soa(amr, seed=13, amp=0.2),
                             dict(hl, UseDivbSource=True), {}),
        "body": (gb, smooth_mhd_soa(gb, seed=14, amp=0.2),
                 dict(hl, UseDivbSource=True), {"body": body_mask(gb, 0.13)}),
        "ensemble persistent": (eg, Ue, hl, {"runs": eg.iBlockFirst_R,
                                             "nWorkgroup": 256}),
        "fp32": (g8, smooth_mhd_soa(g8, seed=15), dict(hl, precision="f32"),
                 {}),
        "steady": (g8, smooth_mhd_soa(g8, seed=16),
                   dict(hl, IsSteadyState=True), {}),
        "low order": (g8, smooth_mhd_soa(g8, seed=17),
                      dict(hl, UseAdaptiveLowOrder=True), {}),
    }


@pytest.mark.parametrize("case", list(_shape_cases().keys()))
def test_every_stage_shape_gives_the_same_bits(case):
    print(f"case stage shapes {case}", flush=True)
    grid, U0, scheme, extra = _shape_cases()[case]
    shapes = [(None, False)] + [(grid.nI >> k, x) for k in range(3)
                                for x in (False, True)]
    ref = None
    for cpt, xcd in shapes:
        with Session(grid, **scheme) as s:
            if "runs" in extra:
                s.set_runs(extra["runs"])
            if "body" in extra:
                s.set_body(*extra["body"], BodyRho=3.0, BodyP=2.0)
            if cpt is not None:
                s.set_stage_shape(cpt, xcd)
                s.set_persistent_grid(extra.get("nWorkgroup", 0))
            s.upload_soa(U0)
            s.init_timestep()
            s.advance(3)
            got = (s.download_soa(), s.get_dtmax_b(), s.time())
        if ref is None:
            ref = got
            continue
        e = rel_err(got[0], ref[0])
        print(f"  cells/thread {cpt} xcd {xcd}: rel err {e:.1e}", flush=True)
        if cpt == grid.nI:
            assert np.array_equal(got[0], ref[0]), (cpt, xcd)
            assert np.array_equal(got[1], ref[1]), (cpt, xcd)
            assert got[2] == ref[2], (cpt, xcd)
        else:
            # a few ulps of the storage type at the part boundaries
            tol = 1e-6 if scheme.get("precision") == "f32" else 1e-14
            assert e <= tol, (cpt, xcd, e)
            np.testing.assert_allclose(got[1], ref[1], rtol=tol)
            np.testing.assert_allclose(got[2], ref[2], rtol=tol)


def test_stage_shape_refusals():
    from batsrus_amd.core import StopMpi
    g = BlockGrid(nRoot=(2, 2, 2), nCell=(8, 8, 8))
    with Session(g, flux="hlld", limiter="mc") as s:
        for cpt, xcd in ((1, False), (3, False), (16, False), (8, 2)):
            with pytest.raises(StopMpi):
                s._check(s.lib.bats_set_stage_shape(s.h, cpt, int(xcd)),
                         "set_stage_shape")
