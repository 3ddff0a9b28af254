"""Option combinations on the GPU: a persistent work-list grid and the step
graph on a two-level grid give the plain launch's bits; the step graph
replays the Rusanov-gated surrogate stage bit for bit; an ensemble with the
adaptive low order faces matches the oracle run by run; FP32 state with a
two-level grid and the low order faces stays close to the FP64 oracle."""
import numpy as np
import pytest

import pyoracle as po
from batsrus_amd.amr import AmrGrid, sphere_refinement
from batsrus_amd.core import Mlp, Session
from batsrus_amd.ensemble import blast_ensemble
from batsrus_amd.workloads import init_weights
from helpers import smooth_mhd_soa, rel_err

pytestmark = pytest.mark.gpu

LO = dict(UseAdaptiveLowOrder=True, nLowOrder=1, pCritLow=1.3,
          pCritHigh=1.0, VelCrit=0.0)


def _amr():
    g = AmrGrid((4, 4, 4), sphere_refinement((4, 4, 4), 0.3))
    return g, smooth_mhd_soa(g, seed=21, amp=0.2)


def test_amr_persistent_and_graph_are_bitwise_equal():
    print("case combo amr persistent graph", flush=True)
    g, U0 = _amr()
    scheme = dict(flux="hlld", limiter="mc", beta=1.5, cfl=0.8)
    out = []
    for nwg, graph in ((0, False), (7, False), (0, True), (7, True)):
        with Session(g, **scheme) as s:
            s.upload_soa(U0)
            s.init_timestep()
            s.set_persistent_grid(nwg)
            s.set_step_graph(graph)
            s.advance(2)
            s.synchronize()
            s.advance(2)
            out.append(s.download_soa())
    for U in out[1:]:
        assert np.array_equal(U, out[0])


def test_graph_replays_the_gated_rusanov_stage():
    """The surrogate stage in Rusanov mode (no network weights) inside the
    step graph: replays, with downloads between them, equal eager steps."""
    print("case combo graph surrogate-rusanov", flush=True)
    g, U0 = _amr()
    scheme = dict(flux="hlld", limiter="mc", beta=1.5, cfl=0.8,
                  precision="f32")
    res = []
    with Mlp(init_weights(0)) as m:
        for graph in (False, True):
            with Session(g, **scheme) as s:
                s.set_surrogate(m, "rusanov")
                s.upload_soa(U0)
                s.init_timestep()
                s.set_step_graph(graph)
                snaps = []
                for _ in range(3):
                    s.advance(2)
                    snaps.append(s.download_soa())
                res.append(snaps)
    for a, b in zip(*res):
        assert np.array_equal(a, b)


def test_ensemble_with_low_order_faces_matches_oracle_per_run():
    print("case combo ensemble low order", flush=True)
    eg, U0 = blast_ensemble(nRun=3, nb_min=8, nb_max=27, seed=5)
    scheme = dict(flux="hlld", limiter="mc", beta=1.5, cfl=0.8, **LO)
    with Session(eg, **scheme) as s:
        s.set_runs(eg.iBlockFirst_R)
        s.upload_soa(U0)
        s.init_timestep()
        s.advance(2)
        U = s.download_soa()
    for r, gr in enumerate(eg.runs):
        cfg = po.config_from_dict(scheme, gr)
        o = po.Session(gr, cfg, gr.soa_to_state_vgb(
            np.ascontiguousarray(U0[:, eg.run_slice(r)])))
        o.init_timestep()
        for _ in range(2):
            assert o.advance_step() == 0
        assert rel_err(U[:, eg.run_slice(r)], o.soa()) <= 1e-12, r


def test_fp32_two_level_low_order_close_to_fp64_oracle():
    print("case combo fp32 amr low order", flush=True)
    g, U0 = _amr()
    scheme = dict(flux="hlld", limiter="mc", beta=1.5, cfl=0.8, **LO)
    cfg = po.config_from_dict(scheme, g)
    o = po.Session(g, cfg, g.soa_to_state_vgb(U0))
    o.exchange()
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
