"""The magnetosphere workload of the benchmark (a dipole B0 at the centre of
an ionosphere body, float sides, a southward-field solar-wind state, HLLD
with mc3 beta 1.5 and the 8-wave source with its B0 terms, FP64).  On a
reduced grid of 2x2x2 blocks of 16^3 (the body reaches all eight blocks)
its initial state advances stage by stage within 1e-12 of the oracle; on
the full grid (4096 blocks of 16^3, 32 of them reaching the body) the body
cells stay bitwise at rest, the state stays finite and the step graph
replays the eager steps bit for bit."""
import numpy as np
import pytest

import pyoracle as po
from batsrus_amd.core import Session
from batsrus_amd.workloads import MAG_BODY, MAG_SCHEME, magnetosphere_256
from helpers import rel_err_vec

pytestmark = pytest.mark.gpu


def test_magnetosphere_reduced_grid_matches_oracle():
    print("case magnetosphere workload reduced grid", flush=True)
    g, U0, B0, used, xyz0 = magnetosphere_256(nRoot=2)
    nb = ~used.reshape(g.nBlock, -1).all(axis=1)
    assert nb.all()
    scheme = dict(MAG_SCHEME)
    cfg = po.config_from_dict(scheme, g)
    o = po.Session(g, cfg, g.soa_to_state_vgb(U0), B0_DGB=B0,
                   UseB0Source=True, Used_GB=used, XyzStart_DB=xyz0,
                   **MAG_BODY)
    o.exchange()
    o.init_timestep()
    with Session(g, **scheme) as s:
        s.set_b0(B0, UseB0Source=True)
        s.set_body(used, xyz0, **MAG_BODY)
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
                e = rel_err_vec(s.download_soa(), o.soa())
                print(f"  step {step} stage {stage} rel err {e:.2e}")
                assert e <= 1e-12, (step, stage, e)


def test_magnetosphere_full_grid_body_at_rest_and_graph_replays():
    print("case magnetosphere workload full grid", flush=True)
    g, U0, B0, used, xyz0 = magnetosphere_256()
    G = g.nG
    inner = used[:, G:-G, G:-G, G:-G]
    assert (~inner).any()
    assert int((~used.reshape(g.nBlock, -1).all(axis=1)).sum()) == 32
    out = []
    for graph in (False, True):
        with Session(g, **MAG_SCHEME) as s:
            s.set_b0(B0, UseB0Source=True)
            s.set_body(used, xyz0, **MAG_BODY)
            s.set_step_graph(graph)
            s.upload_soa(U0)
            s.init_timestep()
            s.advance(3)
            s.check_nan()
            out.append(s.download_soa())
        print(f"  graph={graph} done", flush=True)
    assert np.array_equal(out[0], out[1])
    U = out[0]
    assert np.isfinite(U).all()
    for v in range(8):
        assert np.array_equal(U[v][~inner], U0[v][~inner]), v
    # the run moved the state everywhere else
    assert not np.array_equal(U[:, inner], U0[:, inner])
