"""Ensembles on the GPU: independent runs of uneven block counts advance in
one stage launch, each with its own CFL time step; every run matches the
oracle run alone, and the persistent work-list grid gives the same bits as
one workgroup per block group."""
import numpy as np
import pytest

import pyoracle as po
from batsrus_amd.core import Session
from batsrus_amd.ensemble import blast_ensemble
from helpers import rel_err

pytestmark = pytest.mark.gpu

SCHEME = dict(flux="hlld", limiter="mc", beta=1.5, cfl=0.8)


def _oracle_runs(eg, U0, nstep):
    out, dts = [], []
    for r, g in enumerate(eg.runs):
        cfg = po.config_from_dict(SCHEME, g)
        S = g.soa_to_state_vgb(np.ascontiguousarray(U0[:, eg.run_slice(r)]))
        o = po.Session(g, cfg, S)
        o.init_timestep()
        for _ in range(nstep):
            assert o.advance_step() == 0
        out.append(o.soa())
        dts.append(o.Dt)
    return out, dts


@pytest.mark.parametrize("persistent", [0, 7])
def test_ensemble_matches_oracle_per_run(persistent):
    print(f"case ensemble persistent={persistent}", flush=True)
    eg, U0 = blast_ensemble(nRun=5, nb_min=8, nb_max=40, seed=11)
    nstep = 3
    ref, dts = _oracle_runs(eg, U0, nstep)
    with Session(eg, **SCHEME) as s:
        s.set_runs(eg.iBlockFirst_R)
        s.set_persistent_grid(persistent)
        s.upload_soa(U0)
        s.init_timestep()
        s.advance(nstep)
        U = s.download_soa()
        dt, t, n = s.run_times()
    assert np.all(n == nstep)
    assert len(set(np.round(dt, 15))) > 1      # the runs' steps differ
    for r in range(eg.nRun):
        assert abs(dt[r] - dts[r]) <= 1e-12 * dts[r]
        assert rel_err(U[:, eg.run_slice(r)], ref[r]) <= 1e-12, r


def test_persistent_grid_is_bitwise_identical():
    eg, U0 = blast_ensemble(nRun=4, nb_min=8, nb_max=64, seed=3)
    res = []
    for nwg in (0, 5, 64):
        with Session(eg, **SCHEME) as s:
            s.set_runs(eg.iBlockFirst_R)
            s.set_persistent_grid(nwg)
            s.upload_soa(U0)
            s.init_timestep()
            s.advance(2)
            res.append(s.download_soa())
    assert np.array_equal(res[0], res[1]) and np.array_equal(res[0], res[2])
