"""The launch-shape autotuner: it measures every candidate on a scratch
session, stores the fastest under the workload's key, reloads it without
timing again, and the chosen launch shape gives the same bits as the
default one."""
import os
import tempfile

import numpy as np
import pytest

from batsrus_amd.autotune import Autotuner, apply, compare_block_size
from batsrus_amd.core import Session
from batsrus_amd.ensemble import blast_ensemble
from batsrus_amd.grid import BlockGrid
from helpers import smooth_mhd_soa

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
        print("  measured (nWorkgroup, graph, us):",
              [(m["nWorkgroup"], m["graph"], round(m["step_us"], 1))
               for m in e["measured"]])
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
        assert np.array_equal(out[0], out[1])


def test_compare_block_size():
    print("case autotune block size", flush=True)

    def make_grid(nc):
        n = 64 // nc
        return BlockGrid(nRoot=(n, n, n), nCell=(nc, nc, nc), order="morton")
    r = compare_block_size(make_grid, lambda g: smooth_mhd_soa(g, seed=3),
                           SCHEME)
    print(f"  ns per cell update: 8^3 {r[8]:.3f}, 16^3 {r[16]:.3f}")
    assert r[8] > 0 and r[16] > 0
