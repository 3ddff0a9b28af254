"""Launch-shape autotuner for the stage update.

For one workload (grid, scheme, optional runs) it times the choices that
leave the results bit-identical -- the persistent work-list grid size
(0 = one workgroup per block group, else a multiple of the 256 compute
units) and the per-step hipGraph -- on a scratch session holding the same
state, and keeps the fastest.  Results are stored in a JSON file keyed by
the device, the block size, the scheme and the grid shape, so later runs of
the same workload apply the stored choice without timing again.

The block size (8^3 or 16^3) is a property of the grid; `compare_block_size`
times the same domain on both so a user can choose it before building the
grid.
"""
from __future__ import annotations

import json
import os
import time

from .core import Session

N_CU = 256
DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache",
                             "batsrus_amd", "autotune.json")


def workload_key(grid, scheme, nRun=1):
    """Key of the stored choice: device, block size, scheme, grid shape."""
    s = dict(scheme)
    parts = ["gfx950", "NB%d" % grid.nI, str(s.get("flux", "rusanov")),
             str(s.get("limiter", "mc")), "o%d" % s.get("nOrder", 2),
             "s%d" % s.get("nStage", 2), str(s.get("precision", "f64")),
             "divb%d" % int(bool(s.get("UseDivbSource", False))),
             "low%d" % int(bool(s.get("UseAdaptiveLowOrder", False))),
             "b%d" % grid.nBlock, "r%d" % nRun,
             "amr%d" % int(getattr(grid, "nSlab", 0) > 0)]
    return "/".join(parts)


def candidates(grid):
    """Persistent grid sizes (0 = none) x step graph on/off."""
    groups = -(-grid.nBlock // max(1, 256 // (grid.nI * grid.nJ)))
    sizes = [0] + [k * N_CU for k in (1, 2, 4, 8) if k * N_CU < groups]
    return [dict(nWorkgroup=n, graph=g) for n in sizes for g in (False, True)]


def apply(session, choice):
    session.set_persistent_grid(int(choice["nWorkgroup"]))
    session.set_step_graph(bool(choice["graph"]))


class Autotuner:
    def __init__(self, path=DEFAULT_CACHE):
        self.path = path
        self.table = {}
        if path and os.path.exists(path):
            with open(path) as f:
                self.table = json.load(f)

    def save(self):
        if not self.path:
            return
        os.makedirs(os.path.dirname(self.path), exist_ok=True)
        tmp = self.path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(self.table, f, indent=1, sort_keys=True)
        os.replace(tmp, self.path)

    def lookup(self, grid, scheme, nRun=1):
        return self.table.get(workload_key(grid, scheme, nRun))

    def tune(self, grid, U0, scheme, runs=None, nstep=6, warm=2,
             force=False):
        """Best choice for the workload (stored or measured).  Returns the
        table entry: the choice, and the measured step times (us) of every
        candidate."""
        nRun = 1 if runs is None else len(runs) - 1
        key = workload_key(grid, scheme, nRun)
        if not force and key in self.table:
            return self.table[key]
        times = []
        with Session(grid, **scheme) as s:
            if runs is not None:
                s.set_runs(runs)
            for c in candidates(grid):
                s.upload_soa(U0)
                s.init_timestep()
                apply(s, c)
                s.advance(warm)
                s.synchronize()
                t0 = time.perf_counter()
                s.advance(nstep)
                s.synchronize()
                us = (time.perf_counter() - t0) / nstep * 1e6
                times.append(dict(c, step_us=us))
        best = min(times, key=lambda t: t["step_us"])
        entry = dict(choice=dict(nWorkgroup=best["nWorkgroup"],
                                 graph=best["graph"]),
                     step_us=best["step_us"], measured=times)
        self.table[key] = entry
        self.save()
        return entry


def compare_block_size(make_grid, make_state, scheme, nstep=6, warm=2):
    """Step time per cell update for the same domain on 8^3 and 16^3
    blocks: make_grid(nCell) builds the grid, make_state(grid) the state.
    Returns {nCell: ns per cell update}."""
    out = {}
    for nc in (8, 16):
        g = make_grid(nc)
        with Session(g, **scheme) as s:
            s.upload_soa(make_state(g))
            s.init_timestep()
            s.advance(warm)
            s.synchronize()
            t0 = time.perf_counter()
            s.advance(nstep)
            s.synchronize()
            el = time.perf_counter() - t0
        out[nc] = el / (nstep * g.nBlock * nc ** 3) * 1e9
    return out
