"""Launch-shape and block-size autotuner for the stage update.

For one workload (grid, scheme, optional runs) `Autotuner.tune` times every
launch shape of the stage kernel, on a scratch session holding the same
state, and keeps the fastest:
  * cells per thread: each thread marches NB, NB/2 or NB/4 planes of its
    column (the block's planes split over 1, 2 or 4 workgroups, which
    fills more of the chip on small grids at the price of one more z-face
    row per part; the face at each part boundary is evaluated by the
    window's prologue, which rounds within an ulp of the in-loop face);
  * work order: linear, or XCD-contiguous (each of the 8 XCDs takes a
    contiguous range of the Morton-ordered work list, so neighbouring
    blocks share its L2);
  * the persistent work-list grid size (0 = one workgroup per work item,
    else a multiple of the 256 compute units);
  * the per-step hipGraph.
All shapes but the cells per thread leave the results bit-identical.
Results are stored in a JSON file keyed by the device, the block size, the
scheme and the grid shape, so later runs of the same workload apply the
stored choice without timing again.

The block size (8^3 or 16^3) is a property of the grid: `tune_block_size`
tunes the launch shape of the same domain cut into 8^3 and into 16^3
blocks and stores the faster block size (per cell update) under the
domain's key, for the caller to build its grid with.

Run as `python -m batsrus_amd.autotune --workload NAME` to tune one of the
benchmark workloads and print the measured table.
"""
from __future__ import annotations

import argparse
import json
import os
import time

from .core import Session

N_CU = 256
DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache",
                             "batsrus_amd", "autotune.json")


def _scheme_parts(scheme):
    s = dict(scheme)
    return [str(s.get("flux", "rusanov")), str(s.get("limiter", "mc")),
            "o%d" % s.get("nOrder", 2), "s%d" % s.get("nStage", 2),
            str(s.get("precision", "f64")),
            "divb%d" % int(bool(s.get("UseDivbSource", False))),
            "low%d" % int(bool(s.get("UseAdaptiveLowOrder", False)))]


def workload_key(grid, scheme, nRun=1):
    """Key of the stored choice: device, block size, scheme, grid shape."""
    parts = ["gfx950", "NB%d" % grid.nI] + _scheme_parts(scheme) + [
        "b%d" % grid.nBlock, "r%d" % nRun,
        "amr%d" % int(getattr(grid, "nSlab", 0) > 0)]
    return "/".join(parts)


def domain_key(ncell, scheme):
    """Key of a stored block size: device, cell count, scheme."""
    return "/".join(["gfx950", "domain", "c%d" % ncell] +
                    _scheme_parts(scheme))


def _items(grid, cells_per_thread):
    bpw = max(1, 256 // (grid.nI * grid.nJ))
    groups = -(-grid.nBlock // bpw)
    return groups * (grid.nI // cells_per_thread)


def candidates(grid):
    """Cells per thread x work order x persistent grid size (0 = none)
    x step graph on/off."""
    out = []
    for c in (grid.nI, grid.nI // 2, grid.nI // 4):
        items = _items(grid, c)
        sizes = [0] + [k * N_CU for k in (1, 2, 4, 8) if k * N_CU < items]
        out += [dict(cells_per_thread=c, xcd=x, nWorkgroup=n, graph=g)
                for x in (False, True) for n in sizes for g in (False, True)]
    return out


def apply(session, choice):
    """Set a stored choice on a session (missing keys: the defaults)."""
    session.set_persistent_grid(int(choice.get("nWorkgroup", 0)))
    session.set_step_graph(bool(choice.get("graph", False)))
    session.set_stage_shape(choice.get("cells_per_thread"),
                            bool(choice.get("xcd", False)))


def _choice(c):
    return {k: c[k] for k in ("cells_per_thread", "xcd", "nWorkgroup",
                              "graph")}


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
             force=False, progress=None):
        """Best launch shape for the workload (stored or measured).
        Returns the table entry: the choice, its step time (us) and the
        measured step times of every candidate."""
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
                if progress:
                    progress(times[-1])
        best = min(times, key=lambda t: t["step_us"])
        entry = dict(choice=_choice(best), step_us=best["step_us"],
                     measured=times)
        self.table[key] = entry
        self.save()
        return entry

    def tune_block_size(self, make_grid, make_state, scheme, nstep=6,
                        warm=2, force=False, progress=None):
        """Block size of a domain: make_grid(nCell) cuts the same domain
        into nCell^3 blocks, make_state(grid) gives its state.  Tunes the
        launch shape for 8^3 and 16^3 blocks and stores the block size with
        the lower time per cell update (and that size's launch shape) under
        the domain's key."""
        grids = {nc: make_grid(nc) for nc in (8, 16)}
        ncell = grids[8].nBlock * 8 ** 3
        assert grids[16].nBlock * 16 ** 3 == ncell, "not the same domain"
        key = domain_key(ncell, scheme)
        if not force and key in self.table:
            return self.table[key]
        per, shapes = {}, {}
        for nc, g in grids.items():
            e = self.tune(g, make_state(g), scheme, nstep=nstep, warm=warm,
                          force=force, progress=progress)
            per[nc] = e["step_us"] * 1e3 / ncell
            shapes[nc] = e["choice"]
        best = min(per, key=per.get)
        entry = dict(choice=dict(NB=best, **shapes[best]),
                     ns_per_cell_update={str(k): v for k, v in per.items()})
        self.table[key] = entry
        self.save()
        return entry


def compare_block_size(make_grid, make_state, scheme, nstep=6, warm=2):
    """Step time per cell update for the same domain on 8^3 and 16^3
    blocks with the default launch shape: make_grid(nCell) builds the
    grid, make_state(grid) the state.  Returns {nCell: ns per cell
    update}."""
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


def _workload(name, precision):
    from . import workloads as W
    table = {"orszag_tang_256": (W.orszag_tang_256, W.OT_SCHEME),
             "kh_two_level": (W.kh_two_level, W.KH_SCHEME),
             "shock_tube_64": (W.shock_tube_64, W.TUBE_SCHEME)}
    make, scheme = table[name]
    grid, U0 = make()
    return grid, U0, dict(scheme, precision=precision)


def markdown(name, grid, scheme, entry):
    """The measured table of one workload as markdown."""
    lines = ["| cells/thread | XCD order | persistent WGs | graph | step (us) |",
             "|---|---|---|---|---|"]
    for m in sorted(entry["measured"], key=lambda m: m["step_us"]):
        lines.append("| %d | %s | %d | %s | %.1f |" % (
            m["cells_per_thread"], "yes" if m["xcd"] else "no",
            m["nWorkgroup"], "yes" if m["graph"] else "no", m["step_us"]))
    return "\n".join(["### %s (%s)" % (name, workload_key(grid, scheme)), "",
                      "choice: %s, %.1f us per step" % (
                          json.dumps(entry["choice"]), entry["step_us"]),
                      ""] + lines + [""])


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--workload", action="append", required=True,
                    choices=["orszag_tang_256", "kh_two_level",
                             "shock_tube_64"])
    ap.add_argument("--precision", default="f64", choices=["f64", "f32"])
    ap.add_argument("--table", default=DEFAULT_CACHE)
    ap.add_argument("--steps", type=int, default=6)
    ap.add_argument("--markdown", default=None,
                    help="append the measured tables to this file")
    args = ap.parse_args(argv)
    at = Autotuner(args.table)
    for name in args.workload:
        grid, U0, scheme = _workload(name, args.precision)
        print("tuning %s (%s)" % (name, workload_key(grid, scheme)),
              flush=True)
        e = at.tune(grid, U0, scheme, nstep=args.steps, force=True,
                    progress=lambda m: print("  ", json.dumps(m), flush=True))
        md = markdown(name, grid, scheme, e)
        print(md, flush=True)
        if args.markdown:
            with open(args.markdown, "a") as f:
                f.write(md + "\n")


if __name__ == "__main__":
    main()
