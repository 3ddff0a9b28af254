#!/usr/bin/env python3
"""Benchmark of the plain FP64 explicit MHD stage update on one MI355X.

Workload (the reference's HBM-roofline case): 3-D ideal MHD on 4096 blocks
of 16^3 cells (256^3 = 16.8 M cells, periodic box [0,1]^3), Orszag-Tang
vortex extended in z with seeded phase shifts, Koren limiter (mc3) with
beta = 1.5, HLLD flux with the Linde normal-field fix, 2-stage midpoint
scheme, CFL 0.8, FP64.  One "step" is one full time step (global dt
reduction + both stages) over all blocks; the state stays in HBM.

Reported value: cell-updates/s (interior cells x time steps / s).  The
roofline object prices the fused stage kernel at its algorithmic HBM bytes
(stage 1 reads U^n and writes U*: 2 x 64 B per cell; stage 2 reads U* and
U^n and writes U^{n+1}: 3 x 64 B; 160 B per cell per launch on average)
against the launch durations measured with HIP events on the kernel's own
stream.  The CPU baseline times the CPU oracle (a C restatement of the
reference's Fortran path) on a bounded sample of the same workload.

usage: python bench.py [--gpus 1] [--steps 50] [--warmup 3]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL_UPDATE = 320.0      # 5 x nVar x 8 B, 2-stage midpoint
BYTES_PER_CELL_LAUNCH = 160.0      # average over the two stage launches


def workload_grid():
    from batsrus_amd.grid import BlockGrid
    return BlockGrid(nRoot=(16, 16, 16), nCell=(16, 16, 16), order="morton")


SCHEME = dict(flux="hlld", limiter="mc3", beta=1.5, nStage=2, cfl=0.8,
              nOrder=2)


def cpu_baseline(target_seconds):
    """Time the CPU oracle on a bounded sample of the same workload
    (a slab of the 256^3 box), all host threads."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    from batsrus_amd.grid import BlockGrid
    from batsrus_amd import ics
    nthreads = min(os.cpu_count() or 1, 16)
    g = BlockGrid(nRoot=(16, 16, 2), nCell=(16, 16, 16), order="morton")
    U = ics.orszag_tang_z(g)
    cfg = po.config_from_dict(dict(SCHEME), g)
    s = po.Session(g, cfg, g.soa_to_state_vgb(U), nThread=nthreads)
    s.init_timestep()
    n = 0
    t0 = time.perf_counter()
    while True:
        s.advance_step()
        n += 1
        el = time.perf_counter() - t0
        if n >= 2 and el >= target_seconds:
            break
    return {"value": g.nCellTotal * n / el, "unit": "cell-updates/s",
            "cores": nthreads, "kind": "port",
            "sample": "CPU oracle (C restatement of the reference Fortran "
                      "orig path, OpenMP over blocks) on 512 of the 4096 "
                      "blocks (16x16x2 slab, 2.1M cells), %d steps, "
                      "%.1f s" % (n, el)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a plain device copy of known size "
                         "(1.07 GB, for scaling HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU only; multi-GPU sharding "
              "is designed but not built (see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    from batsrus_amd import ics
    from batsrus_amd.core import Session

    grid = workload_grid()
    U0 = ics.orszag_tang_z(grid)
    ncell = grid.nCellTotal
    with Session(grid, device=0, **SCHEME) as s:
        s.upload_soa(U0)
        del U0
        s.init_timestep()
        if args.warmup:
            s.advance(args.warmup)
        s.synchronize()
        t0 = time.perf_counter()
        s.advance(args.steps)          # synchronizes at the end
        s.synchronize()
        el = time.perf_counter() - t0
        kms, nlaunch = s.last_stage_kernel_ms()
        dt, tsim, nstep = s.time()
        s.check_nan()
    value = ncell * args.steps / el
    achieved = ncell * BYTES_PER_CELL_LAUNCH / (kms * 1e-3) / 1e9
    out = {
        "metric": "cell-updates/s: plain FP64 vs surrogate FP32 at stated "
                  "L1 err; % HBM peak",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "plain FP64 path: 3-D ideal MHD, 4096 blocks of "
                        "16^3 (256^3 cells, periodic), Orszag-Tang-z with "
                        "seeded phases, Koren (mc3) beta=1.5, HLLD + Linde "
                        "fix, 2-stage midpoint, CFL 0.8",
            "blocks": grid.nBlock, "block_cells": "16x16x16",
            "cells": ncell, "flux": "HLLD", "limiter": "mc3",
            "stages": 2, "state_layout": "SoA, interior only, 2 buffers",
            "bytes_per_cell_update": BYTES_PER_CELL_UPDATE,
            "hbm_frac_of_step": value * BYTES_PER_CELL_UPDATE / 1e9 /
            HBM_PEAK_GBS,
            "sim_time": tsim, "n_step": nstep,
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "kernel": "k_stage<16,HLLD> (fused face value + flux + update)",
            "kernel_ms_avg": kms, "launches": nlaunch,
            "bytes_per_launch": ncell * BYTES_PER_CELL_LAUNCH,
        },
    }
    if args.copy_calibration:
        # a plain device-to-device copy of known size (the full state,
        # 8 x 16.8M doubles) through the library's halo snapshot, for
        # scaling the HBM byte counters
        with Session(grid, device=0, **SCHEME) as s2:
            s2.exchange_messages()
            s2.synchronize()
            s2.exchange_messages()
            s2.synchronize()
        out["copy_calibration_bytes"] = {"read": ncell * 64,
                                         "write": ncell * 64}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.cpu_seconds)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
