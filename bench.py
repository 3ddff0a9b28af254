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

The surrogate leg (on by default, --no-surrogate skips it) runs the
two-level Kelvin-Helmholtz workload (32,256 blocks of 8^3, 200 steps) as
plain FP64 HLLD (the reference for the errors), FP32 state with the trained
network on the smooth faces, and FP32 state with the Rusanov flux on the
same faces (the method without the network); it reports each rate, the
worst per-variable relative L1 and L-infinity errors against the FP64 run
(and all eight), and the training statistics.  The headline network is
trained on the GPU as the surrogate configuration states (faces of the
Orszag-Tang and blast-ensemble runs, 2e7 samples, minibatch 65536, lr 1e-3
cosine, 10 % held out); a second network, trained with faces of a KH run of
another seed added, is reported as a labelled ablation.

usage: python bench.py [--gpus 1] [--steps 50] [--warmup 3]
"""
import argparse
import csv
import glob
import json
import os
import shutil
import subprocess
import sys
import tempfile
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


def _time_oracle(nroot_z, nthreads, target_seconds):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    from batsrus_amd.grid import BlockGrid
    from batsrus_amd import ics
    g = BlockGrid(nRoot=(16, 16, nroot_z), nCell=(16, 16, 16),
                  order="morton")
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
    return g, n, el


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_threads():
    """Host threads this process may use: the CPUs it is allowed to run on,
    capped by OMP_NUM_THREADS when the environment sets the job's CPU share
    (a one-GPU job on a shared host gets a share of the host's cores)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    share = os.environ.get("OMP_NUM_THREADS")
    if share and share.isdigit() and int(share) > 0:
        n = min(n, int(share))
    return n


def cpu_baseline(target_seconds):
    """Time the CPU oracle on a bounded sample of the same workload
    (a slab of the 256^3 box) on every host thread this job may use, and on
    one thread for a smaller slab."""
    nthreads = cpu_threads()
    g, n, el = _time_oracle(2, nthreads, target_seconds)
    g1, n1, el1 = _time_oracle(1, 1, target_seconds / 2)
    return {"value": g.nCellTotal * n / el, "unit": "cell-updates/s",
            "cores": nthreads, "kind": "port",
            "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
            "threads_note": "OpenMP threads = the CPUs this job may use "
                            "(affinity, capped by OMP_NUM_THREADS, the "
                            "job's share of the host)",
            "sample": "CPU oracle (C restatement of the reference Fortran "
                      "orig path, OpenMP over blocks) on 512 of the 4096 "
                      "blocks (16x16x2 slab, 2.1M cells), %d steps, "
                      "%.1f s" % (n, el),
            "single_thread": {"value": g1.nCellTotal * n1 / el1,
                              "sample": "256 blocks (16x16x1 slab), %d "
                                        "steps, %.1f s" % (n1, el1)}}


def copy_bandwidth(session, nbytes, reps=10):
    """Achievable HBM bandwidth of a plain device-to-device copy of the
    state (the library's halo snapshot), read + write bytes per second."""
    session.exchange_messages()           # allocates the snapshot
    session.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        session.exchange_messages()
    session.synchronize()
    el = time.perf_counter() - t0
    return 2.0 * nbytes * reps / el / 1e9


def measure_traffic(timeout=150):
    """HBM bytes per stage launch: two rocprofv3 --pmc child runs of this
    script (2 steps, FETCH_SIZE then WRITE_SIZE, each counter in a run of its
    own), each counter scaled by what it reads for the plain 1.07 GB device
    copy of the same run.  Returns (bytes per launch, details) or (None,
    reason)."""
    exe = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not os.path.exists(exe):
        return None, "rocprofv3 not found"
    known_kib = 16777216 * 64 / 1024.0        # the copy: 8 x 16.8M doubles
    per = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        d = tempfile.mkdtemp(prefix="bench_pmc_")
        try:
            cmd = [exe, "--pmc", counter, "--output-format", "csv", "-d", d,
                   "-o", "run", "--", sys.executable,
                   os.path.abspath(__file__), "--steps", "2", "--warmup",
                   "0", "--no-cpu", "--no-surrogate", "--no-ensemble",
                   "--no-magnetosphere", "--no-traffic",
                   "--copy-calibration"]
            r = subprocess.run(cmd, stdout=subprocess.DEVNULL,
                               stderr=subprocess.DEVNULL, timeout=timeout)
            if r.returncode != 0:
                return None, "%s pass exited with %d" % (counter,
                                                         r.returncode)
            files = glob.glob(os.path.join(d, "**",
                                           "*counter_collection.csv"),
                              recursive=True)
            if not files:
                return None, "%s pass wrote no counters" % counter
            with open(files[0]) as f:
                rows = [x for x in csv.DictReader(f)
                        if x["Counter_Name"] == counter]
        except subprocess.TimeoutExpired:
            return None, "%s pass timed out" % counter
        finally:
            shutil.rmtree(d, ignore_errors=True)
        copy = [float(x["Counter_Value"]) for x in rows
                if "copyBuffer" in x["Kernel_Name"]
                and int(x["Grid_Size"]) == 131072]
        stage = [float(x["Counter_Value"]) for x in rows
                 if "k_stage" in x["Kernel_Name"]]
        if not copy or len(stage) < 2:
            return None, "%s: calibration or stage dispatches missing" % (
                counter)
        scale = known_kib / (sum(copy) / len(copy))
        launches = stage[1:]          # the first is the dt-only pass
        per[counter] = (scale, sum(launches) / len(launches) * 1024.0 * scale)
    total = per["FETCH_SIZE"][1] + per["WRITE_SIZE"][1]
    return total, {"fetch_bytes": per["FETCH_SIZE"][1],
                   "write_bytes": per["WRITE_SIZE"][1],
                   "fetch_scale": per["FETCH_SIZE"][0],
                   "write_scale": per["WRITE_SIZE"][0],
                   "source": "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, "
                             "4 stage launches, scaled by a 1.07 GB copy"}


def ensemble_leg(steps=20):
    """The batched ensemble: 48 independent blast-wave runs of 8 to 4096
    blocks of 8^3 (seeded log-uniform counts), MC beta 1.5, HLLD, per-run
    CFL time steps, all runs advanced by one launch per stage; FP64 and
    FP32 state."""
    from batsrus_amd.core import Session
    from batsrus_amd.ensemble import blast_ensemble
    from batsrus_amd.workloads import ENS_SCHEME
    g, U0 = blast_ensemble(nRun=48)
    out = {"workload": "48 blast-wave runs, %d blocks of 8^3 in all (%d "
                       "cells), MC beta=1.5, HLLD+Linde, 2-stage, CFL 0.8, "
                       "per-run dt, %d steps" % (g.nBlock, g.nCellTotal,
                                                 steps),
           "unit": "cell-updates/s", "runs": g.nRun}
    for prec in ("f64", "f32"):
        with Session(g, precision=prec, **ENS_SCHEME) as s:
            s.set_runs(g.iBlockFirst_R)
            s.upload_soa(U0)
            s.init_timestep()
            s.advance(3)
            s.synchronize()
            t0 = time.perf_counter()
            s.advance(steps)
            s.synchronize()
            el = time.perf_counter() - t0
            kms, _ = s.last_stage_kernel_ms()
            s.check_nan()
        rate = g.nCellTotal * steps / el
        bpc = 160.0 if prec == "f64" else 80.0
        out[prec] = {"value": rate, "ms_per_step": el * 1e3 / steps,
                     "stage_kernel_ms": kms,
                     "hbm_frac": g.nCellTotal * bpc / (kms * 1e-3) / 1e9 /
                     HBM_PEAK_GBS}
    return out


def magnetosphere_leg(steps=20):
    """B0 splitting and the inner body on the 256^3 box: a dipole B0 at the
    centre of an ionosphere body, float sides, HLLD with the 8-wave source
    and its B0 terms, FP64 (the B0 + body + 8-wave stage variants on the
    groups of blocks that reach the body, the B0 + 8-wave variant on the
    rest)."""
    from batsrus_amd.core import Session
    from batsrus_amd import workloads as W
    g, U0, B0, used, xyz0 = W.magnetosphere_256()
    with Session(g, **W.MAG_SCHEME) as s:
        s.set_b0(B0, UseB0Source=True)
        del B0
        s.set_body(used, xyz0, **W.MAG_BODY)
        s.upload_soa(U0)
        s.init_timestep()
        s.advance(3)
        s.synchronize()
        t0 = time.perf_counter()
        s.advance(steps)
        s.synchronize()
        el = time.perf_counter() - t0
        kms, _ = s.last_stage_kernel_ms()
        s.check_nan()
    rate = g.nCellTotal * steps / el
    return {"workload": "4096 blocks of 16^3 (256^3 cells), float sides, "
                        "dipole B0 at the centre of an ionosphere body of "
                        "radius 0.06, solar-wind state, mc3 beta=1.5, "
                        "HLLD+Linde, 8-wave source with the B0 terms, "
                        "2-stage, CFL 0.8, FP64, %d steps" % steps,
            "unit": "cell-updates/s", "value": rate,
            "ms_per_step": el * 1e3 / steps, "stage_kernel_ms": kms,
            "hbm_frac": g.nCellTotal * BYTES_PER_CELL_LAUNCH /
            (kms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "body_blocks": int((~used.reshape(g.nBlock, -1).all(1)).sum())}


def _kh_result(W, np, g, U0, steps, ref, r64, m, mode):
    U, r, frac = W.run_kh(g, U0, steps, "f32", model=m, mode=mode)
    l1 = W.rel_l1(U, ref)
    linf = (np.abs(U - ref).reshape(8, -1).max(1) /
            np.abs(ref).reshape(8, -1).max(1))
    return {
        "value": r, "vs_plain_fp64": r / r64,
        "hbm_frac": r * 160.0 / 1e9 / HBM_PEAK_GBS,
        "network_face_fraction": frac,
        "rel_l1_worst_var": float(l1.max()),
        "rel_linf_worst_var": float(linf.max()),
        "rel_l1_per_var": [float(x) for x in l1],
        "rel_linf_per_var": [float(x) for x in linf],
        "rel_l1_all_vars": float(np.abs(U - ref).sum() / np.abs(ref).sum()),
    }


def _training(sources, st):
    return {"sources": list(sources), "samples": st["samples"],
            "epochs": st["epochs"], "batch": st["batch"], "lr": st["lr"],
            "held_out": st["hold"],
            "harvest_s": st["harvest_s"], "train_s": st["train_s"],
            "heldout_flux_rel_l2": st["heldout_rel_l2"],
            "heldout_correction_rel_l2": st["heldout_correction_rel_l2"]}


def surrogate_leg(steps=200, samples=20_000_000):
    """Train the face-flux network on the GPU, then run the two-level KH
    workload in FP64 (plain HLLD, the reference for the errors) and in FP32
    with the network and with the Rusanov flux on the same faces.

    Headline network: the training set and schedule of the surrogate
    configuration -- faces harvested from the Orszag-Tang and blast-ensemble
    runs only, 2e7 samples, minibatch 65536, Adam lr 1e-3 with cosine
    decay, 10 % held out.  Ablation: the same plus faces of a KH run of
    another seed (the evaluated workload's family), minibatch 8192, lr
    2e-3."""
    import numpy as np
    from batsrus_amd import workloads as W
    g, U0 = W.kh_two_level()
    ref, r64, _ = W.run_kh(g, U0, steps, "f64")
    out = {
        "workload": "two-level Kelvin-Helmholtz: %d blocks of 8^3 (%d "
                    "cells), fine sphere r<0.25, MC beta=1.5, HLLD+Linde, "
                    "2-stage, CFL 0.8, %d steps" % (g.nBlock, g.nCellTotal,
                                                    steps),
        "unit": "cell-updates/s",
        "plain_fp64": r64,
        "error_note": "rel_l1 = sum|U-U_fp64| / sum|U_fp64| per conserved "
                      "variable (rho, rho u, B, p); rel_linf = max|U-U_fp64| "
                      "/ max|U_fp64| per variable; worst_var = the largest "
                      "over the 8 variables",
    }
    head_src = ("orszag_tang", "blast_ensemble")
    m, st = W.build_surrogate(n_samples=samples, epochs=100, batch=65536,
                              lr=1e-3, sources=head_src, jump_weight=True)
    try:
        out["surrogate_fp32"] = _kh_result(W, np, g, U0, steps, ref, r64, m,
                                           "network")
        out["surrogate_fp32"]["training"] = _training(head_src, st)
        out["without_network_fp32"] = _kh_result(W, np, g, U0, steps, ref,
                                                 r64, m, "rusanov")
    finally:
        m.close()
    abl_src = ("orszag_tang", "blast_ensemble", "kh_family")
    m, st = W.build_surrogate(n_samples=samples, epochs=100, batch=8192,
                              lr=2e-3, sources=abl_src, jump_weight=True)
    try:
        out["ablation_kh_augmented_fp32"] = _kh_result(
            W, np, g, U0, steps, ref, r64, m, "network")
        out["ablation_kh_augmented_fp32"]["training"] = _training(abl_src, st)
        out["ablation_kh_augmented_fp32"]["note"] = (
            "training set augmented with a KH run of another seed (same "
            "family as the evaluated workload), minibatch 8192, lr 2e-3")
    finally:
        m.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-surrogate", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--no-magnetosphere", action="store_true")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the two rocprofv3 --pmc child runs that "
                         "measure HBM bytes per stage launch")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a plain device copy of known size "
                         "(1.07 GB, for scaling HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: runs on one GPU only; multi-GPU sharding is "
              "designed but not built", file=sys.stderr)
        sys.exit(2)

    # HBM traffic first, in child processes, before this process touches
    # the GPU
    traffic, traffic_info = (None, "skipped") if args.no_traffic else \
        measure_traffic()

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
        copy_gbs = copy_bandwidth(s, ncell * 64)
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
            "traffic": traffic,
            "traffic_detail": traffic_info,
            "kernel": "k_stage<16,HLLD> (fused face value + flux + update)",
            "kernel_ms_avg": kms, "launches": nlaunch,
            "bytes_per_launch": ncell * BYTES_PER_CELL_LAUNCH,
            "copy_peak_measured": copy_gbs,
            "frac_of_measured_copy": achieved / copy_gbs,
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
    if not args.no_ensemble:
        try:
            out["ensemble"] = ensemble_leg()
        except Exception as e:            # reported, not fatal to the line
            out["ensemble"] = {"error": "%s: %s" % (type(e).__name__, e)}
    if not args.no_magnetosphere:
        try:
            out["magnetosphere"] = magnetosphere_leg()
        except Exception as e:            # reported, not fatal to the line
            out["magnetosphere"] = {"error": "%s: %s" % (type(e).__name__,
                                                         e)}
    if not args.no_surrogate:
        try:
            out["surrogate"] = surrogate_leg()
        except Exception as e:            # reported, not fatal to the line
            out["surrogate"] = {"error": "%s: %s" % (type(e).__name__, e)}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.cpu_seconds)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
