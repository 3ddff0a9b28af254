"""The benchmark workloads and the surrogate's training pipeline.

  shock_tube_64     64 blocks of 8^3, Brio-Wu along x, MC beta 1.5,
                    Rusanov, 2 stages, CFL 0.8, 20 steps
  orszag_tang_256   4096 blocks of 16^3 (256^3 cells), Orszag-Tang-z with
                    seeded phases, Koren (mc3) beta 1.5, HLLD + Linde, 2 stages
  blast_ensemble    48 independent blast-wave runs, 8..4096 blocks of 8^3
                    each (seeded log-uniform), MC, HLLD, per-run dt
  kh_two_level      Kelvin-Helmholtz shear layer on a 28^3 root grid of 8^3
                    blocks refined inside r < 0.25 of the centre (32,256
                    blocks), MC beta 1.5, HLLD + Linde, 200 steps
  magnetosphere_256 4096 blocks of 16^3 with float sides, a dipole B0 split
                    off the state, an ionosphere body of radius 0.06 at the
                    centre, solar-wind inflow state, Koren (mc3) beta 1.5,
                    HLLD + Linde, the 8-wave source with the B0 terms
The surrogate is trained on faces harvested from FP64 runs of the
Orszag-Tang and ensemble workloads, then runs in the two-level
Kelvin-Helmholtz workload on FP32 state against the plain FP64 HLLD run.
"""
from __future__ import annotations

import time

import numpy as np

from . import ics
from .amr import AmrGrid, sphere_refinement
from .core import Mlp, Session
from .ensemble import blast_ensemble
from .grid import BlockGrid

OT_SCHEME = dict(flux="hlld", limiter="mc3", beta=1.5, nStage=2, cfl=0.8)
ENS_SCHEME = dict(flux="hlld", limiter="mc", beta=1.5, nStage=2, cfl=0.8)
KH_SCHEME = dict(flux="hlld", limiter="mc", beta=1.5, nStage=2, cfl=0.8)
TUBE_SCHEME = dict(flux="rusanov", limiter="mc", beta=1.5, nStage=2, cfl=0.8)
MAG_SCHEME = dict(flux="hlld", limiter="mc3", beta=1.5, nStage=2, cfl=0.8,
                  UseDivbSource=True)
MAG_BODY = dict(BodyRho=3.0, BodyP=0.5, B1rCoef=-1.0)


def shock_tube_64():
    g = BlockGrid(nRoot=(4, 4, 4), nCell=(8, 8, 8), order="morton")
    return g, ics.brio_wu(g)


def orszag_tang_256(nRoot=16):
    g = BlockGrid(nRoot=(nRoot,) * 3, nCell=(16, 16, 16), order="morton")
    return g, ics.orszag_tang_z(g)


def kh_two_level(nRoot=28, radius=0.25):
    g = AmrGrid((nRoot,) * 3, sphere_refinement((nRoot,) * 3, radius))
    return g, ics.kelvin_helmholtz(g)


def magnetosphere_256(nRoot=16, rBody=0.06, moment=2e-4):
    """A magnetosphere-shaped run on the 256^3 box: returns the grid, the
    state U[var, block, k, j, i] (B1 = B - B0), B0_DGB (nBlock, nKg, nJg,
    nIg, 3) of a dipole at the centre with moment (0, 0, -moment), Used_GB
    (nBlock, nKg, nJg, nIg) outside the body of radius rBody and
    XyzStart_DB (nBlock, 3) relative to the centre.  The state is a
    southward-field solar wind, slowed and compressed around the body."""
    g = BlockGrid(nRoot=(nRoot,) * 3, nCell=(16, 16, 16), order="morton",
                  periodic=(False, False, False))
    n, nG = 16, g.nG
    dx = np.asarray(g.CellSize_DB, dtype=np.float64).reshape(-1, 3)
    org = g.xyzMin[None, :] + g.iRoot_DB * n * dx
    ax = [org[:, d, None] + (np.arange(-nG, n + nG) + 0.5)[None, :] *
          dx[:, d, None] - 0.5 for d in range(3)]
    x = ax[0][:, None, None, :]
    y = ax[1][:, None, :, None]
    z = ax[2][:, :, None, None]
    r2 = x * x + y * y + z * z
    used = np.ascontiguousarray(r2 > rBody ** 2)
    r5 = r2 ** 2.5
    mz = -moment
    B0 = np.empty(r2.shape + (3,))
    B0[..., 0] = 3 * mz * z * x / r5
    B0[..., 1] = 3 * mz * z * y / r5
    B0[..., 2] = (3 * mz * z * z - mz * r2) / r5
    del r5
    xyz0 = np.ascontiguousarray(org + 0.5 * dx - 0.5)
    xc, yc, zc = g.cell_centers()
    r = np.sqrt((xc - 0.5) ** 2 + (yc - 0.5) ** 2 + (zc - 0.5) ** 2)
    f = 1.0 - 0.5 * np.exp(-((r - 2 * rBody) / (2 * rBody)) ** 2)
    U = np.zeros((8,) + xc.shape)
    U[0] = 2.0 - f
    U[1] = U[0] * (-0.5) * f
    U[6] = -0.1
    U[7] = 0.3 * (2.0 - f)
    return g, U, B0, used, xyz0


def kh_uniform(nRoot=28, seed=7):
    """A Kelvin-Helmholtz layer like kh_two_level's on a uniform grid of 8^3
    blocks with its own noise seed (a run of the same family, not the
    evaluated one)."""
    g = BlockGrid(nRoot=(nRoot,) * 3, nCell=(8, 8, 8), order="morton")
    return g, ics.kelvin_helmholtz(g, seed=seed)


def harvest_training_set(model, n_target, seed=1, betaMin=0.0, jumpTol=0.05,
                         log=None, ot_root=16, ens_runs=48,
                         sources=("orszag_tang", "blast_ensemble")):
    """Faces passing the gate, with their FP64 HLLD fluxes, from FP64 runs of
    the named workloads (Orszag-Tang, the blast ensemble and optionally a
    Kelvin-Helmholtz run of the same family as kh_two_level, "kh_family")
    at several times, until about n_target samples are held.  Each source
    is asked for an equal share of what is still missing; the Orszag-Tang
    run, whose 256^3 box holds the most gated faces, comes last and makes
    up what the smaller runs could not give."""
    got = 0
    t0 = time.perf_counter()
    wanted = tuple(sources)
    sources = []
    if "blast_ensemble" in wanted:
        eg, Ue = blast_ensemble(nRun=ens_runs)
        sources.append(("blast_ensemble", eg, Ue, ENS_SCHEME,
                        eg.iBlockFirst_R))
    if "kh_family" in wanted:
        g, U = kh_uniform()
        sources.append(("kh_family", g, U, KH_SCHEME, None))
    if "orszag_tang" in wanted:
        g, U = orszag_tang_256(ot_root)
        sources.append(("orszag_tang", g, U, OT_SCHEME, None))
    for si, (name, grid, U0, scheme, runs) in enumerate(sources):
        per = (n_target - got) / (len(sources) - si)
        with Session(grid, **scheme) as s:
            if runs is not None:
                s.set_runs(runs)
            s.upload_soa(U0)
            s.init_timestep()
            nsnap = 6
            # estimate the smooth fraction on the first snapshot
            frac = 0.05
            have = 0
            for k in range(nsnap):
                want = (per - have) / (nsnap - k)
                n = model.harvest(s, frac, BetaMin=betaMin, JumpTol=jumpTol,
                                  seed=seed * 1000 + si * 100 + k)
                have += n
                if n > 0:
                    frac = min(1.0, frac * want / n) if k == 0 else frac
                s.advance(5)
            got += have
            if log:
                log("harvest %s: %d samples" % (name, have))
    if log:
        log("harvest: %d samples in %.1f s" % (got, time.perf_counter() - t0))
    return got


def init_weights(seed=0):
    """Uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) weights, zero biases, flat
    layout W1, b1, ..., W4, b4 (27,464 floats)."""
    dims = (16, 128, 128, 64, 8)
    rng = np.random.default_rng(seed)
    parts = []
    for i in range(4):
        lim = 1.0 / np.sqrt(dims[i])
        parts.append(rng.uniform(-lim, lim, dims[i] * dims[i + 1]))
        parts.append(np.zeros(dims[i + 1]))
    return np.concatenate(parts).astype(np.float32)


def build_surrogate(n_samples=20_000_000, epochs=4, batch=65536, lr=1e-3,
                    hold=0.1, seed=1, log=None, standardize=False,
                    jump_weight=True, graph=True, **harvest_kw):
    """Harvest n_samples faces and train the network on the GPU; returns the
    model and a dict of statistics (samples, held-out relative L2 error
    overall and per output, seconds)."""
    m = Mlp(init_weights(seed))
    t0 = time.perf_counter()
    n = harvest_training_set(m, n_samples, seed=seed, log=log, **harvest_kw)
    t1 = time.perf_counter()
    loss, err = m.train(epochs, batch=batch, lr=lr, hold=hold, seed=seed,
                        standardize=standardize, jump_weight=jump_weight,
                        graph=graph)
    t2 = time.perf_counter()
    stats = dict(samples=n, harvest_s=t1 - t0, train_s=t2 - t1,
                 epochs=epochs, batch=batch, lr=lr, hold=hold,
                 train_loss=loss, heldout_rel_l2=float(err[0]),
                 heldout_rel_l2_per_output=[float(e) for e in err[1:9]],
                 heldout_correction_rel_l2=float(err[9]),
                 heldout_correction_rel_l2_per_output=[float(e)
                                                       for e in err[10:]])
    if log:
        log("train: %d epochs in %.1f s, held-out rel L2 %.4f" %
            (epochs, t2 - t1, err[0]))
    return m, stats


def run_kh(grid, U0, steps, precision="f64", model=None, mode="off",
           betaMin=0.0, jumpTol=0.05, warm=0):
    """Advance the two-level Kelvin-Helmholtz workload; returns (state,
    cell-updates/s over the timed steps, network face fraction)."""
    with Session(grid, precision=precision, **KH_SCHEME) as s:
        if mode != "off":
            s.set_surrogate(model, mode, BetaMin=betaMin, JumpTol=jumpTol)
        s.upload_soa(U0)
        s.init_timestep()
        if warm:
            s.advance(warm)
        s.synchronize()
        t = time.perf_counter()
        s.advance(steps)
        s.synchronize()
        el = time.perf_counter() - t
        U = s.download_soa()
        frac = None
        if mode != "off":
            nf, nn = s.surrogate_counts()
            frac = nn / nf if nf else 0.0
    return U, grid.nCellTotal * steps / el, frac


def rel_l1(U, ref):
    """Relative L1 error per variable: sum|U-ref| / sum|ref|."""
    d = np.abs(U - ref).reshape(U.shape[0], -1).sum(1)
    return d / np.abs(ref).reshape(U.shape[0], -1).sum(1)


def main(argv=None):
    """python -m batsrus_amd.workloads [--samples N] [--epochs E]
    [--steps S]: train the surrogate, then run the two-level
    Kelvin-Helmholtz workload plain FP64, plain FP32, FP32 with the Rusanov
    flux on the smooth faces and FP32 with the network, printing the rates
    and the relative L1 error against the FP64 run."""
    import argparse
    ap = argparse.ArgumentParser(prog="python -m batsrus_amd.workloads")
    ap.add_argument("--samples", type=float, default=2e7)
    ap.add_argument("--epochs", type=int, default=100)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warm", type=int, default=0)
    ap.add_argument("--no-graph", action="store_true",
                    help="train without replaying a captured step (for "
                         "profilers that do not follow graph launches)")
    args = ap.parse_args(argv)

    def log(s):
        print(s, flush=True)
    m, st = build_surrogate(n_samples=int(args.samples), epochs=args.epochs,
                            batch=args.batch, lr=args.lr,
                            graph=not args.no_graph,
                            sources=("orszag_tang", "blast_ensemble",
                                     "kh_family"), log=log)
    try:
        g, U0 = kh_two_level()
        ref, r64, _ = run_kh(g, U0, args.steps, "f64", warm=args.warm)
        log("f64 plain: %.3e cell-updates/s" % r64)
        for mode in ("off", "rusanov", "network"):
            U, r, frac = run_kh(g, U0, args.steps, "f32", model=m,
                                mode=mode, warm=args.warm)
            log("f32 %s: %.3e cell-updates/s, network faces %s, rel L1 "
                "max %.2e" % (mode, r, frac, rel_l1(U, ref).max()))
    finally:
        m.close()


if __name__ == "__main__":
    main()
