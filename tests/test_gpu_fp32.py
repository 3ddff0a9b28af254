"""FP32-state path: state and face arithmetic in FP32, stage update, pressure
fix and time step in FP64.  Its error against the FP64 oracle is stated as
relative L1 and L-infinity norms per variable (printed), and bounded."""
import numpy as np
import pytest

import pyoracle as po
from batsrus_amd.core import Session, StopMpi
from batsrus_amd.grid import BlockGrid
from batsrus_amd.ensemble import blast_ensemble
from batsrus_amd import ics
from helpers import smooth_mhd_soa

pytestmark = pytest.mark.gpu


def l1_linf(a, b):
    """Per-variable relative L1 (sum|a-b|/sum|b|) and L-inf (max|a-b|/max|b|)."""
    a = a.reshape(8, -1)
    b = b.reshape(8, -1)
    d = np.abs(a - b)
    return (d.sum(1) / np.abs(b).sum(1)), (d.max(1) / np.abs(b).max(1))


def _ghosted(grid, U):
    S = grid.soa_to_state_vgb(U)
    grid.fill_ghosts(S)
    return S


@pytest.mark.parametrize("NB", [8, 16])
@pytest.mark.parametrize("flux,limiter", [("hlld", "mc3"), ("rusanov", "mc"),
                                          ("linde", "minmod")])
def test_fp32_state_vs_fp64_oracle(flux, limiter, NB):
    print(f"case fp32 flux={flux} lim={limiter} NB={NB}", flush=True)
    grid = BlockGrid(nRoot=(2, 2, 2), nCell=(NB, NB, NB), order="morton")
    U0 = smooth_mhd_soa(grid, seed=50 + NB)
    scheme = dict(flux=flux, limiter=limiter, beta=1.5, cfl=0.8)
    o = po.Session(grid, po.config_from_dict(scheme, grid), _ghosted(grid, U0))
    o.init_timestep()
    nstep = 4
    for _ in range(nstep):
        assert o.advance_step() == 0
    with Session(grid, precision="f32", **scheme) as s:
        s.upload_soa(U0)
        s.init_timestep()
        s.advance(nstep)
        U = s.download_soa()
        dt, t, n = s.time()
    l1, linf = l1_linf(U, o.soa())
    print("  rel L1  ", np.array2string(l1, precision=2))
    print("  rel Linf", np.array2string(linf, precision=2))
    assert abs(t - o.tSimulation) <= 1e-6 * o.tSimulation
    assert l1.max() < 2e-6 and linf.max() < 5e-5


def test_fp32_conservation_and_uniform():
    grid = BlockGrid(nRoot=(2, 2, 2), nCell=(8, 8, 8))
    U0 = smooth_mhd_soa(grid, seed=5)
    with Session(grid, precision="f32", flux="hlld", limiter="mc3") as s:
        s.upload_soa(U0)
        s.init_timestep()
        t0 = s.integrate_domain()
        s.advance(5)
        t1 = s.integrate_domain()
    scale = np.abs(t0)
    scale[1:7] = np.maximum(scale[1:7], abs(t0[0]))
    assert (np.abs(t1 - t0) / scale).max() < 1e-6
    Uu = ics.uniform_state(grid, rho=1.2, u=(0.1, 0.2, -0.3),
                           b=(0.4, -0.1, 0.2), p=0.9)
    with Session(grid, precision="f32", flux="hlld", limiter="mc3") as s:
        s.upload_soa(Uu)
        s.init_timestep()
        s.advance(3)
        U = s.download_soa()
    assert np.array_equal(U, Uu.astype(np.float32).astype(np.float64))


def test_fp32_ensemble_vs_fp64():
    eg, U0 = blast_ensemble(nRun=6, nb_min=8, nb_max=64, seed=21)
    out = {}
    for prec in ("f64", "f32"):
        with Session(eg, precision=prec, flux="hlld", limiter="mc",
                     beta=1.5, cfl=0.8) as s:
            s.set_runs(eg.iBlockFirst_R)
            s.upload_soa(U0)
            s.init_timestep()
            s.advance(5)
            out[prec] = s.download_soa()
    for r in range(eg.nRun):
        sl = eg.run_slice(r)
        l1, linf = l1_linf(out["f32"][:, sl], out["f64"][:, sl])
        print(f"  run {r}: rel L1 max {l1.max():.2e}, rel Linf max "
              f"{linf.max():.2e}")
        assert l1.max() < 1e-5


def test_fp32_per_block_routines_refused():
    grid = BlockGrid(nRoot=(1, 1, 1), nCell=(8, 8, 8))
    with Session(grid, precision="f32") as s:
        s.upload_soa(smooth_mhd_soa(grid))
        with pytest.raises(StopMpi):
            s.calc_face_value(1)
