"""The Fortran host (ISO_C_BINDING bridge ModHipBridge) drives the GPU
library with the reference's routine names and time-loop shape, with B0
and an inner body among the options; its result matches the CPU
oracle."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import pyoracle as po
from batsrus_amd.grid import BlockGrid
from helpers import b0_dgb, body_mask, smooth_mhd_soa, rel_err
from test_gpu_body import dipole

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "batsrus_amd", "fortran", "bats_fortran_driver")


# adaptive low order faces with thresholds that give the smooth state
# weights between 0 and 1 on many faces
LOW = dict(UseAdaptiveLowOrder=True, nLowOrder=1, pCritLow=1.3,
           pCritHigh=1.0, VelCrit=0.0)


# fixed x- and inflow x+ sides (iTypeCellBc 9, 10) with their ghost states
BC_STATES = np.array([[1.0 + 0.1 * s, 0.2, -0.1 * s, 0.05, 0.4, -0.3, 0.1,
                       0.8 + 0.05 * s] for s in range(6)])


@pytest.mark.parametrize("option", ["none", "powell", "loworder", "fixedbc",
                                    "reschange", "outflow", "steady"])
@pytest.mark.parametrize("loop", ["blocks", "fused"])
def test_fortran_driver_matches_oracle(loop, option):
    print(f"case fortran_driver loop={loop} option={option}", flush=True)
    powell = option == "powell"
    low = LOW if option == "loworder" else {}
    bc = {}
    extra = ""
    beta = 1.5
    if option == "reschange":
        # the sharper limiter at the float x sides
        beta = 1.9
        bc = dict(TypeCellBc=[5, 5, 4, 4, 4, 4], BetaLimiterResChange=1.0,
                  nFaceLimiterResChange=2)
        extra = ", BetaLimiterResChange=1.0, nFaceLimiterResChange=2"
    if option == "steady":
        bc = dict(IsSteadyState=True)
        extra = ", IsTimeAccurate=.false."
    if option == "outflow":
        # reflecting x- side, outflow x+ side with a fixed pressure
        bc = dict(TypeCellBc=[7, 6, 4, 4, 4, 4], UseOutflowPressure=True,
                  pOutflow=0.9)
        extra = (", iTypeCellBc=7,6,4,4,4,4, UseOutflowPressure=.true., "
                 "pOutflow=0.9")
    if option == "fixedbc":
        bc = dict(TypeCellBc=[9, 10, 4, 4, 4, 4], CellState_VI=BC_STATES)
        extra = (", iTypeCellBc=9,10,4,4,4,4, CellState_VI=" +
                 ",".join("%.17g" % v for v in BC_STATES.reshape(-1)))
    if not os.path.exists(EXE):
        pytest.fail("Fortran driver not built (build() compiles it)")
    grid = BlockGrid(nRoot=(2, 2, 2), nCell=(8, 8, 8),
                     periodic=(not bc or option == "steady", True, True))
    U0 = smooth_mhd_soa(grid, seed=21)
    S0 = grid.soa_to_state_vgb(U0)
    grid.fill_ghosts(S0)
    nstep = 3
    with tempfile.TemporaryDirectory() as d:
        with open(os.path.join(d, "PARAM.nml"), "w") as f:
            f.write("&RUN nRootI=2, nRootJ=2, nRootK=2, nCell=8, nStep=%d, "
                    "nOrder=2, nStage=2, iLimiter=3, iFlux=3, "
                    "LimiterBeta=%g, Cfl=0.8, TypeLoop='%s', "
                    "UseDivbSource=%s%s%s /\n"
                    % (nstep, beta, loop, ".true." if powell else ".false.",
                       (", UseAdaptiveLowOrder=.true., nLowOrder=1, "
                        "pCritLow=1.3, pCritHigh=1.0, VelCrit=0.0")
                       if low else "", extra))
        with open(os.path.join(d, "grid.bin"), "wb") as f:
            f.write(np.ascontiguousarray(grid.iNeighbor_EB, np.int32).tobytes())
            f.write(np.ascontiguousarray(grid.CellSize_DB).tobytes())
        with open(os.path.join(d, "state_in.bin"), "wb") as f:
            f.write(S0.tobytes())
        r = subprocess.run([EXE, d], capture_output=True, text=True,
                           timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        S = np.fromfile(os.path.join(d, "state_out.bin")).reshape(S0.shape)
    cfg = po.config_from_dict(dict(flux="hlld", limiter="mc3", beta=beta,
                                   cfl=0.8, UseDivbSource=powell, **low,
                                   **bc),
                              grid)
    o = po.Session(grid, cfg, S0)
    o.init_timestep()
    for _ in range(nstep):
        assert o.advance_step() == 0
    assert rel_err(grid.state_vgb_to_soa(S), o.soa()) <= 1e-12


@pytest.mark.parametrize("loop", ["blocks", "fused"])
def test_fortran_driver_with_b0_and_body_matches_oracle(loop):
    """A magnetosphere-shaped run from the Fortran host: a dipole B0 at the
    centre of an inner body (B0_DGB, Used_GB and XyzStart_DB passed as the
    reference holds them), float sides, HLLD with the 8-wave source and
    the B0 source terms, through the per-block loop and the fused entry
    points."""
    print(f"case fortran_driver b0+body loop={loop}", flush=True)
    if not os.path.exists(EXE):
        pytest.fail("Fortran driver not built (build() compiles it)")
    grid = BlockGrid(nRoot=(3, 3, 3), nCell=(8, 8, 8),
                     periodic=(False, False, False))
    used, xyz0 = body_mask(grid, 0.13)
    B0 = b0_dgb(grid, dipole)
    body = dict(BodyRho=3.0, BodyP=2.0, B1rCoef=-1.0)
    U0 = smooth_mhd_soa(grid, seed=22)
    S0 = grid.soa_to_state_vgb(U0)
    grid.fill_ghosts(S0)
    nstep = 3
    with tempfile.TemporaryDirectory() as d:
        with open(os.path.join(d, "PARAM.nml"), "w") as f:
            f.write("&RUN nRootI=3, nRootJ=3, nRootK=3, nCell=8, nStep=%d, "
                    "nOrder=2, nStage=2, iLimiter=3, iFlux=3, "
                    "LimiterBeta=1.5, Cfl=0.8, TypeLoop='%s', "
                    "UseDivbSource=.true., UseB0=.true., "
                    "UseB0Source=.true., UseBody=.true., BodyRho=3.0, "
                    "BodyP=2.0, B1rCoef=-1.0 /\n" % (nstep, loop))
        with open(os.path.join(d, "grid.bin"), "wb") as f:
            f.write(np.ascontiguousarray(grid.iNeighbor_EB, np.int32).tobytes())
            f.write(np.ascontiguousarray(grid.CellSize_DB).tobytes())
        with open(os.path.join(d, "state_in.bin"), "wb") as f:
            f.write(S0.tobytes())
        with open(os.path.join(d, "b0.bin"), "wb") as f:
            f.write(np.ascontiguousarray(B0).tobytes())
        with open(os.path.join(d, "body.bin"), "wb") as f:
            f.write(np.ascontiguousarray(used, np.int8).tobytes())
            f.write(np.ascontiguousarray(xyz0).tobytes())
        r = subprocess.run([EXE, d], capture_output=True, text=True,
                           timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        S = np.fromfile(os.path.join(d, "state_out.bin")).reshape(S0.shape)
    cfg = po.config_from_dict(dict(flux="hlld", limiter="mc3", beta=1.5,
                                   cfl=0.8, UseDivbSource=True), grid)
    o = po.Session(grid, cfg, S0, B0_DGB=B0, UseB0Source=True, Used_GB=used,
                   XyzStart_DB=xyz0, **body)
    o.init_timestep()
    for _ in range(nstep):
        assert o.advance_step() == 0
    assert rel_err(grid.state_vgb_to_soa(S), o.soa()) <= 1e-12
