"""B0 splitting on the GPU (bats_set_b0): the fused stage kernel and the
per-block routines with a static B0 against the oracle's stage update
(which follows src/ModB0.f90, src/ModFaceFlux.f90 and
src/ModCalcSource.f90 with UseB0), stage by stage over two steps within
1e-12, with and without the 8-wave source and its B0 terms; and the
combinations the B0 variants do not carry are refused."""
import numpy as np
import pytest

import pyoracle as po
from batsrus_amd.core import Session, StopMpi
from batsrus_amd.grid import BlockGrid
from helpers import smooth_mhd_soa, rel_err, b0_dgb

pytestmark = pytest.mark.gpu

TP = 2 * np.pi


def b0_periodic(x, y, z):
    """A smooth B0 that is periodic on the unit box and has a curl."""
    return (0.3 + 0.1 * np.sin(TP * y) + 0.05 * np.cos(TP * z),
            0.2 + 0.08 * np.cos(TP * x),
            -0.15 + 0.1 * np.sin(TP * x + 1.0) * np.cos(TP * y))


def b0_dipole(x, y, z, c=(0.5, 0.5, -0.6), m=(0.0, 0.3, 1.0)):
    """A dipole centred outside the box (non-periodic B0)."""
    rx, ry, rz = x - c[0], y - c[1], z - c[2]
    r2 = rx * rx + ry * ry + rz * rz
    r5 = r2 ** 2.5
    mr = m[0] * rx + m[1] * ry + m[2] * rz
    return ((3 * mr * rx - m[0] * r2) / r5 * 0.05,
            (3 * mr * ry - m[1] * r2) / r5 * 0.05,
            (3 * mr * rz - m[2] * r2) / r5 * 0.05)


CASES = [
    ("rusanov", "mc", False, 8, "periodic", True, False),
    ("hlld", "mc3", True, 16, "periodic", True, False),
    ("linde", "minmod", True, 8, "dipole", True, False),
    ("hlld", "mc", True, 8, "dipole", False, False),
    ("hlld", "mc", True, 8, "dipole", True, True),
]


@pytest.mark.parametrize("flux,limiter,powell,NB,field,b0src,steady", CASES)
def test_fused_stages_with_b0_match_oracle(flux, limiter, powell, NB, field,
                                           b0src, steady):
    """steady: local time stepping (#TIMEACCURATE F), every cell with its
    own step from the first stage's fluxes."""
    print(f"case b0 fused flux={flux} lim={limiter} powell={powell} NB={NB} "
          f"field={field} b0source={b0src} steady={steady}", flush=True)
    periodic = field == "periodic"
    grid = BlockGrid(nRoot=(2, 2, 2) if NB == 16 else (3, 2, 2),
                     nCell=(NB, NB, NB),
                     periodic=(True, True, True) if periodic
                     else (False, True, True), order="morton")
    U0 = smooth_mhd_soa(grid, seed=90 + NB)
    B0 = b0_dgb(grid, b0_periodic if periodic else b0_dipole)
    scheme = dict(flux=flux, limiter=limiter, beta=1.5, cfl=0.8,
                  UseDivbSource=powell, IsSteadyState=steady)
    cfg = po.config_from_dict(scheme, grid)
    S = grid.soa_to_state_vgb(U0)
    grid.fill_ghosts(S)
    o = po.Session(grid, cfg, S, B0_DGB=B0, UseB0Source=b0src)
    o.init_timestep()
    with Session(grid, **scheme) as s:
        s.set_b0(B0, UseB0Source=b0src)
        s.upload_soa(U0)
        s.init_timestep()
        for step in range(2):
            o.s.nStep += 1
            if steady:
                s.set_global_timestep()     # counts the step
            else:
                dto = o.set_global_timestep()
                dtg = s.set_global_timestep()
                assert abs(dtg - dto) <= 1e-12 * dto, (step, dtg, dto)
            for stage in (1, 2):
                assert o.advance_stage(stage) == 0
                if stage == 2:
                    o.exchange()
                s.advance_stage(stage)
                e = rel_err(s.download_soa(), o.soa())
                print(f"  step {step} stage {stage} rel err {e:.2e}")
                assert e <= 1e-12, (step, stage, e)


def test_b0_changes_the_result():
    """The same run without B0 differs (the B0 enters the fluxes)."""
    print("case b0 changes the result", flush=True)
    grid = BlockGrid(nRoot=(2, 2, 2), nCell=(8, 8, 8), order="morton")
    U0 = smooth_mhd_soa(grid, seed=95)
    B0 = b0_dgb(grid, b0_periodic)
    out = []
    for use in (True, False):
        with Session(grid, flux="hlld", limiter="mc", beta=1.5) as s:
            if use:
                s.set_b0(B0)
            s.upload_soa(U0)
            s.init_timestep()
            s.advance(2)
            out.append(s.download_soa())
    assert rel_err(out[0], out[1]) > 1e-4


def test_per_block_routines_with_b0_match_oracle():
    print("case b0 per-block", flush=True)
    grid = BlockGrid(nRoot=(2, 2, 2), nCell=(8, 8, 8), order="morton")
    U0 = smooth_mhd_soa(grid, seed=96)
    B0 = b0_dgb(grid, b0_periodic)
    scheme = dict(flux="hlld", limiter="mc3", beta=1.5, cfl=0.8,
                  UseDivbSource=True)
    cfg = po.config_from_dict(scheme, grid)
    S = grid.soa_to_state_vgb(U0)
    grid.fill_ghosts(S)
    o = po.Session(grid, cfg, S, B0_DGB=B0)
    o.init_timestep()
    with Session(grid, **scheme) as s:
        s.set_b0(B0)
        s.upload_soa(U0)
        s.init_timestep()
        for step in range(2):
            o.s.nStep += 1
            o.set_global_timestep()
            s.set_global_timestep()
            for stage in (1, 2):
                assert o.advance_stage(stage) == 0
                if stage == 2:
                    o.exchange()
                s.exchange_messages()
                s.set_stage(stage)
                for b in range(1, grid.nBlock + 1):
                    s.calc_face_value(b)
                    s.calc_face_flux(False, b)
                    s.calc_source(b)
                    s.update_state(b)
                    if stage == 2:
                        s.calc_timestep(b)
                e = rel_err(s.download_soa(), o.soa())
                print(f"  step {step} stage {stage} rel err {e:.2e}")
                assert e <= 1e-12, (step, stage, e)


def test_b0_refusals():
    print("case b0 refusals", flush=True)
    grid = BlockGrid(nRoot=(2, 2, 2), nCell=(8, 8, 8), order="morton")
    B0 = b0_dgb(grid, b0_periodic)
    with Session(grid, flux="hlld", limiter="mc", precision="f32") as s:
        with pytest.raises(StopMpi):
            s.set_b0(B0)
    # fixed sides with B0 are accepted (their ghosts hold CellState - B0)
    g2 = BlockGrid(nRoot=(2, 2, 2), nCell=(8, 8, 8),
                   periodic=(False, True, True))
    with Session(g2, flux="hlld", limiter="mc") as s:
        s.set_b0(b0_dgb(g2, b0_dipole))
        cs = np.tile(np.array([1.0, 0, 0, 0, 0, 0, 0, 1.0]), (6, 1))
        s.set_cell_boundary(["fixed"] * 2 + ["periodic"] * 4, cs)
    with Session(grid, flux="hlld", limiter="mc",
                 UseAdaptiveLowOrder=True) as s:
        with pytest.raises(StopMpi):
            s.set_b0(B0)
