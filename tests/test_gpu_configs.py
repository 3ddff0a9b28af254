"""The benchmark workloads under the GPU tests: the Orszag-Tang-z vortex
(4096 blocks of 16^3, Koren beta 1.5, HLLD + Linde fix, 2 stages) and the
two-level Kelvin-Helmholtz layer (32,256 blocks of 8^3 refined inside
r < 0.25, MC beta 1.5, HLLD + Linde fix).  On reduced grids their initial
states advance stage by stage within 1e-12 of the oracle; on the full grids
every conserved quantity (mass, momentum, magnetic flux, energy) is kept to
rounding (FP64, FP32 state, and the FP32 surrogate stage with the network
on the smooth faces).

The stage differences are measured per physical quantity (rho, momentum
vector, field vector, p): the vortex's B_z has an exactly cancelling flux
divergence and stays at rounding level, so against its own size any two
correct evaluations differ by O(1)."""
import numpy as np
import pytest

import mlp_oracle as mo
import pyoracle as po
from batsrus_amd import ics
from batsrus_amd.amr import AmrGrid, sphere_refinement
from batsrus_amd.core import Mlp, Session
from batsrus_amd.grid import BlockGrid
from batsrus_amd.workloads import KH_SCHEME, OT_SCHEME, kh_two_level, \
    orszag_tang_256
from helpers import rel_err_vec

pytestmark = pytest.mark.gpu


def _stage_parity(grid, U0, scheme, nstep=2, tol=1e-12):
    cfg = po.config_from_dict(scheme, grid)
    o = po.Session(grid, cfg, grid.soa_to_state_vgb(U0))
    o.init_timestep()
    worst = 0.0
    with Session(grid, **scheme) as s:
        s.upload_soa(U0)
        s.init_timestep()
        for step in range(nstep):
            o.s.nStep += 1
            dto = o.set_global_timestep()
            dtg = s.set_global_timestep()
            assert abs(dtg - dto) <= tol * dto, (step, dtg, dto)
            for stage in range(1, cfg.nStage + 1):
                assert o.advance_stage(stage) == 0
                if stage == cfg.nStage:
                    o.exchange()
                s.advance_stage(stage)
                e = rel_err_vec(s.download_soa(), o.soa())
                worst = max(worst, e)
                assert e <= tol, (step, stage, e)
    print(f"  worst stage difference {worst:.2e}")


def test_orszag_tang_ic_reduced_grid():
    """The Orszag-Tang-z initial state on 2x2x2 blocks of 16^3."""
    print("case configs orszag-tang reduced", flush=True)
    g = BlockGrid(nRoot=(2, 2, 2), nCell=(16, 16, 16), order="morton")
    _stage_parity(g, ics.orszag_tang_z(g), dict(OT_SCHEME))


def test_kelvin_helmholtz_two_level_reduced_grid():
    """The two-level Kelvin-Helmholtz state on a 4^3 root grid whose eight
    central roots are refined (r < 0.25)."""
    print("case configs kelvin-helmholtz reduced", flush=True)
    g = AmrGrid((4, 4, 4), sphere_refinement((4, 4, 4), 0.25))
    assert g.nBlock == 56 + 64
    _stage_parity(g, ics.kelvin_helmholtz(g), dict(KH_SCHEME))


def _conserved_scale(grid, U):
    """sum over cells of |U| * volume per conserved quantity (energy from
    the pressure; summed over the components for the momentum and field
    vectors), the scale the integrals' drift is measured against."""
    V = np.prod(np.asarray(grid.CellSize_DB), axis=1)
    w = V[None, :, None, None, None]
    gm1 = 5.0 / 3.0 - 1
    E = U[7] / gm1 + 0.5 * (U[1] ** 2 + U[2] ** 2 + U[3] ** 2) / U[0] + \
        0.5 * (U[4] ** 2 + U[5] ** 2 + U[6] ** 2)
    sc = np.array([np.sum(np.abs(U[v]) * w[0]) for v in range(7)] +
                  [np.sum(E * w[0])])
    # a vector's components share the vector's scale (B_z of the vortex and
    # B_y, B_z of the shear layer start at zero)
    sc[1:4] = sc[1:4].sum()
    sc[4:7] = sc[4:7].sum()
    return sc


def _drift(s, t0, scale, nstep):
    s.advance(nstep)
    s.check_nan()                     # raises StopMpi on a NaN
    t1 = s.integrate_domain()
    return np.abs(t1 - t0) / scale


def test_orszag_tang_full_grid_conserves():
    """4096 blocks of 16^3 for 50 steps in FP64."""
    print("case configs orszag-tang full 50 steps", flush=True)
    g, U0 = orszag_tang_256()
    scale = _conserved_scale(g, U0)
    with Session(g, **OT_SCHEME) as s:
        s.upload_soa(U0)
        s.init_timestep()
        t0 = s.integrate_domain()
        d = _drift(s, t0, scale, 50)
        _, t, n = s.time()
    print(f"  t = {t:.4e} after {n} steps, drift per quantity "
          f"{np.array2string(d, formatter={'float': '{:.1e}'.format})}")
    assert n == 50 and t > 0
    assert d.max() < 1e-12


@pytest.fixture(scope="module")
def kh_full():
    g, U0 = kh_two_level()
    return g, U0, _conserved_scale(g, U0)


@pytest.mark.parametrize("mode", ["f64", "f32", "network"])
def test_kelvin_helmholtz_full_grid_conserves(mode, kh_full):
    """32,256 blocks of 8^3 on two levels; the coarse faces of resolution
    changes take the fine fluxes, so the update stays conservative in FP64,
    with FP32 state, and with the network surrogate on the smooth faces
    (one flux per face)."""
    print(f"case configs kelvin-helmholtz full {mode}", flush=True)
    g, U0, scale = kh_full
    assert g.nBlock == 32256
    prec = "f64" if mode == "f64" else "f32"
    nstep = 20 if mode != "network" else 4
    with Session(g, precision=prec, **KH_SCHEME) as s, \
            Mlp(mo.init_params(seed=3)) as m:
        if mode == "network":
            s.set_surrogate(m, "network", BetaMin=0.0, JumpTol=0.05)
        s.upload_soa(U0)
        s.init_timestep()
        t0 = s.integrate_domain()
        d = _drift(s, t0, scale, nstep)
        if mode == "network":
            nf, nn = s.surrogate_counts()
            print(f"  network faces {nn} of {nf}")
            assert nn > 0.5 * nf
    print(f"  drift per quantity "
          f"{np.array2string(d, formatter={'float': '{:.1e}'.format})}")
    assert d.max() < (1e-12 if mode == "f64" else 1e-7)
