"""A small magnetosphere-like run that combines the pieces of the fast
update a global-magnetosphere setup uses at once: a dipole B0 centred in an
inner body with the ionosphere face boundary, a time-varying solar-wind
inflow on the x+ side (the table crosses its points during the run), float
sides elsewhere, HLLD with the MC limiter and the 8-wave source with the
B0 source terms; time accurate and steady state.  Stage by stage against
the oracle within 1e-12; body cells stay at rest."""
import numpy as np
import pytest

import pyoracle as po
from batsrus_amd.core import Session
from batsrus_amd.grid import BlockGrid
from helpers import body_mask, rel_err
from test_gpu_b0 import b0_dgb

pytestmark = pytest.mark.gpu


def _dipole(x, y, z, c=(0.5, 0.5, 0.5), m=(0.0, 0.0, -1.0), scale=2e-3):
    rx, ry, rz = x - c[0], y - c[1], z - c[2]
    r2 = rx * rx + ry * ry + rz * rz
    r5 = r2 ** 2.5
    mr = m[0] * rx + m[1] * ry + m[2] * rz
    return ((3 * mr * rx - m[0] * r2) / r5 * scale,
            (3 * mr * ry - m[1] * r2) / r5 * scale,
            (3 * mr * rz - m[2] * r2) / r5 * scale)


def _wind(rho=1.0, ux=-0.8, bz=-0.1, p=0.3):
    return np.array([rho, ux, 0.0, 0.0, 0.0, 0.0, bz, p])


@pytest.mark.parametrize("steady", [False, True])
def test_magnetosphere_like_run_matches_oracle(steady):
    print(f"case magnetosphere steady={steady}", flush=True)
    g = BlockGrid(nRoot=(3, 3, 3), nCell=(8, 8, 8),
                  periodic=(False, False, False), order="morton")
    used, xyz0 = body_mask(g, 0.12)
    G = g.nG
    inner = used[:, G:-G, G:-G, G:-G]
    # the solar wind everywhere, slowed and compressed near the body
    x, y, z = g.cell_centers()
    r = np.sqrt((x - 0.5) ** 2 + (y - 0.5) ** 2 + (z - 0.5) ** 2)
    U0 = np.zeros((8,) + x.shape)
    w = _wind()
    f = 1.0 - 0.5 * np.exp(-((r - 0.15) / 0.08) ** 2)
    U0[0] = w[0] * (2.0 - f)
    U0[1] = U0[0] * w[1] * f
    U0[6] = w[6]
    U0[7] = w[7] * (2.0 - f)
    # solar-wind table: a density pulse and a southward turning
    T = np.array([0.0, 0.006, 0.012, 0.03])
    S = np.array([_wind(), _wind(rho=1.4, p=0.4), _wind(bz=-0.3),
                  _wind(rho=0.9)])
    B0 = b0_dgb(g, _dipole)
    cs = np.zeros((6, 8))
    cs[1] = S[0] * np.array([1, S[0][0], S[0][0], S[0][0], 1, 1, 1, 1])
    sides = ["float", "vary", "float", "float", "float", "float"]
    scheme = dict(flux="hlld", limiter="mc", beta=1.5, cfl=0.8,
                  UseDivbSource=True, IsSteadyState=steady)
    cfg = po.config_from_dict(dict(scheme, TypeCellBc=sides,
                                   CellState_VI=cs), g)
    body = dict(BodyRho=3.0, BodyP=0.5, B1rCoef=-1.0)
    o = po.Session(g, cfg, g.soa_to_state_vgb(U0), B0_DGB=B0,
                   UseB0Source=True, Used_GB=used, XyzStart_DB=xyz0,
                   VaryTable=(T, S), **body)
    o.exchange()
    o.init_timestep()
    with Session(g, **scheme) as s:
        s.set_cell_boundary(sides, cs)
        s.set_b0(B0, UseB0Source=True)
        s.set_body(used, xyz0, **body)
        s.set_inflow_table(T, S)
        s.upload_soa(U0)
        s.init_timestep()
        for step in range(4):
            o.s.nStep += 1
            if steady:
                s.set_global_timestep()
            else:
                dto = o.set_global_timestep()
                dtg = s.set_global_timestep()
                assert abs(dtg - dto) <= 1e-12 * dto, (step, dtg, dto)
            for stage in (1, 2):
                assert o.advance_stage(stage) == 0
                if stage == 2:
                    o.exchange()
                s.advance_stage(stage)
                Ug = s.download_soa()
                e = rel_err(Ug, o.soa())
                print(f"  step {step} stage {stage} rel err {e:.2e}")
                assert e <= 1e-12, (step, stage, e)
                np.testing.assert_array_equal(Ug[:, ~inner], U0[:, ~inner])
        if not steady:
            assert o.s.tSimulation > T[1]     # the table's points crossed
