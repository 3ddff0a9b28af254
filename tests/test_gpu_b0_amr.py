"""B0 splitting and the inner body on two-level grids on the GPU: the
fused stages (the fine-face pass with the fine blocks' B0, the coarse
faces at resolution changes with the mean of the fine faces' B0,
src/ModB0.f90:338-351,366-455) and the per-block routines in the
reference's advance_explicit order, stage by stage over two steps against
the oracle's stage update within 1e-12; time accurate and steady state,
with and without the 8-wave source and its B0 terms; an ionosphere body
with a dipole B0, inside the fine level and across resolution changes
(the fine-face pass with the body's face treatment); and the TVD
resolution-change faces with B0, the 8-wave source and the body, time
accurate and steady state."""
import numpy as np
import pytest

import pyoracle as po
from batsrus_amd.amr import AmrGrid, sphere_refinement
from batsrus_amd.core import Session
from helpers import b0_dgb, body_mask, smooth_mhd_soa, rel_err
from test_gpu_reschange_blocks import _per_block_stage

pytestmark = pytest.mark.gpu


def _grid(nCell=8):
    refine = np.zeros((2, 2, 2), bool)
    refine[0, 0, 0] = True
    refine[1, 1, 0] = True
    return AmrGrid((2, 2, 2), refine, nCell=nCell)


def _dipole(x, y, z, c=(0.35, 0.6, -0.4), m=(0.2, -0.3, 1.0), scale=0.05):
    rx, ry, rz = x - c[0], y - c[1], z - c[2]
    r2 = rx * rx + ry * ry + rz * rz
    r5 = r2 ** 2.5
    mr = m[0] * rx + m[1] * ry + m[2] * rz
    return ((3 * mr * rx - m[0] * r2) / r5 * scale,
            (3 * mr * ry - m[1] * r2) / r5 * scale,
            (3 * mr * rz - m[2] * r2) / r5 * scale)


def _compare(g, scheme, U0, B0, steps=2, per_block=False, body=None,
             steady=False):
    cfg = po.config_from_dict(scheme, g)
    kw = dict(B0_DGB=B0, UseB0Source=True)
    if body:
        used, xyz0, bp = body
        kw.update(Used_GB=used, XyzStart_DB=xyz0, **bp)
    o = po.Session(g, cfg, g.soa_to_state_vgb(U0), **kw)
    o.exchange()
    o.init_timestep()
    with Session(g, **scheme) as s:
        s.set_b0(B0, UseB0Source=True)
        if body:
            s.set_body(used, xyz0, **bp)
        s.upload_soa(U0)
        s.init_timestep()
        for step in range(steps):
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
                if per_block:
                    _per_block_stage(s, g, stage, last=stage == 2)
                else:
                    s.advance_stage(stage)
                e = rel_err(s.download_soa(), o.soa())
                print(f"  step {step} stage {stage} rel err {e:.2e}")
                assert e <= 1e-12, (step, stage, e)


FUSED = [
    dict(flux="hlld", limiter="mc", UseDivbSource=True),
    dict(flux="rusanov", limiter="mc3"),
    dict(flux="linde", limiter="beta", beta=1.8, BetaLimiterResChange=1.2,
         nFaceLimiterResChange=2, UseDivbSource=True),
    dict(flux="hlld", limiter="mc", UseDivbSource=True, IsSteadyState=True),
    # with the TVD resolution-change faces
    dict(flux="hlld", limiter="mc", UseDivbSource=True, UseTvdResChange=True),
    dict(flux="hlld", limiter="mc3", UseDivbSource=True, UseTvdResChange=True,
         IsSteadyState=True),
    dict(flux="linde", limiter="beta", beta=1.8, BetaLimiterResChange=1.2,
         nFaceLimiterResChange=2, UseTvdResChange=True),
]


@pytest.mark.parametrize("scheme", FUSED,
                         ids=lambda d: "-".join(str(v) for v in d.values()))
def test_fused_two_level_b0_matches_oracle(scheme):
    print(f"case fused two-level b0 {scheme}", flush=True)
    g = _grid()
    U0 = smooth_mhd_soa(g, seed=91, amp=0.2)
    sch = dict(dict(beta=1.5, cfl=0.8), **scheme)
    _compare(g, sch, U0, b0_dgb(g, _dipole),
             steady=scheme.get("IsSteadyState", False))


def test_fused_two_level_b0_16cube_matches_oracle():
    print("case fused two-level b0 16^3", flush=True)
    g = _grid(16)
    U0 = smooth_mhd_soa(g, seed=92, amp=0.2)
    _compare(g, dict(flux="hlld", limiter="mc3", beta=1.5, cfl=0.8,
                     UseDivbSource=True), U0, b0_dgb(g, _dipole), steps=1)


@pytest.mark.parametrize("flux,limiter,powell,tvd", [
    ("hlld", "mc", True, False), ("linde", "minmod", False, False),
    ("hlld", "mc3", True, True)])
def test_per_block_two_level_b0_matches_oracle(flux, limiter, powell, tvd):
    print(f"case per-block two-level b0 flux={flux} powell={powell} "
          f"tvd={tvd}", flush=True)
    g = AmrGrid((4, 4, 4), sphere_refinement((4, 4, 4), 0.3))
    U0 = smooth_mhd_soa(g, seed=93, amp=0.2)
    _compare(g, dict(flux=flux, limiter=limiter, beta=1.5, cfl=0.8,
                     UseDivbSource=powell, UseTvdResChange=tvd), U0,
             b0_dgb(g, _dipole), per_block=True)


def _body_grid():
    """The fine level fills [0.25, 0.75]^3; a body of radius 0.09 at the
    centre reaches only the eight central fine blocks, whose neighbours
    are all fine."""
    g = AmrGrid((4, 4, 4), sphere_refinement((4, 4, 4), 0.3))
    used, xyz0 = body_mask(g, 0.09)
    return g, used, xyz0


@pytest.mark.parametrize("tvd", [False, True])
def test_per_block_two_level_body_matches_oracle(tvd):
    print(f"case per-block two-level body across resolution changes "
          f"tvd={tvd}", flush=True)
    g, used, xyz0 = _across_grid()
    U0 = smooth_mhd_soa(g, seed=96, amp=0.2)
    B0 = b0_dgb(g, lambda x, y, z: _dipole(x, y, z, c=(0.5, 0.5, 0.5),
                                           m=(0.0, 0.0, -1.0), scale=2e-4))
    _compare(g, dict(flux="hlld", limiter="mc", beta=1.5, cfl=0.8,
                     UseDivbSource=True, UseTvdResChange=tvd), U0, B0,
             per_block=True,
             body=(used, xyz0, dict(BodyRho=3.0, BodyP=0.5, B1rCoef=-1.0)))


def _across_grid():
    """A body of radius 0.1 at the corner the refined and unrefined roots
    share: fine blocks with coarser sides and coarse blocks with finer
    sides reach it."""
    g = _grid()
    used, xyz0 = body_mask(g, 0.1)
    return g, used, xyz0


BODY_CASES = [("inside", "hlld", "mc", False, False),
              ("inside", "hlld", "mc", True, False),
              ("across", "hlld", "mc3", False, False),
              ("across", "hlld", "mc", True, False),
              ("across", "linde", "minmod", False, False),
              ("across", "hlld", "mc", False, True),
              ("across", "hlld", "mc3", True, True)]


@pytest.mark.parametrize("where,flux,limiter,steady,tvd", BODY_CASES)
def test_fused_two_level_body_with_b0_matches_oracle(where, flux, limiter,
                                                     steady, tvd):
    """tvd: with the TVD resolution-change faces (the fine faces that
    replace coarse ones and the faces one cell in take the TVD values,
    then the body's treatment)."""
    print(f"case fused two-level body {where} {flux} {limiter} "
          f"steady={steady} tvd={tvd}", flush=True)
    g, used, xyz0 = _body_grid() if where == "inside" else _across_grid()
    nb = ~used.reshape(g.nBlock, -1).all(axis=1)
    nei = g.iNeighbor_EB.reshape(-1, 6)
    kinds = {int(g.iSlab_IS[-2 - n][2]) for n in nei[nb].ravel() if n <= -2}
    if where == "inside":
        assert nb.sum() == 8 and not kinds
    else:
        assert kinds == {0, 1}
    U0 = smooth_mhd_soa(g, seed=94, amp=0.2)
    B0 = b0_dgb(g, lambda x, y, z: _dipole(x, y, z, c=(0.5, 0.5, 0.5),
                                           m=(0.0, 0.0, -1.0), scale=2e-4))
    _compare(g, dict(flux=flux, limiter=limiter, beta=1.5, cfl=0.8,
                     UseDivbSource=True, IsSteadyState=steady,
                     UseTvdResChange=tvd), U0, B0,
             body=(used, xyz0, dict(BodyRho=3.0, BodyP=0.5, B1rCoef=-1.0)),
             steady=steady)
