"""Pins the oracle's sharper limiter next to sides whose DiLevel_EB is not
zero (#RESOLUTIONCHANGE BetaLimiterResChange / nFaceLimiterResChange,
src/ModFaceValue.f90:1982-2015): which cells of each pencil take the
sharper beta (hand-listed cell sets for several nFace), that a face value
depends only on its own cell's beta, the side and direction mapping, the
gate LimiterBeta > BetaLimiterResChange, and whole two-level runs."""
import numpy as np
import pytest

import pyoracle as po
from batsrus_amd.amr import AmrGrid, sphere_refinement
from helpers import smooth_mhd_soa

NC = 8


def _block(seed):
    rng = np.random.default_rng(seed)
    b = rng.uniform(0.5, 1.5, size=(NC + 4,) * 3 + (8,))
    b[..., 1:4] *= rng.uniform(-1, 1, size=(NC + 4,) * 3 + (3,))
    b[..., 4:7] = rng.uniform(-1, 1, size=(NC + 4,) * 3 + (3,))
    return b


# cells (1-based pencil index 0..nI+1, 0 and nI+1 are ghosts) whose slopes
# use BetaLimiterResChange, for nI = 8
CELLS = {  # nFace: (low side, high side)
    1: ({0, 1}, {8, 9}),
    2: ({0, 1, 2}, {7, 8, 9}),
    3: ({0, 1, 2, 3}, {6, 7, 8, 9}),
    12: (set(range(10)), set(range(10))),
}


@pytest.mark.parametrize("limiter", ["mc", "mc3", "beta"])
@pytest.mark.parametrize("nface", sorted(CELLS))
@pytest.mark.parametrize("side", range(6))
def test_sharp_cells_per_side(limiter, nface, side):
    blk = _block(side + 7 * nface)
    beta, brc = 1.9, 1.1
    kw = dict(nI=NC, nJ=NC, nK=NC, limiter=limiter,
              BetaLimiterResChange=brc, nFaceLimiterResChange=nface)
    plain = po.calc_face_value(po.make_config(beta=beta, **kw), blk)
    sharp = po.calc_face_value(po.make_config(beta=brc, **kw), blk)
    rc = [0] * 6
    rc[side] = 1
    got = po.calc_face_value(po.make_config(beta=beta, **kw), blk,
                             res_change=rc)
    d = side // 2
    cells = CELLS[nface][side % 2]
    # face arrays (k, j, i, v); the face index along d is axis 2 - d
    ax = 2 - d
    for w, (p, s_, g) in enumerate(zip(plain, sharp, got)):
        if w // 2 != d:               # other directions are untouched
            assert np.array_equal(g, p), w
            continue
        left = w % 2 == 0
        for f in range(NC + 1):       # face f+1 (1-based) between cells f, f+1
            cell = f if left else f + 1
            want = s_ if cell in cells else p
            sl = [slice(None)] * 4
            sl[ax] = f
            assert np.array_equal(g[tuple(sl)], want[tuple(sl)]), \
                (w, f, cell)
    # the sharper beta does change these faces
    assert not np.array_equal(plain[2 * d], got[2 * d]) or \
        not np.array_equal(plain[2 * d + 1], got[2 * d + 1])


def test_gate_and_off_switches():
    blk = _block(3)
    base = dict(nI=NC, nJ=NC, nK=NC, limiter="mc")
    ref = po.calc_face_value(po.make_config(beta=1.5, **base), blk)
    for kw in (dict(beta=1.5, BetaLimiterResChange=2.0),   # beta <= BetaRC
               dict(beta=1.5, BetaLimiterResChange=1.5),   # equal: off
               dict(beta=1.5, BetaLimiterResChange=1.0,
                    nFaceLimiterResChange=0)):              # no faces
        got = po.calc_face_value(po.make_config(**base, **kw), blk,
                                 res_change=[1] * 6)
        for a, b in zip(got, ref):
            assert np.array_equal(a, b)
    # minmod ignores beta
    mm = dict(nI=NC, nJ=NC, nK=NC, limiter="minmod")
    a = po.calc_face_value(po.make_config(beta=1.9, **mm), blk)
    b = po.calc_face_value(po.make_config(beta=1.9, BetaLimiterResChange=1.0,
                                          **mm), blk, res_change=[1] * 6)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_two_level_run_conserves_and_differs():
    g = AmrGrid((4, 4, 4), sphere_refinement((4, 4, 4), 0.3))
    U0 = smooth_mhd_soa(g, seed=11, amp=0.3)
    res = []
    for brc in (2.0, 1.0):
        cfg = po.make_config(nI=g.nI, nJ=g.nJ, nK=g.nK, flux="hlld",
                             limiter="mc", beta=1.8,
                             BetaLimiterResChange=brc)
        o = po.Session(g, cfg, g.soa_to_state_vgb(U0))
        o.init_timestep()
        for _ in range(2):
            assert o.advance_step() == 0
        res.append(o.soa())
    assert not np.array_equal(res[0], res[1])
    vol = np.prod(g.CellSize_DB, axis=1)[None, :, None, None, None]
    for U in res:
        tot0 = (U0[[0, 1, 2, 3, 4, 5, 6]] * vol).sum(axis=(1, 2, 3, 4))
        tot1 = (U[[0, 1, 2, 3, 4, 5, 6]] * vol).sum(axis=(1, 2, 3, 4))
        scale = (np.abs(U0[:7]) * vol).sum(axis=(1, 2, 3, 4))
        assert (np.abs(tot1 - tot0) / scale).max() < 1e-13
