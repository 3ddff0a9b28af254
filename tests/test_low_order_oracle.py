"""Pins the oracle's adaptive low order faces (#ADAPTIVELOWORDER with
nLowOrder = 1 and the second order scheme): the face criterion on stencils
whose answer is fixed by hand (thresholds, the minimum sound speed, gamma,
the normal velocity, the four-cell stencil), its use as a weight between
the second and first order face values, and whole runs where it is off
everywhere (plain second order) or on everywhere (first order)."""
import numpy as np
import pytest

import pyoracle as po
from batsrus_amd.grid import BlockGrid
from helpers import random_mhd_soa

GAMMA = 5.0 / 3.0
NC = 8
LO = dict(UseAdaptiveLowOrder=True, nLowOrder=1, pCritLow=2.0,
          pCritHigh=1.5, VelCrit=2.0)


def _step_block(i0, left, right, axis=0):
    """Ghosted block (NC+4)^3 x 8 whose cells with index < i0 along axis
    (ghost cells -1, 0 included; interior 1..NC) hold the conserved state
    left and the others right."""
    c = np.arange(-1, NC + 3)
    idx = [None, None, None]
    idx[2 - axis] = c
    shp = [1, 1, 1]
    shp[2 - axis] = NC + 4
    lo = (c < i0).reshape(shp)
    blk = np.where(lo[..., None], np.asarray(left, float),
                   np.asarray(right, float))
    return np.ascontiguousarray(np.broadcast_to(blk, (NC + 4,) * 3 + (8,)))


def _state(rho=1.0, u=(0.0, 0.0, 0.0), b=(0.0, 0.0, 0.0), p=1.0):
    return [rho, rho * u[0], rho * u[1], rho * u[2], *b, p]


def _crit(left, right, i0=4, axis=0, **kw):
    cfg = po.make_config(nI=NC, nJ=NC, nK=NC, **dict(LO, **kw))
    return po.set_low_order_face(cfg, _step_block(i0, left, right, axis))


# sound speed of rho = 1, p = 1: sqrt(5/3) = 1.291
@pytest.mark.parametrize("pl,du,expect", [
    (1.75, 5.0, 0.5),   # pRatio 1.75 between pCritHigh and pCritLow
    (2.5, 5.0, 1.0),    # pRatio >= pCritLow
    (2.0, 5.0, 1.0),    # pRatio == pCritLow counts as low order
    (1.2, 5.0, 0.0),    # pRatio < pCritHigh
    (2.5, 3.0, 1.0),    # 3/c_min = 2.32 > 2; with c_max (2.04) it would not
    (2.5, 2.4, 0.0),    # 2.4/c_min = 1.86 <= 2; without gamma it would be 2.4
    (2.5, 1.0, 0.0),    # small velocity jump: high order whatever pRatio
])
def test_criterion_closed_form(pl, du, expect):
    cx, cy, cz = _crit(_state(p=pl), _state(u=(-du, 0, 0), p=1.0))
    # x faces 1..NC+1 (index f = face - 1): stencil cells f-1 .. f+2 cover
    # both states only for faces i0-1, i0, i0+1
    i0 = 4
    want = np.zeros(NC + 1)
    want[[i0 - 2, i0 - 1, i0]] = expect
    assert np.array_equal(cx, np.broadcast_to(want, cx.shape))
    assert not cy.any() and not cz.any()


def test_criterion_total_pressure_includes_b():
    """pTotal = p + B^2/2: with B = (1, 0, 0) on both sides the pressure
    ratio 3.0/1.5 = 2 becomes 3.5/2.0 = 1.75, i.e. weight 0.5, not 1."""
    cx, _, _ = _crit(_state(p=3.0, b=(1, 0, 0)),
                     _state(u=(-9.0, 0, 0), p=1.5, b=(1, 0, 0)))
    assert np.allclose(cx[..., 3], 0.5, rtol=0, atol=1e-15)


def test_criterion_uses_the_normal_velocity():
    """A jump of u_y across x faces (shear) is not a normal velocity jump:
    x faces stay high order; the same jump across y faces is flagged."""
    left, right = _state(p=2.5), _state(u=(0, -5.0, 0), p=1.0)
    cx, cy, cz = _crit(left, right, axis=0)
    assert not cx.any() and not cy.any() and not cz.any()
    cx, cy, cz = _crit(left, right, axis=1)
    assert not cx.any() and not cz.any()
    want = np.zeros(NC + 1)
    want[[2, 3, 4]] = 1.0
    assert np.array_equal(cy, np.broadcast_to(want[None, :, None], cy.shape))


def test_criterion_z_faces():
    cx, cy, cz = _crit(_state(p=1.75), _state(u=(0, 0, 5.0), p=1.0), i0=2,
                       axis=2)
    want = np.zeros(NC + 1)
    want[[0, 1, 2]] = 0.5
    assert not cx.any() and not cy.any()
    assert np.array_equal(cz, np.broadcast_to(want[:, None, None], cz.shape))


def _random_block(seed):
    g = BlockGrid(nRoot=(1, 1, 1), nCell=(NC, NC, NC))
    S = g.soa_to_state_vgb(random_mhd_soa(g, seed=seed))
    g.fill_ghosts(S)
    return np.ascontiguousarray(S[0])


@pytest.mark.parametrize("lim", ["mc", "mc3", "minmod", "beta"])
def test_weight_interpolates_first_and_second_order(lim):
    """Left = P + dL(1-w) = (1-w)(P + dL) + w P: the weighted face value is
    the blend of the second order and the first order face values."""
    blk = _random_block(7)
    cfg2 = po.make_config(nI=NC, nJ=NC, nK=NC, limiter=lim, beta=1.5, **LO)
    cfg1 = po.make_config(nI=NC, nJ=NC, nK=NC, nOrder=1, limiter=lim)
    rng = np.random.default_rng(3)
    crit = [rng.uniform(0, 1, s) for s in po.face_shapes(cfg2)]
    crit[0][..., 2] = 0.0
    crit[1][..., 5, :] = 1.0
    lo = po.calc_face_value(cfg2, blk, crit)
    f2 = po.calc_face_value(cfg2, blk)
    f1 = po.calc_face_value(cfg1, blk)
    for n in range(6):
        w = crit[n // 2][..., None]
        want = (1 - w) * f2[n] + w * f1[n]
        assert np.allclose(lo[n], want, rtol=1e-13, atol=1e-13), n
    # weight 1: exactly the first order value; weight 0: the second order
    assert np.array_equal(lo[2][:, 5, :], f1[2][:, 5, :])
    assert np.array_equal(lo[0][..., 2, :], f2[0][..., 2, :])


def test_nlow_order_2_leaves_second_order_faces():
    """nLowOrder = 2 selects the second order scheme as the low order one,
    which the second order scheme already is (:2017-2019)."""
    blk = _random_block(8)
    cfg = po.make_config(nI=NC, nJ=NC, nK=NC, **dict(LO, nLowOrder=2))
    ones = [np.ones(s) for s in po.face_shapes(cfg)]
    a = po.calc_face_value(cfg, blk, ones)
    b = po.calc_face_value(cfg, blk)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def _run(grid, U0, nstep, **kw):
    cfg = po.config_from_dict(dict(dict(limiter="mc", beta=1.5,
                                        flux="hlld"), **kw), grid)
    s = po.Session(grid, cfg, grid.soa_to_state_vgb(U0))
    s.exchange()
    assert s.init_timestep() == 0
    for _ in range(nstep):
        assert s.advance_step() == 0
    return s.soa()


def test_session_low_order_everywhere_is_first_order():
    """pCritLow = pCritHigh = 1 and VelCrit = 0 flag every face with any
    normal velocity variation as low order: the run equals the first order
    run."""
    g = BlockGrid(nRoot=(2, 2, 2), nCell=(NC, NC, NC), order="morton")
    U0 = random_mhd_soa(g, seed=11)
    lo = _run(g, U0, 3, **dict(LO, pCritLow=1.0, pCritHigh=1.0, VelCrit=0.0))
    first = _run(g, U0, 3, nOrder=1)
    assert np.allclose(lo, first, rtol=1e-14, atol=1e-14)


def test_session_low_order_never_is_second_order():
    g = BlockGrid(nRoot=(2, 2, 2), nCell=(NC, NC, NC), order="morton")
    U0 = random_mhd_soa(g, seed=12)
    lo = _run(g, U0, 3, **dict(LO, VelCrit=1e30))
    plain = _run(g, U0, 3)
    assert np.array_equal(lo, plain)


def test_session_low_order_conserves():
    """A converging-flow shock problem with the criterion active on part of
    the faces conserves mass, momentum, field and energy to rounding on a
    periodic box, and differs from the plain second order run."""
    g = BlockGrid(nRoot=(2, 2, 2), nCell=(NC, NC, NC), order="morton")
    U0 = random_mhd_soa(g, seed=13)
    x = g.cell_centers()[0]
    U0 = U0.copy()
    U0[1] = U0[0] * 3.0 * np.sign(0.5 - x) * (np.abs(x - 0.5) < 0.3)
    U0[7] = np.where(np.abs(x - 0.5) < 0.15, 3.0, 1.0) * U0[7]
    lo = _run(g, U0, 3, **LO)
    plain = _run(g, U0, 3)
    assert not np.array_equal(lo, plain)
    V = np.prod(g.CellSize_DB[0])

    def totals(U):
        E = U[7] / (GAMMA - 1) + 0.5 * (U[1] ** 2 + U[2] ** 2 + U[3] ** 2) / \
            U[0] + 0.5 * (U[4] ** 2 + U[5] ** 2 + U[6] ** 2)
        t = U.reshape(8, -1).sum(axis=1) * V
        t[7] = E.sum() * V
        return t
    t0, t1 = totals(U0), totals(lo)
    scale = np.abs(U0).reshape(8, -1).sum(axis=1) * V
    scale[7] = t0[7]
    assert (np.abs(t1 - t0) / scale).max() < 1e-13
