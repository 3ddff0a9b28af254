"""Outer boundaries of the fast update that depend on time or B0, in the
oracle (set_boundary_fast and set_boundary_for_cell,
src/ModUpdateStateFast.f90:1151-1158,1262-1268): the x+ side of type vary
takes the inflow table's state at the current time (get_solar_wind_point
without a satellite offset, src/ModSolarwind.f90:591-740: linear in time,
the end values outside the table), and with B0 splitting the ghost cells
of fixed, inflow and vary sides hold B1 = CellState B - B0.  Pinned by:

* the table state is exact at the nodes, linear between them, clamped
  outside, and turns velocity into momentum;
* a constant table is the fixed boundary bit for bit;
* after an exchange at time t the x+ ghost cells hold the table's state
  at t;
* with a uniform B0 the split run (fixed sides, CellState with the full
  field) and the unsplit run of the full field B1 + B0 are the same
  equations: their states agree to rounding over two steps, with and
  without the 8-wave source."""
import numpy as np
import pytest

import pyoracle as po
from batsrus_amd.grid import BlockGrid
from helpers import smooth_mhd_soa

RNG = np.random.default_rng(20261206)


def _table():
    T = np.array([0.0, 0.01, 0.03, 0.05])
    S = np.array([[1.0, -0.5, 0.1, 0.0, 0.2, 0.3, -0.1, 0.6],
                  [1.2, -0.6, 0.0, 0.1, 0.25, 0.3, -0.1, 0.7],
                  [0.9, -0.4, 0.05, 0.0, 0.2, 0.35, 0.0, 0.5],
                  [1.1, -0.5, 0.1, -0.05, 0.2, 0.3, -0.05, 0.65]])
    return T, S


def _cons(P):
    C = np.array(P, dtype=np.float64)
    C[1:4] = C[1:4] * C[0]
    return C


def test_vary_state_interpolates_the_table():
    T, S = _table()
    for i in range(4):
        np.testing.assert_array_equal(po.vary_state(T, S, T[i]), _cons(S[i]))
    for t in RNG.uniform(0, 0.05, 20):
        i = min(np.searchsorted(T, t, side="right") - 1, 2)
        w = (t - T[i]) / (T[i + 1] - T[i])
        P = S[i] + w * (S[i + 1] - S[i])
        np.testing.assert_allclose(po.vary_state(T, S, t), _cons(P),
                                   rtol=1e-14, atol=1e-15)
    np.testing.assert_array_equal(po.vary_state(T, S, -1.0), _cons(S[0]))
    np.testing.assert_array_equal(po.vary_state(T, S, 9.0), _cons(S[-1]))
    np.testing.assert_array_equal(po.vary_state(T[:1], S[:1], 3.0),
                                  _cons(S[0]))


def _grid():
    return BlockGrid(nRoot=(3, 2, 2), nCell=(8, 8, 8),
                     periodic=(False, True, True), order="morton")


def _run(grid, cfg, U0, nstep, **kw):
    S = grid.soa_to_state_vgb(U0)
    o = po.Session(grid, cfg, S, **kw)
    o.exchange()
    o.init_timestep()
    for _ in range(nstep):
        o.s.nStep += 1
        o.set_global_timestep()
        for st in (1, 2):
            assert o.advance_stage(st) == 0
        o.exchange()
    return o


def _bc_cfg(grid, sides, state, **kw):
    cs = np.zeros((6, 8))
    cs[:] = state
    return po.config_from_dict(dict(flux="hlld", limiter="mc", beta=1.5,
                                    TypeCellBc=sides, CellState_VI=cs, **kw),
                               grid)


def test_constant_table_is_the_fixed_boundary():
    g = _grid()
    U0 = smooth_mhd_soa(g, seed=61, amp=0.2)
    P = np.array([1.0, -0.3, 0.1, 0.0, 0.2, 0.3, -0.1, 0.6])
    sides = ["float", "fixed", "periodic", "periodic", "periodic", "periodic"]
    a = _run(g, _bc_cfg(g, sides, _cons(P)), U0, 2)
    sides[1] = "vary"
    T = np.array([0.0, 1.0])
    b = _run(g, _bc_cfg(g, sides, np.zeros(8)), U0, 2,
             VaryTable=(T, np.tile(P, (2, 1))))
    np.testing.assert_array_equal(a.soa(), b.soa())


def test_ghosts_hold_the_table_state_at_the_time():
    g = _grid()
    U0 = smooth_mhd_soa(g, seed=62, amp=0.2)
    T, S = _table()
    sides = ["float", "vary", "periodic", "periodic", "periodic", "periodic"]
    cfg = _bc_cfg(g, sides, np.zeros(8))
    o = po.Session(g, cfg, g.soa_to_state_vgb(U0), VaryTable=(T, S))
    G = g.nG
    nei = g.iNeighbor_EB.reshape(-1, 6)
    for t in (0.004, 0.02, 0.07):
        o.s.tSimulation = t
        o.exchange()
        want = po.vary_state(T, S, t)
        for b in np.nonzero(nei[:, 1] == -1)[0]:
            gh = o.State[b][G:-G, G:-G, -G:]        # both ghost layers
            np.testing.assert_array_equal(
                gh.reshape(-1, 8), np.tile(want, (gh.size // 8, 1)))


@pytest.mark.parametrize("powell", [False, True])
def test_uniform_b0_with_fixed_sides_is_the_unsplit_run(powell):
    g = _grid()
    U0 = smooth_mhd_soa(g, seed=63, amp=0.2)
    b0 = np.array([0.35, -0.2, 0.15])
    nb = g.nBlock
    shp = (nb, g.nK + 2 * g.nG, g.nJ + 2 * g.nG, g.nI + 2 * g.nG, 3)
    B0 = np.broadcast_to(b0, shp).copy()
    sides = ["fixed", "inflow", "periodic", "periodic", "periodic",
             "periodic"]
    full = _cons([1.1, 0.2, -0.1, 0.05, 0.5, 0.1, 0.2, 0.8])
    # the split run holds B1 = B - B0 and gets the full field's CellState
    U1 = U0.copy()
    U1[4:7] = U0[4:7] - b0[:, None, None, None, None]
    kw = dict(UseDivbSource=powell)
    split = _run(g, _bc_cfg(g, sides, full, **kw), U1, 2, B0_DGB=B0,
                 UseB0Source=True)
    plain = _run(g, _bc_cfg(g, sides, full, **kw), U0, 2)
    W = split.soa()
    W[4:7] += b0[:, None, None, None, None]
    ref = plain.soa()
    scale = np.abs(ref).reshape(8, -1).max(axis=1)[:, None, None, None,
                                                   None]
    assert (np.abs(W - ref) / scale).max() < 1e-12
