"""Pins the oracle's fixed and inflow outer boundaries: the ghost cells of
such a side hold the side's state CellState_VI (set_boundary_for_cell,
src/ModUpdateStateFast.f90:1262-1268), a uniform flow equal to that state
stays uniform, the first order boundary update is the flux difference with
the numerical flux of (boundary state, interior cell), and a fixed state
equal to the float copy of the boundary cells gives the float run."""
import numpy as np

import pyoracle as po
from batsrus_amd.grid import BlockGrid
from helpers import random_mhd_soa

GAMMA = 5.0 / 3.0
NC = 8


def _cons(rho, u, b, p):
    return np.array([rho, rho * u[0], rho * u[1], rho * u[2], *b, p])


def _prim(U):
    return np.array([U[0], U[1] / U[0], U[2] / U[0], U[3] / U[0], *U[4:8]])


STATES = np.array([_cons(1.0 + 0.1 * s, (0.3 - 0.1 * s, 0.05 * s, -0.2),
                         (0.4, 0.2 * s - 0.3, 0.1), 0.8 + 0.05 * s)
                   for s in range(6)])


def test_ghost_cells_hold_the_side_state():
    g = BlockGrid(nRoot=(2, 2, 1), nCell=(NC,) * 3, periodic=(False,) * 3)
    types = ["fixed", "float", "float", "inflow", "float", "fixed"]
    cfg = po.make_config(nI=NC, nJ=NC, nK=NC, TypeCellBc=types,
                         CellState_VI=STATES)
    U0 = random_mhd_soa(g, seed=1)
    o = po.Session(g, cfg, g.soa_to_state_vgb(U0))
    o.exchange()
    S = o.State                       # (b, k, j, i, v), 2 ghost layers
    inner = slice(2, 2 + NC)
    nei = g.iNeighbor_EB
    for b in range(g.nBlock):
        if nei[b, 0] == -1:
            assert np.array_equal(S[b, inner, inner, 0:2],
                                  np.broadcast_to(STATES[0], (NC, NC, 2, 8)))
        if nei[b, 1] == -1:           # float: copies of the last cell
            for layer in (NC + 2, NC + 3):
                assert np.array_equal(S[b, inner, inner, layer],
                                      S[b, inner, inner, NC + 1])
        if nei[b, 3] == -1:
            assert np.array_equal(S[b, inner, NC + 2:NC + 4, inner],
                                  np.broadcast_to(STATES[3], (NC, 2, NC, 8)))
        if nei[b, 5] == -1:
            assert np.array_equal(S[b, NC + 2:NC + 4, inner, inner],
                                  np.broadcast_to(STATES[5], (2, NC, NC, 8)))


def test_uniform_flow_equal_to_the_boundary_state_is_kept():
    g = BlockGrid(nRoot=(2, 2, 2), nCell=(NC,) * 3, periodic=(False,) * 3)
    U = STATES[2]
    cfg = po.make_config(nI=NC, nJ=NC, nK=NC, flux="hlld", limiter="mc3",
                         TypeCellBc=["fixed"] * 6,
                         CellState_VI=np.tile(U, (6, 1)))
    U0 = np.broadcast_to(U[:, None, None, None, None],
                         (8, g.nBlock, NC, NC, NC)).copy()
    o = po.Session(g, cfg, g.soa_to_state_vgb(U0))
    o.init_timestep()
    for _ in range(3):
        assert o.advance_step() == 0
    assert np.abs(o.soa() - U0).max() <= 1e-14 * np.abs(U0).max()


def test_first_order_boundary_update_closed_form():
    """Uniform interior state Ui, fixed x- side Ub: after one first order
    stage the x- boundary cells hold Ui - Dt/dx (F(Ui,Ui) - F*(Ub,Ui)) (in
    the energy for the pressure) and every other cell keeps Ui."""
    g = BlockGrid(nRoot=(2, 1, 1), nCell=(NC,) * 3,
                  periodic=(False, True, True))
    Ui, Ub = STATES[1], STATES[4]
    for flux in ("rusanov", "linde", "hlld"):
        cfg = po.make_config(nI=NC, nJ=NC, nK=NC, nOrder=1, nStage=1,
                             flux=flux, UseDtFixed=True, DtFixed=0.02,
                             TypeCellBc=["fixed", "float"] + ["periodic"] * 4,
                             CellState_VI=np.tile(Ub, (6, 1)))
        U0 = np.broadcast_to(Ui[:, None, None, None, None],
                             (8, g.nBlock, NC, NC, NC)).copy()
        o = po.Session(g, cfg, g.soa_to_state_vgb(U0))
        o.init_timestep()
        dt = o.new_step()                  # DtFixed (Cfl forced to 1)
        assert dt == 0.02
        assert o.advance_stage(1) == 0
        U1 = o.soa()
        dx = g.CellSize_DB[0, 0]
        _, Fin = po.numerical_flux(cfg, 0, 1.0, _prim(Ui), _prim(Ui))
        _, Fb = po.numerical_flux(cfg, 0, 1.0, _prim(Ub), _prim(Ui))
        cons = (Fin[:9] - Fb[:9]) * (dt / dx)
        want = Ui.copy()
        want[:7] -= cons[:7]
        E = Ui[7] / (GAMMA - 1) + 0.5 * (Ui[1:4] @ Ui[1:4]) / Ui[0] + \
            0.5 * (Ui[4:7] @ Ui[4:7])
        E1 = E - cons[8]
        want[7] = (GAMMA - 1) * (E1 - 0.5 * (want[1:4] @ want[1:4]) /
                                 want[0] - 0.5 * (want[4:7] @ want[4:7]))
        left = g.iNeighbor_EB[:, 0] == -1
        got = U1[:, left][..., 0]          # (8, blocks, k, j) at i = 0
        err = np.abs(got - want[:, None, None, None]).max()
        assert err <= 1e-13, (flux, err)
        # the boundary state differs, so the update is not trivial
        assert np.abs(want - Ui).max() > 1e-3
        rest = U1.copy()
        for b in np.flatnonzero(left):
            rest[:, b, :, :, 0] = Ui[:, None, None]
        assert np.abs(rest - U0).max() <= 1e-14, flux


def test_fixed_equal_to_float_copy_gives_the_float_stage():
    """A state that varies along x only: the float ghosts of the x sides are
    copies of the (constant over the face) boundary cells, so fixed sides
    with exactly those states give the same first stage, bit for bit, in
    the second order scheme (both ghost layers are read)."""
    g = BlockGrid(nRoot=(3, 1, 1), nCell=(NC,) * 3,
                  periodic=(False, True, True))
    x = (np.arange(3 * NC) + 0.5) / (3 * NC)
    prof = np.stack([1 + 0.3 * np.sin(2 * np.pi * x), 0.2 + 0.1 * x,
                     0.1 * np.cos(3 * x), -0.05 + 0 * x, 0.5 + 0 * x,
                     0.3 * np.sin(5 * x), 0.2 + 0.1 * x * x,
                     1 + 0.2 * np.cos(4 * x)])          # (8, 3*NC)
    U0 = np.zeros((8, 3, NC, NC, NC))
    for b in range(3):
        U0[:, b] = prof[:, b * NC:(b + 1) * NC][:, None, None, :]
    st = np.zeros((6, 8))
    st[0], st[1] = prof[:, 0], prof[:, -1]
    scheme = dict(nI=NC, nJ=NC, nK=NC, flux="hlld", limiter="mc3", beta=1.5)
    res = []
    for types in (["float", "float"], ["fixed", "fixed"]):
        cfg = po.make_config(**scheme, TypeCellBc=types + ["periodic"] * 4,
                             CellState_VI=st)
        o = po.Session(g, cfg, g.soa_to_state_vgb(U0))
        o.init_timestep()
        o.new_step()
        assert o.advance_stage(1) == 0
        res.append(o.soa())
    assert np.array_equal(res[0], res[1])
    # a wrong state on one side is seen
    st[1, 0] *= 1.01
    cfg = po.make_config(**scheme, TypeCellBc=["fixed"] * 2 + ["periodic"] * 4,
                         CellState_VI=st)
    o = po.Session(g, cfg, g.soa_to_state_vgb(U0))
    o.init_timestep()
    o.new_step()
    assert o.advance_stage(1) == 0
    assert not np.array_equal(res[0], o.soa())


def _mirror_x(U):
    """Mirror image of an SoA state about the high x side of the domain:
    cells reversed along x, x momentum and B_x negated (blocks reversed
    too, for a 1-block-high row of blocks)."""
    M = U[:, ::-1, :, :, ::-1].copy()
    M[1] *= -1
    M[4] *= -1
    return M


def test_reflecting_walls_equal_a_mirrored_periodic_domain():
    """[0,1] with reflecting x sides against [0,2] periodic whose second
    half is the mirror image of the first: the halves agree bit for bit
    after a stage (the ghosts are exact mirrors) and to rounding after
    several steps; the ghost layers hold the mirrored cells."""
    nb = 2
    half = BlockGrid(nRoot=(nb, 1, 1), nCell=(NC,) * 3,
                     xyzMax=(1.0, 1.0 / nb, 1.0 / nb),
                     periodic=(False, True, True))
    full = BlockGrid(nRoot=(2 * nb, 1, 1), nCell=(NC,) * 3,
                     xyzMax=(2.0, 1.0 / nb, 1.0 / nb),
                     periodic=(True, True, True))
    U = random_mhd_soa(half, seed=31)
    Uf = np.concatenate([U, _mirror_x(U)], axis=1)
    scheme = dict(nI=NC, nJ=NC, nK=NC, flux="hlld", limiter="mc3", beta=1.5)
    cfg_h = po.make_config(**scheme, TypeCellBc=["reflect", "reflect"] +
                           ["periodic"] * 4)
    cfg_f = po.make_config(**scheme)
    oh = po.Session(half, cfg_h, half.soa_to_state_vgb(U))
    of = po.Session(full, cfg_f, full.soa_to_state_vgb(Uf))
    oh.exchange()
    S = oh.State
    for b, side, ghost, src in ((0, 0, (1, 0), (2, 3)),
                                (nb - 1, 1, (NC + 2, NC + 3), (NC + 1, NC))):
        for gl, sl in zip(ghost, src):
            want = S[b, 2:2 + NC, 2:2 + NC, sl].copy()
            want[..., [1, 4]] *= -1
            assert np.array_equal(S[b, 2:2 + NC, 2:2 + NC, gl], want)
    for o in (oh, of):
        o.init_timestep()
        o.new_step()
    # the same time step: the mirrored half has the same speeds
    assert oh.Dt == of.Dt
    assert oh.advance_stage(1) == 0 and of.advance_stage(1) == 0
    assert np.array_equal(oh.soa(), of.soa()[:, :nb])
    assert oh.advance_stage(2) == 0 and of.advance_stage(2) == 0
    oh.exchange()
    of.exchange()
    for _ in range(3):
        assert oh.advance_step() == 0 and of.advance_step() == 0
    a, b = oh.soa(), of.soa()[:, :nb]
    assert (np.abs(a - b).max(axis=(1, 2, 3, 4)) /
            np.abs(b).max(axis=(1, 2, 3, 4))).max() < 1e-13


def test_outflow_sides_float_with_a_fixed_pressure():
    """OutFlowBC_ is float; with UseOutflowPressure the ghost pressure is
    pOutflow in both layers and every other variable still floats
    (src/ModCellBoundary.f90:273-277)."""
    g = BlockGrid(nRoot=(2, 1, 1), nCell=(NC,) * 3,
                  periodic=(False, True, True))
    U0 = random_mhd_soa(g, seed=32)
    types = ["outflow", "outflow"] + ["periodic"] * 4
    res = {}
    for usep in (False, True):
        cfg = po.make_config(nI=NC, nJ=NC, nK=NC, TypeCellBc=types,
                             UseOutflowPressure=usep, pOutflow=0.37)
        o = po.Session(g, cfg, g.soa_to_state_vgb(U0))
        o.exchange()
        res[usep] = o.State.copy()
    cfgf = po.make_config(nI=NC, nJ=NC, nK=NC)      # all float
    of = po.Session(g, cfgf, g.soa_to_state_vgb(U0))
    of.exchange()
    assert np.array_equal(res[False], of.State)
    inner = slice(2, 2 + NC)
    for b, layers in ((0, (0, 1)), (1, (NC + 2, NC + 3))):
        for gl in layers:
            got = res[True][b, inner, inner, gl]
            want = of.State[b, inner, inner, gl].copy()
            want[..., 7] = 0.37
            assert np.array_equal(got, want)
