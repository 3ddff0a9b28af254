"""The inner body in the oracle (true cells Used_GB and the 'ionosphere' face
boundary condition, as the fast update applies them:
src/ModUpdateStateFast.f90:230-265,343-345,924-937,1277-1327;
src/ModFaceBoundary.f90:621-648,876-884; the time step mask of
src/ModTimeStepControl.f90:158-159,255), pinned by the definitions:

* the ghost face state: density moves toward the body density by at most
  10 % of the true value (and reaches it when closer), the pressure floats,
  the velocity reverses, the radial field becomes B1rCoef times the true
  one (-1: reflected, 0: removed, 1: unchanged) with the tangential field
  kept;
* body cells never change and hold DtMax_CB = 0 outside the block minimum;
* cells whose stencils never reach the body evolve exactly as without it;
* a body at the centre of a mirror-symmetric state keeps the symmetry;
* a face whose limiter stencil meets the body takes first-order values."""
import numpy as np
import pytest

import pyoracle as po
from batsrus_amd.grid import BlockGrid
from helpers import body_mask

RNG = np.random.default_rng(20261204)


def _true_state():
    return np.array([RNG.uniform(0.5, 1.5), *RNG.uniform(-1, 1, 3),
                     *RNG.uniform(-1, 1, 3), RNG.uniform(0.3, 1.5)])


@pytest.mark.parametrize("brho", [0.2, 0.97, 5.0])
def test_ghost_face_state(brho):
    T = _true_state()
    T[0] = 1.0
    r = RNG.uniform(-1, 1, 3)
    g = po.body_face_state(T, r, brho, 7.0, -1.0)
    step = min(abs(brho - 1.0), 0.1)
    assert g[0] == pytest.approx(1.0 + np.sign(brho - 1.0) * step, abs=1e-15)
    assert g[7] == T[7]                         # pressure floats
    np.testing.assert_array_equal(g[1:4], -T[1:4])
    rh = r / np.linalg.norm(r)
    # radial component reflected, tangential kept
    assert g[4:7] @ rh == pytest.approx(-(T[4:7] @ rh), abs=1e-14)
    t_true = T[4:7] - (T[4:7] @ rh) * rh
    t_ghost = g[4:7] - (g[4:7] @ rh) * rh
    np.testing.assert_allclose(t_ghost, t_true, atol=1e-14)
    g0 = po.body_face_state(T, r, brho, 7.0, 0.0)
    assert abs(g0[4:7] @ rh) < 1e-14
    g1 = po.body_face_state(T, r, brho, 7.0, 1.0)
    np.testing.assert_array_equal(g1[4:7], T[4:7])


def _body_setup(nRoot=(2, 2, 2), nCell=8, rBody=0.18, center=(0.5, 0.5, 0.5)):
    """A grid on [0,1]^3 with float sides, the true cells outside a sphere
    at `center` (ghost cells included) and each block's first cell centre
    relative to the sphere's centre."""
    grid = BlockGrid(nRoot=nRoot, nCell=(nCell,) * 3,
                     periodic=(False, False, False), order="morton")
    used, xyz0 = body_mask(grid, rBody, center)
    return grid, used, xyz0


def _run(grid, cfg, U, nstep, **body):
    S = grid.soa_to_state_vgb(U)
    grid.fill_ghosts(S)
    o = po.Session(grid, cfg, S, **body)
    o.init_timestep()
    for _ in range(nstep):
        o.s.nStep += 1
        o.set_global_timestep()
        for st in (1, 2):
            assert o.advance_stage(st) == 0
            o.exchange()
    return o


def _smooth(grid, amp=0.2, sym=False):
    x, y, z = grid.cell_centers()
    tp = 2 * np.pi
    U = np.zeros((8,) + x.shape)
    xs = (x - 0.5) if not sym else np.abs(x - 0.5)
    U[0] = 1.0 + amp * np.cos(tp * xs) * np.cos(tp * (y - 0.5))
    U[1] = U[0] * amp * np.sin(tp * (y - 0.3))
    U[2] = U[0] * amp * np.cos(tp * xs)
    U[4] = 0.3 + 0.1 * np.cos(tp * (z - 0.5))
    U[5] = 0.1 * np.cos(tp * xs)
    U[6] = 0.2
    U[7] = 1.0 + amp * np.cos(tp * (y - 0.5)) * np.cos(tp * xs)
    return U


def test_body_cells_stay_and_far_cells_do_not_notice():
    grid, used, xyz0 = _body_setup()
    cfg = po.config_from_dict(dict(flux="hlld", limiter="mc", beta=1.5), grid)
    U = _smooth(grid)
    body = dict(Used_GB=used, XyzStart_DB=xyz0, BodyRho=3.0, BodyP=2.0,
                B1rCoef=-1.0)
    o = _run(grid, cfg, U, 1, **body)
    ref = _run(grid, cfg, U, 1)
    g = grid.nG
    inner = used[:, g:-g, g:-g, g:-g]
    Ub, Ur = o.soa(), ref.soa()
    np.testing.assert_array_equal(Ub[:, ~inner], U[:, ~inner])
    assert np.all(o.DtMax_CB.reshape(inner.shape)[~inner] == 0)
    # cells more than 2 stages x 2 cells away from every body cell
    x, y, z = grid.cell_centers()
    d = np.sqrt((x - 0.5) ** 2 + (y - 0.5) ** 2 + (z - 0.5) ** 2)
    far = d > 0.18 + 5 * grid.CellSize_DB[0][0]
    np.testing.assert_array_equal(Ub[:, far], Ur[:, far])
    assert np.abs(Ub - Ur)[:, ~far & inner].max() > 1e-6


def test_symmetric_state_keeps_its_symmetry():
    grid, used, xyz0 = _body_setup()
    cfg = po.config_from_dict(dict(flux="rusanov", limiter="mc3", beta=1.5),
                              grid)
    U = _smooth(grid, sym=True)
    # mirror x -> 1 - x: rho, p, u_y, u_z, B_y, B_z even; u_x, B_x odd
    U[1] = U[0] * 0.1 * np.sin(2 * np.pi * (grid.cell_centers()[0] - 0.5))
    U[4] = 0.0
    o = _run(grid, cfg, U, 2, Used_GB=used, XyzStart_DB=xyz0, BodyRho=3.0,
             BodyP=2.0, B1rCoef=-1.0)
    W = o.soa()
    x = grid.cell_centers()[0]
    # gather the state on the global grid to compare x with 1 - x
    n = grid.nRoot[0] * grid.nI
    ix = np.floor(x / grid.CellSize_DB[0][0]).astype(int)
    gi = grid.global_cell_index().astype(np.int64)
    glob = np.zeros((8, n ** 3))
    glob[:, gi.ravel()] = W.reshape(8, -1)
    glob = glob.reshape(8, n, n, n)              # (var, k, j, i)
    sgn = np.array([1, -1, 1, 1, -1, 1, 1, 1])[:, None, None, None]
    np.testing.assert_allclose(glob, sgn * glob[:, :, :, ::-1], rtol=0,
                               atol=1e-12)


def test_face_states_next_to_the_body():
    """Faces of a pencil through the body: a face whose limiter stencil
    meets a body cell takes the first-order value on that side, a face
    between a true and a body cell takes the ghost state of the true side's
    face value, and faces between two body cells are flagged."""
    grid, used, xyz0 = _body_setup(nRoot=(1, 1, 1), nCell=8, rBody=0.2)
    cfg = po.config_from_dict(dict(flux="rusanov", limiter="mc", beta=1.5),
                              grid)
    U = _smooth(grid)
    S = grid.soa_to_state_vgb(U)
    grid.fill_ghosts(S)
    body = dict(Used_GB=used, XyzStart_DB=xyz0, BodyRho=3.0, BodyP=2.0,
                B1rCoef=-1.0)
    o = po.Session(grid, cfg, S, **body)
    (LX, RX, *_), (SX, *_) = o.face_values(0)
    plain = po.calc_face_value(cfg, S[0])
    g = grid.nG
    n = grid.nI
    dx = grid.CellSize_DB[0]
    checked = set()
    for k in range(n):
        for j in range(n):
            u = used[0, k + g, j + g, :]          # ghosted x pencil
            for f in range(n + 1):                # face between f-1 and f
                c = [u[f + g + t - 2] for t in range(4)]
                P = [S[0, k + g, j + g, f + g - 1].copy(),
                     S[0, k + g, j + g, f + g].copy()]
                for q in P:
                    q[1:4] = (1.0 / q[0]) * q[1:4]
                L, R = LX[k, j, f], RX[k, j, f]
                if not c[1] and not c[2]:
                    assert SX[k, j, f] == 1
                    checked.add("skip")
                    continue
                assert SX[k, j, f] == 0
                exp_L = plain[0][k, j, f] if all(c[:3]) else P[0]
                exp_R = plain[1][k, j, f] if all(c[1:]) else P[1]
                if c[1] and not c[2]:
                    xf = xyz0[0] + np.array([f - 0.5, j, k]) * dx
                    exp_R = po.body_face_state(exp_L, xf, 3.0, 2.0, -1.0)
                    checked.add("ghostR")
                elif c[2] and not c[1]:
                    xf = xyz0[0] + np.array([f - 0.5, j, k]) * dx
                    exp_L = po.body_face_state(exp_R, xf, 3.0, 2.0, -1.0)
                    checked.add("ghostL")
                elif not all(c):
                    checked.add("first")
                np.testing.assert_allclose(L, exp_L, rtol=0, atol=1e-15)
                np.testing.assert_allclose(R, exp_R, rtol=0, atol=1e-15)
    assert checked == {"skip", "ghostR", "ghostL", "first"}


def test_body_in_ghost_cells_only():
    """A block whose own cells are all true but whose ghost cells reach a
    body cell of its neighbour is a body block (IsBody_B covers the ghost
    cells): its face next to that cell takes the ghost face state."""
    dx = 1.0 / 16
    c = (0.5 - 0.5 * dx, 0.25 + 0.5 * dx, 0.25 + 0.5 * dx)
    grid, used, xyz0 = _body_setup(nCell=8, rBody=0.3 * dx, center=c)
    g = grid.nG
    inner_true = used[:, g:-g, g:-g, g:-g].all(axis=(1, 2, 3))
    ghost_body = ~used.all(axis=(1, 2, 3))
    blocks = np.nonzero(inner_true & ghost_body)[0]
    assert len(blocks) >= 1
    cfg = po.config_from_dict(dict(flux="rusanov", limiter="mc", beta=1.5),
                              grid)
    S = grid.soa_to_state_vgb(_smooth(grid))
    grid.fill_ghosts(S)
    o = po.Session(grid, cfg, S, Used_GB=used, XyzStart_DB=xyz0, BodyRho=3.0,
                   BodyP=2.0, B1rCoef=-1.0)
    hits = 0
    for b in blocks:
        (LX, RX, *_), _ = o.face_values(int(b))
        plain = po.calc_face_value(cfg, S[b])
        u = used[b]
        for k in range(grid.nI):
            for j in range(grid.nI):
                if u[k + g, j + g, g - 1]:
                    continue
                # face 0 of this row: a body ghost cell on its left
                R = plain[1][k, j, 0] if u[k + g, j + g, g:g + 2].all() \
                    else None
                xf = xyz0[b] + np.array([-0.5, j, k]) * dx
                exp = po.body_face_state(RX[k, j, 0], xf, 3.0, 2.0, -1.0)
                np.testing.assert_allclose(LX[k, j, 0], exp, rtol=0,
                                           atol=1e-15)
                if R is not None:
                    np.testing.assert_array_equal(RX[k, j, 0], R)
                assert np.abs(LX[k, j, 0] - plain[0][k, j, 0]).max() > 1e-3
                hits += 1
    assert hits >= 1
