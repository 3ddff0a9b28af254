"""The TVD treatment at resolution changes in the oracle (UseTvdResChange
with first-order prolongation, the reference's default:
src/ModFaceValue.f90:233-408 correct_monotone_restrict, :680-852
get_face_tvd, :2269-2343 tvd_reschange; ghost layers of
src/ModMessagePass.f90:105-109), pinned by what the construction fixes
(coarse cells twice the fine size, face at x = 0, coarse centres at -1 and
-3, fine centres at 0.5 and 1.5, subface centres at +-0.5 transverse):

* a linear field is reproduced exactly: the coarse-side and fine-side
  values at each subface and the fine value at the next face are the
  field's values there, with a transverse gradient no steeper than the
  normal one;
* the limited slope uses Beta = min(BetaLimiterResChange, LimiterBeta):
  with Beta = 0.8 the linear field's face values move by the known
  fractions of the gradient;
* a coarse cell at a local extremum gets no slope (TVD);
* correct_monotone_restrict moves a face-restricted ghost from the fine
  distance to the coarse centre exactly for a linear field, and floors
  density and pressure;
* in a two-level session the fine block's faces at the coarse side are
  tvd_reschange of its two coarse ghost layers and first two cells, the
  coarse block's face one cell in reads the corrected ghost, and the
  grid conserves mass, momentum, energy and field to rounding."""
import numpy as np
import pytest

import pyoracle as po
from batsrus_amd.amr import AmrGrid, sphere_refinement
from helpers import smooth_mhd_soa

RNG = np.random.default_rng(20261205)
SUB = np.array([[-0.5, -0.5], [0.5, -0.5], [-0.5, 0.5], [0.5, 0.5]])


def _linear(a, g, gt1=None, gt2=None):
    gt1 = np.zeros(8) if gt1 is None else gt1
    gt2 = np.zeros(8) if gt2 is None else gt2

    def u(x, t1=0.0, t2=0.0):
        return a + g * x + gt1 * t1 + gt2 * t2
    C2, C1 = u(-3.0), u(-1.0)
    F1 = np.array([u(0.5, *t) for t in SUB])
    F2 = np.array([u(1.5, *t) for t in SUB])
    return u, C2, C1, F1, F2


@pytest.mark.parametrize("beta", [1.0, 1.5, 2.0])
def test_linear_field_is_exact(beta):
    cfg = po.make_config(limiter="mc", beta=beta, BetaLimiterResChange=2.0)
    for _ in range(20):
        a, g = RNG.uniform(0.5, 2, 8), RNG.uniform(-1, 1, 8)
        # transverse gradients below the normal one
        gt1 = RNG.uniform(-0.4, 0.4, 8) * np.abs(g)
        gt2 = RNG.uniform(-0.4, 0.4, 8) * np.abs(g)
        u, C2, C1, F1, F2 = _linear(a, g, gt1, gt2)
        ctf, ftc, ff = po.tvd_reschange(cfg, C2, C1, F1, F2)
        for f, t in enumerate(SUB):
            np.testing.assert_allclose(ctf[f], u(0.0, *t), rtol=0,
                                       atol=1e-14)
            np.testing.assert_allclose(ftc[f], u(0.0, *t), rtol=0,
                                       atol=1e-14)
            np.testing.assert_allclose(ff[f], u(1.0, *t), rtol=0, atol=1e-14)


def test_beta_is_the_smaller_of_the_two():
    a, g = RNG.uniform(0.5, 2, 8), RNG.uniform(-1, 1, 8)
    _, C2, C1, F1, F2 = _linear(a, g)
    for lb, rb in ((1.5, 0.8), (0.8, 2.0)):
        cfg = po.make_config(limiter="mc", beta=lb, BetaLimiterResChange=rb)
        ctf, ftc, ff = po.tvd_reschange(cfg, C2, C1, F1, F2)
        # Beta = 0.8: coarse slope 0.8 g, fine half-slope 0.4 g
        np.testing.assert_allclose(ctf, np.tile(a - 0.2 * g, (4, 1)),
                                   atol=1e-14)
        np.testing.assert_allclose(ftc, np.tile(a + 0.1 * g, (4, 1)),
                                   atol=1e-14)
        np.testing.assert_allclose(ff, np.tile(a + 0.9 * g, (4, 1)),
                                   atol=1e-14)


def test_extremum_gets_no_slope():
    cfg = po.make_config(limiter="mc", beta=1.5)
    for _ in range(20):
        C1 = RNG.uniform(1, 2, 8)
        s = np.sign(RNG.uniform(-1, 1, 8))
        # C1 above (s > 0) or below both neighbours
        C2 = C1 - s * RNG.uniform(0.1, 0.5, 8)
        F1 = np.tile(C1 - s * RNG.uniform(0.1, 0.5, 8), (4, 1)) + \
            RNG.uniform(-0.01, 0.01, (4, 8))
        F2 = F1 + RNG.uniform(-0.1, 0.1, (4, 8))
        ctf, ftc, ff = po.tvd_reschange(cfg, C2, C1, F1, F2)
        np.testing.assert_array_equal(ctf, np.tile(C1, (4, 1)))
        # the fine values stay between the coarse cell and the fine cell
        lo, hi = np.minimum(F1, C1), np.maximum(F1, C1)
        assert np.all(ftc >= lo - 1e-15) and np.all(ftc <= hi + 1e-15)


def test_monotone_restrict_moves_the_ghost_to_the_coarse_centre():
    for _ in range(20):
        a, g = RNG.uniform(0.5, 2, 8), RNG.uniform(-0.2, 0.2, 8)
        # coarse block's cell at x = +1, face restriction at x = -0.5
        got = po.correct_monotone_restrict(a + g, a - 0.5 * g)
        np.testing.assert_allclose(got, a - g, rtol=0, atol=1e-15)
    # density and pressure are floored
    cell = np.array([1.0, 0, 0, 0, 0, 0, 0, 1.0])
    ghost = np.array([0.2, 0, 0, 0, 0, 0, 0, 0.1])
    got = po.correct_monotone_restrict(cell, ghost)
    assert got[0] == 1e-30 and got[7] == 1e-30


def _prim(S):
    P = S.copy()
    P[..., 1:4] = (1.0 / S[..., :1]) * S[..., 1:4]
    return P


def _session(tvd=True, seed=31):
    g = AmrGrid((4, 4, 4), sphere_refinement((4, 4, 4), 0.3))
    U0 = smooth_mhd_soa(g, seed=seed, amp=0.2)
    cfg = po.config_from_dict(dict(flux="hlld", limiter="mc", beta=1.5,
                                   UseTvdResChange=tvd), g)
    o = po.Session(g, cfg, g.soa_to_state_vgb(U0))
    o.exchange()
    return g, cfg, o


def _sides(g, kind):
    nei = g.iNeighbor_EB.reshape(-1, 6)
    kinds = g.iSlab_IS.reshape(-1, 8)[:, 2]
    return [(b, s) for b in range(g.nBlock) for s in range(6)
            if nei[b, s] <= -2 and kinds[-2 - nei[b, s]] == kind]


def _cell(S, d, nrm, t1, t2, G):
    """Ghosted state S (k, j, i, var) at normal index nrm (ghosted) and
    interior transverse (t1, t2): x (j, k), y (i, k), z (i, j)."""
    if d == 0:
        return S[t2 + G, t1 + G, nrm]
    if d == 1:
        return S[t2 + G, nrm, t1 + G]
    return S[nrm, t2 + G, t1 + G]


def test_session_fine_faces_are_tvd_of_the_ghost_layers():
    g, cfg, o = _session()
    G, n = g.nG, g.nI
    checked = set()
    for b, side in _sides(g, 0):
        d, hi = side // 2, side % 2
        (LX, RX, LY, RY, LZ, RZ), _ = o.face_values(b)
        L, R = (LX, RX) if d == 0 else (LY, RY) if d == 1 else (LZ, RZ)
        S = o.State[b]
        # ghosted normal indices of C1, C2, F1, F2
        idx = (n + G, n + G + 1, n + G - 1, n + G - 2) if hi else \
            (G - 1, G - 2, G, G + 1)
        for t2 in range(0, n, 2):
            for t1 in range(0, n, 2):
                C1 = _prim(_cell(S, d, idx[0], t1, t2, G))
                C2 = _prim(_cell(S, d, idx[1], t1, t2, G))
                F1 = np.array([_prim(_cell(S, d, idx[2], t1 + f % 2,
                                           t2 + f // 2, G)) for f in range(4)])
                F2 = np.array([_prim(_cell(S, d, idx[3], t1 + f % 2,
                                           t2 + f // 2, G)) for f in range(4)])
                ctf, ftc, ff = po.tvd_reschange(cfg, C2, C1, F1, F2)
                for f in range(4):
                    u1, u2 = t1 + f % 2, t2 + f // 2
                    a0 = _face(L, R, d, n if hi else 0, u1, u2)
                    a1 = _face(L, R, d, n - 1 if hi else 1, u1, u2)
                    # coarse side, fine side, fine value one face in
                    np.testing.assert_array_equal(a0[hi], ctf[f])
                    np.testing.assert_array_equal(a0[1 - hi], ftc[f])
                    np.testing.assert_array_equal(a1[hi], ff[f])
        checked.add(side)
    assert len(checked) == 6


def _face(L, R, d, f, t1, t2):
    """(left, right) states of face f normal to d at transverse (t1, t2):
    x faces (k, j, f), y faces (k, f, i), z faces (f, j, i)."""
    if d == 0:
        return L[t2, t1, f], R[t2, t1, f]
    if d == 1:
        return L[t2, f, t1], R[t2, f, t1]
    return L[f, t2, t1], R[f, t2, t1]


def test_coarse_face_reads_the_corrected_ghost():
    """The coarse block's face one cell in from a finer side: its left
    (right) state is cell 1's value plus the MC slope taken with the
    monotone-corrected ghost."""
    g, cfg, o = _session()
    G, n = g.nG, g.nI
    b, side = next((b, s) for b, s in _sides(g, 1) if s == 0)
    (LX, RX, *_), _ = o.face_values(b)
    S = o.State[b]
    for k in range(0, n, 3):
        for j in range(0, n, 3):
            ghost, c1, c2 = (_prim(S[k + G, j + G, G + t]) for t in (-1, 0, 1))
            gc = po.correct_monotone_restrict(c1, ghost)
            d1, d2 = c1 - gc, c2 - c1
            sg = np.copysign(0.25, d1) + np.copysign(0.25, d2)
            slope = sg * np.minimum(np.minimum(1.5 * np.abs(d1),
                                               1.5 * np.abs(d2)),
                                    0.5 * np.abs(d1 + d2))
            np.testing.assert_allclose(LX[k, j, 1], c1 + slope, rtol=0,
                                       atol=1e-15)


def _totals(o, g):
    W = o.soa()
    vol = np.prod(g.CellSize_DB, axis=1)[None, :, None, None, None]
    cons = W.copy()
    u2 = (W[1] ** 2 + W[2] ** 2 + W[3] ** 2) / W[0]
    b2 = W[4] ** 2 + W[5] ** 2 + W[6] ** 2
    cons[7] = W[7] / (5.0 / 3.0 - 1) + 0.5 * u2 + 0.5 * b2
    return (cons * vol).sum(axis=(1, 2, 3, 4))


def test_two_level_tvd_conserves_and_differs_from_default():
    g, cfg, o = _session()
    t0 = _totals(o, g)
    o.init_timestep()
    for _ in range(3):
        o.s.nStep += 1
        o.set_global_timestep()
        for st in (1, 2):
            assert o.advance_stage(st) == 0
        o.exchange()
    t1 = _totals(o, g)
    scale = np.abs(t0) + 1.0
    assert np.all(np.abs(t1 - t0) <= 1e-13 * scale), (t1 - t0) / scale
    g2, cfg2, ref = _session(tvd=False)
    ref.init_timestep()
    for _ in range(3):
        ref.s.nStep += 1
        ref.set_global_timestep()
        for st in (1, 2):
            assert ref.advance_stage(st) == 0
        ref.exchange()
    assert np.abs(o.soa() - ref.soa()).max() > 1e-6


def test_fine_ghost_layers_are_the_two_coarse_cells():
    """With the second-order scheme the fine block's ghost layers 0 and -1
    hold the coarse cells next to the face and behind it (first-order
    prolongation with two coarse layers), shared by each 2x2 group."""
    g, cfg, o = _session()
    G, n = g.nG, g.nI
    slabs = g.iSlab_IS.reshape(-1, 8)
    nei = g.iNeighbor_EB.reshape(-1, 6)
    for b, side in _sides(g, 0):
        d, hi = side // 2, side % 2
        C, p0, q0 = slabs[-2 - nei[b, side], 3:6]
        for L in (0, 1):
            gn = n + G + L if hi else G - 1 - L
            cn = (L if hi else n - 1 - L) + G
            for t2 in range(n):
                for t1 in range(n):
                    got = _cell(o.State[b], d, gn, t1, t2, G)
                    want = _cell(o.State[C], d, cn, p0 + t1 // 2,
                                 q0 + t2 // 2, G)
                    np.testing.assert_array_equal(got, want)


def test_eight_wave_source_reads_the_tvd_faces():
    """With the TVD faces the 8-wave source sees the divergence of a field
    linear along its normal, Bx = 0.3 + 0.4 x (div B = 0.4), exactly in the
    cells next to the resolution changes as everywhere else: the TVD face
    values of such a field are exact (no jump of the normal field at any
    face, the internal divergence the field's own), so Source(Bx) =
    -div B u_x = -0.04 there.  Faces taken without the TVD treatment on the
    TVD ghost layers (which hold the coarse cells unchanged) miss it next
    to the x resolution changes.  Cells near the periodic seam, where the
    field jumps, are left out."""
    g = AmrGrid((4, 4, 4), sphere_refinement((4, 4, 4), 0.3))
    x, y, z = g.cell_centers()
    U = np.zeros((8,) + x.shape)
    U[0] = 1.0
    U[1] = 0.1
    U[7] = 1.0
    U[4] = 0.3 + 0.4 * x
    U[5] = -0.2
    U[6] = 0.1
    cfg = po.config_from_dict(dict(flux="hlld", limiter="mc", beta=1.5,
                                   UseDivbSource=True, UseTvdResChange=True),
                              g)
    o = po.Session(g, cfg, g.soa_to_state_vgb(U))
    o.exchange()
    src = np.moveaxis(o.sources(), -1, 0)     # (9, nBlock, nK, nJ, nI)
    far = (x > 3 / 32) & (x < 1 - 3 / 32)
    # the cells of blocks at an x resolution change
    nei = g.iNeighbor_EB.reshape(-1, 6)
    rc = (nei[:, :2] <= -2).any(axis=1)
    sel = far & rc[:, None, None, None]
    assert sel.sum() > 10000
    np.testing.assert_allclose(src[4][sel], -0.04, rtol=0, atol=1e-12)
    np.testing.assert_allclose(src[1][sel], -0.4 * U[4][sel], rtol=0,
                               atol=1e-12)
