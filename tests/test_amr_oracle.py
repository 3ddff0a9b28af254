"""Two-level grids: the slab table, the oracle's resolution-change ghost
cells (closed forms on linear data) and flux correction (conservation)."""
import numpy as np
import pytest

import pyoracle as po
from batsrus_amd.amr import AmrGrid, sphere_refinement
from batsrus_amd.grid import BlockGrid
from batsrus_amd import ics
from helpers import smooth_mhd_soa


def _grid(n=4, radius=0.3):
    return AmrGrid((n, n, n), sphere_refinement((n, n, n), radius))


def test_slab_table_consistent():
    g = _grid()
    sl = g.iSlab_IS
    assert g.nBlock == 64 - 8 + 64 and g.nSlab == 96 + 24
    for t, r in enumerate(sl):
        b, side, kind = r[:3]
        assert g.iNeighbor_EB[b, side] == -2 - t
        if kind == 0:
            C, tgt = r[3], r[7]
            assert g.level[b] == 1 and g.level[C] == 0
            assert sl[tgt][0] == C and sl[tgt][1] == side ^ 1
            assert sl[tgt][2] == 1
            quad = (r[4] // 4) + 2 * (r[5] // 4)
            assert sl[tgt][3 + quad] == b
        else:
            assert g.level[b] == 0 and all(g.level[r[3:7]] == 1)
    # same-level links are symmetric
    for b in range(g.nBlock):
        for s in range(6):
            n = g.iNeighbor_EB[b, s]
            if n >= 0:
                assert g.iNeighbor_EB[n, s ^ 1] == b


def _linear(x, y, z):
    a = np.array([[1.0, 0.3, -0.2, 0.5], [0.1, 0.2, 0.4, -0.3],
                  [-0.2, 0.1, 0.3, 0.2], [0.3, -0.4, 0.1, 0.1],
                  [0.5, 0.2, -0.1, 0.3], [-0.1, 0.3, 0.2, -0.2],
                  [0.2, -0.2, 0.1, 0.4], [2.0, 0.1, 0.2, -0.3]])
    return np.stack([a[v, 0] + a[v, 1] * x + a[v, 2] * y + a[v, 3] * z
                     for v in range(8)])


def test_resolution_change_ghosts_exact_on_linear_data():
    """Restriction (mean of 8 fine cells) and the limited linear
    prolongation both reproduce a linear field exactly in every ghost cell
    of every resolution-change slab."""
    g = _grid()
    x, y, z = g.cell_centers()
    U = _linear(x, y, z)
    cfg = po.make_config(nI=8, nJ=8, nK=8)
    o = po.Session(g, cfg, g.soa_to_state_vgb(U))
    o.exchange()
    S = o.State
    n = 8
    dxf = 1.0 / (4 * 2 * n)
    for r in g.iSlab_IS:
        b, side = r[0], r[1]
        d, hi = side // 2, side % 2
        dx = g.CellSize_DB[b]
        org = g.pos[b] * n * dxf
        for L in range(2):
            nrm = n + L if hi else -1 - L
            for q in range(n):
                for p in range(n):
                    idx = [p, q]
                    idx.insert(d, nrm)
                    c = [org[a] + (idx[a] + 0.5) * dx[a] for a in range(3)]
                    want = _linear(*c)
                    got = S[b, idx[2] + 2, idx[1] + 2, idx[0] + 2]
                    assert np.allclose(got, want, rtol=0, atol=1e-13), (r, L, p, q)


def test_unrefined_grid_equals_uniform_grid():
    """With nothing refined the two-level grid is the uniform grid."""
    g = AmrGrid((2, 2, 2), np.zeros((2, 2, 2), bool))
    u = BlockGrid(nRoot=(2, 2, 2), nCell=(8, 8, 8), order="morton")
    assert np.array_equal(g.iNeighbor_EB, u.iNeighbor_EB)
    assert g.nSlab == 0


@pytest.mark.parametrize("flux", ["rusanov", "hlld"])
def test_two_level_conservation(flux):
    """The coarse faces at resolution changes take the sum of the fine face
    fluxes, so the totals of a periodic box are conserved to rounding."""
    g = _grid()
    U0 = smooth_mhd_soa(g, seed=8, amp=0.2) if hasattr(g, "cell_centers") else None
    cfg = po.config_from_dict(dict(flux=flux, limiter="mc", beta=1.5), g)
    o = po.Session(g, cfg, g.soa_to_state_vgb(U0))
    o.init_timestep()
    V = np.prod(g.CellSize_DB, axis=1)[None, :, None, None, None]

    def totals(U):
        E = U[7] / (5 / 3 - 1) + 0.5 * (U[1] ** 2 + U[2] ** 2 + U[3] ** 2) / \
            U[0] + 0.5 * (U[4] ** 2 + U[5] ** 2 + U[6] ** 2)
        W = U.copy()
        W[7] = E
        return (W * V).reshape(8, -1).sum(1)
    t0 = totals(U0)
    for _ in range(4):
        assert o.advance_step() == 0
    t1 = totals(o.soa())
    scale = np.abs(U0 * V).reshape(8, -1).sum(1)
    scale[7] = t0[7]
    assert (np.abs(t1 - t0) / scale).max() < 1e-13


def test_two_level_uniform_state():
    g = _grid()
    U0 = ics.uniform_state(g, rho=1.1, u=(0.2, -0.1, 0.3), b=(0.3, 0.2, -0.4),
                           p=0.8)
    cfg = po.config_from_dict(dict(flux="hlld", limiter="mc3", beta=1.5), g)
    o = po.Session(g, cfg, g.soa_to_state_vgb(U0))
    o.init_timestep()
    for _ in range(2):
        o.advance_step()
    assert np.allclose(o.soa(), U0, rtol=1e-13, atol=1e-15)


def test_powell_source_takes_fine_face_fields():
    """The 8-wave source of a coarse cell next to finer blocks uses, at the
    resolution-change face, the mean of the four fine faces' normal field
    (left and right states separately, src/ModConserveFlux.f90:136-140,
    355-372): recomputed here from the fine blocks' own face values and the
    plain per-block source routine, it equals the session's source; every
    other block's source is the plain per-block one; and the correction
    changes the coarse cells' source (the check is not vacuous)."""
    g = _grid()
    U0 = smooth_mhd_soa(g, seed=31, amp=0.3)
    cfg = po.config_from_dict(dict(flux="hlld", limiter="mc", beta=1.5,
                                   UseDivbSource=True), g)
    o = po.Session(g, cfg, g.soa_to_state_vgb(U0))
    o.exchange()
    src = o.sources()
    S = o.State
    nb = g.nI
    H = nb // 2

    def rc_flags(b):
        return [int(g.iNeighbor_EB[b, s] < -1) for s in range(6)]

    def faces_of(b):
        return po.calc_face_value(cfg, np.ascontiguousarray(S[b]),
                                  res_change=rc_flags(b))

    def face_view(F, d, f):
        # the face layer f normal to d, indexed [q, p] as the slab table
        return F[:, :, f] if d == 0 else F[:, f, :] if d == 1 else F[f]

    n_corrected = 0
    for b in range(g.nBlock):
        faces = faces_of(b)
        plain = po.calc_source(cfg, g.CellSize_DB[b], S[b],
                               [f.copy() for f in faces])
        for side in range(6):
            t = -2 - g.iNeighbor_EB[b, side]
            if t < 0 or g.iSlab_IS[t][2] != 1:
                continue
            d, hi = side // 2, side % 2
            kids = g.iSlab_IS[t][3:7]
            L = face_view(faces[2 * d], d, nb if hi else 0)
            R = face_view(faces[2 * d + 1], d, nb if hi else 0)
            for q in range(nb):
                for p in range(nb):
                    kid = kids[(p // H) + 2 * (q // H)]
                    kf = faces_of(kid)
                    KL = face_view(kf[2 * d], d, 0 if hi else nb)
                    KR = face_view(kf[2 * d + 1], d, 0 if hi else nb)
                    pf, qf = 2 * (p % H), 2 * (q % H)
                    for X, KX in ((L, KL), (R, KR)):
                        v = [KX[qf, pf, 4 + d], KX[qf, pf + 1, 4 + d],
                             KX[qf + 1, pf, 4 + d], KX[qf + 1, pf + 1, 4 + d]]
                        X[q, p, 4 + d] = ((0.25 * v[0] + 0.25 * v[1]) +
                                          0.25 * v[2]) + 0.25 * v[3]
            n_corrected += 1
        want = po.calc_source(cfg, g.CellSize_DB[b], S[b],
                              [f.copy() for f in faces])
        assert np.array_equal(src[b], want), b
        if any(g.iNeighbor_EB[b, s] < -1 and
               g.iSlab_IS[-2 - g.iNeighbor_EB[b, s]][2] == 1
               for s in range(6)):
            assert np.abs(want - plain).max() > 1e-6
    assert n_corrected > 0
