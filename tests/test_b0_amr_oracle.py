"""B0 splitting on two-level grids in the oracle (the reference's general
update: set_b0_face with set_b0_reschange, src/ModB0.f90:338-351,366-455,
and the fine blocks' face fluxes with their B0 before save_cons_flux,
src/ModAdvanceExplicit.f90:95-118).  Pinned by:

* the discrete divergence and curl theorems: summed over every cell of the
  grid, DivB0 V and CurlB0 V equal the B0 flux and n x B0 through the
  faces on the domain boundary, to rounding, for a dipole B0.  The
  interior faces cancel only where the coarse face at a resolution change
  carries the mean of the four fine faces' B0 (all three components); the
  coarse block's own face average misses by far more than the tolerance;
* the coarse face equals the mean of the fine blocks' own faces it covers
  (the sums above do not see which fine faces a coarse face takes; this
  does);
* with a uniform B0 the split run (B1 = B - B0) and the unsplit run of the
  full field are the same equations on the two-level grid: their states
  agree to rounding over two steps, with and without the 8-wave source
  (a fine-face flux without the fine block's B0 breaks this)."""
import numpy as np
import pytest

import pyoracle as po
from batsrus_amd.amr import AmrGrid
from helpers import b0_dgb, block_geometry, smooth_mhd_soa


def _grid():
    refine = np.zeros((2, 2, 2), bool)
    refine[0, 0, 0] = True
    refine[1, 1, 0] = True
    return AmrGrid((2, 2, 2), refine, nCell=8)


def _dipole(x, y, z, c=(0.35, 0.6, -0.4), m=(0.2, -0.3, 1.0)):
    rx, ry, rz = x - c[0], y - c[1], z - c[2]
    r2 = rx * rx + ry * ry + rz * rz
    r5 = r2 ** 2.5
    mr = m[0] * rx + m[1] * ry + m[2] * rz
    return ((3 * mr * rx - m[0] * r2) / r5 * 0.05,
            (3 * mr * ry - m[1] * r2) / r5 * 0.05,
            (3 * mr * rz - m[2] * r2) / r5 * 0.05)


def _cfg(g, **kw):
    return po.config_from_dict(dict(flux="hlld", limiter="mc", beta=1.5,
                                    cfl=0.8, **kw), g)


def _faces(X, Y, Z):
    """The six sides' face B0 of a block, (low, high) per direction, each
    (n, n, 3) in the side's in-plane order."""
    return [(X[:, :, 0], X[:, :, -1]), (Y[:, 0], Y[:, -1]), (Z[0], Z[-1])]


def test_b0_div_and_curl_telescope_across_resolution_changes():
    g = _grid()
    n = g.nI
    B0 = b0_dgb(g, _dipole)
    U0 = smooth_mhd_soa(g, seed=71, amp=0.1)
    o = po.Session(g, _cfg(g), g.soa_to_state_vgb(U0), B0_DGB=B0)
    org, dx = block_geometry(g)
    eps = 1e-9 * (g.xyzMax - g.xyzMin)
    tot = np.zeros(4)     # sum DivB0 V, sum CurlB0 V
    bnd = np.zeros(4)     # the boundary faces' B0 . n A, n x B0 A
    scale = 0.0
    unc = 0.0             # the residual with the coarse blocks' own faces
    for b in range(g.nBlock):
        X, Y, Z, D, Cu = o.b0(b)
        V = np.prod(dx[b])
        tot[0] += D.sum() * V
        tot[1:] += Cu.reshape(-1, 3).sum(0) * V
        A = [dx[b, 1] * dx[b, 2], dx[b, 0] * dx[b, 2], dx[b, 0] * dx[b, 1]]
        for d, pair in enumerate(_faces(X, Y, Z)):
            for hi, F in enumerate(pair):
                Fs = F.reshape(-1, 3).sum(0) * A[d]
                scale += np.abs(F[..., d]).sum() * A[d]
                sgn = 1.0 if hi else -1.0
                nb = g.iNeighbor_EB[b, 2 * d + hi]
                if nb <= -2 and g.iSlab_IS[-2 - nb][2] == 1:
                    # the coarse block's own average of its ghost and first
                    # cell at this side, against the corrected face
                    G = B0[b]
                    sl = [slice(2, 2 + n)] * 3
                    sg, sc = list(sl), list(sl)
                    sg[2 - d] = n + 2 if hi else 1
                    sc[2 - d] = n + 1 if hi else 2
                    own = 0.5 * (G[tuple(sg)] + G[tuple(sc)])
                    unc += sgn * (own[..., d].sum() - F[..., d].sum()) * A[d]
                edge = org[b, d] + hi * n * dx[b, d]
                if not (abs(edge - g.xyzMin[d]) < eps[d] or
                        abs(edge - g.xyzMax[d]) < eps[d]):
                    continue
                nv = np.zeros(3)
                nv[d] = sgn
                bnd[0] += sgn * Fs[d]
                bnd[1:] += np.cross(nv, Fs)
    res = np.abs(tot - bnd) / scale
    assert res.max() < 1e-13, res
    # without the correction the interior faces would not cancel
    assert abs(unc) / scale > 1e-8, unc / scale


def test_coarse_face_is_the_mean_of_the_fine_faces():
    g = _grid()
    n, h = g.nI, g.nI // 2
    B0 = b0_dgb(g, _dipole)
    U0 = smooth_mhd_soa(g, seed=72, amp=0.1)
    o = po.Session(g, _cfg(g), g.soa_to_state_vgb(U0), B0_DGB=B0)
    faces = [_faces(*o.b0(b)[:3]) for b in range(g.nBlock)]
    checked = 0
    for sl in g.iSlab_IS:
        b, side, kind = sl[0], sl[1], sl[2]
        if kind != 1:
            continue
        d, hi = side // 2, side % 2
        C = faces[b][d][hi]                 # (q, p, 3)
        for qq in range(2):
            for qp in range(2):
                Ff = faces[sl[3 + qp + 2 * qq]][d][1 - hi]
                mean = 0.25 * (Ff[0::2, 0::2] + Ff[0::2, 1::2] +
                               Ff[1::2, 0::2] + Ff[1::2, 1::2])
                got = C[qq * h:(qq + 1) * h, qp * h:(qp + 1) * h]
                np.testing.assert_allclose(got, mean, rtol=1e-14,
                                           atol=1e-15 * np.abs(mean).max())
                checked += 1
    assert checked >= 24


@pytest.mark.parametrize("powell", [False, True])
def test_uniform_b0_two_level_split_is_the_unsplit_run(powell):
    g = _grid()
    U0 = smooth_mhd_soa(g, seed=73, amp=0.2)
    b0 = np.array([0.35, -0.2, 0.15])
    shp = (g.nBlock, g.nK + 2 * g.nG, g.nJ + 2 * g.nG, g.nI + 2 * g.nG, 3)
    B0 = np.broadcast_to(b0, shp).copy()
    U1 = U0.copy()
    U1[4:7] = U0[4:7] - b0[:, None, None, None, None]

    def run(U, **kw):
        o = po.Session(g, _cfg(g, UseDivbSource=powell),
                       g.soa_to_state_vgb(U), **kw)
        o.exchange()
        o.init_timestep()
        for _ in range(2):
            o.s.nStep += 1
            o.set_global_timestep()
            for st in (1, 2):
                assert o.advance_stage(st) == 0
            o.exchange()
        return o.soa()

    W = run(U1, B0_DGB=B0, UseB0Source=True)
    W[4:7] += b0[:, None, None, None, None]
    ref = run(U0)
    scale = np.abs(ref).reshape(8, -1).max(axis=1)[:, None, None, None,
                                                   None]
    assert (np.abs(W - ref) / scale).max() < 1e-12
