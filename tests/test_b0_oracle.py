"""B0 splitting in the oracle (UseB0: the state's field is B1 = B - B0 with a
static B0, src/ModB0.f90, src/ModFaceFlux.f90:1166-1899,2826-3160,
3853-4252, src/ModCalcSource.f90:862-977,1694-1889), pinned by what the
mathematics fixes:

* for a face with a uniform B0 the split and the unsplit systems are the
  same equations: the split face flux (Rusanov, Linde, HLLD) equals the
  flux of the full field B0 + B1 minus the B0 Maxwell stress
  (-B0n B0 + |B0|^2/2 n) in the momentum, and minus F_B . B0 in the energy
  (E = E1 + B0.B1 + |B0|^2/2), with the same mass and induction fluxes;
  the fast speeds of the split state equal those of the full state;
* with L = R every solver returns the physical flux of the split state;
* centred differences of a linear B0 are exact: div B0 and curl B0 of
  (a x, b y, c z) and of w (-y, x, 0) come out as a + b + c and (0, 0, 2w);
* with a uniform B0 the 8-wave momentum source of the split state is the
  unsplit one, -div(B1) (B1 + B0), and the induction and energy sources
  are those without B0.
A dropped or sign-flipped B0 term in any of these fails one of them."""
import numpy as np
import pytest

import pyoracle as po

RNG = np.random.default_rng(20261203)


def _state():
    return np.array([RNG.uniform(0.5, 1.5), *RNG.uniform(-1, 1, 3),
                     *RNG.uniform(-1, 1, 3), RNG.uniform(0.3, 1.5)])


def _stress(b0, d):
    s = -b0[d] * b0
    s[d] += 0.5 * b0 @ b0
    return s


def test_physical_flux_split_equals_full_minus_b0_stress():
    cfg = po.make_config()
    for t in range(60):
        S, B0, d = _state(), RNG.uniform(-1, 1, 3), t % 3
        cons, Fs, un = po.physical_flux_b0(cfg, d, S, B0)
        Sf = S.copy()
        Sf[4:7] += B0
        consf, Ff, unf = po.physical_flux(cfg, d, Sf)
        assert un == unf
        assert Ff[0] == Fs[0] and Ff[7] == Fs[7]
        np.testing.assert_allclose(Ff[1:4] - Fs[1:4], _stress(B0, d),
                                   rtol=0, atol=1e-14)
        np.testing.assert_allclose(Ff[4:7], Fs[4:7], rtol=0, atol=1e-14)
        np.testing.assert_allclose(Ff[8] - Fs[8], Fs[4:7] @ B0, rtol=0,
                                   atol=1e-14)
        # the conserved energy of the split state holds B1 only
        assert abs(cons[8] - (consf[8] - B0 @ S[4:7] - 0.5 * B0 @ B0)) \
            < 1e-14


def test_fast_speeds_use_the_full_field():
    cfg = po.make_config()
    for t in range(30):
        S, B0, d = _state(), RNG.uniform(-1, 1, 3), t % 3
        Sf = S.copy()
        Sf[4:7] = S[4:7] + B0
        assert po.speed_max_b0(cfg, d, S, B0) == po.speed_max(cfg, d, Sf)
        # and not the split field alone (unless B0 is negligible)
        assert po.speed_max_b0(cfg, d, S, B0)[1] != po.speed_max(cfg, d, S)[1]


@pytest.mark.parametrize("flux", ["rusanov", "linde", "hlld"])
def test_numerical_flux_split_equals_full_minus_b0_stress(flux):
    cfg = po.make_config(flux=flux)
    worst = 0.0
    for t in range(150):
        L = _state()
        R = L * (1 + 0.05 * RNG.standard_normal(8))
        B0, d = RNG.uniform(-1, 1, 3), t % 3
        e1, Fs = po.numerical_flux_b0(cfg, d, 1.0, L, R, B0)
        Lf, Rf = L.copy(), R.copy()
        Lf[4:7] += B0
        Rf[4:7] += B0
        e2, Ff = po.numerical_flux(cfg, d, 1.0, Lf, Rf)
        assert e1 == 0 and e2 == 0
        dev = [Ff[0] - Fs[0],
               *((Ff[1:4] - Fs[1:4]) - _stress(B0, d)),
               *(Ff[4:7] - Fs[4:7]),
               (Ff[8] - Fs[8]) - Fs[4:7] @ B0,
               Ff[11] - Fs[11]]          # Vdt: the same wave speeds
        worst = max(worst, np.abs(dev).max())
    assert worst < 1e-13, worst


@pytest.mark.parametrize("flux", ["rusanov", "linde", "hlld"])
def test_equal_states_give_the_physical_flux(flux):
    cfg = po.make_config(flux=flux)
    for t in range(30):
        S, B0, d = _state(), RNG.uniform(-1, 1, 3), t % 3
        err, F = po.numerical_flux_b0(cfg, d, 0.7, S, S, B0)
        assert err == 0
        _, Fp, _ = po.physical_flux_b0(cfg, d, S, B0)
        np.testing.assert_allclose(F[[0, 1, 2, 3, 4, 5, 6, 8]],
                                   0.7 * Fp[[0, 1, 2, 3, 4, 5, 6, 8]],
                                   rtol=1e-13, atol=1e-14)


def _block_b0(cfg, field, cell=(0.1, 0.2, 0.15), org=(0.3, -0.2, 0.5)):
    """Ghosted B0_DGB (nKg, nJg, nIg, 3) of one block from field(x, y, z)."""
    g = cfg.nG
    i = (np.arange(-g, cfg.nI + g) + 0.5) * cell[0] + org[0]
    j = (np.arange(-g, cfg.nJ + g) + 0.5) * cell[1] + org[1]
    k = (np.arange(-g, cfg.nK + g) + 0.5) * cell[2] + org[2]
    z, y, x = np.meshgrid(k, j, i, indexing="ij")
    return np.stack(field(x, y, z), axis=-1), np.asarray(cell)


def test_div_and_curl_of_linear_b0_are_exact():
    cfg = po.make_config()
    a, b, c, w = 0.7, -0.3, 1.1, 0.45
    B0, cs = _block_b0(cfg, lambda x, y, z: (a * x, b * y, c * z))
    div, curl = po.set_b0_source(cfg, cs, po.set_b0_face(cfg, B0))
    np.testing.assert_allclose(div, a + b + c, rtol=1e-12)
    np.testing.assert_allclose(curl, 0.0, atol=1e-12)
    B0, cs = _block_b0(cfg, lambda x, y, z: (-w * y, w * x, 0 * z))
    div, curl = po.set_b0_source(cfg, cs, po.set_b0_face(cfg, B0))
    np.testing.assert_allclose(div, 0.0, atol=1e-12)
    np.testing.assert_allclose(curl[..., :2], 0.0, atol=1e-12)
    np.testing.assert_allclose(curl[..., 2], 2 * w, rtol=1e-12)
    # face B0 is the mean of the two cells' B0 (exact for a linear field):
    # the x face between cells i-1 and i sits at the cell edge
    fx, fy, fz = po.set_b0_face(cfg, B0)
    xe = 0.3 + np.arange(cfg.nI + 1) * cs[0]
    np.testing.assert_allclose(fx[0, 0, :, 1], w * xe, rtol=1e-12)


def _block_state(cfg):
    g = cfg.nG
    shp = (cfg.nK + 2 * g, cfg.nJ + 2 * g, cfg.nI + 2 * g)
    S = np.zeros(shp + (8,))
    S[..., 0] = RNG.uniform(0.8, 1.2, shp)
    S[..., 1:4] = RNG.uniform(-0.3, 0.3, shp + (3,))
    S[..., 4:7] = RNG.uniform(-0.5, 0.5, shp + (3,))
    S[..., 7] = RNG.uniform(0.5, 1.0, shp)
    return S


def test_powell_source_with_uniform_b0_is_the_unsplit_one():
    cfg = po.make_config(flux="hlld", limiter="mc3", UseDivbSource=True)
    S = _block_state(cfg)
    faces = po.calc_face_value(cfg, S)
    b = np.array([0.4, -0.9, 0.25])
    B0, cs = _block_b0(cfg, lambda x, y, z: (0 * x + b[0], 0 * y + b[1],
                                             0 * z + b[2]))
    f0 = po.set_b0_face(cfg, B0)
    div, curl = po.set_b0_source(cfg, cs, f0)
    assert np.abs(div).max() == 0 and np.abs(curl).max() == 0
    src_b0 = po.calc_source_b0(cfg, cs, S, faces, B0, f0, div, curl)
    src = po.calc_source(cfg, cs, S, faces)
    # div B1 of each cell from the no-B0 source: S_B = -div(B1) u
    g = cfg.nG
    inner = S[g:-g, g:-g, g:-g]
    u = inner[..., 1:4] / inner[..., :1]
    divb = -src[..., 4] / u[..., 0]
    np.testing.assert_allclose(src_b0[..., 4:7], src[..., 4:7], rtol=0,
                               atol=1e-12)
    np.testing.assert_allclose(src_b0[..., 8], src[..., 8], rtol=0,
                               atol=1e-12)
    np.testing.assert_allclose(src_b0[..., 1:4],
                               src[..., 1:4] - divb[..., None] * b,
                               rtol=1e-9, atol=1e-11)


def test_b0_source_terms_follow_div_and_curl():
    """With a curl-carrying linear B0 the B0 source adds
    -B1 div(B0) - curl(B0) x B1 to the momentum; without UseB0Source it
    does not."""
    cfg = po.make_config(flux="rusanov", limiter="mc", UseDivbSource=True)
    S = _block_state(cfg)
    faces = po.calc_face_value(cfg, S)
    a, w = 0.3, 0.8
    B0, cs = _block_b0(cfg, lambda x, y, z: (a * x - w * y, w * x, 0 * z))
    f0 = po.set_b0_face(cfg, B0)
    div, curl = po.set_b0_source(cfg, cs, f0)
    on = po.calc_source_b0(cfg, cs, S, faces, B0, f0, div, curl, True)
    off = po.calc_source_b0(cfg, cs, S, faces, B0, f0, div, curl, False)
    g = cfg.nG
    B1 = S[g:-g, g:-g, g:-g, 4:7]
    expect = -B1 * a - np.cross(np.array([0, 0, 2 * w]), B1)
    np.testing.assert_allclose(on[..., 1:4] - off[..., 1:4], expect,
                               rtol=1e-10, atol=1e-12)
    np.testing.assert_array_equal(on[..., 4:], off[..., 4:])
