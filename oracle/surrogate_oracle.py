"""CPU oracle (TEST INFRASTRUCTURE ONLY) of the surrogate stage: one stage of
the explicit update in which the smooth faces take the network's flux.

The reference has no learned flux; what this restates is the stage built
around it, face by face:

  1. face states of the limited MUSCL reconstruction (the oracle's
     calc_face_value, src/ModFaceValue.f90:1970-2163);
  2. the smoothness gate: for every variable and for both cells next to the
     face, the MC limiter leaves the slope unclipped (the two one-sided
     differences agree in sign and their mean is at most beta times the
     smaller one) or the two one-sided differences, scaled like the network
     inputs, add up to at most JumpTol; and the plasma beta 2p/B^2 of both
     face states is at least BetaMin.  (The reference's closest analogue is
     the low order face criterion, src/ModFaceValue.f90:2913-2963, which
     also decides per face between two flux paths.)
  3. network faces (gate passed, non-zero jump): the Rusanov flux
     (src/ModFaceFlux.f90:2513-2559) plus the network's correction
     y * |j| * scale * area in the rotated slots, with the face's time-step
     speed from the HLLD wave speeds of the Linde-corrected states
     (src/ModFaceFlux.f90:2862-2872, 2302-2320); in the Rusanov mode the
     same faces take the Rusanov flux and its own speed; every other face
     the HLLD flux with the Linde fix (src/ModFaceFlux.f90:2826-3160,
     2403-2418);
  4. at resolution changes the coarse face takes the sum of the four fine
     face fluxes as the fine blocks use them (src/ModConserveFlux.f90);
  5. the FP64 stage update (update_state, src/ModUpdateState.f90:191-965)
     and calc_timestep (src/ModTimeStepControl.f90:157-168).

Network inputs (for a face normal to d, states rotated to (rho, u_n,
u_t1, u_t2, B_n, B_t1, B_t2, p) and scaled by rho_s = (rho_L + rho_R)/2
and c_s = max(c_fL, c_fR)): the mean of the scaled states and the unit
direction of their difference j; the outputs, times |j| and the output
scales rho_s c_s, rho_s c_s^2, c_s^2 sqrt(rho_s), rho_s c_s^3, are the
flux correction per unit area in the rotated frame.  The network itself is
mlp_oracle.forward with the GPU's BF16 rounding points.

Arithmetic is FP64 throughout; the GPU stage computes its faces in FP32,
so the two agree to FP32 rounding except where a face's gate quantities lie
within that rounding of a threshold; gate_margin reports those faces.

Only tests/ import this module.
"""
from __future__ import annotations

import numpy as np

import mlp_oracle as mo
import pyoracle as po

OB_VDT = 11
GAMMA_DEFAULT = 5.0 / 3.0
MODES = ("network", "rusanov")


def _rot(d):
    """rotated component c -> original variable (and energy slot 8 last)"""
    t1, t2 = (d + 1) % 3, (d + 2) % 3
    return [0, 1 + d, 1 + t1, 1 + t2, 4 + d, 4 + t1, 4 + t2]


def fast_speed(S, d, gamma):
    """normal fast magnetosonic speed of primitive states S[..., 8]"""
    rho, p = S[..., 0], S[..., 7]
    B = S[..., 4:7]
    s2 = gamma * p / rho
    a2 = (B * B).sum(-1) / rho
    an2 = B[..., d] ** 2 / rho
    f2 = s2 + a2
    return np.sqrt(0.5 * (f2 + np.sqrt(np.maximum(0.0, f2 * f2 -
                                                   4 * s2 * an2))))


def out_scale(rhos, cs):
    """(n, 8) output scales in the rotated order (rho, m_n, m_t1, m_t2,
    B_n, B_t1, B_t2, E)"""
    sq = np.sqrt(rhos)
    return np.stack([rhos * cs] + [rhos * cs * cs] * 3 +
                    [cs * cs * sq] * 3 + [rhos * cs ** 3], axis=-1)


def face_inputs(L, R, d, gamma):
    """network inputs x (n, 16), rho_s, c_s and |j| of face states L, R"""
    rhos = 0.5 * (L[:, 0] + R[:, 0])
    cs = np.maximum(fast_speed(L, d, gamma), fast_speed(R, d, gamma))
    rot = _rot(d) + [7]
    sc = np.stack([rhos] + [cs] * 3 + [cs * np.sqrt(rhos)] * 3 +
                  [rhos * cs * cs], axis=-1)
    sL, sR = L[:, rot] / sc, R[:, rot] / sc
    j = sR - sL
    jn = np.sqrt((j * j).sum(-1))
    ij = np.where(jn > 0, 1.0 / np.where(jn > 0, jn, 1.0), 0.0)
    x = np.concatenate([0.5 * (sL + sR), j * ij[:, None]], axis=1)
    return x, rhos, cs, jn


def gate(q, L, R, d, beta, betaMin, jumpTol, rhos, cs):
    """Smooth faces and, for each face, the smallest relative distance of
    any gate quantity to its threshold (the margin).  q: the four stencil
    cells' primitive states (4, n, 8)."""
    tol = jumpTol * np.stack([rhos] + [cs] * 3 + [cs * np.sqrt(rhos)] * 3 +
                             [rhos * cs * cs], axis=-1)
    ok = np.ones(L.shape[0], bool)
    margin = np.full(L.shape[0], np.inf)

    def rel(a, b):
        return np.abs(a - b) / np.maximum(np.maximum(np.abs(a), np.abs(b)),
                                          1e-300)
    for c in (0, 1):
        d1, d2 = q[c + 1] - q[c], q[c + 2] - q[c + 1]
        small = np.abs(d1) + np.abs(d2)
        mean = 0.5 * np.abs(d1 + d2)
        lim = beta * np.minimum(np.abs(d1), np.abs(d2))
        v_ok = (small <= tol) | ((d1 * d2 > 0) & (mean <= lim))
        ok &= v_ok.all(axis=1)
        # margins only for the comparison that decided each variable
        m_small = rel(small, tol)
        m_mc = np.minimum(rel(mean, lim),
                          np.minimum(np.abs(d1), np.abs(d2)) /
                          np.maximum(np.abs(d1) + np.abs(d2), 1e-300))
        m = np.where(small <= tol, m_small, np.minimum(m_small, m_mc))
        margin = np.minimum(margin, m.min(axis=1))
    for S in (L, R):
        b2 = (S[:, 4:7] ** 2).sum(-1)
        ok &= 2 * S[:, 7] >= betaMin * b2
        if betaMin > 0:
            margin = np.minimum(margin, rel(2 * S[:, 7], betaMin * b2))
    return ok, margin


def _primitive(blk):
    P = blk.copy()
    with np.errstate(invalid="ignore", divide="ignore"):
        P[..., 1:4] = blk[..., 1:4] / blk[..., :1]
    return P


def _stencils(P, d, nI, nJ, nK):
    """primitive states of the four cells f-2..f+1 of every face normal to
    d, in the face-array order of calc_face_value (k, j, i), flattened"""
    out = []
    for s in range(4):
        if d == 0:
            a = P[2:2 + nK, 2:2 + nJ, s:s + nI + 1]
        elif d == 1:
            a = P[2:2 + nK, s:s + nJ + 1, 2:2 + nI]
        else:
            a = P[s:s + nK + 1, 2:2 + nJ, 2:2 + nI]
        out.append(a.reshape(-1, 8))
    return np.stack(out)


def block_faces(cfg, blk, cell_size, params, mode, betaMin, jumpTol,
                mistake=None):
    """The face fluxes of one block (calc_face_flux layout, 14 slots, times
    area) with the network on the smooth faces; returns the three flux
    arrays, the smooth-face mask and gate margins per direction, and the
    number of network faces."""
    gamma = cfg.Gamma
    cs_ = np.ascontiguousarray(cell_size, dtype=np.float64)
    faces = po.calc_face_value(cfg, blk)
    cfg_r = po.make_config(nI=cfg.nI, nJ=cfg.nJ, nK=cfg.nK, flux="rusanov",
                           limiter=cfg.iLimiter, beta=cfg.LimiterBeta,
                           gamma=gamma)
    cfg_h = po.make_config(nI=cfg.nI, nJ=cfg.nJ, nK=cfg.nK, flux="hlld",
                           limiter=cfg.iLimiter, beta=cfg.LimiterBeta,
                           gamma=gamma, cTiny=cfg.cTiny)
    _, FR = po.calc_face_flux(cfg_r, cs_, [f.copy() for f in faces])
    _, FH = po.calc_face_flux(cfg_h, cs_, [f.copy() for f in faces])
    P = _primitive(blk)
    out, smooth_all, margin_all = [], [], []
    nNet = 0
    for d in range(3):
        L = faces[2 * d].reshape(-1, 8)
        R = faces[2 * d + 1].reshape(-1, 8)
        shp = faces[2 * d].shape[:3]
        fr = FR[d].reshape(-1, 14)
        fh = FH[d].reshape(-1, 14)
        area = [cs_[1] * cs_[2], cs_[0] * cs_[2], cs_[0] * cs_[1]][d]
        x, rhos, csf, jn = face_inputs(L, R, d, gamma)
        q = _stencils(P, d, cfg.nI, cfg.nJ, cfg.nK)
        smooth, margin = gate(q, L, R, d, cfg.LimiterBeta, betaMin, jumpTol,
                              rhos, csf)
        smooth &= csf > 0
        F = fh.copy()
        F[smooth] = fr[smooth]
        if mode == "network":
            # network faces: the Rusanov flux plus the correction; the
            # time-step speed is HLLD's
            F[smooth, OB_VDT] = fh[smooth, OB_VDT]
            net = smooth & (jn > 0)
            nNet += int(net.sum())
            if net.any():
                y = mo.forward(params, x[net].astype(np.float32),
                               rounding=True)
                corr = y * jn[net, None] * out_scale(rhos[net], csf[net])
                if mistake == "no_jump":
                    corr = corr / jn[net, None]
                if mistake == "sign":
                    corr = -corr
                a = area if mistake != "no_area" else 1.0
                slots = _rot(d) + [8]
                if mistake == "slot":
                    slots = slots[:1] + slots[2:4] + slots[1:2] + slots[4:]
                for c, sl in enumerate(slots):
                    F[net, sl] += corr[:, c] * a
        out.append(F.reshape(shp + (14,)))
        smooth_all.append(smooth.reshape(shp))
        margin_all.append(margin.reshape(shp))
    return out, smooth_all, margin_all, nNet


def _fine_sum_correction(grid, fluxes):
    """coarse faces of finer sides take the sum of the four fine faces'
    fluxes (sub-face order (0,0), (1,0), (0,1), (1,1) in (p, q))"""
    nb = grid.nI
    H = nb // 2
    for sl in getattr(grid, "iSlab_IS", np.zeros((0, 8), int)):
        b, side, kind = int(sl[0]), int(sl[1]), int(sl[2])
        if kind != 1:
            continue
        d, hi = side // 2, side % 2
        fc = nb if hi else 0
        ff = 0 if hi else nb

        def view(blk, fn):
            F = fluxes[blk][d]   # (k, j, i, 14) with the d index = face
            if d == 0:
                return F[:, :, fn]           # (q=k, p=j)
            if d == 1:
                return F[:, fn, :]           # (q=k, p=i)
            return F[fn]                     # (q=j, p=i)
        C = view(b, fc)
        for q in range(nb):
            for p in range(nb):
                kid = int(sl[3 + (p // H) + 2 * (q // H)])
                Fk = view(kid, ff)
                pf, qf = 2 * (p % H), 2 * (q % H)
                s = ((Fk[qf, pf, :9] + Fk[qf, pf + 1, :9]) +
                     Fk[qf + 1, pf, :9]) + Fk[qf + 1, pf + 1, :9]
                C[q, p, :9] = s
                C[q, p, OB_VDT] = ((Fk[qf, pf, OB_VDT] +
                                    Fk[qf, pf + 1, OB_VDT]) +
                                   Fk[qf + 1, pf, OB_VDT]) + \
                    Fk[qf + 1, pf + 1, OB_VDT]


def surrogate_stage(grid, scheme, U, Uold, iStage, Dt, params, mode="network",
                    betaMin=0.0, jumpTol=0.05, mistake=None):
    """One stage from the interior states U (stage input) and Uold (start of
    the step), SoA (8, nBlock, nK, nJ, nI), FP64.  Dt is the (Cfl-scaled)
    time step.  Returns the updated SoA state, DtMax_B of the face speeds,
    the number of network faces and the per-cell gate margin (smallest
    margin over the cell's six faces)."""
    assert mode in MODES
    cfg = po.config_from_dict(scheme, grid)
    o = po.Session(grid, cfg, grid.soa_to_state_vgb(U))
    o.exchange()                      # ghost cells (and slabs) of U
    S = o.State
    Sold = grid.soa_to_state_vgb(Uold)
    nb = grid.nBlock
    fluxes, margins, nNet = [], [], 0
    for b in range(nb):
        fl, _, mg, n = block_faces(cfg, np.ascontiguousarray(S[b]),
                                   grid.CellSize_DB[b], params, mode,
                                   betaMin, jumpTol, mistake)
        fluxes.append(fl)
        margins.append(mg)
        nNet += n
    _fine_sum_correction(grid, fluxes)
    out = np.empty_like(U)
    dtb = np.empty(nb)
    cm = np.empty((nb, grid.nK, grid.nJ, grid.nI))
    DtCB = Dt / cfg.Cfl
    for b in range(nb):
        blk = np.ascontiguousarray(S[b]).copy()
        old = np.ascontiguousarray(Sold[b]).copy()
        cs_ = np.ascontiguousarray(grid.CellSize_DB[b], dtype=np.float64)
        fl = [np.ascontiguousarray(f) for f in fluxes[b]]
        po.update_state(cfg, iStage, Dt, cs_,
                        np.full(grid.nI * grid.nJ * grid.nK, DtCB), fl, blk,
                        old)
        out[:, b] = np.moveaxis(blk[2:-2, 2:-2, 2:-2], -1, 0)
        _, dtb[b] = po.calc_timestep(cfg, cs_, fl)
        mx, my, mz = margins[b]
        cm[b] = np.minimum.reduce([mx[:, :, :-1], mx[:, :, 1:],
                                   my[:, :-1, :], my[:, 1:, :],
                                   mz[:-1], mz[1:]])
    return out, dtb, nNet, cm
