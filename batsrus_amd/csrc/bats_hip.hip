// bats_hip.hip -- the C ABI (include/bats_hip.h) of the gfx950 explicit
// block-adaptive ideal-MHD stage update, the reference-shaped per-block
// kernels used for parity dumps, and the small grid-level kernels (global
// time step, NaN check, domain integrals).  The fused stage kernel lives in
// stage_kernel.hpp and is instantiated in stage_inst.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/bats_hip.h"
#include "stage_kernel.hpp"
#include "handle.hpp"

using namespace bats;

namespace {

// ---------------------------------------------------------------------
// per-block kernels (reference-shaped routines for parity dumps)
// ---------------------------------------------------------------------
// ---------------------------------------------------------------------
// set_phys_based_low_order_face (src/ModFaceValue.f90:2839-2911): one
// thread per face, the weight from the four cells f-2 .. f+1 along the
// face normal through low_order_face_criteria (:2913-2963), written in
// crit_index order (thread t writes entry t)
// ---------------------------------------------------------------------
struct LowOrderArgs {
  const void *U;          // state (type T) with its halos
  long vs, nCellTot;
  const int *nei;
  int nBlock;
  double Gamma, pCritLow, pCritHigh, VelCrit;
  void *crit;             // type T
};

template <int NB, typename T>
__global__ void __launch_bounds__(256) k_low_order_crit(LowOrderArgs a) {
  constexpr int NFX = (NB + 1) * NB, NFD = NB * NFX;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)a.nBlock * 3 * NFD) return;
  const int b = (int)(t / (3 * NFD)), r = (int)(t % (3 * NFD));
  const int dir = r / NFD, s = r % NFD;
  int i, j, k;  // face coordinates, the dir component is the face index
  if (dir == 0) { k = s / NFX; j = (s % NFX) / (NB + 1); i = s % (NB + 1); }
  else if (dir == 1) { k = s / NFX; i = (s % NFX) / (NB + 1); j = s % (NB + 1); }
  else { k = s / (NB * NB); j = (s / NB) % NB; i = s % NB; }
  const T *U = static_cast<const T *>(a.U);
  double pTotal[4], VelMax = 0, VelMin = 0, SoundMin = 0;
  for (int l = 0; l < 4; l++) {
    int ci = i, cj = j, ck = k;
    if (dir == 0) ci = i - 2 + l;
    else if (dir == 1) cj = j - 2 + l;
    else ck = k - 2 + l;
    const long off = cell_off<NB>(a.nei, b, ci, cj, ck, a.nCellTot);
    const double Rho = double(U[RHO * a.vs + off]);
    const double p = double(U[PP * a.vs + off]);
    const double Vel = double(U[(MX + dir) * a.vs + off]) / Rho;
    const double Sound = sqrt(p / Rho * a.Gamma);
    pTotal[l] = p;
    if (l == 0) {
      VelMax = Vel; VelMin = Vel; SoundMin = Sound;
    } else {
      VelMax = fmax2(VelMax, Vel);
      VelMin = fmin2(VelMin, Vel);
      SoundMin = fmin2(SoundMin, Sound);
    }
    const double bx = double(U[BX * a.vs + off]), by = double(U[BY * a.vs + off]),
                 bz = double(U[BZ * a.vs + off]);
    pTotal[l] = pTotal[l] + 0.5 * (bx * bx + by * by + bz * bz);
  }
  double Crit = 0;
  if (fmax2(0.0, (VelMax - VelMin) / SoundMin) > a.VelCrit) {
    double pMax = pTotal[0], pMin = pTotal[0];
    for (int l = 1; l < 4; l++) {
      pMax = fmax2(pMax, pTotal[l]);
      pMin = fmin2(pMin, pTotal[l]);
    }
    const double pRatio = pMax / pMin;
    if (pRatio >= a.pCritLow) Crit = 1;
    else if (pRatio < a.pCritHigh) Crit = 0;
    else Crit = (pRatio - a.pCritHigh) / (a.pCritLow - a.pCritHigh);
  }
  static_cast<T *>(a.crit)[t] = T(Crit);
}

struct BlockArgs {
  const double *Usnap;   // halo source (exchange_messages snapshot)
  double *State;         // current state
  double *StateOld;
  const int *nei;
  const double *cellsize;
  double *scal;
  double *LX, *RX, *LY, *RY, *LZ, *RZ;  // reference face layouts
  double *FX, *FY, *FZ;                 // Flux_V?I(14, ...)
  double *Src;                          // Source_VC(9, nI, nJ, nK)
  int useSrc;
  // resolution changes: slab list (8 ints per slab), slab kinds, the
  // fine sub-face records of the coarse sides; resOnly restricts the face
  // routines to the faces of sides whose neighbour is coarser
  // (DoResChangeOnly)
  const int *slabs, *slabKind;
  double *corr;
  int resOnly;
  // B0 splitting: face B0 and cell B0 / div / curl (null: no B0)
  const double *b0F, *b0C;
  int useB0Src;
  const double *lowCrit;                // low order weights or null
  double betaRC;                        // BetaLimiterResChange
  int nFaceRC;                          // its faces (0: not used)
  int tvd;                              // UseTvdResChange (second order)
  double tvdBeta;                       // min(BetaLimiterResChange, beta)
  int steady;                           // steady state: per-cell steps
  // the inner body: mask slot per block (-1: no body cell within the
  // block's halo), the true-cell masks over the block and its 2-cell halo,
  // first cell centres, the ghost face state (used = null: no body)
  const int *bodySlot;
  const unsigned char *used;
  const double *xyz0;
  double bodyRho, b1rCoef;
  double *dtmaxCB, *dtmaxB;
  int *err;
  int nBlock, iBlock;
  long vs, nCellTot;
  Phys ph;
  double beta;
  int lim, iFlux;
  double DtFactor;
  int iStage, nStage, halfStep, dbTrick, useDtFixed;
  double DtFixed;
};

// whether face (fi, fj, fk) normal to dir lies on a side of the block whose
// neighbour is coarser (DiLevel_EB = 1: a kind-0 slab), the faces that
// calc_face_value / calc_face_flux visit with DoResChangeOnly
__device__ inline bool coarser_side_face(const BlockArgs &a, int dir, int fi,
                                         int fj, int fk, int NB) {
  const int f = dir == 0 ? fi : dir == 1 ? fj : fk;
  if (f != 0 && f != NB) return false;
  const int n = a.nei[6 * a.iBlock + 2 * dir + (f == NB)];
  return n <= -2 && a.slabKind[-2 - n] == 0;
}

// the true-cell mask of the block (null: no body cell within its halo)
// and the flag of its cell (ii, jj, kk), each index in -2 .. NB+1
template <int NB>
__device__ inline const unsigned char *block_body_mask(const BlockArgs &a) {
  if (!a.used) return nullptr;
  const int sl = a.bodySlot[a.iBlock];
  return sl < 0 ? nullptr : a.used + (long)sl * (NB + 4) * (NB + 4) * (NB + 4);
}
template <int NB>
__device__ inline bool block_true(const unsigned char *m, int ii, int jj,
                                  int kk) {
  constexpr int NG = NB + 4;
  return !m || m[((kk + 2) * NG + (jj + 2)) * NG + (ii + 2)] != 0;
}

template <int NB>
__global__ void k_block_face_value(BlockArgs a) {
  const int nFx = (NB + 1) * NB * NB;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 3 * nFx) return;
  const int dir = t / nFx, r = t % nFx;
  int fi, fj, fk;
  double *L, *R;
  long fo;
  if (dir == 0) {  // LeftState_VX(nVar, nI+1, nJ, nK)
    fi = r % (NB + 1); fj = (r / (NB + 1)) % NB; fk = r / ((NB + 1) * NB);
    fo = r; L = a.LX; R = a.RX;
  } else if (dir == 1) {  // (nVar, nI, nJ+1, nK)
    fi = r % NB; fj = (r / NB) % (NB + 1); fk = r / (NB * (NB + 1));
    fo = r; L = a.LY; R = a.RY;
  } else {  // (nVar, nI, nJ, nK+1)
    fi = r % NB; fj = (r / NB) % NB; fk = r / (NB * NB);
    fo = r; L = a.LZ; R = a.RZ;
  }
  if (a.resOnly && !coarser_side_face(a, dir, fi, fj, fk, NB)) return;
  double q[4][NVAR];
  for (int s = 0; s < 4; s++) {
    int ii = fi, jj = fj, kk = fk;
    if (dir == 0) ii = fi - 2 + s;
    else if (dir == 1) jj = fj - 2 + s;
    else kk = fk - 2 + s;
    load_state(a.Usnap, a.vs, cell_off<NB>(a.nei, a.iBlock, ii, jj, kk, a.nCellTot), q[s]);
    to_primitive(q[s]);
  }
  double Lv[NVAR], Rv[NVAR];
  // betas of the face's two cells (f-1, f along the normal)
  double bL = a.beta, bR = a.beta;
  if (a.nFaceRC > 0) {
    const int f = dir == 0 ? fi : dir == 1 ? fj : fk;
    const bool lo = a.nei[6 * a.iBlock + 2 * dir] < 0;
    const bool hi = a.nei[6 * a.iBlock + 2 * dir + 1] < 0;
    bL = cell_beta(a.beta, a.betaRC, a.nFaceRC, lo, hi, f - 1, NB);
    bR = cell_beta(a.beta, a.betaRC, a.nFaceRC, lo, hi, f, NB);
  }
  if (a.lowCrit)
    face_states_w_rt(a.lim, bL, bR,
                     1.0 - a.lowCrit[crit_index<NB>(a.iBlock, dir, fi, fj, fk)],
                     q[0], q[1], q[2], q[3], Lv, Rv);
  else
    face_states_rt(a.lim, bL, bR, q[0], q[1], q[2], q[3], Lv, Rv);
  // the TVD treatment (get_face_tvd, src/ModFaceValue.f90:680-852): at a
  // side with a coarser neighbour the face and the fine-side state of the
  // face one in come from tvd_reschange of the two coarse ghost layers and
  // the first two fine layers of the subface's 2x2 group
  const int f = dir == 0 ? fi : dir == 1 ? fj : fk;
  if (a.tvd && (f <= 1 || f >= NB - 1)) {
    const int hi = f >= NB - 1;
    const int n = a.nei[6 * a.iBlock + 2 * dir + hi];
    if (n <= -2 && a.slabKind[-2 - n] == 0) {
      const int t1 = dir == 0 ? fj : fi, t2 = dir == 2 ? fj : fk;
      auto cell = [&](int nrm, int u1, int u2, double *P) {
        const int ii = dir == 0 ? nrm : u1;
        const int jj = dir == 1 ? nrm : dir == 0 ? u1 : u2;
        const int kk = dir == 2 ? nrm : u2;
        load_state(a.Usnap, a.vs,
                   cell_off<NB>(a.nei, a.iBlock, ii, jj, kk, a.nCellTot), P);
        to_primitive(P);
      };
      double C2[NVAR], C1[NVAR], F1[4][NVAR], F2[NVAR];
      cell(hi ? NB + 1 : -2, t1, t2, C2);
      cell(hi ? NB : -1, t1, t2, C1);
      for (int g = 0; g < 4; g++)
        cell(hi ? NB - 1 : 0, (t1 & ~1) + g % 2, (t2 & ~1) + g / 2, F1[g]);
      cell(hi ? NB - 2 : 1, t1, t2, F2);
      double CtoF[NVAR], FtoC[NVAR], FF[NVAR];
      tvd_subface(a.tvdBeta, C2, C1, F1, F2, (t1 % 2) + 2 * (t2 % 2), CtoF,
                  FtoC, FF);
      if (f == 0) {
        for (int v = 0; v < NVAR; v++) { Lv[v] = CtoF[v]; Rv[v] = FtoC[v]; }
      } else if (f == NB) {
        for (int v = 0; v < NVAR; v++) { Lv[v] = FtoC[v]; Rv[v] = CtoF[v]; }
      } else if (f == 1) {
        for (int v = 0; v < NVAR; v++) Lv[v] = FF[v];
      } else {
        for (int v = 0; v < NVAR; v++) Rv[v] = FF[v];
      }
    }
  }
  // the inner body: the fast update's face treatment, as the fused stage
  // applies it (src/ModUpdateStateFast.f90:924-937,1277-1327)
  if (const unsigned char *m = block_body_mask<NB>(a)) {
    bool u[4];
    for (int s = 0; s < 4; s++)
      u[s] = block_true<NB>(m, fi - 2 * (dir == 0) + s * (dir == 0),
                            fj - 2 * (dir == 1) + s * (dir == 1),
                            fk - 2 * (dir == 2) + s * (dir == 2));
    double r[3] = {0.0, 0.0, 0.0};
    if (u[1] != u[2]) {
      // the face centre: the mean of its two cells' centres
      const int c[3] = {fi - (dir == 0), fj - (dir == 1), fk - (dir == 2)};
      for (int v = 0; v < 3; v++) {
        const double h = a.cellsize[3 * a.iBlock + v];
        const double x0 = a.xyz0[3 * a.iBlock + v];
        r[v] = 0.5 * ((x0 + c[v] * h) + (x0 + (c[v] + (v == dir)) * h));
      }
    }
    // the limiter only decides whether the face values fall back to
    // first order next to body cells
    if (a.lim == LIM_FIRST)
      body_face<LIM_FIRST>(u[0], u[1], u[2], u[3], q[1], q[2], Lv, Rv, r,
                           a.bodyRho, a.b1rCoef);
    else
      body_face<LIM_MC>(u[0], u[1], u[2], u[3], q[1], q[2], Lv, Rv, r,
                        a.bodyRho, a.b1rCoef);
  }
  for (int v = 0; v < NVAR; v++) {
    L[fo * NVAR + v] = Lv[v];
    R[fo * NVAR + v] = Rv[v];
  }
}

template <int DIR>
__device__ int block_flux_dir(const BlockArgs &a, double area, const double *L,
                              const double *R, double *F, double &vdt,
                              double &unA, const double *B0) {
  if (B0) switch (a.iFlux) {
  case FLUX_RUSANOV:
    return numerical_flux<DIR, FLUX_RUSANOV, true, true>(a.ph, area, L, R, F, vdt, unA, B0);
  case FLUX_LINDE:
    return numerical_flux<DIR, FLUX_LINDE, true, true>(a.ph, area, L, R, F, vdt, unA, B0);
  default:
    return numerical_flux<DIR, FLUX_HLLD, true, true>(a.ph, area, L, R, F, vdt, unA, B0);
  }
  switch (a.iFlux) {
  case FLUX_RUSANOV:
    return numerical_flux<DIR, FLUX_RUSANOV, true>(a.ph, area, L, R, F, vdt, unA);
  case FLUX_LINDE:
    return numerical_flux<DIR, FLUX_LINDE, true>(a.ph, area, L, R, F, vdt, unA);
  default:
    return numerical_flux<DIR, FLUX_HLLD, true>(a.ph, area, L, R, F, vdt, unA);
  }
}

template <int NB>
__global__ void k_block_face_flux(BlockArgs a) {
  const int nFx = (NB + 1) * NB * NB;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 3 * nFx) return;
  const int dir = t / nFx, r = t % nFx;
  const int fi = dir == 0 ? r % (NB + 1) : r % NB;
  const int fj = dir == 0 ? (r / (NB + 1)) % NB
               : dir == 1 ? (r / NB) % (NB + 1) : (r / NB) % NB;
  const int fk = dir == 0 ? r / ((NB + 1) * NB)
               : dir == 1 ? r / (NB * (NB + 1)) : r / (NB * NB);
  if (a.resOnly && !coarser_side_face(a, dir, fi, fj, fk, NB)) return;
  const double *cs = a.cellsize + 3 * a.iBlock;
  double L[NVAR], R[NVAR], F[NFLUX], vdt, unA;
  const double *Ls = dir == 0 ? a.LX : (dir == 1 ? a.LY : a.LZ);
  const double *Rs = dir == 0 ? a.RX : (dir == 1 ? a.RY : a.RZ);
  double *Fo = dir == 0 ? a.FX : (dir == 1 ? a.FY : a.FZ);
  for (int v = 0; v < NVAR; v++) { L[v] = Ls[r * NVAR + v]; R[v] = Rs[r * NVAR + v]; }
  // the face's B0 (B0 splitting); r is the face's index within the block
  double B0f[3];
  const double *pB0 = nullptr;
  if (a.b0F) {
    const long nFaceTot = (long)a.nBlock * nFx;
    for (int c = 0; c < 3; c++)
      B0f[c] = a.b0F[(dir * 3 + c) * nFaceTot + (long)a.iBlock * nFx + r];
    pB0 = B0f;
  }
  int err;
  if (dir == 0) err = block_flux_dir<0>(a, cs[1] * cs[2], L, R, F, vdt, unA, pB0);
  else if (dir == 1) err = block_flux_dir<1>(a, cs[0] * cs[2], L, R, F, vdt, unA, pB0);
  else err = block_flux_dir<2>(a, cs[0] * cs[1], L, R, F, vdt, unA, pB0);
  double *o = Fo + (long)r * 14;
  for (int v = 0; v < NFLUX; v++) o[v] = F[v];
  o[9] = unA; o[10] = 0.0; o[11] = vdt; o[12] = 0.0; o[13] = 0.0;
  // a face between two body cells carries nothing
  // (src/ModUpdateStateFast.f90:237-240)
  if (const unsigned char *m = block_body_mask<NB>(a)) {
    if (!block_true<NB>(m, fi - (dir == 0), fj - (dir == 1), fk - (dir == 2)) &&
        !block_true<NB>(m, fi, fj, fk)) {
      for (int v = 0; v < 14; v++) o[v] = 0.0;
      err = 0;
    }
  }
  if (err) atomicOr(a.err, err);
}

// index of the face (p, q) of block side `side` in the per-block face
// arrays (LeftState_VX(nVar, nI+1, nJ, nK) etc.); the in-side order (p, q)
// is (j, k) / (i, k) / (i, j) on x / y / z sides
__device__ inline long side_face(int side, int p, int q, int NB) {
  const int d = side / 2, f = (side % 2) ? NB : 0;
  if (d == 0) return ((long)q * NB + p) * (NB + 1) + f;
  if (d == 1) return ((long)q * (NB + 1) + f) * NB + p;
  return ((long)f * NB + q) * NB + p;
}

// save_cons_flux (src/ModConserveFlux.f90:79-187) for one fine block: the
// fluxes of its faces on coarser sides (slots 0..6, Vdt, energy) and the
// normal field of their left and right states times FaceRatio = 1/4
// (:136-140) go to the coarse side's record of the face's sub-face, the
// store that the fused path's fine-flux kernel writes too
template <int NB>
__global__ void k_block_save_cons_flux(BlockArgs a) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 6 * NB * NB) return;
  const int side = t / (NB * NB), p = t % NB, q = (t / NB) % NB;
  const int n = a.nei[6 * a.iBlock + side];
  if (n > -2 || a.slabKind[-2 - n] != 0) return;
  const int *sl = a.slabs + 8 * (-2 - n);
  const int d = side / 2;
  const long f = side_face(side, p, q, NB);
  const double *F = (d == 0 ? a.FX : d == 1 ? a.FY : a.FZ) + f * 14;
  const double *L = (d == 0 ? a.LX : d == 1 ? a.LY : a.LZ) + f * NVAR;
  const double *R = (d == 0 ? a.RX : d == 1 ? a.RY : a.RZ) + f * NVAR;
  const int pc = sl[4] + p / 2, qc = sl[5] + q / 2, sub = (p % 2) + 2 * (q % 2);
  double *dst = a.corr + (((long)sl[7] * NB + qc) * NB + pc) * 4 * NCORR +
                sub * NCORR;
  for (int v = 0; v < 7; v++) dst[v] = F[v];
  dst[PP] = F[11];   // Vdt
  dst[EN] = F[EN];
  dst[CORR_BNL] = L[BX + d] * 0.25;
  dst[CORR_BNR] = R[BX + d] * 0.25;
}

// apply_cons_flux (src/ModConserveFlux.f90:260-441) for one coarse block:
// the faces of its sides with finer neighbours take the sums of the four
// fine sub-face fluxes and Vdt, and their left and right states the sums
// of the fine normal field (:350-384)
template <int NB>
__global__ void k_block_apply_cons_flux(BlockArgs a) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 6 * NB * NB) return;
  const int side = t / (NB * NB), p = t % NB, q = (t / NB) % NB;
  const int n = a.nei[6 * a.iBlock + side];
  if (n > -2 || a.slabKind[-2 - n] != 1) return;
  const int d = side / 2;
  const long f = side_face(side, p, q, NB);
  double *F = (d == 0 ? a.FX : d == 1 ? a.FY : a.FZ) + f * 14;
  double *L = (d == 0 ? a.LX : d == 1 ? a.LY : a.LZ) + f * NVAR;
  double *R = (d == 0 ? a.RX : d == 1 ? a.RY : a.RZ) + f * NVAR;
  const double *c = a.corr + (((long)(-2 - n) * NB + q) * NB + p) * 4 * NCORR;
  auto sum4 = [&](int v) {
    return ((c[v] + c[NCORR + v]) + c[2 * NCORR + v]) + c[3 * NCORR + v];
  };
  for (int v = 0; v < 7; v++) F[v] = sum4(v);
  F[EN] = sum4(EN);
  F[11] = sum4(PP);
  L[BX + d] = sum4(CORR_BNL);
  R[BX + d] = sum4(CORR_BNR);
}

// calc_source for one block: the 8-wave div B source from the block's face
// states (src/ModCalcSource.f90:1694-1889,930-951,1016-1022)
template <int NB>
__global__ void k_block_source(BlockArgs a) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= NB * NB * NB) return;
  const int i = c % NB, j = (c / NB) % NB, k = c / (NB * NB);
  double *Src = a.Src + (long)c * NFLUX;
  if (!a.useSrc) {
    for (int v = 0; v < NFLUX; v++) Src[v] = 0.0;
    return;
  }
  const long off = (((long)a.iBlock * NB + k) * NB + j) * NB + i;
  double St[NVAR];
  load_state(a.State, a.vs, off, St);
  const double *cs = a.cellsize + 3 * a.iBlock;
  const double InvHalf[3] = {0.5 / cs[0], 0.5 / cs[1], 0.5 / cs[2]};
  const long fx = ((long)(k * NB + j) * (NB + 1) + i) * NVAR;
  const long fy = ((long)(k * (NB + 1) + j) * NB + i) * NVAR;
  const long fy1 = fy + (long)NB * NVAR;
  const long fz = ((long)(k * NB + j) * NB + i) * NVAR;
  const long fz1 = fz + (long)NB * NB * NVAR;
  CellB0<double> cb;
  if (a.b0F) {
    const long nFaceTot = (long)a.nBlock * (NB + 1) * NB * NB;
    auto face = [&](int dir, int fi, int fj, int fk, double *o) {
      const long f = b0_face_index<NB>(a.iBlock, dir, fi, fj, fk);
      for (int q = 0; q < 3; q++) o[q] = a.b0F[(dir * 3 + q) * nFaceTot + f];
    };
    face(0, i, j, k, cb.face[0]);
    face(0, i + 1, j, k, cb.face[1]);
    face(1, i, j, k, cb.face[2]);
    face(1, i, j + 1, k, cb.face[3]);
    face(2, i, j, k, cb.face[4]);
    face(2, i, j, k + 1, cb.face[5]);
    for (int q = 0; q < 3; q++) cb.cell[q] = a.b0C[q * a.nCellTot + off];
    cb.useSource = a.useB0Src;
  }
  powell_source(InvHalf, a.LX[fx + BX], a.RX[fx + BX], a.LX[fx + NVAR + BX],
                a.RX[fx + NVAR + BX], a.LY[fy + BY], a.RY[fy + BY],
                a.LY[fy1 + BY], a.RY[fy1 + BY], a.LZ[fz + BZ], a.RZ[fz + BZ],
                a.LZ[fz1 + BZ], a.RZ[fz1 + BZ], St, Src,
                a.b0F ? &cb : nullptr);
}

template <int NB>
__global__ void k_block_update(BlockArgs a) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= NB * NB * NB) return;
  const int i = c % NB, j = (c / NB) % NB, k = c / (NB * NB);
  const long off = (((long)a.iBlock * NB + k) * NB + j) * NB + i;
  double St[NVAR], So[NVAR], Un[NVAR];
  load_state(a.State, a.vs, off, St);
  if (a.iStage == 1) {  // StateOld = State (src/ModUpdateState.f90:161-163)
    for (int v = 0; v < NVAR; v++) a.StateOld[v * a.vs + off] = St[v];
  }
  load_state(a.StateOld, a.vs, off, So);
  const double *cs = a.cellsize + 3 * a.iBlock;
  UpdateCtl u;
  // steady state: the cell's own step, set by calc_timestep at stage 1
  u.DtLocal = a.DtFactor * (a.steady ? a.dtmaxCB[off] : a.scal[S_DTCB]);
  u.InvVolume = 1.0 / (cs[0] * cs[1] * cs[2]);
  u.halfStep = a.halfStep;
  u.iStage = a.iStage;
  u.nStage = a.nStage;
  u.dbTrick = a.dbTrick;
  u.dtZero = !a.steady && (a.scal[S_DT] == 0.0);
  u.roundTripOld = 0;  // StateOld is round-tripped in memory below
  const long fx0 = ((long)(k * NB + j) * (NB + 1) + i) * 14;
  const long fy0 = ((long)(k * (NB + 1) + j) * NB + i) * 14;
  const long fz0 = ((long)(k * NB + j) * NB + i) * 14;
  double dF[NFLUX];
  for (int v = 0; v < NFLUX; v++) {
    double d = a.FX[fx0 + v] - a.FX[fx0 + 14 + v];
    d = d + a.FY[fy0 + v] - a.FY[fy0 + (long)NB * 14 + v];
    d = d + a.FZ[fz0 + v] - a.FZ[fz0 + (long)NB * NB * 14 + v];
    dF[v] = d;
  }
  update_cell(a.ph, u, dF, St, So, Un,
              a.useSrc ? a.Src + (long)c * NFLUX : nullptr);
  // body cells are not updated (src/ModUpdateStateFast.f90:343-345)
  if (block_true<NB>(block_body_mask<NB>(a), i, j, k))
    for (int v = 0; v < NVAR; v++) a.State[v * a.vs + off] = Un[v];
  if (!u.dtZero)  // p -> E -> p of StateOld (src/ModUpdateState.f90:606,965)
    a.StateOld[PP * a.vs + off] =
        energy_to_pressure(a.ph, So, pressure_to_energy(a.ph, So));
}

template <int NB>
__global__ void __launch_bounds__(512) k_block_timestep(BlockArgs a) {
  __shared__ double sRed[16];
  const double *cs = a.cellsize + 3 * a.iBlock;
  const double V = cs[0] * cs[1] * cs[2];
  double m = 1.7976931348623157e308;
  for (int c = threadIdx.x; c < NB * NB * NB; c += blockDim.x) {
    const int i = c % NB, j = (c / NB) % NB, k = c / (NB * NB);
    const long fx0 = ((long)(k * NB + j) * (NB + 1) + i) * 14;
    const long fy0 = ((long)(k * (NB + 1) + j) * NB + i) * 14;
    const long fz0 = ((long)(k * NB + j) * NB + i) * 14;
    const double Vdt = fmax2(a.FX[fx0 + 11], a.FX[fx0 + 14 + 11]) +
                       fmax2(a.FY[fy0 + 11], a.FY[fy0 + (long)NB * 14 + 11]) +
                       fmax2(a.FZ[fz0 + 11], a.FZ[fz0 + (long)NB * NB * 14 + 11]);
    const double dtc = V / Vdt;
    const long off = (((long)a.iBlock * NB + k) * NB + j) * NB + i;
    // body cells hold a zero step outside the block minimum
    // (src/ModTimeStepControl.f90:158-159,255)
    if (!block_true<NB>(block_body_mask<NB>(a), i, j, k)) {
      a.dtmaxCB[off] = 0.0;
      continue;
    }
    m = fmin2(m, dtc);
    a.dtmaxCB[off] = a.useDtFixed ? a.DtFixed : dtc;
  }
  m = wave_min(m);
  if ((threadIdx.x & 63) == 0) sRed[threadIdx.x / 64] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = sRed[0];
    for (int w = 1; w < (int)(blockDim.x / 64); w++) r = fmin2(r, sRed[w]);
    a.dtmaxB[a.iBlock] = r;
  }
}

// set_global_timestep (src/ModTimeStepControl.f90:306-524), one workgroup
// per run; run r owns blocks runFirst[r] .. runFirst[r+1]-1 and the scalars
// scal[r*S_NSCAL ...]
__global__ void __launch_bounds__(1024)
k_global_dt(const double *dtmaxB, const int *runFirst, double *scalAll,
            double tLimit, double Cfl, int useDtFixed, double DtFixed,
            int incStep) {
  __shared__ double sRed[16];
  const int r = blockIdx.x;
  double *scal = scalAll + (long)r * S_NSCAL;
  double m = 1.7976931348623157e308;  // huge()
  for (int b = runFirst[r] + threadIdx.x; b < runFirst[r + 1]; b += blockDim.x)
    m = fmin2(m, dtmaxB[b]);
  m = wave_min(m);
  if ((threadIdx.x & 63) == 0) sRed[threadIdx.x / 64] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double DtMin = sRed[0];
    for (int w = 1; w < (int)(blockDim.x / 64); w++) DtMin = fmin2(DtMin, sRed[w]);
    const double nStep = scal[S_NSTEP] + (incStep ? 1.0 : 0.0);
    scal[S_NSTEP] = nStep;
    double Dt;
    if (useDtFixed)
      Dt = DtFixed;
    else if (nStep < 1.0 || (nStep == 1.0 && tLimit > 0.0))
      Dt = 0.0;
    else
      Dt = DtMin;
    if (tLimit > 0.0 && scal[S_TSIM] + Cfl * Dt > tLimit)
      Dt = (tLimit - scal[S_TSIM]) / Cfl;
    scal[S_DTMIN] = DtMin;
    scal[S_DTCB] = Dt;
    scal[S_DT] = Cfl * Dt;
  }
}

// fill n doubles with v (a kernel rather than a memset, so that a captured
// step graph holds no memset node)
__global__ void __launch_bounds__(256) k_fill(double *p, long n, double v) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = v;
}

template <typename T>
__global__ void k_check_nan(const T *U, long n, int *err) {
  int bad = 0;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n;
       t += (long)gridDim.x * blockDim.x)
    if (!isfinite(U[t])) bad = 1;
  if (bad) atomicOr(err, ERRB_NAN);
}

// per-block sums of U*V (rho, rho*u, B, total energy) in a fixed order
template <int NB, typename T>
__global__ void __launch_bounds__(256)
k_integrate(const T *U, long vs, const double *cellsize, Phys ph,
            int nBlock, double *partial) {
  __shared__ double sRed[256];
  const int b = blockIdx.x;
  if (b >= nBlock) return;  // one workgroup per block
  const double V = cellsize[3 * b] * cellsize[3 * b + 1] * cellsize[3 * b + 2];
  double acc[NVAR];
  for (int v = 0; v < NVAR; v++) acc[v] = 0.0;
  for (int c = threadIdx.x; c < NB * NB * NB; c += 256) {
    T qt[NVAR];
    double q[NVAR];
    load_state(U, vs, (long)b * NB * NB * NB + c, qt);
    for (int v = 0; v < NVAR; v++) q[v] = double(qt[v]);
    q[PP] = pressure_to_energy(ph, q);
    for (int v = 0; v < NVAR; v++) acc[v] += q[v] * V;
  }
  for (int v = 0; v < NVAR; v++) {
    sRed[threadIdx.x] = acc[v];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) sRed[threadIdx.x] += sRed[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) partial[(long)b * NVAR + v] = sRed[0];
    __syncthreads();
  }
}

} // namespace

// =====================================================================
// host side
// =====================================================================
extern "C" const char *bats_strerror(int code) {
  switch (code) {
  case BATS_OK: return "ok";
  case BATS_EINVAL: return "invalid argument or configuration";
  case BATS_ENEGFAST2: return "negative fast speed squared";
  case BATS_EHLLD: return "negative HLLD density ratio";
  case BATS_ENAN: return "NaN or Inf in the state";
  case BATS_EBC: return "unknown boundary condition";
  case BATS_EHIP: return "HIP runtime error";
  case BATS_ENOMEM: return "device allocation failed";
  case BATS_ENODEV: return "no gfx950 device available";
  case BATS_ELIMIT: return "request beyond a fixed buffer limit";
  default: return "unknown status";
  }
}

extern "C" int bats_init(const bats_config *cfg, bats_handle **out) {
  if (!cfg || !out) return BATS_EINVAL;
  *out = nullptr;
  const bats_config &c = *cfg;
  if (c.nI != c.nJ || c.nI != c.nK || (c.nI != 8 && c.nI != 16)) return BATS_EINVAL;
  if (c.nDim != 3 || c.nG != 2 || c.nBlock <= 0) return BATS_EINVAL;
  if (c.nOrder != 1 && c.nOrder != 2) return BATS_EINVAL;
  if (c.nOrder == 2 && (c.iLimiter < BATS_LIM_MINMOD || c.iLimiter > BATS_LIM_NO))
    return BATS_EINVAL;
  if (c.iFlux < BATS_FLUX_RUSANOV || c.iFlux > BATS_FLUX_HLLD) return BATS_EINVAL;
  // one or two stages (half step or SSP-RK2), or SSP-RK3, which the
  // reference runs without the half step scheme
  // (src/ModSetParameters.f90:628-633)
  if (c.nStage < 1 || c.nStage > 3) return BATS_EINVAL;
  if (c.nStage == 3 && c.UseHalfStep) return BATS_EINVAL;
  if (!(c.Gamma > 1.0) || !(c.Cfl > 0.0)) return BATS_EINVAL;
  if (c.UseDivbSource != 0 && c.UseDivbSource != 1) return BATS_EINVAL;
  if (c.StatePrecision != 0 && c.StatePrecision != 4 && c.StatePrecision != 8)
    return BATS_EINVAL;
  // #ADAPTIVELOWORDER checks (src/ModFaceValue.f90:222-225)
  if (c.UseAdaptiveLowOrder != 0 && c.UseAdaptiveLowOrder != 1)
    return BATS_EINVAL;
  if (c.UseAdaptiveLowOrder &&
      ((c.nLowOrder != 1 && c.nLowOrder != 2) ||
       c.pCritLow < c.pCritHigh - 1e-15 || !(c.VelCrit >= 0.0)))
    return BATS_EINVAL;
  if (c.UseOutflowPressure != 0 && c.UseOutflowPressure != 1)
    return BATS_EINVAL;
  if (c.IsSteadyState != 0 && c.IsSteadyState != 1) return BATS_EINVAL;
  if (c.UseOutflowPressure && !std::isfinite(c.pOutflow)) return BATS_EINVAL;
  // #RESOLUTIONCHANGE limiter parameters
  if (c.nFaceLimiterResChange < 0 ||
      (c.nFaceLimiterResChange > 0 && !std::isfinite(c.BetaLimiterResChange)))
    return BATS_EINVAL;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= c.iDevice || c.iDevice < 0)
    return BATS_ENODEV;
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, c.iDevice));
  if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
    fprintf(stderr, "bats_hip: device %d is %s, this library is built for gfx950\n",
            c.iDevice, prop.gcnArchName);
    return BATS_ENODEV;
  }
  HIPCHK(hipSetDevice(c.iDevice));
  bats_handle *h = new bats_handle();
  h->cfg = c;
  // a fixed time step runs with Cfl = 1 (src/ModSetParameters.f90:4314-4321)
  if (c.UseDtFixed) h->cfg.Cfl = 1.0;
  h->NB = c.nI;
  h->nCellBlk = (long)c.nI * c.nJ * c.nK;
  h->nCell = h->nCellBlk * c.nBlock;
  h->vs = h->nCell;
  h->cur = 0;
  h->stateBuf = 0;
  h->iStageBlock = 1;
  h->nTimed = 0;
  h->lastStageMs = 0.0;
  h->lastStageLaunches = 0;
  hipError_t e = hipStreamCreate(&h->stream);
  h->elemSize = (c.StatePrecision == 4) ? 4 : 8;
  const size_t sz = (size_t)h->elemSize * NVAR * h->nCell;
  h->nBuf = c.nStage == 3 ? 3 : 2;
  for (int u = 0; u < h->nBuf; u++)
    if (e == hipSuccess) e = hipMalloc(&h->U[u], sz);
  if (e == hipSuccess) e = hipMalloc(&h->nei, sizeof(int) * 6 * c.nBlock);
  if (e == hipSuccess) e = hipMalloc(&h->cellsize, sizeof(double) * 3 * c.nBlock);
  if (e == hipSuccess) e = hipMalloc(&h->dtmaxB, sizeof(double) * c.nBlock);
  if (e == hipSuccess) e = hipMalloc(&h->scal, sizeof(double) * S_NSCAL);
  if (e == hipSuccess) e = hipMalloc(&h->runFirst, sizeof(int) * 2);
  if (e == hipSuccess) e = hipMalloc(&h->runOfBlock, sizeof(int) * c.nBlock);
  if (e == hipSuccess) e = hipMalloc(&h->err, sizeof(int));
  if (e == hipSuccess) e = hipMalloc(&h->partial, sizeof(double) * NVAR * c.nBlock);
  h->lowCrit = nullptr;
  const size_t critSz =
      (size_t)h->elemSize * 3 * (c.nI + 1) * c.nI * c.nI * c.nBlock;
  if (e == hipSuccess && c.UseAdaptiveLowOrder) e = hipMalloc(&h->lowCrit, critSz);
  if (e == hipSuccess && h->lowCrit) e = hipMemset(h->lowCrit, 0, critSz);
  if (e != hipSuccess) {
    fprintf(stderr, "bats_hip: allocation failed: %s\n", hipGetErrorString(e));
    bats_finalize(h);
    return e == hipErrorOutOfMemory ? BATS_ENOMEM : BATS_EHIP;
  }
  h->Usnap = nullptr;
  h->dtmaxCB = nullptr;
  if (c.IsSteadyState) {  // the per-cell steps carry over between stages
    HIPCHK(hipMalloc(&h->dtmaxCB, sizeof(double) * h->nCell));
    HIPCHK(hipMemset(h->dtmaxCB, 0, sizeof(double) * h->nCell));
  }
  h->LX = h->RX = h->LY = h->RY = h->LZ = h->RZ = nullptr;
  h->FX = h->FY = h->FZ = nullptr;
  double s0[S_NSCAL] = {0};
  HIPCHK(hipMemcpy(h->scal, s0, sizeof s0, hipMemcpyHostToDevice));
  h->nRun = 1;
  h->hRunFirst = {0, c.nBlock};
  h->hRunOfBlock.assign(c.nBlock, 0);
  h->nWorkgroupMax = 0;
  HIPCHK(hipMemcpy(h->runFirst, h->hRunFirst.data(), sizeof(int) * 2,
                   hipMemcpyHostToDevice));
  HIPCHK(hipMemset(h->runOfBlock, 0, sizeof(int) * c.nBlock));
  HIPCHK(hipMemset(h->err, 0, sizeof(int)));
  HIPCHK(hipMemset(h->dtmaxB, 0, sizeof(double) * c.nBlock));
  // outer boundaries float until bats_set_cell_boundary
  for (int side = 0; side < 6; side++) h->bcType[side] = BATS_BC_FLOAT;
  HIPCHK(hipMalloc(&h->dBcState, sizeof h->bcState));
  HIPCHK(hipMemset(h->dBcState, 0, sizeof h->bcState));
  // default grid until bats_set_grid: every block its own domain of unit
  // cells with six outer sides; kept as the caller's grid, so that
  // bats_set_cell_boundary applies its types to it like to any other grid
  h->hNei.assign(6 * c.nBlock, -1);
  h->anySideRC = true;
  h->userNei = h->hNei;
  h->userCellSize.assign(3 * c.nBlock, 1.0);
  h->userSlabs.clear();
  HIPCHK(hipMemcpy(h->nei, h->hNei.data(), sizeof(int) * 6 * c.nBlock,
                   hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(h->cellsize, h->userCellSize.data(),
                   sizeof(double) * 3 * c.nBlock, hipMemcpyHostToDevice));
  *out = h;
  return BATS_OK;
}

extern "C" int bats_finalize(bats_handle *h) {
  if (!h) return BATS_EINVAL;
  (void)hipStreamSynchronize(h->stream);
  void *ptrs[] = {h->U[0], h->U[1], h->U[2], h->Usnap, h->nei, h->cellsize, h->dtmaxB,
                  h->dtmaxCB, h->scal, h->runFirst, h->runOfBlock, h->err,
                  h->partial, h->LX, h->RX,
                  h->LY, h->RY, h->LZ, h->RZ, h->FX, h->FY, h->FZ, h->Src,
                  h->sur.face, h->sur.rowX, h->sur.rowFace, h->sur.rowScale,
                  h->sur.counters, h->slabs, h->slabKind, h->fineSlabs,
                  h->corr, h->lowCrit, h->dBcState, h->b0F, h->b0C,
                  h->bodySlot, h->bodyUsed, h->bodyXyz0, h->groupList,
                  h->b0Side, h->b0SideSlot, h->varyT, h->varyS};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  drop_graphs(h);
  for (auto &ev : h->evStart) (void)hipEventDestroy(ev);
  for (auto &ev : h->evStop) (void)hipEventDestroy(ev);
  (void)hipStreamDestroy(h->stream);
  delete h;
  return BATS_OK;
}

static int apply_grid(bats_handle *h, const std::vector<int> &nei,
                      const std::vector<double> &cs,
                      const std::vector<int> &slabs);

// a new grid drops the B0 field (it belongs to the old block geometry)
static void drop_b0(bats_handle *h) {
  void *ptrs[] = {h->b0F, h->b0C, h->b0Side, h->b0SideSlot};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  h->b0F = h->b0C = h->b0Side = nullptr;
  h->b0SideSlot = nullptr;
  h->useB0 = 0;
}

// and the inner body (its masks belong to the old blocks)
static void drop_body(bats_handle *h) {
  void *ptrs[] = {h->bodySlot, h->bodyUsed, h->bodyXyz0, h->groupList};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  h->bodySlot = nullptr;
  h->bodyUsed = nullptr;
  h->bodyXyz0 = nullptr;
  h->groupList = nullptr;
  h->useBody = 0;
  h->nPlainGroup = h->nBodyGroup = 0;
}

extern "C" int bats_set_grid(bats_handle *h, const int *iNeighbor_EB,
                             const double *CellSize_DB) {
  if (!h || !iNeighbor_EB || !CellSize_DB) return BATS_EINVAL;
  drop_graphs(h);
  (void)hipStreamSynchronize(h->stream);
  drop_b0(h);
  drop_body(h);
  const int nB = h->cfg.nBlock;
  for (int t = 0; t < 6 * nB; t++)
    if (iNeighbor_EB[t] < -1 || iNeighbor_EB[t] >= nB) return BATS_EINVAL;
  for (int t = 0; t < 3 * nB; t++)
    if (!(CellSize_DB[t] > 0.0)) return BATS_EINVAL;
  for (int t = 0; t < 6 * nB; t++)  // runs stay independent domains
    if (iNeighbor_EB[t] >= 0 &&
        h->hRunOfBlock[iNeighbor_EB[t]] != h->hRunOfBlock[t / 6])
      return BATS_EINVAL;
  return apply_grid(h, std::vector<int>(iNeighbor_EB, iNeighbor_EB + 6 * nB),
                    std::vector<double>(CellSize_DB, CellSize_DB + 3 * nB),
                    std::vector<int>());
}

// ---- layout conversions ---------------------------------------------
static void aos_to_soa(const bats_config &c, const double *S, double *U) {
  const int NB = c.nI, g = c.nG, NG = NB + 2 * g;
  const long nB = c.nBlock, vs = nB * NB * NB * NB;
  for (long b = 0; b < nB; b++)
    for (int k = 0; k < NB; k++)
      for (int j = 0; j < NB; j++)
        for (int i = 0; i < NB; i++) {
          const double *src =
              S + ((((b * NG + (k + g)) * NG + (j + g)) * NG) + (i + g)) * NVAR;
          const long off = ((b * NB + k) * NB + j) * NB + i;
          for (int v = 0; v < NVAR; v++) U[v * vs + off] = src[v];
        }
}

static void soa_to_aos(const bats_config &c, const std::vector<int> &nei,
                       const std::vector<int> &slabs, const double *bcState,
                       const double *U, double *S) {
  const int NB = c.nI, g = c.nG, NG = NB + 2 * g;
  const long nB = c.nBlock, vs = nB * NB * NB * NB;
  auto at = [&](long b, int i, int j, int k) {
    return S + ((((b * NG + (k + g)) * NG + (j + g)) * NG) + (i + g)) * NVAR;
  };
  for (long b = 0; b < nB; b++)
    for (int k = 0; k < NB; k++)
      for (int j = 0; j < NB; j++)
        for (int i = 0; i < NB; i++) {
          const long off = ((b * NB + k) * NB + j) * NB + i;
          double *dst = at(b, i, j, k);
          for (int v = 0; v < NVAR; v++) dst[v] = U[v * vs + off];
        }
  // face ghost slabs, as exchange_messages + float or fixed BCs would fill
  // them (resolution-change sides as float)
  for (long b = 0; b < nB; b++)
    for (int side = 0; side < 6; side++) {
      const int n = nei[6 * b + side], dim = side / 2, hi = side % 2;
      const int kindS = n <= -2 ? slabs[8 * (-2 - n) + 2] : -1;
      const bool fixedSide = kindS == 2, reflectSide = kindS == 3;
      const bool outflowP = kindS == 4;
      for (int layer = 1; layer <= g; layer++) {
        const int dst = hi ? NB - 1 + layer : -layer;
        const int src = n >= 0 ? (hi ? layer - 1 : NB - layer)
                        : reflectSide ? (hi ? NB - layer : layer - 1)
                                      : (hi ? NB - 1 : 0);
        const long sb = n >= 0 ? n : b;
        for (int p = 0; p < NB; p++)
          for (int q = 0; q < NB; q++) {
            int di, dj, dk, si, sj, sk;
            if (dim == 0) { di = dst; si = src; dj = sj = q; dk = sk = p; }
            else if (dim == 1) { dj = dst; sj = src; di = si = q; dk = sk = p; }
            else { dk = dst; sk = src; di = si = q; dj = sj = p; }
            const long off = ((sb * NB + sk) * NB + sj) * NB + si;
            double *d = at(b, di, dj, dk);
            for (int v = 0; v < NVAR; v++) {
              const double sg =
                  reflectSide && (v == MX + dim || v == BX + dim) ? -1.0 : 1.0;
              d[v] = fixedSide ? bcState[side * NVAR + v]
                               : sg * U[v * vs + off];
            }
            if (outflowP) d[PP] = c.pOutflow;
          }
      }
    }
}

// host SoA (double) <-> device state buffer, converting to the state type
static int put_soa(bats_handle *h, const double *U) {
  HIPCHK(hipStreamSynchronize(h->stream));
  const size_t n = (size_t)h->nCell, es = h->elemSize;
  char *dst = static_cast<char *>(h->U[h->cur]);
  std::vector<float> f(h->elemSize == 4 ? n : 0);
  for (int v = 0; v < NVAR; v++) {
    const double *src = U + v * n;
    const void *p = src;
    if (es == 4) {
      for (size_t t = 0; t < n; t++) f[t] = (float)src[t];
      p = f.data();
    }
    HIPCHK(hipMemcpy(dst + (size_t)v * h->vs * es, p, es * n,
                     hipMemcpyHostToDevice));
  }
  h->stateBuf = h->cur;
  return BATS_OK;
}
static int get_soa(bats_handle *h, double *U) {
  HIPCHK(hipStreamSynchronize(h->stream));
  const size_t n = (size_t)h->nCell, es = h->elemSize;
  const char *src = static_cast<const char *>(h->U[h->stateBuf]);
  std::vector<float> f(h->elemSize == 4 ? n : 0);
  for (int v = 0; v < NVAR; v++) {
    double *dst = U + v * n;
    if (es == 8) {
      HIPCHK(hipMemcpy(dst, src + (size_t)v * h->vs * es, es * n,
                       hipMemcpyDeviceToHost));
    } else {
      HIPCHK(hipMemcpy(f.data(), src + (size_t)v * h->vs * es, es * n,
                       hipMemcpyDeviceToHost));
      for (size_t t = 0; t < n; t++) dst[t] = f[t];
    }
  }
  return BATS_OK;
}

extern "C" int bats_upload_state(bats_handle *h, const double *State_VGB) {
  if (!h || !State_VGB) return BATS_EINVAL;
  std::vector<double> tmp((size_t)NVAR * h->nCell);
  aos_to_soa(h->cfg, State_VGB, tmp.data());
  return put_soa(h, tmp.data());
}
extern "C" int bats_download_state(bats_handle *h, double *State_VGB) {
  if (!h || !State_VGB) return BATS_EINVAL;
  std::vector<double> tmp((size_t)NVAR * h->nCell);
  int st = get_soa(h, tmp.data());
  if (st) return st;
  soa_to_aos(h->cfg, h->hNei, h->hSlabs, h->bcState, tmp.data(), State_VGB);
  return BATS_OK;
}
extern "C" int bats_upload_state_soa(bats_handle *h, const double *U) {
  if (!h || !U) return BATS_EINVAL;
  return put_soa(h, U);
}
extern "C" int bats_download_state_soa(bats_handle *h, double *U) {
  if (!h || !U) return BATS_EINVAL;
  return get_soa(h, U);
}

// ---- per-block routines ---------------------------------------------
static int ensure_block_buffers(bats_handle *h) {
  // the exchange snapshot follows the state layout (a grid change frees
  // it); the face arrays depend on the block size only
  if (!h->Usnap) {
    HIPCHK(hipMalloc(&h->Usnap, (size_t)h->elemSize * NVAR * h->vs));
    // ghost slabs of the current state (reflecting sides mirror it)
    if (h->nSlab > 0) {
      const int st = amr_fill_ghosts(h, h->U[h->stateBuf]);
      if (st) return st;
    }
    HIPCHK(hipMemcpyAsync(h->Usnap, h->U[h->stateBuf],
                          (size_t)h->elemSize * NVAR * h->vs,
                          hipMemcpyDeviceToDevice, h->stream));
  }
  if (h->Src) return BATS_OK;
  const int NB = h->NB;
  const size_t nF = (size_t)(NB + 1) * NB * NB;
  double **fs[] = {&h->LX, &h->RX, &h->LY, &h->RY, &h->LZ, &h->RZ};
  for (double **p : fs) {
    HIPCHK(hipMalloc(p, sizeof(double) * NVAR * nF));
    HIPCHK(hipMemsetAsync(*p, 0, sizeof(double) * NVAR * nF, h->stream));
  }
  double **fl[] = {&h->FX, &h->FY, &h->FZ};
  for (double **p : fl) {
    HIPCHK(hipMalloc(p, sizeof(double) * 14 * nF));
    HIPCHK(hipMemsetAsync(*p, 0, sizeof(double) * 14 * nF, h->stream));
  }
  HIPCHK(hipMalloc(&h->Src, sizeof(double) * NFLUX * NB * NB * NB));
  h->srcBlock = -1;
  if (!h->dtmaxCB) {
    HIPCHK(hipMalloc(&h->dtmaxCB, sizeof(double) * h->nCell));
    HIPCHK(hipMemsetAsync(h->dtmaxCB, 0, sizeof(double) * h->nCell, h->stream));
  }
  return BATS_OK;
}

// the adaptive low order weights are set (second order scheme) and change
// the face values (nLowOrder = 1), src/ModFaceValue.f90:2726-2735,2017-2019
static bool low_order_set(const bats_config &c) {
  return c.UseAdaptiveLowOrder && c.nOrder == 2;
}
static bool low_order_used(const bats_config &c) {
  return low_order_set(c) && c.nLowOrder == 1;
}
// the sharper BetaLimiterResChange applies next to sides with DiLevel_EB
// /= 0 (src/ModFaceValue.f90:1982-2015): the second order scheme with a
// limiter that reads beta, LimiterBeta above it, and a grid with such sides
static bool res_beta_used(const bats_handle *h) {
  const bats_config &c = h->cfg;
  return c.nOrder == 2 && c.nFaceLimiterResChange > 0 &&
         c.LimiterBeta > c.BetaLimiterResChange &&
         (c.iLimiter == BATS_LIM_MC || c.iLimiter == BATS_LIM_MC3 ||
          c.iLimiter == BATS_LIM_BETA) &&
         h->anySideRC;
}

static int launch_low_order_crit(bats_handle *h, const void *U) {
  LowOrderArgs a;
  a.U = U;
  a.vs = h->vs;
  a.nCellTot = h->nCell;
  a.nei = h->nei;
  a.nBlock = h->cfg.nBlock;
  a.Gamma = h->cfg.Gamma;
  a.pCritLow = h->cfg.pCritLow;
  a.pCritHigh = h->cfg.pCritHigh;
  a.VelCrit = h->cfg.VelCrit;
  a.crit = h->lowCrit;
  const int NB = h->NB;
  const long n = 3L * (NB + 1) * NB * NB * h->cfg.nBlock;
  const dim3 grid((unsigned)((n + 255) / 256)), block(256);
  if (NB == 8) {
    if (h->elemSize == 4) hipLaunchKernelGGL((k_low_order_crit<8, float>), grid, block, 0, h->stream, a);
    else hipLaunchKernelGGL((k_low_order_crit<8, double>), grid, block, 0, h->stream, a);
  } else {
    if (h->elemSize == 4) hipLaunchKernelGGL((k_low_order_crit<16, float>), grid, block, 0, h->stream, a);
    else hipLaunchKernelGGL((k_low_order_crit<16, double>), grid, block, 0, h->stream, a);
  }
  return hip_status(hipGetLastError());
}

static BlockArgs block_args(bats_handle *h, int iBlock) {
  BlockArgs a;
  const bats_config &c = h->cfg;
  a.Usnap = h->Usnap;
  a.State = static_cast<double *>(h->U[h->stateBuf]);
  a.StateOld = static_cast<double *>(h->U[h->stateBuf == 0 ? 1 : 0]);
  a.nei = h->nei;
  a.cellsize = h->cellsize;
  a.scal = h->scal + (long)h->hRunOfBlock[iBlock] * S_NSCAL;
  a.LX = h->LX; a.RX = h->RX; a.LY = h->LY; a.RY = h->RY; a.LZ = h->LZ; a.RZ = h->RZ;
  a.FX = h->FX; a.FY = h->FY; a.FZ = h->FZ;
  a.Src = h->Src;
  a.useSrc = c.UseDivbSource;
  a.slabs = h->slabs;
  a.slabKind = h->slabKind;
  a.corr = static_cast<double *>(h->corr);
  a.resOnly = 0;
  a.b0F = h->useB0 ? h->b0F : nullptr;
  a.b0C = h->b0C;
  a.useB0Src = h->useB0Src;
  a.lowCrit = low_order_used(c) ? static_cast<const double *>(h->lowCrit)
                                : nullptr;
  a.betaRC = c.BetaLimiterResChange;
  a.nFaceRC = res_beta_used(h) ? c.nFaceLimiterResChange : 0;
  a.tvd = c.UseTvdResChange && h->nResSlab > 0 && c.nOrder > 1;
  a.tvdBeta = std::min(c.BetaLimiterResChange, c.LimiterBeta);
  a.steady = c.IsSteadyState;
  a.used = h->useBody ? h->bodyUsed : nullptr;
  a.bodySlot = h->bodySlot;
  a.xyz0 = h->bodyXyz0;
  a.bodyRho = h->bodyRho;
  a.b1rCoef = h->b1rCoef;
  a.dtmaxCB = h->dtmaxCB;
  a.dtmaxB = h->dtmaxB;
  a.err = h->err;
  a.nBlock = c.nBlock;
  a.iBlock = iBlock;
  a.vs = h->vs;
  a.nCellTot = h->nCell;
  a.ph = make_phys(c);
  a.beta = c.LimiterBeta;
  a.lim = limiter_code(c);
  a.iFlux = c.iFlux;
  const int halfStep = c.UseHalfStep || c.nStage == 1;
  a.DtFactor = c.UseHalfStep ? (c.Cfl * h->iStageBlock) / c.nStage : c.Cfl;
  a.iStage = h->iStageBlock;
  a.nStage = c.nStage;
  a.halfStep = halfStep;
  a.dbTrick = db_trick_at(c, h->iStageBlock);
  a.useDtFixed = c.UseDtFixed;
  a.DtFixed = c.DtFixed;
  return a;
}

extern "C" int bats_exchange_messages(bats_handle *h) {
  if (!h) return BATS_EINVAL;
  int st = ensure_block_buffers(h);
  if (st) return st;
  // outer boundaries and resolution changes of the current state
  // (exchange_messages applies the cell boundaries, src/ModMessagePass.f90:
  // 120-140), then the snapshot the per-block routines read
  if (h->nSlab > 0) {
    st = amr_fill_ghosts(h, h->U[h->stateBuf]);
    if (st) return st;
  }
  HIPCHK(hipMemcpyAsync(h->Usnap, h->U[h->stateBuf],
                        (size_t)h->elemSize * NVAR * h->vs,
                        hipMemcpyDeviceToDevice, h->stream));
  return BATS_OK;
}

extern "C" int bats_set_stage(bats_handle *h, int iStage) {
  if (!h || iStage < 1 || iStage > h->cfg.nStage) return BATS_EINVAL;
  h->iStageBlock = iStage;
  return BATS_OK;
}

static int check_block(bats_handle *h, int iBlock1) {
  if (!h || iBlock1 < 1 || iBlock1 > h->cfg.nBlock) return BATS_EINVAL;
  // the reference-shaped per-block routines run on FP64 state only; the
  // TVD resolution-change faces not with low order faces
  if (h->elemSize != 8) return BATS_EINVAL;
  const bats_config &c = h->cfg;
  if (c.UseTvdResChange && h->nResSlab > 0 && low_order_set(c))
    return BATS_EINVAL;
  return ensure_block_buffers(h);
}

extern "C" int bats_calc_face_value(bats_handle *h, int iBlock,
                                    int DoResChangeOnly, int DoMonotoneRestrict) {
  // the ghost slabs of resolution changes are filled by exchange_messages
  // (mean of the fine cells, limited prolongation of the coarse ones)
  (void)DoMonotoneRestrict;
  int st = check_block(h, iBlock);
  if (st) return st;
  // DoResChangeOnly: only the faces of sides with a coarser neighbour
  if (DoResChangeOnly && h->nResSlab == 0) return BATS_OK;
  BlockArgs a = block_args(h, iBlock - 1);
  a.resOnly = DoResChangeOnly != 0;
  const int n = 3 * (h->NB + 1) * h->NB * h->NB;
  if (h->NB == 8)
    hipLaunchKernelGGL(k_block_face_value<8>, dim3((n + 255) / 256), dim3(256), 0, h->stream, a);
  else
    hipLaunchKernelGGL(k_block_face_value<16>, dim3((n + 255) / 256), dim3(256), 0, h->stream, a);
  return hip_status(hipGetLastError());
}

extern "C" int bats_set_low_order_face(bats_handle *h) {
  if (!h || !h->cfg.UseAdaptiveLowOrder || !h->lowCrit) return BATS_EINVAL;
  int st = check_block(h, 1);
  if (st) return st;
  // first order scheme: every face is low order, no weights (:2727-2728)
  if (!low_order_set(h->cfg)) return BATS_OK;
  return launch_low_order_crit(h, h->Usnap);
}

extern "C" int bats_get_low_order_crit(bats_handle *h, int iBlock,
                                       double *CritX, double *CritY,
                                       double *CritZ) {
  if (!h || !h->lowCrit || !CritX || !CritY || !CritZ || iBlock < 1 ||
      iBlock > h->cfg.nBlock)
    return BATS_EINVAL;
  const int NB = h->NB;
  const long n = 3L * (NB + 1) * NB * NB;
  std::vector<double> w(n);
  HIPCHK(hipStreamSynchronize(h->stream));
  if (h->elemSize == 8) {
    HIPCHK(hipMemcpy(w.data(), static_cast<double *>(h->lowCrit) + n * (iBlock - 1),
                     sizeof(double) * n, hipMemcpyDeviceToHost));
  } else {
    std::vector<float> f(n);
    HIPCHK(hipMemcpy(f.data(), static_cast<float *>(h->lowCrit) + n * (iBlock - 1),
                     sizeof(float) * n, hipMemcpyDeviceToHost));
    for (long t = 0; t < n; t++) w[t] = f[t];
  }
  auto at = [&](int dir, int i, int j, int k) {
    return NB == 8 ? w[crit_index<8>(0, dir, i, j, k)]
                   : w[crit_index<16>(0, dir, i, j, k)];
  };
  // X(i,j,k) i = 0..NB; Y(i,j,k) j = 0..NB; Z(i,j,k) k = 0..NB
  for (int a = 0; a < NB; a++)
    for (int b = 0; b < NB; b++)
      for (int f = 0; f <= NB; f++) {
        CritX[((long)a * NB + b) * (NB + 1) + f] = at(0, f, b, a);
        CritY[((long)a * (NB + 1) + f) * NB + b] = at(1, b, f, a);
        CritZ[((long)f * NB + a) * NB + b] = at(2, b, a, f);
      }
  return BATS_OK;
}

extern "C" int bats_calc_face_flux(bats_handle *h, int DoResChangeOnly, int iBlock) {
  int st = check_block(h, iBlock);
  if (st) return st;
  if (DoResChangeOnly && h->nResSlab == 0) return BATS_OK;
  BlockArgs a = block_args(h, iBlock - 1);
  a.resOnly = DoResChangeOnly != 0;
  const int n = 3 * (h->NB + 1) * h->NB * h->NB;
  if (h->NB == 8)
    hipLaunchKernelGGL(k_block_face_flux<8>, dim3((n + 127) / 128), dim3(128), 0, h->stream, a);
  else
    hipLaunchKernelGGL(k_block_face_flux<16>, dim3((n + 127) / 128), dim3(128), 0, h->stream, a);
  return hip_status(hipGetLastError());
}

extern "C" int bats_save_cons_flux(bats_handle *h, int iBlock) {
  int st = check_block(h, iBlock);
  if (st) return st;
  if (h->nResSlab == 0) return BATS_OK;
  BlockArgs a = block_args(h, iBlock - 1);
  const int n = 6 * h->NB * h->NB;
  if (h->NB == 8)
    hipLaunchKernelGGL(k_block_save_cons_flux<8>, dim3((n + 255) / 256), dim3(256), 0, h->stream, a);
  else
    hipLaunchKernelGGL(k_block_save_cons_flux<16>, dim3((n + 255) / 256), dim3(256), 0, h->stream, a);
  return hip_status(hipGetLastError());
}

extern "C" int bats_message_pass_face(bats_handle *h) {
  // save_cons_flux already writes every fine sub-face record into the
  // coarse side's store, in stream order before any apply_cons_flux, and
  // apply_cons_flux adds the four sub-faces: nothing is left to move
  if (!h) return BATS_EINVAL;
  return BATS_OK;
}

extern "C" int bats_apply_cons_flux(bats_handle *h, int iBlock) {
  int st = check_block(h, iBlock);
  if (st) return st;
  if (h->nResSlab == 0) return BATS_OK;
  BlockArgs a = block_args(h, iBlock - 1);
  const int n = 6 * h->NB * h->NB;
  if (h->NB == 8)
    hipLaunchKernelGGL(k_block_apply_cons_flux<8>, dim3((n + 255) / 256), dim3(256), 0, h->stream, a);
  else
    hipLaunchKernelGGL(k_block_apply_cons_flux<16>, dim3((n + 255) / 256), dim3(256), 0, h->stream, a);
  return hip_status(hipGetLastError());
}

extern "C" int bats_calc_source(bats_handle *h, int iBlock) {
  int st = check_block(h, iBlock);
  if (st) return st;
  BlockArgs a = block_args(h, iBlock - 1);
  const int n = h->NB * h->NB * h->NB;
  if (h->NB == 8)
    hipLaunchKernelGGL(k_block_source<8>, dim3((n + 255) / 256), dim3(256), 0, h->stream, a);
  else
    hipLaunchKernelGGL(k_block_source<16>, dim3((n + 255) / 256), dim3(256), 0, h->stream, a);
  h->srcBlock = iBlock - 1;
  return hip_status(hipGetLastError());
}

extern "C" int bats_update_state(bats_handle *h, int iBlock) {
  int st = check_block(h, iBlock);
  if (st) return st;
  if (h->cfg.UseDivbSource && h->srcBlock != iBlock - 1) return BATS_EINVAL;
  h->srcBlock = -1;
  BlockArgs a = block_args(h, iBlock - 1);
  const int n = h->NB * h->NB * h->NB;
  if (h->NB == 8)
    hipLaunchKernelGGL(k_block_update<8>, dim3((n + 255) / 256), dim3(256), 0, h->stream, a);
  else
    hipLaunchKernelGGL(k_block_update<16>, dim3((n + 255) / 256), dim3(256), 0, h->stream, a);
  return hip_status(hipGetLastError());
}

extern "C" int bats_calc_timestep(bats_handle *h, int iBlock, int IsPartLocal) {
  if (IsPartLocal) return BATS_EINVAL;  // local time stepping is not supported
  int st = check_block(h, iBlock);
  if (st) return st;
  BlockArgs a = block_args(h, iBlock - 1);
  if (h->NB == 8)
    hipLaunchKernelGGL(k_block_timestep<8>, dim3(1), dim3(512), 0, h->stream, a);
  else
    hipLaunchKernelGGL(k_block_timestep<16>, dim3(1), dim3(512), 0, h->stream, a);
  return hip_status(hipGetLastError());
}

extern "C" int bats_get_face_values(bats_handle *h, double *LeftX, double *RightX,
                                    double *LeftY, double *RightY, double *LeftZ,
                                    double *RightZ) {
  if (!h || !h->LX) return BATS_EINVAL;
  const size_t n = sizeof(double) * NVAR * (h->NB + 1) * h->NB * h->NB;
  HIPCHK(hipStreamSynchronize(h->stream));
  double *dst[] = {LeftX, RightX, LeftY, RightY, LeftZ, RightZ};
  double *src[] = {h->LX, h->RX, h->LY, h->RY, h->LZ, h->RZ};
  for (int t = 0; t < 6; t++)
    if (dst[t]) HIPCHK(hipMemcpy(dst[t], src[t], n, hipMemcpyDeviceToHost));
  return BATS_OK;
}

extern "C" int bats_get_face_fluxes(bats_handle *h, double *FluxX, double *FluxY,
                                    double *FluxZ) {
  if (!h || !h->FX) return BATS_EINVAL;
  const size_t n = sizeof(double) * 14 * (h->NB + 1) * h->NB * h->NB;
  HIPCHK(hipStreamSynchronize(h->stream));
  double *dst[] = {FluxX, FluxY, FluxZ};
  double *src[] = {h->FX, h->FY, h->FZ};
  for (int t = 0; t < 3; t++)
    if (dst[t]) HIPCHK(hipMemcpy(dst[t], src[t], n, hipMemcpyDeviceToHost));
  int bits = 0;
  HIPCHK(hipMemcpy(&bits, h->err, sizeof(int), hipMemcpyDeviceToHost));
  return err_from_bits(bits);
}

extern "C" int bats_get_dtmax(bats_handle *h, double *DtMax_CB, double *DtMax_B) {
  if (!h) return BATS_EINVAL;
  HIPCHK(hipStreamSynchronize(h->stream));
  if (DtMax_CB) {
    if (!h->dtmaxCB) return BATS_EINVAL;
    // device order (b, k, j, i) equals DtMax_CB(nI,nJ,nK,nBlock)
    HIPCHK(hipMemcpy(DtMax_CB, h->dtmaxCB, sizeof(double) * h->nCell,
                     hipMemcpyDeviceToHost));
  }
  if (DtMax_B)
    HIPCHK(hipMemcpy(DtMax_B, h->dtmaxB, sizeof(double) * h->cfg.nBlock,
                     hipMemcpyDeviceToHost));
  return BATS_OK;
}

// ---- production path ------------------------------------------------
extern "C" int bats_set_global_timestep(bats_handle *h, double TimeSimulationLimit,
                                        double *DtOut) {
  if (!h) return BATS_EINVAL;
  const bats_config &c = h->cfg;
  hipLaunchKernelGGL(k_global_dt, dim3(h->nRun), dim3(1024), 0, h->stream,
                     h->dtmaxB, h->runFirst, h->scal, TimeSimulationLimit,
                     c.Cfl, c.UseDtFixed, c.DtFixed, 1);
  HIPCHK(hipGetLastError());
  if (DtOut) {
    double s[S_NSCAL];
    HIPCHK(hipMemcpyAsync(s, h->scal, sizeof s, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    *DtOut = s[S_DT];
  }
  return BATS_OK;
}

// calcDt: DoCalcTimestep of the stage loop.  The local time steps are set
// after the last stage's update (time accurate, :217-219), before the first
// stage's update (steady state, :184-195), and always by the no-update pass.
static void tvd_args(const bats_handle *h, AmrArgs &m);

static int launch_stage(bats_handle *h, int iStage, int calcDt, int noUpdate,
                        int advanceTime, bool timed) {
  const bats_config &c = h->cfg;
  StageArgs a{};
  const int last = (iStage == c.nStage);
  const bool steady = c.IsSteadyState != 0;
  const int doDt = noUpdate ? 1 : steady ? (calcDt && iStage == 1)
                                         : (calcDt && last);
  // stage buffers: the step's start state U[cur] stays the old state of
  // every stage; stage s reads stage_in(s) and writes stage_in(s+1); the
  // last stage of a multi-stage step writes U[cur] in place (each cell
  // reads its own old value before writing it, no halo reads U[cur])
  const int o1 = (h->cur + 1) % h->nBuf, o2 = (h->cur + 2) % h->nBuf;
  const int inBuf = iStage == 1 ? h->cur : iStage == 2 ? o1 : o2;
  const int outBuf = c.nStage == 1 ? o1
                   : last ? h->cur
                   : iStage == 1 ? o1 : o2;
  a.Uin = h->U[inBuf]; a.Uold = h->U[h->cur]; a.Uout = h->U[outBuf];
  if (noUpdate) {  // face-speed pass on the current state only
    a.Uin = h->U[h->stateBuf]; a.Uold = a.Uin; a.Uout = nullptr;
  }
  a.nei = h->nei;
  a.cellsize = h->cellsize;
  a.scal = h->scal;
  a.runOfBlock = h->nRun > 1 ? h->runOfBlock : nullptr;
  a.nRun = h->nRun;
  a.dtmaxB = h->dtmaxB;
  a.dtmaxCB = h->dtmaxCB;
  a.err = h->err;
  a.nBlock = c.nBlock;
  a.vs = h->vs;
  a.nCellTot = h->nCell;
  a.slabKind = h->slabKind;
  a.corr = nullptr;
  a.ph = make_phys(c);
  a.beta = c.LimiterBeta;
  a.lim = limiter_code(c);
  a.DtFactor = c.UseHalfStep ? (c.Cfl * iStage) / c.nStage : c.Cfl;
  a.iStage = iStage;
  a.nStage = c.nStage;
  a.halfStep = c.UseHalfStep || c.nStage == 1;
  a.dbTrick = db_trick_at(c, iStage);
  a.doDt = doDt;
  a.roundTripOld = 1;
  a.noUpdate = noUpdate;
  a.advanceTime = advanceTime && !steady;
  const int NB = h->NB;
  a.nGroup = (c.nBlock + kThreads / (NB * NB) - 1) / (kThreads / (NB * NB));
  a.splitShift = h->splitShift;
  a.xcdOrder = h->xcdOrder;
  a.nWorkgroup = h->nWorkgroupMax > 0 ? h->nWorkgroupMax
                                      : a.nGroup << a.splitShift;
  int ev = -1;
  if (timed) {
    ev = h->nTimed++;
    while ((int)h->evStart.size() <= ev) {
      hipEvent_t e0, e1;
      HIPCHK(hipEventCreate(&e0));
      HIPCHK(hipEventCreate(&e1));
      h->evStart.push_back(e0);
      h->evStop.push_back(e1);
    }
    HIPCHK(hipEventRecord(h->evStart[ev], h->stream));
  }
  const bool lowUse = low_order_used(c);
  const bool resBeta = res_beta_used(h);
  // the body variant carries no low order faces or resolution-change beta
  if (h->useBody && (lowUse || resBeta)) return BATS_EINVAL;
  // the TVD treatment at resolution changes: second-order faces of the
  // fine sides from the records k_amr_fineflux leaves; not with low order
  // faces or the surrogate
  const bool tvd = c.UseTvdResChange && h->nResSlab > 0;
  if (tvd && (low_order_set(c) || h->sur.mode != BATS_SUR_OFF))
    return BATS_EINVAL;
  // the surrogate's face pass has no low order weights, one beta and no
  // 8-wave source
  if ((lowUse || resBeta || steady || c.UseDivbSource) &&
      h->sur.mode != BATS_SUR_OFF)
    return BATS_EINVAL;
  a.lowCrit = lowUse ? h->lowCrit : nullptr;
  a.betaRC = c.BetaLimiterResChange;
  a.nFaceRC = resBeta ? c.nFaceLimiterResChange : 0;
  const int opt = (c.UseDivbSource ? OPT_POWELL : 0) |
                  (lowUse ? OPT_LOWORDER : 0) | (resBeta ? OPT_RESBETA : 0) |
                  (steady ? OPT_STEADY : 0) | (h->useB0 ? OPT_B0 : 0) |
                  (tvd && c.nOrder > 1 ? OPT_TVD : 0);
  a.b0F = h->b0F;
  a.b0C = h->b0C;
  a.useB0Src = h->useB0Src;
  a.steady = steady;
  const bool f32 = h->elemSize == 4;
  hipError_t e = hipSuccess;
  // resolution changes: ghost slabs of the stage input
  if (h->nSlab > 0) {
    int st = amr_fill_ghosts(h, a.Uin);
    if (st) return st;
  }
  // low order weights at the first stage of the step (and for the initial
  // time step pass, which stands for a first step with Dt = 0),
  // src/ModAdvanceExplicit.f90:87-91
  if (low_order_set(c) && (iStage == 1 || noUpdate)) {
    int st = launch_low_order_crit(h, a.Uin);
    if (st) return st;
  }
  if (h->nSlab > 0) {
    // (plain path) the fine face fluxes that replace the coarse ones
    if (h->sur.mode == BATS_SUR_OFF && h->nFine > 0) {
      AmrArgs m{};
      m.U = a.Uin; m.vs = h->vs; m.nCellTot = h->nCell; m.nei = h->nei;
      m.cellsize = h->cellsize; m.slabs = h->slabs; m.nSlab = h->nSlab;
      m.fineSlabs = h->fineSlabs; m.nFine = h->nFine; m.ph = a.ph;
      m.beta = a.beta; m.lim = a.lim; m.corr = h->corr; m.err = h->err;
      m.lowCrit = a.lowCrit;
      m.betaRC = a.betaRC;
      m.nFaceRC = a.nFaceRC;
      m.b0F = h->useB0 ? h->b0F : nullptr;
      m.nBlock = c.nBlock;
      if (h->useBody) {
        m.bodySlot = h->bodySlot;
        m.used = h->bodyUsed;
        m.xyz0 = h->bodyXyz0;
        m.bodyRho = h->bodyRho;
        m.b1rCoef = h->b1rCoef;
      }
      tvd_args(h, m);
#define BATS_FPICK(NBv, FL)                                                  \
  (f32 ? launch_fineflux_##NBv##_##FL##_f32(m, h->stream)                    \
       : launch_fineflux_##NBv##_##FL##_f64(m, h->stream))
      if (NB == 8)
        e = c.iFlux == BATS_FLUX_RUSANOV ? BATS_FPICK(8, rusanov)
          : c.iFlux == BATS_FLUX_LINDE   ? BATS_FPICK(8, linde)
                                         : BATS_FPICK(8, hlld);
      else
        e = c.iFlux == BATS_FLUX_RUSANOV ? BATS_FPICK(16, rusanov)
          : c.iFlux == BATS_FLUX_LINDE   ? BATS_FPICK(16, linde)
                                         : BATS_FPICK(16, hlld);
#undef BATS_FPICK
      HIPCHK(e);
      a.corr = h->corr;
    }
  }
  if (h->sur.mode != BATS_SUR_OFF) {
    // gated face fluxes with the network on smooth faces
    const int st = sur_stage(h, a);
    if (st) return st;
  } else {
    // the few groups with a body block would leave most of the GPU idle in
    // their own launch: their blocks are cut into z-parts of at least two
    // planes until about one workgroup per CU runs
    int bodyShift = a.splitShift;
    if (h->useBody)
      while ((h->nBodyGroup << bodyShift) < 256 && (NB >> (bodyShift + 1)) >= 2)
        bodyShift++;
    // the parts of a split block meet in an atomic minimum of dtmaxB
    if (a.doDt && (a.splitShift || bodyShift)) {
      const long n = c.nBlock;
      hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256),
                         0, h->stream, h->dtmaxB, n, 1.7976931348623157e308);
    }
    auto launch = [&](const StageArgs &la, int lopt) {
#define BATS_PICK(NBv, FL)                                                   \
  (f32 ? launch_stage_##NBv##_##FL##_f32(la, lopt, h->stream)                \
       : launch_stage_##NBv##_##FL##_f64(la, lopt, h->stream))
      if (NB == 8)
        return c.iFlux == BATS_FLUX_RUSANOV ? BATS_PICK(8, rusanov)
             : c.iFlux == BATS_FLUX_LINDE   ? BATS_PICK(8, linde)
                                            : BATS_PICK(8, hlld);
      return c.iFlux == BATS_FLUX_RUSANOV ? BATS_PICK(16, rusanov)
           : c.iFlux == BATS_FLUX_LINDE   ? BATS_PICK(16, linde)
                                          : BATS_PICK(16, hlld);
#undef BATS_PICK
    };
    if (h->useBody) {
      // the groups without a body block run the plain variant, then those
      // with one the body variant (which also advances the time)
      const int wg = h->nWorkgroupMax;
      StageArgs ap = a;
      ap.groupList = h->groupList;
      ap.nGroup = h->nPlainGroup;
      ap.nWorkgroup = wg > 0 ? wg : ap.nGroup << ap.splitShift;
      ap.advanceTime = 0;
      if (ap.nGroup > 0) e = launch(ap, opt);
      if (e == hipSuccess) {
        StageArgs ab = a;
        ab.groupList = h->groupList + h->nPlainGroup;
        ab.nGroup = h->nBodyGroup;
        ab.splitShift = bodyShift;
        ab.nWorkgroup = wg > 0 ? wg : ab.nGroup << ab.splitShift;
        ab.bodySlot = h->bodySlot;
        ab.used = h->bodyUsed;
        ab.xyz0 = h->bodyXyz0;
        ab.bodyRho = h->bodyRho;
        ab.b1rCoef = h->b1rCoef;
        e = launch(ab, opt | OPT_BODY);
      }
    } else {
      e = launch(a, opt);
    }
  }
  HIPCHK(e);
  if (timed) HIPCHK(hipEventRecord(h->evStop[ev], h->stream));
  if (!noUpdate) {
    if (c.nStage == 1) h->cur = outBuf;
    h->stateBuf = outBuf;
  }
  return BATS_OK;
}

extern "C" int bats_advance_stage(bats_handle *h, int iStage, int DoCalcTimestep) {
  if (!h || iStage < 1 || iStage > h->cfg.nStage) return BATS_EINVAL;
  return launch_stage(h, iStage, DoCalcTimestep, 0, 1, false);
}

extern "C" int bats_advance_explicit(bats_handle *h, int DoCalcTimestep) {
  if (!h) return BATS_EINVAL;
  for (int s = 1; s <= h->cfg.nStage; s++) {
    int st = launch_stage(h, s, DoCalcTimestep, 0, 1, false);
    if (st) return st;
  }
  return BATS_OK;
}

// capture one time step (global time step, every stage) on the handle's
// stream into the graph of the current buffer parity; the host-side buffer
// bookkeeping the stage launcher did while capturing is undone, since
// nothing ran yet
static int capture_step(bats_handle *h, double TimeSimulationLimit) {
  const int par = h->cur, cur0 = h->cur, buf0 = h->stateBuf;
  if (h->graphValid[par]) {
    (void)hipGraphExecDestroy(h->graphExec[par]);
    h->graphValid[par] = 0;
  }
  hipGraph_t g = nullptr;
  HIPCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
  int st = bats_set_global_timestep(h, TimeSimulationLimit, nullptr);
  for (int s = 1; !st && s <= h->cfg.nStage; s++)
    st = launch_stage(h, s, 1, 0, 1, false);
  const hipError_t e = hipStreamEndCapture(h->stream, &g);
  h->cur = cur0;
  h->stateBuf = buf0;
  if (st || e != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return st ? st : BATS_EHIP;
  }
  const hipError_t ei = hipGraphInstantiate(&h->graphExec[par], g, nullptr,
                                            nullptr, 0);
  (void)hipGraphDestroy(g);
  HIPCHK(ei);
  h->graphValid[par] = 1;
  h->graphTLimit[par] = TimeSimulationLimit;
  return BATS_OK;
}

extern "C" int bats_set_step_graph(bats_handle *h, int UseGraph) {
  if (!h || (UseGraph != 0 && UseGraph != 1)) return BATS_EINVAL;
  drop_graphs(h);
  h->useGraph = UseGraph;
  return BATS_OK;
}

extern "C" int bats_advance(bats_handle *h, int nStep, double TimeSimulationLimit) {
  if (!h || nStep < 0) return BATS_EINVAL;
  h->nTimed = 0;
  if (h->useGraph && nStep > 0) {
    // one graph launch per step; the timing is per step (all kernels)
    // divided by the stages.  The network's weights are synchronized with
    // its own stream at every eager launch, which a replay cannot do.
    if (h->sur.mode == BATS_SUR_NETWORK) return BATS_EINVAL;
    if (h->evStart.empty()) {
      hipEvent_t e0, e1;
      HIPCHK(hipEventCreate(&e0));
      HIPCHK(hipEventCreate(&e1));
      h->evStart.push_back(e0);
      h->evStop.push_back(e1);
    }
    HIPCHK(hipEventRecord(h->evStart[0], h->stream));
    for (int n = 0; n < nStep; n++) {
      const int par = h->cur;
      if (!h->graphValid[par] || h->graphTLimit[par] != TimeSimulationLimit) {
        const int st = capture_step(h, TimeSimulationLimit);
        if (st) return st;
      }
      HIPCHK(hipGraphLaunch(h->graphExec[par], h->stream));
      // the buffer bookkeeping of the step's stages
      if (h->cfg.nStage == 1) h->cur = 1 - h->cur;
      h->stateBuf = h->cur;
    }
    HIPCHK(hipEventRecord(h->evStop[0], h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, h->evStart[0], h->evStop[0]));
    h->lastStageLaunches = nStep * h->cfg.nStage;
    h->lastStageMs = ms / h->lastStageLaunches;
    int bits = 0;
    HIPCHK(hipMemcpy(&bits, h->err, sizeof(int), hipMemcpyDeviceToHost));
    return err_from_bits(bits);
  }
  for (int n = 0; n < nStep; n++) {
    int st = bats_set_global_timestep(h, TimeSimulationLimit, nullptr);
    if (st) return st;
    for (int s = 1; s <= h->cfg.nStage; s++) {
      st = launch_stage(h, s, 1, 0, 1, true);
      if (st) return st;
    }
  }
  HIPCHK(hipStreamSynchronize(h->stream));
  double tot = 0.0;
  for (int e = 0; e < h->nTimed; e++) {
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, h->evStart[e], h->evStop[e]));
    tot += ms;
  }
  h->lastStageLaunches = h->nTimed;
  h->lastStageMs = h->nTimed ? tot / h->nTimed : 0.0;
  int bits = 0;
  HIPCHK(hipMemcpy(&bits, h->err, sizeof(int), hipMemcpyDeviceToHost));
  return err_from_bits(bits);
}

extern "C" int bats_init_timestep(bats_handle *h) {
  if (!h) return BATS_EINVAL;
  return launch_stage(h, h->cfg.nStage, 1, 1, 0, false);
}

extern "C" int bats_get_time(bats_handle *h, double *Dt, double *tSimulation,
                             int *nStep) {
  if (!h) return BATS_EINVAL;
  double s[S_NSCAL];
  HIPCHK(hipMemcpyAsync(s, h->scal, sizeof s, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  if (Dt) *Dt = s[S_DT];
  if (tSimulation) *tSimulation = s[S_TSIM];
  if (nStep) *nStep = (int)s[S_NSTEP];
  return BATS_OK;
}

extern "C" int bats_set_runs(bats_handle *h, int nRun, const int *iBlockFirst_R) {
  if (!h || nRun < 1 || !iBlockFirst_R) return BATS_EINVAL;
  drop_graphs(h);
  const int nB = h->cfg.nBlock;
  if (iBlockFirst_R[0] != 0 || iBlockFirst_R[nRun] != nB) return BATS_EINVAL;
  for (int r = 0; r < nRun; r++)
    if (iBlockFirst_R[r + 1] <= iBlockFirst_R[r]) return BATS_EINVAL;
  std::vector<int> runOf(nB);
  for (int r = 0; r < nRun; r++)
    for (int b = iBlockFirst_R[r]; b < iBlockFirst_R[r + 1]; b++) runOf[b] = r;
  // runs are independent domains: no neighbour across a run boundary
  for (int b = 0; b < nB; b++)
    for (int s = 0; s < 6; s++) {
      const int n = h->hNei[6 * b + s];
      if (n >= 0 && runOf[n] != runOf[b]) return BATS_EINVAL;
    }
  HIPCHK(hipStreamSynchronize(h->stream));
  double *scal = nullptr;
  int *first = nullptr;
  HIPCHK(hipMalloc(&scal, sizeof(double) * S_NSCAL * nRun));
  HIPCHK(hipMalloc(&first, sizeof(int) * (nRun + 1)));
  HIPCHK(hipMemset(scal, 0, sizeof(double) * S_NSCAL * nRun));
  HIPCHK(hipMemcpy(first, iBlockFirst_R, sizeof(int) * (nRun + 1),
                   hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(h->runOfBlock, runOf.data(), sizeof(int) * nB,
                   hipMemcpyHostToDevice));
  HIPCHK(hipFree(h->scal));
  HIPCHK(hipFree(h->runFirst));
  h->scal = scal;
  h->runFirst = first;
  h->nRun = nRun;
  h->hRunFirst.assign(iBlockFirst_R, iBlockFirst_R + nRun + 1);
  h->hRunOfBlock = runOf;
  return BATS_OK;
}

extern "C" int bats_get_run_times(bats_handle *h, double *Dt_R,
                                  double *tSimulation_R, int *nStep_R) {
  if (!h) return BATS_EINVAL;
  std::vector<double> s((size_t)S_NSCAL * h->nRun);
  HIPCHK(hipMemcpyAsync(s.data(), h->scal, sizeof(double) * s.size(),
                        hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  for (int r = 0; r < h->nRun; r++) {
    if (Dt_R) Dt_R[r] = s[(size_t)r * S_NSCAL + S_DT];
    if (tSimulation_R) tSimulation_R[r] = s[(size_t)r * S_NSCAL + S_TSIM];
    if (nStep_R) nStep_R[r] = (int)s[(size_t)r * S_NSCAL + S_NSTEP];
  }
  return BATS_OK;
}

extern "C" int bats_set_persistent_grid(bats_handle *h, int nWorkgroup) {
  if (!h || nWorkgroup < 0) return BATS_EINVAL;
  drop_graphs(h);
  h->nWorkgroupMax = nWorkgroup;
  return BATS_OK;
}

// cells per thread of the stage kernel (NB, NB/2 or NB/4 planes of each
// column per work item) and the XCD-contiguous work order (the same bits;
// a split rounds the part-boundary faces within an ulp)
extern "C" int bats_set_stage_shape(bats_handle *h, int CellsPerThread,
                                    int UseXcdOrder) {
  if (!h) return BATS_EINVAL;
  int shift = -1;
  for (int s = 0; s <= 2; s++)
    if (CellsPerThread == (h->NB >> s)) shift = s;
  if (shift < 0 || (UseXcdOrder != 0 && UseXcdOrder != 1)) return BATS_EINVAL;
  drop_graphs(h);
  h->splitShift = shift;
  h->xcdOrder = UseXcdOrder;
  return BATS_OK;
}

// the TVD treatment at resolution changes (UseTvdResChange, the
// reference's default with nOrderProlong = 1): ghost fill and fine faces
static void tvd_args(const bats_handle *h, AmrArgs &m) {
  const bats_config &c = h->cfg;
  m.tvd = c.UseTvdResChange != 0;
  m.twoLayers = c.nOrder > 1;
  m.tvdBeta = std::min(c.BetaLimiterResChange, c.LimiterBeta);
}

int amr_fill_ghosts(bats_handle *h, const void *U) {
  AmrArgs m{};
  tvd_args(h, m);
  m.b0Side = h->useB0 ? h->b0Side : nullptr;
  m.b0SideSlot = h->useB0 ? h->b0SideSlot : nullptr;
  m.varyT = h->varyT;
  m.varyS = h->varyS;
  m.nVary = h->cfg.IsSteadyState ? 0 : h->nVary;
  m.scal = h->scal;
  m.runOfBlock = h->nRun > 1 ? h->runOfBlock : nullptr;
  m.U = U; m.vs = h->vs; m.nCellTot = h->nCell; m.nei = h->nei;
  m.cellsize = h->cellsize; m.slabs = h->slabs; m.nSlab = h->nSlab;
  m.fineSlabs = h->fineSlabs; m.nFine = h->nFine; m.ph = make_phys(h->cfg);
  m.beta = h->cfg.LimiterBeta; m.lim = limiter_code(h->cfg);
  m.corr = h->corr; m.err = h->err; m.bcState = h->dBcState;
  m.pOutflow = h->cfg.pOutflow;
  const long n = 2L * h->NB * h->NB * h->nSlab;
  const dim3 g((unsigned)((n + 255) / 256)), t(256);
  if (h->elemSize == 8)
    hipLaunchKernelGGL(k_amr_ghosts<double>, g, t, 0, h->stream, m, h->NB);
  else
    hipLaunchKernelGGL(k_amr_ghosts<float>, g, t, 0, h->stream, m, h->NB);
  HIPCHK(hipGetLastError());
  return BATS_OK;
}

extern "C" int bats_set_amr_grid(bats_handle *h, const int *iNeighbor_EB,
                                 const double *CellSize_DB, int nSlab,
                                 const int *iSlab_IS) {
  if (!h || !iNeighbor_EB || !CellSize_DB || nSlab < 0 ||
      (nSlab > 0 && !iSlab_IS))
    return BATS_EINVAL;
  drop_graphs(h);
  (void)hipStreamSynchronize(h->stream);
  drop_b0(h);
  drop_body(h);
  const int nB = h->cfg.nBlock, NB = h->NB, H = NB / 2;
  for (int t = 0; t < 3 * nB; t++)
    if (!(CellSize_DB[t] > 0.0)) return BATS_EINVAL;
  for (int b = 0; b < nB; b++)
    for (int s = 0; s < 6; s++) {
      const int n = iNeighbor_EB[6 * b + s];
      if (n >= nB || (n <= -2 && -2 - n >= nSlab)) return BATS_EINVAL;
      if (n >= 0 && h->hRunOfBlock[n] != h->hRunOfBlock[b]) return BATS_EINVAL;
      if (n <= -2) {
        const int *sl = iSlab_IS + 8 * (-2 - n);
        if (sl[0] != b || sl[1] != s) return BATS_EINVAL;
      }
    }
  for (int t = 0; t < nSlab; t++) {
    const int *sl = iSlab_IS + 8 * t;
    if (sl[0] < 0 || sl[0] >= nB || sl[1] < 0 || sl[1] > 5) return BATS_EINVAL;
    if (iNeighbor_EB[6 * sl[0] + sl[1]] != -2 - t) return BATS_EINVAL;
    if (sl[2] == 0) {
      if (sl[3] < 0 || sl[3] >= nB || (sl[4] != 0 && sl[4] != H) ||
          (sl[5] != 0 && sl[5] != H) || sl[7] < 0 || sl[7] >= nSlab ||
          iSlab_IS[8 * sl[7] + 2] != 1)
        return BATS_EINVAL;
    } else if (sl[2] == 1) {
      for (int q = 0; q < 4; q++)
        if (sl[3 + q] < 0 || sl[3 + q] >= nB) return BATS_EINVAL;
    } else {
      return BATS_EINVAL;
    }
  }
  return apply_grid(
      h, std::vector<int>(iNeighbor_EB, iNeighbor_EB + 6 * nB),
      std::vector<double>(CellSize_DB, CellSize_DB + 3 * nB),
      std::vector<int>(iSlab_IS, iSlab_IS + 8 * (size_t)nSlab));
}

// Allocate the state buffers with nSlab ghost slabs after the interior
// cells and install the (already checked) neighbour and slab lists: kinds
// 0 and 1 at resolution changes, kind 2 at fixed outer sides.
static int install_slab_grid(bats_handle *h, const int *iNeighbor_EB,
                             const double *CellSize_DB, int nSlab,
                             const int *iSlab_IS, int nResSlab) {
  const int nB = h->cfg.nBlock, NB = h->NB;
  std::vector<int> kind(nSlab), fine;
  for (int t = 0; t < nSlab; t++) {
    kind[t] = iSlab_IS[8 * t + 2];
    if (kind[t] == 0) fine.push_back(t);
  }
  HIPCHK(hipStreamSynchronize(h->stream));
  // every new buffer is allocated and filled before the old ones are
  // released, so a failed allocation leaves the installed grid unchanged
  const long vs = h->nCell + 2L * NB * NB * nSlab;
  const size_t stateBytes = (size_t)h->elemSize * NVAR * vs;
  const size_t nc = (size_t)nSlab * NB * NB * 4 * NCORR;
  void *nU[3] = {nullptr, nullptr, nullptr}, *nCorr = nullptr;
  int *nSlabs = nullptr, *nKind = nullptr, *nFine = nullptr;
  auto fill = [&]() -> hipError_t {
    hipError_t e = hipSuccess;
    for (int u = 0; u < h->nBuf && e == hipSuccess; u++) {
      e = hipMalloc(&nU[u], stateBytes);
      if (e == hipSuccess) e = hipMemset(nU[u], 0, stateBytes);
    }
    if (nSlab > 0) {
      if (e == hipSuccess) e = hipMalloc(&nSlabs, sizeof(int) * 8 * nSlab);
      if (e == hipSuccess) e = hipMalloc(&nKind, sizeof(int) * nSlab);
      if (e == hipSuccess)
        e = hipMemcpy(nSlabs, iSlab_IS, sizeof(int) * 8 * nSlab,
                      hipMemcpyHostToDevice);
      if (e == hipSuccess)
        e = hipMemcpy(nKind, kind.data(), sizeof(int) * nSlab,
                      hipMemcpyHostToDevice);
      if (e == hipSuccess) e = hipMalloc(&nCorr, (size_t)h->elemSize * nc);
      if (e == hipSuccess) e = hipMemset(nCorr, 0, (size_t)h->elemSize * nc);
      if (e == hipSuccess && !fine.empty())
        e = hipMalloc(&nFine, sizeof(int) * fine.size());
      if (e == hipSuccess && !fine.empty())
        e = hipMemcpy(nFine, fine.data(), sizeof(int) * fine.size(),
                      hipMemcpyHostToDevice);
    }
    return e;
  };
  const hipError_t e = fill();
  if (e != hipSuccess) {
    void *made[] = {nU[0], nU[1], nU[2], nSlabs, nKind, nCorr, nFine};
    for (void *p : made)
      if (p) (void)hipFree(p);
    fprintf(stderr, "bats_hip: grid allocation failed: %s\n",
            hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? BATS_ENOMEM : BATS_EHIP;
  }
  void *old[] = {h->U[0], h->U[1], h->U[2], h->Usnap, h->slabs,
                 h->slabKind, h->fineSlabs, h->corr};
  for (void *p : old)
    if (p) HIPCHK(hipFree(p));
  for (int u = 0; u < 3; u++) h->U[u] = nU[u];
  h->Usnap = nullptr;
  h->slabs = nSlabs;
  h->slabKind = nKind;
  h->fineSlabs = nFine;
  h->corr = nCorr;
  h->vs = vs;
  h->nSlab = nSlab;
  h->nResSlab = nResSlab;
  h->nFine = (int)fine.size();
  h->hSlabs.assign(iSlab_IS, iSlab_IS + 8 * nSlab);
  h->hNei.assign(iNeighbor_EB, iNeighbor_EB + 6 * nB);
  HIPCHK(hipMemcpy(h->nei, iNeighbor_EB, sizeof(int) * 6 * nB,
                   hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(h->cellsize, CellSize_DB, sizeof(double) * 3 * nB,
                   hipMemcpyHostToDevice));
  return BATS_OK;
}

// Install the caller's grid (outer sides -1, resolution-change slabs) with
// the outer boundary types: every -1 side of a fixed/inflow type becomes a
// kind-2 slab whose two ghost layers hold the side's state, of a reflect
// type a kind-3 slab (the mirrored cells, refilled with every ghost fill).
// A grid with no slabs on a handle that has none keeps the state buffers.
static int apply_grid(bats_handle *h, const std::vector<int> &userNei,
                      const std::vector<double> &cs,
                      const std::vector<int> &userSlabs) {
  const int nB = h->cfg.nBlock;
  std::vector<int> nei = userNei, slabs = userSlabs;
  const int nRes = (int)(userSlabs.size() / 8);
  for (int b = 0; b < nB; b++)
    for (int s = 0; s < 6; s++) {
      if (nei[6 * b + s] != -1) continue;
      const int t = h->bcType[s];
      if (t == BATS_BC_PERIODIC) return BATS_EINVAL;
      const bool outP = t == BATS_BC_OUTFLOW && h->cfg.UseOutflowPressure;
      if (t != BATS_BC_FIXED && t != BATS_BC_INFLOW && t != BATS_BC_REFLECT &&
          t != BATS_BC_VARY && !outP)
        continue;
      nei[6 * b + s] = -2 - (int)(slabs.size() / 8);
      // kind 5: the x+ vary side, refilled from the inflow table
      const int kindS = t == BATS_BC_REFLECT ? 3 : outP ? 4
                      : (t == BATS_BC_VARY && s == 1) ? 5 : 2;
      const int sl[8] = {b, s, kindS, 0, 0, 0, 0, 0};
      slabs.insert(slabs.end(), sl, sl + 8);
    }
  const int nSlab = (int)(slabs.size() / 8);
  if (nSlab == 0 && h->nSlab == 0) {
    h->hNei = nei;
    HIPCHK(hipMemcpyAsync(h->nei, nei.data(), sizeof(int) * 6 * nB,
                          hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(h->cellsize, cs.data(), sizeof(double) * 3 * nB,
                          hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
  } else {
    int st = install_slab_grid(h, nei.data(), cs.data(), nSlab, slabs.data(),
                               nRes);
    if (st) return st;
    // ghost layers of the fixed sides in both state buffers (the stages
    // refill them with every other slab)
    for (int u = 0; u < h->nBuf && nSlab > 0; u++) {
      st = amr_fill_ghosts(h, h->U[u]);
      if (st) return st;
    }
    HIPCHK(hipStreamSynchronize(h->stream));
  }
  h->userNei = userNei;
  h->userCellSize = cs;
  h->userSlabs = userSlabs;
  h->anySideRC = false;
  for (int n : nei) h->anySideRC = h->anySideRC || n < 0;
  return BATS_OK;
}

extern "C" int bats_set_cell_boundary(bats_handle *h, const int *iTypeCellBc_I,
                                      const double *CellState_VI) {
  if (!h || !iTypeCellBc_I) return BATS_EINVAL;
  bool fixed = false;
  for (int s = 0; s < 6; s++) {
    const int t = iTypeCellBc_I[s];
    if (t != BATS_BC_PERIODIC && t != BATS_BC_FLOAT && t != BATS_BC_FIXED &&
        t != BATS_BC_INFLOW && t != BATS_BC_REFLECT && t != BATS_BC_OUTFLOW &&
        t != BATS_BC_VARY)
      return BATS_EBC;
    fixed = fixed || t == BATS_BC_FIXED || t == BATS_BC_INFLOW ||
            t == BATS_BC_VARY;
  }
  if (fixed && !CellState_VI) return BATS_EINVAL;
  if (fixed)
    for (int t = 0; t < 6 * NVAR; t++)
      if (!std::isfinite(CellState_VI[t])) return BATS_EINVAL;
  drop_graphs(h);
  int oldType[6];
  double oldState[6 * NVAR];
  std::memcpy(oldType, h->bcType, sizeof oldType);
  std::memcpy(oldState, h->bcState, sizeof oldState);
  std::memcpy(h->bcType, iTypeCellBc_I, sizeof h->bcType);
  for (int t = 0; t < 6 * NVAR; t++)
    h->bcState[t] = CellState_VI ? CellState_VI[t] : 0.0;
  HIPCHK(hipMemcpy(h->dBcState, h->bcState, sizeof h->bcState,
                   hipMemcpyHostToDevice));
  const int st = apply_grid(h, h->userNei, h->userCellSize, h->userSlabs);
  if (st) {
    std::memcpy(h->bcType, oldType, sizeof oldType);
    std::memcpy(h->bcState, oldState, sizeof oldState);
    HIPCHK(hipMemcpy(h->dBcState, h->bcState, sizeof h->bcState,
                     hipMemcpyHostToDevice));
  }
  return st;
}

// the time-varying inflow table of the x+ side (set_boundary_fast,
// src/ModUpdateStateFast.f90:1151-1158)
extern "C" int bats_set_inflow_table(bats_handle *h, int nPoint,
                                     const double *Time_I,
                                     const double *State_VI) {
  if (!h || nPoint < 0) return BATS_EINVAL;
  drop_graphs(h);
  HIPCHK(hipStreamSynchronize(h->stream));
  if (nPoint == 0) {
    if (h->varyT) (void)hipFree(h->varyT);
    if (h->varyS) (void)hipFree(h->varyS);
    h->varyT = h->varyS = nullptr;
    h->nVary = 0;
    return BATS_OK;
  }
  if (!Time_I || !State_VI) return BATS_EINVAL;
  for (int i = 0; i < nPoint; i++) {
    if (!std::isfinite(Time_I[i]) || (i > 0 && !(Time_I[i] > Time_I[i - 1])))
      return BATS_EINVAL;
    for (int v = 0; v < NVAR; v++)
      if (!std::isfinite(State_VI[(long)i * NVAR + v])) return BATS_EINVAL;
    if (!(State_VI[(long)i * NVAR + RHO] > 0.0)) return BATS_EINVAL;
  }
  double *dT = nullptr, *dS = nullptr;
  hipError_t e = hipMalloc(&dT, sizeof(double) * nPoint);
  if (e == hipSuccess) e = hipMalloc(&dS, sizeof(double) * NVAR * nPoint);
  if (e == hipSuccess)
    e = hipMemcpy(dT, Time_I, sizeof(double) * nPoint, hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(dS, State_VI, sizeof(double) * NVAR * nPoint,
                  hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    if (dT) (void)hipFree(dT);
    if (dS) (void)hipFree(dS);
    return e == hipErrorOutOfMemory ? BATS_ENOMEM : BATS_EHIP;
  }
  if (h->varyT) (void)hipFree(h->varyT);
  if (h->varyS) (void)hipFree(h->varyS);
  h->varyT = dT;
  h->varyS = dS;
  h->nVary = nPoint;
  return BATS_OK;
}

// B0 splitting: face B0 (set_b0_face, src/ModB0.f90:326-335; on two-level
// grids the faces of sides with finer neighbours take the mean of the fine
// faces' B0, set_b0_reschange, :338-351,366-455), and per cell the B0
// itself, div B0 and curl B0 from the face values (set_b0_source,
// :568-590), computed once on the host from B0_DGB
extern "C" int bats_set_b0(bats_handle *h, const double *B0_DGB,
                           int UseB0Source) {
  if (!h) return BATS_EINVAL;
  drop_graphs(h);
  HIPCHK(hipStreamSynchronize(h->stream));
  if (!B0_DGB) {
    drop_b0(h);
    return BATS_OK;
  }
  const bats_config &c = h->cfg;
  // FP64 state, no low order faces (the B0 stage variants do not carry
  // them)
  if (h->elemSize != 8) return BATS_EINVAL;
  if (low_order_used(c)) return BATS_EINVAL;
  const int NB = h->NB, g = c.nG, NG = NB + 2 * g, nB = c.nBlock;
  for (long t = 0; t < 3L * NG * NG * NG * nB; t++)
    if (!std::isfinite(B0_DGB[t])) return BATS_EINVAL;
  const long NF = (long)(NB + 1) * NB * NB, nFaceTot = (long)nB * NF;
  const long nCell = h->nCell;
  std::vector<double> F(9 * nFaceTot), C(3 * nCell);
  // B0_DGB(3, MinI:MaxI, MinJ:MaxJ, MinK:MaxK, nBlock), variable fastest
  auto B0 = [&](int b, int i, int j, int k, int d) {
    return B0_DGB[((((long)b * NG + (k + g)) * NG + (j + g)) * NG + (i + g)) * 3 + d];
  };
  for (int b = 0; b < nB; b++)
    for (int k = 0; k <= NB; k++)
      for (int j = 0; j <= NB; j++)
        for (int i = 0; i <= NB; i++)
          for (int d = 0; d < 3; d++) {
            if (j < NB && k < NB)
              F[(0 * 3 + d) * nFaceTot +
                (NB == 8 ? b0_face_index<8>(b, 0, i, j, k)
                         : b0_face_index<16>(b, 0, i, j, k))] =
                  0.5 * (B0(b, i - 1, j, k, d) + B0(b, i, j, k, d));
            if (i < NB && k < NB)
              F[(1 * 3 + d) * nFaceTot +
                (NB == 8 ? b0_face_index<8>(b, 1, i, j, k)
                         : b0_face_index<16>(b, 1, i, j, k))] =
                  0.5 * (B0(b, i, j - 1, k, d) + B0(b, i, j, k, d));
            if (i < NB && j < NB)
              F[(2 * 3 + d) * nFaceTot +
                (NB == 8 ? b0_face_index<8>(b, 2, i, j, k)
                         : b0_face_index<16>(b, 2, i, j, k))] =
                  0.5 * (B0(b, i, j, k - 1, d) + B0(b, i, j, k, d));
          }
  auto face = [&](int dir, int b, int i, int j, int k, int d) -> double & {
    return F[(dir * 3 + d) * nFaceTot +
             (NB == 8 ? b0_face_index<8>(b, dir, i, j, k)
                      : b0_face_index<16>(b, dir, i, j, k))];
  };
  // a side with finer neighbours (kind-1 slab): the face (p, q) takes 1/8
  // times the sum over the four fine faces it covers of the fine block's
  // B0 in its ghost cell and first cell there, summed in the order of the
  // fine faces' quadrant index
  const int H = NB / 2;
  for (size_t t = 0; t < h->userSlabs.size() / 8; t++) {
    const int *sl = h->userSlabs.data() + 8 * t;
    if (sl[2] != 1) continue;
    const int b = sl[0], d = sl[1] / 2, hi = sl[1] % 2;
    // the fine block's side facing b: ghost -1 and first cell 0 (its low
    // side) or NB and NB-1 (its high side)
    const int ng = hi ? -1 : NB, nc = hi ? 0 : NB - 1;
    auto at = [&](int nrm, int p, int q, int &i, int &j, int &k) {
      i = d == 0 ? nrm : p;
      j = d == 1 ? nrm : d == 0 ? p : q;
      k = d == 2 ? nrm : q;
    };
    for (int q = 0; q < NB; q++)
      for (int p = 0; p < NB; p++) {
        const int f = sl[3 + p / H + 2 * (q / H)];
        double sum[3] = {0.0, 0.0, 0.0};
        for (int a = 0; a < 4; a++) {
          const int pf = 2 * (p % H) + a % 2, qf = 2 * (q % H) + a / 2;
          int ig, jg, kg, ic, jc, kc;
          at(ng, pf, qf, ig, jg, kg);
          at(nc, pf, qf, ic, jc, kc);
          for (int v = 0; v < 3; v++)
            sum[v] = sum[v] + 0.125 * (B0(f, ig, jg, kg, v) + B0(f, ic, jc, kc, v));
        }
        int i, j, k;
        at(hi ? NB : 0, p, q, i, j, k);
        for (int v = 0; v < 3; v++) face(d, b, i, j, k, v) = sum[v];
      }
  }
  // the cell B0 (div B0 and curl B0 are formed from the face values where
  // the source needs them)
  for (int b = 0; b < nB; b++)
    for (int k = 0; k < NB; k++)
      for (int j = 0; j < NB; j++)
        for (int i = 0; i < NB; i++) {
          const long off = (((long)b * NB + k) * NB + j) * NB + i;
          for (int d = 0; d < 3; d++) C[d * nCell + off] = B0(b, i, j, k, d);
        }
  // the ghost B0 of every outer side (the caller's -1 sides), for the
  // fixed, inflow and vary ghost cells: slot per (block, side), layer L =
  // 0, 1 away from the side, in-plane (p, q) as the ghost slabs
  std::vector<int> slot(6L * nB, -1);
  std::vector<double> SB;
  for (int b = 0; b < nB; b++)
    for (int s = 0; s < 6; s++) {
      if (h->userNei[6 * b + s] != -1) continue;
      slot[6 * b + s] = (int)(SB.size() / (6L * NB * NB));
      const int d = s / 2, hi = s % 2;
      for (int L = 0; L < 2; L++)
        for (int q = 0; q < NB; q++)
          for (int p = 0; p < NB; p++) {
            const int nrm = hi ? NB + L : -1 - L;
            const int i = d == 0 ? nrm : p, j = d == 1 ? nrm : d == 0 ? p : q,
                      k = d == 2 ? nrm : q;
            for (int v = 0; v < 3; v++) SB.push_back(B0(b, i, j, k, v));
          }
    }
  double *dF = nullptr, *dC = nullptr, *dSB = nullptr;
  int *dSlot = nullptr;
  hipError_t e = hipMalloc(&dF, sizeof(double) * F.size());
  if (e == hipSuccess) e = hipMalloc(&dC, sizeof(double) * C.size());
  if (e == hipSuccess) e = hipMalloc(&dSlot, sizeof(int) * slot.size());
  if (e == hipSuccess && !SB.empty())
    e = hipMalloc(&dSB, sizeof(double) * SB.size());
  if (e == hipSuccess)
    e = hipMemcpy(dF, F.data(), sizeof(double) * F.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(dC, C.data(), sizeof(double) * C.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(dSlot, slot.data(), sizeof(int) * slot.size(),
                  hipMemcpyHostToDevice);
  if (e == hipSuccess && dSB)
    e = hipMemcpy(dSB, SB.data(), sizeof(double) * SB.size(),
                  hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    void *made[] = {dF, dC, dSB, dSlot};
    for (void *p : made)
      if (p) (void)hipFree(p);
    return e == hipErrorOutOfMemory ? BATS_ENOMEM : BATS_EHIP;
  }
  drop_b0(h);
  h->b0F = dF;
  h->b0C = dC;
  h->b0Side = dSB;
  h->b0SideSlot = dSlot;
  h->useB0 = 1;
  h->useB0Src = UseB0Source != 0;
  return BATS_OK;
}

// Inner body: the true cells Used_GB (with the ghost cells) and the
// 'ionosphere' face boundary condition of the fast update
// (src/ModUpdateStateFast.f90:230-265,343-345,924-937,1277-1327).  Blocks
// with a body cell among their cells or ghost cells (IsBody_B,
// src/ModBoundaryGeometry.f90:226) get a mask slot; the groups of blocks
// holding one run the body variant of the stage kernel.
extern "C" int bats_set_body(bats_handle *h, const unsigned char *Used_GB,
                             const double *XyzStart_DB, double BodyRho,
                             double BodyP, double B1rCoef) {
  if (!h) return BATS_EINVAL;
  drop_graphs(h);
  HIPCHK(hipStreamSynchronize(h->stream));
  if (!Used_GB) {
    drop_body(h);
    return BATS_OK;
  }
  const bats_config &c = h->cfg;
  // FP64 state, no low order faces or resolution-change beta (the body
  // variant carries neither)
  if (h->elemSize != 8 || low_order_used(c) || res_beta_used(h) ||
      !XyzStart_DB)
    return BATS_EINVAL;
  // the ghost pressure floats, so BodyP only has to be a number
  if (!(BodyRho > 0.0) || !std::isfinite(BodyRho) || !std::isfinite(BodyP) ||
      !std::isfinite(B1rCoef))
    return BATS_EINVAL;
  const int NB = h->NB, g = c.nG, NGc = NB + 2 * g, nB = c.nBlock;
  const int NG = NB + 4;  // the mask over the block and its 2-cell halo
  for (long t = 0; t < 3L * nB; t++)
    if (!std::isfinite(XyzStart_DB[t])) return BATS_EINVAL;
  const long nMask = (long)NG * NG * NG;
  std::vector<int> slot(nB, -1);
  std::vector<unsigned char> masks;
  for (int b = 0; b < nB; b++) {
    const unsigned char *U = Used_GB + (long)b * NGc * NGc * NGc;
    bool any = false;
    for (long q = 0; q < (long)NGc * NGc * NGc && !any; q++) any = !U[q];
    if (!any) continue;
    slot[b] = (int)(masks.size() / nMask);
    for (int k = -2; k < NB + 2; k++)
      for (int j = -2; j < NB + 2; j++)
        for (int i = -2; i < NB + 2; i++)
          masks.push_back(
              U[(((long)(k + g) * NGc) + (j + g)) * NGc + (i + g)] ? 1 : 0);
  }
  drop_body(h);
  if (masks.empty()) return BATS_OK;  // no body cell anywhere
  // work list: groups without a body block first, then those with one
  const int BPW = kThreads / (NB * NB);
  const int nGroup = (nB + BPW - 1) / BPW;
  std::vector<int> plain, body;
  for (int q = 0; q < nGroup; q++) {
    bool any = false;
    for (int b = q * BPW; b < std::min(nB, (q + 1) * BPW); b++)
      any = any || slot[b] >= 0;
    (any ? body : plain).push_back(q);
  }
  std::vector<int> list(plain);
  list.insert(list.end(), body.begin(), body.end());
  int *dSlot = nullptr, *dList = nullptr;
  unsigned char *dUsed = nullptr;
  double *dXyz = nullptr;
  hipError_t e = hipMalloc(&dSlot, sizeof(int) * nB);
  if (e == hipSuccess) e = hipMalloc(&dList, sizeof(int) * list.size());
  if (e == hipSuccess) e = hipMalloc(&dUsed, masks.size());
  if (e == hipSuccess) e = hipMalloc(&dXyz, sizeof(double) * 3 * nB);
  if (e == hipSuccess)
    e = hipMemcpy(dSlot, slot.data(), sizeof(int) * nB, hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(dList, list.data(), sizeof(int) * list.size(),
                  hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(dUsed, masks.data(), masks.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(dXyz, XyzStart_DB, sizeof(double) * 3 * nB,
                  hipMemcpyHostToDevice);
  h->bodySlot = dSlot;
  h->groupList = dList;
  h->bodyUsed = dUsed;
  h->bodyXyz0 = dXyz;
  if (e != hipSuccess) {
    drop_body(h);
    return e == hipErrorOutOfMemory ? BATS_ENOMEM : BATS_EHIP;
  }
  h->nPlainGroup = (int)plain.size();
  h->nBodyGroup = (int)body.size();
  h->bodyRho = BodyRho;
  h->b1rCoef = B1rCoef;
  h->useBody = 1;
  return BATS_OK;
}

extern "C" int bats_synchronize(bats_handle *h) {
  if (!h) return BATS_EINVAL;
  HIPCHK(hipStreamSynchronize(h->stream));
  int bits = 0;
  HIPCHK(hipMemcpy(&bits, h->err, sizeof(int), hipMemcpyDeviceToHost));
  return err_from_bits(bits);
}

extern "C" int bats_check_nan(bats_handle *h) {
  if (!h) return BATS_EINVAL;
  const long n = h->nCell;  // interior cells of each variable
  for (int v = 0; v < NVAR; v++) {
    if (h->elemSize == 8)
      hipLaunchKernelGGL(k_check_nan<double>, dim3(1024), dim3(256), 0,
                         h->stream,
                         static_cast<const double *>(h->U[h->stateBuf]) + v * h->vs,
                         n, h->err);
    else
      hipLaunchKernelGGL(k_check_nan<float>, dim3(1024), dim3(256), 0,
                         h->stream,
                         static_cast<const float *>(h->U[h->stateBuf]) + v * h->vs,
                         n, h->err);
  }
  HIPCHK(hipGetLastError());
  return bats_synchronize(h);
}

extern "C" int bats_integrate_domain(bats_handle *h, double *Sum_V) {
  if (!h || !Sum_V) return BATS_EINVAL;
  const bats_config &c = h->cfg;
  const void *U = h->U[h->stateBuf];
  const double *Ud = static_cast<const double *>(U);
  const float *Uf = static_cast<const float *>(U);
  const dim3 g(c.nBlock), t(256);
  if (h->NB == 8 && h->elemSize == 8)
    hipLaunchKernelGGL((k_integrate<8, double>), g, t, 0, h->stream, Ud,
                       h->vs, h->cellsize, make_phys(c), c.nBlock, h->partial);
  else if (h->NB == 8)
    hipLaunchKernelGGL((k_integrate<8, float>), g, t, 0, h->stream, Uf,
                       h->vs, h->cellsize, make_phys(c), c.nBlock, h->partial);
  else if (h->elemSize == 8)
    hipLaunchKernelGGL((k_integrate<16, double>), g, t, 0, h->stream, Ud,
                       h->vs, h->cellsize, make_phys(c), c.nBlock, h->partial);
  else
    hipLaunchKernelGGL((k_integrate<16, float>), g, t, 0, h->stream, Uf,
                       h->vs, h->cellsize, make_phys(c), c.nBlock, h->partial);
  HIPCHK(hipGetLastError());
  std::vector<double> p((size_t)NVAR * c.nBlock);
  HIPCHK(hipMemcpyAsync(p.data(), h->partial, sizeof(double) * p.size(),
                        hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  for (int v = 0; v < NVAR; v++) Sum_V[v] = 0.0;
  for (int b = 0; b < c.nBlock; b++)
    for (int v = 0; v < NVAR; v++) Sum_V[v] += p[(size_t)b * NVAR + v];
  return BATS_OK;
}

extern "C" int bats_last_stage_kernel_ms(bats_handle *h, double *ms, int *nLaunch) {
  if (!h) return BATS_EINVAL;
  if (ms) *ms = h->lastStageMs;
  if (nLaunch) *nLaunch = h->lastStageLaunches;
  return BATS_OK;
}
