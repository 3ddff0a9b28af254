// stage_kernel.hpp -- the fused stage kernel k_stage and the device helpers
// it shares with the per-block routines.
//
// Device layout: the state is held interior-only, variable-major
// (structure of arrays): U[((v*nBlock + b)*NB + k)*NB + j)*NB + i].  There
// are no ghost cells in HBM; halos are read directly from the neighbouring
// blocks' interiors through the neighbour table, which replaces the
// reference's ghost-cell fill (exchange_messages / BATL message_pass_cell)
// for same-level blocks.  Two state buffers ping-pong between the stages of
// the midpoint scheme, so StateOld is never copied.
//
// k_stage fuses, per block, what the reference does in calc_face_value +
// calc_face_flux + calc_source + update_state (+ calc_timestep at the last
// stage) (src/ModAdvanceExplicit.f90:140-220):
//   * a workgroup of 256 threads owns one 16^3 block or four 8^3 blocks;
//     each thread owns one (i,j) column and marches along k;
//   * z faces: every thread keeps a window of primitive states of its column
//     in registers and computes the z face above its cell; the flux carries
//     over to the next plane, so each z face is computed once;
//   * x/y faces: the plane's primitives plus a 2-cell x/y halo are staged
//     in LDS, all x and y faces of the plane are spread over the 256
//     threads, and their fluxes go back to LDS;
//   * each thread then forms the flux divergence of its cell (and, with
//     the 8-wave scheme, the div B source from the face states), updates it
//     (p -> E, U = Uold + dt*R, E -> p) and writes it; face states and
//     fluxes never reach HBM.
//
// The kernel is instantiated per (block size, flux) in its own translation
// unit (stage_inst.hip) so the instantiations compile in parallel.
#pragma once
#include <hip/hip_runtime.h>

#include "mhd_device.hpp"

namespace bats {

enum { S_DTCB = 0, S_DT = 1, S_TSIM = 2, S_NSTEP = 3, S_DTMIN = 4, S_NSCAL = 8 };
constexpr int kThreads = 256;
// record of one fine sub-face at a resolution change: flux slots 0..6,
// PP <- Vdt, EN, then the normal field of the face's left and right states
// times 1/4 (save_cons_flux, src/ModConserveFlux.f90:129-140)
enum { CORR_BNL = 9, CORR_BNR = 10, NCORR = 11 };

struct StageArgs {
  // state buffers, double or float (the kernel's storage type T)
  const void *Uin;     // stage input (stencils and halos)
  const void *Uold;    // state at the start of the step (StateOld)
  void *Uout;          // updated state (may alias Uold)
  const int *nei;      // 6 per block
  const double *cellsize;  // 3 per block
  double *scal;        // device scalars, S_NSCAL per run (S_*)
  const int *runOfBlock;  // run index of each block (null: one run)
  int nRun;
  double *dtmaxB;      // per block minimum of the local time step
  double *dtmaxCB;     // optional per-cell local time step
  int *err;
  int nBlock;
  long vs;             // variable stride = nBlock*NB^3
  Phys ph;
  double beta;
  int lim;             // LIM_* (LIM_FIRST for nOrder = 1)
  double DtFactor;     // (Cfl*iStage)/nStage or Cfl
  int iStage, nStage, halfStep, dbTrick, doDt, roundTripOld, noUpdate;
  int advanceTime;     // add Dt/nStage to tSimulation when done
  long nCellTot;       // interior cells; ghost slabs follow in each plane
  const int *slabKind; // resolution-change slabs: 0 coarser, 1 finer side
  const void *corr;    // fine sub-face records of the finer sides (type
                       // T), [slab][q][p][sub 0..3][NCORR]; null: no slabs
  int nGroup;          // workgroup-sized groups of blocks (work items)
  int nWorkgroup;      // launched workgroups (< nGroup: persistent grid)
  const void *lowCrit; // low order weight of every face (type T, see
                       // crit_index); read only by the low order variant
  double betaRC;       // BetaLimiterResChange, nFaceLimiterResChange: read
  int nFaceRC;         // only by the OPT_RESBETA variant
  // B0 splitting (OPT_B0 variant, FP64): B0 of every face, [dir][comp]
  // [block][face in the block's LeftState_V? order], and per cell B0,
  // div B0, curl B0 ([7][interior cell]); useB0Src: UseB0Source
  const double *b0F, *b0C;
  int useB0Src;
  // groups of blocks this launch processes (null: groups 0..nGroup-1)
  const int *groupList;
  // steady state in the B0 and body variants (the plain variants carry it
  // as OPT_STEADY)
  int steady;
  // launch shape: each block group splits into 2^splitShift work items of
  // NB >> splitShift planes (cells per thread); xcdOrder hands each XCD a
  // contiguous range of work items (workgroups go round-robin over the 8
  // XCDs), so neighbouring blocks share an L2
  int splitShift, xcdOrder;
  // inner body (OPT_BODY variant, FP64): per block the slot of its true
  // cell mask (-1: no body cell within 2 cells of the block), the masks
  // [slot][k+2][j+2][i+2] over the block and its 2-cell halo (1: true
  // cell), each block's first cell centre relative to the body centre, and
  // the body density and B1rCoef of the ghost face state
  const int *bodySlot;
  const unsigned char *used;
  const double *xyz0;
  double bodyRho, b1rCoef;
};

// the cyclic frame rotation (x, y, z) -> (y, z, x) that turns a y face into
// an x face: rotated component d of a vector is the original component
// (d + 1) % 3; rot_src gives the original variable of each rotated slot
__host__ __device__ constexpr int rot_src(int v) {
  return (v >= MX && v <= MZ) ? MX + (v - MX + 1) % 3
       : (v >= BX && v <= BZ) ? BX + (v - BX + 1) % 3 : v;
}
// flux of the rotated frame back to the original components
template <typename T>
__device__ __forceinline__ void unrotate_flux(T *F) {
  const T m0 = F[MX], m1 = F[MY], m2 = F[MZ];
  F[MY] = m0; F[MZ] = m1; F[MX] = m2;
  const T b0 = F[BX], b1 = F[BY], b2 = F[BZ];
  F[BY] = b0; F[BZ] = b1; F[BX] = b2;
}

// Stage kernel variants: the 8-wave div B source and the adaptive low order
// face values are compile-time options so the plain kernel carries neither.
// OPT_RESBETA: the sharper limiter next to sides with DiLevel_EB /= 0
// (outer boundaries and resolution changes).
// OPT_STEADY: steady state, every cell with its own time step (dtmaxCB,
// from this launch's fluxes when doDt).
// OPT_B0: B0 splitting, the state holds B1 = B - B0 (FP64 only).
// OPT_BODY: the inner body's true cells and face boundary condition (FP64,
// launched only on the groups of blocks that reach a body cell).
// OPT_TVD: the TVD face values at resolution changes (UseTvdResChange): a
// fine block's faces at a coarser side take the flux, and the faces one
// cell in the fine-side state, that k_amr_fineflux left in the side's
// record (the record layout is below).
enum { OPT_POWELL = 1, OPT_LOWORDER = 2, OPT_RESBETA = 4, OPT_STEADY = 8,
       OPT_B0 = 16, OPT_BODY = 32, OPT_TVD = 64 };
// record of one fine face (p, q) of a coarser side with the TVD treatment,
// in the side's slab store: the face flux (slots 0..6, PP <- Vdt, EN), the
// fine-side primitive state of the face one cell in, and the normal field
// of the face's left and right states (for the 8-wave source)
enum { TVD_FLUX = 0, TVD_FINEF = 9, TVD_BN = 17, NTVD = 19 };
static_assert(NTVD <= 4 * NCORR, "the TVD record fits a slab cell's store");

// tvd_reschange (src/ModFaceValue.f90:2269-2343) for subface f of a 2x2
// group, on primitive states: C2, C1 the coarse cells behind the face, F1
// the group's four first-layer fine cells (f = i2 + 2 j2), F2 this
// subface's second-layer fine cell.  Out: the coarse-side and fine-side
// values at the subface and the fine value at the next face.
template <typename T>
__device__ __forceinline__ void tvd_subface(T Beta, const T *C2, const T *C1,
                                            const T (*F1)[NVAR], const T *F2,
                                            int f, T *CtoF, T *FtoC, T *FF) {
  const T cThird = T(1.0 / 3.0), cTwoThird = T(2.0 / 3.0),
          cSixth = T(1.0 / 6.0);
#pragma unroll
  for (int v = 0; v < NVAR; v++) {
    const T Avg = T(0.25) * (((F1[0][v] + F1[1][v]) + F1[2][v]) + F1[3][v]);
    const T Grad = Avg - C1[v];
    const T Sgn = fsign(T(1), Grad);
    const T dC = C1[v] - C2[v];
    const T Ltd =
        Sgn * fmax2(T(0), fmin2(fmin2(Beta * cTwoThird * fabsr(Grad),
                                      Beta * T(0.5) * Sgn * dC),
                                cThird * fabsr(Grad) + T(0.25) * Sgn * dC));
    const T t = F1[f][v] - Avg;
    CtoF[v] = C1[v] + Ltd + fsign(fmin2(fabsr(Ltd), fabsr(t)), t);
    const T dF = F2[v] - F1[f][v];
    const T Ltd2 =
        Sgn * fmax2(T(0), fmin2(fmin2(Sgn * (F1[f][v] - C1[v]),
                                      Beta * cThird * fabsr(Grad)),
                                fmin2(Beta * T(0.5) * Sgn * dF,
                                      cSixth * fabsr(Grad) +
                                          T(0.25) * Sgn * dF)));
    FtoC[v] = F1[f][v] - Ltd2;
    FF[v] = F1[f][v] + Ltd2;
  }
}

// index of a face of block b in the B0 face arrays (per direction, the
// reference's LeftState_VX(nI+1,nJ,nK) / _VY(nI,nJ+1,nK) / _VZ(nI,nJ,nK+1)
// order within a block); (fi, fj, fk) are the face's cell-index coordinates
template <int NB>
__host__ __device__ __forceinline__ long b0_face_index(int b, int dir, int fi,
                                                       int fj, int fk) {
  constexpr long NF = (long)(NB + 1) * NB * NB;
  long l;
  if (dir == 0) l = ((long)fk * NB + fj) * (NB + 1) + fi;
  else if (dir == 1) l = ((long)fk * (NB + 1) + fj) * NB + fi;
  else l = ((long)fk * NB + fj) * NB + fi;
  return (long)b * NF + l;
}

// index of the low order weight of a face of block b: the face normal to
// dir at face coordinates (i, j, k), whose dir component is the face index
// 0..NB (face between cells f-1 and f).  x and y faces are stored plane by
// plane in the stage kernel's face order, z faces by face then column.
template <int NB>
__host__ __device__ __forceinline__ long crit_index(int b, int dir, int i,
                                                    int j, int k) {
  constexpr int NFX = (NB + 1) * NB;
  const long base = ((long)b * 3 + dir) * NB * NFX;
  if (dir == 0) return base + (long)k * NFX + j * (NB + 1) + i;
  if (dir == 1) return base + (long)k * NFX + i * (NB + 1) + j;
  return base + ((long)k * NB + j) * NB + i;
}

// offset (within one variable plane) of cell (i,j,k) of block b; one index
// may lie in the 2-cell halo: it is resolved to the neighbour block, at a
// float outer boundary to the block's own boundary cell
// (src/ModCellBoundary.f90:518-559), and at a resolution change to the
// side's ghost slab stored after the nCellTot interior cells (two layers of
// NB x NB cells per slab, in-plane order (p, q) = (j,k) / (i,k) / (i,j) on
// x / y / z sides).  Only one index is ever outside the block: the stencils
// are dimension-split.
template <int NB>
__device__ __forceinline__ long cell_off(const int *__restrict__ nei, int b,
                                         int i, int j, int k, long nCellTot) {
  auto slab = [&](int n, int layer, int p, int q) {
    return nCellTot + (((long)(-2 - n) * 2 + layer) * NB + q) * NB + p;
  };
  if (i < 0) {
    const int n = nei[6 * b + 0];
    if (n >= 0) { b = n; i += NB; }
    else if (n == -1) { i = 0; }
    else return slab(n, -1 - i, j, k);
  } else if (i >= NB) {
    const int n = nei[6 * b + 1];
    if (n >= 0) { b = n; i -= NB; }
    else if (n == -1) { i = NB - 1; }
    else return slab(n, i - NB, j, k);
  }
  if (j < 0) {
    const int n = nei[6 * b + 2];
    if (n >= 0) { b = n; j += NB; }
    else if (n == -1) { j = 0; }
    else return slab(n, -1 - j, i, k);
  } else if (j >= NB) {
    const int n = nei[6 * b + 3];
    if (n >= 0) { b = n; j -= NB; }
    else if (n == -1) { j = NB - 1; }
    else return slab(n, j - NB, i, k);
  }
  if (k < 0) {
    const int n = nei[6 * b + 4];
    if (n >= 0) { b = n; k += NB; }
    else if (n == -1) { k = 0; }
    else return slab(n, -1 - k, i, j);
  } else if (k >= NB) {
    const int n = nei[6 * b + 5];
    if (n >= 0) { b = n; k -= NB; }
    else if (n == -1) { k = NB - 1; }
    else return slab(n, k - NB, i, j);
  }
  return (((long)b * NB + k) * NB + j) * NB + i;
}

template <typename T>
__device__ __forceinline__ void load_state(const T *__restrict__ U, long vs,
                                           long off, T *q) {
#pragma unroll
  for (int v = 0; v < NVAR; v++) q[v] = U[v * vs + off];
}

// work item of launch slot `item` of n: with the XCD order, the slots that
// land on XCD x (item % 8 == x) take the x-th contiguous eighth of the
// items (a bijection of [0, n))
__device__ __forceinline__ int xcd_item(int item, int n) {
  const int x = item & 7, r = item >> 3, q = n >> 3, rem = n & 7;
  return x * q + (x < rem ? x : rem) + r;
}

__device__ __forceinline__ double wave_min(double x) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) x = fmin2(x, __shfl_xor(x, m, 64));
  return x;
}

// face values and face flux of one face; Bn (when given) receives the
// normal field of the left and right face states, which the 8-wave source
// reads
template <int DIR, int FLUX, int LIM, bool B0F, typename T>
__device__ __forceinline__ int face_flux_w(const PhysT<T> &ph, T area,
                                           T betaL, T betaR, T w,
                                           const T *q0, const T *q1,
                                           const T *q2, const T *q3, T *F,
                                           T &vdt, T &unA, const T *B0,
                                           T *Bn = nullptr) {
  T L[NVAR], R[NVAR];
  face_states_w<LIM>(betaL, betaR, w, q0, q1, q2, q3, L, R);
  if (Bn) { Bn[0] = L[BX + DIR]; Bn[1] = R[BX + DIR]; }
  return numerical_flux<DIR, FLUX, false, B0F>(ph, area, L, R, F, vdt, unA,
                                               B0);
}

template <int DIR, int FLUX, int LIM, bool B0F, typename T>
__device__ __forceinline__ int face_flux(const PhysT<T> &ph, T area, T betaL,
                                         T betaR, const T *q0, const T *q1,
                                         const T *q2, const T *q3, T *F,
                                         T &vdt, T &unA, const T *B0,
                                         T *Bn = nullptr) {
  T L[NVAR], R[NVAR];
  face_states<LIM>(betaL, betaR, q0, q1, q2, q3, L, R);
  if (Bn) { Bn[0] = L[BX + DIR]; Bn[1] = R[BX + DIR]; }
  return numerical_flux<DIR, FLUX, false, B0F>(ph, area, L, R, F, vdt, unA,
                                               B0);
}

// The inner body's treatment of one face (get_face_x/y/z of the fast
// update, src/ModUpdateStateFast.f90:924-937): u0..u3 flag the true cells
// of the stencil f-2 .. f+1 along the normal, P1 / P2 are the primitive
// states of the cells left and right of the face.  A second-order face
// value whose stencil meets a body cell becomes first order; a face between
// a true and a body cell takes the ghost state on the body side (r: the
// face centre relative to the body centre, in the face's frame).  Returns
// true for a face between two body cells, whose flux is zero (:237-240).
template <int LIM, typename T>
__device__ __forceinline__ bool body_face(bool u0, bool u1, bool u2, bool u3,
                                          const T *P1, const T *P2, T *L,
                                          T *R, const T *r, T BodyRho,
                                          T B1rCoef) {
  if (!u1 && !u2) return true;
  if (LIM != LIM_FIRST) {
    if (!(u0 && u1 && u2))
#pragma unroll
      for (int v = 0; v < NVAR; v++) L[v] = P1[v];
    if (!(u1 && u2 && u3))
#pragma unroll
      for (int v = 0; v < NVAR; v++) R[v] = P2[v];
  }
  if (u1 && !u2) body_ghost(L, r, BodyRho, B1rCoef, R);
  else if (u2 && !u1) body_ghost(R, r, BodyRho, B1rCoef, L);
  return false;
}

// ---------------------------------------------------------------------
// fused stage kernel.  T is the storage and face-arithmetic type (double:
// plain path; float: FP32-state path); the stage update, the pressure fix
// and the time step stay in double.
// ---------------------------------------------------------------------
template <int NB, int FLUX, int LIM, int OPT, typename T>
__global__ void __launch_bounds__(kThreads, sizeof(T) == 4 ? 3 : 2)
    k_stage(StageArgs a) {
  constexpr bool POWELL = (OPT & OPT_POWELL) != 0;
  // adaptive low order faces: slopes scaled by 1 - weight of the face
  constexpr bool LOWORD = (OPT & OPT_LOWORDER) != 0 && LIM != LIM_FIRST;
  // per-cell limiter beta near sides with DiLevel_EB /= 0
  constexpr bool RESB = (OPT & OPT_RESBETA) != 0 && LIM != LIM_FIRST;
  constexpr bool STEADY_BIT = (OPT & OPT_STEADY) != 0;
  constexpr bool B0S = (OPT & OPT_B0) != 0;
  static_assert(!B0S || sizeof(T) == 8, "B0 splitting is FP64 only");
  constexpr bool BODY = (OPT & OPT_BODY) != 0;
  static_assert(!BODY || (sizeof(T) == 8 && !LOWORD && !RESB && !STEADY_BIT),
                "the body variant is FP64, without low order faces or the "
                "resolution-change beta (steady state: the run-time flag)");
  constexpr bool TVD = (OPT & OPT_TVD) != 0;
  static_assert(!TVD || (!LOWORD && !STEADY_BIT),
                "the TVD variant carries no low order faces (steady state: "
                "the run-time flag)");
  constexpr int NC = NB * NB;          // columns per block
  constexpr int BPW = kThreads / NC;   // blocks per workgroup
  constexpr int NR = NB + 4;           // rows of the haloed plane
  constexpr int NS = NB + 5;           // odd row stride: no LDS bank conflicts
  constexpr int NFX = (NB + 1) * NB;   // x (or y) faces per plane
  constexpr int NFS = 9;               // flux slots: 0..6, PP<-vdt, EN
  // z-face fluxes of the lower and upper face of the cell; with the 8-wave
  // scheme two more slots carry the normal-field face states Bz_L, Bz_R
  constexpr int NFZ = POWELL ? NFS + 2 : NFS;
  // LDS slots of the x/y faces: with the 8-wave scheme the face loop keeps
  // the normal-field face states for the source when two workgroups per CU
  // (three for FP32) still fit in the 160 KiB of LDS with them, and always
  // in the body variant (which changes them); otherwise the update
  // recomputes them from the plane (and the TVD faces from their records)
  constexpr int LDS_BN = (BPW * NVAR * NR * NS + BPW * 2 * (NFS + 2) * NFX) *
                             (int)sizeof(T) + (kThreads / 64) * 8;
  constexpr bool BNLDS =
      POWELL && (BODY || LDS_BN * (sizeof(T) == 4 ? 3 : 2) <= 160 * 1024);
  constexpr int NFL = BNLDS ? NFS + 2 : NFS;
  __shared__ T sP[BPW * NVAR * NR * NS];
  __shared__ T sF[BPW * 2 * NFL * NFX];
  __shared__ double sRed[kThreads / 64];

  // steady state: a compile-time option of the plain variants, a run-time
  // flag of the B0, body and TVD variants (which would otherwise double)
  const bool STEADY = STEADY_BIT || ((B0S || BODY || TVD) && a.steady != 0);
  const int tid = threadIdx.x;
  // a wave never spans two blocks (NC >= 64): the block index, and with it
  // every per-block value (cell size, run, time step), is wave-uniform and
  // lives in scalar registers
  static_assert(NC % 64 == 0, "one block per wave");
  const int lb = BPW == 1 ? 0 : __builtin_amdgcn_readfirstlane(tid / NC);
  const int col = tid % NC, j = col / NB, i = col % NB;
  // this thread's cell of the 2-deep x/y halo ring of each plane (at most
  // one: the ring of a group's blocks has NH <= 256 cells)
  constexpr int NH = BPW * 8 * NB;
  static_assert(NH <= kThreads, "one halo cell per thread");
  const bool hasH = tid < NH;
  int hlb = 0, hii = 0, hjj = 0;
  if (hasH) {
    hlb = tid / (8 * NB);
    const int r = tid % (8 * NB), side = r / (2 * NB), rem = r % (2 * NB);
    if (side < 2) {
      const int layer = rem & 1, pos = rem >> 1;
      hjj = pos;
      hii = side == 0 ? layer - 2 : NB + layer;
    } else {
      const int layer = rem / NB, pos = rem % NB;
      hii = pos;
      hjj = side == 2 ? layer - 2 : NB + layer;
    }
  }
  const T *__restrict__ U = static_cast<const T *>(a.Uin);
  const T *Uold = static_cast<const T *>(a.Uold);
  T *Uout = static_cast<T *>(a.Uout);
  const long vs = a.vs;
  const PhysT<T> ph = phys_as<T>(a.ph);
  const T beta = T(a.beta);
  const T betaRC = T(a.betaRC);
  const int nFaceRC = a.nFaceRC;
  const T *crit = static_cast<const T *>(a.lowCrit);
  int err = 0;

  auto sPi = [&](int l, int v, int jj, int ii) {
    return ((l * NVAR + v) * NR + jj) * NS + ii;
  };
  // a face on a side with finer neighbours takes the sum of the four fine
  // face fluxes (apply_cons_flux, src/ModConserveFlux.f90:260-441)
  auto fine_record = [&](int blk, int side, int p, int q) -> const T * {
    const int n = a.nei[6 * blk + side];
    if (n > -2 || a.slabKind[-2 - n] != 1) return nullptr;
    return static_cast<const T *>(a.corr) +
           ((((long)(-2 - n)) * NB + q) * NB + p) * 4 * NCORR;
  };
  auto sum4 = [&](const T *c, int v) {
    return ((c[v] + c[NCORR + v]) + c[2 * NCORR + v]) + c[3 * NCORR + v];
  };
  auto corrected = [&](int blk, int side, int p, int q, T *F) {
    const T *c = fine_record(blk, side, p, q);
    if (!c) return;
#pragma unroll
    for (int v = 0; v < NFS; v++) F[v] = sum4(c, v);
  };
  // and, for the 8-wave source, so do the normal field of its left and
  // right face states (:355-372)
  auto corrected_bn = [&](int blk, int side, int p, int q, T &L, T &R) {
    const T *c = fine_record(blk, side, p, q);
    if (!c) return;
    L = sum4(c, CORR_BNL);
    R = sum4(c, CORR_BNR);
  };
  auto sFi = [&](int l, int dir, int v, int q) {
    return ((l * 2 + dir) * NFL + v) * NFX + q;
  };
  // the TVD record of face (p, q) of a side with a coarser neighbour
  auto tvd_rec = [&](int blk, int side, int p, int q) -> const T * {
    if (!TVD) return nullptr;
    const int n = a.nei[6 * blk + side];
    if (n > -2 || a.slabKind[-2 - n] != 0) return nullptr;
    return static_cast<const T *>(a.corr) +
           ((((long)(-2 - n)) * NB + q) * NB + p) * 4 * NCORR;
  };
  // the body: true cell mask of a block (null: all cells true) and the
  // flag of its cell (ii, jj, kk), each index in -2 .. NB+1
  constexpr int NG = NB + 4;
  auto body_mask = [&](int blk) -> const unsigned char * {
    if (!BODY) return nullptr;
    const int sl = a.bodySlot[blk];
    return sl < 0 ? nullptr : a.used + (long)sl * NG * NG * NG;
  };
  auto is_true = [&](const unsigned char *m, int ii, int jj, int kk) {
    return !m || m[((kk + 2) * NG + (jj + 2)) * NG + (ii + 2)] != 0;
  };
  // centre of the face between cells c and c + e_dir of block blk,
  // relative to the body centre (Xyz_DGB of the two cells averaged)
  auto face_xyz = [&](int blk, int dir, int ci, int cj, int ck, T *r) {
    const double *x0 = a.xyz0 + 3 * blk;
    const int c[3] = {ci, cj, ck};
#pragma unroll
    for (int d = 0; d < 3; d++) {
      const double h = a.cellsize[3 * blk + d];
      r[d] = T(0.5 * ((x0[d] + c[d] * h) + (x0[d] + (c[d] + (d == dir)) * h)));
    }
  };
  // B0 of a face (B0 splitting)
  const long nFaceTot = (long)a.nBlock * (NB + 1) * NB * NB;
  auto b0face = [&](int dir, int blk, int fi, int fj, int fk, T *out) {
    const long f = b0_face_index<NB>(blk, dir, fi, fj, fk);
#pragma unroll
    for (int c = 0; c < 3; c++) out[c] = T(a.b0F[(dir * 3 + c) * nFaceTot + f]);
  };

  // a persistent grid walks the flat work list of block groups; with one
  // workgroup per group the loop runs once
  // a work item is one part (planes k0 .. k1-1) of one group of blocks
  const int nItem = a.nGroup << a.splitShift;
  for (int item = blockIdx.x; item < nItem; item += gridDim.x) {
    const int w = a.xcdOrder ? xcd_item(item, nItem) : item;
    const int grp = w >> a.splitShift;
    const int kLen = NB >> a.splitShift;
    const int k0 = (w & ((1 << a.splitShift) - 1)) * kLen, k1 = k0 + kLen;
    const int blk0 = (a.groupList ? a.groupList[grp] : grp) * BPW;
    const int b = blk0 + lb;
    const bool act = b < a.nBlock;
    const int bb = act ? b : 0;
    const int hb = blk0 + hlb;
    const bool hAct = hasH && hb < a.nBlock;

    const double dxb = a.cellsize[3 * bb], dyb = a.cellsize[3 * bb + 1],
                 dzb = a.cellsize[3 * bb + 2];
    const T AreaZ = T(dxb * dyb);
    // z sides of this block with DiLevel_EB /= 0 (wave-uniform)
    const bool zLo = RESB && a.nei[6 * bb + 4] < 0;
    const bool zHi = RESB && a.nei[6 * bb + 5] < 0;
    auto betaZ = [&](int c) {
      return RESB ? cell_beta(beta, betaRC, nFaceRC, zLo, zHi, c, NB) : beta;
    };
    const double CellVolume = dxb * dyb * dzb;
    const unsigned char *Ub = body_mask(bb);
    T InvHalf[3];
    if (POWELL) {
      InvHalf[0] = T(0.5 / dxb);
      InvHalf[1] = T(0.5 / dyb);
      InvHalf[2] = T(0.5 / dzb);
    }

    // this block's run: its own time step (per-run CFL dt)
    const double *rs =
        a.scal + (a.runOfBlock ? a.runOfBlock[bb] : 0) * S_NSCAL;
    UpdateCtl uc;
    uc.DtLocal = a.DtFactor * rs[S_DTCB];   // time accurate: DtMax_CB = Dt
    uc.InvVolume = 1.0 / CellVolume;
    uc.halfStep = a.halfStep;
    uc.iStage = a.iStage;
    uc.nStage = a.nStage;
    uc.dbTrick = a.dbTrick;
    uc.dtZero = !STEADY && (rs[S_DT] == 0.0);
    uc.roundTripOld = a.roundTripOld;

    // z window of the column: Q1 = P(k), Q2 = P(k+1), Q3 = P(k+2) and the
    // left-state slope dLk of cell k; each cell's slopes are computed once
    // and serve the faces below and above it
    T Q1[NVAR], Q2[NVAR], Q3[NVAR], dLk[NVAR];
    T FzLo[NFZ], FzHi[NFZ];
    double dtmin = 1.7976931348623157e308;

    auto load_prim = [&](int kk, T *q) {
      const long off = cell_off<NB>(a.nei, bb, i, j, kk, a.nCellTot);
      load_state(U, vs, off, q);
      to_primitive(q);
    };
    // z face fz (between Q1 = P(fz-1) and Q2 = P(fz)); Q3 = P(fz+1)
    auto zface = [&](T *Fz, int fz) {
      T L[NVAR], R[NVAR], F[NFLUX], vdt, unA;
      T w = T(1);
      if (LOWORD) w = T(1) - crit[crit_index<NB>(bb, 2, i, j, fz)];
      const T bz = betaZ(fz);
#pragma unroll
      for (int v = 0; v < NVAR; v++) {
        T dl, dr;
        slopes<LIM>(bz, Q1[v], Q2[v], Q3[v], dl, dr);
        {
          // the left state adds the slope carried from the previous face
          // as a rounded value (no FMA with the product that formed it,
          // as the oracle rounds it), whether the slope comes from the
          // window's prologue or from the previous plane
#pragma clang fp contract(off)
          L[v] = LOWORD ? Q1[v] + dLk[v] * w : Q1[v] + dLk[v];
        }
        R[v] = LOWORD ? Q2[v] - dr * w : Q2[v] - dr;
        dLk[v] = dl;
      }
      if (TVD && (fz <= 1 || fz >= NB - 1)) {
        // faces at and one cell in from a coarser z side
        const T *rec = tvd_rec(bb, fz <= 1 ? 4 : 5, i, j);
        if (rec && (fz == 0 || fz == NB)) {
#pragma unroll
          for (int v = 0; v < NFS; v++) Fz[v] = rec[TVD_FLUX + v];
          if (POWELL) {
            Fz[NFZ - 2] = rec[TVD_BN];
            Fz[NFZ - 1] = rec[TVD_BN + 1];
          }
          return;
        }
        if (rec && fz == 1)
#pragma unroll
          for (int v = 0; v < NVAR; v++) L[v] = rec[TVD_FINEF + v];
        if (rec && fz == NB - 1)
#pragma unroll
          for (int v = 0; v < NVAR; v++) R[v] = rec[TVD_FINEF + v];
      }
      if (BODY && Ub) {
        T r[3];
        const bool u1 = is_true(Ub, i, j, fz - 1), u2 = is_true(Ub, i, j, fz);
        if (u1 != u2) face_xyz(bb, 2, i, j, fz - 1, r);
        if (body_face<LIM>(is_true(Ub, i, j, fz - 2), u1, u2,
                           is_true(Ub, i, j, fz + 1), Q1, Q2, L, R, r,
                           T(a.bodyRho), T(a.b1rCoef))) {
#pragma unroll
          for (int v = 0; v < NFZ; v++) Fz[v] = T(0);
          return;
        }
      }
      if (B0S) {
        T B0f[3];
        b0face(2, bb, i, j, fz, B0f);
        err |= numerical_flux<2, FLUX, false, true>(ph, AreaZ, L, R, F, vdt,
                                                    unA, B0f);
      } else {
        err |= numerical_flux<2, FLUX, false>(ph, AreaZ, L, R, F, vdt, unA);
      }
#pragma unroll
      for (int v = 0; v < NFLUX; v++) Fz[v] = F[v];
      Fz[PP] = vdt;
      if (POWELL) { Fz[NFZ - 2] = L[BZ]; Fz[NFZ - 1] = R[BZ]; }
    };

    if (act) {
      // prologue: slopes of cell k0-1 and the face below cell k0
      T Pm2[NVAR];
      load_prim(k0 - 2, Pm2);
      load_prim(k0 - 1, Q1);
      load_prim(k0, Q2);
      load_prim(k0 + 1, Q3);
      const T bz = betaZ(k0 - 1);
#pragma unroll
      for (int v = 0; v < NVAR; v++) {
        T dl, dr;
        slopes<LIM>(bz, Pm2[v], Q1[v], Q2[v], dl, dr);
        dLk[v] = dl;
      }
      zface(FzLo, k0);
      if (a.corr && k0 == 0) {
        corrected(bb, 4, i, j, FzLo);
        if (POWELL) corrected_bn(bb, 4, i, j, FzLo[NFZ - 2], FzLo[NFZ - 1]);
      }
#pragma unroll
      for (int v = 0; v < NVAR; v++) { Q1[v] = Q2[v]; Q2[v] = Q3[v]; }
    }

    for (int k = k0; k < k1; k++) {
      // z face k+1 of this column
      if (act) {
        load_prim(k + 2, Q3);
        zface(FzHi, k + 1);
        if (a.corr && k == NB - 1) {
          corrected(bb, 5, i, j, FzHi);
          if (POWELL) corrected_bn(bb, 5, i, j, FzHi[NFZ - 2], FzHi[NFZ - 1]);
        }
        // own primitive of plane k
#pragma unroll
        for (int v = 0; v < NVAR; v++) sP[sPi(lb, v, j + 2, i + 2)] = Q1[v];
      }
      // this thread's x/y halo cell of plane k
      if (hAct) {
        T q[NVAR];
        load_state(U, vs, cell_off<NB>(a.nei, hb, hii, hjj, k, a.nCellTot), q);
        to_primitive(q);
#pragma unroll
        for (int v = 0; v < NVAR; v++)
          sP[sPi(hlb, v, hjj + 2, hii + 2)] = q[v];
      }
      __syncthreads();
      // x and y faces of plane k
      // the wave with a third round of faces holds up the barrier below:
      // it issues first for that round (FP64 HLLD 1-1.5 % faster per
      // launch; Linde and FP32 lose about as much, so only there)
      constexpr bool PRIO3 = FLUX == FLUX_HLLD && sizeof(T) == 8;
      for (int f = tid; f < BPW * 2 * NFX; f += kThreads) {
        if (PRIO3 && f >= 2 * kThreads) __builtin_amdgcn_s_setprio(2);
        const int lbf = BPW == 1 ? 0 : f / (2 * NFX);
        const int bf = blk0 + lbf;
        if (bf >= a.nBlock) continue;
        const int r = f % (2 * NFX), dir = r / NFX, q = r % NFX;
        const int row = q / (NB + 1), fp = q % (NB + 1);
        T c0[NVAR], c1[NVAR], c2[NVAR], c3[NVAR];
        T F[NFLUX], vdt, unA;
        const double dxf = a.cellsize[3 * bf], dyf = a.cellsize[3 * bf + 1],
                     dzf = a.cellsize[3 * bf + 2];
        // y faces are evaluated as x faces of the cyclically rotated frame
        // (x, y, z) -> (y, z, x): the same normal / tangent1 / tangent2
        // order as the y-normal flux, so a wave whose faces mix x and y runs
        // one flux evaluation instead of both
        // the four cells of the stencil along the face normal: a row of the
        // plane for x faces, a column (stride NS) for y faces
        const int po = dir == 0 ? (row + 2) * NS + fp : fp * NS + row + 2;
        const int st = dir == 0 ? 1 : NS;
#pragma unroll
        for (int v = 0; v < NVAR; v++) {
          const int w = dir == 0 ? v : rot_src(v);
          const int o = (lbf * NVAR + w) * NR * NS + po;
          c0[v] = sP[o];
          c1[v] = sP[o + st];
          c2[v] = sP[o + 2 * st];
          c3[v] = sP[o + 3 * st];
        }
        const T area = T(dir == 0 ? dyf * dzf : dxf * dzf);
        T bL = beta, bR = beta;
        if (RESB) {
          const bool lo = a.nei[6 * bf + 2 * dir] < 0;
          const bool hi = a.nei[6 * bf + 2 * dir + 1] < 0;
          bL = cell_beta(beta, betaRC, nFaceRC, lo, hi, fp - 1, NB);
          bR = cell_beta(beta, betaRC, nFaceRC, lo, hi, fp, NB);
        }
        // the face's B0, in the rotated frame for a y face
        T B0r[3];
        if (B0S) {
          T B0f[3];
          if (dir == 0) b0face(0, bf, fp, row, k, B0f);
          else b0face(1, bf, row, fp, k, B0f);
#pragma unroll
          for (int d = 0; d < 3; d++) B0r[d] = dir == 0 ? B0f[d] : B0f[(d + 1) % 3];
        }
        bool skip = false;
        // TVD: the face at a coarser side takes its record's flux (in the
        // original frame), the face one cell in its fine-side state
        const T *trec = nullptr;
        bool recFlux = false;
        if (TVD && (fp <= 1 || fp >= NB - 1)) {
          trec = tvd_rec(bf, 2 * dir + (fp >= NB - 1), row, k);
          recFlux = trec && (fp == 0 || fp == NB);
        }
        if (TVD && trec && recFlux) {
          // the face at a coarser side: its record's flux (with the
          // body's treatment when the fine block reaches the body) and the
          // normal field of its face states for the 8-wave source
#pragma unroll
          for (int v = 0; v < NFLUX; v++) F[v] = trec[TVD_FLUX + v];
          vdt = trec[TVD_FLUX + PP];
          if (BNLDS) {
            sF[sFi(lbf, dir, NFS, q)] = trec[TVD_BN];
            sF[sFi(lbf, dir, NFS + 1, q)] = trec[TVD_BN + 1];
          }
        } else if ((TVD && trec) || BODY) {
          // the face one cell in from a coarser side takes its fine-side
          // state from the record; then the body's treatment (first order
          // next to body cells, the ghost state between a true and a body
          // cell, nothing between two body cells), in that order
          T L[NVAR], R[NVAR];
          face_states<LIM>(bL, bR, c0, c1, c2, c3, L, R);
          if (TVD && trec) {
            T *S = fp == 1 ? L : R;
#pragma unroll
            for (int v = 0; v < NVAR; v++)
              S[v] = trec[TVD_FINEF + (dir == 0 ? v : rot_src(v))];
          }
          const unsigned char *m = body_mask(bf);
          if (BODY && m) {
            // stencil flags and face centre in the original frame; the
            // face centre rotates with the y face
            bool u[4];
            for (int t = 0; t < 4; t++)
              u[t] = dir == 0 ? is_true(m, fp - 2 + t, row, k)
                              : is_true(m, row, fp - 2 + t, k);
            T r[3], rr[3];
            if (u[1] != u[2]) {
              if (dir == 0) face_xyz(bf, 0, fp - 1, row, k, r);
              else face_xyz(bf, 1, row, fp - 1, k, r);
#pragma unroll
              for (int d = 0; d < 3; d++) rr[d] = dir == 0 ? r[d] : r[(d + 1) % 3];
            }
            skip = body_face<LIM>(u[0], u[1], u[2], u[3], c1, c2, L, R, rr,
                                  T(a.bodyRho), T(a.b1rCoef));
          }
          if (skip) {
#pragma unroll
            for (int v = 0; v < NFLUX; v++) F[v] = T(0);
            vdt = T(0);
          } else {
            err |= numerical_flux<0, FLUX, false, B0S>(ph, area, L, R, F, vdt,
                                                       unA, B0r);
          }
          if (BNLDS) {
            sF[sFi(lbf, dir, NFS, q)] = skip ? T(0) : L[BX];
            sF[sFi(lbf, dir, NFS + 1, q)] = skip ? T(0) : R[BX];
          }
        } else {
          T Bn[2];
          if (LOWORD)
            err |= face_flux_w<0, FLUX, LIM, B0S>(
                ph, area, bL, bR,
                T(1) - crit[dir == 0 ? crit_index<NB>(bf, 0, fp, row, k)
                                     : crit_index<NB>(bf, 1, row, fp, k)],
                c0, c1, c2, c3, F, vdt, unA, B0r, BNLDS ? Bn : nullptr);
          else
            err |= face_flux<0, FLUX, LIM, B0S>(ph, area, bL, bR, c0, c1,
                                                c2, c3, F, vdt, unA, B0r,
                                                BNLDS ? Bn : nullptr);
          if (BNLDS) {
            sF[sFi(lbf, dir, NFS, q)] = Bn[0];
            sF[sFi(lbf, dir, NFS + 1, q)] = Bn[1];
          }
        }
        if (dir != 0 && !recFlux) unrotate_flux(F);
        F[PP] = vdt;
        if (a.corr && (fp == 0 || fp == NB))
          corrected(bf, 2 * dir + (fp == NB), row, k, F);
#pragma unroll
        for (int v = 0; v < NFS; v++) sF[sFi(lbf, dir, v, q)] = F[v];
      }
      if (PRIO3) __builtin_amdgcn_s_setprio(0);
      __syncthreads();
      if (act) {
        const int qx0 = j * (NB + 1) + i, qy0 = i * (NB + 1) + j;
        // flux divergence in T, handed to the update in double
        double dF[NFLUX];
#pragma unroll
        for (int v = 0; v < NFLUX; v++) {
          if (v == PP) { dF[v] = 0.0; continue; }
          T d = sF[sFi(lb, 0, v, qx0)] - sF[sFi(lb, 0, v, qx0 + 1)];
          d = d + sF[sFi(lb, 1, v, qy0)] - sF[sFi(lb, 1, v, qy0 + 1)];
          d = d + FzLo[v] - FzHi[v];
          dF[v] = double(d);
        }
        const long off = (((long)b * NB + k) * NB + j) * NB + i;
        // local time step of the cell (calc_timestep :157-168); in steady
        // state it is also this update's step (from this launch's fluxes
        // when doDt, else from the launch that set dtmaxCB), so it is taken
        // before the update there and after it otherwise
        double dtc = 0.0;
        // body cells are not updated and hold a zero local time step that
        // stays out of the block minimum (src/ModUpdateStateFast.f90:343-345,
        // src/ModTimeStepControl.f90:158-159,255)
        const bool trueCell = !BODY || is_true(Ub, i, j, k);
        auto cell_dt = [&]() {
          if (!trueCell) {
            dtc = 0.0;
            if (a.dtmaxCB) a.dtmaxCB[off] = 0.0;
            return;
          }
          const T Vdt =
              fmax2(sF[sFi(lb, 0, PP, qx0)], sF[sFi(lb, 0, PP, qx0 + 1)]) +
              fmax2(sF[sFi(lb, 1, PP, qy0)], sF[sFi(lb, 1, PP, qy0 + 1)]) +
              fmax2(FzLo[PP], FzHi[PP]);
          dtc = CellVolume / double(Vdt);
          if (a.dtmaxCB) a.dtmaxCB[off] = dtc;
          dtmin = fmin2(dtmin, dtc);
        };
        if (STEADY && a.doDt) cell_dt();
        UpdateCtl ucc = uc;
        if (STEADY && !a.noUpdate)
          ucc.DtLocal = a.DtFactor * (a.doDt ? dtc : a.dtmaxCB[off]);
        if (!a.noUpdate && !trueCell) {
          // the body cell keeps its stage-input state
#pragma unroll
          for (int v = 0; v < NVAR; v++) Uout[v * vs + off] = U[v * vs + off];
        } else if (!a.noUpdate) {
          T Uc[NVAR], Uo[NVAR];
          double Ucd[NVAR], Uod[NVAR], Un[NVAR];
          if (!POWELL) load_state(Uold, vs, off, Uo);
          if (POWELL) {
            // calc_source: the div B source from the normal-field face
            // states of the six faces (x and y from the face loop, or
            // recomputed from the LDS plane exactly as the face loop
            // computed them) and the stage-input cell state
            T p[5], LxLo, RxLo, LxHi, RxHi, LyLo, RyLo, LyHi, RyHi;
            if (BNLDS) {
              // the face states as the face loop formed them (and as the
              // body left them)
              LxLo = sF[sFi(lb, 0, NFS, qx0)];
              RxLo = sF[sFi(lb, 0, NFS + 1, qx0)];
              LxHi = sF[sFi(lb, 0, NFS, qx0 + 1)];
              RxHi = sF[sFi(lb, 0, NFS + 1, qx0 + 1)];
              LyLo = sF[sFi(lb, 1, NFS, qy0)];
              RyLo = sF[sFi(lb, 1, NFS + 1, qy0)];
              LyHi = sF[sFi(lb, 1, NFS, qy0 + 1)];
              RyHi = sF[sFi(lb, 1, NFS + 1, qy0 + 1)];
            } else {
              // betas of the cells i-1, i, i+1 (x) and j-1, j, j+1 (y)
              T bx[3] = {beta, beta, beta}, by[3] = {beta, beta, beta};
              if (RESB) {
                const int *n6 = a.nei + 6 * b;
                for (int d = 0; d < 3; d++) {
                  bx[d] = cell_beta(beta, betaRC, nFaceRC, n6[0] < 0,
                                    n6[1] < 0, i - 1 + d, NB);
                  by[d] = cell_beta(beta, betaRC, nFaceRC, n6[2] < 0,
                                    n6[3] < 0, j - 1 + d, NB);
                }
              }
#pragma unroll
              for (int d = 0; d < 5; d++) p[d] = sP[sPi(lb, BX, j + 2, i + d)];
              if (LOWORD) {
                face_lr_w<LIM>(bx[0], bx[1],
                               T(1) - crit[crit_index<NB>(b, 0, i, j, k)],
                               p[0], p[1], p[2], p[3], LxLo, RxLo);
                face_lr_w<LIM>(
                    bx[1], bx[2], T(1) - crit[crit_index<NB>(b, 0, i + 1, j, k)],
                    p[1], p[2], p[3], p[4], LxHi, RxHi);
              } else {
                face_lr<LIM>(bx[0], bx[1], p[0], p[1], p[2], p[3], LxLo, RxLo);
                face_lr<LIM>(bx[1], bx[2], p[1], p[2], p[3], p[4], LxHi, RxHi);
              }
#pragma unroll
              for (int d = 0; d < 5; d++) p[d] = sP[sPi(lb, BY, j + d, i + 2)];
              if (LOWORD) {
                face_lr_w<LIM>(by[0], by[1],
                               T(1) - crit[crit_index<NB>(b, 1, i, j, k)],
                               p[0], p[1], p[2], p[3], LyLo, RyLo);
                face_lr_w<LIM>(
                    by[1], by[2], T(1) - crit[crit_index<NB>(b, 1, i, j + 1, k)],
                    p[1], p[2], p[3], p[4], LyHi, RyHi);
              } else {
                face_lr<LIM>(by[0], by[1], p[0], p[1], p[2], p[3], LyLo, RyLo);
                face_lr<LIM>(by[1], by[2], p[1], p[2], p[3], p[4], LyHi, RyHi);
              }
              if (TVD) {
                // the TVD faces of coarser sides as the face loop took
                // them: the face at the side from its record, the face one
                // cell in with the record's fine-side state (x faces: the
                // record (j, k), y faces: (i, k))
                const T *rx0 = i <= 1 ? tvd_rec(b, 0, j, k) : nullptr;
                const T *rx1 = i >= NB - 2 ? tvd_rec(b, 1, j, k) : nullptr;
                const T *ry0 = j <= 1 ? tvd_rec(b, 2, i, k) : nullptr;
                const T *ry1 = j >= NB - 2 ? tvd_rec(b, 3, i, k) : nullptr;
                if (rx0 && i == 0) {
                  LxLo = rx0[TVD_BN];
                  RxLo = rx0[TVD_BN + 1];
                  LxHi = rx0[TVD_FINEF + BX];
                }
                if (rx0 && i == 1) LxLo = rx0[TVD_FINEF + BX];
                if (rx1 && i == NB - 1) {
                  LxHi = rx1[TVD_BN];
                  RxHi = rx1[TVD_BN + 1];
                  RxLo = rx1[TVD_FINEF + BX];
                }
                if (rx1 && i == NB - 2) RxHi = rx1[TVD_FINEF + BX];
                if (ry0 && j == 0) {
                  LyLo = ry0[TVD_BN];
                  RyLo = ry0[TVD_BN + 1];
                  LyHi = ry0[TVD_FINEF + BY];
                }
                if (ry0 && j == 1) LyLo = ry0[TVD_FINEF + BY];
                if (ry1 && j == NB - 1) {
                  LyHi = ry1[TVD_BN];
                  RyHi = ry1[TVD_BN + 1];
                  RyLo = ry1[TVD_FINEF + BY];
                }
                if (ry1 && j == NB - 2) RyHi = ry1[TVD_FINEF + BY];
              }
            }
            if (a.corr) {
              if (i == 0) corrected_bn(b, 0, j, k, LxLo, RxLo);
              if (i == NB - 1) corrected_bn(b, 1, j, k, LxHi, RxHi);
              if (j == 0) corrected_bn(b, 2, i, k, LyLo, RyLo);
              if (j == NB - 1) corrected_bn(b, 3, i, k, LyHi, RyHi);
            }
            T Src[NFLUX];
            // the B0 of the cell's faces (x-, x+, y-, y+, z-, z+) and the
            // cell's B0, div B0, curl B0, read as the source asks for them
            auto faceB0 = [&](int f, T *o) {
              const int d = f / 2, u = f % 2;
              b0face(d, b, i + (d == 0) * u, j + (d == 1) * u,
                     k + (d == 2) * u, o);
            };
            auto cellB0 = [&](int q) { return T(a.b0C[q * a.nCellTot + off]); };
            // the cell's stage-input state in primitive form: its entry of
            // the LDS plane (to_primitive of the stage input; the cell's
            // own entry, which no thread rewrites before this thread's next
            // plane)
            T Pc[NVAR];
#pragma unroll
            for (int v = 0; v < NVAR; v++) Pc[v] = sP[sPi(lb, v, j + 2, i + 2)];
            powell_source_with<B0S>(
                InvHalf, LxLo, RxLo, LxHi, RxHi, LyLo, RyLo, LyHi, RyHi,
                FzLo[NFZ - 2], FzLo[NFZ - 1], FzHi[NFZ - 2], FzHi[NFZ - 1],
                Pc, Src, faceB0, cellB0, a.useB0Src);
            // the old state only now, past the source's working set
            load_state(Uold, vs, off, Uo);
            double Srcd[NFLUX];
#pragma unroll
            for (int v = 0; v < NVAR; v++) Uod[v] = double(Uo[v]);
#pragma unroll
            for (int v = 0; v < NFLUX; v++) Srcd[v] = double(Src[v]);
            // the stage input itself only for the SSP-RK2 interpolation and
            // a zero time step
            if (uc.dtZero || !uc.halfStep) {
              load_state(U, vs, off, Uc);
#pragma unroll
              for (int v = 0; v < NVAR; v++) Ucd[v] = double(Uc[v]);
              update_cell(a.ph, ucc, dF, Ucd, Uod, Un, Srcd);
            } else {
              update_cell(a.ph, ucc, dF, Uod, Uod, Un, Srcd);
            }
          } else {
            // the stage input at the cell is used only by the SSP-RK2
            // interpolation and for a zero time step
#pragma unroll
            for (int v = 0; v < NVAR; v++) Uod[v] = double(Uo[v]);
            if (uc.dtZero || !uc.halfStep) {
              load_state(U, vs, off, Uc);
#pragma unroll
              for (int v = 0; v < NVAR; v++) Ucd[v] = double(Uc[v]);
              update_cell(a.ph, ucc, dF, Ucd, Uod, Un);
            } else {
              update_cell(a.ph, ucc, dF, Uod, Uod, Un);
            }
          }
#pragma unroll
          for (int v = 0; v < NVAR; v++) Uout[v * vs + off] = T(Un[v]);
        }
        if (!STEADY && a.doDt) cell_dt();
        // shift the z window
#pragma unroll
        for (int v = 0; v < NVAR; v++) { Q1[v] = Q2[v]; Q2[v] = Q3[v]; }
#pragma unroll
        for (int v = 0; v < NFZ; v++) FzLo[v] = FzHi[v];
      }
      // the div B source that recomputes the face states reads
      // neighbouring cells of the LDS plane, which the next plane
      // overwrites (the LDS face states are rewritten only after the next
      // plane's first barrier)
      if (POWELL && !BNLDS) __syncthreads();
    }

    if (a.doDt) {
      // block minimum of the local time step (calc_timestep :236-256)
      double m = wave_min(act ? dtmin : 1.7976931348623157e308);
      const int w = tid / 64;
      if ((tid & 63) == 0) sRed[w] = m;
      __syncthreads();
      if (act && col == 0) {
        constexpr int WPB = NC >= 64 ? NC / 64 : 1;  // waves per block
        double r = sRed[(lb * NC) / 64];
        for (int q = 1; q < WPB; q++) r = fmin2(r, sRed[(lb * NC) / 64 + q]);
        // the parts of a split block meet in an atomic minimum over the
        // bit patterns (an order-preserving map for non-negative doubles;
        // the launcher fills dtmaxB with DBL_MAX first)
        if (a.splitShift)
          atomicMin(reinterpret_cast<unsigned long long *>(a.dtmaxB + b),
                    (unsigned long long)__double_as_longlong(r));
        else
          a.dtmaxB[b] = r;
      }
    }
    __syncthreads();  // LDS is reused by the next group
  }  // group loop
  if (err) atomicOr(a.err, err);
  if (a.advanceTime && blockIdx.x == 0)
    for (int r = tid; r < a.nRun; r += kThreads) {
      double *rsr = a.scal + r * S_NSCAL;
      rsr[S_TSIM] = rsr[S_TSIM] + rsr[S_DT] / a.nStage;
    }
}

// ---------------------------------------------------------------------
// two-level grids: ghost slabs at resolution changes and the fine face
// fluxes that replace the coarse ones
// ---------------------------------------------------------------------
struct AmrArgs {
  const void *U;        // stage input (type T), slabs filled in place
  long vs, nCellTot;
  const int *nei;
  const double *cellsize;
  const int *slabs;     // 8 ints per slab (block, side, kind, n0..n3, target)
  int nSlab;
  const int *fineSlabs; // the kind-0 slabs
  int nFine;
  Phys ph;
  double beta;
  int lim;
  void *corr;           // type T, [slab][q][p][sub][NCORR]
  int *err;
  const void *lowCrit;  // low order face weights (type T) or null
  const double *bcState; // CellState_VI(8, 6) of the fixed outer sides
  double betaRC;        // sharper limiter next to sides with DiLevel_EB /= 0
  int nFaceRC;          // (0: off)
  double pOutflow;      // ghost pressure of outflow sides (kind-4 slabs)
  // TVD treatment at resolution changes (UseTvdResChange): first-order
  // prolongation, with the two coarse layers when twoLayers (second-order
  // scheme); face restriction and correct_monotone_restrict on the coarse
  // side; tvdBeta = min(BetaLimiterResChange, LimiterBeta)
  int tvd, twoLayers;
  double tvdBeta;
  // B0 at the outer sides' ghost cells (fixed/inflow/vary: B1 = CellState
  // B - B0), [slot][layer][q][p][3] with a slot per (block, side) or null
  const double *b0Side;
  const int *b0SideSlot;
  // time-varying inflow of the x+ side (kind-5 slabs): nVary table points
  // (times, primitive states), read at the slab block's run time
  const double *varyT, *varyS;
  int nVary;
  const double *scal;
  const int *runOfBlock;
  // B0 splitting: the face B0 of every block (as StageArgs::b0F) or null
  const double *b0F;
  int nBlock;
  // the inner body (as StageArgs: mask slot per block, masks, first cell
  // centres, ghost face state) or used = null
  const int *bodySlot;
  const unsigned char *used;
  const double *xyz0;
  double bodyRho, b1rCoef;
};

__device__ __forceinline__ void side_xyz(int d, int n, int p, int q, int &x,
                                         int &y, int &z) {
  if (d == 0) { x = n; y = p; z = q; }
  else if (d == 1) { x = p; y = n; z = q; }
  else { x = p; y = q; z = n; }
}

template <typename T>
__device__ __forceinline__ T minmod2(T a, T b) {
  if (a * b <= T(0)) return T(0);
  return fabsr(a) < fabsr(b) ? a : b;
}

// Ghost slabs of the resolution-change sides: the coarse ghost is the mean
// of the 2x2x2 fine cells it covers; the fine ghost is the coarse cell plus
// its slopes (minmod of the coarse block's one-sided differences, the one
// available difference at the block edge) times the +-1/4 offsets of the
// fine centre.  Kind-2 slabs (fixed outer sides) take the side's state,
// as set_boundary_for_cell does (src/ModUpdateStateFast.f90:1262-1268);
// kind-3 slabs (reflecting sides) mirror the block's first two cells with
// the normal momentum and field negated (set_symm_bc,
// src/ModCellBoundary.f90:308-331,624-666); kind-4 slabs (outflow sides
// with a fixed pressure) copy the boundary cell and set the pressure
// (src/ModCellBoundary.f90:273-277).  One thread per ghost cell.
template <typename T>
__global__ void __launch_bounds__(256) k_amr_ghosts(AmrArgs a, int NB) {
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const long per = 2L * NB * NB;
  if (t >= per * a.nSlab) return;
  const int slab = (int)(t / per), r = (int)(t % per);
  const int L = r / (NB * NB), q = (r / NB) % NB, p = r % NB;
  const int *sl = a.slabs + 8 * slab;
  const int side = sl[1], kind = sl[2], d = side / 2, hi = side % 2, H = NB / 2;
  const T *U = static_cast<const T *>(a.U);
  T val[NVAR];
  auto cell = [&](int blk, int x, int y, int z) {
    return (((long)blk * NB + z) * NB + y) * NB + x;
  };
  if (kind == 2 || kind == 5) {
    if (kind == 5 && a.nVary > 0) {
      // the inflow table at the run's time: V_i + w (V_i+1 - V_i), the end
      // values outside, then momentum (set_boundary_fast,
      // src/ModUpdateStateFast.f90:1151-1158)
      const double t =
          a.scal[(a.runOfBlock ? a.runOfBlock[sl[0]] : 0) * S_NSCAL + S_TSIM];
      const int n = a.nVary;
      double w[NVAR];
      if (t <= a.varyT[0] || n == 1) {
        for (int v = 0; v < NVAR; v++) w[v] = a.varyS[v];
      } else if (t >= a.varyT[n - 1]) {
        for (int v = 0; v < NVAR; v++) w[v] = a.varyS[(n - 1) * NVAR + v];
      } else {
        int lo = 0, hi2 = n - 1;  // varyT[lo] <= t < varyT[hi2]
        while (hi2 - lo > 1) {
          const int mid = (lo + hi2) / 2;
          if (a.varyT[mid] <= t) lo = mid; else hi2 = mid;
        }
        const double Dx = (t - a.varyT[lo]) / (a.varyT[lo + 1] - a.varyT[lo]);
        for (int v = 0; v < NVAR; v++)
          w[v] = a.varyS[lo * NVAR + v] +
                 Dx * (a.varyS[(lo + 1) * NVAR + v] - a.varyS[lo * NVAR + v]);
      }
      for (int v = MX; v <= MZ; v++) w[v] = w[v] * w[RHO];
      for (int v = 0; v < NVAR; v++) val[v] = T(w[v]);
    } else {
      for (int v = 0; v < NVAR; v++) val[v] = T(a.bcState[side * NVAR + v]);
    }
    if (a.b0SideSlot) {
      const int slot = a.b0SideSlot[6 * sl[0] + side];
      if (slot >= 0) {
        const double *b0 = a.b0Side + (((long)slot * 2 + L) * NB * NB +
                                       (long)q * NB + p) * 3;
        for (int c = 0; c < 3; c++)
          val[BX + c] = T(double(val[BX + c]) - b0[c]);
      }
    }
  } else if (kind == 4) {
    int x, y, z;
    side_xyz(d, hi ? NB - 1 : 0, p, q, x, y, z);
    const long o = cell(sl[0], x, y, z);
    for (int v = 0; v < NVAR; v++) val[v] = U[v * a.vs + o];
    val[PP] = T(a.pOutflow);
  } else if (kind == 3) {
    int x, y, z;
    side_xyz(d, hi ? NB - 1 - L : L, p, q, x, y, z);
    const long o = cell(sl[0], x, y, z);
    for (int v = 0; v < NVAR; v++) {
      const T u = U[v * a.vs + o];
      val[v] = (v == MX + d || v == BX + d) ? -u : u;
    }
  } else if (kind == 1 && a.tvd) {
    // face restriction: the 2x2 fine cells of the fine layer nearest to
    // the ghost layer (DoRestrictFace, src/ModMessagePass.f90:105); layer 0
    // moves to the coarse cell distance from the block's own cell in
    // primitive variables (correct_monotone_restrict,
    // src/ModFaceValue.f90:267-294,390-403).  Layer 1 enters no result: the
    // coarse face it serves takes the fine fluxes.
    const int F = sl[3 + (p / H) + 2 * (q / H)];
    const int nf = hi ? 2 * L : NB - 1 - 2 * L;
    for (int v = 0; v < NVAR; v++) val[v] = T(0);
    for (int bq = 0; bq < 2; bq++)
      for (int cp = 0; cp < 2; cp++) {
        int x, y, z;
        side_xyz(d, nf, 2 * (p % H) + cp, 2 * (q % H) + bq, x, y, z);
        const long o = cell(F, x, y, z);
        for (int v = 0; v < NVAR; v++) val[v] += U[v * a.vs + o];
      }
    for (int v = 0; v < NVAR; v++) val[v] = T(0.25) * val[v];
    if (L == 0 && a.twoLayers) {
      int x, y, z;
      side_xyz(d, hi ? NB - 1 : 0, p, q, x, y, z);
      T c[NVAR];
      load_state(U, a.vs, cell(sl[0], x, y, z), c);
      to_primitive(c);
      to_primitive(val);
      const T cThird = T(1.0 / 3.0);
      for (int v = 0; v < NVAR; v++) val[v] = val[v] + cThird * (val[v] - c[v]);
      val[RHO] = fmax2(val[RHO], T(1e-30));
      val[PP] = fmax2(val[PP], T(1e-30));
      for (int v = MX; v <= MZ; v++) val[v] = val[v] * val[RHO];
    }
  } else if (kind == 0 && a.tvd) {
    // first-order prolongation: the coarse cell itself; with the
    // second-order scheme the two ghost layers take the two coarse layers
    // behind the face (DoTwoCoarseLayers, src/ModMessagePass.f90:108-109)
    const int Lc = a.twoLayers ? L : 0;
    int x, y, z;
    side_xyz(d, hi ? Lc : NB - 1 - Lc, sl[4] + p / 2, sl[5] + q / 2, x, y, z);
    load_state(U, a.vs, cell(sl[3], x, y, z), val);
  } else if (kind == 1) {
    const int F = sl[3 + (p / H) + 2 * (q / H)];
    for (int v = 0; v < NVAR; v++) val[v] = T(0);
    for (int ai = 0; ai < 2; ai++)
      for (int bq = 0; bq < 2; bq++)
        for (int cp = 0; cp < 2; cp++) {
          const int nf = hi ? 2 * L + ai : NB - 1 - 2 * L - ai;
          int x, y, z;
          side_xyz(d, nf, 2 * (p % H) + cp, 2 * (q % H) + bq, x, y, z);
          const long o = cell(F, x, y, z);
          for (int v = 0; v < NVAR; v++) val[v] += U[v * a.vs + o];
        }
    for (int v = 0; v < NVAR; v++) val[v] = T(0.125) * val[v];
  } else {
    const int C = sl[3];
    int cc[3];
    side_xyz(d, hi ? 0 : NB - 1, sl[4] + p / 2, sl[5] + q / 2, cc[0], cc[1],
             cc[2]);
    const T dn = hi ? (L == 0 ? T(-0.25) : T(0.25)) : (L == 0 ? T(0.25) : T(-0.25));
    const T dp = (p % 2 == 0) ? T(-0.25) : T(0.25);
    const T dq = (q % 2 == 0) ? T(-0.25) : T(0.25);
    T delta[3];
    if (d == 0) { delta[0] = dn; delta[1] = dp; delta[2] = dq; }
    else if (d == 1) { delta[0] = dp; delta[1] = dn; delta[2] = dq; }
    else { delta[0] = dp; delta[1] = dq; delta[2] = dn; }
    const long o0 = cell(C, cc[0], cc[1], cc[2]);
    for (int v = 0; v < NVAR; v++) val[v] = U[v * a.vs + o0];
    for (int dd = 0; dd < 3; dd++) {
      int lo[3] = {cc[0], cc[1], cc[2]}, up[3] = {cc[0], cc[1], cc[2]};
      lo[dd] -= 1;
      up[dd] += 1;
      const bool hasLo = cc[dd] > 0, hasUp = cc[dd] < NB - 1;
      const long ol = hasLo ? cell(C, lo[0], lo[1], lo[2]) : o0;
      const long ou = hasUp ? cell(C, up[0], up[1], up[2]) : o0;
      for (int v = 0; v < NVAR; v++) {
        const T u0 = U[v * a.vs + o0];
        T g = T(0);
        if (hasLo && hasUp) g = minmod2(u0 - U[v * a.vs + ol], U[v * a.vs + ou] - u0);
        else if (hasLo) g = u0 - U[v * a.vs + ol];
        else if (hasUp) g = U[v * a.vs + ou] - u0;
        val[v] = val[v] + g * delta[dd];
      }
    }
  }
  T *W = const_cast<T *>(U);
  const long o = a.nCellTot + ((long)slab * 2 + L) * NB * NB + (long)q * NB + p;
  for (int v = 0; v < NVAR; v++) W[v * a.vs + o] = val[v];
}

// save_cons_flux (src/ModConserveFlux.f90:79-187): the face fluxes of the
// fine blocks at their coarser sides, into the coarse slab's store at
// (coarse face, sub-face) -- one thread per fine face
template <int NB, int FLUX, int LIM, typename T>
__global__ void __launch_bounds__(256) k_amr_fineflux(AmrArgs a) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.nFine * NB * NB) return;
  const int slab = a.fineSlabs[t / (NB * NB)];
  const int q = (t / NB) % NB, p = t % NB;
  const int *sl = a.slabs + 8 * slab;
  const int b = sl[0], side = sl[1], d = side / 2, hi = side % 2;
  const T *U = static_cast<const T *>(a.U);
  const PhysT<T> ph = phys_as<T>(a.ph);
  T c[4][NVAR];
  for (int s4 = 0; s4 < 4; s4++) {
    // face between normal cells f-1 and f; f = 0 (low side) or NB (high)
    const int nrm = (hi ? NB : 0) + s4 - 2;
    int x, y, z;
    side_xyz(d, nrm, p, q, x, y, z);
    load_state(U, a.vs, cell_off<NB>(a.nei, b, x, y, z, a.nCellTot), c[s4]);
    to_primitive(c[s4]);
  }
  const double *cz = a.cellsize + 3 * b;
  const T area = T(d == 0 ? cz[1] * cz[2] : d == 1 ? cz[0] * cz[2] : cz[0] * cz[1]);
  T L[NVAR], R[NVAR], F[NFLUX], vdt, unA;
  // the face's two cells, f-1 and f along the normal, with their betas
  // (the fine block's side here is a resolution change)
  const int f0 = hi ? NB : 0;
  T bL = T(a.beta), bR = T(a.beta);
  if (a.nFaceRC > 0) {
    const bool lo = a.nei[6 * b + 2 * d] < 0, up = a.nei[6 * b + 2 * d + 1] < 0;
    bL = cell_beta(T(a.beta), T(a.betaRC), a.nFaceRC, lo, up, f0 - 1, NB);
    bR = cell_beta(T(a.beta), T(a.betaRC), a.nFaceRC, lo, up, f0, NB);
  }
  T FF[NVAR];
  if (a.tvd && LIM != LIM_FIRST) {
    // get_face_tvd (src/ModFaceValue.f90:680-852): c holds C2, C1, F1, F2
    // (low side) or F2, F1, C1, C2 (high side) of this face; the group's
    // other three first-layer fine cells complete the 2x2 average
    const int f = (p % 2) + 2 * (q % 2);
    T F1[4][NVAR];
    for (int g = 0; g < 4; g++) {
      if (g == f) {
        for (int v = 0; v < NVAR; v++) F1[g][v] = c[hi ? 1 : 2][v];
        continue;
      }
      int x, y, z;
      side_xyz(d, hi ? NB - 1 : 0, (p & ~1) + g % 2, (q & ~1) + g / 2, x, y,
               z);
      load_state(U, a.vs, cell_off<NB>(a.nei, b, x, y, z, a.nCellTot),
                 F1[g]);
      to_primitive(F1[g]);
    }
    T CtoF[NVAR], FtoC[NVAR];
    tvd_subface(T(a.tvdBeta), c[hi ? 3 : 0], c[hi ? 2 : 1], F1,
                c[hi ? 0 : 3], f, CtoF, FtoC, FF);
    for (int v = 0; v < NVAR; v++) {
      L[v] = hi ? FtoC[v] : CtoF[v];
      R[v] = hi ? CtoF[v] : FtoC[v];
    }
  } else if (a.lowCrit) {
    int x, y, z;
    side_xyz(d, f0, p, q, x, y, z);
    const T w = T(1) - static_cast<const T *>(a.lowCrit)[crit_index<NB>(
                           b, d, x, y, z)];
    face_states_w<LIM>(bL, bR, w, c[0], c[1], c[2], c[3], L, R);
  } else {
    face_states<LIM>(bL, bR, c[0], c[1], c[2], c[3], L, R);
  }
  // the inner body: the fine block's face as the stage treats it (first
  // order next to body cells, the ghost face state between a true and a
  // body cell, no flux between two body cells); the record keeps the face
  // states' normal field either way
  bool skip = false;
  if (a.used && a.bodySlot[b] >= 0) {
    constexpr int NG = NB + 4;
    const unsigned char *m = a.used + (long)a.bodySlot[b] * NG * NG * NG;
    bool u[4];
    for (int s4 = 0; s4 < 4; s4++) {
      int x, y, z;
      side_xyz(d, (hi ? NB : 0) + s4 - 2, p, q, x, y, z);
      u[s4] = m[((z + 2) * NG + (y + 2)) * NG + (x + 2)] != 0;
    }
    T r[3] = {T(0), T(0), T(0)};
    if (u[1] != u[2]) {
      // centre of the face between cells f0-1 and f0 along the normal
      int x, y, z;
      side_xyz(d, f0 - 1, p, q, x, y, z);
      const int ci[3] = {x, y, z};
      for (int v = 0; v < 3; v++) {
        const double h = a.cellsize[3 * b + v], x0 = a.xyz0[3 * b + v];
        r[v] = T(0.5 * ((x0 + ci[v] * h) + (x0 + (ci[v] + (v == d)) * h)));
      }
    }
    skip = body_face<LIM>(u[0], u[1], u[2], u[3], c[1], c[2], L, R, r,
                          T(a.bodyRho), T(a.b1rCoef));
  }
  int err = 0;
  bool withB0 = false;
  if constexpr (sizeof(T) == 8) {
    if (a.b0F) {
      // the fine block's own face B0 (B0 splitting)
      withB0 = true;
      int x, y, z;
      side_xyz(d, f0, p, q, x, y, z);
      const long nFaceTot = (long)a.nBlock * (NB + 1) * NB * NB;
      const long fi = b0_face_index<NB>(b, d, x, y, z);
      T B0f[3];
      for (int v = 0; v < 3; v++) B0f[v] = a.b0F[(d * 3 + v) * nFaceTot + fi];
      if (d == 0) err = numerical_flux<0, FLUX, false, true>(ph, area, L, R, F, vdt, unA, B0f);
      else if (d == 1) err = numerical_flux<1, FLUX, false, true>(ph, area, L, R, F, vdt, unA, B0f);
      else err = numerical_flux<2, FLUX, false, true>(ph, area, L, R, F, vdt, unA, B0f);
    }
  }
  if (!withB0) {
    if (d == 0) err = numerical_flux<0, FLUX, false>(ph, area, L, R, F, vdt, unA);
    else if (d == 1) err = numerical_flux<1, FLUX, false>(ph, area, L, R, F, vdt, unA);
    else err = numerical_flux<2, FLUX, false>(ph, area, L, R, F, vdt, unA);
  }
  if (skip) {
    for (int v = 0; v < NFLUX; v++) F[v] = T(0);
    vdt = T(0);
    err = 0;
  }
  F[PP] = vdt;
  const int pc = sl[4] + p / 2, qc = sl[5] + q / 2, sub = (p % 2) + 2 * (q % 2);
  T *dst = static_cast<T *>(a.corr) +
           ((((long)sl[7]) * NB + qc) * NB + pc) * 4 * NCORR + sub * NCORR;
  for (int v = 0; v < 9; v++) dst[v] = F[v];
  // the normal field of the face states times FaceRatio = 1/4
  // (src/ModConserveFlux.f90:136-140)
  dst[CORR_BNL] = L[BX + d] * T(0.25);
  dst[CORR_BNR] = R[BX + d] * T(0.25);
  if (a.tvd && LIM != LIM_FIRST) {
    // the fine block's record: this face's flux and the fine value of the
    // face one cell in
    T *rec = static_cast<T *>(a.corr) +
             ((((long)slab) * NB + q) * NB + p) * 4 * NCORR;
    for (int v = 0; v < 9; v++) rec[TVD_FLUX + v] = F[v];
    for (int v = 0; v < NVAR; v++) rec[TVD_FINEF + v] = FF[v];
    // and the face states' normal field, for the 8-wave source
    rec[TVD_BN] = L[BX + d];
    rec[TVD_BN + 1] = R[BX + d];
  }
  if (err) atomicOr(a.err, err);
}

template <int NB, int FLUX, typename T>
inline hipError_t launch_amr_fineflux(const AmrArgs &a, hipStream_t stream) {
  const dim3 grid((a.nFine * NB * NB + 255) / 256), block(256);
  switch (a.lim) {
  case LIM_MINMOD: hipLaunchKernelGGL((k_amr_fineflux<NB, FLUX, LIM_MINMOD, T>), grid, block, 0, stream, a); break;
  case LIM_MC: hipLaunchKernelGGL((k_amr_fineflux<NB, FLUX, LIM_MC, T>), grid, block, 0, stream, a); break;
  case LIM_MC3: hipLaunchKernelGGL((k_amr_fineflux<NB, FLUX, LIM_MC3, T>), grid, block, 0, stream, a); break;
  case LIM_BETA: hipLaunchKernelGGL((k_amr_fineflux<NB, FLUX, LIM_BETA, T>), grid, block, 0, stream, a); break;
  case LIM_NO: hipLaunchKernelGGL((k_amr_fineflux<NB, FLUX, LIM_NO, T>), grid, block, 0, stream, a); break;
  default: hipLaunchKernelGGL((k_amr_fineflux<NB, FLUX, LIM_FIRST, T>), grid, block, 0, stream, a); break;
  }
  return hipGetLastError();
}

// Launch k_stage for one block size and flux; the limiter and the options
// (OPT_* bits) are dispatched here.
template <int NB, int FLUX, typename T>
inline hipError_t launch_stage_kernel(const StageArgs &a, int opt,
                                      hipStream_t stream) {
  const int nItem = a.nGroup << a.splitShift;
  const dim3 grid(nItem < a.nWorkgroup ? nItem : a.nWorkgroup),
      block(kThreads);
#define BATS_K(LM, O)                                                        \
  hipLaunchKernelGGL((k_stage<NB, FLUX, LM, O, T>), grid, block, 0, stream, a)
  // the limiters without beta never take OPT_RESBETA (no instances)
#define BATS_LAUNCH3(LM)                                                     \
  switch (opt & 11) {                                                        \
  case 0: BATS_K(LM, 0); break;                                              \
  case 1: BATS_K(LM, 1); break;                                              \
  case 2: BATS_K(LM, 2); break;                                              \
  case 3: BATS_K(LM, 3); break;                                              \
  case 8: BATS_K(LM, 8); break;                                              \
  case 9: BATS_K(LM, 9); break;                                              \
  case 10: BATS_K(LM, 10); break;                                            \
  default: BATS_K(LM, 11); break;                                            \
  }
#define BATS_LAUNCH7(LM)                                                     \
  switch (opt) {                                                             \
  case 0: BATS_K(LM, 0); break;                                              \
  case 1: BATS_K(LM, 1); break;                                              \
  case 2: BATS_K(LM, 2); break;                                              \
  case 3: BATS_K(LM, 3); break;                                              \
  case 4: BATS_K(LM, 4); break;                                              \
  case 5: BATS_K(LM, 5); break;                                              \
  case 6: BATS_K(LM, 6); break;                                              \
  case 7: BATS_K(LM, 7); break;                                              \
  case 8: BATS_K(LM, 8); break;                                              \
  case 9: BATS_K(LM, 9); break;                                              \
  case 10: BATS_K(LM, 10); break;                                            \
  case 11: BATS_K(LM, 11); break;                                            \
  case 12: BATS_K(LM, 12); break;                                            \
  case 13: BATS_K(LM, 13); break;                                            \
  case 14: BATS_K(LM, 14); break;                                            \
  default: BATS_K(LM, 15); break;                                            \
  }
  // B0 splitting (FP64): with or without the 8-wave source and the sharper
  // resolution-change limiter; no low order faces, no steady state
#define BATS_LAUNCH_B0(LM, RB)                                               \
  switch (opt & (RB ? 5 : 1)) {                                              \
  case 0: BATS_K(LM, 16); break;                                             \
  case 1: BATS_K(LM, 17); break;                                             \
  case 4: BATS_K(LM, 20); break;                                             \
  default: BATS_K(LM, 21); break;                                            \
  }
  // the TVD treatment at resolution changes (second-order limiters), with
  // or without the resolution-change beta
#define BATS_LAUNCH_TVD(LM, RB)                                              \
  switch (opt & ((RB ? OPT_RESBETA : 0) | OPT_POWELL)) {                     \
  case 0: BATS_K(LM, 64); break;                                             \
  case 1: BATS_K(LM, 65); break;                                             \
  case 4: BATS_K(LM, 68); break;                                             \
  default: BATS_K(LM, 69); break;                                            \
  }
  // and with B0 splitting (FP64)
#define BATS_LAUNCH_TVD_B0(LM, RB)                                           \
  switch (opt & ((RB ? OPT_RESBETA : 0) | OPT_POWELL)) {                     \
  case 0: BATS_K(LM, 80); break;                                             \
  case 1: BATS_K(LM, 81); break;                                             \
  case 4: BATS_K(LM, 84); break;                                             \
  default: BATS_K(LM, 85); break;                                            \
  }
  // and with the inner body (FP64; no resolution-change beta)
#define BATS_LAUNCH_TVD_BODY(LM)                                             \
  switch (opt & (OPT_POWELL | OPT_B0)) {                                     \
  case 0: BATS_K(LM, 96); break;                                             \
  case 1: BATS_K(LM, 97); break;                                             \
  case 16: BATS_K(LM, 112); break;                                           \
  default: BATS_K(LM, 113); break;                                           \
  }
  if (opt & OPT_TVD) {
    // steady state travels as the run-time flag
    if (opt & OPT_LOWORDER) return hipErrorInvalidValue;
    if (opt & OPT_BODY) {
      if constexpr (sizeof(T) == 8) {
        if (opt & OPT_RESBETA) return hipErrorInvalidValue;
        switch (a.lim) {
        case LIM_MINMOD: BATS_LAUNCH_TVD_BODY(LIM_MINMOD); break;
        case LIM_MC: BATS_LAUNCH_TVD_BODY(LIM_MC); break;
        case LIM_MC3: BATS_LAUNCH_TVD_BODY(LIM_MC3); break;
        case LIM_BETA: BATS_LAUNCH_TVD_BODY(LIM_BETA); break;
        case LIM_NO: BATS_LAUNCH_TVD_BODY(LIM_NO); break;
        default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
      } else {
        return hipErrorInvalidValue;
      }
    }
    if (opt & OPT_B0) {
      if constexpr (sizeof(T) == 8) {
        switch (a.lim) {
        case LIM_MINMOD: BATS_LAUNCH_TVD_B0(LIM_MINMOD, false); break;
        case LIM_MC: BATS_LAUNCH_TVD_B0(LIM_MC, true); break;
        case LIM_MC3: BATS_LAUNCH_TVD_B0(LIM_MC3, true); break;
        case LIM_BETA: BATS_LAUNCH_TVD_B0(LIM_BETA, true); break;
        case LIM_NO: BATS_LAUNCH_TVD_B0(LIM_NO, false); break;
        default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
      } else {
        return hipErrorInvalidValue;
      }
    }
    switch (a.lim) {
    case LIM_MINMOD: BATS_LAUNCH_TVD(LIM_MINMOD, false); break;
    case LIM_MC: BATS_LAUNCH_TVD(LIM_MC, true); break;
    case LIM_MC3: BATS_LAUNCH_TVD(LIM_MC3, true); break;
    case LIM_BETA: BATS_LAUNCH_TVD(LIM_BETA, true); break;
    case LIM_NO: BATS_LAUNCH_TVD(LIM_NO, false); break;
    default: return hipErrorInvalidValue;  // first order needs no record
    }
    return hipGetLastError();
  }
  // the inner body (FP64): with or without the 8-wave source and B0
#define BATS_LAUNCH_BODY(LM)                                                 \
  switch (opt & (OPT_POWELL | OPT_B0)) {                                     \
  case 0: BATS_K(LM, 32); break;                                             \
  case 1: BATS_K(LM, 33); break;                                             \
  case 16: BATS_K(LM, 48); break;                                            \
  default: BATS_K(LM, 49); break;                                            \
  }
  if constexpr (sizeof(T) == 8) {
    if (opt & OPT_BODY) {
      if (opt & (OPT_LOWORDER | OPT_RESBETA)) return hipErrorInvalidValue;
      switch (a.lim) {
      case LIM_MINMOD: BATS_LAUNCH_BODY(LIM_MINMOD); break;
      case LIM_MC: BATS_LAUNCH_BODY(LIM_MC); break;
      case LIM_MC3: BATS_LAUNCH_BODY(LIM_MC3); break;
      case LIM_BETA: BATS_LAUNCH_BODY(LIM_BETA); break;
      case LIM_NO: BATS_LAUNCH_BODY(LIM_NO); break;
      default: BATS_LAUNCH_BODY(LIM_FIRST); break;
      }
      return hipGetLastError();
    }
    if (opt & OPT_B0) {
      if (opt & OPT_LOWORDER) return hipErrorInvalidValue;
      switch (a.lim) {
      case LIM_MINMOD: BATS_LAUNCH_B0(LIM_MINMOD, false); break;
      case LIM_MC: BATS_LAUNCH_B0(LIM_MC, true); break;
      case LIM_MC3: BATS_LAUNCH_B0(LIM_MC3, true); break;
      case LIM_BETA: BATS_LAUNCH_B0(LIM_BETA, true); break;
      case LIM_NO: BATS_LAUNCH_B0(LIM_NO, false); break;
      default: BATS_LAUNCH_B0(LIM_FIRST, false); break;
      }
      return hipGetLastError();
    }
  } else {
    if (opt & (OPT_B0 | OPT_BODY)) return hipErrorInvalidValue;
  }
  switch (a.lim) {
  case LIM_MINMOD: BATS_LAUNCH3(LIM_MINMOD); break;
  case LIM_MC: BATS_LAUNCH7(LIM_MC); break;
  case LIM_MC3: BATS_LAUNCH7(LIM_MC3); break;
  case LIM_BETA: BATS_LAUNCH7(LIM_BETA); break;
  case LIM_NO: BATS_LAUNCH3(LIM_NO); break;
  default: BATS_LAUNCH3(LIM_FIRST); break;
  }
#undef BATS_K
#undef BATS_LAUNCH3
#undef BATS_LAUNCH7
#undef BATS_LAUNCH_B0
#undef BATS_LAUNCH_BODY
#undef BATS_LAUNCH_TVD
#undef BATS_LAUNCH_TVD_B0
#undef BATS_LAUNCH_TVD_BODY
  return hipGetLastError();
}

// one launcher per (block size, flux, storage type) translation unit
// (stage_inst.hip)
#define BATS_DECLARE_LAUNCH(NB, FL, TY)                                     \
  hipError_t launch_stage_##NB##_##FL##_##TY(const StageArgs &, int,        \
                                             hipStream_t);
#define BATS_DECLARE_LAUNCH_FLUX(NB, TY)                                    \
  BATS_DECLARE_LAUNCH(NB, rusanov, TY)                                      \
  BATS_DECLARE_LAUNCH(NB, linde, TY)                                        \
  BATS_DECLARE_LAUNCH(NB, hlld, TY)
#define BATS_DECLARE_FINE(NB, FL, TY)                                       \
  hipError_t launch_fineflux_##NB##_##FL##_##TY(const AmrArgs &, hipStream_t);
#define BATS_DECLARE_FINE_FLUX(NB, TY)                                      \
  BATS_DECLARE_FINE(NB, rusanov, TY)                                        \
  BATS_DECLARE_FINE(NB, linde, TY)                                          \
  BATS_DECLARE_FINE(NB, hlld, TY)
BATS_DECLARE_FINE_FLUX(8, f64)
BATS_DECLARE_FINE_FLUX(16, f64)
BATS_DECLARE_FINE_FLUX(8, f32)
BATS_DECLARE_FINE_FLUX(16, f32)
#undef BATS_DECLARE_FINE_FLUX
#undef BATS_DECLARE_FINE
BATS_DECLARE_LAUNCH_FLUX(8, f64)
BATS_DECLARE_LAUNCH_FLUX(16, f64)
BATS_DECLARE_LAUNCH_FLUX(8, f32)
BATS_DECLARE_LAUNCH_FLUX(16, f32)
#undef BATS_DECLARE_LAUNCH_FLUX
#undef BATS_DECLARE_LAUNCH

} // namespace bats
