// surrogate.hip -- the learned face-flux surrogate in the stage update, and
// the harvesting of its training samples from the plain solver.
//
// Face frame and scaling (shared by harvest and inference).  For a face
// normal to axis d the left and right primitive face states are rotated to
// (rho, u_n, u_t1, u_t2, B_n, B_t1, B_t2, p) with (n, t1, t2) = (d, d+1,
// d+2 mod 3), and scaled by the face's density rho_s = (rho_L + rho_R)/2
// and speed c_s = max(c_fL, c_fR) (normal fast magnetosonic speeds):
//   states  rho/rho_s, u/c_s, B/(c_s sqrt(rho_s)), p/(rho_s c_s^2)
//   fluxes  F_rho/(rho_s c_s), F_m/(rho_s c_s^2), F_B/(c_s^2 sqrt(rho_s)),
//           F_E/(rho_s c_s^3)
// The 16 inputs are the mean of the scaled L and R states and the unit
// direction of their difference j = s_R - s_L; the 8 outputs are the
// scaled flux correction divided by |j|.  (An invertible re-encoding of the
// scaled L/R pair: on smooth faces the jump is far below the BF16 input
// resolution of the raw states, and the correction is proportional to it.)
// The network learns the correction that takes the Rusanov flux of the face
// (src/ModFaceFlux.f90:2513-2559) to its HLLD flux with the Linde
// correction of the normal field (src/ModFaceFlux.f90:2826-3160, 2302-2320,
// 2403-2418): target = (F_HLLD - F_Rusanov) in the rotated frame and the
// output scaling above, per unit area.  In the stage, a network face gets
// F_Rusanov + the network's correction, so its error scales with the
// correction rather than with the flux.
//
// Smoothness gate: a face goes to the network when, for every variable and
// for both cells next to the face, either the MC limiter would leave the
// slope unclipped (the two one-sided differences agree in sign and their
// mean is within beta times the smaller one) or the jumps are small (the
// two one-sided differences, scaled like the network inputs, add up to at
// most JumpTol), and the plasma beta 2p/B^2 of both face states is at least
// BetaMin.  Every other face takes the exact flux.
//
// Surrogate stage (FP32 state):
//   k_sur_faces   face states, gate, Vdt; exact flux of the gated-out faces
//                 into the face array; compacted network input rows
//   k_sur_infer   the network on the rows (matrix cores), fluxes scattered
//                 back to the face array in the original frame, times area
//   k_sur_update  flux divergence, stage update in FP64, local time step
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "handle.hpp"
#include "mlp.hpp"
#include "mlp_host.hpp"

namespace {

constexpr int NSLOT = 9;  // face slots: flux 0..6, 7 <- Vdt, 8 energy

template <int NB> struct FaceGeom {
  static constexpr int NF = (NB + 1) * NB * NB;  // faces per direction
  // local face r of direction dir -> the 4-cell stencil start (i,j,k) of
  // cell "f-2" and the unit step
  __device__ static void decode(int dir, int r, int &fi, int &fj, int &fk) {
    if (dir == 0) {
      fi = r % (NB + 1); fj = (r / (NB + 1)) % NB; fk = r / ((NB + 1) * NB);
    } else if (dir == 1) {
      fi = r % NB; fj = (r / NB) % (NB + 1); fk = r / (NB * (NB + 1));
    } else {
      fi = r % NB; fj = (r / NB) % NB; fk = r / (NB * NB);
    }
  }
};

__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Faces whose scaled jump |j| is below this are not harvested: their target
// (F_HLLD - F_Rusanov)/|j| would carry the FP64 rounding of the two fluxes
// magnified by 1/|j|, and their correction is negligible anyway.
constexpr double kHarvestJumpMin = 1e-6;

template <typename R>
__device__ __forceinline__ bool var_smooth(R p0, R p1, R p2, R beta,
                                           R tol) {
  const R d1 = p1 - p0, d2 = p2 - p1;
  if (fabsr(d1) + fabsr(d2) <= tol) return true;
  return d1 * d2 > R(0) &&
         R(0.5) * fabsr(d1 + d2) <= beta * fmin2(fabsr(d1), fabsr(d2));
}

// tol: JumpTol in the network's scaling (rho_s, c_s) of each variable
template <typename R>
__device__ __forceinline__ bool face_smooth(const R *q0, const R *q1,
                                            const R *q2, const R *q3,
                                            const R *L, const R *Rs, R beta,
                                            R betaMin, R jumpTol, R rhos,
                                            R cs) {
  const R sRho = jumpTol * rhos, sU = jumpTol * cs,
          sB = jumpTol * cs * fsqrt(rhos), sP = jumpTol * rhos * cs * cs;
  bool ok = true;
#pragma unroll
  for (int v = 0; v < NVAR; v++) {
    const R tol = v == RHO ? sRho : v < BX ? sU : v < PP ? sB : sP;
    ok = ok && var_smooth(q0[v], q1[v], q2[v], beta, tol) &&
         var_smooth(q1[v], q2[v], q3[v], beta, tol);
  }
  const R b2L = L[BX] * L[BX] + L[BY] * L[BY] + L[BZ] * L[BZ];
  const R b2R = Rs[BX] * Rs[BX] + Rs[BY] * Rs[BY] + Rs[BZ] * Rs[BZ];
  ok = ok && R(2) * L[PP] >= betaMin * b2L && R(2) * Rs[PP] >= betaMin * b2R;
  return ok;
}

// rotated component index c of the 8-vector (rho, m_n, m_t1, m_t2, B_n,
// B_t1, B_t2, E) -> flux slot in the original frame
__device__ __forceinline__ int rot_slot(int dir, int c) {
  const int it1 = (dir + 1) % 3, it2 = (dir + 2) % 3;
  switch (c) {
  case 0: return RHO;
  case 1: return MX + dir;
  case 2: return MX + it1;
  case 3: return MX + it2;
  case 4: return BX + dir;
  case 5: return BX + it1;
  case 6: return BX + it2;
  default: return EN;
  }
}

// The 16 network inputs of a face: with s_L, s_R the scaled, rotated face
// states, the mean m = (s_L + s_R)/2 and the jump direction j/|j| of
// j = s_R - s_L (|j| the Euclidean norm, returned as jn; j/|j| = 0 when
// |j| = 0).  Returns false on a negative fast speed squared.
template <int DIR, typename R>
__device__ __forceinline__ bool face_inputs(const PhysT<R> &ph, const R *L,
                                            const R *Rs, float *x, R &rhos,
                                            R &cs, R &jn) {
  constexpr int it1 = (DIR + 1) % 3, it2 = (DIR + 2) % 3;
  R unL, cfL, unR, cfR;
  const bool okL = fast_speed<DIR>(ph, L, unL, cfL);   // both speeds are
  const bool okR = fast_speed<DIR>(ph, Rs, unR, cfR);  // needed either way
  const bool ok = okL && okR;
  rhos = R(0.5) * (L[RHO] + Rs[RHO]);
  cs = fmax2(cfL, cfR);
  const R iu = R(1) / cs, ib = iu / fsqrt(rhos), ip = iu * iu / rhos,
          ir = R(1) / rhos;
  R sv[2][8];
  const R *S[2] = {L, Rs};
#pragma unroll
  for (int s = 0; s < 2; s++) {
    sv[s][0] = S[s][RHO] * ir;
    sv[s][1] = S[s][MX + DIR] * iu;
    sv[s][2] = S[s][MX + it1] * iu;
    sv[s][3] = S[s][MX + it2] * iu;
    sv[s][4] = S[s][BX + DIR] * ib;
    sv[s][5] = S[s][BX + it1] * ib;
    sv[s][6] = S[s][BX + it2] * ib;
    sv[s][7] = S[s][PP] * ip;
  }
  R j2 = R(0);
#pragma unroll
  for (int c = 0; c < 8; c++) {
    const R d = sv[1][c] - sv[0][c];
    j2 += d * d;
  }
  jn = fsqrt(j2);
  const R ij = jn > R(0) ? R(1) / jn : R(0);
#pragma unroll
  for (int c = 0; c < 8; c++) {
    x[c] = float(R(0.5) * (sv[0][c] + sv[1][c]));
    x[8 + c] = float((sv[1][c] - sv[0][c]) * ij);
  }
  return ok;
}

template <typename R>
__device__ __forceinline__ R out_scale(int c, R rhos, R cs) {
  switch (c) {
  case 0: return rhos * cs;
  case 1: case 2: case 3: return rhos * cs * cs;
  case 4: case 5: case 6: return cs * cs * fsqrt(rhos);
  default: return rhos * cs * cs * cs;
  }
}

// ---------------------------------------------------------------------
// harvest (FP64 state): one workgroup per block, faces in chunks of 256,
// deterministic placement from per-block counts (pass 1) and a workgroup
// scan (pass 2)
// ---------------------------------------------------------------------
struct HarvestArgs {
  const double *U;
  const int *nei;
  const double *cellsize;
  long vs, nCellTot;
  int nBlock;
  Phys ph;
  double beta, betaMin, jumpTol, frac;
  unsigned long long seed;
  long *count;        // pass 1: selected faces per block
  const long *offset; // pass 2: exclusive prefix of count
  float *X, *T, *B;   // pass 2: sample rows (inputs, targets, baseline)
  float *J;           // pass 2: jump norm of each sample
};

template <int NB, int LIM, bool WRITE>
__global__ void __launch_bounds__(256) k_harvest(HarvestArgs a) {
  using G = FaceGeom<NB>;
  __shared__ int sScan[256];
  const int b = blockIdx.x;
  if (b >= a.nBlock) return;
  long base = WRITE ? a.offset[b] : 0;
  long cnt = 0;
  for (int c0 = 0; c0 < 3 * G::NF; c0 += 256) {
    const int f = c0 + threadIdx.x;
    bool sel = false;
    float x[16], t[8], bl[8], wj = 0.f;
    if (f < 3 * G::NF) {
      const int dir = f / G::NF, r = f % G::NF;
      const unsigned long long gid =
          ((unsigned long long)b * 3 + dir) * G::NF + r;
      const double u = (double)(mix64(a.seed ^ mix64(gid + 0x9E3779B97F4A7C15ull)) >> 11) *
                       (1.0 / 9007199254740992.0);
      if (u < a.frac) {
        int fi, fj, fk;
        G::decode(dir, r, fi, fj, fk);
        double q[4][NVAR];
        for (int s = 0; s < 4; s++) {
          int ii = fi, jj = fj, kk = fk;
          if (dir == 0) ii += s - 2; else if (dir == 1) jj += s - 2; else kk += s - 2;
          load_state(a.U, a.vs, cell_off<NB>(a.nei, b, ii, jj, kk, a.nCellTot), q[s]);
          to_primitive(q[s]);
        }
        double L[NVAR], R[NVAR], rhos, cs, jn;
        face_states<LIM>(a.beta, a.beta, q[0], q[1], q[2], q[3], L, R);
        bool ok;
        if (dir == 0) ok = face_inputs<0>(a.ph, L, R, x, rhos, cs, jn);
        else if (dir == 1) ok = face_inputs<1>(a.ph, L, R, x, rhos, cs, jn);
        else ok = face_inputs<2>(a.ph, L, R, x, rhos, cs, jn);
        if (ok && cs > 0.0 && jn >= kHarvestJumpMin &&
            face_smooth(q[0], q[1], q[2], q[3], L, R, a.beta, a.betaMin,
                        a.jumpTol, rhos, cs)) {
          double F[NFLUX], Fr[NFLUX], vdt, unA;
          if (dir == 0) {
            ok = !numerical_flux<0, FLUX_HLLD, false>(a.ph, 1.0, L, R, F, vdt, unA);
            ok = ok && !numerical_flux<0, FLUX_RUSANOV, false>(a.ph, 1.0, L, R, Fr, vdt, unA);
          } else if (dir == 1) {
            ok = !numerical_flux<1, FLUX_HLLD, false>(a.ph, 1.0, L, R, F, vdt, unA);
            ok = ok && !numerical_flux<1, FLUX_RUSANOV, false>(a.ph, 1.0, L, R, Fr, vdt, unA);
          } else {
            ok = !numerical_flux<2, FLUX_HLLD, false>(a.ph, 1.0, L, R, F, vdt, unA);
            ok = ok && !numerical_flux<2, FLUX_RUSANOV, false>(a.ph, 1.0, L, R, Fr, vdt, unA);
          }
          if (ok) {
            sel = true;
            for (int c = 0; c < 8; c++) {
              const double sc = out_scale(c, rhos, cs);
              const int sl = rot_slot(dir, c);
              t[c] = float((F[sl] - Fr[sl]) / (sc * jn));
              bl[c] = float(Fr[sl] / sc);
            }
            wj = float(jn);
          }
        }
      }
    }
    if (!WRITE) {
      cnt += sel ? 1 : 0;
      continue;
    }
    // workgroup exclusive scan of the selection flags
    sScan[threadIdx.x] = sel ? 1 : 0;
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {
      const int v = (int)threadIdx.x >= o ? sScan[threadIdx.x - o] : 0;
      __syncthreads();
      sScan[threadIdx.x] += v;
      __syncthreads();
    }
    const int incl = sScan[threadIdx.x], total = sScan[255];
    if (sel) {
      const long row = base + incl - 1;
      for (int c = 0; c < 16; c++) a.X[row * 16 + c] = x[c];
      for (int c = 0; c < 8; c++) a.T[row * 8 + c] = t[c];
      for (int c = 0; c < 8; c++) a.B[row * 8 + c] = bl[c];
      a.J[row] = wj;
    }
    base += total;
    __syncthreads();
  }
  if (!WRITE) {
    sScan[threadIdx.x] = (int)cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
      long s = 0;
      for (int q = 0; q < 256; q++) s += sScan[q];
      a.count[b] = s;
    }
  }
}

// ---------------------------------------------------------------------
// surrogate stage kernels (FP32 state)
// ---------------------------------------------------------------------
struct SurArgs {
  const float *Uin, *Uold;
  float *Uout;
  const int *nei;
  const double *cellsize;
  double *scal;
  const int *runOfBlock;
  int nRun;
  double *dtmaxB, *dtmaxCB;
  int *err;
  int nBlock;
  long vs, nCellTot;
  Phys ph;
  double beta, betaMin, jumpTol, DtFactor;
  int lim, mode, iStage, nStage, halfStep, dbTrick, doDt, roundTripOld,
      noUpdate, advanceTime;
  float *face;     // NSLOT x nFaceTot
  long nFaceTot;   // 3 * nBlock * NF
  __bf16 *rowX;
  int *rowFace;
  float *rowScale;
  int *counters;
};

template <int NB, int LIM>
__global__ void __launch_bounds__(256) k_sur_faces(SurArgs a) {
  using G = FaceGeom<NB>;
  const PhysT<float> ph = phys_as<float>(a.ph);
  const float beta = float(a.beta), betaMin = float(a.betaMin);
  int err = 0;
  long nNet = 0;
  const long nTot = a.nFaceTot;
  // per-wave row counts and the workgroup's base row, double-buffered by
  // iteration so a wave can start the next iteration while another still
  // reads this one's
  __shared__ int sCountBuf[2][5];
  int par = 0;
  for (long gf = blockIdx.x * (long)blockDim.x + threadIdx.x;
       gf - threadIdx.x < nTot; gf += (long)gridDim.x * blockDim.x, par ^= 1) {
    int *sCount = sCountBuf[par];
    const bool valid = gf < nTot;
    bool net = false;
    float x[16];
    float rhos = 0.f, cs = 0.f, jn = 0.f, area = 0.f;
    if (valid) {
      const int dir = (int)(gf / ((long)a.nBlock * G::NF));
      const long rem = gf % ((long)a.nBlock * G::NF);
      const int b = (int)(rem / G::NF), r = (int)(rem % G::NF);
      int fi, fj, fk;
      G::decode(dir, r, fi, fj, fk);
      float q[4][NVAR];
      for (int s = 0; s < 4; s++) {
        int ii = fi, jj = fj, kk = fk;
        if (dir == 0) ii += s - 2; else if (dir == 1) jj += s - 2; else kk += s - 2;
        load_state(a.Uin, a.vs, cell_off<NB>(a.nei, b, ii, jj, kk, a.nCellTot), q[s]);
        to_primitive(q[s]);
      }
      float L[NVAR], R[NVAR];
      face_states<LIM>(beta, beta, q[0], q[1], q[2], q[3], L, R);
      const double *cz = a.cellsize + 3 * b;
      area = float(dir == 0 ? cz[1] * cz[2]
                            : dir == 1 ? cz[0] * cz[2] : cz[0] * cz[1]);
      bool ok;
      if (dir == 0) ok = face_inputs<0>(ph, L, R, x, rhos, cs, jn);
      else if (dir == 1) ok = face_inputs<1>(ph, L, R, x, rhos, cs, jn);
      else ok = face_inputs<2>(ph, L, R, x, rhos, cs, jn);
      if (!ok) err |= ERRB_NEGFAST2;
      const bool smooth = ok && cs > 0.f &&
                          face_smooth(q[0], q[1], q[2], q[3], L, R, beta,
                                      betaMin, float(a.jumpTol), rhos, cs);
      float F[NFLUX], vdt, unA;
      if (smooth && a.mode == BATS_SUR_NETWORK) {
        // exact Vdt from the HLLD speeds; the flux comes from the network
        // HLLD's Cmax = max(sR, -sL) of the Linde-corrected states
        float Lc[NVAR], Rc[NVAR];
        for (int v = 0; v < NVAR; v++) { Lc[v] = L[v]; Rc[v] = R[v]; }
        const float dB = 0.5f * (Rc[BX + dir] - Lc[BX + dir]);
        Lc[BX + dir] += dB;
        Rc[BX + dir] -= dB;
        float uL, cL, uR, cR;
        if (dir == 0) { fast_speed<0>(ph, Lc, uL, cL); fast_speed<0>(ph, Rc, uR, cR); }
        else if (dir == 1) { fast_speed<1>(ph, Lc, uL, cL); fast_speed<1>(ph, Rc, uR, cR); }
        else { fast_speed<2>(ph, Lc, uL, cL); fast_speed<2>(ph, Rc, uR, cR); }
        const float sL = fmin2(uL - cL, uR - cR), sR = fmax2(uL + cL, uR + cR);
        // the Rusanov flux; the network adds its correction
        if (dir == 0) err |= numerical_flux<0, FLUX_RUSANOV, false>(ph, area, L, R, F, vdt, unA);
        else if (dir == 1) err |= numerical_flux<1, FLUX_RUSANOV, false>(ph, area, L, R, F, vdt, unA);
        else err |= numerical_flux<2, FLUX_RUSANOV, false>(ph, area, L, R, F, vdt, unA);
        F[PP] = fmax2(sR, -sL) * area;
        for (int v = 0; v < NSLOT; v++) a.face[v * nTot + gf] = F[v];
        net = jn > 0.f;  // no correction where the face states agree
      } else {
        if (smooth) {  // the method without the network: Rusanov
          if (dir == 0) err |= numerical_flux<0, FLUX_RUSANOV, false>(ph, area, L, R, F, vdt, unA);
          else if (dir == 1) err |= numerical_flux<1, FLUX_RUSANOV, false>(ph, area, L, R, F, vdt, unA);
          else err |= numerical_flux<2, FLUX_RUSANOV, false>(ph, area, L, R, F, vdt, unA);
        } else {
          if (dir == 0) err |= numerical_flux<0, FLUX_HLLD, false>(ph, area, L, R, F, vdt, unA);
          else if (dir == 1) err |= numerical_flux<1, FLUX_HLLD, false>(ph, area, L, R, F, vdt, unA);
          else err |= numerical_flux<2, FLUX_HLLD, false>(ph, area, L, R, F, vdt, unA);
        }
        F[PP] = vdt;
        for (int v = 0; v < NSLOT; v++) a.face[v * nTot + gf] = F[v];
      }
    }
    // workgroup-aggregated row allocation for the network faces: one
    // atomic per workgroup and iteration (the loop trip count is uniform
    // over the workgroup)
    const unsigned long long mask = __ballot(net);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) sCount[wave] = __popcll(mask);
    __syncthreads();
    if (threadIdx.x == 0) {
      int tot = 0;
      for (int w = 0; w < 4; w++) tot += sCount[w];
      sCount[4] = tot ? atomicAdd(&a.counters[0], tot) : 0;
    }
    __syncthreads();
    if (mask) {
      int base = sCount[4];
      for (int w = 0; w < wave; w++) base += sCount[w];
      if (net) {
        const int row = base + __popcll(mask & ((1ull << lane) - 1ull));
        // the row's inputs in BF16 (the network's input precision; two
        // 16-byte stores), then its scales
        bats::mlp::bf16x8 x0, x1;
#pragma unroll
        for (int c = 0; c < 8; c++) {
          x0[c] = (__bf16)x[c];
          x1[c] = (__bf16)x[8 + c];
        }
        bats::mlp::bf16x8 *xr =
            reinterpret_cast<bats::mlp::bf16x8 *>(a.rowX + (long)row * 16);
        xr[0] = x0;
        xr[1] = x1;
        a.rowFace[row] = (int)gf;
        reinterpret_cast<float4 *>(a.rowScale)[row] =
            make_float4(rhos, cs, jn, area);
      }
      nNet += net ? 1 : 0;
    }
  }
  if (err) atomicOr(a.err, err);
  (void)nNet;
}

// the network on the compacted rows; fluxes back to the face array
template <int NB>
__global__ void __launch_bounds__(bats::mlp::kInferThreads, 8)
k_sur_infer(const __bf16 *frag, const float *w, SurArgs a) {
  using namespace bats::mlp;
  using G = FaceGeom<NB>;
  __shared__ __attribute__((aligned(16))) __bf16 sW[NFRAG];
  __shared__ float sBias[NBIAS];
  for (int t = threadIdx.x; t < NFRAG / 8; t += kInferThreads)
    reinterpret_cast<bf16x8 *>(sW)[t] = reinterpret_cast<const bf16x8 *>(frag)[t];
  load_biases(w, sBias);
  __syncthreads();
  const int rows = a.counters[0];
  const int wave = threadIdx.x >> 6;
  const long nTot = a.nFaceTot;
  auto out = [&](int row, int c, float y) {
    const int gf = a.rowFace[row];
    const int dir = gf / (a.nBlock * G::NF);
    const float4 sc = reinterpret_cast<const float4 *>(a.rowScale)[row];
    const float area = sc.w;
    const float v = y * sc.z * out_scale(c, sc.x, sc.y);
    float &f = a.face[(long)rot_slot(dir, c) * nTot + gf];
    f = f + v * area;
  };
  const int nTile = (rows + 15) / 16;
  for (int tile = blockIdx.x * kInferWaves + wave; tile < nTile;
       tile += gridDim.x * kInferWaves)
    infer_tile(sW, sBias, a.rowX, rows, tile * 16, out);
}

// flux divergence and stage update (FP64), local time step; one workgroup
// per block
template <int NB>
__global__ void __launch_bounds__(256) k_sur_update(SurArgs a) {
  using G = FaceGeom<NB>;
  __shared__ double sRed[4];
  const int b = blockIdx.x;
  if (b >= a.nBlock) return;
  const long nTot = a.nFaceTot;
  const long dOff = (long)a.nBlock * G::NF;
  const double *cz = a.cellsize + 3 * b;
  const double CellVolume = cz[0] * cz[1] * cz[2];
  const double *rs = a.scal + (a.runOfBlock ? a.runOfBlock[b] : 0) * S_NSCAL;
  UpdateCtl uc;
  uc.DtLocal = a.DtFactor * rs[S_DTCB];
  uc.InvVolume = 1.0 / CellVolume;
  uc.halfStep = a.halfStep;
  uc.iStage = a.iStage;
  uc.nStage = a.nStage;
  uc.dbTrick = a.dbTrick;
  uc.dtZero = (rs[S_DT] == 0.0);
  uc.roundTripOld = a.roundTripOld;
  double dtmin = 1.7976931348623157e308;
  for (int c = threadIdx.x; c < NB * NB * NB; c += 256) {
    const int i = c % NB, j = (c / NB) % NB, k = c / (NB * NB);
    const long fx = 0 * dOff + (long)b * G::NF + (k * NB + j) * (NB + 1) + i;
    const long fy = 1 * dOff + (long)b * G::NF + (k * (NB + 1) + j) * NB + i;
    const long fz = 2 * dOff + (long)b * G::NF + (k * NB + j) * NB + i;
    double dF[NFLUX];
    for (int v = 0; v < NFLUX; v++) {
      if (v == PP) { dF[v] = 0.0; continue; }
      const float *Fv = a.face + (long)v * nTot;
      float d = Fv[fx] - Fv[fx + 1];
      d = d + Fv[fy] - Fv[fy + NB];
      d = d + Fv[fz] - Fv[fz + NB * NB];
      dF[v] = double(d);
    }
    const long off = (((long)b * NB + k) * NB + j) * NB + i;
    if (!a.noUpdate) {
      float Uc[NVAR], Uo[NVAR];
      double Ucd[NVAR], Uod[NVAR], Un[NVAR];
      load_state(a.Uold, a.vs, off, Uo);
      load_state(a.Uin, a.vs, off, Uc);
      for (int v = 0; v < NVAR; v++) { Ucd[v] = Uc[v]; Uod[v] = Uo[v]; }
      update_cell(a.ph, uc, dF, Ucd, Uod, Un);
      for (int v = 0; v < NVAR; v++) a.Uout[v * a.vs + off] = float(Un[v]);
    }
    if (a.doDt) {
      const float *Fv = a.face + (long)PP * nTot;
      const float Vdt = fmax2(Fv[fx], Fv[fx + 1]) + fmax2(Fv[fy], Fv[fy + NB]) +
                        fmax2(Fv[fz], Fv[fz + NB * NB]);
      const double dtc = CellVolume / double(Vdt);
      if (a.dtmaxCB) a.dtmaxCB[off] = dtc;
      dtmin = fmin2(dtmin, dtc);
    }
  }
  if (a.doDt) {
    double m = wave_min(dtmin);
    if ((threadIdx.x & 63) == 0) sRed[threadIdx.x / 64] = m;
    __syncthreads();
    if (threadIdx.x == 0)
      a.dtmaxB[b] = fmin2(fmin2(sRed[0], sRed[1]), fmin2(sRed[2], sRed[3]));
  }
  if (a.advanceTime && b == 0)
    for (int r = threadIdx.x; r < a.nRun; r += 256) {
      double *rr = a.scal + r * S_NSCAL;
      rr[S_TSIM] = rr[S_TSIM] + rr[S_DT] / a.nStage;
    }
}

// two-level grids: the coarse faces of the finer sides take the sum of the
// four fine faces' fluxes as the fine blocks use them (network or exact),
// summed in sub-face order -- one thread per coarse face
template <int NB>
__global__ void __launch_bounds__(256)
k_sur_amr_correct(float *face, long nTot, int nBlock, const int *slabs,
                  int nSlab) {
  using G = FaceGeom<NB>;
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= (long)nSlab * NB * NB) return;
  const int slab = (int)(t / (NB * NB)), q = (int)((t / NB) % NB),
            p = (int)(t % NB);
  const int *sl = slabs + 8 * slab;
  if (sl[2] != 1) return;
  const int b = sl[0], side = sl[1], d = side / 2, H = NB / 2;
  const int F = sl[3 + (p / H) + 2 * (q / H)];
  auto gface = [&](int blk, int fn, int fp, int fq) {
    long r;
    if (d == 0) r = ((long)fq * NB + fp) * (NB + 1) + fn;
    else if (d == 1) r = ((long)fq * (NB + 1) + fn) * NB + fp;
    else r = ((long)fn * NB + fq) * NB + fp;
    return (long)d * nBlock * G::NF + (long)blk * G::NF + r;
  };
  const long gc = gface(b, (side % 2) ? NB : 0, p, q);
  const int fn = (side % 2) ? 0 : NB;  // the fine block's face on its other side
  long gf[4];
  for (int sub = 0; sub < 4; sub++)
    gf[sub] = gface(F, fn, 2 * (p % H) + (sub & 1), 2 * (q % H) + (sub >> 1));
  for (int v = 0; v < NSLOT; v++) {
    const float *Fv = face + (long)v * nTot;
    face[(long)v * nTot + gc] = ((Fv[gf[0]] + Fv[gf[1]]) + Fv[gf[2]]) + Fv[gf[3]];
  }
}

// counters: [0] rows of the current stage, [2..3] running total of faces
// and [4..5] of network rows (64-bit), [1] unused
constexpr int NCOUNTER = 6;
__global__ void k_count_faces(int *counters, long nFace) {
  // rows of this stage -> running totals (single thread)
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    atomicAdd(reinterpret_cast<unsigned long long *>(counters + 4),
              (unsigned long long)counters[0]);
    atomicAdd(reinterpret_cast<unsigned long long *>(counters + 2),
              (unsigned long long)nFace);
    counters[0] = 0;
  }
}

}  // namespace

// =====================================================================
// host side
// =====================================================================
template <int NB, bool W>
static void launch_harvest(int lim, int nB, hipStream_t st, const HarvestArgs &a) {
  const dim3 g(nB), t(256);
  switch (lim) {
  case LIM_MINMOD: hipLaunchKernelGGL((k_harvest<NB, LIM_MINMOD, W>), g, t, 0, st, a); break;
  case LIM_MC: hipLaunchKernelGGL((k_harvest<NB, LIM_MC, W>), g, t, 0, st, a); break;
  case LIM_MC3: hipLaunchKernelGGL((k_harvest<NB, LIM_MC3, W>), g, t, 0, st, a); break;
  case LIM_BETA: hipLaunchKernelGGL((k_harvest<NB, LIM_BETA, W>), g, t, 0, st, a); break;
  case LIM_NO: hipLaunchKernelGGL((k_harvest<NB, LIM_NO, W>), g, t, 0, st, a); break;
  default: hipLaunchKernelGGL((k_harvest<NB, LIM_FIRST, W>), g, t, 0, st, a); break;
  }
}

template <int NB>
static void launch_sur_faces(int lim, dim3 g, hipStream_t st, const SurArgs &a) {
  const dim3 t(256);
  switch (lim) {
  case LIM_MINMOD: hipLaunchKernelGGL((k_sur_faces<NB, LIM_MINMOD>), g, t, 0, st, a); break;
  case LIM_MC: hipLaunchKernelGGL((k_sur_faces<NB, LIM_MC>), g, t, 0, st, a); break;
  case LIM_MC3: hipLaunchKernelGGL((k_sur_faces<NB, LIM_MC3>), g, t, 0, st, a); break;
  case LIM_BETA: hipLaunchKernelGGL((k_sur_faces<NB, LIM_BETA>), g, t, 0, st, a); break;
  case LIM_NO: hipLaunchKernelGGL((k_sur_faces<NB, LIM_NO>), g, t, 0, st, a); break;
  default: hipLaunchKernelGGL((k_sur_faces<NB, LIM_FIRST>), g, t, 0, st, a); break;
  }
}

extern "C" int bats_mlp_harvest(bats_mlp *m, bats_handle *h, double Fraction,
                                double BetaMin, double JumpTol,
                                unsigned long long Seed, long *nAdded) {
  if (!m || !h || !(Fraction >= 0.0) || Fraction > 1.0 || !(BetaMin >= 0.0) ||
      !(JumpTol >= 0.0))
    return BATS_EINVAL;
  if (h->elemSize != 8) return BATS_EINVAL;  // samples come from FP64 runs
  const bats_config &c = h->cfg;
  const int nB = c.nBlock, lim = limiter_code(c);
  HarvestArgs a;
  a.U = static_cast<const double *>(h->U[h->stateBuf]);
  a.nei = h->nei;
  a.cellsize = h->cellsize;
  a.vs = h->vs;
  a.nCellTot = h->nCell;
  a.nBlock = nB;
  a.ph = make_phys(c);
  a.beta = c.LimiterBeta;
  a.betaMin = BetaMin;
  a.jumpTol = JumpTol;
  a.frac = Fraction;
  a.seed = Seed;
  a.X = a.T = a.B = a.J = nullptr;
  long *dCount = nullptr, *dOff = nullptr;
  HIPCHK(hipMalloc(&dCount, sizeof(long) * nB));
  HIPCHK(hipMalloc(&dOff, sizeof(long) * nB));
  a.count = dCount;
  a.offset = dOff;
  if (h->NB == 8) launch_harvest<8, false>(lim, nB, h->stream, a);
  else launch_harvest<16, false>(lim, nB, h->stream, a);
  HIPCHK(hipGetLastError());
  std::vector<long> cnt(nB), off(nB);
  HIPCHK(hipMemcpyAsync(cnt.data(), dCount, sizeof(long) * nB,
                        hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  long total = 0;
  for (int b = 0; b < nB; b++) { off[b] = total; total += cnt[b]; }
  int st = BATS_OK;
  if (total > 0) {
    float *X = nullptr, *T = nullptr, *Bl = nullptr, *Jn = nullptr;
    st = bats::mlp::reserve_data(m, total, &X, &T, &Bl, &Jn);
    if (st == BATS_OK) {
      HIPCHK(hipMemcpyAsync(dOff, off.data(), sizeof(long) * nB,
                            hipMemcpyHostToDevice, h->stream));
      a.X = X;
      a.T = T;
      a.B = Bl;
      a.J = Jn;
      if (h->NB == 8) launch_harvest<8, true>(lim, nB, h->stream, a);
      else launch_harvest<16, true>(lim, nB, h->stream, a);
      HIPCHK(hipGetLastError());
      HIPCHK(hipStreamSynchronize(h->stream));
      bats::mlp::commit_data(m, total);
    }
  }
  (void)hipFree(dCount);
  (void)hipFree(dOff);
  if (nAdded) *nAdded = st == BATS_OK ? total : 0;
  return st;
}

extern "C" int bats_set_surrogate(bats_handle *h, bats_mlp *m, int Mode,
                                  double BetaMin, double JumpTol) {
  if (!h || Mode < BATS_SUR_OFF || Mode > BATS_SUR_RUSANOV ||
      !(BetaMin >= 0.0) || !(JumpTol >= 0.0))
    return BATS_EINVAL;
  if (Mode == BATS_SUR_NETWORK && !m) return BATS_EINVAL;
  drop_graphs(h);
  if (Mode != BATS_SUR_OFF && h->elemSize != 4) return BATS_EINVAL;
  // the surrogate stage has no 8-wave div B source
  if (Mode != BATS_SUR_OFF && h->cfg.UseDivbSource) return BATS_EINVAL;
  // the gated face pass carries no adaptive low order weights
  if (Mode != BATS_SUR_OFF && h->cfg.UseAdaptiveLowOrder &&
      h->cfg.nLowOrder == 1 && h->cfg.nOrder == 2)
    return BATS_EINVAL;
  SurrogateState &s = h->sur;
  if (Mode != BATS_SUR_OFF && !s.face) {
    const long nF = (long)(h->NB + 1) * h->NB * h->NB;
    const long nTot = 3 * (long)h->cfg.nBlock * nF;
    // rows and face indices are 32-bit (about 160k blocks of 16^3)
    if (nTot > 0x7fffffffL) return BATS_ELIMIT;
    HIPCHK(hipMalloc(&s.face, sizeof(float) * NSLOT * nTot));
    HIPCHK(hipMalloc(&s.rowX, sizeof(__bf16) * 16 * nTot));
    HIPCHK(hipMalloc(&s.rowFace, sizeof(int) * nTot));
    HIPCHK(hipMalloc(&s.rowScale, sizeof(float) * 4 * nTot));
    HIPCHK(hipMalloc(&s.counters, sizeof(int) * NCOUNTER));
    s.faceCap = nTot;
  }
  if (s.counters) {
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipMemset(s.counters, 0, sizeof(int) * NCOUNTER));
  }
  s.model = m;
  s.mode = Mode;
  s.betaMin = BetaMin;
  s.jumpTol = JumpTol;
  return BATS_OK;
}

extern "C" int bats_surrogate_counts(bats_handle *h, long *nFace, long *nNetwork) {
  if (!h) return BATS_EINVAL;
  int c[NCOUNTER] = {0, 0, 0, 0, 0, 0};
  if (h->sur.counters) {
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipMemcpy(c, h->sur.counters, sizeof c, hipMemcpyDeviceToHost));
  }
  unsigned long long f = 0, n = 0;
  memcpy(&f, c + 2, sizeof f);
  memcpy(&n, c + 4, sizeof n);
  if (nFace) *nFace = (long)f;
  if (nNetwork) *nNetwork = (long)n;
  return BATS_OK;
}

// one stage through the surrogate kernels (called by the stage launcher)
int sur_stage(bats_handle *h, const StageArgs &sa) {
  SurrogateState &s = h->sur;
  const bats_config &c = h->cfg;
  SurArgs a;
  a.Uin = static_cast<const float *>(sa.Uin);
  a.Uold = static_cast<const float *>(sa.Uold);
  a.Uout = static_cast<float *>(sa.Uout);
  a.nei = sa.nei;
  a.cellsize = sa.cellsize;
  a.scal = sa.scal;
  a.runOfBlock = sa.runOfBlock;
  a.nRun = sa.nRun;
  a.dtmaxB = sa.dtmaxB;
  a.dtmaxCB = sa.dtmaxCB;
  a.err = sa.err;
  a.nBlock = sa.nBlock;
  a.vs = sa.vs;
  a.nCellTot = sa.nCellTot;
  a.ph = sa.ph;
  a.beta = sa.beta;
  a.betaMin = s.betaMin;
  a.jumpTol = s.jumpTol;
  a.DtFactor = sa.DtFactor;
  a.lim = sa.lim;
  a.mode = s.mode;
  a.iStage = sa.iStage;
  a.nStage = sa.nStage;
  a.halfStep = sa.halfStep;
  a.dbTrick = sa.dbTrick;
  a.doDt = sa.doDt;
  a.roundTripOld = sa.roundTripOld;
  a.noUpdate = sa.noUpdate;
  a.advanceTime = sa.advanceTime;
  a.face = s.face;
  a.nFaceTot = s.faceCap;
  a.rowX = s.rowX;
  a.rowFace = s.rowFace;
  a.rowScale = s.rowScale;
  a.counters = s.counters;
  const long nTot = s.faceCap;
  long nWg = (nTot + 255) / 256;
  if (nWg > 16384) nWg = 16384;
  const int lim = sa.lim;
  if (h->NB == 8) launch_sur_faces<8>(lim, dim3((unsigned)nWg), h->stream, a);
  else launch_sur_faces<16>(lim, dim3((unsigned)nWg), h->stream, a);
  if (s.mode == BATS_SUR_NETWORK) {
    const bats::mlp::Weights wts = bats::mlp::weights(s.model);
    if (h->NB == 8)
      hipLaunchKernelGGL(k_sur_infer<8>, dim3(bats::mlp::kInferGrid), dim3(bats::mlp::kInferThreads), 0, h->stream,
                         wts.frag, wts.w, a);
    else
      hipLaunchKernelGGL(k_sur_infer<16>, dim3(bats::mlp::kInferGrid), dim3(bats::mlp::kInferThreads), 0, h->stream,
                         wts.frag, wts.w, a);
  }
  if (h->nSlab > 0) {
    const long n = (long)h->nSlab * h->NB * h->NB;
    const dim3 g((unsigned)((n + 255) / 256)), t(256);
    if (h->NB == 8)
      hipLaunchKernelGGL(k_sur_amr_correct<8>, g, t, 0, h->stream, s.face,
                         nTot, c.nBlock, h->slabs, h->nSlab);
    else
      hipLaunchKernelGGL(k_sur_amr_correct<16>, g, t, 0, h->stream, s.face,
                         nTot, c.nBlock, h->slabs, h->nSlab);
  }
  if (h->NB == 8)
    hipLaunchKernelGGL(k_sur_update<8>, dim3(c.nBlock), dim3(256), 0, h->stream, a);
  else
    hipLaunchKernelGGL(k_sur_update<16>, dim3(c.nBlock), dim3(256), 0, h->stream, a);
  hipLaunchKernelGGL(k_count_faces, dim3(1), dim3(64), 0, h->stream, s.counters,
                     nTot);
  HIPCHK(hipGetLastError());
  return BATS_OK;
}
