// mhd_device.hpp -- device arithmetic of the ideal-MHD stage update
// (limited MUSCL face states, Rusanov / Linde / HLLD face fluxes, wave
// speeds, energy <-> pressure, stage update) for gfx950.
//
// The functions follow the reference's formulas and operand order:
//   limiter            src/ModFaceValue.f90:3689-3791
//   face states        src/ModFaceValue.f90:1970-2163
//   physical flux      src/ModFaceFlux.f90:1166-1410,1550-1617,1619-1899
//   wave speeds        src/ModFaceFlux.f90:3545-3722,3853-4252
//   Rusanov            src/ModFaceFlux.f90:2513-2559
//   Linde (HLLE+fix)   src/ModFaceFlux.f90:2302-2320,2403-2418,2561-2645
//   HLLD               src/ModFaceFlux.f90:2826-3160
//   energy/pressure    src/ModEnergy.f90:47-127,243-332,388-451
//   stage update       src/ModUpdateState.f90:191-965
// The face arithmetic (limiter, face states, fluxes, speeds) is templated
// on its real type: double for the plain path, float for the FP32-state
// path; the stage update itself (energy <-> pressure, U = Uold + dt*R,
// pressure floor) always runs in double.
//
// Faces are Cartesian: the unit normal's zero components only add signed
// zeros in the reference, so they are dropped; sums the reference forms in
// x,y,z order (|B|^2, |u|^2, Poynting term) keep that order, sums it forms
// in the rotated n,t1,t2 frame (HLLD) keep that order.
//
// Division and square root use the hardware reciprocal / reciprocal square
// root with two Newton-Raphson (Goldschmidt) refinements instead of the
// IEEE-exact sequences; each result is within about one ulp of the
// correctly rounded one, far inside the 1e-12 parity tolerance, and the
// inputs here never approach the denormal or overflow range the exact
// sequences guard.
#pragma once
#include <hip/hip_runtime.h>

namespace bats {

enum { RHO = 0, MX = 1, MY = 2, MZ = 3, BX = 4, BY = 5, BZ = 6, PP = 7,
       NVAR = 8, EN = 8, NFLUX = 9 };
enum { FLUX_RUSANOV = 1, FLUX_LINDE = 2, FLUX_HLLD = 3 };
enum { LIM_FIRST = 0, LIM_MINMOD = 1, LIM_MC = 2, LIM_MC3 = 3, LIM_BETA = 4,
       LIM_NO = 5 };
// device error bits
enum { ERRB_NEGFAST2 = 1, ERRB_HLLD = 2, ERRB_NAN = 4 };

// physical constants in the arithmetic type of the face computation
template <typename Real> struct PhysT {
  Real gamma, gm1, invgm1, cTiny, pMin;
};
using Phys = PhysT<double>;
template <typename Real>
__device__ __forceinline__ PhysT<Real> phys_as(const Phys &p) {
  return PhysT<Real>{Real(p.gamma), Real(p.gm1), Real(p.invgm1),
                     Real(p.cTiny), Real(p.pMin)};
}

// Fortran SIGN(a, b) for a > 0
__device__ __forceinline__ double fsign(double a, double b) {
  return __builtin_copysign(a, b);
}
__device__ __forceinline__ float fsign(float a, float b) {
  return __builtin_copysignf(a, b);
}
__device__ __forceinline__ double fabsr(double a) { return __builtin_fabs(a); }
__device__ __forceinline__ float fabsr(float a) { return __builtin_fabsf(a); }
// Fortran MIN / MAX of two (finite) arguments
__device__ __forceinline__ double fmin2(double a, double b) {
  return __builtin_fmin(a, b);
}
__device__ __forceinline__ double fmax2(double a, double b) {
  return __builtin_fmax(a, b);
}
__device__ __forceinline__ float fmin2(float a, float b) {
  return __builtin_fminf(a, b);
}
__device__ __forceinline__ float fmax2(float a, float b) {
  return __builtin_fmaxf(a, b);
}
// 1/x: v_rcp_f64 + two Newton-Raphson steps
__device__ __forceinline__ double frcp(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = __builtin_fma(-x, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-x, r, 1.0);
  return __builtin_fma(r, e, r);
}
// 1/x in FP32: v_rcp_f32 + one Newton-Raphson step
__device__ __forceinline__ float frcp(float x) {
  const float r = __builtin_amdgcn_rcpf(x);
  const float e = __builtin_fmaf(-x, r, 1.0f);
  return __builtin_fmaf(r, e, r);
}
template <typename Real> __device__ __forceinline__ Real fdiv(Real a, Real b) {
  return a * frcp(b);
}
// sqrt(x), x >= 0: v_rsq_f64 + Goldschmidt refinement
__device__ __forceinline__ double fsqrt(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y, h = 0.5 * y;
  const double r = __builtin_fma(-g, h, 0.5);
  g = __builtin_fma(g, r, g);
  h = __builtin_fma(h, r, h);
  const double d = __builtin_fma(-g, g, x);
  g = __builtin_fma(d, h, g);
  return x == 0.0 ? 0.0 : g;
}
// sqrt(x) in FP32: v_sqrt_f32
__device__ __forceinline__ float fsqrt(float x) {
  return __builtin_amdgcn_sqrtf(x);
}

// ---------------------------------------------------------------------
// limiter (half slopes, src/ModFaceValue.f90:3689-3791): slopes of the
// cell l from p0=P(l-1), p1=P(l), p2=P(l+1).  dL builds the left state of
// the face to the right of the cell, dR the right state of the face to its
// left.
template <int LIM, typename Real>
__device__ __forceinline__ void slopes(Real beta, Real p0, Real p1,
                                       Real p2, Real &dL, Real &dR) {
  const Real cThird = Real(1.) / Real(3.);
  const Real d1 = p1 - p0, d2 = p2 - p1;
  if (LIM == LIM_MC3) {
    const Real sg = fsign(Real(0.25), d1) + fsign(Real(0.25), d2);
    const Real m = fmin2(beta * fabsr(d1), beta * fabsr(d2));
    dR = sg * fmin2(m, cThird * fabsr(2 * d1 + d2));
    dL = sg * fmin2(m, cThird * fabsr(d1 + 2 * d2));
  } else if (LIM == LIM_MC) {
    const Real s = (fsign(Real(0.25), d1) + fsign(Real(0.25), d2)) *
                     fmin2(fmin2(beta * fabsr(d1), beta * fabsr(d2)),
                           Real(0.5) * fabsr(d1 + d2));
    dL = s; dR = s;
  } else if (LIM == LIM_MINMOD) {
    const Real s =
        (fsign(Real(0.25), d2) + fsign(Real(0.25), d1)) * fmin2(fabsr(d2), fabsr(d1));
    dL = s; dR = s;
  } else if (LIM == LIM_BETA) {
    const Real a1 = fabsr(d1), a2 = fabsr(d2);
    const Real sum = fsign(Real(0.25), d2) + fsign(Real(0.25), d1);
    const Real s = sum * fmax2(fmin2(a2, beta * a1), fmin2(a1, beta * a2));
    dL = s; dR = s;
  } else if (LIM == LIM_NO) {
    const Real s = Real(0.25) * (p2 - p0);
    dL = s; dR = s;
  } else {
    dL = Real(0.0); dR = Real(0.0);
  }
}

// Left/right face states of the face between cells f-1 and f from the
// primitive values q0=P(f-2), q1=P(f-1), q2=P(f), q3=P(f+1):
// L = P(f-1) + dL(f-1), R = P(f) - dR(f)  (src/ModFaceValue.f90:2021-2027)
// left and right states of one variable at the face between cells 1 and 2
// of the 4-cell stencil q0..q3 (src/ModFaceValue.f90:2021-2027)
// The slopes of the two cells use their own beta (betaL for cell 1, betaR
// for cell 2): the sharper BetaLimiterResChange applies per cell next to a
// resolution change or an outer boundary (src/ModFaceValue.f90:1982-2015).
template <int LIM, typename Real>
__device__ __forceinline__ void face_lr(Real betaL, Real betaR, Real q0,
                                        Real q1, Real q2, Real q3, Real &L,
                                        Real &R) {
  if (LIM == LIM_FIRST) {
    L = q1;
    R = q2;
  } else {
    Real dl, dr, t;
    slopes<LIM>(betaL, q0, q1, q2, dl, t);
    slopes<LIM>(betaR, q1, q2, q3, t, dr);
    L = q1 + dl;
    R = q2 - dr;
  }
}

// face_lr with the low order weight of the face: the limited slopes scaled
// by w = 1 - LowOrderCrit (src/ModFaceValue.f90:2021-2027)
template <int LIM, typename Real>
__device__ __forceinline__ void face_lr_w(Real betaL, Real betaR, Real w,
                                          Real q0, Real q1, Real q2, Real q3,
                                          Real &L, Real &R) {
  if (LIM == LIM_FIRST) {
    L = q1;
    R = q2;
  } else {
    Real dl, dr, t;
    slopes<LIM>(betaL, q0, q1, q2, dl, t);
    slopes<LIM>(betaR, q1, q2, q3, t, dr);
    L = q1 + dl * w;
    R = q2 - dr * w;
  }
}

template <int LIM, typename Real>
__device__ __forceinline__ void face_states_w(Real betaL, Real betaR, Real w,
                                              const Real *q0, const Real *q1,
                                              const Real *q2, const Real *q3,
                                              Real *L, Real *R) {
#pragma unroll
  for (int v = 0; v < NVAR; v++)
    face_lr_w<LIM>(betaL, betaR, w, q0[v], q1[v], q2[v], q3[v], L[v], R[v]);
}

template <int LIM, typename Real>
__device__ __forceinline__ void face_states(Real betaL, Real betaR,
                                            const Real *q0, const Real *q1,
                                            const Real *q2, const Real *q3,
                                            Real *L, Real *R) {
#pragma unroll
  for (int v = 0; v < NVAR; v++)
    face_lr<LIM>(betaL, betaR, q0[v], q1[v], q2[v], q3[v], L[v], R[v]);
}

// runtime-limiter variants used by the per-block parity routines
template <typename Real>
__device__ __forceinline__ void face_states_rt(int lim, Real betaL,
                                               Real betaR, const Real *q0,
                                               const Real *q1,
                                               const Real *q2,
                                               const Real *q3, Real *L,
                                               Real *R) {
  switch (lim) {
  case LIM_MINMOD: face_states<LIM_MINMOD>(betaL, betaR, q0, q1, q2, q3, L, R); break;
  case LIM_MC: face_states<LIM_MC>(betaL, betaR, q0, q1, q2, q3, L, R); break;
  case LIM_MC3: face_states<LIM_MC3>(betaL, betaR, q0, q1, q2, q3, L, R); break;
  case LIM_BETA: face_states<LIM_BETA>(betaL, betaR, q0, q1, q2, q3, L, R); break;
  case LIM_NO: face_states<LIM_NO>(betaL, betaR, q0, q1, q2, q3, L, R); break;
  default: face_states<LIM_FIRST>(betaL, betaR, q0, q1, q2, q3, L, R); break;
  }
}
template <typename Real>
__device__ __forceinline__ void face_states_w_rt(int lim, Real betaL,
                                                 Real betaR, Real w,
                                                 const Real *q0,
                                                 const Real *q1,
                                                 const Real *q2,
                                                 const Real *q3, Real *L,
                                                 Real *R) {
  switch (lim) {
  case LIM_MINMOD: face_states_w<LIM_MINMOD>(betaL, betaR, w, q0, q1, q2, q3, L, R); break;
  case LIM_MC: face_states_w<LIM_MC>(betaL, betaR, w, q0, q1, q2, q3, L, R); break;
  case LIM_MC3: face_states_w<LIM_MC3>(betaL, betaR, w, q0, q1, q2, q3, L, R); break;
  case LIM_BETA: face_states_w<LIM_BETA>(betaL, betaR, w, q0, q1, q2, q3, L, R); break;
  case LIM_NO: face_states_w<LIM_NO>(betaL, betaR, w, q0, q1, q2, q3, L, R); break;
  default: face_states_w<LIM_FIRST>(betaL, betaR, w, q0, q1, q2, q3, L, R); break;
  }
}

// beta of the slopes of cell c (0-based, -1..NB along one direction) of a
// block whose low / high side in that direction has DiLevel_EB /= 0: the
// cells within nFace of such a side take betaRC
// (src/ModFaceValue.f90:1982-2015 with iMin = 1, iMax = nI+1)
template <typename Real>
__device__ __forceinline__ Real cell_beta(Real beta, Real betaRC, int nFace,
                                          bool lo, bool hi, int c, int NB) {
  return ((lo && c <= nFace - 1) || (hi && c >= NB - nFace)) ? betaRC : beta;
}

// conserved (p stored) -> primitive, u = (1/rho)*rho*u
// (src/ModFaceValue.f90:1541,1581-1582)
template <typename Real>
__device__ __forceinline__ void to_primitive(Real *q) {
  const Real RhoInv = frcp(q[RHO]);
  q[MX] = RhoInv * q[MX];
  q[MY] = RhoInv * q[MY];
  q[MZ] = RhoInv * q[MZ];
}

// ---------------------------------------------------------------------
// physical flux of a primitive state for a face normal to axis DIR
// (src/ModFaceFlux.f90:1194-1212,1649-1876,1574-1601).  Cons[9] is the
// conserved state with total energy in slot 8.
//
// With B0 splitting (B0F, the face's B0 in B0): S holds B1; the magnetic
// force is -B1n (B1 + B0) - B0n B1 + (B1^2/2 + B0.B1) n, the induction flux
// uses the full field, the energy flux the Poynting term of B1
// (src/ModFaceFlux.f90:1203-1212,1798-1805).
template <int DIR, bool FULL, bool B0F = false, typename Real>
__device__ __forceinline__ void phys_flux(const PhysT<Real> &ph, const Real *S,
                                          Real *Cons, Real *F,
                                          Real &UnOut,
                                          const Real *B0 = nullptr) {
  const Real Rho = S[RHO], p = S[PP];
  const Real u[3] = {S[MX], S[MY], S[MZ]};
  const Real b[3] = {S[BX], S[BY], S[BZ]};
  const Real Un = u[DIR], Bn = b[DIR];
  Real fullB[3] = {b[0], b[1], b[2]};
  Real B0n = Real(0), FullBn = Bn;
  if (B0F) {
#pragma unroll
    for (int d = 0; d < 3; d++) fullB[d] = b[d] + B0[d];
    B0n = B0[DIR];
    FullBn = B0n + Bn;
  }
  F[RHO] = Rho * Un;
  Cons[RHO] = Rho;
  Cons[MX] = Rho * u[0]; Cons[MY] = Rho * u[1]; Cons[MZ] = Rho * u[2];
  Cons[BX] = b[0]; Cons[BY] = b[1]; Cons[BZ] = b[2];
  Cons[PP] = p;
  const Real e = ph.invgm1 * p + Real(0.5) * Rho * (u[0] * u[0] + u[1] * u[1] +
                                                u[2] * u[2]);
  const Real B2 = b[0] * b[0] + b[1] * b[1] + b[2] * b[2];
  Real pTotal = Real(0.5) * B2;
  if (B0F) pTotal = pTotal + (B0[0] * b[0] + B0[1] * b[1] + B0[2] * b[2]);
  const Real UnRho = Un * Rho;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    Real f = UnRho * u[d];
    if (d == DIR) f = f + p;
    Real mf = -Bn * fullB[d];
    if (B0F) mf = mf - B0n * b[d];
    if (d == DIR) mf = mf + pTotal;
    F[MX + d] = f + mf;
  }
  Real fb[3];
#pragma unroll
  for (int d = 0; d < 3; d++)
    fb[d] = (d == DIR) ? Real(0.0) : Un * fullB[d] - u[d] * FullBn;
  F[BX] = fb[0]; F[BY] = fb[1]; F[BZ] = fb[2];
  Real fe = Un * (e + p);
#pragma unroll
  for (int d = 0; d < 3; d++)
    if (d != DIR) fe = fe + fb[d] * b[d];
  F[EN] = fe;
  Cons[EN] = e + Real(0.5) * B2;
  F[PP] = FULL ? Un * p : Real(0.0);
  UnOut = Un;
}

// fast magnetosonic speed of state S for a face normal to DIR
// (src/ModFaceFlux.f90:3900-3921,3990-4134).  Returns false on a negative
// fast speed squared.
// With B0 splitting the speeds use the full field B1 + B0 (:3912-3920).
template <int DIR, bool B0F = false, typename Real>
__device__ __forceinline__ bool fast_speed(const PhysT<Real> &ph, const Real *S,
                                           Real &Un, Real &Fast,
                                           const Real *B0 = nullptr) {
  const Real InvRho = frcp(S[RHO]);
  const Real Sound2 = S[PP] * ph.gamma * InvRho;
  Un = S[MX + DIR];
  Real bx = S[BX], by = S[BY], bz = S[BZ];
  if (B0F) { bx = bx + B0[0]; by = by + B0[1]; bz = bz + B0[2]; }
  const Real Alfven2 = (bx * bx + by * by + bz * bz) * InvRho;
  const Real FullBn = B0F ? (DIR == 0 ? bx : DIR == 1 ? by : bz)
                          : S[BX + DIR];
  const Real Alfven2Normal = InvRho * (FullBn * FullBn);
  const Real Fast2 = Sound2 + Alfven2;
  const Real Discr =
      fsqrt(fmax2(Real(0.0), Fast2 * Fast2 - 4 * Sound2 * Alfven2Normal));
  const bool ok = !(Fast2 + Discr < Real(0.0));
  Fast = fsqrt(Real(0.5) * (Fast2 + Discr));
  return ok;
}

// ---------------------------------------------------------------------
// HLLD (src/ModFaceFlux.f90:2826-3160) on Linde-corrected states, with the
// reference's two mirror-image branches on the sign of the contact speed Un.
// (Evaluating the Un < 0 branch as the reflected Un >= 0 branch removes the
// branch divergence but keeps about 30 more values live; on gfx950 the
// resulting register spills cost more than the divergence.)
// With B0 splitting (B0F) the states hold B1 and the face's B0 enters the
// total field and total pressure (src/ModFaceFlux.f90:2905-2921,2955-2958,
// 3115-3129).
template <int DIR, bool FULL, bool B0F = false, typename Real>
__device__ __forceinline__ int hlld(const PhysT<Real> &ph, const Real *SL,
                                    const Real *SR, Real *F, Real &Cmax,
                                    Real &UnOut, const Real *B0 = nullptr) {
  constexpr int in = DIR, it1 = (DIR + 1) % 3, it2 = (DIR + 2) % 3;
  int err = 0;
  Real UnLs, FastL, UnRs, FastR;
  if (!fast_speed<DIR, B0F>(ph, SL, UnLs, FastL, B0)) err |= ERRB_NEGFAST2;
  if (!fast_speed<DIR, B0F>(ph, SR, UnRs, FastR, B0)) err |= ERRB_NEGFAST2;
  const Real sL = fmin2(UnLs - FastL, UnRs - FastR);
  const Real sR = fmax2(UnLs + FastL, UnRs + FastR);
  Cmax = fmax2(sR, -sL);
  if (sL >= Real(0.) || sR <= Real(0.)) {  // supersonic: upwind physical flux
    const bool left = sL >= Real(0.);
    Real S[NVAR], Cons[NFLUX];
#pragma unroll
    for (int v = 0; v < NVAR; v++) S[v] = left ? SL[v] : SR[v];
    phys_flux<DIR, FULL, B0F>(ph, S, Cons, F, UnOut, B0);
    return err;
  }
  const Real cTiny = ph.cTiny;
  const Real RhoL = SL[RHO], pL = SL[PP], RhoR = SR[RHO], pR = SR[PP];
  const Real UnL = SL[MX + in], Ut1L = SL[MX + it1], Ut2L = SL[MX + it2];
  const Real B1nL = SL[BX + in], B1t1L = SL[BX + it1], B1t2L = SL[BX + it2];
  const Real UnR = SR[MX + in], Ut1R = SR[MX + it1], Ut2R = SR[MX + it2];
  const Real B1nR = SR[BX + in], B1t1R = SR[BX + it1], B1t2R = SR[BX + it2];

  const Real B0n = B0F ? B0[in] : Real(0), B0t1 = B0F ? B0[it1] : Real(0),
             B0t2 = B0F ? B0[it2] : Real(0);
  const Real B1n = Real(0.5) * (B1nL + B1nR);
  const Real B1n2 = B1n * B1n;
  const Real Bn = B0F ? B1n + B0n : B1n;
  const Real Bn2 = Bn * Bn;
  const Real SignBn = fsign(Real(1.), Bn);
  const Real PbL = Real(0.5) * (B1n2 + B1t1L * B1t1L + B1t2L * B1t2L);
  const Real PbR = Real(0.5) * (B1n2 + B1t1R * B1t1R + B1t2R * B1t2R);
  Real PtotL = pL + PbL;
  Real PtotR = pR + PbR;
  if (B0F) {
    PtotL = PtotL + B1n * B0n + B1t1L * B0t1 + B1t2L * B0t2;
    PtotR = PtotR + B1n * B0n + B1t1R * B0t1 + B1t2R * B0t2;
  }
  const Real DsL = sL - UnL, DsR = sR - UnR;
  const Real DsRhoL = DsL * RhoL, DsRhoR = DsR * RhoR;
  const Real Tmp0 = frcp(DsRhoR - DsRhoL);
  const Real Un = (DsRhoR * UnR - DsRhoL * UnL - PtotR + PtotL) * Tmp0;
  const Real Ptot12 =
      (DsRhoR * PtotL - DsRhoL * PtotR + DsRhoR * DsRhoL * (UnR - UnL)) * Tmp0;

  const Real Bt1L = B0F ? B1t1L + B0t1 : B1t1L;
  const Real Bt2L = B0F ? B1t2L + B0t2 : B1t2L;
  const Real Bt1R = B0F ? B1t1R + B0t1 : B1t1R;
  const Real Bt2R = B0F ? B1t2R + B0t2 : B1t2R;
  Real Rho, Ut1, Ut2, B1t1, B1t2, uDotB1, e;
  Real Ut1L1, Ut2L1, B1t1L1, B1t2L1, Ut1R1, Ut2R1, B1t1R1, B1t2R1;
  auto outer_left = [&]() {
    Real T = DsRhoL * (sL - Un) - Bn2;
    if (T < cTiny) {
      Ut1L1 = Ut1L; Ut2L1 = Ut2L; B1t1L1 = B1t1L; B1t2L1 = B1t2L;
    } else {
      T = frcp(T);
      Ut1L1 = Ut1L - Bn * Bt1L * (Un - UnL) * T;
      Ut2L1 = Ut2L - Bn * Bt2L * (Un - UnL) * T;
      B1t1L1 = Bt1L * (DsRhoL * DsL - Bn2) * T;
      B1t2L1 = Bt2L * (DsRhoL * DsL - Bn2) * T;
      if (B0F) { B1t1L1 = B1t1L1 - B0t1; B1t2L1 = B1t2L1 - B0t2; }
    }
  };
  auto outer_right = [&]() {
    Real T = DsRhoR * (sR - Un) - Bn2;
    if (T < cTiny) {
      Ut1R1 = Ut1R; Ut2R1 = Ut2R; B1t1R1 = B1t1R; B1t2R1 = B1t2R;
    } else {
      T = frcp(T);
      Ut1R1 = Ut1R - Bn * Bt1R * (Un - UnR) * T;
      Ut2R1 = Ut2R - Bn * Bt2R * (Un - UnR) * T;
      B1t1R1 = Bt1R * (DsRhoR * DsR - Bn2) * T;
      B1t2R1 = Bt2R * (DsRhoR * DsR - Bn2) * T;
      if (B0F) { B1t1R1 = B1t1R1 - B0t1; B1t2R1 = B1t2R1 - B0t2; }
    }
  };
  if (Un >= Real(0.)) {
    const Real InvSLUn = frcp(sL - Un);
    Rho = DsRhoL * InvSLUn;
    outer_left();
    const Real eL = ph.invgm1 * pL + PbL +
                      Real(0.5) * RhoL * (UnL * UnL + Ut1L * Ut1L + Ut2L * Ut2L);
    const Real uDotB1L = UnL * B1n + Ut1L * B1t1L + Ut2L * B1t2L;
    const Real uDotB1L1 = Un * B1n + Ut1L1 * B1t1L1 + Ut2L1 * B1t2L1;
    e = (DsL * eL - PtotL * UnL + Ptot12 * Un + Bn * (uDotB1L - uDotB1L1)) *
        InvSLUn;
    const Real SqRhoL1 = fsqrt(Rho);
    const Real sL1 = Un - fabsr(Bn) * frcp(SqRhoL1);
    if (sL1 >= Real(0.)) {
      Ut1 = Ut1L1; Ut2 = Ut2L1; B1t1 = B1t1L1; B1t2 = B1t2L1;
      uDotB1 = uDotB1L1;
    } else {
      const Real ratio = DsRhoR * frcp(sR - Un);
      if (ratio < 0) err |= ERRB_HLLD;
      const Real SqRhoR1 = fsqrt(ratio);
      outer_right();
      const Real Ti = frcp(SqRhoL1 + SqRhoR1);
      const Real SqRhoLR1 = SqRhoL1 * SqRhoR1 * SignBn;
      Ut1 = (SqRhoL1 * Ut1L1 + SqRhoR1 * Ut1R1 + (B1t1R1 - B1t1L1) * SignBn) * Ti;
      Ut2 = (SqRhoL1 * Ut2L1 + SqRhoR1 * Ut2R1 + (B1t2R1 - B1t2L1) * SignBn) * Ti;
      B1t1 = (SqRhoL1 * B1t1R1 + SqRhoR1 * B1t1L1 + (Ut1R1 - Ut1L1) * SqRhoLR1) * Ti;
      B1t2 = (SqRhoL1 * B1t2R1 + SqRhoR1 * B1t2L1 + (Ut2R1 - Ut2L1) * SqRhoLR1) * Ti;
      uDotB1 = Un * B1n + Ut1 * B1t1 + Ut2 * B1t2;
      e = e - SqRhoL1 * (uDotB1L1 - uDotB1) * SignBn;
    }
  } else {
    const Real InvSRUn = frcp(sR - Un);
    Rho = DsRhoR * InvSRUn;
    outer_right();
    const Real eR = ph.invgm1 * pR + PbR +
                      Real(0.5) * RhoR * (UnR * UnR + Ut1R * Ut1R + Ut2R * Ut2R);
    const Real uDotB1R = UnR * B1n + Ut1R * B1t1R + Ut2R * B1t2R;
    const Real uDotB1R1 = Un * B1n + Ut1R1 * B1t1R1 + Ut2R1 * B1t2R1;
    e = (DsR * eR - PtotR * UnR + Ptot12 * Un + Bn * (uDotB1R - uDotB1R1)) *
        InvSRUn;
    const Real SqRhoR1 = fsqrt(Rho);
    const Real sR1 = Un + fabsr(Bn) * frcp(SqRhoR1);
    if (sR1 <= Real(0.)) {
      Ut1 = Ut1R1; Ut2 = Ut2R1; B1t1 = B1t1R1; B1t2 = B1t2R1;
      uDotB1 = uDotB1R1;
    } else {
      const Real ratio = DsRhoL * frcp(sL - Un);
      if (ratio < 0) err |= ERRB_HLLD;
      const Real SqRhoL1 = fsqrt(ratio);
      outer_left();
      const Real Ti = frcp(SqRhoL1 + SqRhoR1);
      const Real SqRhoLR1 = SqRhoL1 * SqRhoR1 * SignBn;
      Ut1 = (SqRhoL1 * Ut1L1 + SqRhoR1 * Ut1R1 + (B1t1R1 - B1t1L1) * SignBn) * Ti;
      Ut2 = (SqRhoL1 * Ut2L1 + SqRhoR1 * Ut2R1 + (B1t2R1 - B1t2L1) * SignBn) * Ti;
      B1t1 = (SqRhoL1 * B1t1R1 + SqRhoR1 * B1t1L1 + (Ut1R1 - Ut1L1) * SqRhoLR1) * Ti;
      B1t2 = (SqRhoL1 * B1t2R1 + SqRhoR1 * B1t2L1 + (Ut2R1 - Ut2L1) * SqRhoLR1) * Ti;
      uDotB1 = Un * B1n + Ut1 * B1t1 + Ut2 * B1t2;
      e = e + SqRhoR1 * (uDotB1R1 - uDotB1) * SignBn;
    }
  }
  // flux of the HLLD state in the original frame (:3111-3131)
  const Real RhoUn = Rho * Un;
  const Real Bt1 = B0F ? B1t1 + B0t1 : B1t1, Bt2 = B0F ? B1t2 + B0t2 : B1t2;
  F[RHO] = RhoUn;
  if (B0F) {
    F[MX + in] = RhoUn * Un - B1n * Bn - B0n * B1n + Ptot12;
    F[MX + it1] = RhoUn * Ut1 - B1n * Bt1 - B0n * B1t1;
    F[MX + it2] = RhoUn * Ut2 - B1n * Bt2 - B0n * B1t2;
  } else {
    F[MX + in] = RhoUn * Un - B1n * Bn + Ptot12;
    F[MX + it1] = RhoUn * Ut1 - B1n * Bt1;
    F[MX + it2] = RhoUn * Ut2 - B1n * Bt2;
  }
  F[BX + in] = Real(0.0);
  F[BX + it1] = Un * Bt1 - Ut1 * Bn;
  F[BX + it2] = Un * Bt2 - Ut2 * Bn;
  F[PP] = FULL ? Real(0.5) * (sR * UnL * pL - sL * UnR * pR + sR * sL * (pR - pL)) *
                     frcp(sR - sL)
               : Real(0.0);
  F[EN] = Un * (e + Ptot12) - Bn * uDotB1;
  UnOut = Un;
  return err;
}

// ---------------------------------------------------------------------
// numerical flux for a face normal to DIR from the left/right primitive
// face states (src/ModFaceFlux.f90:2130-2473).  Out: F[0..8] already times
// Area, Vdt = CmaxDt*Area, Un*Area.
// With B0 splitting (B0F) B0 is the face's B0 in the face's frame.
template <int DIR, int FLUX, bool FULL, bool B0F = false, typename Real>
__device__ __forceinline__ int numerical_flux(const PhysT<Real> &ph, Real Area,
                                              const Real *Lin,
                                              const Real *Rin, Real *F,
                                              Real &Vdt, Real &UnArea,
                                              const Real *B0 = nullptr) {
  Real L[NVAR], R[NVAR];
#pragma unroll
  for (int v = 0; v < NVAR; v++) { L[v] = Lin[v]; R[v] = Rin[v]; }
  int err = 0;
  Real Cmax, Un = Real(0.0);
  Real dB = Real(0.0), DiffE = Real(0.0);
  constexpr bool Linde = (FLUX == FLUX_LINDE || FLUX == FLUX_HLLD);
  if (Linde) {
    // remove the jump of the normal field (:2302-2320)
    dB = Real(0.5) * (R[BX + DIR] - L[BX + DIR]);
    L[BX + DIR] = L[BX + DIR] + dB;
    R[BX + DIR] = R[BX + DIR] - dB;
    DiffE = Real(0.5) * ((R[BX + DIR] + L[BX + DIR]) * dB);
  }
  if (FLUX == FLUX_HLLD) {
    err |= hlld<DIR, FULL, B0F>(ph, L, R, F, Cmax, Un, B0);
  } else {
    Real S[NVAR];
#pragma unroll
    for (int v = 0; v < NVAR; v++) S[v] = Real(0.5) * (L[v] + R[v]);
    Real CL[NFLUX], CR[NFLUX], FL[NFLUX], FR[NFLUX], UnL, UnR;
    phys_flux<DIR, FULL, B0F>(ph, L, CL, FL, UnL, B0);
    phys_flux<DIR, FULL, B0F>(ph, R, CR, FR, UnR, B0);
    Real UnHat, FastHat;
    if (!fast_speed<DIR, B0F>(ph, S, UnHat, FastHat, B0)) err |= ERRB_NEGFAST2;
    Cmax = fabsr(UnHat) + FastHat;
    if (FLUX == FLUX_RUSANOV) {
#pragma unroll
      for (int v = 0; v < NFLUX; v++)
        F[v] = Real(0.5) * (FL[v] + FR[v] - Cmax * (CR[v] - CL[v]));
      Un = Real(0.5) * (UnL + UnR);
    } else {
      Real UnLs, FastL, UnRs, FastR;
      if (!fast_speed<DIR, B0F>(ph, L, UnLs, FastL, B0)) err |= ERRB_NEGFAST2;
      if (!fast_speed<DIR, B0F>(ph, R, UnRs, FastR, B0)) err |= ERRB_NEGFAST2;
      const Real Cleft = fmin2(fmin2(Real(0.0), UnLs - FastL), UnHat - FastHat);
      const Real Cright = fmax2(fmax2(Real(0.0), UnRs + FastR), UnHat + FastHat);
      const Real WeightLeft = Cright * frcp(Cright - Cleft);
      const Real WeightRight = Real(1.0) - WeightLeft;
      const Real Diffusion = Cright * WeightRight;
#pragma unroll
      for (int v = 0; v < NFLUX; v++)
        F[v] = (WeightRight * FR[v] + WeightLeft * FL[v] -
                Diffusion * (CR[v] - CL[v]));
      Un = WeightRight * UnR + WeightLeft * UnL;
    }
    if (!FULL) F[PP] = Real(0.0);
  }
  if (Linde) { // Linde's Lax-Friedrichs flux for Bn (:2403-2418)
    F[BX + DIR] = F[BX + DIR] - Cmax * dB;
    F[EN] = F[EN] - Cmax * DiffE;
  }
#pragma unroll
  for (int v = 0; v < NFLUX; v++) F[v] = F[v] * Area;
  Vdt = Cmax * Area;
  UnArea = Un * Area;
  return err;
}

// ---------------------------------------------------------------------
// energy <-> pressure for one cell (src/ModEnergy.f90:88-97,292-319)
__device__ __forceinline__ double pressure_to_energy(const Phys &ph,
                                                     const double *S,
                                                     double InvRho) {
  double E = ph.invgm1 * S[PP] +
             0.5 * (S[MX] * S[MX] + S[MY] * S[MY] + S[MZ] * S[MZ]) * InvRho;
  E = E + 0.5 * (S[BX] * S[BX] + S[BY] * S[BY] + S[BZ] * S[BZ]);
  return E;
}
__device__ __forceinline__ double pressure_to_energy(const Phys &ph,
                                                     const double *S) {
  return pressure_to_energy(ph, S, frcp(S[RHO]));
}
__device__ __forceinline__ double energy_to_pressure(const Phys &ph,
                                                     const double *S,
                                                     double E,
                                                     double InvRho) {
  double p = E - 0.5 * (S[BX] * S[BX] + S[BY] * S[BY] + S[BZ] * S[BZ]);
  p = ph.gm1 *
      (p - 0.5 * (S[MX] * S[MX] + S[MY] * S[MY] + S[MZ] * S[MZ]) * InvRho);
  return p;
}
__device__ __forceinline__ double energy_to_pressure(const Phys &ph,
                                                     const double *S,
                                                     double E) {
  return energy_to_pressure(ph, S, E, frcp(S[RHO]));
}
__device__ __forceinline__ double limit_pressure(const Phys &ph, double p) {
  return ph.pMin < 0.0 ? p : fmax2(ph.pMin, p);
}

// Stage update of one cell (src/ModUpdateState.f90:331-339,606-686,
// 744-773,916-965).
//   dF[9]     : sum of area-weighted face fluxes (in - out)
//   Uin[8]    : stage-input state (p stored)
//   Uold[8]   : state at the start of the time step (StateOld)
//   out Unew  : updated state (p stored)
struct UpdateCtl {
  double DtLocal;     // DtFactor*DtMax_CB
  double InvVolume;   // 1/CellVolume
  int halfStep;       // UseHalfStep or nStage==1
  int iStage, nStage;
  int dbTrick;        // add |dB|^2/2 to the energy at this stage (the host
                      // decides: src/ModSetParameters.f90:3903-3906,
                      // src/ModUpdateState.f90:916-918)
  int dtZero;        // time accurate and Dt == 0: leave the state alone
  int roundTripOld;   // emulate StateOld's p->E->p round trips (one per
                      // earlier stage of the step)
};

__device__ __forceinline__ void update_cell(const Phys &ph, const UpdateCtl &u,
                                            const double *dF,
                                            const double *Uin,
                                            const double *Uold,
                                            double *Unew,
                                            const double *Src = nullptr) {
  if (u.dtZero) {
#pragma unroll
    for (int v = 0; v < NVAR; v++) Unew[v] = Uin[v];
    Unew[PP] = limit_pressure(ph, Unew[PP]);
    return;
  }
  double Source[NFLUX];
#pragma unroll
  for (int v = 0; v < NFLUX; v++)
    Source[v] = Src ? u.DtLocal * (Src[v] + dF[v] * u.InvVolume)
                    : u.DtLocal * (dF[v] * u.InvVolume);
  double Old[NVAR];
#pragma unroll
  for (int v = 0; v < NVAR; v++) Old[v] = Uold[v];
  const double InvRhoOld = frcp(Old[RHO]);
  if (u.roundTripOld)
    for (int r = 1; r < u.iStage; r++)
      Old[PP] = energy_to_pressure(
          ph, Old, pressure_to_energy(ph, Old, InvRhoOld), InvRhoOld);
  const double Eold = pressure_to_energy(ph, Old, InvRhoOld);
  double Base[NVAR];
  double Ebase;
  if (u.halfStep) {
#pragma unroll
    for (int v = 0; v < NVAR; v++) Base[v] = Old[v];
    Ebase = Eold;
  } else {
#pragma unroll
    for (int v = 0; v < NVAR; v++) Base[v] = Uin[v];
    Ebase = pressure_to_energy(ph, Uin);
  }
#pragma unroll
  for (int v = 0; v < NVAR; v++)
    Unew[v] = (v == PP) ? Ebase + Source[EN] : Base[v] + Source[v];
  if (!u.halfStep && u.iStage > 1) {
    // Runge-Kutta interpolation weights (src/ModUpdateState.f90:748-756)
    const double Coeff1 =
        u.nStage == 3 ? (u.iStage == 2 ? 0.75 : 1.0 / 3.0) : 0.5;
    const double Coeff2 = 1 - Coeff1;
#pragma unroll
    for (int v = 0; v < NVAR; v++) {
      const double o = (v == PP) ? Eold : Old[v];
      Unew[v] = Coeff1 * o + Coeff2 * Unew[v];
    }
  }
  if (u.dbTrick)
    Unew[PP] = Unew[PP] + 0.5 * (Source[BX] * Source[BX] +
                                 Source[BY] * Source[BY] +
                                 Source[BZ] * Source[BZ]);
  Unew[PP] = limit_pressure(ph, energy_to_pressure(ph, Unew, Unew[PP]));
}

// Ghost face state of the inner body ('ionosphere' type, set_face in
// src/ModUpdateStateFast.f90:1277-1327): from the true side's face state
// T, the density moves toward the body density by at most 10 % of the true
// value, the pressure floats, the velocity reverses, and 1 - B1rCoef of the
// radial field (along r, the face centre relative to the body centre) is
// removed twice over: B1rCoef = -1 reflects B1r, 0 removes it.
template <typename Real>
__device__ __forceinline__ void body_ghost(const Real *T, const Real *r,
                                           Real BodyRho, Real B1rCoef,
                                           Real *G) {
  const Real DensityJumpLimit = Real(0.1);
  const Real d = BodyRho - T[RHO];
  G[RHO] = T[RHO] + fsign(Real(1), d) *
                        fmin2(fabsr(d), DensityJumpLimit * T[RHO]);
  G[MX] = -T[MX];
  G[MY] = -T[MY];
  G[MZ] = -T[MZ];
  G[PP] = T[PP];
  if (B1rCoef != Real(1)) {
    const Real Coef = (1 - B1rCoef) / (r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    const Real BdotR = T[BX] * r[0] + T[BY] * r[1] + T[BZ] * r[2];
    G[BX] = T[BX] - Coef * r[0] * BdotR;
    G[BY] = T[BY] - Coef * r[1] * BdotR;
    G[BZ] = T[BZ] - Coef * r[2] * BdotR;
  } else {
    G[BX] = T[BX];
    G[BY] = T[BY];
    G[BZ] = T[BZ];
  }
}

// B0 of a cell for the 8-wave source with B0 splitting: the B0 of its six
// faces and its own B0; useSource adds -B1 div(B0) - curl(B0) x B1
template <typename Real> struct CellB0 {
  Real face[6][3];
  Real cell[3];
  int useSource;
};

// 8-wave (Powell) div B source of one cell on a Cartesian grid
// (src/ModCalcSource.f90:1694-1889 with DoCorrectFace false, :930-951,
// :1016-1022), with the B0 terms.  B?lo/B?hi are the normal-field left (L)
// and right (R) face states at the lower and upper faces of the cell in
// each direction; InvHalf = 0.5/CellSize; P is the cell's stage-input
// state in primitive form (velocity u = rho u / rho, field B1).  B0
// splitting: faceB0(f, out) gives the B0 of the cell's face f (x-, x+, y-,
// y+, z-, z+), cellB0(q) the cell's own B0 (q = 0..2).  The faces are
// fetched a direction at a time, each pair consumed before the next is
// asked for (which keeps the fused stage's register use down), and div B0
// and curl B0 are formed from them as set_b0_source forms them
// (src/ModB0.f90:568-590).
template <bool B0S, typename Real, typename FaceB0, typename CellV>
__device__ __forceinline__ void powell_source_with(
    const Real *InvHalf, Real LxLo, Real RxLo, Real LxHi, Real RxHi,
    Real LyLo, Real RyLo, Real LyHi, Real RyHi, Real LzLo, Real RzLo,
    Real LzHi, Real RzHi, const Real *P, Real *Src, FaceB0 &&faceB0,
    CellV &&cellB0, int useB0Source) {
  const Real dBf[6] = {InvHalf[0] * (RxLo - LxLo), InvHalf[0] * (RxHi - LxHi),
                       InvHalf[1] * (RyLo - LyLo), InvHalf[1] * (RyHi - LyHi),
                       InvHalf[2] * (RzLo - LzLo), InvHalf[2] * (RzHi - LzHi)};
  Real Internal = 2 * InvHalf[0] * (LxHi - RxLo);
  Internal = Internal + 2 * InvHalf[1] * (LyHi - RyLo);
  Internal = Internal + 2 * InvHalf[2] * (LzHi - RzLo);
  Real DivB = Internal + dBf[0] + dBf[1];
  DivB = DivB + dBf[2] + dBf[3];
  DivB = DivB + dBf[4] + dBf[5];
  const Real ux = P[MX], uy = P[MY], uz = P[MZ];
  const Real Bx = P[BX], By = P[BY], Bz = P[BZ];
  Src[RHO] = Real(0);
  if (B0S) {
    // the face B0 times each face's normal-field jump, the cell B0 times
    // the internal divergence (src/ModCalcSource.f90:1792-1808,
    // 1880-1885), then -B1 div(B0) - curl(B0) x B1 with UseB0Source
    // (:953-967)
    Real m[3] = {Real(0), Real(0), Real(0)};
    Real div = Real(0), curl[3] = {Real(0), Real(0), Real(0)};
#pragma unroll
    for (int d = 0; d < 3; d++) {
      Real lo[3], hi[3];
      faceB0(2 * d, lo);
      faceB0(2 * d + 1, hi);
#pragma unroll
      for (int v = 0; v < 3; v++) m[v] = m[v] - lo[v] * dBf[2 * d];
#pragma unroll
      for (int v = 0; v < 3; v++) m[v] = m[v] - hi[v] * dBf[2 * d + 1];
      if (useB0Source) {
        // DxInv*(B0X(i+1) - B0X(i)) etc.; curl B0 in its x, y, z terms
        const Real DInv = 2 * InvHalf[d];
        const int a1 = (d + 1) % 3, a2 = (d + 2) % 3;
        div = div + DInv * (hi[d] - lo[d]);
        curl[a2] = curl[a2] + DInv * (hi[a1] - lo[a1]);
        curl[a1] = curl[a1] - DInv * (hi[a2] - lo[a2]);
      }
      __builtin_amdgcn_sched_barrier(0);
    }
    const Real B1[3] = {Bx, By, Bz};
#pragma unroll
    for (int v = 0; v < 3; v++)
      m[v] = (m[v] - Internal * cellB0(v)) - DivB * B1[v];
    if (useB0Source) {
      m[0] = m[0] - Bx * div - (curl[1] * Bz - curl[2] * By);
      m[1] = m[1] - By * div - (curl[2] * Bx - curl[0] * Bz);
      m[2] = m[2] - Bz * div - (curl[0] * By - curl[1] * Bx);
    }
    Src[MX] = m[0];
    Src[MY] = m[1];
    Src[MZ] = m[2];
  } else {
    Src[MX] = -(DivB * Bx);
    Src[MY] = -(DivB * By);
    Src[MZ] = -(DivB * Bz);
  }
  Src[BX] = -(DivB * ux);
  Src[BY] = -(DivB * uy);
  Src[BZ] = -(DivB * uz);
  Src[PP] = Real(0);
  Src[EN] = -(DivB * (Bx * ux + By * uy + Bz * uz));
}

// the same from the cell's stored state St (rho u) and, with B0 splitting,
// its B0 values in a CellB0 (null: no B0)
template <typename Real>
__device__ __forceinline__ void powell_source(
    const Real *InvHalf, Real LxLo, Real RxLo, Real LxHi, Real RxHi,
    Real LyLo, Real RyLo, Real LyHi, Real RyHi, Real LzLo, Real RzLo,
    Real LzHi, Real RzHi, const Real *St, Real *Src,
    const CellB0<Real> *B0s = nullptr) {
  Real P[NVAR];
#pragma unroll
  for (int v = 0; v < NVAR; v++) P[v] = St[v];
  to_primitive(P);
  auto face = [&](int f, Real *o) {
#pragma unroll
    for (int v = 0; v < 3; v++) o[v] = B0s->face[f][v];
  };
  auto cell = [&](int q) { return B0s->cell[q]; };
  if (B0s)
    powell_source_with<true>(InvHalf, LxLo, RxLo, LxHi, RxHi, LyLo, RyLo,
                             LyHi, RyHi, LzLo, RzLo, LzHi, RzHi, P, Src,
                             face, cell, B0s->useSource);
  else
    powell_source_with<false>(InvHalf, LxLo, RxLo, LxHi, RxHi, LyLo, RyLo,
                              LyHi, RyHi, LzLo, RzLo, LzHi, RzHi, P, Src,
                              face, cell, 0);
}

} // namespace bats
