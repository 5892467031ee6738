// cs_physics.h -- pointwise physics of the cell update, templated on the
// arithmetic type R (double for the FP64 drop-in, float for the mixed path).
//
// Variable order everywhere: 0 rho, 1-3 v (or rho v), 4 p (or E), 5-7 B.
// Gaussian units: magnetic pressure B^2 / 8 pi.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

namespace csp {

template <typename R> struct C;
template <> struct C<double> {
  static constexpr double pi = 3.14159265358979323846;
  static constexpr double ob8pi = 1.0 / (8.0 * 3.14159265358979323846);
  static constexpr double ob4pi = 1.0 / (4.0 * 3.14159265358979323846);
  static constexpr double sq4pi = 3.5449077018110320546;  // sqrt(4 pi)
};
template <> struct C<float> {
  static constexpr float pi = 3.14159265358979323846f;
  static constexpr float ob8pi = (float)(1.0 / (8.0 * 3.14159265358979323846));
  static constexpr float ob4pi = (float)(1.0 / (4.0 * 3.14159265358979323846));
  static constexpr float sq4pi = 3.5449077018110320546f;
};

__device__ __forceinline__ double rmax(double a, double b) { return fmax(a, b); }
__device__ __forceinline__ float rmax(float a, float b) { return fmaxf(a, b); }
__device__ __forceinline__ double rmin(double a, double b) { return fmin(a, b); }
__device__ __forceinline__ float rmin(float a, float b) { return fminf(a, b); }
__device__ __forceinline__ double rsqrt_(double a) { return sqrt(a); }
__device__ __forceinline__ float rsqrt_(float a) { return sqrtf(a); }
__device__ __forceinline__ double rpow(double a, double b) { return pow(a, b); }
__device__ __forceinline__ float rpow(float a, float b) { return powf(a, b); }
__device__ __forceinline__ double rabs(double a) { return fabs(a); }
__device__ __forceinline__ float rabs(float a) { return fabsf(a); }
// Reciprocal: IEEE division in FP64; the hardware v_rcp_f32 (1 ulp) in FP32,
// where the whole evaluation is already carried at single precision.
__device__ __forceinline__ double rrcp(double a) { return 1.0 / a; }
__device__ __forceinline__ float rrcp(float a) { return __builtin_amdgcn_rcpf(a); }

template <typename R> struct Gas {
  R gamma, gm1, rhofloor, prsfloor, igm1;  // igm1 = 1 / (gamma - 1)
};

// rho^(gamma - 1) for the entropy p / rho^(gamma - 1): cbrt(rho^2) for the
// monatomic gamma = 5/3 (a uniform branch), exp2((gamma - 1) log2 rho)
// otherwise; both within a few ulps of pow at a fraction of its cost.
__device__ __forceinline__ double rpow_gm1(const Gas<double>& gs, double x) {
  return gs.gamma == 5.0 / 3.0 ? cbrt(x * x) : exp2(gs.gm1 * log2(x));
}
__device__ __forceinline__ float rpow_gm1(const Gas<float>& gs, float x) {
  return gs.gamma == (float)(5.0 / 3.0) ? cbrtf(x * x) : exp2f(gs.gm1 * log2f(x));
}

// PRIM_TO_CONS, MHD branch with PRESSURE_POSITIVITY 1
// (661pde_3d_prim_to_cons.f:67-125).  V may be modified by the floors.
template <typename R>
__device__ __forceinline__ void prim_to_cons(const Gas<R>& gs, R ent, R* V, R* U) {
  V[0] = rmax(V[0], gs.rhofloor);
  if (V[4] <= gs.prsfloor) V[4] = rmax(gs.prsfloor, ent * rpow_gm1(gs, V[0]));
  U[0] = V[0];
  U[1] = V[0] * V[1];
  U[2] = V[0] * V[2];
  U[3] = V[0] * V[3];
  U[5] = V[5];
  U[6] = V[6];
  U[7] = V[7];
  U[4] = V[4] * gs.igm1 + R(0.5) * V[0] * (V[1] * V[1] + V[2] * V[2] + V[3] * V[3]) +
         C<R>::ob8pi * (V[5] * V[5] + V[6] * V[6] + V[7] * V[7]);
}

// CONS_TO_PRIM, MHD branch (662pde_3d_cons_to_prim.f:69-151): density floor,
// v = m rho / (rho^2 + rho_f^2), entropy fallback for the pressure, then the
// entropy and the energy are reset from the result.  Returns the flag.
template <typename R>
__device__ __forceinline__ int cons_to_prim(const Gas<R>& gs, R* U, R* V, R& ent) {
  int flag = 0;
  U[0] = rmax(U[0], gs.rhofloor);
  V[0] = U[0];
  R a = V[0] / (V[0] * V[0] + gs.rhofloor * gs.rhofloor);
  V[1] = U[1] * a;
  V[2] = U[2] * a;
  V[3] = U[3] * a;
  V[5] = U[5];
  V[6] = U[6];
  V[7] = U[7];
  R V2 = V[1] * V[1] + V[2] * V[2] + V[3] * V[3];
  R B2 = V[5] * V[5] + V[6] * V[6] + V[7] * V[7];
  V[4] = gs.gm1 * (U[4] - R(0.5) * V[0] * V2 - C<R>::ob8pi * B2);
  if (V[4] <= gs.prsfloor) {
    V[4] = rmax(gs.prsfloor, ent * rpow_gm1(gs, V[0]));
    flag = 1;
  }
  ent = V[4] / rpow_gm1(gs, V[0]);
  U[4] = V[4] * gs.igm1 + R(0.5) * V[0] * V2 + C<R>::ob8pi * B2;
  return flag;
}

// PRIM_TO_WAVSPD, x-direction (660pde_3d_prim_to_wavspd.f:54-76)
template <typename R>
__device__ __forceinline__ void wavspd_x(const Gas<R>& gs, const R* V, R& smin, R& smax) {
  // one reciprocal of rho serves the sound speed and both Alfven terms
  R irho = rrcp(V[0]);
  R cs2 = gs.gamma * V[4] * irho;
  R ob4pr = C<R>::ob4pi * irho;
  R b2 = (V[5] * V[5] + V[6] * V[6] + V[7] * V[7]) * ob4pr;
  R bb = -(cs2 + b2);
  R cc = cs2 * V[5] * V[5] * ob4pr;
  R mf = rsqrt_((-bb + rsqrt_(rmax(bb * bb - R(4) * cc, R(0)))) * R(0.5));
  smin = V[1] - mf;
  smax = V[1] + mf;
}

// x-direction ideal-MHD flux with B0 = 0.  CONS_PRIM_TO_FLUX_3D is called at
// 610_1d_rs_oned_rs_mhd_ptwise.f:129,133 but is absent from the reference
// tree; this is the flux written inline at 600_1d_rs_oned_rs_mhd_dad.f:194-230.
template <typename R>
__device__ __forceinline__ void flux_x(const R* V, const R* U, R* F) {
  R ptot = V[4] + C<R>::ob8pi * (V[5] * V[5] + V[6] * V[6] + V[7] * V[7]);
  R vdb = V[1] * V[5] + V[2] * V[6] + V[3] * V[7];
  for (int c = 0; c < 8; ++c) F[c] = V[1] * U[c];
  F[1] += ptot - C<R>::ob4pi * V[5] * V[5];
  F[2] -= C<R>::ob4pi * V[5] * V[6];
  F[3] -= C<R>::ob4pi * V[5] * V[7];
  F[4] += ptot * V[1] - C<R>::ob4pi * V[5] * vdb;
  F[5] = R(0);
  F[6] -= V[5] * V[2];
  F[7] -= V[5] * V[3];
}

// OneD_RS_MHD_Ptwise, MHD HLL in fluctuation form with the entropy-based
// positivity fix of U* (610_1d_rs_oned_rs_mhd_ptwise.f:96-220, 653-704).
// VL, VR: face-frame primitives (normal B already common).  Returns the
// resolved flux Fs[8], entropy flux, and the extremal speeds.
template <typename R>
__device__ __forceinline__ void hll(const Gas<R>& gs, R* VL, R* VR, R entL, R entR, R* Fs,
                                    R& entF, R& SL, R& SR) {
  R bx = R(0.5) * (VL[5] + VR[5]);
  VL[5] = bx;
  VR[5] = bx;
  R UL[8], UR[8], FL[8], FR[8], Vm[8];
  prim_to_cons(gs, entL, VL, UL);
  prim_to_cons(gs, entR, VR, UR);
  flux_x(VL, UL, FL);
  flux_x(VR, UR, FR);
  R smin, smax;
  SR = R(-1e30);
  SL = R(1e30);
  wavspd_x(gs, VL, smin, smax);
  SL = rmin(SL, smin); SR = rmax(SR, smax);
  wavspd_x(gs, VR, smin, smax);
  SL = rmin(SL, smin); SR = rmax(SR, smax);
  for (int c = 0; c < 8; ++c) Vm[c] = R(0.5) * (VL[c] + VR[c]);
  wavspd_x(gs, Vm, smin, smax);
  SL = rmin(SL, smin); SR = rmax(SR, smax);
  if (SL >= R(0)) {
    for (int c = 0; c < 8; ++c) Fs[c] = FL[c];
    entF = entL * VL[1];
    return;
  }
  if (SR <= R(0)) {
    for (int c = 0; c < 8; ++c) Fs[c] = FR[c];
    entF = entR * VR[1];
    return;
  }
  R t3 = R(1) / (SR - SL);
  R t1 = SR * t3, t2 = SL * t3;
  R Us[8];
  for (int c = 0; c < 8; ++c) Us[c] = t1 * UR[c] - t2 * UL[c] - t3 * (FR[c] - FL[c]);
  R entUs = t1 * entR - t2 * entL - t3 * (entR * VR[1] - entL * VL[1]);
  R pent = entUs * rpow_gm1(gs, Us[0]);
  R peng = gs.gm1 * (Us[4] - R(0.5) * (Us[1] * Us[1] + Us[2] * Us[2] + Us[3] * Us[3]) / Us[0] -
                     C<R>::ob8pi * (Us[5] * Us[5] + Us[6] * Us[6] + Us[7] * Us[7]));
  if (peng <= gs.prsfloor) Us[4] = Us[4] + (pent - peng) * gs.igm1;
  for (int c = 0; c < 8; ++c)
    Fs[c] = R(0.5) * (FL[c] + FR[c]) + R(0.5) * SL * (Us[c] - UL[c]) - R(0.5) * SR * (UR[c] - Us[c]);
  R ob = R(1) / (SR - SL);
  entF = (SR * entL * VL[1] - SL * entR * VR[1]) * ob + (entR - entL) * (SR * SL * ob);
}

// HLLD of Miyoshi & Kusano (J. Comput. Phys. 208, 315, 2005), written in the
// rationalised field b = B / sqrt(4 pi) and converted back for the induction
// flux.  Outer speeds are the same estimates as the HLL solver above.  The
// entropy is upwinded with the mass flux.
template <typename R>
__device__ __forceinline__ void hlld(const Gas<R>& gs, R* VL, R* VR, R entL, R entR, R* Fs,
                                     R& entF, R& SL, R& SR) {
  R bxB = R(0.5) * (VL[5] + VR[5]);
  VL[5] = bxB;
  VR[5] = bxB;
  R UL[8], UR[8], FL[8], FR[8], Vm[8];
  prim_to_cons(gs, entL, VL, UL);
  prim_to_cons(gs, entR, VR, UR);
  flux_x(VL, UL, FL);
  flux_x(VR, UR, FR);
  R smin, smax;
  SR = R(-1e30);
  SL = R(1e30);
  wavspd_x(gs, VL, smin, smax);
  SL = rmin(SL, smin); SR = rmax(SR, smax);
  wavspd_x(gs, VR, smin, smax);
  SL = rmin(SL, smin); SR = rmax(SR, smax);
  for (int c = 0; c < 8; ++c) Vm[c] = R(0.5) * (VL[c] + VR[c]);
  wavspd_x(gs, Vm, smin, smax);
  SL = rmin(SL, smin); SR = rmax(SR, smax);
  if (SL >= R(0)) {
    for (int c = 0; c < 8; ++c) Fs[c] = FL[c];
    entF = entL * VL[1];
    return;
  }
  if (SR <= R(0)) {
    for (int c = 0; c < 8; ++c) Fs[c] = FR[c];
    entF = entR * VR[1];
    return;
  }
  const R s4 = C<R>::sq4pi, is4 = R(1) / C<R>::sq4pi;
  R bx = bxB * is4;
  R byL = VL[6] * is4, bzL = VL[7] * is4, byR = VR[6] * is4, bzR = VR[7] * is4;
  R rL = VL[0], rR = VR[0], uL = VL[1], uR = VR[1];
  R ptL = VL[4] + R(0.5) * (bx * bx + byL * byL + bzL * bzL);
  R ptR = VR[4] + R(0.5) * (bx * bx + byR * byR + bzR * bzR);
  R EL = UL[4], ER = UR[4];
  R dL = (SL - uL) * rL, dR = (SR - uR) * rR;
  const R idd = rrcp(dR - dL);
  R SM = (dR * uR - dL * uL - ptR + ptL) * idd;
  // the contact speed lies strictly between the outer speeds; in rounding it may not
  {
    R margin = (SR - SL) * (sizeof(R) == 4 ? R(1e-6) : R(1e-12));
    SM = rmin(rmax(SM, SL + margin), SR - margin);
  }
  R pts = (dR * ptL - dL * ptR + dL * dR * (uR - uL)) * idd;
  // star states; the degenerate case (Miyoshi & Kusano sec. 3.3) is detected
  // relative to the rounding of the arithmetic type
  const R degen = sizeof(R) == 4 ? R(1e-5) : R(1e-10);
  const R iLm = rrcp(SL - SM), iRm = rrcp(SR - SM);
  R rLs = dL * iLm, rRs = dR * iRm;
  R vyLs, vzLs, byLs, bzLs, vyRs, vzRs, byRs, bzRs;
  {
    R den = dL * (SL - SM) - bx * bx;
    if (rabs(den) > degen * (rabs(dL * (SL - SM)) + bx * bx) && den != R(0)) {
      R iden = rrcp(den);
      R f1 = bx * (SM - uL) * iden;
      vyLs = VL[2] - byL * f1;
      vzLs = VL[3] - bzL * f1;
      R f2 = (dL * (SL - uL) - bx * bx) * iden;
      byLs = byL * f2;
      bzLs = bzL * f2;
    } else {
      vyLs = VL[2]; vzLs = VL[3]; byLs = byL; bzLs = bzL;
    }
  }
  {
    R den = dR * (SR - SM) - bx * bx;
    if (rabs(den) > degen * (rabs(dR * (SR - SM)) + bx * bx) && den != R(0)) {
      R iden = rrcp(den);
      R f1 = bx * (SM - uR) * iden;
      vyRs = VR[2] - byR * f1;
      vzRs = VR[3] - bzR * f1;
      R f2 = (dR * (SR - uR) - bx * bx) * iden;
      byRs = byR * f2;
      bzRs = bzR * f2;
    } else {
      vyRs = VR[2]; vzRs = VR[3]; byRs = byR; bzRs = bzR;
    }
  }
  R vbL = uL * bx + VL[2] * byL + VL[3] * bzL;
  R vbR = uR * bx + VR[2] * byR + VR[3] * bzR;
  R vbLs = SM * bx + vyLs * byLs + vzLs * bzLs;
  R vbRs = SM * bx + vyRs * byRs + vzRs * bzRs;
  R ELs = ((SL - uL) * EL - ptL * uL + pts * SM + bx * (vbL - vbLs)) * iLm;
  R ERs = ((SR - uR) * ER - ptR * uR + pts * SM + bx * (vbR - vbRs)) * iRm;
  R sqL = rsqrt_(rLs), sqR = rsqrt_(rRs);
  R SLs = SM - rabs(bx) * rrcp(sqL), SRs = SM + rabs(bx) * rrcp(sqR);
  // conserved star states in the rationalised field
  R ULs[8] = {rLs, rLs * SM, rLs * vyLs, rLs * vzLs, ELs, bx, byLs, bzLs};
  R URs[8] = {rRs, rRs * SM, rRs * vyRs, rRs * vzRs, ERs, bx, byRs, bzRs};
  R ULr[8] = {UL[0], UL[1], UL[2], UL[3], UL[4], bx, byL, bzL};
  R URr[8] = {UR[0], UR[1], UR[2], UR[3], UR[4], bx, byR, bzR};
  R FLr[8], FRr[8];
  for (int c = 0; c < 5; ++c) { FLr[c] = FL[c]; FRr[c] = FR[c]; }
  FLr[5] = R(0); FRr[5] = R(0);
  FLr[6] = FL[6] * is4; FLr[7] = FL[7] * is4;
  FRr[6] = FR[6] * is4; FRr[7] = FR[7] * is4;
  R F[8];
  if (SLs >= R(0)) {
    for (int c = 0; c < 8; ++c) F[c] = FLr[c] + SL * (ULs[c] - ULr[c]);
  } else if (SRs <= R(0)) {
    for (int c = 0; c < 8; ++c) F[c] = FRr[c] + SR * (URs[c] - URr[c]);
  } else {
    R sg = bx > R(0) ? R(1) : (bx < R(0) ? R(-1) : R(0));
    R isum = rrcp(sqL + sqR);
    R vyss = (sqL * vyLs + sqR * vyRs + (byRs - byLs) * sg) * isum;
    R vzss = (sqL * vzLs + sqR * vzRs + (bzRs - bzLs) * sg) * isum;
    R byss = (sqL * byRs + sqR * byLs + sqL * sqR * (vyRs - vyLs) * sg) * isum;
    R bzss = (sqL * bzRs + sqR * bzLs + sqL * sqR * (vzRs - vzLs) * sg) * isum;
    R vbss = SM * bx + vyss * byss + vzss * bzss;
    if (SM >= R(0)) {
      R ELss = ELs - sqL * (vbLs - vbss) * sg;
      R ULss[8] = {rLs, rLs * SM, rLs * vyss, rLs * vzss, ELss, bx, byss, bzss};
      for (int c = 0; c < 8; ++c)
        F[c] = FLr[c] + SL * (ULs[c] - ULr[c]) + SLs * (ULss[c] - ULs[c]);
    } else {
      R ERss = ERs + sqR * (vbRs - vbss) * sg;
      R URss[8] = {rRs, rRs * SM, rRs * vyss, rRs * vzss, ERss, bx, byss, bzss};
      for (int c = 0; c < 8; ++c)
        F[c] = FRr[c] + SR * (URs[c] - URr[c]) + SRs * (URss[c] - URs[c]);
    }
  }
  for (int c = 0; c < 5; ++c) Fs[c] = F[c];
  Fs[5] = R(0);
  Fs[6] = F[6] * s4;
  Fs[7] = F[7] * s4;
  const bool up = Fs[0] >= R(0);
  entF = Fs[0] * ((up ? entL : entR) * rrcp(up ? VL[0] : VR[0]));
}

// Jiang & Shu (1996) WENO5 value at the right edge of cell v2 from
// v0..v4 = cells i-2..i+2 (the left state of face i+1/2):
//   w_k = a_k / sum a,  a_k = g_k / (eps + beta_k)^2,  g = (0.1, 0.6, 0.3).
// FP64 multiplies every a_k by d0 d1 d2 (d_k = (eps + beta_k)^2), which
// leaves the weights unchanged and needs one division instead of four; the
// products stay far inside the FP64 range.  FP32, whose range the products
// could leave, uses four hardware reciprocals.
template <typename R>
__device__ __forceinline__ R weno5(R v0, R v1, R v2, R v3, R v4) {
  const R eps = R(1e-6);
  R s0 = v0 - R(2) * v1 + v2, t0 = v0 - R(4) * v1 + R(3) * v2;
  R s1 = v1 - R(2) * v2 + v3, t1 = v1 - v3;
  R s2 = v2 - R(2) * v3 + v4, t2 = R(3) * v2 - R(4) * v3 + v4;
  R e0 = eps + (R(13.0 / 12.0) * s0 * s0 + R(0.25) * t0 * t0);
  R e1 = eps + (R(13.0 / 12.0) * s1 * s1 + R(0.25) * t1 * t1);
  R e2 = eps + (R(13.0 / 12.0) * s2 * s2 + R(0.25) * t2 * t2);
  R d0 = e0 * e0, d1 = e1 * e1, d2 = e2 * e2;
  // 6 x the three candidate values
  R q0 = R(2) * v0 - R(7) * v1 + R(11) * v2;
  R q1 = -v1 + R(5) * v2 + R(2) * v3;
  R q2 = R(2) * v2 + R(5) * v3 - v4;
  R a0, a1, a2;
  if (sizeof(R) == 8) {
    a0 = R(0.1) * d1 * d2; a1 = R(0.6) * d0 * d2; a2 = R(0.3) * d0 * d1;
  } else {
    a0 = R(0.1) * rrcp(d0); a1 = R(0.6) * rrcp(d1); a2 = R(0.3) * rrcp(d2);
  }
  R num = a0 * q0 + a1 * q1 + a2 * q2, den = R(6) * (a0 + a1 + a2);
  return sizeof(R) == 8 ? num / den : num * rrcp(den);
}

// The two WENO5 values at a wall whose ghost zones mirror the zone:
// stencils (a, b, a, b, a) and (b, a, b, a, b) with a the zone and b its
// mirror image, component by component.  Then beta_0 = beta_2 = 25/3 D and
// beta_1 = 13/3 D with D = (a - b)^2, the candidates are (13a - 7b)/6,
// (5a + b)/6, (a + 5b)/6, and the weights depend on D only, so
//   W(a,b) = (e1 (1.6 a + 0.8 b) + e0 (3 a + 0.6 b)) / (2.4 e1 + 3.6 e0),
//   e0 = (eps + 25/3 D)^2, e1 = (eps + 13/3 D)^2,
// and W(b,a) = a + b - W(a,b) (the candidates of the swapped stencil sum
// with those of the original to a + b under the same weights).
template <typename R>
__device__ __forceinline__ void weno5_mirror(R a, R b, R& wab, R& wba) {
  const R eps = R(1e-6);
  R D = (a - b) * (a - b);
  R e0 = eps + R(25.0 / 3.0) * D, e1 = eps + R(13.0 / 3.0) * D;
  e0 = e0 * e0;
  e1 = e1 * e1;
  R num = e1 * (R(1.6) * a + R(0.8) * b) + e0 * (R(3) * a + R(0.6) * b);
  R den = R(2.4) * e1 + R(3.6) * e0;
  wab = sizeof(R) == 8 ? num / den : num * rrcp(den);
  wba = (a + b) - wab;
}

// Two_D_RS_Ez_4State_HLL (630_2d_rs_ez_two_d_rs_ez_4state_hll.f:10-372):
// multidimensional HLL solver for the induction equation on a skew mesh.
template <typename R>
__device__ __forceinline__ R emf4_hll(R S_R, R S_L, R S_U, R S_D, R cos_t, R sin_t, R Ez_RU,
                                      R Ez_LU, R Ez_LD, R Ez_RD, R Bneta_U, R Bneta_D, R By_R,
                                      R By_L) {
  R obS_x = R(1) / (S_R - S_L);
  R obS_y = R(1) / (S_U - S_D);
  R csc_t = R(1) / sin_t;
  R cot_t = cos_t / sin_t;
  R Bx_RU = (Bneta_U + By_R * cos_t) / sin_t, By_RU = By_R;
  R Bx_LU = (Bneta_U + By_L * cos_t) / sin_t, By_LU = By_L;
  R Bx_RD = (Bneta_D + By_R * cos_t) / sin_t, By_RD = By_R;
  R Bx_LD = (Bneta_D + By_L * cos_t) / sin_t, By_LD = By_L;
  R Btau_RU = Bx_RU * cos_t + By_RU * sin_t;
  R Btau_LU = Bx_LU * cos_t + By_LU * sin_t;
  R Btau_RD = Bx_RD * cos_t + By_RD * sin_t;
  R Btau_LD = Bx_LD * cos_t + By_LD * sin_t;
  if (S_L >= R(0) && S_D >= R(0)) return Ez_LD;
  if (S_R <= R(0) && S_D >= R(0)) return Ez_RD;
  if (S_R <= R(0) && S_U <= R(0)) return Ez_RU;
  if (S_L >= R(0) && S_U <= R(0)) return Ez_LU;
  if (S_L >= R(0))
    return (S_U * Ez_LD - S_D * Ez_LU) * obS_y + S_U * S_D * (Bx_LU - Bx_LD) * obS_y;
  if (S_R <= R(0))
    return (S_U * Ez_RD - S_D * Ez_RU) * obS_y + S_U * S_D * (Bx_RU - Bx_RD) * obS_y;
  if (S_D >= R(0))
    return (S_R * Ez_LD - S_L * Ez_RD) * obS_x - S_R * S_L * (Btau_RD - Btau_LD) * obS_x;
  if (S_U <= R(0))
    return (S_R * Ez_LU - S_L * Ez_RU) * obS_x - S_R * S_L * (Btau_RU - Btau_LU) * obS_x;
  // fully subsonic
  R BSx_L = (S_U * Bx_LU - S_D * Bx_LD) * obS_y - (Ez_LU - Ez_LD) * obS_y;
  R BSy_L = By_L;
  R ESz_L = (S_U * Ez_LD - S_D * Ez_LU) * obS_y + S_U * S_D * (Bx_LU - Bx_LD) * obS_y;
  R BSx_R = (S_U * Bx_RU - S_D * Bx_RD) * obS_y - (Ez_RU - Ez_RD) * obS_y;
  R BSy_R = By_R;
  R ESz_R = (S_U * Ez_RD - S_D * Ez_RU) * obS_y + S_U * S_D * (Bx_RU - Bx_RD) * obS_y;
  R BSneta = Bneta_D;
  R BStau = (S_R * Btau_RD - S_L * Btau_LD) * obS_x + (Ez_RD - Ez_LD) * obS_x;
  R BSx_D = BSneta * sin_t + BStau * cos_t;
  R BSy_D = -BSneta * cos_t + BStau * sin_t;
  R ESz_D = (S_R * Ez_LD - S_L * Ez_RD) * obS_x - S_R * S_L * (Btau_RD - Btau_LD) * obS_x;
  BSneta = Bneta_U;
  BStau = (S_R * Btau_RU - S_L * Btau_LU) * obS_x + (Ez_RU - Ez_LU) * obS_x;
  R BSx_U = BSneta * sin_t + BStau * cos_t;
  R BSy_U = -BSneta * cos_t + BStau * sin_t;
  R ESz_U = (S_R * Ez_LU - S_L * Ez_RU) * obS_x - S_R * S_L * (Btau_RU - Btau_LU) * obS_x;
  R x_c = R(0.5) * ((S_U + S_D) * cos_t + (S_R + S_L)) / sin_t;
  R y_c = R(0.5) * (S_U + S_D);
  R BSSx = -R(0.5) * ((-cos_t * ESz_R - S_R * BSx_R) * obS_x + (ESz_U - S_U * BSx_U) * obS_y +
                      (cos_t * ESz_L + S_L * BSx_L) * obS_x + (-ESz_D + S_D * BSx_D) * obS_y);
  R BSSy = -R(0.5) * ((-sin_t * ESz_R - S_R * BSy_R) * obS_x + (-S_U * BSy_U) * obS_y +
                      (sin_t * ESz_L + S_L * BSy_L) * obS_x + (S_D * BSy_D) * obS_y);
  R ESSz_1 = -x_c * BSSy - R(0.5) * (csc_t * (-sin_t * ESz_R - S_R * BSy_R) + cot_t * (-S_U * BSy_U) -
                                    csc_t * (sin_t * ESz_L + S_L * BSy_L) - cot_t * (S_D * BSy_D));
  R ESSz_2 = y_c * BSSx + R(0.5) * ((ESz_U - S_U * BSx_U) - (-ESz_D + S_D * BSx_D));
  return R(0.5) * (ESSz_1 + ESSz_2);
}

// Two_D_RS_Ez_3State_LLF (633_2d_rs_ez_two_d_rs_ez_3state_llf.f:13-86),
// used at the eight cube corners where three sectors meet.
template <typename R>
__device__ __forceinline__ R emf3_llf(R S, R Ez_1, R Ez_2, R Ez_3, R Bneta_1, R Bneta_2,
                                      R Bneta_3) {
  R Sq3 = rsqrt_(R(3));
  R ob3 = R(1) / R(3);
  return ob3 * (Ez_1 + Ez_2 + Ez_3) - (R(4) * Sq3 * S / R(9)) * (Bneta_1 + Bneta_2 + Bneta_3);
}


// The stage update of one zone, u = combination of the stage input ui (ei),
// U^n uz (ez) and the flux divergence L (L[8]: entropy), with the radial
// gravity g = gr along the unit vector n:
//  mode 0: explicit source, Shu-Osher  u = wa uz + wb (ui + dt (L + S(ui))).
//  IMEX-SSP2(2,2,2) (Pareschi & Russo 2005; explicit part = Heun, the
//  reference's RK2), gravity implicit, gi = 1 - 1/sqrt(2):
//  mode 1: ui = U1, uz = Un:  U2 = Un + dt L(U1) + (1 - 2 gi) dt S(U1) + gi dt S(U2).
//          S leaves rho unchanged and is linear in rho v, so U2 is explicit:
//          rho v then E take the implicit term in turn.
//  mode 2: ui = U2, uz = Un:  Un+1 = (Un + U2 + dt L(U2)) / 2 + dt/2 (2 gi S(U1) + (1 - gi) S(U2)),
//          with U1 = Un + gi dt S(U1) rebuilt from Un.
// uz and ez are read only when mode != 0 or wa != 0.
__device__ __forceinline__ void rk_combine(const double* ui, double ei, const double* uz, double ez, double dt,
                                           double wa, double wb, int mode, double gr, const double* n,
                                           const double* L, double* u, double& en) {
  if (mode == 0) {
    double Lg[5] = {L[0], L[1], L[2], L[3], L[4]};
    if (gr != 0.0) {
      const double mr = ui[1] * n[0] + ui[2] * n[1] + ui[3] * n[2];
      for (int q = 0; q < 3; ++q) Lg[1 + q] += ui[0] * gr * n[q];
      Lg[4] += mr * gr;
    }
    for (int q = 0; q < 8; ++q) {
      double un = ui[q] + dt * (q < 5 ? Lg[q] : L[q]);
      u[q] = wa != 0.0 ? wa * uz[q] + wb * un : wb * un;
    }
    double e1 = ei + dt * L[8];
    en = wa != 0.0 ? wa * ez + wb * e1 : wb * e1;
    return;
  }
  const double gi = 1.0 - 0.70710678118654752440;
  if (mode == 1) {
    const double s1m = ui[0] * gr, s1e = gr * (ui[1] * n[0] + ui[2] * n[1] + ui[3] * n[2]);
    u[0] = uz[0] + dt * L[0];
    for (int q = 1; q < 4; ++q) u[q] = uz[q] + dt * L[q] + (1.0 - 2.0 * gi) * dt * s1m * n[q - 1];
    u[4] = uz[4] + dt * L[4] + (1.0 - 2.0 * gi) * dt * s1e;
    for (int q = 1; q < 4; ++q) u[q] += gi * dt * u[0] * gr * n[q - 1];
    u[4] += gi * dt * gr * (u[1] * n[0] + u[2] * n[1] + u[3] * n[2]);
    for (int q = 5; q < 8; ++q) u[q] = uz[q] + dt * L[q];
    en = ez + dt * L[8];
    return;
  }
  const double s2m = ui[0] * gr, s2e = gr * (ui[1] * n[0] + ui[2] * n[1] + ui[3] * n[2]);
  const double s1m = uz[0] * gr, s1e = gr * ((uz[1] * n[0] + uz[2] * n[1] + uz[3] * n[2]) + gi * dt * uz[0] * gr);
  for (int q = 0; q < 8; ++q) u[q] = 0.5 * uz[q] + 0.5 * (ui[q] + dt * L[q]);
  for (int q = 1; q < 4; ++q) u[q] += 0.5 * dt * (2.0 * gi * s1m + (1.0 - gi) * s2m) * n[q - 1];
  u[4] += 0.5 * dt * (2.0 * gi * s1e + (1.0 - gi) * s2e);
  en = 0.5 * ez + 0.5 * (ei + dt * L[8]);
}

// The implicit first stage of IMEX-SSP2(2,2,2): U1 = Un + gi dt S(U1) in
// closed form (rho unchanged, then rho v, then E).
__device__ __forceinline__ void imex_pre(double dt, double gr, const double* n, double* u) {
  const double gi = 1.0 - 0.70710678118654752440;
  for (int q = 1; q < 4; ++q) u[q] += gi * dt * u[0] * gr * n[q - 1];
  u[4] += gi * dt * gr * (u[1] * n[0] + u[2] * n[1] + u[3] * n[2]);
}

}  // namespace csp
