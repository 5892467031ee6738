// cs_slopes_p1.hip -- the reference's second-order zone reconstruction
// LIMIT_ALL_SLOPES_P1 (cubesphere/400riemann_integrate/410limit_all_slopes_p1.f:9)
// for one patch on the GPU, FP64, behind cs_limit_all_slopes_p1.
//
// The patch arrays keep the reference's column-major layouts (first index
// fastest), so a thread per zone or face reads consecutive addresses along x.
// Each kernel is one dependency level of the routine:
//   k_prim        steps 2 and 5: constants, higher modes cleared, tri-corner
//                 zones made safe, CONS_TO_PRIM, magnetosonic speed
//   k_face_zero   step 2 (faces): face B between two inactive zones cleared
//   k_tricorner   step 3: face B synchronised at the tri-corners
//   k_flat_a/b/c  step 6.a-c: velocity-divergence flattener, its pressure-
//                 ratio extension, its spread to the six neighbours
//   k_weno9       step 7.a: multi-D WENO of the 8 primitives and the entropy,
//                 one thread per three fields of a zone
//   k_pos         step 7.b: ENFORCE_PRESSURE_POSITIVITY_PRIM (440...f)
//   k_face_pts    step 8.a-c: normal B at the five B-field points of each face
//   k_clsq_e      step 8.d and 9.a: ADP constrained least squares for the 12
//                 linear B coefficients; zone E = -v x B
//   k_weno9       step 9.b: WENO of E (the same kernel, three fields)
//   k_diag(+final) step 10: min/max of the primitives, face-B divergence
//
// The constrained least squares of 8.d (the reference assembles the 18 x 18
// KKT matrix and factors it with Numerical Recipes' LU) is solved here through
// the Schur complement: M = A^T A (12 x 12, >= identity) by Cholesky, then the
// 6 x 6 system C M^-1 C^T lambda = C M^-1 A^T b - d, all in registers.
//
// The positivity limiter of the reference runs in loop order and, when a zone
// fails every reduction, also zeroes its entropy mean (440...f:553-600), which
// later zones of the loop then see in their entropy range.  k_pos runs all
// zones with the incoming entropy means and flags the zones that fail
// completely; when any do, the zones with such a neighbour earlier in the
// loop order (k_cands) are recomputed (k_weno9, k_pos on that list) with the
// means of the flagged earlier zones taken as zero, repeating until the flags
// no longer change; a zone flagged for the first time in a sweep brings its
// later neighbours onto the list before the next sweep (the dependence only
// runs forward in the loop order, so this reaches the sequential result), and
// the flagged means are then cleared.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "cs_physics.h"
#include "cs_p1_common.h"
#include "../../include/cubesphere_hip.h"

namespace cs {
void set_error(const char* fmt, ...);
}

using namespace csp;

namespace csp1 {

struct P1 {
  cs_p1_patch a;
  int N, NR, g, nz, nzr, nf, nfr;
  int64_t NZ, NXF, NZF, NFC;

  __device__ __forceinline__ int64_t zi(int i, int j, int k) const {
    return (int64_t)(i - 1 + g) + (int64_t)nz * ((j - 1 + g) + (int64_t)nz * (k - 1 + g));
  }
  __device__ __forceinline__ int64_t fci(int i, int j, int k) const {
    return (int64_t)(i + g) + (int64_t)nf * ((j + g) + (int64_t)nf * (k + g));
  }
  __device__ __forceinline__ int64_t xfi(int i, int j, int k) const {
    return (int64_t)(i + g) + (int64_t)nf * ((j - 1 + g) + (int64_t)nz * (k - 1 + g));
  }
  __device__ __forceinline__ int64_t yfi(int i, int j, int k) const {
    return (int64_t)(i - 1 + g) + (int64_t)nz * ((j + g) + (int64_t)nf * (k - 1 + g));
  }
  __device__ __forceinline__ int64_t zfi(int i, int j, int k) const {
    return (int64_t)(i - 1 + g) + (int64_t)nz * ((j - 1 + g) + (int64_t)nz * (k + g));
  }
  // stencil entry n (R L U D RR LL UU DD RU LU LD RD) of zone column (i, j),
  // clamped to the zone range so that no table entry can address outside it
  __device__ __forceinline__ void nbr(int i, int j, int n, int& ni, int& nj) const {
    const int64_t s = (int64_t)(i - 1 + g) + (int64_t)nz * (j - 1 + g), plane = (int64_t)nz * nz;
    ni = min(max(a.stencil[s + plane * (2 * n)], 1 - g), N + g);
    nj = min(max(a.stencil[s + plane * (2 * n + 1)], 1 - g), N + g);
  }
  __device__ __forceinline__ bool active(int i, int j, int k) const { return a.mask[zi(i, j, k)] == 1; }
  __device__ __forceinline__ double V(int c, int m, int64_t z) const { return a.V[z + NZ * (c + 8 * m)]; }
  __device__ __forceinline__ double& Vr(int c, int m, int64_t z) const { return a.V[z + NZ * (c + 8 * m)]; }
  __device__ __forceinline__ double& Ur(int c, int m, int64_t z) const { return a.U[z + NZ * (c + 8 * m)]; }
  __device__ __forceinline__ double obl(int64_t z) const { return 1.0 / fmax(a.scale[z], a.smallnum); }
  __device__ __forceinline__ double cen(int64_t z, int c) const { return a.v_centroid[z + NZ * c]; }
};

// a box of zone or face indices, decoded from a flat thread index
struct Box {
  int i0, j0, k0, ni, nj, nk;
  __device__ __forceinline__ bool decode(int64_t t, int& i, int& j, int& k) const {
    if (t >= (int64_t)ni * nj * nk) return false;
    i = i0 + (int)(t % ni);
    j = j0 + (int)((t / ni) % nj);
    k = k0 + (int)(t / ((int64_t)ni * nj));
    return true;
  }
  __host__ __device__ int64_t count() const { return (int64_t)ni * nj * nk; }
};
static Box make_box(int i0, int i1, int j0, int j1, int k0, int k1) {
  return Box{i0, j0, k0, i1 - i0 + 1, j1 - j0 + 1, k1 - k0 + 1};
}

__device__ __forceinline__ int64_t tid() { return (int64_t)blockIdx.x * blockDim.x + threadIdx.x; }

// ---------------------------------------------------------------------------
// steps 2 and 5
__global__ void __launch_bounds__(256) k_prim(const P1 P, const Box b) {
  int i, j, k;
  if (!b.decode(tid(), i, j, k)) return;
  const cs_p1_patch& a = P.a;
  const int64_t z = P.zi(i, j, k);
  const double gam = a.gamma;
  a.gamma_eos[z] = gam;
  a.gamma_sound[z] = gam;
  a.mean_mol_wt[z] = a.molewt;
#pragma unroll
  for (int c = 0; c < 8; ++c)
#pragma unroll
    for (int m = 1; m < 4; ++m) {
      P.Ur(c, m, z) = 0.0;
      P.Vr(c, m, z) = 0.0;
    }
#pragma unroll
  for (int m = 1; m < 4; ++m) a.ent[z + P.NZ * m] = 0.0;
  double U[8], V[8];
  if (a.mask[z] == 0) {
#pragma unroll
    for (int c = 0; c < 8; ++c) U[c] = 0.0;
    U[0] = a.rhofloor;
    U[4] = a.prsfloor / (gam - 1.0);
    a.scale[z] = a.smallnum;
  } else {
#pragma unroll
    for (int c = 0; c < 8; ++c) U[c] = P.Ur(c, 0, z);
  }
  double e = a.ent[z];
  Gas<double> gs{gam, gam - 1.0, a.rhofloor, a.prsfloor, 1.0 / (gam - 1.0)};
  V[4] = 0.0;
  cons_to_prim(gs, U, V, e);
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    P.Vr(c, 0, z) = V[c];
    P.Ur(c, 0, z) = U[c];
  }
  a.ent[z] = e;
  const double b2 = (V[5] * V[5] + V[6] * V[6] + V[7] * V[7]) * C<double>::ob4pi;
  a.msonic[z] = sqrt((gam * V[4] + b2) / V[0]);
}

// step 2, faces: blockIdx.y = family
__global__ void __launch_bounds__(256) k_face_zero(const P1 P) {
  const int fam = blockIdx.y;
  const int n = P.nf;
  const int64_t t = tid();
  if (t >= P.NFC) return;
  const int i = (int)(t % n) - P.g, j = (int)((t / n) % n) - P.g, k = (int)(t / ((int64_t)n * n)) - P.g;
  const int lo = 1 - P.g, hi = P.N + P.g, hr = P.NR + P.g;
  if (fam == 0) {
    if (j < lo || k < lo || k > hr) return;
    if (!P.active(max(i, lo), j, k) && !P.active(min(i + 1, hi), j, k)) P.a.Bxf[t] = 0.0;
  } else if (fam == 1) {
    if (i < lo || k < lo || k > hr) return;
    if (!P.active(i, max(j, lo), k) && !P.active(i, min(j + 1, hi), k)) P.a.Byf[t] = 0.0;
  } else {
    if (i < lo || j < lo) return;
    if (!P.active(i, j, 1)) P.a.Bzf[t] = 0.0;
  }
}

// step 3: one thread per (corner, ghost column, radial layer)
__global__ void __launch_bounds__(256) k_tricorner(const P1 P) {
  const int64_t t = tid();
  const int g = P.g, nzr = P.nzr;
  if (t >= 4 * (int64_t)g * nzr) return;
  const int corner = (int)(t / ((int64_t)g * nzr)), ii = (int)((t / nzr) % g), k = (int)(t % nzr) + 1 - g;
  if (!P.a.itri[corner]) return;
  const bool left = corner == 0 || corner == 3, down = corner < 2;
  const int i = left ? 1 - g + ii : P.N + 1 + ii, jj = down ? 1 : P.N;
  int a, b;
  P.nbr(i, jj, down ? 3 : 2, a, b);
  double& by = P.a.Byf[P.fci(i, down ? jj - 1 : jj, k)];
  double& bx = P.a.Bxf[P.fci(left ? a - 1 : a, b, k)];
  if (corner == 0 || corner == 2) {
    const double s = 0.5 * (by - bx);
    by = s;
    bx = -s;
  } else {
    const double s = 0.5 * (by + bx);
    by = s;
    bx = s;
  }
}

// ---------------------------------------------------------------------------
// step 6
__device__ __forceinline__ void six(const P1& P, int i, int j, int k, int64_t* z6) {
  int a, b;
  P.nbr(i, j, 0, a, b);
  z6[0] = P.zi(a, b, k);
  P.nbr(i, j, 1, a, b);
  z6[1] = P.zi(a, b, k);
  P.nbr(i, j, 2, a, b);
  z6[2] = P.zi(a, b, k);
  P.nbr(i, j, 3, a, b);
  z6[3] = P.zi(a, b, k);
  z6[4] = P.zi(i, j, k + 1);
  z6[5] = P.zi(i, j, k - 1);
}

__device__ __forceinline__ bool in_flat_range(const P1& P, int i, int j, int k) {
  const int lo = 2 - P.g, hx = P.N + P.g - 1, hr = P.NR + P.g - 1;
  return i >= lo && i <= hx && j >= lo && j <= hx && k >= lo && k <= hr;
}

__global__ void __launch_bounds__(256) k_flat_a(const P1 P, const Box b) {
  int i, j, k;
  if (!b.decode(tid(), i, j, k)) return;
  const cs_p1_patch& a = P.a;
  const int64_t z = P.zi(i, j, k);
  if (!in_flat_range(P, i, j, k)) {
    a.flatten[z] = 0.0;
    return;
  }
  int64_t z6[6];
  six(P, i, j, k, z6);
  const double obl = P.obl(z);
  double wmin = a.msonic[z];
#pragma unroll
  for (int q = 0; q < 6; ++q) wmin = fmin(wmin, a.msonic[z6[q]]);
  // the unit radial direction of the top z-face serves both z-faces
  double nzv[3];
  {
    const int64_t f = P.zfi(i, j, k);
    double s = 0.0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      nzv[c] = a.zf_centroid[f + P.NZF * c];
    }
    s = 1.0 / fmax(sqrt(nzv[0] * nzv[0] + nzv[1] * nzv[1] + nzv[2] * nzv[2]), a.smallnum);
#pragma unroll
    for (int c = 0; c < 3; ++c) nzv[c] = nzv[c] * s;
  }
  double dv = 0.0;
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    const bool up = (q % 2) == 0;
    const int64_t zo = up ? z6[q] : z, zn = up ? z : z6[q];
    double r[3], n[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) r[c] = (P.cen(zo, c) - P.cen(zn, c)) * obl;
    const double a11 = 0.5 / fmax(sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]), a.smallnum);
    if (q < 2) {
      const int64_t f = P.xfi(q == 0 ? i : i - 1, j, k);
#pragma unroll
      for (int c = 0; c < 3; ++c) n[c] = a.xf_normal[f + P.NXF * c];
    } else if (q < 4) {
      const int64_t f = P.yfi(i, q == 2 ? j : j - 1, k);
#pragma unroll
      for (int c = 0; c < 3; ++c) n[c] = a.yf_normal[f + P.NXF * c];
    } else {
#pragma unroll
      for (int c = 0; c < 3; ++c) n[c] = nzv[c];
    }
    dv = dv + n[0] * a11 * (P.V(1, 0, zo) - P.V(1, 0, zn)) + n[1] * a11 * (P.V(2, 0, zo) - P.V(2, 0, zn)) +
         n[2] * a11 * (P.V(3, 0, zo) - P.V(3, 0, zn));
  }
  a.divvel[z] = dv;
  const double a13 = fabs(dv) / (a.flatten_coef * wmin) - 1.0;
  double f = a13 <= 0.0 ? 0.0 : fmin(1.0, a13);
  if (a.mask[z] != 1) f = 0.0;
  a.flatten[z] = f;
}

// flatten (stage a) -> work (stage b)
__global__ void __launch_bounds__(256) k_flat_b(const P1 P, const Box b) {
  int i, j, k;
  if (!b.decode(tid(), i, j, k)) return;
  const cs_p1_patch& a = P.a;
  const int64_t z = P.zi(i, j, k);
  double f = a.flatten[z];
  if (in_flat_range(P, i, j, k)) {
    int64_t z6[6];
    six(P, i, j, k, z6);
    const double p0 = a.prs_flatten_coef * P.V(4, 0, z);
#pragma unroll
    for (int q = 0; q < 6; ++q)
      if (P.V(4, 0, z6[q]) > p0) f = fmax(f, a.flatten[z6[q]]);
  }
  a.work[z] = f;
}

// work (stage b) -> flatten (stage c)
__global__ void __launch_bounds__(256) k_flat_c(const P1 P, const Box b) {
  int i, j, k;
  if (!b.decode(tid(), i, j, k)) return;
  const cs_p1_patch& a = P.a;
  const int64_t z = P.zi(i, j, k);
  double f = a.work[z];
  if (in_flat_range(P, i, j, k)) {
    int64_t z6[6];
    six(P, i, j, k, z6);
#pragma unroll
    for (int q = 0; q < 6; ++q) f = fmax(f, a.work[z6[q]]);
  }
  a.flatten[z] = f;
}

// ---------------------------------------------------------------------------
// WENO (410...f:2421-3761, 5437-7100): stencil s uses the x-neighbour
// (R for s in {0,1,4,5}, else L), the y-neighbour (U for s < 4, else D) and the
// z-neighbour (top for even s, else bottom), each stencil's 3 x 3 offset
// matrix inverted by its adjugate (inv3)
__device__ __forceinline__ bool in_weno_range(const P1& P, int i, int j, int k) {
  return i >= -2 && i <= P.N + 3 && j >= -2 && j <= P.N + 3 && k >= -2 && k <= P.NR + 3;
}

// ---------------------------------------------------------------------------
// step 7.a: WENO of the 8 primitives and the entropy, one thread per three
// fields of a zone of the range (or of a zone list).  Inactive columns keep
// the zero modes k_prim left.
__device__ __forceinline__ void zone_of(const P1& P, int64_t z, int& i, int& j, int& k) {
  i = (int)(z % P.nz) + 1 - P.g;
  j = (int)((z / P.nz) % P.nz) + 1 - P.g;
  k = (int)(z / ((int64_t)P.nz * P.nz)) + 1 - P.g;
}

// the fields of one WENO launch: the mean of field f at src[f][z], its three
// gradient modes at dst[f][z + m * stride[f]], m = 0, 1, 2
struct WenoFields {
  const double* src[9];
  double* dst[9];
  int64_t stride[9];
  int n;
  uint32_t noflat;  // bit f: field f's gradients are not multiplied by (1 - flattener)
};

// The WENO of up to three fields of one zone in one pass over the 8
// stencils: per stencil its 3 x 3 inverse (formed once for the three
// fields), each field's gradient X_s, its weight
// w_s = 1 / (|X_s|^2 + smallnum)^4, and the running sums of w_s X_s, w_s and
// X_s; the central candidate (the mean of the 8 gradients, weight 100) is
// added at the end and the sum normalised.  The stencil's neighbours are
// chosen by selects: indexing the small arrays by the loop counter would put
// them in scratch.
__global__ void __launch_bounds__(128) k_weno9(const P1 P, const Box b, const int64_t* list, int64_t nlist,
                                               const WenoFields W9) {
  // thread t: fields 3 (t / n) .. 3 (t / n) + 2 of zone t % n
  const int64_t t = tid(), n = list ? nlist : b.count();
  const int ngroups = (W9.n + 2) / 3;
  if (t >= ngroups * n) return;
  const int c0 = 3 * (int)(t / n);
  int i, j, k;
  int64_t z;
  if (list) {
    z = list[t % n];
    zone_of(P, z, i, j, k);
  } else {
    b.decode(t % n, i, j, k);
    z = P.zi(i, j, k);
  }
  if (!P.active(i, j, 1)) return;
  const cs_p1_patch& a = P.a;
  int64_t z6[6];
  six(P, i, j, k, z6);
  const double obl = P.obl(z);
  double off[6][3];
#pragma unroll
  for (int q = 0; q < 6; ++q)
#pragma unroll
    for (int c = 0; c < 3; ++c) off[q][c] = (P.cen(z6[q], c) - P.cen(z, c)) * obl;
  const int nf = 3;  // both launches hold 9 or 3 fields: every group has three
  const double smallnum = a.smallnum, omf = 1.0 - a.flatten[z];
  // the group's field pointers by selects among the table's static entries
  // (c0 is 0, 3 or 6; indexing the argument table at run time would copy it
  // to scratch)
  const double* srcs[3];
  double* dsts[3];
  int64_t strides[3];
#pragma unroll
  for (int f = 0; f < 3; ++f) {
    srcs[f] = c0 == 0 ? W9.src[f] : (c0 == 3 ? W9.src[3 + f] : W9.src[6 + f]);
    dsts[f] = c0 == 0 ? W9.dst[f] : (c0 == 3 ? W9.dst[3 + f] : W9.dst[6 + f]);
    strides[f] = c0 == 0 ? W9.stride[f] : (c0 == 3 ? W9.stride[3 + f] : W9.stride[6 + f]);
  }
  // differences to the six neighbours, one array per neighbour (a select
  // between two entries of one array would be turned into an indexed load)
  double dR[3], dL[3], dU[3], dD[3], dT[3], dB[3], num[3][3], den[3], sx[3][3];
#pragma unroll
  for (int f = 0; f < 3; ++f) {
    const double* src = srcs[f];
    const double f0 = src[z];
    dR[f] = src[z6[0]] - f0;
    dL[f] = src[z6[1]] - f0;
    dU[f] = src[z6[2]] - f0;
    dD[f] = src[z6[3]] - f0;
    dT[f] = src[z6[4]] - f0;
    dB[f] = src[z6[5]] - f0;
    den[f] = 0.0;
#pragma unroll
    for (int m = 0; m < 3; ++m) num[f][m] = sx[f][m] = 0.0;
  }
#pragma unroll 1
  for (int s = 0; s < 8; ++s) {
    const bool bx = (s >> 1) & 1, by = (s >> 2) & 1, bz = s & 1;
    double m3[3][3], A[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      m3[0][c] = bx ? off[1][c] : off[0][c];
      m3[1][c] = by ? off[3][c] : off[2][c];
      m3[2][c] = bz ? off[5][c] : off[4][c];
    }
    inv3(m3, A);
#pragma unroll
    for (int f = 0; f < 3; ++f) {
      const double r0 = bx ? dL[f] : dR[f], r1 = by ? dD[f] : dU[f], r2 = bz ? dB[f] : dT[f];
      double X[3];
#pragma unroll
      for (int m = 0; m < 3; ++m) X[m] = A[m][0] * r0 + A[m][1] * r1 + A[m][2] * r2;
      double tt = X[0] * X[0] + X[1] * X[1] + X[2] * X[2] + smallnum;
      tt = tt * tt;
      const double w = 1.0 / (tt * tt);
      den[f] = s == 0 ? w : den[f] + w;
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        num[f][m] += w * X[m];
        sx[f][m] += X[m];
      }
    }
  }
#pragma unroll
  for (int f = 0; f < 3; ++f) {
    if (f >= nf) break;
    double X8[3];
#pragma unroll
    for (int m = 0; m < 3; ++m) X8[m] = 0.125 * sx[f][m];
    double tt = X8[0] * X8[0] + X8[1] * X8[1] + X8[2] * X8[2] + smallnum;
    tt = tt * tt;
    const double w8 = 100.0 / (tt * tt);
    const double os = 1.0 / (den[f] + w8);
    double* dst = dsts[f];
    const int64_t stride = strides[f];
    const double of = ((W9.noflat >> (c0 + f)) & 1u) ? 1.0 : omf;
#pragma unroll
    for (int m = 0; m < 3; ++m) dst[z + stride * m] = (num[f][m] + w8 * X8[m]) * os * of;
  }
}

// ---------------------------------------------------------------------------
// step 7.b: ENFORCE_PRESSURE_POSITIVITY_PRIM (440...f:6-830).
// pass 0: every zone of the range with the incoming entropy means; sets fa and
// counts the zones whose slopes were removed.
// pass 1: the zones of a list (those with an fa zone earlier in the loop order
// among their 27), taking the entropy means of the prev-flagged earlier zones
// as zero; writes next and counts the zones whose flag changed.
struct Flags {
  uint8_t *fa, *prev, *next;
  unsigned int* counter;
};

// the 27 zones around (i, j, k): the column and its 8 stencil neighbours
// (R, L, U, D, RU, LU, LD, RD), each at k-1, k, k+1, as the zone index of
// layer 0 of each column
__device__ __forceinline__ void columns9(const P1& P, int i, int j, int64_t* col) {
  col[0] = P.zi(i, j, 0);
  const int dirs[8] = {0, 1, 2, 3, 8, 9, 10, 11};
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    int ai, bj;
    P.nbr(i, j, dirs[q], ai, bj);
    col[q + 1] = P.zi(ai, bj, 0);
  }
}

// NM = 4: LIMIT_ALL_SLOPES_P1; NM = 10: LIMIT_ALL_SLOPES_P12 (step 7.c), the
// node values carrying the quadratic modes about the zone's own moments mom
// (440...f:594-600) and every mode above the mean reduced.  Two waves per
// SIMD: at NM = 10 the registers are held to 256 with a small spill, which
// the second wave more than pays for (the NM = 4 form fits without one)
template <int NM>
__global__ void __attribute__((amdgpu_waves_per_eu(2))) __launch_bounds__(256)
    k_pos(const P1 P, const Box b, const int64_t* list, int64_t nlist, const Flags F, const double* __restrict__ mom) {
  const int64_t t = tid();
  int i, j, k;
  int64_t z;
  if (list) {
    if (t >= nlist) return;
    z = list[t];
    zone_of(P, z, i, j, k);
  } else {
    if (!b.decode(t, i, j, k)) return;
    z = P.zi(i, j, k);
  }
  const cs_p1_patch& a = P.a;
  int64_t col[9];
  columns9(P, i, j, col);
  const int64_t ks = (int64_t)P.nz * P.nz;
  // ranges of rho, p and the entropy over the 27 zones, widened by 20 % where
  // the flattener is below 0.1 (440...f:169-280)
  double rmin = 1e300, rmax = -1e300, pmin = 1e300, pmax = -1e300, emin = 1e300, emax = -1e300;
#pragma unroll 1
  for (int q = 0; q < 9; ++q)
#pragma unroll
    for (int dk = -1; dk <= 1; ++dk) {
      const int64_t zz = col[q] + ks * (k + dk);
      const double r = P.V(0, 0, zz), p = P.V(4, 0, zz);
      double e = a.ent[zz];
      if (list && zz < z && F.prev[zz]) e = 0.0;
      rmin = fmin(rmin, r);
      rmax = fmax(rmax, r);
      pmin = fmin(pmin, p);
      pmax = fmax(pmax, p);
      emin = fmin(emin, e);
      emax = fmax(emax, e);
    }
  const double expandfac = 0.2;
  const double ff = fmax(fmin((a.flatten[z] - 0.1) / 0.4, 1.0), 0.0), omff = 1.0 - ff;
  rmin = fmax(rmin * (1.0 - expandfac * omff), a.rhofloor);
  rmax = fmax(rmax * (1.0 + expandfac * omff), a.rhofloor);
  pmin = fmax(pmin * (1.0 - expandfac * omff), a.prsfloor);
  pmax = fmax(pmax * (1.0 + expandfac * omff), a.prsfloor);
  emin = emin * (1.0 - expandfac * omff);
  emax = emax * (1.0 + expandfac * omff);
  // the modes checked at the 27 nodes: rho, p and the entropy
  const double r0 = P.V(0, 0, z), p0 = P.V(4, 0, z), e0 = a.ent[z];
  double rm[NM - 1], pm[NM - 1], em[3], q0[6];
#pragma unroll
  for (int m = 0; m < NM - 1; ++m) {
    rm[m] = P.V(0, m + 1, z);
    pm[m] = P.V(4, m + 1, z);
  }
#pragma unroll
  for (int m = 0; m < 3; ++m) em[m] = a.ent[z + P.NZ * (m + 1)];
#pragma unroll
  for (int m = 0; m < 6; ++m) q0[m] = NM == 10 ? mom[z + P.NZ * (4 + m)] : 0.0;
  const double obl = P.obl(z);
  double c0[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) c0[c] = P.cen(z, c);
  // the zone's 8 vertices and 6 face centroids, loaded once; the 27 nodes
  // (vertices, face centroids, edge midpoints, centroid) are formed from them
  double vx[8][3], fcx[6][3];
#pragma unroll
  for (int n = 0; n < 8; ++n) {
    const int vi = (n & 1) ? i : i - 1, vj = (n & 2) ? j : j - 1, vk = (n & 4) ? k : k - 1;
#pragma unroll
    for (int c = 0; c < 3; ++c) vx[n][c] = a.vert[P.fci(vi, vj, vk) + P.NFC * c];
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    fcx[0][c] = a.xf_centroid[P.xfi(i, j, k) + P.NXF * c];
    fcx[1][c] = a.xf_centroid[P.xfi(i - 1, j, k) + P.NXF * c];
    fcx[2][c] = a.yf_centroid[P.yfi(i, j, k) + P.NXF * c];
    fcx[3][c] = a.yf_centroid[P.yfi(i, j - 1, k) + P.NXF * c];
    fcx[4][c] = a.zf_centroid[P.zfi(i, j, k) + P.NZF * c];
    fcx[5][c] = a.zf_centroid[P.zfi(i, j, k - 1) + P.NZF * c];
  }
  const double ratio = 0.85769589859089412;  // 0.1^(1/15)
  // does every node pass with the slopes reduced n times?  The reduced slopes
  // are formed by n multiplications, as the reference's loop forms them
  auto passes = [&](int n) -> bool {
    double r[NM - 1], p[NM - 1], e[3];
#pragma unroll
    for (int m = 0; m < NM - 1; ++m) {
      r[m] = rm[m];
      p[m] = pm[m];
    }
#pragma unroll
    for (int m = 0; m < 3; ++m) e[m] = em[m];
    for (int q = 0; q < n; ++q) {
#pragma unroll
      for (int m = 0; m < NM - 1; ++m) {
        r[m] = ratio * r[m];
        p[m] = ratio * p[m];
      }
#pragma unroll
      for (int m = 0; m < 3; ++m) e[m] = ratio * e[m];
    }
    bool pass = true;
    // the scale opaque to the optimiser at third order: the nodes' offsets and
    // quadratic basis values are formed in each check instead of being
    // hoisted out of the bisection and held for all 27 nodes at once
    double ob = obl;
    if constexpr (NM == 10) asm volatile("" : "+v"(ob));
    // one node: scaled offset from the centroid, then the three checks
    auto node = [&](const double* x0) {
      double x[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) x[c] = (x0[c] - c0[c]) * ob;
      double rn = r0 + r[0] * x[0] + r[1] * x[1] + r[2] * x[2];
      double pn = p0 + p[0] * x[0] + p[1] * x[1] + p[2] * x[2];
      if constexpr (NM == 10) {
        const double b[6] = {x[0] * x[0] - q0[0], x[1] * x[1] - q0[1], x[2] * x[2] - q0[2],
                             x[0] * x[1] - q0[3], x[1] * x[2] - q0[4], x[0] * x[2] - q0[5]};
#pragma unroll
        for (int m = 0; m < 6; ++m) {
          rn = rn + r[3 + m] * b[m];
          pn = pn + p[3 + m] * b[m];
        }
      }
      const double en = e0 + e[0] * x[0] + e[1] * x[1] + e[2] * x[2];
      if (rn < rmin || rn > rmax || pn < pmin || pn > pmax || en < emin || en > emax) pass = false;
    };
#pragma unroll
    for (int v = 0; v < 8; ++v) node(vx[v]);
#pragma unroll
    for (int f = 0; f < 6; ++f) node(fcx[f]);
    // edge midpoints: vertex pairs along x, then y, then z
#pragma unroll
    for (int q = 0; q < 12; ++q) {
      const int va = q < 4 ? 2 * q : (q < 8 ? ((q - 4) & 1) + 4 * ((q - 4) >> 1) : q - 8);
      const int vb = q < 4 ? va + 1 : (q < 8 ? va + 2 : va + 4);
      double m[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) m[c] = (vx[va][c] + vx[vb][c]) * 0.5;
      node(m);
    }
    node(c0);
    return pass;
  };
  // the reference (440...f:553-600) checks 0, 1, ..., 14 reductions in turn and
  // takes the first that passes, removing the slopes when none does.  With
  // the zone's mean inside the bounds, passing is monotone in the number of
  // reductions (in exact arithmetic), so the first passing count is found by
  // bisection: 2 to 6 checks instead of up to 15.  The count found passes and
  // the one below it fails (both are checked); only a node value that rounds
  // across a bound at a smaller count, although the counts' node values
  // differ by 14 % of the slope part, could make the reference's scan stop
  // earlier
  // one call site (the check is large): probe 0, then 14, then bisect; lo is
  // the largest count known to fail, hi the smallest known to pass (15: none)
  int lo = -1, hi = 15, n = 0;
#pragma unroll 1
  while (hi - lo > 1) {
    const bool ok = passes(n);
    if (ok) hi = n;
    else lo = n;
    n = lo < 0 ? 0 : (lo == 0 && hi == 15 ? 14 : (lo + hi) >> 1);
  }
  const bool removed = hi == 15;
  const int nred = removed ? 0 : hi;
  if (!removed)
    for (int q = 0; q < nred; ++q)
#pragma unroll
      for (int m = 0; m < 3; ++m) em[m] = ratio * em[m];
  // the same reductions on every component (repeated, as the reference
  // multiplies once per iteration); B_zc copies the B modes
  if (nred > 0 || removed) {
    for (int c = 0; c < 8; ++c)
#pragma unroll
      for (int m = 0; m < NM - 1; ++m) {
        double v = P.V(c, m + 1, z);
        if (removed) {
          v = 0.0;
        } else {
          for (int r = 0; r < nred; ++r) v = ratio * v;
        }
        P.Vr(c, m + 1, z) = v;
      }
#pragma unroll
    for (int m = 0; m < 3; ++m) a.ent[z + P.NZ * (m + 1)] = removed ? 0.0 : em[m];
  }
#pragma unroll
  for (int m = 0; m < NM; ++m) {
    a.B1_zc[z + P.NZ * m] = P.V(5, m, z);
    a.B2_zc[z + P.NZ * m] = P.V(6, m, z);
    a.B3_zc[z + P.NZ * m] = P.V(7, m, z);
  }
  const uint8_t flag = removed ? 1 : 0;
  if (!list) {
    F.fa[z] = flag;
    if (removed) atomicAdd(F.counter, 1u);
  } else {
    F.next[z] = flag;
    if (flag != F.prev[z]) atomicAdd(F.counter, 1u);
  }
}

// the zones of the range with a zone ever flagged earlier in the loop order
// among their 27 and not yet on the list: appended to it (order immaterial;
// the list only grows, so a zone whose result depends on a flag is
// recomputed in every later sweep)
__global__ void __launch_bounds__(256) k_cands(const P1 P, const Box b, const uint8_t* ever, uint8_t* inlist,
                                               int64_t* list, unsigned int* count) {
  int i, j, k;
  if (!b.decode(tid(), i, j, k)) return;
  const int64_t z = P.zi(i, j, k), ks = (int64_t)P.nz * P.nz;
  if (inlist[z]) return;
  int64_t col[9];
  columns9(P, i, j, col);
  bool cand = false;
  for (int q = 0; q < 9; ++q)
#pragma unroll
    for (int dk = -1; dk <= 1; ++dk) {
      const int64_t zz = col[q] + ks * (k + dk);
      cand = cand || (zz < z && ever[zz]);
    }
  if (cand) {
    inlist[z] = 1;
    list[atomicAdd(count, 1u)] = z;
  }
}

// after a sweep: the list's new flags become the previous ones and join the
// ever-flagged set
__global__ void __launch_bounds__(256) k_commit(const int64_t* list, int64_t n, const uint8_t* next, uint8_t* prev,
                                                uint8_t* ever) {
  const int64_t t = tid();
  if (t >= n) return;
  const int64_t z = list[t];
  prev[z] = next[z];
  ever[z] = ever[z] | next[z];
}

// the entropy means of the zones that lost all their slopes
__global__ void __launch_bounds__(256) k_clear_ent(const P1 P, const Box b, const uint8_t* flags) {
  int i, j, k;
  if (!b.decode(tid(), i, j, k)) return;
  const int64_t z = P.zi(i, j, k);
  if (flags[z]) P.a.ent[z] = 0.0;
}

// ---------------------------------------------------------------------------
// step 8.a-c: blockIdx.y = family
// NM = 10: the third-order form (421...f:4817-4879), each side's field with
// its quadratic modes about that zone's own moments mom
template <int NM>
__global__ void __launch_bounds__(256) k_face_pts(const P1 P, const Box bx, const Box by, const Box bz,
                                                  const double* __restrict__ mom) {
  const int fam = blockIdx.y;
  const Box& b = fam == 0 ? bx : (fam == 1 ? by : bz);
  int i, j, k;
  if (!b.decode(tid(), i, j, k)) return;
  const cs_p1_patch& a = P.a;
  double* Bq = fam == 0 ? a.Bxf_q : (fam == 1 ? a.Byf_q : a.Bzf_q);
  const double* Bf = fam == 0 ? a.Bxf : (fam == 1 ? a.Byf : a.Bzf);
  const double* pts = fam == 0 ? a.xb_pts : (fam == 1 ? a.yb_pts : a.zb_pts);
  const double* wts = fam == 0 ? a.xb_wts : (fam == 1 ? a.yb_wts : a.zb_wts);
  const double* nrm = fam == 0 ? a.xb_normal : (fam == 1 ? a.yb_normal : a.zb_normal);
  const int64_t NF = fam == 2 ? P.NZF : P.NXF;
  const int64_t fb = P.fci(i, j, k), fg = fam == 0 ? P.xfi(i, j, k) : (fam == 1 ? P.yfi(i, j, k) : P.zfi(i, j, k));
  int li = i, lj = j, ri = i + (fam == 0), rj = j + (fam == 1);
  bool skip = false;
  if (fam < 2) {
    const bool al = P.active(li, lj, k), ar = P.active(ri, rj, k);
    if (!al && ar) {
      P.nbr(ri, rj, fam == 0 ? 1 : 3, li, lj);
    } else if (al && !ar) {
      P.nbr(li, lj, fam == 0 ? 0 : 2, ri, rj);
    } else if (!al && !ar) {
      skip = true;
    }
  } else {
    skip = !P.active(i, j, k);
  }
  double q5[5];
  if (skip) {
#pragma unroll
    for (int q = 0; q < 5; ++q) q5[q] = 0.0;
  } else {
    const int64_t zl = P.zi(li, lj, k), zr = P.zi(ri, rj, fam == 2 ? k + 1 : k);
    const double oll = P.obl(zl), olr = P.obl(zr);
    double vl[3][NM], vr[3][NM], ql[6], qr[6];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int m = 0; m < NM; ++m) {
        vl[c][m] = P.V(5 + c, m, zl);
        vr[c][m] = P.V(5 + c, m, zr);
      }
#pragma unroll
    for (int m = 0; m < 6; ++m) {
      ql[m] = NM == 10 ? mom[zl + P.NZ * (4 + m)] : 0.0;
      qr[m] = NM == 10 ? mom[zr + P.NZ * (4 + m)] : 0.0;
    }
    double n[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) n[c] = nrm[fg + NF * c];
#pragma unroll
    for (int q = 0; q < 5; ++q) {
      if (fam == 2) {
#pragma unroll
        for (int c = 0; c < 3; ++c) n[c] = nrm[fg + NF * (c + 3 * q)];
      }
      double rl[3], rr[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const double x = pts[fg + NF * (c + 3 * q)];
        rl[c] = (x - P.cen(zl, c)) * oll;
        rr[c] = (x - P.cen(zr, c)) * olr;
      }
      double s[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        double bl = vl[c][0] + vl[c][1] * rl[0] + vl[c][2] * rl[1] + vl[c][3] * rl[2];
        double br = vr[c][0] + vr[c][1] * rr[0] + vr[c][2] * rr[1] + vr[c][3] * rr[2];
        if constexpr (NM == 10) {
          const double pl[6] = {rl[0] * rl[0] - ql[0], rl[1] * rl[1] - ql[1], rl[2] * rl[2] - ql[2],
                                rl[0] * rl[1] - ql[3], rl[1] * rl[2] - ql[4], rl[0] * rl[2] - ql[5]};
          const double pr[6] = {rr[0] * rr[0] - qr[0], rr[1] * rr[1] - qr[1], rr[2] * rr[2] - qr[2],
                                rr[0] * rr[1] - qr[3], rr[1] * rr[2] - qr[4], rr[0] * rr[2] - qr[5]};
#pragma unroll
          for (int m = 0; m < 6; ++m) {
            bl = bl + vl[c][4 + m] * pl[m];
            br = br + vr[c][4 + m] * pr[m];
          }
        }
        s[c] = bl + br;
      }
      q5[q] = 0.5 * (n[0] * s[0] + n[1] * s[1] + n[2] * s[2]);
    }
  }
  double bd = Bf[fb];
#pragma unroll
  for (int q = 0; q < 5; ++q) bd = bd - wts[fg + NF * q] * q5[q];
#pragma unroll
  for (int q = 0; q < 5; ++q) Bq[fb + P.NFC * q] = q5[q] + bd;
}

// ---------------------------------------------------------------------------

// face f of zone (i, j, k): 0 right x, 1 left x, 2 up y, 3 down y, 4 top z, 5 bottom z
struct FaceRef {
  const double *nrm, *cen, *pts, *Bf, *Bq;
  int64_t fg, fb, NF;
};
__device__ __forceinline__ FaceRef face_of(const P1& P, int f, int i, int j, int k) {
  const cs_p1_patch& a = P.a;
  const int fam = f >> 1;
  const bool hi = (f & 1) == 0;
  const int fi = i - (fam == 0 && !hi), fj = j - (fam == 1 && !hi), fk = k - (fam == 2 && !hi);
  FaceRef r;
  r.fg = fam == 0 ? P.xfi(fi, fj, fk) : (fam == 1 ? P.yfi(fi, fj, fk) : P.zfi(fi, fj, fk));
  r.fb = P.fci(fi, fj, fk);
  r.NF = fam == 2 ? P.NZF : P.NXF;
  r.nrm = fam == 0 ? a.xb_normal : (fam == 1 ? a.yb_normal : a.zb_normal);
  r.cen = fam == 0 ? a.xf_centroid : (fam == 1 ? a.yf_centroid : a.zf_centroid);
  r.pts = fam == 0 ? a.xb_pts : (fam == 1 ? a.yb_pts : a.zb_pts);
  r.Bf = fam == 0 ? a.Bxf : (fam == 1 ? a.Byf : a.Bzf);
  r.Bq = fam == 0 ? a.Bxf_q : (fam == 1 ? a.Byf_q : a.Bzf_q);
  return r;
}

// constraint row f (face normal, at the face centroid) and its right-hand side
__device__ __forceinline__ double crow(const P1& P, int f, int i, int j, int k, int64_t z, double obl, double* row) {
  const FaceRef F = face_of(P, f, i, j, k);
  double n[3], r[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    n[c] = F.nrm[F.fg + F.NF * c];
    r[c] = (F.cen[F.fg + F.NF * c] - P.cen(z, c)) * obl;
  }
  brow(n, r, row);
  return F.Bf[F.fb];
}

__global__ void __launch_bounds__(64) k_clsq_e(const P1 P, const Box b) {
  __shared__ double Zs[72][64];  // each thread's own column: no barrier needed
  int i, j, k;
  if (!b.decode(tid(), i, j, k)) return;
  const cs_p1_patch& a = P.a;
  const int64_t z = P.zi(i, j, k);
  if (in_weno_range(P, i, j, k) && a.mask[z] != 0) {
    const double obl = P.obl(z);
    double L[78], u[12], row[12];
#pragma unroll
    for (int r = 0; r < 78; ++r) L[r] = 0.0;
#pragma unroll
    for (int r = 0; r < 12; ++r) {
      L[LIX(r, r)] = 1.0;
      u[r] = P.V(5 + r / 4, r % 4, z);
    }
    // least-squares rows: points 2 and 4 of each face
#pragma unroll 1
    for (int f = 0; f < 6; ++f) {
      const FaceRef F = face_of(P, f, i, j, k);
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        const int pt = 1 + 2 * t;
        double n[3], r[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          n[c] = (f >> 1) == 2 ? F.nrm[F.fg + F.NF * (c + 3 * pt)] : F.nrm[F.fg + F.NF * c];
          r[c] = (F.pts[F.fg + F.NF * (c + 3 * pt)] - P.cen(z, c)) * obl;
        }
        brow(n, r, row);
        const double bv = F.Bq[F.fb + P.NFC * pt];
#pragma unroll
        for (int p = 0; p < 12; ++p) {
          u[p] += row[p] * bv;
#pragma unroll
          for (int q = 0; q <= p; ++q) L[LIX(p, q)] += row[p] * row[q];
        }
      }
    }
    chol12(L);
    // with Z = L^-1 C^T (kept in LDS) and zb = L^-1 A^T b: S = Z^T Z,
    // s = Z^T zb - d, then x = L^-T (zb - Z lambda)
    fwd12(L, u);
    double S[21], s[6];
#pragma unroll 1
    for (int ci = 0; ci < 6; ++ci) {
      const double dci = crow(P, ci, i, j, k, z, obl, row);
      fwd12(L, row);
      double t = 0.0;
#pragma unroll
      for (int p = 0; p < 12; ++p) {
        t += row[p] * u[p];
        Zs[ci * 12 + p][threadIdx.x] = row[p];
      }
      s[ci] = t - dci;
    }
#pragma unroll
    for (int r = 0; r < 6; ++r)
#pragma unroll
      for (int c = 0; c <= r; ++c) {
        double t = 0.0;
#pragma unroll
        for (int p = 0; p < 12; ++p) t += Zs[r * 12 + p][threadIdx.x] * Zs[c * 12 + p][threadIdx.x];
        S[r * (r + 1) / 2 + c] = t;
      }
    // S lambda = s (6 x 6 Cholesky)
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      double d = S[c * (c + 1) / 2 + c];
#pragma unroll
      for (int q = 0; q < c; ++q) d -= S[c * (c + 1) / 2 + q] * S[c * (c + 1) / 2 + q];
      d = sqrt(d);
      S[c * (c + 1) / 2 + c] = d;
#pragma unroll
      for (int r = c + 1; r < 6; ++r) {
        double t = S[r * (r + 1) / 2 + c];
#pragma unroll
        for (int q = 0; q < c; ++q) t -= S[r * (r + 1) / 2 + q] * S[c * (c + 1) / 2 + q];
        S[r * (r + 1) / 2 + c] = t / d;
      }
    }
#pragma unroll
    for (int r = 0; r < 6; ++r) {
      double t = s[r];
#pragma unroll
      for (int q = 0; q < r; ++q) t -= S[r * (r + 1) / 2 + q] * s[q];
      s[r] = t / S[r * (r + 1) / 2 + r];
    }
#pragma unroll
    for (int r = 5; r >= 0; --r) {
      double t = s[r];
#pragma unroll
      for (int q = r + 1; q < 6; ++q) t -= S[q * (q + 1) / 2 + r] * s[q];
      s[r] = t / S[r * (r + 1) / 2 + r];
    }
    // x = L^-T (zb - Z lambda)
#pragma unroll
    for (int p = 0; p < 12; ++p) {
      double t = u[p];
#pragma unroll
      for (int c = 0; c < 6; ++c) t -= s[c] * Zs[c * 12 + p][threadIdx.x];
      u[p] = t;
    }
    bwd12(L, u);
    const double aa = (P.Ur(5, 0, z) * P.Ur(5, 0, z) + P.Ur(6, 0, z) * P.Ur(6, 0, z) + P.Ur(7, 0, z) * P.Ur(7, 0, z)) *
                      C<double>::ob8pi;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      a.B1_zc[z + P.NZ * m] = u[m];
      a.B2_zc[z + P.NZ * m] = u[4 + m];
      a.B3_zc[z + P.NZ * m] = u[8 + m];
#pragma unroll
      for (int c = 0; c < 3; ++c) P.Vr(5 + c, m, z) = u[4 * c + m];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) P.Ur(5 + c, 0, z) = u[4 * c];
    const double bb = (u[0] * u[0] + u[4] * u[4] + u[8] * u[8]) * C<double>::ob8pi;
    P.Ur(4, 0, z) = P.Ur(4, 0, z) - aa + bb;
  }
  // step 9.a: E = -v x B from the zone means
  const double vx = P.V(1, 0, z), vy = P.V(2, 0, z), vz = P.V(3, 0, z);
  const double Bx = P.V(5, 0, z), By = P.V(6, 0, z), Bz = P.V(7, 0, z);
  a.Ex_zc[z] = -(vy * Bz - vz * By);
  a.Ey_zc[z] = -(vz * Bx - vx * Bz);
  a.Ez_zc[z] = -(vx * By - vy * Bx);
}

// ---------------------------------------------------------------------------
// step 10: per-block minima and maxima of the 8 primitives and the
// normalised divergence over the owned active zones (18 values per block)
constexpr int DIAG_THREADS = 256;
__global__ void __launch_bounds__(DIAG_THREADS) k_diag(const P1 P, const Box b, double* part) {
  __shared__ double sh[18][DIAG_THREADS];
  double mn[9], mx[9];
#pragma unroll
  for (int q = 0; q < 9; ++q) {
    mn[q] = 1.0e30;
    mx[q] = -1.0e30;
  }
  int i, j, k;
  if (b.decode(tid(), i, j, k)) {
    const cs_p1_patch& a = P.a;
    const int64_t z = P.zi(i, j, k);
    a.work[z] = 0.0;
    const bool own = i >= 1 && i <= P.N && j >= 1 && j <= P.N && k >= 1 && k <= P.NR && a.mask[z] == 1;
    if (own) {
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        mn[c] = fmin(mn[c], P.V(c, 0, z));
        mx[c] = fmax(mx[c], P.V(c, 0, z));
      }
      const double bxp = a.Bxf[P.fci(i, j, k)], bxm = a.Bxf[P.fci(i - 1, j, k)];
      const double byp = a.Byf[P.fci(i, j, k)], bym = a.Byf[P.fci(i, j - 1, k)];
      const double bzp = a.Bzf[P.fci(i, j, k)], bzm = a.Bzf[P.fci(i, j, k - 1)];
      double a21 = a.xf_area[P.xfi(i, j, k)] * bxp + a.yf_area[P.yfi(i, j, k)] * byp +
                   a.zf_area[P.zfi(i, j, k)] * bzp - a.xf_area[P.xfi(i - 1, j, k)] * bxm -
                   a.yf_area[P.yfi(i, j - 1, k)] * bym - a.zf_area[P.zfi(i, j, k - 1)] * bzm;
      a.work[z] = a21;
      const double v6 = P.V(5, 0, z), v7 = P.V(6, 0, z), v8 = P.V(7, 0, z), rho = P.V(0, 0, z);
      const double a11 = v6 * v6 + v7 * v7 + v8 * v8;
      const double ms = sqrt((a.gamma_sound[z] * P.V(4, 0, z) + C<double>::ob4pi * a11) / rho);
      a.msonic[z] = ms;
      const double den = fmax(fmax(sqrt(bxp * bxp + byp * byp + bzp * bzp), sqrt(bxm * bxm + bym * bym + bzm * bzm)),
                              ms * sqrt(4.0 * 3.14 * rho));
      a21 = a21 / den;
      mn[8] = a21;
      mx[8] = a21;
    }
  }
#pragma unroll
  for (int q = 0; q < 9; ++q) {
    sh[q][threadIdx.x] = mn[q];
    sh[9 + q][threadIdx.x] = mx[q];
  }
  __syncthreads();
  for (int w = DIAG_THREADS / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
#pragma unroll
      for (int q = 0; q < 9; ++q) {
        sh[q][threadIdx.x] = fmin(sh[q][threadIdx.x], sh[q][threadIdx.x + w]);
        sh[9 + q][threadIdx.x] = fmax(sh[9 + q][threadIdx.x], sh[9 + q][threadIdx.x + w]);
      }
    __syncthreads();
  }
  if (threadIdx.x < 18) part[(int64_t)blockIdx.x * 18 + threadIdx.x] = sh[threadIdx.x][0];
}

// one block per quantity (9 minima, then 9 maxima)
__global__ void __launch_bounds__(256) k_diag_final(const double* part, int nblocks, double* out) {
  __shared__ double sh[256];
  const int q = blockIdx.x;
  double v = q < 9 ? 1.0e30 : -1.0e30;
  for (int b = threadIdx.x; b < nblocks; b += 256)
    v = q < 9 ? fmin(v, part[(int64_t)b * 18 + q]) : fmax(v, part[(int64_t)b * 18 + q]);
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      sh[threadIdx.x] = q < 9 ? fmin(sh[threadIdx.x], sh[threadIdx.x + w]) : fmax(sh[threadIdx.x], sh[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[q] = sh[0];
}

// ---------------------------------------------------------------------------
// per-device work memory: flags (3 x zones), counter, block partials
struct Scratch {
  void* mem = nullptr;
  size_t bytes = 0;
};
static std::mutex g_mu;
static std::vector<Scratch> g_scratch;

}  // namespace csp1

using namespace csp1;

static int p1_check(const cs_p1_patch* p) {
  if (!p) {
    cs::set_error("cs_limit_all_slopes_p1: null patch");
    return CS_E_INVALID;
  }
  if (p->ntp < 1 || p->nr < 1 || p->nghost < 4) {
    cs::set_error("cs_limit_all_slopes_p1: need ntp >= 1, nr >= 1, nghost >= 4 (got %d, %d, %d)", p->ntp, p->nr,
                  p->nghost);
    return CS_E_INVALID;
  }
  if (p->nghost >= p->ntp) {
    cs::set_error("cs_limit_all_slopes_p1: nghost (%d) must be below ntp (%d)", p->nghost, p->ntp);
    return CS_E_BAD_DECOMP;
  }
  if (!(p->gamma > 1.0) || !(p->smallnum > 0.0) || !(p->rhofloor > 0.0) || !(p->prsfloor > 0.0)) {
    cs::set_error("cs_limit_all_slopes_p1: gamma > 1 and positive floors and smallnum required");
    return CS_E_INVALID;
  }
  const void* ptrs[] = {p->vert, p->mask, p->stencil, p->xf_normal, p->xf_area, p->xf_centroid, p->yf_normal,
                        p->yf_area, p->yf_centroid, p->zf_area, p->zf_centroid, p->xb_pts, p->xb_wts,
                        p->xb_normal, p->yb_pts, p->yb_wts, p->yb_normal, p->zb_pts, p->zb_wts, p->zb_normal,
                        p->scale, p->v_centroid, p->U, p->ent, p->Bxf, p->Byf, p->Bzf, p->V, p->B1_zc, p->B2_zc,
                        p->B3_zc, p->Ex_zc, p->Ey_zc, p->Ez_zc, p->gamma_eos, p->gamma_sound, p->mean_mol_wt,
                        p->divvel, p->msonic, p->flatten, p->work, p->Bxf_q, p->Byf_q, p->Bzf_q};
  for (const void* q : ptrs)
    if (!q) {
      cs::set_error("cs_limit_all_slopes_p1: null array in the patch");
      return CS_E_INVALID;
    }
  return 0;
}

#define P1_HIP(call)                                                                  \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess) {                                                           \
      cs::set_error("cs_limit_all_slopes_p1: %s: %s", #call, hipGetErrorString(e_)); \
      return CS_E_HIP;                                                                \
    }                                                                                 \
  } while (0)

static inline unsigned grid_of(int64_t n, int threads) { return (unsigned)((n + threads - 1) / threads); }

// kernel_ms (optional): per-stage times from HIP events on the stream
// ext (optional): caller-provided device work memory of ext_bytes, used
// instead of the library's per-device allocation (no allocation then)
namespace csp1 {
__global__ void __launch_bounds__(256) k_zero_hi(const P1 P, const Box b);
__global__ void __launch_bounds__(64) k_clsq12(const P1 P, const Box b, const double* __restrict__ mom);
__global__ void __launch_bounds__(256) k_e12(const P1 P, const Box b, const double* __restrict__ mom);
__global__ void __launch_bounds__(256) k_ao_order(const P1 P, const Box b, unsigned int* bad);
constexpr int kAoRec = 200;  // doubles per zone record of k_ao_e_pre: G (19 x 10), 18 int32 indices
__global__ void __launch_bounds__(128) k_ao_e_pre(const P1 P, const Box b, const double* __restrict__ mom,
                                                  double* __restrict__ rec);
__global__ void __launch_bounds__(64) k_weno_ao_e(const P1 P, const Box b, const double* __restrict__ mom,
                                                   const double* __restrict__ rec, int level, int klo);
__global__ void __attribute__((amdgpu_waves_per_eu(2))) __launch_bounds__(128)
    k_weno_ao(const P1 P, const Box b, const double* __restrict__ mom, const int64_t* list, int64_t nlist);
}  // namespace csp1

// mom (optional, DEVICE zone x 10): the third-order form, LIMIT_ALL_SLOPES_P12:
// U, V, B_zc and E_zc of ten modes, the WENO gradients left unflattened and
// followed by the WENO-AO hybridisation of step 7.b, the positivity limiter
// and the face-point B with the quadratic modes, the 30-unknown constrained
// least squares, the zone E at the centroid with its own WENO-AO (step 9)
static int run_p1(const cs_p1_patch* p, void* stream_, float* kernel_ms, uint8_t* ext = nullptr,
                  size_t ext_bytes = 0, const double* mom = nullptr) {
  int rc = p1_check(p);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream_;
  P1 P;
  P.a = *p;
  P.N = p->ntp;
  P.NR = p->nr;
  P.g = p->nghost;
  P.nz = P.N + 2 * P.g;
  P.nzr = P.NR + 2 * P.g;
  P.nf = P.nz + 1;
  P.nfr = P.nzr + 1;
  P.NZ = (int64_t)P.nz * P.nz * P.nzr;
  P.NXF = (int64_t)P.nf * P.nz * P.nzr;
  P.NZF = (int64_t)P.nz * P.nz * P.nfr;
  P.NFC = (int64_t)P.nf * P.nf * P.nfr;
  const int N = P.N, NR = P.NR, g = P.g;
  const Box all = make_box(1 - g, N + g, 1 - g, N + g, 1 - g, NR + g);
  const Box wr = make_box(-2, N + 3, -2, N + 3, -2, NR + 3);
  const Box own = make_box(1, N, 1, N, 1, NR);
  const Box fx = make_box(-3, N + 3, -2, N + 3, -2, NR + 3), fy = make_box(-2, N + 3, -3, N + 3, -2, NR + 3),
            fz = make_box(-2, N + 3, -2, N + 3, -3, NR + 3);
  const int nblk_diag = (int)grid_of(all.count(), DIAG_THREADS);

  int dev = 0;
  P1_HIP(hipGetDevice(&dev));
  const size_t flag_bytes = ((size_t)P.NZ + 255) / 256 * 256;
  const size_t list_bytes = (size_t)wr.count() * sizeof(int64_t);
  const Box ao = make_box(-1, N + 2, -1, N + 2, -1, NR + 2);
  const size_t base_bytes = (5 * flag_bytes + 256 + (size_t)nblk_diag * 18 * sizeof(double) + 18 * sizeof(double) +
                             list_bytes + 255) / 256 * 256;
  const size_t need = base_bytes + (mom ? (size_t)ao.count() * kAoRec * sizeof(double) : 0);
  if (mom && P.NZ >= ((int64_t)1 << 31)) {
    cs::set_error("cs_limit_all_slopes_p12: the patch has %lld zones, more than int32 indices reach",
                  (long long)P.NZ);
    return CS_E_INVALID;
  }
  std::unique_lock<std::mutex> lock(g_mu, std::defer_lock);
  uint8_t* base = nullptr;
  if (ext) {
    if (need > ext_bytes) {
      cs::set_error("cs_limit_all_slopes_p1_noalloc: the patch (%d^2 x %d, nghost %d) needs %zu bytes of work "
                    "memory, the entry point holds %zu", N, NR, g, need, ext_bytes);
      return CS_E_INVALID;
    }
    base = ext;
  } else {
    lock.lock();
    if ((int)g_scratch.size() <= dev) g_scratch.resize(dev + 1);
    Scratch& S = g_scratch[dev];
    if (S.bytes < need) {
      if (S.mem) P1_HIP(hipFree(S.mem));
      S.mem = nullptr;
      S.bytes = 0;
      if (hipMalloc(&S.mem, need) != hipSuccess) {
        cs::set_error("cs_limit_all_slopes_p1: cannot allocate %zu bytes of work memory", need);
        return CS_E_NOMEM;
      }
      S.bytes = need;
    }
    base = (uint8_t*)S.mem;
  }
  Flags F;
  F.fa = base;
  F.prev = base + flag_bytes;
  F.next = base + 2 * flag_bytes;
  uint8_t* ever = base + 3 * flag_bytes;    // flagged in the first pass or any sweep
  uint8_t* inlist = base + 4 * flag_bytes;  // on the replay list
  F.counter = (unsigned int*)(base + 5 * flag_bytes);
  double* part = (double*)(base + 5 * flag_bytes + 256);
  double* red = part + (size_t)nblk_diag * 18;
  int64_t* list = (int64_t*)(red + 18);
  unsigned int* lcount = F.counter + 1;
  double* aorec = (double*)(base + base_bytes);  // third order: the records of k_ao_e_pre

  constexpr int NEV = CS_P1_STAGES + 1;
  hipEvent_t ev[NEV];
  int nev = 0;
  if (kernel_ms) {
    for (int q = 0; q < NEV; ++q) P1_HIP(hipEventCreate(&ev[q]));
  }
  auto mark = [&]() -> int {
    if (kernel_ms) P1_HIP(hipEventRecord(ev[nev++], st));
    return 0;
  };
  if ((rc = mark())) return rc;
  if (mom) k_zero_hi<<<grid_of(all.count(), 256), 256, 0, st>>>(P, all);
  k_prim<<<grid_of(all.count(), 256), 256, 0, st>>>(P, all);
  if ((rc = mark())) return rc;
  k_face_zero<<<dim3(grid_of(P.NFC, 256), 3), 256, 0, st>>>(P);
  k_tricorner<<<grid_of(4 * (int64_t)g * P.nzr, 256), 256, 0, st>>>(P);
  if ((rc = mark())) return rc;
  k_flat_a<<<grid_of(all.count(), 256), 256, 0, st>>>(P, all);
  k_flat_b<<<grid_of(all.count(), 256), 256, 0, st>>>(P, all);
  k_flat_c<<<grid_of(all.count(), 256), 256, 0, st>>>(P, all);
  if ((rc = mark())) return rc;
  // flags of zones outside the limiter's range stay clear
  P1_HIP(hipMemsetAsync(F.fa, 0, P.NZ, st));
  P1_HIP(hipMemsetAsync(F.counter, 0, 2 * sizeof(unsigned int), st));
  WenoFields wv, we;
  wv.n = 9;
  wv.noflat = mom ? 0xffu : 0u;  // third order: V flattened after the hybridisation
  we.noflat = mom ? 0x7u : 0u;  // third order: E flattened after its hybridisation
  for (int c = 0; c < 8; ++c) {
    wv.src[c] = p->V + P.NZ * c;
    wv.dst[c] = p->V + P.NZ * (c + 8);
    wv.stride[c] = 8 * P.NZ;
  }
  wv.src[8] = p->ent;
  wv.dst[8] = p->ent + P.NZ;
  wv.stride[8] = P.NZ;
  we.n = 3;
  double* E3[3] = {p->Ex_zc, p->Ey_zc, p->Ez_zc};
  for (int c = 0; c < 3; ++c) {
    we.src[c] = E3[c];
    we.dst[c] = E3[c] + P.NZ;
    we.stride[c] = P.NZ;
  }
  k_weno9<<<grid_of(3 * wr.count(), 128), 128, 0, st>>>(P, wr, nullptr, 0, wv);
  if (mom) {
    k_weno_ao<<<grid_of(ao.count(), 128), 128, 0, st>>>(P, ao, mom, nullptr, 0);
    k_pos<10><<<grid_of(wr.count(), 256), 256, 0, st>>>(P, wr, nullptr, 0, F, mom);
  } else {
    k_pos<4><<<grid_of(wr.count(), 256), 256, 0, st>>>(P, wr, nullptr, 0, F, nullptr);
  }
  P1_HIP(hipGetLastError());
  unsigned int nfail = 0;
  P1_HIP(hipMemcpyAsync(&nfail, F.counter, sizeof(unsigned int), hipMemcpyDeviceToHost, st));
  P1_HIP(hipStreamSynchronize(st));
  if (nfail > 0) {
    // replay the loop-order dependence on the zones it can reach until the
    // flags settle: every sweep recomputes the list with the previous sweep's
    // flags; a zone newly flagged in a sweep brings its later neighbours
    // onto the list before the next one
    P1_HIP(hipMemcpyAsync(F.prev, F.fa, P.NZ, hipMemcpyDeviceToDevice, st));
    P1_HIP(hipMemcpyAsync(F.next, F.fa, P.NZ, hipMemcpyDeviceToDevice, st));
    P1_HIP(hipMemcpyAsync(ever, F.fa, P.NZ, hipMemcpyDeviceToDevice, st));
    P1_HIP(hipMemsetAsync(inlist, 0, P.NZ, st));
    unsigned int ncand = 0;
    bool settled = false;
    for (int it = 0; it < 256 && !settled; ++it) {
      k_cands<<<grid_of(wr.count(), 256), 256, 0, st>>>(P, wr, ever, inlist, list, lcount);
      P1_HIP(hipMemsetAsync(F.counter, 0, sizeof(unsigned int), st));
      P1_HIP(hipMemcpyAsync(&ncand, lcount, sizeof(unsigned int), hipMemcpyDeviceToHost, st));
      P1_HIP(hipStreamSynchronize(st));
      if (ncand == 0) {
        settled = true;
        break;
      }
      k_weno9<<<grid_of(3 * (int64_t)ncand, 128), 128, 0, st>>>(P, wr, list, ncand, wv);
      if (mom) {
        k_weno_ao<<<grid_of(ncand, 128), 128, 0, st>>>(P, ao, mom, list, ncand);
        k_pos<10><<<grid_of(ncand, 256), 256, 0, st>>>(P, wr, list, ncand, F, mom);
      } else {
        k_pos<4><<<grid_of(ncand, 256), 256, 0, st>>>(P, wr, list, ncand, F, nullptr);
      }
      k_commit<<<grid_of(ncand, 256), 256, 0, st>>>(list, ncand, F.next, F.prev, ever);
      P1_HIP(hipGetLastError());
      unsigned int nchg = 0;
      P1_HIP(hipMemcpyAsync(&nchg, F.counter, sizeof(unsigned int), hipMemcpyDeviceToHost, st));
      P1_HIP(hipStreamSynchronize(st));
      settled = nchg == 0;
    }
    if (!settled) {
      cs::set_error("cs_limit_all_slopes_p1: positivity replay did not settle");
      return CS_E_INVALID;
    }
    k_clear_ent<<<grid_of(wr.count(), 256), 256, 0, st>>>(P, wr, F.prev);
  }
  if ((rc = mark())) return rc;
  if (mom)
    k_face_pts<10><<<dim3(grid_of(std::max(fx.count(), std::max(fy.count(), fz.count())), 256), 3), 256, 0, st>>>(
        P, fx, fy, fz, mom);
  else
    k_face_pts<4><<<dim3(grid_of(std::max(fx.count(), std::max(fy.count(), fz.count())), 256), 3), 256, 0, st>>>(
        P, fx, fy, fz, nullptr);
  if ((rc = mark())) return rc;
  if (mom) {
    k_clsq12<<<(unsigned)wr.count(), 64, 0, st>>>(P, wr, mom);
    k_e12<<<grid_of(all.count(), 256), 256, 0, st>>>(P, all, mom);
  } else {
    k_clsq_e<<<grid_of(all.count(), 64), 64, 0, st>>>(P, all);
  }
  if ((rc = mark())) return rc;
  k_weno9<<<grid_of(wr.count(), 128), 128, 0, st>>>(P, wr, nullptr, 0, we);
  if (mom) {
    // 9.c in the loop order of the reference's serial build: level
    // (i+1) + 3 (j+1) + 9 (k+1) of the zone box; k_ao_order flags a stencil
    // table for which a zone's earlier neighbours are not all on lower
    // levels and its later ones all on higher levels
    P1_HIP(hipMemsetAsync(F.counter + 2, 0, sizeof(unsigned int), st));
    k_ao_order<<<grid_of((int64_t)ao.ni * ao.nj, 256), 256, 0, st>>>(P, ao, F.counter + 2);
    k_ao_e_pre<<<grid_of(ao.count(), 128), 128, 0, st>>>(P, ao, mom, aorec);
    const int top = (ao.ni - 1) + 3 * (ao.nj - 1) + 9 * (ao.nk - 1);
    for (int lv = 0; lv <= top; ++lv) {
      const int klo = std::max(0, (lv - (ao.ni - 1) - 3 * (ao.nj - 1) + 8) / 9);
      const int khi = std::min(ao.nk - 1, lv / 9);
      k_weno_ao_e<<<grid_of(3 * (int64_t)ao.nj * (khi - klo + 1), 64), 64, 0, st>>>(P, ao, mom, aorec, lv, klo);
    }
  }
  if ((rc = mark())) return rc;
  k_diag<<<nblk_diag, DIAG_THREADS, 0, st>>>(P, all, part);
  k_diag_final<<<18, 256, 0, st>>>(part, nblk_diag, red);
  if ((rc = mark())) return rc;
  P1_HIP(hipGetLastError());
  if (kernel_ms) {
    P1_HIP(hipEventSynchronize(ev[nev - 1]));
    for (int q = 0; q + 1 < nev; ++q) P1_HIP(hipEventElapsedTime(&kernel_ms[q], ev[q], ev[q + 1]));
    for (int q = 0; q < NEV; ++q) (void)hipEventDestroy(ev[q]);
  }
  if (mom) {
    unsigned int bad = 0;
    P1_HIP(hipMemcpyAsync(&bad, F.counter + 2, sizeof(unsigned int), hipMemcpyDeviceToHost, st));
    P1_HIP(hipStreamSynchronize(st));
    if (bad) {
      cs::set_error("cs_limit_all_slopes_p12: the stencil table breaks the level order of step 9.c (an entry "
                    "beyond two zones in i or j), so it cannot follow the serial loop order level by level");
      return CS_E_INVALID;
    }
  }
  if (p->V_min_local || p->V_max_local) {
    double h[18];
    P1_HIP(hipMemcpyAsync(h, red, sizeof(h), hipMemcpyDeviceToHost, st));
    P1_HIP(hipStreamSynchronize(st));
    for (int q = 0; q < 9; ++q) {
      if (p->V_min_local) p->V_min_local[q] = h[q];
      if (p->V_max_local) p->V_max_local[q] = h[9 + q];
    }
  }
  (void)own;
  return 0;
}

extern "C" int cs_limit_all_slopes_p1(const cs_p1_patch* p, void* stream) { return run_p1(p, stream, nullptr); }

extern "C" int cs_limit_all_slopes_p1_timed(const cs_p1_patch* p, void* stream, float* kernel_ms) {
  if (!kernel_ms) {
    cs::set_error("cs_limit_all_slopes_p1_timed: kernel_ms is null");
    return CS_E_INVALID;
  }
  return run_p1(p, stream, kernel_ms);
}

// Work memory of cs_limit_all_slopes_p1_noalloc: the positivity flags, the
// replay list and counters and the min/max partials live in this fixed
// array instead of a hipMalloc'd buffer.  Each device holds its own copy
// (from the module load until the process ends) at its own address, which
// the host must look up: the array's name is no device address in host code.
constexpr size_t kNoAllocBytes = (size_t)48 << 20;
__device__ __align__(256) uint8_t scratch[kNoAllocBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_na_mu;
uint8_t* g_na_ptr[kMaxDevices] = {};  // device address of scratch, per device
}  // namespace

extern "C" int cs_limit_all_slopes_p1_noalloc(const cs_p1_patch* p, void* stream) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    cs::set_error("cs_limit_all_slopes_p1_noalloc: no usable current device");
    return CS_E_HIP;
  }
  uint8_t* ptr = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_na_mu);
    if (!g_na_ptr[dev]) {
      void* q = nullptr;
      const hipError_t e = hipGetSymbolAddress(&q, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !q) {
        cs::set_error("cs_limit_all_slopes_p1_noalloc: hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        return CS_E_HIP;
      }
      g_na_ptr[dev] = (uint8_t*)q;
    }
    ptr = g_na_ptr[dev];
  }
  return run_p1(p, stream, nullptr, ptr, kNoAllocBytes);
}

// ===========================================================================
// MAKE_FACIAL_FLUXES (400riemann_integrate/450make_facial_fluxes.f:5) for one
// patch at second order, behind cs_make_facial_fluxes_p1.
//
//   k_ff_face   one face of one family (blockIdx.y): the two zones' linear
//               fields at the face centroid, the face frame, the point Riemann
//               solver OneD_RS_MHD_Ptwise; S, fluctuations and entropy flux
//               written, and the two zones' shares of the update kept
//   k_ff_gather one zone: dU_dt from the shares of its six faces, in the
//               reference's order (x-faces, then y, then z; lower face first)
//   k_ff_remap  the few zones that also take the share of a face whose own
//               zone on that side is a tri-corner zone (the face then points
//               at the stencil neighbour): dU_dt rebuilt with those shares in
//               the order of the reference's loops
namespace csp1 {

struct P1F {
  cs_p1_flux a;
  int N, NR, g, nz, nzr, nf, nfr;
  int64_t NZ, NXF, NZF, NFC;
  __device__ __forceinline__ int64_t zi(int i, int j, int k) const {
    return (int64_t)(i - 1 + g) + (int64_t)nz * ((j - 1 + g) + (int64_t)nz * (k - 1 + g));
  }
  __device__ __forceinline__ int64_t fci(int i, int j, int k) const {
    return (int64_t)(i + g) + (int64_t)nf * ((j + g) + (int64_t)nf * (k + g));
  }
  __device__ __forceinline__ int64_t xfi(int i, int j, int k) const {
    return (int64_t)(i + g) + (int64_t)nf * ((j - 1 + g) + (int64_t)nz * (k - 1 + g));
  }
  __device__ __forceinline__ int64_t yfi(int i, int j, int k) const {
    return (int64_t)(i - 1 + g) + (int64_t)nz * ((j + g) + (int64_t)nf * (k - 1 + g));
  }
  __device__ __forceinline__ int64_t zfi(int i, int j, int k) const {
    return (int64_t)(i - 1 + g) + (int64_t)nz * ((j - 1 + g) + (int64_t)nz * (k + g));
  }
  __device__ __forceinline__ void nbr(int i, int j, int n, int& ni, int& nj) const {
    const int64_t s = (int64_t)(i - 1 + g) + (int64_t)nz * (j - 1 + g), plane = (int64_t)nz * nz;
    ni = min(max(a.stencil[s + plane * (2 * n)], 1 - g), N + g);
    nj = min(max(a.stencil[s + plane * (2 * n + 1)], 1 - g), N + g);
  }
  __device__ __forceinline__ bool active(int i, int j, int k) const { return a.mask[zi(i, j, k)] == 1; }
  __device__ __forceinline__ double obl(int64_t z) const { return 1.0 / fmax(a.scale[z], a.smallnum); }
  // face range of family fam (the reference's loops with iexpand = 1)
  __device__ __forceinline__ bool in_range(int fam, int i, int j, int k) const {
    return i >= (fam == 0 ? -1 : 0) && i <= N + 1 && j >= (fam == 1 ? -1 : 0) && j <= N + 1 &&
           k >= (fam == 2 ? -1 : 0) && k <= NR + 1;
  }
};

// OneD_RS_MHD_Ptwise with every output (610...f:96-220, 644-704), face frame
__device__ __forceinline__ void rs_ptwise(const Gas<double>& gs, double* VL, double* VR, double entL, double entR,
                                          double& SL, double& SR, double& entF, double* FlL, double* FlR,
                                          double* FL, double* FR) {
  const double bx = 0.5 * (VL[5] + VR[5]);
  VL[5] = bx;
  VR[5] = bx;
  double UL[8], UR[8], Vm[8];
  prim_to_cons(gs, entL, VL, UL);
  prim_to_cons(gs, entR, VR, UR);
  flux_x(VL, UL, FL);
  flux_x(VR, UR, FR);
  double smin, smax;
  SR = -1e30;
  SL = 1e30;
  wavspd_x(gs, VL, smin, smax);
  SL = fmin(SL, smin);
  SR = fmax(SR, smax);
  wavspd_x(gs, VR, smin, smax);
  SL = fmin(SL, smin);
  SR = fmax(SR, smax);
#pragma unroll
  for (int c = 0; c < 8; ++c) Vm[c] = 0.5 * (VL[c] + VR[c]);
  wavspd_x(gs, Vm, smin, smax);
  SL = fmin(SL, smin);
  SR = fmax(SR, smax);
  if (SL >= 0.0) {
    // U* = U_L before the one-sided fluctuation is formed (610...f:653-658)
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      FlR[c] = SR * (UR[c] - UL[c]);
      FlL[c] = 0.0;
    }
    entF = entL * VL[1];
    return;
  }
  if (SR <= 0.0) {
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      FlL[c] = SL * (UR[c] - UL[c]);
      FlR[c] = 0.0;
    }
    entF = entR * VR[1];
    return;
  }
  const double t3 = 1.0 / (SR - SL), t1 = SR * t3, t2 = SL * t3;
  double Us[8];
#pragma unroll
  for (int c = 0; c < 8; ++c) Us[c] = t1 * UR[c] - t2 * UL[c] - t3 * (FR[c] - FL[c]);
  const double entUs = t1 * entR - t2 * entL - t3 * (entR * VR[1] - entL * VL[1]);
  const double pent = entUs * rpow_gm1(gs, Us[0]);
  const double peng = gs.gm1 * (Us[4] - 0.5 * (Us[1] * Us[1] + Us[2] * Us[2] + Us[3] * Us[3]) / Us[0] -
                                C<double>::ob8pi * (Us[5] * Us[5] + Us[6] * Us[6] + Us[7] * Us[7]));
  if (peng <= gs.prsfloor) Us[4] = Us[4] + (pent - peng) * gs.igm1;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    FlL[c] = SL * (Us[c] - UL[c]);
    FlR[c] = SR * (UR[c] - Us[c]);
  }
  const double o = 1.0 / (SR - SL);
  entF = (SR * entL * VL[1] - SL * entR * VR[1]) * o + (entR - entL) * (SR * SL * o);
}

// one face; the shares cl (to the left zone) and cr (to the right) are kept
// in share[fam][c][face] (c < 8 left, c >= 8 right); faces whose side zone was
// re-pointed through the stencil are appended to the remap list
__global__ void __launch_bounds__(128) k_ff_face(const P1F P, double* share, int64_t* rlist, int64_t cap,
                                                 unsigned int* rcount) {
  const int fam = blockIdx.y;
  const int64_t t = tid();
  if (t >= P.NFC) return;
  const int n = P.nf;
  const int i = (int)(t % n) - P.g, j = (int)((t / n) % n) - P.g, k = (int)(t / ((int64_t)n * n)) - P.g;
  const cs_p1_flux& a = P.a;
  const int64_t fb = t;
  double* FPl = fam == 0 ? a.Fluct_x_Pl : (fam == 1 ? a.Fluct_y_Pl : a.Fluct_z_Pl);
  double* FMi = fam == 0 ? a.Fluct_x_Mi : (fam == 1 ? a.Fluct_y_Mi : a.Fluct_z_Mi);
  // faces outside the range or between two tri-corner zones keep zero
  // fluctuations (450...f:370-380)
  auto zero = [&]() {
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      FPl[fb + P.NFC * c] = 0.0;
      FMi[fb + P.NFC * c] = 0.0;
    }
  };
  if (!P.in_range(fam, i, j, k)) {
    zero();
    return;
  }
  int li = i, lj = j, ri = i + (fam == 0), rj = j + (fam == 1);
  const int rk = k + (fam == 2);
  int remapped = 0;  // 1: left side re-pointed, 2: right side
  if (fam < 2) {
    const bool al = P.active(li, lj, k), ar = P.active(ri, rj, k);
    if (!al && ar) {
      P.nbr(ri, rj, fam == 0 ? 1 : 3, li, lj);
      remapped = 1;
    } else if (al && !ar) {
      P.nbr(li, lj, fam == 0 ? 0 : 2, ri, rj);
      remapped = 2;
    } else if (!al && !ar) {
      zero();
      return;
    }
  } else if (!P.active(i, j, k)) {
    zero();
    return;
  }
  const int64_t zl = P.zi(li, lj, k), zr = P.zi(ri, rj, fam == 2 ? rk : k);
  const int64_t fg = fam == 0 ? P.xfi(i, j, k) : (fam == 1 ? P.yfi(i, j, k) : P.zfi(i, j, k));
  const int64_t NF = fam == 2 ? P.NZF : P.NXF;
  const double* nrm = fam == 0 ? a.xf_normal : (fam == 1 ? a.yf_normal : a.zf_normal);
  const double* fcen = fam == 0 ? a.xf_centroid : (fam == 1 ? a.yf_centroid : a.zf_centroid);
  const double* pts = fam == 0 ? a.xf_pts : (fam == 1 ? a.yf_pts : a.zf_pts);
  const double sm = a.smallnum;
  double A[3], B[3], Cv[3];
  if (fam == 0) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      A[c] = nrm[fg + NF * c];
      Cv[c] = fcen[fg + NF * c];
    }
    const double s = 1.0 / fmax(sqrt(Cv[0] * Cv[0] + Cv[1] * Cv[1] + Cv[2] * Cv[2]), sm);
#pragma unroll
    for (int c = 0; c < 3; ++c) Cv[c] = Cv[c] * s;
    B[0] = Cv[1] * A[2] - Cv[2] * A[1];
    B[1] = Cv[2] * A[0] - Cv[0] * A[2];
    B[2] = Cv[0] * A[1] - Cv[1] * A[0];
  } else if (fam == 1) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      A[c] = nrm[fg + NF * c];
      B[c] = fcen[fg + NF * c];
    }
    const double s = 1.0 / fmax(sqrt(B[0] * B[0] + B[1] * B[1] + B[2] * B[2]), sm);
#pragma unroll
    for (int c = 0; c < 3; ++c) B[c] = B[c] * s;
    Cv[0] = A[1] * B[2] - A[2] * B[1];
    Cv[1] = A[2] * B[0] - A[0] * B[2];
    Cv[2] = A[0] * B[1] - A[1] * B[0];
  } else {
#pragma unroll
    for (int c = 0; c < 3; ++c) A[c] = nrm[fg + NF * c];
    double s = 1.0 / fmax(sqrt(A[0] * A[0] + A[1] * A[1] + A[2] * A[2]), sm);
#pragma unroll
    for (int c = 0; c < 3; ++c) A[c] = A[c] * s;
#pragma unroll
    for (int c = 0; c < 3; ++c)
      B[c] = a.vert[P.fci(i, j, k) + P.NFC * c] + a.vert[P.fci(i, j - 1, k) + P.NFC * c] -
             a.vert[P.fci(i - 1, j, k) + P.NFC * c] - a.vert[P.fci(i - 1, j - 1, k) + P.NFC * c];
    s = B[0] * A[0] + B[1] * A[1] + B[2] * A[2];
#pragma unroll
    for (int c = 0; c < 3; ++c) B[c] = B[c] - s * A[c];
    s = 1.0 / fmax(sqrt(B[0] * B[0] + B[1] * B[1] + B[2] * B[2]), sm);
#pragma unroll
    for (int c = 0; c < 3; ++c) B[c] = B[c] * s;
    Cv[0] = A[1] * B[2] - A[2] * B[1];
    Cv[1] = A[2] * B[0] - A[0] * B[2];
    Cv[2] = A[0] * B[1] - A[1] * B[0];
    s = 1.0 / fmax(sqrt(Cv[0] * Cv[0] + Cv[1] * Cv[1] + Cv[2] * Cv[2]), sm);
#pragma unroll
    for (int c = 0; c < 3; ++c) Cv[c] = Cv[c] * s;
    B[0] = Cv[1] * A[2] - Cv[2] * A[1];
    B[1] = Cv[2] * A[0] - Cv[0] * A[2];
    B[2] = Cv[0] * A[1] - Cv[1] * A[0];
  }
  // both zones' linear fields at the face centroid, into the face frame
  double VL[8], VR[8], eL, eR;
  {
    const double ol = P.obl(zl), orr = P.obl(zr);
    double rl[3], rr[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double x = pts[fg + NF * c];
      rl[c] = (x - a.v_centroid[zl + P.NZ * c]) * ol;
      rr[c] = (x - a.v_centroid[zr + P.NZ * c]) * orr;
    }
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const double* v = a.V + P.NZ * c;
      VL[c] = v[zl] + v[zl + 8 * P.NZ] * rl[0] + v[zl + 16 * P.NZ] * rl[1] + v[zl + 24 * P.NZ] * rl[2];
      VR[c] = v[zr] + v[zr + 8 * P.NZ] * rr[0] + v[zr + 16 * P.NZ] * rr[1] + v[zr + 24 * P.NZ] * rr[2];
    }
    eL = a.ent[zl] + a.ent[zl + P.NZ] * rl[0] + a.ent[zl + 2 * P.NZ] * rl[1] + a.ent[zl + 3 * P.NZ] * rl[2];
    eR = a.ent[zr] + a.ent[zr + P.NZ] * rr[0] + a.ent[zr + 2 * P.NZ] * rr[1] + a.ent[zr + 3 * P.NZ] * rr[2];
#pragma unroll
    for (int o = 1; o <= 5; o += 4) {
      const double la = A[0] * VL[o] + A[1] * VL[o + 1] + A[2] * VL[o + 2];
      const double lb = B[0] * VL[o] + B[1] * VL[o + 1] + B[2] * VL[o + 2];
      const double lc = Cv[0] * VL[o] + Cv[1] * VL[o + 1] + Cv[2] * VL[o + 2];
      VL[o] = la;
      VL[o + 1] = lb;
      VL[o + 2] = lc;
      const double ra = A[0] * VR[o] + A[1] * VR[o + 1] + A[2] * VR[o + 2];
      const double rb = B[0] * VR[o] + B[1] * VR[o + 1] + B[2] * VR[o + 2];
      const double rc = Cv[0] * VR[o] + Cv[1] * VR[o + 1] + Cv[2] * VR[o + 2];
      VR[o] = ra;
      VR[o + 1] = rb;
      VR[o + 2] = rc;
    }
  }
  const Gas<double> gs{a.gamma, a.gamma - 1.0, a.rhofloor, a.prsfloor, 1.0 / (a.gamma - 1.0)};
  double SL, SR, eF, FlL[8], FlR[8], FL[8], FR[8];
  rs_ptwise(gs, VL, VR, eL, eR, SL, SR, eF, FlL, FlR, FL, FR);
  (fam == 0 ? a.S_L : fam == 1 ? a.S_D : a.S_B)[fb] = SL;
  (fam == 0 ? a.S_R : fam == 1 ? a.S_U : a.S_T)[fb] = SR;
  (fam == 0 ? a.EntFlux_x : fam == 1 ? a.EntFlux_y : a.EntFlux_z)[fb] = eF;
  const double ar = (fam == 0 ? a.xf_area : fam == 1 ? a.yf_area : a.zf_area)[fg];
  const double vl = a.v_volume[zl], vr = a.v_volume[zr];
  double* sh = share + (int64_t)fam * 16 * P.NFC;
  // back to the global frame: scalars as they are, vectors through (A, B, C)
  auto back = [&](const double* f, int c) -> double {
    if (c == 0 || c == 4) return f[c];
    const int o = c < 4 ? 1 : 5, q = c - o;
    return A[q] * f[o] + B[q] * f[o + 1] + Cv[q] * f[o + 2];
  };
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const double pl = back(FlR, c), mi = back(FlL, c), fll = back(FL, c), frr = back(FR, c);
    FPl[fb + P.NFC * c] = pl;
    FMi[fb + P.NFC * c] = mi;
    sh[fb + P.NFC * c] = (mi + fll) * ar / vl;
    sh[fb + P.NFC * (8 + c)] = (pl - frr) * ar / vr;
  }
  if (remapped) {
    const unsigned int e = atomicAdd(rcount, 1u);
    // target zone; (family, face, side) packed in the second half
    rlist[e] = remapped == 1 ? zl : zr;
    rlist[cap + e] = (((int64_t)fam * P.NFC + fb) << 1) | (remapped == 1 ? 0 : 1);
  }
}

// dU_dt of one zone from its own faces: x lower (as right zone), x upper (as
// left zone), then y, then z
__global__ void __launch_bounds__(256) k_ff_gather(const P1F P, const double* share) {
  const int64_t z = tid();
  if (z >= P.NZ) return;
  const int i = (int)(z % P.nz) + 1 - P.g, j = (int)((z / P.nz) % P.nz) + 1 - P.g,
            k = (int)(z / ((int64_t)P.nz * P.nz)) + 1 - P.g;
  double d[8];
#pragma unroll
  for (int c = 0; c < 8; ++c) d[c] = 0.0;
  // a tri-corner zone takes no share of any face
  const bool act = P.active(i, j, k);
#pragma unroll
  for (int fam = 0; fam < 3; ++fam) {
    if (!act) break;
    const double* sh = share + (int64_t)fam * 16 * P.NFC;
    const int di = fam == 0, dj = fam == 1, dk = fam == 2;
    if (P.in_range(fam, i - di, j - dj, k - dk)) {
      const int64_t f = P.fci(i - di, j - dj, k - dk);
#pragma unroll
      for (int c = 0; c < 8; ++c) d[c] = d[c] - sh[f + P.NFC * (8 + c)];
    }
    if (P.in_range(fam, i, j, k)) {
      const int64_t f = P.fci(i, j, k);
#pragma unroll
      for (int c = 0; c < 8; ++c) d[c] = d[c] - sh[f + P.NFC * c];
    }
  }
#pragma unroll
  for (int c = 0; c < 8; ++c) P.a.dU_dt[z + P.NZ * c] = d[c];
}

// zones that also take re-pointed shares: dU_dt rebuilt in loop order, one
// thread per such zone.  grp[2 q], grp[2 q + 1]: the zone and the first of
// its entries in codes (the host groups the list by zone); codes packs
// (family, face, side).
__global__ void __launch_bounds__(64) k_ff_remap(const P1F P, const double* share, const int64_t* grp,
                                                 const int64_t* codes, int ngrp, int ncodes) {
  const int q0 = (int)tid();
  if (q0 >= ngrp) return;
  const int64_t T = grp[2 * q0], e0 = grp[2 * q0 + 1], e1 = q0 + 1 < ngrp ? grp[2 * q0 + 3] : ncodes;
  const int i = (int)(T % P.nz) + 1 - P.g, j = (int)((T / P.nz) % P.nz) + 1 - P.g,
            k = (int)(T / ((int64_t)P.nz * P.nz)) + 1 - P.g;
  double d[8];
  for (int c = 0; c < 8; ++c) d[c] = 0.0;
  for (int fam = 0; fam < 3; ++fam) {
    // the shares reaching T in this family: (face, side), sorted by face
    int64_t fc[8];
    int sd[8], m = 0;
    const int di = fam == 0, dj = fam == 1, dk = fam == 2;
    if (P.active(i, j, k)) {
      if (P.in_range(fam, i - di, j - dj, k - dk)) {
        fc[m] = P.fci(i - di, j - dj, k - dk);
        sd[m++] = 1;
      }
      if (P.in_range(fam, i, j, k)) {
        fc[m] = P.fci(i, j, k);
        sd[m++] = 0;
      }
    }
    for (int64_t q = e0; q < e1 && m < 8; ++q) {
      const int64_t code = codes[q], ff = code >> 1;
      if (ff / P.NFC != fam) continue;
      fc[m] = ff % P.NFC;
      sd[m++] = (int)(code & 1);
    }
    for (int a1 = 1; a1 < m; ++a1)  // insertion sort by face index
      for (int b1 = a1; b1 > 0 && fc[b1] < fc[b1 - 1]; --b1) {
        const int64_t tf = fc[b1];
        fc[b1] = fc[b1 - 1];
        fc[b1 - 1] = tf;
        const int ts = sd[b1];
        sd[b1] = sd[b1 - 1];
        sd[b1 - 1] = ts;
      }
    const double* sh = share + (int64_t)fam * 16 * P.NFC;
    for (int q = 0; q < m; ++q)
      for (int c = 0; c < 8; ++c) d[c] = d[c] - sh[fc[q] + P.NFC * (8 * sd[q] + c)];
  }
  for (int c = 0; c < 8; ++c) P.a.dU_dt[T + P.NZ * c] = d[c];
}

__global__ void __launch_bounds__(256) k_ff_zero_e(const P1F P) {
  const int64_t z = tid();
  if (z >= P.NZ) return;
  const int i = (int)(z % P.nz) + 1 - P.g, j = (int)((z / P.nz) % P.nz) + 1 - P.g,
            k = (int)(z / ((int64_t)P.nz * P.nz)) + 1 - P.g;
  const int64_t f = P.fci(i, j, k);
  P.a.Efx[f] = 0.0;
  P.a.Efy[f] = 0.0;
  P.a.Efz[f] = 0.0;
}

struct FfScratch {
  void* mem = nullptr;
  size_t bytes = 0;
};
static std::vector<FfScratch> g_ff_scratch;

}  // namespace csp1

extern "C" int cs_make_facial_fluxes_p1(const cs_p1_flux* p, void* stream_) {
  if (!p) {
    cs::set_error("cs_make_facial_fluxes_p1: null argument");
    return CS_E_INVALID;
  }
  if (p->ntp < 1 || p->nr < 1 || p->nghost < 4) {
    cs::set_error("cs_make_facial_fluxes_p1: need ntp >= 1, nr >= 1, nghost >= 4 (got %d, %d, %d)", p->ntp, p->nr,
                  p->nghost);
    return CS_E_INVALID;
  }
  if (p->nghost >= p->ntp) {
    cs::set_error("cs_make_facial_fluxes_p1: nghost (%d) must be below ntp (%d)", p->nghost, p->ntp);
    return CS_E_BAD_DECOMP;
  }
  if (!(p->gamma > 1.0) || !(p->smallnum > 0.0) || !(p->rhofloor > 0.0) || !(p->prsfloor > 0.0)) {
    cs::set_error("cs_make_facial_fluxes_p1: gamma > 1 and positive floors and smallnum required");
    return CS_E_INVALID;
  }
  const void* ptrs[] = {p->vert, p->mask, p->stencil, p->xf_pts, p->xf_normal, p->xf_area, p->xf_centroid,
                        p->yf_pts, p->yf_normal, p->yf_area, p->yf_centroid, p->zf_pts, p->zf_normal, p->zf_area,
                        p->zf_centroid, p->scale, p->v_volume, p->v_centroid, p->V, p->ent, p->flatten, p->S_L,
                        p->S_R, p->S_D, p->S_U, p->S_B, p->S_T, p->Fluct_x_Pl, p->Fluct_y_Pl, p->Fluct_z_Pl,
                        p->Fluct_x_Mi, p->Fluct_y_Mi, p->Fluct_z_Mi, p->dU_dt, p->EntFlux_x, p->EntFlux_y,
                        p->EntFlux_z, p->Efx, p->Efy, p->Efz};
  for (const void* q : ptrs)
    if (!q) {
      cs::set_error("cs_make_facial_fluxes_p1: null array");
      return CS_E_INVALID;
    }
  hipStream_t st = (hipStream_t)stream_;
  P1F P;
  P.a = *p;
  P.N = p->ntp;
  P.NR = p->nr;
  P.g = p->nghost;
  P.nz = P.N + 2 * P.g;
  P.nzr = P.NR + 2 * P.g;
  P.nf = P.nz + 1;
  P.nfr = P.nzr + 1;
  P.NZ = (int64_t)P.nz * P.nz * P.nzr;
  P.NXF = (int64_t)P.nf * P.nz * P.nzr;
  P.NZF = (int64_t)P.nz * P.nz * P.nfr;
  P.NFC = (int64_t)P.nf * P.nf * P.nfr;
  int dev = 0;
  P1_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g_mu);
  if ((int)g_ff_scratch.size() <= dev) g_ff_scratch.resize(dev + 1);
  FfScratch& S = g_ff_scratch[dev];
  const size_t share_bytes = (size_t)48 * P.NFC * sizeof(double);
  const size_t list_cap = (size_t)2 * P.NFC;  // faces of two families at most
  const size_t need = share_bytes + 256 + list_cap * 2 * sizeof(int64_t);
  if (S.bytes < need) {
    if (S.mem) P1_HIP(hipFree(S.mem));
    S.mem = nullptr;
    S.bytes = 0;
    if (hipMalloc(&S.mem, need) != hipSuccess) {
      cs::set_error("cs_make_facial_fluxes_p1: cannot allocate %zu bytes of work memory", need);
      return CS_E_NOMEM;
    }
    S.bytes = need;
  }
  double* share = (double*)S.mem;
  unsigned int* rcount = (unsigned int*)((uint8_t*)S.mem + share_bytes);
  int64_t* rlist = (int64_t*)((uint8_t*)S.mem + share_bytes + 256);
  // the face kernel writes every fluctuation (zero where no Riemann problem
  // is solved) and the gather every zone's update; the electric fields of
  // the zone range are zeroed (450...f:400-420)
  P1_HIP(hipMemsetAsync(rcount, 0, sizeof(unsigned int), st));
  k_ff_zero_e<<<grid_of(P.NZ, 256), 256, 0, st>>>(P);
  k_ff_face<<<dim3(grid_of(P.NFC, 128), 3), 128, 0, st>>>(P, share, rlist, (int64_t)list_cap, rcount);
  k_ff_gather<<<grid_of(P.NZ, 256), 256, 0, st>>>(P, share);
  P1_HIP(hipGetLastError());
  unsigned int nrm = 0;
  P1_HIP(hipMemcpyAsync(&nrm, rcount, sizeof(unsigned int), hipMemcpyDeviceToHost, st));
  P1_HIP(hipStreamSynchronize(st));
  if (nrm > 0) {
    // group the entries by zone on the host (a few per radial layer), then
    // one thread per zone
    std::vector<int64_t> tz(nrm), code(nrm), order(nrm);
    P1_HIP(hipMemcpy(tz.data(), rlist, sizeof(int64_t) * nrm, hipMemcpyDeviceToHost));
    P1_HIP(hipMemcpy(code.data(), rlist + list_cap, sizeof(int64_t) * nrm, hipMemcpyDeviceToHost));
    for (unsigned int q = 0; q < nrm; ++q) order[q] = q;
    std::sort(order.begin(), order.end(), [&](int64_t x, int64_t y) {
      return tz[x] != tz[y] ? tz[x] < tz[y] : code[x] < code[y];
    });
    std::vector<int64_t> grp, codes(nrm);
    for (unsigned int q = 0; q < nrm; ++q) {
      codes[q] = code[order[q]];
      if (q == 0 || tz[order[q]] != tz[order[q - 1]]) {
        grp.push_back(tz[order[q]]);
        grp.push_back(q);
      }
    }
    const int ngrp = (int)(grp.size() / 2);
    if (grp.size() > list_cap) {
      cs::set_error("cs_make_facial_fluxes_p1: %d zones take re-pointed shares", ngrp);
      return CS_E_INVALID;
    }
    int64_t* dgrp = rlist;  // the list has been read: its space holds the groups
    int64_t* dcodes = rlist + list_cap;
    P1_HIP(hipMemcpyAsync(dgrp, grp.data(), sizeof(int64_t) * grp.size(), hipMemcpyHostToDevice, st));
    P1_HIP(hipMemcpyAsync(dcodes, codes.data(), sizeof(int64_t) * nrm, hipMemcpyHostToDevice, st));
    k_ff_remap<<<grid_of(ngrp, 64), 64, 0, st>>>(P, share, dgrp, dcodes, ngrp, (int)nrm);
    P1_HIP(hipStreamSynchronize(st));  // the host buffers above are released on return
  }
  P1_HIP(hipGetLastError());
  return 0;
}

// ===========================================================================
// UPDATE_IMEX_RK_STEP_P1 (400riemann_integrate/484update_imex_rk_step_p1.f:7),
// MHD, behind cs_update_imex_rk_step_p1:
//   k_upd_zone  every zone: U += dt dU_dt (:483-489); owned zones: entropy from
//               the face entropy fluxes, CONS_TO_PRIM, its flag (:540-628)
//   k_upd_face  owned faces of one family (blockIdx.y): constrained transport
//               with the edges' radial lengths and arc lengths (:630-906)
namespace csp1 {

struct P1U {
  cs_p1_update a;
  int N, NR, g, nz, nf;
  int64_t NZ, NXF, NZF, NFC;
  __device__ __forceinline__ int64_t fci(int i, int j, int k) const {
    return (int64_t)(i + g) + (int64_t)nf * ((j + g) + (int64_t)nf * (k + g));
  }
  __device__ __forceinline__ int64_t xfi(int i, int j, int k) const {
    return (int64_t)(i + g) + (int64_t)nf * ((j - 1 + g) + (int64_t)nz * (k - 1 + g));
  }
  __device__ __forceinline__ int64_t yfi(int i, int j, int k) const {
    return (int64_t)(i - 1 + g) + (int64_t)nz * ((j + g) + (int64_t)nf * (k - 1 + g));
  }
  __device__ __forceinline__ int64_t zfi(int i, int j, int k) const {
    return (int64_t)(i - 1 + g) + (int64_t)nz * ((j - 1 + g) + (int64_t)nz * (k + g));
  }
  __device__ __forceinline__ double vx(int i, int j, int k, int c) const { return a.vert[fci(i, j, k) + NFC * c]; }
  __device__ __forceinline__ double rad(int i, int j, int k) const {
    const double x = vx(i, j, k, 0), y = vx(i, j, k, 1), z = vx(i, j, k, 2);
    return sqrt(x * x + y * y + z * z);
  }
  // arc between vertices (i1, j1) and (i2, j2) of level k: both directions
  // scaled by t = 1 / (radius of the face's own vertex), as the reference does
  __device__ __forceinline__ double angle(int i1, int j1, int i2, int j2, int k, double t) const {
    const double d = (vx(i1, j1, k, 0) * t) * (vx(i2, j2, k, 0) * t) + (vx(i1, j1, k, 1) * t) * (vx(i2, j2, k, 1) * t) +
                     (vx(i1, j1, k, 2) * t) * (vx(i2, j2, k, 2) * t);
    return acos(d);
  }
};

__global__ void __launch_bounds__(256) k_upd_zone(const P1U P) {
  const int64_t z = tid();
  if (z >= P.NZ) return;
  const cs_p1_update& a = P.a;
  const int i = (int)(z % P.nz) + 1 - P.g, j = (int)((z / P.nz) % P.nz) + 1 - P.g,
            k = (int)(z / ((int64_t)P.nz * P.nz)) + 1 - P.g;
  const double dt = a.dtmove;
  double U[8];
#pragma unroll
  for (int c = 0; c < 8; ++c) U[c] = a.U[z + P.NZ * c] + dt * a.dU_dt[z + P.NZ * c];
  double flag = 0.0;
  if (i >= 1 && i <= P.N && j >= 1 && j <= P.N && k >= 1 && k <= P.NR) {
    const double s = a.xf_area[P.xfi(i, j, k)] * a.EntFlux_x[P.fci(i, j, k)] -
                     a.xf_area[P.xfi(i - 1, j, k)] * a.EntFlux_x[P.fci(i - 1, j, k)] +
                     a.yf_area[P.yfi(i, j, k)] * a.EntFlux_y[P.fci(i, j, k)] -
                     a.yf_area[P.yfi(i, j - 1, k)] * a.EntFlux_y[P.fci(i, j - 1, k)] +
                     a.zf_area[P.zfi(i, j, k)] * a.EntFlux_z[P.fci(i, j, k)] -
                     a.zf_area[P.zfi(i, j, k - 1)] * a.EntFlux_z[P.fci(i, j, k - 1)];
    double e = a.ent[z] - dt * s / a.v_volume[z];
    const Gas<double> gs{a.gamma, a.gamma - 1.0, a.rhofloor, a.prsfloor, 1.0 / (a.gamma - 1.0)};
    double V[8];
    flag = (double)cons_to_prim(gs, U, V, e);
    a.ent[z] = e;
  }
#pragma unroll
  for (int c = 0; c < 8; ++c) a.U[z + P.NZ * c] = U[c];
  a.flag_entprs[z] = flag;
}

__global__ void __launch_bounds__(256) k_upd_face(const P1U P) {
  const int fam = blockIdx.y;
  const int64_t t = tid();
  const int N = P.N, NR = P.NR;
  // owned faces: x (0..N, 1..N, 1..NR), y (1..N, 0..N, 1..NR), z (1..N, 1..N, 0..NR)
  const int ni = fam == 0 ? N + 1 : N, nj = fam == 1 ? N + 1 : N, nk = fam == 2 ? NR + 1 : NR;
  if (t >= (int64_t)ni * nj * nk) return;
  const int i = (int)(t % ni) + (fam == 0 ? 0 : 1), j = (int)((t / ni) % nj) + (fam == 1 ? 0 : 1),
            k = (int)(t / ((int64_t)ni * nj)) + (fam == 2 ? 0 : 1);
  const cs_p1_update& a = P.a;
  const double dt = a.dtmove;
  const int64_t f = P.fci(i, j, k);
  const double rout = P.rad(i, j, k), tt = 1.0 / rout;
  if (fam == 0) {
    const double rin = P.rad(i, j, k - 1), dr = rout - rin, th = P.angle(i, j, i, j - 1, k, tt);
    const double lin = rin * th, lout = rout * th;
    a.Bxf[f] = a.Bxf[f] - dt * (dr * a.Efz[f] - dr * a.Efz[P.fci(i, j - 1, k)] - lout * a.Efy[f] +
                                lin * a.Efy[P.fci(i, j, k - 1)]) / a.xf_area[P.xfi(i, j, k)];
  } else if (fam == 1) {
    const double rin = P.rad(i, j, k - 1), dr = rout - rin, th = P.angle(i, j, i - 1, j, k, tt);
    const double lin = rin * th, lout = rout * th;
    a.Byf[f] = a.Byf[f] - dt * (-(dr * a.Efz[f]) + dr * a.Efz[P.fci(i - 1, j, k)] + lout * a.Efx[f] -
                                lin * a.Efx[P.fci(i, j, k - 1)]) / a.yf_area[P.yfi(i, j, k)];
  } else {
    const double lxp = rout * P.angle(i, j, i - 1, j, k, tt), lxm = rout * P.angle(i, j - 1, i - 1, j - 1, k, tt);
    const double lyp = rout * P.angle(i, j, i, j - 1, k, tt), lym = rout * P.angle(i - 1, j, i - 1, j - 1, k, tt);
    a.Bzf[f] = a.Bzf[f] - dt * (lyp * a.Efy[f] - lym * a.Efy[P.fci(i - 1, j, k)] - lxp * a.Efx[f] +
                                lxm * a.Efx[P.fci(i, j - 1, k)]) / a.zf_area[P.zfi(i, j, k)];
  }
}

}  // namespace csp1

extern "C" int cs_update_imex_rk_step_p1(const cs_p1_update* p, void* stream_) {
  if (!p) {
    cs::set_error("cs_update_imex_rk_step_p1: null argument");
    return CS_E_INVALID;
  }
  if (p->ntp < 1 || p->nr < 1 || p->nghost < 1) {
    cs::set_error("cs_update_imex_rk_step_p1: need ntp, nr, nghost >= 1 (got %d, %d, %d)", p->ntp, p->nr,
                  p->nghost);
    return CS_E_INVALID;
  }
  if (!(p->gamma > 1.0) || !(p->rhofloor > 0.0) || !(p->prsfloor > 0.0)) {
    cs::set_error("cs_update_imex_rk_step_p1: gamma > 1 and positive floors required");
    return CS_E_INVALID;
  }
  const void* ptrs[] = {p->vert, p->xf_area, p->yf_area, p->zf_area, p->v_volume, p->U, p->ent, p->Bxf,
                        p->Byf, p->Bzf, p->dU_dt, p->EntFlux_x, p->EntFlux_y, p->EntFlux_z, p->Efx, p->Efy,
                        p->Efz, p->flag_entprs};
  for (const void* q : ptrs)
    if (!q) {
      cs::set_error("cs_update_imex_rk_step_p1: null array");
      return CS_E_INVALID;
    }
  hipStream_t st = (hipStream_t)stream_;
  P1U P;
  P.a = *p;
  P.N = p->ntp;
  P.NR = p->nr;
  P.g = p->nghost;
  P.nz = P.N + 2 * P.g;
  P.nf = P.nz + 1;
  const int nzr = P.NR + 2 * P.g;
  P.NZ = (int64_t)P.nz * P.nz * nzr;
  P.NXF = (int64_t)P.nf * P.nz * nzr;
  P.NZF = (int64_t)P.nz * P.nz * (nzr + 1);
  P.NFC = (int64_t)P.nf * P.nf * (nzr + 1);
  const int64_t nfaces = (int64_t)(P.N + 1) * (P.N + 1) * (P.NR + 1);
  k_upd_zone<<<grid_of(P.NZ, 256), 256, 0, st>>>(P);
  k_upd_face<<<dim3(grid_of(nfaces, 256), 3), 256, 0, st>>>(P);
  P1_HIP(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------------------
// CLEANUP_AFTER_UPDATE (cubesphere/400riemann_integrate/485cleanup_after_update.f:1)
// and EVALUATE_TIMESTEP (490evaluate_timestep.f:1) of one patch, MHD, on the
// fields of cs_p1_patch:
//   k_face_zero, k_tricorner  steps 2 (faces) and 3 of the cleanup: the
//                             reconstruction's kernels, whose ranges and rules
//                             are the same (485...f:366-686)
//   k_cleanup_zone            step 2 (zones), step 4.a: the zone B as the least-
//                             squares fit of the six face values over the face
//                             normals, the energy rebalanced (:691-865), and
//                             step 5: CONS_TO_PRIM of every zone (:1180-1280)
//   k_dt_init, k_dt           CONS_TO_PRIM of the owned active zones, their
//                             magnetosonic speed, and the minimum of
//                             cfl l / (|v| + c_f) with 1.1 dt (490...f:72-163):
//                             a block minimum, folded into the output by a
//                             compare-and-swap minimum (positive values, so
//                             the result does not depend on the block order)
namespace csp1 {

__global__ void __launch_bounds__(256) k_cleanup_zone(const P1 P) {
  const int64_t z = tid();
  if (z >= P.NZ) return;
  const cs_p1_patch& a = P.a;
  const int i = (int)(z % P.nz) + 1 - P.g, j = (int)((z / P.nz) % P.nz) + 1 - P.g,
            k = (int)(z / ((int64_t)P.nz * P.nz)) + 1 - P.g;
  const double gam = a.gamma;
  a.gamma_eos[z] = gam;
  a.gamma_sound[z] = gam;
  a.mean_mol_wt[z] = a.molewt;
#pragma unroll
  for (int c = 0; c < 8; ++c)
#pragma unroll
    for (int m = 1; m < 4; ++m) {
      P.Ur(c, m, z) = 0.0;
      P.Vr(c, m, z) = 0.0;
    }
#pragma unroll
  for (int m = 1; m < 4; ++m) a.ent[z + P.NZ * m] = 0.0;
  double U[8], V[8];
  if (a.mask[z] == 0) {
#pragma unroll
    for (int c = 0; c < 8; ++c) U[c] = 0.0;
    U[0] = a.rhofloor;
    U[4] = a.prsfloor / (gam - 1.0);
    a.scale[z] = a.smallnum;
  } else {
#pragma unroll
    for (int c = 0; c < 8; ++c) U[c] = P.Ur(c, 0, z);
  }
  if (P.active(i, j, 1)) {
    // rows: x+ and x- face normals, y+ and y-, and the radial direction of
    // the zone's upper z-face centroid for both z-faces
    double A[6][3], R[6];
    const int64_t fz = P.zfi(i, j, k), xp = P.xfi(i, j, k), xm = P.xfi(i - 1, j, k), yp = P.yfi(i, j, k),
                  ym = P.yfi(i, j - 1, k);
    double nv[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) nv[c] = a.zf_centroid[fz + P.NZF * c];
    const double t = 1.0 / fmax(a.smallnum, sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]));
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      A[0][c] = a.xf_normal[xp + P.NXF * c];
      A[1][c] = a.xf_normal[xm + P.NXF * c];
      A[2][c] = a.yf_normal[yp + P.NXF * c];
      A[3][c] = a.yf_normal[ym + P.NXF * c];
      A[4][c] = nv[c] * t;
      A[5][c] = A[4][c];
    }
    R[0] = a.Bxf[P.fci(i, j, k)];
    R[1] = a.Bxf[P.fci(i - 1, j, k)];
    R[2] = a.Byf[P.fci(i, j, k)];
    R[3] = a.Byf[P.fci(i, j - 1, k)];
    R[4] = a.Bzf[P.fci(i, j, k)];
    R[5] = a.Bzf[P.fci(i, j, k - 1)];
    double M[3][3], r3[3];
#pragma unroll
    for (int p = 0; p < 3; ++p) {
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        double s = 0.0;
#pragma unroll
        for (int r = 0; r < 6; ++r) s += A[r][p] * A[r][q];
        M[p][q] = s;
      }
      double s = 0.0;
#pragma unroll
      for (int r = 0; r < 6; ++r) s += A[r][p] * R[r];
      r3[p] = s;
    }
    const double det = M[0][0] * (M[1][1] * M[2][2] - M[2][1] * M[1][2]) -
                       M[0][1] * (M[1][0] * M[2][2] - M[2][0] * M[1][2]) +
                       M[0][2] * (M[1][0] * M[2][1] - M[2][0] * M[1][1]);
    const double od = copysign(1.0, det) / fmax(1.0e-32, fabs(det));
    const double b[3][3] = {{(M[1][1] * M[2][2] - M[1][2] * M[2][1]) * od, (M[0][2] * M[2][1] - M[0][1] * M[2][2]) * od,
                             (M[0][1] * M[1][2] - M[1][1] * M[0][2]) * od},
                            {(M[1][2] * M[2][0] - M[1][0] * M[2][2]) * od, (M[0][0] * M[2][2] - M[0][2] * M[2][0]) * od,
                             (M[0][2] * M[1][0] - M[0][0] * M[1][2]) * od},
                            {(M[1][0] * M[2][1] - M[1][1] * M[2][0]) * od, (M[0][1] * M[2][0] - M[0][0] * M[2][1]) * od,
                             (M[0][0] * M[1][1] - M[0][1] * M[1][0]) * od}};
    const double b2old = U[5] * U[5] + U[6] * U[6] + U[7] * U[7];
#pragma unroll
    for (int p = 0; p < 3; ++p) U[5 + p] = b[p][0] * r3[0] + b[p][1] * r3[1] + b[p][2] * r3[2];
    const double ob8pi = 0.5 * C<double>::ob4pi;
    U[4] = U[4] - b2old * ob8pi + (U[5] * U[5] + U[6] * U[6] + U[7] * U[7]) * ob8pi;
    a.B1_zc[z] = U[5];
    a.B2_zc[z] = U[6];
    a.B3_zc[z] = U[7];
  }
  double e = a.ent[z];
  Gas<double> gs{gam, gam - 1.0, a.rhofloor, a.prsfloor, 1.0 / (gam - 1.0)};
  V[4] = 0.0;
  cons_to_prim(gs, U, V, e);
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    P.Vr(c, 0, z) = V[c];
    P.Ur(c, 0, z) = U[c];
  }
  a.ent[z] = e;
}

constexpr int DT_THREADS = 256, DT_BLOCKS = 1024;

__global__ void k_dt_init(double dt0, double* out) { out[0] = dt0; }

// min of a positive double into *out by compare-and-swap (order-independent)
__device__ __forceinline__ void atomic_min_pos(double* out, double v) {
  unsigned long long* w = (unsigned long long*)out;
  unsigned long long cur = __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  while (v < __longlong_as_double((long long)cur)) {
    const unsigned long long seen = atomicCAS(w, cur, (unsigned long long)__double_as_longlong(v));
    if (seen == cur) break;
    cur = seen;
  }
}

__global__ void __launch_bounds__(DT_THREADS) k_dt(const P1 P, const Box own, double cfl, double dt0,
                                                   double* out) {
  __shared__ double red[DT_THREADS];
  const cs_p1_patch& a = P.a;
  const Gas<double> gs{a.gamma, a.gamma - 1.0, a.rhofloor, a.prsfloor, 1.0 / (a.gamma - 1.0)};
  double best = dt0;
  const int64_t n = own.count(), stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t t = tid(); t < n; t += stride) {
    int i, j, k;
    own.decode(t, i, j, k);
    const int64_t z = P.zi(i, j, k);
    if (a.mask[z] == 0) continue;
    double U[8], V[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) U[c] = P.Ur(c, 0, z);
    double e = a.ent[z];  // the routine keeps its entropy array as it is
    V[4] = 0.0;
    cons_to_prim(gs, U, V, e);
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      P.Ur(c, 0, z) = U[c];
      P.Vr(c, 0, z) = V[c];
    }
    const double b2 = V[5] * V[5] + V[6] * V[6] + V[7] * V[7];
    const double ms = sqrt((a.gamma_sound[z] * V[4] + C<double>::ob4pi * b2) / V[0]);
    a.msonic[z] = ms;
    const double s = sqrt(V[1] * V[1] + V[2] * V[2] + V[3] * V[3]) + ms;
    best = fmin(best, cfl * a.scale[z] / s);
  }
  red[threadIdx.x] = best;
  __syncthreads();
  for (int h = DT_THREADS / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + h]);
    __syncthreads();
  }
  if (threadIdx.x == 0) atomic_min_pos(out, red[0]);
}

static int p1_dims_check(const cs_p1_patch* p, const char* who, P1& P) {
  if (!p) {
    cs::set_error("%s: null patch", who);
    return CS_E_INVALID;
  }
  if (p->ntp < 1 || p->nr < 1 || p->nghost < 1) {
    cs::set_error("%s: need ntp, nr, nghost >= 1 (got %d, %d, %d)", who, p->ntp, p->nr, p->nghost);
    return CS_E_INVALID;
  }
  if (p->nghost >= p->ntp) {
    cs::set_error("%s: nghost (%d) must be below ntp (%d)", who, p->nghost, p->ntp);
    return CS_E_BAD_DECOMP;
  }
  if (!(p->gamma > 1.0) || !(p->rhofloor > 0.0) || !(p->prsfloor > 0.0) || !(p->smallnum > 0.0)) {
    cs::set_error("%s: gamma > 1 and positive floors and smallnum required", who);
    return CS_E_INVALID;
  }
  P.a = *p;
  P.N = p->ntp;
  P.NR = p->nr;
  P.g = p->nghost;
  P.nz = P.N + 2 * P.g;
  P.nzr = P.NR + 2 * P.g;
  P.nf = P.nz + 1;
  P.nfr = P.nzr + 1;
  P.NZ = (int64_t)P.nz * P.nz * P.nzr;
  P.NXF = (int64_t)P.nf * P.nz * P.nzr;
  P.NZF = (int64_t)P.nz * P.nz * P.nfr;
  P.NFC = (int64_t)P.nf * P.nf * P.nfr;
  return 0;
}

}  // namespace csp1

extern "C" int cs_cleanup_after_update_p1(const cs_p1_patch* p, void* stream_) {
  P1 P;
  int rc = p1_dims_check(p, "cs_cleanup_after_update_p1", P);
  if (rc) return rc;
  const void* ptrs[] = {p->mask, p->stencil, p->xf_normal, p->yf_normal, p->zf_centroid, p->scale, p->U,
                        p->ent, p->Bxf, p->Byf, p->Bzf, p->V, p->B1_zc, p->B2_zc, p->B3_zc, p->gamma_eos,
                        p->gamma_sound, p->mean_mol_wt};
  for (const void* q : ptrs)
    if (!q) {
      cs::set_error("cs_cleanup_after_update_p1: null array in the patch");
      return CS_E_INVALID;
    }
  hipStream_t st = (hipStream_t)stream_;
  k_face_zero<<<dim3(grid_of(P.NFC, 256), 3), 256, 0, st>>>(P);
  k_tricorner<<<grid_of(4 * (int64_t)P.g * P.nzr, 256), 256, 0, st>>>(P);
  k_cleanup_zone<<<grid_of(P.NZ, 256), 256, 0, st>>>(P);
  P1_HIP(hipGetLastError());
  return 0;
}

extern "C" int cs_evaluate_timestep_p1(const cs_p1_patch* p, double cfl, double dtcur, double* dt_new,
                                       void* stream_) {
  P1 P;
  int rc = p1_dims_check(p, "cs_evaluate_timestep_p1", P);
  if (rc) return rc;
  const void* ptrs[] = {p->mask, p->scale, p->U, p->ent, p->V, p->gamma_sound, p->msonic, dt_new};
  for (const void* q : ptrs)
    if (!q) {
      cs::set_error("cs_evaluate_timestep_p1: null array in the patch or null dt_new");
      return CS_E_INVALID;
    }
  if (!(cfl > 0.0) || !(dtcur > 0.0)) {
    cs::set_error("cs_evaluate_timestep_p1: cfl and dtcur must be positive (got %g, %g)", cfl, dtcur);
    return CS_E_INVALID;
  }
  hipStream_t st = (hipStream_t)stream_;
  const Box own = make_box(1, P.N, 1, P.N, 1, P.NR);
  const int nblk = (int)std::min<int64_t>(grid_of(own.count(), DT_THREADS), DT_BLOCKS);
  k_dt_init<<<1, 1, 0, st>>>(1.1 * dtcur, dt_new);
  k_dt<<<nblk, DT_THREADS, 0, st>>>(P, own, cfl, 1.1 * dtcur, dt_new);
  P1_HIP(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------------------
// MAKE_EDGE_EFIELDS_P1 (cubesphere/400riemann_integrate/460make_edge_fields_p1.f:5)
// of one patch at second order, FP64.  k_edge: one work item per entry of the
// face-B-shaped edge arrays, the family (x, y, z edges) in blockIdx.y; items
// on an edge of the owned faces compute it, all others write the zero the
// routine leaves there, so the arrays need no separate clearing.  An edge's
// field: the mean mode-1 E of its four zones along the edge's tangent, the
// face-point normal B of the four faces around it (LIMIT_ALL_SLOPES_P1) and
// their wave speeds (MAKE_FACIAL_FLUXES) symmetrised to +-max|S|, and the
// 4-state HLL of 630...f (csp::emf4_hll).  z-edges with an inactive zone keep
// zero; at a flagged tri-corner the edge where the three active zones meet
// adds the 3-state HLL of 631...f (emf3_hll below).
namespace csp1 {

struct P1E {
  cs_p1_edge a;
  int N, NR, g, nz, nf;
  int64_t NZ, NFC;
  __device__ __forceinline__ int64_t zi(int i, int j, int k) const {
    return (int64_t)(i - 1 + g) + (int64_t)nz * ((j - 1 + g) + (int64_t)nz * (k - 1 + g));
  }
  __device__ __forceinline__ int64_t fci(int i, int j, int k) const {
    return (int64_t)(i + g) + (int64_t)nf * ((j + g) + (int64_t)nf * (k + g));
  }
  __device__ __forceinline__ double vx(int i, int j, int k, int c) const { return a.vert[fci(i, j, k) + NFC * c]; }
  __device__ __forceinline__ double emean(int c, const int64_t* z, int n) const {
    const double* E = c == 0 ? a.Ex_zc : c == 1 ? a.Ey_zc : a.Ez_zc;
    return n == 4 ? (E[z[0]] + E[z[1]] + E[z[2]] + E[z[3]]) * 0.25 : (E[z[0]] + E[z[1]] + E[z[2]]) * (1.0 / 3.0);
  }
};

__global__ void __launch_bounds__(256) k_edge(const P1E P) {
  const int fam = blockIdx.y;
  const int64_t t = tid();
  if (t >= P.NFC) return;
  const cs_p1_edge& a = P.a;
  const int n = P.nf, N = P.N, NR = P.NR;
  const int i = (int)(t % n) - P.g, j = (int)((t / n) % n) - P.g, k = (int)(t / ((int64_t)n * n)) - P.g;
  double* out = fam == 0 ? a.Efx : fam == 1 ? a.Efy : a.Efz;
  const bool on = fam == 0   ? (i >= 1 && i <= N && j >= 0 && j <= N && k >= 0 && k <= NR)
                  : fam == 1 ? (i >= 0 && i <= N && j >= 1 && j <= N && k >= 0 && k <= NR)
                             : (i >= 0 && i <= N && j >= 0 && j <= N && k >= 1 && k <= NR);
  double val = 0.0;
  if (on) {
    // zones LD, RD, RU, LU around the edge, and the edge's other vertex
    int zc[4][3];
    int pi = i, pj = j, pk = k;
    if (fam == 0) {
      const int q[4][3] = {{i, j, k}, {i, j + 1, k}, {i, j + 1, k + 1}, {i, j, k + 1}};
      memcpy(zc, q, sizeof(zc));
      pi = i - 1;
    } else if (fam == 1) {
      const int q[4][3] = {{i, j, k}, {i, j, k + 1}, {i + 1, j, k + 1}, {i + 1, j, k}};
      memcpy(zc, q, sizeof(zc));
      pj = j - 1;
    } else {
      const int q[4][3] = {{i, j, k}, {i + 1, j, k}, {i + 1, j + 1, k}, {i, j + 1, k}};
      memcpy(zc, q, sizeof(zc));
      pk = k - 1;
    }
    int64_t z[4], f[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      z[q] = P.zi(zc[q][0], zc[q][1], zc[q][2]);
      f[q] = P.fci(zc[q][0], zc[q][1], zc[q][2]);
    }
    const bool active = fam < 2 || (a.mask[z[0]] && a.mask[z[1]] && a.mask[z[2]] && a.mask[z[3]]);
    double ec[3], tg[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) ec[c] = 0.5 * (P.vx(i, j, k, c) + P.vx(pi, pj, pk, c));
    if (active) {
      double ct = 0.0, st = 1.0;
      if (fam < 2) {
        // tangent: the chord, made normal to the radius at the edge centre
        // (the centre lifted to the vertices' sphere)
        const double rv = sqrt(P.vx(i, j, k, 0) * P.vx(i, j, k, 0) + P.vx(i, j, k, 1) * P.vx(i, j, k, 1) +
                               P.vx(i, j, k, 2) * P.vx(i, j, k, 2));
        const double lift = rv / sqrt(ec[0] * ec[0] + ec[1] * ec[1] + ec[2] * ec[2]);
        double r[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) r[c] = ec[c] * lift;
        const double ir = 1.0 / sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          r[c] *= ir;
          tg[c] = P.vx(i, j, k, c) - P.vx(pi, pj, pk, c);
        }
        const double d = tg[0] * r[0] + tg[1] * r[1] + tg[2] * r[2];
#pragma unroll
        for (int c = 0; c < 3; ++c) tg[c] -= d * r[c];
        const double it = 1.0 / sqrt(tg[0] * tg[0] + tg[1] * tg[1] + tg[2] * tg[2]);
#pragma unroll
        for (int c = 0; c < 3; ++c) tg[c] *= it;
      } else {
        // tangent: radial; the angle between the x and y lattice directions
        const double ir = 1.0 / sqrt(ec[0] * ec[0] + ec[1] * ec[1] + ec[2] * ec[2]);
        double n1[3], tau[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          tg[c] = ec[c] * ir;
          n1[c] = P.vx(i + 1, j, k, c) - P.vx(i - 1, j, k, c);
          tau[c] = P.vx(i, j + 1, k, c) - P.vx(i, j - 1, k, c);
        }
        const double d1 = n1[0] * tg[0] + n1[1] * tg[1] + n1[2] * tg[2];
        const double d2 = tau[0] * tg[0] + tau[1] * tg[1] + tau[2] * tg[2];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          n1[c] -= d1 * tg[c];
          tau[c] -= d2 * tg[c];
        }
        const double i1 = 1.0 / sqrt(n1[0] * n1[0] + n1[1] * n1[1] + n1[2] * n1[2]);
        const double i2 = 1.0 / sqrt(tau[0] * tau[0] + tau[1] * tau[1] + tau[2] * tau[2]);
        ct = (tau[0] * i2) * (n1[0] * i1) + (tau[1] * i2) * (n1[1] * i1) + (tau[2] * i2) * (n1[2] * i1);
        st = sqrt(1.0 - ct * ct);
      }
      const double Et = P.emean(0, z, 4) * tg[0] + P.emean(1, z, 4) * tg[1] + P.emean(2, z, 4) * tg[2];
      // face-point normal B (points 2..5 of the faces around the edge) and speeds
      const double* byq = fam == 0 ? a.Bzf_q : fam == 1 ? a.Bxf_q : a.Byf_q;
      const double* bnq = fam == 0 ? a.Byf_q : fam == 1 ? a.Bzf_q : a.Bxf_q;
      const double By_R = byq[f[1] + 4 * P.NFC], By_L = byq[f[0] + 3 * P.NFC];
      const double Bn_U = bnq[f[3] + 2 * P.NFC], Bn_D = bnq[f[0] + P.NFC];
      const double *sHi1, *sLo1, *sHi2, *sLo2;
      int a1, b1, a2, b2;  // zones whose speeds make S_R, S_L (a1, b1) and S_U, S_D (a2, b2)
      if (fam == 0) {
        sHi1 = a.S_U; sLo1 = a.S_D; sHi2 = a.S_T; sLo2 = a.S_B; a1 = 3; b1 = 0; a2 = 1; b2 = 0;
      } else if (fam == 1) {
        sHi1 = a.S_T; sLo1 = a.S_B; sHi2 = a.S_R; sLo2 = a.S_L; a1 = 3; b1 = 0; a2 = 0; b2 = 1;
      } else {
        sHi1 = a.S_R; sLo1 = a.S_L; sHi2 = a.S_U; sLo2 = a.S_D; a1 = 3; b1 = 0; a2 = 1; b2 = 0;
      }
      const double sR = fmax(sHi1[f[a1]], sHi1[f[b1]]), sL = fmin(sLo1[f[a1]], sLo1[f[b1]]);
      const double sU = fmax(sHi2[f[a2]], sHi2[f[b2]]), sD = fmin(sLo2[f[a2]], sLo2[f[b2]]);
      const double S = fmax(fmax(fabs(sR), fabs(sL)), fmax(fabs(sU), fabs(sD)));
      val = emf4_hll<double>(S, -S, S, -S, ct, st, Et, Et, Et, Et, Bn_U, Bn_D, By_R, By_L);
    }
    if (fam == 2) {
      const int corner = (i == 0 && j == 0) ? 0 : (i == N && j == 0) ? 1 : (i == N && j == N) ? 2
                                                : (i == 0 && j == N) ? 3 : -1;
      if (corner >= 0 && a.itri[corner]) {
        // the three active zones 1, 2, 3 of the corner (460...f tri-corner blocks)
        int q1[2], q2[2], q3[2];
        if (corner == 0) {
          q1[0] = i + 1; q1[1] = j + 1; q2[0] = i; q2[1] = j + 1; q3[0] = i + 1; q3[1] = j;
        } else if (corner == 1) {
          q1[0] = i + 1; q1[1] = j + 1; q2[0] = i; q2[1] = j + 1; q3[0] = i; q3[1] = j;
        } else if (corner == 2) {
          q1[0] = i; q1[1] = j + 1; q2[0] = i; q2[1] = j; q3[0] = i + 1; q3[1] = j;
        } else {
          q1[0] = i + 1; q1[1] = j + 1; q2[0] = i; q2[1] = j; q3[0] = i + 1; q3[1] = j;
        }
        const int64_t z3[3] = {P.zi(q1[0], q1[1], k), P.zi(q2[0], q2[1], k), P.zi(q3[0], q3[1], k)};
        const double ir = 1.0 / sqrt(ec[0] * ec[0] + ec[1] * ec[1] + ec[2] * ec[2]);
        const double Et = P.emean(0, z3, 3) * (ec[0] * ir) + P.emean(1, z3, 3) * (ec[1] * ir) +
                          P.emean(2, z3, 3) * (ec[2] * ir);
        const int64_t f1 = P.fci(q1[0], q1[1], k), f2 = P.fci(q2[0], q2[1], k), f3 = P.fci(q3[0], q3[1], k),
                      f2d = P.fci(q2[0], q2[1] - 1, k), f1d = P.fci(q1[0], q1[1] - 1, k);
        const int64_t NF = P.NFC;
        double bn[3], sr[3], sl[3];
        if (corner == 0) {
          bn[0] = -a.Byf_q[f3 + 4 * NF]; bn[1] = a.Bxf_q[f2 + 2 * NF]; bn[2] = a.Byf_q[f2d + 3 * NF];
          sr[0] = -a.S_D[f3]; sl[0] = -a.S_U[f3]; sr[1] = a.S_R[f2]; sl[1] = a.S_L[f2];
          sr[2] = a.S_U[f2d]; sl[2] = a.S_D[f2d];
        } else if (corner == 1) {
          bn[0] = -a.Byf_q[f1d + 4 * NF]; bn[1] = a.Bxf_q[f2 + 2 * NF]; bn[2] = a.Byf_q[f3 + 3 * NF];
          sr[0] = -a.S_L[f3]; sl[0] = -a.S_R[f3]; sr[1] = a.S_R[f2]; sl[1] = a.S_L[f2];
          sr[2] = a.S_U[f3]; sl[2] = a.S_D[f3];
        } else if (corner == 2) {
          bn[0] = a.Bxf_q[f1 + 2 * NF]; bn[1] = a.Byf_q[f2 + 3 * NF]; bn[2] = -a.Bxf_q[f2 + NF];
          sr[0] = a.S_R[f1]; sl[0] = a.S_L[f1]; sr[1] = a.S_U[f2]; sl[1] = a.S_D[f2];
          sr[2] = -a.S_L[f2]; sl[2] = -a.S_R[f2];
        } else {
          bn[0] = -a.Byf_q[f3 + 4 * NF]; bn[1] = a.Byf_q[f2 + 3 * NF]; bn[2] = -a.Bxf_q[f2 + NF];
          sr[0] = -a.S_D[f3]; sl[0] = -a.S_U[f3]; sr[1] = a.S_U[f2]; sl[1] = a.S_D[f2];
          sr[2] = -a.S_L[f2]; sl[2] = -a.S_R[f2];
        }
        const double S = fmax(fmax(fmax(fabs(sr[0]), fabs(sl[0])), fmax(fabs(sr[1]), fabs(sl[1]))),
                              fmax(fabs(sr[2]), fabs(sl[2])));
        const double Sr[3] = {S, S, S}, Sl[3] = {-S, -S, -S}, E3[3] = {Et, Et, Et};
        val += emf3_hll(Sr, Sl, E3, bn);
      }
    }
  }
  out[t] = val;
}

}  // namespace csp1

extern "C" int cs_make_edge_efields_p1(const cs_p1_edge* p, void* stream_) {
  if (!p) {
    cs::set_error("cs_make_edge_efields_p1: null argument");
    return CS_E_INVALID;
  }
  if (p->ntp < 1 || p->nr < 1 || p->nghost < 1) {
    cs::set_error("cs_make_edge_efields_p1: need ntp, nr, nghost >= 1 (got %d, %d, %d)", p->ntp, p->nr,
                  p->nghost);
    return CS_E_INVALID;
  }
  if (p->nghost >= p->ntp) {
    cs::set_error("cs_make_edge_efields_p1: nghost (%d) must be below ntp (%d)", p->nghost, p->ntp);
    return CS_E_BAD_DECOMP;
  }
  const void* ptrs[] = {p->vert, p->mask, p->Ex_zc, p->Ey_zc, p->Ez_zc, p->Bxf_q, p->Byf_q, p->Bzf_q, p->S_L,
                        p->S_R, p->S_D, p->S_U, p->S_B, p->S_T, p->Efx, p->Efy, p->Efz};
  for (const void* q : ptrs)
    if (!q) {
      cs::set_error("cs_make_edge_efields_p1: null array");
      return CS_E_INVALID;
    }
  hipStream_t st = (hipStream_t)stream_;
  P1E P;
  P.a = *p;
  P.N = p->ntp;
  P.NR = p->nr;
  P.g = p->nghost;
  P.nz = P.N + 2 * P.g;
  P.nf = P.nz + 1;
  P.NZ = (int64_t)P.nz * P.nz * (P.NR + 2 * P.g);
  P.NFC = (int64_t)P.nf * P.nf * (P.NR + 2 * P.g + 1);
  k_edge<<<dim3(grid_of(P.NFC, 256), 3), 256, 0, st>>>(P);
  P1_HIP(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------------------
// BUILD_NON_STIFF_SOURCES (cubesphere/400riemann_integrate/470build_non_stiff_sources.f:6)
// of one patch, MHD, adiabatic, no external source.  k_src: one work item per
// zone; every zone's Source is written (zero outside the routine's range), and
// on the active zones of (0..N+1)^2 x (0..nr+1) the point-mass gravity at the
// zone centroid, the centrifugal and the Coriolis terms of omega_passive and
// the energy source v . S go into Source and are added to dU_dt (:150-250).
namespace csp1 {

__global__ void __launch_bounds__(256) k_src(const cs_p1_source a, int N, int NR, int g, int64_t NZ) {
  const int64_t z = tid();
  if (z >= NZ) return;
  const int nz = N + 2 * g;
  const int i = (int)(z % nz) + 1 - g, j = (int)((z / nz) % nz) + 1 - g, k = (int)(z / ((int64_t)nz * nz)) + 1 - g;
  const bool on = i >= 0 && i <= N + 1 && j >= 0 && j <= N + 1 && k >= 0 && k <= NR + 1 && a.mask[z] == 1;
  double S[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  const double ox = a.omega[0], oy = a.omega[1], oz = a.omega[2];
  if (on && (a.mstar > a.smallnum || sqrt(ox * ox + oy * oy + oz * oz) > a.smallnum)) {
    const double x = a.v_centroid[z], y = a.v_centroid[z + NZ], w = a.v_centroid[z + 2 * NZ];
    const double rho = a.V[z], vx = a.V[z + NZ], vy = a.V[z + 2 * NZ], vz = a.V[z + 3 * NZ];
    const double rad = sqrt(x * x + y * y + w * w);
    const double gmr3 = a.newtonsc * a.mstar / (rad * rad * rad);
    const double cx = oy * w - oz * y, cy = oz * x - ox * w, cz = ox * y - oy * x;  // omega x r
    S[1] = -gmr3 * x * rho - rho * (oy * cz - oz * cy) - 2.0 * rho * (oy * vz - oz * vy);
    S[2] = -gmr3 * y * rho - rho * (oz * cx - ox * cz) - 2.0 * rho * (oz * vx - ox * vz);
    S[3] = -gmr3 * w * rho - rho * (ox * cy - oy * cx) - 2.0 * rho * (ox * vy - oy * vx);
    S[4] = vx * S[1] + vy * S[2] + vz * S[3];
  }
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    a.Source[z + NZ * c] = S[c];
    if (on) a.dU_dt[z + NZ * c] += S[c];
  }
}

}  // namespace csp1

extern "C" int cs_build_non_stiff_sources_p1(const cs_p1_source* p, void* stream_) {
  if (!p) {
    cs::set_error("cs_build_non_stiff_sources_p1: null argument");
    return CS_E_INVALID;
  }
  if (p->ntp < 1 || p->nr < 1 || p->nghost < 1) {
    cs::set_error("cs_build_non_stiff_sources_p1: need ntp, nr, nghost >= 1 (got %d, %d, %d)", p->ntp, p->nr,
                  p->nghost);
    return CS_E_INVALID;
  }
  if (!p->mask || !p->v_centroid || !p->V || !p->Source || !p->dU_dt) {
    cs::set_error("cs_build_non_stiff_sources_p1: null array");
    return CS_E_INVALID;
  }
  const int64_t nz = p->ntp + 2 * p->nghost;
  const int64_t NZ = nz * nz * (p->nr + 2 * p->nghost);
  k_src<<<grid_of(NZ, 256), 256, 0, (hipStream_t)stream_>>>(*p, p->ntp, p->nr, p->nghost, NZ);
  P1_HIP(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------------------
// The defect correction of the driver over the P1 patches of one shell
// (FILL_FLUX_SND_PATCH 599..., FETCH_BUFFERS_FILL_IN_FLUX_X/Y/Z 334-336...,
// SYNCHRONIZE_EFZ_QUAD_CORNER 340..., UNPACK_DEFECT_FLUXES 341...,
// UPDATE_DEFECT_IMEX_RK 485update_defect_imex_rk.f): the fluctuations and
// entropy fluxes of every face that two patches share take the mean of the
// two patches' values (334...f:100-116), the electric field of every shared
// edge the mean of its two or three patches' values (334...f:128-160, the
// tri-corner edges a third of three), and each patch's boundary zones get the
// difference of their fluctuations, -(mean - own) A / V, in dU_dt.  Done
// before the update instead of re-applied after it; the update is linear in
// these, so the result is the same.  The face and edge pairings, with their
// orientations, come from the host (matched geometrically once).
namespace csp1 {

struct P1Sync {
  double* Fl_Pl[6][3];  // [patch][family]
  double* Fl_Mi[6][3];
  double* EntF[6][3];
  double* Ef[6][3];
  double* dU_dt[6];
  const double* area[6][3];
  const double* vol[6];
  int N, g, nz, nf;
  int64_t NZ, NFC, NXF, NZF;
};

using FacePair = cs_p1_face_pair;
using EdgeGroup = cs_p1_edge_group;

__global__ void __launch_bounds__(256) k_sync_faces(const P1Sync S, const FacePair* fp, int64_t n) {
  const int64_t t = tid();
  if (t >= n) return;
  const FacePair q = fp[t];
  const int64_t NFC = S.NFC;
  double ePl[2][8], eMi[2][8];
  double ent[2];
  const int pp[2] = {q.pa, q.pb}, ff[2] = {q.fa, q.fb};
  const int64_t ii[2] = {q.ia, q.ib};
#pragma unroll
  for (int s = 0; s < 2; ++s) {
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      ePl[s][c] = S.Fl_Pl[pp[s]][ff[s]][ii[s] + NFC * c];
      eMi[s][c] = S.Fl_Mi[pp[s]][ff[s]][ii[s] + NFC * c];
    }
    ent[s] = S.EntF[pp[s]][ff[s]][ii[s]];
  }
  // the mean in a's orientation: with opposite normals b's plus side is a's
  // minus side and its entropy flux has the other sign
  double mPl[8], mMi[8];
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    mPl[c] = 0.5 * (ePl[0][c] + (q.flip ? eMi[1][c] : ePl[1][c]));
    mMi[c] = 0.5 * (eMi[0][c] + (q.flip ? ePl[1][c] : eMi[1][c]));
  }
  const double mEnt = 0.5 * (ent[0] + (q.flip ? -ent[1] : ent[1]));
  const int64_t zz[2] = {q.za, q.zb}, ar[2] = {q.aa, q.ab};
  const int sd[2] = {q.sa, q.sb};
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    const bool swap = s == 1 && q.flip;
    const double* nPl = swap ? mMi : mPl;
    const double* nMi = swap ? mPl : mMi;
    const double w = S.area[pp[s]][ff[s]][ar[s]] / S.vol[pp[s]][zz[s]];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const double own = sd[s] ? ePl[s][c] : eMi[s][c], mean = sd[s] ? nPl[c] : nMi[c];
      atomicAdd(&S.dU_dt[pp[s]][zz[s] + S.NZ * c], -(mean - own) * w);
      S.Fl_Pl[pp[s]][ff[s]][ii[s] + NFC * c] = nPl[c];
      S.Fl_Mi[pp[s]][ff[s]][ii[s] + NFC * c] = nMi[c];
    }
    S.EntF[pp[s]][ff[s]][ii[s]] = (s == 1 && q.flip) ? -mEnt : mEnt;
  }
}

__global__ void __launch_bounds__(256) k_sync_edges(const P1Sync S, const EdgeGroup* eg, int64_t n) {
  const int64_t t = tid();
  if (t >= n) return;
  const EdgeGroup q = eg[t];
  double sum = 0.0;
  for (int m = 0; m < q.n; ++m) sum += q.sign[m] * S.Ef[q.p[m]][q.f[m]][q.idx[m]];
  const double mean = q.n == 3 ? sum * (1.0 / 3.0) : 0.5 * sum;
  for (int m = 0; m < q.n; ++m) S.Ef[q.p[m]][q.f[m]][q.idx[m]] = q.sign[m] * mean;
}

}  // namespace csp1

extern "C" int cs_sync_shared_p1(const cs_p1_shell_sync* p, void* stream_) {
  if (!p || !p->faces || !p->edges) {
    cs::set_error("cs_sync_shared_p1: null argument");
    return CS_E_INVALID;
  }
  if (p->ntp < 1 || p->nr < 1 || p->nghost < 1 || p->nface < 0 || p->nedge < 0) {
    cs::set_error("cs_sync_shared_p1: bad sizes");
    return CS_E_INVALID;
  }
  P1Sync S;
  for (int s = 0; s < 6; ++s) {
    const cs_p1_patch_fluxes& q = p->patch[s];
    const void* ptrs[] = {q.Fluct_x_Pl, q.Fluct_y_Pl, q.Fluct_z_Pl, q.Fluct_x_Mi, q.Fluct_y_Mi, q.Fluct_z_Mi,
                          q.EntFlux_x, q.EntFlux_y, q.EntFlux_z, q.Efx, q.Efy, q.Efz, q.dU_dt,
                          q.xf_area, q.yf_area, q.zf_area, q.v_volume};
    for (const void* r : ptrs)
      if (!r) {
        cs::set_error("cs_sync_shared_p1: null array of patch %d", s);
        return CS_E_INVALID;
      }
    S.Fl_Pl[s][0] = q.Fluct_x_Pl; S.Fl_Pl[s][1] = q.Fluct_y_Pl; S.Fl_Pl[s][2] = q.Fluct_z_Pl;
    S.Fl_Mi[s][0] = q.Fluct_x_Mi; S.Fl_Mi[s][1] = q.Fluct_y_Mi; S.Fl_Mi[s][2] = q.Fluct_z_Mi;
    S.EntF[s][0] = q.EntFlux_x; S.EntF[s][1] = q.EntFlux_y; S.EntF[s][2] = q.EntFlux_z;
    S.Ef[s][0] = q.Efx; S.Ef[s][1] = q.Efy; S.Ef[s][2] = q.Efz;
    S.dU_dt[s] = q.dU_dt;
    S.area[s][0] = q.xf_area; S.area[s][1] = q.yf_area; S.area[s][2] = q.zf_area;
    S.vol[s] = q.v_volume;
  }
  S.N = p->ntp;
  S.g = p->nghost;
  S.nz = p->ntp + 2 * p->nghost;
  S.nf = S.nz + 1;
  S.NZ = (int64_t)S.nz * S.nz * (p->nr + 2 * p->nghost);
  S.NFC = (int64_t)S.nf * S.nf * (p->nr + 2 * p->nghost + 1);
  hipStream_t st = (hipStream_t)stream_;
  if (p->nface > 0)
    k_sync_faces<<<grid_of(p->nface, 256), 256, 0, st>>>(S, (const FacePair*)p->faces, p->nface);
  if (p->nedge > 0)
    k_sync_edges<<<grid_of(p->nedge, 256), 256, 0, st>>>(S, (const EdgeGroup*)p->edges, p->nedge);
  P1_HIP(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------------------
// PAD_BOUNDARY_R (800applications_cs/826pad_boundary_r.f:1) of one patch,
// outflow pads.  One work item per (column, ghost layer izz); every write
// reads only layer 1 / face 0 (inner) or layer nr / face nr (outer), and the
// faces two columns share get the same value from both, so the items are
// independent.
namespace csp1 {

__global__ void __launch_bounds__(256) k_pad_r(const cs_p1_pad P, int N, int NR, int g, int64_t NZ) {
  const int64_t nz = N + 2 * g, nf = nz + 1;
  const int64_t t = tid();
  if (t >= nz * nz * g) return;
  const int ix = (int)(t % nz) + 1 - g, iy = (int)((t / nz) % nz) + 1 - g, izz = (int)(t / (nz * nz)) + 1;
  auto zi = [&](int i, int j, int k) { return (int64_t)(i - 1 + g) + nz * ((j - 1 + g) + nz * (k - 1 + g)); };
  auto fc = [&](int i, int j, int k) { return (int64_t)(i + g) + nf * ((j + g) + nf * (k + g)); };
  auto xa = [&](int i, int j, int k) { return (int64_t)(i + g) + nf * ((j - 1 + g) + nz * (k - 1 + g)); };
  auto ya = [&](int i, int j, int k) { return (int64_t)(i - 1 + g) + nz * ((j + g) + nf * (k - 1 + g)); };
  auto za = [&](int i, int j, int k) { return (int64_t)(i - 1 + g) + nz * ((j - 1 + g) + nz * (k + g)); };
  if (P.mask[zi(ix, iy, izz)] != 1) return;
  const int64_t col = (int64_t)(ix - 1 + g) + nz * (iy - 1 + g);
  if (P.bcarr_r[col] == 2) {
    const int l = 1 - izz, r = 1;
#pragma unroll
    for (int c = 0; c < 8; ++c) P.U[zi(ix, iy, l) + NZ * c] = P.U[zi(ix, iy, r) + NZ * c];
    P.Bxf[fc(ix, iy, l)] = P.Bxf[fc(ix, iy, r)];
    P.Bxf[fc(ix - 1, iy, l)] = P.Bxf[fc(ix - 1, iy, r)];
    P.Byf[fc(ix, iy, l)] = P.Byf[fc(ix, iy, r)];
    P.Byf[fc(ix, iy - 1, l)] = P.Byf[fc(ix, iy - 1, r)];
    P.Bzf[fc(ix, iy, l - 1)] = P.Bzf[fc(ix, iy, r - 1)];
  }
  if (P.bcarr_r[col + nz * nz] == 2) {
    const int l = NR + izz, r = NR;
    const double vr = P.v_volume[zi(ix, iy, r)], vl = P.v_volume[zi(ix, iy, l)];
#pragma unroll
    for (int c = 0; c < 8; ++c) P.U[zi(ix, iy, l) + NZ * c] = P.U[zi(ix, iy, r) + NZ * c] * vr / vl;
    P.Bzf[fc(ix, iy, l)] = P.Bzf[fc(ix, iy, r)] * P.zf_area[za(ix, iy, r)] / P.zf_area[za(ix, iy, l)];
    P.Bxf[fc(ix, iy, l)] = P.Bxf[fc(ix, iy, r)] * P.xf_area[xa(ix, iy, r)] / P.xf_area[xa(ix, iy, l)];
    P.Byf[fc(ix, iy, l)] = P.Byf[fc(ix, iy, r)] * P.yf_area[ya(ix, iy, r)] / P.yf_area[ya(ix, iy, l)];
    P.Bxf[fc(ix - 1, iy, l)] =
        P.Bxf[fc(ix - 1, iy, r)] * P.xf_area[xa(ix - 1, iy, r)] / P.xf_area[xa(ix - 1, iy, l)];
    P.Byf[fc(ix, iy - 1, l)] =
        P.Byf[fc(ix, iy - 1, r)] * P.yf_area[ya(ix, iy - 1, r)] / P.yf_area[ya(ix, iy - 1, l)];
  }
}

}  // namespace csp1

extern "C" int cs_pad_boundary_r_p1(const cs_p1_pad* p, void* stream_) {
  if (!p) {
    cs::set_error("cs_pad_boundary_r_p1: null argument");
    return CS_E_INVALID;
  }
  if (p->ntp < 1 || p->nr < 1 || p->nghost < 1) {
    cs::set_error("cs_pad_boundary_r_p1: need ntp, nr, nghost >= 1 (got %d, %d, %d)", p->ntp, p->nr, p->nghost);
    return CS_E_INVALID;
  }
  if (!p->mask || !p->bcarr_r || !p->xf_area || !p->yf_area || !p->zf_area || !p->v_volume || !p->U || !p->Bxf ||
      !p->Byf || !p->Bzf) {
    cs::set_error("cs_pad_boundary_r_p1: null array");
    return CS_E_INVALID;
  }
  const int64_t nz = p->ntp + 2 * p->nghost;
  const int64_t NZ = nz * nz * (p->nr + 2 * p->nghost);
  csp1::k_pad_r<<<grid_of(nz * nz * p->nghost, 256), 256, 0, (hipStream_t)stream_>>>(*p, p->ntp, p->nr,
                                                                                       p->nghost, NZ);
  P1_HIP(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------------------
// Third order: LIMIT_ALL_SLOPES_P12 (421limit_all_slopes_p12.f:9) on one patch
// whose U and V hold ten modes.  Steps 2-7.a are the P1 kernels above (they
// address modes 1-4, which a ten-mode array lays out exactly as a four-mode
// one); k_zero_hi clears modes 5-10, k_weno_ao is step 7.b, k_pos<10> 7.c,
// k_face_pts<10> 8.a-c, k_clsq12 8.d, k_e12 9.a, k_weno9 on E 9.b,
// k_ao_order / k_ao_e_pre / k_weno_ao_e 9.c, and the P1 diagnostics step 10.
namespace csp1 {

__global__ void __launch_bounds__(256) k_zero_hi(const P1 P, const Box b) {
  int i, j, k;
  if (!b.decode(tid(), i, j, k)) return;
  const int64_t z = P.zi(i, j, k);
#pragma unroll
  for (int m = 4; m < 10; ++m)
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      P.Ur(c, m, z) = 0.0;
      P.Vr(c, m, z) = 0.0;
    }
}

// the 18 zones of the central quadratic stencil (421...f:3888-4350, in that
// order): stencil direction (R L U D, then RU LU LD RD as 4-7) and k offset
__constant__ int c_ao_dir[18] = {0, 1, 2, 3, -1, -1, 4, 5, 6, 7, 0, 1, 0, 1, 2, 3, 2, 3};
__constant__ int c_ao_dk[18] = {0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};

// packed lower triangle of a symmetric 9 x 9 matrix
__device__ __forceinline__ constexpr int lt(int r, int c) { return r * (r + 1) / 2 + c; }

// Step 7.b (421...f:3811-4557), one thread per zone: the 18 x 9 rows of
// RECONSTRUCTION_COEFS (251reconstruction_coefs.f:3) — the average over the
// stencil zone of this zone's quadratic basis, from the centroid offset, the
// two scale lengths and the ten basis moments — give the normal equations
// (A^T A) X = A^T R; A^T A is factored once (Cholesky, in registers) and
// solved for the eight primitives, four at a time with the rows formed
// again for the second four.  The fit is hybridised with the P1 gradients
// (MR-WENO linear weights 0.95 / 0.05 for smooth flow, 0.6 / 0.4 where the
// flattener is 1; smoothness indicators of the two polynomials; weights
// (1 + tau / IS)), and every mode above the mean is multiplied by
// (1 - flattener).  The reference inverts A^T A by LU (MATINV); the two
// agree to rounding.  Two waves per SIMD, the registers held to 256 with a
// small spill.
__global__ void __attribute__((amdgpu_waves_per_eu(2))) __launch_bounds__(128)
    k_weno_ao(const P1 P, const Box b, const double* __restrict__ mom, const int64_t* list, int64_t nlist) {
  int i, j, k;
  if (list) {  // the positivity replay's zones: those inside the range
    const int64_t t = tid();
    if (t >= nlist) return;
    zone_of(P, list[t], i, j, k);
    if (i < b.i0 || i >= b.i0 + b.ni || j < b.j0 || j >= b.j0 + b.nj || k < b.k0 || k >= b.k0 + b.nk) return;
  } else if (!b.decode(tid(), i, j, k)) {
    return;
  }
  if (!P.active(i, j, 1)) return;
  const cs_p1_patch& a = P.a;
  const int64_t z = P.zi(i, j, k), NZ = P.NZ;
  const double sm = a.smallnum, ell0 = fmax(a.scale[z], sm), obl0 = 1.0 / ell0;
  double c0[6], cz[3];
#pragma unroll
  for (int m = 0; m < 6; ++m) c0[m] = mom[z + NZ * (4 + m)];
#pragma unroll
  for (int c = 0; c < 3; ++c) cz[c] = a.v_centroid[z + NZ * c];
  int ni[8], nj[8];
  {
    const int dirs[8] = {0, 1, 2, 3, 8, 9, 10, 11};
#pragma unroll
    for (int q = 0; q < 8; ++q) P.nbr(i, j, dirs[q], ni[q], nj[q]);
  }
  const double f = a.flatten[z];
  const double betalo = f + (1.0 - f) * 2.0;
  const double gammahi = f * 0.6 + (1.0 - f) * 0.95, gammalo = 1.0 - gammahi;
  const double obgammahi = 1.0 / gammahi, gammaratio = gammalo / gammahi;
  const double ob12 = 1.0 / 12.0, ob180 = 1.0 / 180.0, ob144 = 1.0 / 144.0;
  const double m5 = c0[0] - ob12, m6 = c0[1] - ob12, m7 = c0[2] - ob12;
  const double q4 = m5 * m5 + ob180, q5 = m6 * m6 + ob180, q6 = m7 * m7 + ob180;
  const double q7 = c0[3] * c0[3] + ob144, q8 = c0[4] * c0[4] + ob144, q9 = c0[5] * c0[5] + ob144;
  double L[45];
#pragma unroll
  for (int e = 0; e < 45; ++e) L[e] = 0.0;
#pragma unroll 1
  for (int grp = 0; grp < 2; ++grp) {
    double Y[4][9];
    double v0[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      v0[c] = P.V(4 * grp + c, 0, z);
#pragma unroll
      for (int m = 0; m < 9; ++m) Y[c][m] = 0.0;
    }
#pragma unroll 1
    for (int r = 0; r < 18; ++r) {
      const int dir = c_ao_dir[r];
      int ii = i, jj = j;
      if (dir >= 0) {
        // selects, not an indexed read of ni / nj (which would go to scratch)
#pragma unroll
        for (int q = 0; q < 8; ++q)
          if (q == dir) {
            ii = ni[q];
            jj = nj[q];
          }
      }
      const int64_t zr = P.zi(ii, jj, k + c_ao_dk[r]);
      double x[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) x[c] = (a.v_centroid[zr + NZ * c] - cz[c]) * obl0;
      const double rho = fmax(a.scale[zr], sm) / ell0, r2 = rho * rho;
      double row[9];
      row[0] = x[0];
      row[1] = x[1];
      row[2] = x[2];
      row[3] = x[0] * x[0] - c0[0] + mom[zr + NZ * 4] * r2;
      row[4] = x[1] * x[1] - c0[1] + mom[zr + NZ * 5] * r2;
      row[5] = x[2] * x[2] - c0[2] + mom[zr + NZ * 6] * r2;
      row[6] = x[0] * x[1] - c0[3] + mom[zr + NZ * 7] * r2;
      row[7] = x[1] * x[2] - c0[4] + mom[zr + NZ * 8] * r2;
      row[8] = x[0] * x[2] - c0[5] + mom[zr + NZ * 9] * r2;
      if (grp == 0) {
#pragma unroll
        for (int p = 0; p < 9; ++p)
#pragma unroll
          for (int q = 0; q <= p; ++q) L[lt(p, q)] += row[p] * row[q];
      }
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const double R = P.V(4 * grp + c, 0, zr) - v0[c];
#pragma unroll
        for (int m = 0; m < 9; ++m) Y[c][m] += row[m] * R;
      }
    }
    if (grp == 0) {  // Cholesky factor of A^T A in place
#pragma unroll
      for (int p = 0; p < 9; ++p) {
        double d = L[lt(p, p)];
#pragma unroll
        for (int q = 0; q < p; ++q) d -= L[lt(p, q)] * L[lt(p, q)];
        d = sqrt(fmax(d, 1.0e-300));
        L[lt(p, p)] = d;
        const double od = 1.0 / d;
#pragma unroll
        for (int r = p + 1; r < 9; ++r) {
          double t = L[lt(r, p)];
#pragma unroll
          for (int q = 0; q < p; ++q) t -= L[lt(r, q)] * L[lt(p, q)];
          L[lt(r, p)] = t * od;
        }
      }
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      double X[9];
#pragma unroll
      for (int p = 0; p < 9; ++p) {  // L y = Y
        double t = Y[c][p];
#pragma unroll
        for (int q = 0; q < p; ++q) t -= L[lt(p, q)] * X[q];
        X[p] = t / L[lt(p, p)];
      }
#pragma unroll
      for (int p = 8; p >= 0; --p) {  // L^T x = y
        double t = X[p];
#pragma unroll
        for (int q = p + 1; q < 9; ++q) t -= L[lt(q, p)] * X[q];
        X[p] = t / L[lt(p, p)];
      }
      const int comp = 4 * grp + c;
      const double g1 = P.V(comp, 1, z), g2 = P.V(comp, 2, z), g3 = P.V(comp, 3, z);
      const double ISlo = betalo * (g1 * g1 + g2 * g2 + g3 * g3) * ob12;
      double Phi[9];
      Phi[0] = X[0] * obgammahi - g1 * gammaratio;
      Phi[1] = X[1] * obgammahi - g2 * gammaratio;
      Phi[2] = X[2] * obgammahi - g3 * gammaratio;
#pragma unroll
      for (int m = 3; m < 9; ++m) Phi[m] = X[m] * obgammahi;
      const double IShi = (Phi[0] * Phi[0] + Phi[1] * Phi[1] + Phi[2] * Phi[2]) * ob12 + Phi[3] * Phi[3] * q4 +
                          Phi[4] * Phi[4] * q5 + Phi[5] * Phi[5] * q6 + Phi[6] * Phi[6] * q7 +
                          Phi[7] * Phi[7] * q8 + Phi[8] * Phi[8] * q9;
      const double tau = fabs(IShi - ISlo);
      double wlo = gammalo * (1.0 + tau / (ISlo + sm)), whi = gammahi * (1.0 + tau / (IShi + sm));
      const double aa = 1.0 / (wlo + whi);
      wlo *= aa;
      whi *= aa;
      const double x1 = 1.0 - f;
      P.Vr(comp, 1, z) = (whi * Phi[0] + wlo * g1) * x1;
      P.Vr(comp, 2, z) = (whi * Phi[1] + wlo * g2) * x1;
      P.Vr(comp, 3, z) = (whi * Phi[2] + wlo * g3) * x1;
#pragma unroll
      for (int m = 3; m < 9; ++m) P.Vr(comp, m + 1, z) = whi * Phi[m] * x1;
    }
  }
}

// Step 8.d at third order (421...f:5574-6680), one wave per zone: the 36 x 36
// system of the normal equations of the 64 least-squares rows (n . B at the
// 30 face points against their face-point B, the 30 WENO-AO modes, the
// divergence and its gradient) and the 6 face-mean constraints.  Lanes 0-29
// put one face point each (normal, the ten basis values about the zone's
// moments, B-field weight, face-point B) in LDS; lane r < 36 forms row r of
// the augmented system in registers (A^T A = P^T P + I + D^T D, P the point
// rows, D the divergence rows; P^T P by FP64 MFMA across the wave); Gaussian
// elimination across the lanes in natural order, then back substitution.  The
// reference solves the same system by Numerical Recipes' LU with partial
// pivoting (MATSOLVE): the two
// agree to rounding.
// lane l's value of v, uniform across the wave (l uniform)
__device__ __forceinline__ double rdlane(double v, int l) {
  const long long b = __double_as_longlong(v);
  const unsigned int lo = (unsigned int)__builtin_amdgcn_readlane((int)(unsigned int)b, l);
  const unsigned int hi = (unsigned int)__builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

__constant__ double c_div_rows[4][30] = {
    {0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0},
    {0, 0, 0, 0, 2, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1},
    {0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 2, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0},
    {0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 2, 0, 0, 0}};

__global__ void __launch_bounds__(64) k_clsq12(const P1 P, const Box b, const double* __restrict__ mom) {
  // one LDS block: the face points' words first (pt), then the Gram (gram),
  // which is written after the last read of pt; 8.4 KB per zone instead of
  // 12.3, so more zones fit a CU than its registers allow anyway
  __shared__ double sh[32 * 33];
  double(*pt)[16] = reinterpret_cast<double(*)[16]>(sh);  // n (3), basis (10), weight, face-point B
  const int lane = threadIdx.x;
  int i, j, k;
  if (!b.decode(blockIdx.x, i, j, k)) return;  // one zone per workgroup: uniform
  const cs_p1_patch& a = P.a;
  const int64_t z = P.zi(i, j, k), NZ = P.NZ;
  if (a.mask[z] == 0) return;
  const double obl = P.obl(z);
  double c0[10];
#pragma unroll
  for (int m = 0; m < 10; ++m) c0[m] = mom[z + NZ * m];
  if (lane < 30) {
    const int f = lane / 5, q = lane % 5, fam = f >> 1;
    const bool hi = (f & 1) == 0;
    const int fi = i - (fam == 0 && !hi), fj = j - (fam == 1 && !hi), fk = k - (fam == 2 && !hi);
    const int64_t fg = fam == 0 ? P.xfi(fi, fj, fk) : (fam == 1 ? P.yfi(fi, fj, fk) : P.zfi(fi, fj, fk));
    const int64_t NF = fam == 2 ? P.NZF : P.NXF, fb = P.fci(fi, fj, fk);
    const double* nrm = fam == 0 ? a.xb_normal : (fam == 1 ? a.yb_normal : a.zb_normal);
    const double* pts = fam == 0 ? a.xb_pts : (fam == 1 ? a.yb_pts : a.zb_pts);
    const double* wts = fam == 0 ? a.xb_wts : (fam == 1 ? a.yb_wts : a.zb_wts);
    const double* Bq = fam == 0 ? a.Bxf_q : (fam == 1 ? a.Byf_q : a.Bzf_q);
    double r[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      pt[lane][c] = fam == 2 ? nrm[fg + NF * (c + 3 * q)] : nrm[fg + NF * c];
      r[c] = (pts[fg + NF * (c + 3 * q)] - P.cen(z, c)) * obl;
    }
    pt[lane][3] = 1.0;
    pt[lane][4] = r[0];
    pt[lane][5] = r[1];
    pt[lane][6] = r[2];
    pt[lane][7] = r[0] * r[0] - c0[4];
    pt[lane][8] = r[1] * r[1] - c0[5];
    pt[lane][9] = r[2] * r[2] - c0[6];
    pt[lane][10] = r[0] * r[1] - c0[7];
    pt[lane][11] = r[1] * r[2] - c0[8];
    pt[lane][12] = r[0] * r[2] - c0[9];
    pt[lane][13] = wts[fg + NF * q];
    pt[lane][14] = Bq[fb + P.NFC * q];
  }
  __syncthreads();
  // row `lane` of the augmented system [K | rhs]
  double row[37];
#pragma unroll
  for (int e = 0; e < 37; ++e) row[e] = 0.0;
  if (lane < 30) {
    const int c = lane / 10, m = lane % 10;
#pragma unroll
    for (int e = 0; e < 30; ++e) {
      double t = e == lane ? 1.0 : 0.0;
#pragma unroll
      for (int d = 0; d < 4; ++d) t += c_div_rows[d][lane] * c_div_rows[d][e];
      row[e] = t;
    }
    row[36] = P.V(5 + c, m, z);
#pragma unroll
    for (int f = 0; f < 6; ++f) {
      double cf = 0.0;  // the constraint column of face f
#pragma unroll
      for (int q = 5 * f; q < 5 * f + 5; ++q) {
        const double pa = pt[q][c] * pt[q][3 + m];
        row[36] += pa * pt[q][14];
        cf += pt[q][13] * pa;
      }
      row[30 + f] = cf;
    }
  } else if (lane < 36) {
    const int f = lane - 30;
#pragma unroll
    for (int q5 = 0; q5 < 5; ++q5) {
      const int q = 5 * f + q5;
      const double w = pt[q][13];
#pragma unroll
      for (int cc = 0; cc < 3; ++cc) {
        const double wn = w * pt[q][cc];
#pragma unroll
        for (int mm = 0; mm < 10; ++mm) row[10 * cc + mm] += wn * pt[q][3 + mm];
      }
      row[36] += w * pt[q][14];
    }
  }
  // P^T P, the Gram of the 30 point rows P[q][10 c + m] = n_c(q) basis_m(q),
  // by FP64 MFMA 16 x 16 x 4 over the zero-padded 32 x 32: lane l brings
  // P[4 kc + l / 16][16 I + l % 16] as both operands (A = P^T by columns, B = P
  // by rows), and result register r of lane l is entry (l / 16 + 4 r, l % 16)
  // of its 16 x 16 block.  Blocks (0, 0), (0, 1) and (1, 1) are formed, (1, 0)
  // is (0, 1) transposed; the sum goes through LDS to the row lanes.
  double(*gram)[33] = reinterpret_cast<double(*)[33]>(sh);
  {
    typedef double d4 __attribute__((ext_vector_type(4)));
    d4 g00 = {0.0, 0.0, 0.0, 0.0}, g01 = g00, g11 = g00;
    const int l16 = lane % 16, lq = lane / 16;
#pragma unroll
    for (int kc = 0; kc < 8; ++kc) {
      const int q = 4 * kc + lq;
      double p0 = 0.0, p1 = 0.0;
      if (q < 30) {
        p0 = pt[q][l16 / 10] * pt[q][3 + l16 % 10];
        if (16 + l16 < 30) p1 = pt[q][(16 + l16) / 10] * pt[q][3 + (16 + l16) % 10];
      }
      g00 = __builtin_amdgcn_mfma_f64_16x16x4f64(p0, p0, g00, 0, 0, 0);
      g01 = __builtin_amdgcn_mfma_f64_16x16x4f64(p0, p1, g01, 0, 0, 0);
      g11 = __builtin_amdgcn_mfma_f64_16x16x4f64(p1, p1, g11, 0, 0, 0);
    }
    __syncthreads();  // pt is read no more: the Gram takes its place
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int gr = lq + 4 * r;
      gram[gr][l16] = g00[r];
      gram[gr][16 + l16] = g01[r];
      gram[16 + l16][gr] = g01[r];
      gram[16 + gr][16 + l16] = g11[r];
    }
  }
  __syncthreads();
  if (lane < 30) {
#pragma unroll
    for (int e = 0; e < 30; ++e) row[e] += gram[lane][e];
  }
  // Elimination in natural order, lane jj holding pivot row jj: the leading
  // 30 x 30 block is I + D^T D + P^T P, symmetric positive definite with
  // every eigenvalue at least 1, so its pivots are too; the trailing 6 x 6
  // block then holds -C K^-1 C^T, negative definite for the six independent
  // face-mean rows C.  No pivot search, and the pivot row comes from a known
  // lane, broadcast through scalar registers.
#pragma unroll
  for (int jj = 0; jj < 36; ++jj) {
    // the whole pivot row first, then the updates: the scalar registers the
    // broadcasts write are read well after, without wait states between
    double pr[37];
#pragma unroll
    for (int e = jj; e < 37; ++e) pr[e] = rdlane(row[e], jj);
    const double fct = row[jj] / pr[jj];
    if (lane > jj) {
#pragma unroll
      for (int e = jj + 1; e < 37; ++e) row[e] -= fct * pr[e];
    }
  }
  // back substitution by columns: x[jj] from lane jj, then every lane above
  // takes it out of its right-hand side
  double x[36];
#pragma unroll
  for (int jj = 35; jj >= 0; --jj) {
    x[jj] = rdlane(row[36] / row[jj], jj);
    if (lane < jj) row[36] -= row[jj] * x[jj];
  }
  if (lane < 30) {
    const int c = lane / 10, m = lane % 10;
    double* Bz = c == 0 ? a.B1_zc : (c == 1 ? a.B2_zc : a.B3_zc);
    double mine = x[0];
#pragma unroll
    for (int e = 1; e < 30; ++e)
      if (e == lane) mine = x[e];
    Bz[z + NZ * m] = mine;
    P.Vr(5 + c, m, z) = mine;
  }
  if (lane == 0) {
    const double ob8pi = 1.0 / (8.0 * M_PI);
    const double aa = (P.Ur(5, 0, z) * P.Ur(5, 0, z) + P.Ur(6, 0, z) * P.Ur(6, 0, z) + P.Ur(7, 0, z) * P.Ur(7, 0, z)) * ob8pi;
    P.Ur(5, 0, z) = x[0];
    P.Ur(6, 0, z) = x[10];
    P.Ur(7, 0, z) = x[20];
    const double bb = (x[0] * x[0] + x[10] * x[10] + x[20] * x[20]) * ob8pi;
    P.Ur(4, 0, z) = P.Ur(4, 0, z) - aa + bb;
  }
}

// Step 9.a at third order (421...f:6684-6795): E = -v x B at the zone's
// centroid, v and B the quadratic fields there (the mean minus the quadratic
// modes times the moments), every zone.
__global__ void __launch_bounds__(256) k_e12(const P1 P, const Box b, const double* __restrict__ mom) {
  int i, j, k;
  if (!b.decode(tid(), i, j, k)) return;
  const int64_t z = P.zi(i, j, k), NZ = P.NZ;
  double q[6];
#pragma unroll
  for (int m = 0; m < 6; ++m) q[m] = mom[z + NZ * (4 + m)];
  auto at = [&](int c) {
    double v = P.V(c, 0, z);
#pragma unroll
    for (int m = 0; m < 6; ++m) v = v - P.V(c, 4 + m, z) * q[m];
    return v;
  };
  const double Bx = at(5), By = at(6), Bz = at(7), vx = at(1), vy = at(2), vz = at(3);
  const cs_p1_patch& a = P.a;
  a.Ex_zc[z] = -(vy * Bz - vz * By);
  a.Ey_zc[z] = -(vz * Bx - vx * Bz);
  a.Ez_zc[z] = -(vx * By - vy * Bx);
}

// The reference's step 9.c overwrites each zone's point value of E
// (421...f:9030) while later zones of the same loop read it as a stencil
// value (:8899-8935): its serial build gives a zone the new point values of
// the active box zones before it in (k, j, i) order and the old ones of those
// after it.  Level (i+1) + 3 (j+1) + 9 (k+1) reproduces that order when every
// stencil entry lies within two zones in i and j: a neighbour in the plane
// differs by at most 2 in i, so a row step of 3 outweighs it, and a neighbour
// above or below by at most 2 + 3 * 2 = 8 within the plane, which a plane step
// of 9 outweighs.  One thread per (i, j) of the box sets *bad when a pair of
// zones of the box breaks the order.
__global__ void __launch_bounds__(256) k_ao_order(const P1 P, const Box b, unsigned int* bad) {
  const int64_t t = tid();
  if (t >= (int64_t)b.ni * b.nj) return;
  const int i = b.i0 + (int)(t % b.ni), j = b.j0 + (int)(t / b.ni);
  if (!P.active(i, j, 1)) return;
  int ni[8], nj[8];
  {
    const int dirs[8] = {0, 1, 2, 3, 8, 9, 10, 11};
#pragma unroll
    for (int q = 0; q < 8; ++q) P.nbr(i, j, dirs[q], ni[q], nj[q]);
  }
  bool ok = true;
#pragma unroll
  for (int r = 0; r < 18; ++r) {
    const int dir = c_ao_dir[r], dk = c_ao_dk[r];
    int ii = i, jj = j;
#pragma unroll
    for (int q = 0; q < 8; ++q)
      if (q == dir) {
        ii = ni[q];
        jj = nj[q];
      }
    if (ii < b.i0 || ii >= b.i0 + b.ni || jj < b.j0 || jj >= b.j0 + b.nj || !P.active(ii, jj, 1)) continue;
    const int di = ii - i, dj = jj - j;
    const bool earlier = dk < 0 || (dk == 0 && (dj < 0 || (dj == 0 && di < 0)));
    const int dl = di + 3 * dj + 9 * dk;
    if (dk == 0 && dj == 0 && di == 0) continue;
    if (earlier ? dl >= 0 : dl <= 0) ok = false;
  }
  if (!ok) atomicOr(bad, 1u);
}

// Step 9.c's fit (421...f:8406-8936) depends on the stencil's point values
// of E only through its right-hand side: the ten modes are X = G R, R the 19
// point values, G = (A^T A)^-1 A^T, A the 19 x 10 matrix of
// INTERPOLATION_COEFS (252interpolation_coefs.f:5: this zone's basis at each
// stencil zone's centroid, row 0 the zone itself).  k_ao_e_pre forms G for
// every active zone of the box, one thread per zone, with a Cholesky factor
// of A^T A, and stores it with the 18 stencil zones' indices as one record of
// kAoRec doubles per zone (G row-major by stencil row, then the indices as
// int32), so that the ordered sweep reads one contiguous record per zone.

__global__ void __launch_bounds__(128) k_ao_e_pre(const P1 P, const Box b, const double* __restrict__ mom,
                                                  double* __restrict__ rec) {
  const int64_t t = tid();
  int i, j, k;
  if (!b.decode(t, i, j, k)) return;
  const cs_p1_patch& a = P.a;
  const int64_t z = P.zi(i, j, k), NZ = P.NZ;
  if (!P.active(i, j, 1)) {  // the sweep loads an inactive zone's stencil too, and discards it
    int* idx = (int*)(rec + t * kAoRec + 190);
#pragma unroll
    for (int r = 0; r < 18; ++r) idx[r] = (int)z;
    return;
  }
  const double obl0 = 1.0 / fmax(a.scale[z], a.smallnum);
  double c0[6], cz[3];
#pragma unroll
  for (int m = 0; m < 6; ++m) c0[m] = mom[z + NZ * (4 + m)];
#pragma unroll
  for (int c = 0; c < 3; ++c) cz[c] = a.v_centroid[z + NZ * c];
  int ni[8], nj[8];
  {
    const int dirs[8] = {0, 1, 2, 3, 8, 9, 10, 11};
#pragma unroll
    for (int q = 0; q < 8; ++q) P.nbr(i, j, dirs[q], ni[q], nj[q]);
  }
  double* G = rec + t * kAoRec;
  int* idx = (int*)(G + 190);
  // row r of A: r = 0 the zone itself, r = 1..18 the stencil zones
  auto arow = [&](int r, double* row) {
    if (r == 0) {
      row[0] = 1.0;
      row[1] = row[2] = row[3] = 0.0;
#pragma unroll
      for (int m = 0; m < 6; ++m) row[4 + m] = -c0[m];
      return;
    }
    const int dir = c_ao_dir[r - 1];
    int ii = i, jj = j;
    if (dir >= 0) {
#pragma unroll
      for (int q = 0; q < 8; ++q)
        if (q == dir) {
          ii = ni[q];
          jj = nj[q];
        }
    }
    const int64_t zr = P.zi(ii, jj, k + c_ao_dk[r - 1]);
    double x[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) x[c] = (a.v_centroid[zr + NZ * c] - cz[c]) * obl0;
    row[0] = 1.0;
    row[1] = x[0];
    row[2] = x[1];
    row[3] = x[2];
    row[4] = x[0] * x[0] - c0[0];
    row[5] = x[1] * x[1] - c0[1];
    row[6] = x[2] * x[2] - c0[2];
    row[7] = x[0] * x[1] - c0[3];
    row[8] = x[1] * x[2] - c0[4];
    row[9] = x[0] * x[2] - c0[5];
    idx[r - 1] = (int)zr;
  };
  double L[55];
#pragma unroll
  for (int e = 0; e < 55; ++e) L[e] = 0.0;
#pragma unroll 1
  for (int r = 0; r < 19; ++r) {
    double row[10];
    arow(r, row);
#pragma unroll
    for (int p = 0; p < 10; ++p)
#pragma unroll
      for (int q = 0; q <= p; ++q) L[p * (p + 1) / 2 + q] += row[p] * row[q];
  }
  double od[10];
#pragma unroll
  for (int p = 0; p < 10; ++p) {  // Cholesky factor in place, reciprocal diagonal in od
    double d = L[p * (p + 1) / 2 + p];
#pragma unroll
    for (int q = 0; q < p; ++q) d -= L[p * (p + 1) / 2 + q] * L[p * (p + 1) / 2 + q];
    d = sqrt(fmax(d, 1.0e-300));
    L[p * (p + 1) / 2 + p] = d;
    od[p] = 1.0 / d;
#pragma unroll
    for (int r = p + 1; r < 10; ++r) {
      double t2 = L[r * (r + 1) / 2 + p];
#pragma unroll
      for (int q = 0; q < p; ++q) t2 -= L[r * (r + 1) / 2 + q] * L[p * (p + 1) / 2 + q];
      L[r * (r + 1) / 2 + p] = t2 * od[p];
    }
  }
#pragma unroll 1
  for (int r = 0; r < 19; ++r) {  // column r of G: (A^T A)^-1 times row r
    double X[10];
    arow(r, X);
#pragma unroll
    for (int p = 0; p < 10; ++p) {
      double t2 = X[p];
#pragma unroll
      for (int q = 0; q < p; ++q) t2 -= L[p * (p + 1) / 2 + q] * X[q];
      X[p] = t2 * od[p];
    }
#pragma unroll
    for (int p = 9; p >= 0; --p) {
      double t2 = X[p];
#pragma unroll
      for (int q = p + 1; q < 10; ++q) t2 -= L[q * (q + 1) / 2 + p] * X[q];
      X[p] = t2 * od[p];
    }
#pragma unroll
    for (int m = 0; m < 10; ++m) G[10 * r + m] = X[m];
  }
}

// Step 9.c at third order (421...f:8406-9353) for component c of E at box
// zone (ir, jr, kr): the ten modes of the quadratic interpolant X = G R from
// the zone's record (k_ao_e_pre) and the point values of E at the zone and
// its 18 stencil zones; the hybridisation with the WENO gradients of 9.b as
// in 7.b, then the point value E_pt f + E (1 - f) and the modes times (1 - f).
__device__ __forceinline__ void ao_e_zone(const P1& P, const Box& b, const double* __restrict__ mom,
                                          const double* __restrict__ rec, int ir, int jr, int kr, int c) {
  const int i = b.i0 + ir, j = b.j0 + jr, k = b.k0 + kr;
  const cs_p1_patch& a = P.a;
  // the loads go out before the mask is known: an inactive zone's record
  // points at the zone itself, and only an active zone writes
  const int64_t z = P.zi(i, j, k), NZ = P.NZ;
  const bool act = P.active(i, j, 1);
  const double* G = rec + ((int64_t)ir + (int64_t)b.ni * (jr + (int64_t)b.nj * kr)) * kAoRec;
  const int* idx = (const int*)(G + 190);
  double* E = c == 0 ? a.Ex_zc : (c == 1 ? a.Ey_zc : a.Ez_zc);
  double X[10];
  const double Ept = E[z];
#pragma unroll
  for (int m = 0; m < 10; ++m) X[m] = G[m] * Ept;
#pragma unroll
  for (int r = 1; r < 19; ++r) {
    const double R = E[idx[r - 1]];
#pragma unroll
    for (int m = 0; m < 10; ++m) X[m] += G[10 * r + m] * R;
  }
  double c0[6];
#pragma unroll
  for (int m = 0; m < 6; ++m) c0[m] = mom[z + NZ * (4 + m)];
  const double sm = a.smallnum, f = a.flatten[z];
  const double betalo = f + (1.0 - f) * 2.0;
  const double gammahi = f * 0.6 + (1.0 - f) * 0.95, gammalo = 1.0 - gammahi;
  const double obgammahi = 1.0 / gammahi, gammaratio = gammalo / gammahi;
  const double ob12 = 1.0 / 12.0, ob180 = 1.0 / 180.0, ob144 = 1.0 / 144.0;
  const double m5 = c0[0] - ob12, m6 = c0[1] - ob12, m7 = c0[2] - ob12;
  double e[4];
#pragma unroll
  for (int m = 0; m < 4; ++m) e[m] = E[z + NZ * m];
  const double ISlo = betalo * (e[1] * e[1] + e[2] * e[2] + e[3] * e[3]) * ob12;
  double Phi[10];
#pragma unroll
  for (int m = 0; m < 4; ++m) Phi[m] = X[m] * obgammahi - e[m] * gammaratio;
#pragma unroll
  for (int m = 4; m < 10; ++m) Phi[m] = X[m] * obgammahi;
  const double IShi = (Phi[1] * Phi[1] + Phi[2] * Phi[2] + Phi[3] * Phi[3]) * ob12 +
                      Phi[4] * Phi[4] * (m5 * m5 + ob180) + Phi[5] * Phi[5] * (m6 * m6 + ob180) +
                      Phi[6] * Phi[6] * (m7 * m7 + ob180) + Phi[7] * Phi[7] * (c0[3] * c0[3] + ob144) +
                      Phi[8] * Phi[8] * (c0[4] * c0[4] + ob144) + Phi[9] * Phi[9] * (c0[5] * c0[5] + ob144);
  const double tau = fabs(IShi - ISlo);
  double wlo = gammalo * (1.0 + tau / (ISlo + sm)), whi = gammahi * (1.0 + tau / (IShi + sm));
  const double aa = 1.0 / (wlo + whi);
  wlo *= aa;
  whi *= aa;
  const double x1 = 1.0 - f;
  if (!act) return;
  E[z] = Ept * f + (whi * Phi[0] + wlo * e[0]) * x1;
#pragma unroll
  for (int m = 1; m < 4; ++m) E[z + NZ * m] = (whi * Phi[m] + wlo * e[m]) * x1;
#pragma unroll
  for (int m = 4; m < 10; ++m) E[z + NZ * m] = whi * Phi[m] * x1;
}

// Step 9.c in the serial loop order (k_ao_order): one launch per level
// (i+1) + 3 (j+1) + 9 (k+1) of the box, three lanes per zone (one per
// component of E), the box's (j, k) rows from plane klo on.  The zones of a
// level read none of each other; the launches order the levels.
__global__ void __launch_bounds__(64) k_weno_ao_e(const P1 P, const Box b, const double* __restrict__ mom,
                                                   const double* __restrict__ rec, int level, int klo) {
  const int64_t t = tid();
  const int c = (int)(t % 3);
  const int64_t tz = t / 3;
  const int jr = (int)(tz % b.nj), kr = klo + (int)(tz / b.nj);
  if (kr >= b.nk) return;
  const int ir = level - 3 * jr - 9 * kr;
  if (ir < 0 || ir >= b.ni) return;
  ao_e_zone(P, b, mom, rec, ir, jr, kr, c);
}

}  // namespace csp1

extern "C" int cs_limit_all_slopes_p12(const cs_p12_patch* q, void* stream_) {
  if (!q || !q->moments) {
    cs::set_error("cs_limit_all_slopes_p12: null patch or moments");
    return CS_E_INVALID;
  }
  return run_p1(&q->p1, stream_, nullptr, nullptr, 0, q->moments);
}

extern "C" int cs_limit_all_slopes_p12_timed(const cs_p12_patch* q, void* stream_, float* kernel_ms) {
  if (!q || !q->moments || !kernel_ms) {
    cs::set_error("cs_limit_all_slopes_p12_timed: null patch, moments or kernel_ms");
    return CS_E_INVALID;
  }
  return run_p1(&q->p1, stream_, kernel_ms, nullptr, 0, q->moments);
}
