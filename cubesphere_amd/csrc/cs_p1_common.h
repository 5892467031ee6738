// cs_p1_common.h -- small dense kernels shared by the P1 reconstruction of one
// patch (cs_slopes_p1.hip) and the P1 stage over the six sectors of a shell
// (cs_p1_stage.hip): the adjugate 3 x 3 inverse of the WENO stencils
// (410limit_all_slopes_p1.f:2421-3761) and the packed 12 x 12 Cholesky solves
// and rows of the constrained least squares (410...f:4643-5360), and the
// 3-state HLL edge field of the tri-corners (631...f:1-216).
#pragma once
#include <hip/hip_runtime.h>

namespace csp1 {

// WENO (410...f:2421-3761, 5437-7100): each stencil's 3 x 3 offset matrix
// inverted by its adjugate, with the reference's sign(det) / max(1e-32, |det|)
__device__ __forceinline__ void inv3(const double (&m)[3][3], double (&r)[3][3]) {
  const double a11 = m[0][0], a12 = m[0][1], a13 = m[0][2], a21 = m[1][0], a22 = m[1][1], a23 = m[1][2],
               a31 = m[2][0], a32 = m[2][1], a33 = m[2][2];
  const double det =
      -(a13 * a22 * a31) + a12 * a23 * a31 + a13 * a21 * a32 - a11 * a23 * a32 - a12 * a21 * a33 + a11 * a22 * a33;
  const double od = (signbit(det) ? -1.0 : 1.0) / fmax(1.0e-32, fabs(det));
  r[0][0] = (-(a23 * a32) + a22 * a33) * od;
  r[0][1] = (a13 * a32 - a12 * a33) * od;
  r[0][2] = (-(a13 * a22) + a12 * a23) * od;
  r[1][0] = (a23 * a31 - a21 * a33) * od;
  r[1][1] = (-(a13 * a31) + a11 * a33) * od;
  r[1][2] = (a13 * a21 - a11 * a23) * od;
  r[2][0] = (-(a22 * a31) + a21 * a32) * od;
  r[2][1] = (a12 * a31 - a11 * a32) * od;
  r[2][2] = (-(a12 * a21) + a11 * a22) * od;
}

// step 8.d: packed lower-triangular 12 x 12 Cholesky and solves.  The
// factor keeps the reciprocal of each diagonal entry in the diagonal slot, so
// the solves multiply instead of dividing (one reciprocal square root per
// column in the factorisation, none in the solves).
#define LIX(r, c) ((r) * ((r) + 1) / 2 + (c))

__device__ __forceinline__ void chol12(double* L) {
#pragma unroll
  for (int j = 0; j < 12; ++j) {
    double s = L[LIX(j, j)];
#pragma unroll
    for (int k = 0; k < j; ++k) s -= L[LIX(j, k)] * L[LIX(j, k)];
    const double od = 1.0 / sqrt(s);
    L[LIX(j, j)] = od;
#pragma unroll
    for (int i = j + 1; i < 12; ++i) {
      double t = L[LIX(i, j)];
#pragma unroll
      for (int k = 0; k < j; ++k) t -= L[LIX(i, k)] * L[LIX(j, k)];
      L[LIX(i, j)] = t * od;
    }
  }
}
// x <- L^-1 x
__device__ __forceinline__ void fwd12(const double* L, double* x) {
#pragma unroll
  for (int i = 0; i < 12; ++i) {
    double s = x[i];
#pragma unroll
    for (int k = 0; k < i; ++k) s -= L[LIX(i, k)] * x[k];
    x[i] = s * L[LIX(i, i)];
  }
}
// x <- L^-T x
__device__ __forceinline__ void bwd12(const double* L, double* x) {
#pragma unroll
  for (int i = 11; i >= 0; --i) {
    double s = x[i];
#pragma unroll
    for (int k = i + 1; k < 12; ++k) s -= L[LIX(k, i)] * x[k];
    x[i] = s * L[LIX(i, i)];
  }
}
// x <- (L L^T)^-1 x
__device__ __forceinline__ void cholsolve12(const double* L, double* x) {
  fwd12(L, x);
  bwd12(L, x);
}

// the constraint or LSQ row of a face: n (x) (1, r)
__device__ __forceinline__ void brow(const double* n, const double* r, double* row) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    row[4 * c] = n[c];
    row[4 * c + 1] = n[c] * r[0];
    row[4 * c + 2] = n[c] * r[1];
    row[4 * c + 3] = n[c] * r[2];
  }
}

// Two_D_RS_Ez_3State_HLL (620riemann_2d/631...f:1-216): three zones around an
// edge, the faces 13, 21, 32 between them at 120 degrees in the solver frame;
// the HLL field from the wave triangle A, B, C of the six speeds.
__device__ __forceinline__ double emf3_hll(const double* sr, const double* sl, const double* ez, const double* bn) {
  const double h3 = 0.5 * sqrt(3.0), r3 = 1.0 / sqrt(3.0);
  const double o1 = 1.0 / (sr[0] - sl[0]), o2 = 1.0 / (sr[1] - sl[1]), o3 = 1.0 / (sr[2] - sl[2]);
  const double Ax = (sl[0] + 2.0 * sr[1]) * r3, Ay = -sl[0], Bx = (sl[1] - sr[2]) * r3, By = sl[1] + sr[2];
  const double Cx = -(sr[0] + 2.0 * sl[2]) * r3, Cy = -sr[0];
  const double xc = (Ax + Bx + Cx) / 3.0, yc = (Ay + By + Cy) / 3.0;
  const double lx[3] = {Ax - Cx, Bx - Ax, Cx - Bx}, ly[3] = {Ay - Cy, By - Ay, Cy - By};
  const double px[3] = {Ax, Bx, Cx}, py[3] = {Ay, By, Cy};
  const double mx[3] = {0.5 * (Ax + Cx), 0.5 * (Bx + Ax), 0.5 * (Cx + Bx)},
               my[3] = {0.5 * (Ay + Cy), 0.5 * (By + Ay), 0.5 * (Cy + By)};
  double l[3], nx[3], ny[3], Sn[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    l[q] = sqrt(lx[q] * lx[q] + ly[q] * ly[q]);
    nx[q] = ly[q] / l[q];
    ny[q] = -lx[q] / l[q];
    Sn[q] = nx[q] * px[q] + ny[q] * py[q];
  }
  const double s = 0.5 * (l[0] + l[1] + l[2]);
  const double area = sqrt(s * (s - l[0]) * (s - l[1]) * (s - l[2]));
  // zone fields in the solver frame, their components along each face
  const double Bx1 = (2.0 * bn[1] + bn[0]) * r3, Bx2 = (bn[1] - bn[2]) * r3, Bx3 = -(2.0 * bn[2] + bn[0]) * r3;
  const double By1 = -bn[0], By2 = bn[1] + bn[2];
  const double t21 = -0.5 * Bx1 + h3 * By1, t22 = -0.5 * Bx2 + h3 * By2;
  const double t32 = -0.5 * Bx2 - h3 * By2, t33 = -0.5 * Bx3 - h3 * By1;
  // the 1-D HLL state of each face: tangential B, E, the frame components
  const double Bt13 = (sr[0] * Bx3 - sl[0] * Bx1 + ez[2] - ez[0]) * o1;
  const double E13 = (sr[0] * ez[0] - sl[0] * ez[2] - sr[0] * sl[0] * (Bx3 - Bx1)) * o1;
  const double Bt21 = (sr[1] * t21 - sl[1] * t22 + ez[0] - ez[1]) * o2;
  const double E21 = (sr[1] * ez[1] - sl[1] * ez[0] - sr[1] * sl[1] * (t21 - t22)) * o2;
  const double Bt32 = (sr[2] * t32 - sl[2] * t33 + ez[1] - ez[2]) * o3;
  const double E32 = (sr[2] * ez[2] - sl[2] * ez[1] - sr[2] * sl[2] * (t32 - t33)) * o3;
  const double Es[3] = {E13, E21, E32};
  const double Bsx[3] = {Bt13, -0.5 * Bt21 + h3 * bn[1], -0.5 * Bt32 - h3 * bn[2]};
  const double Bsy[3] = {-bn[0], h3 * Bt21 + 0.5 * bn[1], -h3 * Bt32 + 0.5 * bn[2]};
  if (sl[0] >= 0.0 && sr[1] <= 0.0) return ez[0];
  if (sl[1] >= 0.0 && sr[2] <= 0.0) return ez[1];
  if (sl[2] >= 0.0 && sr[0] <= 0.0) return ez[2];
  if (sl[0] * (sr[0] + 2.0 * sl[2]) + sr[0] * (sl[0] + 2.0 * sr[1]) <= 0.0) return E13;
  if ((sl[1] + sr[2]) * (sl[0] + 2.0 * sr[1]) + (sl[1] - sr[2]) * sl[0] <= 0.0) return E21;
  if ((sl[1] + sr[2]) * (sr[0] + 2.0 * sl[2]) - (sl[1] - sr[2]) * sr[0] <= 0.0) return E32;
  // strongly interacting state: the face states' fluxes around the triangle
  double sx = 0.0, sy = 0.0, mxs = 0.0, mys = 0.0;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const double fx = ny[q] * Es[q] - Sn[q] * Bsx[q], fy = -nx[q] * Es[q] - Sn[q] * Bsy[q];
    sx += fx * l[q];
    sy += fy * l[q];
    mxs += (mx[q] - xc) * fy * l[q];
    mys += (my[q] - yc) * fx * l[q];
  }
  const double BSSx = sx / (-2.0 * area), BSSy = sy / (-2.0 * area);
  return 0.5 * ((-xc * BSSy - mxs / area) + (yc * BSSx + mys / area));
}

}  // namespace csp1
