// cs_geometry.hip -- the quadratic isoparametric zone and face geometry of the
// reference's P1 reconstruction, built on the GPU for the six sectors:
//   EXTRACT_VERT_COORDS_PATCH       (210extract_vert_coords_patch.f:7)
//   COORDS_TO_WT_QUADPTS_FACES_ISO2 (220coords_to_wt_quadpts_faces_iso2.f:7)
//   COORDS_TO_WT_QUADPTS_VOLUME_ISO2 (240coords_to_wt_quadpts_volume_iso2.f:7)
//   BUILD_MODE_INDICES + INIT_BASIS_FUNCTION_COEFS (204…:7, 250…:5)
//
// Each zone and face is an independent work item; the element maps are the
// 20-node (zones) and 8-node (faces) serendipity elements on [-1/2, 1/2]^d,
// whose basis values and derivatives at the quadrature points sit in constant
// memory.  The kernels are store bound (the outputs are 20-40 words per item
// against a few cached vertex reads).  Contraction is off so the sums round
// as the reference's sequential Fortran sums do.
#include <hip/hip_runtime.h>

#include <math.h>
#include <string.h>

#include <vector>

#include "../../include/cubesphere_hip.h"
#include "cs_internal.h"

#pragma clang fp contract(off)

namespace csg {

constexpr double kSmall = 1.0e-12;  // RIEMANN_COM smallnum (090init_riemann.f:12)

// face rule: centre 1/3, the four edge midpoints 1/6 (220…:212-226)
__constant__ double c_fq_w[5];
__constant__ double c_q8[5][8][3];    // [point][node][value, d/dx, d/dy]
__constant__ double c_h20[8][20][4];  // [point][node][value, d/dxi, d/deta, d/dzeta]
__constant__ double c_h27[27][20][4]; // the same at the 27 points of the third-order rule
__constant__ double c_w27[27];        // their tensor weights

// 8-node quadrilateral: corners (--, +-, ++, -+), then the edge nodes
// (0,-), (+,0), (0,+), (-,0)
const int kQ8[8][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1}, {0, -1}, {1, 0}, {0, 1}, {-1, 0}};
// 20-node hexahedron: corners (lower layer, then upper), edge nodes along x,
// along y, along z
const int kH20[20][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1}, {-1, -1, 1}, {1, -1, 1}, {1, 1, 1},
                         {-1, 1, 1},   {0, -1, -1}, {0, 1, -1}, {0, 1, 1},   {0, -1, 1},  {-1, 0, -1}, {-1, 0, 1},
                         {1, 0, 1},    {1, 0, -1},  {-1, -1, 0}, {1, -1, 0}, {1, 1, 0},   {-1, 1, 0}};

// Serendipity basis and gradient at one point of [-1/2, 1/2]^dim for a node
// with signs sg (0 on the axis the node's edge runs along).
void serendipity(int dim, const int* sg, const double* x, double* val, double* grad) {
  double l[3], dl[3];
  int e = -1;
  for (int a = 0; a < dim; ++a) {
    if (sg[a] == 0) {
      e = a;
      l[a] = 1.0 - 4.0 * x[a] * x[a];
      dl[a] = -8.0 * x[a];
    } else {
      l[a] = 0.5 + sg[a] * x[a];
      dl[a] = sg[a];
    }
  }
  double prod = 1.0;
  for (int a = 0; a < dim; ++a) prod *= l[a];
  if (e < 0) {  // corner: prod * (2 s.x - (dim - 1))
    double t = 0.0;
    for (int a = 0; a < dim; ++a) t = t + 2.0 * sg[a] * x[a];
    t = t - (double)(dim - 1);
    *val = prod * t;
    for (int a = 0; a < dim; ++a) {
      double p = 1.0;
      for (int b = 0; b < dim; ++b)
        if (b != a) p *= l[b];
      grad[a] = dl[a] * p * t + prod * 2.0 * sg[a];
    }
  } else {
    *val = prod;
    for (int a = 0; a < dim; ++a) {
      double p = 1.0;
      for (int b = 0; b < dim; ++b)
        if (b != a) p *= l[b];
      grad[a] = dl[a] * p;
    }
  }
}

int load_tables() {
  static bool done[64] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return CS_E_HIP;
  if (dev >= 0 && dev < 64 && done[dev]) return CS_OK;
  double w[5] = {1.0 / 3.0, 1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0};
  const double pq[5][2] = {{0.0, 0.0}, {0.5, 0.0}, {-0.5, 0.0}, {0.0, 0.5}, {0.0, -0.5}};
  double q8[5][8][3], h20[8][20][4];
  for (int p = 0; p < 5; ++p)
    for (int n = 0; n < 8; ++n) {
      double g[2];
      serendipity(2, kQ8[n], pq[p], &q8[p][n][0], g);
      q8[p][n][1] = g[0];
      q8[p][n][2] = g[1];
    }
  // 2x2x2 Gauss points +-1/(2 sqrt 3), x fastest (240…:97-152)
  const double xg = 0.5 * sqrt(1.0 / 3.0);
  for (int p = 0; p < 8; ++p) {
    double x[3] = {(p & 1) ? xg : -xg, (p & 2) ? xg : -xg, (p & 4) ? xg : -xg};
    for (int n = 0; n < 20; ++n) {
      double g[3];
      serendipity(3, kH20[n], x, &h20[p][n][0], g);
      for (int a = 0; a < 3; ++a) h20[p][n][1 + a] = g[a];
    }
  }
  // 3x3x3 Gauss points 0, +-sqrt(3/5)/2 with weights 8/18, 5/18, x fastest;
  // the weight of point (ix, iy, iz) is lin(iz) lin(iy) lin(ix) in that
  // order (240…:134-181: the weight table is filled with x slowest)
  static double h27[27][20][4], w27[27];
  const double xg3[3] = {-0.5 * sqrt(3.0 / 5.0), 0.0, 0.5 * sqrt(3.0 / 5.0)};
  const double lw3[3] = {5.0 / 18.0, 8.0 / 18.0, 5.0 / 18.0};
  for (int p = 0; p < 27; ++p) {
    const int ix = p % 3, iy = (p / 3) % 3, iz = p / 9;
    double x[3] = {xg3[ix], xg3[iy], xg3[iz]};
    for (int n = 0; n < 20; ++n) {
      double g[3];
      serendipity(3, kH20[n], x, &h27[p][n][0], g);
      for (int a = 0; a < 3; ++a) h27[p][n][1 + a] = g[a];
    }
    w27[p] = lw3[iz] * lw3[iy] * lw3[ix];
  }
  if (hipMemcpyToSymbol(HIP_SYMBOL(c_fq_w), w, sizeof(w)) != hipSuccess ||
      hipMemcpyToSymbol(HIP_SYMBOL(c_q8), q8, sizeof(q8)) != hipSuccess ||
      hipMemcpyToSymbol(HIP_SYMBOL(c_h20), h20, sizeof(h20)) != hipSuccess ||
      hipMemcpyToSymbol(HIP_SYMBOL(c_h27), h27, sizeof(h27)) != hipSuccess ||
      hipMemcpyToSymbol(HIP_SYMBOL(c_w27), w27, sizeof(w27)) != hipSuccess)
    return CS_E_HIP;
  if (dev >= 0 && dev < 64) done[dev] = true;
  return CS_OK;
}

// extents: vertices (0-g:N+g)^2 x (0-g:nr+g); zones (1-g:N+g)^2 x (1-g:nr+g);
// x-faces (0-g:N+g, 1-g:N+g, 1-g:nr+g); y-faces (1-g:N+g, 0-g:N+g, 1-g:nr+g);
// z-faces (1-g:N+g, 1-g:N+g, 0-g:nr+g).  Each array is sector-major with the
// trailing dimensions of the reference's per-patch arrays.
struct GDims {
  int N, NR, g, nv, nvr, nc, nrc;
  int64_t vs, zs, xs, zfs;
};
GDims gdims(const cs_config* c) {
  GDims d;
  d.N = c->ntp; d.NR = c->nr; d.g = c->nghost;
  d.nv = d.N + 2 * d.g + 1; d.nvr = d.NR + 2 * d.g + 1;
  d.nc = d.N + 2 * d.g; d.nrc = d.NR + 2 * d.g;
  d.vs = (int64_t)d.nv * d.nv * d.nvr;
  d.zs = (int64_t)d.nc * d.nc * d.nrc;
  d.xs = (int64_t)d.nv * d.nc * d.nrc;
  d.zfs = (int64_t)d.nc * d.nc * d.nvr;
  return d;
}

__device__ __forceinline__ uint32_t tid() { return blockIdx.x * blockDim.x + threadIdx.x; }

__device__ __forceinline__ void vload(const GDims& d, const double* __restrict__ V, int s, int i, int j, int k,
                                      double* x) {
  const int64_t o = (int64_t)(i + d.g) + (int64_t)d.nv * ((j + d.g) + (int64_t)d.nv * (k + d.g));
  for (int c = 0; c < 3; ++c) x[c] = V[(int64_t)(3 * s + c) * d.vs + o];
}
__device__ __forceinline__ double len3(const double* v) { return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }
// midpoint of a lateral edge, pushed onto the sphere of radius r (r < 0: a
// straight radial edge, kept as the plain mean)
__device__ __forceinline__ void mid(const double* a, const double* b, double r, double* m) {
  for (int c = 0; c < 3; ++c) m[c] = 0.5 * (a[c] + b[c]);
  if (r >= 0.0) {
    const double t = r / fmax(kSmall, len3(m));
    for (int c = 0; c < 3; ++c) m[c] = m[c] * t;
  }
}

// ------------------------------------------------------------------ 210
__global__ void k_vert(GDims d, const double* __restrict__ lat, const double* __restrict__ rb, double* vert) {
  const uint32_t t = tid();
  if (t >= (uint64_t)6 * d.vs) return;
  const int s = (int)(t / d.vs);
  const int64_t o = t - (int64_t)s * d.vs;
  const int i = (int)(o % d.nv), j = (int)((o / d.nv) % d.nv), k = (int)(o / ((int64_t)d.nv * d.nv));
  const int64_t lo = (int64_t)i + (int64_t)d.nv * j;  // lattice (-g:N+g)^2 x 3 x 6
  const double r = rb[k];
  for (int c = 0; c < 3; ++c)
    vert[(int64_t)(3 * s + c) * d.vs + o] = r * lat[lo + (int64_t)d.nv * d.nv * (c + 3 * s)];
}
__global__ void k_mask(GDims d, const double* __restrict__ lat, int32_t* mask) {
  const uint32_t t = tid();
  if (t >= (uint64_t)6 * d.zs) return;
  const int s = (int)(t / d.zs);
  const int64_t o = t - (int64_t)s * d.zs;
  const int i = 1 - d.g + (int)(o % d.nc), j = 1 - d.g + (int)((o / d.nc) % d.nc);
  int on = 1;
  for (int a = 0; a < 4; ++a) {
    const int ii = i - (a & 1) + d.g, jj = j - (a >> 1) + d.g;
    double u[3];
    for (int c = 0; c < 3; ++c) u[c] = lat[ii + (int64_t)d.nv * jj + (int64_t)d.nv * d.nv * (c + 3 * s)];
    if (len3(u) < kSmall) on = 0;
  }
  mask[t] = on;
}

// ------------------------------------------------------------------ 220
// One face from its eight nodes: five quadrature points, weights divided by
// the area, the unit normal at the centre (all five on curved z-faces), the
// area and the weighted centroid.  S: element stride of the face arrays.
__device__ void face_element(double (*X)[3], int64_t o, int64_t S, bool all_normals, double* pts, double* wts,
                             double* nrm, double* area, double* cen, double* bwts) {
  double w[5], P[5][3];
  for (int p = 0; p < 5; ++p) {
    double A[3], B[3];
    for (int c = 0; c < 3; ++c) {
      double x = 0.0, a = 0.0, b = 0.0;
      for (int n = 0; n < 8; ++n) {
        x = x + X[n][c] * c_q8[p][n][0];
        a = a + X[n][c] * c_q8[p][n][1];
        b = b + X[n][c] * c_q8[p][n][2];
      }
      P[p][c] = x;
      pts[o + (c + 3 * p) * S] = x;
      A[c] = a;
      B[c] = b;
    }
    const double C[3] = {A[1] * B[2] - A[2] * B[1], A[2] * B[0] - A[0] * B[2], A[0] * B[1] - A[1] * B[0]};
    const double cm = sqrt(C[0] * C[0] + C[1] * C[1] + C[2] * C[2]);
    w[p] = c_fq_w[p] * cm;
    if (p == 0 || all_normals) {
      const double inv = 1.0 / fmax(cm, kSmall);
      for (int c = 0; c < 3; ++c) nrm[o + (c + 3 * (all_normals ? p : 0)) * S] = inv * C[c];
    }
  }
  double a = 0.0;
  for (int p = 0; p < 5; ++p) a = a + w[p];
  // the B-field points of COORDS_TO_WT_QUADPTS_BFLD_ISO2 (230…:8) are these
  // points and normals with the weights times the reciprocal of their sum
  const double rb = 1.0 / fmax(a, kSmall);
  for (int p = 0; p < 5; ++p) bwts[o + p * S] = rb * w[p];
  a = fmax(kSmall, a);
  area[o] = a;
  for (int p = 0; p < 5; ++p) {
    w[p] = w[p] / a;
    wts[o + p * S] = w[p];
  }
  for (int c = 0; c < 3; ++c) {
    double m = 0.0;
    for (int p = 0; p < 5; ++p) m = m + w[p] * P[p][c];
    cen[o + c * S] = m;
  }
}
struct FaceOut {
  double *pts, *wts, *nrm, *area, *cen, *bwts;
};
// fam 0: x-faces, 1: y-faces, 2: z-faces
__global__ void __launch_bounds__(256) k_faces(GDims d, const double* __restrict__ vert, int fam, FaceOut f) {
  const uint32_t t = tid();
  const int64_t S = fam == 2 ? d.zfs : d.xs;
  if (t >= (uint64_t)6 * S) return;
  const int s = (int)(t / S);
  const int64_t o = t - (int64_t)s * S;
  double X[8][3];
  if (fam == 0) {  // (0-g:N+g) x (1-g:N+g) x (1-g:nr+g); j-edges on spheres, k-edges straight
    const int i = -d.g + (int)(o % d.nv), j = 1 - d.g + (int)((o / d.nv) % d.nc);
    const int k = 1 - d.g + (int)(o / ((int64_t)d.nv * d.nc));
    vload(d, vert, s, i, j - 1, k - 1, X[0]); vload(d, vert, s, i, j, k - 1, X[1]);
    vload(d, vert, s, i, j, k, X[2]);         vload(d, vert, s, i, j - 1, k, X[3]);
    mid(X[0], X[1], len3(X[0]), X[4]); mid(X[1], X[2], -1.0, X[5]);
    mid(X[2], X[3], len3(X[2]), X[6]); mid(X[3], X[0], -1.0, X[7]);
  } else if (fam == 1) {  // (1-g:N+g) x (0-g:N+g) x (1-g:nr+g); i-edges on spheres
    const int i = 1 - d.g + (int)(o % d.nc), j = -d.g + (int)((o / d.nc) % d.nv);
    const int k = 1 - d.g + (int)(o / ((int64_t)d.nc * d.nv));
    vload(d, vert, s, i - 1, j, k - 1, X[0]); vload(d, vert, s, i - 1, j, k, X[1]);
    vload(d, vert, s, i, j, k, X[2]);         vload(d, vert, s, i, j, k - 1, X[3]);
    mid(X[0], X[1], -1.0, X[4]); mid(X[1], X[2], len3(X[2]), X[5]);
    mid(X[2], X[3], -1.0, X[6]); mid(X[3], X[0], len3(X[0]), X[7]);
  } else {  // (1-g:N+g)^2 x (0-g:nr+g); every edge on the face's sphere
    const int i = 1 - d.g + (int)(o % d.nc), j = 1 - d.g + (int)((o / d.nc) % d.nc);
    const int k = -d.g + (int)(o / ((int64_t)d.nc * d.nc));
    vload(d, vert, s, i - 1, j - 1, k, X[0]); vload(d, vert, s, i, j - 1, k, X[1]);
    vload(d, vert, s, i, j, k, X[2]);         vload(d, vert, s, i - 1, j, k, X[3]);
    const double r = len3(X[0]);
    for (int n = 0; n < 4; ++n) mid(X[n], X[(n + 1) & 3], r, X[4 + n]);
  }
  const int64_t b = (int64_t)s * S;
  face_element(X, o, S, fam == 2, f.pts + 15 * b, f.wts + 5 * b, f.nrm + (fam == 2 ? 15 : 3) * b, f.area + b,
               f.cen + 3 * b, f.bwts + 5 * b);
  if (fam == 2) {  // the centroid goes back onto the face's sphere
    double* cn = f.cen + 3 * b;
    double v[3] = {cn[o], cn[o + S], cn[o + 2 * S]};
    const double r1 = fmax(kSmall, len3(v)), r2 = len3(X[2]);
    for (int c = 0; c < 3; ++c) cn[o + c * S] = r2 * v[c] / r1;
  }
}
// scale length: the radial edge, then the shorter lateral edges (220…:2212-2290)
__device__ __forceinline__ double dist(const double* a, const double* b) {
  const double x = a[0] - b[0], y = a[1] - b[1], z = a[2] - b[2];
  return sqrt(x * x + y * y + z * z);
}
__global__ void k_scale(GDims d, const double* __restrict__ vert, double* scale) {
  const uint32_t t = tid();
  if (t >= (uint64_t)6 * d.zs) return;
  const int s = (int)(t / d.zs);
  const int64_t o = t - (int64_t)s * d.zs;
  const int i = 1 - d.g + (int)(o % d.nc), j = 1 - d.g + (int)((o / d.nc) % d.nc);
  const int k = 1 - d.g + (int)(o / ((int64_t)d.nc * d.nc));
  double a[3], b[3], c[3], e[3];
  vload(d, vert, s, i, j, k, a);
  vload(d, vert, s, i, j, k - 1, b);
  double L = dist(a, b);
  vload(d, vert, s, i - 1, j, k - 1, a); vload(d, vert, s, i - 1, j - 1, k - 1, c);
  vload(d, vert, s, i, j - 1, k - 1, e);
  L = fmin(fmin(L, dist(a, b)), dist(c, e));  // x-direction edges at j and j-1
  L = fmin(fmin(L, dist(e, b)), dist(c, a));  // y-direction edges at i and i-1
  scale[t] = L;
}

// ------------------------------------------------------------------ 240
struct VolOut {
  double *pts, *wts, *tan, *vol, *cen;
};
__global__ void __launch_bounds__(256) k_volume(GDims d, const double* __restrict__ vert, VolOut v) {
  const uint32_t t = tid();
  if (t >= (uint64_t)6 * d.zs) return;
  const int s = (int)(t / d.zs);
  const int64_t o = t - (int64_t)s * d.zs, S = d.zs, b = (int64_t)s * S;
  const int i = 1 - d.g + (int)(o % d.nc), j = 1 - d.g + (int)((o / d.nc) % d.nc);
  const int k = 1 - d.g + (int)(o / ((int64_t)d.nc * d.nc));
  double X[20][3];
  for (int n = 0; n < 8; ++n)
    vload(d, vert, s, i - 1 + ((n + 1) >> 1 & 1), j - 1 + ((n >> 1) & 1), k - 1 + (n >> 2), X[n]);
  // nodes 1, 2 of each layer alternate in i as 0 1 1 0: (n + 1) >> 1 & 1
  const double r1 = len3(X[0]), r2 = len3(X[4]);
  mid(X[0], X[1], r1, X[8]);  mid(X[2], X[3], r1, X[9]);  mid(X[6], X[7], r2, X[10]); mid(X[4], X[5], r2, X[11]);
  mid(X[0], X[3], r1, X[12]); mid(X[4], X[7], r2, X[13]); mid(X[5], X[6], r2, X[14]); mid(X[1], X[2], r1, X[15]);
  for (int n = 0; n < 4; ++n) mid(X[n], X[n + 4], -1.0, X[16 + n]);
  double w[8], vol = 0.0;
  for (int p = 0; p < 8; ++p) {
    double T[3][3];
    for (int c = 0; c < 3; ++c) {
      double x = 0.0, t0 = 0.0, t1 = 0.0, t2 = 0.0;
      for (int n = 0; n < 20; ++n) {
        x = x + X[n][c] * c_h20[p][n][0];
        t0 = t0 + X[n][c] * c_h20[p][n][1];
        t1 = t1 + X[n][c] * c_h20[p][n][2];
        t2 = t2 + X[n][c] * c_h20[p][n][3];
      }
      v.pts[24 * b + o + (c + 3 * p) * S] = x;
      T[0][c] = t0; T[1][c] = t1; T[2][c] = t2;
    }
    const double D0 = T[0][1] * T[1][2] - T[0][2] * T[1][1], D1 = T[0][2] * T[1][0] - T[0][0] * T[1][2],
                 D2 = T[0][0] * T[1][1] - T[0][1] * T[1][0];
    w[p] = 0.125 * fabs(T[2][0] * D0 + T[2][1] * D1 + T[2][2] * D2);
    for (int q = 0; q < 3; ++q)
      for (int c = 0; c < 3; ++c) v.tan[72 * b + o + (c + 3 * (q + 3 * p)) * S] = T[q][c];
  }
  for (int p = 0; p < 8; ++p) vol = vol + w[p];
  vol = fmax(kSmall, vol);
  v.vol[b + o] = vol;
  for (int p = 0; p < 8; ++p) {
    w[p] = w[p] / vol;
    v.wts[8 * b + o + p * S] = w[p];
  }
  for (int c = 0; c < 3; ++c) {
    const double c0 = 0.125 * (X[0][c] + X[1][c] + X[2][c] + X[3][c] + X[4][c] + X[5][c] + X[6][c] + X[7][c]);
    double m = 0.0;
    for (int p = 0; p < 8; ++p) m = m + w[p] * (v.pts[24 * b + o + (c + 3 * p) * S] - c0);
    v.cen[3 * b + o + c * S] = c0 + m;
  }
}

// The same at NUMORDER 3: 27 points (no tangents are kept).
__global__ void __launch_bounds__(256) k_volume3(GDims d, const double* __restrict__ vert, VolOut v) {
  const uint32_t t = tid();
  if (t >= (uint64_t)6 * d.zs) return;
  const int s = (int)(t / d.zs);
  const int64_t o = t - (int64_t)s * d.zs, S = d.zs, b = (int64_t)s * S;
  const int i = 1 - d.g + (int)(o % d.nc), j = 1 - d.g + (int)((o / d.nc) % d.nc);
  const int k = 1 - d.g + (int)(o / ((int64_t)d.nc * d.nc));
  double X[20][3];
  for (int n = 0; n < 8; ++n)
    vload(d, vert, s, i - 1 + ((n + 1) >> 1 & 1), j - 1 + ((n >> 1) & 1), k - 1 + (n >> 2), X[n]);
  const double r1 = len3(X[0]), r2 = len3(X[4]);
  mid(X[0], X[1], r1, X[8]);  mid(X[2], X[3], r1, X[9]);  mid(X[6], X[7], r2, X[10]); mid(X[4], X[5], r2, X[11]);
  mid(X[0], X[3], r1, X[12]); mid(X[4], X[7], r2, X[13]); mid(X[5], X[6], r2, X[14]); mid(X[1], X[2], r1, X[15]);
  for (int n = 0; n < 4; ++n) mid(X[n], X[n + 4], -1.0, X[16 + n]);
  double vol = 0.0;
  for (int p = 0; p < 27; ++p) {
    double T[3][3];
    for (int c = 0; c < 3; ++c) {
      double x = 0.0, t0 = 0.0, t1 = 0.0, t2 = 0.0;
      for (int n = 0; n < 20; ++n) {
        x = x + X[n][c] * c_h27[p][n][0];
        t0 = t0 + X[n][c] * c_h27[p][n][1];
        t1 = t1 + X[n][c] * c_h27[p][n][2];
        t2 = t2 + X[n][c] * c_h27[p][n][3];
      }
      v.pts[81 * b + o + (c + 3 * p) * S] = x;
      T[0][c] = t0; T[1][c] = t1; T[2][c] = t2;
    }
    const double D0 = T[0][1] * T[1][2] - T[0][2] * T[1][1], D1 = T[0][2] * T[1][0] - T[0][0] * T[1][2],
                 D2 = T[0][0] * T[1][1] - T[0][1] * T[1][0];
    const double w = c_w27[p] * fabs(T[2][0] * D0 + T[2][1] * D1 + T[2][2] * D2);
    v.wts[27 * b + o + p * S] = w;
    vol = vol + w;
  }
  vol = fmax(kSmall, vol);
  v.vol[b + o] = vol;
  double m[3] = {0.0, 0.0, 0.0}, c0[3];
  for (int c = 0; c < 3; ++c)
    c0[c] = 0.125 * (X[0][c] + X[1][c] + X[2][c] + X[3][c] + X[4][c] + X[5][c] + X[6][c] + X[7][c]);
  for (int p = 0; p < 27; ++p) {
    const double w = v.wts[27 * b + o + p * S] / vol;
    v.wts[27 * b + o + p * S] = w;
    for (int c = 0; c < 3; ++c) m[c] = m[c] + w * (v.pts[81 * b + o + (c + 3 * p) * S] - c0[c]);
  }
  for (int c = 0; c < 3; ++c) v.cen[3 * b + o + c * S] = c0[c] + m[c];
}

// ------------------------------------------------------------------ 250
// Third-order modes (204…:55-95): 1, x, y, z, x^2, y^2, z^2, xy, yz, xz; the
// moments of modes 2-10 over the 27 points about the centroid in units of
// the scale length (inactive zones keep zeros).
__global__ void k_moments10(GDims d, const int32_t* __restrict__ mask, const double* __restrict__ pts,
                            const double* __restrict__ wts, const double* __restrict__ cen,
                            const double* __restrict__ scale, double* mom) {
  const uint32_t t = tid();
  if (t >= (uint64_t)6 * d.zs) return;
  const int s = (int)(t / d.zs);
  const int64_t o = t - (int64_t)s * d.zs, S = d.zs, b = (int64_t)s * S;
  double m[10] = {};
  if (mask[t] == 1) {
    const double aa = 1.0 / fmax(kSmall, fabs(scale[t]));
    double ce[3];
    for (int c = 0; c < 3; ++c) ce[c] = cen[3 * b + o + c * S];
    for (int p = 0; p < 27; ++p) {
      const double w = wts[27 * b + o + p * S];
      double x[3];
      for (int c = 0; c < 3; ++c) x[c] = (pts[81 * b + o + (c + 3 * p) * S] - ce[c]) * aa;
      const double f[9] = {x[0], x[1], x[2], x[0] * x[0], x[1] * x[1], x[2] * x[2], x[0] * x[1], x[1] * x[2],
                           x[0] * x[2]};
      for (int q = 0; q < 9; ++q) m[1 + q] = m[1 + q] + w * f[q];
    }
  }
  for (int q = 0; q < 10; ++q) mom[10 * b + o + q * S] = m[q];
}

// ------------------------------------------------------------------ 250
// Second-order modes: 1, x, y, z; the moments of modes 2-4 about the centroid
// in units of the scale length (inactive zones keep zeros).
__global__ void k_moments(GDims d, const int32_t* __restrict__ mask, const double* __restrict__ pts,
                          const double* __restrict__ wts, const double* __restrict__ cen,
                          const double* __restrict__ scale, double* mom) {
  const uint32_t t = tid();
  if (t >= (uint64_t)6 * d.zs) return;
  const int s = (int)(t / d.zs);
  const int64_t o = t - (int64_t)s * d.zs, S = d.zs, b = (int64_t)s * S;
  double m[4] = {0.0, 0.0, 0.0, 0.0};
  if (mask[t] == 1) {
    const double aa = 1.0 / fmax(kSmall, fabs(scale[t]));
    for (int c = 0; c < 3; ++c) {
      const double ce = cen[3 * b + o + c * S];
      double acc = 0.0;
      for (int p = 0; p < 8; ++p) acc = acc + wts[8 * b + o + p * S] * ((pts[24 * b + o + (c + 3 * p) * S] - ce) * aa);
      m[1 + c] = 0.0 + acc;
    }
  }
  for (int q = 0; q < 4; ++q) mom[4 * b + o + q * S] = m[q];
}

unsigned nb(uint64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace csg

using namespace csg;

extern "C" int cs_geometry_sizes(const cs_config* cfg, int64_t* sizes) {
  cs_sizes sz;
  int rc = cs_sizes_of(cfg, &sz);
  if (rc) return rc;
  if (!sizes) return CS_E_INVALID;
  GDims d = gdims(cfg);
  sizes[0] = 6 * d.vs;   // vertices (x3 components)
  sizes[1] = 6 * d.zs;   // zones
  sizes[2] = 6 * d.xs;   // x-faces
  sizes[3] = 6 * d.xs;   // y-faces
  sizes[4] = 6 * d.zfs;  // z-faces
  return CS_OK;
}

extern "C" int cs_mode_indices(int32_t order, int32_t* n_modes, int32_t* ii, int32_t* jj, int32_t* kk) {
  // 204build_mode_indices.f:40-150: 1, x, y, z; then x^2, y^2, z^2, xy, yz, xz;
  // then the ten cubic monomials
  static const int8_t tab[20][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {2, 0, 0}, {0, 2, 0}, {0, 0, 2},
                                    {1, 1, 0}, {0, 1, 1}, {1, 0, 1}, {3, 0, 0}, {0, 3, 0}, {0, 0, 3}, {2, 1, 0},
                                    {2, 0, 1}, {1, 2, 0}, {0, 2, 1}, {1, 0, 2}, {0, 1, 2}, {1, 1, 1}};
  if (order < 1 || order > 4 || !n_modes) return CS_E_INVALID;
  const int n = order == 1 ? 1 : (order == 2 ? 4 : (order == 3 ? 10 : 20));
  *n_modes = n;
  for (int m = 0; m < n; ++m) {
    if (ii) ii[m] = tab[m][0];
    if (jj) jj[m] = tab[m][1];
    if (kk) kk[m] = tab[m][2];
  }
  return CS_OK;
}

namespace {
struct DevBuf {
  std::vector<void*> p;
  ~DevBuf() {
    for (void* q : p) hipFree(q);
  }
  template <typename T>
  T* get(int64_t n) {  // null on failure; every request is kept for the check
    void* q = nullptr;
    if (hipMalloc(&q, sizeof(T) * (size_t)n) != hipSuccess) q = nullptr;
    p.push_back(q);
    return (T*)q;
  }
};
int copy_out(void* host, const void* dev, size_t bytes, hipStream_t st) {
  if (!host) return CS_OK;
  return hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, st) == hipSuccess ? CS_OK : CS_E_HIP;
}
}  // namespace

// the vertices, mask and scale lengths of the six sectors on the device
// (201-203 inputs, 210, 220…:2212-2290), shared by the geometry builders
static int build_vertices(const cs_config* cfg, const GDims& d, DevBuf& B, hipStream_t st, double** vert,
                          int32_t** mask, double** scale) {
  std::vector<double> lat((size_t)d.nv * d.nv * 18), rbh((size_t)d.nvr);
  int rc = cs_build_unit_lattice(cfg, lat.data());
  if (rc) return rc;
  for (int s = 0; s < 6; ++s)
    for (int c = 0; c < 3; ++c)
      for (int j = 0; j < d.nv; ++j)
        for (int i = 0; i < d.nv; ++i) {
          const bool ci = i < d.g || i > d.g + d.N, cj = j < d.g || j > d.g + d.N;
          if (ci && cj) lat[(size_t)i + (size_t)d.nv * (j + (size_t)d.nv * (c + 3 * (size_t)s))] = 0.0;
        }
  if ((rc = cs_build_radial_mesh(cfg, rbh.data()))) return rc;
  double* dlat = B.get<double>((int64_t)lat.size());
  double* drb = B.get<double>(d.nvr);
  *vert = B.get<double>(18 * d.vs);
  *mask = B.get<int32_t>(6 * d.zs);
  *scale = B.get<double>(6 * d.zs);
  if (!dlat || !drb || !*vert || !*mask || !*scale) return CS_E_NOMEM;
  if (hipMemcpyAsync(dlat, lat.data(), sizeof(double) * lat.size(), hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(drb, rbh.data(), sizeof(double) * rbh.size(), hipMemcpyHostToDevice, st) != hipSuccess)
    return CS_E_HIP;
  k_vert<<<nb(6 * d.vs), 256, 0, st>>>(d, dlat, drb, *vert);
  k_mask<<<nb(6 * d.zs), 256, 0, st>>>(d, dlat, *mask);
  k_scale<<<nb(6 * d.zs), 256, 0, st>>>(d, *vert, *scale);
  // the lattice and radii must outlive the copies: wait before they go
  return hipStreamSynchronize(st) == hipSuccess ? CS_OK : CS_E_HIP;
}

extern "C" int cs_build_volume_p12(const cs_config* cfg, const cs_volume_p12* out) {
  cs_sizes sz;
  int rc = cs_sizes_of(cfg, &sz);
  if (rc) return rc;
  if (!out) return CS_E_INVALID;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return CS_E_HIP;
  if ((rc = load_tables())) return rc;
  const GDims d = gdims(cfg);
  DevBuf B;
  hipStream_t st;
  if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) return CS_E_HIP;
  double *vert = nullptr, *scale = nullptr;
  int32_t* mask = nullptr;
  rc = build_vertices(cfg, d, B, st, &vert, &mask, &scale);
  VolOut Vo{};
  double* mom = nullptr;
  if (!rc) {
    Vo.pts = B.get<double>(6 * 81 * d.zs);
    Vo.wts = B.get<double>(6 * 27 * d.zs);
    Vo.vol = B.get<double>(6 * d.zs);
    Vo.cen = B.get<double>(18 * d.zs);
    mom = B.get<double>(60 * d.zs);
    if (!Vo.pts || !Vo.wts || !Vo.vol || !Vo.cen || !mom) rc = CS_E_NOMEM;
  }
  if (!rc) {
    k_volume3<<<nb(6 * d.zs), 256, 0, st>>>(d, vert, Vo);
    k_moments10<<<nb(6 * d.zs), 256, 0, st>>>(d, mask, Vo.pts, Vo.wts, Vo.cen, scale, mom);
    if (hipGetLastError() != hipSuccess) rc = CS_E_HIP;
  }
  const size_t zb = sizeof(double) * 6 * d.zs;
  if (!rc) rc = copy_out(out->v_pts, Vo.pts, 81 * zb, st);
  if (!rc) rc = copy_out(out->v_wts, Vo.wts, 27 * zb, st);
  if (!rc) rc = copy_out(out->v_vol, Vo.vol, zb, st);
  if (!rc) rc = copy_out(out->v_cen, Vo.cen, 3 * zb, st);
  if (!rc) rc = copy_out(out->moments, mom, 10 * zb, st);
  if (hipStreamSynchronize(st) != hipSuccess && !rc) rc = CS_E_HIP;
  hipStreamDestroy(st);
  return rc;
}

extern "C" int cs_build_geometry(const cs_config* cfg, const cs_geometry* out, float* kernel_ms) {
  cs_sizes sz;
  int rc = cs_sizes_of(cfg, &sz);
  if (rc) return rc;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return CS_E_HIP;
  if ((rc = load_tables())) return rc;
  const GDims d = gdims(cfg);
  // inputs: the unit lattice (201/203) and the radial vertices (202)
  std::vector<double> lat((size_t)d.nv * d.nv * 18), rbh((size_t)d.nvr);
  if ((rc = cs_build_unit_lattice(cfg, lat.data()))) return rc;
  // FILL_GHOST2D_GLOBAL (203fill_ghost2d_global.f:7) fills the four ghost
  // strips only: the tri-corner blocks stay zero, which is what marks their
  // zones inactive (210…:300-330)
  for (int s = 0; s < 6; ++s)
    for (int c = 0; c < 3; ++c)
      for (int j = 0; j < d.nv; ++j)
        for (int i = 0; i < d.nv; ++i) {
          const bool ci = i < d.g || i > d.g + d.N, cj = j < d.g || j > d.g + d.N;
          if (ci && cj) lat[(size_t)i + (size_t)d.nv * (j + (size_t)d.nv * (c + 3 * (size_t)s))] = 0.0;
        }
  if ((rc = cs_build_radial_mesh(cfg, rbh.data()))) return rc;
  DevBuf B;
  double* dlat = B.get<double>((int64_t)lat.size());
  double* drb = B.get<double>(d.nvr);
  double* vert = B.get<double>(18 * d.vs);
  int32_t* mask = B.get<int32_t>(6 * d.zs);
  double* scale = B.get<double>(6 * d.zs);
  FaceOut F[3];
  for (int f = 0; f < 3; ++f) {
    const int64_t S = 6 * (f == 2 ? d.zfs : d.xs);
    F[f].pts = B.get<double>(15 * S);
    F[f].wts = B.get<double>(5 * S);
    F[f].nrm = B.get<double>((f == 2 ? 15 : 3) * S);
    F[f].area = B.get<double>(S);
    F[f].cen = B.get<double>(3 * S);
    F[f].bwts = B.get<double>(5 * S);
  }
  VolOut Vo;
  Vo.pts = B.get<double>(144 * d.zs);
  Vo.wts = B.get<double>(48 * d.zs);
  Vo.tan = B.get<double>(432 * d.zs);
  Vo.vol = B.get<double>(6 * d.zs);
  Vo.cen = B.get<double>(18 * d.zs);
  double* mom = B.get<double>(24 * d.zs);
  for (void* q : B.p)
    if (!q) return CS_E_NOMEM;
  hipStream_t st;
  if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) return CS_E_HIP;
  hipEvent_t ev[5];
  for (int e = 0; e < 5; ++e) hipEventCreate(&ev[e]);
  bool ok = hipMemcpyAsync(dlat, lat.data(), sizeof(double) * lat.size(), hipMemcpyHostToDevice, st) == hipSuccess &&
            hipMemcpyAsync(drb, rbh.data(), sizeof(double) * rbh.size(), hipMemcpyHostToDevice, st) == hipSuccess;
  if (ok) {
    hipEventRecord(ev[0], st);
    k_vert<<<nb(6 * d.vs), 256, 0, st>>>(d, dlat, drb, vert);
    k_mask<<<nb(6 * d.zs), 256, 0, st>>>(d, dlat, mask);
    hipEventRecord(ev[1], st);
    for (int f = 0; f < 3; ++f) k_faces<<<nb(6 * (f == 2 ? d.zfs : d.xs)), 256, 0, st>>>(d, vert, f, F[f]);
    k_scale<<<nb(6 * d.zs), 256, 0, st>>>(d, vert, scale);
    hipEventRecord(ev[2], st);
    k_volume<<<nb(6 * d.zs), 256, 0, st>>>(d, vert, Vo);
    hipEventRecord(ev[3], st);
    k_moments<<<nb(6 * d.zs), 256, 0, st>>>(d, mask, Vo.pts, Vo.wts, Vo.cen, scale, mom);
    hipEventRecord(ev[4], st);
    ok = hipGetLastError() == hipSuccess;
  }
  if (ok && out) {
    const size_t zb = sizeof(double) * 6 * d.zs;
    rc = copy_out(out->vert, vert, sizeof(double) * 18 * d.vs, st);
    if (!rc) rc = copy_out(out->mask, mask, sizeof(int32_t) * 6 * d.zs, st);
    const cs_face_quad* fo[3] = {&out->xf, &out->yf, &out->zf};
    const cs_face_quad* bo[3] = {&out->xb, &out->yb, &out->zb};
    for (int f = 0; f < 3 && !rc; ++f) {
      const size_t fb = sizeof(double) * 6 * (f == 2 ? d.zfs : d.xs);
      rc = copy_out(fo[f]->pts, F[f].pts, 15 * fb, st);
      if (!rc) rc = copy_out(fo[f]->wts, F[f].wts, 5 * fb, st);
      if (!rc) rc = copy_out(fo[f]->normal, F[f].nrm, (f == 2 ? 15 : 3) * fb, st);
      if (!rc) rc = copy_out(fo[f]->area, F[f].area, fb, st);
      if (!rc) rc = copy_out(fo[f]->centroid, F[f].cen, 3 * fb, st);
      if (!rc) rc = copy_out(bo[f]->pts, F[f].pts, 15 * fb, st);
      if (!rc) rc = copy_out(bo[f]->wts, F[f].bwts, 5 * fb, st);
      if (!rc) rc = copy_out(bo[f]->normal, F[f].nrm, (f == 2 ? 15 : 3) * fb, st);
    }
    if (!rc) rc = copy_out(out->scale, scale, zb, st);
    if (!rc) rc = copy_out(out->v_pts, Vo.pts, 24 * zb, st);
    if (!rc) rc = copy_out(out->v_wts, Vo.wts, 8 * zb, st);
    if (!rc) rc = copy_out(out->v_tan, Vo.tan, 72 * zb, st);
    if (!rc) rc = copy_out(out->v_vol, Vo.vol, zb, st);
    if (!rc) rc = copy_out(out->v_cen, Vo.cen, 3 * zb, st);
    if (!rc) rc = copy_out(out->moments, mom, 4 * zb, st);
    ok = !rc;
  }
  ok = (hipStreamSynchronize(st) == hipSuccess) && ok;
  if (ok && kernel_ms)
    for (int e = 0; e < 4; ++e) hipEventElapsedTime(&kernel_ms[e], ev[e], ev[e + 1]);
  for (int e = 0; e < 5; ++e) hipEventDestroy(ev[e]);
  hipStreamDestroy(st);
  return ok ? CS_OK : CS_E_HIP;
}
