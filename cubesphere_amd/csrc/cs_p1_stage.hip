// cs_p1_stage.hip -- the reference's second-order P1 scheme as the RK stage of
// a cs_handle (recon CS_RECON_P1ADP): the multi-D WENO of the primitives and
// the entropy with the flattener and the pressure-positivity limiter, the
// divergence-preserving constrained least squares for the zone B
// (LIMIT_ALL_SLOPES_P1, 410limit_all_slopes_p1.f:9 with 440...f:6), the face
// fluxes at the face centroids (MAKE_FACIAL_FLUXES, 450make_facial_fluxes.f:5,
// with the HLL of OneD_RS_MHD_Ptwise or HLLD), the edge fields
// (MAKE_EDGE_EFIELDS_P1, 460make_edge_fields_p1.f:5), the update with the
// arc-length constrained transport (UPDATE_IMEX_RK_STEP_P1, 484...f:7),
// CLEANUP_AFTER_UPDATE (485...f:4) and EVALUATE_TIMESTEP (490...f:7).
//
// Differences from the per-patch entry points of cs_slopes_p1.hip, which keep
// the reference's routines argument for argument:
//  * all six sectors in each launch (the sector in blockIdx.y, or packed in a
//    list entry), no host synchronisation anywhere: the positivity limiter's
//    loop-order replay is a bounded number of device-side sweeps that return
//    at once when there is nothing to replay, so a whole step is one graph;
//  * the face geometry comes from separable tables (cs_p1geo.h) instead of
//    ~90 stored words per zone; the zone centroid, volume and scale length
//    are stored (5 words);
//  * the state is the handle's (mode 1 only); the slopes of the 8 primitives
//    and the entropy live in V (4 modes x 9 fields); the WENO of the zone E
//    (410...f:5437-7100) is not formed: at second order nothing reads it;
//  * the face fluxes are stored once per face as flux x area (the resolved
//    flux of the HLL or HLLD solve in the face frame, rotated back), so both
//    zones of a face take the same flux; the handle averages the two copies
//    of each panel-edge face and edge before the update (the reference's
//    defect correction in its before-the-update form) and the update gathers
//    the six faces of each owned zone.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "../../include/cubesphere_hip.h"
#include "cs_internal.h"
#include "cs_p1_common.h"
#include "cs_p1_stage.h"
#include "cs_p1geo.h"
#include "cs_problems.h"
#include "cs_p1s_host.h"
#include "cs_physics.h"

using namespace csp;
using namespace csp1;

namespace cs {
void set_error(const char* fmt, ...);
}

namespace p1s {

constexpr int kSweeps = 12;          // bound on the positivity replay sweeps per stage
constexpr int kListBlocks = 1024;    // grid of the list kernels (grid-stride)
constexpr int kCflBlocks = 1024;     // grid of the time-step reduction
// counters: [0] zones whose slopes pass 0 removed, [1] list length,
// [2] settled, [3 .. 3 + kSweeps) flag changes per sweep
constexpr int kCtr = 3 + kSweeps;

struct Dev {
  int N, NR, g, nz, nzr, nf, nfr;
  int64_t NZ, NFC, NZ2;
  csg1::P1Geo G;
  const double* zg;        // [s][zone / 64][5][64]: centroid (3), volume, scale length
  const int32_t* mask;     // [nz * nz] active zone columns (the same in every sector)
  const int32_t* stencil;  // [24][nz * nz] (STENCIL_INDEXING_TRICORNERS)
  double* V;               // [s][zone / 64][4 x 9][64]: mode m of the 8 primitives and the entropy
  int64_t NZB;             // zone blocks of 64 per sector
  double *flat, *work, *ms;  // [s][zone]
  double* Bq;              // [s][fam][5][face]: normal B at the B-field points
  uint8_t *fa, *prev, *next, *ever, *inlist;  // [s][zone]
  uint32_t* list;          // packed s * NZ + zone
  unsigned int* ctr;
  double gamma, rhofloor, prsfloor, smallnum, flatten_coef, prs_flatten_coef, grav;
  int wall, riemann;

  __device__ __forceinline__ int64_t zi(int i, int j, int k) const {
    return (int64_t)(i - 1 + g) + (int64_t)nz * ((j - 1 + g) + (int64_t)nz * (k - 1 + g));
  }
  __device__ __forceinline__ int64_t fci(int i, int j, int k) const {
    return (int64_t)(i + g) + (int64_t)nf * ((j + g) + (int64_t)nf * (k + g));
  }
  __device__ __forceinline__ int64_t col(int i, int j) const { return (int64_t)(i - 1 + g) + (int64_t)nz * (j - 1 + g); }
  __device__ __forceinline__ bool active(int i, int j) const { return mask[col(i, j)] == 1; }
  // stencil entry n (R L U D RR LL UU DD RU LU LD RD) of zone column (i, j),
  // clamped to the zone range
  __device__ __forceinline__ void nbr(int i, int j, int n, int& ni, int& nj) const {
    const int64_t c = col(i, j);
    ni = min(max(stencil[c + NZ2 * (2 * n)], 1 - g), N + g);
    nj = min(max(stencil[c + NZ2 * (2 * n + 1)], 1 - g), N + g);
  }
  // Zone-blocked layouts: the words of 64 consecutive zones lie together
  // (a block's field c is 64 contiguous doubles), so a zone's slopes and
  // geometry share a few pages and a neighbour's words arrive together,
  // while a wave reading one field of 64 consecutive zones stays coalesced.
  __device__ __forceinline__ double& v(int m, int s, int c, int64_t z) const {
    return V[(((int64_t)s * NZB + (z >> 6)) * 36 + m * 9 + c) * 64 + (z & 63)];
  }
  __device__ __forceinline__ double zgw(int s, int64_t z, int q) const {
    return zg[(((int64_t)s * NZB + (z >> 6)) * 5 + q) * 64 + (z & 63)];
  }
  __device__ __forceinline__ double cen(int s, int64_t z, int c) const { return zgw(s, z, c); }
  __device__ __forceinline__ double vol(int s, int64_t z) const { return zgw(s, z, 3); }
  __device__ __forceinline__ double obl(int s, int64_t z) const { return 1.0 / fmax(zgw(s, z, 4), smallnum); }
  __device__ __forceinline__ double scale(int s, int64_t z) const { return zgw(s, z, 4); }
  __device__ __forceinline__ double* bq(int s, int fam, int p) const {
    return Bq + (((int64_t)s * 3 + fam) * 5 + p) * NFC;
  }
  __device__ __forceinline__ void zone_of(int64_t z, int& i, int& j, int& k) const {
    i = (int)(z % nz) + 1 - g;
    j = (int)((z / nz) % nz) + 1 - g;
    k = (int)(z / ((int64_t)nz * nz)) + 1 - g;
  }
  __device__ __forceinline__ Gas<double> gas() const {
    return Gas<double>{gamma, gamma - 1.0, rhofloor, prsfloor, 1.0 / (gamma - 1.0)};
  }
};

struct Buf6 {
  double *U, *ent, *Bx, *By, *Bz;
  __device__ __forceinline__ double& u(const Dev& D, int s, int c, int64_t z) const {
    return U[((int64_t)s * 8 + c) * D.NZ + z];
  }
  __device__ __forceinline__ double& e(const Dev& D, int s, int64_t z) const { return ent[(int64_t)s * D.NZ + z]; }
  __device__ __forceinline__ double* b(int fam) const { return fam == 0 ? Bx : (fam == 1 ? By : Bz); }
};

// a box of zone, face or edge indices of one sector
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
__host__ __device__ static inline Box make_box(int i0, int i1, int j0, int j1, int k0, int k1) {
  return Box{i0, j0, k0, i1 - i0 + 1, j1 - j0 + 1, k1 - k0 + 1};
}

__device__ __forceinline__ int64_t tid() { return (int64_t)blockIdx.x * blockDim.x + threadIdx.x; }

// XCD-aware workgroup order.  The chip hands workgroup b (row-major over the
// grid) to XCD b % 8, each XCD with its own L2; renumbering the workgroups
// so that every XCD receives one contiguous run of the logical grid keeps
// the workgroups that read each other's zones (the next row, the next
// plane) on one L2 and close in time.
constexpr int kXcds = 8;
struct LBlock {
  int64_t x;  // logical blockIdx.x
  int y;      // logical blockIdx.y
};
__device__ __forceinline__ LBlock lblock() {
  const int64_t G = (int64_t)gridDim.x * gridDim.y, b = blockIdx.x + (int64_t)blockIdx.y * gridDim.x;
  const int64_t x = b % kXcds, q = G / kXcds, r = G % kXcds;
  const int64_t lb = x * q + (x < r ? x : r) + b / kXcds;
  return LBlock{lb % gridDim.x, (int)(lb / gridDim.x)};
}
__device__ __forceinline__ int64_t ltid(const LBlock& L) { return L.x * blockDim.x + threadIdx.x; }

// Zone work items in column tiles: workgroup w takes a TI x TJ tile of
// columns at one level, and consecutive workgroups (one XCD's run of lblock
// numbers) step through the levels of one tile before the next tile, so a
// zone's neighbours in the level below and above, and in the tile's other
// rows, were read moments before through the same L2.
template <int TI, int TJ>
__device__ __forceinline__ bool tile_decode(const Box& b, int64_t w, int lane, int& i, int& j, int& k) {
  const int nti = (b.ni + TI - 1) / TI, ntj = (b.nj + TJ - 1) / TJ;
  const int64_t tile = w / b.nk;
  if (tile >= (int64_t)nti * ntj) return false;
  const int ii = (int)(tile % nti) * TI + lane % TI, jj = (int)(tile / nti) * TJ + lane / TI;
  if (ii >= b.ni || jj >= b.nj) return false;
  i = b.i0 + ii;
  j = b.j0 + jj;
  k = b.k0 + (int)(w % b.nk);
  return true;
}
template <int TI, int TJ>
static unsigned tile_blocks(const Box& b) {
  return (unsigned)((int64_t)((b.ni + TI - 1) / TI) * ((b.nj + TJ - 1) / TJ) * b.nk);
}

__device__ __forceinline__ bool in_weno_range(const Dev& D, int i, int j, int k) {
  return i >= -2 && i <= D.N + 3 && j >= -2 && j <= D.N + 3 && k >= -2 && k <= D.NR + 3;
}
__device__ __forceinline__ bool in_flat_range(const Dev& D, int i, int j, int k) {
  const int lo = 2 - D.g, hx = D.N + D.g - 1, hr = D.NR + D.g - 1;
  return i >= lo && i <= hx && j >= lo && j <= hx && k >= lo && k <= hr;
}
__device__ __forceinline__ void six(const Dev& D, int i, int j, int k, int64_t* z6) {
  int a, b;
  D.nbr(i, j, 0, a, b);
  z6[0] = D.zi(a, b, k);
  D.nbr(i, j, 1, a, b);
  z6[1] = D.zi(a, b, k);
  D.nbr(i, j, 2, a, b);
  z6[2] = D.zi(a, b, k);
  D.nbr(i, j, 3, a, b);
  z6[3] = D.zi(a, b, k);
  z6[4] = D.zi(i, j, k + 1);
  z6[5] = D.zi(i, j, k - 1);
}

// ---------------------------------------------------------------------------
// set-up: zone geometry (once)
__global__ void k_zone_geo(Dev D, Box b, double* zg) {
  int i, j, k;
  if (!b.decode(tid(), i, j, k)) return;
  const int s = blockIdx.y;
  const int64_t z = D.zi(i, j, k);
  double c[3], v, l;
  csg1::zone_geometry(D.G, s, i, j, k, c, &v, &l);
  // the inactive tri-corner zones get smallnum (410...f step 2, 485...f)
  if (!D.active(i, j)) l = D.smallnum;
  double* o = zg + ((int64_t)s * D.NZB + (z >> 6)) * 5 * 64 + (z & 63);
  for (int q = 0; q < 3; ++q) o[q * 64] = c[q];
  o[3 * 64] = v;
  o[4 * 64] = l;
}

// ---------------------------------------------------------------------------
// steps 2 and 5 of 410: inactive zones made safe, CONS_TO_PRIM, magnetosonic
// speed; the slopes of zones outside the WENO range or in inactive columns
// are cleared (the WENO writes the others)
__global__ void __launch_bounds__(256) k_prim(Dev D, Box b, Buf6 in) {
  int i, j, k;
  const LBlock wg = lblock();
  if (!b.decode(ltid(wg), i, j, k)) return;
  const int s = wg.y;
  const int64_t z = D.zi(i, j, k);
  const bool act = D.active(i, j);
  double U[8], Vv[8];
  if (!act) {
#pragma unroll
    for (int c = 0; c < 8; ++c) U[c] = 0.0;
    U[0] = D.rhofloor;
    U[4] = D.prsfloor / (D.gamma - 1.0);
  } else {
#pragma unroll
    for (int c = 0; c < 8; ++c) U[c] = in.u(D, s, c, z);
  }
  double e = in.e(D, s, z);
  const Gas<double> gs = D.gas();
  Vv[4] = 0.0;
  cons_to_prim(gs, U, Vv, e);
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    in.u(D, s, c, z) = U[c];
    D.v(0, s, c, z) = Vv[c];
  }
  in.e(D, s, z) = e;
  D.v(0, s, 8, z) = e;
  const double b2 = (Vv[5] * Vv[5] + Vv[6] * Vv[6] + Vv[7] * Vv[7]) * C<double>::ob4pi;
  D.ms[(int64_t)s * D.NZ + z] = sqrt((D.gamma * Vv[4] + b2) / Vv[0]);
  if (!act || !in_weno_range(D, i, j, k)) {
#pragma unroll
    for (int m = 1; m < 4; ++m)
#pragma unroll
      for (int c = 0; c < 9; ++c) D.v(m, s, c, z) = 0.0;
  }
}

// step 2 (faces): face B between two inactive zones cleared; blockIdx.y = fam + 3 s
__global__ void __launch_bounds__(256) k_face_zero(Dev D, Buf6 in) {
  const int fam = blockIdx.y % 3, s = blockIdx.y / 3;
  const int64_t t = tid();
  if (t >= D.NFC) return;
  const int n = D.nf;
  const int i = (int)(t % n) - D.g, j = (int)((t / n) % n) - D.g, k = (int)(t / ((int64_t)n * n)) - D.g;
  const int lo = 1 - D.g, hi = D.N + D.g, hr = D.NR + D.g;
  double* B = in.b(fam) + (int64_t)s * D.NFC;
  if (fam == 0) {
    if (j < lo || k < lo || k > hr) return;
    if (!D.active(max(i, lo), j) && !D.active(min(i + 1, hi), j)) B[t] = 0.0;
  } else if (fam == 1) {
    if (i < lo || k < lo || k > hr) return;
    if (!D.active(i, max(j, lo)) && !D.active(i, min(j + 1, hi))) B[t] = 0.0;
  } else {
    if (i < lo || j < lo) return;
    if (!D.active(i, j)) B[t] = 0.0;
  }
}

// step 3: face B synchronised at the four tri-corners of every sector
__global__ void __launch_bounds__(256) k_tricorner(Dev D, Buf6 in) {
  const int64_t t = tid();
  const int g = D.g, nzr = D.nzr, s = blockIdx.y;
  if (t >= 4 * (int64_t)g * nzr) return;
  const int corner = (int)(t / ((int64_t)g * nzr)), ii = (int)((t / nzr) % g), k = (int)(t % nzr) + 1 - g;
  const bool left = corner == 0 || corner == 3, down = corner < 2;
  const int i = left ? 1 - g + ii : D.N + 1 + ii, jj = down ? 1 : D.N;
  int a, b;
  D.nbr(i, jj, down ? 3 : 2, a, b);
  double& by = in.By[(int64_t)s * D.NFC + D.fci(i, down ? jj - 1 : jj, k)];
  double& bx = in.Bx[(int64_t)s * D.NFC + D.fci(left ? a - 1 : a, b, k)];
  if (corner == 0 || corner == 2) {
    const double h = 0.5 * (by - bx);
    by = h;
    bx = -h;
  } else {
    const double h = 0.5 * (by + bx);
    by = h;
    bx = h;
  }
}

// ---------------------------------------------------------------------------
// step 6: the velocity-divergence flattener (a), its pressure-ratio
// extension (b) and its spread to the six neighbours (c)
__global__ void __launch_bounds__(256) k_flat_a(Dev D, Box b) {
  int i, j, k;
  const LBlock wg = lblock();
  if (!tile_decode<32, 8>(b, wg.x, threadIdx.x, i, j, k)) return;
  const int s = wg.y;
  const int64_t z = D.zi(i, j, k), so = (int64_t)s * D.NZ;
  if (!in_flat_range(D, i, j, k)) {
    D.flat[so + z] = 0.0;
    return;
  }
  int64_t z6[6];
  six(D, i, j, k, z6);
  const double obl = D.obl(s, z);
  double wmin = D.ms[so + z];
#pragma unroll
  for (int q = 0; q < 6; ++q) wmin = fmin(wmin, D.ms[so + z6[q]]);
  // the unit radial direction of the top z-face centroid serves both z-faces
  double nzv[3];
  {
    csg1::zcen_dir(D.G, s, i, j, nzv);
    const double r = D.G.r(k);
#pragma unroll
    for (int c = 0; c < 3; ++c) nzv[c] = r * nzv[c];
    const double t = 1.0 / fmax(sqrt(nzv[0] * nzv[0] + nzv[1] * nzv[1] + nzv[2] * nzv[2]), D.smallnum);
#pragma unroll
    for (int c = 0; c < 3; ++c) nzv[c] = nzv[c] * t;
  }
  double dv = 0.0;
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    const bool up = (q % 2) == 0;
    const int64_t zo = up ? z6[q] : z, zn = up ? z : z6[q];
    double r[3], n[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) r[c] = (D.cen(s, zo, c) - D.cen(s, zn, c)) * obl;
    const double a11 = 0.5 / fmax(sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]), D.smallnum);
    if (q < 2) csg1::normal_x(D.G, s, q == 0 ? i : i - 1, j, n);
    else if (q < 4) csg1::normal_y(D.G, s, i, q == 2 ? j : j - 1, n);
    else {
#pragma unroll
      for (int c = 0; c < 3; ++c) n[c] = nzv[c];
    }
    dv = dv + n[0] * a11 * (D.v(0, s, 1, zo) - D.v(0, s, 1, zn)) + n[1] * a11 * (D.v(0, s, 2, zo) - D.v(0, s, 2, zn)) +
         n[2] * a11 * (D.v(0, s, 3, zo) - D.v(0, s, 3, zn));
  }
  const double a13 = fabs(dv) / (D.flatten_coef * wmin) - 1.0;
  double f = a13 <= 0.0 ? 0.0 : fmin(1.0, a13);
  if (!D.active(i, j)) f = 0.0;
  D.flat[so + z] = f;
}
__global__ void __launch_bounds__(256) k_flat_b(Dev D, Box b) {
  int i, j, k;
  const LBlock wg = lblock();
  if (!b.decode(ltid(wg), i, j, k)) return;
  const int s = wg.y;
  const int64_t z = D.zi(i, j, k), so = (int64_t)s * D.NZ;
  double f = D.flat[so + z];
  if (in_flat_range(D, i, j, k)) {
    int64_t z6[6];
    six(D, i, j, k, z6);
    const double p0 = D.prs_flatten_coef * D.v(0, s, 4, z);
#pragma unroll
    for (int q = 0; q < 6; ++q)
      if (D.v(0, s, 4, z6[q]) > p0) f = fmax(f, D.flat[so + z6[q]]);
  }
  D.work[so + z] = f;
}
__global__ void __launch_bounds__(256) k_flat_c(Dev D, Box b) {
  int i, j, k;
  const LBlock wg = lblock();
  if (!b.decode(ltid(wg), i, j, k)) return;
  const int s = wg.y;
  const int64_t z = D.zi(i, j, k), so = (int64_t)s * D.NZ;
  double f = D.work[so + z];
  if (in_flat_range(D, i, j, k)) {
    int64_t z6[6];
    six(D, i, j, k, z6);
#pragma unroll
    for (int q = 0; q < 6; ++q) f = fmax(f, D.work[so + z6[q]]);
  }
  D.flat[so + z] = f;
}

// ---------------------------------------------------------------------------
// step 7.a: WENO of three of the nine fields (group c0 = 0, 3, 6) of one
// zone in one pass over the 8 stencils (stencil s: the x-neighbour R for s in
// {0,1,4,5}, else L; the y-neighbour U for s < 4, else D; the z-neighbour
// top for even s, else bottom); per stencil its 3 x 3 inverse (formed once
// for the three fields), each field's gradient X_s and weight
// w_s = 1 / (|X_s|^2 + smallnum)^4 as running sums; the central candidate
// (the mean of the 8 gradients, weight 100) at the end; times 1 - flatten.
__device__ __forceinline__ void weno3(const Dev& D, int s, int i, int j, int k, int64_t z, int c0) {
  if (!D.active(i, j)) return;
  int64_t z6[6];
  six(D, i, j, k, z6);
  const double obl = D.obl(s, z);
  double off[6][3];
#pragma unroll
  for (int q = 0; q < 6; ++q)
#pragma unroll
    for (int c = 0; c < 3; ++c) off[q][c] = (D.cen(s, z6[q], c) - D.cen(s, z, c)) * obl;
  const double smallnum = D.smallnum, omf = 1.0 - D.flat[(int64_t)s * D.NZ + z];
  double dR[3], dL[3], dU[3], dD[3], dT[3], dB[3], num[3][3], den[3], sx[3][3];
#pragma unroll
  for (int f = 0; f < 3; ++f) {
    const double f0 = D.v(0, s, c0 + f, z);
    dR[f] = D.v(0, s, c0 + f, z6[0]) - f0;
    dL[f] = D.v(0, s, c0 + f, z6[1]) - f0;
    dU[f] = D.v(0, s, c0 + f, z6[2]) - f0;
    dD[f] = D.v(0, s, c0 + f, z6[3]) - f0;
    dT[f] = D.v(0, s, c0 + f, z6[4]) - f0;
    dB[f] = D.v(0, s, c0 + f, z6[5]) - f0;
    den[f] = 0.0;
#pragma unroll
    for (int m = 0; m < 3; ++m) num[f][m] = sx[f][m] = 0.0;
  }
#pragma unroll 1
  for (int st = 0; st < 8; ++st) {
    const bool bx = (st >> 1) & 1, by = (st >> 2) & 1, bz = st & 1;
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
      den[f] = st == 0 ? w : den[f] + w;
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        num[f][m] += w * X[m];
        sx[f][m] += X[m];
      }
    }
  }
#pragma unroll
  for (int f = 0; f < 3; ++f) {
    double X8[3];
#pragma unroll
    for (int m = 0; m < 3; ++m) X8[m] = 0.125 * sx[f][m];
    double tt = X8[0] * X8[0] + X8[1] * X8[1] + X8[2] * X8[2] + smallnum;
    tt = tt * tt;
    const double w8 = 100.0 / (tt * tt);
    const double os = 1.0 / (den[f] + w8);
#pragma unroll
    for (int m = 0; m < 3; ++m) D.v(m + 1, s, c0 + f, z) = (num[f][m] + w8 * X8[m]) * os * omf;
  }
}
// every zone of the WENO range, one zone per thread: its eight stencil
// inverses once, in registers, then the nine fields in turn over the eight
// stencils, with weno3's arithmetic
__global__ void __launch_bounds__(128) k_weno(Dev D, Box b) {
  int i, j, k;
  const LBlock wg = lblock();
  if (!b.decode(ltid(wg), i, j, k) || !D.active(i, j)) return;
  const int s = wg.y;
  const int64_t z = D.zi(i, j, k);
  int64_t z6[6];
  six(D, i, j, k, z6);
  double Ainv[8][9];
  {
    const double obl = D.obl(s, z);
    double off[6][3];
#pragma unroll
    for (int q = 0; q < 6; ++q)
#pragma unroll
      for (int c = 0; c < 3; ++c) off[q][c] = (D.cen(s, z6[q], c) - D.cen(s, z, c)) * obl;
#pragma unroll
    for (int st = 0; st < 8; ++st) {
      const bool bx = (st >> 1) & 1, by = (st >> 2) & 1, bz = st & 1;
      double m3[3][3], A[3][3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        m3[0][c] = bx ? off[1][c] : off[0][c];
        m3[1][c] = by ? off[3][c] : off[2][c];
        m3[2][c] = bz ? off[5][c] : off[4][c];
      }
      inv3(m3, A);
#pragma unroll
      for (int q = 0; q < 9; ++q) Ainv[st][q] = A[q / 3][q % 3];
    }
  }
  const double smallnum = D.smallnum, omf = 1.0 - D.flat[(int64_t)s * D.NZ + z];
#pragma unroll 1
  for (int f = 0; f < 9; ++f) {
    const double f0 = D.v(0, s, f, z);
    const double dR = D.v(0, s, f, z6[0]) - f0, dL = D.v(0, s, f, z6[1]) - f0, dU = D.v(0, s, f, z6[2]) - f0,
                 dD = D.v(0, s, f, z6[3]) - f0, dT = D.v(0, s, f, z6[4]) - f0, dB = D.v(0, s, f, z6[5]) - f0;
    double num[3] = {0.0, 0.0, 0.0}, sx[3] = {0.0, 0.0, 0.0}, den = 0.0;
#pragma unroll
    for (int st = 0; st < 8; ++st) {
      const bool bx = (st >> 1) & 1, by = (st >> 2) & 1, bz = st & 1;
      const double r0 = bx ? dL : dR, r1 = by ? dD : dU, r2 = bz ? dB : dT;
      double X[3];
#pragma unroll
      for (int m = 0; m < 3; ++m) X[m] = Ainv[st][3 * m] * r0 + Ainv[st][3 * m + 1] * r1 + Ainv[st][3 * m + 2] * r2;
      double tt = X[0] * X[0] + X[1] * X[1] + X[2] * X[2] + smallnum;
      tt = tt * tt;
      const double w = 1.0 / (tt * tt);
      den = st == 0 ? w : den + w;
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        num[m] += w * X[m];
        sx[m] += X[m];
      }
    }
    double X8[3];
#pragma unroll
    for (int m = 0; m < 3; ++m) X8[m] = 0.125 * sx[m];
    double tt = X8[0] * X8[0] + X8[1] * X8[1] + X8[2] * X8[2] + smallnum;
    tt = tt * tt;
    const double w8 = 100.0 / (tt * tt);
    const double os = 1.0 / (den + w8);
#pragma unroll
    for (int m = 0; m < 3; ++m) D.v(m + 1, s, f, z) = (num[m] + w8 * X8[m]) * os * omf;
  }
}
// the zones of the replay list (grid-stride)
__global__ void __launch_bounds__(128) k_weno_list(Dev D) {
  if (D.ctr[0] == 0 || D.ctr[2]) return;
  const int64_t n = D.ctr[1];
  for (int64_t t = tid(); t < 3 * n; t += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t p = D.list[t % n];
    const int s = (int)(p / D.NZ);
    const int64_t z = p - (int64_t)s * D.NZ;
    int i, j, k;
    D.zone_of(z, i, j, k);
    weno3(D, s, i, j, k, z, 3 * (int)(t / n));
  }
}

// ---------------------------------------------------------------------------
// step 7.b: ENFORCE_PRESSURE_POSITIVITY_PRIM (440...f:6-830).  pass < 0: the
// first pass over the range with the incoming entropy means; flags of every
// zone set, zones whose slopes were removed counted.  pass >= 0: replay sweep
// `pass` on the list, taking the entropy means of earlier zones whose prev
// flag is set as zero (440...f:553-560), flag changes counted.
__device__ __forceinline__ void columns9(const Dev& D, int i, int j, int64_t* c9) {
  c9[0] = D.zi(i, j, 0);
  const int dirs[8] = {0, 1, 2, 3, 8, 9, 10, 11};
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    int ai, bj;
    D.nbr(i, j, dirs[q], ai, bj);
    c9[q + 1] = D.zi(ai, bj, 0);
  }
}

__device__ __forceinline__ void positivity(const Dev& D, int s, int i, int j, int k, int64_t z, int pass) {
  const int64_t so = (int64_t)s * D.NZ;
  int64_t c9[9];
  columns9(D, i, j, c9);
  const int64_t ks = (int64_t)D.nz * D.nz;
  double rmin = 1e300, rmax = -1e300, pmin = 1e300, pmax = -1e300, emin = 1e300, emax = -1e300;
#pragma unroll 1
  for (int q = 0; q < 9; ++q)
#pragma unroll
    for (int dk = -1; dk <= 1; ++dk) {
      const int64_t zz = c9[q] + ks * (k + dk);
      const double r = D.v(0, s, 0, zz), p = D.v(0, s, 4, zz);
      double e = D.v(0, s, 8, zz);
      if (pass >= 0 && zz < z && D.prev[so + zz]) e = 0.0;
      rmin = fmin(rmin, r);
      rmax = fmax(rmax, r);
      pmin = fmin(pmin, p);
      pmax = fmax(pmax, p);
      emin = fmin(emin, e);
      emax = fmax(emax, e);
    }
  const double expandfac = 0.2;
  const double ff = fmax(fmin((D.flat[so + z] - 0.1) / 0.4, 1.0), 0.0), omff = 1.0 - ff;
  rmin = fmax(rmin * (1.0 - expandfac * omff), D.rhofloor);
  rmax = fmax(rmax * (1.0 + expandfac * omff), D.rhofloor);
  pmin = fmax(pmin * (1.0 - expandfac * omff), D.prsfloor);
  pmax = fmax(pmax * (1.0 + expandfac * omff), D.prsfloor);
  emin = emin * (1.0 - expandfac * omff);
  emax = emax * (1.0 + expandfac * omff);
  const double r0 = D.v(0, s, 0, z), p0 = D.v(0, s, 4, z), e0 = D.v(0, s, 8, z);
  double rm[3], pm[3], em[3];
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    rm[m] = D.v(m + 1, s, 0, z);
    pm[m] = D.v(m + 1, s, 4, z);
    em[m] = D.v(m + 1, s, 8, z);
  }
  const double obl = D.obl(s, z);
  double c0[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) c0[c] = D.cen(s, z, c);
  // the 27 nodes: vertices, face centroids, edge midpoints, the centroid
  double vx[8][3], fcx[6][3];
#pragma unroll
  for (int n = 0; n < 8; ++n) D.G.vert(s, (n & 1) ? i : i - 1, (n & 2) ? j : j - 1, (n & 4) ? k : k - 1, vx[n]);
  {
    csg1::FaceG F;
    csg1::face_x(D.G, s, i, j, k, F);
#pragma unroll
    for (int c = 0; c < 3; ++c) fcx[0][c] = F.cen[c];
    csg1::face_x(D.G, s, i - 1, j, k, F);
#pragma unroll
    for (int c = 0; c < 3; ++c) fcx[1][c] = F.cen[c];
    csg1::face_y(D.G, s, i, j, k, F);
#pragma unroll
    for (int c = 0; c < 3; ++c) fcx[2][c] = F.cen[c];
    csg1::face_y(D.G, s, i, j - 1, k, F);
#pragma unroll
    for (int c = 0; c < 3; ++c) fcx[3][c] = F.cen[c];
    csg1::face_z(D.G, s, i, j, k, F);
#pragma unroll
    for (int c = 0; c < 3; ++c) fcx[4][c] = F.cen[c];
    csg1::face_z(D.G, s, i, j, k - 1, F);
#pragma unroll
    for (int c = 0; c < 3; ++c) fcx[5][c] = F.cen[c];
  }
  const double ratio = 0.85769589859089412;  // 0.1^(1/15)
  auto passes = [&](int n) -> bool {
    double r[3], p[3], e[3];
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      r[m] = rm[m];
      p[m] = pm[m];
      e[m] = em[m];
    }
    for (int q = 0; q < n; ++q)
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        r[m] = ratio * r[m];
        p[m] = ratio * p[m];
        e[m] = ratio * e[m];
      }
    bool ok = true;
    auto node = [&](const double* x0) {
      double x[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) x[c] = (x0[c] - c0[c]) * obl;
      const double rn = r0 + r[0] * x[0] + r[1] * x[1] + r[2] * x[2];
      const double pn = p0 + p[0] * x[0] + p[1] * x[1] + p[2] * x[2];
      const double en = e0 + e[0] * x[0] + e[1] * x[1] + e[2] * x[2];
      if (rn < rmin || rn > rmax || pn < pmin || pn > pmax || en < emin || en > emax) ok = false;
    };
#pragma unroll
    for (int v = 0; v < 8; ++v) node(vx[v]);
#pragma unroll
    for (int f = 0; f < 6; ++f) node(fcx[f]);
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
    return ok;
  };
  // the reference checks 0, 1, ..., 14 reductions in turn and takes the first
  // that passes (removing the slopes when none does); the zone's mean lies in
  // the bounds, so in exact arithmetic passing is monotone in the count and
  // the first passing count is found by probing 0, then 14, then bisecting.
  // The count found passes and the one below it fails (both are checked);
  // only a node value that rounds across a bound at a smaller count, with
  // cancellation at least as large as the bound's own rounding, could make
  // the reference's scan stop earlier (neighbouring counts' node values
  // differ by 14 % of the slope part: each reduction scales it by 0.1^(1/15))
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
  if (nred > 0 || removed) {
    for (int c = 0; c < 9; ++c)
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        double v = D.v(m + 1, s, c, z);
        if (removed) v = 0.0;
        else
          for (int r = 0; r < nred; ++r) v = ratio * v;
        D.v(m + 1, s, c, z) = v;
      }
  }
  const uint8_t flag = removed ? 1 : 0;
  if (pass < 0) {
    D.fa[so + z] = flag;
    D.prev[so + z] = flag;
    D.next[so + z] = flag;
    D.ever[so + z] = flag;
    D.inlist[so + z] = 0;
    if (removed) atomicAdd(&D.ctr[0], 1u);
  } else {
    D.next[so + z] = flag;
    if (flag != D.prev[so + z]) atomicAdd(&D.ctr[3 + pass], 1u);
  }
}
__global__ void __launch_bounds__(256) k_pos(Dev D, Box b) {
  int i, j, k;
  const LBlock wg = lblock();
  if (!b.decode(ltid(wg), i, j, k)) return;
  positivity(D, wg.y, i, j, k, D.zi(i, j, k), -1);
}
__global__ void __launch_bounds__(256) k_pos_list(Dev D, int sweep) {
  if (D.ctr[0] == 0 || D.ctr[2]) return;
  const int64_t n = D.ctr[1];
  for (int64_t t = tid(); t < n; t += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t p = D.list[t];
    const int s = (int)(p / D.NZ);
    const int64_t z = p - (int64_t)s * D.NZ;
    int i, j, k;
    D.zone_of(z, i, j, k);
    positivity(D, s, i, j, k, z, sweep);
  }
}
__global__ void k_ctr_reset(Dev D) {
  if (threadIdx.x < kCtr) D.ctr[threadIdx.x] = 0;
}
// the zones of the range with a zone ever flagged earlier in the loop order
// among their 27 join the list (each once: the list only grows, so a zone
// that depends on a flag is recomputed in every later sweep)
__global__ void __launch_bounds__(256) k_cands(Dev D, Box b) {
  if (D.ctr[0] == 0 || D.ctr[2]) return;
  int i, j, k;
  if (!b.decode(tid(), i, j, k)) return;
  const int s = blockIdx.y;
  const int64_t z = D.zi(i, j, k), so = (int64_t)s * D.NZ, ks = (int64_t)D.nz * D.nz;
  if (D.inlist[so + z]) return;
  int64_t c9[9];
  columns9(D, i, j, c9);
  bool cand = false;
  for (int q = 0; q < 9; ++q)
#pragma unroll
    for (int dk = -1; dk <= 1; ++dk) {
      const int64_t zz = c9[q] + ks * (k + dk);
      cand = cand || (zz < z && D.ever[so + zz]);
    }
  if (cand) {
    D.inlist[so + z] = 1;
    D.list[atomicAdd(&D.ctr[1], 1u)] = (uint32_t)(so + z);
  }
}
// end of a sweep: the new flags become the previous ones; no change: settled
__global__ void __launch_bounds__(256) k_commit(Dev D, int sweep) {
  if (D.ctr[0] == 0) return;
  const int64_t n = D.ctr[1];
  for (int64_t t = tid(); t < n; t += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t p = D.list[t];
    const uint8_t f = D.next[p];
    D.prev[p] = f;
    D.ever[p] = D.ever[p] | f;
  }
  if (tid() == 0 && D.ctr[3 + sweep] == 0) D.ctr[2] = 1;
}
// after the sweeps: the entropy means of the zones that lost their slopes
// are zero (440...f:553-560); a replay that has not settled is an error
__global__ void __launch_bounds__(256) k_clear_ent(Dev D, Box b, Buf6 in, int* err) {
  if (D.ctr[0] == 0) return;
  if (tid() == 0 && blockIdx.y == 0 && !D.ctr[2]) atomicOr(err, 1);
  int i, j, k;
  if (!b.decode(tid(), i, j, k)) return;
  const int s = blockIdx.y;
  const int64_t z = D.zi(i, j, k);
  if (D.prev[(int64_t)s * D.NZ + z]) {
    in.e(D, s, z) = 0.0;
    D.v(0, s, 8, z) = 0.0;
  }
}

// ---------------------------------------------------------------------------
// step 8.a-c: normal B at the five B-field points of each face from the two
// zones' linear B, shifted so that the weighted mean is the face's Bf;
// blockIdx.y = fam + 3 s
__device__ __forceinline__ void face_points(const Dev& D, int fam, int s, int i, int j, int k, const Buf6& in) {
  const int64_t fb = D.fci(i, j, k);
  int li = i, lj = j, ri = i + (fam == 0), rj = j + (fam == 1);
  bool skip = false;
  if (fam < 2) {
    const bool al = D.active(li, lj), ar = D.active(ri, rj);
    if (!al && ar) D.nbr(ri, rj, fam == 0 ? 1 : 3, li, lj);
    else if (al && !ar) D.nbr(li, lj, fam == 0 ? 0 : 2, ri, rj);
    else if (!al && !ar) skip = true;
  } else {
    skip = !D.active(i, j);
  }
  csg1::FaceG F;
  csg1::face_of(D.G, fam, s, i, j, k, F);
  double q5[5];
  if (skip) {
#pragma unroll
    for (int q = 0; q < 5; ++q) q5[q] = 0.0;
  } else {
    const int64_t zl = D.zi(li, lj, k), zr = D.zi(ri, rj, fam == 2 ? k + 1 : k);
    const double oll = D.obl(s, zl), olr = D.obl(s, zr);
    double vl[3][4], vr[3][4];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        vl[c][m] = D.v(m, s, 5 + c, zl);
        vr[c][m] = D.v(m, s, 5 + c, zr);
      }
#pragma unroll
    for (int q = 0; q < 5; ++q) {
      const double* n = fam == 2 ? F.n[q] : F.n[0];
      double rl[3], rr[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        rl[c] = (F.p[q][c] - D.cen(s, zl, c)) * oll;
        rr[c] = (F.p[q][c] - D.cen(s, zr, c)) * olr;
      }
      double sm[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const double bl = vl[c][0] + vl[c][1] * rl[0] + vl[c][2] * rl[1] + vl[c][3] * rl[2];
        const double br = vr[c][0] + vr[c][1] * rr[0] + vr[c][2] * rr[1] + vr[c][3] * rr[2];
        sm[c] = bl + br;
      }
      q5[q] = 0.5 * (n[0] * sm[0] + n[1] * sm[1] + n[2] * sm[2]);
    }
  }
  double bd = in.b(fam)[(int64_t)s * D.NFC + fb];
#pragma unroll
  for (int q = 0; q < 5; ++q) bd = bd - F.bw[q] * q5[q];
#pragma unroll
  for (int q = 0; q < 5; ++q) D.bq(s, fam, q)[fb] = q5[q] + bd;
}
// A zone's linear B in LDS (kBWords words per lane: the 3 x 4 modes, the
// centroid, the reciprocal scale length) and its value at a point, with the
// arithmetic of face_points.
constexpr int kFpTI = 16, kFpTJ = 8;  // columns per workgroup along i and j
constexpr int kFpLevels = 16;         // levels per workgroup at most
constexpr int kBWords = 16;
__device__ __forceinline__ void b_at(const double* w, const double* p, double* b) {
  double r[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) r[c] = (p[c] - w[12 + c]) * w[15];
#pragma unroll
  for (int c = 0; c < 3; ++c) b[c] = w[4 * c] + w[4 * c + 1] * r[0] + w[4 * c + 2] * r[1] + w[4 * c + 3] * r[2];
}
__device__ __forceinline__ void lds_b(const double (*Bs)[kFpTI * kFpTJ], int slot, const double* p, double* b) {
  double w[kBWords];
#pragma unroll
  for (int q = 0; q < kBWords; ++q) w[q] = Bs[q][slot];
  b_at(w, p, b);
}
// The face-point B of the three families' ranges (x: i -3 .. N+3, j -2 ..
// N+3, k -2 .. NR+3; y likewise; z: i, j -2 .. N+3, k -3 .. NR+3), one
// column (i, j) per thread marching up in k, a workgroup holding a 16 x 8
// block of columns whose last column and row are the next blocks' first
// ones: at each level every lane copies its zone's linear B (16 words) into
// LDS once, the lanes inside the 15 x 7 computed columns take both sides of
// their x- and y-faces from LDS and the lower side of their z-face from the
// zone below's words, kept in registers from the previous level.  Tri-corner faces with one inactive side read the zones from
// memory (face_points).
//
// The faces' geometry (about 60 words of angular tables per column) is the
// same at every level, and one workgroup reads it again at each level; the
// march is split into chunks of nlev levels, and the chunks of one block of
// columns are consecutive workgroups of one XCD, so they run together and
// read the tables through that XCD's L2 instead of HBM.  A chunk after the
// first takes the zone below its first z-face from memory.
__global__ void __launch_bounds__(kFpTI * kFpTJ) k_face_pts(Dev D, Buf6 in, int nch, int nlev) {
  __shared__ double Bs[kBWords][kFpTI * kFpTJ];
  const LBlock wg = lblock();
  const int s = wg.y, N = D.N, NR = D.NR, nti = (N + 7 + kFpTI - 2) / (kFpTI - 1);
  const int t = threadIdx.x, li = t % kFpTI, lj = t / kFpTI;
  const int64_t blk = wg.x / nch;
  const int ch = (int)(wg.x % nch);
  const int i = -3 + (kFpTI - 1) * (int)(blk % nti) + li;
  const int j = -3 + (kFpTJ - 1) * (int)(blk / nti) + lj;
  const bool have = i <= N + D.g && j <= N + D.g;  // the column's zones exist
  const bool mine = have && li < kFpTI - 1 && lj < kFpTJ - 1 && i <= N + 3 && j <= N + 3;
  const bool act = have && D.active(i, j);
  const bool fx = mine && j >= -2, fy = mine && i >= -2, fz = mine && i >= -2 && j >= -2;
  const bool actx = fx && D.active(i + 1, j), acty = fy && D.active(i, j + 1);
  const bool fastx = act && actx, fasty = act && acty;
  double below[kBWords];  // the words of the zone below the z-face about to be made
#pragma unroll
  for (int q = 0; q < kBWords; ++q) below[q] = 0.0;
  // levels k0 .. k1 - 1 of -3 .. NR + 4
  const int k0 = -3 + ch * nlev, k1 = min(k0 + nlev, NR + 5);
  if (k0 >= k1) return;  // the whole workgroup: an empty last chunk
  if (ch > 0 && fz && act) {
    const int64_t z = D.zi(i, j, k0 - 1);
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int m = 0; m < 4; ++m) below[4 * c + m] = D.v(m, s, 5 + c, z);
#pragma unroll
    for (int c = 0; c < 3; ++c) below[12 + c] = D.cen(s, z, c);
    below[15] = D.obl(s, z);
  }
  // one flat loop over (level, family) with one call site per piece
  const int nq = 3 * (k1 - k0);
#pragma unroll 1
  for (int q = 0; q < nq; ++q) {
    const int k = k0 + q / 3, fam = q % 3;
    int ii = i, jj = j;  // opaque, as in k_flux
    asm volatile("" : "+v"(ii), "+v"(jj));
    if (fam == 0) {
      __syncthreads();  // the previous level's LDS words are no longer read
      if (have) {
        const int64_t z = D.zi(ii, jj, k);
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
          for (int m = 0; m < 4; ++m) Bs[4 * c + m][t] = D.v(m, s, 5 + c, z);
#pragma unroll
        for (int c = 0; c < 3; ++c) Bs[12 + c][t] = D.cen(s, z, c);
        Bs[15][t] = D.obl(s, z);
      }
      __syncthreads();
    }
    // the face of this family at this level: x and y (i, j, k), z (i, j, k - 1)
    const bool on = fam == 0 ? (k >= -2 && k <= NR + 3 && fx)
                             : fam == 1 ? (k >= -2 && k <= NR + 3 && fy) : (k >= -2 && fz);
    if (on) {
      const bool fast = fam == 0 ? fastx : fam == 1 ? fasty : act;
      const bool some = fam == 0 ? (act || actx) : fam == 1 ? (act || acty) : act;
      const int fk = fam == 2 ? k - 1 : k;
      if (fast || !some) {
        csg1::FaceG F;
        csg1::face_of(D.G, fam, s, ii, jj, fk, F);
        const int64_t fb = D.fci(ii, jj, fk);
        double q5[5];
#pragma unroll
        for (int pt = 0; pt < 5; ++pt) {
          double bl[3], br[3];
          if (!some) {
#pragma unroll
            for (int c = 0; c < 3; ++c) bl[c] = br[c] = 0.0;
          } else if (fam == 2) {
            b_at(below, F.p[pt], bl);
            lds_b(Bs, t, F.p[pt], br);
          } else {
            lds_b(Bs, t, F.p[pt], bl);
            lds_b(Bs, t + (fam == 0 ? 1 : kFpTI), F.p[pt], br);
          }
          const double* n = fam == 2 ? F.n[pt] : F.n[0];
          double sm[3];
#pragma unroll
          for (int c = 0; c < 3; ++c) sm[c] = bl[c] + br[c];
          q5[pt] = some ? 0.5 * (n[0] * sm[0] + n[1] * sm[1] + n[2] * sm[2]) : 0.0;
        }
        double bd = in.b(fam)[(int64_t)s * D.NFC + fb];
#pragma unroll
        for (int pt = 0; pt < 5; ++pt) bd = bd - F.bw[pt] * q5[pt];
#pragma unroll
        for (int pt = 0; pt < 5; ++pt) D.bq(s, fam, pt)[fb] = q5[pt] + bd;
      } else {
        face_points(D, fam, s, ii, jj, fk, in);  // a tri-corner face: one side re-pointed
      }
    }
    if (fam == 2 && fz && act) {  // this level's zone is the next z-face's lower side
#pragma unroll
      for (int q = 0; q < kBWords; ++q) below[q] = Bs[q][t];
    }
  }
}

// ---------------------------------------------------------------------------
// step 8.d: the ADP constrained least squares for the 12 linear B
// coefficients of a zone (410...f:4643-5360): the six face means are matched
// exactly (constraint rows n (x) (1, r) at the face centroids, the normal of
// the face's first point), the face-point B at points 2 and 4 of each face
// and the 12 WENO modes in least squares; then the energy takes the change
// of magnetic energy (:5336-5357).  The reference solves the 18 x 18 KKT
// system by LU; here six unknowns are eliminated.  In a frame of the zone
// (the mean of its x-face normals, the mean of its y-face normals made
// orthogonal to it, their cross product: a rotation R), the six
// face-mean constraints fix each B component's mean and its derivative along
// its own axis, the pivots (b0x, Gxx, b0y, Gyy, b0z, Gzz) = e - W x_f of the
// six other coefficients x_f (Gaussian elimination of the 6 x 6 pivot block,
// which in this frame is close to block diagonal, [[1, h], [1, -h]] per
// component).  The least-squares rows (the face points 2 and 4 of each face,
// the 12 WENO modes) in x_f are then a 6 x 6 normal system.  B' = R B and
// rho' = R rho, so n . B(rho) = n' . B'(rho') and the WENO rows' squared
// distance is unchanged: the minimiser is the one of the 12-unknown form.
// About 100 live words instead of a 12 x 12 Cholesky factor and a 6 x 12
// block in LDS, so two waves fit per SIMD, and no LDS.
__device__ __forceinline__ void put6(double* a, int f, double v) {
#pragma unroll
  for (int q = 0; q < 6; ++q) a[q] = f == q ? v : a[q];
}
__device__ __forceinline__ void rot3(const double* R, const double* v, double* o) {
#pragma unroll
  for (int r = 0; r < 3; ++r) o[r] = R[3 * r] * v[0] + R[3 * r + 1] * v[1] + R[3 * r + 2] * v[2];
}
// the row n' (x) (1, rho') split into the pivot and the free coefficients
__device__ __forceinline__ void split_row(const double* n, const double* r, double* rp, double* rf) {
  rp[0] = n[0];
  rp[1] = n[0] * r[0];
  rp[2] = n[1];
  rp[3] = n[1] * r[1];
  rp[4] = n[2];
  rp[5] = n[2] * r[2];
  rf[0] = n[0] * r[1];
  rf[1] = n[0] * r[2];
  rf[2] = n[1] * r[0];
  rf[3] = n[1] * r[2];
  rf[4] = n[2] * r[0];
  rf[5] = n[2] * r[1];
}
__global__ void __launch_bounds__(64) k_clsq(Dev D, Box b, Buf6 in) {
  int i, j, k;
  const LBlock wg = lblock();
  if (!b.decode(ltid(wg), i, j, k)) return;
  const int s = wg.y;
  const int64_t z = D.zi(i, j, k);
  if (!in_weno_range(D, i, j, k) || !D.active(i, j)) return;
  const double obl = D.obl(s, z);
  double c0[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) c0[c] = D.cen(s, z, c);
  // face f: 0 right x, 1 left x, 2 up y, 3 down y, 4 top z, 5 bottom z
  // pass 1: the constraint rows (the first point's normal at the centroid) and the face B
  double cn[3][6], cr[3][6], cd[6];
#pragma unroll 1
  for (int f = 0; f < 6; ++f) {
    const int fam = f >> 1;
    const bool hi = (f & 1) == 0;
    const int fi = i - (fam == 0 && !hi), fj = j - (fam == 1 && !hi), fk = k - (fam == 2 && !hi);
    csg1::FaceG F;
    csg1::face_of(D.G, fam, s, fi, fj, fk, F);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      put6(cn[c], f, F.n[0][c]);
      put6(cr[c], f, (F.cen[c] - c0[c]) * obl);
    }
    put6(cd, f, in.b(fam)[(int64_t)s * D.NFC + D.fci(fi, fj, fk)]);
  }
  // the zone's frame
  double R[9];
  {
    double a[3], e2[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      a[c] = cn[c][0] + cn[c][1];
      e2[c] = cn[c][2] + cn[c][3];
    }
    double t = 1.0 / sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
#pragma unroll
    for (int c = 0; c < 3; ++c) a[c] *= t;
    t = e2[0] * a[0] + e2[1] * a[1] + e2[2] * a[2];
#pragma unroll
    for (int c = 0; c < 3; ++c) e2[c] -= t * a[c];
    t = 1.0 / sqrt(e2[0] * e2[0] + e2[1] * e2[1] + e2[2] * e2[2]);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      e2[c] *= t;
      R[c] = a[c];
      R[3 + c] = e2[c];
    }
    R[6] = a[1] * e2[2] - a[2] * e2[1];
    R[7] = a[2] * e2[0] - a[0] * e2[2];
    R[8] = a[0] * e2[1] - a[1] * e2[0];
  }
  // the pivot block and its right-hand sides [free | d], eliminated in place
  double Ap[6][6], Wd[6][7];
#pragma unroll
  for (int f = 0; f < 6; ++f) {
    double n[3], r[3], nn[3], rr[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      n[c] = cn[c][f];
      r[c] = cr[c][f];
    }
    rot3(R, n, nn);
    rot3(R, r, rr);
    double rf[6];
    split_row(nn, rr, Ap[f], rf);
#pragma unroll
    for (int q = 0; q < 6; ++q) Wd[f][q] = rf[q];
    Wd[f][6] = cd[f];
  }
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    const double od = 1.0 / Ap[c][c];
#pragma unroll
    for (int r = c + 1; r < 6; ++r) {
      const double m = Ap[r][c] * od;
#pragma unroll
      for (int q = c + 1; q < 6; ++q) Ap[r][q] -= m * Ap[c][q];
#pragma unroll
      for (int q = 0; q < 7; ++q) Wd[r][q] -= m * Wd[c][q];
    }
  }
#pragma unroll
  for (int r = 5; r >= 0; --r) {
    const double od = 1.0 / Ap[r][r];
#pragma unroll
    for (int q = 0; q < 7; ++q) {
      double t = Wd[r][q];
#pragma unroll
      for (int c = r + 1; c < 6; ++c) t -= Ap[r][c] * Wd[c][q];
      Wd[r][q] = t * od;
    }
  }
  // pass 2: the least-squares rows in the free coefficients
  double N[21], h[6];
#pragma unroll
  for (int q = 0; q < 21; ++q) N[q] = 0.0;
#pragma unroll
  for (int q = 0; q < 6; ++q) h[q] = 0.0;
#pragma unroll 1
  for (int f = 0; f < 6; ++f) {
    const int fam = f >> 1;
    const bool hi = (f & 1) == 0;
    const int fi = i - (fam == 0 && !hi), fj = j - (fam == 1 && !hi), fk = k - (fam == 2 && !hi);
    csg1::FaceG F;
    csg1::face_of(D.G, fam, s, fi, fj, fk, F);
    const int64_t fb = D.fci(fi, fj, fk);
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      const int pt = 1 + 2 * t;
      const double* n = fam == 2 ? F.n[pt] : F.n[0];
      double r[3], nn[3], rr[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) r[c] = (F.p[pt][c] - c0[c]) * obl;
      rot3(R, n, nn);
      rot3(R, r, rr);
      double rp[6], red[6];
      split_row(nn, rr, rp, red);
      double tgt = D.bq(s, fam, pt)[fb];
#pragma unroll
      for (int q = 0; q < 6; ++q) {
        tgt -= rp[q] * Wd[q][6];
#pragma unroll
        for (int c = 0; c < 6; ++c) red[c] -= rp[q] * Wd[q][c];
      }
#pragma unroll
      for (int r2 = 0; r2 < 6; ++r2) {
        h[r2] += red[r2] * tgt;
#pragma unroll
        for (int c = 0; c <= r2; ++c) N[LIX(r2, c)] += red[r2] * red[c];
      }
    }
  }
  // the WENO modes u, in the zone's frame: b0' = R b0, G' = R G R^T
  double up[6], uf[6];
  {
    double b0[3], G[3][3], t[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      b0[c] = D.v(0, s, 5 + c, z);
#pragma unroll
      for (int q = 0; q < 3; ++q) G[c][q] = D.v(1 + q, s, 5 + c, z);
    }
    double bp[3];
    rot3(R, b0, bp);
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int q = 0; q < 3; ++q) t[r][q] = R[3 * r] * G[0][q] + R[3 * r + 1] * G[1][q] + R[3 * r + 2] * G[2][q];
    double Gp[3][3];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int q = 0; q < 3; ++q) Gp[r][q] = t[r][0] * R[3 * q] + t[r][1] * R[3 * q + 1] + t[r][2] * R[3 * q + 2];
    up[0] = bp[0];
    up[1] = Gp[0][0];
    up[2] = bp[1];
    up[3] = Gp[1][1];
    up[4] = bp[2];
    up[5] = Gp[2][2];
    uf[0] = Gp[0][1];
    uf[1] = Gp[0][2];
    uf[2] = Gp[1][0];
    uf[3] = Gp[1][2];
    uf[4] = Gp[2][0];
    uf[5] = Gp[2][1];
  }
  // the WENO rows: free ones directly, pivot ones through x_p = e - W x_f
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    N[LIX(q, q)] += 1.0;
    h[q] += uf[q];
  }
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    const double tgt = up[q] - Wd[q][6];
#pragma unroll
    for (int r2 = 0; r2 < 6; ++r2) {
      h[r2] -= Wd[q][r2] * tgt;
#pragma unroll
      for (int c = 0; c <= r2; ++c) N[LIX(r2, c)] += Wd[q][r2] * Wd[q][c];
    }
  }
  // 6 x 6 Cholesky with reciprocal diagonals, then the two solves
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    double d = N[LIX(c, c)];
#pragma unroll
    for (int q = 0; q < c; ++q) d -= N[LIX(c, q)] * N[LIX(c, q)];
    const double od = 1.0 / sqrt(d);
    N[LIX(c, c)] = od;
#pragma unroll
    for (int r = c + 1; r < 6; ++r) {
      double t = N[LIX(r, c)];
#pragma unroll
      for (int q = 0; q < c; ++q) t -= N[LIX(r, q)] * N[LIX(c, q)];
      N[LIX(r, c)] = t * od;
    }
  }
#pragma unroll
  for (int r = 0; r < 6; ++r) {
    double t = h[r];
#pragma unroll
    for (int q = 0; q < r; ++q) t -= N[LIX(r, q)] * h[q];
    h[r] = t * N[LIX(r, r)];
  }
#pragma unroll
  for (int r = 5; r >= 0; --r) {
    double t = h[r];
#pragma unroll
    for (int q = r + 1; q < 6; ++q) t -= N[LIX(q, r)] * h[q];
    h[r] = t * N[LIX(r, r)];
  }
  // the pivots, then back to the global frame: b0 = R^T b0', G = R^T G' R
  double xp[6];
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    double t = Wd[q][6];
#pragma unroll
    for (int c = 0; c < 6; ++c) t -= Wd[q][c] * h[c];
    xp[q] = t;
  }
  double x[12];
  {
    const double bp[3] = {xp[0], xp[2], xp[4]};
    const double Gp[3][3] = {{xp[1], h[0], h[1]}, {h[2], xp[3], h[3]}, {h[4], h[5], xp[5]}};
    double t[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c) x[4 * c] = R[c] * bp[0] + R[3 + c] * bp[1] + R[6 + c] * bp[2];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int q = 0; q < 3; ++q) t[r][q] = R[r] * Gp[0][q] + R[3 + r] * Gp[1][q] + R[6 + r] * Gp[2][q];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int q = 0; q < 3; ++q) x[4 * r + 1 + q] = t[r][0] * R[q] + t[r][1] * R[3 + q] + t[r][2] * R[6 + q];
  }
  const double b5 = in.u(D, s, 5, z), b6 = in.u(D, s, 6, z), b7 = in.u(D, s, 7, z);
  const double aa = (b5 * b5 + b6 * b6 + b7 * b7) * C<double>::ob8pi;
#pragma unroll
  for (int m = 0; m < 4; ++m)
#pragma unroll
    for (int c = 0; c < 3; ++c) D.v(m, s, 5 + c, z) = x[4 * c + m];
#pragma unroll
  for (int c = 0; c < 3; ++c) in.u(D, s, 5 + c, z) = x[4 * c];
  const double bb = (x[0] * x[0] + x[4] * x[4] + x[8] * x[8]) * C<double>::ob8pi;
  in.u(D, s, 4, z) = in.u(D, s, 4, z) - aa + bb;
}

// ---------------------------------------------------------------------------
// MAKE_FACIAL_FLUXES at second order: per face of the routine's ranges (one
// layer beyond the owned zones), the two zones' linear fields at the face's
// central point, the face frame (x: the normal, the centroid's radial
// direction, their cross product; y: the same with the second and third axes
// swapped; z: the point's normal, the vertex-difference direction made
// orthogonal), the Riemann solve, and the resolved flux rotated back, times
// the area.  Tri-corner faces take the stencil neighbour as the inactive
// side's zone (450...f:581-622).  With conducting walls the radial faces on
// the walls see the inner zone's state mirrored about the face normal.
// The face frame (450...f:890-1030): A the unit normal; for x-faces C the
// radial direction of the centroid and B = C x A, for y-faces B the radial
// direction and C = A x B, for z-faces B along the face's i diagonal made
// orthogonal to A, C = A x B, B = C x A.
__device__ __forceinline__ void face_frame(const Dev& D, int fam, int s, int i, int j, int k, const csg1::FaceG& F,
                                           double* A, double* B, double* Cv) {
  const double sm = D.smallnum;
  if (fam == 0) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      A[c] = F.n[0][c];
      Cv[c] = F.cen[c];
    }
    const double t = 1.0 / fmax(sqrt(Cv[0] * Cv[0] + Cv[1] * Cv[1] + Cv[2] * Cv[2]), sm);
#pragma unroll
    for (int c = 0; c < 3; ++c) Cv[c] = Cv[c] * t;
    B[0] = Cv[1] * A[2] - Cv[2] * A[1];
    B[1] = Cv[2] * A[0] - Cv[0] * A[2];
    B[2] = Cv[0] * A[1] - Cv[1] * A[0];
  } else if (fam == 1) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      A[c] = F.n[0][c];
      B[c] = F.cen[c];
    }
    const double t = 1.0 / fmax(sqrt(B[0] * B[0] + B[1] * B[1] + B[2] * B[2]), sm);
#pragma unroll
    for (int c = 0; c < 3; ++c) B[c] = B[c] * t;
    Cv[0] = A[1] * B[2] - A[2] * B[1];
    Cv[1] = A[2] * B[0] - A[0] * B[2];
    Cv[2] = A[0] * B[1] - A[1] * B[0];
  } else {
#pragma unroll
    for (int c = 0; c < 3; ++c) A[c] = F.n[0][c];
    double t = 1.0 / fmax(sqrt(A[0] * A[0] + A[1] * A[1] + A[2] * A[2]), sm);
#pragma unroll
    for (int c = 0; c < 3; ++c) A[c] = A[c] * t;
    double p00[3], p10[3], p01[3], p11[3];
    D.G.vert(s, i, j, k, p11);
    D.G.vert(s, i, j - 1, k, p10);
    D.G.vert(s, i - 1, j, k, p01);
    D.G.vert(s, i - 1, j - 1, k, p00);
#pragma unroll
    for (int c = 0; c < 3; ++c) B[c] = p11[c] + p10[c] - p01[c] - p00[c];
    t = B[0] * A[0] + B[1] * A[1] + B[2] * A[2];
#pragma unroll
    for (int c = 0; c < 3; ++c) B[c] = B[c] - t * A[c];
    t = 1.0 / fmax(sqrt(B[0] * B[0] + B[1] * B[1] + B[2] * B[2]), sm);
#pragma unroll
    for (int c = 0; c < 3; ++c) B[c] = B[c] * t;
    Cv[0] = A[1] * B[2] - A[2] * B[1];
    Cv[1] = A[2] * B[0] - A[0] * B[2];
    Cv[2] = A[0] * B[1] - A[1] * B[0];
    t = 1.0 / fmax(sqrt(Cv[0] * Cv[0] + Cv[1] * Cv[1] + Cv[2] * Cv[2]), sm);
#pragma unroll
    for (int c = 0; c < 3; ++c) Cv[c] = Cv[c] * t;
    B[0] = Cv[1] * A[2] - Cv[2] * A[1];
    B[1] = Cv[2] * A[0] - Cv[0] * A[2];
    B[2] = Cv[0] * A[1] - Cv[1] * A[0];
  }
}

// A zone's linear reconstruction: the four modes of the nine fields, its
// centroid and its reciprocal scale length; state_at evaluates it at a
// point (the 8 primitives and the entropy).
struct ZoneSlopes {
  double m[4][9], c[3], o;
};
__device__ __forceinline__ void load_slopes(const Dev& D, int s, int64_t z, ZoneSlopes& Z) {
  Z.o = D.obl(s, z);
#pragma unroll
  for (int c = 0; c < 3; ++c) Z.c[c] = D.cen(s, z, c);
#pragma unroll
  for (int m = 0; m < 4; ++m)
#pragma unroll
    for (int c = 0; c < 9; ++c) Z.m[m][c] = D.v(m, s, c, z);
}
__device__ __forceinline__ void state_at(const ZoneSlopes& Z, const double* p, double* w) {
  double r[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) r[c] = (p[c] - Z.c[c]) * Z.o;
#pragma unroll
  for (int c = 0; c < 9; ++c) w[c] = Z.m[0][c] + Z.m[1][c] * r[0] + Z.m[2][c] * r[1] + Z.m[3][c] * r[2];
}

// zone z's linear fields at point p, straight from memory (the arithmetic of
// state_at)
__device__ __forceinline__ void zone_state(const Dev& D, int s, int64_t z, const double* p, double* w) {
  const double o = D.obl(s, z);
  double r[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) r[c] = (p[c] - D.cen(s, z, c)) * o;
#pragma unroll
  for (int c = 0; c < 9; ++c) w[c] = D.v(0, s, c, z) + D.v(1, s, c, z) * r[0] + D.v(2, s, c, z) * r[1] + D.v(3, s, c, z) * r[2];
}

// The two states of a lateral face from memory: the left and right zones
// (a tri-corner face takes the stencil neighbour as its inactive side's
// zone, 450...f:581-622) evaluated at the face's first point.
__device__ __forceinline__ void lateral_states(const Dev& D, int fam, int s, int i, int j, int k, const double* p,
                                               double* VL, double* VR) {
  int li = i, lj = j, ri = i + (fam == 0), rj = j + (fam == 1);
  const bool al = D.active(li, lj), ar = D.active(ri, rj);
  if (!al && ar) D.nbr(ri, rj, fam == 0 ? 1 : 3, li, lj);
  else if (al && !ar) D.nbr(li, lj, fam == 0 ? 0 : 2, ri, rj);
  zone_state(D, s, D.zi(li, lj, k), p, VL);
  zone_state(D, s, D.zi(ri, rj, k), p, VR);
}

// One face from its two states at the face's first point: rotated into the
// face frame, the facial B in the normal slot, the wall mirror for radial
// faces on conducting walls, HLLD or HLL, the resolved flux rotated back to
// the global frame and times the area (450...f:890-1030), with the entropy
// flux as the ninth component and the wave speeds S_L, S_R.
__device__ __forceinline__ void face_solve(const Dev& D, int fam, int s, int k, int64_t fb, double area,
                                           const double* A, const double* B, const double* Cv, double* VL,
                                           double* VR, double* F0, double* F1, double* F2, double* S0, double* S1,
                                           double* S2) {
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
  if (D.wall && fam == 2 && (k == 0 || k == D.NR)) {
    // the wall: the inner zone's state mirrored (normal velocity and normal B
    // reversed), so no mass crosses the wall and its normal B is zero
    const bool lo = k == 0;
#pragma unroll
    for (int c = 0; c < 9; ++c) {
      const double m = (c == 1 || c == 5) ? -1.0 : 1.0, l = VL[c], r = VR[c];
      VL[c] = lo ? m * r : l;
      VR[c] = lo ? r : m * l;
    }
  }
  const Gas<double> gs = D.gas();
  double Fs[8], eF, SL, SR;
  if (D.riemann == CS_RS_HLLD) hlld(gs, VL, VR, VL[8], VR[8], Fs, eF, SL, SR);
  else hll(gs, VL, VR, VL[8], VR[8], Fs, eF, SL, SR);
  double* Fo = (fam == 0 ? F0 : fam == 1 ? F1 : F2) + (int64_t)s * 9 * D.NFC + fb;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    double v;
    if (c == 0 || c == 4) {
      v = Fs[c];
    } else {
      const int o = c < 4 ? 1 : 5, q = c - o;
      v = A[q] * Fs[o] + B[q] * Fs[o + 1] + Cv[q] * Fs[o + 2];
    }
    Fo[(int64_t)c * D.NFC] = v * area;
  }
  Fo[(int64_t)8 * D.NFC] = eF * area;
  double* So = (fam == 0 ? S0 : fam == 1 ? S1 : S2) + (int64_t)s * 2 * D.NFC + fb;
  So[0] = SL;
  So[D.NFC] = SR;
}

// The face fluxes of a sector (the x-faces i -1 .. N+1, j 0 .. N+1, k 0 ..
// NR+1; y likewise; z-faces i, j 0 .. N+1, k -1 .. NR+1), one column (i, j)
// per thread marching up in k.  A workgroup holds a 16 x 8 block of
// columns whose last column and last row are the next blocks' first ones
// (blocks step by 15 and 7): at each level every lane copies its zone's
// slopes, centroid and scale (40 words) from HBM into LDS, and the lanes
// inside the block's 15 x 7 columns evaluate both states of their x- and
// y-faces from their own and their neighbour's LDS words; the lower state of
// a z-face is the column's previous top state (kept in registers).  Only
// tri-corner faces whose neighbour is an inactive column read the zones from
// memory.  So a zone's words come from HBM about 1.2 times per launch
// (16 x 8 loaded for 15 x 7 computed), not once per face side.  As in
// k_face_pts the march is split into chunks of nlev levels whose workgroups
// run together on one XCD and share the faces' geometry tables through its
// L2; a chunk after the first starts one level early with its faces off, so
// that level leaves the bottom state of the chunk's first z-face.
constexpr int kFluxTI = 16, kFluxTJ = 8;  // columns per workgroup along i and j
constexpr int kFluxLevels = 16;           // levels per workgroup at most
constexpr int kSlopeWords = 40;           // 4 x 9 modes, the centroid, the reciprocal scale length
__device__ __forceinline__ void lds_state(const double (*Zs)[kFluxTI * kFluxTJ], int slot, const double* p,
                                          double* w) {
  double r[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) r[c] = (p[c] - Zs[36 + c][slot]) * Zs[39][slot];
#pragma unroll
  for (int c = 0; c < 9; ++c)
    w[c] = Zs[c][slot] + Zs[9 + c][slot] * r[0] + Zs[18 + c][slot] * r[1] + Zs[27 + c][slot] * r[2];
}
__global__ void __launch_bounds__(kFluxTI * kFluxTJ) k_flux(Dev D, double* F0, double* F1, double* F2, double* S0,
                                                           double* S1, double* S2, int nch, int nlev) {
  __shared__ double Zs[kSlopeWords][kFluxTI * kFluxTJ];
  const LBlock wg = lblock();
  const int s = wg.y, N = D.N, NR = D.NR, nti = (N + 3 + kFluxTI - 2) / (kFluxTI - 1);
  const int t = threadIdx.x, li = t % kFluxTI, lj = t / kFluxTI;
  const int64_t blk = wg.x / nch;
  const int ch = (int)(wg.x % nch);
  const int i = -1 + (kFluxTI - 1) * (int)(blk % nti) + li;
  const int j = -1 + (kFluxTJ - 1) * (int)(blk / nti) + lj;
  const bool have = i <= N + 2 && j <= N + 2;  // the column's zones exist
  const bool mine = have && li < kFluxTI - 1 && lj < kFluxTJ - 1;  // the block computes this column's faces
  const bool act = have && D.active(i, j);
  const bool inx = mine && i <= N + 1 && j >= 0 && j <= N + 1, iny = mine && j <= N + 1 && i >= 0 && i <= N + 1;
  const bool actx = inx && D.active(i + 1, j), acty = iny && D.active(i, j + 1);
  const bool onx = inx && (act || actx), ony = iny && (act || acty);
  const bool fastx = act && actx, fasty = act && acty;
  const bool onz = mine && act && i >= 0 && i <= N + 1 && j >= 0 && j <= N + 1;
  double zb[9];  // the column's state at the bottom point of the next z-face (the previous level's top)
#pragma unroll
  for (int c = 0; c < 9; ++c) zb[c] = 0.0;
  // faces of levels k0 .. k1 - 1 of -1 .. NR + 2, the loop from kb
  const int k0 = -1 + ch * nlev, k1 = min(k0 + nlev, NR + 3), kb = ch > 0 ? k0 - 1 : k0;
  if (k0 >= k1) return;  // the whole workgroup: an empty last chunk
  // one flat loop over (level, family): a single call site, so the solver is
  // inlined once (nested loops cost a second wave of registers)
  const int nq = 3 * (k1 - kb);
#pragma unroll 1
  for (int q = 0; q < nq; ++q) {
    const int k = kb + q / 3, fam = q % 3;
    // the column's indices opaque to the optimiser: the k-invariant table
    // loads stay in the loop body instead of being hoisted and held live
    int ii = i, jj = j;
    asm volatile("" : "+v"(ii), "+v"(jj));
    if (fam == 0) {
      __syncthreads();  // the previous level's LDS words are no longer read
      if (have) {
        const int64_t z = D.zi(ii, jj, k);
#pragma unroll
        for (int m = 0; m < 4; ++m)
#pragma unroll
          for (int c = 0; c < 9; ++c) Zs[9 * m + c][t] = D.v(m, s, c, z);
#pragma unroll
        for (int c = 0; c < 3; ++c) Zs[36 + c][t] = D.cen(s, z, c);
        Zs[39][t] = D.obl(s, z);
      }
      __syncthreads();
    }
    const bool on = k >= k0 && (fam == 0 ? (k >= 0 && k <= NR + 1 && onx)
                                          : fam == 1 ? (k >= 0 && k <= NR + 1 && ony) : (k >= 0 && onz));
    if (on) {
      const int fk = fam == 2 ? k - 1 : k;
      csg1::FaceG F;
      csg1::face_of(D.G, fam, s, ii, jj, fk, F);
      double A[3], B[3], Cv[3];
      face_frame(D, fam, s, ii, jj, fk, F, A, B, Cv);
      double VL[9], VR[9];
      if (fam == 2) {
#pragma unroll
        for (int c = 0; c < 9; ++c) VL[c] = zb[c];
        lds_state(Zs, t, F.p[0], VR);
      } else if (fam == 0 ? fastx : fasty) {
        lds_state(Zs, t, F.p[0], VL);
        lds_state(Zs, t + (fam == 0 ? 1 : kFluxTI), F.p[0], VR);
      } else {
        lateral_states(D, fam, s, ii, jj, k, F.p[0], VL, VR);
      }
      face_solve(D, fam, s, fk, D.fci(ii, jj, fk), F.area, A, B, Cv, VL, VR, F0, F1, F2, S0, S1, S2);
    }
    if (fam == 2 && onz) {  // the top state of this level's zone: the bottom of the next z-face
      double p[3];
      csg1::face_p0(D.G, 2, s, ii, jj, k, p);
      lds_state(Zs, t, p, zb);
    }
  }
}

// ---------------------------------------------------------------------------
// MAKE_EDGE_EFIELDS_P1 at second order (460...f:393-3721 with the tri-corner
// blocks :3725-7035): every edge of the owned faces takes the mean zone
// E = -v x B of its four zones along the edge tangent, the normal B at the
// edge points of the four faces around it, and their wave speeds symmetrised
// to +-max|S|, into Two_D_RS_Ez_4State_HLL; radial edges with an inactive
// zone take none, and the radial edge at each tri-corner adds
// Two_D_RS_Ez_3State_HLL of its three active zones.  With conducting walls
// the lateral edges on the walls carry no field.  blockIdx.y = fam + 3 s.
__device__ __forceinline__ double zoneE(const Dev& D, int s, int64_t z, int c) {
  const double vx = D.v(0, s, 1, z), vy = D.v(0, s, 2, z), vz = D.v(0, s, 3, z);
  const double Bx = D.v(0, s, 5, z), By = D.v(0, s, 6, z), Bz = D.v(0, s, 7, z);
  return c == 0 ? -(vy * Bz - vz * By) : (c == 1 ? -(vz * Bx - vx * Bz) : -(vx * By - vy * Bx));
}
__device__ __forceinline__ double emean(const Dev& D, int s, int c, const int64_t* z, int n) {
  return n == 4 ? (zoneE(D, s, z[0], c) + zoneE(D, s, z[1], c) + zoneE(D, s, z[2], c) + zoneE(D, s, z[3], c)) * 0.25
                : (zoneE(D, s, z[0], c) + zoneE(D, s, z[1], c) + zoneE(D, s, z[2], c)) * (1.0 / 3.0);
}
__global__ void __launch_bounds__(256) k_edge(Dev D, Box bx, Box by, Box bz, const double* S0, const double* S1,
                                              const double* S2, double* E0, double* E1, double* E2) {
  const LBlock wg = lblock();
  const int fam = wg.y % 3, s = wg.y / 3;
  const Box& b = fam == 0 ? bx : (fam == 1 ? by : bz);
  int i, j, k;
  if (!tile_decode<32, 8>(b, wg.x, threadIdx.x, i, j, k)) return;
  const int N = D.N;
  const int64_t t = D.fci(i, j, k), NF = D.NFC;
  double* out = (fam == 0 ? E0 : fam == 1 ? E1 : E2) + (int64_t)s * NF;
  if (D.wall && fam < 2 && (k == 0 || k == D.NR)) {
    out[t] = 0.0;
    return;
  }
  const double* SL = S0 + (int64_t)s * 2 * NF;  // x faces: S_L, S_R
  const double* SD = S1 + (int64_t)s * 2 * NF;  // y faces: S_D, S_U
  const double* SB = S2 + (int64_t)s * 2 * NF;  // z faces: S_B, S_T
  const double *S_L = SL, *S_R = SL + NF, *S_D = SD, *S_U = SD + NF, *S_B = SB, *S_T = SB + NF;
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
  bool act = true;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    z[q] = D.zi(zc[q][0], zc[q][1], zc[q][2]);
    f[q] = D.fci(zc[q][0], zc[q][1], zc[q][2]);
    act = act && D.active(zc[q][0], zc[q][1]);
  }
  act = fam < 2 || act;
  double v1[3], v0[3], ec[3], tg[3];
  D.G.vert(s, i, j, k, v1);
  D.G.vert(s, pi, pj, pk, v0);
#pragma unroll
  for (int c = 0; c < 3; ++c) ec[c] = 0.5 * (v1[c] + v0[c]);
  double val = 0.0;
  if (act) {
    double ct = 0.0, st = 1.0;
    if (fam < 2) {
      // tangent: the chord, made normal to the radius at the edge centre
      // (the centre lifted to the vertices' sphere)
      const double rv = sqrt(v1[0] * v1[0] + v1[1] * v1[1] + v1[2] * v1[2]);
      const double lift = rv / sqrt(ec[0] * ec[0] + ec[1] * ec[1] + ec[2] * ec[2]);
      double r[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) r[c] = ec[c] * lift;
      const double ir = 1.0 / sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        r[c] *= ir;
        tg[c] = v1[c] - v0[c];
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
      double n1[3], tau[3], xp[3], xm[3], yp[3], ym[3];
      D.G.vert(s, i + 1, j, k, xp);
      D.G.vert(s, i - 1, j, k, xm);
      D.G.vert(s, i, j + 1, k, yp);
      D.G.vert(s, i, j - 1, k, ym);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        tg[c] = ec[c] * ir;
        n1[c] = xp[c] - xm[c];
        tau[c] = yp[c] - ym[c];
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
    const double Et =
        emean(D, s, 0, z, 4) * tg[0] + emean(D, s, 1, z, 4) * tg[1] + emean(D, s, 2, z, 4) * tg[2];
    const int fy = fam == 0 ? 2 : fam == 1 ? 0 : 1, fn = fam == 0 ? 1 : fam == 1 ? 2 : 0;
    const double By_R = D.bq(s, fy, 4)[f[1]], By_L = D.bq(s, fy, 3)[f[0]];
    const double Bn_U = D.bq(s, fn, 2)[f[3]], Bn_D = D.bq(s, fn, 1)[f[0]];
    const double *sHi1, *sLo1, *sHi2, *sLo2;
    int a1, b1, a2, b2;
    if (fam == 0) {
      sHi1 = S_U; sLo1 = S_D; sHi2 = S_T; sLo2 = S_B; a1 = 3; b1 = 0; a2 = 1; b2 = 0;
    } else if (fam == 1) {
      sHi1 = S_T; sLo1 = S_B; sHi2 = S_R; sLo2 = S_L; a1 = 3; b1 = 0; a2 = 0; b2 = 1;
    } else {
      sHi1 = S_R; sLo1 = S_L; sHi2 = S_U; sLo2 = S_D; a1 = 3; b1 = 0; a2 = 1; b2 = 0;
    }
    const double sR = fmax(sHi1[f[a1]], sHi1[f[b1]]), sL = fmin(sLo1[f[a1]], sLo1[f[b1]]);
    const double sU = fmax(sHi2[f[a2]], sHi2[f[b2]]), sD = fmin(sLo2[f[a2]], sLo2[f[b2]]);
    const double S = fmax(fmax(fabs(sR), fabs(sL)), fmax(fabs(sU), fabs(sD)));
    val = emf4_hll<double>(S, -S, S, -S, ct, st, Et, Et, Et, Et, Bn_U, Bn_D, By_R, By_L);
  }
  if (fam == 2) {
    const int corner = (i == 0 && j == 0) ? 0 : (i == N && j == 0) ? 1 : (i == N && j == N) ? 2
                                              : (i == 0 && j == N) ? 3 : -1;
    if (corner >= 0) {
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
      const int64_t z3[3] = {D.zi(q1[0], q1[1], k), D.zi(q2[0], q2[1], k), D.zi(q3[0], q3[1], k)};
      const double ir = 1.0 / sqrt(ec[0] * ec[0] + ec[1] * ec[1] + ec[2] * ec[2]);
      const double Et = emean(D, s, 0, z3, 3) * (ec[0] * ir) + emean(D, s, 1, z3, 3) * (ec[1] * ir) +
                        emean(D, s, 2, z3, 3) * (ec[2] * ir);
      const int64_t f1 = D.fci(q1[0], q1[1], k), f2 = D.fci(q2[0], q2[1], k), f3 = D.fci(q3[0], q3[1], k),
                    f2d = D.fci(q2[0], q2[1] - 1, k), f1d = D.fci(q1[0], q1[1] - 1, k);
      const double *bx1 = D.bq(s, 0, 1), *bx2 = D.bq(s, 0, 2), *by3 = D.bq(s, 1, 3), *by4 = D.bq(s, 1, 4);
      double bn[3], sr[3], sl[3];
      if (corner == 0) {
        bn[0] = -by4[f3]; bn[1] = bx2[f2]; bn[2] = by3[f2d];
        sr[0] = -S_D[f3]; sl[0] = -S_U[f3]; sr[1] = S_R[f2]; sl[1] = S_L[f2];
        sr[2] = S_U[f2d]; sl[2] = S_D[f2d];
      } else if (corner == 1) {
        bn[0] = -by4[f1d]; bn[1] = bx2[f2]; bn[2] = by3[f3];
        sr[0] = -S_L[f3]; sl[0] = -S_R[f3]; sr[1] = S_R[f2]; sl[1] = S_L[f2];
        sr[2] = S_U[f3]; sl[2] = S_D[f3];
      } else if (corner == 2) {
        bn[0] = bx2[f1]; bn[1] = by3[f2]; bn[2] = -bx1[f2];
        sr[0] = S_R[f1]; sl[0] = S_L[f1]; sr[1] = S_U[f2]; sl[1] = S_D[f2];
        sr[2] = -S_L[f2]; sl[2] = -S_R[f2];
      } else {
        bn[0] = -by4[f3]; bn[1] = by3[f2]; bn[2] = -bx1[f2];
        sr[0] = -S_D[f3]; sl[0] = -S_U[f3]; sr[1] = S_U[f2]; sl[1] = S_D[f2];
        sr[2] = -S_L[f2]; sl[2] = -S_R[f2];
      }
      const double S = fmax(fmax(fmax(fabs(sr[0]), fabs(sl[0])), fmax(fabs(sr[1]), fabs(sl[1]))),
                            fmax(fabs(sr[2]), fabs(sl[2])));
      const double Sr[3] = {S, S, S}, Sl[3] = {-S, -S, -S}, E3[3] = {Et, Et, Et};
      val += emf3_hll(Sr, Sl, E3, bn);
    }
  }
  out[t] = val;
}

// ---------------------------------------------------------------------------
// UPDATE_IMEX_RK_STEP_P1 for the owned faces (484...f:630-906): constrained
// transport with the edges' radial lengths and arc lengths r acos(u1 . u2)
// (both directions scaled by 1 / r of the face's own vertex), the stage
// weights applied: out = wa B^n + wb (B_in - dt circ / A).  blockIdx.y = fam + 3 s.
__device__ __forceinline__ double vrad(const Dev& D, int s, int i, int j, int k) {
  double x[3];
  D.G.vert(s, i, j, k, x);
  return sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
}
__device__ __forceinline__ double arc(const Dev& D, int s, int i1, int j1, int i2, int j2, int k, double t) {
  double a[3], b[3];
  D.G.vert(s, i1, j1, k, a);
  D.G.vert(s, i2, j2, k, b);
  return acos((a[0] * t) * (b[0] * t) + (a[1] * t) * (b[1] * t) + (a[2] * t) * (b[2] * t));
}
__global__ void __launch_bounds__(256) k_upd_face(Dev D, Buf6 in, Buf6 u0, Buf6 out, const double* E0,
                                                  const double* E1, const double* E2, const double* dtp, double wa,
                                                  double wb) {
  const LBlock wg = lblock();
  const int fam = wg.y % 3, s = wg.y / 3;
  const int N = D.N, NR = D.NR;
  // the owned faces of the family: x i 0 .. N, j 1 .. N, k 1 .. NR; y, z likewise
  const Box fb = make_box(fam == 0 ? 0 : 1, N, fam == 1 ? 0 : 1, N, fam == 2 ? 0 : 1, NR);
  int i, j, k;
  if (!tile_decode<32, 8>(fb, wg.x, threadIdx.x, i, j, k)) return;
  const double dt = dtp[0];
  const int64_t so = (int64_t)s * D.NFC, f = D.fci(i, j, k);
  const double *Ex = E0 + so, *Ey = E1 + so, *Ez = E2 + so;
  const double rout = vrad(D, s, i, j, k), tt = 1.0 / rout;
  double circ, area;
  if (fam == 0) {
    const double rin = vrad(D, s, i, j, k - 1), dr = rout - rin, th = arc(D, s, i, j, i, j - 1, k, tt);
    const double lin = rin * th, lout = rout * th;
    circ = dr * Ez[f] - dr * Ez[D.fci(i, j - 1, k)] - lout * Ey[f] + lin * Ey[D.fci(i, j, k - 1)];
    area = csg1::area_x(D.G, s, i, j, k);
  } else if (fam == 1) {
    const double rin = vrad(D, s, i, j, k - 1), dr = rout - rin, th = arc(D, s, i, j, i - 1, j, k, tt);
    const double lin = rin * th, lout = rout * th;
    circ = -(dr * Ez[f]) + dr * Ez[D.fci(i - 1, j, k)] + lout * Ex[f] - lin * Ex[D.fci(i, j, k - 1)];
    area = csg1::area_y(D.G, s, i, j, k);
  } else {
    const double lxp = rout * arc(D, s, i, j, i - 1, j, k, tt), lxm = rout * arc(D, s, i, j - 1, i - 1, j - 1, k, tt);
    const double lyp = rout * arc(D, s, i, j, i, j - 1, k, tt), lym = rout * arc(D, s, i - 1, j, i - 1, j - 1, k, tt);
    circ = lyp * Ey[f] - lym * Ey[D.fci(i - 1, j, k)] - lxp * Ex[f] + lxm * Ex[D.fci(i, j - 1, k)];
    area = csg1::area_z(D.G, s, i, j, k);
  }
  const double bn = in.b(fam)[so + f] - dt * circ / area;
  out.b(fam)[so + f] = wa != 0.0 ? wa * u0.b(fam)[so + f] + wb * bn : wb * bn;
}

// UPDATE_IMEX_RK_STEP_P1 for the owned zones (484...f:483-628) with the stage
// weights and the implicit gravity of the IMEX stages, then
// CLEANUP_AFTER_UPDATE (485...f:691-865, 1180-1280): the zone B as the
// least-squares fit of its six new face values over the face normals (both
// z-faces: the radial direction of the upper z-face centroid), the energy
// changed by the change of magnetic energy, CONS_TO_PRIM.
__global__ void __launch_bounds__(256) k_upd_zone(Dev D, Box b, Buf6 in, Buf6 u0, Buf6 out, const double* F0,
                                                  const double* F1, const double* F2, const double* dtp, double wa,
                                                  double wb, int mode) {
  int i, j, k;
  const LBlock wg = lblock();
  if (!tile_decode<32, 8>(b, wg.x, threadIdx.x, i, j, k)) return;
  const int s = wg.y;
  const int64_t z = D.zi(i, j, k), so = (int64_t)s * 9 * D.NFC;
  const double dt = dtp[0];
  const int64_t fxp = D.fci(i, j, k), fxm = D.fci(i - 1, j, k), fym = D.fci(i, j - 1, k), fzm = D.fci(i, j, k - 1);
  const double ov = 1.0 / D.vol(s, z);
  double L[9];
#pragma unroll
  for (int q = 0; q < 9; ++q) {
    const int64_t o = so + (int64_t)q * D.NFC;
    L[q] = -ov * ((F0[o + fxp] - F0[o + fxm]) + (F1[o + fxp] - F1[o + fym]) + (F2[o + fxp] - F2[o + fzm]));
  }
  double n[3] = {0.0, 0.0, 0.0}, gr = 0.0;
  if (D.grav != 0.0) {
    double c[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) c[q] = D.cen(s, z, q);
    const double r2 = c[0] * c[0] + c[1] * c[1] + c[2] * c[2], ir = 1.0 / sqrt(r2);
#pragma unroll
    for (int q = 0; q < 3; ++q) n[q] = c[q] * ir;
    gr = -D.grav / r2;
  }
  double ui[8], uz[8], u[8], en;
#pragma unroll
  for (int q = 0; q < 8; ++q) ui[q] = in.u(D, s, q, z);
  const double ei = in.e(D, s, z);
  double ez = 0.0;
  if (mode != 0 || wa != 0.0) {
#pragma unroll
    for (int q = 0; q < 8; ++q) uz[q] = u0.u(D, s, q, z);
    ez = u0.e(D, s, z);
  }
  rk_combine(ui, ei, uz, ez, dt, wa, wb, mode, gr, n, L, u, en);
  const Gas<double> gs = D.gas();
  double Vv[8];
  cons_to_prim(gs, u, Vv, en);  // the update's CONS_TO_PRIM (484...f:620-628)
  // the cleanup's least-squares zone B
  double A[6][3], R[6];
  csg1::normal_x(D.G, s, i, j, A[0]);
  csg1::normal_x(D.G, s, i - 1, j, A[1]);
  csg1::normal_y(D.G, s, i, j, A[2]);
  csg1::normal_y(D.G, s, i, j - 1, A[3]);
  {
    double nv[3];
    csg1::zcen_dir(D.G, s, i, j, nv);
    const double r = D.G.r(k);
#pragma unroll
    for (int c = 0; c < 3; ++c) nv[c] = r * nv[c];
    const double t = 1.0 / fmax(D.smallnum, sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]));
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      A[4][c] = nv[c] * t;
      A[5][c] = A[4][c];
    }
  }
  const int64_t fo = (int64_t)s * D.NFC;
  R[0] = out.Bx[fo + fxp];
  R[1] = out.Bx[fo + fxm];
  R[2] = out.By[fo + fxp];
  R[3] = out.By[fo + fym];
  R[4] = out.Bz[fo + fxp];
  R[5] = out.Bz[fo + fzm];
  double M[3][3], r3[3];
#pragma unroll
  for (int p = 0; p < 3; ++p) {
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      double t = 0.0;
#pragma unroll
      for (int r = 0; r < 6; ++r) t += A[r][p] * A[r][q];
      M[p][q] = t;
    }
    double t = 0.0;
#pragma unroll
    for (int r = 0; r < 6; ++r) t += A[r][p] * R[r];
    r3[p] = t;
  }
  const double det = M[0][0] * (M[1][1] * M[2][2] - M[2][1] * M[1][2]) - M[0][1] * (M[1][0] * M[2][2] - M[2][0] * M[1][2]) +
                     M[0][2] * (M[1][0] * M[2][1] - M[2][0] * M[1][1]);
  const double od = copysign(1.0, det) / fmax(1.0e-32, fabs(det));
  const double bi[3][3] = {{(M[1][1] * M[2][2] - M[1][2] * M[2][1]) * od, (M[0][2] * M[2][1] - M[0][1] * M[2][2]) * od,
                            (M[0][1] * M[1][2] - M[1][1] * M[0][2]) * od},
                           {(M[1][2] * M[2][0] - M[1][0] * M[2][2]) * od, (M[0][0] * M[2][2] - M[0][2] * M[2][0]) * od,
                            (M[0][2] * M[1][0] - M[0][0] * M[1][2]) * od},
                           {(M[1][0] * M[2][1] - M[1][1] * M[2][0]) * od, (M[0][1] * M[2][0] - M[0][0] * M[2][1]) * od,
                            (M[0][0] * M[1][1] - M[0][1] * M[1][0]) * od}};
  const double b2old = u[5] * u[5] + u[6] * u[6] + u[7] * u[7];
#pragma unroll
  for (int p = 0; p < 3; ++p) u[5 + p] = bi[p][0] * r3[0] + bi[p][1] * r3[1] + bi[p][2] * r3[2];
  const double ob8pi = 0.5 * C<double>::ob4pi;
  u[4] = u[4] - b2old * ob8pi + (u[5] * u[5] + u[6] * u[6] + u[7] * u[7]) * ob8pi;
  Vv[4] = 0.0;
  cons_to_prim(gs, u, Vv, en);
#pragma unroll
  for (int q = 0; q < 8; ++q) out.u(D, s, q, z) = u[q];
  out.e(D, s, z) = en;
}

// IMEX-SSP2(2,2,2) first stage: U1 = Un + gi dt S(U1) in closed form, every
// zone (the exchange then refreshes the ghosts)
__global__ void __launch_bounds__(256) k_imex(Dev D, Box b, Buf6 un, Buf6 u1, const double* dtp) {
  int i, j, k;
  const LBlock wg = lblock();
  if (!b.decode(ltid(wg), i, j, k)) return;
  const int s = wg.y;
  const int64_t z = D.zi(i, j, k);
  double u[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) u[q] = un.u(D, s, q, z);
  if (D.active(i, j) && D.grav != 0.0) {
    double c[3], n[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) c[q] = D.cen(s, z, q);
    const double r2 = c[0] * c[0] + c[1] * c[1] + c[2] * c[2], ir = 1.0 / sqrt(r2);
#pragma unroll
    for (int q = 0; q < 3; ++q) n[q] = c[q] * ir;
    imex_pre(dtp[0], -D.grav / r2, n, u);
  }
#pragma unroll
  for (int q = 0; q < 8; ++q) u1.u(D, s, q, z) = u[q];
  u1.e(D, s, z) = un.e(D, s, z);
}

// EVALUATE_TIMESTEP (490...f:72-163): CONS_TO_PRIM of the owned active zones
// (U rewritten, the entropy array left as it is), the magnetosonic speed and
// min cfl l / (|v| + c_f), one partial minimum per workgroup
constexpr int kCflThreads = 256;
__global__ void __launch_bounds__(kCflThreads) k_cfl(Dev D, Buf6 cur, double cfl, double* part) {
  __shared__ double red[kCflThreads];
  const Gas<double> gs = D.gas();
  const int64_t per = (int64_t)D.N * D.N * D.NR, n = 6 * per, stride = (int64_t)gridDim.x * blockDim.x;
  double best = 1e300;
  for (int64_t t = tid(); t < n; t += stride) {
    const int s = (int)(t / per);
    const int64_t r = t - (int64_t)s * per;
    const int i = 1 + (int)(r % D.N), j = 1 + (int)((r / D.N) % D.N), k = 1 + (int)(r / ((int64_t)D.N * D.N));
    const int64_t z = D.zi(i, j, k);
    double U[8], Vv[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) U[c] = cur.u(D, s, c, z);
    double e = cur.e(D, s, z);
    Vv[4] = 0.0;
    cons_to_prim(gs, U, Vv, e);
#pragma unroll
    for (int c = 0; c < 8; ++c) cur.u(D, s, c, z) = U[c];
    const double b2 = Vv[5] * Vv[5] + Vv[6] * Vv[6] + Vv[7] * Vv[7];
    const double ms = sqrt((D.gamma * Vv[4] + C<double>::ob4pi * b2) / Vv[0]);
    const double sp = sqrt(Vv[1] * Vv[1] + Vv[2] * Vv[2] + Vv[3] * Vv[3]) + ms;
    best = fmin(best, cfl * D.scale(s, z) / sp);
  }
  red[threadIdx.x] = best;
  __syncthreads();
  for (int h = kCflThreads / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + h]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// ---------------------------------------------------------------------------
// PAD_BOUNDARY_R (826pad_boundary_r.f:201-420) with outflow at both radial
// ends, after the exchange: every active column's inner ghost layers copy
// layer 1 (U, the lateral face B of the layer, the radial face B of face 0),
// its outer ghost layers take layer nr scaled by the volume ratio V(nr) /
// V(nr + l) and the face B of layer nr (radial: face nr) scaled by the ratio
// of the face areas.  The entropy is not written (the exchange's copy stays).
__global__ void __launch_bounds__(256) k_pad6(Dev D, Buf6 b) {
  const int s = blockIdx.y, g = D.g, N = D.N, NR = D.NR, nz = D.nz;
  const int64_t t = tid();
  if (t >= (int64_t)nz * nz * g) return;
  const int ix = (int)(t % nz) + 1 - g, iy = (int)((t / nz) % nz) + 1 - g, izz = (int)(t / ((int64_t)nz * nz)) + 1;
  if (!D.active(ix, iy)) return;
  const int64_t so = (int64_t)s * D.NFC;
  double *Bx = b.Bx + so, *By = b.By + so, *Bz = b.Bz + so;
  {
    const int l = 1 - izz, r = 1;
    const int64_t zl = D.zi(ix, iy, l), zr = D.zi(ix, iy, r);
#pragma unroll
    for (int c = 0; c < 8; ++c) b.u(D, s, c, zl) = b.u(D, s, c, zr);
    Bx[D.fci(ix, iy, l)] = Bx[D.fci(ix, iy, r)];
    Bx[D.fci(ix - 1, iy, l)] = Bx[D.fci(ix - 1, iy, r)];
    By[D.fci(ix, iy, l)] = By[D.fci(ix, iy, r)];
    By[D.fci(ix, iy - 1, l)] = By[D.fci(ix, iy - 1, r)];
    Bz[D.fci(ix, iy, l - 1)] = Bz[D.fci(ix, iy, r - 1)];
  }
  {
    const int l = NR + izz, r = NR;
    const int64_t zl = D.zi(ix, iy, l), zr = D.zi(ix, iy, r);
    const double vr = D.vol(s, zr), vl = D.vol(s, zl);
#pragma unroll
    for (int c = 0; c < 8; ++c) b.u(D, s, c, zl) = b.u(D, s, c, zr) * vr / vl;
    using namespace csg1;
    Bz[D.fci(ix, iy, l)] = Bz[D.fci(ix, iy, r)] * area_z(D.G, s, ix, iy, r) / area_z(D.G, s, ix, iy, l);
    Bx[D.fci(ix, iy, l)] = Bx[D.fci(ix, iy, r)] * area_x(D.G, s, ix, iy, r) / area_x(D.G, s, ix, iy, l);
    By[D.fci(ix, iy, l)] = By[D.fci(ix, iy, r)] * area_y(D.G, s, ix, iy, r) / area_y(D.G, s, ix, iy, l);
    Bx[D.fci(ix - 1, iy, l)] = Bx[D.fci(ix - 1, iy, r)] * area_x(D.G, s, ix - 1, iy, r) / area_x(D.G, s, ix - 1, iy, l);
    By[D.fci(ix, iy - 1, l)] = By[D.fci(ix, iy - 1, r)] * area_y(D.G, s, ix, iy - 1, r) / area_y(D.G, s, ix, iy - 1, l);
  }
  (void)N;
}

// ---------------------------------------------------------------------------
// INIT_PATCH_CUBE_SPHERE (820init_patch_cube_sphere.f:303-1177) and, with
// layers (1 inner, 2 outer), ADD_BOUNDARY (825add_boundary.f:320-1267): U the
// average of PROBLEM_GENERATOR_U_A's U over the zone's 8 Gauss points, the
// entropy of that mean (:431-455); face B the circulation of A over the
// face's four edges over its area.  Zones and faces of inactive columns keep
// their values (the whole-state form clears the buffer first).
__device__ __forceinline__ bool in_layers(const Dev& D, int k, int layers, bool zface) {
  if (!layers) return true;
  const bool inner = zface ? (k >= -D.g && k <= -1) : (k >= 1 - D.g && k <= 0);
  const bool outer = k >= D.NR + 1 && k <= D.NR + D.g;
  return ((layers & 1) && inner) || ((layers & 2) && outer);
}
__global__ void __launch_bounds__(128) k_init_zone(Dev D, Box b, Buf6 out, cspg::Problem P, int layers) {
  int i, j, k;
  if (!b.decode(tid(), i, j, k)) return;
  const int s = blockIdx.y;
  if (!in_layers(D, k, layers, false)) return;
  const int64_t z = D.zi(i, j, k);
  if (layers) {
#pragma unroll
    for (int c = 0; c < 8; ++c) out.u(D, s, c, z) = 0.0;
  }
  if (!D.active(i, j)) return;
  double X[20][3], Pq[8][3], w[8], U[8];
  const double v = csg1::zone_gauss(D.G, s, i, j, k, X, Pq, w);
#pragma unroll
  for (int c = 0; c < 8; ++c) U[c] = 0.0;
  for (int q = 0; q < 8; ++q) {
    double Up[8], A[3];
    cspg::point(P, Pq[q][0], Pq[q][1], Pq[q][2], Up, A);
    const double wq = w[q] / v;
#pragma unroll
    for (int c = 0; c < 8; ++c) U[c] = U[c] + Up[c] * wq;
  }
  if (P.id == 103) U[0] = U[0] * (1.0 + 1.0e-2 * cspg::normal(P.seed, (uint64_t)s * (uint64_t)D.NZ + (uint64_t)z));
  const double rho = fmax(U[0], P.rhofloor), v0 = rho / (rho * rho + P.rhofloor * P.rhofloor);
  const double vx = U[1] * v0, vy = U[2] * v0, vz = U[3] * v0;
  double prs = (P.gamma - 1.0) * (U[4] - 0.5 * rho * (vx * vx + vy * vy + vz * vz) -
                                  (U[5] * U[5] + U[6] * U[6] + U[7] * U[7]) * (0.5 * C<double>::ob4pi));
  prs = fmax(prs, P.prsfloor);
#pragma unroll
  for (int c = 0; c < 8; ++c) out.u(D, s, c, z) = U[c];
  out.e(D, s, z) = prs / pow(rho, P.gamma - 1.0);
}
// blockIdx.y = fam + 3 s; every face of the family's index range
__global__ void __launch_bounds__(128) k_init_face(Dev D, Buf6 out, cspg::Problem P, int layers) {
  const int fam = blockIdx.y % 3, s = blockIdx.y / 3;
  const int64_t t = tid();
  if (t >= D.NFC) return;
  const int n = D.nf, g = D.g, N = D.N;
  const int i = (int)(t % n) - g, j = (int)((t / n) % n) - g, k = (int)(t / ((int64_t)n * n)) - g;
  double v1[3], v2[3], v3[3], v4[3];
  double sum = 0.0, area;
  if (fam == 0) {
    if (j < 1 - g || k < 1 - g || !in_layers(D, k, layers, false)) return;
    if (!D.active(max(i, 1 - g), j) && !D.active(min(i + 1, N + g), j)) return;
    D.G.vert(s, i, j - 1, k - 1, v1);
    D.G.vert(s, i, j, k - 1, v2);
    D.G.vert(s, i, j - 1, k, v3);
    D.G.vert(s, i, j, k, v4);
    sum = sum + cspg::edge(P, v2, v4, false, 0.0);
    sum = sum + cspg::edge(P, v3, v1, false, 0.0);
    sum = sum + cspg::edge(P, v4, v3, true, csg1::len3(v3));
    sum = sum + cspg::edge(P, v1, v2, true, csg1::len3(v1));
    area = csg1::area_x(D.G, s, i, j, k);
  } else if (fam == 1) {
    if (i < 1 - g || k < 1 - g || !in_layers(D, k, layers, false)) return;
    if (!D.active(i, max(j, 1 - g)) && !D.active(i, min(j + 1, N + g))) return;
    D.G.vert(s, i - 1, j, k - 1, v1);
    D.G.vert(s, i - 1, j, k, v2);
    D.G.vert(s, i, j, k - 1, v3);
    D.G.vert(s, i, j, k, v4);
    sum = sum + cspg::edge(P, v2, v4, true, csg1::len3(v2));
    sum = sum + cspg::edge(P, v3, v1, true, csg1::len3(v1));
    sum = sum + cspg::edge(P, v4, v3, false, 0.0);
    sum = sum + cspg::edge(P, v1, v2, false, 0.0);
    area = csg1::area_y(D.G, s, i, j, k);
  } else {
    if (i < 1 - g || j < 1 - g || !in_layers(D, k, layers, true)) return;
    if (!D.active(i, j)) return;
    D.G.vert(s, i - 1, j - 1, k, v1);
    D.G.vert(s, i, j - 1, k, v2);
    D.G.vert(s, i - 1, j, k, v3);
    D.G.vert(s, i, j, k, v4);
    sum = sum + cspg::edge(P, v2, v4, true, csg1::len3(v2));
    sum = sum + cspg::edge(P, v3, v1, true, csg1::len3(v1));
    sum = sum + cspg::edge(P, v4, v3, true, csg1::len3(v3));
    sum = sum + cspg::edge(P, v1, v2, true, csg1::len3(v1));
    area = csg1::area_z(D.G, s, i, j, k);
  }
  out.b(fam)[(int64_t)s * D.NFC + t] = sum / area;
}

// ---------------------------------------------------------------------------
// diagnostics on the P1 geometry: per owned active zone the relative discrete
// divergence |sum A_f B_f| / sum |A_f B_f| (maximum per workgroup), and the
// volume integrals of U and the entropy (two-sum partials per workgroup)
constexpr int kDiagThreads = 256;
__global__ void __launch_bounds__(kDiagThreads) k_divb(Dev D, Buf6 cur, double* part) {
  __shared__ double sm[kDiagThreads / 64];
  const int64_t per = (int64_t)D.N * D.N * D.NR, t = tid();
  double v = 0.0;
  if (t < 6 * per) {
    const int s = (int)(t / per);
    const int64_t r = t - (int64_t)s * per;
    const int i = 1 + (int)(r % D.N), j = 1 + (int)((r / D.N) % D.N), k = 1 + (int)(r / ((int64_t)D.N * D.N));
    if (D.active(i, j)) {
      const int64_t so = (int64_t)s * D.NFC;
      const double p[6] = {csg1::area_x(D.G, s, i, j, k) * cur.Bx[so + D.fci(i, j, k)],
                           -csg1::area_x(D.G, s, i - 1, j, k) * cur.Bx[so + D.fci(i - 1, j, k)],
                           csg1::area_y(D.G, s, i, j, k) * cur.By[so + D.fci(i, j, k)],
                           -csg1::area_y(D.G, s, i, j - 1, k) * cur.By[so + D.fci(i, j - 1, k)],
                           csg1::area_z(D.G, s, i, j, k) * cur.Bz[so + D.fci(i, j, k)],
                           -csg1::area_z(D.G, s, i, j, k - 1) * cur.Bz[so + D.fci(i, j, k - 1)]};
      double sum = 0.0, sa = 0.0;
#pragma unroll
      for (int q = 0; q < 6; ++q) {
        sum += p[q];
        sa += fabs(p[q]);
      }
      v = sa > 0.0 ? fabs(sum) / sa : 0.0;
    }
  }
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double x = 0.0;
    for (int q = 0; q < kDiagThreads / 64; ++q) x = fmax(x, sm[q]);
    part[blockIdx.x] = x;
  }
}
__device__ __forceinline__ void two_sum(double a, double b, double& s, double& e) {
  s = a + b;
  const double bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}
__global__ void __launch_bounds__(kDiagThreads) k_sums(Dev D, Buf6 cur, double* part) {
  __shared__ double ss[kDiagThreads][2];
  const int64_t per = (int64_t)D.N * D.N * D.NR, t = tid();
  int s = 0;
  int64_t z = 0;
  bool on = false;
  if (t < 6 * per) {
    s = (int)(t / per);
    const int64_t r = t - (int64_t)s * per;
    const int i = 1 + (int)(r % D.N), j = 1 + (int)((r / D.N) % D.N), k = 1 + (int)(r / ((int64_t)D.N * D.N));
    z = D.zi(i, j, k);
    on = D.active(i, j);
  }
  for (int q = 0; q < 9; ++q) {
    const double v = on ? D.vol(s, z) * (q < 8 ? cur.u(D, s, q, z) : cur.e(D, s, z)) : 0.0;
    ss[threadIdx.x][0] = v;
    ss[threadIdx.x][1] = 0.0;
    __syncthreads();
    for (int o = kDiagThreads / 2; o > 0; o >>= 1) {
      if ((int)threadIdx.x < o) {
        double sa, ea;
        two_sum(ss[threadIdx.x][0], ss[threadIdx.x + o][0], sa, ea);
        ss[threadIdx.x][0] = sa;
        ss[threadIdx.x][1] = ss[threadIdx.x][1] + ss[threadIdx.x + o][1] + ea;
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      part[(blockIdx.x * 9 + q) * 2 + 0] = ss[0][0];
      part[(blockIdx.x * 9 + q) * 2 + 1] = ss[0][1];
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// host side
struct Stage {
  Dev D;
  double cfl, rmin, rmax;
  std::vector<void*> mem;
  Box all, wr, own, fx, fy, fz, ffx, ffy, ffz, ex, ey, ez;
  Marks* marks = nullptr;
};

void set_marks(Stage* S, Marks* m) { S->marks = m; }
void mark_begin(Stage* S, hipStream_t st) {
  Marks* m = S->marks;
  if (m && m->n < m->cap) (void)hipEventRecord(m->ev[2 * m->n], st);
}
void mark_end(Stage* S, hipStream_t st, int kid) {
  Marks* m = S->marks;
  if (m && m->n < m->cap) {
    (void)hipEventRecord(m->ev[2 * m->n + 1], st);
    m->kid[m->n++] = kid;
  }
}

static inline unsigned nb(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

#define P1S_HIP(call)                                                              \
  do {                                                                            \
    hipError_t e_ = (call);                                                       \
    if (e_ != hipSuccess) {                                                       \
      cs::set_error("P1 stage: %s: %s", #call, hipGetErrorString(e_));            \
      return CS_E_HIP;                                                            \
    }                                                                             \
  } while (0)

template <typename T>
static int dalloc(Stage* S, T*& p, int64_t n, size_t* bytes) {
  void* q = nullptr;
  const size_t b = sizeof(T) * (size_t)n;
  if (hipMalloc(&q, b) != hipSuccess) {
    cs::set_error("P1 stage: hipMalloc of %.3f GB failed", (double)b / 1e9);
    return CS_E_NOMEM;
  }
  S->mem.push_back(q);
  if (hipMemset(q, 0, b) != hipSuccess) {
    cs::set_error("P1 stage: hipMemset failed");
    return CS_E_HIP;
  }
  p = (T*)q;
  *bytes += b;
  return CS_OK;
}

void destroy(Stage* S) {
  if (!S) return;
  for (void* q : S->mem) (void)hipFree(q);
  delete S;
}

int create(const cs_config* cfg, Stage** out, size_t* bytes) {
  *out = nullptr;
  *bytes = 0;
  if (cfg->nghost < 4) {
    cs::set_error("CS_RECON_P1ADP needs nghost >= 4 (the P1 stencils reach four ghost layers)");
    return CS_E_INVALID;
  }
  if (cfg->nghost >= cfg->ntp) {
    cs::set_error("CS_RECON_P1ADP: nghost (%d) must be below ntp (%d)", cfg->nghost, cfg->ntp);
    return CS_E_BAD_DECOMP;
  }
  if (cfg->precision != CS_FP64) {
    cs::set_error("CS_RECON_P1ADP runs in FP64");
    return CS_E_INVALID;
  }
  if (cfg->emf != CS_EMF_LLF) {
    cs::set_error("CS_RECON_P1ADP uses the reference's LLF edge fields (CS_EMF_LLF)");
    return CS_E_INVALID;
  }
  if ((double)6 * (cfg->ntp + 2 * cfg->nghost) * (cfg->ntp + 2 * cfg->nghost) * (cfg->nr + 2 * cfg->nghost) >=
      4294967295.0) {
    cs::set_error("CS_RECON_P1ADP: the shell has too many zones for 32-bit zone lists");
    return CS_E_INVALID;
  }
  cs::P1GeoHost H;
  int rc = cs::p1geo_host(cfg, H);
  if (rc) return rc;
  Stage* S = new Stage();
  Dev& D = S->D;
  const int N = cfg->ntp, NR = cfg->nr, g = cfg->nghost;
  D.N = N;
  D.NR = NR;
  D.g = g;
  D.nz = N + 2 * g;
  D.nzr = NR + 2 * g;
  D.nf = D.nz + 1;
  D.nfr = D.nzr + 1;
  D.NZ = (int64_t)D.nz * D.nz * D.nzr;
  D.NFC = (int64_t)D.nf * D.nf * D.nfr;
  D.NZ2 = (int64_t)D.nz * D.nz;
  D.gamma = cfg->gamma;
  D.rhofloor = cfg->rhofloor;
  D.prsfloor = cfg->prsfloor;
  D.smallnum = csg1::kSmall;
  // the driver's settings (1000riemann_geomesh_cs_mhd.f:314, 319)
  D.flatten_coef = 0.4;
  D.prs_flatten_coef = 1.5;
  D.grav = cfg->grav;
  D.wall = cfg->radial_bc == CS_RBC_WALL ? 1 : 0;
  D.riemann = cfg->riemann;
  S->cfl = cfg->cfl;
  S->rmin = cfg->rmin;
  S->rmax = cfg->rmax;
  auto fail = [&](int r) {
    destroy(S);
    return r;
  };
  // geometry tables, zone geometry, mask, stencil
  double *lat, *rbd, *TX, *TY, *TZ, *zg;
  int32_t *mask, *sten;
  D.NZB = (D.NZ + 63) / 64;
  if ((rc = dalloc(S, lat, (int64_t)H.lat.size(), bytes)) || (rc = dalloc(S, rbd, (int64_t)H.rb.size(), bytes)) ||
      (rc = dalloc(S, TX, (int64_t)H.TX.size(), bytes)) || (rc = dalloc(S, TY, (int64_t)H.TY.size(), bytes)) ||
      (rc = dalloc(S, TZ, (int64_t)H.TZ.size(), bytes)) || (rc = dalloc(S, zg, 6 * 5 * 64 * D.NZB, bytes)) ||
      (rc = dalloc(S, mask, D.NZ2, bytes)) || (rc = dalloc(S, sten, 24 * D.NZ2, bytes)))
    return fail(rc);
  std::vector<int32_t> hm((size_t)D.NZ2), hs((size_t)D.NZ2 * 24);
  for (int j = 1 - g; j <= N + g; ++j)
    for (int i = 1 - g; i <= N + g; ++i) {
      int on = 1;
      for (int a = 0; a < 4; ++a) {
        double u[3];
        H.G.u(0, i - (a & 1), j - (a >> 1), u);
        if (csg1::len3(u) < csg1::kSmall) on = 0;
      }
      hm[(size_t)(i - 1 + g) + (size_t)D.nz * (j - 1 + g)] = on;
    }
  if ((rc = cs_stencil_indexing(cfg, hs.data()))) return fail(rc);
  if (hipMemcpy(lat, H.lat.data(), sizeof(double) * H.lat.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(rbd, H.rb.data(), sizeof(double) * H.rb.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(TX, H.TX.data(), sizeof(double) * H.TX.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(TY, H.TY.data(), sizeof(double) * H.TY.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(TZ, H.TZ.data(), sizeof(double) * H.TZ.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(mask, hm.data(), sizeof(int32_t) * hm.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(sten, hs.data(), sizeof(int32_t) * hs.size(), hipMemcpyHostToDevice) != hipSuccess) {
    cs::set_error("P1 stage: geometry upload failed");
    return fail(CS_E_HIP);
  }
  D.G = H.G;
  D.G.lat = lat;
  D.G.rb = rbd;
  D.G.TX = TX;
  D.G.TY = TY;
  D.G.TZ = TZ;
  D.zg = zg;
  D.mask = mask;
  D.stencil = sten;
  // work arrays
  const int64_t Z6 = 6 * D.NZ;
  if ((rc = dalloc(S, D.V, 6 * 36 * 64 * D.NZB, bytes)) || (rc = dalloc(S, D.flat, Z6, bytes)) ||
      (rc = dalloc(S, D.work, Z6, bytes)) || (rc = dalloc(S, D.ms, Z6, bytes)) ||
      (rc = dalloc(S, D.Bq, 6 * 3 * 5 * D.NFC, bytes)) || (rc = dalloc(S, D.fa, Z6, bytes)) ||
      (rc = dalloc(S, D.prev, Z6, bytes)) || (rc = dalloc(S, D.next, Z6, bytes)) ||
      (rc = dalloc(S, D.ever, Z6, bytes)) || (rc = dalloc(S, D.inlist, Z6, bytes)) ||
      (rc = dalloc(S, D.list, Z6, bytes)) || (rc = dalloc(S, D.ctr, 64, bytes)))
    return fail(rc);
  S->all = make_box(1 - g, N + g, 1 - g, N + g, 1 - g, NR + g);
  S->wr = make_box(-2, N + 3, -2, N + 3, -2, NR + 3);
  S->own = make_box(1, N, 1, N, 1, NR);
  S->fx = make_box(-3, N + 3, -2, N + 3, -2, NR + 3);
  S->fy = make_box(-2, N + 3, -3, N + 3, -2, NR + 3);
  S->fz = make_box(-2, N + 3, -2, N + 3, -3, NR + 3);
  S->ffx = make_box(-1, N + 1, 0, N + 1, 0, NR + 1);
  S->ffy = make_box(0, N + 1, -1, N + 1, 0, NR + 1);
  S->ffz = make_box(0, N + 1, 0, N + 1, -1, NR + 1);
  S->ex = make_box(1, N, 0, N, 0, NR);
  S->ey = make_box(0, N, 1, N, 0, NR);
  S->ez = make_box(0, N, 0, N, 1, NR);
  k_zone_geo<<<dim3(nb(S->all.count(), 128), 6), 128>>>(D, S->all, zg);
  if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
    cs::set_error("P1 stage: zone geometry kernel failed");
    return fail(CS_E_HIP);
  }
  *out = S;
  return CS_OK;
}

static Buf6 b6(const Buf& b) { return Buf6{b.U, b.ent, b.Bx, b.By, b.Bz}; }

int front(Stage* S, const Buf& in_, double* const* F, double* const* Sp, double* const* E, int* err,
          hipStream_t st) {
  const Dev& D = S->D;
  const Buf6 in = b6(in_);
  const int64_t nall = S->all.count(), nwr = S->wr.count();
  mark_begin(S, st);
  k_prim<<<dim3(nb(nall, 256), 6), 256, 0, st>>>(D, S->all, in);
  mark_end(S, st, CS_P1K_PRIM);
  mark_begin(S, st);
  k_face_zero<<<dim3(nb(D.NFC, 256), 18), 256, 0, st>>>(D, in);
  k_tricorner<<<dim3(nb(4 * (int64_t)D.g * D.nzr, 256), 6), 256, 0, st>>>(D, in);
  mark_end(S, st, CS_P1K_FACE_SYNC);
  mark_begin(S, st);
  k_flat_a<<<dim3(tile_blocks<32, 8>(S->all), 6), 256, 0, st>>>(D, S->all);
  k_flat_b<<<dim3(nb(nall, 256), 6), 256, 0, st>>>(D, S->all);
  k_flat_c<<<dim3(nb(nall, 256), 6), 256, 0, st>>>(D, S->all);
  k_ctr_reset<<<1, 64, 0, st>>>(D);
  mark_end(S, st, CS_P1K_FLATTEN);
  mark_begin(S, st);
  k_weno<<<dim3(nb(nwr, 128), 6), 128, 0, st>>>(D, S->wr);
  mark_end(S, st, CS_P1K_WENO);
  mark_begin(S, st);
  k_pos<<<dim3(nb(nwr, 256), 6), 256, 0, st>>>(D, S->wr);
  mark_end(S, st, CS_P1K_POSITIVITY);
  mark_begin(S, st);
  for (int sw = 0; sw < kSweeps; ++sw) {
    k_cands<<<dim3(nb(nwr, 256), 6), 256, 0, st>>>(D, S->wr);
    k_weno_list<<<kListBlocks, 128, 0, st>>>(D);
    k_pos_list<<<kListBlocks, 256, 0, st>>>(D, sw);
    k_commit<<<kListBlocks, 256, 0, st>>>(D, sw);
  }
  k_clear_ent<<<dim3(nb(nwr, 256), 6), 256, 0, st>>>(D, S->wr, in, err);
  mark_end(S, st, CS_P1K_REPLAY);
  const int64_t nfpi = (D.N + 7 + kFpTI - 2) / (kFpTI - 1), nfpj = (D.N + 7 + kFpTJ - 2) / (kFpTJ - 1);
  const int fpl = D.NR + 8, fpch = (fpl + kFpLevels - 1) / kFpLevels, fpnl = (fpl + fpch - 1) / fpch;
  mark_begin(S, st);
  k_face_pts<<<dim3((unsigned)(nfpi * nfpj * fpch), 6), kFpTI * kFpTJ, 0, st>>>(D, in, fpch, fpnl);
  mark_end(S, st, CS_P1K_FACE_POINTS);
  mark_begin(S, st);
  k_clsq<<<dim3(nb(nall, 64), 6), 64, 0, st>>>(D, S->all, in);
  mark_end(S, st, CS_P1K_CLSQ);
  const int64_t nfi = (D.N + 3 + kFluxTI - 2) / (kFluxTI - 1), nfj = (D.N + 3 + kFluxTJ - 2) / (kFluxTJ - 1);
  const int fll = D.NR + 4, flch = (fll + kFluxLevels - 1) / kFluxLevels, flnl = (fll + flch - 1) / flch;
  mark_begin(S, st);
  k_flux<<<dim3((unsigned)(nfi * nfj * flch), 6), kFluxTI * kFluxTJ, 0, st>>>(D, F[0], F[1], F[2], Sp[0], Sp[1],
                                                                              Sp[2], flch, flnl);
  mark_end(S, st, CS_P1K_FLUX);
  const unsigned ned = std::max(tile_blocks<32, 8>(S->ex), std::max(tile_blocks<32, 8>(S->ey), tile_blocks<32, 8>(S->ez)));
  mark_begin(S, st);
  k_edge<<<dim3(ned, 18), 256, 0, st>>>(D, S->ex, S->ey, S->ez, Sp[0], Sp[1], Sp[2], E[0], E[1], E[2]);
  mark_end(S, st, CS_P1K_EDGE);
  P1S_HIP(hipGetLastError());
  return CS_OK;
}

int update(Stage* S, const Buf& in, const Buf& u0, const Buf& out, double* const* F, double* const* E,
           const double* dtp, double wa, double wb, int mode, hipStream_t st) {
  const Dev& D = S->D;
  const unsigned nfc = tile_blocks<32, 8>(make_box(0, D.N, 0, D.N, 0, D.NR));  // covers every family's box
  mark_begin(S, st);
  k_upd_face<<<dim3(nfc, 18), 256, 0, st>>>(D, b6(in), b6(u0), b6(out), E[0], E[1], E[2], dtp, wa, wb);
  mark_end(S, st, CS_P1K_UPDATE_FACES);
  mark_begin(S, st);
  k_upd_zone<<<dim3(tile_blocks<32, 8>(S->own), 6), 256, 0, st>>>(D, S->own, b6(in), b6(u0), b6(out), F[0], F[1], F[2],
                                                                dtp, wa, wb, mode);
  mark_end(S, st, CS_P1K_UPDATE_ZONES);
  P1S_HIP(hipGetLastError());
  return CS_OK;
}

int imex_source(Stage* S, const Buf& un, const Buf& u1, const double* dtp, hipStream_t st) {
  const Dev& D = S->D;
  mark_begin(S, st);
  k_imex<<<dim3(nb(S->all.count(), 256), 6), 256, 0, st>>>(D, S->all, b6(un), b6(u1), dtp);
  const size_t fb = sizeof(double) * (size_t)D.NFC * 6;
  P1S_HIP(hipMemcpyAsync(u1.Bx, un.Bx, fb, hipMemcpyDeviceToDevice, st));
  P1S_HIP(hipMemcpyAsync(u1.By, un.By, fb, hipMemcpyDeviceToDevice, st));
  P1S_HIP(hipMemcpyAsync(u1.Bz, un.Bz, fb, hipMemcpyDeviceToDevice, st));
  mark_end(S, st, CS_P1K_IMEX);
  P1S_HIP(hipGetLastError());
  return CS_OK;
}

int nblocks_cfl() { return kCflBlocks; }

int pad_radial(Stage* S, const Buf& b, hipStream_t st) {
  const Dev& D = S->D;
  k_pad6<<<dim3(nb((int64_t)D.nz * D.nz * D.g, 256), 6), 256, 0, st>>>(D, b6(b));
  P1S_HIP(hipGetLastError());
  return CS_OK;
}

int init_problem(Stage* S, const Buf& out, int problem, uint64_t seed, double time, int layers, hipStream_t st) {
  if (!cspg::known(problem)) {
    cs::set_error("cs_init_problem_p1: unknown problem %d (0, 2, 102, 103)", problem);
    return CS_E_INVALID;
  }
  if (layers < 0 || layers > 3) {
    cs::set_error("cs_add_boundary_p1: layers must be 1 (inner), 2 (outer) or 3");
    return CS_E_INVALID;
  }
  const Dev& D = S->D;
  const cspg::Problem P{problem, time, S->rmin, S->rmax, D.gamma, D.rhofloor, D.prsfloor, D.grav, seed};
  if (!layers) {
    const size_t zb = sizeof(double) * (size_t)D.NZ * 6, fb = sizeof(double) * (size_t)D.NFC * 6;
    P1S_HIP(hipMemsetAsync(out.U, 0, zb * 8, st));
    P1S_HIP(hipMemsetAsync(out.ent, 0, zb, st));
    P1S_HIP(hipMemsetAsync(out.Bx, 0, fb, st));
    P1S_HIP(hipMemsetAsync(out.By, 0, fb, st));
    P1S_HIP(hipMemsetAsync(out.Bz, 0, fb, st));
  }
  k_init_zone<<<dim3(nb(S->all.count(), 128), 6), 128, 0, st>>>(D, S->all, b6(out), P, layers);
  k_init_face<<<dim3(nb(D.NFC, 128), 18), 128, 0, st>>>(D, b6(out), P, layers);
  P1S_HIP(hipGetLastError());
  return CS_OK;
}

int diag_blocks(Stage* S) { return (int)nb(6 * (int64_t)S->D.N * S->D.N * S->D.NR, kDiagThreads); }

int divb_partial(Stage* S, const Buf& cur, double* part, hipStream_t st) {
  k_divb<<<diag_blocks(S), kDiagThreads, 0, st>>>(S->D, b6(cur), part);
  P1S_HIP(hipGetLastError());
  return CS_OK;
}

int sums_partial(Stage* S, const Buf& cur, double* part, hipStream_t st) {
  k_sums<<<diag_blocks(S), kDiagThreads, 0, st>>>(S->D, b6(cur), part);
  P1S_HIP(hipGetLastError());
  return CS_OK;
}

int cfl_partial(Stage* S, const Buf& cur, double* part, hipStream_t st) {
  mark_begin(S, st);
  k_cfl<<<kCflBlocks, kCflThreads, 0, st>>>(S->D, b6(cur), S->cfl, part);
  mark_end(S, st, CS_P1K_CFL);
  P1S_HIP(hipGetLastError());
  return CS_OK;
}

}  // namespace p1s
