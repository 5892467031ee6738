// cs_kernels.hip -- HIP kernels for gfx950 and the device half of the C-ABI.
//
// One handle holds all six cube sectors of one shell resident in HBM in the
// reference's panel-major Fortran layout (x fastest).  A stage is
//   halo gather -> prims -> face fluxes (x,y,z) -> edge EMFs (x,y,z)
//   -> cell update -> face (CT) update -> cleanup,
// and a step ends with the CFL min-reduction; dt stays on the device so a
// whole step can be captured in one hipGraph.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/cubesphere_hip.h"
#include "cs_halo_maps.h"
#include "cs_p1_stage.h"
#include "cs_internal.h"
#include "cs_dump.h"
#include "cs_physics.h"

using namespace csp;

namespace cs {

static thread_local std::string g_err;
void set_error(const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
}
int validate_config(const cs_config* c);
int host_metric(const cs_config* c, std::vector<double>& rb, std::vector<double>& T);

#define CS_HIP(call)                                                                    \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess) {                                                             \
      set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, __LINE__); \
      return CS_E_HIP;                                                                  \
    }                                                                                   \
  } while (0)

// ------------------------------------------------------------------ ranges
// Work-item order: i fastest (coalesced rows), then j then k, or, with kj
// set, k before j.  The k-before-j order suits kernels whose neighbours lie
// along i and k only and that read per-column 2-D metric entries: the radial
// levels of one column follow each other a row apart, so the column's metric
// (and the neighbour row in k) is still in L2 when the next level reads it.
struct Rng {
  int i0, ni, j0, nj, k0, nk;
  uint32_t total;
  int kj;
};
static Rng make_rng(int i0, int i1, int j0, int j1, int k0, int k1, int kj = 0) {
  Rng r;
  r.i0 = i0; r.ni = i1 - i0 + 1;
  r.j0 = j0; r.nj = j1 - j0 + 1;
  r.k0 = k0; r.nk = k1 - k0 + 1;
  r.total = (r.ni > 0 && r.nj > 0 && r.nk > 0) ? (uint32_t)((uint64_t)r.ni * r.nj * r.nk * 6) : 0u;
  r.kj = kj;
  return r;
}
__device__ __forceinline__ bool decode(const Rng& r, uint32_t t, int& i, int& j, int& k, int& s) {
  if (t >= r.total) return false;
  uint32_t q = t / (uint32_t)r.ni;
  i = r.i0 + (int)(t - q * (uint32_t)r.ni);
  if (r.kj) {
    uint32_t q2 = q / (uint32_t)r.nk;
    k = r.k0 + (int)(q - q2 * (uint32_t)r.nk);
    uint32_t q3 = q2 / (uint32_t)r.nj;
    j = r.j0 + (int)(q2 - q3 * (uint32_t)r.nj);
    s = (int)q3;
  } else {
    uint32_t q2 = q / (uint32_t)r.nj;
    j = r.j0 + (int)(q - q2 * (uint32_t)r.nj);
    uint32_t q3 = q2 / (uint32_t)r.nk;
    k = r.k0 + (int)(q2 - q3 * (uint32_t)r.nk);
    s = (int)q3;
  }
  return true;
}
// Workgroups are handed to the 8 XCDs of an MI355X round robin (workgroup b
// runs on XCD b mod 8), and each XCD has its own L2.  Renumbering them so that
// XCD x takes one contiguous range of the index space keeps the rows a
// stencil shares (j-2..j+3 of a y-sweep, the two rows of a face pair) inside
// one L2 instead of fetching them from HBM once per XCD.  The map is a
// bijection of [0, gridDim.x) for any grid size.
constexpr uint32_t kXcds = 8;
__device__ __forceinline__ uint32_t xcd_block() {
  const uint32_t nb = gridDim.x, b = blockIdx.x;
  const uint32_t q0 = nb / kXcds, r = nb % kXcds;
  const uint32_t x = b % kXcds, q = b / kXcds;
  return x * q0 + (x < r ? x : r) + q;
}
__device__ __forceinline__ uint32_t gtid() { return xcd_block() * blockDim.x + threadIdx.x; }

// ------------------------------------------------------------------ device views
struct Geo {
  const double* T;   // 2-D tables
  const double* rb;  // rb[k + g]
  int64_t plane;     // t2 * 6
  int64_t t2;
  __device__ __forceinline__ double t(const CsDims& d, int f, int s, int i, int j) const {
    return T[f * plane + s * t2 + d.t(i, j)];
  }
  __device__ __forceinline__ double r(const CsDims& d, int k) const { return rb[k + d.g]; }
};

struct State {
  double* U;    // [s][8][cell]
  double* ent;  // [s][cell]
  double* Bx;   // [s][face]
  double* By;
  double* Bz;
};

struct Params {
  double gamma, gm1, rhofloor, prsfloor, cfl, grav, igm1;
  int recon, riemann, emf, wall, imex, surf;
  double om[3];  // rotating frame (omega_passive)
  int rot;       // any component of om nonzero
};

// sector neighbours right, left, up, down (203fill_ghost2d_global.f:37-75), 0-based
__constant__ int c_nbr[6][4] = {{2, 5, 1, 4}, {0, 3, 2, 5}, {1, 4, 0, 3},
                                {4, 1, 5, 2}, {5, 2, 3, 0}, {3, 0, 4, 1}};

// ------------------------------------------------------------------ halo exchange
// Coinciding panel-edge faces: one thread per physical face writes both copies
// with the averaging of 330fetch_buffers_fill_in_ux.f:337-350,
// new = (mine + (-theirs)) * 0.5, which is negation-consistent bit for bit.
__global__ void k_avg_coinciding(CsDims d, State st) {
  uint32_t t = gtid();
  const int N = d.N, NR = d.NR;
  uint32_t per = (uint32_t)N * NR;
  if (t >= 12u * per) return;
  int pair = t / per;
  uint32_t rem = t - pair * per;
  int jj = 1 + (int)(rem % N);
  int k = 1 + (int)(rem / N);
  double *pa, *pb;
  if (pair < 6) {  // right edge of sector `pair` with the top edge of its right neighbour
    int s = pair, o = c_nbr[s][0];
    pa = st.Bx + s * d.fpp + d.face(N, jj, k);
    pb = st.By + o * d.fpp + d.face(jj, N, k);
  } else if (pair < 9) {  // left edges 1-6, 2-4, 3-5 (1-based)
    int s = pair - 6, o = c_nbr[s][1];
    pa = st.Bx + s * d.fpp + d.face(0, jj, k);
    pb = st.Bx + o * d.fpp + d.face(0, N + 1 - jj, k);
  } else {  // bottom edges 1-5, 2-6, 3-4
    int s = pair - 9, o = c_nbr[s][3];
    pa = st.By + s * d.fpp + d.face(jj, 0, k);
    pb = st.By + o * d.fpp + d.face(N + 1 - jj, 0, k);
  }
  double a = *pa, b = *pb;
  *pa = (a + (-b)) * 0.5;
  *pb = (b + (-a)) * 0.5;
}

// Zone-centred ghost strips (330-333 index flips): right, left, up, down.
template <typename P>
__device__ __forceinline__ void exchange_cells(const CsDims& d, const State& st, P* V, uint32_t t) {
  const int N = d.N, NR = d.NR, g = d.g;
  uint32_t per = (uint32_t)g * N * NR;
  int s = t / (4 * per);
  uint32_t r = t - s * 4 * per;
  int strip = r / per;
  r -= strip * per;
  int a = 1 + (int)(r % g);
  r /= g;
  int tt = 1 + (int)(r % N);
  int k = 1 + (int)(r / N);
  int o = c_nbr[s][strip];
  int di, dj, si, sj;
  ghost_zone_map(N, strip, a, tt, di, dj, si, sj);
  int64_t dc = d.cell(di, dj, k), sc = d.cell(si, sj, k);
  double v[9];  // all loads in flight before the first store
  P w[9];       // primitives: the same function of the same zone
  for (int c = 0; c < 8; ++c) v[c] = st.U[(o * 8 + c) * d.cpp + sc];
  v[8] = st.ent[o * d.cpp + sc];
  if (V)
    for (int c = 0; c < 9; ++c) w[c] = V[((int64_t)o * 9 + c) * d.cpp + sc];
  for (int c = 0; c < 8; ++c) st.U[(s * 8 + c) * d.cpp + dc] = v[c];
  st.ent[s * d.cpp + dc] = v[8];
  if (V)
    for (int c = 0; c < 9; ++c) V[((int64_t)s * 9 + c) * d.cpp + dc] = w[c];
}

// Face-centred ghost strips with the face-B swap and sign rules
// (330fetch_buffers_fill_in_ux.f:232-294 and the y/z counterparts).
__device__ __forceinline__ void exchange_faces(const CsDims& d, const State& st, uint32_t t) {
  const int N = d.N, NR = d.NR, g = d.g;
  // per (panel, strip, a): 3 families x (N+1) x (NR+1) slots
  uint32_t per = 3u * (N + 1) * (NR + 1);
  int s = t / (4u * g * per);
  uint32_t r = t - s * (4u * g * per);
  int strip = r / (g * per);
  r -= strip * (g * per);
  int a = 1 + (int)(r / per);
  r -= (a - 1) * per;
  int fam = r / ((N + 1) * (NR + 1));
  r -= fam * ((N + 1) * (NR + 1));
  int tt = (int)(r % (N + 1));
  int k = (int)(r / (N + 1));
  int o = c_nbr[s][strip];
  // fam 0: the face family normal to the strip direction, t in [1,N], k in [1,NR]
  // fam 1: the tangential family, t in [0,N], k in [1,NR]
  // fam 2: Bzf, t in [1,N], k in [0,NR]
  if (fam == 0 && (tt < 1 || k < 1)) return;
  if (fam == 1 && k < 1) return;
  if (fam == 2 && tt < 1) return;
  int dfam, di, dj, sfam, si, sj;
  double sign;
  ghost_face_map(N, strip, a, fam, tt, dfam, di, dj, sfam, si, sj, sign);
  const double* src = sfam == 0 ? st.Bx : (sfam == 1 ? st.By : st.Bz);
  double* dst = dfam == 0 ? st.Bx : (dfam == 1 ? st.By : st.Bz);
  dst[s * d.fpp + d.face(di, dj, k)] = sign * src[o * d.fpp + d.face(si, sj, k)];
}

// Corner ghost blocks (no sector owns them: the tri-corners) take the value of
// the panel's own corner zone.
template <typename P>
__device__ __forceinline__ void corner_fill(const CsDims& d, const State& st, P* V, uint32_t t) {
  const int N = d.N, NR = d.NR, g = d.g;
  uint32_t per = (uint32_t)g * g * NR;
  int s = t / (4 * per);
  uint32_t r = t - s * 4 * per;
  int corner = r / per;
  r -= corner * per;
  int a = 1 + (int)(r % g);
  r /= g;
  int b = 1 + (int)(r % g);
  int k = 1 + (int)(r / g);
  int ci = (corner & 1) ? N : 1, cj = (corner & 2) ? N : 1;
  int di = (corner & 1) ? N + a : 1 - a, dj = (corner & 2) ? N + b : 1 - b;
  int64_t dc = d.cell(di, dj, k), sc = d.cell(ci, cj, k);
  double v[9];
  P w[9];
  for (int c = 0; c < 8; ++c) v[c] = st.U[(s * 8 + c) * d.cpp + sc];
  v[8] = st.ent[s * d.cpp + sc];
  if (V)
    for (int c = 0; c < 9; ++c) w[c] = V[((int64_t)s * 9 + c) * d.cpp + sc];
  for (int c = 0; c < 8; ++c) st.U[(s * 8 + c) * d.cpp + dc] = v[c];
  st.ent[s * d.cpp + dc] = v[8];
  if (V)
    for (int c = 0; c < 9; ++c) V[((int64_t)s * 9 + c) * d.cpp + dc] = w[c];
}

// Thread counts of the three parts of the lateral exchange.
CS_HD void exchange_counts(const CsDims& d, uint32_t& nc, uint32_t& nf, uint32_t& nk) {
  nc = 24u * d.g * d.N * d.NR;
  nf = 6u * 4u * d.g * 3u * (d.N + 1) * (d.NR + 1);
  nk = 24u * d.g * d.g * d.NR;
}

// The lateral ghost exchange in one launch: zone strips, face strips and the
// corner blocks read only owned zones and faces and write only ghosts, so
// the three parts are independent.  Runs after the coinciding-face average.
// The zone primitives V travel with U, so no ghost zone is converted again.
template <typename P>
__global__ void k_exchange(CsDims d, State st, P* V) {
  uint32_t t = gtid(), nc, nf, nk;
  exchange_counts(d, nc, nf, nk);
  if (t < nc) { exchange_cells<P>(d, st, V, t); return; }
  t -= nc;
  if (t < nf) { exchange_faces(d, st, t); return; }
  t -= nf;
  if (t < nk) corner_fill<P>(d, st, V, t);
}

// Radial ghost layers: zero-gradient pad (826pad_boundary_r.f:201-214 inner
// form, used at both ends) or a reflecting, perfectly conducting wall.
template <typename P>
__device__ __forceinline__ void radial_bc(const CsDims& d, const Geo& geo, const State& st, int wall, P* V,
                                          const Params& pr, uint32_t t) {
  // no FMA contraction: the wall mirror rounds exactly as written, so the whole
  // halo (exchange + boundary) is reproducible bit for bit on any host
#pragma clang fp contract(off)
  const int N = d.N, NR = d.NR, g = d.g;
  const int nl = N + 2 * g + 1;  // lateral index range [-g, N+g]
  uint32_t per = (uint32_t)nl * nl;
  if (t >= 6u * 2u * g * per) return;
  int s = t / (2u * g * per);
  uint32_t r = t - s * 2u * g * per;
  int side = r / (g * per);
  r -= side * g * per;
  int a = 1 + (int)(r / per);
  r -= (a - 1) * per;
  int i = -g + (int)(r % nl), j = -g + (int)(r / nl);
  // faces: Bxf (x-vertex i, cell j), Byf (cell i, y-vertex j), Bzf (cell i, cell j)
  // Sources always lie in the interior, so no thread reads a ghost another
  // thread writes; a wall mirror of a thin shell (nr < g) reflects repeatedly
  // between the two walls.
  int kg = side == 0 ? 1 - a : NR + a;  // ghost cell layer
  int ks = side == 0 ? 1 : NR;
  int sgn = 1;
  if (wall) {
    ks = kg;
    while (ks < 1 || ks > NR) { ks = ks < 1 ? 1 - ks : 2 * NR + 1 - ks; sgn = -sgn; }
  }
  if (j >= 1 - g) st.Bx[s * d.fpp + d.face(i, j, kg)] = st.Bx[s * d.fpp + d.face(i, j, ks)];
  if (i >= 1 - g) st.By[s * d.fpp + d.face(i, j, kg)] = st.By[s * d.fpp + d.face(i, j, ks)];
  if (i >= 1 - g && j >= 1 - g) {
    int kzg = side == 0 ? -a : NR + a;  // ghost z-vertex
    int kzs = side == 0 ? 0 : NR, zs = 1;
    if (wall) {
      kzs = kzg;
      while (kzs < 0 || kzs > NR) { kzs = kzs < 0 ? -kzs : 2 * NR - kzs; zs = -zs; }
    }
    double bz = st.Bz[s * d.fpp + d.face(i, j, kzs)];
    st.Bz[s * d.fpp + d.face(i, j, kzg)] = zs < 0 ? -bz : bz;
    int64_t dc = d.cell(i, j, kg), sc = d.cell(i, j, ks);
    double u[8];
    for (int c = 0; c < 8; ++c) u[c] = st.U[(s * 8 + c) * d.cpp + sc];
    if (sgn < 0) {
      // mirror about the unit normal of the radial faces, so a wall face sees
      // exactly mirrored states (no mass flux); the zone-centre direction
      // stands in for it in the degenerate tri-corner ghost blocks
      double zn = geo.t(d, T_ZSN, s, i, j);
      double n0, n1, n2;
      if (zn > 1e-300 && zn < 1e300) {
        n0 = geo.t(d, T_ZSX, s, i, j) / zn; n1 = geo.t(d, T_ZSY, s, i, j) / zn; n2 = geo.t(d, T_ZSZ, s, i, j) / zn;
      } else {
        n0 = geo.t(d, T_ZCX, s, i, j); n1 = geo.t(d, T_ZCY, s, i, j); n2 = geo.t(d, T_ZCZ, s, i, j);
      }
      double m = u[1] * n0 + u[2] * n1 + u[3] * n2;
      u[1] -= 2.0 * m * n0; u[2] -= 2.0 * m * n1; u[3] -= 2.0 * m * n2;
      double b = u[5] * n0 + u[6] * n1 + u[7] * n2;
      u[5] -= 2.0 * b * n0; u[6] -= 2.0 * b * n1; u[7] -= 2.0 * b * n2;
    }
    for (int c = 0; c < 8; ++c) st.U[(s * 8 + c) * d.cpp + dc] = u[c];
    double ent = st.ent[s * d.cpp + sc];
    st.ent[s * d.cpp + dc] = ent;
    if (!V) return;  // no primitives kept (the P1 stage converts every zone itself)
    // primitives of the ghost zone, as k_prims converts
    Gas<double> gs{pr.gamma, pr.gm1, pr.rhofloor, pr.prsfloor, pr.igm1};
    double v[8];
    cons_to_prim(gs, u, v, ent);
    for (int c = 0; c < 8; ++c) V[((int64_t)s * 9 + c) * d.cpp + dc] = (P)v[c];
    V[((int64_t)s * 9 + 8) * d.cpp + dc] = (P)ent;
  }
}
template <typename P>
__global__ void k_radial_bc(CsDims d, Geo geo, State st, int wall, P* V, Params pr) {
  radial_bc<P>(d, geo, st, wall, V, pr, gtid());
}

// ------------------------------------------------------------------ primitives
template <typename P>
__global__ void k_prims(CsDims d, Rng rg, State st, P* V, Params pr) {
  int i, j, k, s;
  if (!decode(rg, gtid(), i, j, k, s)) return;
  int64_t c = d.cell(i, j, k);
  double u[8], v[8];
  for (int q = 0; q < 8; ++q) u[q] = st.U[(s * 8 + q) * d.cpp + c];
  double ent = st.ent[s * d.cpp + c];
  Gas<double> gs{pr.gamma, pr.gm1, pr.rhofloor, pr.prsfloor, pr.igm1};
  cons_to_prim(gs, u, v, ent);
  for (int q = 0; q < 8; ++q) V[(s * 9 + q) * d.cpp + c] = (P)v[q];
  V[(s * 9 + 8) * d.cpp + c] = (P)ent;
}

// One face: rotate the reconstructed states into the face frame (A normal,
// Bv, Cv), put the facial B in the normal slot, Riemann-solve, rotate the
// resolved flux back to the global frame (450make_facial_fluxes.f:890-1030).
template <typename R>
__device__ __forceinline__ void solve_face(const R* VL, const R* VR, const double* A, const double* Bv,
                                           const double* Cv, double bface, const Params& pr, double* out,
                                           R& SL, R& SR) {
  R a0 = (R)A[0], a1 = (R)A[1], a2 = (R)A[2], b0 = (R)Bv[0], b1 = (R)Bv[1], b2 = (R)Bv[2],
    c0 = (R)Cv[0], c1 = (R)Cv[1], c2 = (R)Cv[2];
  R wl[8], wr[8];
  wl[0] = VL[0]; wr[0] = VR[0];
  wl[4] = VL[4]; wr[4] = VR[4];
  wl[1] = a0 * VL[1] + a1 * VL[2] + a2 * VL[3];
  wl[2] = b0 * VL[1] + b1 * VL[2] + b2 * VL[3];
  wl[3] = c0 * VL[1] + c1 * VL[2] + c2 * VL[3];
  wr[1] = a0 * VR[1] + a1 * VR[2] + a2 * VR[3];
  wr[2] = b0 * VR[1] + b1 * VR[2] + b2 * VR[3];
  wr[3] = c0 * VR[1] + c1 * VR[2] + c2 * VR[3];
  wl[5] = (R)bface; wr[5] = (R)bface;
  wl[6] = b0 * VL[5] + b1 * VL[6] + b2 * VL[7];
  wl[7] = c0 * VL[5] + c1 * VL[6] + c2 * VL[7];
  wr[6] = b0 * VR[5] + b1 * VR[6] + b2 * VR[7];
  wr[7] = c0 * VR[5] + c1 * VR[6] + c2 * VR[7];
  Gas<R> gs{(R)pr.gamma, (R)pr.gm1, (R)pr.rhofloor, (R)pr.prsfloor, (R)pr.igm1};
  R Fs[8], entF;
  if (pr.riemann == CS_RS_HLLD)
    hlld<R>(gs, wl, wr, VL[8], VR[8], Fs, entF, SL, SR);
  else
    hll<R>(gs, wl, wr, VL[8], VR[8], Fs, entF, SL, SR);
  // rotate back (450make_facial_fluxes.f:995-1030); the caller scales by the area
  out[0] = (double)Fs[0];
  out[4] = (double)Fs[4];
  out[1] = A[0] * (double)Fs[1] + Bv[0] * (double)Fs[2] + Cv[0] * (double)Fs[3];
  out[2] = A[1] * (double)Fs[1] + Bv[1] * (double)Fs[2] + Cv[1] * (double)Fs[3];
  out[3] = A[2] * (double)Fs[1] + Bv[2] * (double)Fs[2] + Cv[2] * (double)Fs[3];
  out[5] = A[0] * (double)Fs[5] + Bv[0] * (double)Fs[6] + Cv[0] * (double)Fs[7];
  out[6] = A[1] * (double)Fs[5] + Bv[1] * (double)Fs[6] + Cv[1] * (double)Fs[7];
  out[7] = A[2] * (double)Fs[5] + Bv[2] * (double)Fs[6] + Cv[2] * (double)Fs[7];
  out[8] = (double)entF;
}

// ------------------------------------------------------------------ face fluxes
// DIR 0: x-faces (x-vertex i, cells i and i+1), 1: y, 2: z.
// Output per face: flux(9) * area in the global frame, speeds SL, SR.
template <int DIR, typename R, typename P>
__device__ __forceinline__ void face_flux(const CsDims& d, int i, int j, int k, int s, const Geo& geo,
                                          const State& st, const P* __restrict__ V, double* __restrict__ F,
                                          double* __restrict__ S, const Params& pr) {
  if (DIR == 2 && (i == 0 || i == d.N + 1) && (j == 0 || j == d.N + 1)) {
    // radial faces of the tri-corner ghost columns have zero area (two of the
    // four vertices coincide); nothing reads them
    int64_t fo = (int64_t)s * 9 * d.fpp + d.face(i, j, k);
    for (int q = 0; q < 9; ++q) F[fo + q * d.fpp] = 0.0;
    int64_t so = (int64_t)s * 2 * d.fpp + d.face(i, j, k);
    S[so] = 0.0;
    S[so + d.fpp] = 0.0;
    return;
  }
  const int64_t stride = DIR == 0 ? 1 : (DIR == 1 ? d.nc : (int64_t)d.nc * d.nc);
  // left cell of the face
  int64_t cl = DIR == 0 ? d.cell(i, j, k) : (DIR == 1 ? d.cell(i, j, k) : d.cell(i, j, k));
  const P* Vs = V + (int64_t)s * 9 * d.cpp;
  // Faces in the ghost rows/columns (outside [1,N] transversally) only supply
  // wave speeds to the EMFs of the panel-edge edges; their stencils would
  // reach into the tri-corner ghost blocks, so they use the first-order states.
  const bool weno = pr.recon == CS_RECON_WENO5 &&
                    (DIR == 0 ? (j >= 1 && j <= d.N)
                              : DIR == 1 ? (i >= 1 && i <= d.N) : (i >= 1 && i <= d.N && j >= 1 && j <= d.N));
  R VL[9], VR[9];
  for (int q = 0; q < 9; ++q) {
    const P* p = Vs + q * d.cpp + cl;
    if (weno) {
      R m2 = (R)p[-2 * stride], m1 = (R)p[-stride], c0 = (R)p[0], p1 = (R)p[stride],
        p2 = (R)p[2 * stride], p3 = (R)p[3 * stride];
      VL[q] = weno5<R>(m2, m1, c0, p1, p2);
      VR[q] = weno5<R>(p3, p2, p1, c0, m1);
    } else {
      VL[q] = (R)p[0];
      VR[q] = (R)p[stride];
    }
  }
  if (!(VL[0] > R(0)) || !(VL[4] > R(0)) || !(VL[8] > R(0)))
    for (int q = 0; q < 9; ++q) VL[q] = (R)Vs[q * d.cpp + cl];
  if (!(VR[0] > R(0)) || !(VR[4] > R(0)) || !(VR[8] > R(0)))
    for (int q = 0; q < 9; ++q) VR[q] = (R)Vs[q * d.cpp + cl + stride];
  // face frame (A normal, B, C) and area
  double A[3], Bv[3], Cv[3], area, bface;
  int64_t f = s * d.fpp + d.face(i, j, k);
  if (DIR == 0) {
    for (int c = 0; c < 3; ++c) { A[c] = geo.t(d, T_XNX + c, s, i, j); Cv[c] = geo.t(d, T_XCX + c, s, i, j); }
    Bv[0] = Cv[1] * A[2] - Cv[2] * A[1];
    Bv[1] = Cv[2] * A[0] - Cv[0] * A[2];
    Bv[2] = Cv[0] * A[1] - Cv[1] * A[0];
    double r1 = geo.r(d, k), r0 = geo.r(d, k - 1);
    area = 0.5 * (r1 * r1 - r0 * r0) * geo.t(d, T_THY, s, i, j);
    bface = st.Bx[f];
  } else if (DIR == 1) {
    for (int c = 0; c < 3; ++c) { A[c] = geo.t(d, T_YNX + c, s, i, j); Cv[c] = geo.t(d, T_YCX + c, s, i, j); }
    Bv[0] = Cv[1] * A[2] - Cv[2] * A[1];
    Bv[1] = Cv[2] * A[0] - Cv[0] * A[2];
    Bv[2] = Cv[0] * A[1] - Cv[1] * A[0];
    double r1 = geo.r(d, k), r0 = geo.r(d, k - 1);
    area = 0.5 * (r1 * r1 - r0 * r0) * geo.t(d, T_THX, s, i, j);
    bface = st.By[f];
  } else {
    double zn = geo.t(d, T_ZSN, s, i, j);
    for (int c = 0; c < 3; ++c) { A[c] = geo.t(d, T_ZSX + c, s, i, j) / zn; Bv[c] = geo.t(d, T_ZTX + c, s, i, j); }
    Cv[0] = A[1] * Bv[2] - A[2] * Bv[1];
    Cv[1] = A[2] * Bv[0] - A[0] * Bv[2];
    Cv[2] = A[0] * Bv[1] - A[1] * Bv[0];
    double r1 = geo.r(d, k);
    area = r1 * r1 * zn;
    bface = st.Bz[f];
  }
  double out[9];
  R SL, SR;
  solve_face<R>(VL, VR, A, Bv, Cv, bface, pr, out, SL, SR);
  int64_t fo = (int64_t)s * 9 * d.fpp + d.face(i, j, k);
  for (int q = 0; q < 9; ++q) F[fo + q * d.fpp] = out[q] * area;
  int64_t so = (int64_t)s * 2 * d.fpp + d.face(i, j, k);
  S[so] = (double)SL;
  S[so + d.fpp] = (double)SR;
}

template <int DIR, typename R, typename P>
__global__ void __launch_bounds__(256, 2) k_face_flux(CsDims d, Rng rg, Geo geo, State st, const P* __restrict__ V,
                                                      double* __restrict__ F, double* __restrict__ S, Params pr) {
  int i, j, k, s;
  if (!decode(rg, gtid(), i, j, k, s)) return;
  face_flux<DIR, R, P>(d, i, j, k, s, geo, st, V, F, S, pr);
}

// x- and y-faces of a stage in one launch (the two sweeps are independent)
template <typename R, typename P>
__global__ void __launch_bounds__(256, 2) k_face_flux_xy(CsDims d, Rng rx, Rng ry, Geo geo, State st,
                                                         const P* __restrict__ V, double* __restrict__ Fx,
                                                         double* __restrict__ Sx, double* __restrict__ Fy,
                                                         double* __restrict__ Sy, Params pr) {
  const uint32_t t = gtid();
  int i, j, k, s;
  if (t < rx.total) {
    decode(rx, t, i, j, k, s);
    face_flux<0, R, P>(d, i, j, k, s, geo, st, V, Fx, Sx, pr);
  } else if (decode(ry, t - rx.total, i, j, k, s)) {
    face_flux<1, R, P>(d, i, j, k, s, geo, st, V, Fy, Sy, pr);
  }
}

// ------------------------------------------------------------------ edge EMFs
template <typename P>
__device__ __forceinline__ void cell_E(const CsDims& d, const P* V, int s, int64_t c, double* E) {
  const P* p = V + (int64_t)s * 9 * d.cpp + c;
  double vx = p[1 * d.cpp], vy = p[2 * d.cpp], vz = p[3 * d.cpp];
  double bx = p[5 * d.cpp], by = p[6 * d.cpp], bz = p[7 * d.cpp];
  E[0] = -(vy * bz - vz * by);
  E[1] = -(vz * bx - vx * bz);
  E[2] = -(vx * by - vy * bx);
}
__device__ __forceinline__ double sp(const CsDims& d, const double* S, int s, int which, int i, int j, int k) {
  return S[(int64_t)s * 2 * d.fpp + which * d.fpp + d.face(i, j, k)];
}

// z-edges (MAKE_EDGE_EFIELDS_P1 STEP 3, 460make_edge_fields_p1.f:2591-3721,
// and the tri-corners of STEP 4, :3725-7035).  Stored as E * dr.
template <typename P>
__device__ __forceinline__ void emf_z(const CsDims& d, int i, int j, int k, int s, const Geo& geo, const State& st,
                                      const P* V, const double* Sx, const double* Sy, double* Ez,
                                      const Params& pr) {
  const int N = d.N;
  double n3[3] = {geo.t(d, T_UX, s, i, j), geo.t(d, T_UY, s, i, j), geo.t(d, T_UZ, s, i, j)};
  double dr = geo.r(d, k) - geo.r(d, k - 1);
  const double* Bx = st.Bx + s * d.fpp;
  const double* By = st.By + s * d.fpp;
  bool ci = (i == 0 || i == N), cj = (j == 0 || j == N);
  double ESSz;
  if (ci && cj) {
    // tri-corner: three active zones, 3-state LLF solver
    int c1i, c1j, c2i, c2j, c3i, c3j;
    double bn;
    double S = 0.0;
    auto sabs = [&](const double* SS, int ii, int jj) {
      return fmax(fabs(sp(d, SS, s, 0, ii, jj, k)), fabs(sp(d, SS, s, 1, ii, jj, k)));
    };
    if (i == 0 && j == 0) {
      c1i = 1; c1j = 1; c2i = 0; c2j = 1; c3i = 1; c3j = 0;
      bn = -By[d.face(1, 0, k)] + Bx[d.face(0, 1, k)] + By[d.face(0, 0, k)];
      S = fmax(sabs(Sx, 0, 1), sabs(Sy, 1, 0));
    } else if (i == N && j == 0) {
      c1i = N; c1j = 0; c2i = N; c2j = 1; c3i = N + 1; c3j = 1;
      bn = Bx[d.face(N, 1, k)] + By[d.face(N, 0, k)] - By[d.face(N + 1, 0, k)];
      S = fmax(sabs(Sx, N, 1), sabs(Sy, N, 0));
    } else if (i == N && j == N) {
      c1i = N; c1j = N; c2i = N + 1; c2j = N; c3i = N; c3j = N + 1;
      bn = By[d.face(N, N, k)] - Bx[d.face(N, N, k)] - By[d.face(N + 1, N, k)];
      S = fmax(sabs(Sx, N, N), sabs(Sy, N, N));
    } else {
      c1i = 0; c1j = N; c2i = 1; c2j = N; c3i = 1; c3j = N + 1;
      bn = -By[d.face(1, N, k)] - Bx[d.face(0, N, k)] + By[d.face(0, N, k)];
      S = fmax(sabs(Sx, 0, N), sabs(Sy, 1, N));
    }
    double e1[3], e2[3], e3[3];
    cell_E(d, V, s, d.cell(c1i, c1j, k), e1);
    cell_E(d, V, s, d.cell(c2i, c2j, k), e2);
    cell_E(d, V, s, d.cell(c3i, c3j, k), e3);
    double ob3 = 1.0 / 3.0;
    double ex = (e1[0] + e2[0] + e3[0]) * ob3, ey = (e1[1] + e2[1] + e3[1]) * ob3,
           ez = (e1[2] + e2[2] + e3[2]) * ob3;
    double e = ex * n3[0] + ey * n3[1] + ez * n3[2];
    ESSz = emf3_llf<double>(S, e, e, e, bn, 0.0, 0.0);
  } else {
    double eLD[3], eRD[3], eRU[3], eLU[3];
    cell_E(d, V, s, d.cell(i, j, k), eLD);
    cell_E(d, V, s, d.cell(i + 1, j, k), eRD);
    cell_E(d, V, s, d.cell(i + 1, j + 1, k), eRU);
    cell_E(d, V, s, d.cell(i, j + 1, k), eLU);
    double zLD, zRD, zRU, zLU;
    if (pr.emf == CS_EMF_LLF) {
      double ex = (eLD[0] + eRD[0] + eRU[0] + eLU[0]) * 0.25;
      double ey = (eLD[1] + eRD[1] + eRU[1] + eLU[1]) * 0.25;
      double ez = (eLD[2] + eRD[2] + eRU[2] + eLU[2]) * 0.25;
      zLD = zRD = zRU = zLU = ex * n3[0] + ey * n3[1] + ez * n3[2];
    } else {
      zLD = eLD[0] * n3[0] + eLD[1] * n3[1] + eLD[2] * n3[2];
      zRD = eRD[0] * n3[0] + eRD[1] * n3[1] + eRD[2] * n3[2];
      zRU = eRU[0] * n3[0] + eRU[1] * n3[1] + eRU[2] * n3[2];
      zLU = eLU[0] * n3[0] + eLU[1] * n3[1] + eLU[2] * n3[2];
    }
    double By_R = By[d.face(i + 1, j, k)], By_L = By[d.face(i, j, k)];
    double Bn_U = Bx[d.face(i, j + 1, k)], Bn_D = Bx[d.face(i, j, k)];
    double S_R = fmax(sp(d, Sx, s, 1, i, j + 1, k), sp(d, Sx, s, 1, i, j, k));
    double S_L = fmin(sp(d, Sx, s, 0, i, j + 1, k), sp(d, Sx, s, 0, i, j, k));
    double S_U = fmax(sp(d, Sy, s, 1, i + 1, j, k), sp(d, Sy, s, 1, i, j, k));
    double S_D = fmin(sp(d, Sy, s, 0, i + 1, j, k), sp(d, Sy, s, 0, i, j, k));
    if (pr.emf == CS_EMF_LLF) {
      double t = fmax(fmax(fabs(S_R), fabs(S_L)), fmax(fabs(S_U), fabs(S_D)));
      S_R = t; S_L = -t; S_U = t; S_D = -t;
    }
    double ct = geo.t(d, T_ECOS, s, i, j), st_ = geo.t(d, T_ESIN, s, i, j);
    ESSz = emf4_hll<double>(S_R, S_L, S_U, S_D, ct, st_, zRU, zLU, zLD, zRD, Bn_U, Bn_D, By_R, By_L);
  }
  Ez[s * d.fpp + d.face(i, j, k)] = ESSz * dr;
}

// x-edges (STEP 1, 460make_edge_fields_p1.f:393-1493) and y-edges (STEP 2,
// :1497-2587): edges normal to the radial direction.  Stored as E * arc length.
template <int DIR, typename P>
__device__ __forceinline__ void emf_xy(const CsDims& d, int i, int j, int k, int s, const Geo& geo,
                                       const State& st, const P* V, const double* Sa, const double* Sb,
                                       double* Eo, const Params& pr) {
  int64_t fo = s * d.fpp + d.face(i, j, k);
  if (pr.wall && (k == 0 || k == d.NR)) { Eo[fo] = 0.0; return; }
  double n3[3];
  int64_t cLD, cRD, cRU, cLU;
  double By_R, By_L, Bn_U, Bn_D, S_R, S_L, S_U, S_D, len;
  const double* Bx = st.Bx + s * d.fpp;
  const double* By = st.By + s * d.fpp;
  const double* Bz = st.Bz + s * d.fpp;
  if (DIR == 0) {  // Efx: cell-x i, y-vertex j, z-vertex k; Sa = y-face speeds, Sb = z-face speeds
    for (int c = 0; c < 3; ++c) n3[c] = geo.t(d, T_XETX + c, s, i, j);
    cLD = d.cell(i, j, k); cRD = d.cell(i, j + 1, k); cRU = d.cell(i, j + 1, k + 1); cLU = d.cell(i, j, k + 1);
    By_R = Bz[d.face(i, j + 1, k)]; By_L = Bz[d.face(i, j, k)];
    Bn_U = By[d.face(i, j, k + 1)]; Bn_D = By[d.face(i, j, k)];
    S_R = fmax(sp(d, Sa, s, 1, i, j, k + 1), sp(d, Sa, s, 1, i, j, k));
    S_L = fmin(sp(d, Sa, s, 0, i, j, k + 1), sp(d, Sa, s, 0, i, j, k));
    S_U = fmax(sp(d, Sb, s, 1, i, j + 1, k), sp(d, Sb, s, 1, i, j, k));
    S_D = fmin(sp(d, Sb, s, 0, i, j + 1, k), sp(d, Sb, s, 0, i, j, k));
    len = geo.r(d, k) * geo.t(d, T_THX, s, i, j);
  } else {  // Efy: x-vertex i, cell-y j, z-vertex k; Sa = z-face speeds, Sb = x-face speeds
    for (int c = 0; c < 3; ++c) n3[c] = geo.t(d, T_YETX + c, s, i, j);
    cLD = d.cell(i, j, k); cRD = d.cell(i, j, k + 1); cLU = d.cell(i + 1, j, k); cRU = d.cell(i + 1, j, k + 1);
    By_R = Bx[d.face(i, j, k + 1)]; By_L = Bx[d.face(i, j, k)];
    Bn_U = Bz[d.face(i + 1, j, k)]; Bn_D = Bz[d.face(i, j, k)];
    S_R = fmax(sp(d, Sa, s, 1, i + 1, j, k), sp(d, Sa, s, 1, i, j, k));
    S_L = fmin(sp(d, Sa, s, 0, i + 1, j, k), sp(d, Sa, s, 0, i, j, k));
    S_U = fmax(sp(d, Sb, s, 1, i, j, k), sp(d, Sb, s, 1, i, j, k + 1));
    S_D = fmin(sp(d, Sb, s, 0, i, j, k), sp(d, Sb, s, 0, i, j, k + 1));
    len = geo.r(d, k) * geo.t(d, T_THY, s, i, j);
  }
  double eLD[3], eRD[3], eRU[3], eLU[3];
  cell_E(d, V, s, cLD, eLD);
  cell_E(d, V, s, cRD, eRD);
  cell_E(d, V, s, cRU, eRU);
  cell_E(d, V, s, cLU, eLU);
  double zLD, zRD, zRU, zLU;
  if (pr.emf == CS_EMF_LLF) {
    double ex = (eLD[0] + eRD[0] + eRU[0] + eLU[0]) * 0.25;
    double ey = (eLD[1] + eRD[1] + eRU[1] + eLU[1]) * 0.25;
    double ez = (eLD[2] + eRD[2] + eRU[2] + eLU[2]) * 0.25;
    zLD = zRD = zRU = zLU = ex * n3[0] + ey * n3[1] + ez * n3[2];
    double t = fmax(fmax(fabs(S_R), fabs(S_L)), fmax(fabs(S_U), fabs(S_D)));
    S_R = t; S_L = -t; S_U = t; S_D = -t;
  } else {
    zLD = eLD[0] * n3[0] + eLD[1] * n3[1] + eLD[2] * n3[2];
    zRD = eRD[0] * n3[0] + eRD[1] * n3[1] + eRD[2] * n3[2];
    zRU = eRU[0] * n3[0] + eRU[1] * n3[1] + eRU[2] * n3[2];
    zLU = eLU[0] * n3[0] + eLU[1] * n3[1] + eLU[2] * n3[2];
  }
  double e = emf4_hll<double>(S_R, S_L, S_U, S_D, 0.0, 1.0, zRU, zLU, zLD, zRD, Bn_U, Bn_D, By_R, By_L);
  Eo[fo] = e * len;
}

template <typename P>
__global__ void k_emf_z(CsDims d, Rng rg, Geo geo, State st, const P* V, const double* Sx,
                        const double* Sy, double* Ez, Params pr) {
  int i, j, k, s;
  if (!decode(rg, gtid(), i, j, k, s)) return;
  emf_z<P>(d, i, j, k, s, geo, st, V, Sx, Sy, Ez, pr);
}
template <int DIR, typename P>
__global__ void k_emf_xy(CsDims d, Rng rg, Geo geo, State st, const P* V, const double* Sa,
                         const double* Sb, double* Eo, Params pr) {
  int i, j, k, s;
  if (!decode(rg, gtid(), i, j, k, s)) return;
  emf_xy<DIR, P>(d, i, j, k, s, geo, st, V, Sa, Sb, Eo, pr);
}
// the three edge families of a stage in one launch; S: face speeds x, y, z
template <typename P>
__global__ void k_emf_all(CsDims d, Rng rz, Rng rx, Rng ry, Geo geo, State st, const P* V, const double* S0,
                          const double* S1, const double* S2, double* Ex, double* Ey, double* Ez, Params pr) {
  uint32_t t = gtid();
  int i, j, k, s;
  if (t < rz.total) {
    decode(rz, t, i, j, k, s);
    emf_z<P>(d, i, j, k, s, geo, st, V, S0, S1, Ez, pr);
    return;
  }
  t -= rz.total;
  if (t < rx.total) {
    decode(rx, t, i, j, k, s);
    emf_xy<0, P>(d, i, j, k, s, geo, st, V, S1, S2, Ex, pr);
    return;
  }
  t -= rx.total;
  if (decode(ry, t, i, j, k, s)) emf_xy<1, P>(d, i, j, k, s, geo, st, V, S2, S0, Ey, pr);
}

// ------------------------------------------------------------------ shared faces and edges
// Both sectors that meet at a panel edge compute its faces and edges, from
// states that differ only where a stencil reaches a ghost row.  The reference
// makes them unique after the update by its defect synchronisation: the two
// copies of a coinciding face flux or edge EMF are averaged
// (334fetch_buffers_fill_in_flux_x.f:290-400, 335/336 for y and z) and the
// radial edge at a cube corner takes the mean of its three copies
// (334...:371-380 accumulation, 340synchronize_efz_quad_corner.f).  Here the
// same averages are taken before the update, so the defect is zero and div B
// is kept at roundoff.  One thread per physical face/edge writes every copy:
//   flux pairs   (a + (-b)) * 0.5: the two face normals are opposite;
//   z-edge pairs (a + b) * 0.5: both store E . rhat * dr;
//   x/y-edge pairs (a + sg b) * 0.5 with sg = -1 where the two sectors run
//   the edge in opposite directions (left-left and down-down panel edges);
//   cube corners (e0 + e1 + e2) * (1/3) in ascending sector order.
// The eight cube corners, from the grouping of the 24 panel-corner vertices
// by position at cs_create (the same for every N): c_corner[c][m] is copy m
// (ascending sector) of corner c as 4 * sector + corner bits (bit 0: i = N,
// bit 1: j = N); c_cornerof[s][bits] is the corner a panel corner belongs to.
__constant__ int c_corner[8][3];
__constant__ int c_cornerof[6][4];
__device__ __forceinline__ int64_t corner_edge(const CsDims& d, int code, int k) {
  const int sc = code >> 2, b = code & 3;
  return (int64_t)sc * d.fpp + d.face((b & 1) ? d.N : 0, (b & 2) ? d.N : 0, k);
}
__device__ __forceinline__ double corner_mean(const CsDims& d, const double* Ez, int c, int k) {
  return (Ez[corner_edge(d, c_corner[c][0], k)] + Ez[corner_edge(d, c_corner[c][1], k)] +
          Ez[corner_edge(d, c_corner[c][2], k)]) * (1.0 / 3.0);
}

__device__ __forceinline__ void avg_pair(double* pa, double* pb, double sg) {
  double a = *pa, b = *pb;
  *pa = (a + sg * b) * 0.5;
  *pb = (b + sg * a) * 0.5;
}

template <bool FLUX, bool EDGE>
__global__ void k_sync_shared(CsDims d, double* Fx, double* Fy, double* Ex, double* Ey, double* Ez,
                              int kh0, int nkh) {
  uint32_t t = gtid();
  const int N = d.N, NR = d.NR;
  const uint32_t nF = FLUX ? 12u * N * NR : 0u;
  const uint32_t nZ = EDGE ? 12u * (N - 1) * NR : 0u;
  const uint32_t nH = EDGE ? 12u * N * nkh : 0u;
  const uint32_t nC = EDGE ? 8u * NR : 0u;
  if (t < nF) {
    uint32_t per = (uint32_t)N * NR;
    int p = t / per;
    uint32_t rem = t - p * per;
    int tt = 1 + (int)(rem % N), k = 1 + (int)(rem / N);
    double *pa, *pb;
    if (p < 6) {         // right edge of sector p with the top edge of its right neighbour
      int s = p, o = c_nbr[s][0];
      pa = Fx + (int64_t)s * 9 * d.fpp + d.face(N, tt, k);
      pb = Fy + (int64_t)o * 9 * d.fpp + d.face(tt, N, k);
    } else if (p < 9) {  // left edges 1-6, 2-4, 3-5 (1-based)
      int s = p - 6, o = c_nbr[s][1];
      pa = Fx + (int64_t)s * 9 * d.fpp + d.face(0, tt, k);
      pb = Fx + (int64_t)o * 9 * d.fpp + d.face(0, N + 1 - tt, k);
    } else {             // bottom edges 1-5, 2-6, 3-4
      int s = p - 9, o = c_nbr[s][3];
      pa = Fy + (int64_t)s * 9 * d.fpp + d.face(tt, 0, k);
      pb = Fy + (int64_t)o * 9 * d.fpp + d.face(N + 1 - tt, 0, k);
    }
    double va[9], vb[9];  // all loads in flight before the first store
    for (int q = 0; q < 9; ++q) { va[q] = pa[q * d.fpp]; vb[q] = pb[q * d.fpp]; }
    for (int q = 0; q < 9; ++q) {
      pa[q * d.fpp] = (va[q] + (-vb[q])) * 0.5;
      pb[q * d.fpp] = (vb[q] + (-va[q])) * 0.5;
    }
    return;
  }
  t -= nF;
  if (t < nZ) {          // radial edges along panel edges, vertices 1..N-1
    uint32_t per = (uint32_t)(N - 1) * NR;
    int p = t / per;
    uint32_t rem = t - p * per;
    int tt = 1 + (int)(rem % (N - 1)), k = 1 + (int)(rem / (N - 1));
    int64_t a, b;
    if (p < 6) { int s = p, o = c_nbr[s][0]; a = s * d.fpp + d.face(N, tt, k); b = o * d.fpp + d.face(tt, N, k); }
    else if (p < 9) { int s = p - 6, o = c_nbr[s][1]; a = s * d.fpp + d.face(0, tt, k); b = o * d.fpp + d.face(0, N - tt, k); }
    else { int s = p - 9, o = c_nbr[s][3]; a = s * d.fpp + d.face(tt, 0, k); b = o * d.fpp + d.face(N - tt, 0, k); }
    avg_pair(Ez + a, Ez + b, 1.0);
    return;
  }
  t -= nZ;
  if (t < nH) {          // x/y-edges along panel edges, levels kh0..kh0+nkh-1
    uint32_t per = (uint32_t)N * nkh;
    int p = t / per;
    uint32_t rem = t - p * per;
    int tt = 1 + (int)(rem % N), k = kh0 + (int)(rem / N);
    double *pa, *pb, sg;
    if (p < 6) {
      int s = p, o = c_nbr[s][0];
      pa = Ey + s * d.fpp + d.face(N, tt, k); pb = Ex + o * d.fpp + d.face(tt, N, k); sg = 1.0;
    } else if (p < 9) {
      int s = p - 6, o = c_nbr[s][1];
      pa = Ey + s * d.fpp + d.face(0, tt, k); pb = Ey + o * d.fpp + d.face(0, N + 1 - tt, k); sg = -1.0;
    } else {
      int s = p - 9, o = c_nbr[s][3];
      pa = Ex + s * d.fpp + d.face(tt, 0, k); pb = Ex + o * d.fpp + d.face(N + 1 - tt, 0, k); sg = -1.0;
    }
    avg_pair(pa, pb, sg);
    return;
  }
  t -= nH;
  if (t < nC) {          // radial edges at the eight cube corners
    int c = t / NR, k = 1 + (int)(t % NR);
    double v = corner_mean(d, Ez, c, k);
    Ez[corner_edge(d, c_corner[c][0], k)] = v;
    Ez[corner_edge(d, c_corner[c][1], k)] = v;
    Ez[corner_edge(d, c_corner[c][2], k)] = v;
  }
}

// ------------------------------------------------------------------ unique on read
// The stage path makes panel-edge fluxes and EMFs unique where they are
// read, with the averages of k_sync_shared written from the reader's side:
// own' = (own + sg * other) / 2 is the same expression for both copies (the
// addition commutes), so both sectors see one value bit for bit, and the
// separate synchronisation launch disappears.  Interior items return the
// stored value.
// Zone fluxes: the four lateral faces of zone (i, j, k) read with their
// panel-edge copies averaged.  The partner copies are located once per zone
// (null inside the panel); the loads are predicated on the boundary lanes.
struct ZoneFaces {
  int64_t xr, xl, yr, yl;                 // own faces
  const double *pxr, *pxl, *pyr, *pyl;    // partner copies (component 0) or null
};
__device__ __forceinline__ ZoneFaces zone_faces(const CsDims& d, const double* Fx, const double* Fy, int s, int i,
                                                int j, int k) {
  const int N = d.N;
  ZoneFaces z;
  z.xr = (int64_t)s * 9 * d.fpp + d.face(i, j, k);
  z.xl = (int64_t)s * 9 * d.fpp + d.face(i - 1, j, k);
  z.yr = z.xr;
  z.yl = (int64_t)s * 9 * d.fpp + d.face(i, j - 1, k);
  z.pxr = i == N ? Fy + (int64_t)c_nbr[s][0] * 9 * d.fpp + d.face(j, N, k) : nullptr;
  z.pxl = i == 1 ? Fx + (int64_t)c_nbr[s][1] * 9 * d.fpp + d.face(0, N + 1 - j, k) : nullptr;
  z.pyr = j == N ? Fx + (int64_t)c_nbr[s][2] * 9 * d.fpp + d.face(N, i, k) : nullptr;
  z.pyl = j == 1 ? Fy + (int64_t)c_nbr[s][3] * 9 * d.fpp + d.face(N + 1 - i, 0, k) : nullptr;
  return z;
}
// (fx(i) - fx(i-1)) + (fy(j) - fy(j-1)) of component q
__device__ __forceinline__ double lateral_div(const CsDims& d, const ZoneFaces& z, const double* __restrict__ Fx,
                                              const double* __restrict__ Fy, int q) {
  const int64_t qo = (int64_t)q * d.fpp;
  double fxr = Fx[z.xr + qo], fxl = Fx[z.xl + qo], fyr = Fy[z.yr + qo], fyl = Fy[z.yl + qo];
  if (z.pxr) fxr = (fxr + (-z.pxr[qo])) * 0.5;
  if (z.pxl) fxl = (fxl + (-z.pxl[qo])) * 0.5;
  if (z.pyr) fyr = (fyr + (-z.pyr[qo])) * 0.5;
  if (z.pyl) fyl = (fyl + (-z.pyl[qo])) * 0.5;
  return (fxr - fxl) + (fyr - fyl);
}
// radial edge at vertex (i, j): two copies on a panel edge, three at a cube corner
__device__ __forceinline__ double ez_unique(const CsDims& d, const double* __restrict__ Ez, int s, int i, int j,
                                            int k) {
  const int N = d.N;
  const double a = Ez[(int64_t)s * d.fpp + d.face(i, j, k)];
  const bool ci = i == 0 || i == N, cj = j == 0 || j == N;
  if (!ci && !cj) return a;
  if (ci && cj) return corner_mean(d, Ez, c_cornerof[s][(i == N ? 1 : 0) | (j == N ? 2 : 0)], k);
  int64_t o;
  if (i == N) o = (int64_t)c_nbr[s][0] * d.fpp + d.face(j, N, k);
  else if (i == 0) o = (int64_t)c_nbr[s][1] * d.fpp + d.face(0, N - j, k);
  else if (j == N) o = (int64_t)c_nbr[s][2] * d.fpp + d.face(N, i, k);
  else o = (int64_t)c_nbr[s][3] * d.fpp + d.face(N - i, 0, k);
  return (a + Ez[o]) * 0.5;
}
// y-edge (x-vertex i, cell j) and x-edge (cell i, y-vertex j) on panel edges
__device__ __forceinline__ double ey_unique(const CsDims& d, const double* __restrict__ Ex,
                                            const double* __restrict__ Ey, int s, int i, int j, int k) {
  const int N = d.N;
  const double a = Ey[(int64_t)s * d.fpp + d.face(i, j, k)];
  if (i != 0 && i != N) return a;
  if (i == N) return (a + Ex[(int64_t)c_nbr[s][0] * d.fpp + d.face(j, N, k)]) * 0.5;
  return (a + (-Ey[(int64_t)c_nbr[s][1] * d.fpp + d.face(0, N + 1 - j, k)])) * 0.5;
}
__device__ __forceinline__ double ex_unique(const CsDims& d, const double* __restrict__ Ex,
                                            const double* __restrict__ Ey, int s, int i, int j, int k) {
  const int N = d.N;
  const double a = Ex[(int64_t)s * d.fpp + d.face(i, j, k)];
  if (j != 0 && j != N) return a;
  if (j == N) return (a + Ey[(int64_t)c_nbr[s][2] * d.fpp + d.face(N, i, k)]) * 0.5;
  return (a + (-Ex[(int64_t)c_nbr[s][3] * d.fpp + d.face(N + 1 - i, 0, k)])) * 0.5;
}

// ------------------------------------------------------------------ update
// CLEANUP_AFTER_UPDATE (485cleanup_after_update.f:4) of one zone of the stage
// output, run by the zone-update kernels after the faces are updated.
// Shells: the zone-centred B is replaced by the least-squares fit to the
// zone's six face values over their unit normals (:692-860,
// B = (A^T A)^-1 A^T b, kept as per-face weights W_f = (A^T A)^-1 n_f), and
// the total energy takes the change of magnetic energy so the thermal and
// kinetic parts are kept (:845-851).  Surface mode (one layer between walls):
// the fit would couple the lateral face values into a radial B, which the
// wall mirror turns into mass flux through the walls, so each face family's
// average is reproduced along its mean normal instead and the total energy
// is kept.  Then CONS_TO_PRIM (662pde_3d_cons_to_prim.f:69-151) with its
// floors; U and the entropy are stored back and the same conversion gives
// the primitives V the next stage reconstructs from.  u[0..7] and ent hold
// the updated zone (u[5..7] advanced by the induction fluxes, as the
// reference's dU_dt).
template <typename P, bool PUSH = true>
__device__ __forceinline__ void cleanup_zone(const CsDims& d, const Geo& geo, const State& st, P* V,
                                             const Params& pr, int* flags, int s, int i, int j, int k,
                                             double* u, double ent) {
  const int64_t c = d.cell(i, j, k);
  const int64_t so = (int64_t)s * d.fpp;
  const double bxp = st.Bx[so + d.face(i, j, k)], bxm = st.Bx[so + d.face(i - 1, j, k)];
  const double byp = st.By[so + d.face(i, j, k)], bym = st.By[so + d.face(i, j - 1, k)];
  const double bzp = st.Bz[so + d.face(i, j, k)], bzm = st.Bz[so + d.face(i, j, k - 1)];
  if (pr.surf) {
    const double b0 = 0.5 * (bxm + bxp), b1 = 0.5 * (bym + byp), b2 = 0.5 * (bzm + bzp);
    for (int r = 0; r < 3; ++r)
      u[5 + r] = geo.t(d, T_M00 + 3 * r, s, i, j) * b0 + geo.t(d, T_M00 + 3 * r + 1, s, i, j) * b1 +
                 geo.t(d, T_M00 + 3 * r + 2, s, i, j) * b2;
  } else {
    double bn[3];
    for (int r = 0; r < 3; ++r)
      bn[r] = geo.t(d, T_BWXP + r, s, i, j) * bxp + geo.t(d, T_BWXM + r, s, i, j) * bxm +
              geo.t(d, T_BWYP + r, s, i, j) * byp + geo.t(d, T_BWYM + r, s, i, j) * bym +
              geo.t(d, T_BWZ + r, s, i, j) * (bzp + bzm);
    u[4] = u[4] - (u[5] * u[5] + u[6] * u[6] + u[7] * u[7]) * C<double>::ob8pi +
           (bn[0] * bn[0] + bn[1] * bn[1] + bn[2] * bn[2]) * C<double>::ob8pi;
    u[5] = bn[0];
    u[6] = bn[1];
    u[7] = bn[2];
  }
  Gas<double> gs{pr.gamma, pr.gm1, pr.rhofloor, pr.prsfloor, pr.igm1};
  double v[8];
  int fl = cons_to_prim(gs, u, v, ent);
  // the zone and, in a shell, every lateral ghost slot that copies it
  // (neighbour strips, and the corner block for a panel corner zone): the
  // exchange is done here.  Surface mode keeps the exchange launch: its zone
  // update is register bound and the scattered ghost writes cost it more than
  // the launch they replace.
  auto put = [&](int ds, int64_t dc) {
    for (int q = 0; q < 8; ++q) st.U[(ds * 8 + q) * d.cpp + dc] = u[q];
    st.ent[ds * d.cpp + dc] = ent;
    for (int q = 0; q < 8; ++q) V[((int64_t)ds * 9 + q) * d.cpp + dc] = (P)v[q];
    V[((int64_t)ds * 9 + 8) * d.cpp + dc] = (P)ent;
  };
  put(s, c);
  if (PUSH && !pr.surf && (i <= d.g || i > d.N - d.g || j <= d.g || j > d.N - d.g)) {
    zone_pushes(c_nbr, d.N, d.g, s, i, j, [&](int ds, int di, int dj) { put(ds, d.cell(di, dj, k)); });
    corner_pushes(d.N, d.g, i, j, [&](int di, int dj) { put(s, d.cell(di, dj, k)); });
  }
  if (fl && flags) flags[0] = 1;
  if (!(u[0] == u[0]) || !(u[4] == u[4]) || isinf(u[4])) { if (flags) flags[1] = 1; }
}

// Rotating-frame sources of BUILD_NON_STIFF_SOURCES (470build_non_stiff_sources.f:
// 150-250), explicit in every scheme: centrifugal -rho Om x (Om x r) and
// Coriolis -2 rho Om x v on the momentum, their work v . S on the energy, at
// the zone position r (mid radius along the radial normal, where the gravity
// acts) with v from the stage input as CONS_TO_PRIM forms it.  Added to L.
__device__ __forceinline__ void rotation_source(const Params& pr, const State& in, const CsDims& d, int s,
                                                int64_t c, const double* r, double* L) {
  const double rho = in.U[(s * 8) * d.cpp + c];
  const double a = rho / (rho * rho + pr.rhofloor * pr.rhofloor);
  double v[3];
  for (int q = 0; q < 3; ++q) v[q] = in.U[(s * 8 + 1 + q) * d.cpp + c] * a;
  const double* w = pr.om;
  const double wr[3] = {w[1] * r[2] - w[2] * r[1], w[2] * r[0] - w[0] * r[2], w[0] * r[1] - w[1] * r[0]};
  const double S[3] = {-rho * (w[1] * wr[2] - w[2] * wr[1]) - 2.0 * rho * (w[1] * v[2] - w[2] * v[1]),
                       -rho * (w[2] * wr[0] - w[0] * wr[2]) - 2.0 * rho * (w[2] * v[0] - w[0] * v[2]),
                       -rho * (w[0] * wr[1] - w[1] * wr[0]) - 2.0 * rho * (w[0] * v[1] - w[1] * v[0])};
  for (int q = 0; q < 3; ++q) L[1 + q] += S[q];
  L[4] += v[0] * S[0] + v[1] * S[1] + v[2] * S[2];
}

// Combination of a zone's stage update with the radial gravity source
// S(U) = (0, rho gr n, gr (rho v . n)), gr = -grav / rc^2 at the zone's mid
// radius, n the unit normal of its radial faces (the point-mass gravity of
// 470build_non_stiff_sources.f:150-250).  The entropy carries no source.
//  mode 0: explicit source, Shu-Osher  u = wa u0 + wb (in + dt (L + S(in))).
//  IMEX-SSP2(2,2,2) (Pareschi & Russo 2005; explicit part = Heun, the
//  reference's RK2), gravity implicit, gi = 1 - 1/sqrt(2):
//  mode 1: in = U1, u0 = Un:  U2 = Un + dt L(U1) + (1 - 2 gi) dt S(U1) + gi dt S(U2).
//          S leaves rho unchanged and is linear in rho v, so U2 is explicit:
//          rho v then E take the implicit term in turn.
//  mode 2: in = U2, u0 = Un:  Un+1 = (Un + U2 + dt L(U2)) / 2 + dt/2 (2 gi S(U1) + (1 - gi) S(U2)),
//          with U1 = Un + gi dt S(U1) rebuilt from Un.
__device__ __forceinline__ void finish_zone(const CsDims& d, const State& in, const State& u0, int s, int64_t c,
                                            double dt, double wa, double wb, int mode, double gr, const double* n,
                                            const double* L, double* u, double& en) {
  // L[0..7]: rho, rho v, E, B (zone B advanced by the induction fluxes, as the
  // reference's dU_dt; the cleanup replaces it); L[8]: entropy
  double ui[8], uz[8];
  for (int q = 0; q < 8; ++q) ui[q] = in.U[(s * 8 + q) * d.cpp + c];
  const double ei = in.ent[s * d.cpp + c];
  double ez = 0.0;
  if (mode != 0 || wa != 0.0) {
    for (int q = 0; q < 8; ++q) uz[q] = u0.U[(s * 8 + q) * d.cpp + c];
    ez = u0.ent[s * d.cpp + c];
  }
  rk_combine(ui, ei, uz, ez, dt, wa, wb, mode, gr, n, L, u, en);
}

// First, purely implicit stage of IMEX-SSP2(2,2,2): U1 = Un + gi dt S(U1) in
// closed form (rho unchanged, then rho v, then E), written to the stage buffer
// whose faces hold a copy of Un's; the cleanup then gives U1 its zone B and
// primitives, like any stage output.
template <typename P>
__device__ __forceinline__ void imex_source(const CsDims& d, int i, int j, int k, int s, const Geo& geo,
                                            const State& un, const State& u1, P* V, double dt, const Params& pr,
                                            int* flags) {
  const double gi = 1.0 - 0.70710678118654752440;
  const int64_t c = d.cell(i, j, k);
  const double r1 = geo.r(d, k), r0 = geo.r(d, k - 1), rc = 0.5 * (r0 + r1), gr = -pr.grav / (rc * rc);
  const double zn = geo.t(d, T_ZSN, s, i, j);
  double n[3], u[8];
  for (int q = 0; q < 3; ++q) n[q] = geo.t(d, T_ZSX + q, s, i, j) / zn;
  for (int q = 0; q < 8; ++q) u[q] = un.U[(s * 8 + q) * d.cpp + c];
  for (int q = 1; q < 4; ++q) u[q] += gi * dt * u[0] * gr * n[q - 1];
  u[4] += gi * dt * gr * (u[1] * n[0] + u[2] * n[1] + u[3] * n[2]);
  cleanup_zone<P>(d, geo, u1, V, pr, flags, s, i, j, k, u, un.ent[s * d.cpp + c]);
}
template <typename P>
__global__ void k_imex_source(CsDims d, Rng rg, Geo geo, State un, State u1, P* V, const double* dtp, Params pr,
                              int* flags) {
  int i, j, k, s;
  if (!decode(rg, gtid(), i, j, k, s)) return;
  imex_source<P>(d, i, j, k, s, geo, un, u1, V, dtp[0], pr, flags);
}

// U_out = a U0 + b (U_in + dt L(U_in)) for rho, rho v, E and entropy
// (UPDATE_IMEX_RK_STEP_P1 484update_imex_rk_step_p1.f:483-497, 547-593 with the
// Shu-Osher stage weights; the RK2 weights reproduce DRV:2661-2773).
template <typename P>
__device__ __forceinline__ void update_cells(const CsDims& d, int i, int j, int k, int s, const Geo& geo,
                                             const State& in, const State& u0, const State& out,
                                             const double* __restrict__ Fx, const double* __restrict__ Fy,
                                             const double* __restrict__ Fz, double dt, double wa, double wb,
                                             const Params& pr, P* Vout, int* flags, int mode) {
  int64_t c = d.cell(i, j, k);
  double r1 = geo.r(d, k), r0 = geo.r(d, k - 1);
  double vol = geo.t(d, T_OM, s, i, j) * (r1 * r1 * r1 - r0 * r0 * r0) / 3.0;
  double ov = 1.0 / vol;
  const int64_t fxp = d.face(i, j, k), fzm = d.face(i, j, k - 1);
  const ZoneFaces zf = zone_faces(d, Fx, Fy, s, i, j, k);
  double L[9];
  for (int q = 0; q < 9; ++q) {
    int64_t o = ((int64_t)s * 9 + q) * d.fpp;
    L[q] = -ov * (lateral_div(d, zf, Fx, Fy, q) + (Fz[o + fxp] - Fz[o + fzm]));
  }
  double n[3] = {0.0, 0.0, 0.0}, gr = 0.0;
  if (pr.grav != 0.0 || pr.rot) {
    const double zn = geo.t(d, T_ZSN, s, i, j), rc = 0.5 * (r0 + r1);
    for (int q = 0; q < 3; ++q) n[q] = geo.t(d, T_ZSX + q, s, i, j) / zn;
    if (pr.grav != 0.0) gr = -pr.grav / (rc * rc);
    if (pr.rot) {
      const double rp[3] = {rc * n[0], rc * n[1], rc * n[2]};
      rotation_source(pr, in, d, s, c, rp, L);
    }
  }
  double u[8], en;
  finish_zone(d, in, u0, s, c, dt, wa, wb, mode, gr, n, L, u, en);
  cleanup_zone<P>(d, geo, out, Vout, pr, flags, s, i, j, k, u, en);
}
// at most 256 work-items per workgroup: the register budget that keeps the
// zone's fluxes, states and cleanup weights out of scratch
template <typename P>
__global__ void __launch_bounds__(256, 2) k_update_cells(CsDims d, Rng rg, Geo geo, State in, State u0, State out,
                               const double* __restrict__ Fx, const double* __restrict__ Fy,
                               const double* __restrict__ Fz, const double* dtp, double wa,
                               double wb, Params pr, P* Vout, int* flags, int mode) {
  int i, j, k, s;
  if (!decode(rg, gtid(), i, j, k, s)) return;
  update_cells<P>(d, i, j, k, s, geo, in, u0, out, Fx, Fy, Fz, dtp[0], wa, wb, pr, Vout, flags, mode);
}

// Surface mode (one radial layer between conducting walls): the radial ghost
// column of a zone is its own state mirrored about the radial face normal,
// alternating (..., M(u), u, M(u), u | M(u), u, ...), so the two wall fluxes
// are functions of the zone alone.  They are evaluated here, with exactly the
// reconstruction and Riemann solve k_face_flux<2> applies to that column, and
// neither the radial ghost fill nor the z-face pass is needed.
template <typename R, typename P>
__global__ void __launch_bounds__(256, 2) k_update_cells_surf(CsDims d, Rng rg, Geo geo, State in, State u0, State out,
                                    P* V, const double* __restrict__ Fx,
                                    const double* __restrict__ Fy, const double* dtp, double wa,
                                    double wb, Params pr, int* flags, int mode) {
  int i, j, k, s;
  if (!decode(rg, gtid(), i, j, k, s)) return;
  const double dt = dtp[0];
  const int64_t c = d.cell(i, j, k);
  const double r1 = geo.r(d, k), r0 = geo.r(d, k - 1);
  const double ov = 1.0 / (geo.t(d, T_OM, s, i, j) * (r1 * r1 * r1 - r0 * r0 * r0) / 3.0);
  const ZoneFaces zf = zone_faces(d, Fx, Fy, s, i, j, k);
  double L[9];
  // the surface cleanup rebuilds the zone B from the faces and keeps the
  // total energy, so the induction components (5-7) are not needed here
  L[5] = L[6] = L[7] = 0.0;
  for (int q = 0; q < 9; ++q)
    if (q < 5 || q == 8) L[q] = -ov * lateral_div(d, zf, Fx, Fy, q);
  // the zone's primitives and their mirror image about the radial face
  // normal: rho, p and the entropy are unchanged, v and B lose twice their
  // normal part (the primitives of the mirrored ghost zone of k_radial_bc)
  const double zn = geo.t(d, T_ZSN, s, i, j);
  double A[3], Bv[3], Cv[3];
  for (int q = 0; q < 3; ++q) { A[q] = geo.t(d, T_ZSX + q, s, i, j) / zn; Bv[q] = geo.t(d, T_ZTX + q, s, i, j); }
  Cv[0] = A[1] * Bv[2] - A[2] * Bv[1];
  Cv[1] = A[2] * Bv[0] - A[0] * Bv[2];
  Cv[2] = A[0] * Bv[1] - A[1] * Bv[0];
  R Vc[9], Vm[9];
  for (int q = 0; q < 9; ++q) Vc[q] = (R)V[((int64_t)s * 9 + q) * d.cpp + c];
  {
    const R n0 = (R)A[0], n1 = (R)A[1], n2 = (R)A[2];
    R m = Vc[1] * n0 + Vc[2] * n1 + Vc[3] * n2;
    R b = Vc[5] * n0 + Vc[6] * n1 + Vc[7] * n2;
    Vm[0] = Vc[0]; Vm[4] = Vc[4]; Vm[8] = Vc[8];
    Vm[1] = Vc[1] - R(2) * m * n0; Vm[2] = Vc[2] - R(2) * m * n1; Vm[3] = Vc[3] - R(2) * m * n2;
    Vm[5] = Vc[5] - R(2) * b * n0; Vm[6] = Vc[6] - R(2) * b * n1; Vm[7] = Vc[7] - R(2) * b * n2;
  }
  // states of the outer wall face (k = 1: zone | ghost) and the inner one (ghost | zone)
  R Wp[9], Wm[9];
  const bool weno = pr.recon == CS_RECON_WENO5;
  for (int q = 0; q < 9; ++q) {
    if (weno) {
      weno5_mirror<R>(Vc[q], Vm[q], Wp[q], Wm[q]);  // weno5(c,m,c,m,c), weno5(m,c,m,c,m)
    } else {
      Wp[q] = Vc[q];
      Wm[q] = Vm[q];
    }
  }
  const bool okp = (Wp[0] > R(0)) && (Wp[4] > R(0)) && (Wp[8] > R(0));
  const bool okm = (Wm[0] > R(0)) && (Wm[4] > R(0)) && (Wm[8] > R(0));
  const int64_t so = (int64_t)s * d.fpp;
  double fo[9], fi[9];
  R sl, sr;
  solve_face<R>(okp ? Wp : Vc, okm ? Wm : Vm, A, Bv, Cv, in.Bz[so + d.face(i, j, k)], pr, fo, sl, sr);
  solve_face<R>(okm ? Wm : Vm, okp ? Wp : Vc, A, Bv, Cv, in.Bz[so + d.face(i, j, k - 1)], pr, fi, sl, sr);
  const double a1 = r1 * r1 * zn, a0 = r0 * r0 * zn;
  for (int q = 0; q < 9; ++q)
    if (q < 5 || q == 8) L[q] -= ov * (fo[q] * a1 - fi[q] * a0);
  const double rcz = 0.5 * (r0 + r1), gr = pr.grav != 0.0 ? -pr.grav / (rcz * rcz) : 0.0;
  if (pr.rot) {
    const double rp[3] = {rcz * A[0], rcz * A[1], rcz * A[2]};
    rotation_source(pr, in, d, s, c, rp, L);
  }
  double u[8], en;
  finish_zone(d, in, u0, s, c, dt, wa, wb, mode, gr, A, L, u, en);
  // V was read above for this zone only, so it is overwritten in place
  cleanup_zone<P, false>(d, geo, out, V, pr, flags, s, i, j, k, u, en);
}

// Constrained transport (484update_imex_rk_step_p1.f:646-908) with every edge
// circulation element E*l computed once per edge, so the discrete div B of
// each zone telescopes.
// With conducting walls the tangential EMFs on the walls vanish, so the wall
// z-faces keep their value; they are left untouched, which also lets the
// surface-mode zone update read them after this kernel has run in place.
__device__ __forceinline__ void update_faces(const CsDims& d, int i, int j, int k, int s, const Geo& geo,
                                             const State& in, const State& u0, const State& out,
                                             const double* __restrict__ Ex, const double* __restrict__ Ey,
                                             const double* __restrict__ Ez, double dt, double wa, double wb,
                                             int wall, int push) {
  const int N = d.N, NR = d.NR;
  const int64_t so = (int64_t)s * d.fpp;
  int64_t f = so + d.face(i, j, k);
  double r1 = geo.r(d, k), r0 = geo.r(d, k - 1);
  const bool xf = j >= 1 && k >= 1, yf = i >= 1 && k >= 1, zf = i >= 1 && j >= 1;
  // edge circulation elements, each read once (unique across sectors)
  const double ez = (xf || yf) ? ez_unique(d, Ez, s, i, j, k) : 0.0;
  // a face value and (push: shells) every lateral ghost face that copies it:
  // the exchange of the faces is done here
  auto put = [&](int fam, int fs, int fi, int fj, double v) {
    double* dst = fam == 0 ? out.Bx : (fam == 1 ? out.By : out.Bz);
    dst[(int64_t)fs * d.fpp + d.face(fi, fj, k)] = v;
    if (push) face_pushes(c_nbr, N, d.g, fam, fs, fi, fj, [&](int gf, int gs, int gi, int gj, double sg) {
      double* gd = gf == 0 ? out.Bx : (gf == 1 ? out.By : out.Bz);
      gd[(int64_t)gs * d.fpp + d.face(gi, gj, k)] = sg * v;
    });
  };
  // A face on a panel edge has two copies (k_avg_coinciding's pairs).  One
  // sector, the one that holds the pair's first copy there, updates it and
  // writes the negated value into the other copy, so the pair stays exactly
  // negation-consistent and the stage needs no averaging at the exchange.
  // x-face: x-vertex i in [0,N], cell j in [1,N], cell k in [1,NR]
  if (xf && !(i == 0 && s >= 3)) {
    double circ = ez - ez_unique(d, Ez, s, i, j - 1, k) - ey_unique(d, Ex, Ey, s, i, j, k) +
                  ey_unique(d, Ex, Ey, s, i, j, k - 1);
    double A = 0.5 * (r1 * r1 - r0 * r0) * geo.t(d, T_THY, s, i, j);
    double bn = in.Bx[f] - dt * circ / A;
    double v = wa != 0.0 ? wa * u0.Bx[f] + wb * bn : wb * bn;
    put(0, s, i, j, v);
    if (i == N) put(1, c_nbr[s][0], j, N, -v);
    else if (i == 0) put(0, c_nbr[s][1], 0, N + 1 - j, -v);
  }
  // y-face: cell i in [1,N], y-vertex j in [0,N], cell k in [1,NR]
  if (yf && j != N && !(j == 0 && s >= 3)) {
    double circ = -ez + ez_unique(d, Ez, s, i - 1, j, k) + ex_unique(d, Ex, Ey, s, i, j, k) -
                  ex_unique(d, Ex, Ey, s, i, j, k - 1);
    double A = 0.5 * (r1 * r1 - r0 * r0) * geo.t(d, T_THX, s, i, j);
    double bn = in.By[f] - dt * circ / A;
    double v = wa != 0.0 ? wa * u0.By[f] + wb * bn : wb * bn;
    put(1, s, i, j, v);
    if (j == 0) put(1, c_nbr[s][3], N + 1 - i, 0, -v);
  }
  // z-face: cells i,j in [1,N], z-vertex k in [0,NR]
  if (zf && !(wall && (k == 0 || k == NR))) {
    double circ = ey_unique(d, Ex, Ey, s, i, j, k) - ey_unique(d, Ex, Ey, s, i - 1, j, k) -
                  ex_unique(d, Ex, Ey, s, i, j, k) + ex_unique(d, Ex, Ey, s, i, j - 1, k);
    double A = r1 * r1 * geo.t(d, T_ZSN, s, i, j);
    double bn = in.Bz[f] - dt * circ / A;
    put(2, s, i, j, wa != 0.0 ? wa * u0.Bz[f] + wb * bn : wb * bn);
  }
}
__global__ void k_update_faces(CsDims d, Rng rg, Geo geo, State in, State u0, State out,
                               const double* __restrict__ Ex, const double* __restrict__ Ey,
                               const double* __restrict__ Ez, const double* dtp, double wa,
                               double wb, int wall, int push) {
  int i, j, k, s;
  if (!decode(rg, gtid(), i, j, k, s)) return;
  update_faces(d, i, j, k, s, geo, in, u0, out, Ex, Ey, Ez, dtp[0], wa, wb, wall, push);
}


// ------------------------------------------------------------------ reductions
// CFL (490evaluate_timestep.f:118-278): per zone cfl * L / (|v| + c_ms),
// reduced with a wavefront shuffle tree then one partial per block.
__device__ __forceinline__ double wave_min(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
  return v;
}
// EVALUATE_TIMESTEP (490evaluate_timestep.f:118-278) for one zone:
// cfl * L / (|v| + c_fast), L the shortest of its radial and lateral extents
__device__ __forceinline__ double cfl_zone(const CsDims& d, int i, int j, int k, int s, const Geo& geo,
                                           const State& st, const Params& pr) {
  int64_t c = d.cell(i, j, k);
  double u[8], w[8];
  for (int q = 0; q < 8; ++q) u[q] = st.U[(s * 8 + q) * d.cpp + c];
  double ent = st.ent[s * d.cpp + c];
  Gas<double> gs{pr.gamma, pr.gm1, pr.rhofloor, pr.prsfloor, pr.igm1};
  cons_to_prim(gs, u, w, ent);
  double b2 = w[5] * w[5] + w[6] * w[6] + w[7] * w[7];
  double ms = sqrt((pr.gamma * w[4] + C<double>::ob4pi * b2) / w[0]);
  double a21 = sqrt(w[1] * w[1] + w[2] * w[2] + w[3] * w[3]) + ms;
  double r1 = geo.r(d, k), r0 = geo.r(d, k - 1);
  double L = fmin(r1 - r0, r0 * geo.t(d, T_CHMIN, s, i, j));
  return pr.cfl * L / a21;
}
// minimum over the block's waves, written by thread 0
__device__ __forceinline__ void block_min_store(double v, double* out) {
  __shared__ double sm[16];
  v = wave_min(v);
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sm[w] = v;
  __syncthreads();
  if (threadIdx.x < 64) {
    double x = threadIdx.x < (blockDim.x >> 6) ? sm[threadIdx.x] : 1e300;
    x = wave_min(x);
    if (threadIdx.x == 0) *out = x;
  }
}
__global__ void k_cfl_partial(CsDims d, Rng rg, Geo geo, State st, Params pr, double* part) {
  int i, j, k, s;
  double v = 1e300;
  if (decode(rg, gtid(), i, j, k, s)) v = cfl_zone(d, i, j, k, s, geo, st, pr);
  block_min_store(v, part + blockIdx.x);
}
__global__ void k_cfl_final(const double* part, int n, double* dtp, double growth) {
  __shared__ double sm[16];
  double v = 1e300;
  for (int t = threadIdx.x; t < n; t += blockDim.x) v = fmin(v, part[t]);
  v = wave_min(v);
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sm[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double x = 1e300;
    for (int q = 0; q < (int)(blockDim.x >> 6); ++q) x = fmin(x, sm[q]);
    double cur = dtp[0];
    double nxt = cur > 0.0 ? fmin(growth * cur, x) : x;
    dtp[0] = nxt;
  }
}

// Compensated volume integrals: Neumaier two-sum per thread, two-sum tree.
__device__ __forceinline__ void two_sum(double a, double b, double& s, double& e) {
  s = a + b;
  double bp = s - a;
  e = (a - (s - bp)) + (b - bp);
}
__global__ void k_sums_partial(CsDims d, Rng rg, Geo geo, State st, double* part) {
  // part[(blk * 9 + q) * 2 + {0,1}]
  __shared__ double ss[256][2];
  for (int q = 0; q < 9; ++q) {
    int i, j, k, s;
    double v = 0.0;
    if (decode(rg, gtid(), i, j, k, s)) {
      int64_t c = d.cell(i, j, k);
      double r1 = geo.r(d, k), r0 = geo.r(d, k - 1);
      double vol = geo.t(d, T_OM, s, i, j) * (r1 * r1 * r1 - r0 * r0 * r0) / 3.0;
      v = vol * (q < 8 ? st.U[(s * 8 + q) * d.cpp + c] : st.ent[s * d.cpp + c]);
    }
    ss[threadIdx.x][0] = v;
    ss[threadIdx.x][1] = 0.0;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
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
// one workgroup of 256 per variable (blockIdx.x): strided two-sum partials,
// then a two-sum tree
__global__ void __launch_bounds__(256) k_sums_final(const double* part, int n, double* out) {
  __shared__ double ss[256][2];
  const int q = blockIdx.x;
  double s = 0.0, e = 0.0;
  for (int b = threadIdx.x; b < n; b += blockDim.x) {
    double t, te;
    two_sum(s, part[(b * 9 + q) * 2], t, te);
    s = t;
    e += te + part[(b * 9 + q) * 2 + 1];
  }
  ss[threadIdx.x][0] = s;
  ss[threadIdx.x][1] = e;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      double sa, ea;
      two_sum(ss[threadIdx.x][0], ss[threadIdx.x + o][0], sa, ea);
      ss[threadIdx.x][0] = sa;
      ss[threadIdx.x][1] = ss[threadIdx.x][1] + ss[threadIdx.x + o][1] + ea;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[q] = ss[0][0] + ss[0][1];
}

// relative discrete divergence |sum A_f B_f| / sum |A_f B_f| per zone, max
__global__ void k_divb_partial(CsDims d, Rng rg, Geo geo, State st, double* part) {
  __shared__ double sm[16];
  int i, j, k, s;
  double v = 0.0;
  if (decode(rg, gtid(), i, j, k, s)) {
    const int64_t so = (int64_t)s * d.fpp;
    double r1 = geo.r(d, k), r0 = geo.r(d, k - 1);
    double hr = 0.5 * (r1 * r1 - r0 * r0);
    double ax1 = hr * geo.t(d, T_THY, s, i, j), ax0 = hr * geo.t(d, T_THY, s, i - 1, j);
    double ay1 = hr * geo.t(d, T_THX, s, i, j), ay0 = hr * geo.t(d, T_THX, s, i, j - 1);
    double zn = geo.t(d, T_ZSN, s, i, j);
    double az1 = r1 * r1 * zn, az0 = r0 * r0 * zn;
    double p[6] = {ax1 * st.Bx[so + d.face(i, j, k)], -ax0 * st.Bx[so + d.face(i - 1, j, k)],
                   ay1 * st.By[so + d.face(i, j, k)], -ay0 * st.By[so + d.face(i, j - 1, k)],
                   az1 * st.Bz[so + d.face(i, j, k)], -az0 * st.Bz[so + d.face(i, j, k - 1)]};
    double sum = 0.0, sa = 0.0;
    for (int q = 0; q < 6; ++q) { sum += p[q]; sa += fabs(p[q]); }
    v = sa > 0.0 ? fabs(sum) / sa : 0.0;
  }
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sm[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double x = 0.0;
    for (int q = 0; q < (int)(blockDim.x >> 6); ++q) x = fmax(x, sm[q]);
    part[blockIdx.x] = x;
  }
}

// ------------------------------------------------------------------ batched shells
// Many independent shells of different sizes advance together: each phase of
// a stage is one launch over all shells.  Every shell's work of a phase starts
// on a workgroup boundary (boff: prefix sums of its workgroup counts), so a
// workgroup belongs to one shell, found by a binary search of boff; per-shell
// reductions (the CFL step) stay per workgroup.
struct ShellDev {
  CsDims d;
  Geo geo;
  Params pr;
  State st[2];       // [0] U^n, [1] stage scratch
  void* V;
  double *F[3], *S[3], *E[3];
  double *dt, *part;
  int* flags;
  Rng all, fx, fy, fz, ez, ex, ey, cells, faces;
};
enum BatchPhase { BP_EXCH = 0, BP_RAD, BP_FLUXXY, BP_FLUXZ, BP_EMF, BP_FACES, BP_CELLS, BP_IMEX, BP_CFL, BP_COUNT };

__device__ __forceinline__ int shell_of(const uint32_t* boff, int nsh, uint32_t blk) {
  int lo = 0, hi = nsh - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (boff[mid] <= blk) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// bin / bout / b0: stage input, output and U^n buffers; mode: stage combination
template <int PH, typename R, typename P>
__global__ void __launch_bounds__(256, 2) k_batch(const ShellDev* __restrict__ sh, const uint32_t* __restrict__ boff,
                                                  int nsh, int bin, int bout, double wa, double wb, int mode) {
  const uint32_t lb = xcd_block();
  const int b = shell_of(boff, nsh, lb);
  const ShellDev& S = sh[b];
  const uint32_t t = (lb - boff[b]) * blockDim.x + threadIdx.x;
  const State& in = S.st[bin];
  P* V = (P*)S.V;
  int i, j, k, s;
  if (PH == BP_EXCH) {
    uint32_t nc, nf, nk;
    exchange_counts(S.d, nc, nf, nk);
    if (t < nc) exchange_cells<P>(S.d, in, V, t);
    else if (t < nc + nf) exchange_faces(S.d, in, t - nc);
    else if (t < nc + nf + nk) corner_fill<P>(S.d, in, V, t - nc - nf);
  } else if (PH == BP_RAD) {
    const uint32_t nl = S.d.N + 2 * S.d.g + 1;
    if (t < 12u * S.d.g * nl * nl) radial_bc<P>(S.d, S.geo, in, S.pr.wall, V, S.pr, t);
  } else if (PH == BP_FLUXXY) {
    if (t < S.fx.total) {
      decode(S.fx, t, i, j, k, s);
      face_flux<0, R, P>(S.d, i, j, k, s, S.geo, in, V, S.F[0], S.S[0], S.pr);
    } else if (decode(S.fy, t - S.fx.total, i, j, k, s)) {
      face_flux<1, R, P>(S.d, i, j, k, s, S.geo, in, V, S.F[1], S.S[1], S.pr);
    }
  } else if (PH == BP_FLUXZ) {
    if (decode(S.fz, t, i, j, k, s)) face_flux<2, R, P>(S.d, i, j, k, s, S.geo, in, V, S.F[2], S.S[2], S.pr);
  } else if (PH == BP_EMF) {
    if (t < S.ez.total) {
      decode(S.ez, t, i, j, k, s);
      emf_z<P>(S.d, i, j, k, s, S.geo, in, V, S.S[0], S.S[1], S.E[2], S.pr);
    } else if (t < S.ez.total + S.ex.total) {
      decode(S.ex, t - S.ez.total, i, j, k, s);
      emf_xy<0, P>(S.d, i, j, k, s, S.geo, in, V, S.S[1], S.S[2], S.E[0], S.pr);
    } else if (decode(S.ey, t - S.ez.total - S.ex.total, i, j, k, s)) {
      emf_xy<1, P>(S.d, i, j, k, s, S.geo, in, V, S.S[2], S.S[0], S.E[1], S.pr);
    }
  } else if (PH == BP_FACES) {
    if (decode(S.faces, t, i, j, k, s))
      update_faces(S.d, i, j, k, s, S.geo, in, S.st[0], S.st[bout], S.E[0], S.E[1], S.E[2], S.dt[0], wa, wb,
                   S.pr.wall, !S.pr.surf);
  } else if (PH == BP_CELLS) {
    if (decode(S.cells, t, i, j, k, s))
      update_cells<P>(S.d, i, j, k, s, S.geo, in, S.st[0], S.st[bout], S.F[0], S.F[1], S.F[2], S.dt[0], wa, wb,
                      S.pr, V, S.flags, mode);
  } else if (PH == BP_IMEX) {
    if (decode(S.cells, t, i, j, k, s)) imex_source<P>(S.d, i, j, k, s, S.geo, S.st[0], S.st[1], V, S.dt[0], S.pr, S.flags);
  } else if (PH == BP_CFL) {
    double v = 1e300;
    if (decode(S.cells, t, i, j, k, s)) v = cfl_zone(S.d, i, j, k, s, S.geo, S.st[0], S.pr);
    block_min_store(v, S.part + (lb - boff[b]));
  }
}

// per shell: dt = min(1.1 dt, min over its CFL workgroups)
__global__ void k_batch_cfl_final(const ShellDev* __restrict__ sh, const uint32_t* __restrict__ boff) {
  const ShellDev& S = sh[blockIdx.x];
  const int n = (int)(boff[blockIdx.x + 1] - boff[blockIdx.x]);
  __shared__ double sm[16];
  double v = 1e300;
  for (int t = threadIdx.x; t < n; t += blockDim.x) v = fmin(v, S.part[t]);
  v = wave_min(v);
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sm[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double x = 1e300;
    for (int q = 0; q < (int)(blockDim.x >> 6); ++q) x = fmin(x, sm[q]);
    double cur = S.dt[0];
    S.dt[0] = cur > 0.0 ? fmin(1.1 * cur, x) : x;
  }
}
}  // namespace cs

// ====================================================================== host
using namespace cs;

struct cs_handle {
  cs_config cfg;
  CsDims d;
  Params pr;
  hipStream_t stream = nullptr;
  double* T = nullptr;
  double* rb = nullptr;
  double* buf[2][5] = {};  // [stage buffer][U, ent, Bx, By, Bz]
  void* V = nullptr;
  double* F[3] = {};
  double* S[3] = {};
  double* E[3] = {};
  double* dt = nullptr;
  double* part = nullptr;
  int* flags = nullptr;
  int64_t npart = 0;
  hipGraphExec_t graph = nullptr;
  int cur = 0;  // which buffer holds U^n
  int surf = 0; // one radial layer between walls: surface-mode stage kernels
  int vbuf = -1;  // stage buffer whose owned-zone primitives V holds (-1: none)
  // stage buffer whose lateral ghost zones (with their primitives in V) and
  // ghost faces were written by the zone and face updates that produced it,
  // so the next stage needs no exchange (-1: none)
  int vghost = -1;
  struct Prof* prof = nullptr;  // per-launch event pairs while cs_profile_steps runs
  p1s::Stage* p1 = nullptr;     // the P1 stage's buffers (recon CS_RECON_P1ADP)
  int bs[CS_K_COUNT];           // workgroup size per kernel id (cs_autotune)
};

// Event pairs around the launches of one eagerly run step (cs_profile_steps).
constexpr int kMaxProf = 64;
struct Prof {
  hipEvent_t ev[2 * kMaxProf];
  int kid[kMaxProf];
  int n = 0;
};

namespace {

const int kBlock = 256;
inline unsigned nblk(uint64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }
inline unsigned nblk_b(uint64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

State state_of(cs_handle* h, int b) {
  return State{h->buf[b][0], h->buf[b][1], h->buf[b][2], h->buf[b][3], h->buf[b][4]};
}
Geo geo_of(cs_handle* h) { return Geo{h->T, h->rb, h->d.t2 * 6, h->d.t2}; }
p1s::Buf p1buf(cs_handle* h, int b) {
  return p1s::Buf{h->buf[b][0], h->buf[b][1], h->buf[b][2], h->buf[b][3], h->buf[b][4]};
}

void launch_avg(cs_handle* h, int b) {
  const CsDims& d = h->d;
  k_avg_coinciding<<<nblk(12ull * d.N * d.NR), kBlock, 0, h->stream>>>(d, state_of(h, b));
}
// avg: average the two copies of every panel-edge face first.  The stage
// path leaves it out: its face update writes both copies from one value.
// lateral: the exchange between sectors; radial: the radial ghost layers
// (after the exchange: they pad the lateral ghost columns too).
template <typename P>
int launch_halo_t(cs_handle* h, int b, bool lateral, bool radial, bool avg) {
  const CsDims& d = h->d;
  State st = state_of(h, b);
  P* V = (P*)h->V;
  if (avg) launch_avg(h, b);
  uint32_t nc, nf, nk;
  exchange_counts(d, nc, nf, nk);
  const int K = h->bs[CS_K_HALO];
  if (lateral) k_exchange<P><<<nblk_b((uint64_t)nc + nf + nk, K), K, 0, h->stream>>>(d, st, V);
  uint64_t nl = (uint64_t)(d.N + 2 * d.g + 1);
  uint64_t n5 = 12ull * d.g * nl * nl;
  if (radial) k_radial_bc<P><<<nblk_b(n5, K), K, 0, h->stream>>>(d, geo_of(h), st, h->pr.wall, V, h->pr);
  CS_HIP(hipGetLastError());
  return CS_OK;
}
int launch_halo(cs_handle* h, int b, bool lateral = true, bool radial = true, bool avg = true) {
  return h->cfg.precision == CS_MIXED ? launch_halo_t<float>(h, b, lateral, radial, avg)
                                      : launch_halo_t<double>(h, b, lateral, radial, avg);
}

void stage_weights(int order, int irk, double& wa, double& wb, int& in, int& out, int imex = 0) {
  // buffers: 0 = U^n (relative to h->cur), 1 = scratch.  IMEX-SSP2(2,2,2)
  // reads its first flux stage from U1 = Un + gi dt S(U1) in the scratch buffer.
  if (order == 2 && imex) {
    if (irk == 1) { wa = 0.0; wb = 1.0; in = 1; out = 1; }
    else { wa = 0.5; wb = 0.5; in = 1; out = 0; }
    return;
  }
  if (order == 3) {
    if (irk == 1) { wa = 0.0; wb = 1.0; in = 0; out = 1; }
    else if (irk == 2) { wa = 0.75; wb = 0.25; in = 1; out = 1; }
    else { wa = 1.0 / 3.0; wb = 2.0 / 3.0; in = 1; out = 0; }
  } else {
    if (irk == 1) { wa = 0.0; wb = 1.0; in = 0; out = 1; }
    else { wa = 0.5; wb = 0.5; in = 1; out = 0; }
  }
}

int stage_input(cs_handle* h, int irk) {
  double wa, wb;
  int bi, bo;
  stage_weights(h->cfg.rk_order, irk, wa, wb, bi, bo, h->pr.imex);
  return (h->cur + bi) & 1;
}

// index ranges each kernel covers (faces and edges needed by the update and
// by the EMF speed look-ups, one layer into the ghost zones)
struct Ranges {
  Rng all, fx, fy, fz, ez, ex, ey, cells, faces;
};
// With conducting walls (surface mode) the x/y-edge EMFs on the walls are
// zero, so neither they nor the lateral faces of the radial ghost layers that
// would feed their speeds are computed.
Ranges ranges_of(const CsDims& d, int wall) {
  const int N = d.N, NR = d.NR, g = d.g;
  const int kl0 = wall ? 1 : 0, kl1 = wall ? NR : NR + 1;
  const int ke0 = wall ? 1 : 0, ke1 = wall ? NR - 1 : NR;
  Ranges r;
  r.all = make_rng(1 - g, N + g, 1 - g, N + g, 1 - g, NR + g);
  // z-faces (stencil along k), y-edges (zones at i-1, i and k-1, k) and zones
  // (update, implicit source, CFL) go k before j
  r.fx = make_rng(0, N, 0, N + 1, kl0, kl1);
  r.fy = make_rng(0, N + 1, 0, N, kl0, kl1);
  r.fz = make_rng(0, N + 1, 0, N + 1, 0, NR, 1);
  r.ez = make_rng(0, N, 0, N, 1, NR);
  r.ex = make_rng(1, N, 0, N, ke0, ke1);
  r.ey = make_rng(0, N, 1, N, ke0, ke1, 1);
  r.cells = make_rng(1, N, 1, N, 1, NR, 1);
  r.faces = make_rng(0, N, 0, N, 0, NR);
  return r;
}

// Number of x/y-edge levels along panel edges that carry an EMF: walls hold
// zero EMF on their tangential edges.
void sync_levels(const cs_handle* h, int& kh0, int& nkh) {
  kh0 = h->pr.wall ? 1 : 0;
  nkh = h->pr.wall ? h->d.NR - 1 : h->d.NR + 1;
}
uint64_t sync_units(const cs_handle* h, bool flux, bool edge) {
  const CsDims& d = h->d;
  int kh0, nkh;
  sync_levels(h, kh0, nkh);
  uint64_t n = 0;
  if (flux) n += 12ull * d.N * d.NR;
  if (edge) n += 12ull * (d.N - 1) * d.NR + 12ull * d.N * nkh + 8ull * d.NR;
  return n;
}
int launch_sync(cs_handle* h, bool flux, bool edge) {
  int kh0, nkh;
  sync_levels(h, kh0, nkh);
  uint64_t n = sync_units(h, flux, edge);
  if (flux && edge)
    k_sync_shared<true, true><<<nblk(n), kBlock, 0, h->stream>>>(h->d, h->F[0], h->F[1], h->E[0], h->E[1], h->E[2],
                                                                 kh0, nkh);
  else if (flux)
    k_sync_shared<true, false><<<nblk(n), kBlock, 0, h->stream>>>(h->d, h->F[0], h->F[1], h->E[0], h->E[1], h->E[2],
                                                                  kh0, nkh);
  else
    k_sync_shared<false, true><<<nblk(n), kBlock, 0, h->stream>>>(h->d, h->F[0], h->F[1], h->E[0], h->E[1], h->E[2],
                                                                  kh0, nkh);
  CS_HIP(hipGetLastError());
  return CS_OK;
}

// Launch one kernel of the stage on buffer b (update kernels use irk).
template <typename R, typename P>
int launch_kernel_t(cs_handle* h, int kid, int b, int irk, bool surf, int halo) {
  const CsDims& d = h->d;
  const Ranges r = ranges_of(d, h->pr.wall);
  State st = state_of(h, b);
  Geo geo = geo_of(h);
  P* V = (P*)h->V;
  hipStream_t sm = h->stream;
  double wa = 0, wb = 1;
  int bi = 0, bo = 1;
  stage_weights(h->cfg.rk_order, irk, wa, wb, bi, bo, h->pr.imex);
  const int K = h->bs[kid];
  State in = state_of(h, (h->cur + bi) & 1), u0 = state_of(h, h->cur), out = state_of(h, (h->cur + bo) & 1);
  switch (kid) {
    case CS_K_HALO: return launch_halo(h, b, halo & 1, (halo & 2) != 0, false);  // the stage's halo
    case CS_K_PRIMS:
      if (surf) {
        Rng r1 = make_rng(1 - d.g, d.N + d.g, 1 - d.g, d.N + d.g, 1, 1);
        k_prims<P><<<nblk(r1.total), kBlock, 0, sm>>>(d, r1, st, V, h->pr);
      } else {
        k_prims<P><<<nblk(r.all.total), kBlock, 0, sm>>>(d, r.all, st, V, h->pr);
      }
      break;
    case CS_K_FLUX_X:
      k_face_flux<0, R, P><<<nblk_b(r.fx.total, K), K, 0, sm>>>(d, r.fx, geo, st, V, h->F[0], h->S[0], h->pr);
      break;
    case CS_K_FLUX_Y:
      k_face_flux<1, R, P><<<nblk_b(r.fy.total, K), K, 0, sm>>>(d, r.fy, geo, st, V, h->F[1], h->S[1], h->pr);
      break;
    case CS_K_FLUX_Z:
      k_face_flux<2, R, P><<<nblk_b(r.fz.total, K), K, 0, sm>>>(d, r.fz, geo, st, V, h->F[2], h->S[2], h->pr);
      break;
    case CS_K_FLUX_XY:
      k_face_flux_xy<R, P><<<nblk_b((uint64_t)r.fx.total + r.fy.total, K), K, 0, sm>>>(
          d, r.fx, r.fy, geo, st, V, h->F[0], h->S[0], h->F[1], h->S[1], h->pr);
      break;
    case CS_K_EMF:
      k_emf_all<P><<<nblk_b((uint64_t)r.ez.total + r.ex.total + r.ey.total, K), K, 0, sm>>>(
          d, r.ez, r.ex, r.ey, geo, st, V, h->S[0], h->S[1], h->S[2], h->E[0], h->E[1], h->E[2], h->pr);
      break;
    case CS_K_EMF_Z:
      k_emf_z<P><<<nblk_b(r.ez.total, K), K, 0, sm>>>(d, r.ez, geo, st, V, h->S[0], h->S[1], h->E[2], h->pr);
      break;
    case CS_K_EMF_X:
      if (r.ex.total) k_emf_xy<0, P><<<nblk_b(r.ex.total, K), K, 0, sm>>>(d, r.ex, geo, st, V, h->S[1], h->S[2], h->E[0], h->pr);
      break;
    case CS_K_EMF_Y:
      if (r.ey.total) k_emf_xy<1, P><<<nblk_b(r.ey.total, K), K, 0, sm>>>(d, r.ey, geo, st, V, h->S[2], h->S[0], h->E[1], h->pr);
      break;
    case CS_K_UPDATE_CELLS:  // zone update + cleanup + primitives of the output
      if (surf)
        k_update_cells_surf<R, P><<<nblk_b(r.cells.total, K), K, 0, sm>>>(d, r.cells, geo, in, u0, out, V, h->F[0],
                                                                          h->F[1], h->dt, wa, wb, h->pr, h->flags,
                                                                          h->pr.imex ? irk : 0);
      else
        k_update_cells<P><<<nblk_b(r.cells.total, K), K, 0, sm>>>(d, r.cells, geo, in, u0, out, h->F[0], h->F[1],
                                                                    h->F[2], h->dt, wa, wb, h->pr, V, h->flags,
                                                                    h->pr.imex ? irk : 0);
      break;
    case CS_K_UPDATE_FACES:
      k_update_faces<<<nblk_b(r.faces.total, K), K, 0, sm>>>(d, r.faces, geo, in, u0, out, h->E[0], h->E[1],
                                                               h->E[2], h->dt, wa, wb, h->pr.wall, !h->surf);
      break;
    case CS_K_CLEANUP:
      set_error("the cleanup runs inside the zone update (CS_K_UPDATE_CELLS)");
      return CS_E_INVALID;
    case CS_K_SYNC:
      return launch_sync(h, true, true);
    case CS_K_CFL: {
      unsigned nb = nblk(r.cells.total);
      k_cfl_partial<<<nb, kBlock, 0, sm>>>(d, r.cells, geo, state_of(h, h->cur), h->pr, h->part);
      k_cfl_final<<<1, 1024, 0, sm>>>(h->part, (int)nb, h->dt, 1.1);
      break;
    }
    default:
      set_error("unknown kernel id %d", kid);
      return CS_E_INVALID;
  }
  CS_HIP(hipGetLastError());
  return CS_OK;
}
// surf: use the surface-mode stage kernels (only from the stage path);
// halo: parts of CS_K_HALO, bit 0 the lateral exchange, bit 1 the radial layers
int launch_kernel(cs_handle* h, int kid, int b, int irk, bool surf = false, int halo = 3) {
  Prof* p = h->prof;
  const bool rec = p && p->n < kMaxProf;
  if (rec) CS_HIP(hipEventRecord(p->ev[2 * p->n], h->stream));
  int rc = h->cfg.precision == CS_MIXED ? launch_kernel_t<float, float>(h, kid, b, irk, surf, halo)
                                        : launch_kernel_t<double, double>(h, kid, b, irk, surf, halo);
  if (rec) {
    CS_HIP(hipEventRecord(p->ev[2 * p->n + 1], h->stream));
    p->kid[p->n++] = kid;
  }
  return rc;
}
int launch_seq(cs_handle* h, const int* ks, int n, int b, int irk, bool surf = false) {
  for (int q = 0; q < n; ++q) {
    int rc = launch_kernel(h, ks[q], b, irk, surf);
    if (rc) return rc;
  }
  return CS_OK;
}
// The API path converts every zone (ghosts included) before the fluxes and
// makes the panel-edge fluxes unique at once; the stage path has the
// primitives already (from the previous zone update and the exchange) and
// defers the flux averaging to the one shared-face/edge launch after the EMFs.
int launch_fluxes(cs_handle* h, int b, bool surf = false, bool api = true) {
  static const int ks[4] = {CS_K_PRIMS, CS_K_FLUX_X, CS_K_FLUX_Y, CS_K_FLUX_Z};
  int rc = api ? launch_seq(h, ks, surf ? 3 : 4, b, 1, surf) : launch_seq(h, ks + 1, surf ? 2 : 3, b, 1, surf);
  if (rc || !api) return rc;
  h->vbuf = b;
  return launch_sync(h, true, false);
}
// with_fluxes: also make the panel-edge fluxes unique in the same launch
int launch_emf(cs_handle* h, int b, bool surf = false, bool with_fluxes = false) {
  static const int ks[3] = {CS_K_EMF_Z, CS_K_EMF_X, CS_K_EMF_Y};
  int rc = launch_seq(h, ks, 3, b, 1, surf);
  if (rc) return rc;
  return launch_sync(h, with_fluxes, true);
}
// Faces first: the zone update finishes with the cleanup, which needs the
// updated faces.  The primitives then belong to the stage output.
int launch_update(cs_handle* h, int irk, bool surf = false) {
  static const int ks[2] = {CS_K_UPDATE_FACES, CS_K_UPDATE_CELLS};
  int rc = launch_seq(h, ks, 2, h->cur, irk, surf);
  if (rc) return rc;
  double wa, wb;
  int bi, bo;
  stage_weights(h->cfg.rk_order, irk, wa, wb, bi, bo, h->pr.imex);
  h->vbuf = (h->cur + bo) & 1;
  h->vghost = h->surf ? -1 : h->vbuf;  // shells: the updates wrote the lateral ghosts
  return CS_OK;
}
// Primitives of the owned zones of buffer b when V does not hold them (after
// an upload or a timing run); the exchange carries them into the ghosts.
template <typename P>
int launch_prims_interior_t(cs_handle* h, int b) {
  const Ranges r = ranges_of(h->d, h->pr.wall);
  k_prims<P><<<nblk(r.cells.total), kBlock, 0, h->stream>>>(h->d, r.cells, state_of(h, b), (P*)h->V, h->pr);
  CS_HIP(hipGetLastError());
  h->vbuf = b;
  h->vghost = -1;
  return CS_OK;
}
int ensure_prims(cs_handle* h, int b) {
  if (h->vbuf == b) return CS_OK;
  return h->cfg.precision == CS_MIXED ? launch_prims_interior_t<float>(h, b) : launch_prims_interior_t<double>(h, b);
}
int launch_cfl(cs_handle* h) {
  if (h->p1) {
    int rc = p1s::cfl_partial(h->p1, p1buf(h, h->cur), h->part, h->stream);
    if (rc) return rc;
    k_cfl_final<<<1, 1024, 0, h->stream>>>(h->part, p1s::nblocks_cfl(), h->dt, 1.1);
    CS_HIP(hipGetLastError());
    return CS_OK;
  }
  return launch_kernel(h, CS_K_CFL, h->cur, 1);
}

// IMEX-SSP2(2,2,2) implicit first stage into the scratch buffer: its faces
// are copied from U^n (the stage leaves B alone), then U1 with its cleanup.
template <typename P>
int launch_imex_source_t(cs_handle* h) {
  const int sb = (h->cur + 1) & 1;
  const size_t fb = sizeof(double) * h->d.fpp * 6;
  for (int q = 2; q < 5; ++q)
    CS_HIP(hipMemcpyAsync(h->buf[sb][q], h->buf[h->cur][q], fb, hipMemcpyDeviceToDevice, h->stream));
  const Ranges r = ranges_of(h->d, h->pr.wall);
  k_imex_source<P><<<nblk(r.cells.total), kBlock, 0, h->stream>>>(h->d, r.cells, geo_of(h), state_of(h, h->cur),
                                                                   state_of(h, sb), (P*)h->V, h->dt, h->pr,
                                                                   h->flags);
  CS_HIP(hipGetLastError());
  h->vbuf = sb;
  h->vghost = h->surf ? -1 : sb;
  return CS_OK;
}
int launch_imex_source(cs_handle* h) {
  return h->cfg.precision == CS_MIXED ? launch_imex_source_t<float>(h) : launch_imex_source_t<double>(h);
}

// The stage of the reference's P1 scheme (recon CS_RECON_P1ADP): the full
// exchange of the stage input, the reconstruction, face fluxes and edge
// fields of the six sectors, the panel-edge averages of the fluxes and edge
// fields (k_sync_shared), the update; the IMEX stages begin with the
// implicit source stage.
int launch_stage_p1(cs_handle* h, int irk) {
  double wa, wb;
  int bi, bo, rc;
  stage_weights(h->cfg.rk_order, irk, wa, wb, bi, bo, h->pr.imex);
  const int b = (h->cur + bi) & 1, bz = h->cur, bt = (h->cur + bo) & 1;
  if (h->pr.imex && irk == 1 &&
      (rc = p1s::imex_source(h->p1, p1buf(h, h->cur), p1buf(h, (h->cur + 1) & 1), h->dt, h->stream)))
    return rc;
  // with the average of the two copies of every panel-edge face, as each
  // exchange of the reference (330fetch_buffers_fill_in_ux.f:337-350)
  p1s::mark_begin(h->p1, h->stream);
  if ((rc = launch_halo(h, b, true, true, true))) return rc;
  // the radial ghost layers as the driver's radial boundary routines leave
  // them: PAD_BOUNDARY_R (outflow pads) or ADD_BOUNDARY (the problem's state)
  if (h->cfg.radial_bc == CS_RBC_PAD && h->cfg.nr > 1 && (rc = p1s::pad_radial(h->p1, p1buf(h, b), h->stream)))
    return rc;
  if (h->cfg.radial_bc == CS_RBC_PROBLEM &&
      (rc = p1s::init_problem(h->p1, p1buf(h, b), h->cfg.bc_problem, 0, 0.0, 3, h->stream)))
    return rc;
  p1s::mark_end(h->p1, h->stream, CS_P1K_EXCHANGE);
  if ((rc = p1s::front(h->p1, p1buf(h, b), h->F, h->S, h->E, h->flags + 2, h->stream))) return rc;
  p1s::mark_begin(h->p1, h->stream);
  if ((rc = launch_sync(h, true, true))) return rc;
  p1s::mark_end(h->p1, h->stream, CS_P1K_SYNC);
  return p1s::update(h->p1, p1buf(h, b), p1buf(h, bz), p1buf(h, bt), h->F, h->E, h->dt, wa, wb,
                     h->pr.imex ? irk : 0, h->stream);
}

int launch_stage(cs_handle* h, int irk) {
  if (h->p1) return launch_stage_p1(h, irk);
  int b = stage_input(h, irk);
  int rc;
  const bool surf = h->surf != 0;
  if (h->pr.imex && irk == 1 && (rc = launch_imex_source(h))) return rc;
  // the updates that produced b wrote its lateral ghosts as they went; after
  // an upload or a timing run the primitives and the exchange are made here.
  // Panel-edge faces are consistent (written once by the previous face update).
  const bool lateral = !(h->vbuf == b && h->vghost == b);
  if (lateral && (rc = ensure_prims(h, b))) return rc;
  const int halo = (lateral ? 1 : 0) | (surf ? 0 : 2);  // surface stages never read the radial ghosts
  if (halo && (rc = launch_kernel(h, CS_K_HALO, b, 1, surf, halo))) return rc;
  if (lateral) h->vghost = b;
  // lateral fluxes in one launch, radial ones (shells), all edge EMFs in one
  // launch; panel-edge fluxes and EMFs are made unique where they are read
  static const int ks[3] = {CS_K_FLUX_XY, CS_K_FLUX_Z, CS_K_EMF};
  if ((rc = launch_kernel(h, ks[0], b, 1, surf))) return rc;
  if (!surf && (rc = launch_kernel(h, ks[1], b, 1, surf))) return rc;
  if ((rc = launch_kernel(h, ks[2], b, 1, surf))) return rc;
  return launch_update(h, irk, surf);
}

int launch_step(cs_handle* h) {
  int rc;
  for (int irk = 1; irk <= h->cfg.rk_order; ++irk)
    if ((rc = launch_stage(h, irk))) return rc;
  return launch_cfl(h);
}

void free_all(cs_handle* h) {
  if (h->graph) hipGraphExecDestroy(h->graph);
  hipFree(h->T); hipFree(h->rb);
  for (int b = 0; b < 2; ++b)
    for (int q = 0; q < 5; ++q) hipFree(h->buf[b][q]);
  hipFree(h->V);
  for (int q = 0; q < 3; ++q) { hipFree(h->F[q]); hipFree(h->S[q]); hipFree(h->E[q]); }
  hipFree(h->dt); hipFree(h->part); hipFree(h->flags);
  p1s::destroy(h->p1);
  h->p1 = nullptr;
  if (h->stream) hipStreamDestroy(h->stream);
}

// The eight cube corners: group the 24 panel-corner vertices by position on
// the unit sphere (each corner is shared by three sectors), ascending sector
// order within a group, and load the tables every kernel reads.
int load_corner_table(const CsDims& d, const std::vector<double>& T) {
  const int64_t plane = d.t2 * 6;
  struct Vx { int s, b; double u[3]; bool used; } v[24];
  int n = 0;
  for (int s = 0; s < 6; ++s)
    for (int b = 0; b < 4; ++b) {
      Vx& x = v[n++];
      x.s = s; x.b = b; x.used = false;
      const int i = (b & 1) ? d.N : 0, j = (b & 2) ? d.N : 0;
      for (int q = 0; q < 3; ++q) x.u[q] = T[(T_UX + q) * plane + s * d.t2 + d.t(i, j)];
    }
  int corner[8][3], cornerof[6][4], nc = 0;
  for (int a = 0; a < 24; ++a) {
    if (v[a].used) continue;
    int m[3] = {a, -1, -1}, nm = 1;
    for (int b = a + 1; b < 24 && nm <= 3; ++b) {
      double dx = v[a].u[0] - v[b].u[0], dy = v[a].u[1] - v[b].u[1], dz = v[a].u[2] - v[b].u[2];
      if (!v[b].used && dx * dx + dy * dy + dz * dz < 1e-20) {
        if (nm == 3) { nm = 4; break; }
        m[nm++] = b;
      }
    }
    if (nm != 3 || nc == 8) { set_error("cube-corner vertices do not group in threes"); return CS_E_INVALID; }
    for (int q = 0; q < 3; ++q) {  // v is in ascending sector order already
      v[m[q]].used = true;
      corner[nc][q] = 4 * v[m[q]].s + v[m[q]].b;
      cornerof[v[m[q]].s][v[m[q]].b] = nc;
    }
    ++nc;
  }
  if (nc != 8) { set_error("found %d cube corners", nc); return CS_E_INVALID; }
  CS_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_corner), corner, sizeof(corner)));
  CS_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_cornerof), cornerof, sizeof(cornerof)));
  return CS_OK;
}

#define ALLOC(ptr, bytes)                                                           \
  do {                                                                             \
    if (hipMalloc((void**)&(ptr), (bytes)) != hipSuccess) {                        \
      set_error("hipMalloc of %.3f GB failed", (double)(bytes) / 1e9);             \
      free_all(h); delete h;                                                       \
      return CS_E_NOMEM;                                                           \
    }                                                                              \
    if (hipMemset((ptr), 0, (bytes)) != hipSuccess) {                              \
      set_error("hipMemset failed"); free_all(h); delete h; return CS_E_HIP;       \
    }                                                                              \
  } while (0)

}  // namespace

extern "C" int cs_create(const cs_config* cfg, cs_handle** out) {
  if (!out) { set_error("null output"); return CS_E_INVALID; }
  *out = nullptr;
  std::vector<double> rb, T;
  int rc = host_metric(cfg, rb, T);
  if (rc) return rc;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    set_error("no HIP device available");
    return CS_E_HIP;
  }
  cs_handle* h = new cs_handle();
  for (int q = 0; q < CS_K_COUNT; ++q) h->bs[q] = kBlock;
  h->cfg = *cfg;
  h->d = cs_make_dims(cfg->ntp, cfg->nr, cfg->nghost);
  h->pr = Params{cfg->gamma, cfg->gamma - 1.0, cfg->rhofloor, cfg->prsfloor, cfg->cfl, cfg->grav,
                 1.0 / (cfg->gamma - 1.0), cfg->recon, cfg->riemann, cfg->emf, cfg->radial_bc == CS_RBC_WALL ? 1 : 0,
                 cfg->imex, (cfg->radial_bc == CS_RBC_WALL && cfg->nr == 1) ? 1 : 0,
                 {cfg->omega[0], cfg->omega[1], cfg->omega[2]},
                 (cfg->omega[0] != 0.0 || cfg->omega[1] != 0.0 || cfg->omega[2] != 0.0) ? 1 : 0};
  h->surf = (cfg->radial_bc == CS_RBC_WALL && cfg->nr == 1) ? 1 : 0;
  if ((rc = load_corner_table(h->d, T))) {
    delete h;
    return rc;
  }
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
    set_error("hipStreamCreate failed");
    delete h;
    return CS_E_HIP;
  }
  const CsDims& d = h->d;
  size_t cb = sizeof(double) * d.cpp * 6, fb = sizeof(double) * d.fpp * 6;
  ALLOC(h->T, sizeof(double) * T.size());
  ALLOC(h->rb, sizeof(double) * rb.size());
  for (int b = 0; b < 2; ++b) {
    ALLOC(h->buf[b][0], cb * 8);
    ALLOC(h->buf[b][1], cb);
    for (int q = 2; q < 5; ++q) ALLOC(h->buf[b][q], fb);
  }
  if (cfg->recon == CS_RECON_P1ADP) {
    size_t pb = 0;
    if ((rc = p1s::create(cfg, &h->p1, &pb))) {
      free_all(h);
      delete h;
      return rc;
    }
  } else {
    size_t vb = (cfg->precision == CS_MIXED ? sizeof(float) : sizeof(double)) * d.cpp * 6 * 9;
    ALLOC(h->V, vb);
  }
  for (int q = 0; q < 3; ++q) {
    ALLOC(h->F[q], fb * 9);
    ALLOC(h->S[q], fb * 2);
    ALLOC(h->E[q], fb);
  }
  ALLOC(h->dt, sizeof(double) * 2);
  h->npart = std::max<int64_t>((int64_t)nblk((uint64_t)6 * d.N * d.N * d.NR) + 16, p1s::nblocks_cfl());
  ALLOC(h->part, sizeof(double) * h->npart * 18);
  ALLOC(h->flags, sizeof(int) * 4);
  if (hipMemcpy(h->T, T.data(), sizeof(double) * T.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(h->rb, rb.data(), sizeof(double) * rb.size(), hipMemcpyHostToDevice) != hipSuccess) {
    set_error("metric upload failed");
    free_all(h);
    delete h;
    return CS_E_HIP;
  }
  *out = h;
  return CS_OK;
}

extern "C" int cs_destroy(cs_handle* h) {
  if (!h) return CS_OK;
  hipStreamSynchronize(h->stream);
  free_all(h);
  delete h;
  return CS_OK;
}

#define CHECK_H(h)                                  \
  do {                                              \
    if (!(h)) { set_error("null handle"); return CS_E_INVALID; } \
  } while (0)
// entry points of the WENO5/PCM stage path only
#define STAGE_PATH_ONLY(h, name)                                                             \
  do {                                                                                      \
    if ((h)->p1) { set_error("%s: not available with CS_RECON_P1ADP", name); return CS_E_INVALID; } \
  } while (0)

namespace {
// After U^n has been copied into stage buffer 0 (cs_upload_state,
// cs_read_dump): the two copies of each panel-edge face are made one
// (negation-consistent) value; from then on the face update keeps them so.
// The ghost zones and faces are filled once: faces the updates never write
// (conducting walls) keep these ghost copies in both stage buffers.
int finish_upload(cs_handle* h) {
  const CsDims& d = h->d;
  size_t cb = sizeof(double) * d.cpp * 6, fb = sizeof(double) * d.fpp * 6;
  int rc = launch_halo(h, 0, true, true, true);
  if (rc) return rc;
  // the scratch stage buffer starts as a copy so its ghost zones are defined
  for (int q = 0; q < 5; ++q)
    CS_HIP(hipMemcpyAsync(h->buf[1][q], h->buf[0][q], q == 0 ? cb * 8 : (q == 1 ? cb : fb),
                          hipMemcpyDeviceToDevice, h->stream));
  CS_HIP(hipStreamSynchronize(h->stream));
  return CS_OK;
}
}  // namespace

extern "C" int cs_upload_state(cs_handle* h, const double* U, const double* ent, const double* Bxf,
                               const double* Byf, const double* Bzf) {
  CHECK_H(h);
  if (!U || !ent || !Bxf || !Byf || !Bzf) { set_error("null array"); return CS_E_INVALID; }
  const CsDims& d = h->d;
  size_t cb = sizeof(double) * d.cpp * 6, fb = sizeof(double) * d.fpp * 6;
  h->cur = 0;
  h->vbuf = h->vghost = -1;
  CS_HIP(hipMemcpyAsync(h->buf[0][0], U, cb * 8, hipMemcpyHostToDevice, h->stream));
  CS_HIP(hipMemcpyAsync(h->buf[0][1], ent, cb, hipMemcpyHostToDevice, h->stream));
  CS_HIP(hipMemcpyAsync(h->buf[0][2], Bxf, fb, hipMemcpyHostToDevice, h->stream));
  CS_HIP(hipMemcpyAsync(h->buf[0][3], Byf, fb, hipMemcpyHostToDevice, h->stream));
  CS_HIP(hipMemcpyAsync(h->buf[0][4], Bzf, fb, hipMemcpyHostToDevice, h->stream));
  return finish_upload(h);
}

extern "C" int cs_download_state(cs_handle* h, double* U, double* ent, double* Bxf, double* Byf,
                                 double* Bzf) {
  CHECK_H(h);
  const CsDims& d = h->d;
  size_t cb = sizeof(double) * d.cpp * 6, fb = sizeof(double) * d.fpp * 6;
  int b = h->cur;
  if (U) CS_HIP(hipMemcpyAsync(U, h->buf[b][0], cb * 8, hipMemcpyDeviceToHost, h->stream));
  if (ent) CS_HIP(hipMemcpyAsync(ent, h->buf[b][1], cb, hipMemcpyDeviceToHost, h->stream));
  if (Bxf) CS_HIP(hipMemcpyAsync(Bxf, h->buf[b][2], fb, hipMemcpyDeviceToHost, h->stream));
  if (Byf) CS_HIP(hipMemcpyAsync(Byf, h->buf[b][3], fb, hipMemcpyDeviceToHost, h->stream));
  if (Bzf) CS_HIP(hipMemcpyAsync(Bzf, h->buf[b][4], fb, hipMemcpyDeviceToHost, h->stream));
  CS_HIP(hipStreamSynchronize(h->stream));
  return CS_OK;
}

// ---------------------------------------------------------------- checkpoints
// The dump format and its checks are in cs_dump.cpp; the state moves between
// device memory and the file through one pinned staging buffer, a chunk at a
// time, so no host copy of the whole state is made.
namespace {
constexpr uint64_t kDumpChunk = 8u << 20;  // float64 words per staged transfer (64 MiB)
struct PinnedStage {
  double* p = nullptr;
  ~PinnedStage() {
    if (p) hipHostFree(p);
  }
};
}  // namespace

extern "C" int cs_write_dump(cs_handle* h, const char* path, const cs_dump_meta* meta) {
  CHECK_H(h);
  if (!path || !meta) { set_error("cs_write_dump: null argument"); return CS_E_INVALID; }
  PinnedStage st;
  CS_HIP(hipHostMalloc((void**)&st.p, sizeof(double) * kDumpChunk, hipHostMallocDefault));
  cs_dump_meta m = *meta;
  CS_HIP(hipMemcpyAsync(&m.dtcur, h->dt, sizeof(double), hipMemcpyDeviceToHost, h->stream));
  CS_HIP(hipStreamSynchronize(h->stream));
  const int b = h->cur;
  DumpIo io;
  io.chunk = st.p;
  io.chunk_n = kDumpChunk;
  io.get = [&](int q, uint64_t off, uint64_t n, double* dst) -> int {
    CS_HIP(hipMemcpyAsync(dst, h->buf[b][q] + off, sizeof(double) * n, hipMemcpyDeviceToHost, h->stream));
    CS_HIP(hipStreamSynchronize(h->stream));
    return CS_OK;
  };
  return dump_write(path, &h->cfg, &m, io);
}

extern "C" int cs_read_dump(cs_handle* h, const char* path, cs_dump_meta* meta) {
  CHECK_H(h);
  if (!path) { set_error("cs_read_dump: null path"); return CS_E_INVALID; }
  cs_dump_meta local;
  if (!meta) meta = &local;
  // header and grid first: a refused file leaves the handle's state as it was
  DumpIo probe;
  int rc = dump_read(path, &h->cfg, meta, probe);
  if (rc) return rc;
  PinnedStage st;
  CS_HIP(hipHostMalloc((void**)&st.p, sizeof(double) * kDumpChunk, hipHostMallocDefault));
  h->cur = 0;
  h->vbuf = h->vghost = -1;
  DumpIo io;
  io.chunk = st.p;
  io.chunk_n = kDumpChunk;
  io.put = [&](int q, uint64_t off, uint64_t n, const double* src) -> int {
    CS_HIP(hipMemcpyAsync(h->buf[0][q] + off, src, sizeof(double) * n, hipMemcpyHostToDevice, h->stream));
    CS_HIP(hipStreamSynchronize(h->stream));
    return CS_OK;
  };
  if ((rc = dump_read(path, &h->cfg, meta, io))) return rc;
  if ((rc = finish_upload(h))) return rc;
  return cs_set_dt(h, meta->dtcur);
}

extern "C" int cs_halo_exchange(cs_handle* h) {
  CHECK_H(h);
  int rc = launch_halo(h, h->cur);
  if (rc) return rc;
  CS_HIP(hipStreamSynchronize(h->stream));
  return CS_OK;
}

// Ghost exchange of six P1 patches (one per cube sector, the layouts of
// cs_p1_patch) through the handle's exchange: mode 1 of each patch's U and
// entropy and its face B are gathered into the handle's current state (the
// per-sector layouts are the same, so these are plain copies), exchanged
// (FETCH_BUFFERS_FILL_IN_* and PAD_BOUNDARY_R, as cs_halo_exchange), and
// scattered back.  The work is ordered after the caller's stream and the
// caller's stream after it, through events.
extern "C" int cs_halo_exchange_p1(cs_handle* h, double* const* U, double* const* ent, double* const* Bxf,
                                   double* const* Byf, double* const* Bzf, void* stream) {
  CHECK_H(h);
  if (!U || !ent || !Bxf || !Byf || !Bzf) {
    set_error("cs_halo_exchange_p1: null array list");
    return CS_E_INVALID;
  }
  for (int s = 0; s < 6; ++s)
    if (!U[s] || !ent[s] || !Bxf[s] || !Byf[s] || !Bzf[s]) {
      set_error("cs_halo_exchange_p1: null array for sector %d", s);
      return CS_E_INVALID;
    }
  const CsDims& d = h->d;
  const int b = h->cur;
  double* const* dst = h->buf[b];
  hipStream_t cst = (hipStream_t)stream;
  hipEvent_t ev[2] = {nullptr, nullptr};
  // the events are destroyed on every path out
  auto run = [&]() -> int {
    for (int q = 0; q < 2; ++q) CS_HIP(hipEventCreateWithFlags(&ev[q], hipEventDisableTiming));
    CS_HIP(hipEventRecord(ev[0], cst));
    CS_HIP(hipStreamWaitEvent(h->stream, ev[0], 0));
    const size_t zb = (size_t)d.cpp * sizeof(double), fb = (size_t)d.fpp * sizeof(double);
    double* const* src[5] = {U, ent, Bxf, Byf, Bzf};
    const size_t per[5] = {8 * zb, zb, fb, fb, fb};
    for (int dir = 0; dir < 2; ++dir) {
      if (dir == 1) {
        int rc = launch_halo(h, b);
        if (rc) return rc;
      }
      for (int s = 0; s < 6; ++s)
        for (int a = 0; a < 5; ++a) {
          char* hs = (char*)dst[a] + per[a] * s;
          char* ps = (char*)src[a][s];
          CS_HIP(hipMemcpyAsync(dir == 0 ? hs : ps, dir == 0 ? ps : hs, per[a], hipMemcpyDeviceToDevice,
                                h->stream));
        }
    }
    CS_HIP(hipEventRecord(ev[1], h->stream));
    CS_HIP(hipStreamWaitEvent(cst, ev[1], 0));
    return CS_OK;
  };
  const int rc = run();
  for (int q = 0; q < 2; ++q)
    if (ev[q]) (void)hipEventDestroy(ev[q]);
  return rc;
}

extern "C" int cs_make_facial_fluxes(cs_handle* h) {
  CHECK_H(h);
  STAGE_PATH_ONLY(h, "cs_make_facial_fluxes");
  int rc = launch_fluxes(h, h->cur);
  if (rc) return rc;
  CS_HIP(hipStreamSynchronize(h->stream));
  return CS_OK;
}

extern "C" int cs_make_edge_efields(cs_handle* h) {
  CHECK_H(h);
  STAGE_PATH_ONLY(h, "cs_make_edge_efields");
  int rc = launch_emf(h, h->cur);
  if (rc) return rc;
  CS_HIP(hipStreamSynchronize(h->stream));
  return CS_OK;
}

extern "C" int cs_update_stage(cs_handle* h, int32_t irk) {
  CHECK_H(h);
  STAGE_PATH_ONLY(h, "cs_update_stage");
  if (irk < 1 || irk > h->cfg.rk_order) { set_error("irk %d out of range", irk); return CS_E_INVALID; }
  if (h->pr.imex) {
    set_error("implicit gravity needs the implicit first stage: advance with cs_rk_stage or cs_advance");
    return CS_E_INVALID;
  }
  int rc = launch_update(h, irk);
  if (rc) return rc;
  CS_HIP(hipStreamSynchronize(h->stream));
  return CS_OK;
}

extern "C" int cs_evaluate_timestep(cs_handle* h) {
  CHECK_H(h);
  int rc = launch_cfl(h);
  if (rc) return rc;
  CS_HIP(hipStreamSynchronize(h->stream));
  return CS_OK;
}

extern "C" int cs_set_dt(cs_handle* h, double dt) {
  CHECK_H(h);
  double v[2] = {dt, 0.0};
  CS_HIP(hipMemcpyAsync(h->dt, v, sizeof(v), hipMemcpyHostToDevice, h->stream));
  CS_HIP(hipStreamSynchronize(h->stream));
  return CS_OK;
}

extern "C" int cs_get_dt(cs_handle* h, double* dt) {
  CHECK_H(h);
  if (!dt) { set_error("null output"); return CS_E_INVALID; }
  CS_HIP(hipMemcpyAsync(dt, h->dt, sizeof(double), hipMemcpyDeviceToHost, h->stream));
  CS_HIP(hipStreamSynchronize(h->stream));
  return CS_OK;
}

extern "C" int cs_rk_stage(cs_handle* h, int32_t irk) {
  CHECK_H(h);
  if (irk < 1 || irk > h->cfg.rk_order) { set_error("irk %d out of range", irk); return CS_E_INVALID; }
  int rc = launch_stage(h, irk);
  if (rc) return rc;
  CS_HIP(hipStreamSynchronize(h->stream));
  return CS_OK;
}

static int check_flags(cs_handle* h) {
  int fl[4];
  CS_HIP(hipMemcpyAsync(fl, h->flags, sizeof(fl), hipMemcpyDeviceToHost, h->stream));
  CS_HIP(hipStreamSynchronize(h->stream));
  if (fl[1]) { set_error("NaN/Inf in the state"); return CS_E_NAN; }
  if (fl[2]) { set_error("the positivity limiter's loop-order replay did not settle within its sweeps"); return CS_E_INVALID; }
  return CS_OK;
}

static int run_steps(cs_handle* h, int nsteps, int use_graph) {
  int rc;
  if (!use_graph) {
    for (int n = 0; n < nsteps; ++n)
      if ((rc = launch_step(h))) return rc;
    return CS_OK;
  }
  // the captured step expects the primitives of U^n in V and leaves them there
  if (!h->p1 && (rc = ensure_prims(h, h->cur))) return rc;
  if (!h->graph) {
    hipGraph_t gr;
    CS_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    rc = launch_step(h);
    hipError_t e = hipStreamEndCapture(h->stream, &gr);
    if (rc) return rc;
    if (e != hipSuccess) { set_error("capture failed: %s", hipGetErrorString(e)); return CS_E_HIP; }
    CS_HIP(hipGraphInstantiate(&h->graph, gr, nullptr, nullptr, 0));
    hipGraphDestroy(gr);
  }
  for (int n = 0; n < nsteps; ++n) CS_HIP(hipGraphLaunch(h->graph, h->stream));
  return CS_OK;
}

extern "C" int cs_advance(cs_handle* h, int32_t nsteps, int32_t use_graph) {
  CHECK_H(h);
  if (nsteps < 0) { set_error("nsteps < 0"); return CS_E_INVALID; }
  int rc = run_steps(h, nsteps, use_graph);
  if (rc) return rc;
  CS_HIP(hipStreamSynchronize(h->stream));
  return check_flags(h);
}

extern "C" int cs_download_fluxes(cs_handle* h, int32_t dir, double* flux, double* speeds) {
  CHECK_H(h);
  if (dir < 0 || dir > 2) { set_error("dir must be 0..2"); return CS_E_INVALID; }
  size_t fb = sizeof(double) * h->d.fpp * 6;
  if (flux) CS_HIP(hipMemcpyAsync(flux, h->F[dir], fb * 9, hipMemcpyDeviceToHost, h->stream));
  if (speeds) CS_HIP(hipMemcpyAsync(speeds, h->S[dir], fb * 2, hipMemcpyDeviceToHost, h->stream));
  CS_HIP(hipStreamSynchronize(h->stream));
  return CS_OK;
}

extern "C" int cs_download_emf(cs_handle* h, double* Efx, double* Efy, double* Efz) {
  CHECK_H(h);
  size_t fb = sizeof(double) * h->d.fpp * 6;
  if (Efx) CS_HIP(hipMemcpyAsync(Efx, h->E[0], fb, hipMemcpyDeviceToHost, h->stream));
  if (Efy) CS_HIP(hipMemcpyAsync(Efy, h->E[1], fb, hipMemcpyDeviceToHost, h->stream));
  if (Efz) CS_HIP(hipMemcpyAsync(Efz, h->E[2], fb, hipMemcpyDeviceToHost, h->stream));
  CS_HIP(hipStreamSynchronize(h->stream));
  return CS_OK;
}

extern "C" int cs_conservation_sums(cs_handle* h, double* sums9) {
  CHECK_H(h);
  if (!sums9) { set_error("null output"); return CS_E_INVALID; }
  const CsDims& d = h->d;
  Rng rc = make_rng(1, d.N, 1, d.N, 1, d.NR);
  unsigned nb = nblk(rc.total);
  if (h->p1) {  // on the P1 geometry the stage conserves
    nb = (unsigned)p1s::diag_blocks(h->p1);
    int r = p1s::sums_partial(h->p1, p1buf(h, h->cur), h->part, h->stream);
    if (r) return r;
  } else {
    k_sums_partial<<<nb, kBlock, 0, h->stream>>>(d, rc, geo_of(h), state_of(h, h->cur), h->part);
  }
  k_sums_final<<<9, 256, 0, h->stream>>>(h->part, (int)nb, h->part + (int64_t)h->npart * 18 - 16);
  CS_HIP(hipGetLastError());
  CS_HIP(hipMemcpyAsync(sums9, h->part + (int64_t)h->npart * 18 - 16, sizeof(double) * 9,
                        hipMemcpyDeviceToHost, h->stream));
  CS_HIP(hipStreamSynchronize(h->stream));
  return CS_OK;
}

extern "C" int cs_max_divb(cs_handle* h, double* out) {
  CHECK_H(h);
  if (!out) { set_error("null output"); return CS_E_INVALID; }
  const CsDims& d = h->d;
  Rng rc = make_rng(1, d.N, 1, d.N, 1, d.NR);
  unsigned nb = nblk(rc.total);
  if (h->p1) {  // on the P1 geometry the constrained transport keeps
    nb = (unsigned)p1s::diag_blocks(h->p1);
    int r = p1s::divb_partial(h->p1, p1buf(h, h->cur), h->part, h->stream);
    if (r) return r;
  } else {
    k_divb_partial<<<nb, kBlock, 0, h->stream>>>(d, rc, geo_of(h), state_of(h, h->cur), h->part);
  }
  CS_HIP(hipGetLastError());
  std::vector<double> p(nb);
  CS_HIP(hipMemcpyAsync(p.data(), h->part, sizeof(double) * nb, hipMemcpyDeviceToHost, h->stream));
  CS_HIP(hipStreamSynchronize(h->stream));
  double m = 0.0;
  for (double v : p) m = v > m ? v : m;
  *out = m;
  return CS_OK;
}

extern "C" int cs_time_steps(cs_handle* h, int32_t nsteps, int32_t use_graph, float* ms_total) {
  CHECK_H(h);
  if (!ms_total) { set_error("null output"); return CS_E_INVALID; }
  hipEvent_t e0, e1;
  CS_HIP(hipEventCreate(&e0));
  CS_HIP(hipEventCreate(&e1));
  if (use_graph) {  // build the graph outside the timed region
    int rc = run_steps(h, 0, 1);
    if (rc) return rc;
  }
  CS_HIP(hipStreamSynchronize(h->stream));
  CS_HIP(hipEventRecord(e0, h->stream));
  int rc = run_steps(h, nsteps, use_graph);
  if (rc) return rc;
  CS_HIP(hipEventRecord(e1, h->stream));
  CS_HIP(hipEventSynchronize(e1));
  CS_HIP(hipEventElapsedTime(ms_total, e0, e1));
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  return check_flags(h);
}

extern "C" int cs_profile_steps(cs_handle* h, int32_t nsteps, float* ms_sum, int32_t* calls) {
  CHECK_H(h);
  STAGE_PATH_ONLY(h, "cs_profile_steps");
  if (!ms_sum || !calls || nsteps < 1) { set_error("bad arguments"); return CS_E_INVALID; }
  for (int q = 0; q < CS_K_COUNT; ++q) { ms_sum[q] = 0.f; calls[q] = 0; }
  Prof* p = new Prof();
  int made = 0, rc = CS_OK;
  for (; made < 2 * kMaxProf; ++made)
    if (hipEventCreate(&p->ev[made]) != hipSuccess) { rc = CS_E_HIP; set_error("hipEventCreate failed"); break; }
  if (rc == CS_OK) rc = ensure_prims(h, h->cur);
  if (rc == CS_OK) rc = hipStreamSynchronize(h->stream) == hipSuccess ? CS_OK : CS_E_HIP;
  h->prof = p;
  for (int n = 0; n < nsteps && rc == CS_OK; ++n) {
    p->n = 0;
    rc = launch_step(h);
    if (rc == CS_OK && hipStreamSynchronize(h->stream) != hipSuccess) { rc = CS_E_HIP; set_error("step failed"); }
    for (int q = 0; q < p->n && rc == CS_OK; ++q) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, p->ev[2 * q], p->ev[2 * q + 1]) != hipSuccess) {
        rc = CS_E_HIP;
        set_error("hipEventElapsedTime failed");
        break;
      }
      ms_sum[p->kid[q]] += ms;
      calls[p->kid[q]] += 1;
    }
  }
  h->prof = nullptr;
  for (int q = 0; q < made; ++q) hipEventDestroy(p->ev[q]);
  delete p;
  if (rc) return rc;
  return check_flags(h);
}

extern "C" int cs_init_problem_p1(cs_handle* h, int32_t problem, uint64_t seed, double time) {
  CHECK_H(h);
  if (!h->p1) { set_error("cs_init_problem_p1: the handle's recon is not CS_RECON_P1ADP"); return CS_E_INVALID; }
  int rc = p1s::init_problem(h->p1, p1buf(h, h->cur), problem, seed, time, 0, h->stream);
  if (rc) return rc;
  h->vbuf = -1;
  h->vghost = -1;
  CS_HIP(hipStreamSynchronize(h->stream));
  return CS_OK;
}

extern "C" int cs_add_boundary_p1(cs_handle* h, int32_t ibdy_inner, int32_t ibdy_outer, int32_t problem,
                                  uint64_t seed, double time) {
  CHECK_H(h);
  if (!h->p1) { set_error("cs_add_boundary_p1: the handle's recon is not CS_RECON_P1ADP"); return CS_E_INVALID; }
  const int layers = (ibdy_inner ? 1 : 0) | (ibdy_outer ? 2 : 0);
  if (!layers) return CS_OK;
  int rc = p1s::init_problem(h->p1, p1buf(h, h->cur), problem, seed, time, layers, h->stream);
  if (rc) return rc;
  CS_HIP(hipStreamSynchronize(h->stream));
  return CS_OK;
}

extern "C" int cs_profile_steps_p1(cs_handle* h, int32_t nsteps, float* ms_sum, int32_t* calls) {
  CHECK_H(h);
  if (!h->p1) { set_error("cs_profile_steps_p1: the handle's recon is not CS_RECON_P1ADP"); return CS_E_INVALID; }
  if (!ms_sum || !calls || nsteps < 1) { set_error("bad arguments"); return CS_E_INVALID; }
  for (int q = 0; q < CS_P1K_COUNT; ++q) { ms_sum[q] = 0.f; calls[q] = 0; }
  constexpr int cap = 256;  // launch groups per step (about 16 per stage)
  std::vector<hipEvent_t> ev(2 * cap);
  std::vector<int> kid(cap);
  int made = 0, rc = CS_OK;
  for (; made < 2 * cap; ++made)
    if (hipEventCreate(&ev[made]) != hipSuccess) { rc = CS_E_HIP; set_error("hipEventCreate failed"); break; }
  p1s::Marks m{ev.data(), kid.data(), 0, cap};
  if (rc == CS_OK) rc = hipStreamSynchronize(h->stream) == hipSuccess ? CS_OK : CS_E_HIP;
  p1s::set_marks(h->p1, &m);
  for (int n = 0; n < nsteps && rc == CS_OK; ++n) {
    m.n = 0;
    rc = launch_step(h);
    if (rc == CS_OK && hipStreamSynchronize(h->stream) != hipSuccess) { rc = CS_E_HIP; set_error("step failed"); }
    for (int q = 0; q < m.n && rc == CS_OK; ++q) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, ev[2 * q], ev[2 * q + 1]) != hipSuccess) {
        rc = CS_E_HIP;
        set_error("hipEventElapsedTime failed");
        break;
      }
      ms_sum[kid[q]] += ms;
      calls[kid[q]] += 1;
    }
  }
  p1s::set_marks(h->p1, nullptr);
  for (int q = 0; q < made; ++q) hipEventDestroy(ev[q]);
  if (rc) return rc;
  return check_flags(h);
}

extern "C" int cs_time_kernel(cs_handle* h, int32_t kernel, int32_t iters, float* ms_avg) {
  CHECK_H(h);
  STAGE_PATH_ONLY(h, "cs_time_kernel");
  if (!ms_avg || iters < 1 || kernel < 0 || kernel >= CS_K_COUNT) {
    set_error("bad arguments");
    return CS_E_INVALID;
  }
  hipEvent_t e0, e1;
  CS_HIP(hipEventCreate(&e0));
  CS_HIP(hipEventCreate(&e1));
  // the update kernels are timed as stage 2 of SSP-RK3 (scratch buffer in
  // place) or stage 1 of RK2 (scratch output), so U^n is never modified
  int irk = h->cfg.rk_order == 3 ? 2 : 1;
  CS_HIP(hipStreamSynchronize(h->stream));
  CS_HIP(hipEventRecord(e0, h->stream));
  for (int it = 0; it < iters; ++it) {
    // the halo is timed as the full exchange of a stage whose input was not
    // produced by the updates (with the radial layers of a shell)
    int rc = kernel == CS_K_STEP ? launch_step(h)
                                 : launch_kernel(h, kernel, h->cur, irk, h->surf != 0, h->surf ? 1 : 3);
    if (rc) return rc;
  }
  CS_HIP(hipEventRecord(e1, h->stream));
  CS_HIP(hipEventSynchronize(e1));
  h->vbuf = h->vghost = -1;  // a timed zone update or conversion leaves V holding another state
  float ms = 0.f;
  CS_HIP(hipEventElapsedTime(&ms, e0, e1));
  *ms_avg = ms / iters;
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  return CS_OK;
}

extern "C" int cs_autotune(cs_handle* h, int32_t iters, int32_t* block_sizes) {
  CHECK_H(h);
  STAGE_PATH_ONLY(h, "cs_autotune");
  if (iters < 1) { set_error("bad arguments"); return CS_E_INVALID; }
  // the launches of a stage; the CFL reduction keeps 256 (its partial sums
  // are sized for it)
  static const int ks[6] = {CS_K_HALO, CS_K_FLUX_XY, CS_K_FLUX_Z, CS_K_EMF, CS_K_UPDATE_FACES, CS_K_UPDATE_CELLS};
  static const int cand[3] = {64, 128, 256};
  for (int q = 0; q < 6; ++q) {
    if (ks[q] == CS_K_FLUX_Z && h->surf) continue;
    int best = kBlock;
    float tbest = 1e30f;
    for (int c = 0; c < 3; ++c) {
      h->bs[ks[q]] = cand[c];
      float ms = 0.f;
      int rc = cs_time_kernel(h, ks[q], iters, &ms);
      if (rc) return rc;
      if (ms < tbest) { tbest = ms; best = cand[c]; }
    }
    h->bs[ks[q]] = best;
  }
  if (h->graph) { hipGraphExecDestroy(h->graph); h->graph = nullptr; }
  if (block_sizes)
    for (int q = 0; q < CS_K_COUNT; ++q) block_sizes[q] = h->bs[q];
  return CS_OK;
}

extern "C" int cs_kernel_units(cs_handle* h, int32_t kernel, int64_t* units) {
  CHECK_H(h);
  STAGE_PATH_ONLY(h, "cs_kernel_units");
  if (!units || kernel < 0 || kernel >= CS_K_COUNT) { set_error("bad arguments"); return CS_E_INVALID; }
  const CsDims& d = h->d;
  const Ranges r = ranges_of(d, h->pr.wall);
  int64_t n = 0;
  switch (kernel) {
    case CS_K_HALO: n = 24ll * d.g * d.N * d.NR; break;
    case CS_K_PRIMS: n = h->surf ? 6ll * d.nc * d.nc : r.all.total; break;
    case CS_K_CLEANUP: n = 0; break;  // part of CS_K_UPDATE_CELLS
    case CS_K_FLUX_X: n = r.fx.total; break;
    case CS_K_FLUX_Y: n = r.fy.total; break;
    case CS_K_FLUX_Z: n = r.fz.total; break;
    case CS_K_EMF_Z: n = r.ez.total; break;
    case CS_K_EMF_X: n = r.ex.total; break;
    case CS_K_EMF_Y: n = r.ey.total; break;
    case CS_K_UPDATE_FACES: n = r.faces.total; break;
    case CS_K_SYNC: n = (int64_t)sync_units(h, true, true); break;
    case CS_K_FLUX_XY: n = (int64_t)r.fx.total + r.fy.total; break;
    case CS_K_EMF: n = (int64_t)r.ez.total + r.ex.total + r.ey.total; break;
    default: n = r.cells.total; break;
  }
  *units = n;
  return CS_OK;
}

extern "C" int cs_last_error(char* buf, size_t len) {
  if (!buf || len == 0) return CS_E_INVALID;
  snprintf(buf, len, "%s", g_err.c_str());
  return CS_OK;
}

extern "C" int cs_device_info(char* buf, size_t len) {
  if (!buf || len == 0) return CS_E_INVALID;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    snprintf(buf, len, "no device");
    return CS_E_HIP;
  }
  hipDeviceProp_t p;
  CS_HIP(hipGetDeviceProperties(&p, 0));
  snprintf(buf, len, "%s %s CUs=%d mem=%.1fGB", p.name, p.gcnArchName, p.multiProcessorCount,
           p.totalGlobalMem / 1e9);
  return CS_OK;
}

// ====================================================================== batched shells (host)
// A batch owns one cs_handle per shell (its buffers and metric) and advances
// them together: each phase of a stage is one launch over all shells
// (k_batch), the whole step one hipGraph.  The shells share the precision and
// the time integrator; sizes, schemes and radii are free.
struct cs_batch {
  std::vector<cs_handle*> sh;
  int n = 0, precision = 0, rk_order = 3, imex = 0;
  int nsurf = 0;  // surface-mode shells: their stages need the exchange phase
  hipStream_t stream = nullptr;
  ShellDev* dsh = nullptr;
  uint32_t* dboff[BP_COUNT] = {};
  std::vector<uint32_t> hoff[BP_COUNT];
  int bs[BP_COUNT];
  hipGraphExec_t graph = nullptr;
};

namespace {

uint64_t batch_work(const cs_handle* h, int ph) {
  const CsDims& d = h->d;
  const Ranges r = ranges_of(d, h->pr.wall);
  uint32_t nc, nf, nk;
  switch (ph) {
    case BP_EXCH: exchange_counts(d, nc, nf, nk); return (uint64_t)nc + nf + nk;
    case BP_RAD: { uint64_t nl = d.N + 2 * d.g + 1; return 12ull * d.g * nl * nl; }
    case BP_FLUXXY: return (uint64_t)r.fx.total + r.fy.total;
    case BP_FLUXZ: return r.fz.total;
    case BP_EMF: return (uint64_t)r.ez.total + r.ex.total + r.ey.total;
    case BP_FACES: return r.faces.total;
    default: return r.cells.total;  // BP_CELLS, BP_IMEX, BP_CFL
  }
}

int batch_offsets(cs_batch* B, int ph, int bs) {
  std::vector<uint32_t>& o = B->hoff[ph];
  o.assign(B->n + 1, 0);
  for (int b = 0; b < B->n; ++b) {
    uint64_t nb = (batch_work(B->sh[b], ph) + bs - 1) / bs;
    if ((uint64_t)o[b] + nb > 0x7fffffffull) { set_error("batch too large for one launch"); return CS_E_INVALID; }
    o[b + 1] = o[b] + (uint32_t)nb;
  }
  B->bs[ph] = bs;
  CS_HIP(hipMemcpy(B->dboff[ph], o.data(), sizeof(uint32_t) * (B->n + 1), hipMemcpyHostToDevice));
  return CS_OK;
}

ShellDev shell_desc(cs_handle* h) {
  ShellDev s;
  const Ranges r = ranges_of(h->d, h->pr.wall);
  s.d = h->d;
  s.geo = geo_of(h);
  s.pr = h->pr;
  s.st[0] = state_of(h, h->cur);
  s.st[1] = state_of(h, (h->cur + 1) & 1);
  s.V = h->V;
  for (int q = 0; q < 3; ++q) { s.F[q] = h->F[q]; s.S[q] = h->S[q]; s.E[q] = h->E[q]; }
  s.dt = h->dt;
  s.part = h->part;
  s.flags = h->flags;
  s.all = r.all; s.fx = r.fx; s.fy = r.fy; s.fz = r.fz; s.ez = r.ez; s.ex = r.ex; s.ey = r.ey;
  s.cells = r.cells; s.faces = r.faces;
  return s;
}

template <int PH, typename R, typename P>
int batch_launch_t(cs_batch* B, int bin, int bout, double wa, double wb, int mode) {
  const unsigned nb = B->hoff[PH][B->n];
  if (nb == 0) return CS_OK;
  k_batch<PH, R, P><<<nb, B->bs[PH], 0, B->stream>>>(B->dsh, B->dboff[PH], B->n, bin, bout, wa, wb, mode);
  CS_HIP(hipGetLastError());
  return CS_OK;
}
template <typename R, typename P>
int batch_phase_t(cs_batch* B, int ph, int bin, int bout, double wa, double wb, int mode) {
  switch (ph) {
    case BP_EXCH: return batch_launch_t<BP_EXCH, R, P>(B, bin, bout, wa, wb, mode);
    case BP_RAD: return batch_launch_t<BP_RAD, R, P>(B, bin, bout, wa, wb, mode);
    case BP_FLUXXY: return batch_launch_t<BP_FLUXXY, R, P>(B, bin, bout, wa, wb, mode);
    case BP_FLUXZ: return batch_launch_t<BP_FLUXZ, R, P>(B, bin, bout, wa, wb, mode);
    case BP_EMF: return batch_launch_t<BP_EMF, R, P>(B, bin, bout, wa, wb, mode);
    case BP_FACES: return batch_launch_t<BP_FACES, R, P>(B, bin, bout, wa, wb, mode);
    case BP_CELLS: return batch_launch_t<BP_CELLS, R, P>(B, bin, bout, wa, wb, mode);
    case BP_IMEX: return batch_launch_t<BP_IMEX, R, P>(B, bin, bout, wa, wb, mode);
    case BP_CFL: return batch_launch_t<BP_CFL, R, P>(B, bin, bout, wa, wb, mode);
  }
  set_error("bad batch phase %d", ph);
  return CS_E_INVALID;
}
int batch_phase(cs_batch* B, int ph, int bin, int bout, double wa, double wb, int mode) {
  return B->precision == CS_MIXED ? batch_phase_t<float, float>(B, ph, bin, bout, wa, wb, mode)
                                  : batch_phase_t<double, double>(B, ph, bin, bout, wa, wb, mode);
}

int batch_stage(cs_batch* B, int irk) {
  double wa, wb;
  int bi, bo, rc;
  stage_weights(B->rk_order, irk, wa, wb, bi, bo, B->imex);
  const int mode = B->imex ? irk : 0;
  if (B->imex && irk == 1) {
    for (cs_handle* h : B->sh) {
      const size_t fb = sizeof(double) * h->d.fpp * 6;
      for (int q = 2; q < 5; ++q)
        CS_HIP(hipMemcpyAsync(h->buf[(h->cur + 1) & 1][q], h->buf[h->cur][q], fb, hipMemcpyDeviceToDevice,
                              B->stream));
    }
    if ((rc = batch_phase(B, BP_IMEX, 0, 1, wa, wb, mode))) return rc;
  }
  // the updates of shells write the lateral ghosts they feed, so the exchange
  // phase runs only when the batch holds surface-mode shells
  static const int ph[7] = {BP_EXCH, BP_RAD, BP_FLUXXY, BP_FLUXZ, BP_EMF, BP_FACES, BP_CELLS};
  for (int q = B->nsurf ? 0 : 1; q < 7; ++q)
    if ((rc = batch_phase(B, ph[q], bi, bo, wa, wb, mode))) return rc;
  return CS_OK;
}

int batch_step(cs_batch* B) {
  int rc;
  for (int irk = 1; irk <= B->rk_order; ++irk)
    if ((rc = batch_stage(B, irk))) return rc;
  if ((rc = batch_phase(B, BP_CFL, 0, 0, 0.0, 1.0, 0))) return rc;
  k_batch_cfl_final<<<B->n, 1024, 0, B->stream>>>(B->dsh, B->dboff[BP_CFL]);
  CS_HIP(hipGetLastError());
  return CS_OK;
}

// the primitives of U^n of every shell and its lateral ghosts, before a run
// (its step leaves them: the updates write the ghosts they feed)
int batch_prepare(cs_batch* B) {
  for (cs_handle* h : B->sh) {
    if (h->vbuf == h->cur && h->vghost == h->cur) continue;
    int rc = ensure_prims(h, h->cur);
    if (!rc) rc = launch_halo(h, h->cur, true, false, false);
    if (rc) return rc;
    h->vghost = h->cur;
    CS_HIP(hipStreamSynchronize(h->stream));
  }
  return CS_OK;
}

int batch_run(cs_batch* B, int nsteps, int use_graph) {
  int rc;
  if ((rc = batch_prepare(B))) return rc;
  if (!use_graph) {
    for (int s = 0; s < nsteps; ++s)
      if ((rc = batch_step(B))) return rc;
  } else {
    if (!B->graph) {
      hipGraph_t gr;
      CS_HIP(hipStreamBeginCapture(B->stream, hipStreamCaptureModeThreadLocal));
      rc = batch_step(B);
      hipError_t e = hipStreamEndCapture(B->stream, &gr);
      if (rc) return rc;
      if (e != hipSuccess) { set_error("capture failed: %s", hipGetErrorString(e)); return CS_E_HIP; }
      CS_HIP(hipGraphInstantiate(&B->graph, gr, nullptr, nullptr, 0));
      hipGraphDestroy(gr);
    }
    for (int s = 0; s < nsteps; ++s) CS_HIP(hipGraphLaunch(B->graph, B->stream));
  }
  for (cs_handle* h : B->sh) {
    h->vbuf = h->cur;
    h->vghost = h->surf ? -1 : h->cur;
  }
  return CS_OK;
}

void batch_free(cs_batch* B) {
  if (B->graph) hipGraphExecDestroy(B->graph);
  for (int q = 0; q < BP_COUNT; ++q) hipFree(B->dboff[q]);
  hipFree(B->dsh);
  if (B->stream) hipStreamDestroy(B->stream);
  for (cs_handle* h : B->sh) cs_destroy(h);
}

}  // namespace

extern "C" int cs_batch_create(const cs_config* cfgs, int32_t n, cs_batch** out) {
  if (!out || !cfgs || n < 1) { set_error("bad arguments"); return CS_E_INVALID; }
  *out = nullptr;
  for (int b = 0; b < n; ++b)
    if (cfgs[b].recon == CS_RECON_P1ADP) {
      set_error("shell %d: batches run the WENO5/PCM stage path (CS_RECON_P1ADP shells step on their own handles)", b);
      return CS_E_INVALID;
    }
  for (int b = 1; b < n; ++b)
    if (cfgs[b].precision != cfgs[0].precision || cfgs[b].rk_order != cfgs[0].rk_order ||
        cfgs[b].imex != cfgs[0].imex) {
      set_error("shell %d: the shells of a batch share precision, rk_order and imex", b);
      return CS_E_INVALID;
    }
  cs_batch* B = new cs_batch();
  B->n = n;
  B->precision = cfgs[0].precision;
  B->rk_order = cfgs[0].rk_order;
  B->imex = cfgs[0].imex;
  int rc = CS_OK;
  for (int b = 0; b < n && rc == CS_OK; ++b) {
    cs_handle* h = nullptr;
    rc = cs_create(&cfgs[b], &h);
    if (rc == CS_OK) B->sh.push_back(h);
  }
  std::vector<ShellDev> desc;
  for (cs_handle* h : B->sh) {
    desc.push_back(shell_desc(h));
    B->nsurf += h->surf;
  }
  if (rc == CS_OK && (hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking) != hipSuccess ||
                      hipMalloc((void**)&B->dsh, sizeof(ShellDev) * n) != hipSuccess ||
                      hipMemcpy(B->dsh, desc.data(), sizeof(ShellDev) * n, hipMemcpyHostToDevice) != hipSuccess)) {
    set_error("batch set-up failed");
    rc = CS_E_HIP;
  }
  for (int q = 0; q < BP_COUNT && rc == CS_OK; ++q) {
    if (hipMalloc((void**)&B->dboff[q], sizeof(uint32_t) * (n + 1)) != hipSuccess) {
      set_error("batch set-up failed");
      rc = CS_E_NOMEM;
      break;
    }
    rc = batch_offsets(B, q, kBlock);
  }
  if (rc) {
    batch_free(B);
    delete B;
    return rc;
  }
  *out = B;
  return CS_OK;
}

extern "C" int cs_batch_destroy(cs_batch* B) {
  if (!B) return CS_OK;
  if (B->stream) hipStreamSynchronize(B->stream);
  batch_free(B);
  delete B;
  return CS_OK;
}

extern "C" int cs_batch_shell(cs_batch* B, int32_t i, cs_handle** h) {
  if (!B || !h || i < 0 || i >= B->n) { set_error("bad arguments"); return CS_E_INVALID; }
  *h = B->sh[i];
  return CS_OK;
}

extern "C" int cs_batch_evaluate_timestep(cs_batch* B) {
  if (!B) { set_error("null batch"); return CS_E_INVALID; }
  for (cs_handle* h : B->sh) {
    int rc = cs_evaluate_timestep(h);
    if (rc) return rc;
  }
  return CS_OK;
}

extern "C" int cs_batch_advance(cs_batch* B, int32_t nsteps, int32_t use_graph) {
  if (!B || nsteps < 0) { set_error("bad arguments"); return CS_E_INVALID; }
  int rc = batch_run(B, nsteps, use_graph);
  if (rc) return rc;
  CS_HIP(hipStreamSynchronize(B->stream));
  for (cs_handle* h : B->sh)
    if ((rc = check_flags(h))) return rc;
  return CS_OK;
}

extern "C" int cs_batch_time_steps(cs_batch* B, int32_t nsteps, float* ms_total) {
  if (!B || !ms_total || nsteps < 1) { set_error("bad arguments"); return CS_E_INVALID; }
  int rc = batch_run(B, 0, 1);  // graph built outside the timed region
  if (rc) return rc;
  hipEvent_t e0, e1;
  CS_HIP(hipEventCreate(&e0));
  CS_HIP(hipEventCreate(&e1));
  CS_HIP(hipStreamSynchronize(B->stream));
  CS_HIP(hipEventRecord(e0, B->stream));
  if ((rc = batch_run(B, nsteps, 1))) return rc;
  CS_HIP(hipEventRecord(e1, B->stream));
  CS_HIP(hipEventSynchronize(e1));
  CS_HIP(hipEventElapsedTime(ms_total, e0, e1));
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  for (cs_handle* h : B->sh)
    if ((rc = check_flags(h))) return rc;
  return CS_OK;
}

extern "C" int cs_batch_autotune(cs_batch* B, int32_t iters, int32_t* block_sizes) {
  if (!B || iters < 1) { set_error("bad arguments"); return CS_E_INVALID; }
  static const int cand[3] = {64, 128, 256};
  int rc;
  if ((rc = batch_prepare(B))) return rc;
  // a stage that writes only the scratch buffer: RK3 stage 2, RK2 / IMEX stage 1
  double wa, wb;
  int bi, bo;
  const int irk = B->rk_order == 3 ? 2 : 1;
  stage_weights(B->rk_order, irk, wa, wb, bi, bo, B->imex);
  const int mode = B->imex ? irk : 0;
  hipEvent_t e0, e1;
  CS_HIP(hipEventCreate(&e0));
  CS_HIP(hipEventCreate(&e1));
  for (int ph = 0; ph < BP_COUNT; ++ph) {
    if (ph == BP_CFL) continue;  // its partial buffers are sized for 256-thread workgroups
    int best = kBlock;
    float tbest = 1e30f;
    for (int c = 0; c < 3; ++c) {
      if ((rc = batch_offsets(B, ph, cand[c]))) return rc;
      if ((rc = batch_phase(B, ph, ph == BP_IMEX ? 0 : bi, ph == BP_IMEX ? 1 : bo, wa, wb, mode))) return rc;
      CS_HIP(hipEventRecord(e0, B->stream));
      for (int it = 0; it < iters; ++it)
        if ((rc = batch_phase(B, ph, ph == BP_IMEX ? 0 : bi, ph == BP_IMEX ? 1 : bo, wa, wb, mode))) return rc;
      CS_HIP(hipEventRecord(e1, B->stream));
      CS_HIP(hipEventSynchronize(e1));
      float ms = 0.f;
      CS_HIP(hipEventElapsedTime(&ms, e0, e1));
      if (ms < tbest) { tbest = ms; best = cand[c]; }
    }
    if ((rc = batch_offsets(B, ph, best))) return rc;
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (B->graph) { hipGraphExecDestroy(B->graph); B->graph = nullptr; }
  for (cs_handle* h : B->sh) h->vbuf = h->vghost = -1;  // the timed zone updates rewrote V
  if (block_sizes)
    for (int ph = 0; ph < BP_COUNT; ++ph) block_sizes[ph] = B->bs[ph];
  return CS_OK;
}
