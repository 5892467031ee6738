// cs_mesh.cpp -- host-side geodesic mesh, metric tables and initial states.
//
// The unit-sphere lattice restates SWIVEL_N / Mesh_Sector_General
// (pkg/inc/cubesphere/200mesh/201mesh_sector_general.f:5,79) for one patch per
// sector and FILL_GHOST2D_GLOBAL (203fill_ghost2d_global.f:7) operation for
// operation, so the vertices are bit-identical to the reference's.  The metric
// is this library's own: exact spherical geometry (planar annular-sector
// x/y faces, spherical z faces, cone volumes), stored separably as 2-D panel
// tables times radial factors because vert = rb(k) * u(i,j)
// (210extract_vert_coords_patch.f:300-308).
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <thread>
#include <vector>

#include "../../include/cubesphere_hip.h"
#include "cs_internal.h"

namespace cs {

void set_error(const char* fmt, ...);

// ------------------------------------------------------------------ lattice
// SWIVEL_N (201mesh_sector_general.f:5-59): subdivide the arc a1 -> a2 into n
// equal angles.  arr has entries for i in [lo, hi] (arr[(i-lo)*3 + c]); the
// end points 0 and n are copied exactly, indices outside [0, n] continue the
// same great circle (used only for the corner ghost blocks).
static void swivel(const double a1[3], const double a2[3], int n, int lo, int hi,
                   double* arr) {
  double dot = a1[0] * a2[0] + a1[1] * a2[1] + a1[2] * a2[2];
  double theta = acos(dot);
  double d_theta = theta / (double)n;
  double x[3], y[3];
  for (int c = 0; c < 3; ++c) x[c] = a1[c];
  for (int c = 0; c < 3; ++c) y[c] = a2[c] - dot * x[c];
  double t = sqrt(y[0] * y[0] + y[1] * y[1] + y[2] * y[2]);
  for (int c = 0; c < 3; ++c) y[c] = y[c] / t;
  for (int i = lo; i <= hi; ++i) {
    double* o = arr + (size_t)(i - lo) * 3;
    if (i == 0) {
      for (int c = 0; c < 3; ++c) o[c] = a1[c];
    } else if (i == n) {
      for (int c = 0; c < 3; ++c) o[c] = a2[c];
    } else {
      double tt = d_theta * (double)i;
      double ct = cos(tt), st = sin(tt);
      for (int c = 0; c < 3; ++c) o[c] = ct * x[c] + st * y[c];
    }
  }
}

// Sector corner points v_endpts (201mesh_sector_general.f:175-330).
static void sector_endpoints(int isector, double v[4][3]) {
  const double s = 1.0 / sqrt(3.0);
  static const int sg[6][4][3] = {
      {{-1, -1, 1}, {1, -1, 1}, {-1, 1, 1}, {1, 1, 1}},
      {{-1, 1, -1}, {-1, 1, 1}, {1, 1, -1}, {1, 1, 1}},
      {{1, -1, -1}, {1, 1, -1}, {1, -1, 1}, {1, 1, 1}},
      {{1, 1, -1}, {1, -1, -1}, {-1, 1, -1}, {-1, -1, -1}},
      {{1, -1, 1}, {-1, -1, 1}, {1, -1, -1}, {-1, -1, -1}},
      {{-1, 1, 1}, {-1, 1, -1}, {-1, -1, 1}, {-1, -1, -1}}};
  for (int a = 0; a < 4; ++a)
    for (int c = 0; c < 3; ++c) v[a][c] = sg[isector - 1][a][c] > 0 ? s : -s;
}

// Great-circle intersection of Mesh_Sector_General's inner loop
// (201mesh_sector_general.f:735-815): plane normals A (x lattice line) and
// B (y lattice line), C = B x A, sign fixed per sector, normalised.
static void lattice_point(int isector, const double* p1, const double* p2,
                          const double* p3, const double* p4, double out[3]) {
  double Ax = p1[1] * p2[2] - p1[2] * p2[1];
  double Ay = p1[2] * p2[0] - p1[0] * p2[2];
  double Az = p1[0] * p2[1] - p1[1] * p2[0];
  double Bx = p3[1] * p4[2] - p3[2] * p4[1];
  double By = p3[2] * p4[0] - p3[0] * p4[2];
  double Bz = p3[0] * p4[1] - p3[1] * p4[0];
  double Cx = By * Az - Bz * Ay;
  double Cy = Bz * Ax - Bx * Az;
  double Cz = Bx * Ay - By * Ax;
  bool flip = (isector == 1 && Cz < 0.0) || (isector == 4 && Cz > 0.0) ||
              (isector == 2 && Cy < 0.0) || (isector == 5 && Cy > 0.0) ||
              (isector == 3 && Cx < 0.0) || (isector == 6 && Cx > 0.0);
  if (flip) { Cx = -Cx; Cy = -Cy; Cz = -Cz; }
  double t = 1.0 / sqrt(Cx * Cx + Cy * Cy + Cz * Cz);
  out[0] = t * Cx;
  out[1] = t * Cy;
  out[2] = t * Cz;
}

// Sector neighbours (203fill_ghost2d_global.f:37-75): right, left, up, down.
static const int kNbr[6][4] = {{3, 6, 2, 5}, {1, 4, 3, 6}, {2, 5, 1, 4},
                               {5, 2, 6, 3}, {6, 3, 4, 1}, {4, 1, 5, 2}};

// coords layout: (-g:N+g, -g:N+g, 3, 6), Fortran order.
static void build_lattice(int N, int g, double* coords) {
  const int nf = N + 2 * g + 1;
  auto at = [&](int i, int j, int c, int s) -> double& {
    return coords[(size_t)(i + g) + (size_t)nf * ((j + g) + (size_t)nf * (c + 3 * (size_t)(s - 1)))];
  };
  std::vector<double> a1((size_t)(N + 2 * g + 1) * 3), a2(a1.size()), a3(a1.size()), a4(a1.size());
  for (int s = 1; s <= 6; ++s) {
    double v[4][3];
    sector_endpoints(s, v);
    // STEP 3 swivels over the extended index range [-g, N+g]
    swivel(v[0], v[1], N, -g, N + g, a1.data());
    swivel(v[2], v[3], N, -g, N + g, a2.data());
    swivel(v[0], v[2], N, -g, N + g, a3.data());
    swivel(v[1], v[3], N, -g, N + g, a4.data());
    for (int j = -g; j <= N + g; ++j)
      for (int i = -g; i <= N + g; ++i) {
        double o[3];
        lattice_point(s, &a1[(size_t)(i + g) * 3], &a2[(size_t)(i + g) * 3],
                      &a3[(size_t)(j + g) * 3], &a4[(size_t)(j + g) * 3], o);
        for (int c = 0; c < 3; ++c) at(i, j, c, s) = o[c];
      }
  }
  // FILL_GHOST2D_GLOBAL: ghost strips copied from the neighbouring sectors.
  // The corner blocks keep the extended lattice from above.
  std::vector<double> src(coords, coords + (size_t)nf * nf * 3 * 6);
  auto sat = [&](int i, int j, int c, int s) -> double {
    return src[(size_t)(i + g) + (size_t)nf * ((j + g) + (size_t)nf * (c + 3 * (size_t)(s - 1)))];
  };
  for (int s = 1; s <= 6; ++s) {
    int isrt = kNbr[s - 1][0], islf = kNbr[s - 1][1], isup = kNbr[s - 1][2], isdn = kNbr[s - 1][3];
    for (int ixx = N + 1; ixx <= N + g; ++ixx) {
      int iyy_you = N - 1 - (ixx - N - 1);
      for (int iyy = 0; iyy <= N; ++iyy)
        for (int c = 0; c < 3; ++c) at(ixx, iyy, c, s) = sat(iyy, iyy_you, c, isrt);
    }
    for (int ixx = -1; ixx >= -g; --ixx) {
      int ixx_you = 1 - (ixx + 1);
      for (int iyy = 0; iyy <= N; ++iyy)
        for (int c = 0; c < 3; ++c) at(ixx, iyy, c, s) = sat(ixx_you, N - iyy, c, islf);
    }
    for (int iyy = N + 1; iyy <= N + g; ++iyy) {
      int ixx_you = N - 1 - (iyy - N - 1);
      for (int ixx = 0; ixx <= N; ++ixx)
        for (int c = 0; c < 3; ++c) at(ixx, iyy, c, s) = sat(ixx_you, ixx, c, isup);
    }
    for (int iyy = -1; iyy >= -g; --iyy) {
      int iyy_you = 1 - (iyy + 1);
      for (int ixx = 0; ixx <= N; ++ixx)
        for (int c = 0; c < 3; ++c) at(ixx, iyy, c, s) = sat(N - ixx, iyy_you, c, isdn);
    }
  }
}

static void build_radii(const cs_config* c, double* rb) {
  const int g = c->nghost, NR = c->nr;
  for (int i = -g; i <= NR + g; ++i) {
    double pw = (double)i / (double)NR;
    rb[i + g] = c->radial_mesh == CS_RMESH_LOG ? c->rmin * pow(c->rmax / c->rmin, pw)
                                               : c->rmin + (c->rmax - c->rmin) * pw;
  }
}

// ------------------------------------------------------------------ vectors
static inline void cross(const double* a, const double* b, double* o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}
static inline double dot(const double* a, const double* b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
static inline double norm(const double* a) { return sqrt(dot(a, a)); }
static inline void normalize(double* a) {
  double n = norm(a);
  if (n > 0) { a[0] /= n; a[1] /= n; a[2] /= n; }
}
// angle between unit vectors, accurate for small angles
static inline double angle(const double* a, const double* b) {
  double x[3];
  cross(a, b, x);
  return atan2(norm(x), dot(a, b));
}
// orthonormalise v against unit n
static inline void perp(double* v, const double* n) {
  double t = dot(v, n);
  for (int c = 0; c < 3; ++c) v[c] -= t * n[c];
  normalize(v);
}
static inline double tri_solid_angle(const double* a, const double* b, const double* c) {
  double bc[3];
  cross(b, c, bc);
  double num = dot(a, bc);
  double den = 1.0 + dot(a, b) + dot(b, c) + dot(c, a);
  return 2.0 * atan2(num, den);
}

// 2-D metric tables: T_COUNT planes of t2*6 doubles (plane stride t2*6,
// panel stride t2).
static void build_tables(const CsDims& d, const double* coords, double* T) {
  const int N = d.N, g = d.g, nf = d.nf;
  const int64_t plane = d.t2 * 6;
  memset(T, 0, sizeof(double) * plane * T_COUNT);
  for (int s = 0; s < 6; ++s) {
    auto U = [&](int i, int j, double* o) {
      for (int c = 0; c < 3; ++c)
        o[c] = coords[(size_t)(i + g) + (size_t)nf * ((j + g) + (size_t)nf * (c + 3 * (size_t)s))];
    };
    auto put = [&](int f, int i, int j, double v) { T[f * plane + s * d.t2 + d.t(i, j)] = v; };
    auto put3 = [&](int f, int i, int j, const double* v) {
      for (int c = 0; c < 3; ++c) put(f + c, i, j, v[c]);
    };
    auto get = [&](int f, int i, int j) { return T[f * plane + s * d.t2 + d.t(i, j)]; };
    for (int j = -g; j <= N + g; ++j)
      for (int i = -g; i <= N + g; ++i) {
        double u[3];
        U(i, j, u);
        put3(T_UX, i, j, u);
      }
    // x-faces (vertex i, cell j) and y-edges: need u(i,j-1), u(i,j)
    for (int j = 1 - g; j <= N + g; ++j)
      for (int i = -g; i <= N + g; ++i) {
        double a[3], b[3], n[3], cdir[3];
        U(i, j - 1, a);
        U(i, j, b);
        cross(b, a, n);
        normalize(n);
        for (int c = 0; c < 3; ++c) cdir[c] = a[c] + b[c];
        normalize(cdir);
        put3(T_XNX, i, j, n);
        put3(T_XCX, i, j, cdir);
        put(T_THY, i, j, angle(a, b));
        double t[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
        perp(t, cdir);
        put3(T_YETX, i, j, t);
        put3(T_YERX, i, j, cdir);
      }
    // y-faces (cell i, vertex j) and x-edges: need u(i-1,j), u(i,j)
    for (int j = -g; j <= N + g; ++j)
      for (int i = 1 - g; i <= N + g; ++i) {
        double a[3], b[3], n[3], cdir[3];
        U(i - 1, j, a);
        U(i, j, b);
        cross(a, b, n);
        normalize(n);
        for (int c = 0; c < 3; ++c) cdir[c] = a[c] + b[c];
        normalize(cdir);
        put3(T_YNX, i, j, n);
        put3(T_YCX, i, j, cdir);
        put(T_THX, i, j, angle(a, b));
        double t[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
        perp(t, cdir);
        put3(T_XETX, i, j, t);
        put3(T_XERX, i, j, cdir);
      }
    // cells (i,j): area vector, solid angle, chords, zone-B operator
    for (int j = 1 - g; j <= N + g; ++j)
      for (int i = 1 - g; i <= N + g; ++i) {
        double a[3], b[3], c3[3], e[3];
        U(i - 1, j - 1, a);
        U(i, j - 1, b);
        U(i, j, c3);
        U(i - 1, j, e);
        double zs[3];
        for (int c = 0; c < 3; ++c)
          zs[c] = 0.5 * (get(T_THX, i, j - 1) * get(T_YNX + c, i, j - 1) -
                         get(T_THY, i, j) * get(T_XNX + c, i, j) -
                         get(T_THX, i, j) * get(T_YNX + c, i, j) +
                         get(T_THY, i - 1, j) * get(T_XNX + c, i - 1, j));
        double zsn = norm(zs);
        put3(T_ZSX, i, j, zs);
        put(T_ZSN, i, j, zsn);
        double sh[3] = {zs[0] / zsn, zs[1] / zsn, zs[2] / zsn};
        double t1[3];
        for (int c = 0; c < 3; ++c) t1[c] = (b[c] + c3[c]) - (a[c] + e[c]);
        perp(t1, sh);
        put3(T_ZTX, i, j, t1);
        put(T_OM, i, j, tri_solid_angle(a, b, c3) + tri_solid_angle(a, c3, e));
        double zc[3];
        for (int c = 0; c < 3; ++c) zc[c] = a[c] + b[c] + c3[c] + e[c];
        normalize(zc);
        put3(T_ZCX, i, j, zc);
        double ch = 1e300, dd[3];
        const double* pairs[4][2] = {{e, c3}, {a, b}, {b, c3}, {a, e}};
        for (auto& p : pairs) {
          for (int c = 0; c < 3; ++c) dd[c] = p[0][c] - p[1][c];
          double l = norm(dd);
          ch = l < ch ? l : ch;
        }
        put(T_CHMIN, i, j, ch);
        // surface mode: the inverse of the face-family mean normals, so each
        // family's face average is reproduced along its mean normal
        double M[3][3];
        for (int c = 0; c < 3; ++c) {
          M[0][c] = 0.5 * (get(T_XNX + c, i - 1, j) + get(T_XNX + c, i, j));
          M[1][c] = 0.5 * (get(T_YNX + c, i, j - 1) + get(T_YNX + c, i, j));
          M[2][c] = sh[c];
        }
        double det = M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) -
                     M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0]) +
                     M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
        double Mi[3][3];
        if (fabs(det) > 1e-300) {
          double id = 1.0 / det;
          Mi[0][0] = (M[1][1] * M[2][2] - M[1][2] * M[2][1]) * id;
          Mi[0][1] = (M[0][2] * M[2][1] - M[0][1] * M[2][2]) * id;
          Mi[0][2] = (M[0][1] * M[1][2] - M[0][2] * M[1][1]) * id;
          Mi[1][0] = (M[1][2] * M[2][0] - M[1][0] * M[2][2]) * id;
          Mi[1][1] = (M[0][0] * M[2][2] - M[0][2] * M[2][0]) * id;
          Mi[1][2] = (M[0][2] * M[1][0] - M[0][0] * M[1][2]) * id;
          Mi[2][0] = (M[1][0] * M[2][1] - M[1][1] * M[2][0]) * id;
          Mi[2][1] = (M[0][1] * M[2][0] - M[0][0] * M[2][1]) * id;
          Mi[2][2] = (M[0][0] * M[1][1] - M[0][1] * M[1][0]) * id;
        } else {
          memset(Mi, 0, sizeof(Mi));
        }
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 3; ++c) put(T_M00 + 3 * r + c, i, j, Mi[r][c]);
        // zone-B least squares of CLEANUP_AFTER_UPDATE (485cleanup_after_update.f:740-800):
        // rows = unit normals of the x+, x-, y+, y- faces and the radial unit
        // vector (for both z faces); A33 = A^T A, inverted by cofactors with
        // the reference's guard sign(det) / max(1e-32, |det|)
        double rows[5][3];
        for (int c = 0; c < 3; ++c) {
          rows[0][c] = get(T_XNX + c, i, j);
          rows[1][c] = get(T_XNX + c, i - 1, j);
          rows[2][c] = get(T_YNX + c, i, j);
          rows[3][c] = get(T_YNX + c, i, j - 1);
          rows[4][c] = zsn > 0.0 ? sh[c] : zc[c];
        }
        double m33[3][3];
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 3; ++c)
            m33[r][c] = rows[0][r] * rows[0][c] + rows[1][r] * rows[1][c] + rows[2][r] * rows[2][c] +
                      rows[3][r] * rows[3][c] + rows[4][r] * rows[4][c] + rows[4][r] * rows[4][c];
        double d33 = m33[0][0] * (m33[1][1] * m33[2][2] - m33[2][1] * m33[1][2]) -
                     m33[0][1] * (m33[1][0] * m33[2][2] - m33[2][0] * m33[1][2]) +
                     m33[0][2] * (m33[1][0] * m33[2][1] - m33[2][0] * m33[1][1]);
        double od = (d33 < 0.0 ? -1.0 : 1.0) / std::max(1e-32, fabs(d33));
        double bi[3][3] = {{(m33[1][1] * m33[2][2] - m33[1][2] * m33[2][1]) * od, (m33[0][2] * m33[2][1] - m33[0][1] * m33[2][2]) * od,
                            (m33[0][1] * m33[1][2] - m33[1][1] * m33[0][2]) * od},
                           {(m33[1][2] * m33[2][0] - m33[1][0] * m33[2][2]) * od, (m33[0][0] * m33[2][2] - m33[0][2] * m33[2][0]) * od,
                            (m33[0][2] * m33[1][0] - m33[0][0] * m33[1][2]) * od},
                           {(m33[1][0] * m33[2][1] - m33[1][1] * m33[2][0]) * od, (m33[0][1] * m33[2][0] - m33[0][0] * m33[2][1]) * od,
                            (m33[0][0] * m33[1][1] - m33[0][1] * m33[1][0]) * od}};
        const int slot[5] = {T_BWXP, T_BWXM, T_BWYP, T_BWYM, T_BWZ};
        for (int f = 0; f < 5; ++f)
          for (int r = 0; r < 3; ++r)
            put(slot[f] + r, i, j, bi[r][0] * rows[f][0] + bi[r][1] * rows[f][1] + bi[r][2] * rows[f][2]);
      }
    // z-edges at vertex (i,j): frame of 460make_edge_fields_p1.f:2700-2760
    for (int j = 1 - g; j <= N + g - 1; ++j)
      for (int i = 1 - g; i <= N + g - 1; ++i) {
        double n3[3], n1[3], tau[3], p[3], m[3];
        U(i, j, n3);
        U(i + 1, j, p);
        U(i - 1, j, m);
        for (int c = 0; c < 3; ++c) n1[c] = p[c] - m[c];
        perp(n1, n3);
        U(i, j + 1, p);
        U(i, j - 1, m);
        for (int c = 0; c < 3; ++c) tau[c] = p[c] - m[c];
        perp(tau, n3);
        double ct = dot(tau, n1);
        put3(T_EN1X, i, j, n1);
        put3(T_ETAX, i, j, tau);
        put(T_ECOS, i, j, ct);
        put(T_ESIN, i, j, sqrt(1.0 - ct * ct));
      }
  }
}

}  // namespace cs

using namespace cs;

static int check_cfg(const cs_config* c) {
  if (!c) { set_error("null config"); return CS_E_INVALID; }
  if (c->ntp < 4 || c->nr < 1 || c->nghost < 3 || c->nghost > 6) {
    set_error("bad grid ntp=%d nr=%d nghost=%d (need ntp>=4, nr>=1, 3<=nghost<=6)", c->ntp, c->nr, c->nghost);
    return CS_E_INVALID;
  }
  if (c->nghost >= c->ntp) {
    set_error("NGhost (%d) must be smaller than ntp (%d)", c->nghost, c->ntp);
    return CS_E_BAD_DECOMP;
  }
  if (!(c->rmin > 0.0) || !(c->rmax > c->rmin) || !(c->gamma > 1.0) || !(c->cfl > 0.0)) {
    set_error("bad physical parameters");
    return CS_E_INVALID;
  }
  if (c->rk_order != 2 && c->rk_order != 3) { set_error("rk_order must be 2 or 3"); return CS_E_INVALID; }
  if (c->imex != 0 && (c->imex != 1 || c->rk_order != 2)) {
    set_error("implicit gravity (imex = 1) is the IMEX-SSP2(2,2,2) scheme: it needs rk_order = 2");
    return CS_E_INVALID;
  }
  if (c->recon < 0 || c->recon > 2 || c->riemann < 0 || c->riemann > 1 || c->emf < 0 || c->emf > 1 ||
      c->precision < 0 || c->precision > 1 || c->radial_bc < 0 || c->radial_bc > 2 ||
      c->radial_mesh < 0 || c->radial_mesh > 1) {
    set_error("bad scheme selector");
    return CS_E_INVALID;
  }
  if (c->radial_bc == CS_RBC_PROBLEM) {
    const int p = c->bc_problem;
    if (c->recon != CS_RECON_P1ADP || c->nr < 2 || !(p == 0 || p == 2 || p == 102 || p == 103)) {
      set_error("CS_RBC_PROBLEM: a CS_RECON_P1ADP shell (nr > 1) and bc_problem 0, 2, 102 or 103");
      return CS_E_INVALID;
    }
  }
  return CS_OK;
}

namespace cs {
int validate_config(const cs_config* c) { return check_cfg(c); }

// Host metric for a handle: radii (NR+2g+1) and 2-D tables.
int host_metric(const cs_config* c, std::vector<double>& rb, std::vector<double>& T) {
  int rc = check_cfg(c);
  if (rc) return rc;
  CsDims d = cs_make_dims(c->ntp, c->nr, c->nghost);
  std::vector<double> coords((size_t)d.t2 * 3 * 6);
  build_lattice(d.N, d.g, coords.data());
  rb.assign((size_t)(d.NR + 2 * d.g + 1), 0.0);
  build_radii(c, rb.data());
  T.assign((size_t)d.t2 * 6 * T_COUNT, 0.0);
  build_tables(d, coords.data(), T.data());
  return CS_OK;
}
}  // namespace cs

extern "C" int cs_sizes_of(const cs_config* cfg, cs_sizes* out) {
  int rc = check_cfg(cfg);
  if (rc) return rc;
  if (!out) { set_error("null output"); return CS_E_INVALID; }
  CsDims d = cs_make_dims(cfg->ntp, cfg->nr, cfg->nghost);
  out->n_cell = d.cpp * 6;
  out->n_face = d.fpp * 6;
  out->n_vert2d = d.t2 * 6;
  out->n_interior = (int64_t)6 * d.N * d.N * d.NR;
  return CS_OK;
}

extern "C" int cs_build_unit_lattice(const cs_config* cfg, double* coords) {
  int rc = check_cfg(cfg);
  if (rc) return rc;
  if (!coords) { set_error("null output"); return CS_E_INVALID; }
  build_lattice(cfg->ntp, cfg->nghost, coords);
  return CS_OK;
}

// STENCIL_INDEXING_TRICORNERS (211stencil_indexing_tricorners.f:6) for a
// sector that is one patch, so all four corners are tri-corners.  Each zone's
// twelve stencil neighbours (1 R, 2 L, 3 U, 4 D, 5 RR, 6 LL, 7 UU, 8 DD,
// 9 RU, 10 LU, 11 LD, 12 RD) are first the index neighbours clamped to the
// patch; around each tri-corner ghost block they are re-pointed to the zones
// that are the neighbours on the sphere.  The four corners' remaps are one
// pattern under the reflections of the patch: written once for the left-down
// corner and mirrored (x -> N+1-x swaps R/L, RR/LL, RU/LU, LD/RD; y -> N+1-y
// swaps U/D, UU/DD, RU/RD, LU/LD), corners applied in the reference's order
// LD, RD, RU, LU (:84, 127, 173, 219).
extern "C" int cs_stencil_indexing(const cs_config* cfg, int32_t* st) {
  int rc = check_cfg(cfg);
  if (rc) return rc;
  if (!st) { set_error("null output"); return CS_E_INVALID; }
  const int N = cfg->ntp, g = cfg->nghost, nc = N + 2 * g, lo = 1 - g, hi = N + g;
  auto set = [&](int x, int y, int dir, int vx, int vy) {
    const size_t o = (size_t)(x - lo) + (size_t)nc * ((size_t)(y - lo) + (size_t)nc * 2 * (size_t)(dir - 1));
    st[o] = vx;
    st[o + (size_t)nc * nc] = vy;
  };
  static const int dx[13] = {0, 1, -1, 0, 0, 2, -2, 0, 0, 1, -1, -1, 1};
  static const int dy[13] = {0, 0, 0, 1, -1, 0, 0, 2, -2, 1, 1, -1, -1};
  for (int y = lo; y <= hi; ++y)
    for (int x = lo; x <= hi; ++x)
      for (int dir = 1; dir <= 12; ++dir)
        set(x, y, dir, std::min(std::max(x + dx[dir], lo), hi), std::min(std::max(y + dy[dir], lo), hi));
  static const int mirror_x[13] = {0, 2, 1, 3, 4, 6, 5, 7, 8, 10, 9, 12, 11};
  static const int mirror_y[13] = {0, 1, 2, 4, 3, 5, 6, 8, 7, 12, 11, 10, 9};
  static const int corner[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};  // LD, RD, RU, LU
  for (const auto& c : corner) {
    auto fx = [&](int x) { return c[0] ? N + 1 - x : x; };
    auto fy = [&](int y) { return c[1] ? N + 1 - y : y; };
    auto put = [&](int x, int y, int dir, int vx, int vy) {
      if (c[0]) dir = mirror_x[dir];
      if (c[1]) dir = mirror_y[dir];
      set(fx(x), fy(y), dir, fx(vx), fy(vy));
    };
    // left-down pattern: the ghost row below zone row 1 left of the corner
    // wraps onto the zones left of column 1, and vice versa
    for (int x = lo; x <= 0; ++x) {
      put(x, 1, 4, 1, x);
      put(x, 1, 8, 2, x);
      put(x, 1, 11, 1, std::max(x - 1, lo));
      if (x == 0) put(x, 1, 12, 2, x - 1);
      else put(x, 1, 12, 1, x + 1);
      put(x, 2, 8, 1, x);
    }
    for (int y = lo; y <= 0; ++y) {
      put(1, y, 2, y, 1);
      put(1, y, 6, y, 2);
      if (y == 0) put(1, y, 10, y - 1, 2);
      else put(1, y, 10, y + 1, 1);
      put(1, y, 11, std::max(y - 1, lo), 1);
      put(2, y, 6, y, 1);
    }
    put(1, 1, 11, 2, -1);
  }
  return CS_OK;
}

extern "C" int cs_build_radial_mesh(const cs_config* cfg, double* rb) {
  int rc = check_cfg(cfg);
  if (rc) return rc;
  if (!rb) { set_error("null output"); return CS_E_INVALID; }
  build_radii(cfg, rb);
  return CS_OK;
}

extern "C" int cs_build_metric(const cs_config* cfg, double* vol, double* xarea, double* yarea,
                               double* zarea) {
  std::vector<double> rb, T;
  int rc = host_metric(cfg, rb, T);
  if (rc) return rc;
  CsDims d = cs_make_dims(cfg->ntp, cfg->nr, cfg->nghost);
  const int g = d.g;
  const int64_t plane = d.t2 * 6;
  auto tab = [&](int f, int s, int i, int j) { return T[f * plane + s * d.t2 + d.t(i, j)]; };
  auto r = [&](int k) { return rb[(size_t)(k + g)]; };
  for (int s = 0; s < 6; ++s) {
    if (vol)
      for (int k = 1 - g; k <= d.NR + g; ++k)
        for (int j = 1 - g; j <= d.N + g; ++j)
          for (int i = 1 - g; i <= d.N + g; ++i)
            vol[s * d.cpp + d.cell(i, j, k)] =
                tab(T_OM, s, i, j) * (r(k) * r(k) * r(k) - r(k - 1) * r(k - 1) * r(k - 1)) / 3.0;
    for (int k = -g; k <= d.NR + g; ++k)
      for (int j = -g; j <= d.N + g; ++j)
        for (int i = -g; i <= d.N + g; ++i) {
          int64_t f = s * d.fpp + d.face(i, j, k);
          bool kc = k > -g;  // cell layer k exists
          if (xarea) xarea[f] = (kc && j > -g) ? 0.5 * (r(k) * r(k) - r(k - 1) * r(k - 1)) * tab(T_THY, s, i, j) : 0.0;
          if (yarea) yarea[f] = (kc && i > -g) ? 0.5 * (r(k) * r(k) - r(k - 1) * r(k - 1)) * tab(T_THX, s, i, j) : 0.0;
          if (zarea) zarea[f] = (i > -g && j > -g) ? r(k) * r(k) * tab(T_ZSN, s, i, j) : 0.0;
        }
  }
  return CS_OK;
}

// ------------------------------------------------------------------ problems
namespace {

inline uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
// uniform in [-1, 1) from (seed, id)
inline double urand(uint64_t seed, uint64_t id, uint64_t stream) {
  uint64_t h = splitmix64(splitmix64(seed ^ (stream * 0xD1B54A32D192ED03ull)) + id);
  return ((double)(h >> 11) * (1.0 / 9007199254740992.0)) * 2.0 - 1.0;
}
inline double grand(uint64_t seed, uint64_t id) {
  double u1 = 0.5 * (urand(seed, id, 11) + 1.0), u2 = 0.5 * (urand(seed, id, 12) + 1.0);
  if (u1 < 1e-300) u1 = 1e-300;
  return sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
}

struct Prim { double rho, vx, vy, vz, p; };

// vector potential A(x) of each problem
void vecpot(int problem, const double* x, double* A) {
  double r = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
  A[0] = A[1] = A[2] = 0.0;
  if (problem == 0) {  // A = (yz, 2zx, 3xy)  (801problem_generator_u_a.f:105-229)
    A[0] = x[1] * x[2];
    A[1] = 2.0 * x[2] * x[0];
    A[2] = 3.0 * x[0] * x[1];
  } else if (problem == 1) {  // A = 0.05 z rhat  ->  B_phi = 0.05 cos(lat)
    double ar = 0.05 * x[2];
    for (int c = 0; c < 3; ++c) A[c] = ar * x[c] / r;
  } else if (problem == 2) {  // A = -0.5 r theta rhat  ->  B = 0.5 e_phi
    double th = atan2(sqrt(x[0] * x[0] + x[1] * x[1]), x[2]);
    double ar = -0.5 * r * th;
    for (int c = 0; c < 3; ++c) A[c] = ar * x[c] / r;
  } else if (problem == 3) {  // smooth non-dipolar potential, radial + tangential B
    double s = 0.05 / (r * r);
    A[0] = s * x[1] * x[2];
    A[1] = s * (x[2] * x[0] + 0.5 * x[0] * x[1]);
    A[2] = s * x[0] * x[1];
  }
}

Prim prim_at(int problem, const cs_config* c, const double* x, uint64_t seed, uint64_t gid) {
  Prim P{1.0, 0.0, 0.0, 0.0, 1.0};
  double r = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
  double u[3] = {x[0] / r, x[1] / r, x[2] / r};
  if (problem == 0) {
    P.rho = 1.0 + fabs(x[0]);
    P.vx = x[0]; P.vy = x[1]; P.vz = x[2];
    P.p = 2.0 + fabs(x[2]);
  } else if (problem == 1) {
    double d = atan2(sqrt(u[1] * u[1] + u[2] * u[2]), u[0]);  // distance to (lon 0, lat 0)
    P.rho = 1.0 + 0.1 * exp(-d * d / 0.05) + 1.0e-3 * urand(seed, gid, 1);
    double vs = 0.1 * sqrt(c->gamma);  // 0.1 of the sound speed of rho = p = 1
    P.vx = -vs * u[1]; P.vy = vs * u[0]; P.vz = 0.0;
    P.p = 1.0;
  } else if (problem == 2) {
    double d = atan2(sqrt(u[1] * u[1] + u[2] * u[2]), u[0]);
    P.rho = 1.0;
    P.p = d < 0.1 ? 10.0 : 1.0;
  } else if (problem == 3) {
    // isothermal hydrostatic: dp/dr = -rho g0 / r^2 with p = rho
    double g0 = c->grav > 0 ? c->grav : 1.0;
    double rho0 = exp(g0 * (1.0 / r - 1.0 / c->rmin));
    P.rho = rho0 * (1.0 + 1.0e-2 * grand(seed, gid));
    P.p = rho0;
  }
  return P;
}

// 5-point Gauss-Legendre on [0,1]
const double kGx[5] = {0.04691007703066800, 0.23076534494715845, 0.5, 0.76923465505284155, 0.95308992296933200};
const double kGw[5] = {0.11846344252809454, 0.23931433524968324, 0.28444444444444444, 0.23931433524968324, 0.11846344252809454};

double radial_integral(int problem, const double* u, double r0, double r1) {
  double s = 0.0;
  for (int q = 0; q < 5; ++q) {
    double r = r0 + (r1 - r0) * kGx[q], x[3] = {r * u[0], r * u[1], r * u[2]}, A[3];
    vecpot(problem, x, A);
    s += kGw[q] * (A[0] * u[0] + A[1] * u[1] + A[2] * u[2]);
  }
  return s * (r1 - r0);
}
// integral of A.dl along the great-circle arc a -> b at radius r
double arc_integral(int problem, const double* a, const double* b, double r) {
  double dt = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
  double w[3] = {b[0] - dt * a[0], b[1] - dt * a[1], b[2] - dt * a[2]};
  double wn = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
  if (wn <= 0) return 0.0;
  for (int c = 0; c < 3; ++c) w[c] /= wn;
  double th = atan2(wn, dt), s = 0.0;
  for (int q = 0; q < 5; ++q) {
    double ph = th * kGx[q], cp = cos(ph), sp = sin(ph);
    double x[3], t[3], A[3];
    for (int c = 0; c < 3; ++c) {
      x[c] = r * (cp * a[c] + sp * w[c]);
      t[c] = r * (-sp * a[c] + cp * w[c]);
    }
    vecpot(problem, x, A);
    s += kGw[q] * (A[0] * t[0] + A[1] * t[1] + A[2] * t[2]);
  }
  return s * th;
}

}  // namespace

extern "C" int cs_init_problem(const cs_config* cfg, int32_t problem, uint64_t seed, double* Uo,
                               double* ento, double* Bxf, double* Byf, double* Bzf) {
  std::vector<double> rb, T;
  int rc = host_metric(cfg, rb, T);
  if (rc) return rc;
  if (problem < 0 || problem > 3) { set_error("unknown problem %d", problem); return CS_E_INVALID; }
  if (!Uo || !ento || !Bxf || !Byf || !Bzf) { set_error("null output"); return CS_E_INVALID; }
  CsDims d = cs_make_dims(cfg->ntp, cfg->nr, cfg->nghost);
  const int N = d.N, NR = d.NR, g = d.g;
  const int64_t plane = d.t2 * 6;
  const bool surface = cfg->radial_bc == CS_RBC_WALL && cfg->nr == 1;
  memset(Uo, 0, sizeof(double) * d.cpp * 6 * 8);
  memset(ento, 0, sizeof(double) * d.cpp * 6);
  memset(Bxf, 0, sizeof(double) * d.fpp * 6);
  memset(Byf, 0, sizeof(double) * d.fpp * 6);
  memset(Bzf, 0, sizeof(double) * d.fpp * 6);
  const double gam = cfg->gamma, ob8pi = 1.0 / (8.0 * (4.0 * atan(1.0)));
  auto r = [&](int k) { return rb[(size_t)(k + g)]; };
  // the six sectors are independent (each writes only its own arrays): one
  // host thread per sector
  auto sector = [&](int s) {
    auto tab = [&](int f, int i, int j) { return T[f * plane + s * d.t2 + d.t(i, j)]; };
    auto uvec = [&](int i, int j, double* o) {
      for (int c = 0; c < 3; ++c) o[c] = tab(T_UX + c, i, j);
    };
    // edge integrals of A, stored in face-shaped scratch
    std::vector<double> Iz(d.fpp, 0.0), Ix(d.fpp, 0.0), Iy(d.fpp, 0.0);
    for (int k = 1; k <= NR; ++k)
      for (int j = 0; j <= N; ++j)
        for (int i = 0; i <= N; ++i) {
          double u[3];
          uvec(i, j, u);
          Iz[d.face(i, j, k)] = radial_integral(problem, u, r(k - 1), r(k));
        }
    for (int k = 0; k <= NR; ++k)
      for (int j = 0; j <= N; ++j)
        for (int i = 1; i <= N; ++i) {
          double a[3], b[3];
          uvec(i - 1, j, a);
          uvec(i, j, b);
          Ix[d.face(i, j, k)] = arc_integral(problem, a, b, r(k));
        }
    for (int k = 0; k <= NR; ++k)
      for (int j = 1; j <= N; ++j)
        for (int i = 0; i <= N; ++i) {
          double a[3], b[3];
          uvec(i, j - 1, a);
          uvec(i, j, b);
          Iy[d.face(i, j, k)] = arc_integral(problem, a, b, r(k));
        }
    // face fluxes by Stokes with the circulation of UPDATE_IMEX_RK_STEP_P1
    // (484update_imex_rk_step_p1.f:672-676, 748-752, 898-902)
    for (int k = 1; k <= NR; ++k)
      for (int j = 1; j <= N; ++j)
        for (int i = 0; i <= N; ++i) {
          double circ = Iz[d.face(i, j, k)] - Iz[d.face(i, j - 1, k)] - Iy[d.face(i, j, k)] + Iy[d.face(i, j, k - 1)];
          double A = 0.5 * (r(k) * r(k) - r(k - 1) * r(k - 1)) * tab(T_THY, i, j);
          Bxf[s * d.fpp + d.face(i, j, k)] = circ / A;
        }
    for (int k = 1; k <= NR; ++k)
      for (int j = 0; j <= N; ++j)
        for (int i = 1; i <= N; ++i) {
          double circ = -Iz[d.face(i, j, k)] + Iz[d.face(i - 1, j, k)] + Ix[d.face(i, j, k)] - Ix[d.face(i, j, k - 1)];
          double A = 0.5 * (r(k) * r(k) - r(k - 1) * r(k - 1)) * tab(T_THX, i, j);
          Byf[s * d.fpp + d.face(i, j, k)] = circ / A;
        }
    for (int k = 0; k <= NR; ++k)
      for (int j = 1; j <= N; ++j)
        for (int i = 1; i <= N; ++i) {
          double circ = Iy[d.face(i, j, k)] - Iy[d.face(i - 1, j, k)] - Ix[d.face(i, j, k)] + Ix[d.face(i, j - 1, k)];
          double A = r(k) * r(k) * tab(T_ZSN, i, j);
          Bzf[s * d.fpp + d.face(i, j, k)] = circ / A;
        }
    // zone states
    for (int k = 1; k <= NR; ++k)
      for (int j = 1; j <= N; ++j)
        for (int i = 1; i <= N; ++i) {
          double uc[3] = {0, 0, 0}, t[3];
          uvec(i - 1, j - 1, t); for (int c = 0; c < 3; ++c) uc[c] += t[c];
          uvec(i, j - 1, t);     for (int c = 0; c < 3; ++c) uc[c] += t[c];
          uvec(i, j, t);         for (int c = 0; c < 3; ++c) uc[c] += t[c];
          uvec(i - 1, j, t);     for (int c = 0; c < 3; ++c) uc[c] += t[c];
          double un = sqrt(uc[0] * uc[0] + uc[1] * uc[1] + uc[2] * uc[2]);
          double rc2 = 0.5 * (r(k - 1) + r(k));
          double x[3] = {rc2 * uc[0] / un, rc2 * uc[1] / un, rc2 * uc[2] / un};
          uint64_t gid = (uint64_t)s * N * N * NR + (uint64_t)(k - 1) * N * N + (uint64_t)(j - 1) * N + (uint64_t)(i - 1);
          Prim P = prim_at(problem, cfg, x, seed, gid);
          // zone B from the faces, with the operator of the stage cleanup
          const double bxp = Bxf[s * d.fpp + d.face(i, j, k)], bxm = Bxf[s * d.fpp + d.face(i - 1, j, k)];
          const double byp = Byf[s * d.fpp + d.face(i, j, k)], bym = Byf[s * d.fpp + d.face(i, j - 1, k)];
          const double bzp = Bzf[s * d.fpp + d.face(i, j, k)], bzm = Bzf[s * d.fpp + d.face(i, j, k - 1)];
          double B[3];
          if (surface) {
            const double b0 = 0.5 * (bxm + bxp), b1 = 0.5 * (bym + byp), b2 = 0.5 * (bzm + bzp);
            for (int rr = 0; rr < 3; ++rr)
              B[rr] = tab(T_M00 + 3 * rr, i, j) * b0 + tab(T_M00 + 3 * rr + 1, i, j) * b1 +
                      tab(T_M00 + 3 * rr + 2, i, j) * b2;
          } else {
            for (int rr = 0; rr < 3; ++rr)
              B[rr] = tab(T_BWXP + rr, i, j) * bxp + tab(T_BWXM + rr, i, j) * bxm + tab(T_BWYP + rr, i, j) * byp +
                      tab(T_BWYM + rr, i, j) * bym + tab(T_BWZ + rr, i, j) * (bzp + bzm);
          }
          int64_t cidx = s * d.cpp * 8 + d.cell(i, j, k);
          Uo[cidx + 0 * d.cpp] = P.rho;
          Uo[cidx + 1 * d.cpp] = P.rho * P.vx;
          Uo[cidx + 2 * d.cpp] = P.rho * P.vy;
          Uo[cidx + 3 * d.cpp] = P.rho * P.vz;
          Uo[cidx + 4 * d.cpp] = P.p / (gam - 1.0) + 0.5 * P.rho * (P.vx * P.vx + P.vy * P.vy + P.vz * P.vz) +
                                 ob8pi * (B[0] * B[0] + B[1] * B[1] + B[2] * B[2]);
          Uo[cidx + 5 * d.cpp] = B[0];
          Uo[cidx + 6 * d.cpp] = B[1];
          Uo[cidx + 7 * d.cpp] = B[2];
          ento[s * d.cpp + d.cell(i, j, k)] = P.p / pow(P.rho, gam - 1.0);
        }
  };
  std::vector<std::thread> th;
  for (int s = 0; s < 6; ++s) th.emplace_back(sector, s);
  for (auto& t : th) t.join();
  return CS_OK;
}
