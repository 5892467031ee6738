// cs_p1s_host.cpp -- host side of the separable P1 geometry (cs_p1geo.h): the
// per-column tables of the six sectors, and host evaluations of the face and
// zone geometry in the reference's per-patch layouts (used by the tests to
// compare the tables with the oracle's restatement of 220/230/240).
//
// Compiled by g++ with -ffp-contract=off, like the lattice it starts from.
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <vector>

#include "../../include/cubesphere_hip.h"
#include "cs_p1geo.h"
#include "cs_p1s_host.h"

namespace cs {
void set_error(const char* fmt, ...);
int validate_config(const cs_config* c);
}  // namespace cs

namespace {

inline void cross3(const double* a, const double* b, double* o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}

// lateral midpoint at unit radius, pushed onto the sphere of radius |a| as the
// reference pushes the node of a lateral edge onto the sphere of its first
// vertex (220...f:196-300)
void lateral_mid(const double* a, const double* b, double* m) {
  csg1::node_mid(a, b, true, csg1::len3(a), m);
}

// one lateral-face table entry from its two vertices a (s = -1/2) and b
// (s = +1/2); xface: normal along u' x u (x-faces) or u x u' (y-faces)
void lateral_entry(const double* ua, const double* ub, bool xface, double* e) {
  double m[3], d0[3], dp[3], dm[3], c[3];
  lateral_mid(ua, ub, m);
  for (int q = 0; q < 3; ++q) {
    d0[q] = ub[q] - ua[q];
    dp[q] = 3.0 * ub[q] + ua[q] - 4.0 * m[q];
    dm[q] = -3.0 * ua[q] - ub[q] + 4.0 * m[q];
  }
  if (xface) cross3(d0, m, c);
  else cross3(m, d0, c);
  const double g0 = csg1::len3(c), inv = 1.0 / fmax(g0, csg1::kSmall);
  for (int q = 0; q < 3; ++q) {
    e[csg1::TL_M + q] = m[q];
    e[csg1::TL_N + q] = inv * c[q];
  }
  e[csg1::TL_G0] = g0;
  cross3(dp, ub, c);
  e[csg1::TL_GP] = csg1::len3(c);
  cross3(dm, ua, c);
  e[csg1::TL_GM] = csg1::len3(c);
}

// 8-node quadrilateral on [-1/2, 1/2]^2, nodes: corners (--, +-, ++, -+),
// then the edge nodes (0,-), (+,0), (0,+), (-,0); value and gradient
void q8(int n, double x, double y, double& v, double& dx, double& dy) {
  static const int sg[8][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1}, {0, -1}, {1, 0}, {0, 1}, {-1, 0}};
  const double sx = sg[n][0], sy = sg[n][1];
  if (n < 4) {
    const double a = 0.5 + sx * x, b = 0.5 + sy * y, e = 2.0 * sx * x + 2.0 * sy * y - 1.0;
    v = a * b * e;
    dx = sx * b * e + a * b * 2.0 * sx;
    dy = sy * a * e + a * b * 2.0 * sy;
  } else if (sx == 0.0) {
    const double b = 0.5 + sy * y;
    v = (1.0 - 4.0 * x * x) * b;
    dx = -8.0 * x * b;
    dy = (1.0 - 4.0 * x * x) * sy;
  } else {
    const double a = 0.5 + sx * x;
    v = a * (1.0 - 4.0 * y * y);
    dx = sx * (1.0 - 4.0 * y * y);
    dy = -8.0 * y * a;
  }
}

// one radial-face table entry from the four lattice vertices (i-1, j-1),
// (i, j-1), (i, j), (i-1, j)
void radial_entry(const double* u00, const double* u10, const double* u11, const double* u01, double* e) {
  static const double pq[5][2] = {{0.0, 0.0}, {0.5, 0.0}, {-0.5, 0.0}, {0.0, 0.5}, {0.0, -0.5}};
  double X[8][3];
  for (int q = 0; q < 3; ++q) {
    X[0][q] = u00[q];
    X[1][q] = u10[q];
    X[2][q] = u11[q];
    X[3][q] = u01[q];
  }
  const double r0 = csg1::len3(X[0]);
  for (int n = 0; n < 4; ++n) csg1::node_mid(X[n], X[(n + 1) & 3], true, r0, X[4 + n]);
  double w[5], P[5][3];
  for (int p = 0; p < 5; ++p) {
    double A[3], B[3];
    for (int q = 0; q < 3; ++q) {
      double x = 0.0, a = 0.0, b = 0.0;
      for (int n = 0; n < 8; ++n) {
        double v, dx, dy;
        q8(n, pq[p][0], pq[p][1], v, dx, dy);
        x = x + X[n][q] * v;
        a = a + X[n][q] * dx;
        b = b + X[n][q] * dy;
      }
      P[p][q] = x;
      A[q] = a;
      B[q] = b;
    }
    double C[3];
    cross3(A, B, C);
    const double cm = csg1::len3(C), inv = 1.0 / fmax(cm, csg1::kSmall);
    for (int q = 0; q < 3; ++q) {
      e[csg1::TZ_U + 3 * p + q] = P[p][q];
      e[csg1::TZ_N + 3 * p + q] = inv * C[q];
    }
    e[csg1::TZ_A + p] = cm;
    w[p] = csg1::face_w(p) * cm;
  }
  double a = 0.0;
  for (int p = 0; p < 5; ++p) a = a + w[p];
  a = fmax(csg1::kSmall, a);
  double c[3];
  for (int q = 0; q < 3; ++q) {
    double s = 0.0;
    for (int p = 0; p < 5; ++p) s = s + (w[p] / a) * P[p][q];
    c[q] = s;
  }
  // the centroid is put back on the sphere of the face's (i, j) vertex
  const double t = csg1::len3(X[2]) / fmax(csg1::kSmall, csg1::len3(c));
  for (int q = 0; q < 3; ++q) e[csg1::TZ_C + q] = c[q] * t;
}

}  // namespace

namespace cs {

int p1geo_host(const cs_config* cfg, P1GeoHost& H) {
  int rc = validate_config(cfg);
  if (rc) return rc;
  const int N = cfg->ntp, NR = cfg->nr, g = cfg->nghost;
  csg1::P1Geo& G = H.G;
  G.N = N;
  G.NR = NR;
  G.g = g;
  G.nz = N + 2 * g;
  G.nf = N + 2 * g + 1;
  G.t2 = (int64_t)G.nf * G.nf;
  G.tl = (int64_t)G.nf * G.nz;
  G.tz = (int64_t)G.nz * G.nz;
  H.lat.assign((size_t)G.t2 * 18, 0.0);
  H.rb.assign((size_t)(NR + 2 * g + 1), 0.0);
  if ((rc = cs_build_unit_lattice(cfg, H.lat.data()))) return rc;
  if ((rc = cs_build_radial_mesh(cfg, H.rb.data()))) return rc;
  // FILL_GHOST2D_GLOBAL (203fill_ghost2d_global.f:7) fills the four ghost
  // strips only: the tri-corner blocks stay zero, which marks their zones
  // inactive (210extract_vert_coords_patch.f:300-330)
  for (int s = 0; s < 6; ++s)
    for (int c = 0; c < 3; ++c)
      for (int j = 0; j < G.nf; ++j)
        for (int i = 0; i < G.nf; ++i) {
          const bool ci = i < g || i > g + N, cj = j < g || j > g + N;
          if (ci && cj) H.lat[(size_t)i + (size_t)G.nf * (j + (size_t)G.nf * (c + 3 * (size_t)s))] = 0.0;
        }
  G.lat = H.lat.data();
  G.rb = H.rb.data();
  H.TX.assign((size_t)csg1::TL_COUNT * 6 * G.tl, 0.0);
  H.TY.assign((size_t)csg1::TL_COUNT * 6 * G.tl, 0.0);
  H.TZ.assign((size_t)csg1::TZ_COUNT * 6 * G.tz, 0.0);
  double e[csg1::TZ_COUNT], ua[3], ub[3], uc[3], ud[3];
  for (int s = 0; s < 6; ++s) {
    for (int j = 1 - g; j <= N + g; ++j)
      for (int i = -g; i <= N + g; ++i) {
        G.u(s, i, j - 1, ua);
        G.u(s, i, j, ub);
        lateral_entry(ua, ub, true, e);
        for (int f = 0; f < csg1::TL_COUNT; ++f) H.TX[((size_t)f * 6 + s) * G.tl + G.ix(i, j)] = e[f];
      }
    for (int j = -g; j <= N + g; ++j)
      for (int i = 1 - g; i <= N + g; ++i) {
        G.u(s, i - 1, j, ua);
        G.u(s, i, j, ub);
        lateral_entry(ua, ub, false, e);
        for (int f = 0; f < csg1::TL_COUNT; ++f) H.TY[((size_t)f * 6 + s) * G.tl + G.iy(i, j)] = e[f];
      }
    for (int j = 1 - g; j <= N + g; ++j)
      for (int i = 1 - g; i <= N + g; ++i) {
        G.u(s, i - 1, j - 1, ua);
        G.u(s, i, j - 1, ub);
        G.u(s, i, j, uc);
        G.u(s, i - 1, j, ud);
        radial_entry(ua, ub, uc, ud, e);
        for (int f = 0; f < csg1::TZ_COUNT; ++f) H.TZ[((size_t)f * 6 + s) * G.tz + G.iz(i, j)] = e[f];
      }
  }
  G.TX = H.TX.data();
  G.TY = H.TY.data();
  G.TZ = H.TZ.data();
  return CS_OK;
}

}  // namespace cs

// Host evaluation of the separable face geometry in the layouts of
// cs_geometry (cs_face_quad): six sectors, sector slowest.
extern "C" int cs_p1s_face_geometry(const cs_config* cfg, int32_t fam, double* pts, double* wts, double* normal,
                                    double* area, double* centroid, double* bwts) {
  if (!cfg || fam < 0 || fam > 2) {
    cs::set_error("cs_p1s_face_geometry: bad arguments");
    return CS_E_INVALID;
  }
  cs::P1GeoHost H;
  int rc = cs::p1geo_host(cfg, H);
  if (rc) return rc;
  const csg1::P1Geo& G = H.G;
  const int N = G.N, NR = G.NR, g = G.g, nz = G.nz, nf = G.nf, nzr = NR + 2 * g;
  const int64_t S = fam == 2 ? (int64_t)nz * nz * (nzr + 1) : (int64_t)nf * nz * nzr;
  const int i0 = fam == 0 ? -g : 1 - g, j0 = fam == 1 ? -g : 1 - g, k0 = fam == 2 ? -g : 1 - g;
  const int ni = fam == 0 ? nf : nz, nj = fam == 1 ? nf : nz;
  csg1::FaceG F;
  for (int s = 0; s < 6; ++s)
    for (int k = k0; k <= NR + g; ++k)
      for (int j = j0; j <= N + g; ++j)
        for (int i = i0; i <= N + g; ++i) {
          csg1::face_of(G, fam, s, i, j, k, F);
          const int64_t o = (int64_t)s * S + (i - i0) + (int64_t)ni * ((j - j0) + (int64_t)nj * (k - k0));
          const int64_t b = (int64_t)s * S;  // the trailing dimensions follow each sector's block
          const int64_t oo = o - b;
          for (int p = 0; p < 5; ++p) {
            for (int c = 0; c < 3; ++c) {
              if (pts) pts[b * 15 + oo + (c + 3 * p) * S] = F.p[p][c];
              if (normal && (fam == 2 || p == 0)) normal[b * (fam == 2 ? 15 : 3) + oo + (c + 3 * p) * S] = F.n[p][c];
            }
            if (wts) wts[b * 5 + oo + p * S] = F.w[p];
            if (bwts) bwts[b * 5 + oo + p * S] = F.bw[p];
          }
          if (area) area[o] = F.area;
          for (int c = 0; c < 3; ++c)
            if (centroid) centroid[b * 3 + oo + c * S] = F.cen[c];
        }
  return CS_OK;
}

// Host evaluation of the zone geometry (centroid (zone, 3), volume, scale
// length) of every zone of the six sectors, sector slowest.
extern "C" int cs_p1s_zone_geometry(const cs_config* cfg, double* centroid, double* volume, double* scale) {
  if (!cfg) {
    cs::set_error("cs_p1s_zone_geometry: null config");
    return CS_E_INVALID;
  }
  cs::P1GeoHost H;
  int rc = cs::p1geo_host(cfg, H);
  if (rc) return rc;
  const csg1::P1Geo& G = H.G;
  const int N = G.N, NR = G.NR, g = G.g, nz = G.nz;
  const int64_t S = (int64_t)nz * nz * (NR + 2 * g);
  for (int s = 0; s < 6; ++s)
    for (int k = 1 - g; k <= NR + g; ++k)
      for (int j = 1 - g; j <= N + g; ++j)
        for (int i = 1 - g; i <= N + g; ++i) {
          double c[3], v, l;
          csg1::zone_geometry(G, s, i, j, k, c, &v, &l);
          const int64_t o = (i - 1 + g) + (int64_t)nz * ((j - 1 + g) + (int64_t)nz * (k - 1 + g));
          for (int q = 0; q < 3; ++q)
            if (centroid) centroid[(int64_t)s * 3 * S + o + q * S] = c[q];
          if (volume) volume[(int64_t)s * S + o] = v;
          if (scale) scale[(int64_t)s * S + o] = l;
        }
  return CS_OK;
}
