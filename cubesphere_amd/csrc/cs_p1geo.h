// cs_p1geo.h -- the quadratic isoparametric face geometry of the reference's
// P1 reconstruction, evaluated from separable tables instead of being stored
// per face.
//
// The reference stores, per face, five quadrature points, weights, normals,
// the area and the centroid (COORDS_TO_WT_QUADPTS_FACES_ISO2,
// 220coords_to_wt_quadpts_faces_iso2.f:178-878) and the B-field weights of the
// same points (COORDS_TO_WT_QUADPTS_BFLD_ISO2, 230...f:8): about 90 words per
// zone.  Every vertex is rb(k) u(i, j) (210extract_vert_coords_patch.f:300-308),
// and the 8-node serendipity face nodes are rb u at the corners, plain
// midpoints along the radial edges and midpoints pushed back onto the sphere
// along the lateral edges.  With r linear along a radial edge and u quadratic
// along a lateral one, the element map of a lateral face is exactly
// x(s, t) = r(t) u(s) (the 8-node space holds r u), and that of a radial face
// x(s, t) = rb(k) u(s, t).  Hence
//   * lateral (x, y) faces: points r(t_p) u(s_p); the cross product of the
//     tangents is r(t) r' (u'(s) x u(s)); the normal is that of the plane of
//     the face (one per face); only |u' x u| at s = 0, +-1/2 is needed;
//   * radial (z) faces: points rb u_p, normals and normalised weights do not
//     depend on k; the raw weights scale with rb^2.
// The tables below hold the angular factors per (i, j) column of each sector
// (O(N^2)); the radial factors are rb.  The results equal the reference's
// per-face arrays in exact arithmetic and to a few units of the last place
// in floating point (the tests compare them with the oracle's restatement of
// 220/230).
#pragma once
#include <math.h>
#include <stdint.h>

#include "cs_internal.h"

namespace csg1 {

constexpr double kSmall = 1.0e-12;  // RIEMANN_COM smallnum (090init_riemann.f:12)

// lateral-face table fields: the pushed-back lateral midpoint at unit radius
// (3), the unit normal (3), |u'(s) x u(s)| at s = 0, +1/2 (the b vertex side),
// -1/2 (the a vertex side)
enum { TL_M = 0, TL_N = 3, TL_G0 = 6, TL_GP = 7, TL_GM = 8, TL_COUNT = 9 };
// radial-face table fields: the five points at unit radius (15), the unit
// normals there (15), |A x B| there (5), the centroid direction scaled by
// |u(i, j)| (3)
enum { TZ_U = 0, TZ_N = 15, TZ_A = 30, TZ_C = 35, TZ_COUNT = 38 };

// face rule of 220...f:212-226: centre 1/3, the four edge midpoints 1/6
CS_HD double face_w(int p) { return p == 0 ? (1.0 / 3.0) : (1.0 / 6.0); }

struct P1Geo {
  int N, NR, g, nz, nf;
  int64_t t2;       // lattice entries per sector, nf * nf
  int64_t tl;       // lateral-face table entries per sector, nf * nz
  int64_t tz;       // radial-face table entries per sector, nz * nz
  const double* lat;  // [s][c][t2]: unit lattice, tri-corner blocks zero
  const double* rb;   // rb[k + g], k = -g .. nr + g
  const double* TX;   // [field][s][tl], x-face (vertex i, zone j)
  const double* TY;   // [field][s][tl], y-face (zone i, vertex j)
  const double* TZ;   // [field][s][tz], z-face (zone i, zone j)

  CS_HD double r(int k) const { return rb[k + g]; }
  CS_HD void u(int s, int i, int j, double* o) const {
    const int64_t q = (int64_t)(i + g) + (int64_t)nf * (j + g);
    for (int c = 0; c < 3; ++c) o[c] = lat[(int64_t)(3 * s + c) * t2 + q];
  }
  CS_HD void vert(int s, int i, int j, int k, double* o) const {
    const double rr = r(k);
    const int64_t q = (int64_t)(i + g) + (int64_t)nf * (j + g);
    for (int c = 0; c < 3; ++c) o[c] = rr * lat[(int64_t)(3 * s + c) * t2 + q];
  }
  CS_HD int64_t ix(int i, int j) const { return (int64_t)(i + g) + (int64_t)nf * (j - 1 + g); }
  CS_HD int64_t iy(int i, int j) const { return (int64_t)(i - 1 + g) + (int64_t)nz * (j + g); }
  CS_HD int64_t iz(int i, int j) const { return (int64_t)(i - 1 + g) + (int64_t)nz * (j - 1 + g); }
  CS_HD double tx(int f, int s, int64_t q) const { return TX[((int64_t)f * 6 + s) * tl + q]; }
  CS_HD double ty(int f, int s, int64_t q) const { return TY[((int64_t)f * 6 + s) * tl + q]; }
  CS_HD double tzv(int f, int s, int64_t q) const { return TZ[((int64_t)f * 6 + s) * tz + q]; }
};

// One face: the five points, the normal(s) (n[0] on lateral faces, all five
// on radial faces), the weights divided by the area (w), the B-field weights
// (bw: the raw weights times the reciprocal of their sum), the area and the
// centroid.  fam 0: x-face (vertex i, zone j, zone k); 1: y-face (zone i,
// vertex j, zone k); 2: z-face (zone i, zone j, vertex k).
struct FaceG {
  double p[5][3];
  double n[5][3];
  double w[5], bw[5];
  double area;
  double cen[3];
};

CS_HD void lateral_face(const P1Geo& G, const double* T, int64_t stride, int64_t q, const double* ua,
                        const double* ub, double r0, double r1, bool xface, FaceG& F) {
  double m[3];
  for (int c = 0; c < 3; ++c) {
    m[c] = T[(TL_M + c) * stride + q];
    F.n[0][c] = T[(TL_N + c) * stride + q];
  }
  const double g0 = T[TL_G0 * stride + q], gp = T[TL_GP * stride + q], gm = T[TL_GM * stride + q];
  // the radial midpoint node is the plain mean of the two vertices
  const double rm = 0.5 * (r0 + r1), rp = r1 - r0;
  double w[5];
  if (xface) {  // s along j (a -> b), t along k: p0 (0,0), p1 (1/2,0), p2 (-1/2,0), p3 (0,1/2), p4 (0,-1/2)
    for (int c = 0; c < 3; ++c) {
      F.p[0][c] = rm * m[c];
      F.p[1][c] = rm * ub[c];
      F.p[2][c] = rm * ua[c];
      F.p[3][c] = r1 * m[c];
      F.p[4][c] = r0 * m[c];
    }
    w[0] = face_w(0) * (rm * rp * g0);
    w[1] = face_w(1) * (rm * rp * gp);
    w[2] = face_w(2) * (rm * rp * gm);
    w[3] = face_w(3) * (r1 * rp * g0);
    w[4] = face_w(4) * (r0 * rp * g0);
  } else {  // s along k, t along i (a -> b)
    for (int c = 0; c < 3; ++c) {
      F.p[0][c] = rm * m[c];
      F.p[1][c] = r1 * m[c];
      F.p[2][c] = r0 * m[c];
      F.p[3][c] = rm * ub[c];
      F.p[4][c] = rm * ua[c];
    }
    w[0] = face_w(0) * (rp * rm * g0);
    w[1] = face_w(1) * (rp * r1 * g0);
    w[2] = face_w(2) * (rp * r0 * g0);
    w[3] = face_w(3) * (rp * rm * gp);
    w[4] = face_w(4) * (rp * rm * gm);
  }
  double a = 0.0;
  for (int p = 0; p < 5; ++p) a = a + w[p];
  const double ob = 1.0 / fmax(a, kSmall);
  for (int p = 0; p < 5; ++p) F.bw[p] = ob * w[p];
  a = fmax(kSmall, a);
  F.area = a;
  for (int p = 0; p < 5; ++p) F.w[p] = w[p] / a;
  for (int c = 0; c < 3; ++c) {
    double s = 0.0;
    for (int p = 0; p < 5; ++p) s = s + F.w[p] * F.p[p][c];
    F.cen[c] = s;
  }
}

CS_HD void face_x(const P1Geo& G, int s, int i, int j, int k, FaceG& F) {
  double ua[3], ub[3];
  G.u(s, i, j - 1, ua);
  G.u(s, i, j, ub);
  lateral_face(G, G.TX + (int64_t)s * G.tl, 6 * G.tl, G.ix(i, j), ua, ub, G.r(k - 1), G.r(k), true, F);
}
CS_HD void face_y(const P1Geo& G, int s, int i, int j, int k, FaceG& F) {
  double ua[3], ub[3];
  G.u(s, i - 1, j, ua);
  G.u(s, i, j, ub);
  lateral_face(G, G.TY + (int64_t)s * G.tl, 6 * G.tl, G.iy(i, j), ua, ub, G.r(k - 1), G.r(k), false, F);
}
CS_HD void face_z(const P1Geo& G, int s, int i, int j, int k, FaceG& F) {
  const int64_t q = G.iz(i, j), st = 6 * G.tz;
  const double* T = G.TZ + (int64_t)s * G.tz;
  const double r = G.r(k), r2 = r * r;
  double w[5];
  for (int p = 0; p < 5; ++p) {
    for (int c = 0; c < 3; ++c) {
      F.p[p][c] = r * T[(TZ_U + 3 * p + c) * st + q];
      F.n[p][c] = T[(TZ_N + 3 * p + c) * st + q];
    }
    w[p] = face_w(p) * (r2 * T[(TZ_A + p) * st + q]);
  }
  double a = 0.0;
  for (int p = 0; p < 5; ++p) a = a + w[p];
  const double ob = 1.0 / fmax(a, kSmall);
  for (int p = 0; p < 5; ++p) F.bw[p] = ob * w[p];
  a = fmax(kSmall, a);
  F.area = a;
  for (int p = 0; p < 5; ++p) F.w[p] = w[p] / a;
  for (int c = 0; c < 3; ++c) F.cen[c] = r * T[(TZ_C + c) * st + q];
}
CS_HD void face_of(const P1Geo& G, int fam, int s, int i, int j, int k, FaceG& F) {
  if (fam == 0) face_x(G, s, i, j, k, F);
  else if (fam == 1) face_y(G, s, i, j, k, F);
  else face_z(G, s, i, j, k, F);
}
// the first point alone, computed as face_of computes F.p[0]
CS_HD void face_p0(const P1Geo& G, int fam, int s, int i, int j, int k, double* p) {
  if (fam == 2) {
    const int64_t q = G.iz(i, j), st = 6 * G.tz;
    const double* T = G.TZ + (int64_t)s * G.tz;
    const double r = G.r(k);
    for (int c = 0; c < 3; ++c) p[c] = r * T[(TZ_U + c) * st + q];
    return;
  }
  const double* T = (fam == 0 ? G.TX : G.TY) + (int64_t)s * G.tl;
  const int64_t q = fam == 0 ? G.ix(i, j) : G.iy(i, j), st = 6 * G.tl;
  const double rm = 0.5 * (G.r(k - 1) + G.r(k));
  for (int c = 0; c < 3; ++c) p[c] = rm * T[(TL_M + c) * st + q];
}

// Only the normal (lateral faces) or the centroid direction (radial faces):
// the pieces the flattener and the cleanup read.
CS_HD void normal_x(const P1Geo& G, int s, int i, int j, double* n) {
  for (int c = 0; c < 3; ++c) n[c] = G.tx(TL_N + c, s, G.ix(i, j));
}
CS_HD void normal_y(const P1Geo& G, int s, int i, int j, double* n) {
  for (int c = 0; c < 3; ++c) n[c] = G.ty(TL_N + c, s, G.iy(i, j));
}
CS_HD void zcen_dir(const P1Geo& G, int s, int i, int j, double* n) {
  for (int c = 0; c < 3; ++c) n[c] = G.tzv(TZ_C + c, s, G.iz(i, j));
}
CS_HD double area_x(const P1Geo& G, int s, int i, int j, int k) {
  const int64_t q = G.ix(i, j);
  const double r0 = G.r(k - 1), r1 = G.r(k), rm = 0.5 * (r0 + r1), rp = r1 - r0;
  const double g0 = G.tx(TL_G0, s, q), gp = G.tx(TL_GP, s, q), gm = G.tx(TL_GM, s, q);
  double a = face_w(0) * (rm * rp * g0) + face_w(1) * (rm * rp * gp) + face_w(2) * (rm * rp * gm) +
             face_w(3) * (r1 * rp * g0) + face_w(4) * (r0 * rp * g0);
  return fmax(kSmall, a);
}
CS_HD double area_y(const P1Geo& G, int s, int i, int j, int k) {
  const int64_t q = G.iy(i, j);
  const double r0 = G.r(k - 1), r1 = G.r(k), rm = 0.5 * (r0 + r1), rp = r1 - r0;
  const double g0 = G.ty(TL_G0, s, q), gp = G.ty(TL_GP, s, q), gm = G.ty(TL_GM, s, q);
  double a = face_w(0) * (rp * rm * g0) + face_w(1) * (rp * r1 * g0) + face_w(2) * (rp * r0 * g0) +
             face_w(3) * (rp * rm * gp) + face_w(4) * (rp * rm * gm);
  return fmax(kSmall, a);
}
CS_HD double area_z(const P1Geo& G, int s, int i, int j, int k) {
  const int64_t q = G.iz(i, j);
  const double r = G.r(k), r2 = r * r;
  double a = 0.0;
  for (int p = 0; p < 5; ++p) a = a + face_w(p) * (r2 * G.tzv(TZ_A + p, s, q));
  return fmax(kSmall, a);
}

}  // namespace csg1

namespace csg1 {

// 20-node serendipity hexahedron on [-1/2, 1/2]^3: corners (lower layer,
// then upper, counter-clockwise), then the edge nodes along x, along y and
// along z (0 on the edge's own axis); value and gradient of node n at x.
CS_HD void h20(int n, const double* x, double& v, double* d) {
  // node signs, three per node, packed as (s + 1) in two bits each
  const int packed[20] = {0x00, 0x02, 0x0a, 0x08, 0x20, 0x22, 0x2a, 0x28, 0x01, 0x09, 0x29, 0x21,
                          0x04, 0x24, 0x26, 0x06, 0x10, 0x12, 0x1a, 0x18};
  double sg[3], l[3], dl[3];
  int e = -1;
  for (int a = 0; a < 3; ++a) {
    sg[a] = (double)(((packed[n] >> (2 * a)) & 3) - 1);
    if (sg[a] == 0.0) {
      e = a;
      l[a] = 1.0 - 4.0 * x[a] * x[a];
      dl[a] = -8.0 * x[a];
    } else {
      l[a] = 0.5 + sg[a] * x[a];
      dl[a] = sg[a];
    }
  }
  const double prod = l[0] * l[1] * l[2];
  if (e < 0) {
    const double t = 2.0 * sg[0] * x[0] + 2.0 * sg[1] * x[1] + 2.0 * sg[2] * x[2] - 2.0;
    v = prod * t;
    d[0] = dl[0] * (l[1] * l[2]) * t + prod * 2.0 * sg[0];
    d[1] = dl[1] * (l[0] * l[2]) * t + prod * 2.0 * sg[1];
    d[2] = dl[2] * (l[0] * l[1]) * t + prod * 2.0 * sg[2];
  } else {
    v = prod;
    d[0] = dl[0] * (l[1] * l[2]);
    d[1] = dl[1] * (l[0] * l[2]);
    d[2] = dl[2] * (l[0] * l[1]);
  }
}

CS_HD double len3(const double* v) { return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }
// midpoint of two nodes; lateral edges are pushed onto the sphere of radius r
CS_HD void node_mid(const double* a, const double* b, bool onto_sphere, double r, double* m) {
  for (int c = 0; c < 3; ++c) m[c] = 0.5 * (a[c] + b[c]);
  if (onto_sphere) {
    const double t = r / fmax(kSmall, len3(m));
    for (int c = 0; c < 3; ++c) m[c] = m[c] * t;
  }
}

// The 8 Gauss points of zone (i, j, k) (P) and their weights det J / 8 (w);
// returns the volume, their sum.  X: the 20 nodes, filled on return.
CS_HD double zone_gauss(const P1Geo& G, int s, int i, int j, int k, double (*X)[3], double (*P)[3], double* w) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
  for (int n = 0; n < 8; ++n) {
    const int di = (n == 1 || n == 2 || n == 5 || n == 6) ? 1 : 0, dj = (n == 2 || n == 3 || n == 6 || n == 7) ? 1 : 0;
    G.vert(s, i - 1 + di, j - 1 + dj, k - 1 + (n >= 4 ? 1 : 0), X[n]);
  }
  const double r1 = len3(X[0]), r2 = len3(X[4]);
  node_mid(X[0], X[1], true, r1, X[8]);
  node_mid(X[2], X[3], true, r1, X[9]);
  node_mid(X[6], X[7], true, r2, X[10]);
  node_mid(X[4], X[5], true, r2, X[11]);
  node_mid(X[0], X[3], true, r1, X[12]);
  node_mid(X[4], X[7], true, r2, X[13]);
  node_mid(X[5], X[6], true, r2, X[14]);
  node_mid(X[1], X[2], true, r1, X[15]);
  node_mid(X[0], X[4], false, 0.0, X[16]);
  node_mid(X[1], X[5], false, 0.0, X[17]);
  node_mid(X[2], X[6], false, 0.0, X[18]);
  node_mid(X[3], X[7], false, 0.0, X[19]);
  const double xg = 0.5 * sqrt(1.0 / 3.0);
  double v = 0.0;
  for (int p = 0; p < 8; ++p) {
    const double x[3] = {(p & 1) ? xg : -xg, (p & 2) ? xg : -xg, (p & 4) ? xg : -xg};
    double T[3][3];
    for (int q = 0; q < 3; ++q) {
      double a = 0.0, t0 = 0.0, t1 = 0.0, t2 = 0.0;
      for (int n = 0; n < 20; ++n) {
        double nv, nd[3];
        h20(n, x, nv, nd);
        a = a + X[n][q] * nv;
        t0 = t0 + X[n][q] * nd[0];
        t1 = t1 + X[n][q] * nd[1];
        t2 = t2 + X[n][q] * nd[2];
      }
      P[p][q] = a;
      T[0][q] = t0;
      T[1][q] = t1;
      T[2][q] = t2;
    }
    const double Dx = T[0][1] * T[1][2] - T[0][2] * T[1][1];
    const double Dy = T[0][2] * T[1][0] - T[0][0] * T[1][2];
    const double Dz = T[0][0] * T[1][1] - T[0][1] * T[1][0];
    w[p] = 0.125 * fabs(T[2][0] * Dx + T[2][1] * Dy + T[2][2] * Dz);
  }
  for (int p = 0; p < 8; ++p) v = v + w[p];
  return fmax(kSmall, v);
}

// Zone (i, j, k) of sector s: volume and centroid of the 20-node element by
// the 2 x 2 x 2 Gauss rule (COORDS_TO_WT_QUADPTS_VOLUME_ISO2,
// 240coords_to_wt_quadpts_volume_iso2.f:97-300: weights det J / 8, centroid
// as the vertex mean plus the weighted offsets) and the scale length (the
// radial edge and the four lateral edges of the lower face,
// 220...f:2212-2290).
CS_HD void zone_geometry(const P1Geo& G, int s, int i, int j, int k, double* cen, double* vol, double* scale) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
  double X[20][3], P[8][3], w[8];
  const double v = zone_gauss(G, s, i, j, k, X, P, w);
  *vol = v;
  for (int q = 0; q < 3; ++q) {
    const double c0 = 0.125 * (X[0][q] + X[1][q] + X[2][q] + X[3][q] + X[4][q] + X[5][q] + X[6][q] + X[7][q]);
    double t = 0.0;
    for (int p = 0; p < 8; ++p) t = t + (w[p] / v) * (P[p][q] - c0);
    cen[q] = c0 + t;
  }
  // scale length
  double a[3], b[3];
  G.vert(s, i, j, k, a);
  G.vert(s, i, j, k - 1, b);
  double d3[3] = {a[0] - b[0], a[1] - b[1], a[2] - b[2]};
  double L = len3(d3);
  const int e[4][2][2] = {{{i - 1, j}, {i, j}}, {{i - 1, j - 1}, {i, j - 1}}, {{i, j - 1}, {i, j}},
                          {{i - 1, j - 1}, {i - 1, j}}};
  for (int q = 0; q < 4; ++q) {
    G.vert(s, e[q][0][0], e[q][0][1], k - 1, a);
    G.vert(s, e[q][1][0], e[q][1][1], k - 1, b);
    for (int c = 0; c < 3; ++c) d3[c] = a[c] - b[c];
    L = fmin(L, len3(d3));
  }
  *scale = L;
}

}  // namespace csg1
