/* cs_oracle.c -- CPU oracle of the CubeSphere cell-update path.
 *
 * TEST INFRASTRUCTURE ONLY (see cs_oracle.h).  Plain C, double precision,
 * loop-per-zone.  The algorithm restated here is the one the GPU path
 * implements: reference routines where the reference has them, and the
 * documented new numerics (WENO5, HLLD, SSP-RK3, exact spherical metric)
 * where it does not.  Ghost-exchange maps are NOT taken from index formulas:
 * every ghost zone/face is matched geometrically to the zone/face of the
 * neighbouring sector that occupies the same place on the sphere, so the
 * index flips and face-B sign rules of 330fetch_buffers_fill_in_ux.f are
 * checked against geometry.
 */
#include "cs_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define OB8PI (1.0 / (8.0 * 3.14159265358979323846))
#define OB4PI (1.0 / (4.0 * 3.14159265358979323846))
#define PI_ 3.14159265358979323846

/* ------------------------------------------------------------- lattice */
/* SWIVEL_N, 201mesh_sector_general.f:5-59 */
static void swivel(const double* a1, const double* a2, int n, int lo, int hi, double* arr) {
  double dot = a1[0] * a2[0] + a1[1] * a2[1] + a1[2] * a2[2];
  double theta = acos(dot), dth = theta / (double)n, x[3], y[3], t;
  int i, c;
  for (c = 0; c < 3; ++c) x[c] = a1[c];
  for (c = 0; c < 3; ++c) y[c] = a2[c] - dot * x[c];
  t = sqrt(y[0] * y[0] + y[1] * y[1] + y[2] * y[2]);
  for (c = 0; c < 3; ++c) y[c] = y[c] / t;
  for (i = lo; i <= hi; ++i) {
    double* o = arr + 3 * (i - lo);
    if (i == 0) { for (c = 0; c < 3; ++c) o[c] = a1[c]; }
    else if (i == n) { for (c = 0; c < 3; ++c) o[c] = a2[c]; }
    else {
      double tt = dth * (double)i, ct = cos(tt), st = sin(tt);
      for (c = 0; c < 3; ++c) o[c] = ct * x[c] + st * y[c];
    }
  }
}

/* sector corners, 201mesh_sector_general.f:175-330 */
static const int kCorner[6][4][3] = {
    {{-1, -1, 1}, {1, -1, 1}, {-1, 1, 1}, {1, 1, 1}},   {{-1, 1, -1}, {-1, 1, 1}, {1, 1, -1}, {1, 1, 1}},
    {{1, -1, -1}, {1, 1, -1}, {1, -1, 1}, {1, 1, 1}},   {{1, 1, -1}, {1, -1, -1}, {-1, 1, -1}, {-1, -1, -1}},
    {{1, -1, 1}, {-1, -1, 1}, {1, -1, -1}, {-1, -1, -1}}, {{-1, 1, 1}, {-1, 1, -1}, {-1, -1, 1}, {-1, -1, -1}}};
/* neighbours right, left, up, down (1-based), 203fill_ghost2d_global.f:37-75 */
static const int kNb[6][4] = {{3, 6, 2, 5}, {1, 4, 3, 6}, {2, 5, 1, 4}, {5, 2, 6, 3}, {6, 3, 4, 1}, {4, 1, 5, 2}};

static void cross3(const double* a, const double* b, double* o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}
static double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static void unit3(double* a) {
  double n = sqrt(dot3(a, a));
  if (n > 0) { a[0] /= n; a[1] /= n; a[2] /= n; }
}

void or_unit_lattice(int N, int g, double* coords) {
  int nf = N + 2 * g + 1, s, i, j, c;
  double* a1 = malloc(sizeof(double) * 3 * nf);
  double* a2 = malloc(sizeof(double) * 3 * nf);
  double* a3 = malloc(sizeof(double) * 3 * nf);
  double* a4 = malloc(sizeof(double) * 3 * nf);
  double* src;
  double q = 1.0 / sqrt(3.0);
#define CO(ii, jj, cc, ss) coords[(size_t)((ii) + g) + (size_t)nf * (((jj) + g) + (size_t)nf * ((cc) + 3 * (size_t)(ss)))]
  memset(coords, 0, sizeof(double) * nf * nf * 18);
  for (s = 0; s < 6; ++s) {
    double v[4][3];
    int a;
    for (a = 0; a < 4; ++a)
      for (c = 0; c < 3; ++c) v[a][c] = kCorner[s][a][c] > 0 ? q : -q;
    swivel(v[0], v[1], N, 0, N, a1);
    swivel(v[2], v[3], N, 0, N, a2);
    swivel(v[0], v[2], N, 0, N, a3);
    swivel(v[1], v[3], N, 0, N, a4);
    for (j = 0; j <= N; ++j)
      for (i = 0; i <= N; ++i) {
        /* 201mesh_sector_general.f:735-815 */
        const double *p1 = a1 + 3 * i, *p2 = a2 + 3 * i, *p3 = a3 + 3 * j, *p4 = a4 + 3 * j;
        double Ax = p1[1] * p2[2] - p1[2] * p2[1], Ay = p1[2] * p2[0] - p1[0] * p2[2],
               Az = p1[0] * p2[1] - p1[1] * p2[0];
        double Bx = p3[1] * p4[2] - p3[2] * p4[1], By = p3[2] * p4[0] - p3[0] * p4[2],
               Bz = p3[0] * p4[1] - p3[1] * p4[0];
        double Cx = By * Az - Bz * Ay, Cy = Bz * Ax - Bx * Az, Cz = Bx * Ay - By * Ax, t;
        int sec = s + 1;
        if ((sec == 1 && Cz < 0.0) || (sec == 4 && Cz > 0.0) || (sec == 2 && Cy < 0.0) ||
            (sec == 5 && Cy > 0.0) || (sec == 3 && Cx < 0.0) || (sec == 6 && Cx > 0.0)) {
          Cx = -Cx; Cy = -Cy; Cz = -Cz;
        }
        t = 1.0 / sqrt(Cx * Cx + Cy * Cy + Cz * Cz);
        CO(i, j, 0, s) = t * Cx;
        CO(i, j, 1, s) = t * Cy;
        CO(i, j, 2, s) = t * Cz;
      }
  }
  /* FILL_GHOST2D_GLOBAL, 203fill_ghost2d_global.f:78-150 */
  src = malloc(sizeof(double) * nf * nf * 18);
  memcpy(src, coords, sizeof(double) * nf * nf * 18);
#define SR(ii, jj, cc, ss) src[(size_t)((ii) + g) + (size_t)nf * (((jj) + g) + (size_t)nf * ((cc) + 3 * (size_t)(ss)))]
  for (s = 0; s < 6; ++s) {
    int rt = kNb[s][0] - 1, lf = kNb[s][1] - 1, up = kNb[s][2] - 1, dn = kNb[s][3] - 1, x, y;
    for (x = N + 1; x <= N + g; ++x)
      for (y = 0; y <= N; ++y)
        for (c = 0; c < 3; ++c) CO(x, y, c, s) = SR(y, N - 1 - (x - N - 1), c, rt);
    for (x = -1; x >= -g; --x)
      for (y = 0; y <= N; ++y)
        for (c = 0; c < 3; ++c) CO(x, y, c, s) = SR(1 - (x + 1), N - y, c, lf);
    for (y = N + 1; y <= N + g; ++y)
      for (x = 0; x <= N; ++x)
        for (c = 0; c < 3; ++c) CO(x, y, c, s) = SR(N - 1 - (y - N - 1), x, c, up);
    for (y = -1; y >= -g; --y)
      for (x = 0; x <= N; ++x)
        for (c = 0; c < 3; ++c) CO(x, y, c, s) = SR(N - x, 1 - (y + 1), c, dn);
  }
#undef SR
  free(src); free(a1); free(a2); free(a3); free(a4);
}

/* continue the sector's own swivel lattice into the corner ghost blocks */
static void corner_blocks(int N, int g, double* coords) {
  int nf = N + 2 * g + 1, s, i, j, c;
  double* a1 = malloc(sizeof(double) * 3 * nf);
  double* a2 = malloc(sizeof(double) * 3 * nf);
  double* a3 = malloc(sizeof(double) * 3 * nf);
  double* a4 = malloc(sizeof(double) * 3 * nf);
  double q = 1.0 / sqrt(3.0);
  for (s = 0; s < 6; ++s) {
    double v[4][3];
    int a;
    for (a = 0; a < 4; ++a)
      for (c = 0; c < 3; ++c) v[a][c] = kCorner[s][a][c] > 0 ? q : -q;
    swivel(v[0], v[1], N, -g, N + g, a1);
    swivel(v[2], v[3], N, -g, N + g, a2);
    swivel(v[0], v[2], N, -g, N + g, a3);
    swivel(v[1], v[3], N, -g, N + g, a4);
    for (j = -g; j <= N + g; ++j)
      for (i = -g; i <= N + g; ++i) {
        if ((i >= 0 && i <= N) || (j >= 0 && j <= N)) continue;
        {
          const double *p1 = a1 + 3 * (i + g), *p2 = a2 + 3 * (i + g), *p3 = a3 + 3 * (j + g), *p4 = a4 + 3 * (j + g);
          double A[3], B[3], C[3];
          int sec = s + 1;
          cross3(p1, p2, A);
          cross3(p3, p4, B);
          cross3(B, A, C);
          if ((sec == 1 && C[2] < 0.0) || (sec == 4 && C[2] > 0.0) || (sec == 2 && C[1] < 0.0) ||
              (sec == 5 && C[1] > 0.0) || (sec == 3 && C[0] < 0.0) || (sec == 6 && C[0] > 0.0)) {
            C[0] = -C[0]; C[1] = -C[1]; C[2] = -C[2];
          }
          {
            /* normalised as in 201mesh_sector_general.f:810-813 */
            double t = 1.0 / sqrt(C[0] * C[0] + C[1] * C[1] + C[2] * C[2]);
            for (c = 0; c < 3; ++c) CO(i, j, c, s) = t * C[c];
          }
        }
      }
  }
  free(a1); free(a2); free(a3); free(a4);
}
#undef CO

/* GENERATE_GLOBAL_RADIAL_UNIFORM_MESH / _LOG_MESH, 202:10-60, 71-125 */
void or_radial_mesh(const or_config* c, double* rb) {
  int i;
  for (i = -c->nghost; i <= c->nr + c->nghost; ++i) {
    double p = (double)i / (double)c->nr;
    rb[i + c->nghost] = c->radial_mesh == 1 ? c->rmin * pow(c->rmax / c->rmin, p)
                                            : c->rmin + (c->rmax - c->rmin) * p;
  }
}

/* ------------------------------------------------------------- grid */
typedef struct { int fam, s, i, j, sfam, ss, si, sj; double sgn; } fmap_t;
typedef struct { int s, i, j, ss, si, sj; } cmap_t;

struct or_grid {
  or_config c;
  int N, NR, g, nc, nrc, nf, nrf;
  int64_t cpp, fpp, t2;
  double* lat;  /* lattice */
  double* rb;
  /* 2-D tables, each t2*6 */
  double *u[3], *xn[3], *xc[3], *thy, *yn[3], *yc[3], *thx, *zs[3], *zsn, *zt[3], *om, *chmin, *mi[9], *bw[15];
  double *en1[3], *etau[3], *ecos, *esin, *xet[3], *yet[3];
  cmap_t* cmap; int ncmap;
  fmap_t* fmap; int nfmap;
  fmap_t* pairs; int npairs;
  fmap_t* epairs; int nepairs;   /* panel-edge edges: fam 0 Efx, 1 Efy, 2 Efz */
  int corner[8][3][3];           /* cube corners: 3 x (sector, i, j), ascending sector */
  int ncorner;
  int fails;
};

#define CELL(G, i, j, k) ((int64_t)((i) - 1 + (G)->g) + (int64_t)(G)->nc * (((j) - 1 + (G)->g) + (int64_t)(G)->nc * ((k) - 1 + (G)->g)))
#define FACE(G, i, j, k) ((int64_t)((i) + (G)->g) + (int64_t)(G)->nf * (((j) + (G)->g) + (int64_t)(G)->nf * ((k) + (G)->g)))
#define T2(G, s, i, j) ((int64_t)(s) * (G)->t2 + (int64_t)((i) + (G)->g) + (int64_t)(G)->nf * ((j) + (G)->g))

static void uv(const or_grid* G, int s, int i, int j, double* o) {
  int c;
  for (c = 0; c < 3; ++c) o[c] = G->lat[(size_t)(i + G->g) + (size_t)G->nf * ((j + G->g) + (size_t)G->nf * (c + 3 * (size_t)s))];
}
static double ang(const double* a, const double* b) {
  double x[3];
  cross3(a, b, x);
  return atan2(sqrt(dot3(x, x)), dot3(a, b));
}
static void perp_unit(double* v, const double* n) {
  double t = dot3(v, n);
  v[0] -= t * n[0]; v[1] -= t * n[1]; v[2] -= t * n[2];
  unit3(v);
}
static double tri_om(const double* a, const double* b, const double* c) {
  double bc[3];
  cross3(b, c, bc);
  return 2.0 * atan2(dot3(a, bc), 1.0 + dot3(a, b) + dot3(b, c) + dot3(c, a));
}

static double* al(int64_t n) { return calloc((size_t)n, sizeof(double)); }

static void build_geometry(or_grid* G) {
  int N = G->N, g = G->g, s, i, j, c;
  for (s = 0; s < 6; ++s) {
    for (j = -g; j <= N + g; ++j)
      for (i = -g; i <= N + g; ++i) {
        double a[3];
        uv(G, s, i, j, a);
        for (c = 0; c < 3; ++c) G->u[c][T2(G, s, i, j)] = a[c];
      }
    for (j = 1 - g; j <= N + g; ++j) /* x-faces and y-edges */
      for (i = -g; i <= N + g; ++i) {
        double a[3], b[3], n[3], m[3], t[3];
        int64_t q = T2(G, s, i, j);
        uv(G, s, i, j - 1, a);
        uv(G, s, i, j, b);
        cross3(b, a, n);
        unit3(n);
        for (c = 0; c < 3; ++c) m[c] = a[c] + b[c];
        unit3(m);
        for (c = 0; c < 3; ++c) t[c] = b[c] - a[c];
        perp_unit(t, m);
        for (c = 0; c < 3; ++c) { G->xn[c][q] = n[c]; G->xc[c][q] = m[c]; G->yet[c][q] = t[c]; }
        G->thy[q] = ang(a, b);
      }
    for (j = -g; j <= N + g; ++j) /* y-faces and x-edges */
      for (i = 1 - g; i <= N + g; ++i) {
        double a[3], b[3], n[3], m[3], t[3];
        int64_t q = T2(G, s, i, j);
        uv(G, s, i - 1, j, a);
        uv(G, s, i, j, b);
        cross3(a, b, n);
        unit3(n);
        for (c = 0; c < 3; ++c) m[c] = a[c] + b[c];
        unit3(m);
        for (c = 0; c < 3; ++c) t[c] = b[c] - a[c];
        perp_unit(t, m);
        for (c = 0; c < 3; ++c) { G->yn[c][q] = n[c]; G->yc[c][q] = m[c]; G->xet[c][q] = t[c]; }
        G->thx[q] = ang(a, b);
      }
    for (j = 1 - g; j <= N + g; ++j) /* zones */
      for (i = 1 - g; i <= N + g; ++i) {
        double a[3], b[3], cc[3], e[3], zs[3], sh[3], t1[3], zc3[3], ch = 1e300, dd[3];
        int64_t q = T2(G, s, i, j), qL = T2(G, s, i - 1, j), qD = T2(G, s, i, j - 1);
        const double* pr[4][2];
        int p;
        uv(G, s, i - 1, j - 1, a);
        uv(G, s, i, j - 1, b);
        uv(G, s, i, j, cc);
        uv(G, s, i - 1, j, e);
        for (c = 0; c < 3; ++c) zc3[c] = a[c] + b[c] + cc[c] + e[c];
        unit3(zc3);
        /* area vector of the spherical zone on the unit sphere, by closing
           the cone: sum over its four edges of (theta/2) * unit normal */
        for (c = 0; c < 3; ++c)
          zs[c] = 0.5 * (G->thx[qD] * G->yn[c][qD] - G->thy[q] * G->xn[c][q] - G->thx[q] * G->yn[c][q] +
                         G->thy[qL] * G->xn[c][qL]);
        G->zsn[q] = sqrt(dot3(zs, zs));
        for (c = 0; c < 3; ++c) { G->zs[c][q] = zs[c]; sh[c] = zs[c] / G->zsn[q]; }
        for (c = 0; c < 3; ++c) t1[c] = (b[c] + cc[c]) - (a[c] + e[c]);
        perp_unit(t1, sh);
        for (c = 0; c < 3; ++c) G->zt[c][q] = t1[c];
        G->om[q] = tri_om(a, b, cc) + tri_om(a, cc, e);
        pr[0][0] = e; pr[0][1] = cc; pr[1][0] = a; pr[1][1] = b; pr[2][0] = b; pr[2][1] = cc; pr[3][0] = a; pr[3][1] = e;
        for (p = 0; p < 4; ++p) {
          double l;
          for (c = 0; c < 3; ++c) dd[c] = pr[p][0][c] - pr[p][1][c];
          l = sqrt(dot3(dd, dd));
          if (l < ch) ch = l;
        }
        G->chmin[q] = ch;
        /* surface mode: inverse of the face-family mean normals */
        {
          double M[3][3], det;
        for (c = 0; c < 3; ++c) {
            M[0][c] = 0.5 * (G->xn[c][qL] + G->xn[c][q]);
            M[1][c] = 0.5 * (G->yn[c][qD] + G->yn[c][q]);
            M[2][c] = sh[c];
          }
          det = M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) - M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0]) +
                M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
          if (fabs(det) > 1e-300) {
            /* inverse by cofactors */
            int r, k2;
            for (r = 0; r < 3; ++r)
              for (k2 = 0; k2 < 3; ++k2) {
                int r1 = (k2 + 1) % 3, r2 = (k2 + 2) % 3, c1 = (r + 1) % 3, c2 = (r + 2) % 3;
                G->mi[3 * r + k2][q] = (M[r1][c1] * M[r2][c2] - M[r1][c2] * M[r2][c1]) / det;
              }
          }
        }
        { /* zone-B least squares, 485cleanup_after_update.f:740-800: rows are the
             unit normals of the x+, x-, y+, y- faces and the radial unit vector
             (both z faces); W_f = (A^T A)^-1 n_f */
          double rows[5][3], m3[3][3], bi[3][3], det, od;
          int r, f;
          for (c = 0; c < 3; ++c) {
            rows[0][c] = G->xn[c][q];
            rows[1][c] = G->xn[c][qL];
            rows[2][c] = G->yn[c][q];
            rows[3][c] = G->yn[c][qD];
            rows[4][c] = G->zsn[q] > 0.0 ? sh[c] : zc3[c];
          }
          for (r = 0; r < 3; ++r)
            for (c = 0; c < 3; ++c)
              m3[r][c] = rows[0][r] * rows[0][c] + rows[1][r] * rows[1][c] + rows[2][r] * rows[2][c] +
                         rows[3][r] * rows[3][c] + rows[4][r] * rows[4][c] + rows[4][r] * rows[4][c];
          det = m3[0][0] * (m3[1][1] * m3[2][2] - m3[2][1] * m3[1][2]) - m3[0][1] * (m3[1][0] * m3[2][2] - m3[2][0] * m3[1][2]) +
                m3[0][2] * (m3[1][0] * m3[2][1] - m3[2][0] * m3[1][1]);
          od = (det < 0.0 ? -1.0 : 1.0) / fmax(1e-32, fabs(det));
          bi[0][0] = (m3[1][1] * m3[2][2] - m3[1][2] * m3[2][1]) * od;
          bi[0][1] = (m3[0][2] * m3[2][1] - m3[0][1] * m3[2][2]) * od;
          bi[0][2] = (m3[0][1] * m3[1][2] - m3[1][1] * m3[0][2]) * od;
          bi[1][0] = (m3[1][2] * m3[2][0] - m3[1][0] * m3[2][2]) * od;
          bi[1][1] = (m3[0][0] * m3[2][2] - m3[0][2] * m3[2][0]) * od;
          bi[1][2] = (m3[0][2] * m3[1][0] - m3[0][0] * m3[1][2]) * od;
          bi[2][0] = (m3[1][0] * m3[2][1] - m3[1][1] * m3[2][0]) * od;
          bi[2][1] = (m3[0][1] * m3[2][0] - m3[0][0] * m3[2][1]) * od;
          bi[2][2] = (m3[0][0] * m3[1][1] - m3[0][1] * m3[1][0]) * od;
          for (f = 0; f < 5; ++f)
            for (r = 0; r < 3; ++r)
              G->bw[3 * f + r][q] = bi[r][0] * rows[f][0] + bi[r][1] * rows[f][1] + bi[r][2] * rows[f][2];
        }
      }
    for (j = 1 - g; j <= N + g - 1; ++j) /* z-edge frames, 460make_edge_fields_p1.f:2700-2760 */
      for (i = 1 - g; i <= N + g - 1; ++i) {
        double n3[3], n1[3], tau[3], p[3], m[3], ct;
        int64_t q = T2(G, s, i, j);
        uv(G, s, i, j, n3);
        uv(G, s, i + 1, j, p);
        uv(G, s, i - 1, j, m);
        for (c = 0; c < 3; ++c) n1[c] = p[c] - m[c];
        perp_unit(n1, n3);
        uv(G, s, i, j + 1, p);
        uv(G, s, i, j - 1, m);
        for (c = 0; c < 3; ++c) tau[c] = p[c] - m[c];
        perp_unit(tau, n3);
        ct = dot3(tau, n1);
        for (c = 0; c < 3; ++c) { G->en1[c][q] = n1[c]; G->etau[c][q] = tau[c]; }
        G->ecos[q] = ct;
        G->esin[q] = sqrt(1.0 - ct * ct);
      }
  }
}

/* --- geometric matching of ghost zones / faces to owners --------------- */
typedef struct { double p[3]; double n[3]; int fam, s, i, j; } gobj;

static void zone_center(const or_grid* G, int s, int i, int j, double* o) {
  double a[3], b[3], c3[3], d[3];
  int c;
  uv(G, s, i - 1, j - 1, a); uv(G, s, i, j - 1, b); uv(G, s, i, j, c3); uv(G, s, i - 1, j, d);
  for (c = 0; c < 3; ++c) o[c] = a[c] + b[c] + c3[c] + d[c];
  unit3(o);
}
/* face of family fam (0 x-face: vertex i, cell j; 1 y-face: cell i, vertex j):
   midpoint and unit normal oriented toward increasing index.  The panel's
   (x, y, r) index frame is right handed, so an x-face, whose edge runs along
   +y, has normal (edge) x r, and a y-face, whose edge runs along +x, has
   normal r x (edge); only the face's own two vertices are used. */
static void face_geo(const or_grid* G, int fam, int s, int i, int j, double* mid, double* nrm) {
  double a[3], b[3], e[3];
  int c;
  if (fam == 0) { uv(G, s, i, j - 1, a); uv(G, s, i, j, b); }
  else { uv(G, s, i - 1, j, a); uv(G, s, i, j, b); }
  for (c = 0; c < 3; ++c) { mid[c] = a[c] + b[c]; e[c] = b[c] - a[c]; }
  unit3(mid);
  if (fam == 0) cross3(e, mid, nrm);
  else cross3(mid, e, nrm);
  unit3(nrm);
}

static int nearest(const gobj* cand, int n, const double* p, int excl_s, double tol) {
  int best = -1, q;
  double bd = tol * tol;
  for (q = 0; q < n; ++q) {
    double dx, dy, dz, dd;
    if (cand[q].s == excl_s) continue;
    dx = cand[q].p[0] - p[0]; dy = cand[q].p[1] - p[1]; dz = cand[q].p[2] - p[2];
    dd = dx * dx + dy * dy + dz * dz;
    if (dd < bd) { bd = dd; best = q; }
  }
  return best;
}

/* sort candidates along x so the search is a window */
static int cmp_gx(const void* a, const void* b) {
  double x = ((const gobj*)a)->p[0], y = ((const gobj*)b)->p[0];
  return x < y ? -1 : (x > y ? 1 : 0);
}
static int nearest_sorted(const gobj* cand, int n, const double* p, int excl_s, double tol) {
  int lo = 0, hi = n, q, best = -1;
  double bd = tol * tol;
  while (lo < hi) { int m = (lo + hi) / 2; if (cand[m].p[0] < p[0] - tol) lo = m + 1; else hi = m; }
  for (q = lo; q < n && cand[q].p[0] <= p[0] + tol; ++q) {
    double dx, dy, dz, dd;
    if (cand[q].s == excl_s) continue;
    dx = cand[q].p[0] - p[0]; dy = cand[q].p[1] - p[1]; dz = cand[q].p[2] - p[2];
    dd = dx * dx + dy * dy + dz * dz;
    if (dd < bd || (dd == bd && best >= 0 && cand[q].s < cand[best].s)) { bd = dd; best = q; }
  }
  return best;
}

static void build_maps(or_grid* G) {
  int N = G->N, g = G->g, s, i, j, a, t, nz = 0, nfc = 0;
  double tol = 0.25 * PI_ / (2.0 * N);
  gobj* zones = malloc(sizeof(gobj) * 6 * N * N);
  gobj* faces = malloc(sizeof(gobj) * 6 * 2 * (N + 1) * N);
  (void)nearest;
  for (s = 0; s < 6; ++s)
    for (j = 1; j <= N; ++j)
      for (i = 1; i <= N; ++i) {
        gobj* o = &zones[nz++];
        zone_center(G, s, i, j, o->p);
        o->s = s; o->i = i; o->j = j; o->fam = 2;
      }
  for (s = 0; s < 6; ++s) {
    for (j = 1; j <= N; ++j)
      for (i = 0; i <= N; ++i) {
        gobj* o = &faces[nfc++];
        face_geo(G, 0, s, i, j, o->p, o->n);
        o->fam = 0; o->s = s; o->i = i; o->j = j;
      }
    for (j = 0; j <= N; ++j)
      for (i = 1; i <= N; ++i) {
        gobj* o = &faces[nfc++];
        face_geo(G, 1, s, i, j, o->p, o->n);
        o->fam = 1; o->s = s; o->i = i; o->j = j;
      }
  }
  qsort(zones, nz, sizeof(gobj), cmp_gx);
  qsort(faces, nfc, sizeof(gobj), cmp_gx);
  G->cmap = malloc(sizeof(cmap_t) * 6 * 4 * g * N);
  G->fmap = malloc(sizeof(fmap_t) * 6 * 4 * g * (3 * N + 2));
  G->pairs = malloc(sizeof(fmap_t) * 6 * 4 * N);
  G->ncmap = G->nfmap = G->npairs = 0;
  G->fails = 0;
  for (s = 0; s < 6; ++s)
    for (a = 1; a <= g; ++a)
      for (t = 1; t <= N; ++t) {
        int gi[4] = {N + a, 1 - a, t, t}, gj[4] = {t, t, N + a, 1 - a}, st;
        for (st = 0; st < 4; ++st) {
          double p[3];
          int q;
          zone_center(G, s, gi[st], gj[st], p);
          q = nearest_sorted(zones, nz, p, s, tol);
          if (q < 0) { G->fails++; continue; }
          {
            cmap_t* m = &G->cmap[G->ncmap++];
            m->s = s; m->i = gi[st]; m->j = gj[st]; m->ss = zones[q].s; m->si = zones[q].i; m->sj = zones[q].j;
          }
        }
      }
  /* ghost faces: normal family on t in [1,N], tangential family on t in [0,N] */
  for (s = 0; s < 6; ++s)
    for (a = 1; a <= g; ++a)
      for (t = 0; t <= N; ++t) {
        int fam[4][2] = {{0, 1}, {0, 1}, {1, 0}, {1, 0}};
        int ni[4] = {N + a, -a, t, t}, nj[4] = {t, t, N + a, -a};        /* normal family */
        int ti[4] = {N + a, 1 - a, t, t}, tj[4] = {t, t, N + a, 1 - a};  /* tangential family */
        int st;
        for (st = 0; st < 4; ++st) {
          int w;
          for (w = 0; w < 2; ++w) {
            int f = fam[st][w], fi = w == 0 ? ni[st] : ti[st], fj = w == 0 ? nj[st] : tj[st], q;
            double p[3], n[3];
            if (w == 0 && t == 0) continue;
            face_geo(G, f, s, fi, fj, p, n);
            q = nearest_sorted(faces, nfc, p, s, tol);
            if (q < 0) { G->fails++; continue; }
            {
              fmap_t* m = &G->fmap[G->nfmap++];
              m->fam = f; m->s = s; m->i = fi; m->j = fj;
              m->sfam = faces[q].fam; m->ss = faces[q].s; m->si = faces[q].i; m->sj = faces[q].j;
              m->sgn = dot3(n, faces[q].n) < 0 ? -1.0 : 1.0;
            }
          }
        }
      }
  /* coinciding boundary faces: each physical panel-edge face once */
  for (s = 0; s < 6; ++s)
    for (t = 1; t <= N; ++t) {
      int fi[4] = {0, N, t, t}, fj[4] = {t, t, 0, N}, ff[4] = {0, 0, 1, 1}, e;
      for (e = 0; e < 4; ++e) {
        double p[3], n[3];
        int q;
        face_geo(G, ff[e], s, fi[e], fj[e], p, n);
        q = nearest_sorted(faces, nfc, p, s, tol);
        if (q < 0) { G->fails++; continue; }
        if (faces[q].s < s) continue; /* the pair is recorded from the lower sector */
        {
          fmap_t* m = &G->pairs[G->npairs++];
          m->fam = ff[e]; m->s = s; m->i = fi[e]; m->j = fj[e];
          m->sfam = faces[q].fam; m->ss = faces[q].s; m->si = faces[q].i; m->sj = faces[q].j;
          m->sgn = dot3(n, faces[q].n) < 0 ? -1.0 : 1.0;
        }
      }
    }
  free(zones);
  free(faces);
}

/* Panel-edge edges shared by two sectors and the cube-corner vertices shared
   by three, found by position on the unit sphere (no index formulas).  An
   x-edge (i, j) runs from vertex (i-1, j) to (i, j), a y-edge (i, j) from
   (i, j-1) to (i, j); two copies of one edge that run in opposite directions
   pair with sign -1.  A radial edge lies on the panel-edge vertex it starts
   from and always points along +r. */
typedef struct { double a[3], b[3]; int fam, s, i, j; } eobj;
static void build_edge_maps(or_grid* G) {
  int N = G->N, s, t, e, q, nv = 0, ne = 0;
  double tol2 = 1e-20;
  eobj* v = malloc(sizeof(eobj) * 6 * 4 * (N + 1));
  eobj* ed = malloc(sizeof(eobj) * 6 * 4 * N);
  for (s = 0; s < 6; ++s)
    for (t = 0; t <= N; ++t) {
      int vi[4] = {0, N, t, t}, vj[4] = {t, t, 0, N};
      for (e = 0; e < 4; ++e) {
        if (e >= 2 && (t == 0 || t == N)) continue; /* corners listed once */
        v[nv].fam = 2; v[nv].s = s; v[nv].i = vi[e]; v[nv].j = vj[e];
        uv(G, s, vi[e], vj[e], v[nv].a);
        ++nv;
      }
    }
  for (s = 0; s < 6; ++s)
    for (t = 1; t <= N; ++t) {
      /* y-edges on i = 0, N; x-edges on j = 0, N */
      int fi[4] = {0, N, t, t}, fj[4] = {t, t, 0, N}, ff[4] = {1, 1, 0, 0};
      for (e = 0; e < 4; ++e) {
        eobj* o = &ed[ne++];
        o->fam = ff[e]; o->s = s; o->i = fi[e]; o->j = fj[e];
        if (ff[e] == 1) { uv(G, s, fi[e], fj[e] - 1, o->a); uv(G, s, fi[e], fj[e], o->b); }
        else { uv(G, s, fi[e] - 1, fj[e], o->a); uv(G, s, fi[e], fj[e], o->b); }
      }
    }
  G->epairs = malloc(sizeof(fmap_t) * (nv + ne));
  G->nepairs = 0;
  G->ncorner = 0;
  for (q = 0; q < nv; ++q) {
    int m[4], nm = 0, r;
    for (r = 0; r < nv; ++r) {
      double dx = v[r].a[0] - v[q].a[0], dy = v[r].a[1] - v[q].a[1], dz = v[r].a[2] - v[q].a[2];
      if (v[r].s != v[q].s && dx * dx + dy * dy + dz * dz < tol2 && nm < 4) m[nm++] = r;
    }
    if (nm == 1) {
      if (v[m[0]].s < v[q].s) continue;
      {
        fmap_t* p = &G->epairs[G->nepairs++];
        p->fam = 2; p->s = v[q].s; p->i = v[q].i; p->j = v[q].j;
        p->sfam = 2; p->ss = v[m[0]].s; p->si = v[m[0]].i; p->sj = v[m[0]].j; p->sgn = 1.0;
      }
    } else if (nm == 2) {
      /* record from the lowest sector of the three; v is in sector order */
      if (v[m[0]].s < v[q].s || v[m[1]].s < v[q].s || G->ncorner == 8) continue;
      {
        int (*c)[3] = G->corner[G->ncorner++];
        int lo = v[m[0]].s < v[m[1]].s ? 0 : 1;
        c[0][0] = v[q].s; c[0][1] = v[q].i; c[0][2] = v[q].j;
        c[1][0] = v[m[lo]].s; c[1][1] = v[m[lo]].i; c[1][2] = v[m[lo]].j;
        c[2][0] = v[m[1 - lo]].s; c[2][1] = v[m[1 - lo]].i; c[2][2] = v[m[1 - lo]].j;
      }
    } else {
      G->fails++;
    }
  }
  for (q = 0; q < ne; ++q) {
    int r, found = 0;
    for (r = 0; r < ne; ++r) {
      double d1 = 0, d2 = 0, d3 = 0, d4 = 0;
      int c;
      if (ed[r].s == ed[q].s) continue;
      for (c = 0; c < 3; ++c) {
        d1 += (ed[r].a[c] - ed[q].a[c]) * (ed[r].a[c] - ed[q].a[c]);
        d2 += (ed[r].b[c] - ed[q].b[c]) * (ed[r].b[c] - ed[q].b[c]);
        d3 += (ed[r].a[c] - ed[q].b[c]) * (ed[r].a[c] - ed[q].b[c]);
        d4 += (ed[r].b[c] - ed[q].a[c]) * (ed[r].b[c] - ed[q].a[c]);
      }
      if ((d1 < tol2 && d2 < tol2) || (d3 < tol2 && d4 < tol2)) {
        found = 1;
        if (ed[r].s < ed[q].s) break;
        {
          fmap_t* p = &G->epairs[G->nepairs++];
          p->fam = ed[q].fam; p->s = ed[q].s; p->i = ed[q].i; p->j = ed[q].j;
          p->sfam = ed[r].fam; p->ss = ed[r].s; p->si = ed[r].i; p->sj = ed[r].j;
          p->sgn = (d1 < tol2 && d2 < tol2) ? 1.0 : -1.0;
        }
        break;
      }
    }
    if (!found) G->fails++;
  }
  if (G->ncorner != 8) G->fails++;
  free(v);
  free(ed);
}

/* Shared-face fluxes and shared-edge EMFs made unique across sectors: the
   averages of the reference's defect synchronisation
   (334fetch_buffers_fill_in_flux_x.f:290-400, 335, 336;
   340synchronize_efz_quad_corner.f), taken before the update. */
static void sync_fluxes(const or_grid* G, double** F) {
  int m, k, q;
  for (m = 0; m < G->npairs; ++m) {
    const fmap_t* p = &G->pairs[m];
    for (k = 1; k <= G->NR; ++k)
      for (q = 0; q < 9; ++q) {
        double* pa = &F[p->fam][(p->s * 9 + q) * G->fpp + FACE(G, p->i, p->j, k)];
        double* pb = &F[p->sfam][(p->ss * 9 + q) * G->fpp + FACE(G, p->si, p->sj, k)];
        double va = *pa, vb = *pb;
        *pa = (va + p->sgn * vb) * 0.5;
        *pb = (vb + p->sgn * va) * 0.5;
      }
  }
}
static void sync_edges(const or_grid* G, double** E) {
  int m, k;
  for (m = 0; m < G->nepairs; ++m) {
    const fmap_t* p = &G->epairs[m];
    int k0 = p->fam == 2 ? 1 : 0;
    for (k = k0; k <= G->NR; ++k) {
      double* pa = &E[p->fam][p->s * G->fpp + FACE(G, p->i, p->j, k)];
      double* pb = &E[p->sfam][p->ss * G->fpp + FACE(G, p->si, p->sj, k)];
      double va = *pa, vb = *pb;
      *pa = (va + p->sgn * vb) * 0.5;
      *pb = (vb + p->sgn * va) * 0.5;
    }
  }
  for (m = 0; m < G->ncorner; ++m)
    for (k = 1; k <= G->NR; ++k) {
      const int(*c)[3] = G->corner[m];
      double* p0 = &E[2][c[0][0] * G->fpp + FACE(G, c[0][1], c[0][2], k)];
      double* p1 = &E[2][c[1][0] * G->fpp + FACE(G, c[1][1], c[1][2], k)];
      double* p2 = &E[2][c[2][0] * G->fpp + FACE(G, c[2][1], c[2][2], k)];
      double val = (*p0 + *p1 + *p2) * (1.0 / 3.0);
      *p0 = *p1 = *p2 = val;
    }
}

/* the two averages above, exported: F[fam] and E[fam] in the layouts of
 * the stage state, six sectors */
void or_sync_shared(const or_grid* G, double* const* F, double* const* E) {
  double* f[3] = {F[0], F[1], F[2]};
  double* e[3] = {E[0], E[1], E[2]};
  sync_fluxes(G, f);
  sync_edges(G, e);
}

or_grid* or_grid_new(const or_config* c) {
  or_grid* G;
  int q;
  int64_t n;
  if (c->ntp < 4 || c->nr < 1 || c->nghost < 3 || c->nghost >= c->ntp) return NULL;
  G = calloc(1, sizeof(or_grid));
  G->c = *c;
  G->N = c->ntp; G->NR = c->nr; G->g = c->nghost;
  G->nc = G->N + 2 * G->g; G->nrc = G->NR + 2 * G->g;
  G->nf = G->nc + 1; G->nrf = G->nrc + 1;
  G->cpp = (int64_t)G->nc * G->nc * G->nrc;
  G->fpp = (int64_t)G->nf * G->nf * G->nrf;
  G->t2 = (int64_t)G->nf * G->nf;
  G->lat = al(G->t2 * 18);
  or_unit_lattice(G->N, G->g, G->lat);
  corner_blocks(G->N, G->g, G->lat);
  G->rb = al(G->NR + 2 * G->g + 1);
  or_radial_mesh(c, G->rb);
  n = G->t2 * 6;
  for (q = 0; q < 3; ++q) {
    G->u[q] = al(n); G->xn[q] = al(n); G->xc[q] = al(n); G->yn[q] = al(n); G->yc[q] = al(n);
    G->zs[q] = al(n); G->zt[q] = al(n); G->en1[q] = al(n); G->etau[q] = al(n); G->xet[q] = al(n); G->yet[q] = al(n);
  }
  G->thy = al(n); G->thx = al(n); G->zsn = al(n); G->om = al(n); G->chmin = al(n);
  G->ecos = al(n); G->esin = al(n);
  for (q = 0; q < 9; ++q) G->mi[q] = al(n);
  for (q = 0; q < 15; ++q) G->bw[q] = al(n);
  build_geometry(G);
  build_maps(G);
  build_edge_maps(G);
  return G;
}

void or_grid_free(or_grid* G) {
  int q;
  if (!G) return;
  free(G->lat); free(G->rb);
  for (q = 0; q < 3; ++q) {
    free(G->u[q]); free(G->xn[q]); free(G->xc[q]); free(G->yn[q]); free(G->yc[q]);
    free(G->zs[q]); free(G->zt[q]); free(G->en1[q]); free(G->etau[q]); free(G->xet[q]); free(G->yet[q]);
  }
  free(G->thy); free(G->thx); free(G->zsn); free(G->om); free(G->chmin); free(G->ecos); free(G->esin);
  for (q = 0; q < 9; ++q) free(G->mi[q]);
  for (q = 0; q < 15; ++q) free(G->bw[q]);
  free(G->cmap); free(G->fmap); free(G->pairs); free(G->epairs);
  free(G);
}

int or_grid_dims(const or_grid* G, int64_t* o) {
  o[0] = G->cpp; o[1] = G->fpp; o[2] = G->t2; o[3] = G->ncmap + G->nfmap + G->npairs;
  return 0;
}
int or_map_failures(const or_grid* G) { return G->fails; }

static double R(const or_grid* G, int k) { return G->rb[k + G->g]; }
static double vol_of(const or_grid* G, int s, int i, int j, int k) {
  double r1 = R(G, k), r0 = R(G, k - 1);
  return G->om[T2(G, s, i, j)] * (r1 * r1 * r1 - r0 * r0 * r0) / 3.0;
}
static double xarea(const or_grid* G, int s, int i, int j, int k) {
  double r1 = R(G, k), r0 = R(G, k - 1);
  return 0.5 * (r1 * r1 - r0 * r0) * G->thy[T2(G, s, i, j)];
}
static double yarea(const or_grid* G, int s, int i, int j, int k) {
  double r1 = R(G, k), r0 = R(G, k - 1);
  return 0.5 * (r1 * r1 - r0 * r0) * G->thx[T2(G, s, i, j)];
}
static double zarea(const or_grid* G, int s, int i, int j, int k) {
  double r = R(G, k);
  return r * r * G->zsn[T2(G, s, i, j)];
}

void or_metric(const or_grid* G, double* vol, double* xa, double* ya, double* za) {
  int s, i, j, k, g = G->g;
  for (s = 0; s < 6; ++s) {
    if (vol)
      for (k = 1 - g; k <= G->NR + g; ++k)
        for (j = 1 - g; j <= G->N + g; ++j)
          for (i = 1 - g; i <= G->N + g; ++i) vol[s * G->cpp + CELL(G, i, j, k)] = vol_of(G, s, i, j, k);
    for (k = -g; k <= G->NR + g; ++k)
      for (j = -g; j <= G->N + g; ++j)
        for (i = -g; i <= G->N + g; ++i) {
          int64_t f = s * G->fpp + FACE(G, i, j, k);
          int kc = k > -g;
          if (xa) xa[f] = (kc && j > -g) ? xarea(G, s, i, j, k) : 0.0;
          if (ya) ya[f] = (kc && i > -g) ? yarea(G, s, i, j, k) : 0.0;
          if (za) za[f] = (i > -g && j > -g) ? zarea(G, s, i, j, k) : 0.0;
        }
  }
}

/* ------------------------------------------------------------- halo */
#define UU(U, G, s, q, c) (U)[((int64_t)(s) * 8 + (q)) * (G)->cpp + (c)]

void or_halo(const or_grid* G, double* U, double* ent, double* Bx, double* By, double* Bz) {
  int N = G->N, NR = G->NR, g = G->g, m, k, s, i, j, a, q;
  double* fam[2] = {Bx, By};
  /* 1) coinciding faces, averaged from the pre-exchange values
        (330fetch_buffers_fill_in_ux.f:337-350) */
  for (m = 0; m < G->npairs; ++m) {
    const fmap_t* p = &G->pairs[m];
    for (k = 1; k <= NR; ++k) {
      double* pa = &fam[p->fam][p->s * G->fpp + FACE(G, p->i, p->j, k)];
      double* pb = &fam[p->sfam][p->ss * G->fpp + FACE(G, p->si, p->sj, k)];
      double va = *pa, vb = *pb;
      *pa = (va + p->sgn * vb) * 0.5;
      *pb = (vb + p->sgn * va) * 0.5;
    }
  }
  /* 2) ghost zones and their z-faces */
  for (m = 0; m < G->ncmap; ++m) {
    const cmap_t* p = &G->cmap[m];
    for (k = 1; k <= NR; ++k) {
      int64_t d = CELL(G, p->i, p->j, k), sidx = CELL(G, p->si, p->sj, k);
      for (q = 0; q < 8; ++q) UU(U, G, p->s, q, d) = UU(U, G, p->ss, q, sidx);
      ent[p->s * G->cpp + d] = ent[p->ss * G->cpp + sidx];
    }
    for (k = 0; k <= NR; ++k)
      Bz[p->s * G->fpp + FACE(G, p->i, p->j, k)] = Bz[p->ss * G->fpp + FACE(G, p->si, p->sj, k)];
  }
  /* 3) ghost lateral faces */
  for (m = 0; m < G->nfmap; ++m) {
    const fmap_t* p = &G->fmap[m];
    for (k = 1; k <= NR; ++k)
      fam[p->fam][p->s * G->fpp + FACE(G, p->i, p->j, k)] =
          p->sgn < 0 ? -fam[p->sfam][p->ss * G->fpp + FACE(G, p->si, p->sj, k)]
                     : fam[p->sfam][p->ss * G->fpp + FACE(G, p->si, p->sj, k)];
  }
  /* 4) corner ghost blocks take the sector's corner zone */
  for (s = 0; s < 6; ++s)
    for (k = 1; k <= NR; ++k)
      for (a = 1; a <= g; ++a)
        for (j = 1; j <= g; ++j) {
          int cc;
          for (cc = 0; cc < 4; ++cc) {
            int ci = (cc & 1) ? N : 1, cj = (cc & 2) ? N : 1;
            int di = (cc & 1) ? N + a : 1 - a, dj = (cc & 2) ? N + j : 1 - j;
            int64_t d = CELL(G, di, dj, k), sidx = CELL(G, ci, cj, k);
            for (q = 0; q < 8; ++q) UU(U, G, s, q, d) = UU(U, G, s, q, sidx);
            ent[s * G->cpp + d] = ent[s * G->cpp + sidx];
          }
        }
  /* 5) radial ghosts: zero-gradient pad or reflecting conducting wall */
  for (s = 0; s < 6; ++s)
    for (a = 1; a <= g; ++a) {
      int side;
      for (side = 0; side < 2; ++side) {
        int kg = side == 0 ? 1 - a : NR + a, ks = side == 0 ? 1 : NR, sg = 1;
        int kzg = side == 0 ? -a : NR + a, kzs = side == 0 ? 0 : NR, zsg = 1;
        if (G->c.radial_bc == 1) {
          ks = kg;
          while (ks < 1 || ks > NR) { ks = ks < 1 ? 1 - ks : 2 * NR + 1 - ks; sg = -sg; }
          kzs = kzg;
          while (kzs < 0 || kzs > NR) { kzs = kzs < 0 ? -kzs : 2 * NR - kzs; zsg = -zsg; }
        }
        for (j = -g; j <= N + g; ++j)
          for (i = -g; i <= N + g; ++i) {
            if (j >= 1 - g) Bx[s * G->fpp + FACE(G, i, j, kg)] = Bx[s * G->fpp + FACE(G, i, j, ks)];
            if (i >= 1 - g) By[s * G->fpp + FACE(G, i, j, kg)] = By[s * G->fpp + FACE(G, i, j, ks)];
            if (i >= 1 - g && j >= 1 - g) {
              double u[8], bz = Bz[s * G->fpp + FACE(G, i, j, kzs)];
              int64_t d = CELL(G, i, j, kg), sidx = CELL(G, i, j, ks);
              Bz[s * G->fpp + FACE(G, i, j, kzg)] = zsg < 0 ? -bz : bz;
              for (q = 0; q < 8; ++q) u[q] = UU(U, G, s, q, sidx);
              if (sg < 0) {
                double n[3];
                double mm, bb;
                int64_t t = T2(G, s, i, j);
                if (G->zsn[t] > 1e-300 && G->zsn[t] < 1e300) {
                  for (q = 0; q < 3; ++q) n[q] = G->zs[q][t] / G->zsn[t];
                } else {
                  zone_center(G, s, i, j, n);
                }
                mm = u[1] * n[0] + u[2] * n[1] + u[3] * n[2];
                bb = u[5] * n[0] + u[6] * n[1] + u[7] * n[2];
                for (q = 0; q < 3; ++q) { u[1 + q] -= 2.0 * mm * n[q]; u[5 + q] -= 2.0 * bb * n[q]; }
              }
              for (q = 0; q < 8; ++q) UU(U, G, s, q, d) = u[q];
              ent[s * G->cpp + d] = ent[s * G->cpp + sidx];
            }
          }
      }
    }
}

/* ------------------------------------------------------------- pointwise */
/* PRIM_TO_CONS (661pde_3d_prim_to_cons.f:67-125) */
void or_prim_to_cons(double gam, double rhof, double prsf, double ent, double* V, double* U) {
  V[0] = V[0] > rhof ? V[0] : rhof;
  if (V[4] <= prsf) { double pe = ent * pow(V[0], gam - 1.0); V[4] = prsf > pe ? prsf : pe; }
  U[0] = V[0]; U[1] = V[0] * V[1]; U[2] = V[0] * V[2]; U[3] = V[0] * V[3];
  U[5] = V[5]; U[6] = V[6]; U[7] = V[7];
  U[4] = V[4] / (gam - 1.0) + 0.5 * V[0] * (V[1] * V[1] + V[2] * V[2] + V[3] * V[3]) +
         OB8PI * (V[5] * V[5] + V[6] * V[6] + V[7] * V[7]);
}
/* CONS_TO_PRIM (662pde_3d_cons_to_prim.f:69-151) */
int or_cons_to_prim(double gam, double rhof, double prsf, double* ent, double* U, double* V) {
  double a, V2, B2;
  int flag = 0;
  U[0] = U[0] > rhof ? U[0] : rhof;
  V[0] = U[0];
  a = V[0] / (V[0] * V[0] + rhof * rhof);
  V[1] = U[1] * a; V[2] = U[2] * a; V[3] = U[3] * a;
  V[5] = U[5]; V[6] = U[6]; V[7] = U[7];
  V2 = V[1] * V[1] + V[2] * V[2] + V[3] * V[3];
  B2 = V[5] * V[5] + V[6] * V[6] + V[7] * V[7];
  V[4] = (gam - 1.0) * (U[4] - 0.5 * V[0] * V2 - OB8PI * B2);
  if (V[4] <= prsf) {
    double pe = *ent * pow(V[0], gam - 1.0);
    V[4] = prsf > pe ? prsf : pe;
    flag = 1;
  }
  *ent = V[4] / pow(V[0], gam - 1.0);
  U[4] = V[4] / (gam - 1.0) + 0.5 * V[0] * V2 + OB8PI * B2;
  return flag;
}
/* PRIM_TO_WAVSPD, x (660pde_3d_prim_to_wavspd.f:54-76) */
void or_wavspd_x(double gam, const double* V, double* smin, double* smax) {
  double cs = sqrt(gam * V[4] / V[0]);
  double b2 = (V[5] * V[5] + V[6] * V[6] + V[7] * V[7]) / (4.0 * PI_ * V[0]);
  double bb = -(cs * cs + b2), cc = cs * cs * V[5] * V[5] / (4.0 * PI_ * V[0]);
  double disc = bb * bb - 4.0 * cc;
  double mf = sqrt((-bb + sqrt(disc > 0 ? disc : 0.0)) / 2.0);
  *smin = V[1] - mf;
  *smax = V[1] + mf;
}
/* x-flux; 600_1d_rs_oned_rs_mhd_dad.f:194-230 with B0 = 0 */
void or_flux_x(const double* V, const double* U, double* F) {
  double pt = V[4] + OB8PI * (V[5] * V[5] + V[6] * V[6] + V[7] * V[7]);
  double vb = V[1] * V[5] + V[2] * V[6] + V[3] * V[7];
  int c;
  for (c = 0; c < 8; ++c) F[c] = V[1] * U[c];
  F[1] += pt - OB4PI * V[5] * V[5];
  F[2] -= OB4PI * V[5] * V[6];
  F[3] -= OB4PI * V[5] * V[7];
  F[4] += pt * V[1] - OB4PI * V[5] * vb;
  F[5] = 0.0;
  F[6] -= V[5] * V[2];
  F[7] -= V[5] * V[3];
}
static void speeds3(double gam, const double* VL, const double* VR, double* SL, double* SR) {
  double a, b, Vm[8];
  int c;
  *SL = 1e30; *SR = -1e30;
  or_wavspd_x(gam, VL, &a, &b); if (a < *SL) *SL = a; if (b > *SR) *SR = b;
  or_wavspd_x(gam, VR, &a, &b); if (a < *SL) *SL = a; if (b > *SR) *SR = b;
  for (c = 0; c < 8; ++c) Vm[c] = 0.5 * (VL[c] + VR[c]);
  or_wavspd_x(gam, Vm, &a, &b); if (a < *SL) *SL = a; if (b > *SR) *SR = b;
}
/* OneD_RS_MHD_Ptwise (610_1d_rs_oned_rs_mhd_ptwise.f:96-220, 653-704) */
void or_hll(double gam, double rhof, double prsf, double* VL, double* VR, double entL, double entR,
            double* Fs, double* entF, double* SL, double* SR) {
  double UL[8], UR[8], FL[8], FR[8], Us[8], t1, t2, t3, eU, pe, pg, ob;
  int c;
  double bx = 0.5 * (VL[5] + VR[5]);
  VL[5] = bx; VR[5] = bx;
  or_prim_to_cons(gam, rhof, prsf, entL, VL, UL);
  or_prim_to_cons(gam, rhof, prsf, entR, VR, UR);
  or_flux_x(VL, UL, FL);
  or_flux_x(VR, UR, FR);
  speeds3(gam, VL, VR, SL, SR);
  if (*SL >= 0.0) { for (c = 0; c < 8; ++c) Fs[c] = FL[c]; *entF = entL * VL[1]; return; }
  if (*SR <= 0.0) { for (c = 0; c < 8; ++c) Fs[c] = FR[c]; *entF = entR * VR[1]; return; }
  t3 = 1.0 / (*SR - *SL); t1 = *SR * t3; t2 = *SL * t3;
  for (c = 0; c < 8; ++c) Us[c] = t1 * UR[c] - t2 * UL[c] - t3 * (FR[c] - FL[c]);
  eU = t1 * entR - t2 * entL - t3 * (entR * VR[1] - entL * VL[1]);
  pe = eU * pow(Us[0], gam - 1.0);
  pg = (gam - 1.0) * (Us[4] - 0.5 * (Us[1] * Us[1] + Us[2] * Us[2] + Us[3] * Us[3]) / Us[0] -
                      OB8PI * (Us[5] * Us[5] + Us[6] * Us[6] + Us[7] * Us[7]));
  if (pg <= prsf) Us[4] = Us[4] + (pe - pg) / (gam - 1.0);
  for (c = 0; c < 8; ++c) Fs[c] = 0.5 * (FL[c] + FR[c]) + 0.5 * *SL * (Us[c] - UL[c]) - 0.5 * *SR * (UR[c] - Us[c]);
  ob = 1.0 / (*SR - *SL);
  *entF = (*SR * entL * VL[1] - *SL * entR * VR[1]) * ob + (entR - entL) * (*SR * *SL * ob);
}
/* HLLD, Miyoshi & Kusano (2005), rationalised field b = B / sqrt(4 pi) */
void or_hlld(double gam, double rhof, double prsf, double* VL, double* VR, double entL, double entR,
             double* Fs, double* entF, double* SL, double* SR) {
  double UL[8], UR[8], FL[8], FR[8], s4 = sqrt(4.0 * PI_), F[8];
  double bx, bL[3], bR[3], ptL, ptR, dL, dR, SM, pts, rLs, rRs, vLs[2], vRs[2], bLs[2], bRs[2];
  double vbL, vbR, vbLs, vbRs, ELs, ERs, sqL, sqR, SLs, SRs, ULs[8], URs[8], ULr[8], URr[8], FLr[8], FRr[8];
  int c, side;
  double B = 0.5 * (VL[5] + VR[5]);
  VL[5] = B; VR[5] = B;
  or_prim_to_cons(gam, rhof, prsf, entL, VL, UL);
  or_prim_to_cons(gam, rhof, prsf, entR, VR, UR);
  or_flux_x(VL, UL, FL);
  or_flux_x(VR, UR, FR);
  speeds3(gam, VL, VR, SL, SR);
  if (*SL >= 0.0) { for (c = 0; c < 8; ++c) Fs[c] = FL[c]; *entF = entL * VL[1]; return; }
  if (*SR <= 0.0) { for (c = 0; c < 8; ++c) Fs[c] = FR[c]; *entF = entR * VR[1]; return; }
  bx = B / s4;
  for (c = 0; c < 3; ++c) { bL[c] = (c == 0 ? B : VL[5 + c]) / s4; bR[c] = (c == 0 ? B : VR[5 + c]) / s4; }
  ptL = VL[4] + 0.5 * (bL[0] * bL[0] + bL[1] * bL[1] + bL[2] * bL[2]);
  ptR = VR[4] + 0.5 * (bR[0] * bR[0] + bR[1] * bR[1] + bR[2] * bR[2]);
  dL = (*SL - VL[1]) * VL[0];
  dR = (*SR - VR[1]) * VR[0];
  SM = (dR * VR[1] - dL * VL[1] - ptR + ptL) / (dR - dL);
  { /* keep the contact strictly inside the fan, as rounding may not */
    double margin = (*SR - *SL) * 1e-12;
    if (SM < *SL + margin) SM = *SL + margin;
    if (SM > *SR - margin) SM = *SR - margin;
  }
  pts = (dR * ptL - dL * ptR + dL * dR * (VR[1] - VL[1])) / (dR - dL);
  rLs = dL / (*SL - SM);
  rRs = dR / (*SR - SM);
  for (side = 0; side < 2; ++side) {
    const double* V = side ? VR : VL;
    const double* bb = side ? bR : bL;
    double d = side ? dR : dL, S = side ? *SR : *SL;
    double* vs = side ? vRs : vLs;
    double* bs = side ? bRs : bLs;
    double den = d * (S - SM) - bx * bx;
    if (fabs(den) > 1e-10 * (fabs(d * (S - SM)) + bx * bx) && den != 0.0) {
      double f1 = bx * (SM - V[1]) / den, f2 = (d * (S - V[1]) - bx * bx) / den;
      vs[0] = V[2] - bb[1] * f1; vs[1] = V[3] - bb[2] * f1;
      bs[0] = bb[1] * f2; bs[1] = bb[2] * f2;
    } else {
      vs[0] = V[2]; vs[1] = V[3]; bs[0] = bb[1]; bs[1] = bb[2];
    }
  }
  vbL = VL[1] * bx + VL[2] * bL[1] + VL[3] * bL[2];
  vbR = VR[1] * bx + VR[2] * bR[1] + VR[3] * bR[2];
  vbLs = SM * bx + vLs[0] * bLs[0] + vLs[1] * bLs[1];
  vbRs = SM * bx + vRs[0] * bRs[0] + vRs[1] * bRs[1];
  ELs = ((*SL - VL[1]) * UL[4] - ptL * VL[1] + pts * SM + bx * (vbL - vbLs)) / (*SL - SM);
  ERs = ((*SR - VR[1]) * UR[4] - ptR * VR[1] + pts * SM + bx * (vbR - vbRs)) / (*SR - SM);
  sqL = sqrt(rLs); sqR = sqrt(rRs);
  SLs = SM - fabs(bx) / sqL;
  SRs = SM + fabs(bx) / sqR;
  ULs[0] = rLs; ULs[1] = rLs * SM; ULs[2] = rLs * vLs[0]; ULs[3] = rLs * vLs[1]; ULs[4] = ELs; ULs[5] = bx; ULs[6] = bLs[0]; ULs[7] = bLs[1];
  URs[0] = rRs; URs[1] = rRs * SM; URs[2] = rRs * vRs[0]; URs[3] = rRs * vRs[1]; URs[4] = ERs; URs[5] = bx; URs[6] = bRs[0]; URs[7] = bRs[1];
  for (c = 0; c < 5; ++c) { ULr[c] = UL[c]; URr[c] = UR[c]; FLr[c] = FL[c]; FRr[c] = FR[c]; }
  ULr[5] = bx; ULr[6] = bL[1]; ULr[7] = bL[2];
  URr[5] = bx; URr[6] = bR[1]; URr[7] = bR[2];
  FLr[5] = 0.0; FRr[5] = 0.0;
  FLr[6] = FL[6] / s4; FLr[7] = FL[7] / s4; FRr[6] = FR[6] / s4; FRr[7] = FR[7] / s4;
  if (SLs >= 0.0) {
    for (c = 0; c < 8; ++c) F[c] = FLr[c] + *SL * (ULs[c] - ULr[c]);
  } else if (SRs <= 0.0) {
    for (c = 0; c < 8; ++c) F[c] = FRr[c] + *SR * (URs[c] - URr[c]);
  } else {
    double sg = bx > 0 ? 1.0 : (bx < 0 ? -1.0 : 0.0), is = 1.0 / (sqL + sqR);
    double vy = (sqL * vLs[0] + sqR * vRs[0] + (bRs[0] - bLs[0]) * sg) * is;
    double vz = (sqL * vLs[1] + sqR * vRs[1] + (bRs[1] - bLs[1]) * sg) * is;
    double by = (sqL * bRs[0] + sqR * bLs[0] + sqL * sqR * (vRs[0] - vLs[0]) * sg) * is;
    double bz = (sqL * bRs[1] + sqR * bLs[1] + sqL * sqR * (vRs[1] - vLs[1]) * sg) * is;
    double vbss = SM * bx + vy * by + vz * bz, Uss[8];
    if (SM >= 0.0) {
      Uss[0] = rLs; Uss[1] = rLs * SM; Uss[2] = rLs * vy; Uss[3] = rLs * vz;
      Uss[4] = ELs - sqL * (vbLs - vbss) * sg; Uss[5] = bx; Uss[6] = by; Uss[7] = bz;
      for (c = 0; c < 8; ++c) F[c] = FLr[c] + *SL * (ULs[c] - ULr[c]) + SLs * (Uss[c] - ULs[c]);
    } else {
      Uss[0] = rRs; Uss[1] = rRs * SM; Uss[2] = rRs * vy; Uss[3] = rRs * vz;
      Uss[4] = ERs + sqR * (vbRs - vbss) * sg; Uss[5] = bx; Uss[6] = by; Uss[7] = bz;
      for (c = 0; c < 8; ++c) F[c] = FRr[c] + *SR * (URs[c] - URr[c]) + SRs * (Uss[c] - URs[c]);
    }
  }
  for (c = 0; c < 5; ++c) Fs[c] = F[c];
  Fs[5] = 0.0;
  Fs[6] = F[6] * s4;
  Fs[7] = F[7] * s4;
  *entF = Fs[0] >= 0.0 ? Fs[0] * (entL / VL[0]) : Fs[0] * (entR / VR[0]);
}
/* Jiang & Shu (1996) WENO5, left state at the right edge of v2 */
double or_weno5(double v0, double v1, double v2, double v3, double v4) {
  double e = 1e-6;
  double b0 = 13.0 / 12.0 * (v0 - 2 * v1 + v2) * (v0 - 2 * v1 + v2) + 0.25 * (v0 - 4 * v1 + 3 * v2) * (v0 - 4 * v1 + 3 * v2);
  double b1 = 13.0 / 12.0 * (v1 - 2 * v2 + v3) * (v1 - 2 * v2 + v3) + 0.25 * (v1 - v3) * (v1 - v3);
  double b2 = 13.0 / 12.0 * (v2 - 2 * v3 + v4) * (v2 - 2 * v3 + v4) + 0.25 * (3 * v2 - 4 * v3 + v4) * (3 * v2 - 4 * v3 + v4);
  double a0 = 0.1 / ((e + b0) * (e + b0)), a1 = 0.6 / ((e + b1) * (e + b1)), a2 = 0.3 / ((e + b2) * (e + b2));
  double q0 = (2 * v0 - 7 * v1 + 11 * v2) / 6.0, q1 = (-v1 + 5 * v2 + 2 * v3) / 6.0, q2 = (2 * v2 + 5 * v3 - v4) / 6.0;
  return (a0 * q0 + a1 * q1 + a2 * q2) / (a0 + a1 + a2);
}
/* Two_D_RS_Ez_4State_HLL (630_2d_rs_ez_two_d_rs_ez_4state_hll.f:10-372);
   s4 = (S_R,S_L,S_U,S_D), ez4 = (RU,LU,LD,RD), bn2 = (U,D), by2 = (R,L) */
double or_emf4_hll(const double* s4, double ct, double st, const double* ez, const double* bn, const double* by) {
  double SR = s4[0], SL = s4[1], SU = s4[2], SD = s4[3];
  double ERU = ez[0], ELU = ez[1], ELD = ez[2], ERD = ez[3], BnU = bn[0], BnD = bn[1], ByR = by[0], ByL = by[1];
  double ox = 1.0 / (SR - SL), oy = 1.0 / (SU - SD), csc = 1.0 / st, cot = ct / st;
  double BxRU = (BnU + ByR * ct) / st, BxLU = (BnU + ByL * ct) / st, BxRD = (BnD + ByR * ct) / st, BxLD = (BnD + ByL * ct) / st;
  double BtRU = BxRU * ct + ByR * st, BtLU = BxLU * ct + ByL * st, BtRD = BxRD * ct + ByR * st, BtLD = BxLD * ct + ByL * st;
  double BSxL, BSyL, ESzL, BSxR, BSyR, ESzR, BSn, BSt, BSxD, BSyD, ESzD, BSxU, BSyU, ESzU, xc, yc, BSSx, BSSy, E1, E2;
  if (SL >= 0.0 && SD >= 0.0) return ELD;
  if (SR <= 0.0 && SD >= 0.0) return ERD;
  if (SR <= 0.0 && SU <= 0.0) return ERU;
  if (SL >= 0.0 && SU <= 0.0) return ELU;
  if (SL >= 0.0) return (SU * ELD - SD * ELU) * oy + SU * SD * (BxLU - BxLD) * oy;
  if (SR <= 0.0) return (SU * ERD - SD * ERU) * oy + SU * SD * (BxRU - BxRD) * oy;
  if (SD >= 0.0) return (SR * ELD - SL * ERD) * ox - SR * SL * (BtRD - BtLD) * ox;
  if (SU <= 0.0) return (SR * ELU - SL * ERU) * ox - SR * SL * (BtRU - BtLU) * ox;
  BSxL = (SU * BxLU - SD * BxLD) * oy - (ELU - ELD) * oy; BSyL = ByL;
  ESzL = (SU * ELD - SD * ELU) * oy + SU * SD * (BxLU - BxLD) * oy;
  BSxR = (SU * BxRU - SD * BxRD) * oy - (ERU - ERD) * oy; BSyR = ByR;
  ESzR = (SU * ERD - SD * ERU) * oy + SU * SD * (BxRU - BxRD) * oy;
  BSn = BnD; BSt = (SR * BtRD - SL * BtLD) * ox + (ERD - ELD) * ox;
  BSxD = BSn * st + BSt * ct; BSyD = -BSn * ct + BSt * st;
  ESzD = (SR * ELD - SL * ERD) * ox - SR * SL * (BtRD - BtLD) * ox;
  BSn = BnU; BSt = (SR * BtRU - SL * BtLU) * ox + (ERU - ELU) * ox;
  BSxU = BSn * st + BSt * ct; BSyU = -BSn * ct + BSt * st;
  ESzU = (SR * ELU - SL * ERU) * ox - SR * SL * (BtRU - BtLU) * ox;
  xc = 0.5 * ((SU + SD) * ct + (SR + SL)) / st;
  yc = 0.5 * (SU + SD);
  BSSx = -0.5 * ((-ct * ESzR - SR * BSxR) * ox + (ESzU - SU * BSxU) * oy + (ct * ESzL + SL * BSxL) * ox + (-ESzD + SD * BSxD) * oy);
  BSSy = -0.5 * ((-st * ESzR - SR * BSyR) * ox + (-SU * BSyU) * oy + (st * ESzL + SL * BSyL) * ox + (SD * BSyD) * oy);
  E1 = -xc * BSSy - 0.5 * (csc * (-st * ESzR - SR * BSyR) + cot * (-SU * BSyU) - csc * (st * ESzL + SL * BSyL) - cot * (SD * BSyD));
  E2 = yc * BSSx + 0.5 * ((ESzU - SU * BSxU) - (-ESzD + SD * BSxD));
  return 0.5 * (E1 + E2);
}
/* Two_D_RS_Ez_3State_LLF (633_2d_rs_ez_two_d_rs_ez_3state_llf.f:59-84) */
double or_emf3_llf(double S, const double* ez, const double* bn) {
  double Sq3 = sqrt(3.0), ob3 = 1.0 / 3.0;
  return ob3 * (ez[0] + ez[1] + ez[2]) - (4.0 * Sq3 * S / 9.0) * (bn[0] + bn[1] + bn[2]);
}

/* ------------------------------------------------------------- stage */
static void prims_of(const or_grid* G, const double* U, const double* ent, int s, int64_t c, double* V) {
  double u[8], e = ent[s * G->cpp + c];
  int q;
  for (q = 0; q < 8; ++q) u[q] = UU(U, G, s, q, c);
  or_cons_to_prim(G->c.gamma, G->c.rhofloor, G->c.prsfloor, &e, u, V);
  V[8] = e;
}

/* reconstruction + face Riemann solve for one face (450make_facial_fluxes.f:507-1152) */
static void one_face(const or_grid* G, int dir, int s, int i, int j, int k, const double* Vc, int64_t stride,
                     int64_t cl, double bface, double* out, double* SL, double* SR) {
  double VL[9], VR[9], A[3], Bv[3], Cv[3], area, wl[8], wr[8], Fs[8], eF;
  int q, c, N = G->N;
  int64_t t = T2(G, s, i, j);
  /* faces of the ghost rows/columns only feed EMF speeds: first-order states */
  int weno = G->c.recon == 1 && (dir == 0 ? (j >= 1 && j <= N)
                                 : dir == 1 ? (i >= 1 && i <= N) : (i >= 1 && i <= N && j >= 1 && j <= N));
  for (q = 0; q < 9; ++q) {
    const double* p = Vc + q * G->cpp + cl;
    if (weno) {
      VL[q] = or_weno5(p[-2 * stride], p[-stride], p[0], p[stride], p[2 * stride]);
      VR[q] = or_weno5(p[3 * stride], p[2 * stride], p[stride], p[0], p[-stride]);
    } else {
      VL[q] = p[0];
      VR[q] = p[stride];
    }
  }
  if (!(VL[0] > 0) || !(VL[4] > 0) || !(VL[8] > 0)) for (q = 0; q < 9; ++q) VL[q] = Vc[q * G->cpp + cl];
  if (!(VR[0] > 0) || !(VR[4] > 0) || !(VR[8] > 0)) for (q = 0; q < 9; ++q) VR[q] = Vc[q * G->cpp + cl + stride];
  if (dir == 0) {
    for (c = 0; c < 3; ++c) { A[c] = G->xn[c][t]; Cv[c] = G->xc[c][t]; }
    cross3(Cv, A, Bv);
    area = xarea(G, s, i, j, k);
  } else if (dir == 1) {
    for (c = 0; c < 3; ++c) { A[c] = G->yn[c][t]; Cv[c] = G->yc[c][t]; }
    cross3(Cv, A, Bv);
    area = yarea(G, s, i, j, k);
  } else {
    for (c = 0; c < 3; ++c) { A[c] = G->zs[c][t] / G->zsn[t]; Bv[c] = G->zt[c][t]; }
    cross3(A, Bv, Cv);
    area = zarea(G, s, i, j, k);
  }
  wl[0] = VL[0]; wr[0] = VR[0]; wl[4] = VL[4]; wr[4] = VR[4];
  wl[1] = dot3(A, VL + 1); wl[2] = dot3(Bv, VL + 1); wl[3] = dot3(Cv, VL + 1);
  wr[1] = dot3(A, VR + 1); wr[2] = dot3(Bv, VR + 1); wr[3] = dot3(Cv, VR + 1);
  wl[5] = bface; wr[5] = bface;
  wl[6] = dot3(Bv, VL + 5); wl[7] = dot3(Cv, VL + 5);
  wr[6] = dot3(Bv, VR + 5); wr[7] = dot3(Cv, VR + 5);
  if (G->c.riemann == 1) or_hlld(G->c.gamma, G->c.rhofloor, G->c.prsfloor, wl, wr, VL[8], VR[8], Fs, &eF, SL, SR);
  else or_hll(G->c.gamma, G->c.rhofloor, G->c.prsfloor, wl, wr, VL[8], VR[8], Fs, &eF, SL, SR);
  out[0] = Fs[0]; out[4] = Fs[4];
  for (c = 0; c < 3; ++c) {
    out[1 + c] = A[c] * Fs[1] + Bv[c] * Fs[2] + Cv[c] * Fs[3];
    out[5 + c] = A[c] * Fs[5] + Bv[c] * Fs[6] + Cv[c] * Fs[7];
  }
  out[8] = eF;
  for (q = 0; q < 9; ++q) out[q] *= area;
}

static double* prims_all(const or_grid* G, const double* U, const double* ent) {
  double* V = malloc(sizeof(double) * 9 * 6 * G->cpp);
  int s, i, j, k, q, g = G->g;
  for (s = 0; s < 6; ++s)
    for (k = 1 - g; k <= G->NR + g; ++k)
      for (j = 1 - g; j <= G->N + g; ++j)
        for (i = 1 - g; i <= G->N + g; ++i) {
          int64_t c = CELL(G, i, j, k);
          double v[9];
          prims_of(G, U, ent, s, c, v);
          for (q = 0; q < 9; ++q) V[((int64_t)s * 9 + q) * G->cpp + c] = v[q];
        }
  return V;
}

static void fluxes_from_V(const or_grid* G, const double* V, const double* Bx, const double* By, const double* Bz,
                          double** F, double** S) {
  int N = G->N, NR = G->NR, dir, s, i, j, k, q;
  /* conducting walls: lateral faces of the radial ghost layers only fed the
     (zero) wall EMFs, so they are not computed */
  int wall = G->c.radial_bc == 1;
  int lo[3][3] = {{0, 0, wall}, {0, 0, wall}, {0, 0, 0}};
  int hi[3][3] = {{N, N + 1, wall ? NR : NR + 1}, {N + 1, N, wall ? NR : NR + 1}, {N + 1, N + 1, NR}};
  int64_t strides[3] = {1, G->nc, (int64_t)G->nc * G->nc};
  const double* fb[3] = {Bx, By, Bz};
  for (dir = 0; dir < 3; ++dir)
    for (s = 0; s < 6; ++s)
      for (k = lo[dir][2]; k <= hi[dir][2]; ++k)
        for (j = lo[dir][1]; j <= hi[dir][1]; ++j)
          for (i = lo[dir][0]; i <= hi[dir][0]; ++i) {
            double out[9], sl, sr;
            int64_t f = FACE(G, i, j, k);
            if (dir == 2 && (i == 0 || i == N + 1) && (j == 0 || j == N + 1)) continue; /* zero-area tri-corner column */
            one_face(G, dir, s, i, j, k, V + (int64_t)s * 9 * G->cpp, strides[dir], CELL(G, i, j, k),
                     fb[dir][s * G->fpp + f], out, &sl, &sr);
            for (q = 0; q < 9; ++q) F[dir][((int64_t)s * 9 + q) * G->fpp + f] = out[q];
            S[dir][((int64_t)s * 2) * G->fpp + f] = sl;
            S[dir][((int64_t)s * 2 + 1) * G->fpp + f] = sr;
          }
}

void or_fluxes(const or_grid* G, const double* U, const double* ent, const double* Bx, const double* By,
               const double* Bz, double** F, double** S) {
  double* V = prims_all(G, U, ent);
  fluxes_from_V(G, V, Bx, By, Bz, F, S);
  sync_fluxes(G, F);
  free(V);
}

static void cellE(const or_grid* G, const double* V, int s, int64_t c, double* E) {
  const double* p = V + (int64_t)s * 9 * G->cpp + c;
  double vx = p[G->cpp], vy = p[2 * G->cpp], vz = p[3 * G->cpp], bx = p[5 * G->cpp], by = p[6 * G->cpp], bz = p[7 * G->cpp];
  E[0] = -(vy * bz - vz * by);
  E[1] = -(vz * bx - vx * bz);
  E[2] = -(vx * by - vy * bx);
}
static double spd(const or_grid* G, double* const* S, int dir, int s, int w, int i, int j, int k) {
  return S[dir][((int64_t)s * 2 + w) * G->fpp + FACE(G, i, j, k)];
}
static double emf_quad(const or_grid* G, const double* V, int s, const int64_t* cells, const double* n3, double S_R,
                       double S_L, double S_U, double S_D, double ct, double st, double BnU, double BnD, double ByR,
                       double ByL) {
  /* cells: LD, RD, RU, LU.  460make_edge_fields_p1.f:3530-3700 */
  double E[4][3], z[4], s4[4], ez4[4], bn2[2], by2[2];
  int q;
  for (q = 0; q < 4; ++q) cellE(G, V, s, cells[q], E[q]);
  if (G->c.emf == 0) {
    double ex = (E[0][0] + E[1][0] + E[2][0] + E[3][0]) * 0.25, ey = (E[0][1] + E[1][1] + E[2][1] + E[3][1]) * 0.25,
           ez = (E[0][2] + E[1][2] + E[2][2] + E[3][2]) * 0.25, m;
    z[0] = z[1] = z[2] = z[3] = ex * n3[0] + ey * n3[1] + ez * n3[2];
    m = fmax(fmax(fabs(S_R), fabs(S_L)), fmax(fabs(S_U), fabs(S_D)));
    S_R = m; S_L = -m; S_U = m; S_D = -m;
  } else {
    for (q = 0; q < 4; ++q) z[q] = dot3(E[q], n3);
  }
  s4[0] = S_R; s4[1] = S_L; s4[2] = S_U; s4[3] = S_D;
  ez4[0] = z[2]; ez4[1] = z[3]; ez4[2] = z[0]; ez4[3] = z[1];
  bn2[0] = BnU; bn2[1] = BnD; by2[0] = ByR; by2[1] = ByL;
  return or_emf4_hll(s4, ct, st, ez4, bn2, by2);
}

static void emf_from_V(const or_grid* G, const double* V, const double* Bx, const double* By, const double* Bz,
                       double* const* S, double** E) {
  int N = G->N, NR = G->NR, s, i, j, k;
  for (s = 0; s < 6; ++s) {
    const double *bx = Bx + s * G->fpp, *by = By + s * G->fpp, *bz = Bz + s * G->fpp;
    /* z-edges */
    for (k = 1; k <= NR; ++k)
      for (j = 0; j <= N; ++j)
        for (i = 0; i <= N; ++i) {
          int64_t t = T2(G, s, i, j);
          double n3[3] = {G->u[0][t], G->u[1][t], G->u[2][t]}, e, dr = R(G, k) - R(G, k - 1);
          int corner = (i == 0 || i == N) && (j == 0 || j == N);
          if (corner) {
            /* tri-corner, 460make_edge_fields_p1.f:3725-4540 and 633 */
            int c[3][2];
            double bn, Sx, Sy, E1[3], E2[3], E3[3], ez3[3], bn3[3] = {0, 0, 0}, em;
            int sxi, sxj, syi, syj;
            if (i == 0 && j == 0) {
              c[0][0] = 1; c[0][1] = 1; c[1][0] = 0; c[1][1] = 1; c[2][0] = 1; c[2][1] = 0;
              bn = -by[FACE(G, 1, 0, k)] + bx[FACE(G, 0, 1, k)] + by[FACE(G, 0, 0, k)];
              sxi = 0; sxj = 1; syi = 1; syj = 0;
            } else if (i == N && j == 0) {
              c[0][0] = N; c[0][1] = 0; c[1][0] = N; c[1][1] = 1; c[2][0] = N + 1; c[2][1] = 1;
              bn = bx[FACE(G, N, 1, k)] + by[FACE(G, N, 0, k)] - by[FACE(G, N + 1, 0, k)];
              sxi = N; sxj = 1; syi = N; syj = 0;
            } else if (i == N && j == N) {
              c[0][0] = N; c[0][1] = N; c[1][0] = N + 1; c[1][1] = N; c[2][0] = N; c[2][1] = N + 1;
              bn = by[FACE(G, N, N, k)] - bx[FACE(G, N, N, k)] - by[FACE(G, N + 1, N, k)];
              sxi = N; sxj = N; syi = N; syj = N;
            } else {
              c[0][0] = 0; c[0][1] = N; c[1][0] = 1; c[1][1] = N; c[2][0] = 1; c[2][1] = N + 1;
              bn = -by[FACE(G, 1, N, k)] - bx[FACE(G, 0, N, k)] + by[FACE(G, 0, N, k)];
              sxi = 0; sxj = N; syi = 1; syj = N;
            }
            Sx = fmax(fabs(spd(G, S, 0, s, 0, sxi, sxj, k)), fabs(spd(G, S, 0, s, 1, sxi, sxj, k)));
            Sy = fmax(fabs(spd(G, S, 1, s, 0, syi, syj, k)), fabs(spd(G, S, 1, s, 1, syi, syj, k)));
            cellE(G, V, s, CELL(G, c[0][0], c[0][1], k), E1);
            cellE(G, V, s, CELL(G, c[1][0], c[1][1], k), E2);
            cellE(G, V, s, CELL(G, c[2][0], c[2][1], k), E3);
            {
              double ob3 = 1.0 / 3.0, ex = (E1[0] + E2[0] + E3[0]) * ob3, ey = (E1[1] + E2[1] + E3[1]) * ob3,
                     ez = (E1[2] + E2[2] + E3[2]) * ob3;
              em = ex * n3[0] + ey * n3[1] + ez * n3[2];
            }
            ez3[0] = ez3[1] = ez3[2] = em;
            bn3[0] = bn;
            e = or_emf3_llf(fmax(Sx, Sy), ez3, bn3);
          } else {
            int64_t cells[4] = {CELL(G, i, j, k), CELL(G, i + 1, j, k), CELL(G, i + 1, j + 1, k), CELL(G, i, j + 1, k)};
            double SRv = fmax(spd(G, S, 0, s, 1, i, j + 1, k), spd(G, S, 0, s, 1, i, j, k));
            double SLv = fmin(spd(G, S, 0, s, 0, i, j + 1, k), spd(G, S, 0, s, 0, i, j, k));
            double SUv = fmax(spd(G, S, 1, s, 1, i + 1, j, k), spd(G, S, 1, s, 1, i, j, k));
            double SDv = fmin(spd(G, S, 1, s, 0, i + 1, j, k), spd(G, S, 1, s, 0, i, j, k));
            e = emf_quad(G, V, s, cells, n3, SRv, SLv, SUv, SDv, G->ecos[t], G->esin[t], bx[FACE(G, i, j + 1, k)],
                         bx[FACE(G, i, j, k)], by[FACE(G, i + 1, j, k)], by[FACE(G, i, j, k)]);
          }
          E[2][s * G->fpp + FACE(G, i, j, k)] = e * dr;
        }
    /* x-edges (460:393-1493) */
    for (k = 0; k <= NR; ++k)
      for (j = 0; j <= N; ++j)
        for (i = 1; i <= N; ++i) {
          int64_t t = T2(G, s, i, j), f = s * G->fpp + FACE(G, i, j, k);
          double n3[3] = {G->xet[0][t], G->xet[1][t], G->xet[2][t]};
          int64_t cells[4] = {CELL(G, i, j, k), CELL(G, i, j + 1, k), CELL(G, i, j + 1, k + 1), CELL(G, i, j, k + 1)};
          if (G->c.radial_bc == 1 && (k == 0 || k == NR)) { E[0][f] = 0.0; continue; }
          E[0][f] = emf_quad(G, V, s, cells, n3, fmax(spd(G, S, 1, s, 1, i, j, k + 1), spd(G, S, 1, s, 1, i, j, k)),
                             fmin(spd(G, S, 1, s, 0, i, j, k + 1), spd(G, S, 1, s, 0, i, j, k)),
                             fmax(spd(G, S, 2, s, 1, i, j + 1, k), spd(G, S, 2, s, 1, i, j, k)),
                             fmin(spd(G, S, 2, s, 0, i, j + 1, k), spd(G, S, 2, s, 0, i, j, k)), 0.0, 1.0,
                             by[FACE(G, i, j, k + 1)], by[FACE(G, i, j, k)], bz[FACE(G, i, j + 1, k)], bz[FACE(G, i, j, k)]) *
                    (R(G, k) * G->thx[t]);
        }
    /* y-edges (460:1497-2587) */
    for (k = 0; k <= NR; ++k)
      for (j = 1; j <= N; ++j)
        for (i = 0; i <= N; ++i) {
          int64_t t = T2(G, s, i, j), f = s * G->fpp + FACE(G, i, j, k);
          double n3[3] = {G->yet[0][t], G->yet[1][t], G->yet[2][t]};
          int64_t cells[4] = {CELL(G, i, j, k), CELL(G, i, j, k + 1), CELL(G, i + 1, j, k + 1), CELL(G, i + 1, j, k)};
          if (G->c.radial_bc == 1 && (k == 0 || k == NR)) { E[1][f] = 0.0; continue; }
          E[1][f] = emf_quad(G, V, s, cells, n3, fmax(spd(G, S, 2, s, 1, i + 1, j, k), spd(G, S, 2, s, 1, i, j, k)),
                             fmin(spd(G, S, 2, s, 0, i + 1, j, k), spd(G, S, 2, s, 0, i, j, k)),
                             fmax(spd(G, S, 0, s, 1, i, j, k), spd(G, S, 0, s, 1, i, j, k + 1)),
                             fmin(spd(G, S, 0, s, 0, i, j, k), spd(G, S, 0, s, 0, i, j, k + 1)), 0.0, 1.0,
                             bz[FACE(G, i + 1, j, k)], bz[FACE(G, i, j, k)], bx[FACE(G, i, j, k + 1)], bx[FACE(G, i, j, k)]) *
                    (R(G, k) * G->thy[t]);
        }
  }
}

void or_emf(const or_grid* G, const double* U, const double* ent, const double* Bx, const double* By, const double* Bz,
            double** S, double** E) {
  double* V = prims_all(G, U, ent);
  emf_from_V(G, V, Bx, By, Bz, S, E);
  sync_edges(G, E);
  free(V);
}

static void weights(int order, int irk, double* wa, double* wb) {
  if (order == 3) {
    if (irk == 1) { *wa = 0.0; *wb = 1.0; }
    else if (irk == 2) { *wa = 0.75; *wb = 0.25; }
    else { *wa = 1.0 / 3.0; *wb = 2.0 / 3.0; }
  } else {
    if (irk == 1) { *wa = 0.0; *wb = 1.0; } else { *wa = 0.5; *wb = 0.5; }
  }
}

/* UPDATE_IMEX_RK_STEP_P1 (484:483-497, 547-593, CT 646-908) and
   CLEANUP_AFTER_UPDATE (485cleanup_after_update.f:4) */
/* CLEANUP_AFTER_UPDATE (485cleanup_after_update.f:4): zone B replaced by the
   least-squares fit to the six face values over their unit normals
   (:692-860), the total energy corrected by the change of magnetic energy
   (:845-851), then CONS_TO_PRIM with its floors (662pde_3d_cons_to_prim.f:69-151);
   U and the entropy are stored back. */
static void cleanup(const or_grid* G, double* const* out) {
  int N = G->N, NR = G->NR, s, i, j, k, q, surf = G->c.radial_bc == 1 && G->NR == 1;
  for (s = 0; s < 6; ++s)
    for (k = 1; k <= NR; ++k)
      for (j = 1; j <= N; ++j)
        for (i = 1; i <= N; ++i) {
          int64_t c = CELL(G, i, j, k), t = T2(G, s, i, j);
          double b[6], bn[3], u[8], v[8], e;
          b[0] = out[2][s * G->fpp + FACE(G, i, j, k)];
          b[1] = out[2][s * G->fpp + FACE(G, i - 1, j, k)];
          b[2] = out[3][s * G->fpp + FACE(G, i, j, k)];
          b[3] = out[3][s * G->fpp + FACE(G, i, j - 1, k)];
          b[4] = out[4][s * G->fpp + FACE(G, i, j, k)];
          b[5] = out[4][s * G->fpp + FACE(G, i, j, k - 1)];
          for (q = 0; q < 8; ++q) u[q] = UU(out[0], G, s, q, c);
          if (surf) { /* face-family averages along their mean normals; energy kept */
            double m0 = 0.5 * (b[1] + b[0]), m1 = 0.5 * (b[3] + b[2]), m2 = 0.5 * (b[5] + b[4]);
            for (q = 0; q < 3; ++q) u[5 + q] = G->mi[3 * q][t] * m0 + G->mi[3 * q + 1][t] * m1 + G->mi[3 * q + 2][t] * m2;
          } else {
            for (q = 0; q < 3; ++q)
              bn[q] = G->bw[q][t] * b[0] + G->bw[3 + q][t] * b[1] + G->bw[6 + q][t] * b[2] + G->bw[9 + q][t] * b[3] +
                      G->bw[12 + q][t] * (b[4] + b[5]);
            /* the total energy takes the change of magnetic energy (:845-851) */
            u[4] = u[4] - (u[5] * u[5] + u[6] * u[6] + u[7] * u[7]) * OB8PI +
                   (bn[0] * bn[0] + bn[1] * bn[1] + bn[2] * bn[2]) * OB8PI;
            for (q = 0; q < 3; ++q) u[5 + q] = bn[q];
          }
          e = out[1][s * G->cpp + c];
          or_cons_to_prim(G->c.gamma, G->c.rhofloor, G->c.prsfloor, &e, u, v);
          for (q = 0; q < 8; ++q) UU(out[0], G, s, q, c) = u[q];
          out[1][s * G->cpp + c] = e;
        }
}

void or_cleanup(const or_grid* G, double* U, double* ent, double* Bx, double* By, double* Bz) {
  double* st[5] = {U, ent, Bx, By, Bz};
  cleanup(G, st);
}

/* Radial gravity at zone (s, i, j, k): gr = -grav / rc^2 along the unit
   normal n of the radial faces (470build_non_stiff_sources.f:150-250). */
static double gravity_at(const or_grid* G, int s, int i, int j, int k, double* n) {
  int64_t t = T2(G, s, i, j);
  double rc = 0.5 * (R(G, k - 1) + R(G, k));
  int q;
  for (q = 0; q < 3; ++q) n[q] = G->zs[q][t] / G->zsn[t];
  return -G->c.grav / (rc * rc);
}
#define IMEX_GI (1.0 - 0.70710678118654752440) /* 1 - 1/sqrt(2) */

/* IMEX-SSP2(2,2,2) (Pareschi & Russo 2005; explicit part Heun, the
   reference's RK2 of DRV:2661-2773) with the gravity source
   S(U) = (0, rho gr n, gr (rho v . n)) implicit.  u = (rho, rho v, E).
   First stage U1 = Un + gi dt S(U1): rho is unchanged and S is linear in
   rho v, so rho v and then E follow in closed form. */
void or_imex_pre(double dt, double gr, const double* n, double* u) {
  int q;
  for (q = 1; q < 4; ++q) u[q] += IMEX_GI * dt * u[0] * gr * n[q - 1];
  u[4] += IMEX_GI * dt * gr * (u[1] * n[0] + u[2] * n[1] + u[3] * n[2]);
}
/* stage 1 (in = U1, uz = Un): U2 = Un + dt L(U1) + (1-2gi) dt S(U1) + gi dt S(U2);
   stage 2 (in = U2, uz = Un): Un+1 = (Un + U2 + dt L(U2))/2 + dt/2 (2gi S(U1) + (1-gi) S(U2)),
   with S(U1) rebuilt from Un. */
void or_imex_combine(int stage, double dt, double gr, const double* n, const double* ui, const double* uz,
                     const double* L, double* u) {
  double gi = IMEX_GI;
  int q;
  if (stage == 1) {
    double s1m = ui[0] * gr, s1e = gr * (ui[1] * n[0] + ui[2] * n[1] + ui[3] * n[2]);
    u[0] = uz[0] + dt * L[0];
    for (q = 1; q < 4; ++q) u[q] = uz[q] + dt * L[q] + (1.0 - 2.0 * gi) * dt * s1m * n[q - 1];
    u[4] = uz[4] + dt * L[4] + (1.0 - 2.0 * gi) * dt * s1e;
    for (q = 1; q < 4; ++q) u[q] += gi * dt * u[0] * gr * n[q - 1];
    u[4] += gi * dt * gr * (u[1] * n[0] + u[2] * n[1] + u[3] * n[2]);
  } else {
    double s2m = ui[0] * gr, s2e = gr * (ui[1] * n[0] + ui[2] * n[1] + ui[3] * n[2]);
    double s1m = uz[0] * gr, s1e = gr * ((uz[1] * n[0] + uz[2] * n[1] + uz[3] * n[2]) + gi * dt * uz[0] * gr);
    for (q = 0; q < 5; ++q) u[q] = 0.5 * uz[q] + 0.5 * (ui[q] + dt * L[q]);
    for (q = 1; q < 4; ++q) u[q] += 0.5 * dt * (2.0 * gi * s1m + (1.0 - gi) * s2m) * n[q - 1];
    u[4] += 0.5 * dt * (2.0 * gi * s1e + (1.0 - gi) * s2e);
  }
}

/* Rotating frame (470build_non_stiff_sources.f:150-250): centrifugal
   -rho Om x (Om x r) and Coriolis -2 rho Om x v added to the momentum rows of
   L, their work v . S to the energy; r the zone position, v = m rho / (rho^2 + rho_f^2). */
void or_rotation_source(const double* om, double rhofloor, const double* r, const double* u, double* L) {
  double rho = u[0], a = rho / (rho * rho + rhofloor * rhofloor), v[3], oxr[3], S[3], cen, cor;
  int q;
  for (q = 0; q < 3; ++q) v[q] = u[1 + q] * a;
  for (q = 0; q < 3; ++q) oxr[q] = om[(q + 1) % 3] * r[(q + 2) % 3] - om[(q + 2) % 3] * r[(q + 1) % 3];
  for (q = 0; q < 3; ++q) {
    cen = -rho * (om[(q + 1) % 3] * oxr[(q + 2) % 3] - om[(q + 2) % 3] * oxr[(q + 1) % 3]);
    cor = -2.0 * rho * (om[(q + 1) % 3] * v[(q + 2) % 3] - om[(q + 2) % 3] * v[(q + 1) % 3]);
    S[q] = cen + cor;
    L[1 + q] += S[q];
  }
  L[4] += v[0] * S[0] + v[1] * S[1] + v[2] * S[2];
}

void or_update(const or_grid* G, int irk, double dt, const double* const* in, const double* const* u0,
               double* const* out, double** F, double** E) {
  int N = G->N, NR = G->NR, s, i, j, k, q;
  double wa, wb;
  weights(G->c.rk_order, irk, &wa, &wb);
  for (s = 0; s < 6; ++s)
    for (k = 1; k <= NR; ++k)
      for (j = 1; j <= N; ++j)
        for (i = 1; i <= N; ++i) {
          int64_t c = CELL(G, i, j, k);
          double ov = 1.0 / vol_of(G, s, i, j, k), L[9]; /* rho, rho v, E, B (advanced as the reference's dU_dt), entropy */
          int64_t fp = FACE(G, i, j, k), fxm = FACE(G, i - 1, j, k), fym = FACE(G, i, j - 1, k), fzm = FACE(G, i, j, k - 1);
          for (q = 0; q < 9; ++q) {
            int64_t o = ((int64_t)s * 9 + q) * G->fpp;
            L[q] = -ov * ((F[0][o + fp] - F[0][o + fxm]) + (F[1][o + fp] - F[1][o + fym]) + (F[2][o + fp] - F[2][o + fzm]));
          }
          if (G->c.omega[0] != 0.0 || G->c.omega[1] != 0.0 || G->c.omega[2] != 0.0) {
            double n[3], r[3], u[4], rc = 0.5 * (R(G, k - 1) + R(G, k));
            gravity_at(G, s, i, j, k, n);
            for (q = 0; q < 3; ++q) r[q] = rc * n[q];
            for (q = 0; q < 4; ++q) u[q] = UU(in[0], G, s, q, c);
            or_rotation_source(G->c.omega, G->c.rhofloor, r, u, L);
          }
          if (!G->c.imex) {
            if (G->c.grav != 0.0) { /* explicit gravity */
              double n[3], gr = gravity_at(G, s, i, j, k, n), rho, mr;
              rho = UU(in[0], G, s, 0, c);
              mr = UU(in[0], G, s, 1, c) * n[0] + UU(in[0], G, s, 2, c) * n[1] + UU(in[0], G, s, 3, c) * n[2];
              for (q = 0; q < 3; ++q) L[1 + q] += rho * gr * n[q];
              L[4] += mr * gr;
            }
            for (q = 0; q < 8; ++q) {
              double un = UU(in[0], G, s, q, c) + dt * L[q];
              UU(out[0], G, s, q, c) = wa != 0.0 ? wa * UU(u0[0], G, s, q, c) + wb * un : wb * un;
            }
            {
              double en = in[1][s * G->cpp + c] + dt * L[8];
              out[1][s * G->cpp + c] = wa != 0.0 ? wa * u0[1][s * G->cpp + c] + wb * en : wb * en;
            }
          } else {
            /* IMEX-SSP2(2,2,2), gravity implicit; the entropy has no source */
            double n[3], gr = gravity_at(G, s, i, j, k, n), ui[5], uz[5], u[5];
            for (q = 0; q < 5; ++q) { ui[q] = UU(in[0], G, s, q, c); uz[q] = UU(u0[0], G, s, q, c); }
            or_imex_combine(irk, dt, gr, n, ui, uz, L, u);
            if (irk == 1) out[1][s * G->cpp + c] = u0[1][s * G->cpp + c] + dt * L[8];
            else out[1][s * G->cpp + c] = 0.5 * u0[1][s * G->cpp + c] + 0.5 * (in[1][s * G->cpp + c] + dt * L[8]);
            for (q = 0; q < 5; ++q) UU(out[0], G, s, q, c) = u[q];
            for (q = 5; q < 8; ++q)
              UU(out[0], G, s, q, c) = irk == 1 ? UU(u0[0], G, s, q, c) + dt * L[q]
                                                : 0.5 * UU(u0[0], G, s, q, c) + 0.5 * (UU(in[0], G, s, q, c) + dt * L[q]);
          }
        }
  for (s = 0; s < 6; ++s) {
    const double *Ex = E[0] + s * G->fpp, *Ey = E[1] + s * G->fpp, *Ez = E[2] + s * G->fpp;
    for (k = 1; k <= NR; ++k)
      for (j = 1; j <= N; ++j)
        for (i = 0; i <= N; ++i) {
          int64_t f = FACE(G, i, j, k);
          double circ = Ez[f] - Ez[FACE(G, i, j - 1, k)] - Ey[f] + Ey[FACE(G, i, j, k - 1)];
          double bn = in[2][s * G->fpp + f] - dt * circ / xarea(G, s, i, j, k);
          out[2][s * G->fpp + f] = wa != 0.0 ? wa * u0[2][s * G->fpp + f] + wb * bn : wb * bn;
        }
    for (k = 1; k <= NR; ++k)
      for (j = 0; j <= N; ++j)
        for (i = 1; i <= N; ++i) {
          int64_t f = FACE(G, i, j, k);
          double circ = -Ez[f] + Ez[FACE(G, i - 1, j, k)] + Ex[f] - Ex[FACE(G, i, j, k - 1)];
          double bn = in[3][s * G->fpp + f] - dt * circ / yarea(G, s, i, j, k);
          out[3][s * G->fpp + f] = wa != 0.0 ? wa * u0[3][s * G->fpp + f] + wb * bn : wb * bn;
        }
    for (k = 0; k <= NR; ++k)
      for (j = 1; j <= N; ++j)
        for (i = 1; i <= N; ++i) {
          int64_t f = FACE(G, i, j, k);
          double circ = Ey[f] - Ey[FACE(G, i - 1, j, k)] - Ex[f] + Ex[FACE(G, i, j - 1, k)];
          double bn = in[4][s * G->fpp + f] - dt * circ / zarea(G, s, i, j, k);
          out[4][s * G->fpp + f] = wa != 0.0 ? wa * u0[4][s * G->fpp + f] + wb * bn : wb * bn;
        }
  }
  cleanup(G, out);
}

/* EVALUATE_TIMESTEP (490evaluate_timestep.f:118-278) */
double or_cfl(const or_grid* G, const double* U, const double* ent, double dt_prev) {
  int s, i, j, k;
  double m = 1e300;
  for (s = 0; s < 6; ++s)
    for (k = 1; k <= G->NR; ++k)
      for (j = 1; j <= G->N; ++j)
        for (i = 1; i <= G->N; ++i) {
          double v[9], b2, ms, a21, L, r1 = R(G, k), r0 = R(G, k - 1), d;
          prims_of(G, U, ent, s, CELL(G, i, j, k), v);
          b2 = v[5] * v[5] + v[6] * v[6] + v[7] * v[7];
          ms = sqrt((G->c.gamma * v[4] + OB4PI * b2) / v[0]);
          a21 = sqrt(v[1] * v[1] + v[2] * v[2] + v[3] * v[3]) + ms;
          L = fmin(r1 - r0, r0 * G->chmin[T2(G, s, i, j)]);
          d = G->c.cfl * L / a21;
          if (d < m) m = d;
        }
  return dt_prev > 0.0 ? fmin(1.1 * dt_prev, m) : m;
}

static double** alloc_set(const or_grid* G, int n, int64_t per) {
  double** p = malloc(sizeof(double*) * n);
  int q;
  for (q = 0; q < n; ++q) p[q] = calloc((size_t)per, sizeof(double));
  return p;
}
static void free_set(double** p, int n) {
  int q;
  for (q = 0; q < n; ++q) free(p[q]);
  free(p);
}

void or_stage(const or_grid* G, int irk, double dt, double* const* in, double* const* u0, double* const* out) {
  double **F = alloc_set(G, 3, 9 * 6 * G->fpp), **S = alloc_set(G, 3, 2 * 6 * G->fpp), **E = alloc_set(G, 3, 6 * G->fpp);
  double* V;
  or_halo(G, in[0], in[1], in[2], in[3], in[4]);
  V = prims_all(G, in[0], in[1]);
  fluxes_from_V(G, V, in[2], in[3], in[4], F, S);
  sync_fluxes(G, F);
  emf_from_V(G, V, in[2], in[3], in[4], S, E);
  sync_edges(G, E);
  free(V);
  or_update(G, irk, dt, (const double* const*)in, (const double* const*)u0, out, F, E);
  free_set(F, 3); free_set(S, 3); free_set(E, 3);
}

void or_advance(const or_grid* G, int nsteps, double* U, double* ent, double* Bx, double* By, double* Bz, double* dt) {
  int64_t cs = 6 * G->cpp, fs = 6 * G->fpp;
  double* A[5] = {U, ent, Bx, By, Bz};
  double* B[5];
  int n, q, irk;
  B[0] = malloc(sizeof(double) * cs * 8); B[1] = malloc(sizeof(double) * cs);
  for (q = 2; q < 5; ++q) B[q] = malloc(sizeof(double) * fs);
  memcpy(B[0], U, sizeof(double) * cs * 8); memcpy(B[1], ent, sizeof(double) * cs);
  for (q = 2; q < 5; ++q) memcpy(B[q], A[q], sizeof(double) * fs);
  for (n = 0; n < nsteps; ++n) {
    if (G->c.imex) { /* U1 = Un + gi dt S(U1), faces copied, then its cleanup */
      int s, i, j, k;
      for (q = 0; q < 5; ++q) memcpy(B[q], A[q], sizeof(double) * (q == 0 ? cs * 8 : (q == 1 ? cs : fs)));
      for (s = 0; s < 6; ++s)
        for (k = 1; k <= G->NR; ++k)
          for (j = 1; j <= G->N; ++j)
            for (i = 1; i <= G->N; ++i) {
              int64_t c = CELL(G, i, j, k);
              double nn[3], gr = gravity_at(G, s, i, j, k, nn), u[5];
              for (q = 0; q < 5; ++q) u[q] = UU(A[0], G, s, q, c);
              or_imex_pre(*dt, gr, nn, u);
              for (q = 0; q < 5; ++q) UU(B[0], G, s, q, c) = u[q];
            }
      cleanup(G, B);
    }
    for (irk = 1; irk <= G->c.rk_order; ++irk) {
      double* const* in = (irk == 1 && !G->c.imex) ? A : B;
      double* const* out = (irk == G->c.rk_order) ? A : B;
      or_stage(G, irk, *dt, in, A, out);
    }
    *dt = or_cfl(G, U, ent, *dt);
  }
  for (q = 0; q < 5; ++q) free(B[q]);
}

/* ------------------------------------------------------------- quadrature geometry
 * The quadratic isoparametric zone and face geometry of the reference's
 * P1 reconstruction, for the six sectors (one patch each).  Layouts are the
 * reference's per-patch Fortran arrays with the sector as the slowest index:
 *   vert  (0-g:N+g, 0-g:N+g, 0-g:nr+g, 3)          mask (1-g:N+g)^2 x (1-g:nr+g)
 *   zone arrays (1-g:N+g)^2 x (1-g:nr+g) x trailing dims
 *   x-faces (0-g:N+g, 1-g:N+g, 1-g:nr+g), y-faces (1-g:N+g, 0-g:N+g, 1-g:nr+g),
 *   z-faces (1-g:N+g, 1-g:N+g, 0-g:nr+g), each x trailing dims.
 * The element maps are the 20-node (zones) and 8-node (faces) serendipity
 * elements on [-1/2, 1/2]^d, written generically from their node tables. */
#define SMALLNUM 1.0e-12

typedef struct { int N, NR, g, nv, nvr, nc, nrc; size_t vs, zs, xs, ys, zfs; } geo_dims;
static geo_dims geo_dims_of(const or_config* c) {
  geo_dims d;
  d.N = c->ntp; d.NR = c->nr; d.g = c->nghost;
  d.nv = d.N + 2 * d.g + 1; d.nvr = d.NR + 2 * d.g + 1;
  d.nc = d.N + 2 * d.g; d.nrc = d.NR + 2 * d.g;
  d.vs = (size_t)d.nv * d.nv * d.nvr;          /* vertices per sector (one component) */
  d.zs = (size_t)d.nc * d.nc * d.nrc;          /* zones per sector */
  d.xs = (size_t)d.nv * d.nc * d.nrc;          /* x-faces per sector */
  d.ys = d.xs;
  d.zfs = (size_t)d.nc * d.nc * d.nvr;         /* z-faces per sector */
  return d;
}
static const double* vtx(const geo_dims* d, const double* vert, int s, int i, int j, int k, double* out) {
  size_t o = (size_t)(i + d->g) + (size_t)d->nv * ((j + d->g) + (size_t)d->nv * (k + d->g));
  int c;
  for (c = 0; c < 3; ++c) out[c] = vert[(size_t)(3 * s + c) * d->vs + o];
  return out;
}
static double norm3(const double* v) { return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }
/* midpoint of two vertices, optionally pushed onto the sphere of radius r */
static void midpoint(const double* a, const double* b, int project, double r, double* m) {
  int c;
  for (c = 0; c < 3; ++c) m[c] = 0.5 * (a[c] + b[c]);
  if (project) {
    double t = r / fmax(SMALLNUM, norm3(m));
    for (c = 0; c < 3; ++c) m[c] = m[c] * t;
  }
}

/* EXTRACT_VERT_COORDS_PATCH (210extract_vert_coords_patch.f:7, :285-343):
 * vert(i, j, k) = rb(k) u(i, j); a zone is inactive when one of its four
 * lateral vertices is the zero of a tri-corner ghost block. */
void or_vert_coords(const or_config* c, double* vert, int32_t* mask) {
  geo_dims d = geo_dims_of(c);
  int N = d.N, g = d.g, nl = N + 2 * g + 1, s, i, j, k, q;
  double* lat = malloc(sizeof(double) * (size_t)nl * nl * 18);
  double* rb = malloc(sizeof(double) * (size_t)d.nvr);
  or_unit_lattice(N, g, lat);
  or_radial_mesh(c, rb);
#define LAT(ii, jj, cc, ss) lat[(size_t)((ii) + g) + (size_t)nl * (((jj) + g) + (size_t)nl * ((cc) + 3 * (size_t)(ss)))]
  for (s = 0; s < 6; ++s) {
    for (k = -g; k <= d.NR + g; ++k)
      for (j = -g; j <= N + g; ++j)
        for (i = -g; i <= N + g; ++i)
          for (q = 0; q < 3; ++q)
            vert[(size_t)(3 * s + q) * d.vs + (size_t)(i + g) + (size_t)d.nv * ((j + g) + (size_t)d.nv * (k + g))] =
                rb[k + g] * LAT(i, j, q, s);
    for (j = 1 - g; j <= N + g; ++j)
      for (i = 1 - g; i <= N + g; ++i) {
        int on = 1, a;
        for (a = 0; a < 4; ++a) {
          int ii = i - (a & 1), jj = j - (a >> 1);
          double u[3] = {LAT(ii, jj, 0, s), LAT(ii, jj, 1, s), LAT(ii, jj, 2, s)};
          if (norm3(u) < SMALLNUM) on = 0;
        }
        for (k = 1 - g; k <= d.NR + g; ++k)
          mask[(size_t)s * d.zs + (size_t)(i - 1 + g) + (size_t)d.nc * ((j - 1 + g) + (size_t)d.nc * (k - 1 + g))] = on;
      }
  }
#undef LAT
  free(lat);
  free(rb);
}

/* 8-node serendipity quadrilateral on [-1/2,1/2]^2: corner nodes (sx, sy)
 * first (--, +-, ++, -+), then the edge nodes (0,-), (+,0), (0,+), (-,0) */
static const int kQ8[8][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1}, {0, -1}, {1, 0}, {0, 1}, {-1, 0}};
static void q8_basis(double x, double y, double* N8, double* dx, double* dy) {
  int n;
  for (n = 0; n < 8; ++n) {
    double sx = kQ8[n][0], sy = kQ8[n][1];
    if (n < 4) {
      double a = 0.5 + sx * x, b = 0.5 + sy * y, e = 2.0 * sx * x + 2.0 * sy * y - 1.0;
      N8[n] = a * b * e;
      dx[n] = sx * b * e + a * b * 2.0 * sx;
      dy[n] = sy * a * e + a * b * 2.0 * sy;
    } else if (sx == 0) {  /* edge node along x at y = sy / 2 */
      double b = 0.5 + sy * y;
      N8[n] = (1.0 - 4.0 * x * x) * b;
      dx[n] = -8.0 * x * b;
      dy[n] = (1.0 - 4.0 * x * x) * sy;
    } else {               /* edge node along y at x = sx / 2 */
      double a = 0.5 + sx * x;
      N8[n] = a * (1.0 - 4.0 * y * y);
      dx[n] = sx * (1.0 - 4.0 * y * y);
      dy[n] = -8.0 * y * a;
    }
  }
}
/* 5-point face rule (220coords_to_wt_quadpts_faces_iso2.f:212-226): centre 1/3, edge midpoints 1/6 */
static const double kFq[5][3] = {{0.0, 0.0, 1.0 / 3.0}, {0.5, 0.0, 1.0 / 6.0}, {-0.5, 0.0, 1.0 / 6.0},
                                 {0.0, 0.5, 1.0 / 6.0}, {0.0, -0.5, 1.0 / 6.0}};

/* One face from its eight nodes: quadrature points, weights / area, normal(s),
 * area, centroid.  nall: normals at all five points (curved z-faces). */
static void face_quad(double X[8][3], double* pts, double* wts, double* nrm, double* area, double* cen,
                      int nall, size_t st, int bnorm) {
  int p, n, c;
  double C[5][3], a = 0.0;
  for (p = 0; p < 5; ++p) {
    double N8[8], dx[8], dy[8], A[3], B[3], cm;
    q8_basis(kFq[p][0], kFq[p][1], N8, dx, dy);
    for (c = 0; c < 3; ++c) {
      double x = 0.0, ax = 0.0, bx = 0.0;
      for (n = 0; n < 8; ++n) { x = x + X[n][c] * N8[n]; ax = ax + X[n][c] * dx[n]; bx = bx + X[n][c] * dy[n]; }
      pts[(c + 3 * p) * st] = x;
      A[c] = ax;
      B[c] = bx;
    }
    C[p][0] = A[1] * B[2] - A[2] * B[1];
    C[p][1] = A[2] * B[0] - A[0] * B[2];
    C[p][2] = A[0] * B[1] - A[1] * B[0];
    cm = sqrt(C[p][0] * C[p][0] + C[p][1] * C[p][1] + C[p][2] * C[p][2]);
    wts[p * st] = kFq[p][2] * cm;
    if (p == 0 || nall) {
      double inv = 1.0 / fmax(cm, SMALLNUM);
      for (c = 0; c < 3; ++c) nrm[(c + 3 * (nall ? p : 0)) * st] = inv * C[p][c];
    }
  }
  for (p = 0; p < 5; ++p) a = a + wts[p * st];
  if (bnorm) {  /* 230coords_to_wt_quadpts_bfld_iso2.f: weights times the reciprocal sum */
    double r = 1.0 / fmax(a, SMALLNUM);
    for (p = 0; p < 5; ++p) wts[p * st] = r * wts[p * st];
    return;
  }
  a = fmax(SMALLNUM, a);
  *area = a;
  for (p = 0; p < 5; ++p) wts[p * st] = wts[p * st] / a;
  for (c = 0; c < 3; ++c) {
    double s = 0.0;
    for (p = 0; p < 5; ++p) s = s + wts[p * st] * pts[(c + 3 * p) * st];
    cen[c * st] = s;
  }
}

/* COORDS_TO_WT_QUADPTS_FACES_ISO2 (220coords_to_wt_quadpts_faces_iso2.f:7):
 * x-faces (:178-348), y-faces (:426-599), z-faces (:677-878, centroid
 * projected back onto the sphere), scale length (:2212-2290).
 * xf/yf/zf = {pts, wts, normal, area, centroid}. */
void or_quad_faces(const or_config* c, const double* vert, double* const* xf, double* const* yf,
                   double* const* zf, double* scl) {
  geo_dims d = geo_dims_of(c);
  int N = d.N, NR = d.NR, g = d.g, s, i, j, k, n, q;
  for (s = 0; s < 6; ++s) {
    /* x-faces: nodes v(i, j-1, k-1), v(i, j, k-1), v(i, j, k), v(i, j-1, k);
       the edges along j lie on the spheres of their radius, those along k are straight */
    for (k = 1 - g; k <= NR + g; ++k)
      for (j = 1 - g; j <= N + g; ++j)
        for (i = -g; i <= N + g; ++i) {
          double X[8][3];
          size_t o = (size_t)(i + g) + (size_t)d.nv * ((j - 1 + g) + (size_t)d.nc * (k - 1 + g));
          size_t S = d.xs, b = (size_t)s * S;
          vtx(&d, vert, s, i, j - 1, k - 1, X[0]); vtx(&d, vert, s, i, j, k - 1, X[1]);
          vtx(&d, vert, s, i, j, k, X[2]); vtx(&d, vert, s, i, j - 1, k, X[3]);
          midpoint(X[0], X[1], 1, norm3(X[0]), X[4]);
          midpoint(X[1], X[2], 0, 0.0, X[5]);
          midpoint(X[2], X[3], 1, norm3(X[2]), X[6]);
          midpoint(X[3], X[0], 0, 0.0, X[7]);
          face_quad(X, xf[0] + 15 * b + o, xf[1] + 5 * b + o, xf[2] + 3 * b + o, xf[3] + b + o,
                    xf[4] + 3 * b + o, 0, S, 0);
        }
    /* y-faces: nodes v(i-1, j, k-1), v(i-1, j, k), v(i, j, k), v(i, j, k-1) */
    for (k = 1 - g; k <= NR + g; ++k)
      for (j = -g; j <= N + g; ++j)
        for (i = 1 - g; i <= N + g; ++i) {
          double X[8][3];
          size_t o = (size_t)(i - 1 + g) + (size_t)d.nc * ((j + g) + (size_t)d.nv * (k - 1 + g));
          size_t S = d.ys, b = (size_t)s * S;
          vtx(&d, vert, s, i - 1, j, k - 1, X[0]); vtx(&d, vert, s, i - 1, j, k, X[1]);
          vtx(&d, vert, s, i, j, k, X[2]); vtx(&d, vert, s, i, j, k - 1, X[3]);
          midpoint(X[0], X[1], 0, 0.0, X[4]);
          midpoint(X[1], X[2], 1, norm3(X[2]), X[5]);
          midpoint(X[2], X[3], 0, 0.0, X[6]);
          midpoint(X[3], X[0], 1, norm3(X[0]), X[7]);
          face_quad(X, yf[0] + 15 * b + o, yf[1] + 5 * b + o, yf[2] + 3 * b + o, yf[3] + b + o,
                    yf[4] + 3 * b + o, 0, S, 0);
        }
    /* z-faces: nodes v(i-1, j-1, k), v(i, j-1, k), v(i, j, k), v(i-1, j, k), all
       edges on the sphere; normals at every point; the centroid is put back
       on the sphere of the face */
    for (k = -g; k <= NR + g; ++k)
      for (j = 1 - g; j <= N + g; ++j)
        for (i = 1 - g; i <= N + g; ++i) {
          double X[8][3], r0, r1, r2;
          size_t o = (size_t)(i - 1 + g) + (size_t)d.nc * ((j - 1 + g) + (size_t)d.nc * (k + g));
          size_t S = d.zfs, b = (size_t)s * S;
          vtx(&d, vert, s, i - 1, j - 1, k, X[0]); vtx(&d, vert, s, i, j - 1, k, X[1]);
          vtx(&d, vert, s, i, j, k, X[2]); vtx(&d, vert, s, i - 1, j, k, X[3]);
          r0 = norm3(X[0]);
          for (n = 0; n < 4; ++n) midpoint(X[n], X[(n + 1) & 3], 1, r0, X[4 + n]);
          face_quad(X, zf[0] + 15 * b + o, zf[1] + 5 * b + o, zf[2] + 15 * b + o, zf[3] + b + o,
                    zf[4] + 3 * b + o, 1, S, 0);
          {
            double* cn = zf[4] + 3 * b + o;
            double cv[3] = {cn[0], cn[S], cn[2 * S]};
            r1 = fmax(SMALLNUM, norm3(cv));
            r2 = norm3(X[2]);
            for (q = 0; q < 3; ++q) cn[q * S] = r2 * cv[q] / r1;
          }
        }
    /* scale length: the radial edge, then the shorter lateral edges */
    for (k = 1 - g; k <= NR + g; ++k)
      for (j = 1 - g; j <= N + g; ++j)
        for (i = 1 - g; i <= N + g; ++i) {
          double a[3], b[3], e[4][3], L;
          vtx(&d, vert, s, i, j, k, a);
          vtx(&d, vert, s, i, j, k - 1, b);
          L = sqrt((a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) + (a[2] - b[2]) * (a[2] - b[2]));
          vtx(&d, vert, s, i - 1, j, k - 1, e[0]); vtx(&d, vert, s, i, j, k - 1, e[1]);
          vtx(&d, vert, s, i - 1, j - 1, k - 1, e[2]); vtx(&d, vert, s, i, j - 1, k - 1, e[3]);
          for (n = 0; n < 2; ++n) {
            double* p0 = e[2 * n], * p1 = e[2 * n + 1];
            double t = sqrt((p0[0] - p1[0]) * (p0[0] - p1[0]) + (p0[1] - p1[1]) * (p0[1] - p1[1]) +
                            (p0[2] - p1[2]) * (p0[2] - p1[2]));
            L = fmin(L, t);
          }
          vtx(&d, vert, s, i, j - 1, k - 1, e[0]); vtx(&d, vert, s, i, j, k - 1, e[1]);
          vtx(&d, vert, s, i - 1, j - 1, k - 1, e[2]); vtx(&d, vert, s, i - 1, j, k - 1, e[3]);
          for (n = 0; n < 2; ++n) {
            double* p0 = e[2 * n], * p1 = e[2 * n + 1];
            double t = sqrt((p0[0] - p1[0]) * (p0[0] - p1[0]) + (p0[1] - p1[1]) * (p0[1] - p1[1]) +
                            (p0[2] - p1[2]) * (p0[2] - p1[2]));
            L = fmin(L, t);
          }
          scl[(size_t)s * d.zs + (size_t)(i - 1 + g) + (size_t)d.nc * ((j - 1 + g) + (size_t)d.nc * (k - 1 + g))] = L;
        }
  }
}

/* 20-node serendipity hexahedron on [-1/2,1/2]^3: corners (x fastest within
 * a layer, lower layer first, counter-clockwise), then the edge nodes along x,
 * along y and along z (coordinate 0 on the edge's own axis) */
static const int kH20[20][3] = {
    {-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1}, {-1, -1, 1}, {1, -1, 1}, {1, 1, 1}, {-1, 1, 1},
    {0, -1, -1}, {0, 1, -1}, {0, 1, 1}, {0, -1, 1},
    {-1, 0, -1}, {-1, 0, 1}, {1, 0, 1}, {1, 0, -1},
    {-1, -1, 0}, {1, -1, 0}, {1, 1, 0}, {-1, 1, 0}};
static void h20_basis(const double* x, double* N, double (*D)[3]) {
  int n, a;
  for (n = 0; n < 20; ++n) {
    double sg[3] = {kH20[n][0], kH20[n][1], kH20[n][2]}, l[3], dl[3];
    int e = -1;
    for (a = 0; a < 3; ++a)
      if (sg[a] == 0.0) e = a;
    for (a = 0; a < 3; ++a) {
      if (a == e) { l[a] = 1.0 - 4.0 * x[a] * x[a]; dl[a] = -8.0 * x[a]; }
      else { l[a] = 0.5 + sg[a] * x[a]; dl[a] = sg[a]; }
    }
    if (e < 0) {
      double t = 2.0 * sg[0] * x[0] + 2.0 * sg[1] * x[1] + 2.0 * sg[2] * x[2] - 2.0;
      N[n] = l[0] * l[1] * l[2] * t;
      for (a = 0; a < 3; ++a) {
        double p = 1.0;
        int b;
        for (b = 0; b < 3; ++b) if (b != a) p *= l[b];
        D[n][a] = dl[a] * p * t + l[0] * l[1] * l[2] * 2.0 * sg[a];
      }
    } else {
      N[n] = l[0] * l[1] * l[2];
      for (a = 0; a < 3; ++a) {
        double p = 1.0;
        int b;
        for (b = 0; b < 3; ++b) if (b != a) p *= l[b];
        D[n][a] = dl[a] * p;
      }
    }
  }
}

/* COORDS_TO_WT_QUADPTS_VOLUME_ISO2 (240coords_to_wt_quadpts_volume_iso2.f:7)
 * at NUMORDER nq (2 or 3): nq^3 Gauss points, x fastest (:184-200), at
 * +-1/(2 sqrt 3) with linear weights 1/2 (:121-132), or at 0, +-sqrt(3/5)/2
 * with 8/18, 5/18 (:134-147); the tensor weight of point (ix, iy, iz) is
 * lin(iz) lin(iy) lin(ix), multiplied in that order (the weight table is
 * filled with x slowest, :169-181).  Outputs: quadrature points, weights
 * times |det J| normalised by the volume, the three tangent vectors
 * d x/d(xi, eta, zeta) at each point, volume, and the centroid (vertex mean
 * plus the weighted offsets). */
static void quad_volume_n(const or_config* c, int nq, const double* vert, double* pts, double* wts, double* tan,
                          double* vol, double* cen) {
  geo_dims d = geo_dims_of(c);
  int N = d.N, NR = d.NR, g = d.g, s, i, j, k, p, n, q;
  const int np = nq * nq * nq;
  double xg[3], lw[3], Nb[27][20], Db[27][20][3], wq[27];
  if (nq == 2) {
    xg[0] = -0.5 * sqrt(1.0 / 3.0);
    xg[1] = 0.5 * sqrt(1.0 / 3.0);
    lw[0] = lw[1] = 0.5;
  } else {
    xg[0] = -0.5 * sqrt(3.0 / 5.0);
    xg[1] = 0.0;
    xg[2] = 0.5 * sqrt(3.0 / 5.0);
    lw[0] = lw[2] = 5.0 / 18.0;
    lw[1] = 8.0 / 18.0;
  }
  for (p = 0; p < np; ++p) {
    int ix = p % nq, iy = (p / nq) % nq, iz = p / (nq * nq);
    double x[3] = {xg[ix], xg[iy], xg[iz]};
    h20_basis(x, Nb[p], Db[p]);
    wq[p] = lw[iz] * lw[iy] * lw[ix];
  }
  for (s = 0; s < 6; ++s)
    for (k = 1 - g; k <= NR + g; ++k)
      for (j = 1 - g; j <= N + g; ++j)
        for (i = 1 - g; i <= N + g; ++i) {
          double X[20][3], r1, r2, v = 0.0, c0[3], tmp[3] = {0.0, 0.0, 0.0};
          size_t S = d.zs, b = (size_t)s * S;
          size_t o = (size_t)(i - 1 + g) + (size_t)d.nc * ((j - 1 + g) + (size_t)d.nc * (k - 1 + g));
          for (n = 0; n < 8; ++n)
            vtx(&d, vert, s, i - 1 + ((n == 1 || n == 2 || n == 5 || n == 6) ? 1 : 0),
                j - 1 + ((n == 2 || n == 3 || n == 6 || n == 7) ? 1 : 0), k - 1 + (n >= 4 ? 1 : 0), X[n]);
          r1 = norm3(X[0]);
          r2 = norm3(X[4]);
          midpoint(X[0], X[1], 1, r1, X[8]);  midpoint(X[2], X[3], 1, r1, X[9]);
          midpoint(X[6], X[7], 1, r2, X[10]); midpoint(X[4], X[5], 1, r2, X[11]);
          midpoint(X[0], X[3], 1, r1, X[12]); midpoint(X[4], X[7], 1, r2, X[13]);
          midpoint(X[5], X[6], 1, r2, X[14]); midpoint(X[1], X[2], 1, r1, X[15]);
          midpoint(X[0], X[4], 0, 0.0, X[16]); midpoint(X[1], X[5], 0, 0.0, X[17]);
          midpoint(X[2], X[6], 0, 0.0, X[18]); midpoint(X[3], X[7], 0, 0.0, X[19]);
          for (p = 0; p < np; ++p) {
            double T[3][3], Dx, Dy, Dz, det;
            for (q = 0; q < 3; ++q) {
              double x = 0.0, t0 = 0.0, t1 = 0.0, t2 = 0.0;
              for (n = 0; n < 20; ++n) {
                x = x + X[n][q] * Nb[p][n];
                t0 = t0 + X[n][q] * Db[p][n][0];
                t1 = t1 + X[n][q] * Db[p][n][1];
                t2 = t2 + X[n][q] * Db[p][n][2];
              }
              pts[b * 3 * np + o + (q + 3 * p) * S] = x;
              T[0][q] = t0; T[1][q] = t1; T[2][q] = t2;
            }
            Dx = T[0][1] * T[1][2] - T[0][2] * T[1][1];
            Dy = T[0][2] * T[1][0] - T[0][0] * T[1][2];
            Dz = T[0][0] * T[1][1] - T[0][1] * T[1][0];
            det = fabs(T[2][0] * Dx + T[2][1] * Dy + T[2][2] * Dz);
            wts[b * np + o + p * S] = wq[p] * det;
            if (tan)
              for (q = 0; q < 3; ++q)
                for (n = 0; n < 3; ++n) tan[b * 9 * np + o + (q + 3 * (n + 3 * p)) * S] = T[n][q];
          }
          for (p = 0; p < np; ++p) v = v + wts[b * np + o + p * S];
          v = fmax(SMALLNUM, v);
          vol[b + o] = v;
          for (p = 0; p < np; ++p) wts[b * np + o + p * S] = wts[b * np + o + p * S] / v;
          for (q = 0; q < 3; ++q)
            c0[q] = 0.125 * (X[0][q] + X[1][q] + X[2][q] + X[3][q] + X[4][q] + X[5][q] + X[6][q] + X[7][q]);
          for (p = 0; p < np; ++p)
            for (q = 0; q < 3; ++q)
              tmp[q] = tmp[q] + wts[b * np + o + p * S] * (pts[b * 3 * np + o + (q + 3 * p) * S] - c0[q]);
          for (q = 0; q < 3; ++q) cen[b * 3 + o + q * S] = c0[q] + tmp[q];
        }
}
void or_quad_volume(const or_config* c, const double* vert, double* pts, double* wts, double* tan,
                    double* vol, double* cen) {
  quad_volume_n(c, 2, vert, pts, wts, tan, vol, cen);
}
void or_quad_volume3(const or_config* c, const double* vert, double* pts, double* wts, double* tan,
                     double* vol, double* cen) {
  quad_volume_n(c, 3, vert, pts, wts, tan, vol, cen);
}

/* BUILD_MODE_INDICES (204build_mode_indices.f:7) and INIT_BASIS_FUNCTION_COEFS
 * (250init_basis_function_coefs.f:5, :150-267): moments <x^a y^b z^c> of the
 * zone's np quadrature points about the centroid in units of the scale
 * length, summed over the points in order, mode 1 zero; inactive zones keep
 * zeros.  Second order: modes 1, x, y, z; third order adds x^2, y^2, z^2,
 * xy, yz, xz (204...f:55-95). */
static const int kModePow[10][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {2, 0, 0},
                                    {0, 2, 0}, {0, 0, 2}, {1, 1, 0}, {0, 1, 1}, {1, 0, 1}};
void or_mode_indices(int32_t* ii, int32_t* jj, int32_t* kk) {
  int m;
  for (m = 0; m < 4; ++m) { ii[m] = kModePow[m][0]; jj[m] = kModePow[m][1]; kk[m] = kModePow[m][2]; }
}
static double ipow(double x, int e) { return e == 0 ? 1.0 : (e == 1 ? x : x * x); }
static void basis_moments_n(const or_config* c, int np, int nm, const int32_t* mask, const double* pts,
                            const double* wts, const double* cen, const double* scl, double* coef) {
  geo_dims d = geo_dims_of(c);
  size_t S = d.zs, z, s;
  int m, p, q;
  memset(coef, 0, sizeof(double) * S * 6 * nm);
  for (s = 0; s < 6; ++s)
    for (z = 0; z < S; ++z) {
      size_t o = s * S + z;
      double aa;
      if (mask[o] != 1) continue;
      aa = 1.0 / fmax(SMALLNUM, fabs(scl[o]));
      for (m = 1; m < nm; ++m) {
        double t = 0.0;
        for (p = 0; p < np; ++p) {
          double x[3];
          for (q = 0; q < 3; ++q) x[q] = (pts[s * S * 3 * np + z + (q + 3 * p) * S] - cen[s * S * 3 + z + q * S]) * aa;
          const double v = ipow(x[0], kModePow[m][0]) * ipow(x[1], kModePow[m][1]) * ipow(x[2], kModePow[m][2]);
          t = t + wts[s * S * np + z + p * S] * v;
        }
        coef[s * S * nm + z + m * S] = coef[s * S * nm + z + m * S] + t;
      }
    }
}
void or_basis_moments(const or_config* c, const int32_t* mask, const double* pts, const double* wts,
                      const double* cen, const double* scl, double* coef) {
  basis_moments_n(c, 8, 4, mask, pts, wts, cen, scl, coef);
}
void or_basis_moments10(const or_config* c, const int32_t* mask, const double* pts, const double* wts,
                        const double* cen, const double* scl, double* coef) {
  basis_moments_n(c, 27, 10, mask, pts, wts, cen, scl, coef);
}

/* STENCIL_INDEXING_TRICORNERS (211stencil_indexing_tricorners.f:6) for a
 * sector that is one patch (all four corners are tri-corners):
 * st(1-g:N+g, 1-g:N+g, 2, 12), the (x, y) zone each of the twelve stencil
 * directions reaches (1 R, 2 L, 3 U, 4 D, 5 RR, 6 LL, 7 UU, 8 DD, 9 RU, 10 LU,
 * 11 LD, 12 RD), clamped to the patch, then re-pointed around the tri-corner
 * blocks (:84-271, each corner's assignments in the reference's order). */
void or_stencil_indexing(int N, int g, int32_t* st) {
  const int nc = N + 2 * g, lo = 1 - g, hi = N + g;
  int x, y;
#define SET(xx, yy, d, vx, vy)                                                            \
  do {                                                                                    \
    size_t o_ = (size_t)((xx) - lo) + (size_t)nc * ((size_t)((yy) - lo) + (size_t)nc * 2 * ((d) - 1)); \
    st[o_] = (vx);                                                                        \
    st[o_ + (size_t)nc * nc] = (vy);                                                      \
  } while (0)
#define MN(a, b) ((a) < (b) ? (a) : (b))
#define MX(a, b) ((a) > (b) ? (a) : (b))
  for (x = lo; x <= hi; ++x)
    for (y = lo; y <= hi; ++y) {
      int xp1 = MN(x + 1, hi), yp1 = MN(y + 1, hi), xp2 = MN(x + 2, hi), yp2 = MN(y + 2, hi);
      int xm1 = MX(x - 1, lo), ym1 = MX(y - 1, lo), xm2 = MX(x - 2, lo), ym2 = MX(y - 2, lo);
      SET(x, y, 1, xp1, y); SET(x, y, 2, xm1, y); SET(x, y, 3, x, yp1); SET(x, y, 4, x, ym1);
      SET(x, y, 5, xp2, y); SET(x, y, 6, xm2, y); SET(x, y, 7, x, yp2); SET(x, y, 8, x, ym2);
      SET(x, y, 9, xp1, yp1); SET(x, y, 10, xm1, yp1); SET(x, y, 11, xm1, ym1); SET(x, y, 12, xp1, ym1);
    }
  /* left-down (:84-122) */
  y = 1;
  for (x = lo; x <= 0; ++x) {
    SET(x, y, 4, 1, x); SET(x, y, 8, 2, x); SET(x, y, 11, 1, MX(x - 1, lo));
    if (x == 0) SET(x, y, 12, 2, x - 1); else SET(x, y, 12, 1, x + 1);
    SET(x, y + 1, 8, 1, x);
  }
  x = 1;
  for (y = lo; y <= 0; ++y) {
    SET(x, y, 2, y, 1); SET(x, y, 6, y, 2);
    if (y == 0) SET(x, y, 10, y - 1, 2); else SET(x, y, 10, y + 1, 1);
    SET(x, y, 11, MX(y - 1, lo), 1);
    SET(x + 1, y, 6, y, 1);
  }
  SET(1, 1, 11, 2, -1);
  /* right-down (:127-168) */
  y = 1;
  for (x = N + 1; x <= hi; ++x) {
    SET(x, y, 4, N, N - x + 1); SET(x, y, 8, N - 1, N - x + 1);
    if (x == N + 1) SET(x, y, 11, N - 1, N - x); else SET(x, y, 11, N, N - x + 2);
    SET(x, y, 12, N, MX(N - x, lo));
    SET(x, y + 1, 8, N, N - x + 1);
  }
  x = N;
  for (y = lo; y <= 0; ++y) {
    SET(x, y, 1, N + 1 - y, 1); SET(x, y, 5, N + 1 - y, 2);
    if (y == 0) SET(x, y, 9, N + 2 - y, 2); else SET(x, y, 9, N - y, 1);
    SET(x, y, 12, MN(N + 2 - y, hi), 1);
    SET(x - 1, y, 5, N + 1 - y, 1);
  }
  SET(N, 1, 12, N - 1, -1);
  /* right-up (:173-214) */
  y = N;
  for (x = N + 1; x <= hi; ++x) {
    SET(x, y, 3, N, x); SET(x, y, 7, N - 1, x);
    if (x == N + 1) SET(x, y, 10, N - 1, x + 1); else SET(x, y, 10, N, x - 1);
    SET(x, y, 9, N, MN(x + 1, hi));
    SET(x, y - 1, 7, N, x);
  }
  x = N;
  for (y = N + 1; y <= hi; ++y) {
    SET(x, y, 1, y, N); SET(x, y, 5, y, N - 1);
    if (y == N + 1) SET(x, y, 12, y + 1, N - 1); else SET(x, y, 12, y - 1, N);
    SET(x, y, 9, MN(y + 1, hi), N);
    SET(x - 1, y, 5, y, N);
  }
  SET(N, N, 9, N - 1, N + 2);
  /* left-up (:219-260) */
  x = 1;
  for (y = N + 1; y <= hi; ++y) {
    SET(x, y, 2, N - y + 1, N); SET(x, y, 6, N - y + 1, N - 1);
    if (y == N + 1) SET(x, y, 11, N - y, N - 1); else SET(x, y, 11, N - y + 2, N);
    SET(x, y, 10, MX(N - y, lo), N);
    SET(x + 1, y, 6, N - y + 1, N);
  }
  y = N;
  for (x = lo; x <= 0; ++x) {
    SET(x, y, 3, 1, N - x + 1); SET(x, y, 7, 2, N - x + 1);
    if (x == 0) SET(x, y, 9, 2, N - x + 2); else SET(x, y, 9, 1, N - x);
    SET(x, y, 10, 1, MN(N - x + 2, hi));
    SET(x, y - 1, 7, 1, N - x + 1);
  }
  SET(1, N, 10, 2, N + 2);
#undef SET
#undef MN
#undef MX
}

/* COORDS_TO_WT_QUADPTS_BFLD_ISO2 (230coords_to_wt_quadpts_bfld_iso2.f:8): the
 * face points and normals of 220 with the weights scaled by the reciprocal of
 * their sum; xb/yb/zb = {pts, wts, normal}. */
void or_quad_bfield(const or_config* c, const double* vert, double* const* xb, double* const* yb,
                    double* const* zb) {
  geo_dims d = geo_dims_of(c);
  int N = d.N, NR = d.NR, g = d.g, s, i, j, k, n;
  double area, cen[3];
  for (s = 0; s < 6; ++s) {
    for (k = 1 - g; k <= NR + g; ++k)
      for (j = 1 - g; j <= N + g; ++j)
        for (i = -g; i <= N + g; ++i) {
          double X[8][3];
          size_t o = (size_t)(i + g) + (size_t)d.nv * ((j - 1 + g) + (size_t)d.nc * (k - 1 + g));
          size_t S = d.xs, b = (size_t)s * S;
          vtx(&d, vert, s, i, j - 1, k - 1, X[0]); vtx(&d, vert, s, i, j, k - 1, X[1]);
          vtx(&d, vert, s, i, j, k, X[2]); vtx(&d, vert, s, i, j - 1, k, X[3]);
          midpoint(X[0], X[1], 1, norm3(X[0]), X[4]);
          midpoint(X[1], X[2], 0, 0.0, X[5]);
          midpoint(X[2], X[3], 1, norm3(X[2]), X[6]);
          midpoint(X[3], X[0], 0, 0.0, X[7]);
          face_quad(X, xb[0] + 15 * b + o, xb[1] + 5 * b + o, xb[2] + 3 * b + o, &area, cen, 0, S, 1);
        }
    for (k = 1 - g; k <= NR + g; ++k)
      for (j = -g; j <= N + g; ++j)
        for (i = 1 - g; i <= N + g; ++i) {
          double X[8][3];
          size_t o = (size_t)(i - 1 + g) + (size_t)d.nc * ((j + g) + (size_t)d.nv * (k - 1 + g));
          size_t S = d.ys, b = (size_t)s * S;
          vtx(&d, vert, s, i - 1, j, k - 1, X[0]); vtx(&d, vert, s, i - 1, j, k, X[1]);
          vtx(&d, vert, s, i, j, k, X[2]); vtx(&d, vert, s, i, j, k - 1, X[3]);
          midpoint(X[0], X[1], 0, 0.0, X[4]);
          midpoint(X[1], X[2], 1, norm3(X[2]), X[5]);
          midpoint(X[2], X[3], 0, 0.0, X[6]);
          midpoint(X[3], X[0], 1, norm3(X[0]), X[7]);
          face_quad(X, yb[0] + 15 * b + o, yb[1] + 5 * b + o, yb[2] + 3 * b + o, &area, cen, 0, S, 1);
        }
    for (k = -g; k <= NR + g; ++k)
      for (j = 1 - g; j <= N + g; ++j)
        for (i = 1 - g; i <= N + g; ++i) {
          double X[8][3], r0;
          size_t o = (size_t)(i - 1 + g) + (size_t)d.nc * ((j - 1 + g) + (size_t)d.nc * (k + g));
          size_t S = d.zfs, b = (size_t)s * S;
          vtx(&d, vert, s, i - 1, j - 1, k, X[0]); vtx(&d, vert, s, i, j - 1, k, X[1]);
          vtx(&d, vert, s, i, j, k, X[2]); vtx(&d, vert, s, i - 1, j, k, X[3]);
          r0 = norm3(X[0]);
          for (n = 0; n < 4; ++n) midpoint(X[n], X[(n + 1) & 3], 1, r0, X[4 + n]);
          face_quad(X, zb[0] + 15 * b + o, zb[1] + 5 * b + o, zb[2] + 15 * b + o, &area, cen, 1, S, 1);
        }
  }
}
