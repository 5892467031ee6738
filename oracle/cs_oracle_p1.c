/* cs_oracle_p1.c -- CPU restatement of the reference's second-order zone
 * reconstruction for one patch: LIMIT_ALL_SLOPES_P1
 * (cubesphere/400riemann_integrate/410limit_all_slopes_p1.f:9) with
 * ENFORCE_PRESSURE_POSITIVITY_PRIM (440enforce_pressure_positivity_prim.f:6)
 * and MATSOLVE (100math/120matsolve.f, LU of Numerical Recipes' ludcmp/lubksb,
 * 100math/110numerical_recipes.f).
 *
 * TEST INFRASTRUCTURE ONLY: the checker of the HIP path, never linked by the
 * product.  Configuration: ideal MHD (8 components), adiabatic EOS, P1 modes
 * (mean + 3 gradients), PRESSURE_POSITIVITY = 1, no B0 field, no helio
 * application branches, 5 B-field points per face.
 *
 * Arrays are the reference's, column-major, for one patch with NGhost = g:
 *   zone   (1-g:N+g, 1-g:N+g, 1-g:nr+g)       [, comp][, mode]
 *   faceB  (0-g:N+g, 0-g:N+g, 0-g:nr+g)       Bxf, Byf, Bzf [, point]
 *   x-face (0-g:N+g, 1-g:N+g, 1-g:nr+g) y-face (1-g:N+g, 0-g:N+g, 1-g:nr+g)
 *   z-face (1-g:N+g, 1-g:N+g, 0-g:nr+g)       [, xyz][, point]
 *   vert   (0-g:N+g, 0-g:N+g, 0-g:nr+g, 3)
 *   stencil (1-g:N+g, 1-g:N+g, 2, 12)
 * The ranges below are the reference's loop bounds; they need g >= 4.
 */
#include <math.h>
#include <stddef.h>
#include <string.h>

#include "cs_oracle.h"

typedef struct {
  int N, NR, g, nz, nzr, nf, nfr;
  long NZ, NXF, NYF, NZF, NFC, NV;
} p1dims;

static p1dims p1_dims(const or_p1_patch* p) {
  p1dims d;
  d.N = p->ntp;
  d.NR = p->nr;
  d.g = p->nghost;
  d.nz = d.N + 2 * d.g;
  d.nzr = d.NR + 2 * d.g;
  d.nf = d.nz + 1;
  d.nfr = d.nzr + 1;
  d.NZ = (long)d.nz * d.nz * d.nzr;
  d.NXF = (long)d.nf * d.nz * d.nzr;
  d.NYF = d.NXF;
  d.NZF = (long)d.nz * d.nz * d.nfr;
  d.NFC = (long)d.nf * d.nf * d.nfr;
  d.NV = d.NFC;
  return d;
}

/* element offsets of the layouts above (first index fastest) */
static long zi(const p1dims* d, int i, int j, int k) {
  return (i - 1 + d->g) + (long)d->nz * ((j - 1 + d->g) + (long)d->nz * (k - 1 + d->g));
}
static long fci(const p1dims* d, int i, int j, int k) {
  return (i + d->g) + (long)d->nf * ((j + d->g) + (long)d->nf * (k + d->g));
}
static long xfi(const p1dims* d, int i, int j, int k) {
  return (i + d->g) + (long)d->nf * ((j - 1 + d->g) + (long)d->nz * (k - 1 + d->g));
}
static long yfi(const p1dims* d, int i, int j, int k) {
  return (i - 1 + d->g) + (long)d->nz * ((j + d->g) + (long)d->nf * (k - 1 + d->g));
}
static long zfi(const p1dims* d, int i, int j, int k) {
  return (i - 1 + d->g) + (long)d->nz * ((j - 1 + d->g) + (long)d->nz * (k + d->g));
}
/* stencil entry n (0-based: R L U D RR LL UU DD RU LU LD RD) of zone column (i, j) */
static void nbr(const or_p1_patch* p, const p1dims* d, int i, int j, int n, int* ni, int* nj) {
  long s = (i - 1 + d->g) + (long)d->nz * (j - 1 + d->g), plane = (long)d->nz * d->nz;
  *ni = p->stencil[s + plane * (2 * n)];
  *nj = p->stencil[s + plane * (2 * n + 1)];
}
static int active(const or_p1_patch* p, const p1dims* d, int i, int j, int k) {
  return p->mask[zi(d, i, j, k)] == 1;
}

#define VV(p, d, c, m, z) (p)->V[(z) + (d)->NZ * ((c) + 8 * (m))]
#define UU(p, d, c, m, z) (p)->U[(z) + (d)->NZ * ((c) + 8 * (m))]

static double omax(double a, double b) { return a > b ? a : b; }
static double omin(double a, double b) { return a < b ? a : b; }

/* relative position of point x to zone centroid, scaled: (x - c) / l */
static void rel(const or_p1_patch* p, const p1dims* d, long z, const double* x, long xs, double obl, double* r) {
  for (int c = 0; c < 3; ++c) r[c] = (x[xs * c] - p->v_centroid[z + d->NZ * c]) * obl;
}

static double obl_of(const or_p1_patch* p, long z) { return 1.0 / omax(p->scale[z], p->smallnum); }

/* ------------------------------------------------------------------------ */
/* Second-order multi-D WENO of 410...f:2421-3761 (and :5437-7100): eight
 * one-sided three-zone stencils {R,L} x {U,D} x {z+1, z-1}, each a 3x3 solve
 * for the gradient; the ninth, central stencil averages the eight and weighs
 * central_weight = 100; weights 1/(IS + smallnum)^4 with IS = |grad|^2. */
typedef struct {
  double A[8][3][3];
} weno_mats;

/* stencil j uses (x-neighbour, y-neighbour, z-neighbour):
 * 1 (R,U,+) 2 (R,U,-) 3 (L,U,+) 4 (L,U,-) 5 (R,D,+) 6 (R,D,-) 7 (L,D,+) 8 (L,D,-) */
static const int c_sx[8] = {0, 0, 1, 1, 0, 0, 1, 1};  /* 0: R, 1: L */
static const int c_sy[8] = {0, 0, 0, 0, 1, 1, 1, 1};  /* 0: U, 1: D */
static const int c_sz[8] = {0, 1, 0, 1, 0, 1, 0, 1};  /* 0: +1, 1: -1 */

/* neighbour zones of (i, j, k) in the order R, L, U, D, top, bottom */
static void six_nbrs(const or_p1_patch* p, const p1dims* d, int i, int j, int k, long* z6) {
  int a, b;
  nbr(p, d, i, j, 0, &a, &b); z6[0] = zi(d, a, b, k);
  nbr(p, d, i, j, 1, &a, &b); z6[1] = zi(d, a, b, k);
  nbr(p, d, i, j, 2, &a, &b); z6[2] = zi(d, a, b, k);
  nbr(p, d, i, j, 3, &a, &b); z6[3] = zi(d, a, b, k);
  z6[4] = zi(d, i, j, k + 1);
  z6[5] = zi(d, i, j, k - 1);
}

/* adjugate / determinant inverse of the 3x3 matrix of scaled centroid offsets */
static void inv3(const double a[3][3], double tiny, double r[3][3]) {
  double a11 = a[0][0], a12 = a[0][1], a13 = a[0][2], a21 = a[1][0], a22 = a[1][1], a23 = a[1][2], a31 = a[2][0],
         a32 = a[2][1], a33 = a[2][2];
  double det = -(a13 * a22 * a31) + a12 * a23 * a31 + a13 * a21 * a32 - a11 * a23 * a32 - a12 * a21 * a33 + a11 * a22 * a33;
  double od = (signbit(det) ? -1.0 : 1.0) / omax(tiny, fabs(det));
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

static void weno_build(const or_p1_patch* p, const p1dims* d, long z0, const long* z6, weno_mats* W) {
  double obl = obl_of(p, z0), off[6][3];
  for (int n = 0; n < 6; ++n)
    for (int c = 0; c < 3; ++c) off[n][c] = (p->v_centroid[z6[n] + d->NZ * c] - p->v_centroid[z0 + d->NZ * c]) * obl;
  for (int s = 0; s < 8; ++s) {
    double a[3][3];
    for (int c = 0; c < 3; ++c) {
      a[0][c] = off[c_sx[s]][c];
      a[1][c] = off[2 + c_sy[s]][c];
      a[2][c] = off[4 + c_sz[s]][c];
    }
    inv3(a, 1.0e-32, W->A[s]);
  }
}

/* gradient modes of a field with value f0 at the zone and f6 at its six
 * neighbours (R, L, U, D, top, bottom), times (1 - flatten) */
static void weno_apply(const weno_mats* W, double f0, const double* f6, double smallnum, double flat, double* g3) {
  double X[9][3], IS[9], w[9], sum;
  for (int s = 0; s < 8; ++s) {
    double r0 = f6[c_sx[s]] - f0, r1 = f6[2 + c_sy[s]] - f0, r2 = f6[4 + c_sz[s]] - f0;
    for (int m = 0; m < 3; ++m) X[s][m] = W->A[s][m][0] * r0 + W->A[s][m][1] * r1 + W->A[s][m][2] * r2;
  }
  for (int m = 0; m < 3; ++m)
    X[8][m] = 0.125 * (X[0][m] + X[1][m] + X[2][m] + X[3][m] + X[4][m] + X[5][m] + X[6][m] + X[7][m]);
  for (int s = 0; s < 9; ++s) {
    IS[s] = X[s][0] * X[s][0] + X[s][1] * X[s][1] + X[s][2] * X[s][2];
    double t = IS[s] + smallnum;
    t = t * t;
    w[s] = (s == 8 ? 100.0 : 1.0) / (t * t);
  }
  sum = 1.0 / (w[0] + w[1] + w[2] + w[3] + w[4] + w[5] + w[6] + w[7] + w[8]);
  for (int s = 0; s < 9; ++s) w[s] = w[s] * sum;
  for (int m = 0; m < 3; ++m) {
    double v = w[0] * X[0][m];
    for (int s = 1; s < 9; ++s) v = v + w[s] * X[s][m];
    g3[m] = v * (1.0 - flat);
  }
}

/* ------------------------------------------------------------------------ */
/* Numerical Recipes' LU decomposition with implicit (row-scaled) partial
 * pivoting (Crout), then forward/back substitution: Press et al., Numerical
 * Recipes, 2nd ed., sec. 2.3, as called through MATSOLVE (120matsolve.f). */
#define LUMAX 36
static int lu_dcmp(int n, double a[LUMAX][LUMAX], int* indx) {
  int imax = 0;
  double vv[LUMAX];
  for (int i = 0; i < n; ++i) {
    double big = 0.0;
    for (int j = 0; j < n; ++j)
      if (fabs(a[i][j]) > big) big = fabs(a[i][j]);
    if (fabs(big) < 1.0e-30) return 1;
    vv[i] = 1.0 / big;
  }
  for (int j = 0; j < n; ++j) {
    for (int i = 0; i < j; ++i) {
      double s = a[i][j];
      for (int k = 0; k < i; ++k) s = s - a[i][k] * a[k][j];
      a[i][j] = s;
    }
    double big = 0.0;
    for (int i = j; i < n; ++i) {
      double s = a[i][j];
      for (int k = 0; k < j; ++k) s = s - a[i][k] * a[k][j];
      a[i][j] = s;
      double t = vv[i] * fabs(s);
      if (t >= big) {
        big = t;
        imax = i;
      }
    }
    if (j != imax) {
      for (int k = 0; k < n; ++k) {
        double t = a[imax][k];
        a[imax][k] = a[j][k];
        a[j][k] = t;
      }
      vv[imax] = vv[j];
    }
    indx[j] = imax;
    if (fabs(a[j][j]) < 1.0e-30) return 1;
    if (j != n - 1) {
      double r = 1.0 / a[j][j];
      for (int i = j + 1; i < n; ++i) a[i][j] = a[i][j] * r;
    }
  }
  return 0;
}
static void lu_bksb(int n, double a[LUMAX][LUMAX], const int* indx, double* b) {
  int ii = -1;
  for (int i = 0; i < n; ++i) {
    int ip = indx[i];
    double s = b[ip];
    b[ip] = b[i];
    if (ii >= 0)
      for (int j = ii; j < i; ++j) s = s - a[i][j] * b[j];
    else if (s != 0.0)
      ii = i;
    b[i] = s;
  }
  for (int i = n - 1; i >= 0; --i) {
    double s = b[i];
    for (int j = i + 1; j < n; ++j) s = s - a[i][j] * b[j];
    b[i] = s / a[i][i];
  }
}
static int lu_solve(int n, double a[LUMAX][LUMAX], double* b) {
  int indx[LUMAX];
  if (lu_dcmp(n, a, indx)) return 1;
  lu_bksb(n, a, indx, b);
  return 0;
}
/* MATINV (100math/121matinv.f:8): the LU above, then one back-substitution
 * per column of the identity; m (row-major n x n) is destroyed.  On a
 * singular matrix the reference continues with its partial factors; here
 * the inverse is left zero and 1 is returned. */
int or_matinv(int n, double* m, double* minv) {
  double a[LUMAX][LUMAX], col[LUMAX];
  int indx[LUMAX];
  if (n < 1 || n > LUMAX) return 1;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) a[i][j] = m[i * n + j];
  for (int i = 0; i < n * n; ++i) minv[i] = 0.0;
  if (lu_dcmp(n, a, indx)) return 1;
  for (int j = 0; j < n; ++j) {
    for (int i = 0; i < n; ++i) col[i] = i == j ? 1.0 : 0.0;
    lu_bksb(n, a, indx, col);
    for (int i = 0; i < n; ++i) minv[i * n + j] = col[i];
  }
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) m[i * n + j] = a[i][j];
  return 0;
}

/* one row [n_x p, n_y p, n_z p], p = (1, x, y, z) */
static void bfield_row(const double* n, const double* r, double* row) {
  for (int c = 0; c < 3; ++c) {
    row[4 * c + 0] = n[c];
    row[4 * c + 1] = n[c] * r[0];
    row[4 * c + 2] = n[c] * r[1];
    row[4 * c + 3] = n[c] * r[2];
  }
}

/* zones of the last call whose slopes the positivity limiter reduced, and
   those it flattened completely */
static long g_p1_stats[2];
void or_p1_positivity_stats(long* out) {
  out[0] = g_p1_stats[0];
  out[1] = g_p1_stats[1];
}

/* ------------------------------------------------------------------------ */
/* Steps 2 to 7.a of LIMIT_ALL_SLOPES_P1 (410...f:529-3761), shared with
 * the third-order LIMIT_ALL_SLOPES_P12 (421...f:573-3810): U and V hold nm
 * modes (4 or 10); the third-order routine leaves the WENO gradients of V
 * unflattened (421...f:3435-3448: the flattener comes after the quadratic
 * hybridisation) and flattens only the entropy's (flat_v 0). */
static void p1_front(const or_p1_patch* p, const p1dims* d, int nm, int flat_v) {
  const int N = d->N, NR = d->NR, g = d->g;
  const double pi = 4.0 * atan(1.0), ob8pi = 1.0 / (8.0 * pi), ob4pi = 1.0 / (4.0 * pi), gam = p->gamma;
  int i, j, k, c, m, q;

  /* STEP 2 (410...f:529-720): constants, zero the higher modes, make the
     tri-corner (inactive) zones safe; zero face B between two inactive zones */
  for (k = 1 - g; k <= NR + g; ++k)
    for (j = 1 - g; j <= N + g; ++j)
      for (i = 1 - g; i <= N + g; ++i) {
        long z = zi(d, i, j, k);
        p->gamma_eos[z] = gam;
        p->gamma_sound[z] = gam;
        p->mean_mol_wt[z] = p->molewt;
        p->msonic[z] = p->smallnum;
        for (c = 0; c < 8; ++c)
          for (m = 1; m < nm; ++m) UU(p, d, c, m, z) = 0.0;
        for (m = 1; m < 4; ++m) p->ent[z + d->NZ * m] = 0.0;
        for (c = 0; c < 8; ++c)
          for (m = 0; m < nm; ++m) VV(p, d, c, m, z) = 0.0;
        if (p->mask[z] == 0) {
          for (c = 0; c < 8; ++c) UU(p, d, c, 0, z) = 0.0;
          UU(p, d, 0, 0, z) = p->rhofloor;
          UU(p, d, 4, 0, z) = p->prsfloor / (gam - 1.0);
          p->scale[z] = p->smallnum;
        }
      }
  for (k = 1 - g; k <= NR + g; ++k)
    for (j = 1 - g; j <= N + g; ++j)
      for (i = -g; i <= N + g; ++i) {
        int a = i < 1 - g ? 1 - g : i, b = i + 1 > N + g ? N + g : i + 1;
        if (!active(p, d, a, j, k) && !active(p, d, b, j, k)) p->Bxf[fci(d, i, j, k)] = 0.0;
      }
  for (k = 1 - g; k <= NR + g; ++k)
    for (j = -g; j <= N + g; ++j)
      for (i = 1 - g; i <= N + g; ++i) {
        int a = j < 1 - g ? 1 - g : j, b = j + 1 > N + g ? N + g : j + 1;
        if (!active(p, d, i, a, k) && !active(p, d, i, b, k)) p->Byf[fci(d, i, j, k)] = 0.0;
      }
  for (k = -g; k <= NR + g; ++k)
    for (j = 1 - g; j <= N + g; ++j)
      for (i = 1 - g; i <= N + g; ++i)
        if (!active(p, d, i, j, 1)) p->Bzf[fci(d, i, j, k)] = 0.0;

  /* STEP 3 (410...f:722-920): at a tri-corner, the y-face B of the ghost row
     next to the panel edge and the x-face B it coincides with (the Down or Up
     neighbour's) are made equal, with the sign each corner's orientation gives */
  for (int corner = 0; corner < 4; ++corner) {
    if (!p->itri[corner]) continue;
    int jj = (corner < 2) ? 1 : N, lo = (corner == 0 || corner == 3) ? 1 - g : N + 1,
        hi = (corner == 0 || corner == 3) ? 0 : N + g;
    for (i = lo; i <= hi; ++i) {
      int a, b;
      nbr(p, d, i, jj, corner < 2 ? 3 : 2, &a, &b);
      for (k = 1 - g; k <= NR + g; ++k) {
        double* by = &p->Byf[fci(d, i, corner < 2 ? jj - 1 : jj, k)];
        double* bx = &p->Bxf[fci(d, (corner == 0 || corner == 3) ? a - 1 : a, b, k)];
        if (corner == 0 || corner == 2) { /* LD, RU: opposite signs */
          double t = 0.5 * (*by - *bx);
          *by = t;
          *bx = -t;
        } else { /* RD, LU: same signs */
          double t = 0.5 * (*by + *bx);
          *by = t;
          *bx = t;
        }
      }
    }
  }

  /* STEP 5 (410...f:1421-1740): primitives of every zone, magnetosonic speed */
  for (k = 1 - g; k <= NR + g; ++k)
    for (j = 1 - g; j <= N + g; ++j)
      for (i = 1 - g; i <= N + g; ++i) {
        long z = zi(d, i, j, k);
        double Up[8], Vp[8], e = p->ent[z];
        for (c = 0; c < 8; ++c) Up[c] = UU(p, d, c, 0, z);
        or_cons_to_prim(gam, p->rhofloor, p->prsfloor, &e, Up, Vp);
        for (c = 0; c < 8; ++c) {
          VV(p, d, c, 0, z) = Vp[c];
          UU(p, d, c, 0, z) = Up[c];
        }
        p->ent[z] = e;
        for (m = 1; m < 4; ++m) p->ent[z + d->NZ * m] = 0.0;
        double b2 = (Vp[5] * Vp[5] + Vp[6] * Vp[6] + Vp[7] * Vp[7]) * ob4pi;
        double ms = p->gamma_sound[z] * Vp[4] + b2;
        p->msonic[z] = sqrt(ms / Vp[0]);
      }

  /* STEP 6.a (410...f:1748-2090): divergence of the velocity through the six
     faces, scaled to the zone; flattener from |div v| / (0.4 min c_f) */
  const int f0 = 2 - g, fx1 = N + g - 1, fz1 = NR + g - 1;
  for (long z = 0; z < d->NZ; ++z) p->flatten[z] = 0.0;
  for (k = f0; k <= fz1; ++k)
    for (j = f0; j <= fx1; ++j)
      for (i = f0; i <= fx1; ++i) {
        long z = zi(d, i, j, k), z6[6];
        six_nbrs(p, d, i, j, k, z6);
        double obl = obl_of(p, z), dv = 0.0, wmin = p->msonic[z];
        for (q = 0; q < 6; ++q) wmin = omin(wmin, p->msonic[z6[q]]);
        p->work[z] = wmin;
        for (q = 0; q < 6; ++q) {
          /* faces R, L, U, D use the face normal; top and bottom the unit z-face centroid */
          long zo = (q % 2 == 0) ? z6[q] : z, zi_ = (q % 2 == 0) ? z : z6[q];
          double r[3], n[3], s2 = 0.0;
          for (c = 0; c < 3; ++c) r[c] = (p->v_centroid[zo + d->NZ * c] - p->v_centroid[zi_ + d->NZ * c]) * obl;
          double a11 = 0.5 / omax(sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]), p->smallnum);
          if (q < 2) {
            long f = xfi(d, q == 0 ? i : i - 1, j, k);
            for (c = 0; c < 3; ++c) n[c] = p->xf_normal[f + d->NXF * c];
          } else if (q < 4) {
            long f = yfi(d, i, q == 2 ? j : j - 1, k);
            for (c = 0; c < 3; ++c) n[c] = p->yf_normal[f + d->NYF * c];
          } else {
            long f = zfi(d, i, j, k);
            for (c = 0; c < 3; ++c) n[c] = p->zf_centroid[f + d->NZF * c];
            double a12 = 1.0 / omax(sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]), p->smallnum);
            for (c = 0; c < 3; ++c) n[c] = n[c] * a12;
          }
          (void)s2;
          dv = dv + n[0] * a11 * (VV(p, d, 1, 0, zo) - VV(p, d, 1, 0, zi_)) +
               n[1] * a11 * (VV(p, d, 2, 0, zo) - VV(p, d, 2, 0, zi_)) +
               n[2] * a11 * (VV(p, d, 3, 0, zo) - VV(p, d, 3, 0, zi_));
        }
        p->divvel[z] = dv;
        double a13 = fabs(dv) / (p->flatten_coef * p->work[z]) - 1.0;
        p->flatten[z] = a13 <= 0.0 ? 0.0 : omin(1.0, a13);
        if (p->mask[z] != 1) p->flatten[z] = 0.0;
      }
  /* STEP 6.b (:2094-2310): take a neighbour's flattener where its pressure
     exceeds prs_flatten_coef times this zone's */
  memcpy(p->work, p->flatten, sizeof(double) * d->NZ);
  for (k = f0; k <= fz1; ++k)
    for (j = f0; j <= fx1; ++j)
      for (i = f0; i <= fx1; ++i) {
        long z = zi(d, i, j, k), z6[6];
        six_nbrs(p, d, i, j, k, z6);
        for (q = 0; q < 6; ++q)
          if (VV(p, d, 4, 0, z6[q]) > p->prs_flatten_coef * VV(p, d, 4, 0, z))
            p->flatten[z] = omax(p->flatten[z], p->work[z6[q]]);
      }
  /* STEP 6.c (:2317-2418): spread the flattener to the six neighbours */
  memcpy(p->work, p->flatten, sizeof(double) * d->NZ);
  for (k = f0; k <= fz1; ++k)
    for (j = f0; j <= fx1; ++j)
      for (i = f0; i <= fx1; ++i) {
        long z = zi(d, i, j, k), z6[6];
        six_nbrs(p, d, i, j, k, z6);
        for (q = 0; q < 6; ++q) p->flatten[z] = omax(p->flatten[z], p->work[z6[q]]);
      }

  /* STEP 7.a (:2421-3761): WENO gradients of the 8 primitives and the entropy */
  for (k = -2; k <= NR + 3; ++k)
    for (j = -2; j <= N + 3; ++j)
      for (i = -2; i <= N + 3; ++i) {
        if (!active(p, d, i, j, 1)) continue;
        long z = zi(d, i, j, k), z6[6];
        weno_mats W;
        six_nbrs(p, d, i, j, k, z6);
        weno_build(p, d, z, z6, &W);
        for (c = 0; c < 9; ++c) {
          double f6[6], g3[3], f0v;
          if (c < 8) {
            f0v = VV(p, d, c, 0, z);
            for (q = 0; q < 6; ++q) f6[q] = VV(p, d, c, 0, z6[q]);
          } else {
            f0v = p->ent[z];
            for (q = 0; q < 6; ++q) f6[q] = p->ent[z6[q]];
          }
          weno_apply(&W, f0v, f6, p->smallnum, (c < 8 && !flat_v) ? 0.0 : p->flatten[z], g3);
          for (m = 0; m < 3; ++m) {
            if (c < 8)
              VV(p, d, c, m + 1, z) = g3[m];
            else
              p->ent[z + d->NZ * (m + 1)] = g3[m];
          }
        }
      }

}

/* STEP 7.b of LIMIT_ALL_SLOPES_P1, 7.c of LIMIT_ALL_SLOPES_P12:
 * ENFORCE_PRESSURE_POSITIVITY_PRIM (440...f:6-830), zone by zone in the
 * loop order; nm modes (4, or 10 with the zone's moments mom: the node
 * values then carry the quadratic modes about the zone's own moments,
 * 440...f:594-600, and every mode above the mean is reduced) */
static void p1_positivity(const or_p1_patch* p, const p1dims* d, int nm, const double* mom) {
  const int N = d->N, NR = d->NR;
  int i, j, k, c, m, q;
    const int n_iter = 15;
    const double ratio = pow(0.10, 1.0 / (double)n_iter), expandfac = 0.2;
    for (k = -2; k <= NR + 3; ++k)
      for (j = -2; j <= N + 3; ++j)
        for (i = -2; i <= N + 3; ++i) {
          long z = zi(d, i, j, k), col[9];
          int a, b;
          /* the 27 zones around: the column itself and its 8 stencil neighbours
             (R, L, U, D, RU, LU, LD, RD), each at k-1, k, k+1 */
          static const int dirs[8] = {0, 1, 2, 3, 8, 9, 10, 11};
          col[0] = zi(d, i, j, 0);
          for (q = 0; q < 8; ++q) {
            nbr(p, d, i, j, dirs[q], &a, &b);
            col[q + 1] = zi(d, a, b, 0);
          }
          double rmin = 1e300, rmax = -1e300, pmin = 1e300, pmax = -1e300, emin = 1e300, emax = -1e300;
          const long kstride = (long)d->nz * d->nz;
          for (q = 0; q < 9; ++q)
            for (int dk = -1; dk <= 1; ++dk) {
              long zz = col[q] + kstride * (k - 0 + dk);
              rmin = omin(rmin, VV(p, d, 0, 0, zz));
              rmax = omax(rmax, VV(p, d, 0, 0, zz));
              pmin = omin(pmin, VV(p, d, 4, 0, zz));
              pmax = omax(pmax, VV(p, d, 4, 0, zz));
              emin = omin(emin, p->ent[zz]);
              emax = omax(emax, p->ent[zz]);
            }
          double ff = omax(omin((p->flatten[z] - 0.1) / 0.4, 1.0), 0.0), omff = 1.0 - ff;
          rmin = omax(rmin * (1.0 - expandfac * omff), p->rhofloor);
          rmax = omax(rmax * (1.0 + expandfac * omff), p->rhofloor);
          pmin = omax(pmin * (1.0 - expandfac * omff), p->prsfloor);
          pmax = omax(pmax * (1.0 + expandfac * omff), p->prsfloor);
          emin = emin * (1.0 - expandfac * omff);
          emax = emax * (1.0 + expandfac * omff);
          /* 27 nodes: 8 vertices, 6 face centroids, 12 edge midpoints, centroid */
          double node[27][3], cen[3], obl = obl_of(p, z);
          for (c = 0; c < 3; ++c) cen[c] = p->v_centroid[z + d->NZ * c];
          for (int v = 0; v < 8; ++v) {
            int vi = (v & 1) ? i : i - 1, vj = (v & 2) ? j : j - 1, vk = (v & 4) ? k : k - 1;
            for (c = 0; c < 3; ++c) node[v][c] = p->vert[fci(d, vi, vj, vk) + d->NV * c];
          }
          for (c = 0; c < 3; ++c) {
            node[8][c] = p->xf_centroid[xfi(d, i, j, k) + d->NXF * c];
            node[9][c] = p->xf_centroid[xfi(d, i - 1, j, k) + d->NXF * c];
            node[10][c] = p->yf_centroid[yfi(d, i, j, k) + d->NYF * c];
            node[11][c] = p->yf_centroid[yfi(d, i, j - 1, k) + d->NYF * c];
            node[12][c] = p->zf_centroid[zfi(d, i, j, k) + d->NZF * c];
            node[13][c] = p->zf_centroid[zfi(d, i, j, k - 1) + d->NZF * c];
          }
          static const int edge[12][2] = {{0, 1}, {2, 3}, {4, 5}, {6, 7}, {0, 2}, {1, 3},
                                          {4, 6}, {5, 7}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
          for (int e = 0; e < 12; ++e)
            for (c = 0; c < 3; ++c) node[14 + e][c] = (node[edge[e][0]][c] + node[edge[e][1]][c]) * 0.5;
          for (c = 0; c < 3; ++c) node[26][c] = cen[c];
          for (int n = 0; n < 27; ++n)
            for (c = 0; c < 3; ++c) node[n][c] = (node[n][c] - cen[c]) * obl;
          double Vm[8][10], em[4], q0[10] = {0};
          for (c = 0; c < 8; ++c)
            for (m = 0; m < nm; ++m) Vm[c][m] = VV(p, d, c, m, z);
          if (mom)
            for (m = 4; m < 10; ++m) q0[m] = mom[z + d->NZ * m];
          for (m = 0; m < 4; ++m) em[m] = p->ent[z + d->NZ * m];
          for (int it = 1; it <= n_iter + 1; ++it) {
            if (it == 2) ++g_p1_stats[0];
            if (it == n_iter + 1) {
              ++g_p1_stats[1];
              for (c = 0; c < 8; ++c)
                for (m = 1; m < nm; ++m) Vm[c][m] = 0.0;
              for (m = 0; m < 4; ++m) em[m] = 0.0;
              break;
            }
            int pass = 1;
            for (int n = 0; n < 27; ++n) {
              double x = node[n][0], y = node[n][1], zz = node[n][2];
              double rn = Vm[0][0] + Vm[0][1] * x + Vm[0][2] * y + Vm[0][3] * zz;
              double pn = Vm[4][0] + Vm[4][1] * x + Vm[4][2] * y + Vm[4][3] * zz;
              if (nm == 10) { /* 440...f:594-600: the quadratic modes about the zone's moments */
                const double b5 = x * x - q0[4], b6 = y * y - q0[5], b7 = zz * zz - q0[6];
                const double b8 = x * y - q0[7], b9 = y * zz - q0[8], b10 = x * zz - q0[9];
                rn = rn + Vm[0][4] * b5 + Vm[0][5] * b6 + Vm[0][6] * b7 + Vm[0][7] * b8 + Vm[0][8] * b9 + Vm[0][9] * b10;
                pn = pn + Vm[4][4] * b5 + Vm[4][5] * b6 + Vm[4][6] * b7 + Vm[4][7] * b8 + Vm[4][8] * b9 + Vm[4][9] * b10;
              }
              double en = em[0] + em[1] * x + em[2] * y + em[3] * zz;
              if (rn < rmin || rn > rmax) pass = 0;
              if (pn < pmin || pn > pmax) pass = 0;
              if (en < emin || en > emax) pass = 0;
            }
            if (pass) break;
            for (c = 0; c < 8; ++c)
              for (m = 1; m < nm; ++m) Vm[c][m] = ratio * Vm[c][m];
            for (m = 1; m < 4; ++m) em[m] = ratio * em[m];
          }
          for (c = 0; c < 8; ++c)
            for (m = 0; m < nm; ++m) VV(p, d, c, m, z) = Vm[c][m];
          for (m = 0; m < 4; ++m) p->ent[z + d->NZ * m] = em[m];
          for (m = 0; m < nm; ++m) {
            p->B1_zc[z + d->NZ * m] = Vm[5][m];
            p->B2_zc[z + d->NZ * m] = Vm[6][m];
            p->B3_zc[z + d->NZ * m] = Vm[7][m];
          }
        }
}

/* STEP 8.a-c of LIMIT_ALL_SLOPES_P1 (410...f:3874-4640) and of
 * LIMIT_ALL_SLOPES_P12 (421...f:4668-5573): normal B at the five B-field
 * points of each face from the two zones' fields (linear, or with nm 10 the
 * quadratic modes about each zone's moments mom), averaged, then shifted so
 * the weighted mean is the face's own Bf */
static void p1_face_points(const or_p1_patch* p, const p1dims* d, int nm, const double* mom) {
  const int N = d->N, NR = d->NR;
  int i, j, k, c, q;
  for (int fam = 0; fam < 3; ++fam) {
    const int ilo = fam == 0 ? -3 : -2, jlo = fam == 1 ? -3 : -2, klo = fam == 2 ? -3 : -2;
    double* Bq = fam == 0 ? p->Bxf_q : (fam == 1 ? p->Byf_q : p->Bzf_q);
    const double* Bf = fam == 0 ? p->Bxf : (fam == 1 ? p->Byf : p->Bzf);
    const double* pts = fam == 0 ? p->xb_pts : (fam == 1 ? p->yb_pts : p->zb_pts);
    const double* wts = fam == 0 ? p->xb_wts : (fam == 1 ? p->yb_wts : p->zb_wts);
    const double* nrm = fam == 0 ? p->xb_normal : (fam == 1 ? p->yb_normal : p->zb_normal);
    const long NF = fam == 0 ? d->NXF : (fam == 1 ? d->NYF : d->NZF);
    for (k = klo; k <= NR + 3; ++k)
      for (j = jlo; j <= N + 3; ++j)
        for (i = ilo; i <= N + 3; ++i) {
          long fb = fci(d, i, j, k), fg = fam == 0 ? xfi(d, i, j, k) : (fam == 1 ? yfi(d, i, j, k) : zfi(d, i, j, k));
          int lo_ok, zl_i = i, zl_j = j, zr_i = i + (fam == 0), zr_j = j + (fam == 1), zr_k = k + (fam == 2);
          int skip = 0;
          if (fam < 2) {
            int al = active(p, d, zl_i, zl_j, k), ar = active(p, d, zr_i, zr_j, k);
            if (al && ar) {
            } else if (!al && ar) {
              nbr(p, d, zr_i, zr_j, fam == 0 ? 1 : 3, &zl_i, &zl_j);
            } else if (al && !ar) {
              nbr(p, d, zl_i, zl_j, fam == 0 ? 0 : 2, &zr_i, &zr_j);
            } else {
              skip = 1;
            }
          } else {
            skip = !active(p, d, i, j, k);
          }
          (void)lo_ok;
          for (q = 0; q < 5; ++q) {
            if (skip) {
              Bq[fb + d->NFC * q] = 0.0;
              continue;
            }
            long zl = zi(d, zl_i, zl_j, k), zr = zi(d, zr_i, zr_j, fam == 2 ? zr_k : k);
            double n[3], rl[3], rr[3], bl[3], br[3];
            for (c = 0; c < 3; ++c) n[c] = fam == 2 ? nrm[fg + NF * (c + 3 * q)] : nrm[fg + NF * c];
            rel(p, d, zl, pts + fg + NF * (3 * q), NF, obl_of(p, zl), rl);
            rel(p, d, zr, pts + fg + NF * (3 * q), NF, obl_of(p, zr), rr);
            for (c = 0; c < 3; ++c) {
              bl[c] = VV(p, d, 5 + c, 0, zl) + VV(p, d, 5 + c, 1, zl) * rl[0] + VV(p, d, 5 + c, 2, zl) * rl[1] +
                      VV(p, d, 5 + c, 3, zl) * rl[2];
              br[c] = VV(p, d, 5 + c, 0, zr) + VV(p, d, 5 + c, 1, zr) * rr[0] + VV(p, d, 5 + c, 2, zr) * rr[1] +
                      VV(p, d, 5 + c, 3, zr) * rr[2];
              if (nm == 10) { /* 421...f:4817-4879: the quadratic modes about each zone's own moments */
                const double* ml = mom + zl;
                const double* mr = mom + zr;
                const long S = d->NZ;
                bl[c] = bl[c] + VV(p, d, 5 + c, 4, zl) * (rl[0] * rl[0] - ml[4 * S]) +
                        VV(p, d, 5 + c, 5, zl) * (rl[1] * rl[1] - ml[5 * S]) +
                        VV(p, d, 5 + c, 6, zl) * (rl[2] * rl[2] - ml[6 * S]) +
                        VV(p, d, 5 + c, 7, zl) * (rl[0] * rl[1] - ml[7 * S]) +
                        VV(p, d, 5 + c, 8, zl) * (rl[1] * rl[2] - ml[8 * S]) +
                        VV(p, d, 5 + c, 9, zl) * (rl[0] * rl[2] - ml[9 * S]);
                br[c] = br[c] + VV(p, d, 5 + c, 4, zr) * (rr[0] * rr[0] - mr[4 * S]) +
                        VV(p, d, 5 + c, 5, zr) * (rr[1] * rr[1] - mr[5 * S]) +
                        VV(p, d, 5 + c, 6, zr) * (rr[2] * rr[2] - mr[6 * S]) +
                        VV(p, d, 5 + c, 7, zr) * (rr[0] * rr[1] - mr[7 * S]) +
                        VV(p, d, 5 + c, 8, zr) * (rr[1] * rr[2] - mr[8 * S]) +
                        VV(p, d, 5 + c, 9, zr) * (rr[0] * rr[2] - mr[9 * S]);
              }
            }
            Bq[fb + d->NFC * q] = 0.5 * (n[0] * (bl[0] + br[0]) + n[1] * (bl[1] + br[1]) + n[2] * (bl[2] + br[2]));
          }
          double bd = Bf[fb];
          for (q = 0; q < 5; ++q) bd = bd - wts[fg + NF * q] * Bq[fb + d->NFC * q];
          for (q = 0; q < 5; ++q) Bq[fb + d->NFC * q] = Bq[fb + d->NFC * q] + bd;
        }
  }

}

/* STEP 10 of LIMIT_ALL_SLOPES_P1 and _P12 (410...f:7101-7700,
 * 421...f:9354-): the same in both routines */
static void p1_step10(const or_p1_patch* p, const p1dims* d) {
  const int N = d->N, NR = d->NR;
  const double pi = 4.0 * atan(1.0), ob4pi = 1.0 / (4.0 * pi);
  int i, j, k, c, q;
  /* min and max of the primitives over the owned zones, the face-B
     divergence (kept in work) normalised by |B| and the magnetosonic speed
     (recomputed from the new primitives) */
  double vmin[9], vmax[9];
  for (q = 0; q < 9; ++q) {
    vmin[q] = 1.0e30;
    vmax[q] = -1.0e30;
  }
  for (long z = 0; z < d->NZ; ++z) p->work[z] = 0.0;
  for (k = 1; k <= NR; ++k)
    for (j = 1; j <= N; ++j)
      for (i = 1; i <= N; ++i) {
        long z = zi(d, i, j, k);
        if (p->mask[z] != 1) continue;
        for (c = 0; c < 8; ++c) {
          vmin[c] = omin(vmin[c], VV(p, d, c, 0, z));
          vmax[c] = omax(vmax[c], VV(p, d, c, 0, z));
        }
        long xp = xfi(d, i, j, k), xm = xfi(d, i - 1, j, k), yp = yfi(d, i, j, k), ym = yfi(d, i, j - 1, k),
             zp = zfi(d, i, j, k), zm = zfi(d, i, j, k - 1);
        double bxp = p->Bxf[fci(d, i, j, k)], bxm = p->Bxf[fci(d, i - 1, j, k)], byp = p->Byf[fci(d, i, j, k)],
               bym = p->Byf[fci(d, i, j - 1, k)], bzp = p->Bzf[fci(d, i, j, k)], bzm = p->Bzf[fci(d, i, j, k - 1)];
        double a21 = p->xf_area[xp] * bxp + p->yf_area[yp] * byp + p->zf_area[zp] * bzp - p->xf_area[xm] * bxm -
                     p->yf_area[ym] * bym - p->zf_area[zm] * bzm;
        p->work[z] = a21;
        double v6 = VV(p, d, 5, 0, z), v7 = VV(p, d, 6, 0, z), v8 = VV(p, d, 7, 0, z);
        double a11 = (v6 + 0.0) * (v6 + 0.0) + (v7 + 0.0) * (v7 + 0.0) + (v8 + 0.0) * (v8 + 0.0);
        double ms = p->gamma_sound[z] * VV(p, d, 4, 0, z) + ob4pi * a11;
        p->msonic[z] = sqrt(ms / VV(p, d, 0, 0, z));
        double den = omax(omax(sqrt(bxp * bxp + byp * byp + bzp * bzp), sqrt(bxm * bxm + bym * bym + bzm * bzm)),
                          p->msonic[z] * sqrt(4.0 * 3.14 * VV(p, d, 0, 0, z)));
        a21 = a21 / den;
        vmin[8] = omin(vmin[8], a21);
        vmax[8] = omax(vmax[8], a21);
      }
  if (p->V_min_local)
    for (q = 0; q < 9; ++q) p->V_min_local[q] = vmin[q];
  if (p->V_max_local)
    for (q = 0; q < 9; ++q) p->V_max_local[q] = vmax[q];
}

int or_limit_all_slopes_p1(const or_p1_patch* p) {
  if (!p || p->nghost < 4 || p->ntp < 1 || p->nr < 1) return 1;
  const p1dims D = p1_dims(p), *d = &D;
  const int N = d->N, NR = d->NR, g = d->g;
  g_p1_stats[0] = g_p1_stats[1] = 0;
  const double pi = 4.0 * atan(1.0), ob8pi = 1.0 / (8.0 * pi), ob4pi = 1.0 / (4.0 * pi), gam = p->gamma;
  int i, j, k, c, m, q;
  p1_front(p, d, 4, 1);

  p1_positivity(p, d, 4, NULL);

  p1_face_points(p, d, 4, NULL);

  /* STEP 8.d (:4643-5360): ADP constrained least squares for the 12 linear
     coefficients of B: the six face means exactly (constraints), points 2 and
     4 of each face and the WENO modes in the least-squares sense */
  for (k = -2; k <= NR + 3; ++k)
    for (j = -2; j <= N + 3; ++j)
      for (i = -2; i <= N + 3; ++i) {
        long z = zi(d, i, j, k);
        if (p->mask[z] == 0) continue;
        double obl = obl_of(p, z), Cm[6][12], dv[6], Am[24][12], bv[24], r[3], n[3];
        memset(Am, 0, sizeof Am);
        memset(bv, 0, sizeof bv);
        /* faces in the order right x, left x, up y, down y, top z, bottom z */
        for (int f = 0; f < 6; ++f) {
          int fam = f / 2, hi = f % 2 == 0;
          int fi_ = i - (fam == 0 && !hi), fj_ = j - (fam == 1 && !hi), fk_ = k - (fam == 2 && !hi);
          long fg = fam == 0 ? xfi(d, fi_, fj_, fk_) : (fam == 1 ? yfi(d, fi_, fj_, fk_) : zfi(d, fi_, fj_, fk_));
          long NF = fam == 0 ? d->NXF : (fam == 1 ? d->NYF : d->NZF);
          long fb = fci(d, fi_, fj_, fk_);
          const double* nrm = fam == 0 ? p->xb_normal : (fam == 1 ? p->yb_normal : p->zb_normal);
          const double* cen = fam == 0 ? p->xf_centroid : (fam == 1 ? p->yf_centroid : p->zf_centroid);
          const double* pts = fam == 0 ? p->xb_pts : (fam == 1 ? p->yb_pts : p->zb_pts);
          const double* Bf = fam == 0 ? p->Bxf : (fam == 1 ? p->Byf : p->Bzf);
          const double* Bq = fam == 0 ? p->Bxf_q : (fam == 1 ? p->Byf_q : p->Bzf_q);
          for (c = 0; c < 3; ++c) n[c] = nrm[fg + NF * c];  /* z: the first (central) point's normal */
          rel(p, d, z, cen + fg, NF, obl, r);
          bfield_row(n, r, Cm[f]);
          dv[f] = Bf[fb];
          for (int t = 0; t < 2; ++t) {
            int pt = t == 0 ? 1 : 3;  /* points 2 and 4 */
            if (fam == 2)
              for (c = 0; c < 3; ++c) n[c] = nrm[fg + NF * (c + 3 * pt)];
            rel(p, d, z, pts + fg + NF * (3 * pt), NF, obl, r);
            bfield_row(n, r, Am[2 * f + t]);
            bv[2 * f + t] = Bq[fb + d->NFC * pt];
          }
        }
        for (q = 0; q < 12; ++q) {
          Am[12 + q][q] = 1.0;
          bv[12 + q] = VV(p, d, 5 + q / 4, q % 4, z);
        }
        double K[LUMAX][LUMAX], rhs[LUMAX];
        for (int a = 0; a < 12; ++a) {
          for (int b = 0; b < 12; ++b) {
            double s = 0.0;
            for (q = 0; q < 24; ++q) s = s + Am[q][a] * Am[q][b];
            K[a][b] = s;
          }
          double s = 0.0;
          for (q = 0; q < 24; ++q) s = s + Am[q][a] * bv[q];
          rhs[a] = s;
          for (int f = 0; f < 6; ++f) {
            K[a][12 + f] = Cm[f][a];
            K[12 + f][a] = Cm[f][a];
          }
        }
        for (int f = 0; f < 6; ++f) {
          for (int h = 0; h < 6; ++h) K[12 + f][12 + h] = 0.0;
          rhs[12 + f] = dv[f];
        }
        lu_solve(18, K, rhs);
        double aa = (UU(p, d, 5, 0, z) * UU(p, d, 5, 0, z) + UU(p, d, 6, 0, z) * UU(p, d, 6, 0, z) +
                     UU(p, d, 7, 0, z) * UU(p, d, 7, 0, z)) * ob8pi;
        for (m = 0; m < 4; ++m) {
          p->B1_zc[z + d->NZ * m] = rhs[m];
          p->B2_zc[z + d->NZ * m] = rhs[4 + m];
          p->B3_zc[z + d->NZ * m] = rhs[8 + m];
          for (c = 0; c < 3; ++c) VV(p, d, 5 + c, m, z) = rhs[4 * c + m];
        }
        for (c = 0; c < 3; ++c) UU(p, d, 5 + c, 0, z) = VV(p, d, 5 + c, 0, z);
        double bb = (UU(p, d, 5, 0, z) * UU(p, d, 5, 0, z) + UU(p, d, 6, 0, z) * UU(p, d, 6, 0, z) +
                     UU(p, d, 7, 0, z) * UU(p, d, 7, 0, z)) * ob8pi;
        UU(p, d, 4, 0, z) = UU(p, d, 4, 0, z) - aa + bb;
      }

  /* STEP 9.a (:5373-5430): zone electric field E = -v x B */
  for (long z = 0; z < d->NZ; ++z) {
    double vx = VV(p, d, 1, 0, z), vy = VV(p, d, 2, 0, z), vz = VV(p, d, 3, 0, z);
    double Bx = 0.0 + VV(p, d, 5, 0, z), By = 0.0 + VV(p, d, 6, 0, z), Bz = 0.0 + VV(p, d, 7, 0, z);
    p->Ex_zc[z] = -(vy * Bz - vz * By);
    p->Ey_zc[z] = -(vz * Bx - vx * Bz);
    p->Ez_zc[z] = -(vx * By - vy * Bx);
  }
  /* STEP 9.b (:5437-7100): WENO gradients of E, same stencils and weights */
  for (k = -2; k <= NR + 3; ++k)
    for (j = -2; j <= N + 3; ++j)
      for (i = -2; i <= N + 3; ++i) {
        if (!active(p, d, i, j, 1)) continue;
        long z = zi(d, i, j, k), z6[6];
        weno_mats W;
        six_nbrs(p, d, i, j, k, z6);
        weno_build(p, d, z, z6, &W);
        double* E[3] = {p->Ex_zc, p->Ey_zc, p->Ez_zc};
        for (c = 0; c < 3; ++c) {
          double f6[6], g3[3];
          for (q = 0; q < 6; ++q) f6[q] = E[c][z6[q]];
          weno_apply(&W, E[c][z], f6, p->smallnum, p->flatten[z], g3);
          for (m = 0; m < 3; ++m) E[c][z + d->NZ * (m + 1)] = g3[m];
        }
      }

  p1_step10(p, d);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* OneD_RS_MHD_Ptwise (600riemann_1d/610_1d_rs_oned_rs_mhd_ptwise.f:96-220,
 * 644-704), MHD, no eigensystem: every output of the routine.  The x-flux is
 * the ideal-MHD flux with B0 = 0 (the routine calls CONS_PRIM_TO_FLUX_3D, which
 * is absent from the reference tree; or_flux_x writes the flux of
 * 600_1d_rs_oned_rs_mhd_dad.f:194-230).  In the supersonic branches the state
 * U* is set to the upwind state before the one-sided fluctuation is formed
 * (:653-666), as there. */
void or_rs_ptwise(double gam, double rhof, double prsf, double* VL, double* VR, double entL, double entR,
                  double* SL, double* SR, double* Us, double* Fs, double* entF, double* FlL, double* FlR,
                  double* UL, double* UR, double* FL, double* FR) {
  const double pi = 4.0 * atan(1.0), ob8pi = 1.0 / (8.0 * pi);
  double Vm[8], a, b, t1, t2, t3, eU, pe, pg;
  int c;
  double bx = 0.5 * (VL[5] + VR[5]);
  VL[5] = bx;
  VR[5] = bx;
  or_prim_to_cons(gam, rhof, prsf, entL, VL, UL);
  or_prim_to_cons(gam, rhof, prsf, entR, VR, UR);
  or_flux_x(VL, UL, FL);
  or_flux_x(VR, UR, FR);
  *SR = -1.0e30;
  *SL = 1.0e30;
  or_wavspd_x(gam, VL, &a, &b);
  *SL = omin(*SL, a);
  *SR = omax(*SR, b);
  or_wavspd_x(gam, VR, &a, &b);
  *SL = omin(*SL, a);
  *SR = omax(*SR, b);
  for (c = 0; c < 8; ++c) Vm[c] = 0.5 * (VL[c] + VR[c]);
  or_wavspd_x(gam, Vm, &a, &b);
  *SL = omin(*SL, a);
  *SR = omax(*SR, b);
  t3 = 1.0 / (*SR - *SL);
  t1 = *SR * t3;
  t2 = *SL * t3;
  for (c = 0; c < 8; ++c) Us[c] = t1 * UR[c] - t2 * UL[c] - t3 * (FR[c] - FL[c]);
  eU = t1 * entR - t2 * entL - t3 * (entR * VR[1] - entL * VL[1]);
  pe = eU * pow(Us[0], gam - 1.0);
  pg = (gam - 1.0) * (Us[4] - 0.5 * (Us[1] * Us[1] + Us[2] * Us[2] + Us[3] * Us[3]) / Us[0] -
                      ob8pi * (Us[5] * Us[5] + Us[6] * Us[6] + Us[7] * Us[7]));
  if (pg <= prsf) Us[4] = Us[4] + (pe - pg) / (gam - 1.0);
  if (*SL >= 0.0) {
    for (c = 0; c < 8; ++c) {
      Us[c] = UL[c];
      Fs[c] = FL[c];
    }
    *entF = entL * VL[1];
    for (c = 0; c < 8; ++c) {
      FlR[c] = *SL * (Us[c] - UL[c]) + *SR * (UR[c] - Us[c]);
      FlL[c] = 0.0;
    }
  } else if (*SR <= 0.0) {
    for (c = 0; c < 8; ++c) {
      Us[c] = UR[c];
      Fs[c] = FR[c];
    }
    *entF = entR * VR[1];
    for (c = 0; c < 8; ++c) {
      FlL[c] = *SL * (Us[c] - UL[c]) + *SR * (UR[c] - Us[c]);
      FlR[c] = 0.0;
    }
  } else {
    double o = 1.0 / (*SR - *SL), s2 = *SR * *SL * o;
    for (c = 0; c < 8; ++c) {
      Fs[c] = 0.5 * (FL[c] + FR[c]) + 0.5 * *SL * (Us[c] - UL[c]) - 0.5 * *SR * (UR[c] - Us[c]);
      FlL[c] = *SL * (Us[c] - UL[c]);
      FlR[c] = *SR * (UR[c] - Us[c]);
    }
    *entF = (*SR * entL * VL[1] - *SL * entR * VR[1]) * o + (entR - entL) * s2;
  }
}

/* vector components 1-3 and 5-7 from the global frame to (A, B, C) */
static void to_frame(const double* A, const double* B, const double* C, double* v) {
  for (int o = 1; o <= 5; o += 4) {
    double a = A[0] * v[o] + A[1] * v[o + 1] + A[2] * v[o + 2];
    double b = B[0] * v[o] + B[1] * v[o + 1] + B[2] * v[o + 2];
    double c = C[0] * v[o] + C[1] * v[o + 1] + C[2] * v[o + 2];
    v[o] = a;
    v[o + 1] = b;
    v[o + 2] = c;
  }
}
/* and back: out = f with the vector parts rotated into the global frame */
static void from_frame(const double* A, const double* B, const double* C, const double* f, double* out) {
  for (int c = 0; c < 8; ++c) out[c] = f[c];
  for (int o = 1; o <= 5; o += 4)
    for (int q = 0; q < 3; ++q) out[o + q] = A[q] * f[o] + B[q] * f[o + 1] + C[q] * f[o + 2];
}

/* The face frame (A normal, B, C) of 450...f:507-1152 (x), :1212-1883 (y),
 * :1940-2910 (z): x-faces the normal, the centroid's radial direction C and
 * B = C x A; y-faces the normal, the centroid's radial direction B and
 * C = A x B; z-faces the first point's normal, the vertex-difference
 * direction made orthogonal, and their cross product. */
static void face_frame(const or_p1_flux* f, const p1dims* d, int fam, int i, int j, int k, long fg, long NF,
                       const double* nrm, const double* fcen, double* A, double* B, double* C) {
  const double sm = f->smallnum;
  double t;
  int c;
  if (fam == 0) {
    for (c = 0; c < 3; ++c) {
      A[c] = nrm[fg + NF * c];
      C[c] = fcen[fg + NF * c];
    }
    t = 1.0 / omax(sqrt(C[0] * C[0] + C[1] * C[1] + C[2] * C[2]), sm);
    for (c = 0; c < 3; ++c) C[c] = C[c] * t;
    B[0] = C[1] * A[2] - C[2] * A[1];
    B[1] = C[2] * A[0] - C[0] * A[2];
    B[2] = C[0] * A[1] - C[1] * A[0];
  } else if (fam == 1) {
    for (c = 0; c < 3; ++c) {
      A[c] = nrm[fg + NF * c];
      B[c] = fcen[fg + NF * c];
    }
    t = 1.0 / omax(sqrt(B[0] * B[0] + B[1] * B[1] + B[2] * B[2]), sm);
    for (c = 0; c < 3; ++c) B[c] = B[c] * t;
    C[0] = A[1] * B[2] - A[2] * B[1];
    C[1] = A[2] * B[0] - A[0] * B[2];
    C[2] = A[0] * B[1] - A[1] * B[0];
  } else {
    for (c = 0; c < 3; ++c) A[c] = nrm[fg + NF * c]; /* the first (central) point's normal */
    t = 1.0 / omax(sqrt(A[0] * A[0] + A[1] * A[1] + A[2] * A[2]), sm);
    for (c = 0; c < 3; ++c) A[c] = A[c] * t;
    for (c = 0; c < 3; ++c)
      B[c] = f->vert[fci(d, i, j, k) + d->NV * c] + f->vert[fci(d, i, j - 1, k) + d->NV * c] -
             f->vert[fci(d, i - 1, j, k) + d->NV * c] - f->vert[fci(d, i - 1, j - 1, k) + d->NV * c];
    t = B[0] * A[0] + B[1] * A[1] + B[2] * A[2];
    for (c = 0; c < 3; ++c) B[c] = B[c] - t * A[c];
    t = 1.0 / omax(sqrt(B[0] * B[0] + B[1] * B[1] + B[2] * B[2]), sm);
    for (c = 0; c < 3; ++c) B[c] = B[c] * t;
    C[0] = A[1] * B[2] - A[2] * B[1];
    C[1] = A[2] * B[0] - A[0] * B[2];
    C[2] = A[0] * B[1] - A[1] * B[0];
    t = 1.0 / omax(sqrt(C[0] * C[0] + C[1] * C[1] + C[2] * C[2]), sm);
    for (c = 0; c < 3; ++c) C[c] = C[c] * t;
    B[0] = C[1] * A[2] - C[2] * A[1];
    B[1] = C[2] * A[0] - C[0] * A[2];
    B[2] = C[0] * A[1] - C[1] * A[0];
  }
}

/* MAKE_FACIAL_FLUXES (400riemann_integrate/450make_facial_fluxes.f:5) at
 * second order (one Riemann problem per face, at the face centroid, weight 1),
 * MHD, no facial electric fields, no B0. */
int or_make_facial_fluxes_p1(const or_p1_flux* f) {
  if (!f || f->nghost < 4 || f->ntp < 1 || f->nr < 1) return 1;
  or_p1_patch pp;
  memset(&pp, 0, sizeof pp);
  pp.ntp = f->ntp;
  pp.nr = f->nr;
  pp.nghost = f->nghost;
  pp.mask = f->mask;
  pp.stencil = f->stencil;
  pp.scale = (double*)f->scale;
  pp.v_centroid = f->v_centroid;
  pp.V = (double*)f->V;
  const p1dims D = p1_dims(&pp), *d = &D;
  const int N = d->N, NR = d->NR, g = d->g;
  const double gam = f->gamma, sm = f->smallnum;
  int i, j, k, c, q;
  /* initial values (450...f:370-480) */
  for (long z = 0; z < d->NZ * 8; ++z) f->dU_dt[z] = 0.0;
  for (long z = 0; z < d->NFC * 8; ++z) {
    f->Fluct_x_Pl[z] = f->Fluct_y_Pl[z] = f->Fluct_z_Pl[z] = 0.0;
    f->Fluct_x_Mi[z] = f->Fluct_y_Mi[z] = f->Fluct_z_Mi[z] = 0.0;
  }
  for (k = 1 - g; k <= NR + g; ++k)
    for (j = 1 - g; j <= N + g; ++j)
      for (i = 1 - g; i <= N + g; ++i) {
        long fb = fci(d, i, j, k);
        f->Efx[fb] = f->Efy[fb] = f->Efz[fb] = 0.0;
      }
  for (int fam = 0; fam < 3; ++fam) {
    const int ilo = fam == 0 ? -1 : 0, jlo = fam == 1 ? -1 : 0, klo = fam == 2 ? -1 : 0;
    double* SLg = fam == 0 ? f->S_L : (fam == 1 ? f->S_D : f->S_B);
    double* SRg = fam == 0 ? f->S_R : (fam == 1 ? f->S_U : f->S_T);
    double* FPl = fam == 0 ? f->Fluct_x_Pl : (fam == 1 ? f->Fluct_y_Pl : f->Fluct_z_Pl);
    double* FMi = fam == 0 ? f->Fluct_x_Mi : (fam == 1 ? f->Fluct_y_Mi : f->Fluct_z_Mi);
    double* EF = fam == 0 ? f->EntFlux_x : (fam == 1 ? f->EntFlux_y : f->EntFlux_z);
    const double* pts = fam == 0 ? f->xf_pts : (fam == 1 ? f->yf_pts : f->zf_pts);
    const double* nrm = fam == 0 ? f->xf_normal : (fam == 1 ? f->yf_normal : f->zf_normal);
    const double* area = fam == 0 ? f->xf_area : (fam == 1 ? f->yf_area : f->zf_area);
    const double* fcen = fam == 0 ? f->xf_centroid : (fam == 1 ? f->yf_centroid : f->zf_centroid);
    const long NF = fam == 2 ? d->NZF : d->NXF;
    for (k = klo; k <= NR + 1; ++k)
      for (j = jlo; j <= N + 1; ++j)
        for (i = ilo; i <= N + 1; ++i) {
          long fb = fci(d, i, j, k), fg = fam == 0 ? xfi(d, i, j, k) : (fam == 1 ? yfi(d, i, j, k) : zfi(d, i, j, k));
          int li = i, lj = j, ri = i + (fam == 0), rj = j + (fam == 1), rk = k + (fam == 2);
          if (fam < 2) {
            int al = active(&pp, d, li, lj, k), ar = active(&pp, d, ri, rj, k);
            if (al && ar) {
            } else if (!al && ar) {
              nbr(&pp, d, ri, rj, fam == 0 ? 1 : 3, &li, &lj);
            } else if (al && !ar) {
              nbr(&pp, d, li, lj, fam == 0 ? 0 : 2, &ri, &rj);
            } else {
              continue;
            }
          } else if (!active(&pp, d, i, j, k)) {
            continue;
          }
          long zl = zi(d, li, lj, k), zr = zi(d, ri, rj, fam == 2 ? rk : k);
          double A[3], B[3], C[3];
          face_frame(f, d, fam, i, j, k, fg, NF, nrm, fcen, A, B, C);
          /* the two zones' linear fields at the face centroid (point 1) */
          double rl[3], rr[3], VL[8], VR[8], eL, eR;
          rel(&pp, d, zl, pts + fg, NF, obl_of(&pp, zl), rl);
          rel(&pp, d, zr, pts + fg, NF, obl_of(&pp, zr), rr);
          for (c = 0; c < 8; ++c) {
            VL[c] = VV(&pp, d, c, 0, zl) + VV(&pp, d, c, 1, zl) * rl[0] + VV(&pp, d, c, 2, zl) * rl[1] +
                    VV(&pp, d, c, 3, zl) * rl[2];
            VR[c] = VV(&pp, d, c, 0, zr) + VV(&pp, d, c, 1, zr) * rr[0] + VV(&pp, d, c, 2, zr) * rr[1] +
                    VV(&pp, d, c, 3, zr) * rr[2];
          }
          eL = f->ent[zl] + f->ent[zl + d->NZ] * rl[0] + f->ent[zl + 2 * d->NZ] * rl[1] + f->ent[zl + 3 * d->NZ] * rl[2];
          eR = f->ent[zr] + f->ent[zr + d->NZ] * rr[0] + f->ent[zr + 2 * d->NZ] * rr[1] + f->ent[zr + 3 * d->NZ] * rr[2];
          to_frame(A, B, C, VL);
          to_frame(A, B, C, VR);
          double SL, SR, Us[8], Fs[8], eF, FlL[8], FlR[8], UL[8], UR[8], FL[8], FR[8];
          or_rs_ptwise(gam, f->rhofloor, f->prsfloor, VL, VR, eL, eR, &SL, &SR, Us, Fs, &eF, FlL, FlR, UL, UR, FL, FR);
          SLg[fb] = SL;
          SRg[fb] = SR;
          double Pl[8], Mi[8], FLL[8], FRR[8];
          from_frame(A, B, C, FlR, Pl);
          from_frame(A, B, C, FlL, Mi);
          from_frame(A, B, C, FL, FLL);
          from_frame(A, B, C, FR, FRR);
          for (c = 0; c < 8; ++c) {
            FPl[fb + d->NFC * c] = FPl[fb + d->NFC * c] + 1.0 * Pl[c];
            FMi[fb + d->NFC * c] = FMi[fb + d->NFC * c] + 1.0 * Mi[c];
          }
          const double ar_ = area[fg];
          for (c = 0; c < 8; ++c) {
            f->dU_dt[zl + d->NZ * c] = f->dU_dt[zl + d->NZ * c] - 1.0 * (Mi[c] + FLL[c]) * ar_ / f->v_volume[zl];
            f->dU_dt[zr + d->NZ * c] = f->dU_dt[zr + d->NZ * c] - 1.0 * (Pl[c] - FRR[c]) * ar_ / f->v_volume[zr];
          }
          EF[fb] = eF;
          (void)q;
        }
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* UPDATE_IMEX_RK_STEP_P1 (400riemann_integrate/484update_imex_rk_step_p1.f:7),
 * MHD: every zone's U += dt dU_dt (:483-489); the owned zones' entropy from the
 * face entropy fluxes, then CONS_TO_PRIM and its flag (:540-628); constrained
 * transport of the owned faces with the edges' arc lengths (:630-906). */
int or_update_imex_rk_step_p1(const or_p1_update* u) {
  if (!u || u->nghost < 1 || u->ntp < 1 || u->nr < 1) return 1;
  or_p1_patch pp;
  memset(&pp, 0, sizeof pp);
  pp.ntp = u->ntp;
  pp.nr = u->nr;
  pp.nghost = u->nghost;
  const p1dims D = p1_dims(&pp), *d = &D;
  const int N = d->N, NR = d->NR, g = d->g;
  const double dt = u->dtmove;
  int i, j, k, c;
  for (k = 1 - g; k <= NR + g; ++k)
    for (j = 1 - g; j <= N + g; ++j)
      for (i = 1 - g; i <= N + g; ++i) {
        long z = zi(d, i, j, k);
        for (c = 0; c < 8; ++c) u->U[z + d->NZ * c] = u->U[z + d->NZ * c] + dt * u->dU_dt[z + d->NZ * c];
      }
  for (long z = 0; z < d->NZ; ++z) u->flag_entprs[z] = 0.0;
  for (k = 1; k <= NR; ++k)
    for (j = 1; j <= N; ++j)
      for (i = 1; i <= N; ++i) {
        long z = zi(d, i, j, k);
        double s = u->xf_area[xfi(d, i, j, k)] * u->EntFlux_x[fci(d, i, j, k)] -
                   u->xf_area[xfi(d, i - 1, j, k)] * u->EntFlux_x[fci(d, i - 1, j, k)] +
                   u->yf_area[yfi(d, i, j, k)] * u->EntFlux_y[fci(d, i, j, k)] -
                   u->yf_area[yfi(d, i, j - 1, k)] * u->EntFlux_y[fci(d, i, j - 1, k)] +
                   u->zf_area[zfi(d, i, j, k)] * u->EntFlux_z[fci(d, i, j, k)] -
                   u->zf_area[zfi(d, i, j, k - 1)] * u->EntFlux_z[fci(d, i, j, k - 1)];
        u->ent[z] = u->ent[z] - dt * s / u->v_volume[z];
        double Up[8], Vp[8], e = u->ent[z];
        for (c = 0; c < 8; ++c) Up[c] = u->U[z + d->NZ * c];
        int flag = or_cons_to_prim(u->gamma, u->rhofloor, u->prsfloor, &e, Up, Vp);
        for (c = 0; c < 8; ++c) u->U[z + d->NZ * c] = Up[c];
        u->ent[z] = e;
        u->flag_entprs[z] = (double)flag;
      }
  /* vertex (i, j, k) and its unit direction scaled by 1/r of another vertex */
#define VX(ii, jj, kk, cc) u->vert[fci(d, ii, jj, kk) + d->NV * (cc)]
#define RAD(ii, jj, kk) sqrt(VX(ii, jj, kk, 0) * VX(ii, jj, kk, 0) + VX(ii, jj, kk, 1) * VX(ii, jj, kk, 1) + \
                             VX(ii, jj, kk, 2) * VX(ii, jj, kk, 2))
  for (k = 1; k <= NR; ++k)
    for (j = 1; j <= N; ++j)
      for (i = 0; i <= N; ++i) {
        double rout = RAD(i, j, k), rin = RAD(i, j, k - 1), dr = rout - rin, t = 1.0 / rout;
        double dot = (VX(i, j, k, 0) * t) * (VX(i, j - 1, k, 0) * t) + (VX(i, j, k, 1) * t) * (VX(i, j - 1, k, 1) * t) +
                     (VX(i, j, k, 2) * t) * (VX(i, j - 1, k, 2) * t);
        double th = acos(dot), lin = rin * th, lout = rout * th;
        long f = fci(d, i, j, k);
        u->Bxf[f] = u->Bxf[f] - dt * (dr * u->Efz[f] - dr * u->Efz[fci(d, i, j - 1, k)] - lout * u->Efy[f] +
                                      lin * u->Efy[fci(d, i, j, k - 1)]) / u->xf_area[xfi(d, i, j, k)];
      }
  for (k = 1; k <= NR; ++k)
    for (j = 0; j <= N; ++j)
      for (i = 1; i <= N; ++i) {
        double rout = RAD(i, j, k), rin = RAD(i, j, k - 1), dr = rout - rin, t = 1.0 / rout;
        double dot = (VX(i, j, k, 0) * t) * (VX(i - 1, j, k, 0) * t) + (VX(i, j, k, 1) * t) * (VX(i - 1, j, k, 1) * t) +
                     (VX(i, j, k, 2) * t) * (VX(i - 1, j, k, 2) * t);
        double th = acos(dot), lin = rin * th, lout = rout * th;
        long f = fci(d, i, j, k);
        u->Byf[f] = u->Byf[f] - dt * (-(dr * u->Efz[f]) + dr * u->Efz[fci(d, i - 1, j, k)] + lout * u->Efx[f] -
                                      lin * u->Efx[fci(d, i, j, k - 1)]) / u->yf_area[yfi(d, i, j, k)];
      }
  for (k = 0; k <= NR; ++k)
    for (j = 1; j <= N; ++j)
      for (i = 1; i <= N; ++i) {
        double rout = RAD(i, j, k), t = 1.0 / rout;
        /* arc between two vertices of this radial level, both scaled by 1/rout */
#define ARC(a1, b1, a2, b2)                                                                                  \
  (rout * acos((VX(a1, b1, k, 0) * t) * (VX(a2, b2, k, 0) * t) + (VX(a1, b1, k, 1) * t) * (VX(a2, b2, k, 1) * t) + \
               (VX(a1, b1, k, 2) * t) * (VX(a2, b2, k, 2) * t)))
        double lxp = ARC(i, j, i - 1, j), lxm = ARC(i, j - 1, i - 1, j - 1), lyp = ARC(i, j, i, j - 1),
               lym = ARC(i - 1, j, i - 1, j - 1);
#undef ARC
        long f = fci(d, i, j, k);
        u->Bzf[f] = u->Bzf[f] - dt * (lyp * u->Efy[f] - lym * u->Efy[fci(d, i - 1, j, k)] - lxp * u->Efx[f] +
                                      lxm * u->Efx[fci(d, i, j - 1, k)]) / u->zf_area[zfi(d, i, j, k)];
      }
#undef VX
#undef RAD
  return 0;
}

/* ------------------------------------------------------------------------ */
/* CLEANUP_AFTER_UPDATE (400riemann_integrate/485cleanup_after_update.f:4), MHD:
 * the constants and the safe tri-corner zones of STEP 2 (:305-490), the
 * tri-corner face-B sync of STEP 3 (:491-686), the zone B as the least-squares
 * fit of the six face values over the face normals of STEP 4.a (:691-865, the
 * branch the routine compiles) with the energy rebalanced, and CONS_TO_PRIM of
 * every zone (STEP 5, :1180-1280). */
int or_cleanup_after_update_p1(const or_p1_patch* p) {
  if (!p || p->nghost < 1 || p->ntp < 1 || p->nr < 1) return 1;
  const p1dims D = p1_dims(p), *d = &D;
  const int N = d->N, NR = d->NR, g = d->g;
  const double pi = 4.0 * atan(1.0), ob8pi = 1.0 / (8.0 * pi), gam = p->gamma;
  int i, j, k, c, m;
  for (k = 1 - g; k <= NR + g; ++k)
    for (j = 1 - g; j <= N + g; ++j)
      for (i = 1 - g; i <= N + g; ++i) {
        long z = zi(d, i, j, k);
        p->gamma_eos[z] = gam;
        p->gamma_sound[z] = gam;
        p->mean_mol_wt[z] = p->molewt;
        for (c = 0; c < 8; ++c)
          for (m = 1; m < 4; ++m) UU(p, d, c, m, z) = 0.0;
        for (m = 1; m < 4; ++m) p->ent[z + d->NZ * m] = 0.0;
        for (c = 0; c < 8; ++c)
          for (m = 0; m < 4; ++m) VV(p, d, c, m, z) = 0.0;
        if (p->mask[z] == 0) {
          for (c = 0; c < 8; ++c) UU(p, d, c, 0, z) = 0.0;
          UU(p, d, 0, 0, z) = p->rhofloor;
          UU(p, d, 4, 0, z) = p->prsfloor / (gam - 1.0);
          p->scale[z] = p->smallnum;
        }
      }
  for (k = 1 - g; k <= NR + g; ++k)
    for (j = 1 - g; j <= N + g; ++j)
      for (i = -g; i <= N + g; ++i) {
        int a = i < 1 - g ? 1 - g : i, b = i + 1 > N + g ? N + g : i + 1;
        if (!active(p, d, a, j, k) && !active(p, d, b, j, k)) p->Bxf[fci(d, i, j, k)] = 0.0;
      }
  for (k = 1 - g; k <= NR + g; ++k)
    for (j = -g; j <= N + g; ++j)
      for (i = 1 - g; i <= N + g; ++i) {
        int a = j < 1 - g ? 1 - g : j, b = j + 1 > N + g ? N + g : j + 1;
        if (!active(p, d, i, a, k) && !active(p, d, i, b, k)) p->Byf[fci(d, i, j, k)] = 0.0;
      }
  for (k = -g; k <= NR + g; ++k)
    for (j = 1 - g; j <= N + g; ++j)
      for (i = 1 - g; i <= N + g; ++i)
        if (!active(p, d, i, j, 1)) p->Bzf[fci(d, i, j, k)] = 0.0;
  for (int corner = 0; corner < 4; ++corner) {
    if (!p->itri[corner]) continue;
    int jj = (corner < 2) ? 1 : N, lo = (corner == 0 || corner == 3) ? 1 - g : N + 1,
        hi = (corner == 0 || corner == 3) ? 0 : N + g;
    for (i = lo; i <= hi; ++i) {
      int a, b;
      nbr(p, d, i, jj, corner < 2 ? 3 : 2, &a, &b);
      for (k = 1 - g; k <= NR + g; ++k) {
        double* by = &p->Byf[fci(d, i, corner < 2 ? jj - 1 : jj, k)];
        double* bx = &p->Bxf[fci(d, (corner == 0 || corner == 3) ? a - 1 : a, b, k)];
        if (corner == 0 || corner == 2) {
          double t = 0.5 * (*by - *bx);
          *by = t;
          *bx = -t;
        } else {
          double t = 0.5 * (*by + *bx);
          *by = t;
          *bx = t;
        }
      }
    }
  }
  for (k = 1 - g; k <= NR + g; ++k)
    for (j = 1 - g; j <= N + g; ++j)
      for (i = 1 - g; i <= N + g; ++i) {
        if (!active(p, d, i, j, 1)) continue;
        long z = zi(d, i, j, k);
        double nv[3], A6[6][3], R6[6], A33[3][3], R3[3], t;
        long fz = zfi(d, i, j, k);
        for (c = 0; c < 3; ++c) nv[c] = p->zf_centroid[fz + d->NZF * c];
        t = 1.0 / omax(p->smallnum, sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]));
        for (c = 0; c < 3; ++c) nv[c] = nv[c] * t;
        for (c = 0; c < 3; ++c) {
          A6[0][c] = p->xf_normal[xfi(d, i, j, k) + d->NXF * c];
          A6[1][c] = p->xf_normal[xfi(d, i - 1, j, k) + d->NXF * c];
          A6[2][c] = p->yf_normal[yfi(d, i, j, k) + d->NYF * c];
          A6[3][c] = p->yf_normal[yfi(d, i, j - 1, k) + d->NYF * c];
          A6[4][c] = nv[c];
          A6[5][c] = nv[c];
        }
        R6[0] = p->Bxf[fci(d, i, j, k)];
        R6[1] = p->Bxf[fci(d, i - 1, j, k)];
        R6[2] = p->Byf[fci(d, i, j, k)];
        R6[3] = p->Byf[fci(d, i, j - 1, k)];
        R6[4] = p->Bzf[fci(d, i, j, k)];
        R6[5] = p->Bzf[fci(d, i, j, k - 1)];
        for (int a = 0; a < 3; ++a) {
          for (int b = 0; b < 3; ++b) {
            double s = 0.0;
            for (int r = 0; r < 6; ++r) s = s + A6[r][a] * A6[r][b];
            A33[a][b] = s;
          }
          double s = 0.0;
          for (int r = 0; r < 6; ++r) s = s + A6[r][a] * R6[r];
          R3[a] = s;
        }
        double a11 = A33[0][0], a12 = A33[0][1], a13 = A33[0][2], a21 = A33[1][0], a22 = A33[1][1], a23 = A33[1][2],
               a31 = A33[2][0], a32 = A33[2][1], a33 = A33[2][2];
        double det = a11 * (a22 * a33 - a32 * a23) - a12 * (a21 * a33 - a31 * a23) + a13 * (a21 * a32 - a31 * a22);
        double od = (signbit(det) ? -1.0 : 1.0) / omax(1.0e-32, fabs(det));
        double b11 = (a22 * a33 - a23 * a32) * od, b12 = (a13 * a32 - a12 * a33) * od, b13 = (a12 * a23 - a22 * a13) * od;
        double b21 = (a23 * a31 - a21 * a33) * od, b22 = (a11 * a33 - a13 * a31) * od, b23 = (a13 * a21 - a11 * a23) * od;
        double b31 = (a21 * a32 - a22 * a31) * od, b32 = (a12 * a31 - a11 * a32) * od, b33 = (a11 * a22 - a12 * a21) * od;
        double Bx = UU(p, d, 5, 0, z), By = UU(p, d, 6, 0, z), Bz = UU(p, d, 7, 0, z);
        UU(p, d, 5, 0, z) = b11 * R3[0] + b12 * R3[1] + b13 * R3[2];
        UU(p, d, 6, 0, z) = b21 * R3[0] + b22 * R3[1] + b23 * R3[2];
        UU(p, d, 7, 0, z) = b31 * R3[0] + b32 * R3[1] + b33 * R3[2];
        UU(p, d, 4, 0, z) = UU(p, d, 4, 0, z) - (Bx * Bx + By * By + Bz * Bz) * ob8pi +
                            (UU(p, d, 5, 0, z) * UU(p, d, 5, 0, z) + UU(p, d, 6, 0, z) * UU(p, d, 6, 0, z) +
                             UU(p, d, 7, 0, z) * UU(p, d, 7, 0, z)) * ob8pi;
        p->B1_zc[z] = UU(p, d, 5, 0, z);
        p->B2_zc[z] = UU(p, d, 6, 0, z);
        p->B3_zc[z] = UU(p, d, 7, 0, z);
        for (c = 5; c < 8; ++c) VV(p, d, c, 0, z) = UU(p, d, c, 0, z);
      }
  for (k = 1 - g; k <= NR + g; ++k)
    for (j = 1 - g; j <= N + g; ++j)
      for (i = 1 - g; i <= N + g; ++i) {
        long z = zi(d, i, j, k);
        double Up[8], Vp[8], e = p->ent[z];
        for (c = 0; c < 8; ++c) Up[c] = UU(p, d, c, 0, z);
        or_cons_to_prim(gam, p->rhofloor, p->prsfloor, &e, Up, Vp);
        for (c = 0; c < 8; ++c) {
          VV(p, d, c, 0, z) = Vp[c];
          UU(p, d, c, 0, z) = Up[c];
        }
        p->ent[z] = e;
      }
  return 0;
}

/* EVALUATE_TIMESTEP (400riemann_integrate/490evaluate_timestep.f:7-306), MHD:
 * CONS_TO_PRIM of the owned active zones (U, V mode 1 rewritten), their
 * magnetosonic speed, and dt_new = min(1.1 dt, cfl l / (|v| + c_f)). */
int or_evaluate_timestep_p1(const or_p1_patch* p, double cfl, double dtcur, double* dtnew) {
  if (!p || !dtnew) return 1;
  const p1dims D = p1_dims(p), *d = &D;
  const double pi = 4.0 * atan(1.0), ob4pi = 1.0 / (4.0 * pi);
  double dt = 1.1 * dtcur;
  for (int k = 1; k <= d->NR; ++k)
    for (int j = 1; j <= d->N; ++j)
      for (int i = 1; i <= d->N; ++i) {
        long z = zi(d, i, j, k);
        if (p->mask[z] == 0) continue;
        double Up[8], Vp[8], e = p->ent[z];
        int c;
        for (c = 0; c < 8; ++c) Up[c] = UU(p, d, c, 0, z);
        or_cons_to_prim(p->gamma, p->rhofloor, p->prsfloor, &e, Up, Vp);
        for (c = 0; c < 8; ++c) {
          UU(p, d, c, 0, z) = Up[c];
          VV(p, d, c, 0, z) = Vp[c];
        }
        double b2 = (Vp[5] + 0.0) * (Vp[5] + 0.0) + (Vp[6] + 0.0) * (Vp[6] + 0.0) + (Vp[7] + 0.0) * (Vp[7] + 0.0);
        double ms = p->gamma_sound[z] * Vp[4] + ob4pi * b2;
        p->msonic[z] = sqrt(ms / Vp[0]);
        double a21 = sqrt(Vp[1] * Vp[1] + Vp[2] * Vp[2] + Vp[3] * Vp[3]) + p->msonic[z];
        dt = omin(dt, cfl * p->scale[z] / a21);
      }
  *dtnew = dt;
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Two_D_RS_Ez_3State_HLL (620riemann_2d/631_2d_rs_ez_two_d_rs_ez_3state_hll.f:1-216):
 * the 2-D HLL electric field where three zones meet, each pair across a face
 * at 120 degrees in the solver's frame.  sr, sl: the speeds of the three
 * faces (13, 21, 32); ez: the zones' E; bn: the faces' normal B.  The
 * operations keep the routine's order so that results match it bit for bit. */
double or_emf3_hll(const double* sr, const double* sl, const double* ez, const double* bn) {
  const double SR1 = sr[0], SL1 = sl[0], SR2 = sr[1], SL2 = sl[1], SR3 = sr[2], SL3 = sl[2];
  const double E1 = ez[0], E2 = ez[1], E3 = ez[2], Bn1 = bn[0], Bn2 = bn[1], Bn3 = bn[2];
  const double h3 = sqrt(3.0) / 2.0, r3 = 1.0 / sqrt(3.0);
  const double o1 = 1.0 / (SR1 - SL1), o2 = 1.0 / (SR2 - SL2), o3 = 1.0 / (SR3 - SL3);
  /* corners A, B, C of the wave triangle, its centre, sides, outward normals */
  const double Ax = (SL1 + 2.0 * SR2) * r3, Ay = -SL1, Bx = (SL2 - SR3) * r3, By = SL2 + SR3;
  const double Cx = -((SR1 + 2.0 * SL3) * r3), Cy = -SR1;
  const double xc = (Ax + Bx + Cx) / 3.0, yc = (Ay + By + Cy) / 3.0;
  const double l1x = Ax - Cx, l2x = Bx - Ax, l3x = Cx - Bx, l1y = Ay - Cy, l2y = By - Ay, l3y = Cy - By;
  const double l1 = sqrt(l1x * l1x + l1y * l1y), l2 = sqrt(l2x * l2x + l2y * l2y), l3 = sqrt(l3x * l3x + l3y * l3y);
  const double n13x = l1y / l1, n13y = -(l1x / l1), n21x = l2y / l2, n21y = -(l2x / l2), n32x = l3y / l3,
               n32y = -(l3x / l3);
  const double S13 = n13x * Ax + n13y * Ay, S21 = n21x * Bx + n21y * By, S32 = n32x * Cx + n32y * Cy;
  const double xb13 = 0.5 * (Ax + Cx), yb13 = 0.5 * (Ay + Cy), xb21 = 0.5 * (Bx + Ax), yb21 = 0.5 * (By + Ay);
  const double xb32 = 0.5 * (Cx + Bx), yb32 = 0.5 * (Cy + By);
  const double s = (l1 + l2 + l3) / 2.0, area = sqrt(s * (s - l1) * (s - l2) * (s - l3));
  /* the three zones' field in the solver frame and along each face */
  const double Bx1 = (2.0 * Bn2 + Bn1) * r3, Bx2 = (Bn2 - Bn3) * r3, Bx3 = -((2.0 * Bn3 + Bn1) * r3);
  const double By1 = -Bn1, By2 = Bn2 + Bn3;
  const double t1U1 = Bx1, t1U3 = Bx3, t2U1 = -0.5 * Bx1 + h3 * By1, t2U2 = -0.5 * Bx2 + h3 * By2;
  const double t3U2 = -0.5 * Bx2 - h3 * By2, t3U3 = -0.5 * Bx3 - h3 * By1;
  /* the 1-D HLL state of each face */
  const double Bt13 = (SR1 * t1U3 - SL1 * t1U1 + E3 - E1) * o1;
  const double Ez13 = (SR1 * E1 - SL1 * E3 - SR1 * SL1 * (t1U3 - t1U1)) * o1;
  const double Bx13 = Bt13, By13 = -Bn1;
  const double Bt21 = (SR2 * t2U1 - SL2 * t2U2 + E1 - E2) * o2;
  const double Ez21 = (SR2 * E2 - SL2 * E1 - SR2 * SL2 * (t2U1 - t2U2)) * o2;
  const double Bx21 = -0.5 * Bt21 + h3 * Bn2, By21 = h3 * Bt21 + 0.5 * Bn2;
  const double Bt32 = (SR3 * t3U2 - SL3 * t3U3 + E2 - E3) * o3;
  const double Ez32 = (SR3 * E3 - SL3 * E2 - SR3 * SL3 * (t3U2 - t3U3)) * o3;
  const double Bx32 = -0.5 * Bt32 - h3 * Bn3, By32 = -(h3 * Bt32) + 0.5 * Bn3;
  if (SL1 >= 0.0 && SR2 <= 0.0) return E1;
  if (SL2 >= 0.0 && SR3 <= 0.0) return E2;
  if (SL3 >= 0.0 && SR1 <= 0.0) return E3;
  if (SL1 * (SR1 + 2.0 * SL3) + SR1 * (SL1 + 2.0 * SR2) <= 0) return Ez13;
  if ((SL2 + SR3) * (SL1 + 2.0 * SR2) + (SL2 - SR3) * SL1 <= 0) return Ez21;
  if ((SL2 + SR3) * (SR1 + 2.0 * SL3) - (SL2 - SR3) * SR1 <= 0) return Ez32;
  /* the strongly interacting state: the face states' fluxes around the triangle */
  const double fx13 = n13y * Ez13 - S13 * Bx13, fx21 = n21y * Ez21 - S21 * Bx21, fx32 = n32y * Ez32 - S32 * Bx32;
  const double fy13 = -(n13x * Ez13) - S13 * By13, fy21 = -(n21x * Ez21) - S21 * By21,
               fy32 = -(n32x * Ez32) - S32 * By32;
  const double BSSx = (fx13 * l1 + fx21 * l2 + fx32 * l3) / (-2.0 * area);
  const double BSSy = (fy13 * l1 + fy21 * l2 + fy32 * l3) / (-2.0 * area);
  const double e1 = -(xc * BSSy) -
                    ((xb13 - xc) * fy13 * l1 + (xb21 - xc) * fy21 * l2 + (xb32 - xc) * fy32 * l3) / area;
  const double e2 = yc * BSSx + ((yb13 - yc) * fx13 * l1 + (yb21 - yc) * fx21 * l2 + (yb32 - yc) * fx32 * l3) / area;
  return 0.5 * (e1 + e2);
}

/* MAKE_EDGE_EFIELDS_P1 (400riemann_integrate/460make_edge_fields_p1.f:5) at
 * second order (NUMORDER 2, no facial electric fields, no B0): every edge of
 * the owned faces takes the mean E of its four zones along the edge's
 * tangent and the 2-D HLL field of Two_D_RS_Ez_4State_HLL (630) with the
 * face-point normal B of the four faces around it (from the reconstruction)
 * and their wave speeds (from the facial fluxes), all four symmetrised to
 * +-max|S|; z-edges with an inactive zone are skipped, and at the tri-corners
 * the three zones' edge takes the 3-state HLL (631).  Efx, Efy, Efz are set
 * to zero first over their whole arrays. */
int or_make_edge_efields_p1(const or_p1_edge* e) {
  if (!e || e->nghost < 2 || e->ntp < 1 || e->nr < 1) return 1;
  or_p1_patch pp;
  memset(&pp, 0, sizeof(pp));
  pp.ntp = e->ntp;
  pp.nr = e->nr;
  pp.nghost = e->nghost;
  const p1dims D = p1_dims(&pp), *d = &D;
  const int N = d->N, NR = d->NR;
  const long NFC = d->NFC, NZ = d->NZ;
  long q;
  int i, j, k, c;
  for (q = 0; q < NFC; ++q) {
    e->Efx[q] = 0.0;
    e->Efy[q] = 0.0;
    e->Efz[q] = 0.0;
  }
#define EV(ii, jj, kk, cc) e->vert[fci(d, ii, jj, kk) + NFC * (cc)]
#define BQ(a, ii, jj, kk, pt) e->a[fci(d, ii, jj, kk) + NFC * ((pt)-1)]
#define SG(a, ii, jj, kk) e->a[fci(d, ii, jj, kk)]
  /* the mean E of four zones (listed LD, RD, RU, LU) along t */
  double Eq[4], ez4[4], s4[4], bn2[2], by2[2], t3[3], n[3], ec[3], a, b;
  long zq[4];
  for (int fam = 0; fam < 3; ++fam) {
    const int i0 = fam == 0 ? 1 : 0, j0 = fam == 1 ? 1 : 0, k0 = fam == 2 ? 1 : 0;
    for (k = k0; k <= NR; ++k)
      for (j = j0; j <= N; ++j)
        for (i = i0; i <= N; ++i) {
          int zz[4][3];
          if (fam == 0) { /* LD (i,j,k) RD (i,j+1,k) RU (i,j+1,k+1) LU (i,j,k+1) */
            int t[4][3] = {{i, j, k}, {i, j + 1, k}, {i, j + 1, k + 1}, {i, j, k + 1}};
            memcpy(zz, t, sizeof(zz));
          } else if (fam == 1) { /* LD (i,j,k) RD (i,j,k+1) RU (i+1,j,k+1) LU (i+1,j,k) */
            int t[4][3] = {{i, j, k}, {i, j, k + 1}, {i + 1, j, k + 1}, {i + 1, j, k}};
            memcpy(zz, t, sizeof(zz));
          } else { /* LD (i,j,k) RD (i+1,j,k) RU (i+1,j+1,k) LU (i,j+1,k) */
            int t[4][3] = {{i, j, k}, {i + 1, j, k}, {i + 1, j + 1, k}, {i, j + 1, k}};
            memcpy(zz, t, sizeof(zz));
          }
          for (c = 0; c < 4; ++c) zq[c] = zi(d, zz[c][0], zz[c][1], zz[c][2]);
          if (fam == 2 && (e->mask[zq[0]] == 0 || e->mask[zq[1]] == 0 || e->mask[zq[2]] == 0 || e->mask[zq[3]] == 0))
            continue;
          /* edge centre, the radial direction there, the tangent */
          const int pi_ = fam == 0 ? i - 1 : i, pj = fam == 1 ? j - 1 : j, pk = fam == 2 ? k - 1 : k;
          for (c = 0; c < 3; ++c) ec[c] = 0.5 * (EV(i, j, k, c) + EV(pi_, pj, pk, c));
          double ct = 0.0, st = 1.0;
          if (fam < 2) {
            a = sqrt(EV(i, j, k, 0) * EV(i, j, k, 0) + EV(i, j, k, 1) * EV(i, j, k, 1) + EV(i, j, k, 2) * EV(i, j, k, 2)) /
                sqrt(ec[0] * ec[0] + ec[1] * ec[1] + ec[2] * ec[2]);
            for (c = 0; c < 3; ++c) ec[c] = ec[c] * a;
            a = 1.0 / sqrt(ec[0] * ec[0] + ec[1] * ec[1] + ec[2] * ec[2]);
            for (c = 0; c < 3; ++c) n[c] = ec[c] * a;
            for (c = 0; c < 3; ++c) t3[c] = EV(i, j, k, c) - EV(pi_, pj, pk, c);
            a = t3[0] * n[0] + t3[1] * n[1] + t3[2] * n[2];
            for (c = 0; c < 3; ++c) t3[c] = t3[c] - a * n[c];
            a = 1.0 / sqrt(t3[0] * t3[0] + t3[1] * t3[1] + t3[2] * t3[2]);
            for (c = 0; c < 3; ++c) t3[c] = t3[c] * a;
          } else {
            double n1[3], tau[3];
            a = 1.0 / sqrt(ec[0] * ec[0] + ec[1] * ec[1] + ec[2] * ec[2]);
            for (c = 0; c < 3; ++c) t3[c] = ec[c] * a;
            for (c = 0; c < 3; ++c) n1[c] = EV(i + 1, j, k, c) - EV(i - 1, j, k, c);
            a = n1[0] * t3[0] + n1[1] * t3[1] + n1[2] * t3[2];
            for (c = 0; c < 3; ++c) n1[c] = n1[c] - a * t3[c];
            a = 1.0 / sqrt(n1[0] * n1[0] + n1[1] * n1[1] + n1[2] * n1[2]);
            for (c = 0; c < 3; ++c) n1[c] = n1[c] * a;
            for (c = 0; c < 3; ++c) tau[c] = EV(i, j + 1, k, c) - EV(i, j - 1, k, c);
            a = tau[0] * t3[0] + tau[1] * t3[1] + tau[2] * t3[2];
            for (c = 0; c < 3; ++c) tau[c] = tau[c] - a * t3[c];
            a = 1.0 / sqrt(tau[0] * tau[0] + tau[1] * tau[1] + tau[2] * tau[2]);
            for (c = 0; c < 3; ++c) tau[c] = tau[c] * a;
            ct = tau[0] * n1[0] + tau[1] * n1[1] + tau[2] * n1[2];
            st = sqrt(1.0 - ct * ct);
          }
          const double* Ec[3] = {e->Ex_zc, e->Ey_zc, e->Ez_zc};
          for (c = 0; c < 3; ++c) Eq[c] = (Ec[c][zq[0]] + Ec[c][zq[1]] + Ec[c][zq[2]] + Ec[c][zq[3]]) * 0.25;
          b = Eq[0] * t3[0] + Eq[1] * t3[1] + Eq[2] * t3[2];
          ez4[0] = b; /* RU */
          ez4[1] = b; /* LU */
          ez4[2] = b; /* LD */
          ez4[3] = b; /* RD */
          double sR, sL, sU, sD;
          long fz[4];
          for (c = 0; c < 4; ++c) fz[c] = fci(d, zz[c][0], zz[c][1], zz[c][2]);
          if (fam == 0) {
            by2[0] = e->Bzf_q[fz[1] + NFC * 4];
            by2[1] = e->Bzf_q[fz[0] + NFC * 3];
            bn2[0] = e->Byf_q[fz[3] + NFC * 2];
            bn2[1] = e->Byf_q[fz[0] + NFC * 1];
            sR = omax(e->S_U[fz[3]], e->S_U[fz[0]]);
            sL = omin(e->S_D[fz[3]], e->S_D[fz[0]]);
            sU = omax(e->S_T[fz[1]], e->S_T[fz[0]]);
            sD = omin(e->S_B[fz[1]], e->S_B[fz[0]]);
          } else if (fam == 1) {
            by2[0] = e->Bxf_q[fz[1] + NFC * 4];
            by2[1] = e->Bxf_q[fz[0] + NFC * 3];
            bn2[0] = e->Bzf_q[fz[3] + NFC * 2];
            bn2[1] = e->Bzf_q[fz[0] + NFC * 1];
            sR = omax(e->S_T[fz[3]], e->S_T[fz[0]]);
            sL = omin(e->S_B[fz[3]], e->S_B[fz[0]]);
            sU = omax(e->S_R[fz[0]], e->S_R[fz[1]]);
            sD = omin(e->S_L[fz[0]], e->S_L[fz[1]]);
          } else {
            by2[0] = e->Byf_q[fz[1] + NFC * 4];
            by2[1] = e->Byf_q[fz[0] + NFC * 3];
            bn2[0] = e->Bxf_q[fz[3] + NFC * 2];
            bn2[1] = e->Bxf_q[fz[0] + NFC * 1];
            sR = omax(e->S_R[fz[3]], e->S_R[fz[0]]);
            sL = omin(e->S_L[fz[3]], e->S_L[fz[0]]);
            sU = omax(e->S_U[fz[1]], e->S_U[fz[0]]);
            sD = omin(e->S_D[fz[1]], e->S_D[fz[0]]);
          }
          a = omax(omax(fabs(sR), fabs(sL)), omax(fabs(sU), fabs(sD)));
          s4[0] = a;
          s4[1] = -a;
          s4[2] = a;
          s4[3] = -a;
          double* out = fam == 0 ? e->Efx : fam == 1 ? e->Efy : e->Efz;
          q = fci(d, i, j, k);
          out[q] = out[q] + 1.0 * or_emf4_hll(s4, ct, st, ez4, bn2, by2);
        }
  }
  /* tri-corners: the edge where the three active zones meet */
  for (int corner = 0; corner < 4; ++corner) {
    if (!e->itri[corner]) continue;
    /* corner vertex (ixx, iyy); zones 1, 2, 3; per face: the field (array,
     * zone, point, sign) and the speeds (right and left array, zone, sign) */
    const int ix = (corner == 1 || corner == 2) ? N : 0, iy = (corner >= 2) ? N : 0;
    int z1[2], z2[2], z3[2];
    if (corner == 0) {
      z1[0] = ix + 1; z1[1] = iy + 1; z2[0] = ix; z2[1] = iy + 1; z3[0] = ix + 1; z3[1] = iy;
    } else if (corner == 1) {
      z1[0] = ix + 1; z1[1] = iy + 1; z2[0] = ix; z2[1] = iy + 1; z3[0] = ix; z3[1] = iy;
    } else if (corner == 2) {
      z1[0] = ix; z1[1] = iy + 1; z2[0] = ix; z2[1] = iy; z3[0] = ix + 1; z3[1] = iy;
    } else {
      z1[0] = ix + 1; z1[1] = iy + 1; z2[0] = ix; z2[1] = iy; z3[0] = ix + 1; z3[1] = iy;
    }
    for (k = 1; k <= NR; ++k) {
      double sr[3], sl[3], ez3[3], bn3[3];
      for (c = 0; c < 3; ++c) ec[c] = 0.5 * (EV(ix, iy, k, c) + EV(ix, iy, k - 1, c));
      a = 1.0 / sqrt(ec[0] * ec[0] + ec[1] * ec[1] + ec[2] * ec[2]);
      for (c = 0; c < 3; ++c) t3[c] = ec[c] * a;
      const long q1 = zi(d, z1[0], z1[1], k), q2 = zi(d, z2[0], z2[1], k), q3 = zi(d, z3[0], z3[1], k);
      const double* Ec[3] = {e->Ex_zc, e->Ey_zc, e->Ez_zc};
      for (c = 0; c < 3; ++c) Eq[c] = (Ec[c][q1] + Ec[c][q2] + Ec[c][q3]) * (1.0 / 3.0);
      b = Eq[0] * t3[0] + Eq[1] * t3[1] + Eq[2] * t3[2];
      ez3[0] = b;
      ez3[1] = b;
      ez3[2] = b;
      if (corner == 0) { /* 460...f LD block */
        bn3[0] = -BQ(Byf_q, z3[0], z3[1], k, 5);
        bn3[1] = BQ(Bxf_q, z2[0], z2[1], k, 3);
        bn3[2] = BQ(Byf_q, z2[0], z2[1] - 1, k, 4);
        sr[0] = -SG(S_D, z3[0], z3[1], k);
        sl[0] = -SG(S_U, z3[0], z3[1], k);
        sr[1] = SG(S_R, z2[0], z2[1], k);
        sl[1] = SG(S_L, z2[0], z2[1], k);
        sr[2] = SG(S_U, z2[0], z2[1] - 1, k);
        sl[2] = SG(S_D, z2[0], z2[1] - 1, k);
      } else if (corner == 1) { /* RD block */
        bn3[0] = -BQ(Byf_q, z1[0], z1[1] - 1, k, 5);
        bn3[1] = BQ(Bxf_q, z2[0], z2[1], k, 3);
        bn3[2] = BQ(Byf_q, z3[0], z3[1], k, 4);
        sr[0] = -SG(S_L, z3[0], z3[1], k);
        sl[0] = -SG(S_R, z3[0], z3[1], k);
        sr[1] = SG(S_R, z2[0], z2[1], k);
        sl[1] = SG(S_L, z2[0], z2[1], k);
        sr[2] = SG(S_U, z3[0], z3[1], k);
        sl[2] = SG(S_D, z3[0], z3[1], k);
      } else if (corner == 2) { /* RU block */
        bn3[0] = BQ(Bxf_q, z1[0], z1[1], k, 3);
        bn3[1] = BQ(Byf_q, z2[0], z2[1], k, 4);
        bn3[2] = -BQ(Bxf_q, z2[0], z2[1], k, 2);
        sr[0] = SG(S_R, z1[0], z1[1], k);
        sl[0] = SG(S_L, z1[0], z1[1], k);
        sr[1] = SG(S_U, z2[0], z2[1], k);
        sl[1] = SG(S_D, z2[0], z2[1], k);
        sr[2] = -SG(S_L, z2[0], z2[1], k);
        sl[2] = -SG(S_R, z2[0], z2[1], k);
      } else { /* LU block */
        bn3[0] = -BQ(Byf_q, z3[0], z3[1], k, 5);
        bn3[1] = BQ(Byf_q, z2[0], z2[1], k, 4);
        bn3[2] = -BQ(Bxf_q, z2[0], z2[1], k, 2);
        sr[0] = -SG(S_D, z3[0], z3[1], k);
        sl[0] = -SG(S_U, z3[0], z3[1], k);
        sr[1] = SG(S_U, z2[0], z2[1], k);
        sl[1] = SG(S_D, z2[0], z2[1], k);
        sr[2] = -SG(S_L, z2[0], z2[1], k);
        sl[2] = -SG(S_R, z2[0], z2[1], k);
      }
      a = omax(omax(omax(fabs(sr[0]), fabs(sl[0])), omax(fabs(sr[1]), fabs(sl[1]))),
               omax(fabs(sr[2]), fabs(sl[2])));
      for (c = 0; c < 3; ++c) {
        sr[c] = a;
        sl[c] = -a;
      }
      q = fci(d, ix, iy, k);
      e->Efz[q] = e->Efz[q] + 1.0 * or_emf3_hll(sr, sl, ez3, bn3);
    }
  }
#undef EV
#undef BQ
#undef SG
  (void)NZ;
  return 0;
}

/* BUILD_NON_STIFF_SOURCES (400riemann_integrate/470build_non_stiff_sources.f:6)
 * of one patch, MHD, adiabatic, with no external source (the routine's
 * MAKE_EXTERNAL_SOURCE hook is not in the reference tree): Source cleared;
 * on the active zones of (0..N+1)^2 x (0..nr+1), when Mstar or |omega|
 * exceeds smallnum, the point-mass gravity -G Mstar rho r / |r|^3 at the
 * zone centroid plus the centrifugal -rho omega x (omega x r) and Coriolis
 * -2 rho omega x v terms, the energy source v . S, added to Source and
 * Source added to dU_dt (:150-250). */
int or_build_non_stiff_sources_p1(const or_p1_source* s) {
  if (!s || s->nghost < 1 || s->ntp < 1 || s->nr < 1) return 1;
  or_p1_patch pp;
  memset(&pp, 0, sizeof(pp));
  pp.ntp = s->ntp;
  pp.nr = s->nr;
  pp.nghost = s->nghost;
  const p1dims D = p1_dims(&pp), *d = &D;
  const long NZ = d->NZ;
  long q;
  int i, j, k, c;
  for (q = 0; q < 8 * NZ; ++q) s->Source[q] = 0.0;
  const double ox = s->omega[0], oy = s->omega[1], oz = s->omega[2];
  const double aa = sqrt(ox * ox + oy * oy + oz * oz);
  for (k = 0; k <= d->NR + 1; ++k)
    for (j = 0; j <= d->N + 1; ++j)
      for (i = 0; i <= d->N + 1; ++i) {
        const long z = zi(d, i, j, k);
        if (s->mask[z] != 1) continue;
        const double x = s->v_centroid[z], y = s->v_centroid[z + NZ], w = s->v_centroid[z + 2 * NZ];
        const double rho = s->V[z], vx = s->V[z + NZ], vy = s->V[z + 2 * NZ], vz = s->V[z + 3 * NZ];
        double S[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        const double rad = sqrt(x * x + y * y + w * w);
        const double gmr3 = s->newtonsc * s->mstar * (1.0 / (rad * rad * rad));
        if (s->mstar > s->smallnum || aa > s->smallnum) {
          const double cx = oy * w - oz * y, cy = oz * x - ox * w, cz = ox * y - oy * x;
          const double fx = -rho * (oy * cz - oz * cy), fy = -rho * (oz * cx - ox * cz), fz = -rho * (ox * cy - oy * cx);
          const double kx = -2.0 * rho * (oy * vz - oz * vy), ky = -2.0 * rho * (oz * vx - ox * vz),
                       kz = -2.0 * rho * (ox * vy - oy * vx);
          S[1] = -(gmr3 * x * rho) + fx + kx;
          S[2] = -(gmr3 * y * rho) + fy + ky;
          S[3] = -(gmr3 * w * rho) + fz + kz;
          S[4] = vx * S[1] + vy * S[2] + vz * S[3];
        }
        for (c = 0; c < 8; ++c) {
          s->Source[z + NZ * c] = s->Source[z + NZ * c] + S[c];
          s->dU_dt[z + NZ * c] = s->dU_dt[z + NZ * c] + s->Source[z + NZ * c];
        }
      }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* PAD_BOUNDARY_R (800applications_cs/826pad_boundary_r.f:1) of one patch,
 * outflow (bcarr_r = 2) only, no application branch (CAK, helio, B0B1 off).
 * For izz = 1..NGhost and every zone column whose mask at layer izz is 1:
 *   inner (bcarr_r(.,.,1) = 2, :219-268, the "#else" copy): mode 1 of U at
 *     layer 1 - izz copied from layer 1; Bxf at x-faces ixx-1 and ixx, Byf at
 *     y-faces iyy-1 and iyy, copied from layer 1; Bzf at radial face -izz from
 *     face 0;
 *   outer (bcarr_r(.,.,2) = 2, :296-337): mode 1 of U at layer nr + izz from
 *     layer nr scaled by V(nr) / V(nr + izz); Bzf at face nr + izz from face
 *     nr, Bxf and Byf from layer nr, each scaled by the ratio of the face areas.
 * The entropy argument is not written by the routine and is not an argument
 * here.  Every write reads only layers 1 and nr and faces 0 and nr, so the
 * order of the loop does not matter. */
int or_pad_boundary_r_p1(const or_p1_pad* p) {
  const int N = p->ntp, NR = p->nr, g = p->nghost;
  const int64_t nz = N + 2 * g, nf = nz + 1, NZ = nz * nz * (NR + 2 * g);
#define ZI(i, j, k) ((int64_t)((i) - 1 + g) + nz * (((j) - 1 + g) + nz * ((k) - 1 + g)))
#define FC(i, j, k) ((int64_t)((i) + g) + nf * (((j) + g) + nf * ((k) + g)))
#define XA(i, j, k) ((int64_t)((i) + g) + nf * (((j) - 1 + g) + nz * ((k) - 1 + g)))
#define YA(i, j, k) ((int64_t)((i) - 1 + g) + nz * (((j) + g) + nf * ((k) - 1 + g)))
#define ZA(i, j, k) ((int64_t)((i) - 1 + g) + nz * (((j) - 1 + g) + nz * ((k) + g)))
#define BC(i, j, s) ((int64_t)((i) - 1 + g) + nz * (((j) - 1 + g) + nz * (s)))
  for (int izz = 1; izz <= g; ++izz)
    for (int iy = 1 - g; iy <= N + g; ++iy)
      for (int ix = 1 - g; ix <= N + g; ++ix) {
        if (p->mask[ZI(ix, iy, izz)] != 1) continue;
        if (p->bcarr_r[BC(ix, iy, 0)] == 2) {
          const int l = 1 - izz, r = 1;
          for (int c = 0; c < 8; ++c) p->U[ZI(ix, iy, l) + NZ * c] = p->U[ZI(ix, iy, r) + NZ * c];
          p->Bxf[FC(ix, iy, l)] = p->Bxf[FC(ix, iy, r)];
          p->Bxf[FC(ix - 1, iy, l)] = p->Bxf[FC(ix - 1, iy, r)];
          p->Byf[FC(ix, iy, l)] = p->Byf[FC(ix, iy, r)];
          p->Byf[FC(ix, iy - 1, l)] = p->Byf[FC(ix, iy - 1, r)];
          p->Bzf[FC(ix, iy, l - 1)] = p->Bzf[FC(ix, iy, r - 1)];
        }
        if (p->bcarr_r[BC(ix, iy, 1)] == 2) {
          const int l = NR + izz, r = NR;
          for (int c = 0; c < 8; ++c)
            p->U[ZI(ix, iy, l) + NZ * c] =
                p->U[ZI(ix, iy, r) + NZ * c] * p->v_volume[ZI(ix, iy, r)] / p->v_volume[ZI(ix, iy, l)];
          p->Bzf[FC(ix, iy, l)] = p->Bzf[FC(ix, iy, r)] * p->zf_area[ZA(ix, iy, r)] / p->zf_area[ZA(ix, iy, l)];
          p->Bxf[FC(ix, iy, l)] = p->Bxf[FC(ix, iy, r)] * p->xf_area[XA(ix, iy, r)] / p->xf_area[XA(ix, iy, l)];
          p->Byf[FC(ix, iy, l)] = p->Byf[FC(ix, iy, r)] * p->yf_area[YA(ix, iy, r)] / p->yf_area[YA(ix, iy, l)];
          p->Bxf[FC(ix - 1, iy, l)] =
              p->Bxf[FC(ix - 1, iy, r)] * p->xf_area[XA(ix - 1, iy, r)] / p->xf_area[XA(ix - 1, iy, l)];
          p->Byf[FC(ix, iy - 1, l)] =
              p->Byf[FC(ix, iy - 1, r)] * p->yf_area[YA(ix, iy - 1, r)] / p->yf_area[YA(ix, iy - 1, l)];
        }
      }
#undef ZI
#undef FC
#undef XA
#undef YA
#undef ZA
#undef BC
  return 0;
}

/* ------------------------------------------------------------------------ */
/* The face fluxes of the P1 stage of a shell (the product's CS_RECON_P1ADP):
 * MAKE_FACIAL_FLUXES's faces, zones (with the tri-corner re-pointing of
 * 450...f:581-622), frames and face-centre evaluation as above, but each face
 * keeps one resolved flux: riemann 0 the HLL flux of OneD_RS_MHD_Ptwise
 * (flux form: (F_L + F_R)/2 + S_L/2 (U* - U_L) - S_R/2 (U_R - U*), with its
 * entropy fix of U*), 1 HLLD (Miyoshi & Kusano 2005); rotated back to the
 * global frame and times the face area, with the entropy flux as a ninth
 * component.  wall: the radial faces k = 0 and k = nr take the inner zone's
 * face state mirrored (normal velocity and normal B reversed).  F[fam]:
 * (9, face-B cube), S[fam]: (2, face-B cube) = S_L, S_R (x), S_D, S_U (y),
 * S_B, S_T (z); entries outside the routine's range are left as they are. */
int or_p1s_face_fluxes(const or_p1_flux* f, int riemann, int wall, double* const* F, double* const* S) {
  if (!f || f->nghost < 4 || f->ntp < 1 || f->nr < 1) return 1;
  or_p1_patch pp;
  memset(&pp, 0, sizeof pp);
  pp.ntp = f->ntp;
  pp.nr = f->nr;
  pp.nghost = f->nghost;
  pp.mask = f->mask;
  pp.stencil = f->stencil;
  pp.scale = (double*)f->scale;
  pp.v_centroid = f->v_centroid;
  pp.V = (double*)f->V;
  const p1dims D = p1_dims(&pp), *d = &D;
  const int N = d->N, NR = d->NR;
  int i, j, k, c;
  for (int fam = 0; fam < 3; ++fam) {
    const int ilo = fam == 0 ? -1 : 0, jlo = fam == 1 ? -1 : 0, klo = fam == 2 ? -1 : 0;
    const double* pts = fam == 0 ? f->xf_pts : (fam == 1 ? f->yf_pts : f->zf_pts);
    const double* nrm = fam == 0 ? f->xf_normal : (fam == 1 ? f->yf_normal : f->zf_normal);
    const double* area = fam == 0 ? f->xf_area : (fam == 1 ? f->yf_area : f->zf_area);
    const double* fcen = fam == 0 ? f->xf_centroid : (fam == 1 ? f->yf_centroid : f->zf_centroid);
    const long NF = fam == 2 ? d->NZF : d->NXF;
    for (k = klo; k <= NR + 1; ++k)
      for (j = jlo; j <= N + 1; ++j)
        for (i = ilo; i <= N + 1; ++i) {
          long fb = fci(d, i, j, k), fg = fam == 0 ? xfi(d, i, j, k) : (fam == 1 ? yfi(d, i, j, k) : zfi(d, i, j, k));
          int li = i, lj = j, ri = i + (fam == 0), rj = j + (fam == 1), rk = k + (fam == 2);
          if (fam < 2) {
            int al = active(&pp, d, li, lj, k), ar = active(&pp, d, ri, rj, k);
            if (al && ar) {
            } else if (!al && ar) {
              nbr(&pp, d, ri, rj, fam == 0 ? 1 : 3, &li, &lj);
            } else if (al && !ar) {
              nbr(&pp, d, li, lj, fam == 0 ? 0 : 2, &ri, &rj);
            } else {
              continue;
            }
          } else if (!active(&pp, d, i, j, k)) {
            continue;
          }
          long zl = zi(d, li, lj, k), zr = zi(d, ri, rj, fam == 2 ? rk : k);
          double A[3], B[3], C[3];
          face_frame(f, d, fam, i, j, k, fg, NF, nrm, fcen, A, B, C);
          double rl[3], rr[3], VL[8], VR[8], eL, eR;
          rel(&pp, d, zl, pts + fg, NF, obl_of(&pp, zl), rl);
          rel(&pp, d, zr, pts + fg, NF, obl_of(&pp, zr), rr);
          for (c = 0; c < 8; ++c) {
            VL[c] = VV(&pp, d, c, 0, zl) + VV(&pp, d, c, 1, zl) * rl[0] + VV(&pp, d, c, 2, zl) * rl[1] +
                    VV(&pp, d, c, 3, zl) * rl[2];
            VR[c] = VV(&pp, d, c, 0, zr) + VV(&pp, d, c, 1, zr) * rr[0] + VV(&pp, d, c, 2, zr) * rr[1] +
                    VV(&pp, d, c, 3, zr) * rr[2];
          }
          eL = f->ent[zl] + f->ent[zl + d->NZ] * rl[0] + f->ent[zl + 2 * d->NZ] * rl[1] + f->ent[zl + 3 * d->NZ] * rl[2];
          eR = f->ent[zr] + f->ent[zr + d->NZ] * rr[0] + f->ent[zr + 2 * d->NZ] * rr[1] + f->ent[zr + 3 * d->NZ] * rr[2];
          to_frame(A, B, C, VL);
          to_frame(A, B, C, VR);
          if (wall && fam == 2 && (k == 0 || k == NR)) {
            double* src = k == 0 ? VR : VL;
            double* dst = k == 0 ? VL : VR;
            for (c = 0; c < 8; ++c) dst[c] = src[c];
            dst[1] = -src[1];
            dst[5] = -src[5];
            if (k == 0) eL = eR;
            else eR = eL;
          }
          double Fs[8], eF, SL, SR, Fg[8];
          if (riemann == 1) or_hlld(f->gamma, f->rhofloor, f->prsfloor, VL, VR, eL, eR, Fs, &eF, &SL, &SR);
          else or_hll(f->gamma, f->rhofloor, f->prsfloor, VL, VR, eL, eR, Fs, &eF, &SL, &SR);
          from_frame(A, B, C, Fs, Fg);
          for (c = 0; c < 8; ++c) F[fam][fb + d->NFC * c] = Fg[c] * area[fg];
          F[fam][fb + d->NFC * 8] = eF * area[fg];
          S[fam][fb] = SL;
          S[fam][fb + d->NFC] = SR;
        }
  }
  return 0;
}

/* The zone update of the P1 stage of a shell for the owned zones of one
 * sector: L = -(1/V) sum over the six faces of the flux-times-area
 * differences (9 components, the ninth the entropy), the stage combination
 * with U^n (uz, ez) and the radial gravity g = -grav / r^2 along the zone
 * centroid's direction: mode 0 explicit, Shu-Osher u = wa uz + wb (ui + dt
 * (L + S(ui))); modes 1 and 2 the IMEX-SSP2(2,2,2) stages of or_imex_combine
 * (the induction rows and the entropy as Heun); then CONS_TO_PRIM
 * (484...f:620-628).  uo, eo receive the owned zones; the others are left. */
int or_p1s_update_zones(int N, int NR, int g, const double* const* F, const double* vol, const double* cen,
                        const double* ui_, const double* ei_, const double* uz_, const double* ez_, double dt,
                        double wa, double wb, int mode, double grav, double gam, double rhof, double prsf,
                        double* uo, double* eo) {
  const int nz = N + 2 * g, nzr = NR + 2 * g, nf = nz + 1;
  const long NZ = (long)nz * nz * nzr, NFC = (long)nf * nf * (nzr + 1);
  for (int k = 1; k <= NR; ++k)
    for (int j = 1; j <= N; ++j)
      for (int i = 1; i <= N; ++i) {
        const long z = (i - 1 + g) + (long)nz * ((j - 1 + g) + (long)nz * (k - 1 + g));
        const long fp = (i + g) + (long)nf * ((j + g) + (long)nf * (k + g));
        const long fxm = fp - 1, fym = fp - nf, fzm = fp - (long)nf * nf;
        double L[9], ui[8], uz[8], u[8], n[3] = {0.0, 0.0, 0.0}, gr = 0.0, en;
        int q;
        for (q = 0; q < 9; ++q) {
          const long o = (long)q * NFC;
          L[q] = -(1.0 / vol[z]) * ((F[0][o + fp] - F[0][o + fxm]) + (F[1][o + fp] - F[1][o + fym]) +
                                    (F[2][o + fp] - F[2][o + fzm]));
        }
        if (grav != 0.0) {
          double c0 = cen[z], c1 = cen[z + NZ], c2 = cen[z + 2 * NZ], r2 = c0 * c0 + c1 * c1 + c2 * c2;
          double ir = 1.0 / sqrt(r2);
          n[0] = c0 * ir;
          n[1] = c1 * ir;
          n[2] = c2 * ir;
          gr = -grav / r2;
        }
        for (q = 0; q < 8; ++q) {
          ui[q] = ui_[z + NZ * q];
          uz[q] = uz_[z + NZ * q];
        }
        const double ei = ei_[z], ez = ez_[z];
        if (mode == 0) {
          double Lg[5] = {L[0], L[1], L[2], L[3], L[4]};
          if (gr != 0.0) {
            for (q = 0; q < 3; ++q) Lg[1 + q] += ui[0] * gr * n[q];
            Lg[4] += (ui[1] * n[0] + ui[2] * n[1] + ui[3] * n[2]) * gr;
          }
          for (q = 0; q < 8; ++q) {
            double un = ui[q] + dt * (q < 5 ? Lg[q] : L[q]);
            u[q] = wa != 0.0 ? wa * uz[q] + wb * un : wb * un;
          }
          double e1 = ei + dt * L[8];
          en = wa != 0.0 ? wa * ez + wb * e1 : wb * e1;
        } else {
          or_imex_combine(mode, dt, gr, n, ui, uz, L, u);
          if (mode == 1) {
            for (q = 5; q < 8; ++q) u[q] = uz[q] + dt * L[q];
            en = ez + dt * L[8];
          } else {
            for (q = 5; q < 8; ++q) u[q] = 0.5 * uz[q] + 0.5 * (ui[q] + dt * L[q]);
            en = 0.5 * ez + 0.5 * (ei + dt * L[8]);
          }
        }
        double V[8];
        V[4] = 0.0;
        or_cons_to_prim(gam, rhof, prsf, &en, u, V);
        for (q = 0; q < 8; ++q) uo[z + NZ * q] = u[q];
        eo[z] = en;
      }
  return 0;
}

/* ---------------------------------------------------------------------------
 * PROBLEM_GENERATOR_U_A (801problem_generator_u_a.f:10) and
 * INIT_PATCH_CUBE_SPHERE (820init_patch_cube_sphere.f:8) / ADD_BOUNDARY
 * (825add_boundary.f:12) on one P1 patch. */

/* splitmix64 and the standard normal of (seed, id) by Box-Muller: the
 * density noise of problem 103, one draw per zone */
static uint64_t pg_mix(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
static double pg_unit(uint64_t seed, uint64_t id, uint64_t stream) {
  const uint64_t h = pg_mix(pg_mix(seed ^ (stream * 0xD1B54A32D192ED03ull)) + id);
  return ((double)(h >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}
static double pg_normal(uint64_t seed, uint64_t id) {
  const double u1 = pg_unit(seed, id, 21), u2 = pg_unit(seed, id, 22);
  return sqrt(-2.0 * log(u1)) * cos(2.0 * 3.14159265358979323846 * u2);
}

void or_problem_generator_u_a(const or_prob* pb, double x, double y, double z, double* U, double* A) {
  const double gam = pb->gamma;
  int c;
  for (c = 0; c < 8; ++c) U[c] = 0.0;
  A[0] = A[1] = A[2] = 0.0;
  if (pb->iproblem == 0) {
    /* the coordinate fields of the messaging test (:158-196): U directly */
    A[0] = y * z;
    A[1] = z * x * 2.0;
    A[2] = x * y * 3.0;
    U[0] = fabs(x);
    U[1] = x;
    U[2] = y;
    U[3] = z;
    U[4] = fabs(z);
    U[5] = 1.0 * x;
    U[6] = -2.0 * y;
    U[7] = 1.0 * z;
  } else if (pb->iproblem == 2) {
    /* the magnetised blast (:275-364): p 1000 within radius 1 of
     * ((rmax + rmin) / 2, 0, 0), 0.1 outside, B = 100 / sqrt(3) (1, 1, 1) */
    const double b = 100.0 / sqrt(3.0);
    const double dx = x - (pb->rmax + pb->rmin) * 0.5;
    const double dist = sqrt(dx * dx + y * y + z * z);
    double V[8] = {1.0, 0.0, 0.0, 0.0, 0.1, b, b, b};
    if (dist <= 1.0) V[4] = 1000.0;
    A[0] = b * z;
    A[1] = b * x;
    A[2] = b * y;
    const double ent = V[4] / pow(V[0], gam - 1.0);
    or_prim_to_cons(gam, pb->rhofloor, pb->prsfloor, ent, V, U);
  } else if (pb->iproblem == 102) {
    /* the surface blast: p 10 within 0.1 rad of (1, 0, 0), B = 0.5 east,
     * A = -0.5 r theta rhat (theta the colatitude) */
    const double r = sqrt(x * x + y * y + z * z);
    const double d = atan2(sqrt(y * y + z * z), x);
    const double rho_xy = sqrt(x * x + y * y);
    double V[8] = {1.0, 0.0, 0.0, 0.0, d < 0.1 ? 10.0 : 1.0, 0.0, 0.0, 0.0};
    if (rho_xy > 0.0) {
      V[5] = -0.5 * y / rho_xy;
      V[6] = 0.5 * x / rho_xy;
    }
    const double th = atan2(rho_xy, z), ar = -0.5 * th;
    A[0] = ar * x;
    A[1] = ar * y;
    A[2] = ar * z;
    (void)r;
    const double ent = V[4] / pow(V[0], gam - 1.0);
    or_prim_to_cons(gam, pb->rhofloor, pb->prsfloor, ent, V, U);
  } else if (pb->iproblem == 103) {
    /* isothermal hydrostatic: rho = p = exp(g0 (1/r - 1/rmin)), at rest;
     * A = s (yz, zx + xy / 2, xy), s = 0.05 / r^2, whose curl is
     * s (0, 0, y / 2) + grad s x a with grad s = -0.1 x / r^4 */
    const double r2 = x * x + y * y + z * z, r = sqrt(r2);
    const double g0 = pb->grav > 0.0 ? pb->grav : 1.0;
    const double rho = exp(g0 * (1.0 / r - 1.0 / pb->rmin));
    const double s = 0.05 / r2, gs = -0.1 / (r2 * r2);
    const double a[3] = {y * z, z * x + 0.5 * x * y, x * y};
    A[0] = s * a[0];
    A[1] = s * a[1];
    A[2] = s * a[2];
    double V[8] = {rho, 0.0, 0.0, 0.0, rho, 0.0, 0.0, 0.0};
    V[5] = gs * (y * a[2] - z * a[1]);
    V[6] = gs * (z * a[0] - x * a[2]);
    V[7] = s * 0.5 * y + gs * (x * a[1] - y * a[0]);
    const double ent = V[4] / pow(V[0], gam - 1.0);
    or_prim_to_cons(gam, pb->rhofloor, pb->prsfloor, ent, V, U);
  }
}

/* A . dl of edge a -> b of a face: along a radial edge the chord with A at
 * its midpoint; along a lateral edge the arc at the radius of vertex r0 (the
 * midpoint lifted onto that sphere, the unit chord times r0 acos(a . b / |a||b|)) */
static double pg_edge(const or_prob* pb, const double* a, const double* b, int lateral, double rref) {
  double cen[3], dl[3], U[8], A[3];
  int c;
  for (c = 0; c < 3; ++c) cen[c] = (a[c] + b[c]) * 0.5;
  if (!lateral) {
    for (c = 0; c < 3; ++c) dl[c] = b[c] - a[c];
  } else {
    const double cm = sqrt(cen[0] * cen[0] + cen[1] * cen[1] + cen[2] * cen[2]);
    for (c = 0; c < 3; ++c) cen[c] = rref * cen[c] / cm;
    double u[3];
    for (c = 0; c < 3; ++c) u[c] = b[c] - a[c];
    const double um = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    for (c = 0; c < 3; ++c) u[c] = u[c] / um;
    const double am = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    const double bm = sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
    const double ab = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
    const double ang = acos(ab / (am * bm));
    for (c = 0; c < 3; ++c) dl[c] = rref * ang * u[c];
  }
  or_problem_generator_u_a(pb, cen[0], cen[1], cen[2], U, A);
  return A[0] * dl[0] + A[1] * dl[1] + A[2] * dl[2];
}

int or_init_patch_cube_sphere(const or_p1_init* p, const or_prob* pb, int32_t layers) {
  const int N = p->ntp, NR = p->nr, g = p->nghost;
  if (N < 1 || NR < 1 || g < 1 || layers < 0 || layers > 3) return -1;
  const int64_t nz = N + 2 * g, nf = nz + 1, nzr = NR + 2 * g, nfr = nzr + 1;
  const int64_t NZ = nz * nz * nzr, NFC = nf * nf * nfr;
  const double ob8pi = 1.0 / (8.0 * (4.0 * atan(1.0)));
  const double gam = pb->gamma;
#define ZI(i, j, k) ((int64_t)((i) - 1 + g) + nz * (((j) - 1 + g) + nz * ((k) - 1 + g)))
#define FC(i, j, k) ((int64_t)((i) + g) + nf * (((j) + g) + nf * ((k) + g)))
#define XA(i, j, k) ((int64_t)((i) + g) + nf * (((j) - 1 + g) + nz * ((k) - 1 + g)))
#define YA(i, j, k) ((int64_t)((i) - 1 + g) + nz * (((j) + g) + nf * ((k) - 1 + g)))
#define ZA(i, j, k) ((int64_t)((i) - 1 + g) + nz * (((j) - 1 + g) + nz * ((k) + g)))
#define LAYER(k) (((layers & 1) && (k) >= 1 - g && (k) <= 0) || ((layers & 2) && (k) >= NR + 1 && (k) <= NR + g))
#define ZLAYER(k) (((layers & 1) && (k) >= -g && (k) <= -1) || ((layers & 2) && (k) >= NR + 1 && (k) <= NR + g))
  int i, j, k, c, q;
  if (!layers) {
    memset(p->U, 0, sizeof(double) * NZ * 8);
    memset(p->ent, 0, sizeof(double) * NZ);
    memset(p->Bxf, 0, sizeof(double) * NFC);
    memset(p->Byf, 0, sizeof(double) * NFC);
    memset(p->Bzf, 0, sizeof(double) * NFC);
  } else {
    for (k = 1 - g; k <= NR + g; ++k)
      if (LAYER(k))
        for (j = 1 - g; j <= N + g; ++j)
          for (i = 1 - g; i <= N + g; ++i)
            for (c = 0; c < 8; ++c) p->U[ZI(i, j, k) + NZ * c] = 0.0;
  }
  /* zones: the normalised Gauss weights of the 8 points (:303-460) */
  for (k = 1 - g; k <= NR + g; ++k) {
    if (layers && !LAYER(k)) continue;
    for (j = 1 - g; j <= N + g; ++j)
      for (i = 1 - g; i <= N + g; ++i) {
        const int64_t z0 = ZI(i, j, k);
        if (p->mask[z0] != 1) continue;
        for (q = 0; q < 8; ++q) {
          double U[8], A[3];
          or_problem_generator_u_a(pb, p->v_pts[z0 + NZ * (0 + 3 * q)], p->v_pts[z0 + NZ * (1 + 3 * q)],
                                   p->v_pts[z0 + NZ * (2 + 3 * q)], U, A);
          for (c = 0; c < 8; ++c) p->U[z0 + NZ * c] = p->U[z0 + NZ * c] + U[c] * p->v_wts[z0 + NZ * q];
        }
        if (pb->iproblem == 103) {
          /* the zone's density noise (no reference counterpart) */
          const uint64_t gid = (((uint64_t)p->sector * (uint64_t)nzr + (uint64_t)(k - 1 + g)) * (uint64_t)nz +
                                (uint64_t)(j - 1 + g)) * (uint64_t)nz + (uint64_t)(i - 1 + g);
          p->U[z0] = p->U[z0] * (1.0 + 1.0e-2 * pg_normal(pb->seed, gid));
        }
        /* entropy of the zone mean (:431-455) */
        const double rho = omax(p->U[z0], pb->rhofloor);
        const double v0 = rho / (rho * rho + pb->rhofloor * pb->rhofloor);
        const double vx = p->U[z0 + NZ] * v0, vy = p->U[z0 + 2 * NZ] * v0, vz = p->U[z0 + 3 * NZ] * v0;
        const double bx = p->U[z0 + 5 * NZ], by = p->U[z0 + 6 * NZ], bz = p->U[z0 + 7 * NZ];
        double prs = (gam - 1.0) * (p->U[z0 + 4 * NZ] - 0.5 * rho * (vx * vx + vy * vy + vz * vz) -
                                    (bx * bx + by * by + bz * bz) * ob8pi);
        prs = omax(prs, pb->prsfloor);
        p->ent[z0] = prs / pow(rho, gam - 1.0);
      }
  }
  double v1[3], v2[3], v3[3], v4[3];
#define VX(out, ii, jj, kk) \
  for (c = 0; c < 3; ++c) (out)[c] = p->vert[FC(ii, jj, kk) + NFC * c]
#define MAG(v) sqrt((v)[0] * (v)[0] + (v)[1] * (v)[1] + (v)[2] * (v)[2])
  /* x-faces (:464-693): v1 (i, j-1, k-1), v2 (i, j, k-1), v3 (i, j-1, k),
   * v4 (i, j, k); e1 v2 -> v4 and e2 v3 -> v1 radial, e3 v4 -> v3 at |v3|
   * and e4 v1 -> v2 at |v1| lateral */
  for (k = 1 - g; k <= NR + g; ++k) {
    if (layers && !LAYER(k)) continue;
    for (j = 1 - g; j <= N + g; ++j)
      for (i = -g; i <= N + g; ++i) {
        const int ii = i > 1 - g ? i : 1 - g, ip = i + 1 < N + g ? i + 1 : N + g;
        if (!(p->mask[ZI(ii, j, k)] == 1 || p->mask[ZI(ip, j, k)] == 1)) continue;
        VX(v1, i, j - 1, k - 1);
        VX(v2, i, j, k - 1);
        VX(v3, i, j - 1, k);
        VX(v4, i, j, k);
        double s = 0.0;
        s = s + pg_edge(pb, v2, v4, 0, 0.0);
        s = s + pg_edge(pb, v3, v1, 0, 0.0);
        s = s + pg_edge(pb, v4, v3, 1, MAG(v3));
        s = s + pg_edge(pb, v1, v2, 1, MAG(v1));
        p->Bxf[FC(i, j, k)] = s / p->xf_area[XA(i, j, k)];
      }
  }
  /* y-faces (:695-923): v1 (i-1, j, k-1), v2 (i-1, j, k), v3 (i, j, k-1),
   * v4 (i, j, k); e1 v2 -> v4 at |v2| and e2 v3 -> v1 at |v1| lateral, e3
   * v4 -> v3 and e4 v1 -> v2 radial */
  for (k = 1 - g; k <= NR + g; ++k) {
    if (layers && !LAYER(k)) continue;
    for (j = -g; j <= N + g; ++j)
      for (i = 1 - g; i <= N + g; ++i) {
        const int jj = j > 1 - g ? j : 1 - g, jp = j + 1 < N + g ? j + 1 : N + g;
        if (!(p->mask[ZI(i, jj, k)] == 1 || p->mask[ZI(i, jp, k)] == 1)) continue;
        VX(v1, i - 1, j, k - 1);
        VX(v2, i - 1, j, k);
        VX(v3, i, j, k - 1);
        VX(v4, i, j, k);
        double s = 0.0;
        s = s + pg_edge(pb, v2, v4, 1, MAG(v2));
        s = s + pg_edge(pb, v3, v1, 1, MAG(v1));
        s = s + pg_edge(pb, v4, v3, 0, 0.0);
        s = s + pg_edge(pb, v1, v2, 0, 0.0);
        p->Byf[FC(i, j, k)] = s / p->yf_area[YA(i, j, k)];
      }
  }
  /* z-faces (:925-1177): v1 (i-1, j-1, k), v2 (i, j-1, k), v3 (i-1, j, k),
   * v4 (i, j, k); all four edges lateral: e1 v2 -> v4 at |v2|, e2 v3 -> v1
   * at |v1|, e3 v4 -> v3 at |v3|, e4 v1 -> v2 at |v1| */
  for (k = -g; k <= NR + g; ++k) {
    if (layers && !ZLAYER(k)) continue;
    for (j = 1 - g; j <= N + g; ++j)
      for (i = 1 - g; i <= N + g; ++i) {
        if (p->mask[ZI(i, j, 1)] != 1) continue;
        VX(v1, i - 1, j - 1, k);
        VX(v2, i, j - 1, k);
        VX(v3, i - 1, j, k);
        VX(v4, i, j, k);
        double s = 0.0;
        s = s + pg_edge(pb, v2, v4, 1, MAG(v2));
        s = s + pg_edge(pb, v3, v1, 1, MAG(v1));
        s = s + pg_edge(pb, v4, v3, 1, MAG(v3));
        s = s + pg_edge(pb, v1, v2, 1, MAG(v1));
        p->Bzf[FC(i, j, k)] = s / p->zf_area[ZA(i, j, k)];
      }
  }
#undef VX
#undef MAG
#undef ZI
#undef FC
#undef XA
#undef YA
#undef ZA
#undef LAYER
#undef ZLAYER
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Third order: LIMIT_ALL_SLOPES_P12 (421limit_all_slopes_p12.f:9), steps 2
 * to 7.b, for one patch whose U and V hold ten modes (mean, x, y, z, x^2,
 * y^2, z^2, xy, yz, xz: 204build_mode_indices.f:55-95) and whose basis
 * moments (moments, zone x 10) are those of the 27-point volume rule.
 * Steps 7.c (positivity at third order), 8 (face B and the 30-unknown
 * constrained least squares) and 9 (zone E) are not restated. */

/* RECONSTRUCTION_COEFS (200mesh/251reconstruction_coefs.f:3) at order 3: the
 * row of zone j in the quadratic fit about zone 0, (x, y, z) the scaled
 * centroid offset, l0, lj the two scale lengths, c0, cj the two zones' ten
 * basis moments: the average over zone j of zone 0's basis functions */
void or_reconstruction_row3(double x, double y, double z, double l0, double lj, const double* c0, const double* cj,
                            double* row) {
  const double rho = lj / l0, r2 = rho * rho;
  row[0] = x;
  row[1] = y;
  row[2] = z;
  row[3] = x * x - c0[4] + cj[4] * r2;
  row[4] = y * y - c0[5] + cj[5] * r2;
  row[5] = z * z - c0[6] + cj[6] * r2;
  row[6] = x * y - c0[7] + cj[7] * r2;
  row[7] = y * z - c0[8] + cj[8] * r2;
  row[8] = x * z - c0[9] + cj[9] * r2;
}

/* the 18 zones of the central quadratic stencil (421...f:3888-4350, in that
 * order): stencil directions (0-based R L U D .. RU LU LD RD) and k offsets */
static const int kAoDir[18] = {0, 1, 2, 3, -1, -1, 8, 9, 10, 11, 0, 1, 0, 1, 2, 3, 2, 3};
static const int kAoDk[18] = {0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};

/* the least-squares operator of the quadratic stencil of zone (i, j, k):
 * the 18 x 9 rows A and (A^T A)^-1 (421...f:3870-4370); returns the zones */
static void ao_build(const or_p1_patch* p, const p1dims* d, const double* mom, int i, int j, int k, long* z18,
                     double A[18][9], double inv[81]) {
  const long z = zi(d, i, j, k);
  const double ell0 = omax(p->scale[z], p->smallnum), obl0 = 1.0 / ell0;
  double c0[10], cj[10], ata[81];
  for (int m = 0; m < 10; ++m) c0[m] = mom[z + d->NZ * m];
  for (int r = 0; r < 18; ++r) {
    int a = i, b = j;
    if (kAoDir[r] >= 0) nbr(p, d, i, j, kAoDir[r], &a, &b);
    const long zj = zi(d, a, b, k + kAoDk[r]);
    z18[r] = zj;
    double x[3];
    for (int c = 0; c < 3; ++c) x[c] = (p->v_centroid[zj + d->NZ * c] - p->v_centroid[z + d->NZ * c]) * obl0;
    const double ellj = omax(p->scale[zj], p->smallnum);
    for (int m = 0; m < 10; ++m) cj[m] = mom[zj + d->NZ * m];
    or_reconstruction_row3(x[0], x[1], x[2], ell0, ellj, c0, cj, A[r]);
  }
  for (int a = 0; a < 9; ++a)
    for (int b = 0; b < 9; ++b) {
      double t = 0.0;
      for (int r = 0; r < 18; ++r) t = t + A[r][a] * A[r][b];
      ata[a * 9 + b] = t;
    }
  or_matinv(9, ata, inv);
}

/* Step 7.b (421...f:3811-4557): on the active zones of -1 .. N+2 (k -1 ..
 * NR+2), the 9 modes X of the quadratic least-squares fit to the 18 mean
 * differences, hybridised with the WENO gradients of 7.a (MR-WENO: linear
 * weights 0.95 / 0.05 for smooth flow, 0.6 / 0.4 where the flattener is 1,
 * smoothness indicators of the two polynomials, weights (1 + tau / IS)),
 * then every mode above the mean times (1 - flattener).  The eight
 * primitives only; the entropy keeps its flattened gradients. */
int or_weno_ao_p12(const or_p1_patch* p, const double* mom) {
  if (!p || p->nghost < 4 || p->ntp < 1 || p->nr < 1 || !mom) return 1;
  const p1dims D = p1_dims(p), *d = &D;
  const int N = d->N, NR = d->NR;
  const double ob12 = 1.0 / 12.0, ob180 = 1.0 / 180.0, ob144 = 1.0 / 144.0, sm = p->smallnum;
  for (int k = -1; k <= NR + 2; ++k)
    for (int j = -1; j <= N + 2; ++j)
      for (int i = -1; i <= N + 2; ++i) {
        if (!active(p, d, i, j, 1)) continue;
        const long z = zi(d, i, j, k);
        long z18[18];
        double A[18][9], inv[81];
        ao_build(p, d, mom, i, j, k, z18, A, inv);
        const double f = p->flatten[z];
        const double betalo = f + (1.0 - f) * 2.0;
        const double gammahi = f * 0.6 + (1.0 - f) * 0.95, gammalo = 1.0 - gammahi;
        const double obgammahi = 1.0 / gammahi, gammaratio = gammalo / gammahi;
        const double* c0 = mom;
        for (int c = 0; c < 8; ++c) {
          double R[18], Y[9], X[9], Phi[9];
          const double v0 = VV(p, d, c, 0, z);
          for (int r = 0; r < 18; ++r) R[r] = VV(p, d, c, 0, z18[r]) - v0;
          for (int a = 0; a < 9; ++a) {
            double t = 0.0;
            for (int r = 0; r < 18; ++r) t = t + A[r][a] * R[r];
            Y[a] = t;
          }
          for (int a = 0; a < 9; ++a) {
            double t = 0.0;
            for (int b = 0; b < 9; ++b) t = t + inv[a * 9 + b] * Y[b];
            X[a] = t;
          }
          const double g1 = VV(p, d, c, 1, z), g2 = VV(p, d, c, 2, z), g3 = VV(p, d, c, 3, z);
          const double ISlo = betalo * (g1 * g1 + g2 * g2 + g3 * g3) * ob12;
          Phi[0] = X[0] * obgammahi - g1 * gammaratio;
          Phi[1] = X[1] * obgammahi - g2 * gammaratio;
          Phi[2] = X[2] * obgammahi - g3 * gammaratio;
          for (int a = 3; a < 9; ++a) Phi[a] = X[a] * obgammahi;
          const double m5 = c0[z + d->NZ * 4] - ob12, m6 = c0[z + d->NZ * 5] - ob12, m7 = c0[z + d->NZ * 6] - ob12;
          const double m8 = c0[z + d->NZ * 7], m9 = c0[z + d->NZ * 8], m10 = c0[z + d->NZ * 9];
          const double IShi = (Phi[0] * Phi[0] + Phi[1] * Phi[1] + Phi[2] * Phi[2]) * ob12 +
                              (Phi[3] * Phi[3]) * (m5 * m5 + ob180) + (Phi[4] * Phi[4]) * (m6 * m6 + ob180) +
                              (Phi[5] * Phi[5]) * (m7 * m7 + ob180) + (Phi[6] * Phi[6]) * (m8 * m8 + ob144) +
                              (Phi[7] * Phi[7]) * (m9 * m9 + ob144) + (Phi[8] * Phi[8]) * (m10 * m10 + ob144);
          const double tau = fabs(IShi - ISlo);
          double wlo = gammalo * (1.0 + tau / (ISlo + sm)), whi = gammahi * (1.0 + tau / (IShi + sm));
          const double aa = 1.0 / (wlo + whi);
          wlo = wlo * aa;
          whi = whi * aa;
          double out[9];
          out[0] = whi * Phi[0] + wlo * g1;
          out[1] = whi * Phi[1] + wlo * g2;
          out[2] = whi * Phi[2] + wlo * g3;
          for (int a = 3; a < 9; ++a) out[a] = whi * Phi[a];
          const double x = 1.0 - f;
          for (int a = 0; a < 9; ++a) VV(p, d, c, a + 1, z) = out[a] * x;
        }
      }
  return 0;
}

/* the quadratic least-squares modes X (9) of component c of zone (i, j, k)
 * from the means in V (mode 1), before the hybridisation: the fit of step
 * 7.b alone, for tests of its exactness */
int or_ao_fit_p12(const or_p1_patch* p, const double* mom, int i, int j, int k, int c, double* X) {
  if (!p || !mom || c < 0 || c > 7) return 1;
  const p1dims D = p1_dims(p), *d = &D;
  long z18[18];
  double A[18][9], inv[81], R[18], Y[9];
  const long z = zi(d, i, j, k);
  ao_build(p, d, mom, i, j, k, z18, A, inv);
  for (int r = 0; r < 18; ++r) R[r] = VV(p, d, c, 0, z18[r]) - VV(p, d, c, 0, z);
  for (int a = 0; a < 9; ++a) {
    double t = 0.0;
    for (int r = 0; r < 18; ++r) t = t + A[r][a] * R[r];
    Y[a] = t;
  }
  for (int a = 0; a < 9; ++a) {
    double t = 0.0;
    for (int b = 0; b < 9; ++b) t = t + inv[a * 9 + b] * Y[b];
    X[a] = t;
  }
  return 0;
}

/* CMat_Row_Coefs_P2 (200mesh/253cmat_row_coefs_p2.f:7): n (x) the ten
 * basis functions at the scaled point (x, y, z) about a zone with moments c0 */
static void cmat_row_p2(double x, double y, double z, const double* n, const double* c0, double* row) {
  const double b[10] = {1.0, x, y, z, x * x - c0[4], y * y - c0[5], z * z - c0[6], x * y - c0[7], y * z - c0[8],
                        x * z - c0[9]};
  for (int a = 0; a < 3; ++a) {
    row[10 * a] = n[a];
    for (int m = 1; m < 10; ++m) row[10 * a + m] = n[a] * b[m];
  }
}

/* STEP 8.d of LIMIT_ALL_SLOPES_P12 (421...f:5574-6680): the almost
 * divergence-preserving constrained least squares for the 30 quadratic
 * coefficients of B.  Constraints (6): each face's weighted mean of n . B
 * over its five B-field points equals that of its face-point B.  Least
 * squares (64 rows): n . B at every face point against the face-point B (30),
 * the 30 WENO-AO modes, and the divergence and its gradient set to zero (4).
 * The 36 x 36 system of normal equations and constraints is solved by
 * MATSOLVE; B_zc, V's B modes and U's B means take the result and the energy
 * the change of magnetic energy. */
static void p12_clsq(const or_p1_patch* p, const p1dims* d, const double* mom) {
  const int N = d->N, NR = d->NR;
  const double pi = 4.0 * atan(1.0), ob8pi = 1.0 / (8.0 * pi);
  for (int k = -2; k <= NR + 3; ++k)
    for (int j = -2; j <= N + 3; ++j)
      for (int i = -2; i <= N + 3; ++i) {
        const long z = zi(d, i, j, k);
        if (p->mask[z] == 0) continue;
        const double obl = obl_of(p, z);
        double c0[10], Cm[6][30], dv[6], Am[64][30], bv[64];
        for (int m = 0; m < 10; ++m) c0[m] = mom[z + d->NZ * m];
        memset(Cm, 0, sizeof Cm);
        memset(dv, 0, sizeof dv);
        memset(Am, 0, sizeof Am);
        memset(bv, 0, sizeof bv);
        /* faces in the order right x, left x, up y, down y, top z, bottom z */
        for (int f = 0; f < 6; ++f) {
          const int fam = f / 2, hi = f % 2 == 0;
          const int fi_ = i - (fam == 0 && !hi), fj_ = j - (fam == 1 && !hi), fk_ = k - (fam == 2 && !hi);
          const long fg = fam == 0 ? xfi(d, fi_, fj_, fk_) : (fam == 1 ? yfi(d, fi_, fj_, fk_) : zfi(d, fi_, fj_, fk_));
          const long NF = fam == 0 ? d->NXF : (fam == 1 ? d->NYF : d->NZF);
          const long fb = fci(d, fi_, fj_, fk_);
          const double* nrm = fam == 0 ? p->xb_normal : (fam == 1 ? p->yb_normal : p->zb_normal);
          const double* pts = fam == 0 ? p->xb_pts : (fam == 1 ? p->yb_pts : p->zb_pts);
          const double* wts = fam == 0 ? p->xb_wts : (fam == 1 ? p->yb_wts : p->zb_wts);
          const double* Bq = fam == 0 ? p->Bxf_q : (fam == 1 ? p->Byf_q : p->Bzf_q);
          for (int q = 0; q < 5; ++q) {
            double n[3], r[3], row[30];
            for (int c = 0; c < 3; ++c) n[c] = fam == 2 ? nrm[fg + NF * (c + 3 * q)] : nrm[fg + NF * c];
            const double wt = wts[fg + NF * q], bq = Bq[fb + d->NFC * q];
            rel(p, d, z, pts + fg + NF * (3 * q), NF, obl, r);
            cmat_row_p2(r[0], r[1], r[2], n, c0, row);
            for (int u = 0; u < 30; ++u) {
              Cm[f][u] = Cm[f][u] + wt * row[u];
              Am[5 * f + q][u] = row[u];
            }
            dv[f] = dv[f] + wt * bq;
            bv[5 * f + q] = bq;
          }
        }
        for (int q = 0; q < 30; ++q) {  /* the WENO-AO modes */
          Am[30 + q][q] = 1.0;
          bv[30 + q] = VV(p, d, 5 + q / 10, q % 10, z);
        }
        /* div B = 0 and its gradient: unknowns c * 10 + mode (1 x, 2 y, 3 z,
           4 xx, 5 yy, 6 zz, 7 xy, 8 yz, 9 xz) */
        Am[60][1] = 1.0; Am[60][12] = 1.0; Am[60][23] = 1.0;
        Am[61][4] = 2.0; Am[61][17] = 1.0; Am[61][29] = 1.0;
        Am[62][7] = 1.0; Am[62][15] = 2.0; Am[62][28] = 1.0;
        Am[63][9] = 1.0; Am[63][18] = 1.0; Am[63][26] = 2.0;
        double K[LUMAX][LUMAX], rhs[LUMAX];
        for (int a = 0; a < 30; ++a) {
          for (int b = 0; b < 30; ++b) {
            double t = 0.0;
            for (int q = 0; q < 64; ++q) t = t + Am[q][a] * Am[q][b];
            K[a][b] = t;
          }
          double t = 0.0;
          for (int q = 0; q < 64; ++q) t = t + Am[q][a] * bv[q];
          rhs[a] = t;
          for (int f = 0; f < 6; ++f) {
            K[a][30 + f] = Cm[f][a];
            K[30 + f][a] = Cm[f][a];
          }
        }
        for (int f = 0; f < 6; ++f) {
          for (int h = 0; h < 6; ++h) K[30 + f][30 + h] = 0.0;
          rhs[30 + f] = dv[f];
        }
        lu_solve(36, K, rhs);
        const double aa = (UU(p, d, 5, 0, z) * UU(p, d, 5, 0, z) + UU(p, d, 6, 0, z) * UU(p, d, 6, 0, z) +
                           UU(p, d, 7, 0, z) * UU(p, d, 7, 0, z)) * ob8pi;
        for (int m = 0; m < 10; ++m) {
          p->B1_zc[z + d->NZ * m] = rhs[m];
          p->B2_zc[z + d->NZ * m] = rhs[10 + m];
          p->B3_zc[z + d->NZ * m] = rhs[20 + m];
          for (int c = 0; c < 3; ++c) VV(p, d, 5 + c, m, z) = rhs[10 * c + m];
        }
        for (int c = 0; c < 3; ++c) UU(p, d, 5 + c, 0, z) = VV(p, d, 5 + c, 0, z);
        const double bb = (UU(p, d, 5, 0, z) * UU(p, d, 5, 0, z) + UU(p, d, 6, 0, z) * UU(p, d, 6, 0, z) +
                           UU(p, d, 7, 0, z) * UU(p, d, 7, 0, z)) * ob8pi;
        UU(p, d, 4, 0, z) = UU(p, d, 4, 0, z) - aa + bb;
      }
}

/* INTERPOLATION_COEFS (200mesh/252interpolation_coefs.f:5) at order 3: the
 * ten basis functions of a zone with moments c0 at the scaled point
 * (x, y, z) */
void or_interpolation_row3(double x, double y, double z, const double* c0, double* row) {
  row[0] = 1.0;
  row[1] = x;
  row[2] = y;
  row[3] = z;
  row[4] = x * x - c0[4];
  row[5] = y * y - c0[5];
  row[6] = z * z - c0[6];
  row[7] = x * y - c0[7];
  row[8] = y * z - c0[8];
  row[9] = x * z - c0[9];
}

/* STEP 9 of LIMIT_ALL_SLOPES_P12 (421...f:6681-9353): the zone electric
 * field.  9.a: E = -v x B at the centroid, v and B the quadratic fields there
 * (the mean minus the quadratic modes times the moments), every zone; 9.b:
 * the P1 WENO gradients of E, unflattened (:6796-8405); 9.c (:8406-9353): on
 * the active zones of -1 .. N+2 the ten modes of the least-squares quadratic
 * interpolant (INTERPOLATION_COEFS) through the zone's and its 18
 * neighbours' point values, hybridised with the WENO gradients as in 7.b,
 * then E = E_pt f + E (1 - f) for the point value and the modes times
 * (1 - f). */
static void p12_efield(const or_p1_patch* p, const p1dims* d, const double* mom) {
  const int N = d->N, NR = d->NR;
  const long S = d->NZ;
  double* E[3] = {p->Ex_zc, p->Ey_zc, p->Ez_zc};
  for (long z = 0; z < S; ++z) {
    double q[7][1];
    int cs_[6] = {5, 6, 7, 1, 2, 3};
    double val[6];
    for (int t = 0; t < 6; ++t) {
      const int c = cs_[t];
      double v = (t < 3 ? 0.0 + VV(p, d, c, 0, z) : VV(p, d, c, 0, z));
      for (int m = 4; m < 10; ++m) v = v - VV(p, d, c, m, z) * mom[z + S * m];
      val[t] = v;
    }
    (void)q;
    const double Bx = val[0], By = val[1], Bz = val[2], vx = val[3], vy = val[4], vz = val[5];
    p->Ex_zc[z] = -(vy * Bz - vz * By);
    p->Ey_zc[z] = -(vz * Bx - vx * Bz);
    p->Ez_zc[z] = -(vx * By - vy * Bx);
  }
  for (int k = -2; k <= NR + 3; ++k)
    for (int j = -2; j <= N + 3; ++j)
      for (int i = -2; i <= N + 3; ++i) {
        if (!active(p, d, i, j, 1)) continue;
        long z = zi(d, i, j, k), z6[6];
        weno_mats W;
        six_nbrs(p, d, i, j, k, z6);
        weno_build(p, d, z, z6, &W);
        for (int c = 0; c < 3; ++c) {
          double f6[6], g3[3];
          for (int q = 0; q < 6; ++q) f6[q] = E[c][z6[q]];
          weno_apply(&W, E[c][z], f6, p->smallnum, 0.0, g3);
          for (int m = 0; m < 3; ++m) E[c][z + S * (m + 1)] = g3[m];
        }
      }
  const double ob12 = 1.0 / 12.0, ob180 = 1.0 / 180.0, ob144 = 1.0 / 144.0, sm = p->smallnum;
  for (int k = -1; k <= NR + 2; ++k)
    for (int j = -1; j <= N + 2; ++j)
      for (int i = -1; i <= N + 2; ++i) {
        if (!active(p, d, i, j, 1)) continue;
        const long z = zi(d, i, j, k);
        const double ell0 = omax(p->scale[z], sm), obl0 = 1.0 / ell0;
        double c0[10], A[19][10], ata[100], inv[100];
        long z19[19];
        for (int m = 0; m < 10; ++m) c0[m] = mom[z + S * m];
        z19[0] = z;
        or_interpolation_row3(0.0, 0.0, 0.0, c0, A[0]);
        for (int r = 0; r < 18; ++r) {
          int a = i, b = j;
          if (kAoDir[r] >= 0) nbr(p, d, i, j, kAoDir[r], &a, &b);
          const long zj = zi(d, a, b, k + kAoDk[r]);
          z19[r + 1] = zj;
          double x[3];
          for (int c = 0; c < 3; ++c) x[c] = (p->v_centroid[zj + S * c] - p->v_centroid[z + S * c]) * obl0;
          or_interpolation_row3(x[0], x[1], x[2], c0, A[r + 1]);
        }
        for (int a = 0; a < 10; ++a)
          for (int b = 0; b < 10; ++b) {
            double t = 0.0;
            for (int r = 0; r < 19; ++r) t = t + A[r][a] * A[r][b];
            ata[a * 10 + b] = t;
          }
        or_matinv(10, ata, inv);
        const double f = p->flatten[z];
        const double betalo = f + (1.0 - f) * 2.0;
        const double gammahi = f * 0.6 + (1.0 - f) * 0.95, gammalo = 1.0 - gammahi;
        const double obgammahi = 1.0 / gammahi, gammaratio = gammalo / gammahi;
        for (int c = 0; c < 3; ++c) {
          double R[19], Y[10], X[10], Phi[10], e[4];
          const double Ept = E[c][z];
          for (int r = 0; r < 19; ++r) R[r] = E[c][z19[r]];
          for (int a = 0; a < 10; ++a) {
            double t = 0.0;
            for (int r = 0; r < 19; ++r) t = t + A[r][a] * R[r];
            Y[a] = t;
          }
          for (int a = 0; a < 10; ++a) {
            double t = 0.0;
            for (int b = 0; b < 10; ++b) t = t + inv[a * 10 + b] * Y[b];
            X[a] = t;
          }
          for (int m = 0; m < 4; ++m) e[m] = E[c][z + S * m];
          const double ISlo = betalo * (e[1] * e[1] + e[2] * e[2] + e[3] * e[3]) * ob12;
          for (int m = 0; m < 4; ++m) Phi[m] = X[m] * obgammahi - e[m] * gammaratio;
          for (int m = 4; m < 10; ++m) Phi[m] = X[m] * obgammahi;
          const double m5 = c0[4] - ob12, m6 = c0[5] - ob12, m7 = c0[6] - ob12;
          const double IShi = (Phi[1] * Phi[1] + Phi[2] * Phi[2] + Phi[3] * Phi[3]) * ob12 +
                              (Phi[4] * Phi[4]) * (m5 * m5 + ob180) + (Phi[5] * Phi[5]) * (m6 * m6 + ob180) +
                              (Phi[6] * Phi[6]) * (m7 * m7 + ob180) + (Phi[7] * Phi[7]) * (c0[7] * c0[7] + ob144) +
                              (Phi[8] * Phi[8]) * (c0[8] * c0[8] + ob144) + (Phi[9] * Phi[9]) * (c0[9] * c0[9] + ob144);
          const double tau = fabs(IShi - ISlo);
          double wlo = gammalo * (1.0 + tau / (ISlo + sm)), whi = gammahi * (1.0 + tau / (IShi + sm));
          const double aa = 1.0 / (wlo + whi);
          wlo = wlo * aa;
          whi = whi * aa;
          double out[10];
          for (int m = 0; m < 4; ++m) out[m] = whi * Phi[m] + wlo * e[m];
          for (int m = 4; m < 10; ++m) out[m] = whi * Phi[m];
          const double x1 = 1.0 - f;
          E[c][z] = Ept * f + out[0] * x1;
          for (int m = 1; m < 10; ++m) E[c][z + S * m] = out[m] * x1;
        }
      }
}

/* LIMIT_ALL_SLOPES_P12, steps 2 to 10: 2 to 7.a with ten modes and
 * unflattened WENO gradients, 7.b, the positivity limiter of 7.c and the
 * face-point B of 8.a-c with the quadratic modes, the constrained least
 * squares of 8.d, the zone E of 9, and the diagnostics of 10 */
int or_limit_all_slopes_p12(const or_p1_patch* p, const double* mom) {
  if (!p || p->nghost < 4 || p->ntp < 1 || p->nr < 1 || !mom) return 1;
  const p1dims D = p1_dims(p);
  p1_front(p, &D, 10, 0);
  if (or_weno_ao_p12(p, mom)) return 1;
  g_p1_stats[0] = g_p1_stats[1] = 0;
  p1_positivity(p, &D, 10, mom);
  p1_face_points(p, &D, 10, mom);
  p12_clsq(p, &D, mom);
  p12_efield(p, &D, mom);
  p1_step10(p, &D);
  return 0;
}
