// cs_problems.h -- pointwise initial and boundary states of the P1 handles:
// PROBLEM_GENERATOR_U_A (800applications_cs/801problem_generator_u_a.f:10)
// gives the conserved state U(8) and the vector potential A(3) at a point;
// INIT_PATCH_CUBE_SPHERE (820init_patch_cube_sphere.f:8) averages U over a
// zone's Gauss points and takes each face's B as the circulation of A over
// its four edges divided by its area, ADD_BOUNDARY (825add_boundary.f:12)
// does the same on the radial ghost layers only.
//
// Problems: 0 and 2 are the reference's (the coordinate fields of its
// messaging test, 801...f:158-196, and its magnetised blast, :275-364); 102
// and 103 belong to this build's benchmark configurations (the surface blast
// with an east field and the isothermal hydrostatic shell with a field from
// A = 0.05 / r^2 (yz, zx + xy / 2, xy) and seeded 1e-2 density noise).
#pragma once
#include <math.h>
#include <stdint.h>

#include "cs_internal.h"

namespace cspg {

struct Problem {
  int id;
  double time, rmin, rmax, gamma, rhofloor, prsfloor, grav;
  uint64_t seed;
};

CS_HD inline bool known(int id) { return id == 0 || id == 2 || id == 102 || id == 103; }

// PRIM_TO_CONS of the MHD branch for a state with p above the floor
CS_HD inline void to_cons(const Problem& P, const double* V, double* U) {
  const double rho = V[0] > P.rhofloor ? V[0] : P.rhofloor;
  const double ob8pi = 1.0 / (8.0 * 3.14159265358979323846);
  U[0] = rho;
  U[1] = rho * V[1];
  U[2] = rho * V[2];
  U[3] = rho * V[3];
  U[4] = V[4] / (P.gamma - 1.0) + 0.5 * rho * (V[1] * V[1] + V[2] * V[2] + V[3] * V[3]) +
         ob8pi * (V[5] * V[5] + V[6] * V[6] + V[7] * V[7]);
  U[5] = V[5];
  U[6] = V[6];
  U[7] = V[7];
}

CS_HD inline void point(const Problem& P, double x, double y, double z, double* U, double* A) {
  for (int c = 0; c < 8; ++c) U[c] = 0.0;
  A[0] = A[1] = A[2] = 0.0;
  if (P.id == 0) {
    A[0] = y * z;
    A[1] = z * x * 2.0;
    A[2] = x * y * 3.0;
    U[0] = fabs(x);
    U[1] = x;
    U[2] = y;
    U[3] = z;
    U[4] = fabs(z);
    U[5] = x;
    U[6] = -2.0 * y;
    U[7] = z;
  } else if (P.id == 2) {
    const double b = 100.0 / sqrt(3.0), dx = x - (P.rmax + P.rmin) * 0.5;
    const double V[8] = {1.0, 0.0, 0.0, 0.0, sqrt(dx * dx + y * y + z * z) <= 1.0 ? 1000.0 : 0.1, b, b, b};
    A[0] = b * z;
    A[1] = b * x;
    A[2] = b * y;
    to_cons(P, V, U);
  } else if (P.id == 102) {
    const double rxy = sqrt(x * x + y * y), d = atan2(sqrt(y * y + z * z), x);
    double V[8] = {1.0, 0.0, 0.0, 0.0, d < 0.1 ? 10.0 : 1.0, 0.0, 0.0, 0.0};
    if (rxy > 0.0) {
      V[5] = -0.5 * y / rxy;
      V[6] = 0.5 * x / rxy;
    }
    const double ar = -0.5 * atan2(rxy, z);
    A[0] = ar * x;
    A[1] = ar * y;
    A[2] = ar * z;
    to_cons(P, V, U);
  } else if (P.id == 103) {
    const double r2 = x * x + y * y + z * z, r = sqrt(r2);
    const double g0 = P.grav > 0.0 ? P.grav : 1.0, rho = exp(g0 * (1.0 / r - 1.0 / P.rmin));
    const double s = 0.05 / r2, gs = -0.1 / (r2 * r2);
    const double a0 = y * z, a1 = z * x + 0.5 * x * y, a2 = x * y;
    A[0] = s * a0;
    A[1] = s * a1;
    A[2] = s * a2;
    const double V[8] = {rho, 0.0, 0.0, 0.0, rho, gs * (y * a2 - z * a1), gs * (z * a0 - x * a2),
                         s * 0.5 * y + gs * (x * a1 - y * a0)};
    to_cons(P, V, U);
  }
}

// a zone's standard normal draw (splitmix64 hashing of seed, zone and
// stream; Box-Muller) for the density noise of problem 103
CS_HD inline uint64_t mix(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
CS_HD inline double unit(uint64_t seed, uint64_t id, uint64_t stream) {
  const uint64_t h = mix(mix(seed ^ (stream * 0xD1B54A32D192ED03ull)) + id);
  return ((double)(h >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}
CS_HD inline double normal(uint64_t seed, uint64_t id) {
  return sqrt(-2.0 * log(unit(seed, id, 21))) * cos(2.0 * 3.14159265358979323846 * unit(seed, id, 22));
}

// A . dl along edge a -> b: a radial edge as its chord with A at the chord's
// midpoint, a lateral edge as the arc at radius rref (the midpoint lifted to
// that sphere, the unit chord times rref acos(a . b / |a| |b|))
CS_HD inline double edge(const Problem& P, const double* a, const double* b, bool lateral, double rref) {
  double cen[3], dl[3], U[8], A[3];
  for (int c = 0; c < 3; ++c) cen[c] = (a[c] + b[c]) * 0.5;
  if (!lateral) {
    for (int c = 0; c < 3; ++c) dl[c] = b[c] - a[c];
  } else {
    const double cm = sqrt(cen[0] * cen[0] + cen[1] * cen[1] + cen[2] * cen[2]);
    double u[3];
    for (int c = 0; c < 3; ++c) {
      cen[c] = rref * cen[c] / cm;
      u[c] = b[c] - a[c];
    }
    const double um = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    const double am = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    const double bm = sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
    const double ang = acos((a[0] * b[0] + a[1] * b[1] + a[2] * b[2]) / (am * bm));
    for (int c = 0; c < 3; ++c) dl[c] = rref * ang * (u[c] / um);
  }
  point(P, cen[0], cen[1], cen[2], U, A);
  return A[0] * dl[0] + A[1] * dl[1] + A[2] * dl[2];
}

}  // namespace cspg
