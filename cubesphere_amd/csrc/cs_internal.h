// cs_internal.h -- layout helpers and the metric tables shared by the host
// mesh builder (cs_mesh.cpp) and the HIP kernels (cs_kernels.hip).
//
// Index conventions follow the reference's patch arrays
// (pkg/inc/cubesphere/400riemann_integrate/460make_edge_fields_p1.f:40-190):
//   cell  (i,j,k), i,j in [1-g, N+g], k in [1-g, nr+g]
//   x-face (i,j,k): x-vertex i, cell-y j, cell-z k  (between cells i and i+1)
//   y-face (i,j,k): cell-x i, y-vertex j, cell-z k
//   z-face (i,j,k): cell-x i, cell-y j, z-vertex k
//   x-edge (i,j,k): cell-x i, y-vertex j, z-vertex k   (Efx)
//   y-edge (i,j,k): x-vertex i, cell-y j, z-vertex k   (Efy)
//   z-edge (i,j,k): x-vertex i, y-vertex j, cell-z k   (Efz)
// Face and edge arrays share one shape (0-g:N+g)^2 x (0-g:nr+g).
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define CS_HD __host__ __device__ __forceinline__
#else
#define CS_HD inline
#endif

struct CsDims {
  int N, NR, g;
  int nc, nrc;   // cell extents N+2g, nr+2g
  int nf, nrf;   // face extents N+2g+1, nr+2g+1
  int64_t cpp;   // cells per panel  nc*nc*nrc
  int64_t fpp;   // faces per panel  nf*nf*nrf
  int64_t t2;    // 2-D table entries per panel nf*nf

  CS_HD int64_t cell(int i, int j, int k) const {
    return (int64_t)(i - 1 + g) + (int64_t)nc * ((j - 1 + g) + (int64_t)nc * (k - 1 + g));
  }
  CS_HD int64_t face(int i, int j, int k) const {
    return (int64_t)(i + g) + (int64_t)nf * ((j + g) + (int64_t)nf * (k + g));
  }
  // 2-D panel table index, valid for vertex and cell indices in [-g, N+g]
  CS_HD int64_t t(int i, int j) const { return (int64_t)(i + g) + (int64_t)nf * (j + g); }
};

static inline CsDims cs_make_dims(int N, int NR, int g) {
  CsDims d;
  d.N = N; d.NR = NR; d.g = g;
  d.nc = N + 2 * g; d.nrc = NR + 2 * g;
  d.nf = N + 2 * g + 1; d.nrf = NR + 2 * g + 1;
  d.cpp = (int64_t)d.nc * d.nc * d.nrc;
  d.fpp = (int64_t)d.nf * d.nf * d.nrf;
  d.t2 = (int64_t)d.nf * d.nf;
  return d;
}

// Number of doubles per 2-D table entry and their offsets (SoA: one plane of
// t2*6 doubles per field).  The whole 2-D metric is separable from the radial
// direction because every vertex is rb(k) * u(i,j)
// (210extract_vert_coords_patch.f:300-308).
enum CsT2 {
  T_UX = 0, T_UY, T_UZ,          // vertex unit vector u(i,j)
  T_XNX, T_XNY, T_XNZ,           // x-face unit normal (plane through origin)
  T_XCX, T_XCY, T_XCZ,           // x-face radial frame direction
  T_THY,                         // angle between u(i,j-1), u(i,j): x-face / y-edge
  T_YNX, T_YNY, T_YNZ,           // y-face unit normal
  T_YCX, T_YCY, T_YCZ,           // y-face radial frame direction
  T_THX,                         // angle between u(i-1,j), u(i,j): y-face / x-edge
  T_ZSX, T_ZSY, T_ZSZ,           // unit-sphere area vector of cell (i,j): int rhat dOmega
  T_ZSN,                         // |zs|
  T_ZTX, T_ZTY, T_ZTZ,           // z-face tangent frame vector
  T_OM,                          // solid angle of cell (i,j)
  T_ZCX, T_ZCY, T_ZCZ,           // unit direction of the zone centre (vertex sum)
  T_CHMIN,                       // min lateral chord |u_a - u_b| of cell (i,j)
  T_M00, T_M01, T_M02, T_M10, T_M11, T_M12, T_M20, T_M21, T_M22, // surface-mode zone-B operator
  // zone-centred B from the six face values, least squares over the face
  // normals (485cleanup_after_update.f:692-860): B = sum_f W_f b_f with
  // W_f = (A^T A)^-1 n_f for the x+, x-, y+, y- faces and W_z for both z faces
  T_BWXP, T_BWXPY, T_BWXPZ, T_BWXM, T_BWXMY, T_BWXMZ,
  T_BWYP, T_BWYPY, T_BWYPZ, T_BWYM, T_BWYMY, T_BWYMZ,
  T_BWZ, T_BWZY, T_BWZZ,
  T_EN1X, T_EN1Y, T_EN1Z,        // z-edge frame n1 (along mesh x)
  T_ETAX, T_ETAY, T_ETAZ,        // z-edge frame tau (along mesh y)
  T_ECOS, T_ESIN,                // z-edge cos/sin of the angle between n1 and tau
  T_XETX, T_XETY, T_XETZ,        // x-edge direction (unit, normal to radial)
  T_XERX, T_XERY, T_XERZ,        // x-edge radial direction at its midpoint
  T_YETX, T_YETY, T_YETZ,        // y-edge direction
  T_YERX, T_YERY, T_YERZ,        // y-edge radial direction at its midpoint
  T_COUNT
};
