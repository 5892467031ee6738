// cs_probe.hip -- a small program for hardware-counter runs of one kernel of
// the cell-update path.  It builds a workload's initial state, advances it two
// steps, launches the chosen kernel `iters` times through cs_time_kernel, and
// then runs one plain copy kernel of known size (read n bytes, write n bytes)
// whose FETCH_SIZE / WRITE_SIZE readings calibrate the byte counters.
//
//   cs_probe <kernel id> [iters=5] [workload=C2] [precision=0|1] [copy MiB=1024]
//   cs_probe step [steps=1] [workload=C5] [0] [copy MiB=1024]
//
// Workloads: C1 6x32^2 surface HLL, C2 6x256^2 surface HLLD (magnetised
// blast), C3 6x192^2x64 shell with gravity (all WENO5, the stage path's own
// numerics); C3P1 and C5 the 6x192^2x64 and 6x512^2x128 shells with the
// reference's P1 scheme (CS_RECON_P1ADP, NGhost 4) + HLLD + IMEX gravity.
// "step" runs whole steps launch by launch (every kernel of the P1 stage).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/cubesphere_hip.h"

__global__ void k_calibration_copy(const double2* __restrict__ a, double2* __restrict__ b, size_t n) {
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) b[i] = a[i];
}

static void fail(const char* what, int rc) {
  char msg[512] = {0};
  cs_last_error(msg, sizeof(msg));
  fprintf(stderr, "%s failed (%d): %s\n", what, rc, msg);
  exit(1);
}

int main(int argc, char** argv) {
  if (argc < 2) {
    fprintf(stderr, "usage: %s <kernel id> [iters] [C1|C2|C3] [precision 0|1] [copy MiB]\n", argv[0]);
    return 2;
  }
  const bool steps = !strcmp(argv[1], "step");
  int kid = steps ? 0 : atoi(argv[1]);
  int iters = argc > 2 ? atoi(argv[2]) : (steps ? 1 : 5);
  const char* wl = argc > 3 ? argv[3] : (steps ? "C5" : "C2");
  int prec = argc > 4 ? atoi(argv[4]) : 0;
  size_t mib = argc > 5 ? (size_t)atol(argv[5]) : 1024;
  if (kid < 0 || kid >= CS_K_COUNT || iters < 1 || (prec != 0 && prec != 1) || mib < 1 || mib > 8192) {
    fprintf(stderr, "bad arguments\n");
    return 2;
  }
  cs_config c;
  memset(&c, 0, sizeof(c));
  c.nghost = 3; c.recon = CS_RECON_WENO5; c.riemann = CS_RS_HLLD; c.emf = CS_EMF_LLF;
  c.precision = prec; c.rk_order = 3; c.radial_mesh = CS_RMESH_UNIFORM;
  c.gamma = 5.0 / 3.0; c.rhofloor = 1e-12; c.prsfloor = 1e-9; c.cfl = 0.4; c.rmin = 1.0;
  int problem;
  if (!strcmp(wl, "C1") || !strcmp(wl, "C2")) {
    c.ntp = !strcmp(wl, "C1") ? 32 : 256;
    c.nr = 1;
    c.radial_bc = CS_RBC_WALL;
    c.rmax = c.rmin + 0.5 * M_PI / c.ntp;
    if (!strcmp(wl, "C1")) c.riemann = CS_RS_HLL;
    problem = !strcmp(wl, "C1") ? 1 : 2;
  } else if (!strcmp(wl, "C3") || !strcmp(wl, "C3P1") || !strcmp(wl, "C5")) {
    c.ntp = !strcmp(wl, "C5") ? 512 : 192; c.nr = !strcmp(wl, "C5") ? 128 : 64;
    c.radial_bc = CS_RBC_PAD; c.rmax = 1.5; c.grav = 1.0;
    c.rk_order = 2; c.imex = 1;  // IMEX-SSP2(2,2,2), gravity implicit
    if (strcmp(wl, "C3")) {
      c.recon = CS_RECON_P1ADP;
      c.nghost = 4;
    }
    problem = 3;
  } else {
    fprintf(stderr, "unknown workload %s\n", wl);
    return 2;
  }
  cs_sizes sz;
  int rc;
  if ((rc = cs_sizes_of(&c, &sz))) fail("cs_sizes_of", rc);
  std::vector<double> U(sz.n_cell * 8), ent(sz.n_cell), bx(sz.n_face), by(sz.n_face), bz(sz.n_face);
  if ((rc = cs_init_problem(&c, problem, 20261202ull, U.data(), ent.data(), bx.data(), by.data(), bz.data())))
    fail("cs_init_problem", rc);
  cs_handle* h = nullptr;
  if ((rc = cs_create(&c, &h))) fail("cs_create", rc);
  if ((rc = cs_upload_state(h, U.data(), ent.data(), bx.data(), by.data(), bz.data()))) fail("upload", rc);
  if ((rc = cs_set_dt(h, 0.0))) fail("cs_set_dt", rc);
  if ((rc = cs_evaluate_timestep(h))) fail("cs_evaluate_timestep", rc);
  float ms = 0.f;
  int64_t units = 0;
  if (steps) {
    if ((rc = cs_time_steps(h, iters, 0, &ms))) fail("cs_time_steps", rc);
    ms /= iters;
    units = 6ll * c.ntp * c.ntp * c.nr;
  } else {
    if ((rc = cs_advance(h, 2, 0))) fail("cs_advance", rc);
    if ((rc = cs_time_kernel(h, kid, iters, &ms))) fail("cs_time_kernel", rc);
    if ((rc = cs_kernel_units(h, kid, &units))) fail("cs_kernel_units", rc);
  }
  cs_destroy(h);

  size_t bytes = mib << 20, n = bytes / sizeof(double2);
  double2 *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) {
    fprintf(stderr, "hipMalloc failed\n");
    return 1;
  }
  if (hipMemset(a, 0, bytes) != hipSuccess || hipDeviceSynchronize() != hipSuccess) return 1;
  k_calibration_copy<<<4096, 256>>>(a, b, n);
  if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
    fprintf(stderr, "copy kernel failed\n");
    return 1;
  }
  hipFree(a);
  hipFree(b);
  if (steps)
    printf("probe: %d step(s) of %s launch by launch, %.2f ms per step, %lld owned zones; "
           "calibration copy %zu bytes read + %zu written\n",
           iters, wl, ms, (long long)units, bytes, bytes);
  else
    printf("probe kernel %d workload %s precision %d: %d launches, %.2f us avg, %lld units; "
           "calibration copy %zu bytes read + %zu written\n",
           kid, wl, prec, iters, ms * 1e3, (long long)units, bytes, bytes);
  return 0;
}
