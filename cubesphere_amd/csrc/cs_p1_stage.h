// cs_p1_stage.h -- the reference's P1 scheme (LIMIT_ALL_SLOPES_P1, the facial
// fluxes, MAKE_EDGE_EFIELDS_P1, UPDATE_IMEX_RK_STEP_P1, CLEANUP_AFTER_UPDATE,
// EVALUATE_TIMESTEP) as the stage of a cs_handle whose recon is
// CS_RECON_P1ADP: all six sectors per launch, no host synchronisation, so a
// whole step can be captured in one HIP graph.  Host interface used by the
// handle (cs_kernels.hip); the kernels are in cs_p1_stage.hip.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/cubesphere_hip.h"

namespace p1s {

struct Stage;  // device buffers and geometry of one handle's P1 stage

// a stage buffer of the handle: U [s][8][zone], ent [s][zone], face B [s][face]
struct Buf {
  double *U, *ent, *Bx, *By, *Bz;
};

// Allocates the P1 work buffers and builds the separable geometry, the zone
// geometry, the active mask and the stencil table on the device.
int create(const cs_config* cfg, Stage** out, size_t* bytes);
void destroy(Stage* S);

// Steps 1 to 11 of the stage on buffer `in` (the exchange has run):
// primitives, tri-corner face B, flattener, WENO, positivity (with its
// bounded device-side replay), face-point B, constrained least squares, face
// fluxes (times area, F[fam]: [s][9][face]; speeds S[fam]: [s][2][face]),
// edge fields (E[fam]: [s][face]).  err: device int set non-zero when the
// positivity replay has not settled.
int front(Stage* S, const Buf& in, double* const* F, double* const* Sp, double* const* E, int* err,
          hipStream_t st);
// The zone and face update of the stage into `out` (weights wa, wb on U^n in
// u0; mode 0 explicit, 1 and 2 the IMEX stages), with CLEANUP_AFTER_UPDATE.
int update(Stage* S, const Buf& in, const Buf& u0, const Buf& out, double* const* F, double* const* E,
           const double* dtp, double wa, double wb, int mode, hipStream_t st);
// IMEX first stage: U1 = Un + gi dt S(U1) on every zone, face B copied.
int imex_source(Stage* S, const Buf& un, const Buf& u1, const double* dtp, hipStream_t st);
// EVALUATE_TIMESTEP of buffer `cur`: one partial minimum per workgroup into
// part (nblocks() entries); the caller folds them into dt.
int cfl_partial(Stage* S, const Buf& cur, double* part, hipStream_t st);
int nblocks_cfl();
// INIT_PATCH_CUBE_SPHERE (layers 0: the whole buffer, cleared first) or
// ADD_BOUNDARY (layers 1 inner, 2 outer, 3 both radial ghost layers) of a
// problem of PROBLEM_GENERATOR_U_A into buffer `out` (cs_problems.h).
int init_problem(Stage* S, const Buf& out, int problem, uint64_t seed, double time, int layers, hipStream_t st);
// PAD_BOUNDARY_R with outflow at both radial ends on buffer b (after the
// exchange; U and face B of the radial ghost layers)
int pad_radial(Stage* S, const Buf& b, hipStream_t st);
// Diagnostics on the P1 geometry, one partial per workgroup (diag_blocks()
// of them): the largest relative discrete divergence |sum A B| / sum |A B|
// of the owned active zones, and the volume integrals of U and the entropy
// as two-sum pairs (part[(block * 9 + q) * 2 + {0, 1}]).
int diag_blocks(Stage* S);
int divb_partial(Stage* S, const Buf& cur, double* part, hipStream_t st);
int sums_partial(Stage* S, const Buf& cur, double* part, hipStream_t st);

// Per-launch timing while the handle profiles steps: when set, every launch
// (or launch group) of the stage is bracketed by an event pair on its stream
// and tagged with its CS_P1K_* id (include/cubesphere_hip.h).
struct Marks {
  hipEvent_t* ev;  // 2 * cap events
  int* kid;        // cap ids
  int n, cap;
};
void set_marks(Stage* S, Marks* m);
void mark_begin(Stage* S, hipStream_t st);
void mark_end(Stage* S, hipStream_t st, int kid);

}  // namespace p1s
