/*
 * cubesphere_hip.h -- C-ABI of the MI355X (gfx950) cell-update path of the
 * CubeSphere cubed-sphere MHD solver.
 *
 * Array layout: every state array is the reference's per-patch Fortran
 * (column-major) layout, one patch per cube sector, patches stacked last:
 *
 *   U   (1-g:N+g, 1-g:N+g, 1-g:nr+g, 8, 6)   conserved rho, rho*v(3), E, B(3)
 *   ent (1-g:N+g, 1-g:N+g, 1-g:nr+g, 6)      entropy density p / rho^(gamma-1)
 *   Bxf, Byf, Bzf (0-g:N+g, 0-g:N+g, 0-g:nr+g, 6)   facial normal B
 *
 * i.e. exactly U_amr / entropygr_amr / Bxf_amr ... of
 * pkg/inc/cubesphere/300riemann_module/300riemann_module_3d.f:502-552 with
 * Max_patches_per_proc = 6 (all sectors on one GPU).  Because the device keeps
 * the same panel-major SoA layout (x fastest), upload and download are plain
 * copies: no transposition happens at the boundary.
 *
 * Units are the reference's Gaussian units (B^2 / 8 pi).
 *
 * Contract: every entry point returns 0 on success or a negative CS_E_* code;
 * nothing throws across the ABI; cs_last_error() returns the message of the
 * last failure on the calling thread.  The handle owns its device memory;
 * host arrays are borrowed for the duration of the call only.  Calls on one
 * handle are not re-entrant; different handles are independent.
 */
#ifndef CUBESPHERE_HIP_H
#define CUBESPHERE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* error codes (counterparts of the reference's STOP / MPI_ABORT paths) */
#define CS_OK 0
#define CS_E_INVALID (-1)     /* bad argument or configuration               */
#define CS_E_BAD_DECOMP (-2)  /* NGhost >= ntp etc. (203fill_ghost2d_global.f:27-31) */
#define CS_E_HIP (-3)         /* HIP runtime failure                          */
#define CS_E_NONPOSITIVE (-4) /* non-positive density/pressure after floors   */
#define CS_E_NAN (-5)         /* NaN/Inf found in the state                   */
#define CS_E_NOMEM (-6)       /* device allocation failed                     */
#define CS_E_IO (-7)          /* a checkpoint file could not be opened, read or written */
#define CS_E_FORMAT (-8)      /* a checkpoint file is not a dump of this format or is damaged */

/* reconstruction */
#define CS_RECON_PCM 0   /* piecewise constant (first order)               */
#define CS_RECON_WENO5 1 /* Jiang-Shu WENO5 of the primitive variables     */
#define CS_RECON_P1ADP 2 /* the reference's P1 scheme: multi-D WENO of the primitives and
                            the entropy with the flattener and the positivity limiter, the
                            divergence-preserving constrained least squares for the zone B
                            (LIMIT_ALL_SLOPES_P1, 410limit_all_slopes_p1.f:9), face fluxes at
                            the face centroids (MAKE_FACIAL_FLUXES 450...f:5), edge fields
                            (MAKE_EDGE_EFIELDS_P1 460...f:5), arc-length CT and cleanup
                            (484...f:7, 485...f:4), CFL of EVALUATE_TIMESTEP (490...f:7);
                            on the quadratic isoparametric geometry of 220/230/240; FP64,
                            CS_EMF_LLF, nghost >= 4 */
/* face Riemann solver */
#define CS_RS_HLL 0      /* OneD_RS_MHD_Ptwise restated (610_1d_rs_oned_rs_mhd_ptwise.f:13) */
#define CS_RS_HLLD 1     /* Miyoshi & Kusano (2005) HLLD                    */
/* edge EMF solver */
#define CS_EMF_LLF 0     /* 4-state HLL with symmetrised speeds, averaged E (USE_2D_LLF_RS_FOR_EZ 1) */
#define CS_EMF_HLL 1     /* 4-state HLL with the face speeds (630_2d_rs_ez_two_d_rs_ez_4state_hll.f:10) */
/* precision */
#define CS_FP64 0        /* everything in FP64 (the drop-in)               */
#define CS_MIXED 1       /* FP32 reconstruction + Riemann solves; FP64 metric, CT, update, sums */
/* radial boundary */
#define CS_RBC_PAD 0     /* zero-gradient pad (826pad_boundary_r.f:201 form); CS_RECON_P1ADP shells:
                            PAD_BOUNDARY_R with outflow at both ends, the inner layers a copy of
                            layer 1, the outer ones layer nr rescaled by volume and face area */
#define CS_RBC_WALL 1    /* reflecting, perfectly conducting wall (surface mode) */
#define CS_RBC_PROBLEM 2 /* CS_RECON_P1ADP shells: the radial ghost layers from the problem
                            generator after each exchange, ADD_BOUNDARY (825add_boundary.f:12)
                            with problem bc_problem (cs_init_problem_p1) */
/* radial mesh (202generate_global_radial_mesh.f) */
#define CS_RMESH_UNIFORM 0
#define CS_RMESH_LOG 1

typedef struct cs_config {
  int32_t ntp;         /* zones per panel edge (ntp_zones)                 */
  int32_t nr;          /* radial zones (nr_zones)                           */
  int32_t nghost;      /* ghost layers (NGhost), >= 3                       */
  int32_t recon;       /* CS_RECON_*                                        */
  int32_t riemann;     /* CS_RS_*                                           */
  int32_t emf;         /* CS_EMF_*                                          */
  int32_t precision;   /* CS_FP64 / CS_MIXED                                */
  int32_t rk_order;    /* 2 = Heun RK2 (reference DRV:2661-2773), 3 = SSP-RK3 */
  int32_t radial_bc;   /* CS_RBC_*                                          */
  int32_t radial_mesh; /* CS_RMESH_*                                        */
  double rmin, rmax;   /* shell radii                                       */
  double gamma;        /* polytropic index                                  */
  double rhofloor;     /* density floor (DRV:181: 1e-12)                    */
  double prsfloor;     /* pressure floor (DRV:182: 1e-9)                    */
  double cfl;          /* Courant number (cfl_coef)                         */
  double grav;         /* radial gravity g = grav / r^2 (0 = off)           */
  int32_t imex;        /* 1: gravity handled implicitly in the stage solve  */
  int32_t bc_problem;  /* the problem of CS_RBC_PROBLEM (else unused)        */
  double omega[3];     /* angular velocity of the rotating frame (omega_passive,
                          470build_non_stiff_sources.f): centrifugal and
                          Coriolis sources, explicit; 0 = off               */
} cs_config;

/* element counts of the host arrays above (per array, all 6 panels) */
typedef struct cs_sizes {
  int64_t n_cell;  /* (N+2g)^2 (nr+2g) * 6      : ent, and per variable of U */
  int64_t n_face;  /* (N+2g+1)^2 (nr+2g+1) * 6  : each of Bxf, Byf, Bzf, Ef* */
  int64_t n_vert2d;/* (N+2g+1)^2 * 6            : 2-D vertex lattice        */
  int64_t n_interior; /* 6 N^2 nr                                          */
} cs_sizes;

typedef struct cs_handle cs_handle;

/* ---------------------------------------------------------------- host only */
/* No GPU is touched by these. */

int cs_sizes_of(const cs_config* cfg, cs_sizes* out);

/* Unit-sphere vertex lattice of the six sectors including NGhost ghost rows,
 * layout coordinates_gr_global(-g:N+g, -g:N+g, 3, 6): Mesh_Sector_General with
 * ntp_patches = 1 (201mesh_sector_general.f:79) followed by FILL_GHOST2D_GLOBAL
 * (203fill_ghost2d_global.f:7).  Corner ghost blocks, which the reference
 * leaves unset, are filled by continuing the sector's own swivel lattice. */
int cs_build_unit_lattice(const cs_config* cfg, double* coords);

/* Radial vertex radii rb(-g:nr+g) (GENERATE_GLOBAL_RADIAL_{UNIFORM,LOG}_MESH,
 * 202generate_global_radial_mesh.f:10,71). */
int cs_build_radial_mesh(const cs_config* cfg, double* rb);

/* Metric tables the kernels use (exact spherical geometry, separable into a
 * 2-D panel table times 1-D radial factors).  Each output may be NULL.
 * vol(1-g:N+g,1-g:N+g,1-g:nr+g,6), xarea/yarea/zarea face shaped. */
int cs_build_metric(const cs_config* cfg, double* vol, double* xarea,
                    double* yarea, double* zarea);

/* Initial conditions written into host arrays of the layout above.
 * problem: 0 closed-form coordinate fields (801problem_generator_u_a.f:105-229),
 *          1 Gaussian bump + solid-body rotation + azimuthal B (surface test),
 *          2 magnetised blast (p=10 inside a 0.1 rad cap, B = 0.5 east),
 *          3 hydrostatic shell with seeded density noise, B = curl A.
 * Face B comes from a vector potential by Stokes' theorem (820init_patch_cube_sphere.f:464+),
 * so the discrete div B is zero to roundoff.  Ghost zones are left zero. */
int cs_init_problem(const cs_config* cfg, int32_t problem, uint64_t seed,
                    double* U, double* ent, double* Bxf, double* Byf,
                    double* Bzf);

/* ---------------------------------------------------------------- device */
int cs_create(const cs_config* cfg, cs_handle** out);
int cs_destroy(cs_handle* h);

int cs_upload_state(cs_handle* h, const double* U, const double* ent,
                    const double* Bxf, const double* Byf, const double* Bzf);
int cs_download_state(cs_handle* h, double* U, double* ent, double* Bxf,
                      double* Byf, double* Bzf);

/* Ghost exchange between the six sectors: FETCH_BUFFERS_FILL_IN_UX/UY/UY_TOP/UZ
 * (330-333) semantics on one device -- index flips, face-B swap and sign
 * flips, averaging of coinciding faces -- then the radial boundary pad. */
int cs_halo_exchange(cs_handle* h);
/* The same exchange for six P1 patches, one per cube sector, each in the
 * layouts of cs_p1_patch (DEVICE arrays): for sector s, U[s] (mode 1 is
 * exchanged; modes 2-4 untouched), ent[s] (mode 1), Bxf[s], Byf[s], Bzf[s].
 * FETCH_BUFFERS_FILL_IN_UX/UY/UY_TOP/UZ (330-333) + PAD_BOUNDARY_R over the
 * six patches, as the driver does before each stage (DRV:1746-1756); the
 * handle (same ntp, nr, nghost, radial_bc) lends its state buffer and its
 * exchange kernels.  Ordered after the work already on `stream` and before
 * the work put on it afterwards (events); no host wait. */
int cs_halo_exchange_p1(cs_handle* h, double* const* U, double* const* ent, double* const* Bxf,
                        double* const* Byf, double* const* Bzf, void* stream);

/* Reconstruction + face Riemann solves for every x/y/z face
 * (LIMIT_ALL_SLOPES_P1 reconstruction role + MAKE_FACIAL_FLUXES,
 *  450make_facial_fluxes.f:5).  Reads the current stage state. */
int cs_make_facial_fluxes(cs_handle* h);

/* Edge EMFs by the multidimensional Riemann solver
 * (MAKE_EDGE_EFIELDS_P1, 460make_edge_fields_p1.f:5). */
int cs_make_edge_efields(cs_handle* h);

/* Stage update: conserved variables, entropy, constrained-transport face B,
 * then zone B from the faces and the cons->prim floors
 * (UPDATE_IMEX_RK_STEP_P1 484:7 + CLEANUP_AFTER_UPDATE 485:4).
 * irk is the 1-based stage; the stage weights follow cfg->rk_order. */
int cs_update_stage(cs_handle* h, int32_t irk);

/* CFL time step (EVALUATE_TIMESTEP, 490evaluate_timestep.f:7):
 * dt_next = min(1.1 dt, cfl * L / (|v| + c_f)); kept on the device. */
int cs_evaluate_timestep(cs_handle* h);

int cs_set_dt(cs_handle* h, double dt);
int cs_get_dt(cs_handle* h, double* dt);

/* One RK stage: halo -> fluxes -> EMFs -> update (replaces the per-patch
 * body DRV:2007-3120). */
int cs_rk_stage(cs_handle* h, int32_t irk);

/* nsteps full time steps (all stages + CFL), dt stays on the device.
 * use_graph != 0 captures one step in a hipGraph and replays it. */
int cs_advance(cs_handle* h, int32_t nsteps, int32_t use_graph);

/* Intermediate results for parity checks (face shaped, all 6 panels).
 * flux: (9 comps, face) per direction; speeds: (2, face) per direction. */
int cs_download_fluxes(cs_handle* h, int32_t dir, double* flux, double* speeds);
int cs_download_emf(cs_handle* h, double* Efx, double* Efy, double* Efz);

/* Diagnostics on the device: compensated (two-sum) volume integrals of the
 * 8 conserved variables + entropy over the owned zones, and the largest
 * relative discrete divergence |sum_f A_f B_f| / sum_f |A_f B_f| of an owned
 * zone.  With recon CS_RECON_P1ADP both use the P1 geometry (the volumes and
 * face areas the stage conserves on). */
int cs_conservation_sums(cs_handle* h, double* sums9);
int cs_max_divb(cs_handle* h, double* out);

/* Timing with hipEvents recorded on the handle's stream.
 * cs_time_steps: nsteps whole steps (optionally graph replays).
 * cs_time_kernel: average duration of one kernel (CS_K_* id) over iters
 * launches on the current state; CS_K_STEP is a whole step.  Timing the
 * update kernels overwrites the stage scratch buffer only.
 * cs_kernel_units: number of work items (faces, edges, zones) one launch of
 * the kernel processes, for algorithmic-byte accounting. */
#define CS_K_HALO 0
#define CS_K_PRIMS 1
#define CS_K_FLUX_X 2
#define CS_K_FLUX_Y 3
#define CS_K_FLUX_Z 4
#define CS_K_EMF_Z 5
#define CS_K_EMF_X 6
#define CS_K_EMF_Y 7
#define CS_K_UPDATE_CELLS 8
#define CS_K_UPDATE_FACES 9
#define CS_K_CLEANUP 10
#define CS_K_CFL 11
#define CS_K_STEP 12
#define CS_K_SYNC 13   /* panel-edge faces and edges made unique across sectors */
#define CS_K_FLUX_XY 14 /* x- and y-face fluxes in one launch (the stage's form) */
#define CS_K_EMF 15     /* all edge EMFs in one launch (the stage's form)        */
#define CS_K_COUNT 16
int cs_time_steps(cs_handle* h, int32_t nsteps, int32_t use_graph, float* ms_total);
int cs_time_kernel(cs_handle* h, int32_t kernel, int32_t iters, float* ms_avg);
/* cs_profile_steps: nsteps whole steps run launch by launch (no graph) with a
 * hipEvent pair around every launch on the handle's stream; ms_sum[id] and
 * calls[id] (arrays of CS_K_COUNT) accumulate the time and the launches per
 * kernel id, as the kernels run inside the step (caches as the step leaves
 * them).  Advances the state by nsteps. */
int cs_profile_steps(cs_handle* h, int32_t nsteps, float* ms_sum, int32_t* calls);
/* ---- problem generator and radial boundary of a P1 handle -----------------
 * cs_init_problem_p1 replaces INIT_PATCH_CUBE_SPHERE
 * (800applications_cs/820init_patch_cube_sphere.f:8) with
 * PROBLEM_GENERATOR_U_A (801problem_generator_u_a.f:10) on every sector of a
 * handle of recon CS_RECON_P1ADP, on the device: the state (U, entropy, face
 * B) is cleared, every active zone takes the mean of the problem's pointwise
 * U over its 8 Gauss points and the entropy of that mean, every active face
 * the circulation of the problem's vector potential over its four edges
 * (chords radially, arcs along the sphere) over its area, so the discrete
 * divergence of every zone vanishes to rounding.  Problems: 0 and 2 the
 * reference's (coordinate fields :105-229, magnetised blast :249-371), 102
 * and 103 this build's (surface blast with B = 0.5 east; isothermal
 * hydrostatic shell with gravity 1/r^2, A = 0.05 / r^2 (yz, zx + xy/2, xy)
 * and seeded 1e-2 density noise).  cs_add_boundary_p1 replaces ADD_BOUNDARY
 * (825add_boundary.f:12): the same on the inner (ibdy_inner) and / or outer
 * (ibdy_outer) radial ghost layers only, the rest of the state untouched. */
int cs_init_problem_p1(cs_handle* h, int32_t problem, uint64_t seed, double time);
int cs_add_boundary_p1(cs_handle* h, int32_t ibdy_inner, int32_t ibdy_outer, int32_t problem, uint64_t seed,
                       double time);
/* cs_profile_steps_p1: the same for a handle of recon CS_RECON_P1ADP (no
 * counterpart in the reference: measurement only), per launch (or launch
 * group) of its stage, arrays of CS_P1K_COUNT */
#define CS_P1K_EXCHANGE 0       /* ghost exchange + radial ghosts of the stage input  */
#define CS_P1K_PRIM 1           /* CONS_TO_PRIM of every zone, magnetosonic speed      */
#define CS_P1K_FACE_SYNC 2      /* face B cleared between inactive zones, tri-corners  */
#define CS_P1K_FLATTEN 3        /* the three flattener passes                          */
#define CS_P1K_WENO 4           /* multi-D WENO of the 8 primitives and the entropy    */
#define CS_P1K_POSITIVITY 5     /* ENFORCE_PRESSURE_POSITIVITY_PRIM, first pass         */
#define CS_P1K_REPLAY 6         /* its loop-order replay (bounded sweeps)               */
#define CS_P1K_FACE_POINTS 7    /* normal B at the B-field points of every face        */
#define CS_P1K_CLSQ 8           /* ADP constrained least squares, zone B modes         */
#define CS_P1K_FLUX 9           /* face fluxes (HLL / HLLD) x area, speeds              */
#define CS_P1K_EDGE 10          /* edge fields (4-state and tri-corner 3-state HLL)    */
#define CS_P1K_SYNC 11          /* panel-edge faces and edges made one across sectors  */
#define CS_P1K_UPDATE_FACES 12  /* constrained transport of the owned faces            */
#define CS_P1K_UPDATE_ZONES 13  /* zone update, RK weights, IMEX source, cleanup       */
#define CS_P1K_IMEX 14          /* IMEX first (source-only) stage                      */
#define CS_P1K_CFL 15           /* EVALUATE_TIMESTEP partial minima                    */
#define CS_P1K_COUNT 16
int cs_profile_steps_p1(cs_handle* h, int32_t nsteps, float* ms_sum, int32_t* calls);
int cs_kernel_units(cs_handle* h, int32_t kernel, int64_t* units);
/* cs_autotune: time every launch of a stage at 64/128/256-thread workgroups
 * (iters launches each, through cs_time_kernel) and keep the fastest; later
 * steps and graphs use them.  block_sizes[CS_K_COUNT] (may be NULL) receives
 * the workgroup size of every kernel id.  Only the stage scratch buffer and
 * the primitives are modified. */
int cs_autotune(cs_handle* h, int32_t iters, int32_t* block_sizes);

/* ---------------------------------------------------------------- batches */
/* Independent shells of different sizes advanced together (no counterpart
 * in the reference, whose driver runs one shell): each phase of a stage is
 * one launch over all shells, a step one hipGraph.  The shells must share
 * precision, rk_order and imex.  cs_batch_shell lends the handle of shell i
 * for upload, download and diagnostics (the batch owns it).
 * cs_batch_autotune times every phase at 64/128/256-thread workgroups on the
 * batch's current state, keeps the fastest (block_sizes[CS_BP_COUNT] may be
 * NULL), and leaves the stage scratch buffers modified. */
typedef struct cs_batch cs_batch;
#define CS_BP_COUNT 9 /* exchange, radial pad, x/y fluxes, z fluxes, EMFs, faces, zones, implicit source, CFL */
int cs_batch_create(const cs_config* cfgs, int32_t n, cs_batch** out);
int cs_batch_destroy(cs_batch* b);
int cs_batch_shell(cs_batch* b, int32_t i, cs_handle** h);
int cs_batch_evaluate_timestep(cs_batch* b);
int cs_batch_advance(cs_batch* b, int32_t nsteps, int32_t use_graph);
int cs_batch_time_steps(cs_batch* b, int32_t nsteps, float* ms_total);
int cs_batch_autotune(cs_batch* b, int32_t iters, int32_t* block_sizes);

/* ---------------------------------------------------------------- P1 geometry
 * The quadratic isoparametric zone and face geometry the reference's P1
 * reconstruction reads, built on the GPU for the six sectors (one patch each)
 * and returned in the reference's per-patch Fortran layouts, sector slowest:
 *   vert    (0-g:N+g, 0-g:N+g, 0-g:nr+g, 3)     mask  (1-g:N+g)^2 x (1-g:nr+g)
 *   zones   (1-g:N+g)^2 x (1-g:nr+g) x trailing dims
 *   x-faces (0-g:N+g, 1-g:N+g, 1-g:nr+g) x trailing dims
 *   y-faces (1-g:N+g, 0-g:N+g, 1-g:nr+g) x trailing dims
 *   z-faces (1-g:N+g, 1-g:N+g, 0-g:nr+g) x trailing dims
 * Any null pointer skips that output. */
typedef struct cs_face_quad {
  double* pts;      /* (face, 3, 5) quadrature points                      */
  double* wts;      /* (face, 5) weights divided by the area               */
  double* normal;   /* x, y: (face, 3) at the centre; z: (face, 3, 5)       */
  double* area;     /* (face)                                              */
  double* centroid; /* (face, 3)                                           */
} cs_face_quad;
typedef struct cs_geometry {
  double* vert;     /* EXTRACT_VERT_COORDS_PATCH (210extract_vert_coords_patch.f:7) */
  int32_t* mask;
  cs_face_quad xf, yf, zf;  /* COORDS_TO_WT_QUADPTS_FACES_ISO2 (220coords_to_wt_quadpts_faces_iso2.f:7) */
  double* scale;            /* v_scale_length, same routine (:2212-2290)            */
  cs_face_quad xb, yb, zb;  /* COORDS_TO_WT_QUADPTS_BFLD_ISO2 (230coords_to_wt_quadpts_bfld_iso2.f:8):
                               pts, wts (times the reciprocal sum), normal; area and
                               centroid are not used */
  double* v_pts;            /* COORDS_TO_WT_QUADPTS_VOLUME_ISO2 (240coords_to_wt_quadpts_volume_iso2.f:7): (zone, 3, 8) */
  double* v_wts;            /* (zone, 8)                                            */
  double* v_tan;            /* (zone, 3, 3, 8): d x_c / d xi_q at each point         */
  double* v_vol;            /* (zone)                                               */
  double* v_cen;            /* (zone, 3)                                            */
  double* moments;          /* INIT_BASIS_FUNCTION_COEFS (250init_basis_function_coefs.f:5): (zone, 4) */
} cs_geometry;
/* element counts of the six sectors: vertices, zones, x-, y-, z-faces */
int cs_geometry_sizes(const cs_config* cfg, int64_t* sizes5);
/* builds every output on the GPU and copies the requested ones back;
 * kernel_ms (optional, 4 floats): vertices + mask, faces + scale length,
 * volume quadrature, basis moments */
int cs_build_geometry(const cs_config* cfg, const cs_geometry* out, float* kernel_ms);
/* STENCIL_INDEXING_TRICORNERS (211stencil_indexing_tricorners.f:6) for a
 * sector that is one patch: istencil (1-g:N+g, 1-g:N+g, 2, 12), the zone (x, y)
 * reached in each of the twelve directions (R, L, U, D, RR, LL, UU, DD, RU, LU,
 * LD, RD), re-pointed around the tri-corner blocks; the same for every sector */
int cs_stencil_indexing(const cs_config* cfg, int32_t* istencil);
/* BUILD_MODE_INDICES (204build_mode_indices.f:7): monomial powers of the
 * modes of a reconstruction of the given order (1-4) */
int cs_mode_indices(int32_t order, int32_t* n_modes, int32_t* ii, int32_t* jj, int32_t* kk);
/* The geometry CS_RECON_P1ADP evaluates, computed on the host (no GPU):
 * COORDS_TO_WT_QUADPTS_FACES_ISO2 / _BFLD_ISO2 (220...f:7, 230...f:8) of face
 * family fam (0 x, 1 y, 2 z) from its separable form (vertices rb(k) u(i, j),
 * 210...f:300-308) in the layouts of cs_face_quad (pts, wts, normal, area,
 * centroid) plus the B-field weights bwts (face, 5); and the zone centroid
 * (zone, 3), volume and scale length of 240...f:7 / 220...f:2212-2290.  Any
 * output may be NULL. */
int cs_p1s_face_geometry(const cs_config* cfg, int32_t fam, double* pts, double* wts, double* normal,
                         double* area, double* centroid, double* bwts);
int cs_p1s_zone_geometry(const cs_config* cfg, double* centroid, double* volume, double* scale);

/* ---- P1 zone reconstruction of one patch ---------------------------------
 * LIMIT_ALL_SLOPES_P1 (400riemann_integrate/410limit_all_slopes_p1.f:9-160,
 * called from the driver's stage loop) with ENFORCE_PRESSURE_POSITIVITY_PRIM
 * (440enforce_pressure_positivity_prim.f:6): ideal MHD, P1 modes (mean and
 * three gradients), PRESSURE_POSITIVITY 1, no B0.  Every array is a DEVICE
 * pointer in the reference's own column-major layout for one patch of
 * NGhost = nghost >= 4 (the reference's loop bounds reach 4 ghost layers):
 *   zone    (1-g:N+g, 1-g:N+g, 1-g:nr+g) [, component 8][, mode 4]
 *   face B  (0-g:N+g, 0-g:N+g, 0-g:nr+g) Bxf, Byf, Bzf (mode 1 only) [, point 5]
 *   x-face  (0-g:N+g, 1-g:N+g, 1-g:nr+g), y-face (1-g:N+g, 0-g:N+g, 1-g:nr+g),
 *   z-face  (1-g:N+g, 1-g:N+g, 0-g:nr+g) [, xyz][, point]
 *   vert    (0-g:N+g, 0-g:N+g, 0-g:nr+g, 3); stencil (1-g:N+g, 1-g:N+g, 2, 12)
 * (the layouts cs_build_geometry and cs_stencil_indexing produce, one sector
 * per patch).  As in the reference, scale (v_scale_length) is written: the
 * inactive tri-corner zones get smallnum.  V_min_local / V_max_local are HOST
 * arrays of 9 (may be NULL).  stream: a hipStream_t, NULL = default stream.
 * The call waits on the stream once mid-way (the positivity limiter's
 * loop-order dependence) and at the end when the min/max are requested; calls
 * on one device share its work memory and are serialised. */
typedef struct cs_p1_patch {
  int32_t ntp, nr, nghost;
  int32_t itri[4];                 /* itri_patch_LD, _RD, _RU, _LU (1 = tri-corner) */
  int32_t pad;
  double gamma, molewt, rhofloor, prsfloor, smallnum, flatten_coef, prs_flatten_coef;
  const double* vert;              /* vert_coords_patch */
  const int32_t* mask;             /* imask_active_patch */
  const int32_t* stencil;          /* istencil_patch */
  const double *xf_normal, *xf_area, *xf_centroid; /* xf_quadnorml_patch, xf_area_patch, xf_centroid_patch */
  const double *yf_normal, *yf_area, *yf_centroid;
  const double *zf_area, *zf_centroid;
  const double *xb_pts, *xb_wts, *xb_normal;       /* xb_quadpts_patch, xb_quadwt_patch, xb_quadnorml_patch */
  const double *yb_pts, *yb_wts, *yb_normal;
  const double *zb_pts, *zb_wts, *zb_normal;       /* zb normal: (face, 3, 5) */
  double* scale;                   /* v_scale_length (in/out) */
  const double* v_centroid;        /* v_centroid_patch */
  double *U, *ent, *Bxf, *Byf, *Bzf;               /* in/out: U, entropygr, Bxf, Byf, Bzf */
  double *V, *B1_zc, *B2_zc, *B3_zc, *Ex_zc, *Ey_zc, *Ez_zc;
  double *gamma_eos, *gamma_sound, *mean_mol_wt, *divvel, *msonic, *flatten, *work;
  double *Bxf_q, *Byf_q, *Bzf_q;   /* Bxf_quadpts, Byf_quadpts, Bzf_quadpts */
  double *V_min_local, *V_max_local;
} cs_p1_patch;
int cs_limit_all_slopes_p1(const cs_p1_patch* p, void* stream);
/* the same, timing its stages with HIP events on the stream (no counterpart
 * in the reference: measurement only).  kernel_ms[CS_P1_STAGES]: primitives;
 * face-B clearing and tri-corner sync; flattener; WENO + positivity (with any
 * replay); face B points; constrained least squares + zone E; WENO of E;
 * min/max */
#define CS_P1_STAGES 8
int cs_limit_all_slopes_p1_timed(const cs_p1_patch* p, void* stream, float* kernel_ms);
/* The same as cs_limit_all_slopes_p1, allocating nothing: no counterpart in
 * the reference; for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Its work memory (positivity flags, replay list and counters, min/max
 * partials) is a fixed 48 MiB device array per GPU, held from the library's
 * load until the process ends; a patch needing more (about 3.5 M zones with
 * ghosts, e.g. 192^2 x 64 at nghost 4) is refused with CS_E_INVALID before
 * anything runs.  Every call on one GPU shares that GPU's array: calls that
 * overlap on one GPU race. */
int cs_limit_all_slopes_p1_noalloc(const cs_p1_patch* p, void* stream);

/* ---- third order: the zone-centred reconstruction of LIMIT_ALL_SLOPES_P12 --
 * The geometry of NUMORDER 3 (the driver builds it once, as at second
 * order): COORDS_TO_WT_QUADPTS_VOLUME_ISO2 (240coords_to_wt_quadpts_volume_iso2.f:7)
 * with the 3x3x3 Gauss rule (:134-147) and INIT_BASIS_FUNCTION_COEFS
 * (250init_basis_function_coefs.f:5) with the ten modes of
 * BUILD_MODE_INDICES at order 3 (204build_mode_indices.f:55-95: 1, x, y, z,
 * x^2, y^2, z^2, xy, yz, xz), for the six sectors on the GPU, copied back to
 * HOST arrays in the layouts of cs_geometry: v_pts (zone, 3, 27), v_wts
 * (zone, 27), v_vol, v_cen (zone, 3), moments (zone, 10).  Any null output
 * is skipped. */
typedef struct cs_volume_p12 {
  double *v_pts, *v_wts, *v_vol, *v_cen, *moments;
} cs_volume_p12;
int cs_build_volume_p12(const cs_config* cfg, const cs_volume_p12* out);
/* LIMIT_ALL_SLOPES_P12 (400riemann_integrate/421limit_all_slopes_p12.f:9,
 * called from the driver's stage loop at NUMORDER 3, DRV:2395-2398) on one
 * patch, the third-order WENO-AO reconstruction: the steps of
 * LIMIT_ALL_SLOPES_P1 up to its WENO (the P1 gradients of the eight
 * primitives left unflattened, 421…:3435-3448); step 7.b (:3811-4557): on
 * the active zones of -1 .. N+2 (k -1 .. NR+2) the nine modes of the
 * least-squares quadratic fit to the 18-zone central stencil
 * (RECONSTRUCTION_COEFS, 251reconstruction_coefs.f:3), non-linearly
 * hybridised with the P1 gradients, then times (1 - flattener); step 7.c:
 * ENFORCE_PRESSURE_POSITIVITY_PRIM (440enforce_pressure_positivity_prim.f:6)
 * with the quadratic modes in the node values (:594-600), reducing all ten
 * modes, with the loop-order replay of cs_limit_all_slopes_p1; steps 8.a-c:
 * the normal B at the five B-field points of each face from both zones'
 * quadratic fields about their own moments (:4817-4879), shifted to the
 * face's mean B; step 8.d (:5574-6680): the zone's 30 quadratic B
 * coefficients from the constrained least squares (the six face means
 * exactly; n . B at the 30 face points, the WENO-AO modes and a zero
 * divergence and divergence gradient in the least-squares sense), U's B
 * means and energy updated; step 9 (:6681-9353): E = -v x B at the centroid
 * from the quadratic fields, its P1 WENO gradients, and the WENO-AO with
 * the quadratic interpolant through the 19 zones' point values
 * (INTERPOLATION_COEFS, 252interpolation_coefs.f:5), in the (k, j, i) loop
 * order of the reference's serial build, whose loop overwrites each point
 * value that later zones read as stencil values; step 10: the min/max and
 * divergence diagnostics of cs_limit_all_slopes_p1.  The patch is a
 * cs_p1_patch whose U, V (zone, 8, 10), B1..3_zc and Ex..Ez_zc (zone, 10)
 * hold ten modes and whose v_centroid comes from the 27-point rule, plus
 * the ten basis moments (DEVICE, zone x 10).  Work memory, stream and the
 * host min/max as cs_limit_all_slopes_p1 (the call waits on the stream
 * mid-way and at the end).  CS_E_INVALID, with the outputs undefined, when the
 * stencil table breaks that order as the call runs it (level
 * (i+1) + 3 (j+1) + 9 (k+1)), which entries within two zones in i and j never
 * do; the cubed-sphere table reaches two, at the corner strips. */
typedef struct cs_p12_patch {
  cs_p1_patch p1;
  const double* moments; /* Coef_Basis_Moments (zone, 10) */
} cs_p12_patch;
int cs_limit_all_slopes_p12(const cs_p12_patch* p, void* stream);
/* the same, timing its stages with HIP events on the stream (no counterpart
 * in the reference: measurement only), kernel_ms[CS_P1_STAGES] as for
 * cs_limit_all_slopes_p1_timed: primitives; face-B clearing and tri-corner
 * sync; flattener; WENO, WENO-AO and positivity (with any replay); face B
 * points; constrained least squares + zone E; WENO and WENO-AO of E;
 * min/max */
int cs_limit_all_slopes_p12_timed(const cs_p12_patch* p, void* stream, float* kernel_ms);

/* ---- P1 facial fluxes of one patch -----------------------------------------
 * MAKE_FACIAL_FLUXES (400riemann_integrate/450make_facial_fluxes.f:5-60) at
 * second order: per face, the two zones' P1 fields (V, entropygr from
 * cs_limit_all_slopes_p1) at the face centroid, rotated into the face frame,
 * OneD_RS_MHD_Ptwise (600riemann_1d/610...f:13), the fluctuations rotated back
 * and scattered into dU_dt of both zones.  MHD, no facial electric fields
 * (Efx, Efy, Efz are zeroed over the zone range, as the routine does), no B0.
 * Device arrays in the layouts of cs_p1_patch; face outputs (S_*, Fluct_*,
 * EntFlux_*) in the face-B layout (0-g:N+g)^3-shaped [, 8]; dU_dt (zone, 8).
 * The call waits on the stream once (the count of faces re-pointed at the
 * tri-corners). */
typedef struct cs_p1_flux {
  int32_t ntp, nr, nghost, pad;
  double gamma, rhofloor, prsfloor, smallnum;
  const double* vert;
  const int32_t* mask;
  const int32_t* stencil;
  const double *xf_pts, *xf_normal, *xf_area, *xf_centroid; /* x-face area quadrature (220...f) */
  const double *yf_pts, *yf_normal, *yf_area, *yf_centroid;
  const double *zf_pts, *zf_normal, *zf_area, *zf_centroid; /* zf normal: (face, 3, 5) */
  const double *scale, *v_volume, *v_centroid;
  const double *V, *ent, *flatten;                        /* from LIMIT_ALL_SLOPES_P1 */
  double *S_L, *S_R, *S_D, *S_U, *S_B, *S_T;             /* S_L_gr ... S_T_gr */
  double *Fluct_x_Pl, *Fluct_y_Pl, *Fluct_z_Pl, *Fluct_x_Mi, *Fluct_y_Mi, *Fluct_z_Mi;
  double* dU_dt;
  double *EntFlux_x, *EntFlux_y, *EntFlux_z;
  double *Efx, *Efy, *Efz;
} cs_p1_flux;
int cs_make_facial_fluxes_p1(const cs_p1_flux* p, void* stream);

/* ---- P1 non-stiff sources of one patch -------------------------------------
 * BUILD_NON_STIFF_SOURCES (400riemann_integrate/470build_non_stiff_sources.f:6-7),
 * MHD, adiabatic, with no external source (the routine's application hook
 * MAKE_EXTERNAL_SOURCE contributes nothing): Source (zone, 8) is cleared; on
 * the active zones of (0..N+1)^2 x (0..nr+1), when mstar or |omega| exceeds
 * smallnum, the point-mass gravity -newtonsc mstar rho r / |r|^3 at the zone
 * centroid, the centrifugal -rho omega x (omega x r) and the Coriolis
 * -2 rho omega x v (omega = omega_passive, v and rho from V mode 1), the
 * energy source v . S, stored in Source and added to dU_dt.  Device arrays in
 * the layouts of cs_p1_patch.  No wait on the stream. */
typedef struct cs_p1_source {
  int32_t ntp, nr, nghost, pad;
  double newtonsc, mstar, smallnum, omega[3];
  const int32_t* mask;
  const double *v_centroid, *V;    /* v_centroid_patch, V (mode 1 read) */
  double *Source, *dU_dt;          /* (zone, 8) each */
} cs_p1_source;
int cs_build_non_stiff_sources_p1(const cs_p1_source* p, void* stream);

/* PAD_BOUNDARY_R (800applications_cs/826pad_boundary_r.f:1; the driver calls
 * it after the exchange, 1000riemann_geomesh_cs_mhd.f:2238-2275) of one
 * patch, outflow pads (bcarr_r = 2), no application branch.  For the ghost
 * layers 1..NGhost below layer 1 and above layer nr of every zone column whose
 * mask is 1: inner side, mode 1 of U, the column's Bxf (x-faces i-1, i) and
 * Byf (y-faces j-1, j) copied from layer 1 and Bzf at radial faces -1..-NGhost
 * from face 0; outer side, mode 1 of U scaled by V(nr) / V(ghost), Bzf from
 * face nr and Bxf, Byf from layer nr scaled by the ratio of the face areas.
 * bcarr_r is (1-g:N+g, 1-g:N+g, 2), inner then outer; a side whose code is
 * not 2 is left alone.  The entropy is not touched (the reference routine
 * takes it but does not write it).  Device arrays in the layouts of
 * cs_p1_patch; U with any number of modes (only mode 1 is touched).  No wait
 * on the stream. */
typedef struct cs_p1_pad {
  int32_t ntp, nr, nghost, pad;
  const int32_t *mask, *bcarr_r;
  const double *xf_area, *yf_area, *zf_area, *v_volume;
  double *U, *Bxf, *Byf, *Bzf;
} cs_p1_pad;
int cs_pad_boundary_r_p1(const cs_p1_pad* p, void* stream);

/* ---- P1 stage update of one patch ------------------------------------------
 * UPDATE_IMEX_RK_STEP_P1 (400riemann_integrate/484update_imex_rk_step_p1.f:7-26),
 * MHD: U(:, mode 1) += dtmove dU_dt on every zone; on the owned zones the
 * entropy from the face entropy fluxes (EntFlux_* of cs_make_facial_fluxes_p1),
 * then CONS_TO_PRIM (U updated, flag_entprs = 1 where the pressure was taken
 * from the floor or the entropy); on the owned faces constrained transport of
 * Bxf, Byf, Bzf with the edge electric fields Efx, Efy, Efz over the edges'
 * radial lengths and arc lengths r acos(u1 . u2).  Device arrays in the
 * layouts of cs_p1_patch / cs_p1_flux; U and entropy with their 4 modes
 * (only mode 1 is touched).  No wait on the stream. */
typedef struct cs_p1_update {
  int32_t ntp, nr, nghost, pad;
  double dtmove, gamma, rhofloor, prsfloor;
  const double* vert;
  const double *xf_area, *yf_area, *zf_area, *v_volume;
  double *U, *ent, *Bxf, *Byf, *Bzf;
  const double* dU_dt;
  const double *EntFlux_x, *EntFlux_y, *EntFlux_z, *Efx, *Efy, *Efz;
  double* flag_entprs;
} cs_p1_update;
int cs_update_imex_rk_step_p1(const cs_p1_update* p, void* stream);

/* ---- P1 edge electric fields of one patch ----------------------------------
 * MAKE_EDGE_EFIELDS_P1 (400riemann_integrate/460make_edge_fields_p1.f:5-20) at
 * second order (NUMORDER 2, USE_FACIAL_EFIELDS 0, no B0): for every edge of
 * the owned faces (x-edges i 1..N, j 0..N, k 0..nr; y-edges i 0..N, j 1..N;
 * z-edges i, j 0..N, k 1..nr) the mean mode-1 E of its four zones along the
 * edge, the face-point normal B (Bxf_q, Byf_q, Bzf_q of
 * cs_limit_all_slopes_p1) and the face wave speeds (S_* of
 * cs_make_facial_fluxes_p1) of the four faces around it, symmetrised to
 * +-max|S|, into Two_D_RS_Ez_4State_HLL (620riemann_2d/630...f:10); z-edges
 * with an inactive zone get zero, and at a flagged tri-corner the edge of the
 * three active zones adds Two_D_RS_Ez_3State_HLL (631...f:1).  Efx, Efy, Efz
 * (face-B layout) are written over their whole arrays, zero off those edges.
 * Device arrays in the layouts of cs_p1_patch.  No wait on the stream. */
typedef struct cs_p1_edge {
  int32_t ntp, nr, nghost, pad;
  int32_t itri[4];                 /* itri_patch_LD, _RD, _RU, _LU */
  const double* vert;
  const int32_t* mask;
  const double *Ex_zc, *Ey_zc, *Ez_zc;             /* zone E (mode 1 read) */
  const double *Bxf_q, *Byf_q, *Bzf_q;             /* Bxf_quadpts ... (face-B layout, 5 points) */
  const double *S_L, *S_R, *S_D, *S_U, *S_B, *S_T; /* S_L_gr ... S_T_gr */
  double *Efx, *Efy, *Efz;
} cs_p1_edge;
int cs_make_edge_efields_p1(const cs_p1_edge* p, void* stream);

/* ---- P1 defect correction over the six patches of a shell -------------------
 * The driver's defect correction (FILL_FLUX_SND_PATCH 500riemann_main/599...,
 * FETCH_BUFFERS_FILL_IN_FLUX_X/Y/Z 300riemann_module/334-336...,
 * SYNCHRONIZE_EFZ_QUAD_CORNER 340..., UNPACK_DEFECT_FLUXES 341...,
 * UPDATE_DEFECT_IMEX_RK 400riemann_integrate/485update_defect_imex_rk.f) in the
 * form that comes before the update: for every face two patches share, both
 * patches' Fluct_*_Pl, Fluct_*_Mi and EntFlux_* become the mean of the two
 * (334...f:100-116) and the patch's zone next to it gets -(mean - own) A / V
 * added to dU_dt; for every edge two or three patches share, Ef* becomes the
 * mean (334...f:128-160; three at the tri-corners, 340...).  The update is
 * linear in these, so updating after this equals the reference's update plus
 * its defect step.  faces / edges: DEVICE arrays of the pairings (built once
 * by the host from the patches' geometry): face a and face b with their
 * family (0 x, 1 y, 2 z), face-B-cube index, zone and area index, the side of
 * the face the patch's zone is on (0 minus, 1 plus) and whether the two
 * normals are opposite (then b's plus side is a's minus side and its entropy
 * flux changes sign); an edge group lists its 2 or 3 edges with the sign that
 * turns each into the first one's direction.  The table entries are not
 * checked on the device: they must index the patches' arrays (patch 0-5,
 * family 0-2, indices within the layouts), as the host builder's are.  No
 * wait on the stream. */
typedef struct cs_p1_face_pair {
  int32_t pa, fa, sa, pb, fb, sb, flip, pad;
  int64_t ia, za, aa, ib, zb, ab;
} cs_p1_face_pair;
typedef struct cs_p1_edge_group {
  int32_t n, pad;
  int32_t p[3], f[3], sign[3], pad2;
  int64_t idx[3];
} cs_p1_edge_group;
typedef struct cs_p1_patch_fluxes {
  double *Fluct_x_Pl, *Fluct_y_Pl, *Fluct_z_Pl, *Fluct_x_Mi, *Fluct_y_Mi, *Fluct_z_Mi;
  double *EntFlux_x, *EntFlux_y, *EntFlux_z, *Efx, *Efy, *Efz, *dU_dt;
  const double *xf_area, *yf_area, *zf_area, *v_volume;
} cs_p1_patch_fluxes;
typedef struct cs_p1_shell_sync {
  int32_t ntp, nr, nghost, pad;
  cs_p1_patch_fluxes patch[6];
  const cs_p1_face_pair* faces;
  int64_t nface;
  const cs_p1_edge_group* edges;
  int64_t nedge;
} cs_p1_shell_sync;
int cs_sync_shared_p1(const cs_p1_shell_sync* p, void* stream);

/* ---- P1 cleanup and time step of one patch ---------------------------------
 * CLEANUP_AFTER_UPDATE (400riemann_integrate/485cleanup_after_update.f:1-25),
 * MHD, on the fields of cs_p1_patch: gamma_eos, gamma_sound, mean_mol_wt set;
 * modes 2-4 of U, entropygr and V cleared; inactive (tri-corner) zones made
 * safe (scale = smallnum); face B between two inactive zones cleared and
 * synchronised at the tri-corners; on every zone of an active column the zone
 * B as the least-squares fit of its six face values over the face normals
 * (x and y faces: xf_normal, yf_normal; both z faces: the radial direction of
 * zf_centroid), the energy rebalanced by the change of magnetic energy, and
 * B1_zc..B3_zc (mode 1) set; then CONS_TO_PRIM on every zone (V, U, entropy
 * mode 1).  Reads mask, stencil, xf_normal, yf_normal, zf_centroid; the other
 * geometry pointers may be NULL.  nghost >= 1.  No wait on the stream. */
int cs_cleanup_after_update_p1(const cs_p1_patch* p, void* stream);
/* EVALUATE_TIMESTEP (400riemann_integrate/490evaluate_timestep.f:1-23), MHD:
 * CONS_TO_PRIM of the owned active zones (U and V mode 1 rewritten, the
 * entropy array left as it is), msonic = sqrt((gamma_sound p + B^2 / 4pi) /
 * rho), and *dt_new = min(1.1 dtcur, min cfl scale / (|v| + msonic)).  dt_new
 * is a DEVICE pointer to one double, written on the stream (no wait), so the
 * minimum over many patches can stay on the device.  Reads mask, scale,
 * gamma_sound; writes U, V, msonic. */
int cs_evaluate_timestep_p1(const cs_p1_patch* p, double cfl, double dtcur, double* dt_new, void* stream);

/* ------------------------------------------------------------ checkpoints */
/* Checkpoint / restart.  Replaces the driver's riemann.in (time, dtcur,
 * n_cc_components, dump number: read DRV:1527-1547 and checked against the
 * dump DRV:1618-1647, written DRV:1798-1802) and its Silo READDATA /
 * WRITEDATA of U, entropy and face B (DRV:1564-1605, DRV:1845-1930) with one
 * versioned binary file: a 4096-byte header (format version, byte-order
 * mark, the cs_config, the fields below, a checksum per section and one of
 * the header), then U, ent, Bxf, Byf, Bzf in exactly the layouts of
 * cs_upload_state above, ghosts included, each 4096-byte aligned.  Files
 * are written as path.tmp and renamed into place, so a dump that exists is
 * complete.  A reader refuses (CS_E_FORMAT) a bad magic, another version,
 * another byte order, a header whose sections do not fit its grid or the
 * file, and any checksum mismatch; it refuses (CS_E_INVALID) a grid (ntp,
 * nr, nghost, radial mesh, radii) other than the caller's.  The scheme
 * fields of the configuration may differ: a run may restart with another
 * scheme on the same grid. */
typedef struct cs_dump_meta {
  double time;      /* simulation time (riemann.in's time)                */
  double dtcur;     /* current time step (riemann.in's dtcur)             */
  int64_t filenum;  /* dump number (riemann.in's filenum)                 */
  int64_t step;     /* completed time steps                               */
  int32_t n_cc_components; /* 8 (filled by the readers)                   */
  int32_t version;  /* format version (filled by the readers)             */
  cs_config cfg;    /* configuration the state was written with (readers) */
} cs_dump_meta;
/* host arrays (no GPU): write / read a state held in the upload layout */
int cs_dump_write(const char* path, const cs_config* cfg, const cs_dump_meta* meta, const double* U,
                  const double* ent, const double* Bxf, const double* Byf, const double* Bzf);
int cs_dump_read(const char* path, const cs_config* cfg, cs_dump_meta* meta, double* U, double* ent,
                 double* Bxf, double* Byf, double* Bzf);
/* the header alone (checked as by cs_dump_read, no grid comparison) */
int cs_dump_info(const char* path, cs_dump_meta* meta);
/* a handle's state U^n straight from device memory through a pinned
 * 64 MiB staging buffer (no host copy of the state); dtcur is the handle's
 * current time step, meta->dtcur is not used */
int cs_write_dump(cs_handle* h, const char* path, const cs_dump_meta* meta);
/* into a handle, as cs_upload_state (exchange included), and the handle's
 * time step set to the dump's dtcur; meta receives the header.  A file
 * refused by its header or grid leaves the handle as it was; one whose
 * section fails its checksum leaves the handle's state undefined. */
int cs_read_dump(cs_handle* h, const char* path, cs_dump_meta* meta);

int cs_last_error(char* buf, size_t len);
int cs_device_info(char* buf, size_t len);

#ifdef __cplusplus
}
#endif
#endif
