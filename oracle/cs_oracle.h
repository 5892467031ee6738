/* cs_oracle.h -- CPU oracle for the CubeSphere cell-update path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product links, loads or calls
 * this; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * use it, as the checker.
 *
 * Plain C restatement of the reference's algorithm for this path; every
 * function cites the reference file:line it follows.  Arrays use the same
 * Fortran patch layout as include/cubesphere_hip.h.
 */
#ifndef CS_ORACLE_H
#define CS_ORACLE_H
#include <stdint.h>

typedef struct or_config { /* field-for-field the same as cs_config */
  int32_t ntp, nr, nghost, recon, riemann, emf, precision, rk_order, radial_bc, radial_mesh;
  double rmin, rmax, gamma, rhofloor, prsfloor, cfl, grav;
  int32_t imex, bc_problem;
  double omega[3];
} or_config;

typedef struct or_grid or_grid;

or_grid* or_grid_new(const or_config* c);
void or_grid_free(or_grid* g);
int or_grid_dims(const or_grid* g, int64_t* out4); /* cpp, fpp, t2, nmap */

/* 201/203: unit lattice (-g:N+g, -g:N+g, 3, 6), reference corner blocks = 0 */
void or_unit_lattice(int N, int g, double* coords);
/* 202: rb(-g:nr+g) */
void or_radial_mesh(const or_config* c, double* rb);

/* metric (same layout as cs_build_metric) */
void or_metric(const or_grid* g, double* vol, double* xarea, double* yarea, double* zarea);

/* ghost exchange (maps found by geometric matching) + corner fill + radial bc */
void or_halo(const or_grid* g, double* U, double* ent, double* Bx, double* By, double* Bz);
/* number of ghost-map entries whose geometric match failed (0 expected) */
int or_map_failures(const or_grid* g);

/* face fluxes (9 comps * area, global frame) and speeds for dir 0..2 */
void or_fluxes(const or_grid* g, const double* U, const double* ent, const double* Bx,
               const double* By, const double* Bz, double** F, double** S);
void or_emf(const or_grid* g, const double* U, const double* ent, const double* Bx,
            const double* By, const double* Bz, double** S, double** E);
void or_update(const or_grid* g, int irk, double dt, const double* const* in,
               const double* const* u0, double* const* out, double** F, double** E);
double or_cfl(const or_grid* g, const double* U, const double* ent, double dt_prev);

/* whole steps on a state (U, ent, Bx, By, Bz); *dt is updated by the CFL */
void or_advance(const or_grid* g, int nsteps, double* U, double* ent, double* Bx, double* By,
                double* Bz, double* dt);
/* single stage from a given input */
void or_stage(const or_grid* g, int irk, double dt, double* const* in, double* const* u0,
              double* const* out);

/* pointwise (for known-answer tests against the reference build) */
int or_cons_to_prim(double gam, double rhof, double prsf, double* ent, double* U, double* V);
void or_prim_to_cons(double gam, double rhof, double prsf, double ent, double* V, double* U);
void or_wavspd_x(double gam, const double* V, double* smin, double* smax);
void or_flux_x(const double* V, const double* U, double* F);
void or_hll(double gam, double rhof, double prsf, double* VL, double* VR, double entL,
            double entR, double* F, double* entF, double* SL, double* SR);
void or_hlld(double gam, double rhof, double prsf, double* VL, double* VR, double entL,
             double entR, double* F, double* entF, double* SL, double* SR);
double or_weno5(double v0, double v1, double v2, double v3, double v4);
double or_emf4_hll(const double* s4, double cs_t, double sn_t, const double* ez4,
                   const double* bn2, const double* by2);
double or_emf3_llf(double S, const double* ez3, const double* bn3);
/* rotating-frame centrifugal + Coriolis sources added to L (rho, rho v, E rows) */
void or_rotation_source(const double* om, double rhofloor, const double* r, const double* u, double* L);
/* CLEANUP_AFTER_UPDATE on the owned zones of a state */
void or_cleanup(const or_grid* G, double* U, double* ent, double* Bx, double* By, double* Bz);
/* panel-edge face fluxes and edge fields made unique (the averages taken
 * before the update) */
void or_sync_shared(const or_grid* G, double* const* F, double* const* E);
/* IMEX-SSP2(2,2,2) with the implicit gravity source, u = (rho, rho v, E) */
void or_imex_pre(double dt, double gr, const double* n, double* u);
void or_imex_combine(int stage, double dt, double gr, const double* n, const double* ui, const double* uz,
                     const double* L, double* u);

/* quadratic isoparametric geometry of the P1 reconstruction (210, 220, 240,
 * 204/250), six sectors; layouts in cs_oracle.c */
void or_vert_coords(const or_config* c, double* vert, int32_t* mask);
void or_quad_faces(const or_config* c, const double* vert, double* const* xf, double* const* yf,
                   double* const* zf, double* scl);
void or_quad_bfield(const or_config* c, const double* vert, double* const* xb, double* const* yb,
                    double* const* zb);
void or_quad_volume(const or_config* c, const double* vert, double* pts, double* wts, double* tan,
                    double* vol, double* cen);
void or_mode_indices(int32_t* ii, int32_t* jj, int32_t* kk);
/* 211: twelve-direction stencil zone table with the tri-corner remaps,
 * (1-g:N+g, 1-g:N+g, 2, 12) for a one-patch sector */
void or_stencil_indexing(int N, int g, int32_t* st);
void or_basis_moments(const or_config* c, const int32_t* mask, const double* pts, const double* wts,
                      const double* cen, const double* scl, double* coef);
/* the same at third order (NUMORDER 3): 27 volume points, 10 modes */
void or_quad_volume3(const or_config* c, const double* vert, double* pts, double* wts, double* tan,
                     double* vol, double* cen);
void or_basis_moments10(const or_config* c, const int32_t* mask, const double* pts, const double* wts,
                        const double* cen, const double* scl, double* coef);

/* LIMIT_ALL_SLOPES_P1 (410limit_all_slopes_p1.f:9) with the pressure
 * positivity of 440 for one patch (cs_oracle_p1.c).  Field order and types
 * are those of cs_p1_patch in the product's C header, so one structure serves
 * both. */
typedef struct or_p1_patch {
  int32_t ntp, nr, nghost;
  int32_t itri[4]; /* itri_patch LD, RD, RU, LU */
  int32_t pad;
  double gamma, molewt, rhofloor, prsfloor, smallnum, flatten_coef, prs_flatten_coef;
  const double* vert;
  const int32_t* mask;
  const int32_t* stencil;
  const double *xf_normal, *xf_area, *xf_centroid;
  const double *yf_normal, *yf_area, *yf_centroid;
  const double *zf_area, *zf_centroid;
  const double *xb_pts, *xb_wts, *xb_normal;
  const double *yb_pts, *yb_wts, *yb_normal;
  const double *zb_pts, *zb_wts, *zb_normal;
  double* scale;
  const double* v_centroid;
  double *U, *ent, *Bxf, *Byf, *Bzf;
  double *V, *B1_zc, *B2_zc, *B3_zc, *Ex_zc, *Ey_zc, *Ez_zc;
  double *gamma_eos, *gamma_sound, *mean_mol_wt, *divvel, *msonic, *flatten, *work;
  double *Bxf_q, *Byf_q, *Bzf_q;
  double *V_min_local, *V_max_local;
} or_p1_patch;
int or_limit_all_slopes_p1(const or_p1_patch* p);
/* zones of the last call with reduced slopes, and with slopes removed */
void or_p1_positivity_stats(long* out);

/* MAKE_FACIAL_FLUXES (450make_facial_fluxes.f:5) of one patch at second order
 * (cs_oracle_p1.c); the same field order as cs_p1_flux in the product header */
typedef struct or_p1_flux {
  int32_t ntp, nr, nghost, pad;
  double gamma, rhofloor, prsfloor, smallnum;
  const double* vert;
  const int32_t* mask;
  const int32_t* stencil;
  const double *xf_pts, *xf_normal, *xf_area, *xf_centroid;
  const double *yf_pts, *yf_normal, *yf_area, *yf_centroid;
  const double *zf_pts, *zf_normal, *zf_area, *zf_centroid;
  const double *scale, *v_volume, *v_centroid;
  const double *V, *ent, *flatten;
  double *S_L, *S_R, *S_D, *S_U, *S_B, *S_T;
  double *Fluct_x_Pl, *Fluct_y_Pl, *Fluct_z_Pl, *Fluct_x_Mi, *Fluct_y_Mi, *Fluct_z_Mi;
  double* dU_dt;
  double *EntFlux_x, *EntFlux_y, *EntFlux_z;
  double *Efx, *Efy, *Efz;
} or_p1_flux;
int or_make_facial_fluxes_p1(const or_p1_flux* f);
/* the P1 stage of a shell (CS_RECON_P1ADP): one resolved flux x area per face
 * (riemann 0 HLL, 1 HLLD; wall: mirrored states on the radial walls), and the
 * owned zones' update with the stage weights and the IMEX gravity */
int or_p1s_face_fluxes(const or_p1_flux* f, int riemann, int wall, double* const* F, double* const* S);
int or_p1s_update_zones(int N, int NR, int g, const double* const* F, const double* vol, const double* cen,
                        const double* ui, const double* ei, const double* uz, const double* ez, double dt,
                        double wa, double wb, int mode, double grav, double gam, double rhof, double prsf,
                        double* uo, double* eo);
/* OneD_RS_MHD_Ptwise with all its outputs (face frame, x-direction) */
void or_rs_ptwise(double gam, double rhof, double prsf, double* VL, double* VR, double entL, double entR,
                  double* SL, double* SR, double* Us, double* Fs, double* entF, double* FlL, double* FlR,
                  double* UL, double* UR, double* FL, double* FR);

/* UPDATE_IMEX_RK_STEP_P1 (484update_imex_rk_step_p1.f:7), MHD; the same field
 * order as cs_p1_update in the product header */
typedef struct or_p1_update {
  int32_t ntp, nr, nghost, pad;
  double dtmove, gamma, rhofloor, prsfloor;
  const double* vert;
  const double *xf_area, *yf_area, *zf_area, *v_volume;
  double *U, *ent, *Bxf, *Byf, *Bzf;
  const double* dU_dt;
  const double *EntFlux_x, *EntFlux_y, *EntFlux_z, *Efx, *Efy, *Efz;
  double* flag_entprs;
} or_p1_update;
int or_update_imex_rk_step_p1(const or_p1_update* u);

/* CLEANUP_AFTER_UPDATE (485cleanup_after_update.f:4) and EVALUATE_TIMESTEP
 * (490evaluate_timestep.f:7) of one patch, MHD, on the fields of or_p1_patch */
int or_cleanup_after_update_p1(const or_p1_patch* p);
int or_evaluate_timestep_p1(const or_p1_patch* p, double cfl, double dtcur, double* dtnew);

/* MAKE_EDGE_EFIELDS_P1 (460make_edge_fields_p1.f:5) of one patch at second
 * order, and its tri-corner solver Two_D_RS_Ez_3State_HLL (631...f:1); the
 * same field order as cs_p1_edge in the product header */
typedef struct or_p1_edge {
  int32_t ntp, nr, nghost, pad;
  int32_t itri[4];
  const double* vert;
  const int32_t* mask;
  const double *Ex_zc, *Ey_zc, *Ez_zc;
  const double *Bxf_q, *Byf_q, *Bzf_q;
  const double *S_L, *S_R, *S_D, *S_U, *S_B, *S_T;
  double *Efx, *Efy, *Efz;
} or_p1_edge;
int or_make_edge_efields_p1(const or_p1_edge* e);
double or_emf3_hll(const double* sr, const double* sl, const double* ez, const double* bn);

/* BUILD_NON_STIFF_SOURCES (470build_non_stiff_sources.f:6) of one patch with
 * no external source; the same field order as cs_p1_source in the product
 * header */
typedef struct or_p1_source {
  int32_t ntp, nr, nghost, pad;
  double newtonsc, mstar, smallnum, omega[3];
  const int32_t* mask;
  const double *v_centroid, *V;
  double *Source, *dU_dt;
} or_p1_source;
int or_build_non_stiff_sources_p1(const or_p1_source* s);

/* PAD_BOUNDARY_R (826pad_boundary_r.f:1) of one patch, outflow pads; the same
 * field order as cs_p1_pad in the product header.  bcarr_r is
 * (1-g:N+g, 1-g:N+g, 2): inner, outer; 2 pads, anything else leaves the side */
typedef struct or_p1_pad {
  int32_t ntp, nr, nghost, pad;
  const int32_t *mask, *bcarr_r;
  const double *xf_area, *yf_area, *zf_area, *v_volume;
  double *U, *Bxf, *Byf, *Bzf;
} or_p1_pad;
int or_pad_boundary_r_p1(const or_p1_pad* p);

/* PROBLEM_GENERATOR_U_A (800applications_cs/801problem_generator_u_a.f:10):
 * the pointwise conserved state U_pt(8) and vector potential A_pt(3) at x,
 * MHD branch, for iproblem 0 (the coordinate fields of the messaging test,
 * :105-229 with TESTING_OF_MPI_MESSAGING) and 2 (the magnetised blast,
 * :249-371), plus two problems of this build with no reference counterpart:
 * 102 (the surface blast of the benchmark: p 10 in a 0.1 rad cap around
 * (1, 0, 0), B = 0.5 east) and 103 (the isothermal hydrostatic shell with
 * gravity 1/r^2 and a field from A = 0.05 / r^2 (yz, zx + xy / 2, xy)). */
typedef struct or_prob {
  int32_t iproblem, pad;
  double time, rmin, rmax, gamma, rhofloor, prsfloor, grav;
  uint64_t seed;   /* 103: the 1e-2 density noise of zone (sector, i, j, k) */
} or_prob;
void or_problem_generator_u_a(const or_prob* pb, double x, double y, double z, double* U_pt, double* A_pt);

/* INIT_PATCH_CUBE_SPHERE (800applications_cs/820init_patch_cube_sphere.f:8)
 * of one patch, MHD, second order, no B0 field: U (mode 1) the volume
 * average of U_pt over the 8 Gauss points of every active zone (v_pts,
 * normalised v_wts of COORDS_TO_WT_QUADPTS_VOLUME_ISO2), the entropy
 * (mode 1) from it; face B of every active face as the sum of A_pt . dl over
 * its four edges (chords radially, arcs r acos(u1 . u2) along the sphere)
 * over the face area.  U, ent, Bxf, Byf, Bzf are cleared first.  With
 * layers = 1 (inner) | 2 (outer) only the radial ghost layers are written
 * and nothing is cleared elsewhere: ADD_BOUNDARY
 * (800applications_cs/825add_boundary.f:12) with ibdy_inner / ibdy_outer. */
typedef struct or_p1_init {
  int32_t ntp, nr, nghost, sector;
  const double* vert;        /* (face-B layout) x 3 */
  const int32_t* mask;
  const double *v_pts, *v_wts; /* (zone, 3, 8), (zone, 8) */
  const double *xf_area, *yf_area, *zf_area;
  double *U, *ent, *Bxf, *Byf, *Bzf; /* U (zone, 8) mode 1, ent (zone) mode 1 */
} or_p1_init;
int or_init_patch_cube_sphere(const or_p1_init* p, const or_prob* pb, int32_t layers);


/* third order (cs_oracle_p1.c): MATINV (121matinv.f:8) of a row-major
 * n x n matrix (n <= 18; m destroyed), RECONSTRUCTION_COEFS (251...f:3) at
 * order 3, step 7.b of LIMIT_ALL_SLOPES_P12 (421...f:3811-4557) and the
 * whole routine (steps 2 to 10); U and V of the patch hold ten modes, mom the ten basis
 * moments of each zone (or_basis_moments10) */
int or_matinv(int n, double* m, double* minv);
void or_reconstruction_row3(double x, double y, double z, double l0, double lj, const double* c0, const double* cj,
                            double* row);
int or_weno_ao_p12(const or_p1_patch* p, const double* mom);
int or_limit_all_slopes_p12(const or_p1_patch* p, const double* mom);
int or_ao_fit_p12(const or_p1_patch* p, const double* mom, int i, int j, int k, int c, double* X);
/* INTERPOLATION_COEFS (252interpolation_coefs.f:5) at order 3 */
void or_interpolation_row3(double x, double y, double z, const double* c0, double* row);
#endif
