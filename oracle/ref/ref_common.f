! ref_common.f -- sets the RIEMANN_COM values the reference's mesh routines
! read (smallnum, pi, do_on_gpu), as 090init_riemann.f does at start-up.
! Test infrastructure only (see Makefile); compiled with the reference's
! include path so the common block has the reference's own layout.
      SUBROUTINE REF_INIT_COMMON
      IMPLICIT NONE
#include <cubesphere/000com/riemann.com>
      smallnum = 1.0e-12
      pi = 4.0 * ATAN ( 1.0)
      do_on_gpu = .FALSE.
      END SUBROUTINE REF_INIT_COMMON

! Physical constants LIMIT_ALL_SLOPES_P1 reads from RIEMANN_COM (the values
! 090init_riemann.f and the drivers set), for one call.
      SUBROUTINE REF_SET_P1_COMMON ( gam, mw, rhof, prsf, small,
     1                               flatc, prsflatc)
      IMPLICIT NONE
#include <cubesphere/000com/riemann.com>
      REAL gam, mw, rhof, prsf, small, flatc, prsflatc
      gamma = gam
      gamma_euler = gam
      molewt = mw
      gasconst = 1.0
      rhofloor = rhof
      prsfloor = prsf
      tempfloor = 1.0e-5
      smallnum = small
      flatten_coef = flatc
      prs_flatten_coef = prsflatc
      isospeed = 1.0
      barotropic_stiffness = 0.0
      magB_small_cgl = 1.0e-12
      eps_bdry_cgl = 1.0e-12
      END SUBROUTINE REF_SET_P1_COMMON

! CFL number of EVALUATE_TIMESTEP (490evaluate_timestep.f:158)
      SUBROUTINE REF_SET_CFL ( cfl)
      IMPLICIT NONE
#include <cubesphere/000com/riemann.com>
      REAL cfl
      cfl_coef = cfl
      END SUBROUTINE REF_SET_CFL
