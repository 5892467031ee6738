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
