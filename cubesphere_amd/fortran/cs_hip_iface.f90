! cs_hip_iface.f90 -- ISO_C_BINDING interface to libcubesphere_hip.so for the
! reference's Fortran driver (_I/_1000riemann_geomesh_cs_mhd/1000riemann_geomesh_cs_mhd.f).
!
! The driver keeps its module/subroutine API; under one preprocessor switch it
! calls these entry points instead of the per-patch routines:
!   cs_halo_exchange       <- LOAD_LOCAL_BUFFERS_* / FETCH_BUFFERS_FILL_IN_* (DRV:1746-1756)
!   cs_make_facial_fluxes  <- LIMIT_ALL_SLOPES_P1 + MAKE_FACIAL_FLUXES      (DRV:2396, 2485)
!   cs_halo_exchange_p1    <- FETCH_BUFFERS_FILL_IN_* on six P1 patches' arrays (DRV:1746-1756)
!   cs_sync_shared_p1      <- the defect correction over six P1 patches, before the update (DRV:3027-3461)
!   cs_limit_all_slopes_p1 <- LIMIT_ALL_SLOPES_P1 alone, argument for argument (DRV:2396)
!   cs_limit_all_slopes_p12 <- LIMIT_ALL_SLOPES_P12 alone at NUMORDER 3, argument for argument (DRV:2398)
!   cs_build_volume_p12    <- COORDS_TO_WT_QUADPTS_VOLUME_ISO2 + INIT_BASIS_FUNCTION_COEFS at NUMORDER 3
!   cs_make_facial_fluxes_p1 <- MAKE_FACIAL_FLUXES alone, argument for argument (DRV:2485)
!   cs_make_edge_efields   <- MAKE_EDGE_EFIELDS_P1                          (DRV:2548)
!   cs_make_edge_efields_p1 <- MAKE_EDGE_EFIELDS_P1 alone, on the arguments it reads (DRV:2548)
!   cs_build_non_stiff_sources_p1 <- BUILD_NON_STIFF_SOURCES alone, the routine's arguments (DRV:2603)
!   cs_pad_boundary_r_p1      <- PAD_BOUNDARY_R with outflow bcarr_r, one patch (DRV:2262)
!   cs_update_imex_rk_step_p1 <- UPDATE_IMEX_RK_STEP_P1 alone, argument for argument (DRV:2782)
!   cs_cleanup_after_update_p1 <- CLEANUP_AFTER_UPDATE alone, on the cs_p1_patch (DRV:2815)
!   cs_evaluate_timestep_p1 <- EVALUATE_TIMESTEP alone, dt into device memory  (DRV:2887)
!   cs_update_stage        <- UPDATE_IMEX_RK_STEP_P1 + CLEANUP_AFTER_UPDATE (DRV:2782, 2815)
!   cs_evaluate_timestep   <- EVALUATE_TIMESTEP                             (DRV:2887)
!   cs_rk_stage            <- the whole per-patch loop body                (DRV:2007-3120)
!   cs_write_dump / cs_read_dump <- riemann.in + WRITEDATA / READDATA      (DRV:1527-1666, 1783-1967)
!   cs_init_problem_p1     <- INIT_PATCH_CUBE_SPHERE (820...f:303) on the device (DRV:1386)
!   cs_add_boundary_p1     <- ADD_BOUNDARY (825...f:320) on the device (DRV:2304)
! State arrays keep the driver's layout U_amr(1-g:N+g,1-g:N+g,1-g:nr+g,8,6),
! entropygr_amr(...,6), Bxf_amr(0-g:N+g,0-g:N+g,0-g:nr+g,6) (MOD:502-552).
module cs_hip_iface
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: CS_RECON_PCM = 0, CS_RECON_WENO5 = 1, CS_RECON_P1ADP = 2
  integer(c_int), parameter :: CS_RS_HLL = 0, CS_RS_HLLD = 1
  integer(c_int), parameter :: CS_EMF_LLF = 0, CS_EMF_HLL = 1
  integer(c_int), parameter :: CS_FP64 = 0, CS_MIXED = 1
  integer(c_int), parameter :: CS_RBC_PAD = 0, CS_RBC_WALL = 1, CS_RBC_PROBLEM = 2

  type, bind(C) :: cs_config
    integer(c_int32_t) :: ntp, nr, nghost, recon, riemann, emf, precision, rk_order, radial_bc, radial_mesh
    real(c_double) :: rmin, rmax, gamma, rhofloor, prsfloor, cfl, grav
    integer(c_int32_t) :: imex, bc_problem
    real(c_double) :: omega(3)
  end type

  ! cs_p1_patch: the arguments of LIMIT_ALL_SLOPES_P1 (410limit_all_slopes_p1.f:9-60)
  ! for one patch, as device addresses (c_loc of the arrays inside an
  ! "!$acc host_data use_device(...)" region, or of HIP allocations).
  type, bind(C) :: cs_p1_patch
    integer(c_int32_t) :: ntp, nr, nghost, itri(4), pad
    real(c_double) :: gamma, molewt, rhofloor, prsfloor, smallnum, flatten_coef, prs_flatten_coef
    type(c_ptr) :: vert, mask, stencil, xf_normal, xf_area, xf_centroid, yf_normal, yf_area, yf_centroid
    type(c_ptr) :: zf_area, zf_centroid, xb_pts, xb_wts, xb_normal, yb_pts, yb_wts, yb_normal
    type(c_ptr) :: zb_pts, zb_wts, zb_normal, scale, v_centroid, U, ent, Bxf, Byf, Bzf
    type(c_ptr) :: V, B1_zc, B2_zc, B3_zc, Ex_zc, Ey_zc, Ez_zc
    type(c_ptr) :: gamma_eos, gamma_sound, mean_mol_wt, divvel, msonic, flatten, work
    type(c_ptr) :: Bxf_q, Byf_q, Bzf_q, V_min_local, V_max_local
  end type

  ! cs_p12_patch: a cs_p1_patch whose U, V, B1..3_zc and Ex..Ez_zc hold ten
  ! modes (NUMORDER 3), plus the device address of Coef_Basis_Moments (zone, 10)
  type, bind(C) :: cs_p12_patch
    type(cs_p1_patch) :: p1
    type(c_ptr) :: moments
  end type
  ! cs_volume_p12: host addresses of v_quadpts (zone, 3, 27), v_quadwt (zone,
  ! 27), v_volume, v_centroid (zone, 3), Coef_Basis_Moments (zone, 10), six sectors
  type, bind(C) :: cs_volume_p12
    type(c_ptr) :: v_pts, v_wts, v_vol, v_cen, moments
  end type

  ! cs_p1_flux: the arguments of MAKE_FACIAL_FLUXES (450make_facial_fluxes.f:5-60)
  type, bind(C) :: cs_p1_flux
    integer(c_int32_t) :: ntp, nr, nghost, pad
    real(c_double) :: gamma, rhofloor, prsfloor, smallnum
    type(c_ptr) :: vert, mask, stencil, xf_pts, xf_normal, xf_area, xf_centroid
    type(c_ptr) :: yf_pts, yf_normal, yf_area, yf_centroid, zf_pts, zf_normal, zf_area, zf_centroid
    type(c_ptr) :: scale, v_volume, v_centroid, V, ent, flatten
    type(c_ptr) :: S_L, S_R, S_D, S_U, S_B, S_T
    type(c_ptr) :: Fluct_x_Pl, Fluct_y_Pl, Fluct_z_Pl, Fluct_x_Mi, Fluct_y_Mi, Fluct_z_Mi
    type(c_ptr) :: dU_dt, EntFlux_x, EntFlux_y, EntFlux_z, Efx, Efy, Efz
  end type

  ! cs_p1_edge: the arguments MAKE_EDGE_EFIELDS_P1 (460make_edge_fields_p1.f:5-20)
  ! reads at second order
  type, bind(C) :: cs_p1_edge
    integer(c_int32_t) :: ntp, nr, nghost, pad
    integer(c_int32_t) :: itri(4)
    type(c_ptr) :: vert, mask, Ex_zc, Ey_zc, Ez_zc, Bxf_q, Byf_q, Bzf_q
    type(c_ptr) :: S_L, S_R, S_D, S_U, S_B, S_T, Efx, Efy, Efz
  end type

  ! cs_p1_source: the arguments of BUILD_NON_STIFF_SOURCES (470build_non_stiff_sources.f:6-7)
  type, bind(C) :: cs_p1_source
    integer(c_int32_t) :: ntp, nr, nghost, pad
    real(c_double) :: newtonsc, mstar, smallnum, omega(3)
    type(c_ptr) :: mask, v_centroid, V, Source, dU_dt
  end type

  ! cs_p1_pad: the arguments of PAD_BOUNDARY_R that it reads or writes
  ! (826pad_boundary_r.f:1-16); bcarr_r is (1-g:N+g, 1-g:N+g, 2)
  type, bind(C) :: cs_p1_pad
    integer(c_int32_t) :: ntp, nr, nghost, pad
    type(c_ptr) :: mask, bcarr_r, xf_area, yf_area, zf_area, v_volume, U, Bxf, Byf, Bzf
  end type

  ! cs_p1_shell_sync: the six patches' face and edge arrays and the pairings
  ! (device arrays of cs_p1_face_pair / cs_p1_edge_group records)
  type, bind(C) :: cs_p1_patch_fluxes
    type(c_ptr) :: Fluct_x_Pl, Fluct_y_Pl, Fluct_z_Pl, Fluct_x_Mi, Fluct_y_Mi, Fluct_z_Mi
    type(c_ptr) :: EntFlux_x, EntFlux_y, EntFlux_z, Efx, Efy, Efz, dU_dt
    type(c_ptr) :: xf_area, yf_area, zf_area, v_volume
  end type
  type, bind(C) :: cs_p1_shell_sync
    integer(c_int32_t) :: ntp, nr, nghost, pad
    type(cs_p1_patch_fluxes) :: patch(6)
    type(c_ptr) :: faces
    integer(c_int64_t) :: nface
    type(c_ptr) :: edges
    integer(c_int64_t) :: nedge
  end type

  ! cs_p1_update: the arguments of UPDATE_IMEX_RK_STEP_P1 (484update_imex_rk_step_p1.f:7-26)
  type, bind(C) :: cs_p1_update
    integer(c_int32_t) :: ntp, nr, nghost, pad
    real(c_double) :: dtmove, gamma, rhofloor, prsfloor
    type(c_ptr) :: vert, xf_area, yf_area, zf_area, v_volume
    type(c_ptr) :: U, ent, Bxf, Byf, Bzf, dU_dt
    type(c_ptr) :: EntFlux_x, EntFlux_y, EntFlux_z, Efx, Efy, Efz, flag_entprs
  end type

  ! cs_dump_meta: riemann.in's time, dtcur, n_cc_components and dump number
  ! (DRV:1527-1547, 1798-1802), the step count, the format version and the
  ! configuration of a checkpoint
  type, bind(C) :: cs_dump_meta
    real(c_double) :: time, dtcur
    integer(c_int64_t) :: filenum, step
    integer(c_int32_t) :: n_cc_components, version
    type(cs_config) :: cfg
  end type

  interface
    ! checkpoint / restart in place of riemann.in + WRITEDATA / READDATA
    ! (DRV:1564-1605, 1845-1930); path: trim(name)//c_null_char
    integer(c_int) function cs_write_dump(h, path, meta) bind(C, name="cs_write_dump")
      import :: c_int, c_ptr, c_char, cs_dump_meta
      type(c_ptr), value :: h
      character(kind=c_char), intent(in) :: path(*)
      type(cs_dump_meta), intent(in) :: meta
    end function
    integer(c_int) function cs_read_dump(h, path, meta) bind(C, name="cs_read_dump")
      import :: c_int, c_ptr, c_char, cs_dump_meta
      type(c_ptr), value :: h
      character(kind=c_char), intent(in) :: path(*)
      type(cs_dump_meta), intent(out) :: meta
    end function
    integer(c_int) function cs_dump_info(path, meta) bind(C, name="cs_dump_info")
      import :: c_int, c_char, cs_dump_meta
      character(kind=c_char), intent(in) :: path(*)
      type(cs_dump_meta), intent(out) :: meta
    end function
    ! INIT_PATCH_CUBE_SPHERE (820...f:303) and ADD_BOUNDARY (825...f:320) with
    ! PROBLEM_GENERATOR_U_A on a CS_RECON_P1ADP handle's device state
    integer(c_int) function cs_init_problem_p1(h, problem, seed, time) bind(C, name="cs_init_problem_p1")
      import :: c_int, c_ptr, c_int32_t, c_int64_t, c_double
      type(c_ptr), value :: h
      integer(c_int32_t), value :: problem
      integer(c_int64_t), value :: seed
      real(c_double), value :: time
    end function
    integer(c_int) function cs_add_boundary_p1(h, ibdy_inner, ibdy_outer, problem, seed, time) &
        bind(C, name="cs_add_boundary_p1")
      import :: c_int, c_ptr, c_int32_t, c_int64_t, c_double
      type(c_ptr), value :: h
      integer(c_int32_t), value :: ibdy_inner, ibdy_outer, problem
      integer(c_int64_t), value :: seed
      real(c_double), value :: time
    end function
    integer(c_int) function cs_sync_shared_p1(p, stream) bind(C, name="cs_sync_shared_p1")
      import :: c_int, c_ptr, cs_p1_shell_sync
      type(cs_p1_shell_sync), intent(in) :: p
      type(c_ptr), value :: stream
    end function
    ! U, ent, Bxf, Byf, Bzf: arrays of six device addresses (c_loc under host_data)
    integer(c_int) function cs_halo_exchange_p1(h, U, ent, Bxf, Byf, Bzf, stream) &
        bind(C, name="cs_halo_exchange_p1")
      import :: c_int, c_ptr
      type(c_ptr), value :: h, stream
      type(c_ptr), intent(in) :: U(6), ent(6), Bxf(6), Byf(6), Bzf(6)
    end function
    integer(c_int) function cs_build_non_stiff_sources_p1(p, stream) bind(C, name="cs_build_non_stiff_sources_p1")
      import :: c_int, c_ptr, cs_p1_source
      type(cs_p1_source), intent(in) :: p
      type(c_ptr), value :: stream
    end function
    integer(c_int) function cs_pad_boundary_r_p1(p, stream) bind(C, name="cs_pad_boundary_r_p1")
      import :: c_int, c_ptr, cs_p1_pad
      type(cs_p1_pad), intent(in) :: p
      type(c_ptr), value :: stream
    end function
    integer(c_int) function cs_make_edge_efields_p1(p, stream) bind(C, name="cs_make_edge_efields_p1")
      import :: c_int, c_ptr, cs_p1_edge
      type(cs_p1_edge), intent(in) :: p
      type(c_ptr), value :: stream
    end function
    integer(c_int) function cs_update_imex_rk_step_p1(p, stream) bind(C, name="cs_update_imex_rk_step_p1")
      import :: c_int, c_ptr, cs_p1_update
      type(cs_p1_update), intent(in) :: p
      type(c_ptr), value :: stream
    end function
    integer(c_int) function cs_limit_all_slopes_p12(p, stream) bind(C, name="cs_limit_all_slopes_p12")
      import :: c_int, c_ptr, cs_p12_patch
      type(cs_p12_patch), intent(in) :: p
      type(c_ptr), value :: stream
    end function
    integer(c_int) function cs_build_volume_p12(cfg, out) bind(C, name="cs_build_volume_p12")
      import :: c_int, cs_config, cs_volume_p12
      type(cs_config), intent(in) :: cfg
      type(cs_volume_p12), intent(in) :: out
    end function
    integer(c_int) function cs_cleanup_after_update_p1(p, stream) bind(C, name="cs_cleanup_after_update_p1")
      import :: c_int, c_ptr, cs_p1_patch
      type(cs_p1_patch), intent(in) :: p
      type(c_ptr), value :: stream
    end function
    ! dt_new: the device address of one double (c_loc of a device array under
    ! host_data), written on the stream
    integer(c_int) function cs_evaluate_timestep_p1(p, cfl, dtcur, dt_new, stream) &
        bind(C, name="cs_evaluate_timestep_p1")
      import :: c_int, c_ptr, c_double, cs_p1_patch
      type(cs_p1_patch), intent(in) :: p
      real(c_double), value :: cfl, dtcur
      type(c_ptr), value :: dt_new, stream
    end function
    integer(c_int) function cs_make_facial_fluxes_p1(p, stream) bind(C, name="cs_make_facial_fluxes_p1")
      import :: c_int, c_ptr, cs_p1_flux
      type(cs_p1_flux), intent(in) :: p
      type(c_ptr), value :: stream
    end function
    integer(c_int) function cs_limit_all_slopes_p1(p, stream) bind(C, name="cs_limit_all_slopes_p1")
      import :: c_int, c_ptr, cs_p1_patch
      type(cs_p1_patch), intent(in) :: p
      type(c_ptr), value :: stream
    end function
    integer(c_int) function cs_create(cfg, h) bind(C, name="cs_create")
      import :: c_int, c_ptr, cs_config
      type(cs_config), intent(in) :: cfg
      type(c_ptr), intent(out) :: h
    end function
    integer(c_int) function cs_destroy(h) bind(C, name="cs_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function cs_init_problem(cfg, problem, seed, u, ent, bxf, byf, bzf) &
        bind(C, name="cs_init_problem")
      import :: c_int, c_int32_t, c_int64_t, c_double, cs_config
      type(cs_config), intent(in) :: cfg
      integer(c_int32_t), value :: problem
      integer(c_int64_t), value :: seed
      real(c_double), intent(out) :: u(*), ent(*), bxf(*), byf(*), bzf(*)
    end function
    integer(c_int) function cs_upload_state(h, u, ent, bxf, byf, bzf) bind(C, name="cs_upload_state")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: u(*), ent(*), bxf(*), byf(*), bzf(*)
    end function
    integer(c_int) function cs_download_state(h, u, ent, bxf, byf, bzf) bind(C, name="cs_download_state")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: u(*), ent(*), bxf(*), byf(*), bzf(*)
    end function
    integer(c_int) function cs_halo_exchange(h) bind(C, name="cs_halo_exchange")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function cs_make_facial_fluxes(h) bind(C, name="cs_make_facial_fluxes")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function cs_make_edge_efields(h) bind(C, name="cs_make_edge_efields")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function cs_update_stage(h, irk) bind(C, name="cs_update_stage")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: h
      integer(c_int32_t), value :: irk
    end function
    integer(c_int) function cs_evaluate_timestep(h) bind(C, name="cs_evaluate_timestep")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function cs_set_dt(h, dt) bind(C, name="cs_set_dt")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), value :: dt
    end function
    integer(c_int) function cs_get_dt(h, dt) bind(C, name="cs_get_dt")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: dt
    end function
    integer(c_int) function cs_rk_stage(h, irk) bind(C, name="cs_rk_stage")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: h
      integer(c_int32_t), value :: irk
    end function
    integer(c_int) function cs_advance(h, nsteps, use_graph) bind(C, name="cs_advance")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: h
      integer(c_int32_t), value :: nsteps, use_graph
    end function
    integer(c_int) function cs_conservation_sums(h, sums) bind(C, name="cs_conservation_sums")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: sums(9)
    end function
    integer(c_int) function cs_max_divb(h, v) bind(C, name="cs_max_divb")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: v
    end function
    integer(c_int) function cs_last_error(buf, n) bind(C, name="cs_last_error")
      import :: c_int, c_char, c_size_t
      character(kind=c_char), intent(out) :: buf(*)
      integer(c_size_t), value :: n
    end function
  end interface

contains

  ! Stop the way the reference stops (WRITE(6,*) + STOP) on a non-zero status.
  subroutine cs_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: what
    character(kind=c_char) :: buf(512)
    integer :: i, n
    if (rc == 0) return
    i = cs_last_error(buf, int(512, c_size_t))
    n = 0
    do while (n < 511)
      if (buf(n + 1) == c_null_char) exit
      n = n + 1
    end do
    write (6, *) "Stopping in ", what, " status ", rc, ": ", (buf(i), i = 1, n)
    stop 1
  end subroutine

end module cs_hip_iface
