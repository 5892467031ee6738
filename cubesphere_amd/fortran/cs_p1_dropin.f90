! cs_p1_dropin.f90 -- the reference driver's per-patch P1 stage made from
! Fortran through the bind(C) structures of cs_hip_iface, the way the driver
! calls the routines (DRV:2396-2887): one patch's dummy arrays, in the
! reference's layouts, are read from raw files in a directory (one file per
! array, named for it), put in device memory (the driver's "!$acc host_data
! use_device" addresses; here hipMalloc), and the seven routines run in the
! driver's order on the default stream:
!   LIMIT_ALL_SLOPES_P1 (DRV:2396)  -> cs_limit_all_slopes_p1    (cs_p1_patch)
!   MAKE_FACIAL_FLUXES (DRV:2485)   -> cs_make_facial_fluxes_p1  (cs_p1_flux)
!   MAKE_EDGE_EFIELDS_P1 (DRV:2548) -> cs_make_edge_efields_p1   (cs_p1_edge)
!   BUILD_NON_STIFF_SOURCES (DRV:2603) -> cs_build_non_stiff_sources_p1 (cs_p1_source)
!   UPDATE_IMEX_RK_STEP_P1 (DRV:2782)  -> cs_update_imex_rk_step_p1 (cs_p1_update)
!   CLEANUP_AFTER_UPDATE (DRV:2815) -> cs_cleanup_after_update_p1 (cs_p1_patch)
!   EVALUATE_TIMESTEP (DRV:2887)    -> cs_evaluate_timestep_p1   (cs_p1_patch)
! Every file's size is checked against its layout before anything reaches
! the device.  The arrays the stage writes are stored as out_<name>.bin
! beside the inputs, the new time step as out_dt.bin and the reconstruction's
! V_min_local, V_max_local as out_vminmax.bin.
!
! Usage: cs_p1_dropin <dir> [p12], where <dir>/scalars.bin holds int32 ntp,
! nr, nghost, 0 and the float64 gamma, molewt, rhofloor, prsfloor, smallnum,
! flatten_coef, prs_flatten_coef, dtmove, newtonsc, mstar, omega(3), cfl,
! dtcur.  With p12 the driver's third-order reconstruction runs instead
! (DRV:2398): U, V, B1..3_zc and Ex..Ez_zc hold ten modes, moments.bin the
! basis moments (zone, 10), and LIMIT_ALL_SLOPES_P12 -> cs_limit_all_slopes_p12
! (cs_p12_patch) is the only call; the arrays it writes are stored as
! out_<name>.bin and its V_min_local, V_max_local as out_vminmax.bin.
program cs_p1_dropin
  use iso_c_binding
  use cs_hip_iface
  implicit none
  interface
    integer(c_int) function hipMalloc(p, n) bind(C, name="hipMalloc")
      import :: c_int, c_ptr, c_size_t
      type(c_ptr) :: p
      integer(c_size_t), value :: n
    end function
    integer(c_int) function hipFree(p) bind(C, name="hipFree")
      import :: c_int, c_ptr
      type(c_ptr), value :: p
    end function
    integer(c_int) function hipMemcpy(dst, src, n, kind) bind(C, name="hipMemcpy")
      import :: c_int, c_ptr, c_size_t
      type(c_ptr), value :: dst, src
      integer(c_size_t), value :: n
      integer(c_int), value :: kind
    end function
    integer(c_int) function hipDeviceSynchronize() bind(C, name="hipDeviceSynchronize")
      import :: c_int
    end function
  end interface
  integer(c_int), parameter :: H2D = 1, D2H = 2
  integer, parameter :: nn = 69
  character(len=12), parameter :: names(nn) = [character(len=12) :: &
      "vert", "mask", "stencil", "xf_normal", "xf_area", "xf_centroid", "yf_normal", "yf_area", &
      "yf_centroid", "zf_area", "zf_centroid", "xb_pts", "xb_wts", "xb_normal", "yb_pts", "yb_wts", &
      "yb_normal", "zb_pts", "zb_wts", "zb_normal", "scale", "v_centroid", "U", "ent", "Bxf", "Byf", "Bzf", &
      "V", "B1_zc", "B2_zc", "B3_zc", "Ex_zc", "Ey_zc", "Ez_zc", "gamma_eos", "gamma_sound", "mean_mol_wt", &
      "divvel", "msonic", "flatten", "work", "Bxf_q", "Byf_q", "Bzf_q", "xf_pts", "yf_pts", "zf_pts", &
      "zf_normal", "v_volume", "S_L", "S_R", "S_D", "S_U", "S_B", "S_T", "Fluct_x_Pl", "Fluct_y_Pl", &
      "Fluct_z_Pl", "Fluct_x_Mi", "Fluct_y_Mi", "Fluct_z_Mi", "dU_dt", "EntFlux_x", "EntFlux_y", "EntFlux_z", &
      "Efx", "Efy", "Efz", "Source"]
  integer, parameter :: nout = 13
  character(len=12), parameter :: outs(nout) = [character(len=12) :: &
      "U", "ent", "V", "Bxf", "Byf", "Bzf", "Efx", "Efy", "Efz", "Source", "dU_dt", "flag_entprs", "msonic"]
  integer, parameter :: nout12 = 14
  character(len=12), parameter :: outs12(nout12) = [character(len=12) :: &
      "U", "ent", "V", "B1_zc", "B2_zc", "B3_zc", "Ex_zc", "Ey_zc", "Ez_zc", "Bxf_q", "Byf_q", "Bzf_q", &
      "flatten", "work"]
  character(len=4096) :: dir
  character(len=8) :: mode
  logical :: p12
  integer(c_int64_t) :: nmode
  type(c_ptr) :: dev(nn + 1), dt_dev, mom_dev
  type(cs_p12_patch) :: pq
  integer(c_size_t) :: nbytes(nn + 1)
  integer(c_int32_t) :: ihdr(4)
  real(c_double) :: rhdr(15)
  real(c_double), target :: vmin(9), vmax(9), dt_new(1)
  integer(c_int64_t) :: zone, xf, zf, fc, nzz
  integer :: ntp, nr, g, i, u
  type(cs_p1_patch) :: p
  type(cs_p1_flux) :: f
  type(cs_p1_edge) :: e
  type(cs_p1_source) :: s
  type(cs_p1_update) :: up

  call get_command_argument(1, dir)
  if (len_trim(dir) == 0) then
    write (6, *) "usage: cs_p1_dropin <dir> [p12]"
    stop 2
  end if
  call get_command_argument(2, mode)
  p12 = trim(mode) == "p12"
  nmode = 4
  if (p12) nmode = 10
  open (newunit=u, file=trim(dir)//"/scalars.bin", access="stream", form="unformatted", status="old", action="read")
  read (u) ihdr, rhdr
  close (u)
  ntp = ihdr(1); nr = ihdr(2); g = ihdr(3)
  nzz = ntp + 2 * g
  zone = nzz * nzz * (nr + 2 * g)
  xf = (nzz + 1) * nzz * (nr + 2 * g)
  zf = nzz * nzz * (nr + 2 * g + 1)
  fc = (nzz + 1) * (nzz + 1) * (nr + 2 * g + 1)

  ! every array: its size from the layout checked against its file, then the
  ! files into device memory
  do i = 1, nn
    nbytes(i) = bytes_of(names(i))
    call check_size(i)
  end do
  do i = 1, nn
    call load(i)
  end do
  ! flag_entprs (one per zone) is written by the update and read by nothing
  ! before it: allocated and cleared here as the driver's array is
  nbytes(nn + 1) = 8_c_size_t * zone
  call hip_check(hipMalloc(dev(nn + 1), nbytes(nn + 1)), "hipMalloc")
  call zero_dev(dev(nn + 1), nbytes(nn + 1))
  call hip_check(hipMalloc(dt_dev, 8_c_size_t), "hipMalloc")
  mom_dev = c_null_ptr
  if (p12) call load_moments()

  ! LIMIT_ALL_SLOPES_P1 (410limit_all_slopes_p1.f:9-60), CLEANUP_AFTER_UPDATE and EVALUATE_TIMESTEP
  p%ntp = ntp; p%nr = nr; p%nghost = g; p%itri = 1; p%pad = 0
  p%gamma = rhdr(1); p%molewt = rhdr(2); p%rhofloor = rhdr(3); p%prsfloor = rhdr(4)
  p%smallnum = rhdr(5); p%flatten_coef = rhdr(6); p%prs_flatten_coef = rhdr(7)
  p%vert = d("vert"); p%mask = d("mask"); p%stencil = d("stencil")
  p%xf_normal = d("xf_normal"); p%xf_area = d("xf_area"); p%xf_centroid = d("xf_centroid")
  p%yf_normal = d("yf_normal"); p%yf_area = d("yf_area"); p%yf_centroid = d("yf_centroid")
  p%zf_area = d("zf_area"); p%zf_centroid = d("zf_centroid")
  p%xb_pts = d("xb_pts"); p%xb_wts = d("xb_wts"); p%xb_normal = d("xb_normal")
  p%yb_pts = d("yb_pts"); p%yb_wts = d("yb_wts"); p%yb_normal = d("yb_normal")
  p%zb_pts = d("zb_pts"); p%zb_wts = d("zb_wts"); p%zb_normal = d("zb_normal")
  p%scale = d("scale"); p%v_centroid = d("v_centroid")
  p%U = d("U"); p%ent = d("ent"); p%Bxf = d("Bxf"); p%Byf = d("Byf"); p%Bzf = d("Bzf"); p%V = d("V")
  p%B1_zc = d("B1_zc"); p%B2_zc = d("B2_zc"); p%B3_zc = d("B3_zc")
  p%Ex_zc = d("Ex_zc"); p%Ey_zc = d("Ey_zc"); p%Ez_zc = d("Ez_zc")
  p%gamma_eos = d("gamma_eos"); p%gamma_sound = d("gamma_sound"); p%mean_mol_wt = d("mean_mol_wt")
  p%divvel = d("divvel"); p%msonic = d("msonic"); p%flatten = d("flatten"); p%work = d("work")
  p%Bxf_q = d("Bxf_q"); p%Byf_q = d("Byf_q"); p%Bzf_q = d("Bzf_q")
  p%V_min_local = c_loc(vmin); p%V_max_local = c_loc(vmax)

  ! MAKE_FACIAL_FLUXES (450make_facial_fluxes.f:5-60)
  f%ntp = ntp; f%nr = nr; f%nghost = g; f%pad = 0
  f%gamma = rhdr(1); f%rhofloor = rhdr(3); f%prsfloor = rhdr(4); f%smallnum = rhdr(5)
  f%vert = d("vert"); f%mask = d("mask"); f%stencil = d("stencil")
  f%xf_pts = d("xf_pts"); f%xf_normal = d("xf_normal"); f%xf_area = d("xf_area"); f%xf_centroid = d("xf_centroid")
  f%yf_pts = d("yf_pts"); f%yf_normal = d("yf_normal"); f%yf_area = d("yf_area"); f%yf_centroid = d("yf_centroid")
  f%zf_pts = d("zf_pts"); f%zf_normal = d("zf_normal"); f%zf_area = d("zf_area"); f%zf_centroid = d("zf_centroid")
  f%scale = d("scale"); f%v_volume = d("v_volume"); f%v_centroid = d("v_centroid")
  f%V = d("V"); f%ent = d("ent"); f%flatten = d("flatten")
  f%S_L = d("S_L"); f%S_R = d("S_R"); f%S_D = d("S_D"); f%S_U = d("S_U"); f%S_B = d("S_B"); f%S_T = d("S_T")
  f%Fluct_x_Pl = d("Fluct_x_Pl"); f%Fluct_y_Pl = d("Fluct_y_Pl"); f%Fluct_z_Pl = d("Fluct_z_Pl")
  f%Fluct_x_Mi = d("Fluct_x_Mi"); f%Fluct_y_Mi = d("Fluct_y_Mi"); f%Fluct_z_Mi = d("Fluct_z_Mi")
  f%dU_dt = d("dU_dt"); f%EntFlux_x = d("EntFlux_x"); f%EntFlux_y = d("EntFlux_y"); f%EntFlux_z = d("EntFlux_z")
  f%Efx = d("Efx"); f%Efy = d("Efy"); f%Efz = d("Efz")

  ! MAKE_EDGE_EFIELDS_P1 (460make_edge_fields_p1.f:5-20)
  e%ntp = ntp; e%nr = nr; e%nghost = g; e%pad = 0; e%itri = 1
  e%vert = d("vert"); e%mask = d("mask")
  e%Ex_zc = d("Ex_zc"); e%Ey_zc = d("Ey_zc"); e%Ez_zc = d("Ez_zc")
  e%Bxf_q = d("Bxf_q"); e%Byf_q = d("Byf_q"); e%Bzf_q = d("Bzf_q")
  e%S_L = d("S_L"); e%S_R = d("S_R"); e%S_D = d("S_D"); e%S_U = d("S_U"); e%S_B = d("S_B"); e%S_T = d("S_T")
  e%Efx = d("Efx"); e%Efy = d("Efy"); e%Efz = d("Efz")

  ! BUILD_NON_STIFF_SOURCES (470build_non_stiff_sources.f:6-7)
  s%ntp = ntp; s%nr = nr; s%nghost = g; s%pad = 0
  s%newtonsc = rhdr(9); s%mstar = rhdr(10); s%smallnum = rhdr(5); s%omega = rhdr(11:13)
  s%mask = d("mask"); s%v_centroid = d("v_centroid"); s%V = d("V"); s%Source = d("Source"); s%dU_dt = d("dU_dt")

  ! UPDATE_IMEX_RK_STEP_P1 (484update_imex_rk_step_p1.f:7-26)
  up%ntp = ntp; up%nr = nr; up%nghost = g; up%pad = 0
  up%dtmove = rhdr(8); up%gamma = rhdr(1); up%rhofloor = rhdr(3); up%prsfloor = rhdr(4)
  up%vert = d("vert"); up%xf_area = d("xf_area"); up%yf_area = d("yf_area"); up%zf_area = d("zf_area")
  up%v_volume = d("v_volume")
  up%U = d("U"); up%ent = d("ent"); up%Bxf = d("Bxf"); up%Byf = d("Byf"); up%Bzf = d("Bzf"); up%dU_dt = d("dU_dt")
  up%EntFlux_x = d("EntFlux_x"); up%EntFlux_y = d("EntFlux_y"); up%EntFlux_z = d("EntFlux_z")
  up%Efx = d("Efx"); up%Efy = d("Efy"); up%Efz = d("Efz"); up%flag_entprs = dev(nn + 1)

  if (p12) then
    ! LIMIT_ALL_SLOPES_P12 (421limit_all_slopes_p12.f:9), the patch's own
    ! arrays of ten modes and the basis moments
    pq%p1 = p
    pq%moments = mom_dev
    call cs_check(cs_limit_all_slopes_p12(pq, c_null_ptr), "LIMIT_ALL_SLOPES_P12")
    call hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize")
    do i = 1, nout12
      call store(outs12(i))
    end do
    open (newunit=u, file=trim(dir)//"/out_vminmax.bin", access="stream", form="unformatted", status="replace")
    write (u) vmin, vmax
    close (u)
    do i = 1, nn + 1
      call hip_check(hipFree(dev(i)), "hipFree")
    end do
    call hip_check(hipFree(dt_dev), "hipFree")
    call hip_check(hipFree(mom_dev), "hipFree")
    write (6, '(a, i0, a, i0, a, i0)') "p12 dropin: ntp=", ntp, " nr=", nr, " nghost=", g
    stop
  end if
  call cs_check(cs_limit_all_slopes_p1(p, c_null_ptr), "LIMIT_ALL_SLOPES_P1")
  call cs_check(cs_make_facial_fluxes_p1(f, c_null_ptr), "MAKE_FACIAL_FLUXES")
  call cs_check(cs_make_edge_efields_p1(e, c_null_ptr), "MAKE_EDGE_EFIELDS_P1")
  call cs_check(cs_build_non_stiff_sources_p1(s, c_null_ptr), "BUILD_NON_STIFF_SOURCES")
  call cs_check(cs_update_imex_rk_step_p1(up, c_null_ptr), "UPDATE_IMEX_RK_STEP_P1")
  call cs_check(cs_cleanup_after_update_p1(p, c_null_ptr), "CLEANUP_AFTER_UPDATE")
  call cs_check(cs_evaluate_timestep_p1(p, rhdr(14), rhdr(15), dt_dev, c_null_ptr), "EVALUATE_TIMESTEP")
  call hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize")

  do i = 1, nout
    call store(outs(i))
  end do
  call hip_check(hipMemcpy(c_loc(dt_new), dt_dev, 8_c_size_t, D2H), "hipMemcpy")
  open (newunit=u, file=trim(dir)//"/out_dt.bin", access="stream", form="unformatted", status="replace")
  write (u) dt_new
  close (u)
  open (newunit=u, file=trim(dir)//"/out_vminmax.bin", access="stream", form="unformatted", status="replace")
  write (u) vmin, vmax
  close (u)
  do i = 1, nn + 1
    call hip_check(hipFree(dev(i)), "hipFree")
  end do
  call hip_check(hipFree(dt_dev), "hipFree")
  write (6, '(a, i0, a, i0, a, i0, a, es23.15)') "p1 dropin: ntp=", ntp, " nr=", nr, " nghost=", g, " dt=", dt_new(1)

contains

  ! bytes of a dummy array of the patch, from its declaration in the reference
  integer(c_size_t) function bytes_of(name)
    character(*), intent(in) :: name
    integer(c_int64_t) :: n
    select case (trim(name))
    case ("vert"); n = 3 * fc
    case ("mask", "scale", "gamma_eos", "gamma_sound", "mean_mol_wt", "divvel", "msonic", "flatten", "work", &
          "v_volume"); n = zone
    case ("stencil"); n = nzz * nzz * 24
    case ("xf_normal", "xf_centroid", "yf_normal", "yf_centroid", "xb_normal", "yb_normal"); n = 3 * xf
    case ("xf_area", "yf_area"); n = xf
    case ("zf_area"); n = zf
    case ("zf_centroid"); n = 3 * zf
    case ("xb_pts", "yb_pts", "xf_pts", "yf_pts"); n = 15 * xf
    case ("xb_wts", "yb_wts"); n = 5 * xf
    case ("zb_pts", "zb_normal", "zf_pts", "zf_normal"); n = 15 * zf
    case ("zb_wts"); n = 5 * zf
    case ("v_centroid"); n = 3 * zone
    case ("U", "V"); n = 8 * nmode * zone
    case ("ent"); n = 4 * zone
    case ("B1_zc", "B2_zc", "B3_zc", "Ex_zc", "Ey_zc", "Ez_zc"); n = nmode * zone
    case ("Bxf", "Byf", "Bzf", "S_L", "S_R", "S_D", "S_U", "S_B", "S_T", "EntFlux_x", "EntFlux_y", "EntFlux_z", &
          "Efx", "Efy", "Efz"); n = fc
    case ("Bxf_q", "Byf_q", "Bzf_q"); n = 5 * fc
    case ("Fluct_x_Pl", "Fluct_y_Pl", "Fluct_z_Pl", "Fluct_x_Mi", "Fluct_y_Mi", "Fluct_z_Mi"); n = 8 * fc
    case ("dU_dt", "Source"); n = 8 * zone
    case default
      write (6, *) "no layout for ", trim(name)
      stop 2
    end select
    if (trim(name) == "mask" .or. trim(name) == "stencil") then
      bytes_of = 4_c_size_t * n
    else
      bytes_of = 8_c_size_t * n
    end if
  end function

  type(c_ptr) function d(name)
    character(*), intent(in) :: name
    integer :: k
    do k = 1, nn
      if (trim(names(k)) == name) then
        d = dev(k)
        return
      end if
    end do
    write (6, *) "no array ", name
    stop 2
  end function

  subroutine check_size(k)
    integer, intent(in) :: k
    integer :: v, ios
    integer(c_int64_t) :: sz
    open (newunit=v, file=trim(dir)//"/"//trim(names(k))//".bin", access="stream", form="unformatted", &
          status="old", action="read", iostat=ios)
    if (ios /= 0) then
      write (6, *) "cannot open ", trim(names(k)), ".bin"
      stop 2
    end if
    inquire (unit=v, size=sz)
    close (v)
    if (sz /= nbytes(k)) then
      write (6, *) trim(names(k)), ".bin holds ", sz, " bytes, its layout ", nbytes(k)
      stop 2
    end if
  end subroutine

  subroutine load(k)
    integer, intent(in) :: k
    integer :: v
    integer(c_int8_t), allocatable, target :: buf(:)
    open (newunit=v, file=trim(dir)//"/"//trim(names(k))//".bin", access="stream", form="unformatted", &
          status="old", action="read")
    allocate (buf(nbytes(k)))
    read (v) buf
    close (v)
    call hip_check(hipMalloc(dev(k), nbytes(k)), "hipMalloc")
    call hip_check(hipMemcpy(dev(k), c_loc(buf), nbytes(k), H2D), "hipMemcpy")
  end subroutine

  ! the third order's basis moments (zone, 10) from moments.bin
  subroutine load_moments()
    integer :: v, ios
    integer(c_int64_t) :: sz
    integer(c_size_t) :: nb
    real(c_double), allocatable, target :: buf(:)
    nb = 80_c_size_t * zone
    open (newunit=v, file=trim(dir)//"/moments.bin", access="stream", form="unformatted", status="old", &
          action="read", iostat=ios)
    if (ios /= 0) then
      write (6, *) "cannot open moments.bin"
      stop 2
    end if
    inquire (unit=v, size=sz)
    if (sz /= nb) then
      write (6, *) "moments.bin holds ", sz, " bytes, its layout ", nb
      stop 2
    end if
    allocate (buf(10 * zone))
    read (v) buf
    close (v)
    call hip_check(hipMalloc(mom_dev, nb), "hipMalloc")
    call hip_check(hipMemcpy(mom_dev, c_loc(buf), nb, H2D), "hipMemcpy")
  end subroutine
  subroutine zero_dev(ptr, n)
    type(c_ptr), intent(in) :: ptr
    integer(c_size_t), intent(in) :: n
    integer(c_int8_t), allocatable, target :: buf(:)
    allocate (buf(n))
    buf = 0
    call hip_check(hipMemcpy(ptr, c_loc(buf), n, H2D), "hipMemcpy")
  end subroutine

  subroutine store(name)
    character(*), intent(in) :: name
    integer :: k, v
    integer(c_int8_t), allocatable, target :: buf(:)
    k = 0
    if (trim(name) == "flag_entprs") then
      k = nn + 1
    else
      do k = 1, nn
        if (trim(names(k)) == trim(name)) exit
      end do
    end if
    allocate (buf(nbytes(k)))
    call hip_check(hipMemcpy(c_loc(buf), dev(k), nbytes(k), D2H), "hipMemcpy")
    open (newunit=v, file=trim(dir)//"/out_"//trim(name)//".bin", access="stream", form="unformatted", &
          status="replace")
    write (v) buf
    close (v)
  end subroutine

  subroutine hip_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: what
    if (rc /= 0) then
      write (6, *) what, " failed with HIP status ", rc
      stop 1
    end if
  end subroutine

end program cs_p1_dropin
