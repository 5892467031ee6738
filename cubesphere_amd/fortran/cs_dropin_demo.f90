! cs_dropin_demo.f90 -- the per-stage call sequence of the reference driver
! (DRV:1746-3120) made through the C-ABI from Fortran: driver-layout arrays,
! one SSP-RK3 step as exchange -> fluxes -> EMFs -> update for each stage,
! then the CFL step and the diagnostics.  Prints one line and stops with 0.
! With a file name as its argument it also writes the state as a checkpoint
! there (riemann.in + WRITEDATA's replacement, cs_write_dump), reads its header
! back (cs_dump_info) and prints a second line.
program cs_dropin_demo
  use iso_c_binding
  use cs_hip_iface
  implicit none
  integer, parameter :: n = 16, nr = 1, g = 3
  type(cs_config) :: cfg
  type(c_ptr) :: h
  real(c_double), allocatable :: u(:, :, :, :, :), ent(:, :, :, :), bxf(:, :, :, :), byf(:, :, :, :), bzf(:, :, :, :)
  real(c_double) :: dt, sums0(9), sums1(9), divb
  integer(c_int32_t) :: irk
  character(len=4096) :: dumpfile
  type(cs_dump_meta) :: meta, back

  cfg = cs_config(n, nr, g, CS_RECON_WENO5, CS_RS_HLLD, CS_EMF_LLF, CS_FP64, 3, CS_RBC_WALL, 0, &
                  1.0d0, 1.0d0 + 2.0d0 * atan(1.0d0) / n, 5.0d0 / 3.0d0, 1.0d-12, 1.0d-9, 0.4d0, 0.0d0, 0, 0, &
                  [0.0d0, 0.0d0, 0.0d0])
  allocate (u(1 - g:n + g, 1 - g:n + g, 1 - g:nr + g, 8, 6), ent(1 - g:n + g, 1 - g:n + g, 1 - g:nr + g, 6))
  allocate (bxf(0 - g:n + g, 0 - g:n + g, 0 - g:nr + g, 6), byf(0 - g:n + g, 0 - g:n + g, 0 - g:nr + g, 6), &
            bzf(0 - g:n + g, 0 - g:n + g, 0 - g:nr + g, 6))
  call cs_check(cs_init_problem(cfg, 2, 7_c_int64_t, u, ent, bxf, byf, bzf), "CS_INIT_PROBLEM")
  call cs_check(cs_create(cfg, h), "CS_CREATE")
  call cs_check(cs_upload_state(h, u, ent, bxf, byf, bzf), "CS_UPLOAD_STATE")
  call cs_check(cs_conservation_sums(h, sums0), "CS_CONSERVATION_SUMS")
  call cs_check(cs_set_dt(h, 0.0d0), "CS_SET_DT")
  call cs_check(cs_evaluate_timestep(h), "EVALUATE_TIMESTEP")
  do irk = 1, 3
    call cs_check(cs_rk_stage(h, irk), "RK_STAGE")
  end do
  call cs_check(cs_evaluate_timestep(h), "EVALUATE_TIMESTEP")
  call cs_check(cs_get_dt(h, dt), "CS_GET_DT")
  call cs_check(cs_download_state(h, u, ent, bxf, byf, bzf), "CS_DOWNLOAD_STATE")
  call cs_check(cs_conservation_sums(h, sums1), "CS_CONSERVATION_SUMS")
  call cs_check(cs_max_divb(h, divb), "CS_MAX_DIVB")
  call get_command_argument(1, dumpfile)
  if (len_trim(dumpfile) > 0) then
    meta%time = 0.0d0
    meta%dtcur = dt
    meta%filenum = 1
    meta%step = 1
    call cs_check(cs_write_dump(h, trim(dumpfile)//c_null_char, meta), "CS_WRITE_DUMP")
    call cs_check(cs_dump_info(trim(dumpfile)//c_null_char, back), "CS_DUMP_INFO")
  end if
  call cs_check(cs_destroy(h), "CS_DESTROY")
  write (6, '(a, es12.4, a, es12.4, a, es10.2, a, es12.4)') "dropin: rho(1,1,1,panel 3) =", u(1, 1, 1, 1, 3), &
      " mass drift =", (sums1(1) - sums0(1)) / sums0(1), " divB =", divb, " dt =", dt
  if (len_trim(dumpfile) > 0) write (6, '(a, i0, a, i0, a, i0, a, es25.17)') "dump: version ", back%version, &
      " step ", back%step, " ntp ", back%cfg%ntp, " dtcur ", back%dtcur
end program cs_dropin_demo
