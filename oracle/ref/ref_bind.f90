! ref_bind.f90 -- C-callable wrappers around reference routines compiled from
! /root/reference/pkg/inc/cubesphere (see oracle/ref/Makefile).  Test
! infrastructure only: used by the CPU tests to pin the oracle's
! restatement; never linked into the product.
module cs_ref_bind
  use iso_c_binding
  implicit none

  ! mirror of or_p1_patch (oracle/cs_oracle.h) and cs_p1_patch (include/cubesphere_hip.h)
  type, bind(C) :: ref_p1_patch
    integer(c_int32_t) :: ntp, nr, nghost, itri(4), pad
    real(c_double) :: gamma, molewt, rhofloor, prsfloor, smallnum, flatten_coef, prs_flatten_coef
    type(c_ptr) :: vert, mask, stencil, xf_normal, xf_area, xf_centroid, yf_normal, yf_area, yf_centroid
    type(c_ptr) :: zf_area, zf_centroid, xb_pts, xb_wts, xb_normal, yb_pts, yb_wts, yb_normal
    type(c_ptr) :: zb_pts, zb_wts, zb_normal, scale, v_centroid, U, ent, Bxf, Byf, Bzf
    type(c_ptr) :: V, B1_zc, B2_zc, B3_zc, Ex_zc, Ey_zc, Ez_zc
    type(c_ptr) :: gamma_eos, gamma_sound, mean_mol_wt, divvel, msonic, flatten, work
    type(c_ptr) :: Bxf_q, Byf_q, Bzf_q, V_min_local, V_max_local
  end type

  ! mirror of or_p1_pad / cs_p1_pad
  type, bind(C) :: ref_p1_pad
    integer(c_int32_t) :: ntp, nr, nghost, pad
    type(c_ptr) :: mask, bcarr_r, xf_area, yf_area, zf_area, v_volume, U, Bxf, Byf, Bzf
  end type
  ! mirrors of or_p1_init and or_prob
  type, bind(C) :: ref_p1_init
    integer(c_int32_t) :: ntp, nr, nghost, sector
    type(c_ptr) :: vert, mask, v_pts, v_wts, xf_area, yf_area, zf_area, U, ent, Bxf, Byf, Bzf
  end type
  type, bind(C) :: ref_prob
    integer(c_int32_t) :: iproblem, pad
    real(c_double) :: time, rmin, rmax, gamma, rhofloor, prsfloor, grav
    integer(c_int64_t) :: seed
  end type
  ! mirror of or_p1_edge / cs_p1_edge
  type, bind(C) :: ref_p1_edge
    integer(c_int32_t) :: ntp, nr, nghost, pad
    integer(c_int32_t) :: itri(4)
    type(c_ptr) :: vert, mask, Ex_zc, Ey_zc, Ez_zc, Bxf_q, Byf_q, Bzf_q
    type(c_ptr) :: S_L, S_R, S_D, S_U, S_B, S_T, Efx, Efy, Efz
  end type
contains

  ! Mesh_Sector_General (201mesh_sector_general.f:79) for one patch per sector,
  ! then FILL_GHOST2D_GLOBAL (203fill_ghost2d_global.f:7) for all six.
  subroutine ref_unit_lattice(n, g, coords) bind(C, name="ref_unit_lattice")
    integer(c_int), value :: n, g
    real(c_double), intent(inout) :: coords(0-g:n+g, 0-g:n+g, 3, 6)
    real(c_double) :: vertices(4, 3)
    integer :: s
    external Mesh_Sector_General, FILL_GHOST2D_GLOBAL
    coords = 0.0d0
    do s = 1, 6
      call Mesh_Sector_General(s, 1, 1, 1, n, g, vertices, coords(:, :, :, s))
    end do
    do s = 1, 6
      call FILL_GHOST2D_GLOBAL(s, n, g, coords)
    end do
  end subroutine

  subroutine ref_radial_mesh(nr, g, rmin, rmax, ilog, rb) bind(C, name="ref_radial_mesh")
    integer(c_int), value :: nr, g, ilog
    real(c_double), value :: rmin, rmax
    real(c_double), intent(out) :: rb(0-g:nr+g)
    real(c_double) :: dr(1-g:nr+g), rc(1-g:nr+g), smallnum, ratio
    external GENERATE_GLOBAL_RADIAL_UNIFORM_MESH, GENERATE_GLOBAL_RADIAL_LOG_MESH
    smallnum = 1.0d-12
    ratio = 2.0d0
    if (ilog == 0) then
      call GENERATE_GLOBAL_RADIAL_UNIFORM_MESH(nr, g, rmin, rmax, smallnum, ratio, rb, dr, rc)
    else
      call GENERATE_GLOBAL_RADIAL_LOG_MESH(nr, g, rmin, rmax, smallnum, ratio, rb, dr, rc)
    end if
  end subroutine

  ! CONS_TO_PRIM / PRIM_TO_CONS / PRIM_TO_WAVSPD (662, 661, 660), MHD branch.
  subroutine ref_cons_to_prim(gam, rhof, prsf, ent, u, v, flag) bind(C, name="ref_cons_to_prim")
    real(c_double), value :: gam, rhof, prsf
    real(c_double), intent(inout) :: ent, u(8)
    real(c_double), intent(out) :: v(8)
    integer(c_int), intent(out) :: flag
    real(c_double) :: b0(3)
    integer :: iflag
    external CONS_TO_PRIM
    b0 = 0.0d0
    call CONS_TO_PRIM(gam, 1.4d0, 1.0d0, 0.0d0, 1.4d0, 1.0d0, rhof, prsf, 1.0d-5, &
                      1.0d-12, 1.0d-12, iflag, ent, b0, u, v)
    flag = iflag
  end subroutine

  subroutine ref_prim_to_cons(gam, rhof, prsf, ent, v, u) bind(C, name="ref_prim_to_cons")
    real(c_double), value :: gam, rhof, prsf, ent
    real(c_double), intent(inout) :: v(8)
    real(c_double), intent(out) :: u(8)
    real(c_double) :: b0(3), e
    integer :: iflag
    external PRIM_TO_CONS
    b0 = 0.0d0
    e = ent
    call PRIM_TO_CONS(gam, 1.4d0, 1.0d0, 0.0d0, 1.4d0, 1.0d0, rhof, prsf, 1.0d-5, &
                      1.0d-12, 1.0d-12, iflag, e, b0, u, v)
  end subroutine

  subroutine ref_wavspd(idir, gam, v, smin, smax) bind(C, name="ref_wavspd")
    integer(c_int), value :: idir
    real(c_double), value :: gam
    real(c_double), intent(in) :: v(8)
    real(c_double), intent(out) :: smin, smax
    real(c_double) :: b0(3), vv(8)
    external PRIM_TO_WAVSPD
    b0 = 0.0d0
    vv = v
    call PRIM_TO_WAVSPD(idir, gam, b0, vv, smin, smax)
  end subroutine

  ! Two_D_RS_Ez_4State_HLL / _4State_LLF / _3State_LLF (630, 632, 633).
  subroutine ref_emf4_hll(s, cs_t, sn_t, ez, bn, by, out) bind(C, name="ref_emf4_hll")
    real(c_double), intent(in) :: s(4), ez(4), bn(2), by(2)
    real(c_double), value :: cs_t, sn_t
    real(c_double), intent(out) :: out(3)
    external Two_D_RS_Ez_4State_HLL
    ! s = (S_R, S_L, S_U, S_D); ez = (RU, LU, LD, RD); bn = (U, D); by = (R, L)
    call Two_D_RS_Ez_4State_HLL(s(1), s(2), s(3), s(4), cs_t, sn_t, ez(1), ez(2), ez(3), ez(4), &
                                bn(1), bn(2), by(1), by(2), out(1), out(2), out(3))
  end subroutine

  subroutine ref_emf4_llf(s, cs_t, sn_t, ez, bn, by, out) bind(C, name="ref_emf4_llf")
    real(c_double), intent(in) :: s(4), ez(4), bn(2), by(2)
    real(c_double), value :: cs_t, sn_t
    real(c_double), intent(out) :: out(3)
    external Two_D_RS_Ez_4State_LLF
    call Two_D_RS_Ez_4State_LLF(s(1), s(2), s(3), s(4), cs_t, sn_t, ez(1), ez(2), ez(3), ez(4), &
                                bn(1), bn(2), by(1), by(2), out(1), out(2), out(3))
  end subroutine

  subroutine ref_emf3_llf(s6, ez, bn, out) bind(C, name="ref_emf3_llf")
    real(c_double), intent(in) :: s6(6), ez(3), bn(3)
    real(c_double), intent(out) :: out(3)
    external Two_D_RS_Ez_3State_LLF
    call Two_D_RS_Ez_3State_LLF(s6(1), s6(2), s6(3), s6(4), s6(5), s6(6), ez(1), ez(2), ez(3), &
                                bn(1), bn(2), bn(3), out(1), out(2), out(3))
  end subroutine

  ! EXTRACT_VERT_COORDS_PATCH (210extract_vert_coords_patch.f:7) for the six
  ! sectors, one patch each: vertices r(k) * u(i, j) and the active-zone mask.
  subroutine ref_vert_coords(n, nr, g, coords, rb, vert, mask) bind(C, name="ref_vert_coords")
    integer(c_int), value :: n, nr, g
    real(c_double), intent(in) :: coords(0-g:n+g, 0-g:n+g, 3, 6), rb(0-g:nr+g)
    real(c_double), intent(out) :: vert(0-g:n+g, 0-g:n+g, 0-g:nr+g, 3, 6)
    integer(c_int), intent(out) :: mask(1-g:n+g, 1-g:n+g, 1-g:nr+g, 6)
    integer :: s
    external EXTRACT_VERT_COORDS_PATCH, REF_INIT_COMMON
    call REF_INIT_COMMON()
    do s = 1, 6
      call EXTRACT_VERT_COORDS_PATCH(n, nr, g, n, nr, 1, 1, g, s, 1, 1, 1, 1, 1, coords, rb, &
                                     vert(:, :, :, :, s), mask(:, :, :, s))
    end do
  end subroutine

  ! COORDS_TO_WT_QUADPTS_VOLUME_ISO2 (240coords_to_wt_quadpts_volume_iso2.f:7), six sectors.
  subroutine ref_quad_volume(n, nr, g, vert, mask, pts, wts, tang, vol, cen) bind(C, name="ref_quad_volume")
    integer(c_int), value :: n, nr, g
    real(c_double), intent(in) :: vert(0-g:n+g, 0-g:n+g, 0-g:nr+g, 3, 6)
    integer(c_int), intent(in) :: mask(1-g:n+g, 1-g:n+g, 1-g:nr+g, 6)
    real(c_double), intent(out) :: pts(1-g:n+g, 1-g:n+g, 1-g:nr+g, 3, 8, 6), wts(1-g:n+g, 1-g:n+g, 1-g:nr+g, 8, 6)
    real(c_double), intent(out) :: tang(1-g:n+g, 1-g:n+g, 1-g:nr+g, 3, 3, 8, 6)
    real(c_double), intent(out) :: vol(1-g:n+g, 1-g:n+g, 1-g:nr+g, 6), cen(1-g:n+g, 1-g:n+g, 1-g:nr+g, 3, 6)
    integer :: s
    external COORDS_TO_WT_QUADPTS_VOLUME_ISO2, REF_INIT_COMMON
    call REF_INIT_COMMON()
    do s = 1, 6
      call COORDS_TO_WT_QUADPTS_VOLUME_ISO2(nr, n, g, vert(:, :, :, :, s), mask(:, :, :, s), &
           pts(:, :, :, :, :, s), wts(:, :, :, :, s), tang(:, :, :, :, :, :, s), vol(:, :, :, s), cen(:, :, :, :, s))
    end do
  end subroutine

  ! COORDS_TO_WT_QUADPTS_FACES_ISO2 (220coords_to_wt_quadpts_faces_iso2.f:7), six sectors.
  subroutine ref_quad_faces(n, nr, g, vert, mask, xp, xw, xn, xa, xc, yp, yw, yn, ya, yc, &
                            zp, zw, zn, za, zc, scl) bind(C, name="ref_quad_faces")
    integer(c_int), value :: n, nr, g
    real(c_double), intent(in) :: vert(0-g:n+g, 0-g:n+g, 0-g:nr+g, 3, 6)
    integer(c_int), intent(in) :: mask(1-g:n+g, 1-g:n+g, 1-g:nr+g, 6)
    real(c_double), intent(out) :: xp(0-g:n+g, 1-g:n+g, 1-g:nr+g, 3, 5, 6), xw(0-g:n+g, 1-g:n+g, 1-g:nr+g, 5, 6)
    real(c_double), intent(out) :: xn(0-g:n+g, 1-g:n+g, 1-g:nr+g, 3, 6), xa(0-g:n+g, 1-g:n+g, 1-g:nr+g, 6)
    real(c_double), intent(out) :: xc(0-g:n+g, 1-g:n+g, 1-g:nr+g, 3, 6)
    real(c_double), intent(out) :: yp(1-g:n+g, 0-g:n+g, 1-g:nr+g, 3, 5, 6), yw(1-g:n+g, 0-g:n+g, 1-g:nr+g, 5, 6)
    real(c_double), intent(out) :: yn(1-g:n+g, 0-g:n+g, 1-g:nr+g, 3, 6), ya(1-g:n+g, 0-g:n+g, 1-g:nr+g, 6)
    real(c_double), intent(out) :: yc(1-g:n+g, 0-g:n+g, 1-g:nr+g, 3, 6)
    real(c_double), intent(out) :: zp(1-g:n+g, 1-g:n+g, 0-g:nr+g, 3, 5, 6), zw(1-g:n+g, 1-g:n+g, 0-g:nr+g, 5, 6)
    real(c_double), intent(out) :: zn(1-g:n+g, 1-g:n+g, 0-g:nr+g, 3, 5, 6), za(1-g:n+g, 1-g:n+g, 0-g:nr+g, 6)
    real(c_double), intent(out) :: zc(1-g:n+g, 1-g:n+g, 0-g:nr+g, 3, 6)
    real(c_double), intent(out) :: scl(1-g:n+g, 1-g:n+g, 1-g:nr+g, 6)
    integer :: s
    external COORDS_TO_WT_QUADPTS_FACES_ISO2, REF_INIT_COMMON
    call REF_INIT_COMMON()
    do s = 1, 6
      call COORDS_TO_WT_QUADPTS_FACES_ISO2(nr, n, g, vert(:, :, :, :, s), mask(:, :, :, s), &
           xp(:, :, :, :, :, s), xw(:, :, :, :, s), xn(:, :, :, :, s), xa(:, :, :, s), xc(:, :, :, :, s), &
           yp(:, :, :, :, :, s), yw(:, :, :, :, s), yn(:, :, :, :, s), ya(:, :, :, s), yc(:, :, :, :, s), &
           zp(:, :, :, :, :, s), zw(:, :, :, :, s), zn(:, :, :, :, :, s), za(:, :, :, s), zc(:, :, :, :, s), &
           scl(:, :, :, s))
    end do
  end subroutine

  ! BUILD_MODE_INDICES (204build_mode_indices.f:7) and INIT_BASIS_FUNCTION_COEFS
  ! (250init_basis_function_coefs.f:5), second order: four modes.
  subroutine ref_basis_moments(n, nr, g, mask, pts, wts, vol, cen, scl, coef) bind(C, name="ref_basis_moments")
    integer(c_int), value :: n, nr, g
    integer(c_int), intent(in) :: mask(1-g:n+g, 1-g:n+g, 1-g:nr+g, 6)
    real(c_double), intent(in) :: pts(1-g:n+g, 1-g:n+g, 1-g:nr+g, 3, 8, 6), wts(1-g:n+g, 1-g:n+g, 1-g:nr+g, 8, 6)
    real(c_double), intent(in) :: vol(1-g:n+g, 1-g:n+g, 1-g:nr+g, 6), cen(1-g:n+g, 1-g:n+g, 1-g:nr+g, 3, 6)
    real(c_double), intent(in) :: scl(1-g:n+g, 1-g:n+g, 1-g:nr+g, 6)
    real(c_double), intent(out) :: coef(1-g:n+g, 1-g:n+g, 1-g:nr+g, 4, 6)
    integer :: s, ii(4), jj(4), kk(4)
    external BUILD_MODE_INDICES, INIT_BASIS_FUNCTION_COEFS, REF_INIT_COMMON
    call REF_INIT_COMMON()
    call BUILD_MODE_INDICES(4, ii, jj, kk)
    do s = 1, 6
      call INIT_BASIS_FUNCTION_COEFS(nr, n, g, 4, ii, jj, kk, mask(:, :, :, s), pts(:, :, :, :, :, s), &
           wts(:, :, :, :, s), vol(:, :, :, s), cen(:, :, :, :, s), scl(:, :, :, s), coef(:, :, :, :, s))
    end do
  end subroutine

  subroutine ref_mode_indices(nm, ii, jj, kk) bind(C, name="ref_mode_indices")
    integer(c_int), value :: nm
    integer(c_int), intent(out) :: ii(nm), jj(nm), kk(nm)
    external BUILD_MODE_INDICES
    call BUILD_MODE_INDICES(nm, ii, jj, kk)
  end subroutine

  ! STENCIL_INDEXING_TRICORNERS (211stencil_indexing_tricorners.f:6) for a
  ! one-patch sector.
  subroutine ref_stencil_indexing(n, nr, g, st) bind(C, name="ref_stencil_indexing")
    integer(c_int), value :: n, nr, g
    integer(c_int), intent(out) :: st(1-g:n+g, 1-g:n+g, 2, 12)
    integer(c_int), allocatable :: mask(:, :, :)
    integer :: ld, rd, ru, lu
    external STENCIL_INDEXING_TRICORNERS, REF_INIT_COMMON
    call REF_INIT_COMMON()
    allocate(mask(1-g:n+g, 1-g:n+g, 1-g:nr+g))
    mask = 1
    call STENCIL_INDEXING_TRICORNERS(n, nr, 1, 1, g, 1, 1, ld, rd, ru, lu, mask, st)
    deallocate(mask)
  end subroutine

  ! COORDS_TO_WT_QUADPTS_BFLD_ISO2 (230coords_to_wt_quadpts_bfld_iso2.f:8), six sectors.
  subroutine ref_quad_bfield(n, nr, g, vert, mask, xp, xw, xn, yp, yw, yn, zp, zw, zn) &
      bind(C, name="ref_quad_bfield")
    integer(c_int), value :: n, nr, g
    real(c_double), intent(in) :: vert(0-g:n+g, 0-g:n+g, 0-g:nr+g, 3, 6)
    integer(c_int), intent(in) :: mask(1-g:n+g, 1-g:n+g, 1-g:nr+g, 6)
    real(c_double), intent(out) :: xp(0-g:n+g, 1-g:n+g, 1-g:nr+g, 3, 5, 6), xw(0-g:n+g, 1-g:n+g, 1-g:nr+g, 5, 6)
    real(c_double), intent(out) :: xn(0-g:n+g, 1-g:n+g, 1-g:nr+g, 3, 6)
    real(c_double), intent(out) :: yp(1-g:n+g, 0-g:n+g, 1-g:nr+g, 3, 5, 6), yw(1-g:n+g, 0-g:n+g, 1-g:nr+g, 5, 6)
    real(c_double), intent(out) :: yn(1-g:n+g, 0-g:n+g, 1-g:nr+g, 3, 6)
    real(c_double), intent(out) :: zp(1-g:n+g, 1-g:n+g, 0-g:nr+g, 3, 5, 6), zw(1-g:n+g, 1-g:n+g, 0-g:nr+g, 5, 6)
    real(c_double), intent(out) :: zn(1-g:n+g, 1-g:n+g, 0-g:nr+g, 3, 5, 6)
    integer :: s
    external COORDS_TO_WT_QUADPTS_BFLD_ISO2, REF_INIT_COMMON
    call REF_INIT_COMMON()
    do s = 1, 6
      call COORDS_TO_WT_QUADPTS_BFLD_ISO2(nr, n, g, vert(:, :, :, :, s), mask(:, :, :, s), &
           xp(:, :, :, :, :, s), xw(:, :, :, :, s), xn(:, :, :, :, s), &
           yp(:, :, :, :, :, s), yw(:, :, :, :, s), yn(:, :, :, :, s), &
           zp(:, :, :, :, :, s), zw(:, :, :, :, s), zn(:, :, :, :, :, s))
    end do
  end subroutine

  ! LIMIT_ALL_SLOPES_P1 (410limit_all_slopes_p1.f:9) on one patch, from the
  ! fields of a C structure laid out as or_p1_patch / cs_p1_patch.  The arrays
  ! the routine declares but does not read at this order (area quadrature
  ! points and weights, z-face normals, volumes, basis moments, Wzc) are
  ! passed as zero-filled stand-ins of their declared sizes.
  subroutine ref_limit_all_slopes_p1(p) bind(C, name="ref_limit_all_slopes_p1")
    type(ref_p1_patch), intent(in) :: p
    integer :: n, nr, g, nz, nzr, nf, nfr, nzone, nxf, nzf, nfc, nv
    integer :: md(4)
    real(c_double), pointer :: vert(:), xfn(:), xfa(:), xfc(:), yfn(:), yfa(:), yfc(:), zfa(:), zfc(:)
    real(c_double), pointer :: xbp(:), xbw(:), xbn(:), ybp(:), ybw(:), ybn(:), zbp(:), zbw(:), zbn(:)
    real(c_double), pointer :: scl(:), cen(:), u(:), ent(:), bxf(:), byf(:), bzf(:), v(:)
    real(c_double), pointer :: b1(:), b2(:), b3(:), ex(:), ey(:), ez(:), geos(:), gsnd(:), mmw(:)
    real(c_double), pointer :: divv(:), msn(:), flat(:), work(:), bxq(:), byq(:), bzq(:), vmin(:), vmax(:)
    integer(c_int), pointer :: mask(:), st(:)
    real(c_double), allocatable :: xq(:), xw(:), yq(:), yw(:), zq(:), zw(:), zn(:), vol(:), coef(:), wzc(:)
    external LIMIT_ALL_SLOPES_P1, REF_INIT_COMMON, REF_SET_P1_COMMON
    n = p%ntp; nr = p%nr; g = p%nghost
    nz = n + 2 * g; nzr = nr + 2 * g; nf = nz + 1; nfr = nzr + 1
    nzone = nz * nz * nzr; nxf = nf * nz * nzr; nzf = nz * nz * nfr; nfc = nf * nf * nfr; nv = nfc
    md = (/ 0, 1, 0, 0 /)
    call REF_INIT_COMMON()
    call REF_SET_P1_COMMON(p%gamma, p%molewt, p%rhofloor, p%prsfloor, p%smallnum, p%flatten_coef, &
                           p%prs_flatten_coef)
    call c_f_pointer(p%vert, vert, [3 * nv])
    call c_f_pointer(p%mask, mask, [nzone])
    call c_f_pointer(p%stencil, st, [nz * nz * 24])
    call c_f_pointer(p%xf_normal, xfn, [3 * nxf]); call c_f_pointer(p%xf_area, xfa, [nxf])
    call c_f_pointer(p%xf_centroid, xfc, [3 * nxf])
    call c_f_pointer(p%yf_normal, yfn, [3 * nxf]); call c_f_pointer(p%yf_area, yfa, [nxf])
    call c_f_pointer(p%yf_centroid, yfc, [3 * nxf])
    call c_f_pointer(p%zf_area, zfa, [nzf]); call c_f_pointer(p%zf_centroid, zfc, [3 * nzf])
    call c_f_pointer(p%xb_pts, xbp, [15 * nxf]); call c_f_pointer(p%xb_wts, xbw, [5 * nxf])
    call c_f_pointer(p%xb_normal, xbn, [3 * nxf])
    call c_f_pointer(p%yb_pts, ybp, [15 * nxf]); call c_f_pointer(p%yb_wts, ybw, [5 * nxf])
    call c_f_pointer(p%yb_normal, ybn, [3 * nxf])
    call c_f_pointer(p%zb_pts, zbp, [15 * nzf]); call c_f_pointer(p%zb_wts, zbw, [5 * nzf])
    call c_f_pointer(p%zb_normal, zbn, [15 * nzf])
    call c_f_pointer(p%scale, scl, [nzone]); call c_f_pointer(p%v_centroid, cen, [3 * nzone])
    call c_f_pointer(p%U, u, [32 * nzone]); call c_f_pointer(p%ent, ent, [4 * nzone])
    call c_f_pointer(p%Bxf, bxf, [nfc]); call c_f_pointer(p%Byf, byf, [nfc]); call c_f_pointer(p%Bzf, bzf, [nfc])
    call c_f_pointer(p%V, v, [32 * nzone])
    call c_f_pointer(p%B1_zc, b1, [4 * nzone]); call c_f_pointer(p%B2_zc, b2, [4 * nzone])
    call c_f_pointer(p%B3_zc, b3, [4 * nzone])
    call c_f_pointer(p%Ex_zc, ex, [4 * nzone]); call c_f_pointer(p%Ey_zc, ey, [4 * nzone])
    call c_f_pointer(p%Ez_zc, ez, [4 * nzone])
    call c_f_pointer(p%gamma_eos, geos, [nzone]); call c_f_pointer(p%gamma_sound, gsnd, [nzone])
    call c_f_pointer(p%mean_mol_wt, mmw, [nzone]); call c_f_pointer(p%divvel, divv, [nzone])
    call c_f_pointer(p%msonic, msn, [nzone]); call c_f_pointer(p%flatten, flat, [nzone])
    call c_f_pointer(p%work, work, [nzone])
    call c_f_pointer(p%Bxf_q, bxq, [5 * nfc]); call c_f_pointer(p%Byf_q, byq, [5 * nfc])
    call c_f_pointer(p%Bzf_q, bzq, [5 * nfc])
    call c_f_pointer(p%V_min_local, vmin, [9]); call c_f_pointer(p%V_max_local, vmax, [9])
    allocate(xq(15 * nxf), xw(5 * nxf), yq(15 * nxf), yw(5 * nxf), zq(15 * nzf), zw(5 * nzf), zn(15 * nzf))
    allocate(vol(nzone), coef(4 * nzone), wzc(4 * nzone))
    xq = 0; xw = 0; yq = 0; yw = 0; zq = 0; zw = 0; zn = 0; vol = 1; coef = 0; wzc = 0
    call LIMIT_ALL_SLOPES_P1(nr, n, g, 8, 4, 1, 1, 1, md, md, md, vert, mask, &
         xq, xw, xfn, xfa, xfc, yq, yw, yfn, yfa, yfc, zq, zw, zn, zfa, zfc, &
         xbp, xbw, xbn, ybp, ybw, ybn, zbp, zbw, zbn, scl, vol, cen, coef, &
         p%itri(1), p%itri(2), p%itri(3), p%itri(4), st, &
         u, ent, bxf, byf, bzf, v, b1, b2, b3, ex, ey, ez, wzc, &
         geos, gsnd, mmw, divv, msn, flat, work, bxq, byq, bzq, vmin, vmax)
    deallocate(xq, xw, yq, yw, zq, zw, zn, vol, coef, wzc)
  end subroutine

  ! CLEANUP_AFTER_UPDATE (485cleanup_after_update.f:1) on the fields of a
  ! ref_p1_patch; z-face normals are declared but unused in the MHD branch
  subroutine ref_cleanup_after_update_p1(p) bind(C, name="ref_cleanup_after_update_p1")
    type(ref_p1_patch), intent(in) :: p
    integer :: n, nr, g, nz, nzr, nf, nfr, nzone, nxf, nzf, nfc
    real(c_double), pointer :: vert(:), xfn(:), xfc(:), yfn(:), yfc(:), zfc(:), scl(:), cen(:)
    real(c_double), pointer :: u(:), ent(:), bxf(:), byf(:), bzf(:), v(:), b1(:), b2(:), b3(:)
    real(c_double), pointer :: geos(:), gsnd(:), mmw(:)
    integer(c_int), pointer :: mask(:), st(:)
    real(c_double), allocatable :: zn(:)
    external CLEANUP_AFTER_UPDATE, REF_INIT_COMMON, REF_SET_P1_COMMON
    n = p%ntp; nr = p%nr; g = p%nghost
    nz = n + 2 * g; nzr = nr + 2 * g; nf = nz + 1; nfr = nzr + 1
    nzone = nz * nz * nzr; nxf = nf * nz * nzr; nzf = nz * nz * nfr; nfc = nf * nf * nfr
    call REF_INIT_COMMON()
    call REF_SET_P1_COMMON(p%gamma, p%molewt, p%rhofloor, p%prsfloor, p%smallnum, p%flatten_coef, &
                           p%prs_flatten_coef)
    call c_f_pointer(p%vert, vert, [3 * nfc]); call c_f_pointer(p%mask, mask, [nzone])
    call c_f_pointer(p%stencil, st, [nz * nz * 24])
    call c_f_pointer(p%xf_normal, xfn, [3 * nxf]); call c_f_pointer(p%xf_centroid, xfc, [3 * nxf])
    call c_f_pointer(p%yf_normal, yfn, [3 * nxf]); call c_f_pointer(p%yf_centroid, yfc, [3 * nxf])
    call c_f_pointer(p%zf_centroid, zfc, [3 * nzf])
    call c_f_pointer(p%scale, scl, [nzone]); call c_f_pointer(p%v_centroid, cen, [3 * nzone])
    call c_f_pointer(p%U, u, [32 * nzone]); call c_f_pointer(p%ent, ent, [4 * nzone])
    call c_f_pointer(p%Bxf, bxf, [nfc]); call c_f_pointer(p%Byf, byf, [nfc]); call c_f_pointer(p%Bzf, bzf, [nfc])
    call c_f_pointer(p%V, v, [32 * nzone])
    call c_f_pointer(p%B1_zc, b1, [4 * nzone]); call c_f_pointer(p%B2_zc, b2, [4 * nzone])
    call c_f_pointer(p%B3_zc, b3, [4 * nzone])
    call c_f_pointer(p%gamma_eos, geos, [nzone]); call c_f_pointer(p%gamma_sound, gsnd, [nzone])
    call c_f_pointer(p%mean_mol_wt, mmw, [nzone])
    allocate(zn(15 * nzf)); zn = 0
    call CLEANUP_AFTER_UPDATE(nr, n, g, 8, 4, 1, 1, 1, vert, mask, xfn, xfc, yfn, yfc, zn, zfc, &
         scl, cen, p%itri(1), p%itri(2), p%itri(3), p%itri(4), st, u, ent, bxf, byf, bzf, &
         v, b1, b2, b3, geos, gsnd, mmw)
    deallocate(zn)
  end subroutine

  ! EVALUATE_TIMESTEP (490evaluate_timestep.f:1) on the fields of a ref_p1_patch
  subroutine ref_evaluate_timestep_p1(p, cfl, dtcur, dtnew) bind(C, name="ref_evaluate_timestep_p1")
    type(ref_p1_patch), intent(in) :: p
    real(c_double), value :: cfl, dtcur
    real(c_double), intent(out) :: dtnew
    integer :: n, nr, g, nz, nzr, nf, nfr, nzone, nfc
    real(c_double) :: t
    real(c_double), pointer :: vert(:), scl(:), cen(:), u(:), ent(:), bxf(:), byf(:), bzf(:), v(:)
    real(c_double), pointer :: b1(:), b2(:), b3(:), geos(:), gsnd(:), mmw(:), divv(:), msn(:), flat(:)
    integer(c_int), pointer :: mask(:)
    external EVALUATE_TIMESTEP, REF_INIT_COMMON, REF_SET_P1_COMMON, REF_SET_CFL
    n = p%ntp; nr = p%nr; g = p%nghost
    nz = n + 2 * g; nzr = nr + 2 * g; nf = nz + 1; nfr = nzr + 1
    nzone = nz * nz * nzr; nfc = nf * nf * nfr
    call REF_INIT_COMMON()
    call REF_SET_P1_COMMON(p%gamma, p%molewt, p%rhofloor, p%prsfloor, p%smallnum, p%flatten_coef, &
                           p%prs_flatten_coef)
    call REF_SET_CFL(cfl)
    call c_f_pointer(p%vert, vert, [3 * nfc]); call c_f_pointer(p%mask, mask, [nzone])
    call c_f_pointer(p%scale, scl, [nzone]); call c_f_pointer(p%v_centroid, cen, [3 * nzone])
    call c_f_pointer(p%U, u, [32 * nzone]); call c_f_pointer(p%ent, ent, [4 * nzone])
    call c_f_pointer(p%Bxf, bxf, [nfc]); call c_f_pointer(p%Byf, byf, [nfc]); call c_f_pointer(p%Bzf, bzf, [nfc])
    call c_f_pointer(p%V, v, [32 * nzone])
    call c_f_pointer(p%B1_zc, b1, [4 * nzone]); call c_f_pointer(p%B2_zc, b2, [4 * nzone])
    call c_f_pointer(p%B3_zc, b3, [4 * nzone])
    call c_f_pointer(p%gamma_eos, geos, [nzone]); call c_f_pointer(p%gamma_sound, gsnd, [nzone])
    call c_f_pointer(p%mean_mol_wt, mmw, [nzone]); call c_f_pointer(p%divvel, divv, [nzone])
    call c_f_pointer(p%msonic, msn, [nzone]); call c_f_pointer(p%flatten, flat, [nzone])
    t = 0
    call EVALUATE_TIMESTEP(nr, n, g, 8, 4, 1, t, dtcur, dtnew, vert, mask, scl, cen, &
         u, ent, bxf, byf, bzf, v, b1, b2, b3, geos, gsnd, mmw, divv, msn, flat)
  end subroutine

  ! PAD_BOUNDARY_R (826pad_boundary_r.f:1) on the fields of a ref_p1_pad; U
  ! carries mode 1 only (n_cc_modes 1), the arguments the routine does not read
  ! (vertices, centroids, entropy) are zero-filled arrays of their sizes
  subroutine ref_pad_boundary_r_p1(p) bind(C, name="ref_pad_boundary_r_p1")
    type(ref_p1_pad), intent(in) :: p
    integer :: n, nr, g, nz, nzr, nf, nfr, nzone, nxf, nzf, nfc
    real(c_double), pointer :: xa(:), ya(:), za(:), vol(:), u(:), bxf(:), byf(:), bzf(:)
    integer(c_int), pointer :: mask(:), bc(:)
    real(c_double), allocatable :: vert(:), cen(:), ent(:)
    external PAD_BOUNDARY_R
    n = p%ntp; nr = p%nr; g = p%nghost
    nz = n + 2 * g; nzr = nr + 2 * g; nf = nz + 1; nfr = nzr + 1
    nzone = nz * nz * nzr; nxf = nf * nz * nzr; nzf = nz * nz * nfr; nfc = nf * nf * nfr
    call c_f_pointer(p%mask, mask, [nzone]); call c_f_pointer(p%bcarr_r, bc, [2 * nz * nz])
    call c_f_pointer(p%xf_area, xa, [nxf]); call c_f_pointer(p%yf_area, ya, [nxf])
    call c_f_pointer(p%zf_area, za, [nzf]); call c_f_pointer(p%v_volume, vol, [nzone])
    call c_f_pointer(p%U, u, [8 * nzone])
    call c_f_pointer(p%Bxf, bxf, [nfc]); call c_f_pointer(p%Byf, byf, [nfc]); call c_f_pointer(p%Bzf, bzf, [nfc])
    allocate(vert(3 * nfc), cen(3 * nzone), ent(4 * nzone))
    vert = 0; cen = 0; ent = 0
    call PAD_BOUNDARY_R(nr, n, g, 8, 1, 1, vert, mask, xa, ya, za, vol, cen, u, ent, bxf, byf, bzf, bc)
    deallocate(vert, cen, ent)
  end subroutine

  ! MAKE_EDGE_EFIELDS_P1 (460make_edge_fields_p1.f:5) on the fields of a
  ! ref_p1_edge; the arguments the routine declares but does not read at second
  ! order are zero-filled arrays of their declared sizes
  subroutine ref_make_edge_efields_p1(p) bind(C, name="ref_make_edge_efields_p1")
    type(ref_p1_edge), intent(in) :: p
    integer :: n, nr, g, nz, nzr, nf, nfr, nzone, nxf, nzf, nfc
    integer :: md(4)
    real(c_double), pointer :: vert(:), ex(:), ey(:), ez(:), bxq(:), byq(:), bzq(:)
    real(c_double), pointer :: sl(:), sr(:), sd(:), su(:), sb(:), st_(:), efx(:), efy(:), efz(:)
    integer(c_int), pointer :: mask(:)
    real(c_double), allocatable :: fq(:), fw(:), fn(:), fa(:), fc(:), zq(:), zw(:), zn(:), za(:), zc(:)
    real(c_double), allocatable :: zone1(:), zone3(:), zone4(:), zone32(:), ent(:), bf(:)
    integer(c_int), allocatable :: st(:)
    external MAKE_EDGE_EFIELDS_P1, REF_INIT_COMMON
    n = p%ntp; nr = p%nr; g = p%nghost
    nz = n + 2 * g; nzr = nr + 2 * g; nf = nz + 1; nfr = nzr + 1
    nzone = nz * nz * nzr; nxf = nf * nz * nzr; nzf = nz * nz * nfr; nfc = nf * nf * nfr
    md = (/ 0, 1, 0, 0 /)
    call REF_INIT_COMMON()
    call c_f_pointer(p%vert, vert, [3 * nfc]); call c_f_pointer(p%mask, mask, [nzone])
    call c_f_pointer(p%Ex_zc, ex, [4 * nzone]); call c_f_pointer(p%Ey_zc, ey, [4 * nzone])
    call c_f_pointer(p%Ez_zc, ez, [4 * nzone])
    call c_f_pointer(p%Bxf_q, bxq, [5 * nfc]); call c_f_pointer(p%Byf_q, byq, [5 * nfc])
    call c_f_pointer(p%Bzf_q, bzq, [5 * nfc])
    call c_f_pointer(p%S_L, sl, [nfc]); call c_f_pointer(p%S_R, sr, [nfc]); call c_f_pointer(p%S_D, sd, [nfc])
    call c_f_pointer(p%S_U, su, [nfc]); call c_f_pointer(p%S_B, sb, [nfc]); call c_f_pointer(p%S_T, st_, [nfc])
    call c_f_pointer(p%Efx, efx, [nfc]); call c_f_pointer(p%Efy, efy, [nfc]); call c_f_pointer(p%Efz, efz, [nfc])
    allocate(fq(15 * nxf), fw(5 * nxf), fn(3 * nxf), fa(nxf), fc(3 * nxf))
    allocate(zq(15 * nzf), zw(5 * nzf), zn(15 * nzf), za(nzf), zc(3 * nzf))
    allocate(zone1(nzone), zone3(3 * nzone), zone4(4 * nzone), zone32(32 * nzone), ent(4 * nzone), bf(nfc))
    allocate(st(nz * nz * 24))
    fq = 0; fw = 0; fn = 0; fa = 0; fc = 0; zq = 0; zw = 0; zn = 0; za = 0; zc = 0
    zone1 = 1; zone3 = 0; zone4 = 0; zone32 = 0; ent = 0; bf = 0; st = 1
    call MAKE_EDGE_EFIELDS_P1(nr, n, g, 8, 4, 1, md, md, md, vert, mask, &
         fq, fw, fn, fa, fc, fq, fw, fn, fa, fc, zq, zw, zn, za, zc, &
         fq, fw, fn, fq, fw, fn, zq, zw, zn, zone1, zone1, zone3, zone4, &
         p%itri(1), p%itri(2), p%itri(3), p%itri(4), st, &
         zone32, ent, bf, bf, bf, zone32, zone4, zone4, zone4, ex, ey, ez, &
         zone1, zone1, zone1, zone1, zone1, zone1, bxq, byq, bzq, sl, sr, sd, su, sb, st_, efx, efy, efz)
    deallocate(fq, fw, fn, fa, fc, zq, zw, zn, za, zc, zone1, zone3, zone4, zone32, ent, bf, st)
  end subroutine


  ! INIT_PATCH_CUBE_SPHERE (820init_patch_cube_sphere.f:8) with layers = 0, else
  ! ADD_BOUNDARY (825add_boundary.f:12) with ibdy_inner = layers & 1 and
  ! ibdy_outer = layers >> 1, on one patch of a ref_p1_init (U and the face B
  ! carry one mode; the entropy mode 1 is copied in and out); the geometry
  ! arguments the routines do not read at these problems are zero-filled
  subroutine ref_init_patch(p, pb, layers) bind(C, name="ref_init_patch")
    type(ref_p1_init), intent(in) :: p
    type(ref_prob), intent(in) :: pb
    integer(c_int), value :: layers
    integer :: n, nr, g, nz, nzr, nf, nfr, nzone, nxf, nzf, nfc, ib_in, ib_out
    real(c_double), pointer :: vert(:), vpts(:), vwts(:), xa(:), ya(:), za(:), u(:), ent(:)
    real(c_double), pointer :: bxf(:), byf(:), bzf(:)
    integer(c_int), pointer :: mask(:)
    real(c_double), allocatable :: scl(:), xq(:), xw(:), xn(:), xc(:), yq(:), yw(:), yn(:), yc(:)
    real(c_double), allocatable :: zq(:), zw(:), zn(:), zc(:), v(:), entm(:), b0(:), geos(:), vcen(:)
    real(c_double) :: vmin(9), vmax(9)
    integer(c_int), allocatable :: bc(:)
    external INIT_PATCH_CUBE_SPHERE, ADD_BOUNDARY, REF_INIT_COMMON, REF_SET_P1_COMMON
    n = p%ntp; nr = p%nr; g = p%nghost
    nz = n + 2 * g; nzr = nr + 2 * g; nf = nz + 1; nfr = nzr + 1
    nzone = nz * nz * nzr; nxf = nf * nz * nzr; nzf = nz * nz * nfr; nfc = nf * nf * nfr
    call c_f_pointer(p%vert, vert, [3 * nfc]); call c_f_pointer(p%mask, mask, [nzone])
    call c_f_pointer(p%v_pts, vpts, [24 * nzone]); call c_f_pointer(p%v_wts, vwts, [8 * nzone])
    call c_f_pointer(p%xf_area, xa, [nxf]); call c_f_pointer(p%yf_area, ya, [nxf])
    call c_f_pointer(p%zf_area, za, [nzf])
    call c_f_pointer(p%U, u, [8 * nzone]); call c_f_pointer(p%ent, ent, [nzone])
    call c_f_pointer(p%Bxf, bxf, [nfc]); call c_f_pointer(p%Byf, byf, [nfc]); call c_f_pointer(p%Bzf, bzf, [nfc])
    allocate(scl(nzone), xq(15 * nxf), xw(5 * nxf), xn(3 * nxf), xc(3 * nxf), yq(15 * nxf), yw(5 * nxf), &
             yn(3 * nxf), yc(3 * nxf), zq(15 * nzf), zw(5 * nzf), zn(15 * nzf), zc(3 * nzf), v(8 * nzone), &
             entm(4 * nzone), b0(3 * nzone), geos(nzone), vcen(3 * nzone), bc(2 * nz * nz))
    scl = 0; xq = 0; xw = 0; xn = 0; xc = 0; yq = 0; yw = 0; yn = 0; yc = 0
    zq = 0; zw = 0; zn = 0; zc = 0; v = 0; entm = 0; b0 = 0; geos = pb%gamma; vcen = 0; bc = 0
    entm(1:nzone) = ent(1:nzone)
    call REF_INIT_COMMON()
    call REF_SET_P1_COMMON(pb%gamma, 1.0d0, pb%rhofloor, pb%prsfloor, 1.0d-12, 0.4d0, 1.5d0)
    if (layers == 0) then
      call INIT_PATCH_CUBE_SPHERE(nr, n, g, 8, 1, 1, pb%iproblem, pb%time, 0.0d0, pb%rmin, pb%rmax, vert, mask, &
           scl, xq, xw, xn, xa, xc, yq, yw, yn, ya, yc, zq, zw, zn, za, zc, vpts, vwts, &
           u, v, entm, bxf, byf, bzf, b0, geos, vmin, vmax)
    else
      ib_in = iand(layers, 1); ib_out = iand(ishft(layers, -1), 1)
      call ADD_BOUNDARY(ib_in, ib_out, nr, n, g, 8, 1, 1, pb%iproblem, pb%time, 0.0d0, pb%rmin, pb%rmax, &
           vert, mask, scl, xq, xw, xn, xa, xc, yq, yw, yn, ya, yc, zq, zw, zn, za, zc, vpts, vwts, vcen, &
           u, v, entm, bxf, byf, bzf, geos, bc, vmin, vmax)
    end if
    ent(1:nzone) = entm(1:nzone)
    deallocate(scl, xq, xw, xn, xc, yq, yw, yn, yc, zq, zw, zn, zc, v, entm, b0, geos, vcen, bc)
  end subroutine

end module cs_ref_bind
