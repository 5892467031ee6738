! ref_bind3.f90 -- C-callable wrappers around the reference's third-order
! (NUMORDER 3) routines compiled from /root/reference/pkg/inc/cubesphere
! (see oracle/ref/Makefile, libcsref3.so).  Test infrastructure only: used
! by the CPU tests to pin the oracle's restatement; never linked into the
! product.
module cs_ref_bind3
  use iso_c_binding
  implicit none

  ! the layout of or_p1_patch (oracle/cs_oracle.h) and cs_p1_patch
  ! (include/cubesphere_hip.h); at third order U, V, B1..3_zc and E*_zc
  ! hold ten modes
  type, bind(C) :: ref3_patch
    integer(c_int32_t) :: ntp, nr, nghost, itri(4), pad
    real(c_double) :: gamma, molewt, rhofloor, prsfloor, smallnum, flatten_coef, prs_flatten_coef
    type(c_ptr) :: vert, mask, stencil, xf_normal, xf_area, xf_centroid, yf_normal, yf_area, yf_centroid
    type(c_ptr) :: zf_area, zf_centroid, xb_pts, xb_wts, xb_normal, yb_pts, yb_wts, yb_normal
    type(c_ptr) :: zb_pts, zb_wts, zb_normal, scale, v_centroid, U, ent, Bxf, Byf, Bzf
    type(c_ptr) :: V, B1_zc, B2_zc, B3_zc, Ex_zc, Ey_zc, Ez_zc
    type(c_ptr) :: gamma_eos, gamma_sound, mean_mol_wt, divvel, msonic, flatten, work
    type(c_ptr) :: Bxf_q, Byf_q, Bzf_q, V_min_local, V_max_local
  end type
  ! the arrays LIMIT_ALL_SLOPES_P12 declares beyond those: the area
  ! quadrature points and weights, the z-face normals at the points, the
  ! volumes, the ten basis moments and Wzc
  type, bind(C) :: ref3_extra
    type(c_ptr) :: xf_pts, xf_wts, yf_pts, yf_wts, zf_pts, zf_wts, zf_normal, v_volume, moments, Wzc
  end type
contains

  ! COORDS_TO_WT_QUADPTS_VOLUME_ISO2 (240coords_to_wt_quadpts_volume_iso2.f:7)
  ! at NUMORDER 3 (27 points), six sectors.
  subroutine ref3_quad_volume(n, nr, g, vert, mask, pts, wts, tang, vol, cen) bind(C, name="ref3_quad_volume")
    integer(c_int), value :: n, nr, g
    real(c_double), intent(in) :: vert(0-g:n+g, 0-g:n+g, 0-g:nr+g, 3, 6)
    integer(c_int), intent(in) :: mask(1-g:n+g, 1-g:n+g, 1-g:nr+g, 6)
    real(c_double), intent(out) :: pts(1-g:n+g, 1-g:n+g, 1-g:nr+g, 3, 27, 6), wts(1-g:n+g, 1-g:n+g, 1-g:nr+g, 27, 6)
    real(c_double), intent(out) :: tang(1-g:n+g, 1-g:n+g, 1-g:nr+g, 3, 3, 27, 6)
    real(c_double), intent(out) :: vol(1-g:n+g, 1-g:n+g, 1-g:nr+g, 6), cen(1-g:n+g, 1-g:n+g, 1-g:nr+g, 3, 6)
    integer :: s
    external COORDS_TO_WT_QUADPTS_VOLUME_ISO2, REF_INIT_COMMON
    call REF_INIT_COMMON()
    do s = 1, 6
      call COORDS_TO_WT_QUADPTS_VOLUME_ISO2(nr, n, g, vert(:, :, :, :, s), mask(:, :, :, s), &
           pts(:, :, :, :, :, s), wts(:, :, :, :, s), tang(:, :, :, :, :, :, s), vol(:, :, :, s), cen(:, :, :, :, s))
    end do
  end subroutine

  ! BUILD_MODE_INDICES (204build_mode_indices.f:7) and INIT_BASIS_FUNCTION_COEFS
  ! (250init_basis_function_coefs.f:5) at third order: ten modes.
  subroutine ref3_basis_moments(n, nr, g, mask, pts, wts, vol, cen, scl, coef) bind(C, name="ref3_basis_moments")
    integer(c_int), value :: n, nr, g
    integer(c_int), intent(in) :: mask(1-g:n+g, 1-g:n+g, 1-g:nr+g, 6)
    real(c_double), intent(in) :: pts(1-g:n+g, 1-g:n+g, 1-g:nr+g, 3, 27, 6), wts(1-g:n+g, 1-g:n+g, 1-g:nr+g, 27, 6)
    real(c_double), intent(in) :: vol(1-g:n+g, 1-g:n+g, 1-g:nr+g, 6), cen(1-g:n+g, 1-g:n+g, 1-g:nr+g, 3, 6)
    real(c_double), intent(in) :: scl(1-g:n+g, 1-g:n+g, 1-g:nr+g, 6)
    real(c_double), intent(out) :: coef(1-g:n+g, 1-g:n+g, 1-g:nr+g, 10, 6)
    integer :: s, ii(10), jj(10), kk(10)
    external BUILD_MODE_INDICES, INIT_BASIS_FUNCTION_COEFS, REF_INIT_COMMON
    call REF_INIT_COMMON()
    call BUILD_MODE_INDICES(10, ii, jj, kk)
    do s = 1, 6
      call INIT_BASIS_FUNCTION_COEFS(nr, n, g, 10, ii, jj, kk, mask(:, :, :, s), pts(:, :, :, :, :, s), &
           wts(:, :, :, :, s), vol(:, :, :, s), cen(:, :, :, :, s), scl(:, :, :, s), coef(:, :, :, :, s))
    end do
  end subroutine

  ! RECONSTRUCTION_COEFS (251reconstruction_coefs.f:3) at order 3: one row.
  subroutine ref3_reconstruction_row(x, y, z, l0, lj, c0, cj, row) bind(C, name="ref3_reconstruction_row")
    real(c_double), value :: x, y, z, l0, lj
    real(c_double), intent(in) :: c0(10), cj(10)
    real(c_double), intent(out) :: row(9)
    real(c_double) :: a(10), b(10)
    external RECONSTRUCTION_COEFS
    a = c0
    b = cj
    call RECONSTRUCTION_COEFS(3, 10, x, y, z, l0, lj, a, b, row)
  end subroutine

  ! MATINV (121matinv.f:8) of an n x n column-major matrix (destroyed).
  subroutine ref3_matinv(n, m, minv) bind(C, name="ref3_matinv")
    integer(c_int), value :: n
    real(c_double), intent(inout) :: m(n, n)
    real(c_double), intent(out) :: minv(n, n)
    integer :: indx(n)
    real(c_double) :: vv(n)
    external MATINV
    call MATINV(n, m, minv, indx, vv)
  end subroutine

  ! LIMIT_ALL_SLOPES_P12 (421limit_all_slopes_p12.f:9) on one patch.
  subroutine ref3_limit_all_slopes_p12(p, e) bind(C, name="ref3_limit_all_slopes_p12")
    type(ref3_patch), intent(in) :: p
    type(ref3_extra), intent(in) :: e
    integer :: n, nr, g, nz, nzr, nf, nfr, nzone, nxf, nzf, nfc, nv
    integer :: ii(10), jj(10), kk(10)
    real(c_double), pointer :: vert(:), xfn(:), xfa(:), xfc(:), yfn(:), yfa(:), yfc(:), zfa(:), zfc(:)
    real(c_double), pointer :: xbp(:), xbw(:), xbn(:), ybp(:), ybw(:), ybn(:), zbp(:), zbw(:), zbn(:)
    real(c_double), pointer :: scl(:), cen(:), u(:), ent(:), bxf(:), byf(:), bzf(:), v(:)
    real(c_double), pointer :: b1(:), b2(:), b3(:), ex(:), ey(:), ez(:), geos(:), gsnd(:), mmw(:)
    real(c_double), pointer :: divv(:), msn(:), flat(:), work(:), bxq(:), byq(:), bzq(:), vmin(:), vmax(:)
    real(c_double), pointer :: xq(:), xw(:), yq(:), yw(:), zq(:), zw(:), zn(:), vol(:), coef(:), wzc(:)
    integer(c_int), pointer :: mask(:), st(:)
    external LIMIT_ALL_SLOPES_P12, REF_INIT_COMMON, REF_SET_P1_COMMON, BUILD_MODE_INDICES
    n = p%ntp; nr = p%nr; g = p%nghost
    nz = n + 2 * g; nzr = nr + 2 * g; nf = nz + 1; nfr = nzr + 1
    nzone = nz * nz * nzr; nxf = nf * nz * nzr; nzf = nz * nz * nfr; nfc = nf * nf * nfr; nv = nfc
    call REF_INIT_COMMON()
    call REF_SET_P1_COMMON(p%gamma, p%molewt, p%rhofloor, p%prsfloor, p%smallnum, p%flatten_coef, &
                           p%prs_flatten_coef)
    call BUILD_MODE_INDICES(10, ii, jj, kk)
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
    call c_f_pointer(p%U, u, [80 * nzone]); call c_f_pointer(p%ent, ent, [4 * nzone])
    call c_f_pointer(p%Bxf, bxf, [nfc]); call c_f_pointer(p%Byf, byf, [nfc]); call c_f_pointer(p%Bzf, bzf, [nfc])
    call c_f_pointer(p%V, v, [80 * nzone])
    call c_f_pointer(p%B1_zc, b1, [10 * nzone]); call c_f_pointer(p%B2_zc, b2, [10 * nzone])
    call c_f_pointer(p%B3_zc, b3, [10 * nzone])
    call c_f_pointer(p%Ex_zc, ex, [10 * nzone]); call c_f_pointer(p%Ey_zc, ey, [10 * nzone])
    call c_f_pointer(p%Ez_zc, ez, [10 * nzone])
    call c_f_pointer(p%gamma_eos, geos, [nzone]); call c_f_pointer(p%gamma_sound, gsnd, [nzone])
    call c_f_pointer(p%mean_mol_wt, mmw, [nzone]); call c_f_pointer(p%divvel, divv, [nzone])
    call c_f_pointer(p%msonic, msn, [nzone]); call c_f_pointer(p%flatten, flat, [nzone])
    call c_f_pointer(p%work, work, [nzone])
    call c_f_pointer(p%Bxf_q, bxq, [5 * nfc]); call c_f_pointer(p%Byf_q, byq, [5 * nfc])
    call c_f_pointer(p%Bzf_q, bzq, [5 * nfc])
    call c_f_pointer(p%V_min_local, vmin, [9]); call c_f_pointer(p%V_max_local, vmax, [9])
    call c_f_pointer(e%xf_pts, xq, [15 * nxf]); call c_f_pointer(e%xf_wts, xw, [5 * nxf])
    call c_f_pointer(e%yf_pts, yq, [15 * nxf]); call c_f_pointer(e%yf_wts, yw, [5 * nxf])
    call c_f_pointer(e%zf_pts, zq, [15 * nzf]); call c_f_pointer(e%zf_wts, zw, [5 * nzf])
    call c_f_pointer(e%zf_normal, zn, [15 * nzf]); call c_f_pointer(e%v_volume, vol, [nzone])
    call c_f_pointer(e%moments, coef, [10 * nzone]); call c_f_pointer(e%Wzc, wzc, [10 * nzone])
    call LIMIT_ALL_SLOPES_P12(nr, n, g, 8, 10, 1, 1, 1, ii, jj, kk, vert, mask, &
         xq, xw, xfn, xfa, xfc, yq, yw, yfn, yfa, yfc, zq, zw, zn, zfa, zfc, &
         xbp, xbw, xbn, ybp, ybw, ybn, zbp, zbw, zbn, scl, vol, cen, coef, &
         p%itri(1), p%itri(2), p%itri(3), p%itri(4), st, &
         u, ent, bxf, byf, bzf, v, b1, b2, b3, ex, ey, ez, wzc, &
         geos, gsnd, mmw, divv, msn, flat, work, bxq, byq, bzq, vmin, vmax)
  end subroutine
end module cs_ref_bind3
