! ---------------------------------------------------------------------------
! TEST INFRASTRUCTURE ONLY -- never linked into the product library.
!
! C-callable glue around the reference's own compiled hot-path routines
! (hydro/umuscl.f90 `unsplit`, hydro/godunov_utils.f90 `cmpdt`), built from
! the reference sources where they lie by oracle/Makefile into oracle/_ref/.
! The reference keeps its solver parameters in module globals
! (hydro/hydro_parameters.f90:74-85); this file sets them through an explicit
! call so that the C side never binds to compiler-mangled module symbols.
! ---------------------------------------------------------------------------
subroutine ref_set_hydro_params(gamma_in, smallr_in, smallc_in, slope_type_in, &
     & slope_theta_in, riemann_code, courant_in, difmag_in) bind(C, name="ref_set_hydro_params")
  use iso_c_binding
  use amr_parameters
  use hydro_parameters
  implicit none
  real(c_double), value :: gamma_in, smallr_in, smallc_in, slope_theta_in, courant_in, difmag_in
  integer(c_int), value :: slope_type_in, riemann_code
  gamma = gamma_in
  smallr = smallr_in
  smallc = smallc_in
  slope_type = slope_type_in
  slope_theta = slope_theta_in
  courant_factor = courant_in
  difmag = difmag_in
  scheme = 'muscl'
  select case (riemann_code)
  case (0)
     riemann = 'llf'
  case (1)
     riemann = 'hllc'
  case (2)
     riemann = 'hll'
  case (3)
     riemann = 'acoustic'
  case (4)
     riemann = 'exact'
  case default
     riemann = 'llf'
  end select
end subroutine ref_set_hydro_params

! niter_riemann (hydro/hydro_parameters.f90:75): Newton iterations of the
! 'exact' solver.
subroutine ref_set_niter_riemann(n) bind(C, name="ref_set_niter_riemann")
  use iso_c_binding
  use hydro_parameters
  implicit none
  integer(c_int), value :: n
  niter_riemann = n
end subroutine ref_set_niter_riemann

! Shapes of the compiled reference (compile-time constants).
subroutine ref_shape(nvector_out, nvar_out, ndim_out) bind(C, name="ref_shape")
  use iso_c_binding
  use amr_parameters
  use hydro_parameters
  implicit none
  integer(c_int) :: nvector_out, nvar_out, ndim_out
  nvector_out = nvector
  nvar_out = nvar
  ndim_out = ndim
end subroutine ref_shape

! Forwards to the reference `unsplit` (hydro/umuscl.f90:22) with the
! caller's arrays in the reference layout (nvector fastest, column-major).
subroutine ref_unsplit(uin, gravin, pin, flux, tmp, dx, dy, dz, dt, ngrid) bind(C, name="ref_unsplit")
  use iso_c_binding
  use amr_parameters
  use hydro_parameters
  implicit none
  real(c_double), dimension(1:nvector,iu1:iu2,ju1:ju2,ku1:ku2,1:nvar) :: uin
  real(c_double), dimension(1:nvector,iu1:iu2,ju1:ju2,ku1:ku2,1:ndim) :: gravin
  real(c_double), dimension(1:nvector,iu1:iu2,ju1:ju2,ku1:ku2) :: pin
  real(c_double), dimension(1:nvector,if1:if2,jf1:jf2,kf1:kf2,1:nvar,1:ndim) :: flux
  real(c_double), dimension(1:nvector,if1:if2,jf1:jf2,kf1:kf2,1:2,1:ndim) :: tmp
  real(c_double), value :: dx, dy, dz, dt
  integer(c_int), value :: ngrid
  real(dp) :: dxl, dyl, dzl, dtl
  integer :: ng
  dxl = dx; dyl = dy; dzl = dz; dtl = dt; ng = ngrid
  call unsplit(uin, gravin, pin, flux, tmp, dxl, dyl, dzl, dtl, ng)
end subroutine ref_unsplit

! Forwards to the reference `cmpdt` (hydro/godunov_utils.f90:5).  `uu` is
! overwritten in place exactly as the reference does.
subroutine ref_cmpdt(uu, gg, dx, dt, ncell) bind(C, name="ref_cmpdt")
  use iso_c_binding
  use amr_parameters
  use hydro_parameters
  implicit none
  real(c_double), dimension(1:nvector,1:nvar) :: uu
  real(c_double), dimension(1:nvector,1:ndim) :: gg
  real(c_double), value :: dx
  real(c_double) :: dt
  integer(c_int), value :: ncell
  real(dp) :: dxl, dtl
  integer :: nc
  dxl = dx; nc = ncell
  call cmpdt(uu, gg, dxl, dtl, nc)
  dt = dtl
end subroutine ref_cmpdt
