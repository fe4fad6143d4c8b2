! ---------------------------------------------------------------------------
! TEST INFRASTRUCTURE ONLY -- never linked into the product library.
!
! C-callable glue that installs an oct tree built by the test driver into
! the reference's own AMR module globals (amr/amr_commons.f90,
! hydro/hydro_commons.f90) and runs the reference's compiled level routines
! on it: set_unew / godunov_fine / set_uold (hydro/godunov_fine.f90),
! courant_fine (hydro/courant_fine.f90) and upload_fine
! (hydro/interpol_hydro.f90).  Single domain (ncpu=1), nx=ny=nz=1.
! ---------------------------------------------------------------------------
subroutine ref_amr_init(nlevelmax_in, ngridmax_in) bind(C, name="ref_amr_init")
  use iso_c_binding
  use amr_commons
  use hydro_commons
  implicit none
  integer(c_int), value :: nlevelmax_in, ngridmax_in
  integer :: il, ncell
  nx = 1; ny = 1; nz = 1
  icoarse_min = 0; icoarse_max = 0
  jcoarse_min = 0; jcoarse_max = 0
  kcoarse_min = 0; kcoarse_max = 0
  boxlen = 1.0d0
  ncoarse = nx*ny*nz
  ncpu = 1; myid = 1
  nlevelmax = nlevelmax_in
  ngridmax = ngridmax_in
  ncell = ncoarse + twotondim*ngridmax
  if (allocated(son)) deallocate(son, father, nbor, uold, unew, numbtot, active, reception)
  allocate(son(1:ncell), father(1:ngridmax), nbor(1:ngridmax,1:twondim))
  allocate(uold(1:ncell,1:nvar), unew(1:ncell,1:nvar))
  allocate(numbtot(1:10,1:nlevelmax))
  allocate(active(1:nlevelmax), reception(1:ncpu,1:nlevelmax))
  son = 0; father = 0; nbor = 0; uold = 0; unew = 0; numbtot = 0
  do il = 1, nlevelmax
     active(il)%ngrid = 0
     reception(1,il)%ngrid = 0
     nullify(active(il)%igrid)
  end do
  mass_tot = 0; ekin_tot = 0; eint_tot = 0
end subroutine ref_amr_init

subroutine ref_amr_set_tree(son_in, father_in, nbor_in) bind(C, name="ref_amr_set_tree")
  use iso_c_binding
  use amr_commons
  implicit none
  integer(c_int) :: son_in(*), father_in(*), nbor_in(*)
  integer :: i, j, ncell
  ncell = ncoarse + twotondim*ngridmax
  do i = 1, ncell
     son(i) = son_in(i)
  end do
  do i = 1, ngridmax
     father(i) = father_in(i)
  end do
  do j = 1, twondim
     do i = 1, ngridmax
        nbor(i,j) = nbor_in(i + (j-1)*ngridmax)
     end do
  end do
end subroutine ref_amr_set_tree

subroutine ref_amr_set_level(ilevel, ngrid, igrid_in) bind(C, name="ref_amr_set_level")
  use iso_c_binding
  use amr_commons
  implicit none
  integer(c_int), value :: ilevel, ngrid
  integer(c_int) :: igrid_in(*)
  integer :: i
  if (associated(active(ilevel)%igrid)) deallocate(active(ilevel)%igrid)
  allocate(active(ilevel)%igrid(1:max(ngrid,1)))
  do i = 1, ngrid
     active(ilevel)%igrid(i) = igrid_in(i)
  end do
  active(ilevel)%ngrid = ngrid
  numbtot(1,ilevel) = ngrid
end subroutine ref_amr_set_level

subroutine ref_amr_set_interpol(itype, ivar) bind(C, name="ref_amr_set_interpol")
  use iso_c_binding
  use hydro_parameters
  implicit none
  integer(c_int), value :: itype, ivar
  interpol_type = itype
  interpol_var = ivar
end subroutine ref_amr_set_interpol

subroutine ref_amr_put_u(which, u_in) bind(C, name="ref_amr_put_u")
  use iso_c_binding
  use amr_commons
  use hydro_commons
  implicit none
  integer(c_int), value :: which
  real(c_double) :: u_in(*)
  integer :: i, v, ncell
  ncell = ncoarse + twotondim*ngridmax
  do v = 1, nvar
     do i = 1, ncell
        if (which == 0) then
           uold(i,v) = u_in(i + (v-1)*ncell)
        else
           unew(i,v) = u_in(i + (v-1)*ncell)
        end if
     end do
  end do
end subroutine ref_amr_put_u

subroutine ref_amr_get_u(which, u_out) bind(C, name="ref_amr_get_u")
  use iso_c_binding
  use amr_commons
  use hydro_commons
  implicit none
  integer(c_int), value :: which
  real(c_double) :: u_out(*)
  integer :: i, v, ncell
  ncell = ncoarse + twotondim*ngridmax
  do v = 1, nvar
     do i = 1, ncell
        if (which == 0) then
           u_out(i + (v-1)*ncell) = uold(i,v)
        else
           u_out(i + (v-1)*ncell) = unew(i,v)
        end if
     end do
  end do
end subroutine ref_amr_get_u

! set_unew at a level (hydro/godunov_fine.f90:40)
subroutine ref_amr_set_unew(ilevel) bind(C, name="ref_amr_set_unew")
  use iso_c_binding
  implicit none
  integer(c_int), value :: ilevel
  integer :: il
  il = ilevel
  call set_unew(il)
end subroutine ref_amr_set_unew

! godunov_fine at a level with time step dt (hydro/godunov_fine.f90:5)
subroutine ref_amr_godunov(ilevel, dt) bind(C, name="ref_amr_godunov")
  use iso_c_binding
  use amr_commons
  implicit none
  integer(c_int), value :: ilevel
  real(c_double), value :: dt
  integer :: il
  il = ilevel
  dtnew(il) = dt
  call godunov_fine(il)
end subroutine ref_amr_godunov

! set_uold at a level (hydro/godunov_fine.f90:135)
subroutine ref_amr_set_uold(ilevel) bind(C, name="ref_amr_set_uold")
  use iso_c_binding
  implicit none
  integer(c_int), value :: ilevel
  integer :: il
  il = ilevel
  call set_uold(il)
end subroutine ref_amr_set_uold

! upload_fine at a level (hydro/interpol_hydro.f90:5)
subroutine ref_amr_upload(ilevel) bind(C, name="ref_amr_upload")
  use iso_c_binding
  implicit none
  integer(c_int), value :: ilevel
  integer :: il
  il = ilevel
  call upload_fine(il)
end subroutine ref_amr_upload

! courant_fine at a level starting from dt_start (hydro/courant_fine.f90:1);
! returns the level's dt and the mass / total energy / internal energy sums
subroutine ref_amr_courant(ilevel, dt_start, dt, mass, etot, eint) bind(C, name="ref_amr_courant")
  use iso_c_binding
  use amr_commons
  use hydro_commons
  implicit none
  integer(c_int), value :: ilevel
  real(c_double), value :: dt_start
  real(c_double) :: dt, mass, etot, eint
  integer :: il
  il = ilevel
  dtnew(il) = dt_start
  mass_tot = 0; ekin_tot = 0; eint_tot = 0
  call courant_fine(il)
  dt = dtnew(il)
  mass = mass_tot
  etot = ekin_tot
  eint = eint_tot
end subroutine ref_amr_courant
