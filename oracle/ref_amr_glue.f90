! ---------------------------------------------------------------------------
! TEST INFRASTRUCTURE ONLY -- never linked into the product library.
!
! C-callable glue that installs an oct tree built by the test driver into
! the reference's own AMR module globals (amr/amr_commons.f90,
! hydro/hydro_commons.f90) and runs the reference's compiled level routines
! on it: set_unew / godunov_fine / set_uold (hydro/godunov_fine.f90),
! courant_fine (hydro/courant_fine.f90), upload_fine
! (hydro/interpol_hydro.f90), and the mesh refinement flag_fine
! (amr/flag_utils.f90), refine_fine (amr/refine_utils.f90) and build_comm
! (amr/virtual_boundaries.f90).  Single domain (ncpu=1), nx=ny=nz=1; the
! tree arrays are allocated as amr/init_amr.f90 allocates them.
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
  hydro = .true.       ! set by read_params for the hydro solver
  ifout = 1; iout = 1; noutput = 0; t = 0; nstep = 0; nstep_coarse = 0
  nsubcycle = 2
  nlevelmax = nlevelmax_in
  ngridmax = ngridmax_in
  ncell = ncoarse + twotondim*ngridmax
  if (allocated(son)) then
     deallocate(son, father, nbor, uold, unew, numbtot, active, reception)
     deallocate(flag1, flag2, cpu_map, cpu_map2, hilbert_key, xg, next, prev)
     deallocate(headl, taill, numbl, headb, tailb, numbb, boundary, bound_key, bound_key2)
  end if
  allocate(son(1:ncell), father(1:ngridmax), nbor(1:ngridmax,1:twondim))
  allocate(uold(1:ncell,1:nvar), unew(1:ncell,1:nvar))
  allocate(numbtot(1:10,1:nlevelmax))
  allocate(active(1:nlevelmax), reception(1:ncpu,1:nlevelmax))
  allocate(flag1(0:ncell), flag2(0:ncell), cpu_map(1:ncell), cpu_map2(1:ncell), hilbert_key(1:ncell))
  allocate(xg(1:ngridmax,1:ndim), next(1:ngridmax), prev(1:ngridmax))
  allocate(headl(1:ncpu,1:nlevelmax), taill(1:ncpu,1:nlevelmax), numbl(1:ncpu,1:nlevelmax))
  allocate(headb(1:MAXBOUND,1:nlevelmax), tailb(1:MAXBOUND,1:nlevelmax), numbb(1:MAXBOUND,1:nlevelmax))
  allocate(boundary(1:MAXBOUND,1:nlevelmax))
  ndomain = ncpu
  allocate(bound_key(0:ndomain), bound_key2(0:ndomain))
  ordering = 'planar'
  bound_key(0) = -1; bound_key(1) = 1d30
  bound_key2 = bound_key
  son = 0; father = 0; nbor = 0; uold = 0; unew = 0; numbtot = 0
  flag1 = 0; flag2 = 0; cpu_map = 1; cpu_map2 = 0; hilbert_key = 0; xg = 0; next = 0; prev = 0
  headl = 0; taill = 0; numbl = 0; headb = 0; tailb = 0; numbb = 0
  headf = 0; tailf = 0; numbf = 0; used_mem = 0
  do il = 1, nlevelmax
     active(il)%ngrid = 0
     reception(1,il)%ngrid = 0
     nullify(active(il)%igrid)
     boundary(1:MAXBOUND,il)%ngrid = 0
  end do
  mass_tot = 0; ekin_tot = 0; eint_tot = 0
  dtold = 0; dtnew = 0     ! a fresh run's level steps (module zeros)
  ! update_time's conservation baselines and step bookkeeping as a fresh
  ! run starts them (amr/amr_commons.f90:10-22)
  nstep_coarse_old = 0
  epot_tot = 0; epot_tot_old = 0; epot_tot_int = 0; einit = 0; aexp_old = 1
  mass_tot_0 = 0
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
  numbtot(1:4,ilevel) = ngrid   ! as refine_fine leaves them with one domain
  ! the level's linked list in the same order (amr/amr_commons.f90 headl/taill)
  numbl(1,ilevel) = ngrid
  headl(1,ilevel) = 0; taill(1,ilevel) = 0
  if (ngrid > 0) then
     headl(1,ilevel) = igrid_in(1)
     taill(1,ilevel) = igrid_in(ngrid)
     do i = 1, ngrid
        if (i > 1) then
           prev(igrid_in(i)) = igrid_in(i-1)
        else
           prev(igrid_in(i)) = 0
        end if
        if (i < ngrid) then
           next(igrid_in(i)) = igrid_in(i+1)
        else
           next(igrid_in(i)) = 0
        end if
     end do
  end if
end subroutine ref_amr_set_level

! The free-grid list headf..tailf in the given order.
subroutine ref_amr_set_free(nfree, igrid_in) bind(C, name="ref_amr_set_free")
  use iso_c_binding
  use amr_commons
  implicit none
  integer(c_int), value :: nfree
  integer(c_int) :: igrid_in(*)
  integer :: i
  numbf = nfree
  used_mem = ngridmax - numbf
  used_mem_tot = used_mem
  headf = 0; tailf = 0
  if (nfree > 0) then
     headf = igrid_in(1)
     tailf = igrid_in(nfree)
     do i = 1, nfree
        if (i > 1) then
           prev(igrid_in(i)) = igrid_in(i-1)
        else
           prev(igrid_in(i)) = 0
        end if
        if (i < nfree) then
           next(igrid_in(i)) = igrid_in(i+1)
        else
           next(igrid_in(i)) = 0
        end if
     end do
  end if
end subroutine ref_amr_set_free

! Refinement criteria (hydro/hydro_parameters.f90:47-52, amr/amr_parameters.f90:85)
subroutine ref_amr_set_refine(egd, egu, egp, fd, fu, fp, nexp, lmin) bind(C, name="ref_amr_set_refine")
  use iso_c_binding
  use amr_commons
  use hydro_commons
  implicit none
  real(c_double), value :: egd, egu, egp, fd, fu, fp
  integer(c_int), value :: nexp, lmin
  err_grad_d = egd; err_grad_u = egu; err_grad_p = egp
  floor_d = fd; floor_u = fu; floor_p = fp
  nexpand = nexp
  levelmin = lmin
end subroutine ref_amr_set_refine

subroutine ref_amr_flag_fine(ilevel, icount) bind(C, name="ref_amr_flag_fine")
  use iso_c_binding
  implicit none
  integer(c_int), value :: ilevel, icount
  integer :: il, ic
  il = ilevel; ic = icount
  call flag_fine(il, ic)
end subroutine ref_amr_flag_fine

! refine_fine(ilevel), then build_comm(ilevel+1) as amr_step does before it
! touches the next level (amr/amr_step.f90:40-93).
subroutine ref_amr_refine_fine(ilevel) bind(C, name="ref_amr_refine_fine")
  use iso_c_binding
  use amr_commons
  implicit none
  integer(c_int), value :: ilevel
  integer :: il
  il = ilevel
  call refine_fine(il)
  if (il + 1 <= nlevelmax) then
     il = ilevel + 1
     call build_comm(il)
  end if
end subroutine ref_amr_refine_fine

! Current grid list of a level; returns its length (out must hold ngridmax).
function ref_amr_get_active(ilevel, out) bind(C, name="ref_amr_get_active") result(n)
  use iso_c_binding
  use amr_commons
  implicit none
  integer(c_int), value :: ilevel
  integer(c_int) :: out(*)
  integer(c_int) :: n
  integer :: i
  n = active(ilevel)%ngrid
  do i = 1, n
     out(i) = active(ilevel)%igrid(i)
  end do
end function ref_amr_get_active

! The free list from headf; returns numbf.
function ref_amr_get_free(out) bind(C, name="ref_amr_get_free") result(n)
  use iso_c_binding
  use amr_commons
  implicit none
  integer(c_int) :: out(*)
  integer(c_int) :: n
  integer :: i, ig
  n = numbf
  ig = headf
  do i = 1, numbf
     out(i) = ig
     ig = next(ig)
  end do
end function ref_amr_get_free

subroutine ref_amr_get_tree(son_out, father_out, nbor_out, flag1_out) bind(C, name="ref_amr_get_tree")
  use iso_c_binding
  use amr_commons
  implicit none
  integer(c_int) :: son_out(*), father_out(*), nbor_out(*), flag1_out(*)
  integer :: i, j, ncell
  ncell = ncoarse + twotondim*ngridmax
  do i = 1, ncell
     son_out(i) = son(i)
     flag1_out(i) = flag1(i)
  end do
  do i = 1, ngridmax
     father_out(i) = father(i)
  end do
  do j = 1, twondim
     do i = 1, ngridmax
        nbor_out(i + (j-1)*ngridmax) = nbor(i,j)
     end do
  end do
end subroutine ref_amr_get_tree

! flag1 of every cell from the caller (the refinement map to refine from)
subroutine ref_amr_set_flag1(flag1_in) bind(C, name="ref_amr_set_flag1")
  use iso_c_binding
  use amr_commons
  implicit none
  integer(c_int) :: flag1_in(*)
  integer :: i, ncell
  ncell = ncoarse + twotondim*ngridmax
  flag1(0) = 0
  do i = 1, ncell
     flag1(i) = flag1_in(i)
  end do
end subroutine ref_amr_set_flag1

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

! Grid centres xg(igrid,1:3) from the caller (xg_in(3*(igrid-1)+d)), and back.
subroutine ref_amr_set_xg(xg_in) bind(C, name="ref_amr_set_xg")
  use iso_c_binding
  use amr_commons
  implicit none
  real(c_double) :: xg_in(*)
  integer :: i, d
  do i = 1, ngridmax
     do d = 1, ndim
        xg(i,d) = xg_in(3*(i-1)+d)
     end do
  end do
end subroutine ref_amr_set_xg

subroutine ref_amr_get_xg(xg_out) bind(C, name="ref_amr_get_xg")
  use iso_c_binding
  use amr_commons
  implicit none
  real(c_double) :: xg_out(*)
  integer :: i, d
  do i = 1, ngridmax
     do d = 1, ndim
        xg_out(3*(i-1)+d) = xg(i,d)
     end do
  end do
end subroutine ref_amr_get_xg

! The Sedov set-up of namelist/sedov3d.nml (&INIT_PARAMS): a 'square' region
! covering the box with (d_amb, p_amb) and a 'point' region at (xc,yc,zc)
! depositing pressure p_point (hydro/init_flow_fine.f90:475-596).
subroutine ref_amr_set_sedov(d_amb, p_amb, p_point, xc, yc, zc) bind(C, name="ref_amr_set_sedov")
  use iso_c_binding
  use amr_commons
  use hydro_commons
  implicit none
  real(c_double), value :: d_amb, p_amb, p_point, xc, yc, zc
  nregion = 2
  region_type(1) = 'square'
  region_type(2) = 'point'
  x_center(1) = 0.5d0*boxlen; y_center(1) = 0.5d0*boxlen; z_center(1) = 0.5d0*boxlen
  x_center(2) = xc; y_center(2) = yc; z_center(2) = zc
  length_x(1:2) = (/10.0d0*boxlen, 1.0d0/)
  length_y(1:2) = (/10.0d0*boxlen, 1.0d0/)
  length_z(1:2) = (/10.0d0*boxlen, 1.0d0/)
  exp_region(1:2) = 10.0d0
  d_region(1:2) = (/d_amb, 0.0d0/)
  u_region(1:2) = 0.0d0; v_region(1:2) = 0.0d0; w_region(1:2) = 0.0d0
  p_region(1:2) = (/p_amb, p_point/)
end subroutine ref_amr_set_sedov

subroutine ref_amr_set_init(flag) bind(C, name="ref_amr_set_init")
  use iso_c_binding
  use amr_commons
  implicit none
  integer(c_int), value :: flag
  init = (flag /= 0)
end subroutine ref_amr_set_init

! init_flow (hydro/init_flow_fine.f90:5-23): condinit on every level
! >= levelmin, finest first, each followed by upload_fine.
subroutine ref_amr_init_flow() bind(C, name="ref_amr_init_flow")
  implicit none
  call init_flow
end subroutine ref_amr_init_flow

subroutine ref_amr_init_flow_fine(ilevel) bind(C, name="ref_amr_init_flow_fine")
  use iso_c_binding
  implicit none
  integer(c_int), value :: ilevel
  integer :: il
  il = ilevel
  call init_flow_fine(il)
end subroutine ref_amr_init_flow_fine

! Level time steps dtold(1:nlevelmax), dtnew(1:nlevelmax)
subroutine ref_amr_set_dt(dto, dtn) bind(C, name="ref_amr_set_dt")
  use iso_c_binding
  use amr_commons
  implicit none
  real(c_double) :: dto(*), dtn(*)
  integer :: i
  do i = 1, nlevelmax
     dtold(i) = dto(i)
     dtnew(i) = dtn(i)
  end do
end subroutine ref_amr_set_dt

subroutine ref_amr_defrag() bind(C, name="ref_amr_defrag")
  implicit none
  call defrag
end subroutine ref_amr_defrag

! The snapshot part of dump_all (amr/output_amr.f90:5-206) for one domain:
! defrag, then info / amr / hydro files into the existing directory dir_c
! (NUL-terminated) with the counters advanced as dump_all advances them.
subroutine ref_amr_dump(dir_c, tnow, nstep_in, nstep_coarse_in) bind(C, name="ref_amr_dump")
  use iso_c_binding
  use amr_commons
  use hydro_commons
  implicit none
  character(kind=c_char), dimension(*) :: dir_c
  real(c_double), value :: tnow
  integer(c_int), value :: nstep_in, nstep_coarse_in
  character(len=80) :: filedir, filename, filename_desc
  character(len=5) :: nchar
  integer :: i
  filedir = ' '
  i = 1
  do while (dir_c(i) /= c_null_char .and. i < 79)
     filedir(i:i) = dir_c(i)
     i = i + 1
  end do
  filedir(i:i) = '/'
  t = tnow
  nstep = nstep_in
  nstep_coarse = nstep_coarse_in
  call defrag
  call flag_coarse
  call title(ifout, nchar)
  ifout = ifout + 1
  if (t >= tout(iout) .or. aexp >= aout(iout)) iout = iout + 1
  filename = TRIM(filedir)//'info_'//TRIM(nchar)//'.txt'
  call output_info(filename)
  filename = TRIM(filedir)//'amr_'//TRIM(nchar)//'.out'
  call backup_amr(filename)
  filename = TRIM(filedir)//'hydro_'//TRIM(nchar)//'.out'
  filename_desc = TRIM(filedir)//'hydro_file_descriptor.txt'
  call backup_hydro(filename, filename_desc)
end subroutine ref_amr_dump

! used_mem / used_mem_tot (amr/amr_commons.f90)
function ref_amr_get_used_mem() bind(C, name="ref_amr_get_used_mem") result(n)
  use iso_c_binding
  use amr_commons
  implicit none
  integer(c_int) :: n
  n = used_mem
end function ref_amr_get_used_mem

! Regions of &INIT_PARAMS (hydro/hydro_parameters.f90, amr/amr_parameters.f90):
! rtype(k) = 1 'square', 2 'point'; arrays of length n.
subroutine ref_set_regions(n, rtype, xc, yc, zc, lx, ly, lz, ex, d, u, v, w, p) bind(C, name="ref_set_regions")
  use iso_c_binding
  use amr_commons
  use hydro_commons
  implicit none
  integer(c_int), value :: n
  integer(c_int) :: rtype(*)
  real(c_double) :: xc(*), yc(*), zc(*), lx(*), ly(*), lz(*), ex(*), d(*), u(*), v(*), w(*), p(*)
  integer :: k
  nregion = n
  do k = 1, n
     if (rtype(k) == 1) then
        region_type(k) = 'square'
     else
        region_type(k) = 'point'
     end if
     x_center(k) = xc(k); y_center(k) = yc(k); z_center(k) = zc(k)
     length_x(k) = lx(k); length_y(k) = ly(k); length_z(k) = lz(k)
     exp_region(k) = ex(k)
     d_region(k) = d(k); u_region(k) = u(k); v_region(k) = v(k); w_region(k) = w(k); p_region(k) = p(k)
  end do
end subroutine ref_set_regions

! condinit (hydro/condinit.f90) on nn <= nvector points: x(nvector,3) in,
! uu(nvector,nvar) out, both column-major.
subroutine ref_condinit(nn, x, uu, dx) bind(C, name="ref_condinit")
  use iso_c_binding
  use amr_parameters
  use hydro_parameters
  implicit none
  integer(c_int), value :: nn
  real(c_double) :: x(nvector,3), uu(nvector,nvar)
  real(c_double), value :: dx
  real(dp) :: xx(nvector,ndim), u(nvector,nvar), dxl
  integer :: n
  xx = x(:,1:ndim)
  dxl = dx
  n = nn
  u = 0
  call condinit(xx, u, dxl, n)
  uu = u
end subroutine ref_condinit

! nsubcycle(1:n) of &RUN_PARAMS (amr/amr_commons.f90:51)
subroutine ref_amr_set_nsubcycle(n, ns) bind(C, name="ref_amr_set_nsubcycle")
  use iso_c_binding
  use amr_commons
  implicit none
  integer(c_int), value :: n
  integer(c_int) :: ns(*)
  integer :: i
  do i = 1, n
     nsubcycle(i) = ns(i)
  end do
end subroutine ref_amr_set_nsubcycle

subroutine ref_amr_set_boxlen(b) bind(C, name="ref_amr_set_boxlen")
  use iso_c_binding
  use amr_parameters
  implicit none
  real(c_double), value :: b
  boxlen = b
end subroutine ref_amr_set_boxlen

! noutput and tout(1:noutput) of &OUTPUT_PARAMS (amr/amr_parameters.f90:107-119)
subroutine ref_amr_set_outputs(n, t_out) bind(C, name="ref_amr_set_outputs")
  use iso_c_binding
  use amr_parameters
  implicit none
  integer(c_int), value :: n
  real(c_double) :: t_out(*)
  integer :: i
  noutput = n
  do i = 1, n
     tout(i) = t_out(i)
  end do
end subroutine ref_amr_set_outputs

subroutine ref_amr_flag_coarse() bind(C, name="ref_amr_flag_coarse")
  implicit none
  call flag_coarse
end subroutine ref_amr_flag_coarse

! How many octs refine_fine(ilevel) would create (leaf cells with flag1 == 1),
! and the free-grid count: lets a caller refuse a refinement that would make
! the reference stop with "No more free memory" (amr/refine_utils.f90:415-423).
function ref_amr_count_create(ilevel) bind(C, name="ref_amr_count_create") result(n)
  use iso_c_binding
  use amr_commons
  implicit none
  integer(c_int), value :: ilevel
  integer(c_int) :: n
  integer :: i, ind, c
  n = 0
  if (ilevel < 1 .or. ilevel > nlevelmax) return
  do ind = 1, twotondim
     do i = 1, active(ilevel)%ngrid
        c = ncoarse + (ind-1)*ngridmax + active(ilevel)%igrid(i)
        if (flag1(c) == 1 .and. son(c) == 0) n = n + 1
     end do
  end do
end function ref_amr_count_create

function ref_amr_numbf() bind(C, name="ref_amr_numbf") result(n)
  use iso_c_binding
  use amr_commons
  implicit none
  integer(c_int) :: n
  n = numbf
end function ref_amr_numbf

! The restart of amr/adaptive_loop.f90:32-34 for one domain: init_amr reads
! output_NNNNN/amr_NNNNN.out00001 (amr/init_amr.f90:261-551: tree, level
! lists, free list, times, level dts; build_comm per level) and init_hydro
! reads hydro_NNNNN.out00001 (hydro/init_hydro.f90:56-250: primitive ->
! conserved), both relative to the current directory.  The arrays the glue
! or an earlier restart allocated are released first, since both routines
! allocate them.  The caller checks the files first: the reference stops
! the process on an incompatible file.
subroutine ref_amr_restart(nlevelmax_in, ngridmax_in, nrestart_in) bind(C, name="ref_amr_restart")
  use iso_c_binding
  use amr_commons
  use hydro_commons
  implicit none
  integer(c_int), value :: nlevelmax_in, ngridmax_in, nrestart_in
  nx = 1; ny = 1; nz = 1
  icoarse_min = 0; icoarse_max = 0
  jcoarse_min = 0; jcoarse_max = 0
  kcoarse_min = 0; kcoarse_max = 0
  ncpu = 1; myid = 1
  hydro = .true.
  ordering = 'planar'
  nlevelmax = nlevelmax_in
  ngridmax = ngridmax_in
  nrestart = nrestart_in
  if (allocated(son)) then
     deallocate(son, father, nbor, uold, unew, numbtot, active, reception)
     deallocate(flag1, flag2, cpu_map, cpu_map2, hilbert_key, xg, next, prev)
     deallocate(headl, taill, numbl, headb, tailb, numbb, boundary, bound_key, bound_key2)
  end if
  if (allocated(emission)) deallocate(emission)
  call init_amr
  call init_hydro
  nrestart = 0
  nstep_coarse_old = nstep_coarse   ! as adaptive_loop sets it before its loop (amr/adaptive_loop.f90:75)
end subroutine ref_amr_restart

! Run state after a restart: t, nstep, nstep_coarse, ifout, iout, used_mem
! and the level dts dtold/dtnew(1:n).
subroutine ref_amr_get_run(n, tout_, ints, dtold_out, dtnew_out) bind(C, name="ref_amr_get_run")
  use iso_c_binding
  use amr_commons
  implicit none
  integer(c_int), value :: n
  real(c_double) :: tout_
  integer(c_int) :: ints(5)
  real(c_double) :: dtold_out(n), dtnew_out(n)
  tout_ = t
  ints(1) = nstep; ints(2) = nstep_coarse; ints(3) = ifout; ints(4) = iout; ints(5) = used_mem
  dtold_out(1:n) = dtold(1:n)
  dtnew_out(1:n) = dtnew(1:n)
end subroutine ref_amr_get_run

! One coarse step of the reference's own driver for one domain:
! amr/adaptive_loop.f90:84-194 around amr_step(levelmin,1)
! (amr/amr_step.f90, which calls newdt_fine pm/newdt_fine.f90 and
! update_time amr/update_time.f90): the conservation-sum reset (:88-91),
! refine_coarse and the coarse levels' build_comm / refine_fine (:95-133),
! amr_step, then upload_fine and flag_fine of the coarse levels and
! flag_coarse (:143-191), and nstep_coarse+1 (:194).  The driver writes no
! snapshot here (foutput is 0 during the call; the caller writes outputs
! with ref_amr_dump), ncontrol is raised so that update_time prints only
! at step 0, and with no output times set (noutput 0) update_time's end
! test reads tout(1) = HUGE for the call instead of tout(0).  The coarse
! cell's flag must be set (flag_coarse) before the first call, as the
! reference's init_refine leaves it, or refine_coarse removes level 1.
! Returns the conservation sums accumulated by the step's courant_fine
! calls (mass_tot, ekin_tot, eint_tot).
subroutine ref_amr_coarse_step(sums) bind(C, name="ref_amr_coarse_step")
  use iso_c_binding
  use amr_commons
  use hydro_commons
  implicit none
  real(c_double) :: sums(3)
  integer :: ilevel, fout_save, nctl_save, nout_save
  real(dp) :: tout1_save
  fout_save = foutput; nctl_save = ncontrol; nout_save = noutput; tout1_save = tout(1)
  if (noutput < 1) then
     noutput = 1
     tout(1) = huge(1.0d0)
  end if
  foutput = 0
  ncontrol = huge(ncontrol)
  epot_tot = 0; ekin_tot = 0; mass_tot = 0; eint_tot = 0
  if (levelmin < nlevelmax .and. .not. static) then
     call refine_coarse
     do ilevel = 1, levelmin
        call build_comm(ilevel)
        if (ilevel < levelmin) call refine_fine(ilevel)
     end do
  end if
  call amr_step(levelmin, 1)
  if (levelmin < nlevelmax .and. .not. static) then
     do ilevel = levelmin-1, 1, -1
        call upload_fine(ilevel)
     end do
     do ilevel = levelmin-1, 1, -1
        call flag_fine(ilevel, 2)
     end do
     call flag_coarse
  end if
  nstep_coarse = nstep_coarse + 1
  foutput = fout_save; ncontrol = nctl_save; noutput = nout_save; tout(1) = tout1_save
  sums(1) = mass_tot; sums(2) = ekin_tot; sums(3) = eint_tot
end subroutine ref_amr_coarse_step

! Run clock and level time steps as the reference's driver keeps them:
! t, nstep and dtold/dtnew(1:n) in, for a caller that resumes a run.
subroutine ref_amr_set_clock(tnow, nstep_in, nstep_coarse_in) bind(C, name="ref_amr_set_clock")
  use iso_c_binding
  use amr_commons
  implicit none
  real(c_double), value :: tnow
  integer(c_int), value :: nstep_in, nstep_coarse_in
  t = tnow
  nstep = nstep_in
  nstep_coarse = nstep_coarse_in
  nstep_coarse_old = nstep_coarse_in
end subroutine ref_amr_set_clock

! update_time's conservation baselines: out = (einit, mass_tot_0).
subroutine ref_amr_get_baselines(out) bind(C, name="ref_amr_get_baselines")
  use iso_c_binding
  use amr_commons
  use hydro_commons
  implicit none
  real(c_double) :: out(2)
  out(1) = einit
  out(2) = mass_tot_0
end subroutine ref_amr_get_baselines
