! psod_sod.f90 -- Fortran drop-in for the SOD routines of
! sod/sod_profile_rdfof_parallel_hdf5.f90 ("main file").
!
! The reference passes bulk data through module globals (main file:1-24) and
! calls external subroutines with implicit interfaces.  This file keeps the
! module names, the global variables the routines read and write, and the
! subroutine names and argument lists of cic, cicm1, tri_cic and spherover;
! the bodies call the HIP library through the C ABI of include/psod.h.
! Scope: one zone covering the periodic box (single process, one GPU).

module modvar
  implicit none
  integer, parameter :: idprec = 8
  integer(idprec), dimension(:), allocatable :: idf
  real, dimension(:,:), allocatable :: posf, velf
  real, dimension(:), allocatable :: mpf
  real(8), dimension(:), allocatable :: xsort
  integer, dimension(:), allocatable :: isort
end module modvar

module modcic
  implicit none
  real(8), dimension(:,:,:), allocatable :: cube
  real(8), dimension(:), allocatable :: rho
end module modcic

module psod_capi
  use iso_c_binding
  implicit none
  type(c_ptr), save :: psod_handle = c_null_ptr
  ! when not blank, spherover also writes its records to this file (no
  ! reference counterpart: a full-precision copy of the halo lines)
  character(len=512), save :: psod_record_file = ''

  type, bind(C) :: psod_greedy_params
    integer(c_int32_t) :: nx, ncut
    real(c_double) :: seuil, seuil2, mmin
    integer(c_int64_t) :: nsel
    integer(c_int32_t) :: nbins
    real(c_double) :: rmin_frac, rmax_frac
  end type

  interface
    integer(c_int) function psod_create(device, ctx) bind(C, name="psod_create")
      import :: c_int, c_ptr
      integer(c_int), value :: device
      type(c_ptr) :: ctx
    end function
    type(c_ptr) function psod_last_error() bind(C, name="psod_last_error")
      import :: c_ptr
    end function
    integer(c_int) function psod_host_cic(ctx, npart, pos, mass, xxmin, ng, &
        grid) bind(C, name="psod_host_cic")
      import :: c_int, c_ptr, c_int64_t, c_int32_t, c_float, c_double
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: npart
      real(c_float), intent(in) :: pos(*), mass(*)
      real(c_double), value :: xxmin
      integer(c_int32_t), value :: ng
      real(c_double), intent(out) :: grid(*)
    end function
    integer(c_int) function psod_host_cicm1(ctx, npart, pos, xxmin, ng, grid, &
        rho) bind(C, name="psod_host_cicm1")
      import :: c_int, c_ptr, c_int64_t, c_int32_t, c_float, c_double
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: npart
      real(c_float), intent(in) :: pos(*)
      real(c_double), intent(in) :: grid(*)
      real(c_double), value :: xxmin
      integer(c_int32_t), value :: ng
      real(c_double), intent(out) :: rho(*)
    end function
    integer(c_int) function psod_host_tri_cic(ctx, npart, rho, threshold, &
        isort, npart_new) bind(C, name="psod_host_tri_cic")
      import :: c_int, c_ptr, c_int64_t, c_int32_t, c_float, c_double
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: npart
      real(c_double), intent(in) :: rho(*)
      real(c_double), value :: threshold
      integer(c_int32_t), intent(out) :: isort(*)
      integer(c_int64_t), intent(out) :: npart_new
    end function
    integer(c_int) function psod_host_spherover(ctx, npart, pos, ids, isort, &
        npart_new, p, maxrec, nrec, rec_i64, rec_f64, rad_bins, mass_bins, &
        alive, structure) bind(C, name="psod_host_spherover")
      import :: c_int, c_ptr, c_int64_t, c_int32_t, c_float, c_double, &
          c_int8_t, psod_greedy_params
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: npart, npart_new, maxrec
      real(c_float), intent(in) :: pos(*)
      integer(c_int64_t), intent(in) :: ids(*)
      integer(c_int32_t), intent(in) :: isort(*)
      type(psod_greedy_params), intent(in) :: p
      integer(c_int64_t), intent(out) :: nrec
      integer(c_int64_t), intent(out) :: rec_i64(*)
      real(c_double), intent(out) :: rec_f64(*), rad_bins(*)
      integer(c_int32_t), intent(out) :: mass_bins(*)
      integer(c_int8_t), intent(out) :: alive(*)
      integer(c_int64_t), intent(out) :: structure(*)
    end function
  end interface

contains

  subroutine psod_init()
    integer(c_int) :: rc
    if (c_associated(psod_handle)) return
    rc = psod_create(0_c_int, psod_handle)
    call psod_check(rc, 'psod_create')
    if (rc /= 0) error stop 'psod: no device'
  end subroutine

  ! prints the library's message for a non-zero status
  subroutine psod_report(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    integer :: n
    if (rc == 0) return
    call c_f_pointer(psod_last_error(), msg, [1024])
    n = 0
    do while (n < 1024)
      if (msg(n + 1) == c_null_char) exit
      n = n + 1
    end do
    write(*, '(a,a,i0,a,1024a)') what, ': status ', rc, ': ', msg(1:n)
  end subroutine

  ! a failed cic, cicm1 or tri_cic leaves cube, rho or isort unusable, and
  ! every later step of the zone would read them: print and stop
  subroutine psod_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    if (rc == 0) return
    call psod_report(rc, what)
    error stop 'psod: the zone cannot continue'
  end subroutine

end module psod_capi

! cic (main file:1077-1109): CIC of posf/mpf onto cube(0:local_nx,...)
subroutine cic(npart, xxmin, xxmax, yymin, yymax, zzmin, zzmax, xmin, xmax, &
    ymin, ymax, zmin, zmax, local_nx, local_ny, local_nz, nx, ny, nz, ixzone, &
    iyzone, izzone, nxzone, nyzone, nzzone, periodic)
  use iso_c_binding
  use modvar, only: posf, mpf
  use modcic, only: cube
  use psod_capi
  implicit none
  integer :: npart, local_nx, local_ny, local_nz, nx, ny, nz
  integer :: ixzone, iyzone, izzone, nxzone, nyzone, nzzone
  real(8) :: xxmin, xxmax, yymin, yymax, zzmin, zzmax
  real(8) :: xmin, xmax, ymin, ymax, zmin, zmax
  logical :: periodic
  real(c_double), allocatable :: grid(:,:,:)
  integer(c_int) :: rc
  call psod_init()
  if (.not. periodic .or. nx /= ny .or. nx /= nz .or. local_nx < nx - 1) then
    write(*, *) 'psod cic: only one periodic zone covering the box'
    error stop 'psod: unsupported zone layout'
  end if
  allocate(cube(0:local_nx, 0:local_ny, 0:local_nz))
  cube = 0.d0
  allocate(grid(0:nx-1, 0:ny-1, 0:nz-1))
  rc = psod_host_cic(psod_handle, int(npart, c_int64_t), posf, mpf, &
                     real(xxmin, c_double), int(nx, c_int32_t), grid)
  call psod_check(rc, 'cic')
  cube(0:nx-1, 0:ny-1, 0:nz-1) = grid
  deallocate(grid)
end subroutine cic

! cicm1 (main file:1268-1285): rho(npart) from posf and cube
subroutine cicm1(npart, xxmin, xxmax, yymin, yymax, zzmin, zzmax, local_nx, &
    local_ny, local_nz, nx, ny, nz, periodic)
  use iso_c_binding
  use modvar, only: posf
  use modcic, only: cube, rho
  use psod_capi
  implicit none
  integer :: npart, local_nx, local_ny, local_nz, nx, ny, nz
  real(8) :: xxmin, xxmax, yymin, yymax, zzmin, zzmax
  logical :: periodic
  real(c_double), allocatable :: grid(:,:,:)
  integer(c_int) :: rc
  call psod_init()
  allocate(rho(npart))
  allocate(grid(0:nx-1, 0:ny-1, 0:nz-1))
  grid = cube(0:nx-1, 0:ny-1, 0:nz-1)
  rc = psod_host_cicm1(psod_handle, int(npart, c_int64_t), posf, &
                       real(xxmin, c_double), int(nx, c_int32_t), grid, rho)
  call psod_check(rc, 'cicm1')
  deallocate(grid)
end subroutine cicm1

! tri_cic (main file:998-1006): pdens = sorted 1/rho, isort, npart_new;
! frees rho like the reference (:1023)
subroutine tri_cic(pdens, isort, npart, nx, ny, nz, npart_new, seuil)
  use iso_c_binding
  use modcic, only: rho
  use psod_capi
  implicit none
  integer :: npart, nx, ny, nz, npart_new
  integer, dimension(1:npart) :: isort
  real(8), dimension(1:npart) :: pdens
  real(8) :: seuil
  integer(c_int64_t) :: nn
  integer(c_int) :: rc
  integer :: i
  call psod_init()
  rc = psod_host_tri_cic(psod_handle, int(npart, c_int64_t), rho, &
                         100.0_c_double, isort, nn)
  call psod_check(rc, 'tri_cic')
  npart_new = int(nn)
  do i = 1, npart
    pdens(i) = 1.d0 / rho(isort(i))
  end do
  deallocate(rho)
end subroutine tri_cic

! spherover (main file:1396-1400): greedy SOD; positions in posf are already
! scaled by `scale` (main file:689-690); prints one line per accepted halo
! with the fields of the reference's format 997 (main file:2132-2133)
subroutine spherover(nchar, npart, nx, ny, nz, npart_new, ncut, seuil, &
    seuil2, Mmin, scale, nparttotal, nzone, xminzone, xmaxzone, yminzone, &
    ymaxzone, zminzone, zmaxzone, xmincenter, xmaxcenter, ymincenter, &
    ymaxcenter, zmincenter, zmaxcenter, nsel, xxmin, xxmax, yymin, yymax, &
    zzmin, zzmax, local_nx, local_ny, local_nz, periodic)
  use iso_c_binding
  use modvar, only: posf, idf, isort
  use psod_capi
  implicit none
  character(len=5) :: nchar
  integer :: npart, nx, ny, nz, npart_new, ncut, nzone, nsel
  integer :: local_nx, local_ny, local_nz
  integer(8) :: nparttotal
  real(8) :: seuil, seuil2, Mmin, scale
  real(8) :: xminzone, xmaxzone, yminzone, ymaxzone, zminzone, zmaxzone
  real(8) :: xmincenter, xmaxcenter, ymincenter, ymaxcenter, zmincenter
  real(8) :: zmaxcenter, xxmin, xxmax, yymin, yymax, zzmin, zzmax
  logical :: periodic
  type(psod_greedy_params) :: p
  real(c_float), allocatable :: pos(:,:)
  integer(c_int64_t), allocatable :: ri(:,:), structure(:)
  real(c_double), allocatable :: rf(:,:), rb(:,:)
  integer(c_int32_t), allocatable :: mb(:,:)
  integer(c_int8_t), allocatable :: alive(:)
  integer(c_int64_t) :: nrec, maxrec
  integer(c_int) :: rc
  integer :: k, nacc, u
  call psod_init()
  p%nx = nx
  p%ncut = ncut
  p%seuil = seuil
  p%seuil2 = seuil2
  p%mmin = Mmin
  p%nsel = nsel
  p%nbins = 60
  p%rmin_frac = 0.01d0
  p%rmax_frac = 1.d0
  maxrec = max(int(npart_new, c_int64_t), 1_c_int64_t)
  allocate(pos(3, npart), ri(6, maxrec), rf(8, maxrec), rb(60, maxrec), &
           mb(60, maxrec), alive(npart), structure(npart))
  pos = real(posf / scale, c_float)
  rc = psod_host_spherover(psod_handle, int(npart, c_int64_t), pos, idf, &
                           isort, int(npart_new, c_int64_t), p, maxrec, nrec, &
                           ri, rf, rb, mb, alive, structure)
  ! the reference drops the rest of the zone with a message when the gather
  ! overflows Nsel or a sphere does not close (main file:1681-1693,
  ! 2072-2089), after the halos of the earlier seeds were written; any other
  ! failure stops
  if (rc == 3 .or. rc == 4) then
    call psod_report(rc, 'spherover')
  else
    call psod_check(rc, 'spherover')
  end if
  nacc = 0
  do k = 1, int(nrec)
    if (rf(8, k) > 0.5d0) then
      nacc = nacc + 1
      ! ipart, halo id, masse, masse2, M/scale^3, centre/scale, R/scale,
      ! overdensities (main file:2132-2133, 2434)
      write(*, '(i10,i14,i9,i9,e12.5,3f11.6,f11.6,2f10.2)') ri(1, k), &
          ri(4, k), ri(5, k), ri(6, k), rf(1, k) / scale**3, &
          rf(2, k) / scale, rf(3, k) / scale, rf(4, k) / scale, &
          rf(5, k) / scale, rf(6, k), rf(7, k)
    end if
  end do
  write(*, '(a,i0,a,i0,a)') 'psod spherover: ', nacc, ' halos accepted of ', &
      nrec, ' grown spheres'
  ! every grown sphere, full precision, when the caller named a file:
  ! int64 nrec, then per sphere int64 (ipart, seed, centre, iamas, masse,
  ! masse2) and real(8) (amass, xc, yc, zc, rayon, overdens, overdens2,
  ! accepted), positions in cells
  if (len_trim(psod_record_file) > 0) then
    open(newunit=u, file=trim(psod_record_file), access='stream', &
         form='unformatted', status='replace')
    write(u) nrec
    do k = 1, int(nrec)
      write(u) ri(:, k), rf(:, k)
    end do
    close(u)
  end if
  deallocate(pos, ri, rf, rb, mb, alive, structure)
end subroutine spherover
