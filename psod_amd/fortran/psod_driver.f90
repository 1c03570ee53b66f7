! psod_driver.f90 -- single-process driver of the SOD pass with the
! reference's flags and call sequence (main file:26-908, zone loop :333-757),
! for one zone covering the periodic box.
!
! Input (-inp FILE): raw little-endian catalogue
!   int64 npart, real(4) posf(3,npart) in [0,1), int64 idf(npart).
! Flags as the reference (main file:811-852): -nx, -min, -del, -dl2, -sel,
! -nse, -nch.  Output: one line per accepted halo (see spherover); -rec FILE
! also writes every grown sphere's record to FILE at full precision.
program psod_driver
  use modvar
  use modcic
  use psod_capi, only: psod_record_file
  implicit none
  character(len=256) :: inp, opt, arg
  character(len=5) :: nchar
  integer :: nx, nmin, ncut, nsel, i, n, npart, npart_new
  real(8) :: seuil, seuil2, scale, Mmin
  real(8) :: xmin, xmax, xminzone, xmaxzone
  integer(8) :: nparttotal, np8
  integer :: nxbuffer, nxzone, local_nx, nxcic, u
  logical :: periodic

  inp = ''
  nx = 0
  nmin = 100
  ncut = 4
  nsel = 4000000
  seuil = 200.d0
  seuil2 = 800.d0
  nchar = '00001'
  n = command_argument_count()
  i = 1
  do while (i < n)
    call get_command_argument(i, opt)
    call get_command_argument(i + 1, arg)
    select case (trim(opt))
    case ('-inp'); inp = arg
    case ('-nx'); read(arg, *) nx
    case ('-min'); read(arg, *) nmin
    case ('-sel'); read(arg, *) ncut
    case ('-nse'); read(arg, *) nsel
    case ('-del'); read(arg, *) seuil
    case ('-dl2'); read(arg, *) seuil2
    case ('-nch'); nchar = arg(1:5)
    case ('-rec'); psod_record_file = arg
    case default; write(*, *) 'unknown option ', trim(opt), ' ignored'
    end select
    i = i + 2
  end do
  if (nx <= 0 .or. len_trim(inp) == 0) then
    write(*, *) 'usage: psod_driver -inp FILE -nx NX [-min 100 -del 200 ', &
        '-dl2 800 -sel 4 -nse 4000000 -nch 00001 -rec FILE]'
    stop 2
  end if

  open(newunit=u, file=trim(inp), access='stream', form='unformatted', &
       status='old')
  read(u) np8
  npart = int(np8)
  allocate(posf(3, npart), idf(npart), mpf(npart))
  read(u) posf
  read(u) idf
  close(u)
  nparttotal = np8
  mpf = 1.0

  ! one zone with the default buffer nxzone/2 (main file:340-413)
  nxzone = nx
  nxbuffer = nxzone / 2
  xmin = -real(nxbuffer, 8) / nxzone
  xmax = 1.d0 + real(nxbuffer, 8) / nxzone
  xminzone = 0.d0
  xmaxzone = 1.d0
  periodic = .true.

  ! CIC density on the 2*nx grid, back to the particles (main file:659-685)
  allocate(xsort(1:npart), isort(1:npart))
  mpf = mpf / real(nparttotal, kind=8)
  nxcic = 2 * nx
  local_nx = 2 * (nxzone + 2 * nxbuffer)
  call cic(npart, xmin, xmax, xmin, xmax, xmin, xmax, xmin, xmax, xmin, xmax, &
      xmin, xmax, local_nx, local_nx, local_nx, nxcic, nxcic, nxcic, 0, 0, 0, &
      1, 1, 1, periodic)
  call cicm1(npart, xmin, xmax, xmin, xmax, xmin, xmax, local_nx, local_nx, &
      local_nx, nxcic, nxcic, nxcic, periodic)
  deallocate(cube)
  mpf = mpf * real(nparttotal, kind=8)

  ! scale positions by nx (main file:689-698)
  scale = real(nx, kind=8)
  posf = posf * scale
  xminzone = xminzone * scale
  xmaxzone = xmaxzone * scale
  Mmin = dble(nmin)

  call tri_cic(xsort, isort, npart, nxcic, nxcic, nxcic, npart_new, seuil)

  local_nx = nxzone + 2 * nxbuffer
  call spherover(nchar, npart, nx, nx, nx, npart_new, ncut, seuil, seuil2, &
      Mmin, scale, nparttotal, 0, xminzone, xmaxzone, xminzone, xmaxzone, &
      xminzone, xmaxzone, xminzone, xmaxzone, xminzone, xmaxzone, xminzone, &
      xmaxzone, nsel, xmin, xmax, xmin, xmax, xmin, xmax, local_nx, local_nx, &
      local_nx, periodic)
  deallocate(posf, idf, mpf, xsort, isort)
end program psod_driver
