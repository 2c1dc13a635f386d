! fieldio.f90 -- input.inp and raw field files in the reference's layout.
!
! input.inp: 18 list-directed values, one per record, in the order of
! multi/readinput.f90:14-35.  Field files: raw FP64, Fortran order
! (x fastest), nx*ny*nz values, no header, native byte order, as written
! by writefield (multi/readwrite.f90:39-92): ./output/<f>_%08d.dat, and
! read from ./input/<f>.dat (readfield, readwrite.f90:107-207) or from
! ./output/<f>_%08d.dat on restart (readfield_restart, 218-318).
module fieldio
  use iso_c_binding
  use mhit_c
  implicit none
  character(len=8), parameter :: FIELD_TAG(9) = ['u       ', 'v       ', 'w       ', &
      'p       ', 'phi     ', 'rhsu_o  ', 'rhsv_o  ', 'rhsw_o  ', 'rhsphi_o']

contains

  subroutine read_input_inp(path, p)
    character(len=*), intent(in) :: path
    type(mhit_params), intent(inout) :: p
    integer :: u, ios
    integer :: i4
    real(8) :: r8
    integer(8) :: i8
    open(newunit=u, file=path, form='formatted', status='old', action='read', iostat=ios)
    if (ios /= 0) then
      write(*,*) 'ERROR: cannot open ', trim(path)
      error stop 1
    end if
    read(u,*) i4; p%restart = i4
    read(u,*) i4; p%tstart = i4
    read(u,*) i4; p%tfin = i4
    read(u,*) i4; p%dump = i4
    read(u,*) i4; p%inflow = i4
    read(u,*) i4; p%inphi = i4
    read(u,*) r8; p%dt = r8
    read(u,*) r8; p%mu = r8
    read(u,*) r8; p%rho = r8
    read(u,*) r8; p%f1 = r8
    read(u,*) r8; p%f2 = r8
    read(u,*) r8; p%f3 = r8
    read(u,*) r8; p%k0 = r8
    read(u,*) r8; p%radius = r8
    read(u,*) r8; p%sigma = r8
    read(u,*) r8; p%epsr = r8
    read(u,*) i8; p%npart = i8
    read(u,*) i4; p%inpart = i4
    close(u)
  end subroutine

  function field_name(dir, fieldn, t) result(name)
    character(len=*), intent(in) :: dir
    integer, intent(in) :: fieldn, t
    character(len=256) :: name
    if (t < 0) then
      write(name,'(a,a,a,a)') dir, '/', trim(FIELD_TAG(fieldn)), '.dat'
    else
      write(name,'(a,a,a,a,i8.8,a)') dir, '/', trim(FIELD_TAG(fieldn)), '_', t, '.dat'
    end if
  end function

  subroutine write_field(ctx, fieldn, t, n, buf)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: fieldn, t, n
    real(c_double), intent(inout) :: buf(n, n, n)
    integer :: u
    call check(mhit_download_field(ctx, int(fieldn, c_int), buf), 'mhit_download_field')
    open(newunit=u, file=trim(field_name('./output', fieldn, t)), access='stream', &
         form='unformatted', status='replace', action='write')
    write(u) buf
    close(u)
  end subroutine

  ! Reads a field file; a missing or short file is an error (the reference
  ! does not check, readwrite.f90:147-153).
  subroutine read_field(ctx, fieldn, path, n, buf)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: fieldn, n
    character(len=*), intent(in) :: path
    real(c_double), intent(inout) :: buf(n, n, n)
    integer :: u, ios
    integer(8) :: fsize
    inquire(file=path, size=fsize)
    if (fsize /= int(n, 8)**3 * 8_8) then
      write(*,'(a,a,a,i0,a,i0,a)') 'ERROR: ', trim(path), ' has ', fsize, ' bytes, expected ', &
          int(n, 8)**3 * 8_8, ' (missing or wrong nx)'
      error stop 1
    end if
    open(newunit=u, file=path, access='stream', form='unformatted', status='old', action='read', &
         iostat=ios)
    read(u) buf
    close(u)
    call check(mhit_upload_field(ctx, int(fieldn, c_int), buf), 'mhit_upload_field')
  end subroutine

  ! Tracer positions (not written by the reference, which does not restart
  ! tracers): raw FP64 tag, x, y, z columns of npart values each.  With
  ! ab2 = .true. the previous-step velocities follow in partab2_%08d.dat
  ! (up, vp, wp columns) for a restart that continues bit for bit.
  subroutine write_tracers(ctx, t, ab2)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: t
    logical, intent(in) :: ab2
    integer(8) :: np
    real(c_double), allocatable :: tag(:), x(:), y(:), z(:), up(:), vp(:), wp(:)
    character(len=256) :: name
    integer :: u
    np = mhit_tracer_count(ctx)
    if (np <= 0) return
    allocate(tag(np), x(np), y(np), z(np), up(np), vp(np), wp(np))
    call check(mhit_download_tracers(ctx, tag, x, y, z, up, vp, wp), 'mhit_download_tracers')
    write(name,'(a,i8.8,a)') './output/part_', t, '.dat'
    open(newunit=u, file=trim(name), access='stream', form='unformatted', status='replace', &
         action='write')
    write(u) tag, x, y, z
    close(u)
    if (.not. ab2) return
    write(name,'(a,i8.8,a)') './output/partab2_', t, '.dat'
    open(newunit=u, file=trim(name), access='stream', form='unformatted', status='replace', &
         action='write')
    write(u) up, vp, wp
    close(u)
  end subroutine

  ! Reads part_%08d.dat and partab2_%08d.dat back (the tracer count comes
  ! from the file size) and restores the full tracer state.
  subroutine read_tracers(ctx, t)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: t
    integer(8) :: np, fsize, fsize2
    real(c_double), allocatable :: tag(:), x(:), y(:), z(:), up(:), vp(:), wp(:)
    character(len=256) :: name, name2
    integer :: u
    write(name,'(a,i8.8,a)') './output/part_', t, '.dat'
    write(name2,'(a,i8.8,a)') './output/partab2_', t, '.dat'
    inquire(file=name, size=fsize)
    inquire(file=name2, size=fsize2)
    np = fsize / 32_8
    if (fsize <= 0 .or. mod(fsize, 32_8) /= 0 .or. fsize2 /= np * 24_8) then
      write(*,*) 'ERROR: tracer restart files ', trim(name), ' / ', trim(name2), &
          ' are missing or inconsistent'
      error stop 1
    end if
    allocate(tag(np), x(np), y(np), z(np), up(np), vp(np), wp(np))
    open(newunit=u, file=trim(name), access='stream', form='unformatted', status='old', &
         action='read')
    read(u) tag, x, y, z
    close(u)
    open(newunit=u, file=trim(name2), access='stream', form='unformatted', status='old', &
         action='read')
    read(u) up, vp, wp
    close(u)
    call check(mhit_upload_tracer_state(ctx, np, tag, x, y, z, up, vp, wp), &
        'mhit_upload_tracer_state')
  end subroutine

  ! Drop centres in the format of post/unionfind_periodic/main.f90
  ! (write_centers): center_timestep_%08d.dat with "# ID x y z u v w volume"
  ! and one line per drop, plus a line "step count" appended to
  ! drop_count.dat.  Connectivity 6 (faces), threshold 0.5.
  subroutine write_drops(ctx, t)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: t
    integer(c_int64_t), parameter :: maxd = 1000000_c_int64_t
    type(mhit_drop), allocatable :: d(:)
    integer(c_int64_t) :: nd, m
    character(len=256) :: name
    integer :: u
    allocate(d(maxd))
    call check(mhit_drop_stats(ctx, 0.5d0, 6_c_int, maxd, d, nd), 'mhit_drop_stats')
    write(name,'(a,i8.8,a)') './output/center_timestep_', t, '.dat'
    open(newunit=u, file=trim(name), status='replace', action='write')
    write(u,*) '# ID x y z u v w volume'
    do m = 1, min(nd, maxd)
      write(u,'(I8,1X,7E20.10)') m, d(m)%x, d(m)%y, d(m)%z, d(m)%u, d(m)%v, d(m)%w, d(m)%volume
    end do
    close(u)
    open(newunit=u, file='./output/drop_count.dat', status='unknown', position='append', &
         action='write')
    write(u,'(I12,2X,I8)') t, nd
    close(u)
  end subroutine

  ! AB2 coefficients and gamma of the next step: raw FP64 alpha, beta, gumax.
  subroutine write_ab2(t, alpha, beta, gumax)
    integer, intent(in) :: t
    real(c_double), intent(in) :: alpha, beta, gumax
    character(len=256) :: name
    integer :: u
    write(name,'(a,i8.8,a)') './output/ab2_', t, '.dat'
    open(newunit=u, file=trim(name), access='stream', form='unformatted', status='replace', &
         action='write')
    write(u) alpha, beta, gumax
    close(u)
  end subroutine

  logical function read_ab2(t, alpha, beta, gumax)
    integer, intent(in) :: t
    real(c_double), intent(out) :: alpha, beta, gumax
    character(len=256) :: name
    integer :: u
    integer(8) :: fsize
    write(name,'(a,i8.8,a)') './output/ab2_', t, '.dat'
    inquire(file=name, size=fsize)
    read_ab2 = fsize == 24_8
    if (.not. read_ab2) return
    open(newunit=u, file=trim(name), access='stream', form='unformatted', status='old', &
         action='read')
    read(u) alpha, beta, gumax
    close(u)
  end function

end module fieldio
