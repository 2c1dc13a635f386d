! main.f90 -- MHIT36 driver on one MI355X.
!
! Keeps the reference's program flow (multi/main.f90): read input.inp,
! initialise velocity / phase field / tracers (fresh start, input files or
! restart), save the initial fields, then loop t = tstart+1 .. tfin with
! the timestep on the GPU, the CFL print/stop and the periodic dumps.
!
! What the reference fixes at compile time is given on the command line:
!   mhit36 [--nx N] [--phi 0|1] [--part 0|1] [--fp64-poisson] [--k2-modified]
!          [--trilinear] [--seed S] [--exact-restart] [--drops] [--quiet]
! --exact-restart also dumps the AB2 history (rhs*_o fields, tracer
! velocities, alpha/beta/gumax) and, on restart, reads it back so that the
! run continues bit for bit instead of restarting with an Euler step.
! --drops writes the drop statistics of post/unionfind_periodic at every
! dump (output/center_timestep_%08d.dat, output/drop_count.dat).
! (defaults: nx = 256 as multi/module.f90:2, phiflag = partflag = 0 as
! multi/main.f90:55-57, FP32 Poisson as the mixedp branch).
program mhit36
  use iso_c_binding
  use mhit_c
  use fieldio
  implicit none

  type(mhit_params) :: p
  type(mhit_diag) :: d
  type(c_ptr) :: ctx
  integer :: n, t, tstart, rc, verbose, f
  logical :: exact, drops
  real(c_double) :: alpha, beta, gumax
  integer(8) :: seed, count0, count1, rate, countt
  real(c_double), allocatable :: buf(:,:,:)
  character(len=64) :: arg

  call check(mhit_params_default(p), 'mhit_params_default')
  call read_input_inp('input.inp', p)
  seed = 1_8
  verbose = 1
  exact = .false.
  drops = .false.
  call parse_args()
  n = p%nx

  write(*,*) '----------------------------------------------'
  write(*,*) ' MHIT36 on MI355X (one GPU)'
  write(*,*) '----------------------------------------------'
  write(*,*) 'Grid:', n, 'x', n, 'x', n
  write(*,*) 'Restart ', p%restart
  write(*,*) 'Tstart  ', p%tstart
  write(*,*) 'Tfin    ', p%tfin
  write(*,*) 'Dump    ', p%dump
  write(*,*) 'Inflow  ', p%inflow
  write(*,*) 'Deltat  ', p%dt
  write(*,*) 'Mu      ', p%mu
  write(*,*) 'Rho     ', p%rho
  write(*,*) 'f1,f2,f3,k0', p%f1, p%f2, p%f3, p%k0
  write(*,*) 'Radius  ', p%radius
  write(*,*) 'Sigma   ', p%sigma
  write(*,*) 'Eps/dx  ', p%epsr
  write(*,*) 'Phase field, tracers', p%phiflag, p%partflag
  if (p%poisson_precision == MHIT_POISSON_FP32) then
    write(*,*) 'Poisson solve in FP32'
  else
    write(*,*) 'Poisson solve in FP64'
  end if

  call check(mhit_create(p, ctx), 'mhit_create')
  allocate(buf(n, n, n))

  ! initial velocity (main.f90:322-350)
  if (p%restart == 0) then
    if (p%inflow == 0) then
      write(*,*) 'Initialize Taylor-green'
      call check(mhit_init_taylor_green(ctx), 'mhit_init_taylor_green')
    else
      write(*,*) 'Initialize from data'
      call read_field(ctx, FIELD_U, trim(field_name('./input', FIELD_U, -1)), n, buf)
      call read_field(ctx, FIELD_V, trim(field_name('./input', FIELD_V, -1)), n, buf)
      call read_field(ctx, FIELD_W, trim(field_name('./input', FIELD_W, -1)), n, buf)
    end if
  else
    write(*,*) 'Initialize velocity field (from output folder), iteration:', p%tstart
    call read_field(ctx, FIELD_U, trim(field_name('./output', FIELD_U, p%tstart)), n, buf)
    call read_field(ctx, FIELD_V, trim(field_name('./output', FIELD_V, p%tstart)), n, buf)
    call read_field(ctx, FIELD_W, trim(field_name('./output', FIELD_W, p%tstart)), n, buf)
  end if

  ! initial phase field (main.f90:366-397)
  if (p%phiflag == 1) then
    if (p%restart == 0) then
      if (p%inphi == 0) then
        write(*,*) 'Spherical drop'
        call check(mhit_init_drop(ctx), 'mhit_init_drop')
      else
        write(*,*) 'Initialize phase-field from data'
        call read_field(ctx, FIELD_PHI, trim(field_name('./input', FIELD_PHI, -1)), n, buf)
      end if
    else
      write(*,*) 'Initialize phase-field (restart, from output folder), iteration:', p%tstart
      call read_field(ctx, FIELD_PHI, trim(field_name('./output', FIELD_PHI, p%tstart)), n, buf)
    end if
  end if

  ! AB2 history of an exact restart (no reference counterpart)
  if (p%restart /= 0 .and. exact) then
    write(*,*) 'Exact restart: AB2 history from output folder, iteration:', p%tstart
    do f = FIELD_RHSU_O, FIELD_RHSW_O
      call read_field(ctx, f, trim(field_name('./output', f, p%tstart)), n, buf)
    end do
    if (p%phiflag == 1) call read_field(ctx, FIELD_RHSPHI_O, &
        trim(field_name('./output', FIELD_RHSPHI_O, p%tstart)), n, buf)
    if (.not. read_ab2(p%tstart, alpha, beta, gumax)) then
      write(*,*) 'ERROR: ./output/ab2 file of iteration', p%tstart, 'missing'
      error stop 1
    end if
    call check(mhit_set_ab2_state(ctx, alpha, beta, gumax), 'mhit_set_ab2_state')
  end if

  ! tracers (main.f90:399-415): uniform random positions
  if (p%partflag == 1 .and. p%restart /= 0 .and. exact) then
    write(*,*) 'Tracers from output folder, iteration:', p%tstart
    call read_tracers(ctx, p%tstart)
  else if (p%partflag == 1 .and. p%npart > 0) then
    write(*,*) 'Random Position whole Domain, tracers:', p%npart
    call check(mhit_init_tracers_uniform(ctx, p%npart, seed), 'mhit_init_tracers_uniform')
  end if

  ! save the initial fields on a fresh start (main.f90:418-427)
  if (p%restart == 0) then
    write(*,*) 'Save initial fields'
    call dump_fields(p%tstart)
  end if

  ! time loop (main.f90:439-1042); the first step is Euler
  tstart = p%tstart + 1
  call system_clock(count0, rate)
  do t = tstart, p%tfin
    if (verbose > 0) write(*,*) 'Time step', t, 'of', p%tfin
    call system_clock(countt)
    rc = mhit_step(ctx, 1_c_int)
    call check(mhit_diag_compute_light(), 'mhit_diag_compute')
    if (verbose > 0) then
      write(*,*) 'CFL (max among tasks)', d%cou
      call system_clock(count1)
      write(*,'(" Time elapsed = ",f8.1," ms")') 1000.d0 * real(count1 - countt, 8) / real(rate, 8)
    end if
    if (rc == MHIT_ERR_CFL) then
      write(*,*) 'CFL above 7: stop (', mhit_last_error(), ')'
      exit
    end if
    call check(rc, 'mhit_step')
    if (mod(t, p%dump) == 0) then
      write(*,*) 'Saving output files'
      call dump_fields(t)
    end if
  end do
  call system_clock(count1)
  write(*,*) 'Elapsed time (seconds):', real(count1 - count0, 8) / real(rate, 8)
  call check(mhit_diag_compute(ctx, d), 'mhit_diag_compute')
  write(*,'(a,es16.8,a,es16.8,a,es12.4)') ' KE ', d%ke, '  dissipation ', d%dissipation, &
      '  max|div u| ', d%maxdiv
  if (p%phiflag == 1) write(*,'(a,es20.12)') ' phi mass ', d%phi_mass
  call check(mhit_dump_wait(ctx), 'mhit_dump_wait')
  call check(mhit_destroy(ctx), 'mhit_destroy')

contains

  ! The scalars of the step (cou, printed as main.f90:1005 does) without
  ! the field diagnostics pass.
  integer(c_int) function mhit_diag_compute_light()
    d%cou = 0
    mhit_diag_compute_light = MHIT_OK
    if (verbose > 0) mhit_diag_compute_light = mhit_scalars(ctx, d)
  end function

  ! The field files are written by the library's host thread while the
  ! following steps run (the previous dump is finished first); tracers,
  ! drops and the AB2 scalars are small and written here.
  subroutine dump_fields(tt)
    integer, intent(in) :: tt
    integer(c_int32_t) :: mask
    mask = ibset(ibset(ibset(ibset(0_c_int32_t, FIELD_U), FIELD_V), FIELD_W), FIELD_P)
    if (p%phiflag == 1) mask = ibset(mask, FIELD_PHI)
    if (exact) then
      mask = ibset(ibset(ibset(mask, FIELD_RHSU_O), FIELD_RHSV_O), FIELD_RHSW_O)
      if (p%phiflag == 1) mask = ibset(mask, FIELD_RHSPHI_O)
    end if
    call check(mhit_dump_async(ctx, './output' // c_null_char, int(tt, c_int32_t), mask), &
               'mhit_dump_async')
    if (p%partflag == 1) call write_tracers(ctx, tt, exact)
    if (drops .and. p%phiflag == 1) call write_drops(ctx, tt)
    if (.not. exact) return
    call check(mhit_diag_compute(ctx, d), 'mhit_diag_compute')
    call write_ab2(tt, d%alpha, d%beta, d%gumax)
  end subroutine

  subroutine parse_args()
    integer :: i, na
    na = command_argument_count()
    i = 1
    do while (i <= na)
      call get_command_argument(i, arg)
      select case (trim(arg))
      case ('--nx')
        i = i + 1; call get_command_argument(i, arg); read(arg, *) p%nx
      case ('--phi')
        i = i + 1; call get_command_argument(i, arg); read(arg, *) p%phiflag
      case ('--part')
        i = i + 1; call get_command_argument(i, arg); read(arg, *) p%partflag
      case ('--seed')
        i = i + 1; call get_command_argument(i, arg); read(arg, *) seed
      case ('--fp64-poisson')
        p%poisson_precision = MHIT_POISSON_FP64
      case ('--k2-modified')
        p%k2_mode = MHIT_K2_MODIFIED
      case ('--trilinear')
        p%interp_mode = MHIT_INTERP_TRILINEAR
      case ('--exact-restart')
        exact = .true.
      case ('--drops')
        drops = .true.
      case ('--quiet')
        verbose = 0
      case default
        write(*,*) 'unknown argument ', trim(arg)
        error stop 2
      end select
      i = i + 1
    end do
  end subroutine

end program mhit36
