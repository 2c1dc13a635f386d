! mhit_c.f90 -- iso_c_binding interface to libmhit36.so (include/mhit36.h).
!
! Replaces the reference's `use cudafor`, `use cufft`, `use cudecomp` and
! `use mpi` call sites of multi/main.f90 for one GPU: the whole timestep
! (main.f90:456-1007) is one call, mhit_step.
module mhit_c
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: MHIT_OK = 0, MHIT_ERR_ARG = 1, MHIT_ERR_HIP = 2, &
                               MHIT_ERR_CFL = 3, MHIT_ERR_NOMEM = 4, MHIT_ERR_STATE = 5, &
                               MHIT_ERR_NOGPU = 6, MHIT_ERR_IO = 7
  integer(c_int), parameter :: MHIT_POISSON_FP64 = 0, MHIT_POISSON_FP32 = 1
  integer(c_int), parameter :: MHIT_K2_CONTINUOUS = 0, MHIT_K2_MODIFIED = 1
  integer(c_int), parameter :: MHIT_INTERP_REFERENCE = 0, MHIT_INTERP_TRILINEAR = 1
  ! field numbers of writefield/readfield (multi/readwrite.f90:21-25)
  integer(c_int), parameter :: FIELD_U = 1, FIELD_V = 2, FIELD_W = 3, FIELD_P = 4, FIELD_PHI = 5
  ! AB2 history for a restart that continues bit for bit (no reference counterpart)
  integer(c_int), parameter :: FIELD_RHSU_O = 6, FIELD_RHSV_O = 7, FIELD_RHSW_O = 8, &
                               FIELD_RHSPHI_O = 9

  ! input.inp (multi/input.inp:1-18) plus the reference's compile-time switches
  type, bind(c) :: mhit_params
    integer(c_int32_t) :: restart, tstart, tfin, dump, inflow, inphi
    real(c_double)     :: dt, mu, rho, f1, f2, f3, k0, radius, sigma, epsr
    integer(c_int64_t) :: npart
    integer(c_int32_t) :: inpart
    integer(c_int32_t) :: nx, phiflag, partflag, poisson_precision, k2_mode, interp_mode, cfl_abort
  end type mhit_params

  ! one connected drop of phi >= threshold (mhit_drop_stats)
  type, bind(c) :: mhit_drop
    integer(c_int64_t) :: cells, first_cell
    real(c_double)     :: x, y, z, u, v, w, volume, d_eq
  end type mhit_drop

  type, bind(c) :: mhit_diag
    integer(c_int64_t) :: step
    real(c_double)     :: alpha, beta, gumax, cou, umean, vmean, wmean
    real(c_double)     :: ke, dissipation, maxdiv, phi_mass, phi_min, phi_max
    integer(c_int64_t) :: nan_count
  end type mhit_diag

  interface
    integer(c_int) function mhit_params_default(p) bind(c, name='mhit_params_default')
      import :: c_int, mhit_params
      type(mhit_params), intent(out) :: p
    end function
    integer(c_int) function mhit_create(p, ctx) bind(c, name='mhit_create')
      import :: c_int, c_ptr, mhit_params
      type(mhit_params), intent(in) :: p
      type(c_ptr), intent(out) :: ctx
    end function
    integer(c_int) function mhit_destroy(ctx) bind(c, name='mhit_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function mhit_upload_field(ctx, fieldn, host) bind(c, name='mhit_upload_field')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: fieldn
      real(c_double), intent(in) :: host(*)
    end function
    integer(c_int) function mhit_download_field(ctx, fieldn, host) bind(c, name='mhit_download_field')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: fieldn
      real(c_double), intent(out) :: host(*)
    end function
    ! writefield (multi/readwrite.f90:38-98) from a host thread; dir is a
    ! NUL-terminated path, mask has bit fieldn set for each field
    integer(c_int) function mhit_dump_async(ctx, dir, step, mask) bind(c, name='mhit_dump_async')
      import :: c_int, c_ptr, c_char, c_int32_t
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: dir(*)
      integer(c_int32_t), value :: step, mask
    end function
    integer(c_int) function mhit_dump_wait(ctx) bind(c, name='mhit_dump_wait')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function mhit_upload_tracers(ctx, n, x, y, z) bind(c, name='mhit_upload_tracers')
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: n
      real(c_double), intent(in) :: x(*), y(*), z(*)
    end function
    integer(c_int) function mhit_drop_stats(ctx, threshold, connectivity, max_drops, drops, ndrops) &
        bind(c, name='mhit_drop_stats')
      import :: c_int, c_ptr, c_double, c_int64_t, mhit_drop
      type(c_ptr), value :: ctx
      real(c_double), value :: threshold
      integer(c_int), value :: connectivity
      integer(c_int64_t), value :: max_drops
      type(mhit_drop), intent(out) :: drops(*)
      integer(c_int64_t), intent(out) :: ndrops
    end function
    integer(c_int) function mhit_set_ab2_state(ctx, alpha, beta, gumax) &
        bind(c, name='mhit_set_ab2_state')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: alpha, beta, gumax
    end function
    integer(c_int) function mhit_upload_tracer_state(ctx, n, tag, x, y, z, up, vp, wp) &
        bind(c, name='mhit_upload_tracer_state')
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: n
      real(c_double), intent(in) :: tag(*), x(*), y(*), z(*), up(*), vp(*), wp(*)
    end function
    integer(c_int) function mhit_download_tracers(ctx, tag, x, y, z, up, vp, wp) &
        bind(c, name='mhit_download_tracers')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: tag(*), x(*), y(*), z(*), up(*), vp(*), wp(*)
    end function
    integer(c_int64_t) function mhit_tracer_count(ctx) bind(c, name='mhit_tracer_count')
      import :: c_int64_t, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function mhit_step(ctx, nsteps) bind(c, name='mhit_step')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: nsteps
    end function
    integer(c_int) function mhit_diag_compute(ctx, d) bind(c, name='mhit_diag_compute')
      import :: c_int, c_ptr, mhit_diag
      type(c_ptr), value :: ctx
      type(mhit_diag), intent(out) :: d
    end function
    ! the last step's scalars (cou, gumax, ...) without a pass over the fields
    integer(c_int) function mhit_scalars(ctx, d) bind(c, name='mhit_scalars')
      import :: c_int, c_ptr, mhit_diag
      type(c_ptr), value :: ctx
      type(mhit_diag), intent(out) :: d
    end function
    integer(c_int) function mhit_init_taylor_green(ctx) bind(c, name='mhit_init_taylor_green')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function mhit_init_drop(ctx) bind(c, name='mhit_init_drop')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function mhit_init_tracers_uniform(ctx, n, seed) bind(c, name='mhit_init_tracers_uniform')
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: n
      integer(c_int64_t), value :: seed
    end function
    type(c_ptr) function mhit_last_error_c() bind(c, name='mhit_last_error')
      import :: c_ptr
    end function
    integer(c_size_t) function c_strlen(s) bind(c, name='strlen')
      import :: c_size_t, c_ptr
      type(c_ptr), value :: s
    end function
  end interface

contains

  ! Text of the last library error as a Fortran string.
  function mhit_last_error() result(msg)
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: chars(:)
    integer :: i, n
    p = mhit_last_error_c()
    n = int(c_strlen(p))
    allocate(character(len=n) :: msg)
    if (n == 0) return
    call c_f_pointer(p, chars, [n])
    do i = 1, n
      msg(i:i) = chars(i)
    end do
  end function

  ! Stop with the library's message on any failed call (the reference exits
  ! on a cuDecomp error, main.f90:1).
  subroutine check(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    if (rc /= MHIT_OK) then
      write(*,'(a,a,a,i0,a,a)') 'ERROR: ', where, ' returned ', rc, ': ', mhit_last_error()
      error stop 1
    end if
  end subroutine

end module mhit_c
