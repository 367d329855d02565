!==============================================================================!
! hit3d_hip -- ISO_C_BINDING interface to libhit3d_hip.so (include/hit3d.h)
! and drop-in wrappers with the names of the reference's subroutines
! (x_fftw.f90: xFFT3d, xFFT3d_fields, x_derivative; rhs_scalars.f90;
! rhs_velocity.f90; m_force.f90: force_velocity; pressure.f90;
! dealias_all.f90).  Every C entry point takes its integers as int64 and its
! reals as double by reference, which is what a Fortran caller built with
! -fdefault-integer-8 -fdefault-real-8 passes for default kinds.  A non-zero
! status prints the library's message and stops, as my_exit(-1) would.
!==============================================================================!
module hit3d_hip
  use iso_c_binding
  implicit none

  integer(c_int64_t), parameter :: H3D_FIELDS = 1, H3D_WRK = 2, H3D_RHS_OLD = 3

  interface
     integer(c_int) function c_init(nx, ny, nz_all, ns, nles, dealias) bind(C, name="hit3d_init")
       import :: c_int, c_int64_t
       integer(c_int64_t), intent(in) :: nx, ny, nz_all, ns, nles, dealias
     end function c_init
     integer(c_int) function c_set_physics(nu, flow_type, kfmax, famp, pe, stype) &
          bind(C, name="hit3d_set_physics")
       import :: c_int, c_int64_t, c_double
       real(c_double), intent(in) :: nu, famp, pe(*)
       integer(c_int64_t), intent(in) :: flow_type, kfmax, stype(*)
     end function c_set_physics
     integer(c_int) function c_finalize() bind(C, name="hit3d_finalize")
       import :: c_int
     end function c_finalize
     integer(c_int) function c_put(arr, slot, host) bind(C, name="hit3d_put")
       import :: c_int, c_int64_t, c_double
       integer(c_int64_t), intent(in) :: arr, slot
       real(c_double), intent(in) :: host(*)
     end function c_put
     integer(c_int) function c_get(arr, slot, host) bind(C, name="hit3d_get")
       import :: c_int, c_int64_t, c_double
       integer(c_int64_t), intent(in) :: arr, slot
       real(c_double), intent(out) :: host(*)
     end function c_get
     integer(c_int) function c_copy(da, ds, sa, ss) bind(C, name="hit3d_copy")
       import :: c_int, c_int64_t
       integer(c_int64_t), intent(in) :: da, ds, sa, ss
     end function c_copy
     integer(c_int) function c_xfft3d(flag, n) bind(C, name="hit3d_xfft3d")
       import :: c_int, c_int64_t
       integer(c_int64_t), intent(in) :: flag, n
     end function c_xfft3d
     integer(c_int) function c_xfft3d_fields(flag, n) bind(C, name="hit3d_xfft3d_fields")
       import :: c_int, c_int64_t
       integer(c_int64_t), intent(in) :: flag, n
     end function c_xfft3d_fields
     integer(c_int) function c_x_derivative(n, axis, nto) bind(C, name="hit3d_x_derivative")
       import :: c_int, c_int64_t
       integer(c_int64_t), intent(in) :: n, axis, nto
     end function c_x_derivative
     integer(c_int) function c_dealias_all() bind(C, name="hit3d_dealias_all")
       import :: c_int
     end function c_dealias_all
     integer(c_int) function c_rhs_scalars(int_scalars) bind(C, name="hit3d_rhs_scalars")
       import :: c_int, c_int64_t
       integer(c_int64_t), intent(in) :: int_scalars
     end function c_rhs_scalars
     integer(c_int) function c_rhs_velocity(dt, courant) bind(C, name="hit3d_rhs_velocity")
       import :: c_int, c_double
       real(c_double), intent(in) :: dt
       real(c_double), intent(out) :: courant
     end function c_rhs_velocity
     integer(c_int) function c_force_velocity(energy) bind(C, name="hit3d_force_velocity")
       import :: c_int, c_double
       real(c_double), intent(out) :: energy
     end function c_force_velocity
     integer(c_int) function c_advance(which, first, dt) bind(C, name="hit3d_advance")
       import :: c_int, c_int64_t, c_double
       integer(c_int64_t), intent(in) :: which
       integer(c_int64_t), intent(inout) :: first
       real(c_double), intent(in) :: dt
     end function c_advance
     integer(c_int) function c_pressure() bind(C, name="hit3d_pressure")
       import :: c_int
     end function c_pressure
     integer(c_int) function c_step(dt, fov, fos, int_scalars, precision, courant, energy) &
          bind(C, name="hit3d_step")
       import :: c_int, c_int64_t, c_double
       real(c_double), intent(in) :: dt
       integer(c_int64_t), intent(inout) :: fov, fos
       integer(c_int64_t), intent(in) :: int_scalars, precision
       real(c_double), intent(out) :: courant, energy
     end function c_step
     integer(c_int) function c_stats(e_spec, enstrophy) bind(C, name="hit3d_stats")
       import :: c_int, c_double
       real(c_double), intent(out) :: e_spec(*), enstrophy
     end function c_stats
     integer(c_int) function c_stat_scalars(sc_diss, mean, variance, sc_min, sc_max, es_spec) &
          bind(C, name="hit3d_stat_scalars")
       import :: c_int, c_double
       real(c_double), intent(out) :: sc_diss(*), mean(*), variance(*), sc_min(*), sc_max(*)
       real(c_double), intent(out) :: es_spec(*)
     end function c_stat_scalars
     integer(c_int) function c_divergence(dmin, dmax) bind(C, name="hit3d_divergence")
       import :: c_int, c_double
       real(c_double), intent(out) :: dmin, dmax
     end function c_divergence
     integer(c_int) function c_init_scalars(rn2, peak_sc) bind(C, name="hit3d_init_scalars")
       import :: c_int, c_double
       real(c_double), intent(in) :: rn2, peak_sc(*)
     end function c_init_scalars
     integer(c_int) function c_set_reaction(reac_sc) bind(C, name="hit3d_set_reaction")
       import :: c_int, c_double
       real(c_double), intent(in) :: reac_sc(*)
     end function c_set_reaction
     integer(c_int) function c_init_velocity(rn1, isp_type, ir_exp, peak) &
          bind(C, name="hit3d_init_velocity")
       import :: c_int, c_int64_t, c_double
       real(c_double), intent(in) :: rn1, peak
       integer(c_int64_t), intent(in) :: isp_type, ir_exp
     end function c_init_velocity
     integer(c_int) function c_last_error(buf, len) bind(C, name="hit3d_last_error")
       import :: c_int, c_int64_t, c_char
       character(kind=c_char), intent(out) :: buf(*)
       integer(c_int64_t), intent(in) :: len
     end function c_last_error
  end interface

contains

  subroutine h3d_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    character(kind=c_char) :: buf(512)
    character(len=512) :: msg
    integer :: i, ierr
    if (rc == 0) return
    ierr = c_last_error(buf, 512_c_int64_t)
    msg = ' '
    do i = 1, 512
       if (buf(i) == c_null_char) exit
       msg(i:i) = buf(i)
    end do
    write(*, '(A)') '*** '//what//': '//trim(msg)
    stop 1
  end subroutine h3d_check

  ! ---------------- wrappers with the reference's names and arguments --------
  subroutine xFFT3d(flag, n)                 ! x_fftw.f90:852
    integer, intent(in) :: flag, n
    call h3d_check(c_xfft3d(int(flag, c_int64_t), int(n, c_int64_t)), 'xFFT3d')
  end subroutine xFFT3d

  subroutine xFFT3d_fields(flag, n)          ! x_fftw.f90:1098
    integer, intent(in) :: flag, n
    call h3d_check(c_xfft3d_fields(int(flag, c_int64_t), int(n, c_int64_t)), 'xFFT3d_fields')
  end subroutine xFFT3d_fields

  subroutine x_derivative(n, axis, nto)      ! x_fftw.f90:637
    integer, intent(in) :: n, nto
    character, intent(in) :: axis
    integer(c_int64_t) :: ax
    select case (axis)
    case ('x'); ax = 1
    case ('y'); ax = 2
    case ('z'); ax = 3
    case default; ax = 0
    end select
    call h3d_check(c_x_derivative(int(n, c_int64_t), ax, int(nto, c_int64_t)), 'x_derivative')
  end subroutine x_derivative

  subroutine dealias_all                     ! dealias_all.f90:1
    call h3d_check(c_dealias_all(), 'dealias_all')
  end subroutine dealias_all

  subroutine rhs_scalars_gpu(int_scalars)    ! rhs_scalars.f90:1
    logical, intent(in) :: int_scalars
    integer(c_int64_t) :: f
    f = 0
    if (int_scalars) f = 1
    call h3d_check(c_rhs_scalars(f), 'rhs_scalars')
  end subroutine rhs_scalars_gpu

  subroutine rhs_velocity_gpu(dt, courant)   ! rhs_velocity.f90:1
    real(c_double), intent(in) :: dt
    real(c_double), intent(out) :: courant
    call h3d_check(c_rhs_velocity(dt, courant), 'rhs_velocity')
  end subroutine rhs_velocity_gpu

  subroutine force_velocity_gpu(energy)      ! m_force.f90:101
    real(c_double), intent(out) :: energy
    call h3d_check(c_force_velocity(energy), 'force_velocity')
  end subroutine force_velocity_gpu

  subroutine pressure_gpu                    ! pressure.f90:1
    call h3d_check(c_pressure(), 'pressure')
  end subroutine pressure_gpu

  subroutine divergence                      ! pressure.f90:53
    real(c_double) :: dmin, dmax
    call h3d_check(c_divergence(dmin, dmax), 'divergence')
    write(*,*) 'divergence:', dmin, dmax
  end subroutine divergence

end module hit3d_hip
