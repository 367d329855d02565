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
     integer(c_int) function c_init_scalar(n_scalar, rn2, peak) bind(C, name="hit3d_init_scalar")
       import :: c_int, c_int64_t, c_double
       integer(c_int64_t), intent(in) :: n_scalar
       real(c_double), intent(in) :: rn2, peak
     end function c_init_scalar
     integer(c_int) function c_set_les(les_model, c_smag) bind(C, name="hit3d_set_les")
       import :: c_int, c_int64_t, c_double
       integer(c_int64_t), intent(in) :: les_model
       real(c_double), intent(in) :: c_smag
     end function c_set_les
     integer(c_int) function c_turb_visc(out) bind(C, name="hit3d_turb_visc")
       import :: c_int, c_float
       real(c_float), intent(out) :: out(*)
     end function c_turb_visc
     integer(c_int) function c_snapshot(n_field, out) bind(C, name="hit3d_snapshot")
       import :: c_int, c_int64_t, c_float
       integer(c_int64_t), intent(in) :: n_field
       real(c_float), intent(out) :: out(*)
     end function c_snapshot
     integer(c_int) function c_skewness(skew) bind(C, name="hit3d_skewness")
       import :: c_int, c_double
       real(c_double), intent(out) :: skew
     end function c_skewness
     ! particles: uvwp may be c_null_ptr (set: -999 everywhere; get: not copied)
     integer(c_int) function c_particles_set(np, xyzp, uvwp) bind(C, name="hit3d_particles_set")
       import :: c_int, c_int64_t, c_double, c_ptr
       integer(c_int64_t), intent(in) :: np
       real(c_double), intent(in) :: xyzp(*)
       type(c_ptr), value :: uvwp
     end function c_particles_set
     integer(c_int) function c_particles_get(xyzp, uvwp) bind(C, name="hit3d_particles_get")
       import :: c_int, c_ptr
       type(c_ptr), value :: xyzp, uvwp
     end function c_particles_get
     integer(c_int) function c_fields_to_parts() bind(C, name="hit3d_fields_to_parts")
       import :: c_int
     end function c_fields_to_parts
     integer(c_int) function c_particles_move(dt, scheme) bind(C, name="hit3d_particles_move")
       import :: c_int, c_int64_t, c_double
       real(c_double), intent(in) :: dt
       integer(c_int64_t), intent(in) :: scheme
     end function c_particles_move
     integer(c_int) function c_set_fusion(on) bind(C, name="hit3d_set_fusion")
       import :: c_int, c_int64_t
       integer(c_int64_t), intent(in) :: on
     end function c_set_fusion
     integer(c_int) function c_set_nonlinear(form) bind(C, name="hit3d_set_nonlinear")
       import :: c_int, c_int64_t
       integer(c_int64_t), intent(in) :: form
     end function c_set_nonlinear
     integer(c_int) function c_set_graphs(on) bind(C, name="hit3d_set_graphs")
       import :: c_int, c_int64_t
       integer(c_int64_t), intent(in) :: on
     end function c_set_graphs
     integer(c_int) function c_graph_stats(captured, replays) bind(C, name="hit3d_graph_stats")
       import :: c_int, c_int64_t
       integer(c_int64_t), intent(out) :: captured, replays
     end function c_graph_stats
     integer(c_int) function c_get_ialias(out) bind(C, name="hit3d_get_ialias")
       import :: c_int, c_int8_t
       integer(c_int8_t), intent(out) :: out(*)
     end function c_get_ialias
     integer(c_int) function c_rng_state(iy) bind(C, name="hit3d_rng_state")
       import :: c_int, c_int64_t
       integer(c_int64_t), intent(out) :: iy
     end function c_rng_state
     integer(c_int) function c_velocity_rng_state(nx, rn1, iy) &
          bind(C, name="hit3d_velocity_rng_state")
       import :: c_int, c_int64_t, c_double
       integer(c_int64_t), intent(in) :: nx
       real(c_double), intent(in) :: rn1
       integer(c_int64_t), intent(out) :: iy
     end function c_velocity_rng_state
     integer(c_int) function c_scalar_gaussian(nx, rn2, iy_prev, out, iy_next) &
          bind(C, name="hit3d_scalar_gaussian")
       import :: c_int, c_int64_t, c_double
       integer(c_int64_t), intent(in) :: nx, iy_prev
       real(c_double), intent(in) :: rn2
       real(c_double), intent(out) :: out(*)
       integer(c_int64_t), intent(out) :: iy_next
     end function c_scalar_gaussian
     integer(c_int) function c_random_draws(seed, skip, count, out) &
          bind(C, name="hit3d_random_draws")
       import :: c_int, c_int64_t, c_double
       real(c_double), intent(in) :: seed
       integer(c_int64_t), intent(in) :: skip, count
       real(c_double), intent(out) :: out(*)
     end function c_random_draws
     integer(c_int) function c_init_gaussian(nx, rn1, c, out) bind(C, name="hit3d_init_gaussian")
       import :: c_int, c_int64_t, c_double
       integer(c_int64_t), intent(in) :: nx, c
       real(c_double), intent(in) :: rn1
       real(c_double), intent(out) :: out(*)
     end function c_init_gaussian
     ! learned closure (les_model 11)
     integer(c_int) function c_mlp_set_params(params) bind(C, name="hit3d_mlp_set_params")
       import :: c_int, c_float
       real(c_float), intent(in) :: params(*)
     end function c_mlp_set_params
     integer(c_int) function c_mlp_get_params(params) bind(C, name="hit3d_mlp_get_params")
       import :: c_int, c_float
       real(c_float), intent(out) :: params(*)
     end function c_mlp_get_params
     integer(c_int) function c_mlp_set_norm(xmean, xstd, ymean, ystd) &
          bind(C, name="hit3d_mlp_set_norm")
       import :: c_int, c_double
       real(c_double), intent(in) :: xmean(*), xstd(*), ymean(*), ystd(*)
     end function c_mlp_set_norm
     integer(c_int) function c_mlp_set_data(count, x, t) bind(C, name="hit3d_mlp_set_data")
       import :: c_int, c_int64_t, c_float
       integer(c_int64_t), intent(in) :: count
       real(c_float), intent(in) :: x(*), t(*)
     end function c_mlp_set_data
     integer(c_int) function c_mlp_train(steps, batch, lr, losses) bind(C, name="hit3d_mlp_train")
       import :: c_int, c_int64_t, c_double
       integer(c_int64_t), intent(in) :: steps, batch
       real(c_double), intent(in) :: lr
       real(c_double), intent(out) :: losses(*)
     end function c_mlp_train
     integer(c_int) function c_mlp_grads(count, x, t, grads, loss) bind(C, name="hit3d_mlp_grads")
       import :: c_int, c_int64_t, c_float, c_double
       integer(c_int64_t), intent(in) :: count
       real(c_float), intent(in) :: x(*), t(*)
       real(c_float), intent(out) :: grads(*)
       real(c_double), intent(out) :: loss
     end function c_mlp_grads
     integer(c_int) function c_mlp_grads_static(count, x, t, grads, loss) &
          bind(C, name="hit3d_mlp_grads_static")
       import :: c_int, c_int64_t, c_float, c_double
       integer(c_int64_t), intent(in) :: count
       real(c_float), intent(in) :: x(*), t(*)
       real(c_float), intent(out) :: grads(*)
       real(c_double), intent(out) :: loss
     end function c_mlp_grads_static
     integer(c_int) function c_mlp_forward(count, x, y) bind(C, name="hit3d_mlp_forward")
       import :: c_int, c_int64_t, c_float
       integer(c_int64_t), intent(in) :: count
       real(c_float), intent(in) :: x(*)
       real(c_float), intent(out) :: y(*)
     end function c_mlp_forward
     integer(c_int) function c_mlp_fallbacks(count) bind(C, name="hit3d_mlp_fallbacks")
       import :: c_int, c_int64_t
       integer(c_int64_t), intent(out) :: count
     end function c_mlp_fallbacks
     integer(c_int) function c_mlp_set_mix(weight) bind(C, name="hit3d_mlp_set_mix")
       import :: c_int, c_double
       real(c_double), intent(in) :: weight
     end function c_mlp_set_mix
     integer(c_int) function c_box_filter(ratio, ubar, tau) bind(C, name="hit3d_box_filter")
       import :: c_int, c_int64_t, c_double
       integer(c_int64_t), intent(in) :: ratio
       real(c_double), intent(out) :: ubar(*), tau(*)
     end function c_box_filter
     integer(c_int) function c_sgs_features(delta, x) bind(C, name="hit3d_sgs_features")
       import :: c_int, c_double, c_float
       real(c_double), intent(in) :: delta
       real(c_float), intent(out) :: x(*)
     end function c_sgs_features
     ! timing and synchronisation
     integer(c_int) function c_timing_enable(on) bind(C, name="hit3d_timing_enable")
       import :: c_int, c_int64_t
       integer(c_int64_t), intent(in) :: on
     end function c_timing_enable
     integer(c_int) function c_timing_read(cap, ms, count) bind(C, name="hit3d_timing_read")
       import :: c_int, c_int64_t, c_double
       integer(c_int64_t), intent(in) :: cap
       real(c_double), intent(out) :: ms(*)
       integer(c_int64_t), intent(out) :: count(*)
     end function c_timing_read
     integer(c_int) function c_timing_names(buf, len) bind(C, name="hit3d_timing_names")
       import :: c_int, c_int64_t, c_char
       character(kind=c_char), intent(out) :: buf(*)
       integer(c_int64_t), intent(in) :: len
     end function c_timing_names
     ! include/hit3d_measure.h
     integer(c_int) function c_stream_copy(reps, blocks, ms, bytes) &
          bind(C, name="hit3d_stream_copy")
       import :: c_int, c_int64_t, c_double
       integer(c_int64_t), intent(in) :: reps, blocks
       real(c_double), intent(out) :: ms, bytes
     end function c_stream_copy
     integer(c_int) function c_synchronize() bind(C, name="hit3d_synchronize")
       import :: c_int
     end function c_synchronize
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
