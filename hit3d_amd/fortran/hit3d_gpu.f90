!==============================================================================!
! hit3d_gpu -- Fortran driver of the GPU time step.
!
! Reads <run_name>.in (the reference's positional input file,
! m_parameters.f90:172-451) and runs the driver loop of main.f90:56-297 with
! the hydro step on the GPU (hit3d_step: rhs_scalars, rhs_velocity,
! force_velocity, AB2 update, pressure):
!   * start: the seeded initial field (begin_new.f90, init_velocity.f90,
!     init_scalars.f90) and its io_write_4 snapshot, or, for ITMIN > 0, the
!     restart file <run_name>.64.<ITMIN> (restart_read_parallel,
!     restart_io.f90:341-506; missing scalars from init_scalar; dt halved for
!     variable dt, begin_restart.f90:33-38); then divergence and dealias_all
!     (main.f90:62-76);
!   * LES: les_model 1 (Smagorinsky, m_les.f90:1142-1215) and 11, the learned
!     closure, whose parameters come from <run_name>.sgs (6288 float32
!     parameters, then float64 xmean(10), xstd(10), ymean(6), ystd(6));
!   * particles (nptot > 0): positions from pt.<ITMIN>, else <run_name>.pt,
!     else the lattice (m_particles.f90:157-337); moved every step once
!     TIME >= starttime_particles with the velocity the step starts from
!     (fields_to_parts + particles_move, main.f90:235-247), and written to
!     pt.<ITIME> every IWRITE4 steps (particles_restart_write_binary,
!     m_particles.f90:558-648);
!   * output: stat1.gp, stat2.gp, es.gp (and scNN.gp, es_NN.gp) every
!     IPRINT1 steps (m_stats.f90:152-185, :371-437); u, v, w, scNN and, with
!     Smagorinsky, nu_t snapshots every IWRITE4 steps (io_write_4.f90:1-89,
!     write_tmp4.f90:14-30); restart files every IPRINT2 steps and at the end
!     (restart_io.f90:287-314).
!
! Usage:  hit3d_gpu <run_name> [precision]     (precision 0 = FP64, 1 = mixed)
! Build with -fdefault-integer-8 -fdefault-real-8, like the reference.
!==============================================================================!
program hit3d_gpu
  use iso_c_binding
  use hit3d_hip
  implicit none

  character(len=10) :: run_name
  character(len=80) :: str
  integer :: nx, ny, nz_all, itmin, itmax, iprint1, iprint2, iwrite4, nrescale
  integer :: flow_type, isp_type, ir_exp, force_type, kfmax, dealias, det_rand
  integer :: nptot, ptrack, les_model, n_scalars, n, kmax, itime, precision, ios
  real*8 :: tmax, trescale, tscalar, re, dt, peak, famp, rn1, rn2, rn3, tp, pfs
  real*8 :: nu, time, courant, energy, enstrophy
  logical :: variable_dt, int_scalars, int_particles, there
  integer :: np_run, nums_read
  integer*4, allocatable :: ipart(:)
  real*8, allocatable, target :: xyzp(:,:)
  real*8, parameter :: two_pi = 8.d0 * atan(1.d0)
  integer, allocatable :: scalar_type(:)
  real*8, allocatable :: sc(:), pe(:), e_spec(:), ir_sc(:), peak_sc(:), reac_sc(:)
  real*8 :: dummy(3)
  integer(c_int64_t) :: fov, fos, isc, prec

  if (command_argument_count() < 1) stop 'usage: hit3d_gpu <run_name> [precision]'
  call get_command_argument(1, run_name)
  if (len_trim(run_name) /= 10) stop 'Length of run name is not 10'
  precision = 0
  if (command_argument_count() >= 2) then
     call get_command_argument(2, str)
     read(str, *) precision
  end if

  ! ---------------------------------------------------------------- input
  open(10, file=run_name//'.in', form='formatted', status='old')
  read(10,*); read(10,*); read(10,*)
  read(10,*) nx, ny, nz_all
  read(10,*)
  read(10,*) itmin
  read(10,*) itmax
  read(10,*) iprint1
  read(10,*) iprint2
  read(10,*) iwrite4
  read(10,*)
  read(10,*) tmax
  read(10,*, iostat=ios) trescale, nrescale
  if (ios /= 0) then
     nrescale = 0
     if (trescale > 0.d0) nrescale = 1
  end if
  read(10,*) tscalar
  read(10,*)
  read(10,*) flow_type
  read(10,*)
  read(10,*) re
  read(10,*) dt
  variable_dt = dt >= 0.d0
  dt = abs(dt)
  read(10,*)
  read(10,*) isp_type
  read(10,*) ir_exp
  read(10,*) peak
  read(10,*)
  read(10,*) force_type
  read(10,*) kfmax
  read(10,*) famp
  read(10,*)
  read(10,*) dealias
  read(10,*)
  read(10,*) det_rand
  read(10,*) rn1
  read(10,*) rn2
  read(10,*) rn3
  read(10,*)
  read(10,*) nptot
  read(10,*) ptrack
  read(10,*) tp
  read(10,*) pfs
  read(10,*)
  read(10,*) les_model
  read(10,*)
  read(10,*) n_scalars
  read(10,*)
  allocate(scalar_type(max(n_scalars,1)), sc(max(n_scalars,1)), pe(max(n_scalars,1)), &
       ir_sc(max(n_scalars,1)), peak_sc(max(n_scalars,1)), reac_sc(max(n_scalars,1)))
  scalar_type = 0; sc = 1.d0; ir_sc = 0.d0; peak_sc = 1.d0; reac_sc = 0.d0
  nu = 1.d0 / re
  if (n_scalars > 0) then
     read(10,*)
     do n = 1, n_scalars                     ! m_parameters.f90:440-448
        read(10,*) scalar_type(n), sc(n), ir_sc(n), peak_sc(n), reac_sc(n)
        pe(n) = nu / sc(n)
     end do
  end if
  close(10)
  if (les_model /= 0 .and. les_model /= 1 .and. les_model /= 11) &
       stop 'les_model: 0, 1 (Smagorinsky) and 11 (learned closure) are built'
  if (les_model > 0 .and. dealias /= 0) stop 'LES needs dealias = 0 (m_parameters.f90:415-420)'
  if (nptot > 0 .and. pfs > 0.d0) stop 'particles on a locally averaged velocity are not built'

  if (dealias == 0) then
     kmax = nx / 3
  else
     kmax = floor(real(nx, 8) / 3.d0 * sqrt(2.d0))
  end if
  allocate(e_spec(kmax))

  ! ----------------------------------------------------------- start-up
  call h3d_check(c_init(int(nx, c_int64_t), int(ny, c_int64_t), int(nz_all, c_int64_t), &
       int(n_scalars, c_int64_t), 0_c_int64_t, int(dealias, c_int64_t)), 'init')
  call h3d_check(c_set_physics(nu, int(flow_type, c_int64_t), int(kfmax, c_int64_t), famp, &
       pe, int(scalar_type, c_int64_t)), 'set_physics')
  if (n_scalars > 0) call h3d_check(c_set_reaction(reac_sc), 'set_reaction')
  if (les_model == 1) call h3d_check(c_set_les(1_c_int64_t, 0.18d0), 'set_les')   ! m_les.f90:1148
  if (les_model == 11) call closure_in
  if (itmin == 0) then                       ! begin_new.f90:10-31
     itime = 0
     time = 0.d0
     int_scalars = (tscalar <= 0.d0 .and. n_scalars > 0)
     call h3d_check(c_init_velocity(rn1, int(isp_type, c_int64_t), int(ir_exp, c_int64_t), &
          peak), 'init_velocity')
     if (n_scalars > 0) call h3d_check(c_init_scalars(rn2, peak_sc), 'init_scalars')
     call io_write_4
  else                                       ! begin_restart.f90:1-90
     itime = itmin
     call restart_in
     if (variable_dt) dt = 0.5d0 * dt
     int_scalars = (n_scalars > 0 .and. time > tscalar)
  end if
  call divergence                            ! main.f90:66
  fov = 1; fos = 1
  call dealias_all                           ! main.f90:76
  prec = precision
  int_particles = .false.
  if (nptot > 0) call particles_in

  ! ---------------------------------------------------------- main loop
  do itime = itmin + 1, itmax
     ! parts task (main.f90:235-247): the velocity this step starts from
     if (nptot > 0 .and. time >= tp) then
        int_particles = .true.
        call h3d_check(c_fields_to_parts(), 'fields_to_parts')
        call h3d_check(c_particles_move(dt, int(ptrack, c_int64_t)), 'particles_move')
     end if
     isc = 0
     if (int_scalars) isc = 1
     call h3d_check(c_step(dt, fov, fos, isc, prec, courant, energy), 'step')
     time = time + dt
     if (mod(itime, iprint2) == 0) call restart_out           ! main.f90:173
     if (variable_dt) call my_dt_host
     if (mod(itime, iprint1) == 0) then
        write(*, "('ITIME=',i6,3x,'TIME=',f8.4,4x,'DT=',f8.5,3x,'Courn= ',f6.4)") &
             itime, time, dt, courant
        call stats_out
     end if
     if (mod(itime, iwrite4) == 0) then
        call io_write_4
        if (int_particles) call particles_out
     end if
     if (mod(itime, iprint1) == 0 .or. mod(itime, iwrite4) == 0) then
        if (time > tmax) exit
        if (n_scalars > 0 .and. .not. int_scalars .and. time > tscalar) then
           int_scalars = .true.                ! main.f90:199-204
           call h3d_check(c_init_scalars(rn2, peak_sc), 'init_scalars')
        end if
     end if
  end do

  if (itime > itmax) itime = itmax
  if (itmax > itmin .and. mod(itime, iprint2) /= 0) call restart_out   ! main.f90:288-296
  call h3d_check(c_finalize(), 'finalize')

contains

  function ext6(it) result(e)
    integer, intent(in) :: it
    character(len=6) :: e
    write(e, '(i6.6)') it
  end function ext6

  subroutine restart_in                      ! restart_io.f90:341-506 layout
    real*8, allocatable :: slot(:,:,:)
    integer*4 :: hdr(5)
    integer :: u, c
    inquire(file=run_name//'.64.'//ext6(itmin), exist=there)
    if (.not. there) stop 'restart file <run_name>.64.<ITMIN> not found'
    open(newunit=u, file=run_name//'.64.'//ext6(itmin), access='stream', form='unformatted', &
         status='old')
    read(u) hdr, time, dt
    if (hdr(1) /= nx .or. hdr(2) /= ny .or. hdr(3) /= nz_all) &
         stop '*** error: Dimensions are different'
    nums_read = min(n_scalars, int(hdr(4)))
    allocate(slot(nx+2, ny, nz_all))
    slot = 0.d0
    do c = 1, 3 + nums_read
       read(u) slot(1:nx, 1:ny, 1:nz_all)
       call h3d_check(c_put(H3D_FIELDS, int(c, c_int64_t), slot), 'put')
    end do
    close(u)
    deallocate(slot)
    do c = nums_read + 1, n_scalars          ! restart_io.f90:420-431
       call h3d_check(c_init_scalar(int(c, c_int64_t), rn2, peak_sc(c)), 'init_scalar')
    end do
  end subroutine restart_in

  subroutine closure_in                      ! les_model 11: <run_name>.sgs
    real(c_float) :: params(6288)
    real*8 :: xm(10), xs(10), ym(6), ys(6), mix
    integer :: u
    integer(8) :: fsize
    inquire(file=run_name//'.sgs', exist=there, size=fsize)
    if (.not. there) stop 'les_model 11 needs the closure file <run_name>.sgs'
    open(newunit=u, file=run_name//'.sgs', access='stream', form='unformatted', status='old')
    read(u) params, xm, xs, ym, ys
    mix = 0.d0                               ! files without the blend weight
    if (fsize >= 6288*4 + 33*8) read(u) mix
    close(u)
    call h3d_check(c_mlp_set_params(params), 'mlp_set_params')
    call h3d_check(c_mlp_set_norm(xm, xs, ym, ys), 'mlp_set_norm')
    call h3d_check(c_mlp_set_mix(mix), 'mlp_set_mix')
    call h3d_check(c_set_les(11_c_int64_t, 0.18d0), 'set_les')
  end subroutine closure_in

  subroutine particles_in                    ! m_particles.f90:157-337 on one rank
    integer :: u, i, j, k, m, npthird, nrec
    real*8 :: t_file, sctmp
    integer*4 :: n4, id4
    inquire(file='pt.'//ext6(itmin), exist=there)
    if (there) then                          ! particles_restart_read_binary, :655-744
       open(newunit=u, file='pt.'//ext6(itmin), access='stream', form='unformatted', &
            status='old')
       read(u) n4
       np_run = n4
       allocate(ipart(np_run), xyzp(3, np_run))
       do i = 1, np_run
          read(u) id4, xyzp(:, i)
          ipart(i) = id4
       end do
       read(u, iostat=nrec) t_file
       close(u)
    else
       inquire(file=run_name//'.pt', exist=there)
       if (there) then                       ! text positions, physical units
          np_run = nptot
          allocate(ipart(np_run), xyzp(3, np_run))
          open(newunit=u, file=run_name//'.pt', form='formatted', status='old')
          read(u, *) xyzp
          close(u)
       else                                  ! the lattice of the nearest cube, :222-262
          npthird = int(real(nptot, 8)**(1.d0 / 3.d0))
          if (nptot - npthird**3 >= (npthird + 1)**3 - nptot) npthird = npthird + 1
          np_run = min(npthird**3, 2**30)
          allocate(ipart(np_run), xyzp(3, np_run))
          sctmp = two_pi / real(npthird, 8)
          m = 0
          do k = 0, npthird - 1
             do j = 0, npthird - 1
                do i = 0, npthird - 1
                   m = m + 1
                   xyzp(1, m) = (real(i, 8) + 0.5d0) * sctmp
                   xyzp(2, m) = (real(j, 8) + 0.5d0) * sctmp
                   xyzp(3, m) = (real(k, 8) + 0.5d0) * sctmp
                end do
             end do
          end do
       end if
       do i = 1, np_run
          ipart(i) = i
       end do
    end if
    ! wrapped into [0, 2 pi), then grid units x / dx + 1 (:300-324)
    do i = 1, np_run
       do j = 1, 3
          if (xyzp(j, i) >= two_pi .or. xyzp(j, i) < 0.d0) &
               xyzp(j, i) = xyzp(j, i) - two_pi * floor(xyzp(j, i) / two_pi)
          xyzp(j, i) = xyzp(j, i) / (two_pi / real(nx, 8)) + 1.d0
       end do
    end do
    call h3d_check(c_particles_set(int(np_run, c_int64_t), xyzp, c_null_ptr), 'particles_set')
  end subroutine particles_in

  subroutine particles_out                   ! particles_restart_write_binary, :558-648
    integer :: u, i
    real*8 :: dx
    call h3d_check(c_particles_get(c_loc(xyzp), c_null_ptr), 'particles_get')
    dx = two_pi / real(nx, 8)
    open(newunit=u, file='pt.'//ext6(itime), access='stream', form='unformatted', &
         status='replace')
    write(u) int(np_run, 4)
    do i = 1, np_run
       write(u) int(ipart(i), 4), (xyzp(:, i) - 1.d0) * dx
    end do
    write(u) time
    close(u)
  end subroutine particles_out

  subroutine io_write_4                      ! io_write_4.f90:1-89
    real(c_float), allocatable :: tmp4(:,:,:)
    integer :: c
    character(len=16) :: fname
    allocate(tmp4(nx, ny, nz_all))
    do c = 1, 3
       call h3d_check(c_snapshot(int(c, c_int64_t), tmp4), 'snapshot')
       fname = 'u.'//ext6(itime)
       if (c == 2) fname = 'v.'//ext6(itime)
       if (c == 3) fname = 'w.'//ext6(itime)
       call write_tmp4(fname, tmp4)
    end do
    if (int_scalars) then
       do c = 1, n_scalars
          call h3d_check(c_snapshot(int(3 + c, c_int64_t), tmp4), 'snapshot')
          write(fname, "('sc',i2.2,'.',a6)") c, ext6(itime)
          call write_tmp4(fname, tmp4)
       end do
    end if
    if (les_model == 1) then
       call h3d_check(c_turb_visc(tmp4), 'turb_visc')
       call write_tmp4('nu_t.'//ext6(itime), tmp4)
    end if
    deallocate(tmp4)
  end subroutine io_write_4

  subroutine write_tmp4(fname, a)            ! write_tmp4.f90:14-30
    character(len=*), intent(in) :: fname
    real(c_float), intent(in) :: a(:,:,:)
    integer :: u
    open(newunit=u, file=trim(fname), access='stream', form='unformatted', status='replace')
    write(u) int(nx, 4), int(ny, 4), int(nz_all, 4)
    write(u) a
    close(u)
  end subroutine write_tmp4

  subroutine my_dt_host                      ! my_dt.f90:8-30
    real*8, parameter :: courant_max = 0.1d0
    integer :: k
    if (courant < courant_max) then
       dt = dt * 1.01d0
       dt = min(dt, 0.01d0)
    else
       dt = dt / 1.05d0
    end if
    if (int_scalars) then
       do k = 1, n_scalars
          dt = min(dt, dt * courant_max / courant / sc(k))
       end do
    end if
  end subroutine my_dt_host

  subroutine stats_out                       ! m_stats.f90:115-185
    real*8 :: en, eps_v, eta, uvar, xl, lam, rl, rl1, tau, s
    integer :: k
    logical :: there
    call h3d_check(c_stats(e_spec, enstrophy), 'stats')
    en = sum(e_spec(1:kmax))
    eps_v = 0.d0; s = 0.d0
    do k = 1, kmax
       eps_v = eps_v + e_spec(k) * real(k**2, 8) * 2.d0 * nu
       s = s + e_spec(k) / real(k, 8)
    end do
    eta = (nu**3 / eps_v)**0.25d0
    uvar = 2.d0 / 3.d0 * en
    xl = 4.d0 * atan(1.d0) / 2.d0 * s / uvar
    lam = sqrt(15.d0 * uvar * nu / eps_v)
    rl = uvar * sqrt(15.d0 / eps_v * re)
    rl1 = sqrt(uvar) * lam / nu
    tau = xl / sqrt(uvar)
    inquire(file='stat1.gp', exist=there)
    open(69, file='stat1.gp', position='append')
    if (.not. there) write(69,'(A)') ' # 1.itime 2.time         3.energy       4.diss         5.eta'// &
         '          6.enstrophy    7.R_lambda'
    write(69, "(i8,20e15.6)") itime, time, en, eps_v, eta, enstrophy, rl
    close(69)
    inquire(file='stat2.gp', exist=there)
    open(70, file='stat2.gp', position='append')
    if (.not. there) write(70,'(A)') '# 1.itime  2.time         3.int LS       4. lambda      '// &
         '5.R_lambda1    6.tau_e        7.etakmax'
    write(70, "(i8,20e15.6)") itime, time, xl, lam, rl1, tau, eta * real(kmax, 8)
    close(70)
    open(900, file='es.gp', position='append')
    write(900, "()")
    write(900, "()")
    write(900, "('# ITIME=',i7,' TIME=',e17.8)") itime, time
    do k = 1, kmax
       write(900, "(i4,2e15.6)") k, e_spec(k), e_spec(k) * real(k**2, 8) * 2.d0 * nu
    end do
    close(900)
    if (int_scalars) call scalar_stats_out
  end subroutine stats_out

  subroutine scalar_stats_out                ! m_stats.f90:62-80, :371-437
    real*8, allocatable :: d(:), m1(:), m2(:), mn(:), mx(:), es(:,:)
    character(len=16) :: fname
    logical :: there
    integer :: k, n
    allocate(d(n_scalars), m1(n_scalars), m2(n_scalars), mn(n_scalars), mx(n_scalars), &
         es(kmax, n_scalars))
    call h3d_check(c_stat_scalars(d, m1, m2, mn, mx, es), 'stat_scalars')
    do n = 1, n_scalars
       write(fname, "('es_',i2.2,'.gp')") n
       open(900, file=fname, position='append')
       write(900, "()")
       write(900, "()")
       write(900, "('# ITIME=',i7,' TIME=',e17.8)") itime, time
       do k = 1, kmax
          write(900, "(i4,4e15.6)") k, es(k, n)
       end do
       close(900)
       write(fname, "('sc',i2.2,'.gp')") n
       inquire(file=fname, exist=there)
       open(901, file=fname, position='append')
       if (.not. there) write(901, '(A)') '# 1.itime 2.time          3.sc.diss       '// &
            '4. mean      5.variance    6.min     7.max'
       write(901, "(i7,10e15.6)") itime, time, d(n), m1(n), m2(n), mn(n), mx(n)
       close(901)
    end do
    deallocate(d, m1, m2, mn, mx, es)
  end subroutine scalar_stats_out

  subroutine restart_out                     ! restart_io.f90:287-314 layout
    real*8, allocatable :: slot(:,:,:)
    integer(c_int64_t) :: c
    integer :: nout, u
    integer*4 :: hdr(5)
    character(len=6) :: ext
    allocate(slot(nx+2, ny, nz_all))
    nout = 0
    if (int_scalars) nout = n_scalars
    write(ext, '(i6.6)') itime
    open(newunit=u, file=run_name//'.64.'//ext, access='stream', form='unformatted', &
         status='replace')
    hdr = (/ int(nx,4), int(ny,4), int(nz_all,4), int(nout,4), 0_4 /)
    write(u) hdr, time, dt
    do c = 1, 3 + nout
       call h3d_check(c_get(H3D_FIELDS, c, slot), 'get')
       write(u) slot(1:nx, 1:ny, 1:nz_all)
    end do
    close(u)
    deallocate(slot)
  end subroutine restart_out

end program hit3d_gpu
