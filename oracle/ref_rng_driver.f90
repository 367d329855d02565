! ORACLE (test infrastructure only).  A driver of this project's own that
! calls the reference's random(), compiled from /root/reference/RANDu.f with
! the reference's kinds (-i8 -r8), so the CPU restatement in randu.c can be
! checked against the reference generator itself.
!
! Usage:  randu_ref SEED1 COUNT1 [SEED2 COUNT2 ...]
! For each (SEED, COUNT) pair: one call random(-SEED) (re-initialisation,
! RANDu.f:28-37, which keeps iy and iv), then COUNT calls random(SEED).
! Every returned value (the initialising one included) is printed as the
! integer with the same 64 bits, one per line, so comparisons are exact.
program randu_ref
  use RANDu
  implicit none
  integer :: nargs, ia, i, cnt
  character(len=64) :: arg
  real*8 :: seed, x
  nargs = command_argument_count()
  ia = 1
  do while (ia + 1 <= nargs)
     call get_command_argument(ia, arg)
     read(arg, *) seed
     call get_command_argument(ia + 1, arg)
     read(arg, *) cnt
     rseed = seed
     x = random(-rseed)
     print '(I0)', transfer(x, 0_8)
     do i = 1, cnt
        x = random(rseed)
        print '(I0)', transfer(x, 0_8)
     end do
     ia = ia + 2
  end do
end program randu_ref
