! symslv_gpu.f90 -- drop-in replacement of CATHY's SYMSLV
! (solscal-extended.f:4669-4699) that solves on the MI355X.  Same name, same
! dummy arguments, same log line; linked ahead of the reference's
! solscal-extended.o it takes over every call PICARD makes
! (picard.f:165-166).  COEF2 and SCR, the reference's factor and PCG
! scratch, are not touched: the device keeps its own (picard.f re-zeroes
! COEF2 before each assembly, picard.f:88).  A library error ends the run
! with STOP, the reference's own way of failing.
subroutine symslv(iout, n, nterm, numdir, niter, itmxcg, tolcg, rmin, noddir, ja, topol, pnew, tnoti, &
                  coef1, coef2, scr)
  use cathy_gpu
  implicit none
  integer :: iout, n, nterm, numdir, niter, itmxcg
  integer :: noddir(*), ja(*), topol(*)
  real(8) :: tolcg, rmin
  real(8) :: pnew(*), tnoti(*), coef1(*), coef2(*), scr(*)
  real(c_double) :: err, res
  integer(c_int) :: rc

  rc = cathy_gpu_symslv(n, nterm, numdir, niter, itmxcg, tolcg, rmin, noddir, ja, topol, pnew, tnoti, &
                        coef1, err, res)
  if (rc < 0) then
    write (iout, '(a, i0, 2a)') ' SYMSLV on the GPU failed (', rc, '): ', cathy_gpu_error_message()
    stop
  end if
  write (iout, 1000) niter, err, res
  return
1000 format(1x, i4, 2(1pe15.6), '  <<SYMMETRIC SOLVER>>')
end subroutine symslv
