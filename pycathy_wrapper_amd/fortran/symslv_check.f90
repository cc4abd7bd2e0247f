! symslv_check.f90 -- calls cathy_gpu_symslv through the Fortran interface
! module on a small SPD system in the reference's upper-CSR layout
! (diagonal first per row, 1-based): a 3-D 7-point Laplacian plus a mass
! term on an 8 x 8 x 8 grid, with the last grid plane fixed (NODDIR) the
! way BCPIC marks Dirichlet rows.  Checks the returned solution against the
! residual it must satisfy and prints "SYMSLV CHECK OK" on success.
program symslv_check
  use cathy_gpu
  implicit none
  integer, parameter :: m = 8, n = m * m * m
  integer :: topol(n + 1), ja(4 * n), noddir(m * m)
  real(8) :: coef1(4 * n), tnoti(n), pnew(n), r(n)
  real(8) :: err, res, rnorm, bnorm
  integer :: i, j, k, row, nterm, numdir, niter, rc, q, c
  logical :: isdir(n)

  nterm = 0
  numdir = 0
  isdir = .false.
  do k = 1, m
    do j = 1, m
      do i = 1, m
        row = i + m * (j - 1) + m * m * (k - 1)
        topol(row) = nterm + 1
        nterm = nterm + 1
        ja(nterm) = row
        coef1(nterm) = 6.0d0 + 0.1d0
        if (i < m) then
          nterm = nterm + 1; ja(nterm) = row + 1; coef1(nterm) = -1.0d0
        end if
        if (j < m) then
          nterm = nterm + 1; ja(nterm) = row + m; coef1(nterm) = -1.0d0
        end if
        if (k < m) then
          nterm = nterm + 1; ja(nterm) = row + m * m; coef1(nterm) = -1.0d0
        end if
        tnoti(row) = sin(0.3d0 * i) + cos(0.2d0 * j) + 0.1d0 * k
        if (k == m) then
          numdir = numdir + 1
          noddir(numdir) = row
          isdir(row) = .true.
          coef1(topol(row)) = 1.7d91
          tnoti(row) = 0.0d0
        end if
      end do
    end do
  end do
  topol(n + 1) = nterm + 1

  rc = cathy_gpu_symslv(n, nterm, numdir, niter, 500, 1.0d-12, 1.7d-100, noddir, ja, topol, pnew, tnoti, &
                        coef1, err, res)
  if (rc /= CATHY_GPU_OK) then
    print '(a, i0, 2a)', 'SYMSLV CHECK FAILED rc=', rc, ' ', cathy_gpu_error_message()
    stop 1
  end if

  ! r = b - A x over the free rows, with A symmetric from its upper half
  r = tnoti
  do row = 1, n
    do q = topol(row), topol(row + 1) - 1
      c = ja(q)
      if (isdir(row) .or. isdir(c)) cycle
      r(row) = r(row) - coef1(q) * pnew(c)
      if (c /= row) r(c) = r(c) - coef1(q) * pnew(row)
    end do
  end do
  rnorm = 0.0d0
  bnorm = 0.0d0
  do row = 1, n
    if (isdir(row)) cycle
    rnorm = rnorm + r(row) ** 2
    bnorm = bnorm + tnoti(row) ** 2
  end do
  rnorm = sqrt(rnorm / bnorm)
  print '(a, i0, a, es10.3, a, es10.3, a, es10.3)', 'niter=', niter, ' err=', err, ' res=', res, &
        ' recomputed=', rnorm
  do row = 1, n
    if (isdir(row) .and. pnew(row) /= 0.0d0) then
      print '(a)', 'SYMSLV CHECK FAILED: Dirichlet row not zero'
      stop 1
    end if
  end do
  if (err > 1.0d-12 .or. rnorm > 1.0d-10 .or. abs(res - rnorm) > 1.0d-3 * rnorm + 1.0d-14) then
    print '(a)', 'SYMSLV CHECK FAILED'
    stop 1
  end if
  rc = cathy_gpu_symslv_release()
  print '(a)', 'SYMSLV CHECK OK'
end program symslv_check
