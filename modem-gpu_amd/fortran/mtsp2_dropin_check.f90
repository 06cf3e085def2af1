! Fortran-side check of the drop-in: reads one system (the reference's 1-based
! AAii CSR, VomegaMuSig/omega, omega, b, x0) from a stream file, solves it
! through mtsp2_bicg exactly as a reference solver.f90 would call it, and
! writes niter, failed, ierr, a flag, rerr(1:maxIt) and x.  Before that solve
! it solves with a slightly different operator of the same pattern (the
! binding must notice the change and upload the real one); after it, it
! solves once more with the same operator (flag = 1 when that repeat gives
! the same iterates).
!   usage: mtsp2_dropin_check <input> <output>
program mtsp2_dropin_check
   use iso_c_binding
   use mtsp2_fortran
   implicit none
   character(len=512) :: fin, fout
   integer :: u, nx, ny, nz, n, nnz, maxit, niter, niter3, ierr, flag, k, same
   integer(c_int), allocatable :: row(:), col(:)
   real(c_double), allocatable :: val(:), val2(:), vms1(:), rerr(:)
   complex(c_double_complex), allocatable :: b(:), x(:), x0(:), x3(:)
   real(c_double) :: omega, tol
   logical :: failed, failed3
   call get_command_argument(1, fin)
   call get_command_argument(2, fout)
   open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
   read(u) nx, ny, nz, n, nnz
   allocate(row(n + 1), col(nnz), val(nnz), vms1(n), b(n), x(n))
   read(u) row, col
   read(u) val, vms1
   read(u) omega
   read(u) maxit
   read(u) tol
   read(u) b, x
   close(u)
   allocate(rerr(maxit), x0(n), x3(n), val2(nnz))
   x0 = x
   ! 1. an operator that differs from AAii only in two swapped entries nnz/3
   !    and nnz/3 + 9973 (a weighted checksum with period 9973 would not see
   !    the difference): solved and discarded
   val2 = val
   k = nnz / 3
   if (k + 9973 <= nnz) then
      val2(k) = val(k + 9973)
      val2(k + 9973) = val(k)
   else
      val2(k) = 2.0d0 * val(k) + 1.0d0
   endif
   call mtsp2_bicg(0, nx, ny, nz, row, col, val2, vms1, omega, b, x, maxit, tol, &
                   niter, rerr, failed, ierr)
   ! 2. the real operator: must be uploaded again
   x = x0
   if (ierr == MT_OK .or. ierr == MT_E_MAXIT) &
      call mtsp2_bicg(0, nx, ny, nz, row, col, val, vms1, omega, b, x, maxit, tol, &
                      niter, rerr, failed, ierr)
   ! 3. the same operator again: no upload, the same iterates
   same = 0
   if (ierr == MT_OK) then
      x3 = x0
      call mtsp2_bicg(0, nx, ny, nz, row, col, val, vms1, omega, b, x3, maxit, tol, &
                      niter3, rerr, failed3, ierr)
      if (ierr == MT_OK .and. niter3 == niter .and. all(x3 == x)) same = 1
      call mtsp2_release()
   endif
   flag = 0
   if (failed) flag = 1
   open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
   write(u) niter, flag, ierr, same
   write(u) rerr
   write(u) x
   close(u)
end program
