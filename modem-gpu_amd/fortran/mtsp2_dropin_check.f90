! Fortran-side check of the drop-in: reads one system (the reference's 1-based
! AAii CSR, VomegaMuSig/omega, omega, b, x0) from a stream file, solves it
! through mtsp2_bicg exactly as a reference solver.f90 would call it, and
! writes niter, failed, ierr, rerr(1:maxIt) and x.
!   usage: mtsp2_dropin_check <input> <output>
program mtsp2_dropin_check
   use iso_c_binding
   use mtsp2_fortran
   implicit none
   character(len=512) :: fin, fout
   integer :: u, nx, ny, nz, n, nnz, maxit, niter, ierr, flag
   integer(c_int), allocatable :: row(:), col(:)
   real(c_double), allocatable :: val(:), vms1(:), rerr(:)
   complex(c_double_complex), allocatable :: b(:), x(:)
   real(c_double) :: omega, tol
   logical :: failed
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
   allocate(rerr(maxit))
   call mtsp2_bicg(0, nx, ny, nz, row, col, val, vms1, omega, b, x, maxit, tol, &
                   niter, rerr, failed, ierr)
   ! second call with the same operator: no re-upload, same answer expected
   if (ierr == MT_OK) then
      call mtsp2_release()
   endif
   flag = 0
   if (failed) flag = 1
   open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
   write(u) niter, flag, ierr
   write(u) rerr
   write(u) x
   close(u)
end program
