! The drop-in, end to end (test infrastructure): the reference's own SP2 host
! code sets up the operator (ModelDataInit, UpdateCond, UpdateFreq from
! f90/3D_MT/FWD_SP2/modelOperator3D.f90, compiled from the reference sources
! by oracle/ref/Makefile) and the solve goes through `cuBiCG` below -- the
! replacement a maintainer puts into solver.f90, which hands
! the reference's module state (AAii, VomegaMuSig, omega, mGrid) to libmtsp2
! through modem-gpu_amd/fortran/mtsp2_fortran.f90.  The reference CPU BICG
! (solver.f90:720-926) solves the same system for comparison.
!
! Input (stream): int32 flags (unused), nx, ny, nzAir, nzEarth; real64 dx, dy,
! dz, logcond, airLog; int32 nsolve; nsolve x {real64 omega; int32 maxit;
! real64 tol; complex128 b(nInt), x0(nInt)}.
! Output per solve k: dropin_k.bin = int32 niter_gpu, failed, ierr,
! niter_cpu; complex128 x_gpu(nInt), x_cpu(nInt).
module solver_mtsp2
   use math_constants
   use solver, only: solverControl_t
   implicit none
contains
   subroutine cuBiCG(b, x, KSPiter, device_idx)
      use modeloperator3d, only: AAii, VomegaMuSig, omega, mGrid
      use mtsp2_fortran
      complex(kind=prec), intent(in), dimension(:)    :: b
      complex(kind=prec), intent(inout), dimension(:) :: x
      type(solverControl_t), intent(inout)            :: KSPiter
      integer, intent(in)                             :: device_idx
      integer :: ierr
      call mtsp2_bicg(device_idx, mGrid%nx, mGrid%ny, mGrid%nz, AAii%row, AAii%col, AAii%val, &
                      VomegaMuSig / omega, omega, b, x, KSPiter%maxIt, KSPiter%tol, &
                      KSPiter%niter, KSPiter%rerr, KSPiter%failed, ierr)
      if (ierr /= MT_OK) then
         write(0, *) 'libmtsp2 error ', ierr
         KSPiter%failed = .true.
      endif
   end subroutine
end module solver_mtsp2

program ref_dropin
   use math_constants
   use griddef
   use sg_scalar
   use ModelSpace
   use modelOperator3D
   use solver
   use solver_mtsp2, only: gpuBiCG => cuBiCG
   implicit none
   character(len=512) :: infile, outdir
   integer :: fin, uo, s, nx, ny, nzAir, nzEarth, nz, k, nsolve, maxit, flags, fl, ierr, ncpu
   real(kind=prec), allocatable :: dx(:), dy(:), dz(:), logc(:)
   real(kind=prec) :: airLog, w, tol
   type(grid_t), target :: grid
   type(rscalar) :: v
   type(modelParam_t) :: m
   complex(kind=prec), allocatable :: b(:), x(:), xc(:)
   type(solverControl_t) :: ctl
   character(len=16) :: tag
   call get_command_argument(1, infile)
   call get_command_argument(2, outdir)
   open(newunit=fin, file=trim(infile), access='stream', form='unformatted', status='old')
   read(fin) flags
   read(fin) nx, ny, nzAir, nzEarth
   nz = nzAir + nzEarth
   allocate(dx(nx), dy(ny), dz(nz), logc(nx*ny*nzEarth))
   read(fin) dx, dy, dz
   read(fin) logc, airLog
   call create_grid(nx, ny, nzAir, nzEarth, grid)
   grid%dx = dx
   grid%dy = dy
   grid%dz = dz
   call setup_grid(grid)
   call create_rscalar(grid, v, CELL_EARTH)
   v%v = reshape(logc, (/nx, ny, nzEarth/))
   call create_modelParam(grid, LOGE, m, v, airLog)
   call ModelDataInit(grid)
   call UpdateCond(m)
   allocate(b(size(EDGEi)), x(size(EDGEi)), xc(size(EDGEi)))
   read(fin) nsolve
   do k = 1, nsolve
      read(fin) w
      read(fin) maxit
      read(fin) tol
      read(fin) b, x
      xc = x
      call UpdateFreq(w)
      ctl%maxIt = maxit
      ctl%tol = tol
      ctl%failed = .false.
      allocate(ctl%rerr(maxit))
      call gpuBiCG(b, x, ctl, 0)
      fl = 0
      if (ctl%failed) fl = 1
      ierr = 0
      s = ctl%niter
      ctl%failed = .false.
      ctl%rerr = 0.0
      call BICG(b, xc, ctl)
      ncpu = ctl%niter
      deallocate(ctl%rerr)
      write(tag, '(I0)') k
      open(newunit=uo, file=trim(outdir)//'/dropin_'//trim(tag)//'.bin', access='stream', &
           form='unformatted', status='replace')
      write(uo) s, fl, ierr, ncpu
      write(uo) x, xc
      close(uo)
   enddo
   close(fin)
end program
