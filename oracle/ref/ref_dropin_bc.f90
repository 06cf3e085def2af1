! Drop-in check of the boundary fields (test infrastructure, never shipped):
! one Fortran host runs the reference's own ComputeBC -> BC_x0_WS
! (FWD_SP2/modelOperator3D.f90:292-317, FWD/boundary_ws.f90:21-171) and the
! library through its Fortran binding (mtsp2_bc_x0), for the same grid,
! model and (omega, mode) list, and writes both.
!
! Input: as oracle/ref/ref_bcx0.f90.  Output in the output folder:
!   ref.bin  nsys x complex128: E0 with setBC(BC) on the boundary edges
!   lib.bin  nsys x complex128: what mtsp2_bc_x0 returns
!   ierr.bin int32 status of mtsp2_bc_x0
program ref_dropin_bc
   use math_constants
   use griddef
   use sg_scalar
   use sg_vector
   use sg_boundary
   use ModelSpace
   use modelOperator3D
   use mtsp2_fortran, only: mtsp2_bc_x0
   use iso_c_binding
   implicit none

   character(len=512)              :: infile, outdir
   integer                         :: fin, fr, fl, fe, nx, ny, nzAir, nzEarth, nz, nsys, is, ierr
   integer                         :: nXe, nYe, nZe
   real(kind=prec), allocatable    :: dx(:), dy(:), dz(:), logc(:), om(:)
   integer(c_int), allocatable     :: md(:)
   complex(kind=prec), allocatable :: elib(:,:)
   real(kind=prec)                 :: airLog
   type(grid_t), target            :: grid
   type(rscalar)                   :: v
   type(modelParam_t)              :: m
   type(cvector)                   :: E0
   type(cboundary)                 :: BC

   call get_command_argument(1, infile)
   call get_command_argument(2, outdir)
   open(newunit=fin, file=trim(infile), access='stream', form='unformatted', status='old')
   read(fin) nx, ny, nzAir, nzEarth
   nz = nzAir + nzEarth
   allocate(dx(nx), dy(ny), dz(nz), logc(nx*ny*nzEarth))
   read(fin) dx, dy, dz
   read(fin) logc, airLog
   read(fin) nsys
   allocate(om(nsys), md(nsys))
   do is = 1, nsys
      read(fin) om(is), md(is)
   end do
   close(fin)

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

   open(newunit=fr, file=trim(outdir)//'/ref.bin', access='stream', form='unformatted', &
        status='replace')
   do is = 1, nsys
      call create_cvector(grid, E0, EDGE)
      call create_cboundary(grid, BC)
      omega = om(is)
      call ComputeBC(1, int(md(is)), E0, BC)
      call setBC(BC, E0)
      write(fr) E0%x, E0%y, E0%z
      call deall_cvector(E0)
      call deall_cboundary(BC)
   end do
   close(fr)

   nXe = nx*(ny+1)*(nz+1)
   nYe = (nx+1)*ny*(nz+1)
   nZe = (nx+1)*(ny+1)*nz
   allocate(elib(nXe + nYe + nZe, nsys))
   call mtsp2_bc_x0(0, nx, ny, nzAir, nzEarth, dx, dy, dz, logc, airLog, nsys, om, md, elib, ierr)
   open(newunit=fl, file=trim(outdir)//'/lib.bin', access='stream', form='unformatted', &
        status='replace')
   write(fl) elib
   close(fl)
   open(newunit=fe, file=trim(outdir)//'/ierr.bin', access='stream', form='unformatted', &
        status='replace')
   write(fe) int(ierr, 4)
   close(fe)
end program ref_dropin_bc
