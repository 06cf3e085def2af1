! Reference-side driver for the 2-D TE boundary fields (test infrastructure,
! never shipped).
!
! Builds the grid and LOGE model as the forward code does, then runs the
! reference's own ComputeBC (3D_MT/FWD_SP2/modelOperator3D.f90:292-317) ->
! BC_x0_WS (3D_MT/FWD/boundary_ws.f90:21-171) for each requested (omega,
! mode), with the module frequency set as UpdateFreq sets it.  The banded
! LAPACK it calls comes from the reference's vendored LAPACK/wsLAPACK.f90 and
! LAPACK/wsBLAS.f90 (recipe: oracle/ref/Makefile).
!
! Input (stream, native endianness), written by oracle/refdriver.py:
!   int32 nx, ny, nzAir, nzEarth ; real64 dx(nx), dy(ny), dz(nz)
!   real64 logcond(nx*ny*nzEarth) (Fortran order), airLogCond
!   int32 nsys ; nsys x { real64 omega ; int32 imode }
! Output in the output folder:
!   e0.bin  nsys x complex128 E0 (x-, y-, z-edges, Fortran order each)
!   bc.bin  nsys x complex128 the boundary conditions written into a zero
!           edge vector by setBC (what FWDsolve3D copies to e(EDGEb))
program ref_bcx0
   use math_constants
   use griddef
   use sg_scalar
   use sg_vector
   use sg_boundary
   use ModelSpace
   use modelOperator3D
   implicit none

   character(len=512)            :: infile, outdir
   integer                       :: fin, fe, fb, nx, ny, nzAir, nzEarth, nz, nsys, is, imode
   real(kind=prec), allocatable  :: dx(:), dy(:), dz(:), logc(:)
   real(kind=prec)               :: airLog, om
   type(grid_t), target          :: grid
   type(rscalar)                 :: v
   type(modelParam_t)            :: m
   type(cvector)                 :: E0, tvec
   type(cboundary)               :: BC

   call get_command_argument(1, infile)
   call get_command_argument(2, outdir)
   open(newunit=fin, file=trim(infile), access='stream', form='unformatted', status='old')
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

   open(newunit=fe, file=trim(outdir)//'/e0.bin', access='stream', form='unformatted', &
        status='replace')
   open(newunit=fb, file=trim(outdir)//'/bc.bin', access='stream', form='unformatted', &
        status='replace')
   read(fin) nsys
   do is = 1, nsys
      read(fin) om, imode
      call create_cvector(grid, E0, EDGE)
      call create_cvector(grid, tvec, EDGE)
      call create_cboundary(grid, BC)
      omega = om
      call ComputeBC(1, imode, E0, BC)
      call setBC(BC, tvec)
      write(fe) E0%x, E0%y, E0%z
      write(fb) tvec%x, tvec%y, tvec%z
      call deall_cvector(E0)
      call deall_cvector(tvec)
      call deall_cboundary(BC)
   end do
   close(fe)
   close(fb)
   close(fin)
end program ref_bcx0
