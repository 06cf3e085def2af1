! Reference-side driver of the sensitivity chain (test infrastructure, never
! shipped): one MT transmitter (both modes) through the reference's own
! JmultT_TX and through the same steps called one by one, with every stage
! written out, so that the oracle's restatement of the TRN solve (RHS and
! post-processing) and of PmultT, and the library's GPU chain, can be checked
! against the reference's gradient.
!
! Reference routines driven here (paths relative to the reference f90/ tree):
!   grid, model           3D_MT/GridDef.f90 create_grid/setup_grid,
!                         3D_MT/modelParam/ModelSpace.f90 create_modelParam
!   dictionaries          3D_MT/DICT/transmitters.f90, receivers.f90,
!                         dataTypes.f90 setup_typeDict
!   solver control        3D_MT/FWD_SP2/EMsolve3D.f90:638-686 setEMsolveControl
!                         (BICG, one call of at most maxit iterations, tolerance tol)
!   forward               3D_MT/ForwardSolver.f90:176-282 initSolver, :314-462
!                         fwdSetup (ComputeBC -> BC_x0_WS), :465-564 fwdSolve
!                         (FWDsolve3D, EMsolve3D.f90:129-501, forward branch)
!   adjoint source        SENS/DataSens.f90:130-225 LmultT
!   transposed solve      3D_MT/ForwardSolver.f90:567-620 sensSolve(TRN) ->
!                         FWDsolve3D TRN branch (RHS EMsolve3D.f90:217-256,
!                         post-processing :463-478)
!   gradient              3D_MT/SolverSens.f90:63-116 PmultT (-> ModelMap.inc:
!                         281-372 dEdgeToModelParam); SENS/SensComp.f90:408-485
!                         JmultT_TX (the same chain, plus QmultT)
!
! Input file layout (stream, native endianness), written by oracle/refdriver.py:
!   int32 nx, ny, nzAir, nzEarth
!   real64 dx(nx), dy(ny), dz(nzAir+nzEarth), origin(3)
!   real64 logcond(nx*ny*nzEarth) (Fortran order), airLogCond
!   real64 omega ; int32 maxit ; real64 tol
!   int32 nsite ; real64 x(3,nsite) ; real64 d(8,nsite)   (full impedance data)
! Output (stream): fwd.bin = complex128 e0 mode 1, mode 2 (all edges, [x;y;z]);
!   comb.bin = the LmultT sources of modes 1, 2 (all edges);
!   trn.bin = the sensSolve(TRN) solutions of modes 1, 2 (all edges);
!   dsig.bin = real64 PmultT(e0, e) on the cells (nx*ny*nzEarth);
!   dsig_jt.bin = real64 JmultT_TX(d) on the cells.
program ref_jmultT
   use math_constants
   use griddef
   use sg_scalar
   use sg_vector
   use ModelSpace
   use transmitters
   use receivers
   use dataTypes
   use DataSpace
   use SolnSpace
   use DataFunc
   use DataSens
   use SolverSens
   use EMsolve3D
   use ForwardSolver
   use SensComp
   implicit none

   character(len=512)              :: infile, outdir
   integer                         :: fin, fout, nx, ny, nzAir, nzEarth, nz, nsite, s, maxit, k
   real(kind=prec), allocatable    :: dx(:), dy(:), dz(:), logc(:), xs(:,:), d(:,:)
   real(kind=prec)                 :: origin(3), om, airLog, tol
   type(grid_t), target            :: grid
   type(rscalar)                   :: v
   type(modelParam_t)              :: m, dsig, dsig_jt
   type(emsolve_control)           :: ctl
   type(solnVector_t)              :: e0, e
   type(rhsVector_t)               :: b0, comb
   type(dataVector_t)              :: dv

   call get_command_argument(1, infile)
   call get_command_argument(2, outdir)
   open(newunit=fin, file=trim(infile), access='stream', form='unformatted', status='old')
   read(fin) nx, ny, nzAir, nzEarth
   nz = nzAir + nzEarth
   allocate(dx(nx), dy(ny), dz(nz), logc(nx*ny*nzEarth))
   read(fin) dx, dy, dz, origin
   read(fin) logc, airLog
   read(fin) om, maxit, tol
   read(fin) nsite
   allocate(xs(3,nsite), d(8,nsite))
   read(fin) xs, d
   close(fin)

   call create_grid(nx, ny, nzAir, nzEarth, grid)
   grid%dx = dx
   grid%dy = dy
   grid%dz = dz
   call setup_grid(grid, origin)
   call create_rscalar(grid, v, CELL_EARTH)
   v%v = reshape(logc, (/nx, ny, nzEarth/))
   call create_modelParam(grid, LOGE, m, v, airLog)

   ! one BICG call of at most maxit iterations per solve, to tolerance tol
   ctl%UseDefaults = .false.
   ctl%IterPerDivCor = maxit
   ctl%MaxDivCor = 1
   ctl%MaxIterDivCor = 100
   ctl%tolEMfwd = tol
   ctl%tolEMadj = tol
   ctl%tolDivCor = 1d-7
   ctl%solver_name = 'BICG'
   ctl%get_1D_from = 'Geometric_mean'
   call setEMsolveControl(ctl)

   allocate(txDict(1))
   txDict(1)%tx_type = 'MT'
   txDict(1)%nPol = 2
   txDict(1)%omega = om
   txDict(1)%period = 2*PI/om
   allocate(rxDict(nsite))
   do s = 1, nsite
      rxDict(s)%x = xs(:,s)
      rxDict(s)%r = xs(:,s)
   enddo
   call setup_typeDict()

   call create_dataVector(1, dv)
   call create_dataBlock(8, nsite, dv%data(1), .true.)
   dv%tx = 1
   dv%data(1)%tx = 1
   dv%data(1)%dataType = Full_Impedance
   dv%data(1)%value = d
   dv%data(1)%exist = .true.
   do s = 1, nsite
      dv%data(1)%rx(s) = s
      dv%data(1)%orient(s)%azimuth%Hx = 0d0
      dv%data(1)%orient(s)%azimuth%Hy = 90d0
      dv%data(1)%orient(s)%azimuth%Hz = 0d0
      dv%data(1)%orient(s)%azimuth%Ex = 0d0
      dv%data(1)%orient(s)%azimuth%Ey = 90d0
      dv%data(1)%orient(s)%azimuth%Hx_ref = 0d0
      dv%data(1)%orient(s)%azimuth%Hy_ref = 90d0
   enddo
   dv%allocated = .true.

   ! the chain of JmultT_TX, step by step
   call initSolver(1, m, grid, e0, e, comb)
   call fwdSetup(1, e0, b0)
   call fwdSolve(1, e0, b0)
   call LmultT(e0, m, dv, comb)
   call sensSolve(1, TRN, e, comb)
   dsig = m
   call zero(dsig)
   call PmultT(e0, m, e, dsig)

   call wr_open(trim(outdir)//'/fwd.bin', fout)
   do k = 1, 2
      write(fout) e0%pol(k)%x, e0%pol(k)%y, e0%pol(k)%z
   enddo
   close(fout)
   call wr_open(trim(outdir)//'/comb.bin', fout)
   do k = 1, 2
      write(fout) comb%b(k)%s%x, comb%b(k)%s%y, comb%b(k)%s%z
   enddo
   close(fout)
   call wr_open(trim(outdir)//'/trn.bin', fout)
   do k = 1, 2
      write(fout) e%pol(k)%x, e%pol(k)%y, e%pol(k)%z
   enddo
   close(fout)
   call wr_open(trim(outdir)//'/dsig.bin', fout)
   call write_cells(dsig, fout)
   close(fout)

   ! the reference's own JmultT_TX on the same transmitter and data
   call setGrid(grid)
   dsig_jt = m
   call zero(dsig_jt)
   call JmultT_TX(m, dv, dsig_jt)
   call wr_open(trim(outdir)//'/dsig_jt.bin', fout)
   call write_cells(dsig_jt, fout)
   close(fout)

contains

   subroutine wr_open(fname, unit)
      character(len=*), intent(in) :: fname
      integer, intent(out)         :: unit
      open(newunit=unit, file=fname, access='stream', form='unformatted', status='replace')
   end subroutine

   subroutine write_cells(mm, w)
      type(modelParam_t), intent(in) :: mm
      integer, intent(in)            :: w
      type(rscalar)                  :: cv
      character(len=80)              :: ptype
      real(kind=prec)                :: va
      ptype = LOGE
      call getValue_modelParam(mm, ptype, cv, va)
      write(w) cv%v
   end subroutine

end program ref_jmultT
