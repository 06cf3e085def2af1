! Reference-side dump driver (test infrastructure, never shipped).
!
! Links against the reference's own SP2 modules, compiled from the sources
! where they lie (recipe: oracle/ref/Makefile), and writes what they compute
! for one synthetic grid/model to raw binary files.  The oracle restatement
! (oracle/sp2_oracle.c) and the HIP library are checked against these dumps.
!
! Reference routines driven here (paths relative to the reference f90/ tree):
!   grid set-up            3D_MT/GridDef.f90:114-165 create_grid, :301-399 setup_grid
!   model parameter        3D_MT/modelParam/ModelSpace.f90:175-215 create_modelParam
!   sigma -> edge/node     3D_MT/modelParam/ModelMap.inc:88-277
!   operator set-up        3D_MT/FWD_SP2/modelOperator3D.f90:69-117 ModelDataInit,
!                          :167-213 UpdateCond (-> GradDivSetup2 :916-1008),
!                          :155-164 UpdateFreq (-> updateOmegaMuSig, PC_setup :512-611)
!   operator apply         modelOperator3D.f90:415-436 Mult_Aii, :622-644 PC_Lsolve/PC_Usolve
!   CPU BiCGStab           3D_MT/FWD_SP2/solver.f90:720-926 BICG
!   CPU QMR                3D_MT/FWD_SP2/solver.f90:140-370 QMR
!   divergence correction  modelOperator3D.f90:648-730 DivCorInit/DivCorSetup,
!                          solver.f90:57-136 PCG
!   edge -> model adjoint  ModelMap.inc:281-372 dEdgeToModelParam
!
! Input file layout (stream, native endianness), written by oracle/refdriver.py:
!   int32 flags (bit 0: dump index sets, metrics and operators)
!   int32 nx, ny, nzAir, nzEarth
!   real64 dx(nx), dy(ny), dz(nzAir+nzEarth)
!   real64 logcond(nx*ny*nzEarth) (Fortran order), airLogCond
!   int32 nomega ; real64 omega(nomega)
!   int32 nsolve ; nsolve x { int32 iomega, maxit ; real64 tol ;
!                              complex128 b(nInt), x0(nInt) }
!   int32 napply ; napply x { int32 iomega ; complex128 x(nInt) }
!   int32 divcor ; if 1: { int32 maxit ; real64 tol ; complex128 rhs(nNodeInt) }
!   int32 e2m    ; if 1: { real64 e(nEdgeAll) }
!   int32 nqmr   ; nqmr x { int32 iomega, maxit ; real64 tol ;
!                           complex128 b(nInt), x0(nInt) }          (QMR solves)
!   int32 nadj   ; nadj x { int32 iomega ; complex128 x(nInt) }   (A^H x, L^-H x, U^-H x)
program ref_driver
   use math_constants
   use griddef
   use sg_scalar
   use sg_vector
   use gridcalc
   use ModelSpace
   use vecTranslate
   use spOpTools
   use spOpTopology_SG
   use MetricElements_SG
   use modelOperator3D
   use solver
   implicit none

   character(len=512)            :: infile, outdir
   integer                       :: fin, nx, ny, nzAir, nzEarth, nz, i, k, s
   integer                       :: nomega, nsolve, napply, idiv, ie2m, iom, maxit
   integer                       :: nqmr, nadj, ios
   real(kind=prec), allocatable  :: dx(:), dy(:), dz(:), logc(:), omegas(:)
   real(kind=prec)               :: airLog, tol
   type(grid_t), target          :: grid
   type(rscalar)                 :: v
   type(modelParam_t)            :: m, dm
   type(rvector)                 :: sigE, etmp
   type(rscalar)                 :: sigN
   real(kind=prec), pointer      :: pE(:), pN(:)
   complex(kind=prec), allocatable :: b(:), x(:), y(:), t1(:), t2(:), rhs(:), phi(:)
   real(kind=prec), allocatable  :: e(:)
   type(solverControl_t)         :: ctl
   integer                       :: nInt, nAll, nNi, flags
   integer(kind=8)               :: c0, c1, crate
   real(kind=8)                  :: secs
   character(len=16)             :: tag

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

   nInt = size(EDGEi)
   nAll = nInt + size(EDGEb)
   nNi = size(NODEi)

   call wr_open(trim(outdir)//'/sizes.bin', s)
   write(s) nAll, nInt, size(EDGEb), size(NODEi)+size(NODEb), nNi, size(NODEb)
   close(s)
   if (iand(flags, 1) .ne. 0) then
   ! index sets and metric vectors
   call wr_open(trim(outdir)//'/index.bin', s)
   write(s) nAll, nInt, size(EDGEb), size(NODEi)+size(NODEb), nNi, size(NODEb)
   write(s) EDGEi, EDGEb, NODEi, NODEb
   close(s)
   call wr_open(trim(outdir)//'/metric.bin', s)
   write(s) FaceA, EdgeL, DualFaceA, DualEdgeL, Vnode, Vedge
   close(s)

   ! edge / node conductivities exactly as UpdateCond forms them
   call create_rvector(grid, sigE, EDGE)
   call create_rscalar(grid, sigN, CORNER)
   call ModelParamToEdge(m, sigE)
   call ModelParamToNode(m, sigN)
   nullify(pE)
   nullify(pN)
   call getVector(sigE, pE)
   call getScalar(sigN, pN)
   call wr_open(trim(outdir)//'/sigma.bin', s)
   write(s) pE, pN
   close(s)

   call wr_csr_r(trim(outdir)//'/AAii.csr', AAii)
   call wr_csr_r(trim(outdir)//'/CCii.csr', CCii)
   call wr_csr_r(trim(outdir)//'/CCib.csr', CCib)
   call wr_csr_r(trim(outdir)//'/GDii.csr', GDii)
   call wr_csr_r(trim(outdir)//'/T.csr', T)
   call wr_csr_r(trim(outdir)//'/G.csr', G)
   endif

   read(fin) nomega
   allocate(omegas(nomega))
   read(fin) omegas
   do k = 1, nomega
      call UpdateFreq(omegas(k))
      write(tag, '(I0)') k
      call wr_open(trim(outdir)//'/vms_'//trim(tag)//'.bin', s)
      write(s) VomegaMuSig
      close(s)
      if (iand(flags, 1) .ne. 0) then
         call wr_csr_c(trim(outdir)//'/L_'//trim(tag)//'.csr', L)
         call wr_csr_c(trim(outdir)//'/U_'//trim(tag)//'.csr', U)
      endif
   enddo

   allocate(b(nInt), x(nInt), y(nInt), t1(nInt), t2(nInt))

   ! BiCGStab solves (CPU reference path)
   read(fin) nsolve
   do k = 1, nsolve
      read(fin) iom, maxit
      read(fin) tol
      read(fin) b, x
      call UpdateFreq(omegas(iom))
      ctl%maxIt = maxit
      ctl%tol = tol
      ctl%niter = 0
      ctl%failed = .false.
      allocate(ctl%rerr(maxit))
      ctl%rerr = 0.0
      call system_clock(c0, crate)
      call BICG(b, x, ctl)
      call system_clock(c1)
      secs = dble(c1 - c0) / dble(crate)
      write(tag, '(I0)') k
      call wr_open(trim(outdir)//'/bicg_'//trim(tag)//'.bin', s)
      i = 0
      if (ctl%failed) i = 1
      write(s) ctl%niter, i, maxit
      write(s) secs
      write(s) ctl%rerr
      write(s) x
      close(s)
      deallocate(ctl%rerr)
   enddo

   ! operator / preconditioner applications
   read(fin) napply
   do k = 1, napply
      read(fin) iom
      read(fin) x
      call UpdateFreq(omegas(iom))
      call Mult_Aii(x, .false., y)
      call PC_Lsolve(x, .false., t1)
      call PC_Usolve(t1, .false., t2)
      write(tag, '(I0)') k
      call wr_open(trim(outdir)//'/apply_'//trim(tag)//'.bin', s)
      write(s) y, t1, t2
      close(s)
   enddo

   ! divergence correction operators and PCG (compiled but unused in SP2)
   read(fin) idiv
   if (idiv .eq. 1) then
      read(fin) maxit
      read(fin) tol
      allocate(rhs(nNi), phi(nNi))
      read(fin) rhs
      call DivCorInit()
      call DivCorSetup()
      call wr_csr_r(trim(outdir)//'/VDiv.csr', VDiv)
      call wr_csr_r(trim(outdir)//'/VDs.csr', VDs)
      call wr_csr_r(trim(outdir)//'/Gai.csr', Gai)
      call wr_csr_r(trim(outdir)//'/VDsG.csr', VDsG)
      call wr_csr_r(trim(outdir)//'/VDsG_L.csr', VDsG_L)
      call wr_csr_r(trim(outdir)//'/VDsG_U.csr', VDsG_U)
      ctl%maxIt = maxit
      ctl%tol = tol
      ctl%niter = 0
      allocate(ctl%rerr(maxit))
      ctl%rerr = 0.0
      phi = C_ZERO
      call PCG(rhs, phi, ctl)
      call wr_open(trim(outdir)//'/pcg.bin', s)
      write(s) ctl%niter, maxit
      write(s) ctl%rerr
      write(s) phi
      close(s)
      deallocate(ctl%rerr)
   endif

   ! adjoint of the sigma -> edge map (gradient assembly back end)
   read(fin) ie2m
   if (ie2m .eq. 1) then
      allocate(e(nAll))
      read(fin) e
      call create_rvector(grid, etmp, EDGE)
      call setVector(e, etmp)
      call create_modelParam(grid, LOGE, dm)
      call dEdgeToModelParam(etmp, dm, m)
      call wr_open(trim(outdir)//'/e2m.bin', s)
      call getValue_cells(dm, s)
      close(s)
   endif

   ! QMR solves (CPU reference path, solver.f90:140-370) and the adjoint
   ! applications QMR uses (Mult_Aii adjt, PC_Lsolve/PC_Usolve adjt,
   ! modelOperator3D.f90:415-436, 622-644)
   read(fin, iostat=ios) nqmr
   if (ios .ne. 0) nqmr = 0
   do k = 1, nqmr
      read(fin) iom, maxit
      read(fin) tol
      read(fin) b, x
      call UpdateFreq(omegas(iom))
      ctl%maxIt = maxit
      ctl%tol = tol
      ctl%niter = 0
      ctl%failed = .false.
      allocate(ctl%rerr(maxit))
      ctl%rerr = 0.0
      call QMR(b, x, ctl)
      write(tag, '(I0)') k
      call wr_open(trim(outdir)//'/qmr_'//trim(tag)//'.bin', s)
      i = 0
      if (ctl%failed) i = 1
      write(s) ctl%niter, i, maxit
      write(s) ctl%rerr
      write(s) x
      close(s)
      deallocate(ctl%rerr)
   enddo
   read(fin, iostat=ios) nadj
   if (ios .ne. 0) nadj = 0
   do k = 1, nadj
      read(fin) iom
      read(fin) x
      call UpdateFreq(omegas(iom))
      call Mult_Aii(x, .true., y)
      call PC_Lsolve(x, .true., t1)
      call PC_Usolve(x, .true., t2)
      write(tag, '(I0)') k
      call wr_open(trim(outdir)//'/adj_'//trim(tag)//'.bin', s)
      write(s) y, t1, t2
      close(s)
   enddo
   close(fin)

contains

   subroutine wr_open(fname, unit)
      character(len=*), intent(in) :: fname
      integer, intent(out)         :: unit
      open(newunit=unit, file=fname, access='stream', form='unformatted', status='replace')
   end subroutine

   subroutine wr_csr_r(fname, A)
      character(len=*), intent(in)    :: fname
      type(spMatCSR_Real), intent(in) :: A
      integer :: w, nnz
      nnz = A%row(A%nRow+1) - 1
      call wr_open(fname, w)
      write(w) A%nRow, A%nCol, nnz
      write(w) A%row(1:A%nRow+1), A%col(1:nnz), A%val(1:nnz)
      close(w)
   end subroutine

   subroutine wr_csr_c(fname, A)
      character(len=*), intent(in)     :: fname
      type(spMatCSR_Cmplx), intent(in) :: A
      integer :: w, nnz
      nnz = A%row(A%nRow+1) - 1
      call wr_open(fname, w)
      write(w) A%nRow, A%nCol, nnz
      write(w) A%row(1:A%nRow+1), A%col(1:nnz), A%val(1:nnz)
      close(w)
   end subroutine

   subroutine getValue_cells(mm, w)
      type(modelParam_t), intent(in) :: mm
      integer, intent(in)            :: w
      type(rscalar)                  :: cv
      character(len=80)              :: ptype
      real(kind=prec)                :: va
      ptype = LOGE
      call getValue_modelParam(mm, ptype, cv, va)
      write(w) cv%v
   end subroutine

end program ref_driver
