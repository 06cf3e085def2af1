! Reference-side dump driver for the data functionals (test infrastructure,
! never shipped).
!
! Links against the reference's own modules, compiled from the sources where
! they lie (recipe: oracle/ref/Makefile), and writes what they compute for one
! transmitter and one data type to raw binary files, for the oracle
! restatement (oracle/sp2_oracle.c) to be checked against.
!
! Reference routines driven here (paths relative to the reference f90/ tree):
!   grid set-up          3D_MT/GridDef.f90:114-165 create_grid, :301-399 setup_grid
!   dictionaries         3D_MT/DICT/transmitters.f90, receivers.f90,
!                        dataTypes.f90:87-222 setup_typeDict
!   data blocks          SENS/DataSpace.f90:207-287 create_dataBlock
!   impedances           3D_MT/DataFunc.f90:43-462 dataResp
!   adjoint source       SENS/DataSens.f90:130-225 LmultT (-> DataFunc.f90:465-869 Lrows)
!
! Input file layout (stream, native endianness), written by oracle/refdriver.py:
!   int32 nx, ny, nzAir, nzEarth
!   real64 dx(nx), dy(ny), dz(nzAir+nzEarth), origin(3), omega
!   int32 dataType, nsite, ncomp
!   real64 x(3,nsite), r(3,nsite), azimuth(7,nsite)   (Hx,Hy,Hz,Ex,Ey,Hx_ref,Hy_ref)
!   real64 d(ncomp,nsite) ; int32 exist(ncomp,nsite)
!   complex128 e1(nEdgeAll), e2(nEdgeAll)              ([x;y;z] edges, Fortran order)
! Output: resp.bin = real64 resp(ncomp,nsite); comb.bin = complex128 comb1, comb2.
program ref_datafunc
   use math_constants
   use griddef
   use sg_vector
   use ModelSpace
   use transmitters
   use receivers
   use dataTypes
   use DataSpace
   use SolnSpace
   use DataFunc
   use DataSens
   use fields_orientation
   implicit none

   character(len=512)              :: infile, outdir
   integer                         :: fin, fout, nx, ny, nzAir, nzEarth, nz, dt, nsite, ncomp
   integer                         :: s, nXe, nYe, nZe
   real(kind=prec), allocatable    :: dx(:), dy(:), dz(:), xs(:,:), rs(:,:), az(:,:), d(:,:)
   real(kind=prec), allocatable    :: resp(:,:)
   integer, allocatable            :: ex(:,:)
   complex(kind=prec), allocatable :: e1(:), e2(:)
   real(kind=prec)                 :: origin(3), omega, rbuf(12)
   complex(kind=prec)              :: binv(2,2)
   type(grid_t), target            :: grid
   type(modelParam_t)              :: sig0
   type(solnVector_t)              :: e0
   type(rhsVector_t)               :: comb
   type(dataVector_t)              :: dv
   type(orient_t)                  :: orient

   call get_command_argument(1, infile)
   call get_command_argument(2, outdir)
   open(newunit=fin, file=trim(infile), access='stream', form='unformatted', status='old')
   read(fin) nx, ny, nzAir, nzEarth
   nz = nzAir + nzEarth
   allocate(dx(nx), dy(ny), dz(nz))
   read(fin) dx, dy, dz, origin, omega
   read(fin) dt, nsite, ncomp
   allocate(xs(3,nsite), rs(3,nsite), az(7,nsite), d(ncomp,nsite), ex(ncomp,nsite))
   allocate(resp(ncomp,nsite))
   read(fin) xs, rs, az, d, ex
   nXe = nx*(ny+1)*(nz+1)
   nYe = (nx+1)*ny*(nz+1)
   nZe = (nx+1)*(ny+1)*nz
   allocate(e1(nXe+nYe+nZe), e2(nXe+nYe+nZe))
   read(fin) e1, e2
   close(fin)

   call create_grid(nx, ny, nzAir, nzEarth, grid)
   grid%dx = dx
   grid%dy = dy
   grid%dz = dz
   call setup_grid(grid, origin)

   allocate(txDict(1))
   txDict(1)%tx_type = 'MT'
   txDict(1)%nPol = 2
   txDict(1)%omega = omega
   txDict(1)%period = 2*PI/omega
   allocate(rxDict(nsite))
   do s = 1, nsite
      rxDict(s)%x = xs(:,s)
      rxDict(s)%r = rs(:,s)
   enddo
   call setup_typeDict()

   call create_solnVector(grid, 1, e0)
   e0%pol(1)%x = reshape(e1(1:nXe), (/nx, ny+1, nz+1/))
   e0%pol(1)%y = reshape(e1(nXe+1:nXe+nYe), (/nx+1, ny, nz+1/))
   e0%pol(1)%z = reshape(e1(nXe+nYe+1:), (/nx+1, ny+1, nz/))
   e0%pol(2)%x = reshape(e2(1:nXe), (/nx, ny+1, nz+1/))
   e0%pol(2)%y = reshape(e2(nXe+1:nXe+nYe), (/nx+1, ny, nz+1/))
   e0%pol(2)%z = reshape(e2(nXe+nYe+1:), (/nx+1, ny+1, nz/))

   call create_dataVector(1, dv)
   call create_dataBlock(ncomp, nsite, dv%data(1), typeDict(dt)%isComplex)
   dv%tx = 1
   dv%data(1)%tx = 1
   dv%data(1)%dataType = dt
   dv%data(1)%value = d
   dv%data(1)%exist = ex .ne. 0
   do s = 1, nsite
      dv%data(1)%rx(s) = s
      dv%data(1)%orient(s)%azimuth%Hx = az(1,s)
      dv%data(1)%orient(s)%azimuth%Hy = az(2,s)
      dv%data(1)%orient(s)%azimuth%Hz = az(3,s)
      dv%data(1)%orient(s)%azimuth%Ex = az(4,s)
      dv%data(1)%orient(s)%azimuth%Ey = az(5,s)
      dv%data(1)%orient(s)%azimuth%Hx_ref = az(6,s)
      dv%data(1)%orient(s)%azimuth%Hy_ref = az(7,s)
   enddo
   dv%allocated = .true.

   do s = 1, nsite
      orient = dv%data(1)%orient(s)
      rbuf = 0
      call dataResp(e0, sig0, dt, s, rbuf, orient, binv)
      resp(:,s) = rbuf(1:ncomp)
   enddo
   open(newunit=fout, file=trim(outdir)//'/resp.bin', access='stream', form='unformatted', &
        status='replace')
   write(fout) resp
   close(fout)

   comb%nonzero_source = .true.
   comb%sparse_source = .false.
   call create_rhsVector(grid, 1, comb)
   call LmultT(e0, sig0, dv, comb)
   open(newunit=fout, file=trim(outdir)//'/comb.bin', access='stream', form='unformatted', &
        status='replace')
   write(fout) comb%b(1)%s%x, comb%b(1)%s%y, comb%b(1)%s%z
   write(fout) comb%b(2)%s%x, comb%b(2)%s%y, comb%b(2)%s%z
   close(fout)
end program ref_datafunc
