! The data-functional drop-in, end to end (test infrastructure): the
! reference's own dataResp and LmultT (3D_MT/DataFunc.f90, SENS/DataSens.f90,
! compiled from the reference sources by oracle/ref/Makefile) and the
! replacements mtsp2_data_resp / mtsp2_lmultT of
! modem-gpu_amd/fortran/mtsp2_fortran.f90 on the same inputs, from one
! Fortran host.  Input layout: as oracle/ref/ref_datafunc.f90.  Output:
! resp.bin / comb.bin (reference, as ref_datafunc) and resp_lib.bin /
! comb_lib.bin (library), plus ierr.bin (int32 status of the library calls).
program ref_dropin_datafunc
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
   use mtsp2_fortran, only: mtsp2_lmultT, mtsp2_data_resp
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
   complex(kind=prec), allocatable :: g1(:), g2(:)
   real(kind=prec), allocatable    :: gresp(:,:)
   integer                         :: ierr(2)

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

   ! the library, from the same host arrays
   allocate(g1(size(e1)), g2(size(e1)), gresp(ncomp, nsite))
   call mtsp2_data_resp(0, nx, ny, nzAir, nzEarth, dx, dy, dz, origin, xs, rs, az, dt, omega, &
                        e1, e2, gresp, ierr(1))
   call mtsp2_lmultT(0, nx, ny, nzAir, nzEarth, dx, dy, dz, origin, xs, rs, az, dt, omega, &
                     e1, e2, d, ex .ne. 0, g1, g2, ierr(2))
   open(newunit=fout, file=trim(outdir)//'/resp_lib.bin', access='stream', form='unformatted', &
        status='replace')
   write(fout) gresp
   close(fout)
   open(newunit=fout, file=trim(outdir)//'/comb_lib.bin', access='stream', form='unformatted', &
        status='replace')
   write(fout) g1, g2
   close(fout)
   open(newunit=fout, file=trim(outdir)//'/ierr.bin', access='stream', form='unformatted', &
        status='replace')
   write(fout) ierr
   close(fout)
end program ref_dropin_datafunc
