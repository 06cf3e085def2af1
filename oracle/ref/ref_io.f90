! Reference-side driver for the file formats (test infrastructure, never
! shipped): runs the reference's own readers and writers, compiled from the
! sources where they lie (recipe: oracle/ref/Makefile), and dumps what they
! read to raw binary.
!
!   ref_io ws   <model file> <out> [method nz maxheight_km mintopdz_km alpha]
!                                     read_modelParam_WS (3D_MT/modelParam/
!                                     modelParamIO/WS.inc); the air layers
!                                     (GridDef.f90:93-108, 406-488) by default
!                                     or as given, as a control file sets them
!       out: int32 nx, ny, nzAir, nzEarth ; real64 dx, dy, dz (all layers),
!            ox, oy, oz, rotdeg ; real64 LOGE cell values (Fortran order),
!            AirCond ; char(80) the file's parameter type
!   ref_io solnw <in> <soln file>     FileWriteInit + EfileWrite
!                                     (3D_MT/ioMod/ioAscii.f90), the records
!                                     write_solnVectorMTX writes
!       in:  int32 nx, ny, nzAir, nzEarth ; real64 dx, dy, dz, ox, oy, oz, rotdeg ;
!            int32 nPer ; nPer x { real64 omega ; 2 x complex128 E (x, y, z edges) }
!   ref_io zlist <data file> <out>    read_dataVectorMTX = read_Z_list
!                                     (3D_MT/DataIO.f90:552-1209)
!       out: int32 nTx ; real64 periods ; int32 nRx ; nRx x { char(50) id ; real64 x(3) } ;
!            int32 nd ; nd x { int32 tx, nDt ; nDt x { int32 dataType, nComp, nSite,
!            isComplex ; int32 rx(nSite) ; real64 value(nComp, nSite), error ;
!            int32 exist ; real64 (Hx, Hy, Ex, Ey, Hx_ref, Hy_ref)(nSite) } }
!   ref_io solnr <soln file> <out>    FileReadInit + EfileRead (ioAscii.f90)
!       out: int32 nPer, nMode, nx, ny, nz, nzAir ; real64 ox, oy, oz, rotdeg,
!            dx, dy, dz ; nPer*nMode x { real64 omega ; complex128 E }
program ref_io
   use math_constants
   use griddef
   use sg_scalar
   use sg_vector
   use ModelSpace
   use file_units
   use ioAscii
   use dataspace
   use transmitters
   use receivers
   use DataIO
   implicit none

   character(len=512)           :: cmd, a1, a2, a3
   real(kind=prec)              :: rv
   type(dataVectorMTX_t)        :: alld
   integer                      :: i, q
   integer                      :: fin, fo, nx, ny, nza, nze, nper, j, k, ios, fper, fmode
   type(grid_t), target         :: grid
   type(airLayers_t)            :: air
   type(modelParam_t)           :: m
   type(rscalar)                :: v
   type(cvector)                :: e
   real(kind=prec)              :: vair, om
   character(80)                :: ptype, fname
   character(len=20)            :: version
   character(len=20), parameter :: names(2) = (/ 'Ex', 'Ey' /)

   call get_command_argument(1, cmd)
   call get_command_argument(2, a1)
   call get_command_argument(3, a2)
   select case (trim(cmd))
   case ('ws')
      if (command_argument_count() >= 8) then
         call get_command_argument(4, a3)
         air%method = trim(a3)
         call get_command_argument(5, a3)
         read(a3, *) air%Nz
         call get_command_argument(6, a3)
         read(a3, *) rv
         air%MaxHeight = 1000. * rv
         call get_command_argument(7, a3)
         read(a3, *) rv
         air%MinTopDz = 1000. * rv
         call get_command_argument(8, a3)
         read(a3, *) air%alpha
      end if
      call read_modelParam(grid, air, m, trim(a1))
      ptype = 'LOGE'
      call getValue_modelParam(m, ptype, v, vair)
      open(newunit=fo, file=trim(a2), access='stream', form='unformatted', status='replace')
      write(fo) int(grid%nx, 4), int(grid%ny, 4), int(grid%nzAir, 4), int(grid%nzEarth, 4)
      write(fo) grid%dx, grid%dy, grid%dz, grid%ox, grid%oy, grid%oz, grid%rotdeg
      write(fo) v%v, vair
      ptype = userParamType
      write(fo) ptype
      close(fo)
   case ('solnw')
      open(newunit=fin, file=trim(a1), access='stream', form='unformatted', status='old')
      read(fin) nx, ny, nza, nze
      call create_grid(nx, ny, nza, nze, grid)
      read(fin) grid%dx, grid%dy, grid%dz, grid%ox, grid%oy, grid%oz, grid%rotdeg
      call setup_grid(grid)
      read(fin) nper
      fname = a2
      version = ''
      call FileWriteInit(version, fname, ioE, grid, nper, 2, ios)
      call create_cvector(grid, e, EDGE)
      do j = 1, nper
         read(fin) om
         do k = 1, 2
            read(fin) e%x, e%y, e%z
            call EfileWrite(ioE, om, j, k, names(k), e)
         end do
      end do
      close(ioE)
      close(fin)
   case ('solnr')
      fname = a1
      call FileReadInit(fname, ioE, grid, fper, fmode, version, ios)
      call setup_grid(grid)
      call create_cvector(grid, e, EDGE)
      open(newunit=fo, file=trim(a2), access='stream', form='unformatted', status='replace')
      write(fo) int(fper, 4), int(fmode, 4), int(grid%nx, 4), int(grid%ny, 4), int(grid%nz, 4), &
                int(grid%nzAir, 4)
      write(fo) grid%ox, grid%oy, grid%oz, grid%rotdeg, grid%dx, grid%dy, grid%dz
      do j = 1, fper
         do k = 1, fmode
            call EfileRead(ioE, j, k, om, e)
            write(fo) om, e%x, e%y, e%z
         end do
      end do
      close(fo)
      close(ioE)
   case ('zlist')
      call read_dataVectorMTX(alld, trim(a1))
      open(newunit=fo, file=trim(a2), access='stream', form='unformatted', status='replace')
      write(fo) int(size(txDict), 4)
      write(fo) (txDict(i)%period, i = 1, size(txDict))
      write(fo) int(size(rxDict), 4)
      do i = 1, size(rxDict)
         write(fo) rxDict(i)%id, rxDict(i)%x
      end do
      write(fo) int(alld%nTx, 4)
      do i = 1, alld%nTx
         write(fo) int(alld%d(i)%tx, 4), int(alld%d(i)%nDt, 4)
         do j = 1, alld%d(i)%nDt
            associate (bl => alld%d(i)%data(j))
               write(fo) int(bl%dataType, 4), int(bl%nComp, 4), int(bl%nSite, 4), &
                         int(merge(1, 0, bl%isComplex), 4)
               write(fo) (int(bl%rx(k), 4), k = 1, bl%nSite)
               write(fo) bl%value, bl%error
               write(fo) ((int(merge(1, 0, bl%exist(q, k)), 4), q = 1, bl%nComp), k = 1, bl%nSite)
               write(fo) (bl%orient(k)%azimuth%Hx, bl%orient(k)%azimuth%Hy, bl%orient(k)%azimuth%Ex, &
                          bl%orient(k)%azimuth%Ey, bl%orient(k)%azimuth%Hx_ref, &
                          bl%orient(k)%azimuth%Hy_ref, k = 1, bl%nSite)
            end associate
         end do
      end do
      close(fo)
   case default
      write(0, *) 'usage: ref_io ws|solnw|solnr|zlist <in> <out>'
      stop 2
   end select
end program ref_io
