! mtsp2_fortran -- ISO_C_BINDING interface of libmtsp2 for the reference's
! Fortran host (ModEM SP2, f90/3D_MT/FWD_SP2).
!
! It plays the role cudaFortMap.f90 plays for the reference's CUDA path
! (f90/3D_MT/FWD_SP2/cudaFortMap.f90:93-1408): bind(C) interfaces to the
! C ABI of include/mtsp2.h, plus mtsp2_bicg, a subroutine with the argument
! meaning of cuBiCG(b, x, KSPiter, device_idx) (solver.f90:931-965) that
! takes the operator explicitly (AAii as the reference's 1-based CSR, the
! frequency-free diagonal VomegaMuSig/omega and omega), keeps one library
! context per device and re-uploads the operator only when it changes (an
! exact byte comparison against a host copy of the operator last uploaded).
module mtsp2_fortran
   use iso_c_binding
   implicit none
   private

   integer(c_int), parameter, public :: MT_OK = 0, MT_E_ARG = 1, MT_E_DEVICE = 2, &
        MT_E_OOM = 3, MT_E_NAN_RHS = 4, MT_E_BREAKDOWN = 5, MT_E_MAXIT = 6, &
        MT_E_STATE = 7, MT_E_STENCIL = 8, MT_E_PIVOT = 9
   integer(c_int), parameter, public :: MT_PC_DILU = 0, MT_PC_JACOBI = 1, MT_PC_NONE = 2, &
        MT_PC_DILU_RB = 3
   integer(c_int), parameter, public :: MT_PREC_FP64 = 0, MT_PREC_MIXED = 1

   type, bind(C), public :: mt_solve_opts
      integer(c_int) :: maxit
      real(c_double) :: tol
      integer(c_int) :: precond, precision, restart, best_iterate
      real(c_double) :: inner_tol
      integer(c_int) :: max_outer, op_transpose, chunk, use_graph
   end type

   type, bind(C), public :: mt_sys_stats
      integer(c_int) :: niter, nouter, status
      real(c_double) :: relres
   end type

   public :: mt_default_opts, mt_ctx_create, mt_ctx_destroy, mt_device_count, mt_set_grid, &
             mt_operator_set_csr, mt_bicg, mt_solve_batch, mt_apply, mt_precond_apply, &
             mt_divcor, mt_pmultT, mt_last_error_string, mt_receivers_set, mt_data_resp, mt_qmr, &
             mt_lmultT, mt_dev_alloc, mt_dev_free, mt_dev_copy, mt_assemble, mt_update_cond, &
             mt_bc_x0_dev, mt_rhs_fwd_dev
   public :: mtsp2_bicg, mtsp2_qmr, mtsp2_solve_batch, mtsp2_release, mtsp2_lmultT, &
             mtsp2_data_resp, mtsp2_bc_x0, mtsp2_operator_reset

   interface
      subroutine mt_default_opts(opts) bind(C, name='mt_default_opts')
         import :: mt_solve_opts
         type(mt_solve_opts), intent(out) :: opts
      end subroutine
      integer(c_int) function mt_ctx_create(device, ctx) bind(C, name='mt_ctx_create')
         import :: c_int, c_ptr
         integer(c_int), value :: device
         type(c_ptr), intent(out) :: ctx
      end function
      integer(c_int) function mt_ctx_destroy(ctx) bind(C, name='mt_ctx_destroy')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function
      type(c_ptr) function mt_last_error(ctx) bind(C, name='mt_last_error')
         import :: c_ptr
         type(c_ptr), value :: ctx
      end function
      integer(c_int) function mt_device_count(n) bind(C, name='mt_device_count')
         import :: c_int
         integer(c_int), intent(out) :: n
      end function
      integer(c_int) function mt_set_grid(ctx, nx, ny, nz) bind(C, name='mt_set_grid')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int), value :: nx, ny, nz
      end function
      integer(c_int) function mt_operator_set_csr(ctx, n, nnz, rowptr, col, base, val, vms1) &
            bind(C, name='mt_operator_set_csr')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: n, nnz, base
         integer(c_int), intent(in) :: rowptr(*), col(*)
         real(c_double), intent(in) :: val(*), vms1(*)
      end function
      integer(c_int) function mt_assemble(ctx, nx, ny, nzAir, nzEarth, dx, dy, dz, logcond, &
            air_logcond) bind(C, name='mt_assemble')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: nx, ny, nzAir, nzEarth
         real(c_double), intent(in) :: dx(*), dy(*), dz(*), logcond(*)
         real(c_double), value :: air_logcond
      end function
      ! updateCond (ForwardSolver.f90:261) on the device for the assembled grid
      integer(c_int) function mt_update_cond(ctx, logcond, air_logcond) &
            bind(C, name='mt_update_cond')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(in) :: logcond(*)
         real(c_double), value :: air_logcond
      end function
      ! ComputeBC -> BC_x0_WS (FWD_SP2/modelOperator3D.f90:292-317) for a device batch
      integer(c_int) function mt_bc_x0_dev(ctx, nsys, omegas, modes, e_dev) &
            bind(C, name='mt_bc_x0_dev')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx, e_dev
         integer(c_int), value :: nsys
         real(c_double), intent(in) :: omegas(*)
         integer(c_int), intent(in) :: modes(*)
      end function
      ! b = -CCib e_b and x0 = e(EDGEi) (EMsolve3D.f90:217-274) for a device batch
      integer(c_int) function mt_rhs_fwd_dev(ctx, nsys, e_dev, b_dev, x0_dev) &
            bind(C, name='mt_rhs_fwd_dev')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx, e_dev, b_dev, x0_dev
         integer(c_int), value :: nsys
      end function
      integer(c_int) function mt_bicg(ctx, omega, b, x, maxit, tol, precond, niter, rerr, failed) &
            bind(C, name='mt_bicg')
         import :: c_int, c_ptr, c_double, c_double_complex
         type(c_ptr), value :: ctx
         real(c_double), value :: omega, tol
         complex(c_double_complex), intent(in) :: b(*)
         complex(c_double_complex), intent(inout) :: x(*)
         integer(c_int), value :: maxit, precond
         integer(c_int), intent(out) :: niter, failed
         real(c_double), intent(out) :: rerr(*)
      end function
      integer(c_int) function mt_qmr(ctx, omega, b, x, maxit, tol, precond, niter, rerr, failed) &
            bind(C, name='mt_qmr')
         import :: c_int, c_ptr, c_double, c_double_complex
         type(c_ptr), value :: ctx
         real(c_double), value :: omega, tol
         complex(c_double_complex), intent(in) :: b(*)
         complex(c_double_complex), intent(inout) :: x(*)
         integer(c_int), value :: maxit, precond
         integer(c_int), intent(out) :: niter, failed
         real(c_double), intent(out) :: rerr(*)
      end function
      integer(c_int) function mt_solve_batch(ctx, nsys, omegas, b, x, opts, stats) &
            bind(C, name='mt_solve_batch')
         import :: c_int, c_ptr, c_double, c_double_complex, mt_solve_opts, mt_sys_stats
         type(c_ptr), value :: ctx
         integer(c_int), value :: nsys
         real(c_double), intent(in) :: omegas(*)
         complex(c_double_complex), intent(in) :: b(*)
         complex(c_double_complex), intent(inout) :: x(*)
         type(mt_solve_opts), intent(in) :: opts
         type(mt_sys_stats), intent(out) :: stats(*)
      end function
      integer(c_int) function mt_apply(ctx, omega, op, x, y) bind(C, name='mt_apply')
         import :: c_int, c_ptr, c_double, c_double_complex
         type(c_ptr), value :: ctx
         real(c_double), value :: omega
         integer(c_int), value :: op
         complex(c_double_complex), intent(in) :: x(*)
         complex(c_double_complex), intent(out) :: y(*)
      end function
      integer(c_int) function mt_precond_apply(ctx, omega, precond, x, y) &
            bind(C, name='mt_precond_apply')
         import :: c_int, c_ptr, c_double, c_double_complex
         type(c_ptr), value :: ctx
         real(c_double), value :: omega
         integer(c_int), value :: precond
         complex(c_double_complex), intent(in) :: x(*)
         complex(c_double_complex), intent(out) :: y(*)
      end function
      integer(c_int) function mt_divcor(ctx, omega, ei, phi0, maxit, tol, niter, relres) &
            bind(C, name='mt_divcor')
         import :: c_int, c_ptr, c_double, c_double_complex
         type(c_ptr), value :: ctx
         real(c_double), value :: omega, tol
         complex(c_double_complex), intent(inout) :: ei(*)
         type(c_ptr), value :: phi0
         integer(c_int), value :: maxit
         integer(c_int), intent(out) :: niter
         real(c_double), intent(out) :: relres
      end function
      integer(c_int) function mt_pmultT(ctx, omega, npol, e0, lam, dm, acc) bind(C, name='mt_pmultT')
         import :: c_int, c_ptr, c_double, c_double_complex
         type(c_ptr), value :: ctx
         real(c_double), value :: omega
         integer(c_int), value :: npol, acc
         complex(c_double_complex), intent(in) :: e0(*), lam(*)
         real(c_double), intent(inout) :: dm(*)
      end function
      integer(c_int) function mt_receivers_set(ctx, nx, ny, nzAir, nzEarth, dx, dy, dz, origin, &
            nsite, x, r, az) bind(C, name='mt_receivers_set')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: nx, ny, nzAir, nzEarth, nsite
         real(c_double), intent(in) :: dx(*), dy(*), dz(*), origin(3), x(*), r(*), az(*)
      end function
      integer(c_int) function mt_data_resp(ctx, dtype, nfreq, omegas, e_dev, resp) &
            bind(C, name='mt_data_resp')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx, e_dev
         integer(c_int), value :: dtype, nfreq
         real(c_double), intent(in) :: omegas(*)
         real(c_double), intent(out) :: resp(*)
      end function
      integer(c_int) function mt_lmultT(ctx, dtype, nfreq, omegas, e_dev, d, exist, comb_dev) &
            bind(C, name='mt_lmultT')
         import :: c_int, c_ptr, c_double, c_signed_char
         type(c_ptr), value :: ctx, e_dev, comb_dev
         integer(c_int), value :: dtype, nfreq
         real(c_double), intent(in) :: omegas(*), d(*)
         integer(c_signed_char), intent(in) :: exist(*)
      end function
      integer(c_int) function mt_dev_alloc(ctx, nbytes, p) bind(C, name='mt_dev_alloc')
         import :: c_int, c_ptr, c_size_t
         type(c_ptr), value :: ctx
         integer(c_size_t), value :: nbytes
         type(c_ptr), intent(out) :: p
      end function
      integer(c_int) function mt_dev_free(ctx, p) bind(C, name='mt_dev_free')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx, p
      end function
      integer(c_int) function mt_dev_copy(ctx, dst, src, nbytes, kind) bind(C, name='mt_dev_copy')
         import :: c_int, c_ptr, c_size_t
         type(c_ptr), value :: ctx, dst, src
         integer(c_size_t), value :: nbytes
         integer(c_int), value :: kind
      end function
      integer(c_size_t) function c_strlen(s) bind(C, name='strlen')
         import :: c_size_t, c_ptr
         type(c_ptr), value :: s
      end function
      integer(c_int) function c_memcmp(a, b, n) bind(C, name='memcmp')
         import :: c_int, c_ptr, c_size_t
         type(c_ptr), value :: a, b
         integer(c_size_t), value :: n
      end function
   end interface

   ! one context per device, with a host copy of the operator last uploaded
   ! to it (grid, AAii as 1-based CSR, vms1): a solve re-uploads the operator
   ! unless every byte is unchanged
   integer, parameter :: MAXDEV = 16
   type(c_ptr), save :: ctxs(0:MAXDEV-1) = c_null_ptr
   type :: op_copy
      logical :: valid = .false.
      integer :: nx = -1, ny = -1, nz = -1
      integer(c_int), allocatable :: row(:), col(:)
      real(c_double), allocatable :: val(:), vms1(:)
   end type
   type(op_copy), save, target :: ops(0:MAXDEV-1)

contains

   function mt_last_error_string(ctx) result(msg)
      type(c_ptr), intent(in) :: ctx
      character(len=:), allocatable :: msg
      type(c_ptr) :: p
      character(kind=c_char), pointer :: s(:)
      integer :: i, n
      p = mt_last_error(ctx)
      n = int(c_strlen(p))
      call c_f_pointer(p, s, [n])
      allocate(character(len=n) :: msg)
      do i = 1, n
         msg(i:i) = s(i)
      enddo
   end function

   ! Context for device_idx with the operator (AAii, vms1) uploaded.  The
   ! operator is re-uploaded when the grid, the CSR pattern, the values or
   ! vms1 differ in any byte from the copy kept of the last upload.
   subroutine get_ctx(device_idx, nx, ny, nz, row, col, val, vms1, ctx, ierr)
      integer, intent(in)                 :: device_idx, nx, ny, nz
      integer(c_int), intent(in), target  :: row(:), col(:)
      real(c_double), intent(in), target  :: val(:), vms1(:)
      type(c_ptr), intent(out)            :: ctx
      integer, intent(out)                :: ierr
      integer(c_int) :: rc
      ierr = MT_OK
      if (device_idx < 0 .or. device_idx >= MAXDEV) then
         ierr = MT_E_DEVICE
         return
      endif
      if (.not. c_associated(ctxs(device_idx))) then
         rc = mt_ctx_create(int(device_idx, c_int), ctxs(device_idx))
         if (rc /= MT_OK) then
            ierr = rc
            return
         endif
      endif
      ctx = ctxs(device_idx)
      if (same_operator(ops(device_idx), nx, ny, nz, row, col, val, vms1)) return
      ops(device_idx)%valid = .false.
      rc = mt_set_grid(ctx, int(nx, c_int), int(ny, c_int), int(nz, c_int))
      if (rc == MT_OK) rc = mt_operator_set_csr(ctx, int(size(row) - 1, c_int), int(size(col), c_int), &
                                               row, col, 1_c_int, val, vms1)
      if (rc /= MT_OK) then
         ierr = rc
         return
      endif
      ops(device_idx)%nx = nx
      ops(device_idx)%ny = ny
      ops(device_idx)%nz = nz
      ops(device_idx)%row = row
      ops(device_idx)%col = col
      ops(device_idx)%val = val
      ops(device_idx)%vms1 = vms1
      ops(device_idx)%valid = .true.
   end subroutine

   ! .true. when (grid, row, col, val, vms1) equal the kept copy byte for byte.
   logical function same_operator(o, nx, ny, nz, row, col, val, vms1)
      type(op_copy), intent(in), target   :: o
      integer, intent(in)                 :: nx, ny, nz
      integer(c_int), intent(in), target  :: row(:), col(:)
      real(c_double), intent(in), target  :: val(:), vms1(:)
      same_operator = .false.
      if (.not. o%valid) return
      if (o%nx /= nx .or. o%ny /= ny .or. o%nz /= nz) return
      if (size(o%row) /= size(row) .or. size(o%col) /= size(col) .or. &
          size(o%val) /= size(val) .or. size(o%vms1) /= size(vms1)) return
      if (.not. same_bytes(c_loc(o%row), c_loc(row), 4_c_size_t * size(row, kind=c_size_t))) return
      if (.not. same_bytes(c_loc(o%col), c_loc(col), 4_c_size_t * size(col, kind=c_size_t))) return
      if (.not. same_bytes(c_loc(o%val), c_loc(val), 8_c_size_t * size(val, kind=c_size_t))) return
      if (.not. same_bytes(c_loc(o%vms1), c_loc(vms1), 8_c_size_t * size(vms1, kind=c_size_t))) return
      same_operator = .true.
   end function

   logical function same_bytes(a, b, n)
      type(c_ptr), intent(in)       :: a, b
      integer(c_size_t), intent(in) :: n
      same_bytes = .true.
      if (n > 0) same_bytes = c_memcmp(a, b, n) == 0
   end function

   ! Drops the kept operator copy of device_idx: the next solve uploads the
   ! operator it is given (for callers that change AAii in place between
   ! solves and want to say so; the byte comparison catches that as well).
   subroutine mtsp2_operator_reset(device_idx)
      integer, intent(in) :: device_idx
      if (device_idx < 0 .or. device_idx >= MAXDEV) return
      ops(device_idx)%valid = .false.
      if (allocated(ops(device_idx)%row)) deallocate(ops(device_idx)%row)
      if (allocated(ops(device_idx)%col)) deallocate(ops(device_idx)%col)
      if (allocated(ops(device_idx)%val)) deallocate(ops(device_idx)%val)
      if (allocated(ops(device_idx)%vms1)) deallocate(ops(device_idx)%vms1)
   end subroutine

   ! Drop-in for cuBiCG (solver.f90:931-2009).  Arguments follow solverControl_t
   ! (solver.f90:25-36): on entry maxIt/tol, on exit niter, rerr(1:maxIt),
   ! failed.  x holds x0 on entry and the solution (or, when not converged,
   ! the smallest-residual iterate) on exit.  ierr /= 0 replaces the
   ! reference's `stop` on device errors.
   subroutine mtsp2_bicg(device_idx, nx, ny, nz, row, col, val, vms1, omega, b, x, &
                         maxIt, tol, niter, rerr, failed, ierr, precond)
      integer, intent(in)                     :: device_idx, nx, ny, nz, maxIt
      integer(c_int), intent(in)              :: row(:), col(:)
      real(c_double), intent(in)              :: val(:), vms1(:), omega, tol
      complex(c_double_complex), intent(in)   :: b(:)
      complex(c_double_complex), intent(inout):: x(:)
      integer, intent(out)                    :: niter, ierr
      real(c_double), intent(out)             :: rerr(:)
      logical, intent(out)                    :: failed
      integer, intent(in), optional           :: precond
      type(c_ptr) :: ctx
      integer(c_int) :: rc, nit, fl, pc
      call get_ctx(device_idx, nx, ny, nz, row, col, val, vms1, ctx, ierr)
      if (ierr /= MT_OK) return
      pc = MT_PC_DILU
      if (present(precond)) pc = int(precond, c_int)
      rc = mt_bicg(ctx, omega, b, x, int(maxIt, c_int), tol, pc, nit, rerr, fl)
      ierr = rc
      niter = nit
      failed = (fl /= 0)
   end subroutine

   ! Drop-in for QMR (solver.f90:140-370), the reference's default solver,
   ! with the arguments of mtsp2_bicg: x holds x0 on entry and the last
   ! iterate on exit, rerr(niter) the smallest relative residual, niter <
   ! maxIt means converged, failed the reference's breakdown flag.  ierr is
   ! MT_OK, MT_E_MAXIT or MT_E_BREAKDOWN (with failed set) or a device error.
   subroutine mtsp2_qmr(device_idx, nx, ny, nz, row, col, val, vms1, omega, b, x, &
                        maxIt, tol, niter, rerr, failed, ierr, precond)
      integer, intent(in)                     :: device_idx, nx, ny, nz, maxIt
      integer(c_int), intent(in)              :: row(:), col(:)
      real(c_double), intent(in)              :: val(:), vms1(:), omega, tol
      complex(c_double_complex), intent(in)   :: b(:)
      complex(c_double_complex), intent(inout):: x(:)
      integer, intent(out)                    :: niter, ierr
      real(c_double), intent(out)             :: rerr(:)
      logical, intent(out)                    :: failed
      integer, intent(in), optional           :: precond
      type(c_ptr) :: ctx
      integer(c_int) :: rc, nit, fl, pc
      call get_ctx(device_idx, nx, ny, nz, row, col, val, vms1, ctx, ierr)
      if (ierr /= MT_OK) return
      pc = MT_PC_DILU
      if (present(precond)) pc = int(precond, c_int)
      rc = mt_qmr(ctx, omega, b, x, int(maxIt, c_int), tol, pc, nit, rerr, fl)
      ierr = rc
      niter = nit
      failed = (fl /= 0)
   end subroutine

   ! Both polarizations and many frequencies at once (nsys systems sharing
   ! AAii): b, x are (n_int, nsys).
   subroutine mtsp2_solve_batch(device_idx, nx, ny, nz, row, col, val, vms1, omegas, b, x, &
                                opts, stats, ierr)
      integer, intent(in)                     :: device_idx, nx, ny, nz
      integer(c_int), intent(in)              :: row(:), col(:)
      real(c_double), intent(in)              :: val(:), vms1(:), omegas(:)
      complex(c_double_complex), intent(in)   :: b(:, :)
      complex(c_double_complex), intent(inout):: x(:, :)
      type(mt_solve_opts), intent(in)         :: opts
      type(mt_sys_stats), intent(out)         :: stats(:)
      integer, intent(out)                    :: ierr
      type(c_ptr) :: ctx
      call get_ctx(device_idx, nx, ny, nz, row, col, val, vms1, ctx, ierr)
      if (ierr /= MT_OK) return
      ierr = mt_solve_batch(ctx, int(size(omegas), c_int), omegas, b, x, opts, stats)
   end subroutine

   subroutine mtsp2_release()
      integer :: d
      integer(c_int) :: rc
      do d = 0, MAXDEV - 1
         if (c_associated(ctxs(d))) then
            rc = mt_ctx_destroy(ctxs(d))
            ctxs(d) = c_null_ptr
            call mtsp2_operator_reset(d)
         endif
      enddo
   end subroutine


   ! Library context for device_idx, created on first use.
   subroutine dev_ctx(device_idx, ctx, ierr)
      integer, intent(in)       :: device_idx
      type(c_ptr), intent(out)  :: ctx
      integer, intent(out)      :: ierr
      ierr = MT_OK
      if (device_idx < 0 .or. device_idx >= MAXDEV) then
         ierr = MT_E_DEVICE
         return
      endif
      if (.not. c_associated(ctxs(device_idx))) &
         ierr = mt_ctx_create(int(device_idx, c_int), ctxs(device_idx))
      ctx = ctxs(device_idx)
   end subroutine

   ! Fields of both modes onto the device: e_dev = [e1 ; e2].
   subroutine fields_up(ctx, e1, e2, e_dev, ierr)
      type(c_ptr), intent(in)                       :: ctx
      complex(c_double_complex), intent(in), target :: e1(:), e2(:)
      type(c_ptr), intent(out)                      :: e_dev
      integer, intent(out)                          :: ierr
      integer(c_size_t) :: nb
      nb = int(size(e1), c_size_t) * 16_c_size_t
      ierr = mt_dev_alloc(ctx, 2_c_size_t * nb, e_dev)
      if (ierr /= MT_OK) return
      ierr = mt_dev_copy(ctx, e_dev, c_loc(e1), nb, 0_c_int)
      if (ierr == MT_OK) ierr = mt_dev_copy(ctx, transfer(transfer(e_dev, 0_c_intptr_t) + &
                                            int(nb, c_intptr_t), e_dev), c_loc(e2), nb, 0_c_int)
   end subroutine

   ! Drop-in for dataResp (3D_MT/DataFunc.f90:43-462) over all sites of one
   ! data block: grid (cells, spacings, origin), receivers x / r (3, nsite),
   ! azimuths az (7, nsite, degrees), data type dtype, the fields of both
   ! modes on all edges; resp (ncomp, nsite) as dataResp's Resp per site.
   subroutine mtsp2_data_resp(device_idx, nx, ny, nzAir, nzEarth, dx, dy, dz, origin, x, r, az, &
                              dtype, omega, e1, e2, resp, ierr)
      integer, intent(in)                    :: device_idx, nx, ny, nzAir, nzEarth, dtype
      real(c_double), intent(in)             :: dx(:), dy(:), dz(:), origin(3), omega
      real(c_double), intent(in)             :: x(:,:), r(:,:), az(:,:)
      complex(c_double_complex), intent(in)  :: e1(:), e2(:)
      real(c_double), intent(out)            :: resp(:,:)
      integer, intent(out)                   :: ierr
      type(c_ptr) :: ctx, e_dev
      integer :: rc
      call dev_ctx(device_idx, ctx, ierr)
      if (ierr /= MT_OK) return
      ierr = mt_receivers_set(ctx, int(nx, c_int), int(ny, c_int), int(nzAir, c_int), &
                              int(nzEarth, c_int), dx, dy, dz, origin, int(size(x, 2), c_int), &
                              x, r, az)
      if (ierr /= MT_OK) return
      call fields_up(ctx, e1, e2, e_dev, ierr)
      if (ierr == MT_OK) ierr = mt_data_resp(ctx, int(dtype, c_int), 1_c_int, [omega], e_dev, resp)
      rc = mt_dev_free(ctx, e_dev)
   end subroutine

   ! Drop-in for LmultT (SENS/DataSens.f90:130-225) for one data block of one
   ! transmitter: the adjoint sources comb1, comb2 (all edges, modes 1 and 2)
   ! of data d (ncomp, nsite) with presence flags exist (ncomp, nsite), at the
   ! background fields e1, e2 of the two modes.
   subroutine mtsp2_lmultT(device_idx, nx, ny, nzAir, nzEarth, dx, dy, dz, origin, x, r, az, &
                           dtype, omega, e1, e2, d, exist, comb1, comb2, ierr)
      integer, intent(in)                            :: device_idx, nx, ny, nzAir, nzEarth, dtype
      real(c_double), intent(in)                     :: dx(:), dy(:), dz(:), origin(3), omega
      real(c_double), intent(in)                     :: x(:,:), r(:,:), az(:,:), d(:,:)
      logical, intent(in)                            :: exist(:,:)
      complex(c_double_complex), intent(in)          :: e1(:), e2(:)
      complex(c_double_complex), intent(out), target :: comb1(:), comb2(:)
      integer, intent(out)                           :: ierr
      type(c_ptr) :: ctx, e_dev, c_dev
      integer(c_signed_char), allocatable :: ex(:,:)
      integer(c_size_t) :: nb
      integer :: rc
      call dev_ctx(device_idx, ctx, ierr)
      if (ierr /= MT_OK) return
      ierr = mt_receivers_set(ctx, int(nx, c_int), int(ny, c_int), int(nzAir, c_int), &
                              int(nzEarth, c_int), dx, dy, dz, origin, int(size(x, 2), c_int), &
                              x, r, az)
      if (ierr /= MT_OK) return
      allocate(ex(size(exist, 1), size(exist, 2)))
      ex = merge(1_c_signed_char, 0_c_signed_char, exist)
      nb = int(size(e1), c_size_t) * 16_c_size_t
      call fields_up(ctx, e1, e2, e_dev, ierr)
      if (ierr == MT_OK) ierr = mt_dev_alloc(ctx, 2_c_size_t * nb, c_dev)
      if (ierr == MT_OK) ierr = mt_lmultT(ctx, int(dtype, c_int), 1_c_int, [omega], e_dev, d, ex, c_dev)
      if (ierr == MT_OK) ierr = mt_dev_copy(ctx, c_loc(comb1), c_dev, nb, 1_c_int)
      if (ierr == MT_OK) ierr = mt_dev_copy(ctx, c_loc(comb2), transfer(transfer(c_dev, &
                                            0_c_intptr_t) + int(nb, c_intptr_t), c_dev), nb, 1_c_int)
      rc = mt_dev_free(ctx, e_dev)
      rc = mt_dev_free(ctx, c_dev)
   end subroutine


   ! Drop-in for the ComputeBC -> BC_x0_WS calls of fwdSetup
   ! (FWD_SP2/modelOperator3D.f90:292-317, ForwardSolver.f90:333-462) for a
   ! batch of (omega, mode) systems: e(nEdge, nsys) receives what FWDsolve3D
   ! starts from -- E0 on the interior edges, the boundary conditions (setBC
   ! of BC) on the boundary edges.  The model is the LOGE cell model of the
   ! grid (logcond nx*ny*nzEarth, Fortran order, and the air value).
   subroutine mtsp2_bc_x0(device_idx, nx, ny, nzAir, nzEarth, dx, dy, dz, logcond, air_logcond, &
                          nsys, omegas, modes, e, ierr)
      integer, intent(in)                    :: device_idx, nx, ny, nzAir, nzEarth, nsys
      real(c_double), intent(in)             :: dx(:), dy(:), dz(:), logcond(:), air_logcond
      real(c_double), intent(in)             :: omegas(:)
      integer(c_int), intent(in)             :: modes(:)
      complex(c_double_complex), intent(out), target :: e(:,:)
      integer, intent(out)                   :: ierr
      type(c_ptr)       :: ctx, e_dev
      integer(c_size_t) :: nb
      call dev_ctx(device_idx, ctx, ierr)
      if (ierr /= MT_OK) return
      ! the context's operator is replaced: a later solve must upload its own
      ops(device_idx)%valid = .false.
      ierr = mt_assemble(ctx, int(nx, c_int), int(ny, c_int), int(nzAir, c_int), &
                         int(nzEarth, c_int), dx, dy, dz, logcond, air_logcond)
      if (ierr /= MT_OK) return
      nb = int(size(e, 1), c_size_t) * int(nsys, c_size_t) * 16_c_size_t
      ierr = mt_dev_alloc(ctx, nb, e_dev)
      if (ierr /= MT_OK) return
      ierr = mt_bc_x0_dev(ctx, int(nsys, c_int), omegas, modes, e_dev)
      if (ierr == MT_OK .or. ierr == MT_E_PIVOT) then
         if (mt_dev_copy(ctx, c_loc(e), e_dev, nb, 1_c_int) /= MT_OK) ierr = MT_E_DEVICE
      end if
      if (mt_dev_free(ctx, e_dev) /= MT_OK .and. ierr == MT_OK) ierr = MT_E_DEVICE
   end subroutine

end module mtsp2_fortran
