"""Python host binding of libmtsp2 (the MI355X SP2 3D-MT solver library).

The product boundary is the C ABI in include/mtsp2.h; the reference's own host
is Fortran and binds it through modem-gpu_amd/fortran/mtsp2_fortran.f90.  This
module is the same binding for Python callers (tests, benchmark): ctypes over
plain pointers, mirroring the reference's solver interface
(cuBiCG(b, x, KSPiter, device_idx), f90/3D_MT/FWD_SP2/solver.f90:931-965).

The library must be built (``make -C modem-gpu_amd``); loading fails loudly
otherwise -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(HERE), "lib", "libmtsp2.so")

MT_OK, MT_E_ARG, MT_E_DEVICE, MT_E_OOM, MT_E_NAN_RHS = 0, 1, 2, 3, 4
MT_E_BREAKDOWN, MT_E_MAXIT, MT_E_STATE, MT_E_STENCIL, MT_E_PIVOT = 5, 6, 7, 8, 9
PC_DILU, PC_JACOBI, PC_NONE, PC_DILU_RB = 0, 1, 2, 3
PREC_FP64, PREC_MIXED = 0, 1
KT_SPMV, KT_PRECOND = 0, 1
# data types, dataTypes.f90:65-77 numbering
(DT_FULL_IMPEDANCE, DT_OFF_DIAGONAL_IMPEDANCE, DT_FULL_VERTICAL_COMPONENTS,
 DT_FULL_INTERSTATION_TF, DT_OFF_DIAGONAL_RHO_PHASE, DT_PHASE_TENSOR, DT_EX_FIELD, DT_EY_FIELD,
 DT_BX_FIELD, DT_BY_FIELD, DT_BZ_FIELD) = range(1, 12)
DT_NCOMP = {1: 8, 2: 4, 3: 4, 4: 8, 5: 4, 6: 4, 7: 2, 8: 2, 9: 2, 10: 2, 11: 2}


class MTError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"libmtsp2 error {code}: {msg}")
        self.code = code


class SolveOpts(C.Structure):
    _fields_ = [("maxit", C.c_int), ("tol", C.c_double), ("precond", C.c_int),
                ("precision", C.c_int), ("restart", C.c_int), ("best_iterate", C.c_int),
                ("inner_tol", C.c_double), ("max_outer", C.c_int), ("op_transpose", C.c_int),
                ("chunk", C.c_int), ("use_graph", C.c_int)]


class SysStats(C.Structure):
    _fields_ = [("niter", C.c_int), ("nouter", C.c_int), ("status", C.c_int),
                ("relres", C.c_double)]


_lib = None
_P = C.c_void_p
_I = C.c_int
_D = C.c_double
_PI = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_PD = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_PZ = np.ctypeslib.ndpointer(np.complex128, flags="C_CONTIGUOUS")

_SIGS = {
    "mt_default_opts": (None, [C.POINTER(SolveOpts)]),
    "mt_ctx_create": (_I, [_I, C.POINTER(_P)]),
    "mt_ctx_destroy": (_I, [_P]),
    "mt_last_error": (C.c_char_p, [_P]),
    "mt_device_count": (_I, [C.POINTER(_I)]),
    "mt_set_grid": (_I, [_P, _I, _I, _I]),
    "mt_operator_set_csr": (_I, [_P, _I, _I, _PI, _PI, _I, _PD, _PD]),
    "mt_assemble": (_I, [_P, _I, _I, _I, _I, _PD, _PD, _PD, _PD, _D]),
    "mt_update_cond": (_I, [_P, _PD, _D]),
    "mt_host_assemble": (_I, [_I, _I, _I, _I, _PD, _PD, _PD, _PD, _D, C.POINTER(_I),
                              C.POINTER(_I), _P, _P, _P, _P]),
    "mt_host_dilu_pivots": (_I, [_I, _I, _I, _I, _I, _PI, _PI, _I, _PD, _PD, _D, _PZ]),
    "mt_operator_size": (_I, [_P, C.POINTER(_I), C.POINTER(_I), C.POINTER(_I)]),
    "mt_operator_get_csr": (_I, [_P, _PI, _PI, _PD, _PD]),
    "mt_bicg": (_I, [_P, _D, _PZ, _PZ, _I, _D, _I, C.POINTER(_I), _PD, C.POINTER(_I)]),
    "mt_qmr": (_I, [_P, _D, _PZ, _PZ, _I, _D, _I, C.POINTER(_I), _PD, C.POINTER(_I)]),
    "mt_solve_batch": (_I, [_P, _I, _PD, _PZ, _PZ, C.POINTER(SolveOpts), C.POINTER(SysStats)]),
    "mt_solve_batch_dev": (_I, [_P, _I, _PD, _P, _P, C.POINTER(SolveOpts), C.POINTER(SysStats)]),
    "mt_dev_alloc": (_I, [_P, C.c_size_t, C.POINTER(_P)]),
    "mt_dev_free": (_I, [_P, _P]),
    "mt_dev_copy": (_I, [_P, _P, _P, C.c_size_t, _I]),
    "mt_apply": (_I, [_P, _D, _I, _PZ, _PZ]),
    "mt_precond_apply": (_I, [_P, _D, _I, _PZ, _PZ]),
    "mt_rhs_fwd": (_I, [_P, _PZ, _PZ]),
    "mt_rhs_fwd_dev": (_I, [_P, _I, _P, _P, _P]),
    "mt_bc_x0_dev": (_I, [_P, _I, _PD, _PI, _P]),
    "mt_rhs_trn": (_I, [_P, _D, _PZ, _PZ]),
    "mt_post_trn": (_I, [_P, _PZ, _PZ]),
    "mt_post_fwd_dev": (_I, [_P, _I, _P, _P]),
    "mt_rhs_trn_dev": (_I, [_P, _I, _PD, _P, _P, _I]),
    "mt_post_trn_dev": (_I, [_P, _I, _P, _P]),
    "mt_pmultT_dev": (_I, [_P, _I, _PD, _I, _P, _P, _P, _I]),
    "mt_gradient_dev": (_I, [_P, _I, _PD, _I, _P, _P, _P, _I, _I]),
    "mt_divcor": (_I, [_P, _D, _PZ, _P, _I, _D, C.POINTER(_I), C.POINTER(_D)]),
    "mt_pmultT": (_I, [_P, _D, _I, _PZ, _PZ, _PD, _I]),
    "mt_gradient": (_I, [_P, _D, _I, _PZ, _PZ, _PD, _I]),
    "mt_host_gradient": (_I, [_I, _I, _I, _I, _PD, _PD, _PD, _PD, _D, _D, _I, _PZ, _PZ, _PD, _I]),
    "mt_autotune": (_I, [_P, _I, C.POINTER(_I)]),
    "mt_set_tuning": (_I, [_P, _I]),
    "mt_bench_spmv": (_I, [_P, _I, _I, _I, C.POINTER(_D), C.POINTER(_D)]),
    "mt_timing_enable": (_I, [_P, _I]),
    "mt_set_spmv_variant": (_I, [_P, _I, _I]),
    "mt_set_dilu_variant": (_I, [_P, _I]),
    "mt_set_inner_layout": (_I, [_P, _I, _I]),
    "mt_inner_apply": (_I, [_P, _D, _I, _I, _PZ, _PZ]),
    "mt_bench_precond": (_I, [_P, _I, _I, _I, _I, C.POINTER(_D), C.POINTER(_D)]),
    "mt_kernel_timing": (_I, [_P, _I, C.POINTER(C.c_longlong), C.POINTER(_D), C.POINTER(_D)]),
    "mt_receivers_set": (_I, [_P, _I, _I, _I, _I, _PD, _PD, _PD, _P, _I, _PD, _P, _P]),
    "mt_data_resp": (_I, [_P, _I, _I, _PD, _P, _PD]),
    "mt_lmultT": (_I, [_P, _I, _I, _PD, _P, _PD, _P, _P]),
    "mt_lmultT_noalloc": (_I, [_P, _I, _I, _PD, _P, _PD, _P, _P]),
}


def exported_symbols():
    """Names the C ABI exports (include/mtsp2.h)."""
    return sorted(_SIGS)


def lib():
    """Loads libmtsp2.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.isfile(LIB_PATH):
            raise RuntimeError(f"libmtsp2 not built: {LIB_PATH} missing "
                               "(run `make -C modem-gpu_amd` or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def device_count():
    n = _I(0)
    lib().mt_device_count(C.byref(n))
    return n.value


def default_opts(**kw):
    o = SolveOpts()
    lib().mt_default_opts(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def host_assemble(dx, dy, dz, nzAir, logcond, air_log=np.log(1e-10)):
    """Host-only assembly (no GPU): AAii as 0-based interior CSR plus vms1."""
    dx = np.ascontiguousarray(dx, np.float64)
    dy = np.ascontiguousarray(dy, np.float64)
    dz = np.ascontiguousarray(dz, np.float64)
    nze = len(dz) - int(nzAir)
    lc = np.ascontiguousarray(np.asarray(logcond, np.float64).ravel(order="F"))
    ni, nz = _I(), _I()
    args = (len(dx), len(dy), int(nzAir), nze, dx, dy, dz, lc, float(air_log))
    rc = lib().mt_host_assemble(*args, C.byref(ni), C.byref(nz), None, None, None, None)
    if rc:
        raise MTError(rc, "mt_host_assemble failed")
    row = np.zeros(ni.value + 1, np.int32)
    col = np.zeros(nz.value, np.int32)
    val = np.zeros(nz.value)
    vms = np.zeros(ni.value)
    rc = lib().mt_host_assemble(*args, C.byref(ni), C.byref(nz), row.ctypes.data, col.ctypes.data,
                                val.ctypes.data, vms.ctypes.data)
    if rc:
        raise MTError(rc, "mt_host_assemble failed")
    return {"row": row, "col": col, "val": val, "vms1": vms, "nrow": ni.value}


def host_dilu_pivots(grid_cells, row, col, val, vms1, omega, base=0):
    """Host-only inverse block-DILU pivots (interior order)."""
    nx, ny, nz = grid_cells
    n = len(row) - 1
    out = np.zeros(n, np.complex128)
    rc = lib().mt_host_dilu_pivots(nx, ny, nz, n, int(len(col)), np.ascontiguousarray(row, np.int32),
                                   np.ascontiguousarray(col, np.int32), int(base),
                                   np.ascontiguousarray(val, np.float64),
                                   np.ascontiguousarray(vms1, np.float64), float(omega), out)
    if rc:
        raise MTError(rc, "mt_host_dilu_pivots failed")
    return out


GRAD_REFERENCE, GRAD_EXACT = 0, 1


def host_gradient(dx, dy, dz, nzAir, logcond, omega, e_all, lam, mode=GRAD_EXACT,
                  air_log=np.log(1e-10)):
    """Host-only gradient d/dm Re(sum lam^T A e) (see mt_gradient)."""
    dx = np.ascontiguousarray(dx, np.float64)
    dy = np.ascontiguousarray(dy, np.float64)
    dz = np.ascontiguousarray(dz, np.float64)
    lc = np.asarray(logcond, np.float64)
    e = np.ascontiguousarray(np.atleast_2d(e_all), np.complex128)
    la = np.ascontiguousarray(np.atleast_2d(lam), np.complex128)
    dm = np.zeros(lc.size)
    rc = lib().mt_host_gradient(len(dx), len(dy), int(nzAir), len(dz) - int(nzAir), dx, dy, dz,
                                np.ascontiguousarray(lc.ravel(order="F")), float(air_log),
                                float(omega), e.shape[0], e, la, dm, int(mode))
    if rc:
        raise MTError(rc, "mt_host_gradient failed")
    return dm.reshape(lc.shape, order="F")


class Context:
    """One device, one stream, one device-resident operator."""

    def __init__(self, device=0):
        h = _P()
        rc = lib().mt_ctx_create(int(device), C.byref(h))
        if rc != MT_OK:
            raise MTError(rc, f"mt_ctx_create(device={device}) failed")
        self._h = h
        self.n_int = None
        self.n_edge = None

    def close(self):
        if getattr(self, "_h", None):
            lib().mt_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def _chk(self, rc):
        if rc != MT_OK:
            raise MTError(rc, lib().mt_last_error(self._h).decode())

    def _sizes(self):
        ni, ne, nz = _I(), _I(), _I()
        self._chk(lib().mt_operator_size(self._h, C.byref(ni), C.byref(ne), C.byref(nz)))
        self.n_int, self.n_edge, self.nnz = ni.value, ne.value, nz.value

    # -- operator ---------------------------------------------------------
    def set_operator_csr(self, grid_cells, n, rowptr, col, val, vms1, base=1):
        nx, ny, nz = grid_cells
        self._chk(lib().mt_set_grid(self._h, nx, ny, nz))
        rowptr = np.ascontiguousarray(rowptr, np.int32)
        col = np.ascontiguousarray(col, np.int32)
        self._chk(lib().mt_operator_set_csr(self._h, int(n), int(len(col)), rowptr, col, int(base),
                                            np.ascontiguousarray(val, np.float64),
                                            np.ascontiguousarray(vms1, np.float64)))
        self._sizes()

    def assemble(self, dx, dy, dz, nzAir, logcond, air_log=np.log(1e-10)):
        dx = np.ascontiguousarray(dx, np.float64)
        dy = np.ascontiguousarray(dy, np.float64)
        dz = np.ascontiguousarray(dz, np.float64)
        nze = len(dz) - int(nzAir)
        lc = np.asarray(logcond, np.float64)
        assert lc.shape == (len(dx), len(dy), nze)
        self._chk(lib().mt_assemble(self._h, len(dx), len(dy), int(nzAir), nze, dx, dy, dz,
                                    np.ascontiguousarray(lc.ravel(order="F")), float(air_log)))
        self._sizes()

    def update_cond(self, logcond, air_log=np.log(1e-10)):
        """New LOGE model (nx, ny, nzEarth) on the assembled grid, on the GPU."""
        lc = np.ascontiguousarray(np.asarray(logcond, np.float64).ravel(order="F"))
        self._chk(lib().mt_update_cond(self._h, lc, float(air_log)))
        self.nnz = None   # the stored pattern may change; counted again when asked

    def operator_csr(self):
        if self.nnz is None:
            self._sizes()
        row = np.zeros(self.n_int + 1, np.int32)
        col = np.zeros(self.nnz, np.int32)
        val = np.zeros(self.nnz)
        vms = np.zeros(self.n_int)
        self._chk(lib().mt_operator_get_csr(self._h, row, col, val, vms))
        return {"row": row, "col": col, "val": val, "vms1": vms}

    # -- solves -------------------------------------------------------------
    def bicg(self, omega, b, x0=None, maxit=1200, tol=1e-7, precond=PC_DILU):
        """cuBiCG semantics: returns (x, niter, rerr, failed)."""
        x = np.zeros(self.n_int, np.complex128) if x0 is None else np.array(x0, np.complex128)
        rerr = np.zeros(maxit)
        nit, fl = _I(), _I()
        self._chk(lib().mt_bicg(self._h, float(omega), np.ascontiguousarray(b, np.complex128), x,
                                int(maxit), float(tol), int(precond), C.byref(nit), rerr,
                                C.byref(fl)))
        return x, nit.value, rerr, bool(fl.value)

    def qmr(self, omega, b, x0=None, maxit=1200, tol=1e-7, precond=PC_DILU):
        """QMR (the reference's solver.f90:140-370 semantics): returns
        (x, niter, rerr, failed, status); status is MT_OK, MT_E_MAXIT or
        MT_E_BREAKDOWN (other codes raise)."""
        x = np.zeros(self.n_int, np.complex128) if x0 is None else np.array(x0, np.complex128)
        rerr = np.zeros(maxit)
        nit, fl = _I(), _I()
        rc = lib().mt_qmr(self._h, float(omega), np.ascontiguousarray(b, np.complex128), x,
                          int(maxit), float(tol), int(precond), C.byref(nit), rerr, C.byref(fl))
        if rc not in (MT_OK, MT_E_MAXIT, MT_E_BREAKDOWN):
            self._chk(rc)
        return x, nit.value, rerr, bool(fl.value), rc

    def solve_batch(self, omegas, b, x0=None, **opts):
        omegas = np.ascontiguousarray(omegas, np.float64)
        ns = len(omegas)
        b = np.ascontiguousarray(np.reshape(b, (ns, self.n_int)), np.complex128)
        x = np.zeros_like(b) if x0 is None else np.array(np.reshape(x0, b.shape), np.complex128)
        o = default_opts(**opts)
        st = (SysStats * ns)()
        self._chk(lib().mt_solve_batch(self._h, ns, omegas, b, x, C.byref(o), st))
        stats = [{"niter": s.niter, "nouter": s.nouter, "status": s.status, "relres": s.relres}
                 for s in st]
        return x, stats

    def solve_batch_dev(self, omegas, b_ptr, x_ptr, **opts):
        omegas = np.ascontiguousarray(omegas, np.float64)
        ns = len(omegas)
        o = default_opts(**opts)
        st = (SysStats * ns)()
        self._chk(lib().mt_solve_batch_dev(self._h, ns, omegas, C.c_void_p(b_ptr),
                                           C.c_void_p(x_ptr), C.byref(o), st))
        return [{"niter": s.niter, "nouter": s.nouter, "status": s.status, "relres": s.relres}
                for s in st]

    # -- device buffers (benchmark plumbing) ----------------------------------
    def dev_alloc(self, nbytes):
        p = _P()
        self._chk(lib().mt_dev_alloc(self._h, int(nbytes), C.byref(p)))
        return p.value

    def dev_free(self, ptr):
        self._chk(lib().mt_dev_free(self._h, C.c_void_p(ptr)))

    def upload(self, ptr, arr):
        arr = np.ascontiguousarray(arr)
        self._chk(lib().mt_dev_copy(self._h, C.c_void_p(ptr), arr.ctypes.data, arr.nbytes, 0))

    def download(self, ptr, out):
        self._chk(lib().mt_dev_copy(self._h, out.ctypes.data, C.c_void_p(ptr), out.nbytes, 1))
        return out

    def dev_copy(self, dst, src, nbytes):
        self._chk(lib().mt_dev_copy(self._h, C.c_void_p(dst), C.c_void_p(src), int(nbytes), 2))

    # -- operator applications -----------------------------------------------
    def apply(self, omega, x, transpose=0):
        y = np.zeros(self.n_int, np.complex128)
        self._chk(lib().mt_apply(self._h, float(omega), int(transpose),
                                 np.ascontiguousarray(x, np.complex128), y))
        return y

    def precond_apply(self, omega, x, precond=PC_DILU):
        y = np.zeros(self.n_int, np.complex128)
        self._chk(lib().mt_precond_apply(self._h, float(omega), int(precond),
                                         np.ascontiguousarray(x, np.complex128), y))
        return y

    def inner_apply(self, omega, x, what=0, transpose=False):
        """The mixed solver's brick kernels on one vector, FP32: y = A x
        (what 0) or y = M^-1 x (what 1, red-black DILU), of A or A^T."""
        y = np.zeros(self.n_int, np.complex128)
        self._chk(lib().mt_inner_apply(self._h, float(omega), int(what), int(bool(transpose)),
                                       np.ascontiguousarray(x, np.complex128), y))
        return y

    def rhs_fwd(self, e_all):
        b = np.zeros(self.n_int, np.complex128)
        self._chk(lib().mt_rhs_fwd(self._h, np.ascontiguousarray(e_all, np.complex128), b))
        return b

    def rhs_fwd_dev(self, nsys, e_ptr, b_ptr, x0_ptr=None):
        """b = -CCib e_b (and x0 = e(EDGEi)) for nsys device all-edge fields."""
        self._chk(lib().mt_rhs_fwd_dev(self._h, int(nsys), C.c_void_p(e_ptr), C.c_void_p(b_ptr),
                                       None if x0_ptr is None else C.c_void_p(x0_ptr)))

    def bc_x0_dev(self, omegas, modes, e_ptr):
        """BC_x0_WS on the device: the starting fields (E0 inside, boundary
        conditions outside) of len(omegas) forward systems into e_ptr
        ([nsys][n_edge] complex, device)."""
        om = np.ascontiguousarray(omegas, np.float64)
        md = np.ascontiguousarray(modes, np.int32)
        if om.shape != md.shape:
            raise ValueError("omegas and modes differ in length")
        self._chk(lib().mt_bc_x0_dev(self._h, len(om), om, md, C.c_void_p(e_ptr)))

    def bc_x0(self, omegas, modes):
        """bc_x0_dev into a temporary device buffer, returned as (nsys, n_edge)."""
        ns = len(omegas)
        if self.n_edge is None:
            self._sizes()
        p = self.dev_alloc(ns * self.n_edge * 16)
        try:
            self.bc_x0_dev(omegas, modes, p)
            return self.download(p, np.zeros((ns, self.n_edge), np.complex128))
        finally:
            self.dev_free(p)

    def rhs_trn(self, omega, s_all):
        b = np.zeros(self.n_int, np.complex128)
        self._chk(lib().mt_rhs_trn(self._h, float(omega), np.ascontiguousarray(s_all, np.complex128), b))
        return b

    def post_trn(self, ei):
        e = np.zeros(self.n_edge, np.complex128)
        self._chk(lib().mt_post_trn(self._h, np.ascontiguousarray(ei, np.complex128), e))
        return e

    def pmultT(self, omega, e0_all, lam_all, shape, dm=None):
        e0 = np.ascontiguousarray(np.atleast_2d(e0_all), np.complex128)
        lam = np.ascontiguousarray(np.atleast_2d(lam_all), np.complex128)
        acc = dm is not None
        out = np.zeros(int(np.prod(shape))) if dm is None else np.ascontiguousarray(
            np.asarray(dm, np.float64).ravel(order="F"))
        self._chk(lib().mt_pmultT(self._h, float(omega), e0.shape[0], e0, lam, out, int(acc)))
        return out.reshape(shape, order="F")

    def gradient(self, omega, e_all, lam, shape, mode=GRAD_EXACT, dm=None):
        e = np.ascontiguousarray(np.atleast_2d(e_all), np.complex128)
        la = np.ascontiguousarray(np.atleast_2d(lam), np.complex128)
        out = np.zeros(int(np.prod(shape))) if dm is None else np.ascontiguousarray(
            np.asarray(dm, np.float64).ravel(order="F")).copy()
        self._chk(lib().mt_gradient(self._h, float(omega), e.shape[0], e, la, out, int(mode)))
        return out.reshape(shape, order="F")

    # -- sensitivity path on device buffers ------------------------------------
    def post_fwd_dev(self, nsys, x_ptr, e_ptr):
        """e(EDGEi) = x for nsys systems (x [nsys][n_int], e [nsys][n_edge])."""
        self._chk(lib().mt_post_fwd_dev(self._h, int(nsys), C.c_void_p(x_ptr), C.c_void_p(e_ptr)))

    def rhs_trn_dev(self, omegas, s_ptr, b_ptr, gd_correct=True):
        om = np.ascontiguousarray(np.atleast_1d(omegas), np.float64)
        self._chk(lib().mt_rhs_trn_dev(self._h, len(om), om, C.c_void_p(s_ptr), C.c_void_p(b_ptr),
                                       int(bool(gd_correct))))

    def post_trn_dev(self, nsys, ei_ptr, e_ptr):
        self._chk(lib().mt_post_trn_dev(self._h, int(nsys), C.c_void_p(ei_ptr), C.c_void_p(e_ptr)))

    def pmultT_dev(self, omegas, npol, e0_ptr, lam_ptr, dm_ptr, accumulate=False):
        om = np.ascontiguousarray(np.atleast_1d(omegas), np.float64)
        self._chk(lib().mt_pmultT_dev(self._h, len(om), om, int(npol), C.c_void_p(e0_ptr),
                                      C.c_void_p(lam_ptr), C.c_void_p(dm_ptr), int(bool(accumulate))))

    def gradient_dev(self, omegas, npol, e_ptr, lam_ptr, dm_ptr, mode=None, accumulate=False):
        om = np.ascontiguousarray(np.atleast_1d(omegas), np.float64)
        mode = GRAD_EXACT if mode is None else mode
        self._chk(lib().mt_gradient_dev(self._h, len(om), om, int(npol), C.c_void_p(e_ptr),
                                        C.c_void_p(lam_ptr), C.c_void_p(dm_ptr), int(mode),
                                        int(bool(accumulate))))

    def divcor(self, omega, ei, maxit=100, tol=1e-7):
        e = np.array(ei, np.complex128)
        nit, rr = _I(), _D()
        self._chk(lib().mt_divcor(self._h, float(omega), e, None, int(maxit), float(tol),
                                  C.byref(nit), C.byref(rr)))
        return e, nit.value, rr.value

    # -- tuning / measurement ---------------------------------------------------
    def autotune(self, nsys):
        v = _I()
        self._chk(lib().mt_autotune(self._h, int(nsys), C.byref(v)))
        return v.value

    def set_spmv_variant(self, variant, layers_per_segment=0):
        """0 row-parallel SpMV, 2 k-marching, 1 k-marching in FP32 solves and
        row-parallel in FP64 ones, 3 (default) multi-system rows in FP32 solves
        (layers_per_segment = systems per block, 4 or 8) and row-parallel in
        FP64 ones."""
        self._chk(lib().mt_set_spmv_variant(self._h, int(variant), int(layers_per_segment)))

    def set_dilu_variant(self, variant):
        """Launch shape of the red-black DILU sweeps: 1 one system per
        single-wave workgroup, 4 four systems per workgroup, 5 (default) four
        systems per workgroup in the lean LDS layout."""
        self._chk(lib().mt_set_dilu_variant(self._h, int(variant)))

    def set_inner_layout(self, bricks, systems_per_block=0):
        """Inner FP32 vectors of the mixed solver (red-black DILU, default
        SpMV): True or 2 (default) = bricks with p-hat and s-hat interleaved
        by system pair (the pair SpMV); 1 = bricks, system-major; False or 0 =
        the all-edge layout.  systems_per_block: workgroup size in systems of
        the brick SpMV (0, the default: all active systems)."""
        b = (2 if bricks else 0) if isinstance(bricks, bool) else int(bricks)
        self._chk(lib().mt_set_inner_layout(self._h, b, int(systems_per_block)))

    def set_tuning(self, groups):
        self._chk(lib().mt_set_tuning(self._h, int(groups)))

    def timing_enable(self, on=True):
        self._chk(lib().mt_timing_enable(self._h, int(bool(on))))

    def kernel_timing(self, family):
        """(launches, ms, algorithmic bytes) of KT_SPMV / KT_PRECOND since
        timing_enable()."""
        n, ms, by = C.c_longlong(), _D(), _D()
        self._chk(lib().mt_kernel_timing(self._h, int(family), C.byref(n), C.byref(ms),
                                         C.byref(by)))
        return n.value, ms.value, by.value

    def bench_precond(self, nsys, precision=PREC_MIXED, precond=PC_DILU_RB, reps=10):
        ms, by = _D(), _D()
        self._chk(lib().mt_bench_precond(self._h, int(nsys), int(precision), int(precond),
                                         int(reps), C.byref(ms), C.byref(by)))
        return ms.value, by.value

    # -- data functionals ----------------------------------------------------
    def receivers_set(self, dx, dy, dz, nzAir, x, r=None, az=None, origin=(0.0, 0.0, 0.0)):
        """Receivers at x (nsite, 3) with reference sites r and channel
        azimuths az (nsite, 7; degrees); dz lists all layers, air first."""
        dx, dy, dz = (np.ascontiguousarray(v, np.float64) for v in (dx, dy, dz))
        x = np.ascontiguousarray(np.asarray(x, np.float64).reshape(-1, 3))
        self._keep_r = None if r is None else np.ascontiguousarray(
            np.asarray(r, np.float64).reshape(-1, 3))
        self._keep_az = None if az is None else np.ascontiguousarray(
            np.asarray(az, np.float64).reshape(-1, 7))
        org = np.ascontiguousarray(origin, np.float64)
        self._nsite = x.shape[0]
        self._chk(lib().mt_receivers_set(
            self._h, len(dx), len(dy), int(nzAir), len(dz) - int(nzAir), dx, dy, dz,
            org.ctypes.data, self._nsite, x,
            None if self._keep_r is None else self._keep_r.ctypes.data,
            None if self._keep_az is None else self._keep_az.ctypes.data))

    def data_resp(self, dtype, omegas, e_ptr, ncomp):
        """Responses [freq][site][ncomp] from device fields e_ptr
        ([freq][2][n_edge] complex)."""
        om = np.ascontiguousarray(np.atleast_1d(omegas), np.float64)
        out = np.zeros(len(om) * self._nsite * ncomp)
        self._chk(lib().mt_data_resp(self._h, int(dtype), len(om), om, C.c_void_p(e_ptr), out))
        return out.reshape(len(om), self._nsite, ncomp)

    def lmultT(self, dtype, omegas, e_ptr, d, comb_ptr, exist=None):
        """Adjoint sources into device comb_ptr ([freq][2][n_edge] complex) for
        data d [freq][site][ncomp]."""
        om = np.ascontiguousarray(np.atleast_1d(omegas), np.float64)
        d = np.ascontiguousarray(d, np.float64)
        ex = None if exist is None else np.ascontiguousarray(exist, np.uint8)
        self._chk(lib().mt_lmultT(self._h, int(dtype), len(om), om, C.c_void_p(e_ptr), d.ravel(),
                                  None if ex is None else ex.ctypes.data, C.c_void_p(comb_ptr)))

    def lmultT_noalloc(self, dtype, omegas, e_ptr, d, comb_ptr, exist=None):
        """lmultT without device allocations (scratch in a static device
        array; nfreq * nsite <= 16384)."""
        om = np.ascontiguousarray(np.atleast_1d(omegas), np.float64)
        d = np.ascontiguousarray(d, np.float64)
        ex = None if exist is None else np.ascontiguousarray(exist, np.uint8)
        self._chk(lib().mt_lmultT_noalloc(self._h, int(dtype), len(om), om, C.c_void_p(e_ptr),
                                          d.ravel(), None if ex is None else ex.ctypes.data,
                                          C.c_void_p(comb_ptr)))

    def bench_spmv(self, nsys, precision=PREC_FP64, reps=20):
        ms, by = _D(), _D()
        self._chk(lib().mt_bench_spmv(self._h, int(nsys), int(precision), int(reps),
                                      C.byref(ms), C.byref(by)))
        return ms.value, by.value
