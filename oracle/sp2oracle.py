"""ctypes view of the C oracle (oracle/sp2_oracle.c).  TEST INFRASTRUCTURE ONLY:
imported by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg.

All index arrays returned here are 0-based.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libsp2oracle.so")
_lib = None

_P = C.c_void_p
_I = C.c_int
_D = C.c_double
_PI = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_PD = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_PZ = np.ctypeslib.ndpointer(np.complex128, flags="C_CONTIGUOUS")


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.isfile(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        sig = {
            "sp2o_create": (_P, [_I, _I, _I, _I, _PD, _PD, _PD, _PD, _D]),
            "sp2o_destroy": (None, [_P]),
            "sp2o_sizes": (None, [_P, _PI]),
            "sp2o_index": (None, [_P, _PI, _PI, _PI, _PI]),
            "sp2o_metric": (None, [_P, _PD, _PD, _PD, _PD, _PD, _PD]),
            "sp2o_sigma": (None, [_P, _PD, _PD]),
            "sp2o_csr_nnz": (_I, [_P, _I, C.POINTER(_I), C.POINTER(_I)]),
            "sp2o_csr_get": (None, [_P, _I, _PI, _PI, _PD]),
            "sp2o_update_freq": (None, [_P, _D]),
            "sp2o_vms": (None, [_P, _PD]),
            "sp2o_lu_nnz": (_I, [_P, _I]),
            "sp2o_lu_get": (None, [_P, _I, _PI, _PI, _PZ]),
            "sp2o_mult_aii": (None, [_P, _I, _PZ, _PZ]),
            "sp2o_pc_lsolve": (None, [_P, _PZ, _PZ]),
            "sp2o_pc_usolve": (None, [_P, _PZ, _PZ]),
            "sp2o_bicg": (_I, [_P, _PZ, _PZ, _I, _D, _I, _I, C.POINTER(_I), _PD]),
            "sp2o_qmr": (_I, [_P, _PZ, _PZ, _I, _D, C.POINTER(_I), _PD]),
            "sp2o_pc_lsolve_h": (None, [_P, _PZ, _PZ]),
            "sp2o_pc_usolve_h": (None, [_P, _PZ, _PZ]),
            "sp2o_divcor_setup": (None, [_P]),
            "sp2o_pcg": (_I, [_P, _PZ, _PZ, _I, _D, C.POINTER(_I), _PD]),
            "sp2o_divc": (None, [_P, _PZ, _PZ]),
            "sp2o_grad": (None, [_P, _PZ, _PZ]),
            "sp2o_rhs_fwd": (None, [_P, _PZ, _PZ]),
            "sp2o_rhs_trn": (None, [_P, _D, _PZ, _PZ]),
            "sp2o_post_trn": (None, [_P, _PZ, _PZ]),
            "sp2o_edge2model": (None, [_P, _PD, _PD]),
            "sp2o_pmultT": (None, [_P, _D, _I, _PZ, _PZ, _PD]),
            "sp2o_model2edge": (None, [_P, _PD, _PD]),
            "sp2o_data_resp": (_I, [_P, _I, _D, _PD, _I, _PD, _P, _P, _PZ, _PZ, _PD]),
            "sp2o_lmultT": (_I, [_P, _I, _D, _PD, _I, _PD, _P, _P, _PD, _P, _PZ, _PZ,
                                 _PZ, _PZ]),
            "sp2o_bc_x0": (_I, [_I, _I, _I, _I, _PD, _PD, _PD, _PD, _D, _I, _PZ,
                                C.POINTER(_I)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


CSR_NAMES = {"AAii": 0, "CCii": 1, "CCib": 2, "GDii": 3, "T": 4, "G": 5, "VDiv": 6,
             "VDs": 7, "Gai": 8, "VDsG": 9, "VDsG_L": 10, "VDsG_U": 11}


class Problem:
    """One grid + LOGE model, as ModelDataInit + UpdateCond leave it."""

    def __init__(self, dx, dy, dz, nzAir, logcond, air_log=np.log(1e-10)):
        self.dx = np.ascontiguousarray(dx, np.float64)
        self.dy = np.ascontiguousarray(dy, np.float64)
        self.dz = np.ascontiguousarray(dz, np.float64)
        self.nx, self.ny, self.nz = len(self.dx), len(self.dy), len(self.dz)
        self.nzAir = int(nzAir)
        self.nzEarth = self.nz - self.nzAir
        lc = np.asarray(logcond, np.float64)
        assert lc.shape == (self.nx, self.ny, self.nzEarth)
        self.logcond = lc
        self._h = lib().sp2o_create(self.nx, self.ny, self.nzAir, self.nzEarth, self.dx,
                                    self.dy, self.dz,
                                    np.ascontiguousarray(lc.ravel(order="F")), float(air_log))
        s = np.zeros(9, np.int32)
        lib().sp2o_sizes(self._h, s)
        (self.nE, self.nInt, self.nBdry, self.nNode, self.nNi, self.nNb,
         self.nxInt, self.nyInt, self.nzInt) = [int(v) for v in s]
        self.omega = 1.0

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().sp2o_destroy(h)
            self._h = None

    def index(self):
        ei = np.zeros(self.nInt, np.int32); eb = np.zeros(self.nBdry, np.int32)
        ni = np.zeros(self.nNi, np.int32); nb = np.zeros(self.nNb, np.int32)
        lib().sp2o_index(self._h, ei, eb, ni, nb)
        return ei, eb, ni, nb

    def metric(self):
        nF = (self.nx + 1) * self.ny * self.nz + self.nx * (self.ny + 1) * self.nz + \
            self.nx * self.ny * (self.nz + 1)
        out = [np.zeros(nF), np.zeros(self.nE), np.zeros(self.nE), np.zeros(nF),
               np.zeros(self.nNode), np.zeros(self.nE)]
        lib().sp2o_metric(self._h, *out)
        return dict(zip(["FaceA", "EdgeL", "DualFaceA", "DualEdgeL", "Vnode", "Vedge"], out))

    def sigma(self):
        se, sn = np.zeros(self.nE), np.zeros(self.nNode)
        lib().sp2o_sigma(self._h, se, sn)
        return se, sn

    def csr(self, name):
        nr, nc = _I(), _I()
        nnz = lib().sp2o_csr_nnz(self._h, CSR_NAMES[name], C.byref(nr), C.byref(nc))
        if nnz < 0:
            raise KeyError(name)
        row = np.zeros(nr.value + 1, np.int32)
        col = np.zeros(nnz, np.int32)
        val = np.zeros(nnz)
        lib().sp2o_csr_get(self._h, CSR_NAMES[name], row, col, val)
        return {"nrow": nr.value, "ncol": nc.value, "row": row, "col": col, "val": val}

    def update_freq(self, omega):
        lib().sp2o_update_freq(self._h, float(omega))
        self.omega = float(omega)

    def vms(self):
        v = np.zeros(self.nInt)
        lib().sp2o_vms(self._h, v)
        return v

    def lu(self, which):
        nnz = lib().sp2o_lu_nnz(self._h, which)
        row = np.zeros(self.nInt + 1, np.int32)
        col = np.zeros(nnz, np.int32)
        val = np.zeros(nnz, np.complex128)
        lib().sp2o_lu_get(self._h, which, row, col, val)
        return {"nrow": self.nInt, "ncol": self.nInt, "row": row, "col": col, "val": val}

    def mult_aii(self, x, adjt=False):
        y = np.zeros(self.nInt, np.complex128)
        lib().sp2o_mult_aii(self._h, int(adjt), np.ascontiguousarray(x, np.complex128), y)
        return y

    def pc_lsolve(self, x):
        y = np.zeros(self.nInt, np.complex128)
        lib().sp2o_pc_lsolve(self._h, np.ascontiguousarray(x, np.complex128), y)
        return y

    def pc_usolve(self, x):
        y = np.zeros(self.nInt, np.complex128)
        lib().sp2o_pc_usolve(self._h, np.ascontiguousarray(x, np.complex128), y)
        return y

    def bicg(self, b, x0=None, maxit=500, tol=1e-7, gpu_semantics=True, precond=0):
        x = np.zeros(self.nInt, np.complex128) if x0 is None else np.array(x0, np.complex128)
        rerr = np.zeros(maxit)
        nit = _I()
        st = lib().sp2o_bicg(self._h, np.ascontiguousarray(b, np.complex128), x, maxit,
                             float(tol), int(gpu_semantics), int(precond), C.byref(nit), rerr)
        return x, {"status": st, "niter": nit.value, "rerr": rerr}

    def pc_lsolve_h(self, x):
        y = np.zeros(self.nInt, np.complex128)
        lib().sp2o_pc_lsolve_h(self._h, np.ascontiguousarray(x, np.complex128), y)
        return y

    def pc_usolve_h(self, x):
        y = np.zeros(self.nInt, np.complex128)
        lib().sp2o_pc_usolve_h(self._h, np.ascontiguousarray(x, np.complex128), y)
        return y

    def qmr(self, b, x0=None, maxit=500, tol=1e-7):
        """QMR (solver.f90:140-370); returns (x, info) like bicg."""
        x = np.zeros(self.nInt, np.complex128) if x0 is None else np.array(x0, np.complex128)
        rerr = np.zeros(maxit)
        nit = _I()
        st = lib().sp2o_qmr(self._h, np.ascontiguousarray(b, np.complex128), x, maxit, float(tol),
                            C.byref(nit), rerr)
        return x, {"status": st, "niter": nit.value, "rerr": rerr}

    def divcor_setup(self):
        lib().sp2o_divcor_setup(self._h)

    def pcg(self, b, maxit=100, tol=1e-7, x0=None):
        x = np.zeros(self.nNi, np.complex128) if x0 is None else np.array(x0, np.complex128)
        rerr = np.zeros(maxit)
        nit = _I()
        st = lib().sp2o_pcg(self._h, np.ascontiguousarray(b, np.complex128), x, maxit,
                            float(tol), C.byref(nit), rerr)
        return x, {"status": st, "niter": nit.value, "rerr": rerr}

    def rhs_fwd(self, e_all):
        b = np.zeros(self.nInt, np.complex128)
        lib().sp2o_rhs_fwd(self._h, np.ascontiguousarray(e_all, np.complex128), b)
        return b

    def rhs_trn(self, omega, s_all):
        b = np.zeros(self.nInt, np.complex128)
        lib().sp2o_rhs_trn(self._h, float(omega), np.ascontiguousarray(s_all, np.complex128), b)
        return b

    def post_trn(self, ei):
        e = np.zeros(self.nE, np.complex128)
        lib().sp2o_post_trn(self._h, np.ascontiguousarray(ei, np.complex128), e)
        return e

    def edge2model(self, e_all):
        dm = np.zeros(self.nx * self.ny * self.nzEarth)
        lib().sp2o_edge2model(self._h, np.ascontiguousarray(e_all, np.float64), dm)
        return dm.reshape((self.nx, self.ny, self.nzEarth), order="F")

    def pmultT(self, omega, e0_all, e_all):
        e0 = np.ascontiguousarray(np.atleast_2d(e0_all), np.complex128)
        e = np.ascontiguousarray(np.atleast_2d(e_all), np.complex128)
        dm = np.zeros(self.nx * self.ny * self.nzEarth)
        lib().sp2o_pmultT(self._h, float(omega), e0.shape[0], e0, e, dm)
        return dm.reshape((self.nx, self.ny, self.nzEarth), order="F")

    def model2edge(self, logcond):
        se = np.zeros(self.nE)
        lib().sp2o_model2edge(self._h, np.ascontiguousarray(
            np.asarray(logcond, np.float64).ravel(order="F")), se)
        return se

    # ---- data functionals: dataResp / LmultT for one transmitter -------------
    @staticmethod
    def _sites(x, r, az):
        x = np.ascontiguousarray(np.asarray(x, np.float64).reshape(-1, 3))
        rr = None if r is None else np.ascontiguousarray(np.asarray(r, np.float64).reshape(-1, 3))
        aa = None if az is None else np.ascontiguousarray(np.asarray(az, np.float64).reshape(-1, 7))
        return x, rr, aa

    @staticmethod
    def _ptr(a):
        return None if a is None else a.ctypes.data_as(_P)

    def data_resp(self, dtype, omega, origin, x, e1, e2, r=None, az=None):
        """Responses (nsite, ncomp) of data type dtype (dataTypes.f90 numbering)."""
        x, rr, aa = self._sites(x, r, az)
        ns = x.shape[0]
        resp = np.zeros(ns * 8)
        nc = lib().sp2o_data_resp(self._h, int(dtype), float(omega),
                                  np.ascontiguousarray(origin, np.float64), ns, x,
                                  self._ptr(rr), self._ptr(aa),
                                  np.ascontiguousarray(e1, np.complex128),
                                  np.ascontiguousarray(e2, np.complex128), resp)
        if nc < 0:
            raise ValueError(f"unsupported data type {dtype}")
        return resp[:ns * nc].reshape(ns, nc)

    def lmultT(self, dtype, omega, origin, x, d, e1, e2, r=None, az=None, exist=None):
        """Adjoint sources (comb1, comb2) on all edges for data d (nsite, ncomp)."""
        x, rr, aa = self._sites(x, r, az)
        ns = x.shape[0]
        d = np.ascontiguousarray(d, np.float64)
        ex = None if exist is None else np.ascontiguousarray(exist, np.uint8)
        c1 = np.zeros(self.nE, np.complex128)
        c2 = np.zeros(self.nE, np.complex128)
        nc = lib().sp2o_lmultT(self._h, int(dtype), float(omega),
                               np.ascontiguousarray(origin, np.float64), ns, x, self._ptr(rr),
                               self._ptr(aa), d.ravel(), self._ptr(ex),
                               np.ascontiguousarray(e1, np.complex128),
                               np.ascontiguousarray(e2, np.complex128), c1, c2)
        if nc < 0:
            raise ValueError(f"unsupported data type {dtype}")
        return c1, c2


def bc_x0(dx, dy, dz, nza, logcond, omega, imode):
    """BC_x0_WS: the all-edge E0 of the 2-D TE slice solves for mode imode
    (1: Ey polarisation, 2: Ex) and the number of pivot interchanges."""
    dx, dy, dz = (np.ascontiguousarray(a, np.float64) for a in (dx, dy, dz))
    nx, ny, nz = len(dx), len(dy), len(dz)
    lc = np.ascontiguousarray(np.asarray(logcond, np.float64).reshape(nx, ny, nz - nza).ravel(order="F"))
    ne = nx * (ny + 1) * (nz + 1) + (nx + 1) * ny * (nz + 1) + (nx + 1) * (ny + 1) * nz
    e0 = np.zeros(ne, np.complex128)
    nsw = _I()
    rc = lib().sp2o_bc_x0(nx, ny, nza, nz - nza, dx, dy, dz, lc, float(omega), int(imode), e0,
                          C.byref(nsw))
    if rc:
        raise ValueError(f"bad mode {imode}")
    return e0, nsw.value


def bc_form(e0, nx, ny, nz):
    """The field FWDsolve3D starts from: E0 with the bottom-face x and y edges
    zeroed (BC_x0_WS sets BC%xZMax = BC%yZMax = 0 and setBC writes the z
    faces last)."""
    e = np.array(e0, np.complex128)
    nXe = nx * (ny + 1) * (nz + 1)
    ex = e[:nXe].reshape(nz + 1, ny + 1, nx)
    ey = e[nXe:nXe + (nx + 1) * ny * (nz + 1)].reshape(nz + 1, ny, nx + 1)
    ex[nz] = 0
    ey[nz] = 0
    return e
