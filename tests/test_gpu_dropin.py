"""The Fortran side of the drop-in, on the GPU.

1. modem-gpu_amd/lib/mtsp2_dropin_check: a Fortran program calling
   mtsp2_bicg (mtsp2_fortran.f90) exactly as solver.f90 would, on the
   reference's 1-based AAii CSR (from the oracle, bit-exact with the
   reference); checked against the oracle's cuBiCG-semantics solve.
2. oracle/_ref/ref_dropin (where it was built): the reference's own SP2 host
   code (ModelDataInit/UpdateCond/UpdateFreq) with `cuBiCG` replaced by the
   libmtsp2 call, next to the reference CPU BICG on the same system.
"""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import sp2oracle
from conftest import ROOT, rel, small_problem
from mtsp2 import problem as P

pytestmark = pytest.mark.gpu
CHECK = os.path.join(ROOT, "modem-gpu_amd", "lib", "mtsp2_dropin_check")
REF_DROPIN = os.path.join(ROOT, "oracle", "_ref", "ref_dropin")


def test_fortran_binding_solve():
    print("case fortran dropin check", flush=True)
    assert os.path.isfile(CHECK), "Fortran binding not built (make -C modem-gpu_amd/fortran)"
    dx, dy, dz, nza, lc = small_problem(41, nx=12, ny=11, nzAir=4, nzEarth=9)
    o = sp2oracle.Problem(dx, dy, dz, nza, lc)
    w = 2 * np.pi * 0.05
    grid = {"dx": dx, "dy": dy, "dz": dz, "nzAir": nza}
    e = P.plane_wave(grid, w, 1)
    o.update_freq(w)
    b = o.rhs_fwd(e)
    x0 = e[o.index()[0]]
    A = o.csr("AAii")
    maxit, tol = 500, 1e-10
    with tempfile.TemporaryDirectory() as td:
        fi, fo = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fi, "wb") as f:
            np.array([len(dx), len(dy), len(dz), o.nInt, len(A["col"])], np.int32).tofile(f)
            (A["row"] + 1).astype(np.int32).tofile(f)
            (A["col"] + 1).astype(np.int32).tofile(f)
            A["val"].tofile(f)
            (o.vms() / w).tofile(f)
            np.array([w]).tofile(f)
            np.array([maxit], np.int32).tofile(f)
            np.array([tol]).tofile(f)
            b.tofile(f)
            x0.astype(np.complex128).tofile(f)
        subprocess.run([CHECK, fi, fo], check=True, timeout=120)
        with open(fo, "rb") as f:
            niter, failed, ierr = np.fromfile(f, np.int32, 3)
            rerr = np.fromfile(f, np.float64, maxit)
            x = np.fromfile(f, np.complex128, o.nInt)
    assert ierr == 0 and failed == 0
    xo, info = o.bicg(b, x0, maxit=maxit, tol=tol, gpu_semantics=True)
    assert abs(int(niter) - info["niter"]) <= 3
    assert rel(x, xo) <= 1e-8
    assert np.linalg.norm(b - o.mult_aii(x)) / np.linalg.norm(b) <= tol
    assert rerr[niter - 1] <= tol


@pytest.mark.skipif(not os.path.isfile(REF_DROPIN), reason="reference drop-in not built here")
def test_reference_host_with_library_solver():
    print("case reference dropin", flush=True)
    grid = P.make_grid(18, 16, 10, nair=5, npad=4)
    lc = P.boxes_model(grid, nbox=6, seed=3)
    nx, ny, nz = len(grid["dx"]), len(grid["dy"]), len(grid["dz"])
    o = sp2oracle.Problem(grid["dx"], grid["dy"], grid["dz"], grid["nzAir"], lc)
    ei = o.index()[0]
    sols = []
    for f, pol in ((0.01, 0), (10.0, 1)):
        w = 2 * np.pi * f
        e = P.plane_wave(grid, w, pol)
        sols.append((w, o.rhs_fwd(e), e[ei]))
    with tempfile.TemporaryDirectory() as td:
        fi = os.path.join(td, "in.bin")
        with open(fi, "wb") as fh:
            np.array([0, nx, ny, grid["nzAir"], nz - grid["nzAir"]], np.int32).tofile(fh)
            grid["dx"].tofile(fh); grid["dy"].tofile(fh); grid["dz"].tofile(fh)
            lc.ravel(order="F").tofile(fh)
            np.array([np.log(1e-10)]).tofile(fh)
            np.array([len(sols)], np.int32).tofile(fh)
            for w, b, x0 in sols:
                np.array([w]).tofile(fh)
                np.array([800], np.int32).tofile(fh)
                np.array([1e-9]).tofile(fh)
                b.tofile(fh)
                x0.astype(np.complex128).tofile(fh)
        subprocess.run([REF_DROPIN, fi, td], check=True, timeout=300)
        for k, (w, b, x0) in enumerate(sols):
            with open(os.path.join(td, f"dropin_{k+1}.bin"), "rb") as fh:
                ng, failed, ierr, nc = np.fromfile(fh, np.int32, 4)
                xg = np.fromfile(fh, np.complex128, len(b))
                xc = np.fromfile(fh, np.complex128, len(b))
            assert ierr == 0 and failed == 0
            assert ng < 800 and nc < 800
            o.update_freq(w)
            assert np.linalg.norm(b - o.mult_aii(xg)) / np.linalg.norm(b) <= 1e-9
            assert rel(xg, xc) <= 1e-6
