"""The Fortran side of the drop-in, on the GPU.

1. modem-gpu_amd/lib/mtsp2_dropin_check: a Fortran program calling
   mtsp2_bicg (mtsp2_fortran.f90) exactly as solver.f90 would, on the
   reference's 1-based AAii CSR (from the oracle, bit-exact with the
   reference); checked against the oracle's cuBiCG-semantics solve.
2. oracle/_ref/ref_dropin (where it was built): the reference's own SP2 host
   code (ModelDataInit/UpdateCond/UpdateFreq) with `cuBiCG` replaced by the
   libmtsp2 call, next to the reference CPU BICG on the same system.
3. oracle/_ref/ref_dropin_datafunc (where it was built): the reference's
   dataResp/LmultT next to mtsp2_data_resp/mtsp2_lmultT, one Fortran host.
4. oracle/_ref/ref_dropin_bc (where it was built): the reference's
   ComputeBC -> BC_x0_WS next to mtsp2_bc_x0, one Fortran host.
"""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import refdriver
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
            niter, failed, ierr, same = np.fromfile(f, np.int32, 4)
            rerr = np.fromfile(f, np.float64, maxit)
            x = np.fromfile(f, np.complex128, o.nInt)
    assert ierr == 0 and failed == 0
    assert same == 1, "a repeat solve with an unchanged operator gave other iterates"
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


@pytest.mark.skipif(not os.path.isfile(refdriver.DF_DROPIN_BIN),
                    reason="reference data-functional drop-in not built here")
def test_reference_host_with_library_lmultT():
    """The reference's dataResp/LmultT and the library's, called from one
    Fortran host (mtsp2_data_resp, mtsp2_lmultT) on the same inputs: every
    data type to 1e-10 relative."""
    print("case reference dropin LmultT", flush=True)
    dx, dy, dz, nza, lc = small_problem(4)
    o = sp2oracle.Problem(dx, dy, dz, nza, lc)
    rng = np.random.default_rng(9)
    ns = 5
    x = np.column_stack([rng.uniform(-0.3, 0.3, ns) * dx.sum(),
                         rng.uniform(-0.3, 0.3, ns) * dy.sum(), rng.uniform(0, 60, ns)])
    r = x[::-1].copy()
    az = np.tile([0.0, 90.0, 0.0, 0.0, 90.0, 0.0, 90.0], (ns, 1))
    az[2] = [10.0, 95.0, 0.0, 5.0, 100.0, -3.0, 80.0]
    origin = np.array([-dx.sum() / 2, -dy.sum() / 2, 0.0])
    e1 = rng.standard_normal(o.nE) + 1j * rng.standard_normal(o.nE)
    e2 = rng.standard_normal(o.nE) + 1j * rng.standard_normal(o.nE)
    grid = dict(dx=dx, dy=dy, dz=dz, nzAir=nza)
    for dt, nc in {1: 8, 2: 4, 3: 4, 4: 8, 5: 4, 6: 4, 7: 2, 11: 2}.items():
        d = rng.standard_normal((ns, nc))
        ex = np.ones((ns, nc), np.int32)
        ex[1, 0] = 0
        rr, c1, c2, lr, l1, l2, st = refdriver.run_datafunc(
            grid, origin, 2 * np.pi * 0.2, dt, x, e1, e2, d, r=r, az=az, exist=ex,
            binary=refdriver.DF_DROPIN_BIN)
        assert np.all(st == 0), (dt, st)
        assert np.max(np.abs(lr - rr)) <= 1e-10 * np.max(np.abs(rr)), dt
        for a, b in ((l1, c1), (l2, c2)):
            assert np.max(np.abs(a - b)) <= 1e-10 * max(np.max(np.abs(b)), 1e-300), dt


@pytest.mark.skipif(not refdriver.dropin_bc_available(),
                    reason="reference boundary-field drop-in not built here")
def test_reference_host_with_library_bc_x0():
    """The reference's ComputeBC (the 2-D TE slice solves) and the library's
    mtsp2_bc_x0, called from one Fortran host on the same grid, model and
    (omega, mode) list: bit for bit."""
    print("case reference dropin BC_x0_WS", flush=True)
    grid = P.make_grid(11, 8, 12, nair=5, npad=3)
    lc = P.boxes_model(grid, nbox=10, seed=12)
    sy = [(2 * np.pi * 0.01, 1), (2 * np.pi * 0.01, 2), (2 * np.pi * 40.0, 1), (2 * np.pi * 40.0, 2)]
    ref, lib, st = refdriver.run_dropin_bc(grid, lc, np.log(1e-10), sy)
    assert st == 0
    assert np.array_equal(lib, ref)
