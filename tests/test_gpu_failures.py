"""Failure paths of the solvers on the GPU, each forced from the host so that
every kernel still ends normally:
  a right-hand side with a NaN       -> MT_E_NAN_RHS (the reference stops,
                                        solver.f90:1500); in a batch only that
                                        system reports it, the others converge
  maxit exhausted                    -> niter = maxit, the smallest-residual
                                        iterate returned (solver.f90:1911-1919),
                                        equal to the oracle's cuBiCG-semantics
                                        best iterate to 1e-8
  breakdown                          -> failed / MT_E_BREAKDOWN: an operator
                                        with a skew-symmetric coupling makes
                                        <rt, A p> = 0 at the first iteration of
                                        BiCGStab and <q, A p> = 0 for QMR
"""
import numpy as np
import pytest

import mtsp2
import sp2oracle
from conftest import rel, small_problem
from mtsp2 import problem as P

pytestmark = pytest.mark.gpu


def test_nan_rhs(gpu_ctx):
    print("case NaN right-hand side", flush=True)
    dx, dy, dz, nza, lc = small_problem(51)
    ctx = gpu_ctx
    ctx.assemble(dx, dy, dz, nza, lc)
    grid = {"dx": dx, "dy": dy, "dz": dz, "nzAir": nza}
    om, B, X0 = P.forward_systems(ctx, grid, [0.5])
    bad = B[0].copy()
    bad[7] = np.nan
    with pytest.raises(mtsp2.MTError) as ei:
        ctx.bicg(om[0], bad, X0[0], maxit=100, tol=1e-9)
    assert ei.value.code == mtsp2.MT_E_NAN_RHS
    with pytest.raises(mtsp2.MTError) as ei:
        ctx.qmr(om[0], bad, X0[0], maxit=100, tol=1e-9)
    assert ei.value.code == mtsp2.MT_E_NAN_RHS
    Bb = np.array([bad, B[1]])
    # (the mixed solver with the red-black DILU runs on bricks with paired
    # x: the NaN system stops at once, its pair partner runs alone)
    for prec, pc in ((mtsp2.PREC_FP64, None), (mtsp2.PREC_MIXED, None),
                     (mtsp2.PREC_MIXED, mtsp2.PC_DILU_RB)):
        kw = {} if pc is None else {"precond": pc}
        X, st = ctx.solve_batch(om, Bb, X0, tol=1e-9, maxit=800, precision=prec, **kw)
        assert st[0]["status"] == mtsp2.MT_E_NAN_RHS, (prec, pc, st)
        assert st[1]["status"] == mtsp2.MT_OK and st[1]["relres"] <= 1e-9, (prec, pc, st)


def test_maxit_best_iterate(gpu_ctx):
    print("case maxit best iterate", flush=True)
    dx, dy, dz, nza, lc = small_problem(52, nx=12, ny=11, nzAir=4, nzEarth=9)
    o = sp2oracle.Problem(dx, dy, dz, nza, lc)
    ctx = gpu_ctx
    ctx.assemble(dx, dy, dz, nza, lc)
    w = 2 * np.pi * 0.2
    o.update_freq(w)
    grid = {"dx": dx, "dy": dy, "dz": dz, "nzAir": nza}
    e = P.plane_wave(grid, w, 0)
    b = o.rhs_fwd(e)
    x0 = e[o.index()[0]]
    x, nit, rerr, failed = ctx.bicg(w, b, x0, maxit=9, tol=1e-14)
    xo, info = o.bicg(b, x0, maxit=9, tol=1e-14, gpu_semantics=True)
    assert nit == 9 and not failed and info["status"] == 1
    assert rel(x, xo) <= 1e-8
    # the returned iterate is the best one seen: its true residual is the
    # smallest of the history (to the recursive/true residual gap)
    rr = np.linalg.norm(b - o.mult_aii(x)) / np.linalg.norm(b)
    assert rr <= min(rerr[:nit]) * 1.01


def _skew_operator(ctx):
    """3x3x3 cells (36 interior edges): AAii with one skew-symmetric coupling
    between the x-edges (0,1,1) and (1,1,1) (rows 0 and 1), everything else
    zero, vms1 = 0."""
    n = 36
    row = np.zeros(n + 1, np.int32)
    row[1:] = 2
    row[1] = 1
    col = np.array([1, 0], np.int32)
    val = np.array([1.0, -1.0])
    ctx.set_operator_csr((3, 3, 3), n, row, col, val, np.zeros(n), base=0)
    b = np.zeros(n, complex)
    b[0] = 1.0
    return n, b


def test_forced_breakdown(gpu_ctx):
    print("case forced breakdown", flush=True)
    ctx = gpu_ctx
    n, b = _skew_operator(ctx)
    x, nit, rerr, failed = ctx.bicg(1.0, b, None, maxit=50, tol=1e-10, precond=mtsp2.PC_NONE)
    assert failed and nit == 50
    assert np.array_equal(x, np.zeros(n))          # the best iterate is x0
    X, st = ctx.solve_batch([1.0], b[None], None, tol=1e-10, maxit=50, precond=mtsp2.PC_NONE)
    assert st[0]["status"] == mtsp2.MT_E_BREAKDOWN
    xq, nq, _, fq, sq = ctx.qmr(1.0, b, None, maxit=50, tol=1e-10, precond=mtsp2.PC_NONE)
    assert sq == mtsp2.MT_E_BREAKDOWN and fq
