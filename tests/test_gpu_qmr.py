"""QMR on the GPU (mt_qmr: the reference's solver.f90:140-370, its default
solver) against the oracle's QMR (oracle/sp2_oracle.c sp2o_qmr, bit-exact
with the flang build of the reference's QMR, tests/test_oracle_vs_ref.py).

The GPU forms its inner products in another order, so iterates agree to
rounding: iteration counts within max(3, 10 %), x rel L2 <= 1e-7 at tol
1e-10, true relative residual of the returned (last) iterate <= 1e-9; with
maxit cut short, the same niter = maxit, the last iterate within 1e-6 and
rerr[niter-1] (the smallest residual seen, stored as a default real) within
1e-4 relative.
"""
import numpy as np
import pytest

import mtsp2
import sp2oracle
from conftest import rel, small_problem
from mtsp2 import problem as P

pytestmark = pytest.mark.gpu


def _setup(ctx, seed, **kw):
    dx, dy, dz, nza, lc = small_problem(seed, **kw)
    o = sp2oracle.Problem(dx, dy, dz, nza, lc)
    ctx.assemble(dx, dy, dz, nza, lc)
    return o, {"dx": dx, "dy": dy, "dz": dz, "nzAir": nza}


@pytest.mark.parametrize("f_hz", [0.01, 10.0])
def test_qmr_matches_oracle(gpu_ctx, f_hz):
    print(f"case qmr f={f_hz}", flush=True)
    o, grid = _setup(gpu_ctx, 31, nx=13, ny=12, nzAir=4, nzEarth=9)
    w = 2 * np.pi * f_hz
    o.update_freq(w)
    e = P.plane_wave(grid, w, 1)
    b = o.rhs_fwd(e)
    x0 = e[o.index()[0]]
    xg, nit, rerr, failed, st = gpu_ctx.qmr(w, b, x0, maxit=2000, tol=1e-10)
    xo, info = o.qmr(b, x0, maxit=2000, tol=1e-10)
    print(f"qmr iterations gpu {nit} oracle {info['niter']}", flush=True)
    assert st == mtsp2.MT_OK and not failed and info["status"] == 0
    assert abs(nit - info["niter"]) <= max(3, info["niter"] // 10)
    assert rel(xg, xo) <= 1e-7
    assert np.linalg.norm(b - o.mult_aii(xg)) / np.linalg.norm(b) <= 1e-9
    assert rerr[nit - 1] <= 1e-10


def test_qmr_maxit_last_iterate(gpu_ctx):
    print("case qmr maxit", flush=True)
    o, grid = _setup(gpu_ctx, 32, nx=12, ny=11, nzAir=4, nzEarth=8)
    w = 2 * np.pi * 0.3
    o.update_freq(w)
    b = o.rhs_fwd(P.plane_wave(grid, w, 0))
    x0 = np.zeros(o.nInt, complex)
    xg, nit, rerr, failed, st = gpu_ctx.qmr(w, b, x0, maxit=12, tol=1e-12)
    xo, info = o.qmr(b, x0, maxit=12, tol=1e-12)
    assert st == mtsp2.MT_E_MAXIT and info["status"] == 1
    assert nit == info["niter"] == 12
    assert rel(xg, xo) <= 1e-6
    assert abs(rerr[nit - 1] - info["rerr"][nit - 1]) <= 1e-4 * info["rerr"][nit - 1]
    assert rerr[nit - 1] <= min(info["rerr"][:nit - 1]) * (1 + 1e-4)


def test_qmr_unpreconditioned_and_transposed_pivots(gpu_ctx):
    """Without the DILU (M1 = M2 = I) QMR still reaches the tolerance; with it,
    it needs fewer iterations (the L^-H / U^-H solves of the dual recurrence
    use the transposed stencil with conjugated pivots: a wrong slot or sign
    there would slow or stall it)."""
    print("case qmr precond none vs dilu", flush=True)
    o, grid = _setup(gpu_ctx, 33, nx=11, ny=10, nzAir=4, nzEarth=7)
    w = 2 * np.pi * 1.0
    o.update_freq(w)
    b = o.rhs_fwd(P.plane_wave(grid, w, 1))
    _, n_none, _, f1, s1 = gpu_ctx.qmr(w, b, None, maxit=6000, tol=1e-9, precond=mtsp2.PC_NONE)
    x, n_dilu, _, f2, s2 = gpu_ctx.qmr(w, b, None, maxit=6000, tol=1e-9, precond=mtsp2.PC_DILU)
    print(f"qmr iterations: none {n_none}, dilu {n_dilu}", flush=True)
    assert s1 == mtsp2.MT_OK and s2 == mtsp2.MT_OK
    assert n_dilu < n_none
    assert np.linalg.norm(b - o.mult_aii(x)) / np.linalg.norm(b) <= 1e-8
