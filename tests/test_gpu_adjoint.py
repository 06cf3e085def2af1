"""GPU adjoint path: transposed batched solves (A^T lam = s) and the model
gradient assembled from GPU forward and adjoint fields.

Tolerances:
  A^T solve, FP64 and mixed, tol 1e-10   true FP64 relres of A^T lam = s <= tol,
                                         lam vs dense solve rel L2 <= 1e-7
  gradient from GPU fields (tol 1e-11)   fields rel L2 <= 1e-7 and gradient
                                         rel max-norm <= 1e-6 vs dense solves
                                         (the dense-solve
                                         gradient matches central differences,
                                         tests/test_gradient_fd.py)
"""
import numpy as np
import pytest
import scipy.sparse as sp

import mtsp2
import sp2oracle
from conftest import rel, small_problem
from mtsp2 import problem as P

pytestmark = pytest.mark.gpu


def _amat(o, w):
    """Dense A(w) = AAii + i*ISIGN*w*mu0*sigma_E*V_E from the oracle's CSR and
    diagonal (ISIGN = -1)."""
    o.update_freq(w)
    a = o.csr("AAii")
    A = sp.csr_matrix((a["val"], a["col"], a["row"]), shape=(o.nInt, o.nInt)).toarray().astype(complex)
    A -= 1j * np.diag(o.vms())
    return A


@pytest.mark.parametrize("precond", [mtsp2.PC_DILU, mtsp2.PC_DILU_RB])
@pytest.mark.parametrize("precision", [mtsp2.PREC_FP64, mtsp2.PREC_MIXED])
def test_transpose_batch(gpu_ctx, precision, precond):
    """A^T solves with the block-DILU of A^T's blocks in the reference
    ordering and in the red-black tile ordering (transposed stencil, pivots
    of A^T) on a grid of partial 8x8x8 tiles."""
    print(f"case transpose batch precision={precision} precond={precond}", flush=True)
    dx, dy, dz, nza, lc = small_problem(11, nx=11, ny=10, nzAir=4, nzEarth=8)
    o = sp2oracle.Problem(dx, dy, dz, nza, lc)
    gpu_ctx.assemble(dx, dy, dz, nza, lc)
    rng = np.random.default_rng(3)
    om = 2 * np.pi * np.array([1e-2, 1.0, 1.0, 50.0])
    S = rng.standard_normal((len(om), o.nInt)) + 1j * rng.standard_normal((len(om), o.nInt))
    L, stats = gpu_ctx.solve_batch(om, S, None, tol=1e-10, maxit=3000, precision=precision,
                                   op_transpose=1, precond=precond)
    for s, w in enumerate(om):
        assert stats[s]["status"] == 0, stats[s]
        A = _amat(o, w)
        r = S[s] - A.T @ L[s]
        assert np.linalg.norm(r) / np.linalg.norm(S[s]) <= 1e-10 * 1.0001
        assert rel(L[s], np.linalg.solve(A.T, S[s])) <= 1e-7


def test_gradient_from_gpu_fields(gpu_ctx):
    """Forward solves, adjoint (A^T) solves and gradient assembly through the
    library reproduce the gradient computed from dense solves."""
    print("case gradient from GPU fields", flush=True)
    grid = P.make_grid(10, 9, 6, nair=3, npad=2)
    nx, ny, nz = len(grid["dx"]), len(grid["dy"]), len(grid["dz"])
    nze = nz - grid["nzAir"]
    rng = np.random.default_rng(5)
    lc = np.log(0.01) + 0.5 * rng.standard_normal((nx, ny, nze))
    w = 2 * np.pi * 1.0
    o = sp2oracle.Problem(grid["dx"], grid["dy"], grid["dz"], grid["nzAir"], lc)
    gpu_ctx.assemble(grid["dx"], grid["dy"], grid["dz"], grid["nzAir"], lc)
    ei = o.index()[0]
    A = _amat(o, w)
    om, B, X0 = P.forward_systems(gpu_ctx, grid, [1.0])
    E, st = gpu_ctx.solve_batch(om, B, X0, tol=1e-11, maxit=3000)
    assert all(s["status"] == 0 for s in st), st
    sites = rng.choice(o.nInt, 30, replace=False)
    d = [E[k][sites] * (1 + 0.1 * rng.standard_normal(30)) for k in range(2)]
    Sadj = np.zeros((2, o.nInt), complex)
    for k in range(2):
        Sadj[k, sites] = np.conj(d[k] - E[k][sites])
    Lam, st = gpu_ctx.solve_batch(om, Sadj, None, tol=1e-11, maxit=3000, op_transpose=1)
    assert all(s["status"] == 0 for s in st), st
    e_all = np.array([P.plane_wave(grid, w, k) for k in range(2)])
    e_all[:, ei] = E
    g = gpu_ctx.gradient(w, e_all, Lam, lc.shape, mode=mtsp2.GRAD_EXACT)
    # the same gradient from dense solves
    Ed = [np.linalg.solve(A, B[k]) for k in range(2)]
    Ld = [np.linalg.solve(A.T, Sadj[k]) for k in range(2)]
    assert rel(E, np.array(Ed)) <= 1e-7
    e_alld = e_all.copy()
    e_alld[:, ei] = np.array(Ed)
    gd = mtsp2.host_gradient(grid["dx"], grid["dy"], grid["dz"], grid["nzAir"], lc, w, e_alld,
                             np.array(Ld), mode=mtsp2.GRAD_EXACT)
    assert np.max(np.abs(g - gd)) <= 1e-6 * np.max(np.abs(gd))
