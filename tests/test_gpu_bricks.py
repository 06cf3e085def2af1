"""The mixed solver's inner solves on bricks (each 8x8x8 tile of the
red-black plan contiguous): the brick SpMV and the brick red-black DILU on
one vector against the FP64 all-edge kernels and the numpy DILU restatement,
and batched solves on bricks against the all-edge inner layout and the
oracle.  Grids with partial tiles in every direction.

Tolerances: the brick kernels compute in FP32, so one application agrees
with the FP64 operator to FP32 rounding (rel L2 <= 2e-6 for A x; <= 2e-5 for
M^-1 x, whose sweeps accumulate); solves reach the FP64 true residual
tolerance and agree with the oracle's solution to rel L2 <= 1e-7 (tol 1e-10).
"""
import numpy as np
import pytest

import mtsp2
import sp2oracle
from conftest import (block_diag_operator, dilu_factor, dilu_mult, rb_tile_keys, rel,
                      small_problem)
from mtsp2 import problem as P

pytestmark = pytest.mark.gpu


def spla_solve(A, b):
    import scipy.sparse.linalg as spla
    return spla.spsolve(A.tocsc(), b)


def _grid(ctx, seed):
    dx, dy, dz, nza, lc = small_problem(seed, nx=26, ny=17, nzAir=5, nzEarth=12)
    o = sp2oracle.Problem(dx, dy, dz, nza, lc)
    ctx.assemble(dx, dy, dz, nza, lc)
    return o, (dx, dy, dz, nza, lc)


@pytest.mark.parametrize("transpose", [False, True])
def test_brick_spmv_matches_fp64(gpu_ctx, transpose):
    print(f"case brick spmv transpose={transpose}", flush=True)
    o, g = _grid(gpu_ctx, 41)
    rng = np.random.default_rng(3)
    x = rng.standard_normal(gpu_ctx.n_int) + 1j * rng.standard_normal(gpu_ctx.n_int)
    for f in (1e-3, 1.0, 300.0):
        w = 2 * np.pi * f
        y64 = gpu_ctx.apply(w, x, transpose=int(transpose))
        y32 = gpu_ctx.inner_apply(w, x, what=0, transpose=transpose)
        assert rel(y32, y64) <= 2e-6, f


@pytest.mark.parametrize("transpose", [False, True])
def test_brick_dilu_inverts_its_factorisation(gpu_ctx, transpose):
    """M z = x for the red-black tile DILU of A(omega) (or of A^T) applied on
    bricks, M from the numpy restatement pinned to the oracle's
    PC_Lsolve/PC_Usolve (tests/test_oracle_invariants.py)."""
    print(f"case brick dilu transpose={transpose}", flush=True)
    o, g = _grid(gpu_ctx, 42)
    key = rb_tile_keys(len(g[0]), len(g[1]), len(g[2]))
    rng = np.random.default_rng(4)
    x = rng.standard_normal(o.nInt) + 1j * rng.standard_normal(o.nInt)
    for f in (1e-2, 3.0):
        w = 2 * np.pi * f
        B = block_diag_operator(o, w)
        if transpose:
            B = B.T.tocsr()
        piv, lower, upper = dilu_factor(B, key)
        z = gpu_ctx.inner_apply(w, x, what=1, transpose=transpose)
        assert rel(dilu_mult(piv, lower, upper, z), x) <= 2e-5, f
        if not transpose:
            z64 = gpu_ctx.precond_apply(w, x, mtsp2.PC_DILU_RB)
            assert rel(z, z64) <= 2e-5, f


def test_brick_dilu_persistent_matches(gpu_ctx):
    """The persistent brick DILU launches (variant 6: a workgroup walks its
    items and loads the next one while it sweeps) do k_dilu_bk's arithmetic:
    one application and batched solves are bitwise equal."""
    print("case brick dilu persistent", flush=True)
    o, g = _grid(gpu_ctx, 44)
    grid = {"dx": g[0], "dy": g[1], "dz": g[2], "nzAir": g[3]}
    rng = np.random.default_rng(6)
    x = rng.standard_normal(o.nInt) + 1j * rng.standard_normal(o.nInt)
    om, B, X0 = P.forward_systems(gpu_ctx, grid, [1e-2, 0.5, 20.0])
    res = {}
    try:
        for v in (5, 6):
            gpu_ctx.set_dilu_variant(v)
            z = [gpu_ctx.inner_apply(2 * np.pi * 0.7, x, what=1, transpose=t) for t in (False, True)]
            X, st = gpu_ctx.solve_batch(om, B, X0, tol=1e-10, maxit=4000, precision=mtsp2.PREC_MIXED,
                                        precond=mtsp2.PC_DILU_RB)
            res[v] = (z, X, [(s["niter"], s["nouter"]) for s in st])
    finally:
        gpu_ctx.set_dilu_variant(6)   # the library's default
    for t in range(2):
        assert np.array_equal(res[5][0][t], res[6][0][t]), t
    assert res[5][2] == res[6][2]
    assert np.array_equal(res[5][1], res[6][1])


@pytest.mark.parametrize("transpose,nsys", [(False, 5), (True, 5), (False, 1)])
def test_brick_paired_x_matches_unpaired(gpu_ctx, transpose, nsys):
    """Paired x (p-hat and s-hat interleaved by system pair, written so by the
    DILU and gathered 16 bytes at a time for both systems by the pair SpMV)
    does the unpaired kernels' arithmetic: batched mixed solves are bitwise
    equal.  Five systems: a pair with one member, and systems that stop at
    different iterations (pairs whose other member has stopped); one system:
    the only pair's second member absent throughout."""
    print(f"case brick paired x transpose={transpose} nsys={nsys}", flush=True)
    o, g = _grid(gpu_ctx, 45)
    grid = {"dx": g[0], "dy": g[1], "dz": g[2], "nzAir": g[3]}
    om, B, X0 = P.forward_systems(gpu_ctx, grid, [1e-2, 0.5, 20.0])
    om, B, X0 = om[:nsys], B[:nsys], X0[:nsys]
    res = {}
    try:
        for layout in (2, 1):
            gpu_ctx.set_inner_layout(layout)
            X, st = gpu_ctx.solve_batch(om, B, X0, tol=1e-10, maxit=4000, precision=mtsp2.PREC_MIXED,
                                        precond=mtsp2.PC_DILU_RB, op_transpose=int(transpose))
            res[layout] = (X, [(s["status"], s["niter"], s["nouter"]) for s in st])
    finally:
        gpu_ctx.set_inner_layout(True)   # the library's default (paired)
    assert all(s == 0 for s, _, _ in res[2][1]), res[2][1]
    assert res[2][1] == res[1][1]
    if nsys > 1:
        assert len({n for _, n, _ in res[2][1]}) > 1, res[2][1]
    assert np.array_equal(res[2][0], res[1][0])


def test_brick_paired_graph_and_eager_agree(gpu_ctx):
    """The mixed solver with paired x replayed as hipGraphs gives the eager
    launches' iterates: short chunks, so the graphs are captured again as
    systems stop (the pair list and its length change with the active
    list)."""
    print("case brick paired graph-vs-eager", flush=True)
    o, g = _grid(gpu_ctx, 46)
    grid = {"dx": g[0], "dy": g[1], "dz": g[2], "nzAir": g[3]}
    om, B, X0 = P.forward_systems(gpu_ctx, grid, [1e-2, 0.5, 20.0])
    om, B, X0 = om[:5], B[:5], X0[:5]
    res = {}
    for graph in (1, 0):
        X, st = gpu_ctx.solve_batch(om, B, X0, tol=1e-10, maxit=4000, precision=mtsp2.PREC_MIXED,
                                    precond=mtsp2.PC_DILU_RB, use_graph=graph, chunk=3)
        res[graph] = (X, [(s["status"], s["niter"], s["nouter"]) for s in st])
    assert all(s == 0 for s, _, _ in res[1][1]), res[1][1]
    assert res[1][1] == res[0][1]
    assert np.array_equal(res[1][0], res[0][0])


@pytest.mark.parametrize("transpose", [False, True])
def test_brick_solves_match_all_edge_layout(gpu_ctx, transpose):
    """Mixed batched solves (red-black DILU) with the inner vectors on
    bricks and in the all-edge layout: both reach the FP64 tolerance and the
    direct solution of A x = b (A^T x = b); the iteration counts stay alike."""
    print(f"case brick solves transpose={transpose}", flush=True)
    o, g = _grid(gpu_ctx, 43)
    grid = {"dx": g[0], "dy": g[1], "dz": g[2], "nzAir": g[3]}
    om, B, X0 = P.forward_systems(gpu_ctx, grid, [1e-2, 0.5, 20.0])
    res = {}
    try:
        for bricks in (True, False):
            gpu_ctx.set_inner_layout(bricks)
            X, stats = gpu_ctx.solve_batch(om, B, X0, tol=1e-10, maxit=4000,
                                           precision=mtsp2.PREC_MIXED, precond=mtsp2.PC_DILU_RB,
                                           op_transpose=int(transpose))
            res[bricks] = (X, stats)
    finally:
        gpu_ctx.set_inner_layout(True)
    import scipy.sparse as sp
    for s in range(len(om)):
        o.update_freq(om[s])
        # A(w) = AAii + i*ISIGN*w*mu0*sigma_E*V_E (ISIGN = -1)
        a = o.csr("AAii")
        A = sp.csr_matrix((a["val"], a["col"], a["row"]), shape=(o.nInt, o.nInt)).astype(complex)
        A = (A - sp.diags(1j * o.vms())).tocsr()
        Ao = A.T.tocsr() if transpose else A
        xo = spla_solve(Ao, B[s])
        for bricks in (True, False):
            X, stats = res[bricks]
            assert stats[s]["status"] == 0, (bricks, stats[s])
            assert np.linalg.norm(B[s] - Ao @ X[s]) / np.linalg.norm(B[s]) <= 1e-10 * 1.0001, bricks
            assert rel(X[s], xo) <= 1e-7, bricks
        nb, na = res[True][1][s]["niter"], res[False][1][s]["niter"]
        assert abs(nb - na) <= 0.25 * max(nb, na) + 10, (nb, na)
