"""GPU parity of libmtsp2 against the oracle (oracle/sp2_oracle.c, itself
bit-exact against the flang build of the reference, tests/test_oracle_vs_ref.py).

Tolerances (FP64 unless stated):
  operator apply A x, A^T x, A^H x       rel L2 <= 1e-13 (summation order only)
  block-DILU apply U^-1 L^-1 x           rel L2 <= 1e-12
  BiCGStab (cuBiCG semantics, tol 1e-10) rel L2 of x <= 1e-8, true relres <= 1e-10
  batched FP64 / mixed (tol 1e-10)       rel L2 of x <= 1e-7, true FP64 relres <= tol
"""
import numpy as np
import pytest

import mtsp2
import sp2oracle
from conftest import (block_diag_operator, dilu_factor, dilu_mult, rb_tile_keys, rel,
                      small_problem)
from mtsp2 import problem as P

pytestmark = pytest.mark.gpu


def _setup(ctx, seed=0, **kw):
    dx, dy, dz, nza, lc = small_problem(seed, **kw)
    o = sp2oracle.Problem(dx, dy, dz, nza, lc)
    ctx.assemble(dx, dy, dz, nza, lc)
    assert ctx.n_int == o.nInt
    return o, (dx, dy, dz, nza, lc)


def _cvec(rng, n):
    return rng.standard_normal(n) + 1j * rng.standard_normal(n)


@pytest.mark.parametrize("seed", [0, 1])
def test_apply_matches_oracle(gpu_ctx, seed):
    print(f"case apply seed={seed}", flush=True)
    o, _ = _setup(gpu_ctx, seed)
    rng = np.random.default_rng(10 + seed)
    x = _cvec(rng, o.nInt)
    for w in (2 * np.pi * 1e-3, 2 * np.pi * 10.0):
        o.update_freq(w)
        assert rel(gpu_ctx.apply(w, x), o.mult_aii(x)) <= 1e-13
        # A^H = the reference's Mult_Aii(adjt=.true.)
        assert rel(gpu_ctx.apply(w, x, transpose=2), o.mult_aii(x, adjt=True)) <= 1e-13
        # A^T x = conj(A^H conj(x)) because AAii is real
        assert rel(gpu_ctx.apply(w, x, transpose=1), np.conj(o.mult_aii(np.conj(x), adjt=True))) <= 1e-13


def test_apply_from_reference_csr(gpu_ctx):
    """Drop-in path: the reference's 1-based AAii CSR uploaded as-is."""
    print("case apply-from-csr", flush=True)
    dx, dy, dz, nza, lc = small_problem(2)
    o = sp2oracle.Problem(dx, dy, dz, nza, lc)
    A = o.csr("AAii")
    gpu_ctx.set_operator_csr((len(dx), len(dy), len(dz)), o.nInt, A["row"] + 1, A["col"] + 1,
                             A["val"], o.vms(), base=1)
    rng = np.random.default_rng(7)
    x = _cvec(rng, o.nInt)
    w = 2 * np.pi * 0.5
    o.update_freq(w)
    assert rel(gpu_ctx.apply(w, x), o.mult_aii(x)) <= 1e-14


@pytest.mark.parametrize("f_hz", [1e-3, 1.0, 100.0])
def test_dilu_apply_matches_oracle(gpu_ctx, f_hz):
    print(f"case dilu f={f_hz}", flush=True)
    o, _ = _setup(gpu_ctx, 3, nx=19, ny=17, nzAir=4, nzEarth=13)
    w = 2 * np.pi * f_hz
    o.update_freq(w)
    rng = np.random.default_rng(5)
    x = _cvec(rng, o.nInt)
    ref = o.pc_usolve(o.pc_lsolve(x))
    assert rel(gpu_ctx.precond_apply(w, x, mtsp2.PC_DILU), ref) <= 1e-12


@pytest.mark.parametrize("pc", [mtsp2.PC_DILU, mtsp2.PC_JACOBI])
def test_bicg_matches_oracle(gpu_ctx, pc):
    print(f"case bicg pc={pc}", flush=True)
    o, g = _setup(gpu_ctx, 4, nx=14, ny=12, nzAir=4, nzEarth=10)
    w = 2 * np.pi * 0.1
    o.update_freq(w)
    grid = {"dx": g[0], "dy": g[1], "dz": g[2], "nzAir": g[3]}
    e = P.plane_wave(grid, w, 0)
    b = o.rhs_fwd(e)
    assert rel(gpu_ctx.rhs_fwd(e), b) <= 1e-13
    x0 = e[o.index()[0]]
    xg, nit, rerr, failed = gpu_ctx.bicg(w, b, x0, maxit=600, tol=1e-10, precond=pc)
    xo, info = o.bicg(b, x0, maxit=600, tol=1e-10, gpu_semantics=True, precond=pc)
    assert info["status"] == 0 and not failed
    assert abs(nit - info["niter"]) <= max(3, info["niter"] // 10)
    assert rel(xg, xo) <= 1e-8
    r = b - o.mult_aii(xg)
    assert np.linalg.norm(r) / np.linalg.norm(b) <= 1e-10
    assert rerr[nit - 1] <= 1e-10


@pytest.mark.parametrize("precision", [mtsp2.PREC_FP64, mtsp2.PREC_MIXED])
def test_batch_matches_oracle(gpu_ctx, precision):
    print(f"case batch precision={precision}", flush=True)
    o, g = _setup(gpu_ctx, 5, nx=14, ny=12, nzAir=4, nzEarth=10)
    grid = {"dx": g[0], "dy": g[1], "dz": g[2], "nzAir": g[3]}
    om, B, X0 = P.forward_systems(gpu_ctx, grid, [1e-3, 1.0, 100.0])
    X, stats = gpu_ctx.solve_batch(om, B, X0, tol=1e-10, maxit=2000, precision=precision,
                                   precond=mtsp2.PC_DILU)
    for s in range(len(om)):
        o.update_freq(om[s])
        assert stats[s]["status"] == 0, stats[s]
        r = B[s] - o.mult_aii(X[s])
        assert np.linalg.norm(r) / np.linalg.norm(B[s]) <= 1e-10 * 1.0001
        xo, info = o.bicg(B[s], X0[s], maxit=2000, tol=1e-10, gpu_semantics=True)
        assert rel(X[s], xo) <= 1e-7


def test_graph_and_eager_agree(gpu_ctx):
    """The hipGraph-replayed chunks give the same iterates as eager launches."""
    print("case graph-vs-eager", flush=True)
    o, g = _setup(gpu_ctx, 6, nx=12, ny=12, nzAir=4, nzEarth=8)
    grid = {"dx": g[0], "dy": g[1], "dz": g[2], "nzAir": g[3]}
    om, B, X0 = P.forward_systems(gpu_ctx, grid, [0.01, 10.0])
    Xg, sg = gpu_ctx.solve_batch(om, B, X0, tol=1e-9, use_graph=1, chunk=5)
    Xe, se = gpu_ctx.solve_batch(om, B, X0, tol=1e-9, use_graph=0, chunk=5)
    assert [s["niter"] for s in sg] == [s["niter"] for s in se]
    assert np.array_equal(Xg, Xe)


def test_divcor_matches_oracle(gpu_ctx):
    """SdivCorr on the GPU (DivC, node-DILU PCG, Grad update) against the
    oracle's DivC / PCG / Grad, which are bit-exact with the reference."""
    print("case divcor", flush=True)
    o, g = _setup(gpu_ctx, 8, nx=14, ny=12, nzAir=4, nzEarth=10)
    w = 2 * np.pi * 1.0
    o.update_freq(w)
    o.divcor_setup()
    rng = np.random.default_rng(9)
    ei = _cvec(rng, o.nInt)
    phi_rhs = np.zeros(o.nNi, complex)
    import ctypes
    sp2oracle.lib().sp2o_divc(o._h, ei, phi_rhs)
    phi, info = o.pcg(phi_rhs, maxit=200, tol=1e-9)
    phi_all = np.zeros(o.nNode, complex)
    phi_all[o.index()[2]] = phi
    ge = np.zeros(o.nInt, complex)
    sp2oracle.lib().sp2o_grad(o._h, phi_all, ge)
    e_ref = ei - ge
    eg, nit, rr = gpu_ctx.divcor(w, ei, maxit=200, tol=1e-9)
    assert abs(nit - info["niter"]) <= 2
    assert rr <= 1e-9
    assert rel(eg, e_ref) <= 1e-8
    d_after = np.zeros(o.nNi, complex)
    sp2oracle.lib().sp2o_divc(o._h, np.ascontiguousarray(eg), d_after)
    assert np.linalg.norm(d_after) <= 1e-6 * np.linalg.norm(phi_rhs)


def test_kernel_timing_graph_and_eager(gpu_ctx):
    """HIP-event kernel timing: with timing on, use_graph is overridden (eager
    launches), so both settings count the same launches and bytes; the times
    are positive."""
    print("case kernel timing", flush=True)
    o, g = _setup(gpu_ctx, 7, nx=12, ny=11, nzAir=4, nzEarth=8)
    grid = {"dx": g[0], "dy": g[1], "dz": g[2], "nzAir": g[3]}
    om, B, X0 = P.forward_systems(gpu_ctx, grid, [0.1, 10.0])
    res = {}
    for graph in (1, 0):
        gpu_ctx.timing_enable(True)
        try:
            X, st = gpu_ctx.solve_batch(om, B, X0, tol=1e-8, precision=mtsp2.PREC_MIXED,
                                        use_graph=graph, chunk=4)
            res[graph] = [gpu_ctx.kernel_timing(f) for f in (mtsp2.KT_SPMV, mtsp2.KT_PRECOND)]
        finally:
            gpu_ctx.timing_enable(False)
        assert all(s["status"] == 0 for s in st)
    for f in range(2):
        (n1, ms1, b1), (n0, ms0, b0) = res[1][f], res[0][f]
        assert n1 == n0 > 0 and b1 == b0 > 0
        assert ms1 > 0 and ms0 > 0
    # after disabling, nothing more is counted
    gpu_ctx.solve_batch(om, B, X0, tol=1e-8, precision=mtsp2.PREC_MIXED)
    assert gpu_ctx.kernel_timing(mtsp2.KT_SPMV) == (0, 0.0, 0.0)


@pytest.fixture
def dilu_variant(gpu_ctx, request):
    """Launch shape of the red-black DILU for one test (restored after)."""
    gpu_ctx.set_dilu_variant(request.param)
    yield request.param
    gpu_ctx.set_dilu_variant(6)   # the library's default


@pytest.mark.parametrize("dilu_variant", [4, 5, 1], indirect=True)
def test_dilu_rb_apply_inverts_its_factorisation(gpu_ctx, dilu_variant):
    """MT_PC_DILU_RB applies M^-1 with M = (P+E) P^-1 (P+F), E and F the
    strictly lower and upper parts of the x-x, y-y, z-z blocks of A(omega) in
    the red-black tile ordering and P the DILU pivots of that ordering
    (dilu_factor, pinned on the natural ordering against the oracle's
    PC_Lsolve/PC_Usolve in tests/test_oracle_invariants.py): M z = x, for
    every launch shape (1 single-system workgroups, 4 four systems per
    workgroup, 5 the same with the lean LDS layout).  The grid has partial
    tiles in every direction."""
    print(f"case dilu-rb apply variant={dilu_variant}", flush=True)
    dx, dy, dz, nza, lc = small_problem(12, nx=20, ny=18, nzAir=4, nzEarth=10)
    o = sp2oracle.Problem(dx, dy, dz, nza, lc)
    gpu_ctx.assemble(dx, dy, dz, nza, lc)
    w = 2 * np.pi * 3.0
    key = rb_tile_keys(len(dx), len(dy), len(dz))
    piv, lower, upper = dilu_factor(block_diag_operator(o, w), key)
    rng = np.random.default_rng(4)
    x = rng.standard_normal(o.nInt) + 1j * rng.standard_normal(o.nInt)
    z = gpu_ctx.precond_apply(w, x, mtsp2.PC_DILU_RB)
    assert rel(dilu_mult(piv, lower, upper, z), x) <= 1e-12


def test_dilu_rb_lean_layout_matches(gpu_ctx):
    """The lean-LDS launches (variant 5) against the four-systems launches
    (variant 4) on a grid of full and partial tiles: the same operator (the
    black backward sweep rounds differently), FP64."""
    print("case dilu-rb variants 4/5", flush=True)
    dx, dy, dz, nza, lc = small_problem(21, nx=26, ny=17, nzAir=5, nzEarth=12)
    gpu_ctx.assemble(dx, dy, dz, nza, lc)
    rng = np.random.default_rng(5)
    x = rng.standard_normal(gpu_ctx.n_int) + 1j * rng.standard_normal(gpu_ctx.n_int)
    try:
        zs = []
        for v in (4, 5):
            gpu_ctx.set_dilu_variant(v)
            zs.append(gpu_ctx.precond_apply(2 * np.pi * 0.3, x, mtsp2.PC_DILU_RB))
    finally:
        gpu_ctx.set_dilu_variant(6)   # the library's default
    assert rel(zs[1], zs[0]) <= 1e-13


@pytest.mark.parametrize("dilu_variant", [4, 5], indirect=True)
@pytest.mark.parametrize("precision", [mtsp2.PREC_FP64, mtsp2.PREC_MIXED])
def test_batch_dilu_rb_converges(gpu_ctx, precision, dilu_variant):
    """Solves preconditioned with the red-black tile DILU reach the FP64 true
    residual tolerance and agree with the oracle's solution."""
    print(f"case batch dilu-rb precision={precision} variant={dilu_variant}", flush=True)
    o, g = _setup(gpu_ctx, 13, nx=20, ny=18, nzAir=4, nzEarth=10)
    grid = {"dx": g[0], "dy": g[1], "dz": g[2], "nzAir": g[3]}
    om, B, X0 = P.forward_systems(gpu_ctx, grid, [1e-2, 10.0])
    X, stats = gpu_ctx.solve_batch(om, B, X0, tol=1e-10, maxit=3000, precision=precision,
                                   precond=mtsp2.PC_DILU_RB)
    for s in range(len(om)):
        o.update_freq(om[s])
        assert stats[s]["status"] == 0, stats[s]
        r = B[s] - o.mult_aii(X[s])
        assert np.linalg.norm(r) / np.linalg.norm(B[s]) <= 1e-10 * 1.0001
        xo, info = o.bicg(B[s], X0[s], maxit=3000, tol=1e-10, gpu_semantics=True)
        assert rel(X[s], xo) <= 1e-7


@pytest.mark.parametrize("precision", [mtsp2.PREC_FP64, mtsp2.PREC_MIXED])
def test_spmv_variants_agree(gpu_ctx, precision):
    """The k-marching SpMV (LDS layers of an (i,j) tile, partial tiles in i and
    j, several k segments), the multi-system row SpMV (4 or 8 systems per
    block, partly filled groups) and the row-parallel SpMV drive the solver to
    the same solutions; all meet the FP64 true residual."""
    print(f"case spmv variants precision={precision}", flush=True)
    o, g = _setup(gpu_ctx, 14, nx=37, ny=13, nzAir=4, nzEarth=15)
    grid = {"dx": g[0], "dy": g[1], "dz": g[2], "nzAir": g[3]}
    om, B, X0 = P.forward_systems(gpu_ctx, grid, [0.05, 20.0])
    res = {}
    try:
        for var, ks in ((0, 0), (2, 3), (2, 16), (3, 4), (3, 8)):
            gpu_ctx.set_spmv_variant(var, ks)
            X, st = gpu_ctx.solve_batch(om, B, X0, tol=1e-10, maxit=3000, precision=precision)
            for s in range(len(om)):
                o.update_freq(om[s])
                assert st[s]["status"] == 0, (var, ks, st[s])
                r = B[s] - o.mult_aii(X[s])
                assert np.linalg.norm(r) / np.linalg.norm(B[s]) <= 1e-10 * 1.0001
            res[(var, ks)] = X
    finally:
        gpu_ctx.set_spmv_variant(3, 4)
    assert rel(res[(2, 3)], res[(0, 0)]) <= 1e-7
    assert rel(res[(2, 16)], res[(0, 0)]) <= 1e-7
    assert rel(res[(3, 4)], res[(0, 0)]) <= 1e-7
    assert rel(res[(3, 8)], res[(0, 0)]) <= 1e-7

def test_rhs_fwd_dev_matches_oracle(gpu_ctx):
    """Mult_Aib for a batch on the device (b = -CCib e_b, x0 = e(EDGEi)) from
    random all-edge fields, against the oracle's FWDsolve3D RHS."""
    print("case rhs_fwd_dev", flush=True)
    o, _ = _setup(gpu_ctx, 6)
    rng = np.random.default_rng(61)
    nsys = 3
    e = np.ascontiguousarray(np.array([_cvec(rng, o.nE) for _ in range(nsys)]))
    ei = o.index()[0]
    pe = gpu_ctx.dev_alloc(e.nbytes)
    pb = gpu_ctx.dev_alloc(nsys * o.nInt * 16)
    px = gpu_ctx.dev_alloc(nsys * o.nInt * 16)
    try:
        gpu_ctx.upload(pe, e)
        gpu_ctx.rhs_fwd_dev(nsys, pe, pb, px)
        b = gpu_ctx.download(pb, np.zeros((nsys, o.nInt), complex))
        x0 = gpu_ctx.download(px, np.zeros((nsys, o.nInt), complex))
    finally:
        for q in (pe, pb, px):
            gpu_ctx.dev_free(q)
    for s in range(nsys):
        want = o.rhs_fwd(e[s])
        assert rel(b[s], want) <= 1e-13, s
        assert np.array_equal(x0[s], e[s][ei]), s
