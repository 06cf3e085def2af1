"""UpdateCond on the GPU (mt_update_cond): a new conductivity model on the
assembled grid rebuilds sigma_E / sigma_N, AAii = CCii + GDii(sigma) and vms1
on the device.  Checked against a fresh host assembly of the same model
(bit for bit, through the operator the device applies), against the oracle,
through solves, and through the host-side entry points that read the
refreshed host copies."""
import time

import numpy as np
import pytest

import sp2oracle
from conftest import rel, small_problem
from mtsp2 import problem as P

pytestmark = pytest.mark.gpu


def _cvec(rng, n):
    return rng.standard_normal(n) + 1j * rng.standard_normal(n)


def test_update_cond_equals_assembly(gpu_ctx):
    print("case update_cond vs assembly", flush=True)
    dx, dy, dz, nza, lc1 = small_problem(21)
    rng = np.random.default_rng(22)
    lc2 = lc1 + rng.normal(0.0, 1.0, lc1.shape)
    ctx = gpu_ctx
    ctx.assemble(dx, dy, dz, nza, lc2)
    x = _cvec(rng, ctx.n_int)
    ref = {w: ctx.apply(w, x) for w in (2 * np.pi * 0.01, 2 * np.pi * 3.0)}
    ctx.assemble(dx, dy, dz, nza, lc1)
    ctx.update_cond(lc2)
    o = sp2oracle.Problem(dx, dy, dz, nza, lc2)
    for w, y_ref in ref.items():
        y = ctx.apply(w, x)
        assert np.array_equal(y, y_ref), w          # same coefficients, bit for bit
        o.update_freq(w)
        assert rel(y, o.mult_aii(x)) <= 1e-13
    # a second update, back to the first model
    ctx.update_cond(lc1)
    o1 = sp2oracle.Problem(dx, dy, dz, nza, lc1)
    o1.update_freq(2 * np.pi * 3.0)
    assert rel(ctx.apply(2 * np.pi * 3.0, x), o1.mult_aii(x)) <= 1e-13


def test_update_cond_solves_and_host_paths(gpu_ctx):
    print("case update_cond solves", flush=True)
    grid = P.make_grid(14, 12, 8, nair=4, npad=3)
    dx, dy, dz, nza = grid["dx"], grid["dy"], grid["dz"], grid["nzAir"]
    nze = len(dz) - nza
    rng = np.random.default_rng(23)
    lc1 = np.full((len(dx), len(dy), nze), np.log(0.01))
    lc2 = P.boxes_model(grid, nbox=4, seed=5)
    ctx = gpu_ctx
    ctx.assemble(dx, dy, dz, nza, lc1)
    ctx.update_cond(lc2)
    o = sp2oracle.Problem(dx, dy, dz, nza, lc2)
    om, B, X0 = P.forward_systems(ctx, grid, [0.1, 10.0])
    X, st = ctx.solve_batch(om, B, X0, tol=1e-10, maxit=3000)
    assert all(s["status"] == 0 for s in st), st
    for s in range(len(om)):
        o.update_freq(om[s])
        r = B[s] - o.mult_aii(X[s])
        assert np.linalg.norm(r) / np.linalg.norm(B[s]) <= 1e-10 * 1.0001
    # the transposed operator is rebuilt from the new model too
    Xt, st = ctx.solve_batch(om[:1], B[:1], None, tol=1e-10, maxit=3000, op_transpose=1)
    assert st[0]["status"] == 0, st
    o.update_freq(om[0])
    # A^T x = conj(A^H conj(x)) for A = AAii + i*D with AAii and D real
    at = np.conj(o.mult_aii(np.conj(Xt[0]), adjt=True))
    assert np.linalg.norm(B[0] - at) / np.linalg.norm(B[0]) <= 1e-9
    # host-side entry points read host copies rebuilt for the new model
    s_all = _cvec(rng, o.nE)
    assert rel(ctx.rhs_trn(om[0], s_all), o.rhs_trn(om[0], s_all)) <= 1e-12
    import scipy.sparse as sp
    csr = ctx.operator_csr()
    a = o.csr("AAii")
    n = o.nInt
    A_lib = sp.csr_matrix((csr["val"], csr["col"], csr["row"]), shape=(n, n))
    A_or = sp.csr_matrix((a["val"], a["col"], a["row"]), shape=(n, n))
    assert abs(A_lib - A_or).max() <= 1e-13 * abs(A_or).max()
    assert rel(om[0] * csr["vms1"], o.vms()) <= 1e-13      # the oracle holds omega*vms1


def test_update_cond_time_vs_assembly(gpu_ctx):
    """The same model change by host assembly and by mt_update_cond, timed
    (100x100x60+10)."""
    grid = P.config_grid("cfg1")
    lc = P.config_model("cfg1", grid, 0)
    ctx = gpu_ctx
    t0 = time.time()
    ctx.assemble(grid["dx"], grid["dy"], grid["dz"], grid["nzAir"], lc)
    t_asm = time.time() - t0
    x = _cvec(np.random.default_rng(3), ctx.n_int)
    y0 = ctx.apply(1.0, x)
    ctx.update_cond(lc + 0.5)
    t0 = time.time()
    ctx.update_cond(lc)          # second call: set-up data already on the device
    t_upd = time.time() - t0
    print(f"assembly {t_asm * 1e3:.0f} ms, update_cond {t_upd * 1e3:.1f} ms", flush=True)
    assert np.array_equal(ctx.apply(1.0, x), y0)
    assert t_upd < t_asm
