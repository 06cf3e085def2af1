"""BC_x0_WS on the device (mt_bc_x0_dev): the starting fields of a batch of
forward systems from the 2-D TE slice solves, bit for bit against the oracle
(itself bit-exact against the reference) and the reference's own fixture, at
small sizes, on a deep grid (bandwidth 69) and at 100x100x60+10; then as the
start of a forward solve through mt_rhs_fwd_dev."""
import os
import time

import numpy as np
import pytest

import sp2oracle as O
from conftest import rel
from mtsp2 import problem as P

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden", "ref_bcx0.npz")


def _want(dx, dy, dz, nza, lc, omegas, modes):
    return np.array([O.bc_form(O.bc_x0(dx, dy, dz, nza, lc, w, m)[0], len(dx), len(dy), len(dz))
                     for w, m in zip(omegas, modes)])


def test_bc_x0_dev_reference_fixture(gpu_ctx):
    print("case bc_x0 fixture", flush=True)
    g = np.load(GOLD)
    dx, dy, dz, nza, lc = g["dx"], g["dy"], g["dz"], int(g["nzAir"]), g["logcond"]
    ctx = gpu_ctx
    ctx.assemble(dx, dy, dz, nza, lc)
    e = ctx.bc_x0(g["omega"], g["mode"])
    pb = O.Problem(dx, dy, dz, nza, lc)
    ei, eb, _, _ = pb.index()
    for s in range(len(g["omega"])):
        assert np.array_equal(e[s][ei], g["e0"][s][ei]), s     # E0 inside
        assert np.array_equal(e[s][eb], g["bc"][s][eb]), s     # setBC(BC) outside


def test_bc_x0_dev_deep_and_mixed_batch(gpu_ctx):
    print("case bc_x0 deep", flush=True)
    g = np.load(GOLD)
    dx, dy, dz, nza, lc = (g["deep_dx"], g["deep_dy"], g["deep_dz"], int(g["deep_nzAir"]),
                           g["deep_logcond"])
    ctx = gpu_ctx
    ctx.assemble(dx, dy, dz, nza, lc)
    e = ctx.bc_x0(g["deep_omega"], g["deep_mode"])
    nx, ny, nz = len(dx), len(dy), len(dz)
    nxe, nye = nx * (ny + 1) * (nz + 1), (nx + 1) * ny * (nz + 1)
    ey = g["deep_ey"].reshape(nz + 1, ny, nx + 1).copy()
    ex = g["deep_ex"].reshape(nz + 1, ny + 1, nx).copy()
    ey[nz] = 0
    ex[nz] = 0                                                 # the zeroed bottom face
    assert np.array_equal(e[0][nxe:nxe + nye], ey.ravel())
    assert np.array_equal(e[1][:nxe], ex.ravel())
    # a non-square grid, both modes and several frequencies in one batch
    grid = P.make_grid(13, 9, 14, nair=6, npad=3)
    lc2 = P.boxes_model(grid, nbox=12, seed=4)
    ctx.assemble(grid["dx"], grid["dy"], grid["dz"], grid["nzAir"], lc2)
    om = 2 * np.pi * np.array([1e-3, 0.1, 10.0, 0.1, 300.0])
    md = np.array([1, 2, 1, 1, 2])
    e = ctx.bc_x0(om, md)
    want = _want(grid["dx"], grid["dy"], grid["dz"], grid["nzAir"], lc2, om, md)
    for s in range(len(om)):
        assert np.array_equal(e[s], want[s]), s
    # after mt_update_cond the boundary fields follow the new model
    lc3 = lc2 + 0.7
    ctx.update_cond(lc3)
    e = ctx.bc_x0(om[:2], md[:2])
    want = _want(grid["dx"], grid["dy"], grid["dz"], grid["nzAir"], lc3, om[:2], md[:2])
    assert np.array_equal(e, want)


def test_bc_x0_dev_starts_a_forward_solve(gpu_ctx):
    print("case bc_x0 solve", flush=True)
    grid = P.make_grid(14, 12, 10, nair=5, npad=3)
    lc = P.boxes_model(grid, nbox=8, seed=6)
    ctx = gpu_ctx
    ctx.assemble(grid["dx"], grid["dy"], grid["dz"], grid["nzAir"], lc)
    o = O.Problem(grid["dx"], grid["dy"], grid["dz"], grid["nzAir"], lc)
    om = 2 * np.pi * np.array([0.05, 0.05, 5.0, 5.0])
    md = np.array([1, 2, 1, 2])
    ns, n, ne = len(om), ctx.n_int, ctx.n_edge
    pe, pb, px = ctx.dev_alloc(ns * ne * 16), ctx.dev_alloc(ns * n * 16), ctx.dev_alloc(ns * n * 16)
    try:
        ctx.bc_x0_dev(om, md, pe)
        ctx.rhs_fwd_dev(ns, pe, pb, px)
        e = ctx.download(pe, np.zeros((ns, ne), complex))
        b = ctx.download(pb, np.zeros((ns, n), complex))
        x0 = ctx.download(px, np.zeros((ns, n), complex))
    finally:
        for q in (pe, pb, px):
            ctx.dev_free(q)
    X, st = ctx.solve_batch(om, b, x0, tol=1e-10, maxit=3000)
    assert all(s["status"] == 0 for s in st), st
    ei = o.index()[0]
    for s in range(ns):
        assert rel(b[s], o.rhs_fwd(e[s])) <= 1e-13
        assert np.array_equal(x0[s], e[s][ei])
        o.update_freq(om[s])
        r = b[s] - o.mult_aii(X[s])
        assert np.linalg.norm(r) / np.linalg.norm(b[s]) <= 1e-10 * 1.0001


def test_bc_x0_dev_full_size(gpu_ctx):
    """100x100x60+10, the 16 systems of 8 frequencies x 2 modes: timed, and
    one system per mode checked bit for bit against the oracle."""
    print("case bc_x0 100x100x70", flush=True)
    grid = P.config_grid("cfg1")
    lc = P.config_model("cfg1", grid, 0)
    ctx = gpu_ctx
    ctx.assemble(grid["dx"], grid["dy"], grid["dz"], grid["nzAir"], lc)
    fr = P.config_freqs("cfg1")
    om = np.repeat(2 * np.pi * np.asarray(fr, np.float64), 2)
    md = np.tile([1, 2], len(fr))
    ctx.bc_x0(om, md)                                  # first call sizes the scratch
    t0 = time.time()
    e = ctx.bc_x0(om, md)
    dt = time.time() - t0
    print(f"bc_x0 16 systems {dt * 1e3:.1f} ms", flush=True)
    for s in (0, len(om) - 1):
        want = _want(grid["dx"], grid["dy"], grid["dz"], grid["nzAir"], lc, om[s:s + 1], md[s:s + 1])
        assert np.array_equal(e[s], want[0]), s


def test_bc_x0_dev_deep_window(gpu_ctx):
    """nz = 90 (the 120x120x80+10 configuration's depth): the band window
    takes 131 KB of LDS, one slice per CU; bit for bit against the oracle."""
    print("case bc_x0 nz=90", flush=True)
    grid = P.make_grid(5, 4, 80, nair=10, npad=2)
    lc = P.boxes_model(grid, nbox=6, seed=8)
    ctx = gpu_ctx
    ctx.assemble(grid["dx"], grid["dy"], grid["dz"], grid["nzAir"], lc)
    assert len(grid["dz"]) == 90
    om = 2 * np.pi * np.array([3e-3, 20.0])
    md = np.array([1, 2])
    e = ctx.bc_x0(om, md)
    want = _want(grid["dx"], grid["dy"], grid["dz"], grid["nzAir"], lc, om, md)
    assert np.array_equal(e, want)


def test_bc_x0_dev_argument_errors(gpu_ctx):
    """Invalid modes and frequencies are refused before any launch."""
    print("case bc_x0 errors", flush=True)
    import mtsp2
    dx, dy, dz = np.full(4, 300.0), np.full(3, 300.0), np.concatenate([[3000.0, 1000.0], np.full(4, 100.0)])
    lc = np.full((4, 3, 4), np.log(0.01))
    ctx = gpu_ctx
    ctx.assemble(dx, dy, dz, 2, lc)
    for om, md in (([1.0], [3]), ([0.0], [1]), ([-2.0], [2])):
        with pytest.raises(mtsp2.MTError) as ei:
            ctx.bc_x0(np.array(om), np.array(md))
        assert ei.value.code == mtsp2.MT_E_ARG
