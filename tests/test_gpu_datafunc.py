"""Data functionals on the GPU (mt_receivers_set, mt_data_resp, mt_lmultT)
against the oracle's dataResp / LmultT (oracle/sp2_oracle.c, itself bit-exact
against the reference's DataFunc/DataSens): every data type, batched over
frequencies, default, rotated and non-orthogonal channel frames, reference
sites, missing data; and the responses of a GPU forward solve over a
halfspace.  Tolerance: 1e-10 relative to the largest entry (the GPU sums the
same terms in a different order)."""
import numpy as np
import pytest

import mtsp2
import sp2oracle
from conftest import small_problem
from mtsp2 import problem as P

pytestmark = pytest.mark.gpu

NCOMP = {1: 8, 2: 4, 3: 4, 4: 8, 5: 4, 6: 4, 7: 2, 8: 2, 9: 2, 10: 2, 11: 2}
TOL = 1e-10


def _sites(dx, dy, rng, ns):
    x = np.column_stack([rng.uniform(-0.3, 0.3, ns) * dx.sum(),
                         rng.uniform(-0.3, 0.3, ns) * dy.sum(), rng.uniform(0, 80, ns)])
    x[0, 2] = 0.0
    r = x[::-1].copy()
    az = np.tile([0.0, 90.0, 0.0, 0.0, 90.0, 0.0, 90.0], (ns, 1))
    az[1] = [30.0, 120.0, 0.0, 30.0, 120.0, 30.0, 120.0]
    az[2] = [10.0, 95.0, 0.0, 5.0, 100.0, -3.0, 80.0]
    return x, r, az


def _upload(ctx, arr):
    ptr = ctx.dev_alloc(arr.nbytes)
    ctx.upload(ptr, arr)
    return ptr


def test_data_resp_and_lmultT_match_oracle(gpu_ctx):
    print("case data functionals vs oracle, 11 types x 3 frequencies", flush=True)
    dx, dy, dz, nza, lc = small_problem(11)
    p = sp2oracle.Problem(dx, dy, dz, nza, lc)
    rng = np.random.default_rng(12)
    ns, nf = 7, 3
    x, r, az = _sites(dx, dy, rng, ns)
    origin = np.array([-dx.sum() / 2, -dy.sum() / 2, 0.0])
    omegas = 2 * np.pi * np.array([0.01, 1.0, 30.0])
    e = rng.standard_normal((nf, 2, p.nE)) + 1j * rng.standard_normal((nf, 2, p.nE))
    ctx = gpu_ctx
    ctx.receivers_set(dx, dy, dz, nza, x, r=r, az=az, origin=origin)
    e_ptr = _upload(ctx, np.ascontiguousarray(e))
    comb = np.zeros((nf, 2, p.nE), complex)
    c_ptr = ctx.dev_alloc(comb.nbytes)
    try:
        for dt, nc in NCOMP.items():
            got = ctx.data_resp(dt, omegas, e_ptr, nc)
            d = rng.standard_normal((nf, ns, nc))
            ex = np.ones((nf, ns, nc), np.uint8)
            ex[1, 3, 0] = 0
            ctx.lmultT(dt, omegas, e_ptr, d, c_ptr, exist=ex)
            ctx.download(c_ptr, comb)
            for f in range(nf):
                want = p.data_resp(dt, omegas[f], origin, x, e[f, 0], e[f, 1], r=r, az=az)
                scale = np.max(np.abs(want))
                assert np.max(np.abs(got[f] - want)) <= TOL * scale, (dt, f)
                w1, w2 = p.lmultT(dt, omegas[f], origin, x, d[f], e[f, 0], e[f, 1], r=r, az=az,
                                  exist=ex[f])
                for k, w in enumerate((w1, w2)):
                    sc = max(np.max(np.abs(w)), 1e-300)
                    assert np.max(np.abs(comb[f, k] - w)) <= TOL * sc, (dt, f, k)
                    # same support: nothing leaks onto other edges
                    assert np.all(comb[f, k][w == 0] == 0), (dt, f, k)
    finally:
        ctx.dev_free(e_ptr)
        ctx.dev_free(c_ptr)


def test_halfspace_responses_from_gpu_solve(gpu_ctx):
    """Forward solves of a uniform 100 ohm-m halfspace on the GPU, then the
    GPU impedance: rho_a = 100 ohm-m within 2 %, phase 45 degrees within 1,
    diagonal impedances negligible; and the same responses as the oracle's
    dataResp on those fields."""
    print("case halfspace responses from GPU solve", flush=True)
    grid = P.make_grid(20, 20, 40, nair=8, dcore=400.0, npad=5)
    dx, dy, dz, nza = grid["dx"], grid["dy"], grid["dz"], grid["nzAir"]
    nze = len(dz) - nza
    lc = np.full((len(dx), len(dy), nze), np.log(0.01))
    ctx = gpu_ctx
    ctx.assemble(dx, dy, dz, nza, lc)
    om, B, X0 = P.forward_systems(ctx, grid, [1.0])
    E, st = ctx.solve_batch(om, B, X0, tol=1e-10, maxit=3000)
    assert all(s["status"] == 0 for s in st), st
    o = sp2oracle.Problem(dx, dy, dz, nza, lc)
    ei = o.index()[0]
    w = om[0]
    full = np.array([P.plane_wave(grid, w, k) for k in range(2)])
    full[:, ei] = E
    x = np.array([[dx.sum() / 2 + 30.0, dy.sum() / 2 - 20.0, 0.0],
                  [dx.sum() / 2 - 150.0, dy.sum() / 2 + 90.0, 0.0]])
    ctx.receivers_set(dx, dy, dz, nza, x)
    e_ptr = _upload(ctx, np.ascontiguousarray(full[None]))
    try:
        rp = ctx.data_resp(mtsp2.DT_OFF_DIAGONAL_RHO_PHASE, [w], e_ptr, 4)[0]
        z = ctx.data_resp(mtsp2.DT_FULL_IMPEDANCE, [w], e_ptr, 8)[0]
    finally:
        ctx.dev_free(e_ptr)
    print("rho/phase", rp, flush=True)
    for s in range(2):
        for q in (0, 2):
            assert abs(10 ** rp[s, q] - 100.0) <= 2.0, rp
        for q in (1, 3):
            ph = np.degrees(rp[s, q]) % 180.0
            assert min(abs(ph - 45.0), abs(ph - 135.0)) <= 1.0, rp
        zxy = abs(complex(z[s, 2], z[s, 3]))
        assert abs(complex(z[s, 0], z[s, 1])) <= 1e-3 * zxy
        assert abs(complex(z[s, 6], z[s, 7])) <= 1e-3 * zxy
        want = o.data_resp(5, w, np.zeros(3), x, full[0], full[1])
        assert np.max(np.abs(rp - want)) <= TOL * np.max(np.abs(want))


def test_receiver_and_type_errors(gpu_ctx):
    dx, dy, dz, nza, lc = small_problem(2)
    ctx = gpu_ctx
    with pytest.raises(mtsp2.MTError) as ei:
        ctx.receivers_set(dx, dy, dz, nza, np.zeros((0, 3)))
    assert ei.value.code == mtsp2.MT_E_ARG
    ok = np.array([[0.0, 0.0, 10.0]])
    ctx.receivers_set(dx, dy, dz, nza, ok, origin=(-dx.sum() / 2, -dy.sum() / 2, 0.0))
    ne = len(dx) * (len(dy) + 1) * (len(dz) + 1) + (len(dx) + 1) * len(dy) * (len(dz) + 1) + \
        (len(dx) + 1) * (len(dy) + 1) * len(dz)
    e_ptr = ctx.dev_alloc(2 * ne * 16)
    try:
        with pytest.raises(mtsp2.MTError) as ei:
            ctx.data_resp(12, [1.0], e_ptr, 2)
        assert ei.value.code == mtsp2.MT_E_ARG
        with pytest.raises(mtsp2.MTError) as ei:
            ctx.data_resp(1, [-1.0], e_ptr, 8)
        assert ei.value.code == mtsp2.MT_E_ARG
    finally:
        ctx.dev_free(e_ptr)
