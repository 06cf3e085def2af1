"""The largest workload of the survey's configurations, 120x120x80+10 (3.8 M
interior edges, 57 M couplings), with all of its 32 systems (16 frequencies
1e-4..1e3 Hz x 2 modes): assembly, device boundary fields, device
right-hand sides and the mixed batched solve of the 32 systems together,
checked through the FP64 true residual computed by the library's own FP64
operator (mt_apply, itself checked against the oracle elsewhere).  The oracle
is checked bit for bit at 100x100x60+10 elsewhere; here the sizes are the
point (memory, 32-bit indexing, launch shapes)."""
import numpy as np
import pytest

import mtsp2
from mtsp2 import problem as P

pytestmark = pytest.mark.gpu


@pytest.mark.timeout(600)
def test_largest_configuration(gpu_ctx):
    print("case 120x120x80+10, 32 systems", flush=True)
    grid = P.config_grid("cfg3")
    lc = P.config_model("cfg3", grid, 0)
    ctx = gpu_ctx
    ctx.assemble(grid["dx"], grid["dy"], grid["dz"], grid["nzAir"], lc)
    assert ctx.n_int == 3816330
    fr = P.config_freqs("cfg3")
    # all 32 systems: 16 frequencies x modes 1, 2, in one batched solve
    om = 2 * P.PI * np.repeat(fr, 2)
    md = np.tile(np.array([1, 2], np.int32), len(fr))
    ns, n, ne = len(om), ctx.n_int, ctx.n_edge
    pe, pb, px = ctx.dev_alloc(ns * ne * 16), ctx.dev_alloc(ns * n * 16), ctx.dev_alloc(ns * n * 16)
    try:
        ctx.bc_x0_dev(om, md, pe)
        ctx.rhs_fwd_dev(ns, pe, pb, px)
        st = ctx.solve_batch_dev(om, pb, px, tol=1e-7, maxit=4000, precision=mtsp2.PREC_MIXED,
                                 precond=mtsp2.PC_DILU_RB)
        b = ctx.download(pb, np.zeros((ns, n), np.complex128))
        x = ctx.download(px, np.zeros((ns, n), np.complex128))
    finally:
        for q in (pe, pb, px):
            ctx.dev_free(q)
    print([(s["niter"], s["nouter"], s["relres"]) for s in st], flush=True)
    assert all(s["status"] == 0 for s in st), st
    for s in range(ns):
        r = b[s] - ctx.apply(om[s], x[s])
        assert np.linalg.norm(r) / np.linalg.norm(b[s]) <= 1e-7 * 1.001, s
