"""mt_lmultT_noalloc (no device allocation: its scratch is a file-scope
device array) against mt_lmultT on the same inputs, bit for bit (the same
kernels on the same data): two calls in a row, the second with fewer
frequencies and sites, plus the rejection of a batch beyond the array."""
import numpy as np
import pytest

import mtsp2
from conftest import small_problem

pytestmark = pytest.mark.gpu


def _run(ctx, fn, dt, om, e_ptr, d, ex, c_ptr, n):
    getattr(ctx, fn)(dt, om, e_ptr, d, c_ptr, exist=ex)
    return ctx.download(c_ptr, np.zeros(n, complex))


def test_lmultT_noalloc_matches_lmultT(gpu_ctx):
    print("case lmultT_noalloc vs lmultT", flush=True)
    dx, dy, dz, nza, lc = small_problem(61)
    ctx = gpu_ctx
    ctx.assemble(dx, dy, dz, nza, lc)
    ne = ctx.n_edge
    rng = np.random.default_rng(62)
    for call, (nf, ns) in enumerate(((4, 9), (2, 5))):
        x = np.column_stack([rng.uniform(0.3, 0.7, ns) * dx.sum(), rng.uniform(0.3, 0.7, ns) * dy.sum(),
                             np.zeros(ns)])
        ctx.receivers_set(dx, dy, dz, nza, x)
        om = 2 * np.pi * np.geomspace(0.01, 10.0, nf)
        e = rng.standard_normal((nf, 2, ne)) + 1j * rng.standard_normal((nf, 2, ne))
        d = rng.standard_normal((nf, ns, 8))
        ex = (rng.uniform(size=(nf, ns, 8)) > 0.2).astype(np.uint8)
        pe = ctx.dev_alloc(e.nbytes)
        pc = ctx.dev_alloc(nf * 2 * ne * 16)
        try:
            ctx.upload(pe, np.ascontiguousarray(e))
            want = _run(ctx, "lmultT", 1, om, pe, d, ex, pc, nf * 2 * ne)
            # stale values in the output buffer must be cleared by the call
            ctx.upload(pc, np.full(nf * 2 * ne, 7.0 + 3.0j))
            got = _run(ctx, "lmultT_noalloc", 1, om, pe, d, ex, pc, nf * 2 * ne)
        finally:
            ctx.dev_free(pe)
            ctx.dev_free(pc)
        print(f"call {call}: {nf} frequencies x {ns} sites, max |comb| {np.abs(want).max():.3e}",
              flush=True)
        assert np.abs(want).max() > 0
        assert np.array_equal(got, want), call


def test_lmultT_noalloc_rejects_oversize(gpu_ctx):
    print("case lmultT_noalloc oversize", flush=True)
    dx, dy, dz, nza, lc = small_problem(63)
    ctx = gpu_ctx
    ctx.assemble(dx, dy, dz, nza, lc)
    ns = 200
    x = np.column_stack([np.linspace(0.3, 0.7, ns) * dx.sum(), np.full(ns, 0.5 * dy.sum()), np.zeros(ns)])
    ctx.receivers_set(dx, dy, dz, nza, x)
    nf = 100                                      # 100 x 200 > 16384: rejected before any launch
    p = ctx.dev_alloc(16)
    try:
        with pytest.raises(mtsp2.MTError) as ei:
            ctx.lmultT_noalloc(1, 2 * np.pi * np.ones(nf), p, np.zeros((nf, ns, 8)), p)
        assert ei.value.code == mtsp2.MT_E_ARG
    finally:
        ctx.dev_free(p)
