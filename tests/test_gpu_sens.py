"""The sensitivity path on the GPU: FWDsolve3D's TRN right-hand side and
post-processing, PmultT and the gradient assembly as device kernels, and the
reference's JmultT chain run end to end on device buffers.

Checks and tolerances:
  TRN RHS / post, PmultT (device kernels)   vs the oracle (sp2o_rhs_trn,
      sp2o_post_trn, sp2o_pmultT, pinned bit for bit to the reference's
      JmultT chain, tests/test_oracle_jmultT.py): rel <= 1e-13
  gradient kernels (both modes)             vs the host restatement
      mt_host_gradient: rel max-norm <= 1e-12
  JmultT chain on the GPU (BC_x0_WS, forward solves, LmultT, TRN solves with
      the reference's GD-corrected RHS, post-processing, PmultT)   vs the
      reference's own chain (fixture tests/golden/ref_jmultT.npz, written by
      the flang build of SensComp.f90 JmultT_TX): every solve to 1e-10, so
      fields rel L2 <= 1e-7 and the cell gradient rel max-norm <= 1e-6
  exact gradient at 80x80x50+10, 12 frequencies, 144 sites (forward and A^T
      solves batched on the GPU, mixed precision to 3e-14) vs central
      differences of the misfit (step 1e-4 in ln sigma) on 20 seeded cells
      among those whose gradient is at least 5 % of the largest: rel <= 1e-3
"""
import os

import numpy as np
import pytest

import mtsp2
import sp2oracle
from conftest import rel, small_problem
from mtsp2 import problem as P

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def _cvec(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def _up(ctx, arr):
    arr = np.ascontiguousarray(arr)
    p = ctx.dev_alloc(max(arr.nbytes, 16))
    ctx.upload(p, arr)
    return p


def test_trn_rhs_post_pmult_vs_oracle(gpu_ctx):
    print("case TRN rhs/post + PmultT kernels vs oracle", flush=True)
    dx, dy, dz, nza, lc = small_problem(21, nx=11, ny=10, nzAir=4, nzEarth=8)
    o = sp2oracle.Problem(dx, dy, dz, nza, lc)
    ctx = gpu_ctx
    ctx.assemble(dx, dy, dz, nza, lc)
    rng = np.random.default_rng(4)
    om = 2 * np.pi * np.array([1e-2, 1.0, 40.0])
    S = _cvec(rng, len(om), o.nE)
    # host-buffer forms (staged through the device kernels)
    for k, w in enumerate(om):
        assert rel(ctx.rhs_trn(w, S[k]), o.rhs_trn(w, S[k])) <= 1e-13
    ei = _cvec(rng, o.nInt)
    assert rel(ctx.post_trn(ei), o.post_trn(ei)) <= 1e-14
    # device batch forms
    ps, pb, pe = _up(ctx, S), ctx.dev_alloc(len(om) * o.nInt * 16), ctx.dev_alloc(len(om) * o.nE * 16)
    try:
        ctx.rhs_trn_dev(om, ps, pb, gd_correct=True)
        B = ctx.download(pb, np.zeros((len(om), o.nInt), complex))
        for k, w in enumerate(om):
            assert rel(B[k], o.rhs_trn(w, S[k])) <= 1e-13, k
        ctx.rhs_trn_dev(om, ps, pb, gd_correct=False)
        B = ctx.download(pb, np.zeros((len(om), o.nInt), complex))
        assert np.array_equal(B, S[:, o.index()[0]])
        ctx.upload(pb, np.ascontiguousarray(_cvec(rng, len(om), o.nInt)))
        Ei = ctx.download(pb, np.zeros((len(om), o.nInt), complex))
        ctx.post_trn_dev(len(om), pb, pe)
        Eo = ctx.download(pe, np.zeros((len(om), o.nE), complex))
        for k in range(len(om)):
            assert rel(Eo[k], o.post_trn(Ei[k])) <= 1e-14, k
        # forward post-processing: e(EDGEi) = x on top of the boundary data
        base = _cvec(rng, len(om), o.nE)
        ctx.upload(pe, np.ascontiguousarray(base))
        ctx.post_fwd_dev(len(om), pb, pe)
        got = ctx.download(pe, np.zeros((len(om), o.nE), complex))
        want = base.copy()
        want[:, o.index()[0]] = Ei
        assert np.array_equal(got, want)
    finally:
        for p in (ps, pb, pe):
            ctx.dev_free(p)
    # PmultT: two polarisations, three frequencies summed on the device
    E0 = _cvec(rng, len(om), 2, o.nE)
    L = _cvec(rng, len(om), 2, o.nE)
    want = sum(o.pmultT(w, E0[k], L[k]) for k, w in enumerate(om))
    assert rel(ctx.pmultT(om[0], E0[0], L[0], lc.shape), o.pmultT(om[0], E0[0], L[0])) <= 1e-13
    pe0, pl, pd = _up(ctx, E0), _up(ctx, L), ctx.dev_alloc(lc.size * 8)
    try:
        ctx.pmultT_dev(om, 2, pe0, pl, pd)
        got = ctx.download(pd, np.zeros(lc.size)).reshape(lc.shape, order="F")
    finally:
        for p in (pe0, pl, pd):
            ctx.dev_free(p)
    assert rel(got, want) <= 1e-13


@pytest.mark.parametrize("mode", [mtsp2.GRAD_REFERENCE, mtsp2.GRAD_EXACT])
def test_gradient_kernels_vs_host(gpu_ctx, mode):
    print(f"case gradient kernels mode={mode}", flush=True)
    dx, dy, dz, nza, lc = small_problem(23, nx=12, ny=11, nzAir=4, nzEarth=9)
    ctx = gpu_ctx
    ctx.assemble(dx, dy, dz, nza, lc)
    rng = np.random.default_rng(8)
    om = 2 * np.pi * np.array([0.1, 5.0])
    ne, ni = ctx.n_edge, ctx.n_int
    E = _cvec(rng, len(om), 2, ne)
    L = _cvec(rng, len(om), 2, ni)
    want = sum(mtsp2.host_gradient(dx, dy, dz, nza, lc, w, E[k], L[k], mode=mode)
               for k, w in enumerate(om))
    pe, pl, pd = _up(ctx, E), _up(ctx, L), ctx.dev_alloc(lc.size * 8)
    try:
        ctx.gradient_dev(om, 2, pe, pl, pd, mode=mode)
        got = ctx.download(pd, np.zeros(lc.size)).reshape(lc.shape, order="F")
    finally:
        for p in (pe, pl, pd):
            ctx.dev_free(p)
    err = np.max(np.abs(got - want)) / np.max(np.abs(want))
    print(f"gradient kernels vs host: max rel {err:.2e}", flush=True)
    assert err <= 1e-12
    # after a device model update the kernels use the new sigma
    lc2 = lc + 0.3 * rng.standard_normal(lc.shape)
    ctx.update_cond(lc2)
    want2 = mtsp2.host_gradient(dx, dy, dz, nza, lc2, om[0], E[0], L[0], mode=mode)
    got2 = ctx.gradient(om[0], E[0], L[0], lc.shape, mode=mode)
    assert np.max(np.abs(got2 - want2)) <= 1e-12 * np.max(np.abs(want2))


def test_jmultT_chain_vs_reference(gpu_ctx):
    """The reference's JmultT_TX on the GPU, device buffers throughout."""
    print("case JmultT chain vs the reference's fixture", flush=True)
    z = np.load(os.path.join(HERE, "golden", "ref_jmultT.npz"))
    grid = {"dx": z["dx"], "dy": z["dy"], "dz": z["dz"], "nzAir": int(z["nzAir"])}
    lc = z["logcond"]
    ctx = gpu_ctx
    ctx.assemble(grid["dx"], grid["dy"], grid["dz"], grid["nzAir"], lc)
    ctx.receivers_set(grid["dx"], grid["dy"], grid["dz"], grid["nzAir"], z["sites"])
    ne, ni = ctx.n_edge, ctx.n_int
    for k, w in enumerate(z["omegas"]):
        om2 = np.array([w, w])
        pe, pb, px = ctx.dev_alloc(2 * ne * 16), ctx.dev_alloc(2 * ni * 16), ctx.dev_alloc(2 * ni * 16)
        pc, pt, pd = ctx.dev_alloc(2 * ne * 16), ctx.dev_alloc(2 * ne * 16), ctx.dev_alloc(lc.size * 8)
        try:
            ctx.bc_x0_dev(om2, np.array([1, 2], np.int32), pe)
            ctx.rhs_fwd_dev(2, pe, pb, px)
            st = ctx.solve_batch_dev(om2, pb, px, tol=1e-10, maxit=3000, precision=mtsp2.PREC_FP64,
                                     precond=mtsp2.PC_DILU)
            assert all(s["status"] == 0 for s in st), st
            ctx.post_fwd_dev(2, px, pe)
            ctx.lmultT(mtsp2.DT_FULL_IMPEDANCE, [w], pe, z["d"][k][None], pc)
            ctx.rhs_trn_dev(om2, pc, pb, gd_correct=True)
            ctx.upload(px, np.zeros((2, ni), complex))
            st = ctx.solve_batch_dev(om2, pb, px, tol=1e-10, maxit=3000, precision=mtsp2.PREC_FP64,
                                     precond=mtsp2.PC_DILU)
            assert all(s["status"] == 0 for s in st), st
            ctx.post_trn_dev(2, px, pt)
            ctx.pmultT_dev([w], 2, pe, pt, pd)
            fwd = ctx.download(pe, np.zeros((2, ne), complex))
            comb = ctx.download(pc, np.zeros((2, ne), complex))
            trn = ctx.download(pt, np.zeros((2, ne), complex))
            dsig = ctx.download(pd, np.zeros(lc.size)).reshape(lc.shape, order="F")
        finally:
            for p in (pe, pb, px, pc, pt, pd):
                ctx.dev_free(p)
        errs = {"fwd": rel(fwd, z[f"fwd_{k}"]), "comb": rel(comb, z[f"comb_{k}"]),
                "trn": rel(trn, z[f"trn_{k}"]),
                "dsig": float(np.max(np.abs(dsig - z[f"dsig_jt_{k}"])) /
                              np.max(np.abs(z[f"dsig_jt_{k}"])))}
        print(f"f={w / 2 / np.pi:.3g} Hz: " + ", ".join(f"{n} {v:.1e}" for n, v in errs.items()),
              flush=True)
        assert errs["fwd"] <= 1e-7 and errs["comb"] <= 1e-7 and errs["trn"] <= 1e-7
        assert errs["dsig"] <= 1e-6


class Misfit:
    """phi(m) = 1/2 sum |(Z(m) - d) / err|^2 over the full impedance tensor at
    the sites and frequencies, with the forward systems' boundary data held
    at those of the base model (the derivative the discrete adjoint gives)."""

    def __init__(self, ctx, grid, lc, freqs, sites, tol):
        self.ctx, self.grid, self.tol = ctx, grid, tol
        self.om = np.repeat(2 * P.PI * np.asarray(freqs), 2)
        self.nf = len(freqs)
        ns = len(self.om)
        self.ns, self.ne, self.ni = ns, ctx.n_edge, ctx.n_int
        ctx.receivers_set(grid["dx"], grid["dy"], grid["dz"], grid["nzAir"], sites)
        self.nsite = len(sites)
        self.pbc = ctx.dev_alloc(ns * self.ne * 16)
        self.pe = ctx.dev_alloc(ns * self.ne * 16)
        self.pb = ctx.dev_alloc(ns * self.ni * 16)
        self.px = ctx.dev_alloc(ns * self.ni * 16)
        self.pc = ctx.dev_alloc(ns * self.ne * 16)
        self.pl = ctx.dev_alloc(ns * self.ni * 16)
        self.pd = ctx.dev_alloc(lc.size * 8)
        ctx.bc_x0_dev(self.om, np.tile(np.array([1, 2], np.int32), self.nf), self.pbc)
        self.d = None
        self.err = None

    def close(self):
        for p in (self.pbc, self.pe, self.pb, self.px, self.pc, self.pl, self.pd):
            self.ctx.dev_free(p)

    def forward(self, lc):
        """Z(m), each solve started from the boundary fields' interior values
        (as FWDsolve3D starts)."""
        ctx = self.ctx
        ctx.update_cond(lc)
        ctx.rhs_fwd_dev(self.ns, self.pbc, self.pb, self.px)
        st = ctx.solve_batch_dev(self.om, self.pb, self.px, tol=self.tol, maxit=20000,
                                 precision=mtsp2.PREC_MIXED, precond=mtsp2.PC_DILU_RB)
        self.last = st
        assert all(s["status"] == 0 for s in st), st
        ctx.dev_copy(self.pe, self.pbc, self.ns * self.ne * 16)
        ctx.post_fwd_dev(self.ns, self.px, self.pe)
        z = ctx.data_resp(mtsp2.DT_FULL_IMPEDANCE, self.om[::2], self.pe, 8)
        return z[..., 0::2] + 1j * z[..., 1::2]          # (nf, nsite, 4) complex

    def phi(self, lc):
        r = (self.forward(lc) - self.d) / self.err
        return 0.5 * float(np.sum(np.abs(r) ** 2))

    def gradient(self, lc):
        """d phi / d m by the adjoint: lam solves A^T lam = g, g = LmultT of
        the weighted residuals, then -d/dm Re(lam^T A e)."""
        ctx = self.ctx
        Z = self.forward(lc)
        w = (Z - self.d) / self.err ** 2
        dv = np.zeros((self.nf, self.nsite, 8))
        dv[..., 0::2], dv[..., 1::2] = w.real, w.imag
        ctx.lmultT(mtsp2.DT_FULL_IMPEDANCE, self.om[::2], self.pe, dv, self.pc)
        ctx.rhs_trn_dev(self.om, self.pc, self.pb, gd_correct=False)
        ctx.upload(self.pl, np.zeros((self.ns, self.ni), complex))
        st = ctx.solve_batch_dev(self.om, self.pb, self.pl, tol=self.tol, maxit=20000,
                                 precision=mtsp2.PREC_MIXED, precond=mtsp2.PC_DILU_RB,
                                 op_transpose=1)
        assert all(s["status"] == 0 for s in st), st
        ctx.gradient_dev(self.om[::2], 2, self.pe, self.pl, self.pd, mode=mtsp2.GRAD_EXACT)
        g = ctx.download(self.pd, np.zeros(lc.size)).reshape(lc.shape, order="F")
        return -g


@pytest.mark.timeout(600)
def test_gradient_vs_central_differences_80x80x50(gpu_ctx, capsys):
    def say(msg):
        with capsys.disabled():
            print(f"[80x80x50 FD] {msg}", flush=True)

    print("case exact gradient vs central differences, 80x80x50+10, 12 f, 144 sites", flush=True)
    grid = P.config_grid("cfg4")
    lc0 = P.config_model("cfg4", grid, 0)
    ctx = gpu_ctx
    ctx.assemble(grid["dx"], grid["dy"], grid["dz"], grid["nzAir"], lc0)
    sites = P.config_sites("cfg4", grid)
    # every solve to 3e-14, near FP64's floor for this operator (about 1.1e-14):
    # the misfit's change under a step of 1e-4 in one cell's ln sigma must
    # stand above the solves' error in the fields (at 1e-12 it does not)
    M = Misfit(ctx, grid, lc0, P.config_freqs("cfg4"), sites, tol=3e-14)
    try:
        rng = np.random.default_rng(4)
        # synthetic data: a perturbed "true" model, 5 % Gaussian noise on the
        # full impedance tensor; errors 5 % of sqrt|Zxy Zyx| per site
        true = lc0.copy()
        true[30:50, 30:50, 3:10] += np.log(3.0)
        Zt = M.forward(true)
        floor = 0.05 * np.sqrt(np.abs(Zt[..., 1] * Zt[..., 2]))[..., None]
        M.d = Zt + floor * (rng.standard_normal(Zt.shape) + 1j * rng.standard_normal(Zt.shape)) / np.sqrt(2)
        M.err = np.broadcast_to(floor, Zt.shape)
        g = M.gradient(lc0)
        say(f"adjoint gradient done, max |g| {np.max(np.abs(g)):.3e}; forward iters "
            f"{[(q['niter'], q['nouter']) for q in M.last]}, relres {max(q['relres'] for q in M.last):.1e}")
        say(f"phi(m0) = {M.phi(lc0):.10e}")
        # 20 seeded cells among those the data resolve: gradient at least 5 %
        # of the largest (a step of 1e-4 in ln sigma of a cell the data do not
        # see changes the misfit by less than the solves' rounding)
        cand = np.argwhere(np.abs(g) >= 0.05 * np.max(np.abs(g)))
        say(f"{len(cand)} cells with |g| >= 5 % of the largest")
        pick = rng.choice(len(cand), size=min(20, len(cand)), replace=False)
        cells = [tuple(int(v) for v in cand[q]) for q in pick]
        h = 1e-4
        rows = []
        for c in cells:
            lp, lm = lc0.copy(), lc0.copy()
            lp[c] += h
            lm[c] -= h
            fd = (M.phi(lp) - M.phi(lm)) / (2 * h)
            rows.append((c, fd, g[c]))
            say(f"cell {c}: FD {fd:+.6e}  adjoint {g[c]:+.6e}  rel {abs(fd - g[c]) / abs(fd):.1e}")
    finally:
        M.close()
    for c, fd, ga in rows:
        assert abs(fd - ga) <= 1e-3 * abs(fd), (c, fd, ga)
