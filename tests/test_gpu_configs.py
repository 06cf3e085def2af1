"""Parity at the benchmark workloads' own sizes: GPU forward solves against
the oracle (oracle/sp2_oracle.c, bit-exact with the reference's CPU SP2 path)
on the same grids, models, frequencies and boundary data.

  40x40x30+10, one 10 ohm-m block in 100 ohm-m, 1 Hz, modes 1 and 2
      the FP64 drop-in (mt_bicg: cuBiCG semantics, block-DILU in the
      reference ordering) and the same with the Jacobi preconditioner, and the
      mixed batched solver; impedances at 10 x 10 core sites.
  100x100x60+10, 200 seeded boxes, the lowest and the highest of its 8
      frequencies (1e-3 and 1e2 Hz), modes 1 and 2
      the FP64 drop-in and the mixed batched solver (FP32 inner BiCGStab with
      the red-black tile DILU, FP64 defect correction); impedances at 20 x 20
      = 400 core sites; a per-frequency report of the E and Z errors and the
      inner/outer iteration counts.

Every solve, GPU and oracle, runs to a relative residual of 1e-10, the
reference's default (tolEMDef, EMsolve3D.f90:66): two solutions each stopped
at 1e-7 differ by more than the E and Z bounds below (the reference's own
1e-7 and 1e-10 solutions differ by up to 3.3e-5 in E), so the bounds are
asserted between 1e-10 solutions.  Bounds, per (frequency, mode) system:
  true relative residual of the GPU solution (FP64 operator)   <= 1e-10
  E: relative L2 difference of the all-edge fields             <= 1e-5
  Z: per-site relative difference of the full impedance tensor
     (dataResp type 1 at every site, both modes)               <= 1e-4
The oracle's responses are the oracle's own dataResp (sp2o_data_resp).
The GPU solutions at the benchmark tolerance 1e-7 are compared too and their
errors printed, not asserted.
"""
import numpy as np
import pytest

import mtsp2
import sp2oracle
from cfg_oracle import OracleFarm
from mtsp2 import problem as P

pytestmark = pytest.mark.gpu

E_TOL = 1e-5
Z_TOL = 1e-4
TOL = 1e-10


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _z_err(zg, zo):
    """Per-site relative difference of the 2x2 impedance tensors (8 reals)."""
    return np.linalg.norm(zg - zo, axis=-1) / np.linalg.norm(zo, axis=-1)


class GpuForward:
    """Boundary fields, right-hand sides and initial guesses of the systems on
    the device (BC_x0_WS and Mult_Aib, as the reference's fwdSetup and
    FWDsolve3D do), downloaded once for the per-system solves."""

    def __init__(self, ctx, cfg, seed, systems):
        self.ctx = ctx
        self.grid = P.config_grid(cfg)
        g = self.grid
        self.lc = P.config_model(cfg, g, seed)
        ctx.assemble(g["dx"], g["dy"], g["dz"], g["nzAir"], self.lc)
        self.om = np.array([2 * P.PI * f for f, _ in systems])
        self.modes = np.array([m for _, m in systems], np.int32)
        ns, n, ne = len(systems), ctx.n_int, ctx.n_edge
        self.ns, self.n, self.ne = ns, n, ne
        self.ei = P.interior_edges(len(g["dx"]), len(g["dy"]), len(g["dz"]))
        pe = ctx.dev_alloc(ns * ne * 16)
        pb, px = ctx.dev_alloc(ns * n * 16), ctx.dev_alloc(ns * n * 16)
        try:
            ctx.bc_x0_dev(self.om, self.modes, pe)
            ctx.rhs_fwd_dev(ns, pe, pb, px)
            self.E_bc = ctx.download(pe, np.zeros((ns, ne), np.complex128))
            self.B = ctx.download(pb, np.zeros((ns, n), np.complex128))
            self.X0 = ctx.download(px, np.zeros((ns, n), np.complex128))
        finally:
            for q in (pe, pb, px):
                ctx.dev_free(q)

    def fields(self, X):
        E = self.E_bc.copy()
        E[:, self.ei] = X
        return E

    def relres(self, X):
        return [float(np.linalg.norm(self.B[s] - self.ctx.apply(self.om[s], X[s])) /
                      np.linalg.norm(self.B[s])) for s in range(self.ns)]

    def dropin(self, tol, precond=mtsp2.PC_DILU):
        X, its = [], []
        for s in range(self.ns):
            x, nit, _, failed = self.ctx.bicg(self.om[s], self.B[s], self.X0[s], maxit=4000,
                                              tol=tol, precond=precond)
            assert not failed, (s, nit)
            X.append(x)
            its.append(nit)
        return np.array(X), its

    def mixed(self, tol):
        X, st = self.ctx.solve_batch(self.om, self.B, self.X0, tol=tol, maxit=8000,
                                     precision=mtsp2.PREC_MIXED, precond=mtsp2.PC_DILU_RB)
        assert all(s["status"] == 0 for s in st), st
        return X, st

    def impedances(self, E, sites):
        """dataResp type 1 at the sites, fields grouped [frequency][mode 1, 2]."""
        g, ctx = self.grid, self.ctx
        ctx.receivers_set(g["dx"], g["dy"], g["dz"], g["nzAir"], sites)
        nf = self.ns // 2
        ptr = ctx.dev_alloc(E.nbytes)
        try:
            ctx.upload(ptr, np.ascontiguousarray(E.reshape(nf, 2, self.ne)))
            return ctx.data_resp(mtsp2.DT_FULL_IMPEDANCE, self.om[::2], ptr, 8)
        finally:
            ctx.dev_free(ptr)


def _oracle_z(grid, lc, om, E_or, sites):
    o = sp2oracle.Problem(grid["dx"], grid["dy"], grid["dz"], grid["nzAir"], lc)
    return np.array([o.data_resp(1, om[2 * f], np.zeros(3), sites, E_or[2 * f], E_or[2 * f + 1])
                     for f in range(len(om) // 2)])


def _run_workload(ctx, capsys, cfg, systems, extra_precond=()):
    def say(msg):
        with capsys.disabled():
            print(f"[{cfg}] {msg}", flush=True)

    farm = OracleFarm(cfg, 0, systems)
    say(f"oracle started on {len(systems)} systems; GPU solves")
    G = GpuForward(ctx, cfg, 0, systems)
    sites = P.config_sites(cfg, G.grid)
    runs = {}
    X, its = G.dropin(TOL)
    runs["fp64 drop-in"] = (X, [f"{k}" for k in its])
    for pc in extra_precond:
        X, its = G.dropin(TOL, precond=pc)
        runs[f"fp64 drop-in pc={pc}"] = (X, [f"{k}" for k in its])
    X, st = G.mixed(TOL)
    runs["mixed"] = (X, [f"{s['niter']}/{s['nouter']}" for s in st])
    X7, st7 = G.mixed(1e-7)
    say("GPU solves done; waiting for the oracle")
    res = farm.results(progress=say)
    E_or = np.array([r["e_all"] for r in res])
    for s, r in enumerate(res):
        assert r["status"] == 0 and r["relres"] <= TOL * 1.0001, (s, r["status"], r["relres"])
    Z_or = _oracle_z(G.grid, G.lc, G.om, E_or, sites)
    report = []
    for name, (X, its) in list(runs.items()) + [("mixed @1e-7", (X7, [
            f"{s['niter']}/{s['nouter']}" for s in st7]))]:
        rr = G.relres(X)
        E = G.fields(X)
        Z = G.impedances(E, sites)
        ez = _z_err(Z, Z_or)                       # (nf, nsite)
        for s in range(G.ns):
            e_err = _rel(E[s], E_or[s])
            z_err = float(ez[s // 2].max())
            report.append((name, systems[s][0], systems[s][1], its[s], res[s]["niter"],
                           rr[s], e_err, z_err))
    say("system                 f [Hz]  mode  iters(gpu)  iters(oracle)  relres    E rel L2  Z max rel")
    for name, f, m, it, ito, rr, e_err, z_err in report:
        say(f"{name:22s} {f:8.3g}  {m:4d}  {it:>10s}  {ito:13d}  {rr:8.1e}  {e_err:8.1e}  {z_err:8.1e}")
    for name, f, m, it, ito, rr, e_err, z_err in report:
        if name == "mixed @1e-7":
            assert rr <= 1e-7 * 1.0001
            continue
        assert rr <= TOL * 1.0001, (name, f, m, rr)
        assert e_err <= E_TOL, (name, f, m, e_err)
        assert z_err <= Z_TOL, (name, f, m, z_err)


@pytest.mark.timeout(300)
def test_workload_40x40x30(gpu_ctx, capsys):
    print("case 40x40x30+10 parity vs oracle", flush=True)
    systems = [(1.0, 1), (1.0, 2)]
    _run_workload(gpu_ctx, capsys, "cfg0", systems, extra_precond=(mtsp2.PC_JACOBI,))


@pytest.mark.timeout(900)
def test_workload_100x100x60(gpu_ctx, capsys):
    print("case 100x100x60+10 parity vs oracle (lowest and highest frequency)", flush=True)
    f = P.config_freqs("cfg1")
    systems = [(f[0], 1), (f[0], 2), (f[-1], 1), (f[-1], 2)]
    _run_workload(gpu_ctx, capsys, "cfg1", systems)
