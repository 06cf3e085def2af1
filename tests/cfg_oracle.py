"""Oracle side of the workload-size parity tests (test infrastructure).

Each forward system of a workload is solved by the oracle (oracle/sp2_oracle.c,
bit-exact with the reference's CPU SP2 path) in a process of its own, so the
systems of a workload run in parallel on the host cores while the GPU side
runs: boundary fields and initial guess from BC_x0_WS (sp2o_bc_x0, as fwdSetup
-> ComputeBC does), b = -CCib e_b, then the cuBiCG-semantics BiCGStab with the
reference-ordered block-DILU to the reference's default tolerance 1e-10
(tolEMDef, EMsolve3D.f90:66), and the all-edge field FWDsolve3D returns.
"""
from __future__ import annotations

import concurrent.futures as cf
import multiprocessing as mp
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def _paths():
    for p in (os.path.join(ROOT, "modem-gpu_amd"), os.path.join(ROOT, "oracle"), HERE):
        if p not in sys.path:
            sys.path.insert(0, p)


def solve_system(cfg, seed, f_hz, mode, tol=1e-10, maxit=4000):
    """One forward system of workload `cfg` on the oracle.  Returns a dict:
    e_all (all-edge field: the BC_x0 boundary data with the solution inside),
    x (interior solution), niter, status, relres (oracle true residual),
    seconds."""
    _paths()
    import sp2oracle
    from mtsp2 import problem as P
    t0 = time.time()
    grid = P.config_grid(cfg)
    lc = P.config_model(cfg, grid, seed)
    o = sp2oracle.Problem(grid["dx"], grid["dy"], grid["dz"], grid["nzAir"], lc)
    w = 2 * P.PI * f_hz
    e0, _ = sp2oracle.bc_x0(grid["dx"], grid["dy"], grid["dz"], grid["nzAir"], lc, w, mode)
    e = sp2oracle.bc_form(e0, o.nx, o.ny, o.nz)
    ei = o.index()[0]
    o.update_freq(w)
    b = o.rhs_fwd(e)
    x, info = o.bicg(b, e[ei], maxit=maxit, tol=tol, gpu_semantics=True, precond=0)
    r = b - o.mult_aii(x)
    e[ei] = x
    return {"e_all": e, "niter": info["niter"], "status": info["status"],
            "relres": float(np.linalg.norm(r) / np.linalg.norm(b)), "seconds": time.time() - t0}


class OracleFarm:
    """The systems [(f_hz, mode), ...] of a workload submitted to a pool of
    spawned processes (the oracle is single-threaded)."""

    def __init__(self, cfg, seed, systems, workers=None):
        ctx = mp.get_context("spawn")
        n = workers or min(len(systems), 8)
        self.pool = cf.ProcessPoolExecutor(max_workers=n, mp_context=ctx)
        self.systems = list(systems)
        self.futs = [self.pool.submit(solve_system, cfg, seed, f, m) for f, m in self.systems]

    def results(self, progress=None, poll=20.0):
        """Waits for every system; calls progress(msg) every `poll` seconds."""
        t0 = time.time()
        while True:
            done, pending = cf.wait(self.futs, timeout=poll)
            if not pending:
                break
            if progress:
                progress(f"oracle: {len(done)}/{len(self.futs)} systems done after "
                         f"{time.time() - t0:.0f} s")
        out = [f.result() for f in self.futs]
        self.pool.shutdown()
        return out


def jmult_chain(grid, lc, omega, sites, d, tol=1e-10, maxit=2000):
    """The reference's JmultT_TX chain for one MT transmitter (full-impedance
    data d (nsite, 8)) restated with the oracle, step by step as
    SENS/SensComp.f90:408-485 runs it with the CPU solver: forward solves of
    modes 1, 2 from the BC_x0_WS boundary data (fwdSetup/fwdSolve ->
    FWDsolve3D, one BICG call), LmultT, the TRN solves (sensSolve ->
    FWDsolve3D TRN branch: RHS, BICG from zero, post-processing) and PmultT.
    Returns dict(fwd, comb, trn (2, nEdge), dsig (nx, ny, nzEarth))."""
    _paths()
    import sp2oracle
    o = sp2oracle.Problem(grid["dx"], grid["dy"], grid["dz"], grid["nzAir"], lc)
    ei = o.index()[0]
    o.update_freq(omega)
    fwd = []
    for mode in (1, 2):
        e0, _ = sp2oracle.bc_x0(grid["dx"], grid["dy"], grid["dz"], grid["nzAir"], lc, omega, mode)
        e = sp2oracle.bc_form(e0, o.nx, o.ny, o.nz)
        x, _ = o.bicg(o.rhs_fwd(e), e[ei], maxit=maxit, tol=tol, gpu_semantics=False)
        e[ei] = x
        fwd.append(e)
    comb = o.lmultT(1, omega, np.zeros(3), sites, d, fwd[0], fwd[1])
    trn = []
    for c in comb:
        x, _ = o.bicg(o.rhs_trn(omega, c), np.zeros(o.nInt, complex), maxit=maxit, tol=tol,
                      gpu_semantics=False)
        trn.append(o.post_trn(x))
    dsig = o.pmultT(omega, np.array(fwd), np.array(trn))
    return {"fwd": np.array(fwd), "comb": np.array(comb), "trn": np.array(trn), "dsig": dsig}
