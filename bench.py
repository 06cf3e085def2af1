"""Benchmark of the SP2 3D-MT forward + adjoint solve on one MI355X.

Default workload (the config the headline metric is quoted on): 100x100x60
cells + 10 air layers (2,052,270 interior edges), 200 seeded blocky boxes over
100 ohm-m, 8 frequencies 1e-3..1e2 Hz x 2 polarizations, impedances at 20 x 20
= 400 core sites.  A step is one pass of the forward + adjoint hot path:
  * the 16 forward systems solved together by the mixed-precision batched
    solver (FP32 BiCGStab inner iterations with the red-black tile block-DILU,
    FP64 defect correction, DILU pivots computed on the device inside the
    step), right-hand sides and initial guesses (BC_x0_WS + Mult_Aib) already
    resident in HBM;
  * the forward fields, the impedances at the sites, the adjoint sources of
    the weighted data residuals (LmultT), and the 16 adjoint systems
    A^T lam = s solved together by the same mixed batched solver;
  * the gradient of the misfit assembled on the device.
value = (16 forward + 16 adjoint solves) / step time, every system at FP64
true relative residual <= 1e-7.  The timed steps run with HIP events around
every SpMV launch and preconditioner application (the roofline figures come
from them).  Beside it: the same steps replayed as hipGraphs, the same 32
solves by the plain FP64 GPU solver (one system at a time, block-DILU in the
reference ordering: the drop-in cuBiCG semantics) and by the batched FP64
solver in both orderings (what batching, the red-black ordering and FP32
each contribute), the end-to-end rate with the boundary fields included, the
isolated kernels, and the reference CPU path timed on the host's cores.

--config cfg3 runs the 120x120x80 + 10 air workload: 16 frequencies
1e-4..1e3 Hz x 2 polarizations = 32 forward systems in one batched solve,
against the same 32 solves one system at a time (the unbatched loop).
"""
import argparse
import concurrent.futures as cf
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "modem-gpu_amd"))

PEAK_HBM_GBS = 8000.0
WORKER_CAP = 16          # host worker processes at most (the GPU box's CPU share)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(grid, lc, om, B, X0, iters, cfg_label, sample_iters=12):
    """The reference CPU SP2 path on the host's cores, run as the reference
    runs its MPI task farm: one process per (frequency, polarization) system,
    as many at once as there are cores (at most WORKER_CAP), each timing
    `sample_iters` iterations of the reference's CPU BICG (flang build in
    oracle/_ref; the oracle's C restatement, single process, where that build
    is absent).  The per-iteration time under full load, times the FP64
    iteration counts of the same solves, gives the farm's solves/s."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    nsolve = len(iters)
    try:
        import refdriver
        have_ref = refdriver.available()
    except ImportError:
        have_ref = False
    if have_ref:
        P = max(1, min(WORKER_CAP, os.cpu_count() or 1, len(om)))

        def one(s):
            out = refdriver.run(grid, lc, np.log(1e-10), [om[s]],
                                solves=[(0, sample_iters, 1e-30, B[s], X0[s])], dump=False,
                                timeout=1200)
            return out["bicg_0"]["seconds"] / sample_iters

        t0 = time.time()
        with cf.ThreadPoolExecutor(max_workers=P) as ex:
            per_it = list(ex.map(one, range(P)))
        wall = time.time() - t0
        kind, cores = "reference", P
    else:
        import sp2oracle
        t0 = time.time()
        p = sp2oracle.Problem(grid["dx"], grid["dy"], grid["dz"], grid["nzAir"], lc)
        p.update_freq(om[0])
        t1 = time.time()
        p.bicg(B[0], X0[0], maxit=sample_iters, tol=1e-30, gpu_semantics=False)
        per_it = [(time.time() - t1) / sample_iters]
        wall = time.time() - t0
        kind, cores = "port", 1
    t_it = float(np.mean(per_it))
    total_iters = float(np.sum(iters))
    value = nsolve * cores / (total_iters * t_it)
    bc_s = None
    try:
        import refdriver
        if refdriver.bcx0_available():
            t0 = time.time()
            refdriver.run_bcx0(grid, lc, np.log(1e-10), [(om[0], 1)], timeout=600)
            bc_s = time.time() - t0
    except (ImportError, OSError):
        bc_s = None
    return {"value": value, "unit": "solves/s", "cores": cores, "kind": kind,
            "cpu_model": cpu_model(), "farm_wall_s": wall,
            "bc_x0_reference_s_per_system": bc_s,
            "bc_x0_reference_note": ("wall time of the reference's ComputeBC -> BC_x0_WS for one "
                                     "(frequency, mode) of the workload, including the driver's "
                                     "grid/operator set-up"),
            "sample": (f"{cores} concurrent processes of the reference's CPU BICG, one per system "
                       f"(frequency, polarization) of the {cfg_label} workload, {sample_iters} "
                       f"iterations each: {t_it:.3f} s/iteration per process under full load; "
                       f"the {nsolve} solves of a step need {total_iters:.0f} iterations (the FP64 "
                       f"GPU solver's counts), so {cores} cores give {value:.4f} solves/s; "
                       f"farm wall {wall:.1f} s incl. each process's operator set-up")}


TRAFFIC_SUMMARY = os.path.join("profiles", "mixed_fwdadj_100x100x60_step_hbm_counters.md")


def counted_traffic(fam, args):
    """HBM bytes per launch of kernel family `fam` over the launches of one
    step, as counted by two rocprofv3 --pmc passes (FETCH_SIZE, WRITE_SIZE,
    each scaled by its reading for a 1 GiB device copy) of this benchmark.
    The counter summary holds one line per family,
        traffic_bytes_per_launch <family> <bytes> <arg>=<value> ...
    with the arguments of the profiled run; the figure applies only to a run
    with the same ones (they fix the launch sequence), else None."""
    path = os.path.join(ROOT, TRAFFIC_SUMMARY)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            w = line.split()
            if len(w) < 3 or w[0] != "traffic_bytes_per_launch" or w[1] != fam:
                continue
            want = dict(kv.split("=", 1) for kv in w[3:])
            have = {k: str(getattr(args, k)) for k in want}
            return float(w[2]) if have == want else None
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="cfg1", choices=("cfg1", "cfg3", "cfg4"))
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--tol", type=float, default=1e-7)
    ap.add_argument("--inner-tol", type=float, default=3e-4,
                    help="least residual reduction asked of each FP32 inner solve")
    ap.add_argument("--restart", type=int, default=120,
                    help="BiCGStab restart interval of the inner solves (the reference's 120)")
    ap.add_argument("--forward-only", action="store_true",
                    help="time the forward systems alone (the default adds the adjoint systems "
                         "and the gradient, except for cfg3, whose workload is forward)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--precond", type=int, default=3,
                    help="preconditioner of the mixed solver's FP32 inner solves "
                         "(0 block-DILU in the reference ordering, 3 red-black tile block-DILU)")
    ap.add_argument("--kernels-only", action="store_true",
                    help="only the isolated SpMV launches (FP32 x nsys, FP64 x 1) and a 1 GiB "
                         "device copy, for hardware-counter runs")
    ap.add_argument("--spmv-variant", type=int, default=None,
                    help="SpMV kernel of the solves: 0 row, 1 k-marching (FP32), 3 multi-system row")
    ap.add_argument("--spmv-systems", type=int, default=4,
                    help="systems per block of the multi-system row SpMV (4 or 8)")
    ap.add_argument("--dilu-variant", type=int, default=6, choices=(1, 4, 5, 6),
                    help="red-black DILU launch shape: 1 single-system persistent, 4 four systems "
                         "per workgroup, 5 the same with the lean LDS layout, 6 (inner solves on "
                         "bricks) persistent four-system workgroups loading the next item while "
                         "they sweep")
    ap.add_argument("--drop-in-only", type=int, default=0, metavar="NSYS",
                    help="only the plain FP64 solves (one system at a time, reference-ordered "
                         "DILU: the cuBiCG drop-in) of the first NSYS forward systems, for profiling")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch kernels eagerly instead of replaying hipGraphs")
    ap.add_argument("--inner-layout", default="bricks", choices=("bricks", "bricks-unpaired", "all-edge"),
                    help="inner FP32 vectors of the mixed solves: 8x8x8 tiles contiguous with "
                         "p-hat and s-hat interleaved by system pair (bricks), the same with "
                         "every vector system-major (bricks-unpaired), or the all-edge layout "
                         "of the outer solve")
    ap.add_argument("--brick-systems", type=int, default=0,
                    help="systems per workgroup of the brick SpMV (0: all)")
    ap.add_argument("--autotune", action="store_true",
                    help="time and pick the brick SpMV's systems per workgroup and the vector "
                         "updates' row chunks before the run (the default keeps fixed choices, "
                         "so that every run takes the same iterations)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the systems would shard across GPUs by frequency, with no exchange in the solve)",
              file=sys.stderr)
        sys.exit(2)

    import mtsp2
    from mtsp2 import problem as P

    if mtsp2.device_count() < 1:
        raise RuntimeError("no HIP device visible")
    cfg = args.config
    grid = P.config_grid(cfg)
    lc = P.config_model(cfg, grid, args.seed)
    fhz = P.config_freqs(cfg)
    nx, ny, nz = len(grid["dx"]), len(grid["dy"]), len(grid["dz"])
    adjoint = not args.forward_only and cfg != "cfg3"
    ctx = mtsp2.Context(0)
    ctx.set_dilu_variant(args.dilu_variant)
    bricks = args.inner_layout != "all-edge"
    layout = {"bricks": 2, "bricks-unpaired": 1, "all-edge": 0}[args.inner_layout]
    ctx.set_inner_layout(layout, args.brick_systems)
    t0 = time.time()
    ctx.assemble(grid["dx"], grid["dy"], grid["dz"], grid["nzAir"], lc)
    log(f"assembled n_int={ctx.n_int} nnz={ctx.nnz} in {time.time()-t0:.1f}s")
    label = f"{nx}x{ny}x{nz - grid['nzAir']}+{grid['nzAir']}"

    # the reference's start of a forward solve, on the device: fwdSetup ->
    # ComputeBC -> BC_x0_WS (modes 1, 2 per frequency), then FWDsolve3D's
    # b = -CCib e_b and x0 = e(EDGEi)
    om = np.repeat(2 * P.PI * np.asarray(fhz, np.float64), 2)
    modes = np.tile(np.array([1, 2], np.int32), len(fhz))
    nsys, n, ne = len(om), ctx.n_int, ctx.n_edge
    zb = nsys * n * 16
    d_b, d_x0, d_x = ctx.dev_alloc(zb), ctx.dev_alloc(zb), ctx.dev_alloc(zb)
    d_e = ctx.dev_alloc(nsys * ne * 16)
    ctx.bc_x0_dev(om, modes, d_e)                   # sizes the scratch
    t0 = time.time()
    ctx.bc_x0_dev(om, modes, d_e)
    t_bc = time.time() - t0
    t0 = time.time()
    ctx.rhs_fwd_dev(nsys, d_e, d_b, d_x0)
    t_rhs = time.time() - t0
    bc_info = {"bc_x0_ms": 1e3 * t_bc, "rhs_fwd_ms": 1e3 * t_rhs, "systems": nsys}
    log(f"boundary fields of {nsys} systems on the device: BC_x0_WS {1e3 * t_bc:.1f} ms, "
        f"Mult_Aib {1e3 * t_rhs:.1f} ms")

    if args.kernels_only:
        for prec, ns in ((mtsp2.PREC_MIXED, nsys), (mtsp2.PREC_FP64, 1)):
            kms, kby = ctx.bench_spmv(ns, prec, reps=5)
            log(f"SpMV {'FP32' if prec == mtsp2.PREC_MIXED else 'FP64'} x {ns}: {kms * 1e3:.1f} us, "
                f"{kby / 1e6:.1f} MB algorithmic, {kby / (kms * 1e-3) / 1e9:.0f} GB/s")
        for pc in (mtsp2.PC_DILU_RB, mtsp2.PC_DILU):
            kms, kby = ctx.bench_precond(nsys, mtsp2.PREC_MIXED, pc, reps=5)
            log(f"DILU {'red-black' if pc == mtsp2.PC_DILU_RB else 'reference order'} FP32 x {nsys}: "
                f"{kms * 1e3:.1f} us, {kby / 1e6:.1f} MB algorithmic, "
                f"{kby / (kms * 1e-3) / 1e9:.0f} GB/s")
        gib = 1 << 30
        src, dst = ctx.dev_alloc(gib), ctx.dev_alloc(gib)
        for _ in range(3):
            ctx.dev_copy(dst, src, gib)
        log("device copy of 1 GiB x 3 done")
        ctx.close()
        return

    if args.drop_in_only:
        use_graph = 0 if args.no_graph else 1
        ns = min(args.drop_in_only, nsys)
        ctx.timing_enable(True)
        t0 = time.time()
        its = []
        for s in range(ns):
            xs = d_x + s * n * 16
            ctx.dev_copy(xs, d_x0 + s * n * 16, n * 16)
            r = ctx.solve_batch_dev(om[s:s + 1], d_b + s * n * 16, xs, tol=args.tol, maxit=4000,
                                    precision=mtsp2.PREC_FP64, precond=mtsp2.PC_DILU, chunk=10,
                                    use_graph=use_graph)
            its.append(r[0]["niter"])
        el = time.time() - t0
        kt = {name: ctx.kernel_timing(fam) for name, fam in
              (("spmv", mtsp2.KT_SPMV), ("precond", mtsp2.KT_PRECOND))}
        ctx.timing_enable(False)
        log(f"plain FP64 (drop-in): {ns} systems in {el:.3f} s, {ns / el:.2f} solves/s, "
            f"iters {its}, {1e3 * el / max(1, sum(its)):.3f} ms/iteration")
        for name, (nl, kms, kby) in kt.items():
            if nl:
                log(f"  {name}: {nl} launches, {kms:.1f} ms, {1e3 * kms / nl:.1f} us/launch")
        ctx.close()
        return

    tuned = None
    if args.autotune:
        tuned = ctx.autotune(nsys)
        log(f"autotune: row chunks {tuned} over the active systems")
    if args.spmv_variant is not None:
        ctx.set_spmv_variant(args.spmv_variant,
                             args.spmv_systems if args.spmv_variant == 3 else 0)
    use_graph = 0 if args.no_graph else 1
    opts = dict(tol=args.tol, maxit=4000, precision=mtsp2.PREC_MIXED, precond=args.precond,
                inner_tol=args.inner_tol, restart=args.restart, chunk=10, use_graph=use_graph)

    # adjoint side: forward fields, impedances at the sites, data with 5 %
    # noise around the background model's own responses (a seeded synthetic
    # survey), adjoint sources, adjoint solutions, gradient
    nf = len(fhz)
    if adjoint:
        sites = P.config_sites(cfg, grid)
        ctx.receivers_set(grid["dx"], grid["dy"], grid["dz"], grid["nzAir"], sites)
        d_f, d_c = ctx.dev_alloc(nsys * ne * 16), ctx.dev_alloc(nsys * ne * 16)
        d_ba, d_l, d_zero = ctx.dev_alloc(zb), ctx.dev_alloc(zb), ctx.dev_alloc(zb)
        ctx.upload(d_zero, np.zeros((nsys, n), np.complex128))
        ncell = nx * ny * (nz - grid["nzAir"])
        d_dm = ctx.dev_alloc(ncell * 8)
        ctx.dev_copy(d_x, d_x0, zb)
        ctx.solve_batch_dev(om, d_b, d_x, **opts)
        ctx.dev_copy(d_f, d_e, nsys * ne * 16)
        ctx.post_fwd_dev(nsys, d_x, d_f)
        z = ctx.data_resp(mtsp2.DT_FULL_IMPEDANCE, om[::2], d_f, 8)
        Z0 = z[..., 0::2] + 1j * z[..., 1::2]
        rng = np.random.default_rng(args.seed + 100)
        err = 0.05 * np.sqrt(np.abs(Z0[..., 1] * Z0[..., 2]))[..., None] * np.ones(Z0.shape)
        d_obs = Z0 + err * (rng.standard_normal(Z0.shape) + 1j * rng.standard_normal(Z0.shape)) / np.sqrt(2)

    def step():
        ta = time.time()
        ctx.dev_copy(d_x, d_x0, zb)
        st_f = ctx.solve_batch_dev(om, d_b, d_x, **opts)
        tb = time.time()
        st_a = []
        if adjoint:
            ctx.dev_copy(d_f, d_e, nsys * ne * 16)
            ctx.post_fwd_dev(nsys, d_x, d_f)
            zz = ctx.data_resp(mtsp2.DT_FULL_IMPEDANCE, om[::2], d_f, 8)
            w = ((zz[..., 0::2] + 1j * zz[..., 1::2]) - d_obs) / err ** 2
            dv = np.zeros((nf, len(sites), 8))
            dv[..., 0::2], dv[..., 1::2] = w.real, w.imag
            ctx.lmultT(mtsp2.DT_FULL_IMPEDANCE, om[::2], d_f, dv, d_c)
            ctx.rhs_trn_dev(om, d_c, d_ba, gd_correct=False)
            ctx.dev_copy(d_l, d_zero, zb)
            st_a = ctx.solve_batch_dev(om, d_ba, d_l, op_transpose=1, **opts)
            ctx.gradient_dev(om[::2], 2, d_f, d_l, d_dm, mode=mtsp2.GRAD_EXACT)
        tc = time.time()
        return st_f, st_a, tb - ta, tc - tb

    for w in range(args.warmup):
        st_f, st_a, _, _ = step()
        log(f"warmup {w}: forward iters {[s['niter'] for s in st_f]}, adjoint "
            f"{[s['niter'] for s in st_a]}")
    ctx.timing_enable(True)
    t_fwd = t_adj = 0.0
    t0 = time.time()
    for k in range(args.steps):
        st_f, st_a, a, b = step()
        t_fwd += a
        t_adj += b
    el = time.time() - t0
    kt = {name: ctx.kernel_timing(fam) for name, fam in
          (("spmv", mtsp2.KT_SPMV), ("precond", mtsp2.KT_PRECOND))}
    ctx.timing_enable(False)
    for name, (nl, kms, kby) in kt.items():
        if nl:
            log(f"timed region, {name}: {nl} launches, {kms:.1f} ms, {1e3 * kms / nl:.1f} us/launch, "
                f"{kby / (kms * 1e-3) / 1e9:.0f} GB/s algorithmic")
    per_step = nsys * (2 if adjoint else 1)
    ms = 1e3 * el / args.steps
    value = per_step * args.steps / el
    graph_value = None
    if use_graph:
        t0 = time.time()
        for k in range(args.steps):
            step()
        graph_value = per_step * args.steps / (time.time() - t0)
        log(f"hipGraph replay without events: {graph_value:.2f} solves/s")
    allst = list(st_f) + list(st_a)
    ok = all(s["status"] == 0 and s["relres"] <= args.tol for s in allst)
    relres = max(s["relres"] for s in allst)
    log(f"mixed+batched: {ms:.1f} ms/step ({1e3 * t_fwd / args.steps:.1f} forward, "
        f"{1e3 * t_adj / args.steps:.1f} adjoint), {value:.2f} solves/s, max relres {relres:.2e}")
    log(f"  forward inner iters {[s['niter'] for s in st_f]}, outer {[s['nouter'] for s in st_f]}")
    if adjoint:
        log(f"  adjoint inner iters {[s['niter'] for s in st_a]}, outer {[s['nouter'] for s in st_a]}")

    # the same solves by the plain FP64 GPU solver (one system at a time,
    # reference-ordered DILU: the drop-in semantics) and by the batched FP64
    # solver in both orderings
    ablation = {}
    fp64_iters = None
    if not args.no_fp64 and bricks:
        # the same mixed solves with the inner vectors in the all-edge layout
        ctx.set_inner_layout(False)
        t0 = time.time()
        step()
        ablation["mixed_batched_red_black_all_edge_layout"] = per_step / (time.time() - t0)
        ctx.set_inner_layout(layout, args.brick_systems)
        log(f"mixed, all-edge inner layout: {ablation['mixed_batched_red_black_all_edge_layout']:.2f} solves/s")
    if not args.no_fp64:
        sets = [(d_b, d_x0, 0)] + ([(d_ba, d_zero, 1)] if adjoint else [])
        t0 = time.time()
        its = []
        for bb, xx0, tr in sets:
            for s in range(nsys):
                xs = d_x + s * n * 16
                ctx.dev_copy(xs, xx0 + s * n * 16, n * 16)
                r = ctx.solve_batch_dev(om[s:s + 1], bb + s * n * 16, xs, tol=args.tol, maxit=4000,
                                        precision=mtsp2.PREC_FP64, precond=mtsp2.PC_DILU,
                                        chunk=10, use_graph=use_graph, op_transpose=tr)
                its.append(r[0]["niter"])
        plain = per_step / (time.time() - t0)
        fp64_iters = its
        log(f"plain FP64 GPU (one system at a time): {plain:.2f} solves/s, iters {its}")
        ablation["fp64_one_at_a_time_reference_order"] = plain
        for name, pc in (("fp64_batched_reference_order", mtsp2.PC_DILU),
                         ("fp64_batched_red_black", mtsp2.PC_DILU_RB)):
            t0 = time.time()
            for bb, xx0, tr in sets:
                ctx.dev_copy(d_x, xx0, zb)
                ctx.solve_batch_dev(om, bb, d_x, tol=args.tol, maxit=4000, precision=mtsp2.PREC_FP64,
                                    precond=pc, chunk=10, use_graph=use_graph, op_transpose=tr)
            ablation[name] = per_step / (time.time() - t0)
            log(f"{name}: {ablation[name]:.2f} solves/s")
        ablation["mixed_batched_red_black"] = value
        if cfg == "cfg3":
            # the unbatched loop: the same mixed solver, one system at a time
            t0 = time.time()
            ctx.timing_enable(True)
            for s in range(nsys):
                xs = d_x + s * n * 16
                ctx.dev_copy(xs, d_x0 + s * n * 16, n * 16)
                ctx.solve_batch_dev(om[s:s + 1], d_b + s * n * 16, xs, **opts)
            el1 = time.time() - t0
            k1 = [ctx.kernel_timing(f) for f in (mtsp2.KT_SPMV, mtsp2.KT_PRECOND)]
            ctx.timing_enable(False)
            ablation["mixed_unbatched_loop"] = nsys / el1
            ablation["mixed_unbatched_loop_GBps"] = {
                name: (k[2] / (k[1] * 1e-3) / 1e9 if k[1] else None)
                for name, k in zip(("spmv", "precond"), k1)}
            log(f"mixed unbatched loop: {nsys / el1:.2f} solves/s, "
                f"GB/s {ablation['mixed_unbatched_loop_GBps']}")
    fp64 = ablation.get("fp64_one_at_a_time_reference_order")

    dom = max(kt, key=lambda k: kt[k][1])
    nl, kms, kby = kt[dom]
    rb_kernel = {4: "k_dilu_rbm<float,float,RED_FWD|BLACK_FB|RED_BWD,4>",
                 5: "k_dilu_rbl<float,float,RED_FWD|BLACK_FB|RED_BWD,4> (lean LDS)",
                 6: "k_dilu_rbl<float,float,RED_FWD|BLACK_FB|RED_BWD,4> (lean LDS)",
                 1: "k_dilu_rb2<float,float,RED_FWD|BLACK_FB|RED_BWD>"}[args.dilu_variant]
    on_bricks = bricks and args.precond == 3 and args.spmv_variant in (None, 3) and args.spmv_systems == 4
    if on_bricks:
        rb_kernel = ("k_dilu_bkp<float,float,RED_FWD|BLACK_FB|RED_BWD,4> (on bricks, persistent, "
                     "next item's loads in flight during the sweeps)" if args.dilu_variant == 6 else
                     "k_dilu_bk<float,float,RED_FWD|BLACK_FB|RED_BWD,4> (on bricks)")
    pc_kernels = {0: "k_dilu_tile<float,float,8,8,8,true|false> (block-DILU level sweeps)",
                  3: rb_kernel + " (red-black tile block-DILU, 3 launches per application)"}
    spv = 3 if args.spmv_variant is None else args.spmv_variant
    spmv_kernels = {0: "k_spmv_dot<float,float,1|2> (row-parallel v = A x with its dot products)",
                    1: "k_spmv_km<float,float,1|2> (k-marching v = A x with its dot products)",
                    2: "k_spmv_km<float,float,1|2> (k-marching v = A x with its dot products)",
                    3: f"k_spmv_ms<float,float,1|2,{args.spmv_systems}> (v = A x for "
                       f"{args.spmv_systems} systems per block, with the dot products)"}
    if on_bricks:
        spmv_kernels[3] = (f"k_spmv_bt<float,float,1|2> (v = A x on bricks, "
                           f"{args.brick_systems or 'all'} systems per block, with the dot products)")
        if layout == 2:
            spmv_kernels[3] = (f"k_spmv_btp<1|2> (v = A x on bricks, x of system pairs "
                               f"interleaved: one 16-byte gather per neighbour for both, "
                               f"{args.brick_systems or 'all'} systems per block, with the dot products)")
    kernels = {"spmv": spmv_kernels[spv], "precond": pc_kernels.get(args.precond, "preconditioner")}
    achieved = kby / (kms * 1e-3) / 1e9
    traffic = counted_traffic(dom, args)
    roofline = {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": achieved / PEAK_HBM_GBS, "traffic": traffic, "kernel": kernels[dom],
                "traffic_source": (f"{TRAFFIC_SUMMARY}: mean HBM bytes per launch over the "
                                   "launches of one step, counted (FETCH_SIZE x 2.0 + WRITE_SIZE)")
                if traffic else None,
                "traffic_over_algorithmic": traffic / (kby / nl) if traffic else None,
                "launches": nl, "us_per_launch": 1e3 * kms / nl, "bytes_per_launch": kby / nl,
                "share_of_step": kms / (ms * args.steps),
                "timed_families": {k: {"launches": v[0], "ms": v[1],
                                       "GBps": v[2] / (v[1] * 1e-3) / 1e9 if v[1] else None}
                                   for k, v in kt.items()}}
    spmv_ms, spmv_bytes = ctx.bench_spmv(nsys, mtsp2.PREC_MIXED, reps=20)
    log(f"SpMV isolated (FP32, {nsys} systems): {spmv_ms*1e3:.1f} us/launch, "
        f"{spmv_bytes/1e6:.1f} MB, {spmv_bytes / (spmv_ms * 1e-3) / 1e9:.0f} GB/s")
    roofline["spmv_isolated_fp32"] = {"systems": nsys, "us_per_launch": spmv_ms * 1e3,
                                      "bytes_per_launch": spmv_bytes,
                                      "GBps": spmv_bytes / (spmv_ms * 1e-3) / 1e9}
    spmv64_ms, spmv64_bytes = ctx.bench_spmv(1, mtsp2.PREC_FP64, reps=20)
    roofline["spmv_isolated_fp64_1sys"] = {"us_per_launch": spmv64_ms * 1e3,
                                           "bytes_per_launch": spmv64_bytes,
                                           "GBps": spmv64_bytes / (spmv64_ms * 1e-3) / 1e9}
    log(f"SpMV (FP64, 1 system): {spmv64_ms*1e3:.1f} us/launch, "
        f"{spmv64_bytes/spmv64_ms/1e6:.0f} GB/s")
    pc_ms, pc_bytes = ctx.bench_precond(nsys, mtsp2.PREC_MIXED, mtsp2.PC_DILU_RB, reps=10)
    roofline["precond_isolated_fp32"] = {"systems": nsys, "us_per_apply": pc_ms * 1e3,
                                         "bytes_per_apply": pc_bytes,
                                         "GBps": pc_bytes / (pc_ms * 1e-3) / 1e9}
    log(f"red-black DILU isolated (FP32, {nsys} systems): {pc_ms * 1e3:.1f} us/application, "
        f"{pc_bytes / 1e6:.1f} MB, {pc_bytes / (pc_ms * 1e-3) / 1e9:.0f} GB/s")

    e2e = per_step / (ms * 1e-3 + t_bc + t_rhs)
    cpu = None
    if not args.no_cpu:
        B = ctx.download(d_b, np.zeros((nsys, n), np.complex128))
        X0 = ctx.download(d_x0, np.zeros((nsys, n), np.complex128))
        iters = fp64_iters or [s["niter"] for s in allst]
        cpu = cpu_baseline(grid, lc, om, B, X0, iters, label)
        log(f"cpu baseline: {cpu}")

    fwd_adj = "forward + adjoint (A^T)" if adjoint else "forward"
    line = {
        "metric": "solves/s at 100x100x60 grid, tol 1e-7: mixed+batched vs FP64 GPU; SpMV HBM GB/s",
        "value": value, "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "c32 inner / c64 refinement",
        "data": ("synthetic (seeded blocky model; boundary data and initial guess from the "
                 "2-D TE slice solves of BC_x0_WS, on the device" +
                 ("; impedance data at the sites = the model's own responses + 5 % seeded noise)"
                  if adjoint else ")")),
        "boundary_fields": bc_info,
        "config": {"workload": f"{cfg}: {label} air, {n} interior edges, {nf} f x 2 pol "
                               f"{fwd_adj} systems ({per_step} solves per step), tol {args.tol:g} "
                               "FP64 true residual",
                   "nsys": nsys, "seed": args.seed, "inner_tol": args.inner_tol,
                   "inner_layout": args.inner_layout,
                   "tuning": ({"autotune": True, "rows_total": tuned} if tuned else
                              {"autotune": False, "rows_total": 2048,
                               "brick_systems_per_block": args.brick_systems}),
                   "restart": args.restart, "sites": (len(sites) if adjoint else None),
                   "precond": {0: "block-DILU (reference ordering)", 1: "Jacobi", 2: "none",
                               3: "block-DILU, red-black 8x8x8 tile ordering"}[args.precond]},
        "converged": ok, "max_relres": relres,
        "forward_solves_per_s": nsys * args.steps / t_fwd,
        "adjoint_solves_per_s": (nsys * args.steps / t_adj) if adjoint else None,
        "end_to_end_solves_per_s": e2e,
        "end_to_end_note": "the step plus the device boundary fields (BC_x0_WS) and Mult_Aib",
        "launch": "eager, HIP events around the SpMV launches and DILU sweeps",
        "graph_solves_per_s": graph_value,
        "fp64_gpu_solves_per_s": fp64,
        "speedup_vs_fp64_gpu": (value / fp64) if fp64 else None,
        "ablation_solves_per_s": ablation or None,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
