"""Benchmark of the SP2 3D-MT solve on one MI355X.

Workload (the config the headline metric is quoted on): 100x100x60 cells + 10
air layers (2,052,270 interior edges), 200 seeded blocky boxes over 100 ohm-m,
8 frequencies 1e-3..1e2 Hz x 2 polarizations = 16 forward systems, FP64 true
relative residual <= 1e-7 for every system.

A step is one pass of the hot path over the batch: the 16 forward systems
solved together in the mixed-precision batched solver (FP32 BiCGStab inner
iterations preconditioned with the red-black tile block-DILU, FP64 defect
correction, DILU pivots computed on the device inside the step), with
right-hand sides and initial guesses already resident in HBM.  value =
solves/s over the timed steps, which run with HIP events around every SpMV
launch and preconditioner application (the roofline figures come from them).
Beside it: the same steps replayed as hipGraphs, the plain FP64 GPU solver
(one system at a time, block-DILU in the reference ordering: the drop-in
cuBiCG semantics), the isolated SpMV kernel, and the reference CPU path timed
on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "modem-gpu_amd"))

PEAK_HBM_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(grid, lc, om, B, X0, iters_per_solve, sample_iters=12):
    """Reference CPU SP2 BiCG (flang build in oracle/_ref) timed for a bounded
    number of iterations on one system of the same workload; falls back to the
    oracle's C restatement when the reference build is absent."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    try:
        import refdriver
        if refdriver.available():
            t0 = time.time()
            out = refdriver.run(grid, lc, np.log(1e-10), [om[0]],
                                solves=[(0, sample_iters, 1e-30, B[0], X0[0])], dump=False,
                                timeout=900)
            r = out["bicg_0"]
            per_it = r["seconds"] / max(1, sample_iters)
            kind, total = "reference", time.time() - t0
        else:
            raise FileNotFoundError
    except (ImportError, FileNotFoundError, OSError):
        import sp2oracle
        t0 = time.time()
        p = sp2oracle.Problem(grid["dx"], grid["dy"], grid["dz"], grid["nzAir"], lc)
        p.update_freq(om[0])
        t1 = time.time()
        p.bicg(B[0], X0[0], maxit=sample_iters, tol=1e-30, gpu_semantics=False)
        per_it = (time.time() - t1) / sample_iters
        kind, total = "port", time.time() - t0
    solves_per_s = 1.0 / (per_it * iters_per_solve)
    bc_s = None
    try:
        import refdriver
        if refdriver.bcx0_available():
            # the reference's own BC_x0_WS for one system of the workload
            t0 = time.time()
            refdriver.run_bcx0(grid, lc, np.log(1e-10), [(om[0], 1)], timeout=600)
            bc_s = time.time() - t0
    except (ImportError, OSError):
        bc_s = None
    return {"value": solves_per_s, "unit": "solves/s", "cores": 1, "kind": kind,
            "bc_x0_reference_s_per_system": bc_s,
            "bc_x0_reference_note": ("wall time of the reference's ComputeBC -> BC_x0_WS for one "
                                     "(frequency, mode) of the workload, including the driver's "
                                     "grid/operator set-up"),
            "sample": (f"{sample_iters} BiCGStab iterations of one system (f={om[0]/(2*np.pi):.3g} Hz) "
                       f"of the same 100x100x60+10 workload on 1 host core: {per_it:.3f} s/iteration, "
                       f"scaled by the mean FP64 iterations per solve ({iters_per_solve:.1f}); "
                       f"{total:.1f} s wall incl. operator set-up")}


TRAFFIC_SUMMARY = os.path.join("profiles", "mixed_batch16_100x100x60_step_hbm_counters.md")


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
    ap.add_argument("--config", default="cfg1")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--tol", type=float, default=1e-7)
    ap.add_argument("--inner-tol", type=float, default=3e-4,
                    help="least residual reduction asked of each FP32 inner solve")
    ap.add_argument("--restart", type=int, default=120,
                    help="BiCGStab restart interval of the inner solves (the reference's 120)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--precond", type=int, default=3,
                    help="preconditioner of the mixed solver's FP32 inner solves "
                         "(0 block-DILU in the reference ordering, 3 red-black tile block-DILU); "
                         "the plain FP64 baseline always uses the reference ordering")
    ap.add_argument("--kernels-only", action="store_true",
                    help="only the isolated SpMV launches (FP32 x 16 systems, FP64 x 1) and a "
                         "1 GiB device copy, for hardware-counter runs")
    ap.add_argument("--spmv-variant", type=int, default=None,
                    help="SpMV kernel of the solves: 0 row, 1 k-marching (FP32), 3 multi-system row")
    ap.add_argument("--spmv-systems", type=int, default=4,
                    help="systems per block of the multi-system row SpMV (4 or 8)")
    ap.add_argument("--dilu-variant", type=int, default=5, choices=(1, 4, 5),
                    help="red-black DILU launch shape: 1 single-system persistent, 4 four systems per "
                         "workgroup, 5 the same with the lean LDS layout (three workgroups per CU)")
    ap.add_argument("--bc", choices=("2d", "1d"), default="2d",
                    help="boundary data and initial guess: 2d = the reference's 2-D TE slice "
                         "solves (BC_x0_WS) on the device; 1d = a plane wave over a halfspace")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch kernels eagerly instead of replaying hipGraphs")
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
    grid = P.config_grid(args.config)
    lc = P.config_model(args.config, grid, args.seed)
    fhz = P.config_freqs(args.config)
    ctx = mtsp2.Context(0)
    ctx.set_dilu_variant(args.dilu_variant)
    t0 = time.time()
    ctx.assemble(grid["dx"], grid["dy"], grid["dz"], grid["nzAir"], lc)
    log(f"assembled n_int={ctx.n_int} nnz={ctx.nnz} in {time.time()-t0:.1f}s")
    bc_info = None
    if args.bc == "2d" and not args.kernels_only:
        # the reference's start of a forward solve, on the device: fwdSetup ->
        # ComputeBC -> BC_x0_WS (mode 1, 2 per frequency), then FWDsolve3D's
        # b = -CCib e_b and x0 = e(EDGEi)
        om = np.repeat(2 * P.PI * np.asarray(fhz, np.float64), 2)
        modes = np.tile(np.array([1, 2], np.int32), len(fhz))
        nsys, n = len(om), ctx.n_int
        zb = nsys * n * 16
        d_b, d_x0, d_x = ctx.dev_alloc(zb), ctx.dev_alloc(zb), ctx.dev_alloc(zb)
        d_e = ctx.dev_alloc(nsys * ctx.n_edge * 16)
        ctx.bc_x0_dev(om, modes, d_e)                   # sizes the scratch
        t0 = time.time()
        ctx.bc_x0_dev(om, modes, d_e)
        t_bc = time.time() - t0
        t0 = time.time()
        ctx.rhs_fwd_dev(nsys, d_e, d_b, d_x0)
        t_rhs = time.time() - t0
        ctx.dev_free(d_e)
        bc_info = {"bc_x0_ms": 1e3 * t_bc, "rhs_fwd_ms": 1e3 * t_rhs, "systems": nsys}
        log(f"boundary fields of {nsys} systems on the device: BC_x0_WS {1e3 * t_bc:.1f} ms, "
            f"Mult_Aib {1e3 * t_rhs:.1f} ms")
        B = X0 = None
        if not args.no_cpu:
            B = ctx.download(d_b, np.zeros((nsys, n), np.complex128))
            X0 = ctx.download(d_x0, np.zeros((nsys, n), np.complex128))
    else:
        om, B, X0 = P.forward_systems(ctx, grid, fhz)
        nsys, n = B.shape
        zb = nsys * n * 16
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
        # calibration kernel for the counters: 1 GiB read + 1 GiB written
        gib = 1 << 30
        src, dst = ctx.dev_alloc(gib), ctx.dev_alloc(gib)
        for _ in range(3):
            ctx.dev_copy(dst, src, gib)
        log("device copy of 1 GiB x 3 done")
        ctx.dev_free(src)
        ctx.dev_free(dst)
        ctx.close()
        return
    if bc_info is None:
        d_b, d_x0, d_x = ctx.dev_alloc(zb), ctx.dev_alloc(zb), ctx.dev_alloc(zb)
        ctx.upload(d_b, B)
        ctx.upload(d_x0, X0)
    tuned = ctx.autotune(nsys)
    log(f"autotune: {tuned} row-tile groups per system")
    if args.spmv_variant is not None:
        ctx.set_spmv_variant(args.spmv_variant,
                             args.spmv_systems if args.spmv_variant == 3 else 0)
    opts = dict(tol=args.tol, maxit=4000, precision=mtsp2.PREC_MIXED, precond=args.precond,
                inner_tol=args.inner_tol, restart=args.restart, chunk=10,
                use_graph=0 if args.no_graph else 1)

    def step():
        ctx.dev_copy(d_x, d_x0, zb)
        return ctx.solve_batch_dev(om, d_b, d_x, **opts)

    for w in range(args.warmup):
        st = step()
        log(f"warmup {w}: iters {[s['niter'] for s in st]}")
    # HIP events around every SpMV launch and DILU sweep of the timed steps
    # (with events on, the solver launches eagerly: HIP cannot time events
    # recorded by graph nodes)
    ctx.timing_enable(True)
    t0 = time.time()
    for k in range(args.steps):
        st = step()
    el = time.time() - t0
    kt = {name: ctx.kernel_timing(fam) for name, fam in
          (("spmv", mtsp2.KT_SPMV), ("precond", mtsp2.KT_PRECOND))}
    ctx.timing_enable(False)
    for name, (nl, kms, kby) in kt.items():
        if nl:
            log(f"timed region, {name}: {nl} launches, {kms:.1f} ms, {1e3 * kms / nl:.1f} us/launch, "
                f"{kby / (kms * 1e-3) / 1e9:.0f} GB/s algorithmic")
    ms = 1e3 * el / args.steps
    value = nsys * args.steps / el
    graph_value = None
    if not args.no_graph:
        # the same steps replayed as hipGraphs, without events (reported beside)
        t0 = time.time()
        for k in range(args.steps):
            step()
        graph_value = nsys * args.steps / (time.time() - t0)
        log(f"hipGraph replay without events: {graph_value:.2f} solves/s")
    ok = all(s["status"] == 0 and s["relres"] <= args.tol for s in st)
    relres = max(s["relres"] for s in st)
    inner = [s["niter"] for s in st]
    log(f"mixed+batched: {ms:.1f} ms/step, {value:.2f} solves/s, max relres {relres:.2e}, "
        f"inner iters {inner}, outer {[s['nouter'] for s in st]}")

    fp64 = None
    fp64_iters = None
    if not args.no_fp64:
        # plain FP64 GPU solver: one system at a time (drop-in semantics)
        t0 = time.time()
        its = []
        for s in range(nsys):
            bs = d_b + s * n * 16
            xs = d_x + s * n * 16
            ctx.dev_copy(xs, d_x0 + s * n * 16, n * 16)
            r = ctx.solve_batch_dev(om[s:s + 1], bs, xs, tol=args.tol, maxit=4000,
                                    precision=mtsp2.PREC_FP64, precond=mtsp2.PC_DILU,
                                    chunk=10, use_graph=opts["use_graph"])
            its.append(r[0]["niter"])
        fp64 = nsys / (time.time() - t0)
        fp64_iters = float(np.mean(its))
        log(f"plain FP64 GPU: {fp64:.2f} solves/s, iters {its}")

    dom = max(kt, key=lambda k: kt[k][1])
    nl, kms, kby = kt[dom]
    rb_kernel = {4: "k_dilu_rbm<float,float,RED_FWD|BLACK_FB|RED_BWD,4>",
                 5: "k_dilu_rbl<float,float,RED_FWD|BLACK_FB|RED_BWD,4> (lean LDS)",
                 1: "k_dilu_rb2<float,float,RED_FWD|BLACK_FB|RED_BWD>"}[args.dilu_variant]
    pc_kernels = {0: "k_dilu_tile<float,float,8,8,8,true|false> (block-DILU level sweeps)",
                  3: rb_kernel + " (red-black tile block-DILU, 3 launches per application)"}
    spv = 3 if args.spmv_variant is None else args.spmv_variant
    spmv_kernels = {0: "k_spmv_dot<float,float,1|2> (row-parallel v = A x with its dot products)",
                    1: "k_spmv_km<float,float,1|2> (k-marching v = A x with its dot products)",
                    2: "k_spmv_km<float,float,1|2> (k-marching v = A x with its dot products)",
                    3: f"k_spmv_ms<float,float,1|2,{args.spmv_systems}> (v = A x for "
                       f"{args.spmv_systems} systems per block, with the dot products)"}
    kernels = {"spmv": spmv_kernels[spv],
               "precond": pc_kernels.get(args.precond, "preconditioner")}
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
    roofline["spmv_isolated_fp32_16sys"] = {"us_per_launch": spmv_ms * 1e3,
                                            "bytes_per_launch": spmv_bytes,
                                            "GBps": spmv_bytes / (spmv_ms * 1e-3) / 1e9}
    pc_ms, pc_bytes = ctx.bench_precond(nsys, mtsp2.PREC_MIXED, mtsp2.PC_DILU_RB, reps=10)
    roofline["precond_isolated_fp32_16sys"] = {"us_per_apply": pc_ms * 1e3,
                                               "bytes_per_apply": pc_bytes,
                                               "GBps": pc_bytes / (pc_ms * 1e-3) / 1e9}
    log(f"red-black DILU isolated (FP32, {nsys} systems): {pc_ms * 1e3:.1f} us/application, "
        f"{pc_bytes / 1e6:.1f} MB, {pc_bytes / (pc_ms * 1e-3) / 1e9:.0f} GB/s")
    spmv64_ms, spmv64_bytes = ctx.bench_spmv(1, mtsp2.PREC_FP64, reps=20)
    log(f"SpMV (FP64, 1 system): {spmv64_ms*1e3:.1f} us/launch, "
        f"{spmv64_bytes/spmv64_ms/1e6:.0f} GB/s")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(grid, lc, om, B, X0, fp64_iters or float(np.mean(inner)))
        log(f"cpu baseline: {cpu}")

    line = {
        "metric": "solves/s at 100x100x60 grid, tol 1e-7: mixed+batched vs FP64 GPU; SpMV HBM GB/s",
        "value": value, "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "c32 inner / c64 refinement",
        "data": ("synthetic (seeded blocky model; boundary data and initial guess from the "
                 "2-D TE slice solves of BC_x0_WS, on the device)") if bc_info else
                "synthetic (seeded blocky model, plane-wave boundary data)",
        "boundary_fields": bc_info,
        "config": {"workload": f"{args.config}: 100x100x60+10 air, 2052270 interior edges, "
                               "8 f x 2 pol forward systems, tol 1e-7 FP64 true residual",
                   "nsys": nsys, "seed": args.seed, "inner_tol": args.inner_tol,
                   "restart": args.restart,
                   "precond": {0: "block-DILU (reference ordering)", 1: "Jacobi", 2: "none",
                               3: "block-DILU, red-black 8x8x8 tile ordering"}[args.precond]},
        "converged": ok, "max_relres": relres,
        "launch": "eager, HIP events around the SpMV launches and DILU sweeps",
        "graph_solves_per_s": graph_value,
        "fp64_gpu_solves_per_s": fp64,
        "speedup_vs_fp64_gpu": (value / fp64) if fp64 else None,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    for p in (d_b, d_x0, d_x):
        ctx.dev_free(p)
    ctx.close()


if __name__ == "__main__":
    main()
