#!/usr/bin/env python
"""Benchmark of the TT hot path on one MI355X.

Headline: ALS_SPD full sweeps per second on the d=16, n=16 Dirichlet-Laplacian + U(0,1) diagonal potential
operator (rank-3 channel form), random rank-4 right-hand side, solution rank 64; one step = one full sweep
(ALS_SPD with numHalfSweeps=2, convergenceEpsilon=0) followed by round(1e-12). The mixed-precision path
(FP32 or BF16 local matvecs with FP64 refinement) is the reported value; the FP64 device path runs the same
steps from the same start for the ratio, and the relative residual ‖Ax−b‖/‖b‖ of both is reported.
Secondary: ADF gradients per second (d=10, n=10, rank 8, 200k sampled entries, FP32, all cores).

Prints one JSON line. `--gpus N` with N > 1 exits: this build runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

SEED = 0xBAADF00D
PEAK_TFLOPS = {"f64": 78.6, "f32": 157.3, "bf16": 2500.0}
PEAK_HBM_GBS = 8000.0


def build_config2(xe, np):
    from xerus_amd.workloads import laplace_cores
    d, n = 16, 16
    pots = np.random.default_rng(SEED).random((d, n))
    A = xe.TTOperator.from_components(laplace_cores(d, n, list(pots), rank3=True))
    xe.seed(SEED)
    b = xe.TTTensor.random([n] * d, [4] * (d - 1))
    x0 = xe.TTTensor.random([n] * d, [64] * (d - 1))
    return A, b, x0, pots


def run_als(xe, A, b, x0, precision, steps):
    x = x0.copy()
    als = xe.ALS_SPD.copy()
    als.convergenceEpsilon = 0.0
    als.precision = precision
    stats = {"cgIterations": 0, "localSolves": 0, "refinements": 0, "flopsMatvec": 0.0, "flopsEnv": 0.0}
    xe.lib().xt_synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        als(A, x, b, 2)
        x.round(1e-12)
        for k in stats:
            stats[k] += als.last_stats[k]
    xe.lib().xt_synchronize()
    return time.perf_counter() - t0, x, stats


def per_step(xe, A, b, x0, precision, steps):
    """The same steps one at a time, the device synchronised around ALS and around round: wall time of each part, inner
    CG iterations and FP64 refinements per step (a separate pass; the timed value comes from run_als)."""
    x = x0.copy()
    als = xe.ALS_SPD.copy()
    als.convergenceEpsilon = 0.0
    als.precision = precision
    out = []
    sync = xe.lib().xt_synchronize
    for _ in range(steps):
        sync()
        t0 = time.perf_counter()
        als(A, x, b, 2)
        sync()
        t1 = time.perf_counter()
        x.round(1e-12)
        sync()
        t2 = time.perf_counter()
        out.append({"als_ms": round(1e3 * (t1 - t0), 3), "round_ms": round(1e3 * (t2 - t1), 3),
                    "cg_iterations": als.last_stats["cgIterations"], "refinements": als.last_stats["refinements"],
                    "max_rank": max(x.ranks())})
    return out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def adf_workload(xe, np):
    d, n, r, P = 10, 10, 8, 200_000
    rng = np.random.default_rng(SEED + 5)
    xe.seed(SEED)

    def tt(scale):
        t = xe.TTTensor([n] * d)
        for k in range(d):
            rl = 1 if k == 0 else r
            rr = 1 if k == d - 1 else r
            t.set_component(k, xe.normal_draws(rl * n * rr).reshape(rl, n, rr) * scale)
        return t

    truth = tt(1 / np.sqrt(8))
    x = tt(1 / np.sqrt(8))
    m = xe.SinglePointMeasurementSet.random(P, [n] * d)
    m.measure(truth)
    # "N(0, 1e-4)" noise read with the variance convention of the cores: σ = 1e-2
    m.measuredValues = m.measuredValues + rng.normal(0.0, 1e-2, P)
    return d, n, r, P, x, m


def bench_adf(xe, np, steps):
    """Full gradients (every core) of ½‖P_Ω(x) − b‖² through the raw FP32 kernels on device-resident data."""
    import ctypes as C
    from xerus_amd._lib import DeviceArray, call
    d, n, r, P, x, m = adf_workload(xe, np)
    cores = [x.get_component(k) for k in range(d)]
    flat = DeviceArray.from_numpy(np.concatenate([c.ravel() for c in cores]).astype(np.float32))
    vals_b = DeviceArray.from_numpy(m.measuredValues.astype(np.float32))
    resid = DeviceArray(P, np.float32)
    grad = DeviceArray(flat.shape, np.float32)
    ns = (C.c_int * d)(*[n] * d)
    rs = (C.c_int * (d + 1))(*([1] + [r] * (d - 1) + [1]))
    pos = m.positions_device_ptr()
    sums = (C.c_double * 2)()

    def one():  # residual x − b (with its FP64 norms) and the gradient of every core, one pass over the samples
        call("xt_tt_sample_residual_adjoint_f32", d, ns, rs, flat.ptr, pos, P, vals_b.ptr, resid.ptr, sums, grad.ptr, None)

    for _ in range(3):
        one()
    xe.lib().xt_synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        one()
    xe.lib().xt_synchronize()
    dt = time.perf_counter() - t0
    # kernel durations from a separate pass with the launch profiler on (its events are not in the timed loop)
    xe._lib.profile_enable(True)
    for _ in range(min(steps, 50)):
        one()
    xe.lib().xt_synchronize()
    prof = xe._lib.profile_read()
    xe._lib.profile_enable(False)
    # algorithmic bytes per full gradient: int32 indices + FP32 value in + FP32 residual, cores in + grads out
    alg_bytes = P * (4 * d + 4 + 4) + 2 * int(np.prod(flat.shape)) * 4
    g = (prof.get("adf_grad_fast_kernel<float> residual") or prof.get("adf_grad_fast_kernel<float>")
         or prof.get("adf_grad_kernel<float>"))
    rk = prof.get("adf_residual_fast_kernel<float>") or prof.get("adf_residual_kernel<float>")
    return {"grads_per_s": steps / dt, "ms_per_grad": 1e3 * dt / steps, "P": P, "alg_bytes": alg_bytes,
            "residual_norm": float(np.sqrt(sums[0] / sums[1])),
            "grad_kernel_ms": g[1] / g[0] if g else None, "residual_kernel_ms": rk[1] / rk[0] if rk else None}


def bench_adf_solve(xe, np, iterations):
    """ADFVariant(iterations, 0, 1) at d=10, n=10, r=8, P=200k (FP32 kernels, FP64 cores and norms): every iteration is a
    full sweep of residual, gradient, slice norms and update_x over all cores plus the QR moves; ADF iterations/s."""
    d, n, r, P, x, m = adf_workload(xe, np)
    adf = xe.ADFVariant(iterations, 0.0, 1.0)  # no early stop: a fixed number of sweeps
    adf.precision = xe.XT_F32
    y = x.copy()
    xe.ADFVariant(1, 0.0, 1.0)(y, m)  # warm-up (allocations, descriptors)
    y = x.copy()
    xe.lib().xt_synchronize()
    t0 = time.perf_counter()
    res = adf(y, m)
    xe.lib().xt_synchronize()
    dt = time.perf_counter() - t0
    return {"iterations_per_s": adf.last_iterations / dt, "iterations": adf.last_iterations, "seconds": dt,
            "residual": res, "precision": "f32 samples, f64 cores"}


def adf_cpu_baseline(np, iterations):
    """The oracle's ADF (FP64, OpenMP) on the same workload for a bounded number of iterations."""
    from oracle import oracle as O
    d, n, r, P = 10, 10, 8, 200_000
    O.seed(SEED)

    def tt(scale):
        cores = []
        for k in range(d):
            rl = 1 if k == 0 else r
            rr = 1 if k == d - 1 else r
            cores.append(O.normal_draws(rl * n * rr).reshape(rl, n, rr) * scale)
        return O.TT.from_cores(cores)

    truth = tt(1 / np.sqrt(8))
    x = tt(1 / np.sqrt(8))
    m = O.Measurements.random(P, [n] * d)
    m.measure(truth)
    m.set_values(m.values() + np.random.default_rng(SEED + 5).normal(0.0, 1e-2, P))
    t0 = time.perf_counter()
    grads = 3
    for _ in range(grads):
        O.adf_gradient(x, m)
    tg = (time.perf_counter() - t0) / grads
    t0 = time.perf_counter()
    O.adf_solve(x, m, max_iterations=iterations, target=0.0, min_decrease=1.0)
    ts = (time.perf_counter() - t0) / iterations
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"grads_per_s": 1.0 / tg, "iterations_per_s": 1.0 / ts, "cores": threads, "kind": "port", "cpu": cpu_model(),
            "sample": f"{grads} full gradients and {iterations} ADF iterations of the same d=10 n=10 r=8 P=200k workload"}


def bench_batch(xe, np, count, sweeps, threads, precision):
    """Many independent ALS_SPD + round problems of uneven size (d ~ U{6..20}, n ∈ {2,4,8,16}, r ~ U{4..64},
    rank-3 Laplacian + potential operators, rank-4 right-hand sides) through the batched driver: threads=0 is the
    lockstep driver (one grouped launch per step for all problems), threads > 0 worker threads with a stream each. Returns aggregate sweeps/s (count·sweeps / wall time, residuals included) and residual
    statistics."""
    from xerus_amd.workloads import batch_problems
    probs = batch_problems(count)
    als = xe.ALS_SPD.copy()
    als.convergenceEpsilon = 0.0
    als.numHalfSweeps = 2
    als.precision = precision
    xe.lib().xt_synchronize()
    t0 = time.perf_counter()
    res, st = xe.als_round_batch(als, [p[0] for p in probs], [p[1] for p in probs], [p[2] for p in probs], sweeps, 1e-12,
                                 threads=threads)
    dt = time.perf_counter() - t0
    shapes = np.array([p[3] for p in probs])
    res = np.asarray(res)
    worst = [{"problem": int(i), "d": int(shapes[i, 0]), "n": int(shapes[i, 1]), "r": int(shapes[i, 2]), "residual": float(res[i])}
             for i in np.argsort(res)[::-1][:5]]
    return {"problems": count, "sweeps_per_problem": sweeps,
            "driver": ("lockstep (grouped launches)" if precision in (xe.XT_F64, xe.XT_F32) else "16 worker threads (no lockstep for bf16)")
            if threads <= 0 else f"{threads} worker threads",
            "precision": {xe.XT_F64: "f64", xe.XT_F32: "f32+f64-refinement", xe.XT_BF16: "bf16+f64-refinement"}[precision],
            "seconds": dt, "sweeps_per_s": count * sweeps / dt, "residual_max": float(np.max(res)),
            "residual_median": float(np.median(res)), "residual_p90": float(np.percentile(res, 90)),
            "above_1e-10": int((res > 1e-10).sum()), "worst": worst,
            "d_mean": float(shapes[:, 0].mean()), "n_mean": float(shapes[:, 1].mean()), "r_mean": float(shapes[:, 2].mean()),
            "cg_iterations": st["cgIterations"], "local_solves": st["localSolves"], "refinements": st["refinements"]}


def cpu_baseline(np, steps):
    """The oracle (FP64 CPU restatement of xerus's ALS_SPD and round, matrix-free CG local solver, OpenMP) on
    the same workload as the timed GPU run: `steps` × (ALS_SPD with 2 half-sweeps, round(1e-12)) from the same
    start."""
    from oracle import oracle as O
    from xerus_amd.workloads import laplace_cores
    d, n = 16, 16
    pots = np.random.default_rng(SEED).random((d, n))
    A = O.TT.from_cores(laplace_cores(d, n, list(pots), rank3=True), op=True)
    O.seed(SEED)
    b = O.TT.random([n] * d, [4] * (d - 1))
    x = O.TT.random([n] * d, [64] * (d - 1))
    t0 = time.perf_counter()
    for _ in range(steps):
        O.als(A, x, b, num_half_sweeps=2, eps=0.0, local_solver="cg", cg_tol=1e-12)
        x.round(eps=1e-12)
    dt = time.perf_counter() - t0
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": steps / dt, "unit": "sweeps/s", "cores": threads, "kind": "port", "cpu": cpu_model(),
            "sample": f"the full timed workload ({steps} steps of ALS_SPD(2 half-sweeps) + round(1e-12), d16 n16 r64, "
                      f"local CG tol 1e-12) on the host, {dt:.1f} s"}


def measured_traffic(tag):
    """HBM bytes per launch of a kernel (FETCH_SIZE + WRITE_SIZE per dispatch, calibrated against a copy of known
    size) from the committed rocprofv3 counter summary of this workload, or (None, None) when it has no entry."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_als_d16_n16_r64.json")
    try:
        with open(path) as f:
            table = json.load(f)
    except OSError:
        return None, None
    kernels = table.get("kernels", {})
    name = tag.split(" ")[0]
    # the launch profiler prints a bool template argument as 0/1, rocprofv3's kernel names as false/true
    entry = kernels.get(name) or kernels.get(name.replace(",0>", ",false>").replace(",1>", ",true>"))
    if not entry:
        return None, None
    return entry["fetch_bytes"] + entry["write_bytes"], os.path.relpath(path, ROOT)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", choices=["f32", "bf16"], default="f32")
    ap.add_argument("--adf-steps", type=int, default=200)
    ap.add_argument("--adf-iterations", type=int, default=20)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batch-problems", type=int, default=512)
    ap.add_argument("--batch-sweeps", type=int, default=2)
    ap.add_argument("--batch-threads", type=int, default=16)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="before the workload, copy MIB MiB three times with the library's permute kernel (identity "
                         "shuffle: reads and writes exactly MIB MiB), the known-size reference for HBM counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU layout is described in DESIGN.md",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import xerus_amd as xe

    if args.copy_probe > 0:
        import ctypes as C
        from xerus_amd._lib import DeviceArray, call
        count = args.copy_probe * (1 << 20) // 8
        src, dst = DeviceArray.from_numpy(np.ones(count)), DeviceArray((count,), np.float64)
        dims, ident = (C.c_size_t * 1)(count), (C.c_size_t * 1)(0)
        for _ in range(3):
            call("xt_reshuffle", 1, dims, ident, src.ptr, dst.ptr)
        print(f"# copy probe: 3 x permute_kernel over {args.copy_probe} MiB (read {args.copy_probe} MiB, write "
              f"{args.copy_probe} MiB each)", file=sys.stderr, flush=True)
        del src, dst

    prec = xe.XT_F32 if args.precision == "f32" else xe.XT_BF16
    A, b, x0, _ = build_config2(xe, np)
    print(f"# config2 built: x0 ranks {x0.ranks()}", file=sys.stderr, flush=True)

    if args.warmup > 0:
        run_als(xe, A, b, x0, prec, args.warmup)
        print("# warmup done", file=sys.stderr, flush=True)
    t_mixed, x_mixed, st_mixed = run_als(xe, A, b, x0, prec, args.steps)
    res_mixed = xe.residual(A, x_mixed, b)
    # Roofline pass: the timed workload once more with the launch profiler on (HIP events around every GEMM,
    # Householder and Jacobi launch on its stream; launches run eagerly while profiling, so this pass is not part
    # of the timed value).
    xe._lib.profile_enable(True)
    run_als(xe, A, b, x0, prec, args.steps)
    prof = xe._lib.profile_read()
    xe._lib.profile_enable(False)
    print(f"# mixed: {t_mixed:.2f} s, residual {res_mixed:.3e}, ranks {x_mixed.ranks()}", file=sys.stderr, flush=True)

    steps_mixed = per_step(xe, A, b, x0, prec, args.steps)
    print(f"# per step (mixed): {[(p['als_ms'], p['round_ms'], p['cg_iterations']) for p in steps_mixed]}", file=sys.stderr,
          flush=True)

    fp64 = None
    if not args.no_fp64:
        t64, x64, st64 = run_als(xe, A, b, x0, xe.XT_F64, args.steps)
        res64 = xe.residual(A, x64, b)
        diff = (x_mixed - x64).frob_norm() / x64.frob_norm()
        fp64 = {"sweeps_per_s": args.steps / t64, "ms_per_step": 1e3 * t64 / args.steps, "rel_residual": res64,
                "cg_iterations": st64["cgIterations"], "rel_diff_mixed_vs_fp64": diff,
                "per_step": per_step(xe, A, b, x0, xe.XT_F64, args.steps)}
        print(f"# fp64: {t64:.2f} s, residual {res64:.3e}", file=sys.stderr, flush=True)

    # dominant kernel of the timed workload (largest total time among the profiled launches); its bound is the
    # roofline side its algorithmic intensity (flops / compulsory bytes) falls on for its dtype
    dom_tag, (dom_n, dom_ms, dom_flops, dom_bytes) = max(prof.items(), key=lambda kv: kv[1][1])
    dtype_key = "bf16" if "bf16" in dom_tag else ("f32" if "float" in dom_tag else "f64")
    secs = dom_ms * 1e-3
    ridge = PEAK_TFLOPS[dtype_key] * 1e12 / (PEAK_HBM_GBS * 1e9)
    intensity = dom_flops / dom_bytes if dom_bytes else float("inf")
    if intensity < ridge:
        bound, achieved, peak, unit = "hbm", dom_bytes / secs / 1e9, PEAK_HBM_GBS, "GB/s"
    else:
        bound, achieved, peak, unit = "mfma", dom_flops / secs / 1e12, PEAK_TFLOPS[dtype_key], "TFLOP/s"
    traffic, traffic_src = measured_traffic(dom_tag)
    roofline = {"bound": bound, "kernel": dom_tag, "launches": dom_n, "avg_launch_us": 1e3 * dom_ms / dom_n,
                "flops_per_launch": dom_flops / dom_n, "bytes_per_launch": dom_bytes / dom_n,
                "intensity_flop_per_byte": intensity, "achieved": achieved, "peak": peak, "unit": unit,
                "frac": achieved / peak, "traffic": traffic, "traffic_source": traffic_src,
                "kernels_ms": {k: round(v[1], 3) for k, v in sorted(prof.items(), key=lambda kv: -kv[1][1])[:8]}}

    # MFMA use of the contraction kernels (the environment / stack contractions and local matvecs run there): their
    # algorithmic flops over their summed kernel time, against the peak of their arithmetic type
    mm = {"f64": [0.0, 0.0], "f32": [0.0, 0.0], "bf16": [0.0, 0.0]}
    for tag, (_n, ms, fl, _by) in prof.items():
        if tag.startswith("gemm"):
            key = "bf16" if "bf16" in tag else ("f32" if "float" in tag else "f64")
            mm[key][0] += fl
            mm[key][1] += ms
    contraction_mfma = {k: {"tflops": v[0] / (v[1] * 1e-3) / 1e12, "frac_of_peak": v[0] / (v[1] * 1e-3) / 1e12 / PEAK_TFLOPS[k],
                            "kernel_ms": v[1]} for k, v in mm.items() if v[1] > 0}

    adf = bench_adf(xe, np, args.adf_steps) if args.adf_steps > 0 else None
    if adf and adf["grad_kernel_ms"]:
        adf["grad_kernel_GBs"] = adf["alg_bytes"] / (adf["grad_kernel_ms"] * 1e-3) / 1e9
        adf["grad_kernel_hbm_frac"] = adf["grad_kernel_GBs"] / PEAK_HBM_GBS

    adf_solve_r = bench_adf_solve(xe, np, args.adf_iterations) if args.adf_iterations > 0 else None
    if adf_solve_r:
        print(f"# ADF solve: {adf_solve_r['iterations_per_s']:.1f} iterations/s", file=sys.stderr, flush=True)

    batch = None
    if args.batch_problems > 0:
        keys = ("driver", "precision", "seconds", "sweeps_per_s", "residual_max", "residual_median", "cg_iterations")
        if args.warmup > 0:  # one untimed batch: the allocator cache and the scratch reach their working size
            bench_batch(xe, np, args.batch_problems, args.batch_sweeps, 0, prec)
        batch = bench_batch(xe, np, args.batch_problems, args.batch_sweeps, 0, prec)
        print(f"# batch (lockstep, {batch['precision']}): {batch['sweeps_per_s']:.1f} sweeps/s over {args.batch_problems} problems",
              file=sys.stderr, flush=True)
        if prec != xe.XT_F64:
            f64 = bench_batch(xe, np, args.batch_problems, args.batch_sweeps, 0, xe.XT_F64)
            batch["lockstep_f64"] = {k: f64[k] for k in keys}
            print(f"# batch (lockstep, f64): {f64['sweeps_per_s']:.1f} sweeps/s", file=sys.stderr, flush=True)
        # MFMA use of the grouped contractions (stack pushes, right-hand sides and local products of every problem in
        # one launch each): one more lockstep batch under the launch profiler, algorithmic 2mnk of every task of a
        # launch (converged systems' skipped products included, so an upper bound) over the launches' kernel time
        from xerus_amd import _lib as xl
        xl.profile_enable(True)
        bench_batch(xe, np, args.batch_problems, args.batch_sweeps, 0, prec)
        recs = xl.profile_read(4096)
        xl.profile_enable(False)
        gm = {"f64": [0.0, 0.0, 0], "f32": [0.0, 0.0, 0]}
        for tag, (n_, ms, fl, _by) in recs.items():
            if tag.startswith("gemm_group_kernel"):
                k = "f32" if "float" in tag else "f64"
                gm[k][0] += fl
                gm[k][1] += ms
                gm[k][2] += n_
        batch["contraction_mfma"] = {k: {"tflops": v[0] / (v[1] * 1e-3) / 1e12, "frac_of_peak": v[0] / (v[1] * 1e-3) / 1e12 / PEAK_TFLOPS[k],
                                         "kernel_ms": v[1], "launches": v[2]} for k, v in gm.items() if v[1] > 0}
        if args.batch_threads > 0:
            ref = bench_batch(xe, np, args.batch_problems, args.batch_sweeps, args.batch_threads, prec)
            batch["threads_reference"] = {k: ref[k] for k in keys}
            print(f"# batch ({ref['driver']}, {ref['precision']}): {ref['sweeps_per_s']:.1f} sweeps/s", file=sys.stderr, flush=True)

    cpu = None if args.no_cpu else cpu_baseline(np, args.steps)
    adf_cpu = None if (args.no_cpu or not adf_solve_r) else adf_cpu_baseline(np, 2)
    if adf and adf_cpu:
        adf["cpu_grads_per_s"] = adf_cpu["grads_per_s"]
    if adf_solve_r and adf_cpu:
        adf_solve_r["cpu_baseline"] = adf_cpu

    value = args.steps / t_mixed
    out = {
        "metric": "ALS sweep/s d16 n16 r64 mixed vs FP64-GPU @rel.res 1e-10; ADF grads/s",
        "value": value,
        "unit": "sweeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * t_mixed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.precision + "+f64-refinement",
        "data": "synthetic (seeded Laplacian+potential operator, random TT right-hand side and start)",
        "config": {"workload": "ALS_SPD d=16 n=16 x-rank 64, A rank 3 (Laplacian + U(0,1) potential), b rank 4; "
                               "step = 2 half-sweeps + round(1e-12)",
                   "precision": args.precision, "rel_residual": res_mixed, "ranks_final": max(x_mixed.ranks()),
                   "cg_iterations": st_mixed["cgIterations"], "refinements": st_mixed["refinements"],
                   "per_step": steps_mixed},
        "fp64_gpu": fp64,
        "speedup_vs_fp64": (value / fp64["sweeps_per_s"]) if fp64 else None,
        "adf": adf,
        "adf_solve": adf_solve_r,
        "contraction_mfma": contraction_mfma,
        "batch": batch,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
