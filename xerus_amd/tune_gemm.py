"""Measure libxtt's GEMM tile table on this device and write it to xerus_amd/tables/gemm_gfx950.tsv.

    XTT_PROFILE_SHAPES=1 XTT_NO_TUNING=1 python -m xerus_amd.tune_gemm [--iters 50] [--out PATH]

Collects the solo GEMM shapes of three workloads (ALS_SPD + round(1e-12) steps at d=16 n=16 rank 64 in FP64 and in
FP32 refinement; 64 problems of the uneven batch of ALS/round problems; ADF sweeps at d=10 n=10 r=8 P=200k), times the
heuristic and every candidate tile / split-K choice per shape and writes the measured table.
"""
import argparse
import os
import sys
import time

import numpy as np

if not os.environ.get("XTT_PROFILE_SHAPES"):
    sys.exit("set XTT_PROFILE_SHAPES=1 (and XTT_NO_TUNING=1) before running the tuner")

import xerus_amd as xe  # noqa: E402
from xerus_amd import tuning  # noqa: E402
from xerus_amd.workloads import batch_problems, laplace_cores  # noqa: E402

SEED = 0xBAADF00D


def als_steps(precision, steps=6):
    d, n = 16, 16
    pots = np.random.default_rng(SEED).random((d, n))
    A = xe.TTOperator.from_components(laplace_cores(d, n, list(pots), rank3=True))
    xe.seed(SEED)
    b = xe.TTTensor.random([n] * d, [4] * (d - 1))
    x = xe.TTTensor.random([n] * d, [64] * (d - 1))
    als = xe.ALS_SPD.copy()
    als.convergenceEpsilon = 0.0
    als.precision = precision
    for _ in range(steps):
        als(A, x, b, 2)
        x.round(1e-12)


def batch(count=64):
    als = xe.ALS_SPD.copy()
    als.convergenceEpsilon = 0.0
    als.numHalfSweeps = 2
    for A, x, b, _shape in batch_problems(count, SEED):
        als(A, x, b)
        x.round(1e-12)


def adf(iterations=4):
    d, n, r, P = 10, 10, 8, 200_000
    xe.seed(SEED)
    truth = xe.TTTensor.random([n] * d, [r] * (d - 1))
    x = xe.TTTensor.random([n] * d, [r] * (d - 1))
    m = xe.SinglePointMeasurementSet.random(P, [n] * d)
    m.measure(truth)
    a = xe.ADFVariant(iterations, 0.0, 1.0)
    a.precision = xe.XT_F32
    a(x, m)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--iters", type=int, default=50)
    ap.add_argument("--out", default=tuning.DEFAULT_TABLE)
    args = ap.parse_args()
    xe.lib()
    shapes = {}
    for name, run in [("als f64", lambda: als_steps(xe.XT_F64)), ("als f32", lambda: als_steps(xe.XT_F32)), ("batch", batch), ("adf", adf)]:
        t0 = time.perf_counter()
        for s in tuning.collect_shapes(run):
            shapes[s[:7]] = shapes.get(s[:7], 0) + s[7]
        print(f"# {name}: {len(shapes)} shapes so far ({time.perf_counter() - t0:.1f} s)", flush=True)
    ordered = sorted(((*k, c) for k, c in shapes.items()), key=lambda r: -r[-1])
    t0 = time.perf_counter()
    rows = tuning.tune(ordered, args.iters)
    gain = sum(r[7] * (r[13] - r[10]) for r in rows)
    total = sum(r[7] * r[13] for r in rows)
    changed = sum(1 for r in rows if (r[8], r[9]) != (r[11], r[12]))
    print(f"# tuned {len(rows)} shapes in {time.perf_counter() - t0:.1f} s; {changed} differ from the heuristic; "
          f"launch-weighted time {total / 1e3:.2f} -> {(total - gain) / 1e3:.2f} ms", flush=True)
    header = (f"libxtt GEMM tile table, measured by `python -m xerus_amd.tune_gemm --iters {args.iters}` on {xe_device()}\n"
              "precision 0 = FP64, 1 = FP32; launches = count in the collected workloads (ALS_SPD+round d16 n16 r64 FP64/FP32,\n"
              "64 batch problems, ADF d10 n10 r8 P200k); us = mean µs per launch (HIP events); tile 32/64, slices = split-K (0: none)")
    tuning.save(rows, args.out, header)
    print(f"# wrote {args.out}", flush=True)


def xe_device():
    try:
        with open("/proc/driver/amdgpu/0/name") as f:  # not always present
            return f.read().strip()
    except OSError:
        return "gfx950 (MI355X)"


if __name__ == "__main__":
    main()
