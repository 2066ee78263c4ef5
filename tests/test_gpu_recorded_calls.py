"""Recorded calls: the speculative steady ALS call and the ADF sweep replayed as hipGraphs.

The graph path must give what the eager path gives (XTT_NO_GRAPHS=1, read per call) at solver rounding: the same
ranks, core position, statistics and energies, the same tensor to 1e-11 relative. A replay whose device checks fail
(here: b scaled in place after the graphs were recorded, so every local start residual fails) must fall back to the
normal run with x untouched by the replay.
"""
import os

import numpy as np
import pytest

from helpers import SEED

pytestmark = pytest.mark.gpu


def problem(xe, d, n, rx, rb=4):
    from xerus_amd.workloads import laplace_cores
    pots = np.random.default_rng(SEED).random((d, n))
    A = xe.TTOperator.from_components(laplace_cores(d, n, list(pots), rank3=True))
    xe.seed(SEED)
    b = xe.TTTensor.random([n] * d, [rb] * (d - 1))
    x0 = xe.TTTensor.random([n] * d, [rx] * (d - 1))
    return A, b, x0


def steps(xe, A, b, x0, count, graphs, precision):
    if not graphs:
        os.environ["XTT_NO_GRAPHS"] = "1"
    try:
        x = x0.copy()
        als = xe.ALS_SPD.copy()
        als.convergenceEpsilon = 0.0
        als.precision = precision
        energies, stats = [], []
        for _ in range(count):
            energies.append(als(A, x, b, 2))
            x.round(1e-12)
            st = als.last_stats
            stats.append((st["localSolves"], st["cgIterations"], st["halfSweeps"]))
        return x, energies, stats
    finally:
        os.environ.pop("XTT_NO_GRAPHS", None)


def rel_diff(xe, x, y):
    return (x - y).frob_norm() / y.frob_norm()


@pytest.mark.parametrize("d,n,r,prec,count", [(8, 8, 16, "f64", 12), (16, 16, 64, "f32", 8)])
def test_recorded_steady_steps_match_eager(xe, d, n, r, prec, count):
    """The small case needs inner iterations for six steps, the d=16 case for two; after that a converged call, an
    eager speculative one and the recording come before the replays."""
    print(f"recorded calls d={d} n={n} r={r} {prec}", flush=True)
    precision = {"f64": xe.XT_F64, "f32": xe.XT_F32}[prec]
    A, b, x0 = problem(xe, d, n, r)
    before = xe._lib.graph_replays()
    xg, eg, sg = steps(xe, A, b, x0, count, True, precision)
    after = xe._lib.graph_replays()
    xn, en, sn = steps(xe, A, b, x0, count, False, precision)
    print(f"  replays als {after[0] - before[0]}; diff {rel_diff(xe, xg, xn):.3e}", flush=True)
    assert after[0] - before[0] >= 3  # the steps after a converged, an eager speculative and the recording call
    assert xg.ranks() == xn.ranks() and xg.corePosition == xn.corePosition
    assert sg == sn
    assert np.allclose(eg, en, rtol=1e-11, atol=0)
    assert rel_diff(xe, xg, xn) < 1e-11


def test_replay_failure_falls_back(xe):
    """After the graphs are recorded, b is scaled in place (same core addresses, so the recorded call is replayed):
    every site's start residual fails on the device, the replay is rejected and the normal run solves; the result is
    the eager path's."""
    d, n, r = 8, 8, 16
    A, b, x0 = problem(xe, d, n, r)
    results = []
    for graphs in (True, False):
        if not graphs:
            os.environ["XTT_NO_GRAPHS"] = "1"
        try:
            bb = b.copy()
            x = x0.copy()
            als = xe.ALS_SPD.copy()
            als.convergenceEpsilon = 0.0
            for _ in range(10):
                als(A, x, bb, 2)
                x.round(1e-12)
            before = xe._lib.graph_replays()
            xe._lib.call("xt_tt_scale", bb._h, 2.0)
            e = als(A, x, bb, 2)
            after = xe._lib.graph_replays()
            results.append((x, e, als.last_stats["cgIterations"], after[0] - before[0]))
        finally:
            os.environ.pop("XTT_NO_GRAPHS", None)
    (xg, eg, cg, rg), (xn, en, cn, _) = results
    print(f"  energies {eg:.12e} {en:.12e}, cg {cg} {cn}, accepted replays {rg}", flush=True)
    assert rg == 0 and cg == cn and cg > 0
    assert abs(eg - en) <= 1e-11 * abs(en)
    assert rel_diff(xe, xg, xn) < 1e-11


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_recorded_adf_sweeps_match_eager(xe, prec):
    """The stacked ADF solve with its sweeps replayed from a graph gives the eager sweeps' residual and tensor (the
    same kernels in the same order: equal to rounding), and every sweep after the second is a replay."""
    precision = {"f64": xe.XT_F64, "f32": xe.XT_F32}[prec]
    d, n, r, P = 8, 6, 4, 20000
    xe.seed(SEED)
    truth = xe.TTTensor.random([n] * d, [r] * (d - 1))
    x0 = xe.TTTensor.random([n] * d, [r] * (d - 1))
    m = xe.SinglePointMeasurementSet.random(P, [n] * d)
    m.measure(truth)
    out = []
    for graphs in (True, False):
        if not graphs:
            os.environ["XTT_NO_GRAPHS"] = "1"
        try:
            x = x0.copy()
            adf = xe.ADFVariant(12, 0.0, 1.0)
            adf.precision = precision
            before = xe._lib.graph_replays()
            res = adf(x, m)
            after = xe._lib.graph_replays()
            out.append((x, res, adf.last_iterations, after[1] - before[1]))
        finally:
            os.environ.pop("XTT_NO_GRAPHS", None)
    (xg, rg, ig, ng), (xn, rn, in_, nn) = out
    print(f"  residual {rg:.12e} {rn:.12e}; replays {ng} {nn}", flush=True)
    assert ig == in_ == 12 and ng == 10 and nn == 0  # sweep 1 eager, 2 recorded and launched, 3..11 replayed, 12 alone
    assert abs(rg - rn) <= 1e-12 * rn
    assert xg.corePosition == xn.corePosition
    assert rel_diff(xe, xg, xn) < 1e-12
