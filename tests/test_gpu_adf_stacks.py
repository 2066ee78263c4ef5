"""ADF's stack plan (adf.cpp:102-191) built on the device, and the ADF solve that sweeps with the stacks.

- The plan must equal the oracle's bit for bit: per level the forward and backward entry counts and their
  representatives (smallest sample id of each group, ascending), for sorted and unsorted position sets, sets with
  long shared prefixes and suffixes, duplicates, one sample and order one. Every sample's slot must point at an
  entry whose representative shares the sample's prefix (suffix).
- The stacked solve must match the oracle's ADF and the chain-walking device path (the same algorithm without
  stacks, XTT_ADF_NO_STACKS), including the rank-increase loop.
"""
import os

import numpy as np
import pytest

from helpers import SEED

pytestmark = pytest.mark.gpu


def brute_plan(pos):
    P, d = pos.shape
    fwd, bwd = [], []
    for k in range(d):
        groups = {}
        for i in range(P):
            key = tuple(pos[i, : k + 1])
            groups[key] = min(groups.get(key, i), i)
        fwd.append(sorted(groups.values()))
        groups = {}
        for i in range(P):
            key = tuple(pos[i, k:])
            groups[key] = min(groups.get(key, i), i)
        bwd.append(sorted(groups.values()))
    return fwd, bwd


def position_sets():
    rng = np.random.default_rng(SEED)
    out = {}
    # sorted, unique (what SinglePointMeasurementSet::random gives)
    p = np.unique(rng.integers(0, [5, 4, 6, 5], size=(600, 4)), axis=0)
    out["sorted"] = p
    # the same shuffled: groups are no longer runs of the id order
    out["shuffled"] = p[rng.permutation(len(p))]
    # long shared prefixes and suffixes, duplicates included (the plan works with repeated samples, adf.cpp:97)
    q = rng.integers(0, 2, size=(500, 9)).astype(np.uint32)
    q[:, 2:7] = 0
    out["shared"] = q
    out["one sample"] = np.array([[3, 1, 4]], dtype=np.uint32)
    out["order one"] = rng.integers(0, 7, size=(40, 1)).astype(np.uint32)
    out["wide modes"] = rng.integers(0, 70000, size=(300, 3)).astype(np.uint32)
    return out


@pytest.mark.parametrize("name", list(position_sets().keys()))
def test_stack_plan_matches_oracle(xe, oracle, name):
    print(f"stack plan: {name}", flush=True)
    pos = np.ascontiguousarray(position_sets()[name], dtype=np.uint32)
    P, d = pos.shape
    vals = np.zeros(P)
    m = xe.SinglePointMeasurementSet.from_arrays(pos, vals)
    mo = oracle.Measurements.from_arrays(pos, vals)
    plan = m.stack_plan(d)
    fc, bc, fi, bi, nu = oracle.adf_plan(mo, d)
    assert list(plan["fwd_counts"]) == list(fc) and list(plan["bwd_counts"]) == list(bc)
    assert np.array_equal(plan["fwd_ids"], fi) and np.array_equal(plan["bwd_ids"], bi)
    assert plan["num_unique"] == nu
    fwd, bwd = brute_plan(pos)
    assert [len(g) for g in fwd] == list(fc) and [len(g) for g in bwd] == list(bc)
    # slots: the entry's representative shares the sample's prefix (forward) or suffix (backward)
    fo = np.concatenate([[0], np.cumsum(plan["fwd_counts"])])
    bo = np.concatenate([[0], np.cumsum(plan["bwd_counts"])])
    for k in range(d):
        fs, bs = plan["fwd_slots"][k].astype(np.int64), plan["bwd_slots"][k].astype(np.int64)
        assert fs.max() < plan["fwd_counts"][k] and bs.max() < plan["bwd_counts"][k]
        rf = plan["fwd_ids"][fo[k] + fs]
        rb = plan["bwd_ids"][bo[k] + bs]
        assert np.array_equal(pos[rf, : k + 1], pos[:, : k + 1])
        assert np.array_equal(pos[rb, k:], pos[:, k:])
        assert np.all(rf <= np.arange(P)) and np.all(rb <= np.arange(P))  # the representative is the smallest id


def twin_random(xe, oracle, dims, ranks, seed=SEED):
    xe.seed(seed)
    oracle.seed(seed)
    return xe.TTTensor.random(dims, ranks), oracle.TT.random(dims, ranks)


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def run_adf(xe, x, m, iters, prec, max_ranks=None, no_stacks=False):
    adf = xe.ADFVariant(iters, 1e-8, 0.999)
    adf.precision = prec
    if no_stacks:
        os.environ["XTT_ADF_NO_STACKS"] = "1"
    try:
        res = adf(x, m, max_ranks) if max_ranks else adf(x, m)
    finally:
        os.environ.pop("XTT_ADF_NO_STACKS", None)
    return res, adf.last_iterations


@pytest.mark.parametrize("d,n,r,P", [(5, 6, 3, 2000), (8, 4, 5, 6000), (3, 40, 7, 3000)])
def test_stacked_adf_matches_oracle_and_chain_path(xe, oracle, d, n, r, P):
    """FP64: residual and tensor after 6 iterations against the oracle (1e-9) and against the chain-walking device path
    (1e-11: the same arithmetic but for the summation order of each sample's chain)."""
    print(f"stacked ADF d={d} n={n} r={r} P={P}", flush=True)
    truth, trutho = twin_random(xe, oracle, [n] * d, [r] * (d - 1))
    x, xo = twin_random(xe, oracle, [n] * d, [r] * (d - 1), seed=SEED + 3)
    x2 = x.copy()
    xe.seed(SEED + 1); oracle.seed(SEED + 1)
    m = xe.SinglePointMeasurementSet.random(P, [n] * d)
    mo = oracle.Measurements.random(P, [n] * d)
    m.measure(truth)
    mo.measure(trutho)
    res, it = run_adf(xe, x, m, 6, xe.XT_F64)
    res2, it2 = run_adf(xe, x2, m, 6, xe.XT_F64, no_stacks=True)
    reso, ito = oracle.adf_solve(xo, mo, max_iterations=6)
    print(f"  residual {res:.6e} chain {res2:.6e} oracle {reso:.6e}", flush=True)
    assert it == it2 == ito
    assert abs(res - reso) < 1e-9 * reso and abs(res - res2) < 1e-11 * res2
    assert rel(x.full(), xo.full()) < 1e-9
    assert rel(x.full(), x2.full()) < 1e-11


def test_stacked_adf_f32_follows_f64(xe, oracle):
    """FP32 stacks (the throughput configuration's precision) follow the FP64 run's residual trajectory."""
    d, n, r, P = 6, 8, 4, 20000
    truth, _ = twin_random(xe, oracle, [n] * d, [r] * (d - 1))
    x, _ = twin_random(xe, oracle, [n] * d, [r] * (d - 1), seed=SEED + 3)
    x64 = x.copy()
    xe.seed(SEED + 1)
    m = xe.SinglePointMeasurementSet.random(P, [n] * d)
    m.measure(truth)
    res32, _ = run_adf(xe, x, m, 10, xe.XT_F32)
    res64, _ = run_adf(xe, x64, m, 10, xe.XT_F64)
    print(f"  f32 {res32:.6e} f64 {res64:.6e}", flush=True)
    assert abs(res32 - res64) < 1e-4 * res64
    assert rel(x.full(), x64.full()) < 1e-4


def test_stacked_adf_rank_increase(xe, oracle):
    """From TTTensor::ones with maxRanks 3 (the rank-increase loop, adf.cpp:589-602): stacked path, chain path and
    oracle reach the same ranks, iteration count and residual."""
    D, N, R = 5, 4, 3
    xe.seed(SEED); oracle.seed(SEED)
    truth, trutho = xe.TTTensor.random([N] * D, [R] * (D - 1)), oracle.TT.random([N] * D, [R] * (D - 1))
    m = xe.SinglePointMeasurementSet.random(600, [N] * D)
    mo = oracle.Measurements.random(600, [N] * D)
    m.measure(truth)
    mo.measure(trutho)
    X, X2, Xo = xe.TTTensor.ones([N] * D), xe.TTTensor.ones([N] * D), oracle.TT.ones([N] * D)
    xe.seed(SEED + 9)
    res, it = run_adf(xe, X, m, 60, xe.XT_F64, max_ranks=[R] * (D - 1))
    xe.seed(SEED + 9)
    res2, it2 = run_adf(xe, X2, m, 60, xe.XT_F64, max_ranks=[R] * (D - 1), no_stacks=True)
    oracle.seed(SEED + 9)
    reso, ito = oracle.adf_solve(Xo, mo, max_ranks=[R] * (D - 1), max_iterations=60, target=1e-8, min_decrease=0.999)
    print(f"  residual {res:.6e} chain {res2:.6e} oracle {reso:.6e}; iterations {it} {it2} {ito}", flush=True)
    assert X.ranks() == X2.ranks() == Xo.ranks()
    assert it == it2 == ito
    # the residuals are relative to ‖b‖; near the 1e-8 target their differences are rounding of ‖b‖-sized terms
    assert abs(res - reso) < 1e-7 * reso + 1e-13 and abs(res - res2) < 1e-9 * res2 + 1e-15
