"""GPU parity of the TT path (TTTensor/TTOperator, move_core, round, ALS/DMRG, ADF) against the oracle.

Integer results (ranks, core positions, measurement index sets, RNG draws) must be bit-exact; floating point
results are compared through gauge-invariant quantities (full tensors, norms, energies, residuals) at the
tolerance written in each test.
"""
import numpy as np
import pytest

from helpers import SEED, apply_dense_laplace, exact_laplace_solve, laplace_cores

pytestmark = pytest.mark.gpu


def twin_random(xe, oracle, dims, ranks, op=False, seed=SEED):
    xe.seed(seed)
    oracle.seed(seed)
    cls = xe.TTOperator if op else xe.TTTensor
    return cls.random(dims, ranks), oracle.TT.random(dims, ranks, op=op)


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def test_rng_streams_match(xe, oracle):
    print("rng draws", flush=True)
    xe.seed(SEED); oracle.seed(SEED)
    assert np.array_equal(xe.normal_draws(1000), oracle.normal_draws(1000))
    assert np.array_equal(xe.uniform_draws(1000, 9), oracle.uniform_draws(1000, 9))


@pytest.mark.parametrize("dims,ranks", [([4] * 8, [8] * 7), ([16] * 16, [64] * 15), ([2, 3, 4, 5], [3, 10, 2]), ([5], [])])
def test_random_ranks_and_values(xe, oracle, dims, ranks):
    print(f"random {dims} {ranks}", flush=True)
    x, xo = twin_random(xe, oracle, dims, ranks)
    assert x.ranks() == xo.ranks()
    assert x.canonicalized and x.corePosition == 0
    import math
    if math.prod(dims) <= 65536:
        assert rel(x.full(), xo.full()) < 1e-12
    else:
        assert abs(x.frob_norm() - xo.frob_norm()) < 1e-10 * xo.frob_norm()


@pytest.mark.parametrize("keep", [True, False])
def test_move_core(xe, oracle, keep):
    print(f"move_core keep={keep}", flush=True)
    x, xo = twin_random(xe, oracle, [3, 4, 5, 4, 3], [3, 7, 7, 3])
    ref = xo.full()
    for pos in (4, 1, 3, 0, 2):
        x.move_core(pos, keep)
        xo.move_core(pos, keep)
        assert x.ranks() == xo.ranks()
        assert x.corePosition == pos and x.canonicalized
        assert rel(x.full(), ref) < 1e-13
        for k in range(x.order()):  # left- / right-orthogonality of the non-core components
            c = x.get_component(k)
            if k < pos:
                M = c.reshape(-1, c.shape[-1]); assert np.abs(M.T @ M - np.eye(M.shape[1])).max() < 1e-12
            elif k > pos:
                M = c.reshape(c.shape[0], -1); assert np.abs(M @ M.T - np.eye(M.shape[0])).max() < 1e-12


def test_round_eps_and_maxrank(xe, oracle):
    print("round", flush=True)
    # a TT whose sum with itself has exactly half the rank
    a, ao = twin_random(xe, oracle, [4, 4, 4, 4, 4], [4, 8, 8, 4])
    s, so = a + a, ao + ao
    for t, to in ((s, so),):
        t.round(1e-12)
        to.round(eps=1e-12)
        assert t.ranks() == to.ranks()
        assert rel(t.full(), to.full()) < 1e-12
    b, bo = twin_random(xe, oracle, [4, 4, 4, 4, 4], [4, 16, 16, 4])
    b.round(5)
    bo.round(5)
    assert b.ranks() == bo.ranks() == [4, 5, 5, 4]
    assert rel(b.full(), bo.full()) < 1e-10


def test_round_speculation_matches_normal(xe, monkeypatch):
    """A round after one that kept every rank on the same shapes runs with its rank decisions assumed and checked
    once at the end; a tensor of the same ranks whose round does truncate makes the check fail and the round rerun.
    Every result (cores, ranks, core position) must equal the same rounds with speculation off."""
    print("round speculation", flush=True)
    dims, ranks = [4] * 8, [6] * 7
    xe.seed(SEED)
    x1 = xe.TTTensor.random(dims, ranks)                 # generic: round(1e-12) keeps its ranks
    y = xe.TTTensor.random(dims, [2, 3, 3, 3, 3, 3, 2])
    z = xe.TTTensor.random(dims, [2, 3, 3, 3, 3, 3, 2])
    x2 = y + z * 1e-14   # x1's ranks [4, 6, …, 6, 4] (z is above the QC rule's 16ε), but round(1e-12) drops z
    assert x2.ranks() == x1.ranks() and x2.corePosition == x1.corePosition

    def sequence():
        out = []
        for src in (x1, x1, x2, x1):
            z = src.copy()
            z.round(1e-12)
            out.append((z.ranks(), z.corePosition, [z.get_component(k) for k in range(z.order())]))
        return out

    monkeypatch.setenv("XTT_NO_SPECULATION", "1")
    ref = sequence()
    monkeypatch.delenv("XTT_NO_SPECULATION")
    got = sequence()
    assert ref[2][0] == [2, 3, 3, 3, 3, 3, 2]           # the third round truncates: its speculation fails
    for (r1, c1, k1), (r2, c2, k2) in zip(ref, got):
        assert r1 == r2 and c1 == c2
        for a, b in zip(k1, k2):
            assert np.array_equal(a, b)


def test_tt_sum_and_difference(xe, oracle):
    """TT:real_difference pins (ttArithmetic.cxx:74-115): sum ranks {8,8,8,8}; a − a ≈ 0."""
    print("tt sum", flush=True)
    a, _ = twin_random(xe, oracle, [10] * 5, [4] * 4)
    b, _ = twin_random(xe, oracle, [10] * 5, [4] * 4, seed=SEED + 1)
    c = a + b
    assert c.ranks() == [8, 8, 8, 8]
    assert (a - a).frob_norm() < 1e-11
    assert ((a + b) - (b + a)).frob_norm() < 1e-11
    assert ((73 * a + b) - (b + 73 * a)).frob_norm() < 1e-9


def test_operator_apply_and_residual(xe, oracle):
    print("apply/residual", flush=True)
    d, n = 5, 4
    rng = np.random.default_rng(0)
    pots = [rng.random(n) for _ in range(d)]
    cores = laplace_cores(d, n, pots)
    A, Ao = xe.TTOperator.from_components(cores), oracle.TT.from_cores(cores, op=True)
    x, xo = twin_random(xe, oracle, [n] * d, [3] * (d - 1))
    y = A(x)
    ref = apply_dense_laplace(xo.full(), d, n, pots)
    assert rel(y.full(), ref) < 1e-13
    b, bo = twin_random(xe, oracle, [n] * d, [2] * (d - 1), seed=SEED + 5)
    r = xe.residual(A, x, b)
    ro = oracle.residual(Ao, xo, bo)
    dense = np.linalg.norm(ref - bo.full()) / np.linalg.norm(bo.full())
    assert abs(r - ro) < 1e-12 * ro and abs(r - dense) < 1e-12 * dense


@pytest.mark.parametrize("n,rx,rb", [(6, 24, 4), (8, 40, 5), (5, 25, 3)])
def test_residual_high_rank(xe, oracle, n, rx, rb):
    """The stable residual when the TT of Ax − b has ranks above 64 (3·rx + rb columns): the blocked TSQR
    R-only path (64-column panels) against the oracle's Householder residual and a dense evaluation."""
    print(f"residual high rank n={n} rx={rx} rb={rb}", flush=True)
    d = 4
    rng = np.random.default_rng(n)
    pots = [rng.random(n) for _ in range(d)]
    cores = laplace_cores(d, n, pots, rank3=True)
    A, Ao = xe.TTOperator.from_components(cores), oracle.TT.from_cores(cores, op=True)
    x, xo = twin_random(xe, oracle, [n] * d, [rx] * (d - 1), seed=SEED + n)
    b, bo = twin_random(xe, oracle, [n] * d, [rb] * (d - 1), seed=SEED + 2 * n)
    r, ro = xe.residual(A, x, b), oracle.residual(Ao, xo, bo)
    dense = np.linalg.norm(apply_dense_laplace(xo.full(), d, n, pots) - bo.full()) / np.linalg.norm(bo.full())
    assert abs(r - ro) < 1e-11 * ro and abs(r - dense) < 1e-11 * dense


def config1(xe, oracle, d=8, n=4):
    """d=8, n=4 Laplacian (rank 2), b random rank 2 normalised, x0 random rank 8 — both engines, same draws."""
    cores = laplace_cores(d, n)
    A, Ao = xe.TTOperator.from_components(cores), oracle.TT.from_cores(cores, op=True)
    xe.seed(SEED); oracle.seed(SEED)
    b, bo = xe.TTTensor.random([n] * d, [2] * (d - 1)), oracle.TT.random([n] * d, [2] * (d - 1))
    nb = bo.frob_norm()
    b = b / nb
    bo.scale(1.0 / nb)
    x, xo = xe.TTTensor.random([n] * d, [8] * (d - 1)), oracle.TT.random([n] * d, [8] * (d - 1))
    return A, Ao, b, bo, x, xo


def test_als_spd_config1(xe, oracle):
    """ALS_SPD, 5 sweeps (numHalfSweeps=10, eps=0): energy, residual, solution and ranks vs the oracle's
    dense-Cholesky ALS; the residual also checked against the exact (fast-diagonalisation) solution."""
    print("als_spd config1", flush=True)
    A, Ao, b, bo, x, xo = config1(xe, oracle)
    assert x.ranks() == xo.ranks() == [4, 8, 8, 8, 8, 8, 4]
    als = xe.ALS_SPD.copy()
    als.convergenceEpsilon = 0.0
    als.localTolerance = 1e-14
    e = als(A, x, b, 10)
    eo, st = oracle.als(Ao, xo, bo, num_half_sweeps=10, eps=0.0)
    assert als.last_stats["halfSweeps"] == st["half_sweeps"] == 10
    assert als.last_stats["localSolves"] == st["local_solves"] == 51
    assert abs(e - eo) < 1e-12 * abs(eo)
    assert x.ranks() == xo.ranks() and x.corePosition == xo.state()[1]
    assert rel(x.full(), xo.full()) < 1e-9
    r, ro = xe.residual(A, x, b), oracle.residual(Ao, xo, bo)
    assert abs(r - ro) < 1e-6 * ro
    xs = exact_laplace_solve(bo.full(), 8, 4)
    dense = np.linalg.norm(apply_dense_laplace(x.full(), 8, 4) - bo.full()) / np.linalg.norm(bo.full())
    assert abs(dense - r) < 1e-9 * dense
    assert rel(x.full(), xs) < 1e-4  # rank-8 ALS is within its approximation error of the exact solution


@pytest.mark.parametrize("prec,tol", [("f32", 1e-9), ("bf16", 1e-9)])
def test_als_mixed_precision_config3(xe, oracle, prec, tol):
    """Config-1 seed and sizes with FP32 / BF16 local matvecs and FP64 refinement: same fixed point as FP64."""
    print(f"als mixed {prec}", flush=True)
    A, Ao, b, bo, x, xo = config1(xe, oracle)
    als = xe.ALS_SPD.copy()
    als.convergenceEpsilon = 0.0
    als.localTolerance = 1e-13
    als.precision = xe.XT_F32 if prec == "f32" else xe.XT_BF16
    als(A, x, b, 10)
    oracle.als(Ao, xo, bo, num_half_sweeps=10, eps=0.0)
    err = rel(x.full(), xo.full())
    print(f"  ‖x − x_oracle‖/‖x_oracle‖ = {err:.3e}, refinements {als.last_stats['refinements']}", flush=True)
    assert err < tol


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_als_speculative_calls_match_normal(xe, oracle, prec, monkeypatch):
    """An ALS call after one that needed no inner CG iteration (same shapes) first runs with every local solve taken
    as converged and its residual tests on the device; if any test fails, x is restored and the call reruns. Calls
    with b (speculation holds) and then with a changed b (speculation fails, rerun) must give exactly the result,
    energies and statistics of the same calls with speculation off."""
    print(f"als speculation {prec}", flush=True)
    A, Ao, b, bo, x, xo = config1(xe, oracle)
    als = xe.ALS_SPD.copy()
    als.convergenceEpsilon = 0.0
    als.localTolerance = 1e-9
    als.precision = xe.XT_F64 if prec == "f64" else xe.XT_F32
    for _ in range(40):  # converge until a call needs no inner iteration
        als(A, x, b, 2)
        if als.last_stats["cgIterations"] == 0:
            break
    assert als.last_stats["cgIterations"] == 0
    b2 = b * 1.001

    def sequence():
        y = x.copy()
        out = []
        for rhs in (b, b, b2, b2):
            e = als(A, y, rhs, 2)
            out.append((e, y.full(), y.ranks(), y.corePosition, dict(als.last_stats)))
        return out

    monkeypatch.setenv("XTT_NO_SPECULATION", "1")
    ref = sequence()
    monkeypatch.delenv("XTT_NO_SPECULATION")
    got = sequence()
    assert ref[2][4]["cgIterations"] > 0  # the changed b needed iterations: the third call's speculation failed
    for (e1, f1, r1, c1, s1), (e2, f2, r2, c2, s2) in zip(ref, got):
        assert e1 == e2 and r1 == r2 and c1 == c2 and s1 == s2
        assert np.array_equal(f1, f2)


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("shape", ["config1", "d16n16r64"])
def test_als_fused_cg_matches_chunked(xe, oracle, prec, shape, monkeypatch):
    """The fused local CG (one grid-synchronised launch per solve, Chronopoulos–Gear recurrence) and the chunked
    standard CG (16-iteration graph chunks) solve the same local systems to the same tolerance: the ALS results
    agree to that tolerance and the inner iteration counts are close."""
    print(f"als fused cg {prec} {shape}", flush=True)
    if shape == "config1":
        A, _, b, _, x, _ = config1(xe, oracle)
        sweeps, tol, bound = 10, 1e-12, 1e-9
    else:
        from xerus_amd.workloads import laplace_cores
        d, n = 16, 16
        pots = np.random.default_rng(SEED).random((d, n))
        A = xe.TTOperator.from_components(laplace_cores(d, n, list(pots), rank3=True))
        xe.seed(SEED)
        b = xe.TTTensor.random([n] * d, [4] * (d - 1))
        x = xe.TTTensor.random([n] * d, [64] * (d - 1))
        sweeps, tol, bound = 2, 1e-10, 1e-7
    als = xe.ALS_SPD.copy()
    als.convergenceEpsilon = 0.0
    als.localTolerance = tol
    als.precision = xe.XT_F64 if prec == "f64" else xe.XT_F32

    def solve():
        y = x.copy()
        e = als(A, y, b, sweeps)
        return e, y, dict(als.last_stats)

    monkeypatch.setenv("XTT_NO_FUSED_CG", "1")
    e1, y1, s1 = solve()
    monkeypatch.delenv("XTT_NO_FUSED_CG")
    e2, y2, s2 = solve()
    e3, y3, s3 = solve()
    assert e2 == e3 and s2 == s3  # deterministic
    assert all(np.array_equal(y2.get_component(k), y3.get_component(k)) for k in range(y2.order()))
    diff = (y1 - y2).frob_norm() / y1.frob_norm()
    print(f"  rel diff {diff:.3e}, energies {e1:.15e} {e2:.15e}, cg iterations {s1['cgIterations']} {s2['cgIterations']}",
          flush=True)
    assert s1["localSolves"] == s2["localSolves"]
    assert diff < bound
    assert abs(e1 - e2) <= 1e-9 * abs(e1)
    assert abs(s1["cgIterations"] - s2["cgIterations"]) <= 0.1 * s1["cgIterations"] + 2


def test_als_identity(xe, oracle):
    """ALS:identity (als.cxx:28-67): with A = I the solution is b."""
    print("als identity", flush=True)
    d, n = 3, 10
    I = [np.eye(n).reshape(1, n, n, 1)] * d
    A = xe.TTOperator.from_components(I)
    xe.seed(SEED)
    b = xe.TTTensor.random([n] * d, [4, 4])
    x = xe.TTTensor.random([n] * d, [4, 4])
    als = xe.ALS_SPD.copy()
    als.useResidualForEndCriterion = True
    res = als(A, x, b, 1e-4)
    assert res < 0.01
    assert (x - b).frob_norm() < 1e-9


@pytest.mark.parametrize("variant", ["DMRG_SPD", "ALS", "DMRG"])
def test_other_variants(xe, oracle, variant):
    print(f"variant {variant}", flush=True)
    d, n = 5, 3
    rng = np.random.default_rng(7)
    pots = [rng.random(n) for _ in range(d)]
    cores = laplace_cores(d, n, pots)
    A, Ao = xe.TTOperator.from_components(cores), oracle.TT.from_cores(cores, op=True)
    xe.seed(SEED); oracle.seed(SEED)
    b, bo = xe.TTTensor.random([n] * d, [2] * (d - 1)), oracle.TT.random([n] * d, [2] * (d - 1))
    x, xo = xe.TTTensor.random([n] * d, [4] * (d - 1)), oracle.TT.random([n] * d, [4] * (d - 1))
    als = getattr(xe, variant).copy()
    als.convergenceEpsilon = 0.0
    als.localTolerance = 1e-14
    e = als(A, x, b, 4)
    eo, st = oracle.als(Ao, xo, bo, sites=als.sites, spd=als.assumeSPD, num_half_sweeps=4, eps=0.0)
    assert x.ranks() == xo.ranks()
    assert abs(e - eo) < 1e-8 * max(abs(eo), 1e-12)
    assert rel(x.full(), xo.full()) < 1e-7


def test_measurement_sets_bit_exact(xe, oracle):
    print("measurements", flush=True)
    dims = [10] * 10
    xe.seed(SEED); oracle.seed(SEED)
    m = xe.SinglePointMeasurementSet.random(20000, dims)
    mo = oracle.Measurements.random(20000, dims)
    assert np.array_equal(m.positions(10), mo.positions(10))


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_adf_gradient(xe, oracle, prec):
    """The adjoint kernel's full gradient vs the oracle's xerus projected gradient (adf.cpp:314-396)."""
    print(f"adf gradient {prec}", flush=True)
    d, n, r = 6, 5, 4
    truth, trutho = twin_random(xe, oracle, [n] * d, [r] * (d - 1))
    x, xo = twin_random(xe, oracle, [n] * d, [r] * (d - 1), seed=SEED + 9)
    xe.seed(SEED); oracle.seed(SEED)
    m = xe.SinglePointMeasurementSet.random(3000, [n] * d)
    mo = oracle.Measurements.random(3000, [n] * d)
    m.measure(truth)
    mo.measure(trutho)
    assert np.abs(m.measuredValues - mo.values()).max() < 1e-12
    g, rn = xe.adf_gradient(x, m, xe.XT_F64 if prec == "f64" else xe.XT_F32)
    go, rno = oracle.adf_gradient(xo, mo)
    tol = 1e-11 if prec == "f64" else 2e-5
    assert abs(rn - rno) < tol * rno
    for k in range(d):
        assert np.abs(g[k] - go[k]).max() < tol * np.abs(go[k]).max()


@pytest.mark.parametrize("d,n,r,P", [(10, 10, 8, 20000), (6, 4, 12, 2000), (16, 3, 8, 5000)])
def test_adf_gradient_f32_shapes(xe, oracle, d, n, r, P):
    """FP32 adjoint kernels at the ADF benchmark's shape (d=10, n=10, r=8), above the rank-8 fast path (r=12) and
    at the largest order of the fast path (d=16), against the FP64 oracle; repeated calls are bit-identical."""
    print(f"adf gradient f32 d={d} n={n} r={r} P={P}", flush=True)
    truth, trutho = twin_random(xe, oracle, [n] * d, [r] * (d - 1))
    x, xo = twin_random(xe, oracle, [n] * d, [r] * (d - 1), seed=SEED + 4)
    xe.seed(SEED); oracle.seed(SEED)
    m = xe.SinglePointMeasurementSet.random(P, [n] * d)
    mo = oracle.Measurements.random(P, [n] * d)
    m.measure(truth)
    mo.measure(trutho)
    g, rn = xe.adf_gradient(x, m, xe.XT_F32)
    g2, rn2 = xe.adf_gradient(x, m, xe.XT_F32)
    go, rno = oracle.adf_gradient(xo, mo)
    assert abs(rn - rno) < 1e-5 * rno
    for k in range(d):
        assert np.array_equal(g[k], g2[k])
        assert np.abs(g[k] - go[k]).max() < 3e-5 * max(np.abs(go[k]).max(), 1e-30)


def test_adf_solve(xe, oracle):
    print("adf solve", flush=True)
    d, n, r = 5, 6, 3
    truth, trutho = twin_random(xe, oracle, [n] * d, [r] * (d - 1))
    x, xo = twin_random(xe, oracle, [n] * d, [r] * (d - 1), seed=SEED + 3)
    xe.seed(SEED + 1); oracle.seed(SEED + 1)
    m = xe.SinglePointMeasurementSet.random(2000, [n] * d)
    mo = oracle.Measurements.random(2000, [n] * d)
    m.measure(truth)
    mo.measure(trutho)
    adf = xe.ADFVariant(8, 1e-8, 0.999)
    res = adf(x, m)
    reso, ito = oracle.adf_solve(xo, mo, max_iterations=8)
    assert adf.last_iterations == ito
    assert abs(res - reso) < 1e-6 * reso
    assert rel(x.full(), xo.full()) < 1e-6



def test_performance_data(xe, oracle):
    """PerformanceData points where xerus records them (als.cpp:426-480,515-526; adf.cpp:514): one after the initial
    energy, one per local step with flags 1/3 at half-sweep ends, residuals equal to the stable TT residual."""
    print("performance data", flush=True)
    d, n = 6, 4
    from helpers import laplace_cores
    A = xe.TTOperator.from_components(laplace_cores(d, n))
    xe.seed(SEED)
    b = xe.TTTensor.random([n] * d, [2] * (d - 1))
    x = xe.TTTensor.random([n] * d, [3] * (d - 1))
    perf = xe.PerformanceData(errorFunction=lambda t: xe.frob_norm(t))
    als = xe.ALS_SPD.copy()
    als.convergenceEpsilon = 0.0
    als(A, x, b, 2, perf)
    flags = [p.flags for p in perf.data]
    # optimizedRange [1, 5) for d=6, n=4, ranks 3 ≤ 4: 4 steps per half sweep
    assert flags[0] == 3 and flags.count(1) == 1 and flags[-1] == 3
    assert [p.iterationCount for p in perf.data] == list(range(len(perf.data)))
    assert all(b2.elapsedTime >= a2.elapsedTime for a2, b2 in zip(perf.data, perf.data[1:]))
    assert abs(perf.data[-1].residual - xe.residual(A, x, b)) < 1e-12
    assert abs(perf.data[-1].error - xe.frob_norm(x)) < 1e-12 * xe.frob_norm(x)
    assert perf.data[-1].ranks == x.ranks()
    # ADF: one point per iteration
    truth = xe.TTTensor.random([n] * d, [2] * (d - 1))
    m = xe.SinglePointMeasurementSet.random(1500, [n] * d)
    m.measure(truth)
    y = xe.TTTensor.random([n] * d, [2] * (d - 1))
    pa = xe.PerformanceData()
    adf = xe.ADFVariant(5, 1e-12, 0.999999)
    res = adf(y, m, pa)
    assert [p.iterationCount for p in pa.data] == list(range(len(pa.data)))
    assert abs(pa.data[-1].residual - res) < 1e-12 * max(res, 1e-30)


def test_batch_matches_sequential(xe):
    """Problems run by the batched driver (worker threads, one stream each) give the same TTs, ranks and
    residuals as the same problems run one after the other on the library stream. The workers use the chunked
    local CG and the library stream the fused one, so the results agree to the local tolerance, not bitwise."""
    print("batch vs sequential", flush=True)
    from xerus_amd.workloads import batch_problems
    probs = batch_problems(6, d_range=(5, 8), n_choices=(2, 4), r_range=(3, 9))
    twins = batch_problems(6, d_range=(5, 8), n_choices=(2, 4), r_range=(3, 9))
    als = xe.ALS_SPD.copy()
    als.convergenceEpsilon = 0.0
    als.numHalfSweeps = 2
    als.precision = xe.XT_F32
    res, stats = xe.als_round_batch(als, [p[0] for p in probs], [p[1] for p in probs], [p[2] for p in probs], 2, 1e-12,
                                    threads=3)
    assert stats["localSolves"] > 0
    for (A, x, b, shape), (A2, x2, b2, _), r in zip(probs, twins, res):
        for _ in range(2):
            als(A2, x2, b2, 2)
            x2.round(1e-12)
        assert x.ranks() == x2.ranks(), shape
        assert rel(x.full(), x2.full()) < 1e-10, shape
        # both stop their local solves somewhere below the FP64 bound, the FP32 inner solves only as far as that bound
        # needs, so residuals at round-off level differ by a few 1e-14
        assert abs(r - xe.residual(A2, x2, b2)) < 1e-9 * r + 5e-14
