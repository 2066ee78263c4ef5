"""The lockstep batch driver (als_round_batch with threads=0): every step of every problem is one grouped launch.

It runs the same per-problem algorithm as the threaded driver (one-site ALS_SPD with FP64 local CG, or FP32 CG inside
FP64 iterative refinement, then round(eps)), so both drivers and the oracle are compared problem by problem on the
uneven batch of configuration 4 (d ~ U{6..20}, n ∈ {2,4,8,16}, r ~ U{4..64}, rank-3 Laplacian + potential, rank-4
right-hand sides).
"""
import numpy as np
import pytest

from helpers import SEED, laplace_cores
from test_gpu_reference_tests import round_parity

pytestmark = pytest.mark.gpu


def run_batch(xe, probs, threads, precision=None):
    als = xe.ALS_SPD.copy()
    als.convergenceEpsilon = 0.0
    als.numHalfSweeps = 2
    als.precision = xe.XT_F64 if precision is None else precision
    res, st = xe.als_round_batch(als, [p[0] for p in probs], [p[1] for p in probs], [p[2] for p in probs], 2, 1e-12,
                                 threads=threads)
    return np.asarray(res), st


def test_lockstep_small_matches_threads_bitwise(xe):
    """Small problems (d 5..8, n 2 or 4, ranks 3..9): the lockstep driver issues the same kernels on the same shapes as
    the threaded one (the grouped GEMM, Householder, Jacobi and CG kernels share the single-problem tile code), so the
    results agree to rounding: TTs within 1e-14, equal ranks, equal CG iteration counts."""
    print("lockstep small", flush=True)
    from xerus_amd.workloads import batch_problems
    kw = dict(d_range=(5, 8), n_choices=(2, 4), r_range=(3, 9))
    pa, pb = batch_problems(8, **kw), batch_problems(8, **kw)
    ra, sa = run_batch(xe, pa, 0)
    rb, sb = run_batch(xe, pb, 4)
    assert sa["cgIterations"] == sb["cgIterations"] and sa["localSolves"] == sb["localSolves"]
    for (A, x, b, shape), (_, y, _, _), r1, r2 in zip(pa, pb, ra, rb):
        assert x.ranks() == y.ranks(), shape
        assert (x - y).frob_norm() <= 1e-14 * y.frob_norm() + 1e-300, shape
        # the residuals come from different code paths (grouped sweep against tt_residual per problem): each is
        # exact to a few ε·(‖A x‖ + ‖b‖)/‖b‖
        assert abs(r1 - r2) <= 1e-6 * r2 + 1e-13, shape


def test_lockstep_residuals_match_tt_residual(xe):
    """The lockstep driver's residuals (w = A·x − b assembled core by core, grouped blocked R-only QR sweep, ‖b‖ from
    b's transfer matrices) against tt_residual on every problem of a 48-problem batch of configuration 4: widths up to
    3·64 + 4 = 196 columns take the blocked panels (64 columns each) and rows up to 196·16 the TSQR tree. Both are
    backward stable, so they agree to a few ε·(‖A x‖ + ‖b‖)/‖b‖ in absolute terms."""
    print("lockstep residuals", flush=True)
    from xerus_amd.workloads import batch_problems
    probs = batch_problems(48)
    res, _ = run_batch(xe, probs, 0)
    wide = 0
    for p, ((A, x, b, shape), r1) in enumerate(zip(probs, res)):
        r2 = xe.residual(A, x, b)
        wide += max(x.ranks()) * 3 + 4 > 64
        assert abs(r1 - r2) <= 1e-6 * r2 + 1e-13, (p, shape, r1, r2)
    assert wide > 0


def test_lockstep_matches_threads_full_shapes(xe, oracle):
    """24 problems of configuration 4 at full size through both drivers. The CG reductions differ in summation order
    (one workgroup per system against the chunked kernels), so trajectories agree to the local tolerance: TTs within
    1e-8 + 10·ρ (ρ the residual, which bounds how well a rank-r trajectory is defined), residuals within 10% of ρ.
    The rank lists of the two trajectories may differ where a singular value sits within the solvers' tolerance of the
    1e-12·s_0 cut; round of the same TT gives bit-exact ranks against the oracle (checked on every result)."""
    print("lockstep vs threads", flush=True)
    from xerus_amd.workloads import batch_problems
    pa, pb = batch_problems(24), batch_problems(24)
    ra, sa = run_batch(xe, pa, 0)
    rb, sb = run_batch(xe, pb, 8)
    assert sa["localSolves"] == sb["localSolves"]
    for p, ((A, x, b, shape), (_, y, _, _), r1, r2) in enumerate(zip(pa, pb, ra, rb)):
        diff = (x - y).frob_norm() / y.frob_norm()
        print(f"  problem {p} {shape}: residual {r1:.3e} (threads {r2:.3e}), rel diff {diff:.2e}", flush=True)
        assert diff < 1e-8 + 10 * r2, p
        assert abs(r1 - r2) <= 0.1 * r2 + 1e-11, p
        round_parity(xe, oracle, x, 1e-12)


def test_lockstep_config4_subset_matches_oracle(xe, oracle):
    """The first 16 problems of configuration 4 through the lockstep driver, each against the oracle's FP64 ALS_SPD
    (matrix-free CG to the same local tolerance) + round(1e-12) for 2 sweeps from the same start: TTs within
    1e-8 + 10·ρ and residuals within 10% of ρ (ρ the oracle's residual)."""
    print("lockstep config4 subset", flush=True)
    from xerus_amd.workloads import batch_problems
    count = 16
    probs = batch_problems(count)
    res, _ = run_batch(xe, probs, 0)
    for p, (A, x, b, (d, n, r)) in enumerate(probs):
        rng = np.random.default_rng(SEED + p)
        assert (int(rng.integers(6, 21)), int(rng.choice((2, 4, 8, 16))), int(rng.integers(4, 65))) == (d, n, r)
        pots = rng.random((d, n))
        Ao = oracle.TT.from_cores(laplace_cores(d, n, list(pots), rank3=True), op=True)
        oracle.seed(SEED + p)
        bo = oracle.TT.random([n] * d, [4] * (d - 1))
        xo = oracle.TT.random([n] * d, [r] * (d - 1))
        for _ in range(2):
            oracle.als(Ao, xo, bo, num_half_sweeps=2, eps=0.0, local_solver="cg", cg_tol=1e-13)
            xo.round(eps=1e-12)
        ro = oracle.residual(Ao, xo, bo)
        xot = xe.TTTensor.from_components(xo.cores())
        diff = (x - xot).frob_norm() / xot.frob_norm()
        print(f"  problem {p} d={d} n={n} r={r}: residual {res[p]:.3e} (oracle {ro:.3e}), rel diff {diff:.2e}", flush=True)
        assert diff < 1e-8 + 10 * ro, p
        assert abs(res[p] - ro) <= 0.1 * ro + 1e-11, p


def test_lockstep_mixed_small_matches_threads(xe):
    """FP32 CG inside FP64 refinement on small problems through both drivers. The FP32 reductions differ in summation
    order (one workgroup per system against the chunked kernels), so each refinement's correction differs at FP32
    rounding, which the refinement itself removes: TTs within 1e-9, equal ranks, residuals within 1e-6 relative."""
    print("lockstep mixed small", flush=True)
    from xerus_amd.workloads import batch_problems
    kw = dict(d_range=(5, 8), n_choices=(2, 4), r_range=(3, 9))
    pa, pb = batch_problems(8, **kw), batch_problems(8, **kw)
    ra, sa = run_batch(xe, pa, 0, xe.XT_F32)
    rb, sb = run_batch(xe, pb, 4, xe.XT_F32)
    assert sa["localSolves"] == sb["localSolves"]
    assert sa["refinements"] > 0
    for (A, x, b, shape), (_, y, _, _), r1, r2 in zip(pa, pb, ra, rb):
        assert x.ranks() == y.ranks(), shape
        assert (x - y).frob_norm() <= 1e-9 * y.frob_norm(), shape
        assert abs(r1 - r2) <= 1e-6 * r2 + 1e-13, shape


def test_lockstep_mixed_config4_subset_matches_oracle(xe, oracle):
    """Configuration 4 as specified (FP32 MFMA products with FP64 refinement): the first 12 problems through the
    lockstep driver against the oracle's FP64 ALS_SPD + round(1e-12), 2 sweeps from the same start. The refinement
    reaches the same local tolerance as the FP64 CG, so the bar is the FP64 one: TTs within 1e-8 + 10·ρ, residuals
    within 10% of ρ (ρ the oracle's residual)."""
    print("lockstep mixed config4 subset", flush=True)
    from xerus_amd.workloads import batch_problems
    count = 12
    probs = batch_problems(count)
    res, st = run_batch(xe, probs, 0, xe.XT_F32)
    assert st["refinements"] > 0
    for p, (A, x, b, (d, n, r)) in enumerate(probs):
        rng = np.random.default_rng(SEED + p)
        assert (int(rng.integers(6, 21)), int(rng.choice((2, 4, 8, 16))), int(rng.integers(4, 65))) == (d, n, r)
        pots = rng.random((d, n))
        Ao = oracle.TT.from_cores(laplace_cores(d, n, list(pots), rank3=True), op=True)
        oracle.seed(SEED + p)
        bo = oracle.TT.random([n] * d, [4] * (d - 1))
        xo = oracle.TT.random([n] * d, [r] * (d - 1))
        for _ in range(2):
            oracle.als(Ao, xo, bo, num_half_sweeps=2, eps=0.0, local_solver="cg", cg_tol=1e-13)
            xo.round(eps=1e-12)
        ro = oracle.residual(Ao, xo, bo)
        xot = xe.TTTensor.from_components(xo.cores())
        diff = (x - xot).frob_norm() / xot.frob_norm()
        print(f"  problem {p} d={d} n={n} r={r}: residual {res[p]:.3e} (oracle {ro:.3e}), rel diff {diff:.2e}", flush=True)
        assert diff < 1e-8 + 10 * ro, p
        assert abs(res[p] - ro) <= 0.1 * ro + 1e-11, p


def test_lockstep_energy_stop_matches_threads(xe):
    """ALS_SPD as preset (numHalfSweeps = 0: half-sweeps until the energy test with convergenceEpsilon stops each
    problem, so problems stop at different counts) on 16 small problems: both drivers run the same half-sweeps and
    local solves per batch and agree per problem to rounding (1e-12 relative, equal ranks)."""
    print("lockstep energy stop", flush=True)
    from xerus_amd.workloads import batch_problems
    kw = dict(d_range=(4, 9), n_choices=(2, 3, 4), r_range=(2, 12))
    pa, pb = batch_problems(16, **kw), batch_problems(16, **kw)
    als = xe.ALS_SPD.copy()
    assert als.numHalfSweeps == 0 and als.convergenceEpsilon > 0
    als.precision = xe.XT_F64
    ra, sa = xe.als_round_batch(als, [p[0] for p in pa], [p[1] for p in pa], [p[2] for p in pa], 2, 1e-12, threads=0)
    rb, sb = xe.als_round_batch(als, [p[0] for p in pb], [p[1] for p in pb], [p[2] for p in pb], 2, 1e-12, threads=4)
    print(f"  half-sweeps {sa['halfSweeps']} / {sb['halfSweeps']}, local solves {sa['localSolves']} / {sb['localSolves']}", flush=True)
    assert sa["halfSweeps"] == sb["halfSweeps"] and sa["localSolves"] == sb["localSolves"]
    assert sa["halfSweeps"] > 2 * 2 * 16  # some problem ran more than the two half-sweeps of a fixed count
    for (A, x, b, shape), (_, y, _, _), r1, r2 in zip(pa, pb, ra, rb):
        assert x.ranks() == y.ranks(), shape
        assert (x - y).frob_norm() <= 1e-12 * y.frob_norm() + 1e-300, shape
        assert abs(r1 - r2) <= 1e-6 * r2 + 1e-13, shape


def test_lockstep_short_trains_match_threads(xe):
    """Orders 2 and 3 with n in {2, 3, 7} and ranks up to 30 (clipped to the maximal ranks n^k): the boundary cores
    that prepare_x_for_als absorbs, stacks of length one, round with one or two edges and both branches of
    round_edge. Both drivers agree per problem to 1e-12 with equal ranks."""
    print("lockstep short trains", flush=True)
    from xerus_amd.workloads import batch_problems
    kw = dict(d_range=(2, 3), n_choices=(2, 3, 7), r_range=(1, 30))
    pa, pb = batch_problems(24, **kw), batch_problems(24, **kw)
    ra, sa = run_batch(xe, pa, 0)
    rb, sb = run_batch(xe, pb, 4)
    assert sa["localSolves"] == sb["localSolves"]
    for (A, x, b, shape), (_, y, _, _), r1, r2 in zip(pa, pb, ra, rb):
        assert x.ranks() == y.ranks(), shape
        assert (x - y).frob_norm() <= 1e-12 * y.frob_norm() + 1e-300, shape
        assert abs(r1 - r2) <= 1e-6 * r2 + 1e-13, shape


def test_lockstep_edge_batches(xe):
    """An empty batch, a batch of one problem and a batch without residuals: the lockstep driver returns what the
    threaded one does (the single problem to 1e-12, the same solver counters)."""
    print("lockstep edge batches", flush=True)
    from xerus_amd.workloads import batch_problems
    als = xe.ALS_SPD.copy()
    als.convergenceEpsilon = 0.0
    als.numHalfSweeps = 2
    als.precision = xe.XT_F64
    res, st = xe.als_round_batch(als, [], [], [], 2, 1e-12, threads=0)
    assert len(res) == 0 and st["localSolves"] == 0
    kw = dict(d_range=(8, 8), n_choices=(4,), r_range=(16, 16))
    pa, pb = batch_problems(1, **kw), batch_problems(1, **kw)
    ra, sa = run_batch(xe, pa, 0)
    rb, sb = run_batch(xe, pb, 1)
    assert sa["localSolves"] == sb["localSolves"] and sa["cgIterations"] == sb["cgIterations"]
    (_, x, _, _), (_, y, _, _) = pa[0], pb[0]
    assert x.ranks() == y.ranks()
    assert (x - y).frob_norm() <= 1e-12 * y.frob_norm()
    assert abs(ra[0] - rb[0]) <= 1e-6 * rb[0] + 1e-13
    pc = batch_problems(3, **kw)
    rc, _ = xe.als_round_batch(als, [p[0] for p in pc], [p[1] for p in pc], [p[2] for p in pc], 2, 1e-12, threads=0,
                               residuals=False)
    assert rc is None
    for (A, x, b, _) in pc:
        assert xe.residual(A, x, b) < 1e-6


def test_lockstep_falls_back_for_other_options(xe):
    """Options outside the lockstep path (here BF16 local matvecs) run on worker threads with the same results as an
    explicit thread count."""
    print("lockstep fallback", flush=True)
    from xerus_amd.workloads import batch_problems
    kw = dict(d_range=(5, 7), n_choices=(2, 4), r_range=(3, 6))
    pa, pb = batch_problems(4, **kw), batch_problems(4, **kw)
    als = xe.ALS_SPD.copy()
    als.convergenceEpsilon = 0.0
    als.numHalfSweeps = 2
    als.precision = xe.XT_BF16
    ra, _ = xe.als_round_batch(als, [p[0] for p in pa], [p[1] for p in pa], [p[2] for p in pa], 1, 1e-12, threads=0)
    rb, _ = xe.als_round_batch(als, [p[0] for p in pb], [p[1] for p in pb], [p[2] for p in pb], 1, 1e-12, threads=2)
    for (A, x, b, shape), (_, y, _, _), r1, r2 in zip(pa, pb, ra, rb):
        assert x.ranks() == y.ranks(), shape
        assert (x - y).frob_norm() <= 1e-10 * y.frob_norm(), shape
