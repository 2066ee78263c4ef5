"""Pins of the oracle's remaining paths (CPU): the TT constructors, non-SPD ALS, the approximation ALS (A = None),
DMRG, and the single-point ADF solve with its rank increase.

Each pin is a closed form, an exactly known solution or an invariant that a plausible mistake breaks:
- TT-SVD: exact reconstruction, unfolding ranks and the truncation rule on a tensor of known spectrum;
- non-SPD ALS / DMRG: a nonsymmetric system at full rank has the dense solution np.linalg.solve; ALS:real
  (src/unitTests/als.cxx:70-86) recovers a known rank-3 solution;
- approximation ALS: ALS:projectionALS (als.cxx:88-103), every sweep beats plain truncation;
- DMRG: at full rank the Laplacian's closed-form solution, with A = I the right-hand side, ranks within the target;
- ADF: every slice step of update_x is the exact line-search minimiser (the new gradient slice is orthogonal to the
  old one), and Algorithm:adf_random_low_rank (ttCompletion.cxx:107-130) from TTTensor::ones with rank increases.
"""
import numpy as np
import pytest

from helpers import SEED, apply_dense_laplace, exact_laplace_solve, laplace_cores

EPS = 8 * np.finfo(float).eps


def unfold_rank(T, k, tol=1e-10):
    M = T.reshape(int(np.prod(T.shape[: k + 1])), -1)
    s = np.linalg.svd(M, compute_uv=False)
    return int((s > tol * s[0]).sum())


# ------------------------------------------------------------------ TT constructors


def test_from_tensor_exact_and_unfolding_ranks(oracle):
    """TTNetwork(Tensor) with eps = 8ε reproduces the tensor, and its ranks are the unfolding ranks (TT-SVD)."""
    rng = np.random.default_rng(SEED)
    G = [rng.standard_normal(s) for s in [(1, 3, 2), (2, 4, 3), (3, 5, 2), (2, 3, 1)]]
    T = np.einsum("aib,bjc,ckd,dle->ijkl", *G)
    t = oracle.TT.from_tensor(T)
    assert t.ranks() == [unfold_rank(T, k) for k in range(3)] == [2, 3, 2]
    assert np.abs(t.full() - T).max() < 1e-12 * np.abs(T).max()
    assert t.state() == (True, 0)
    # every component but the first is right-orthogonal (the SVD's Vᵀ)
    for c in t.cores()[1:]:
        M = c.reshape(c.shape[0], -1)
        assert np.abs(M @ M.T - np.eye(M.shape[0])).max() < 1e-12
    # full-rank random tensor: ranks are the maximal ones
    R = rng.standard_normal((3, 4, 5, 2))
    assert oracle.TT.from_tensor(R).ranks() == [3, 10, 2]


def test_from_tensor_truncation(oracle):
    """Spectrum 1, 1e-3, 1e-9 on the middle unfolding: eps 1e-6 keeps two values, eps 1e-10 three, maxRank 1 one;
    the truncated TT is the best approximation of that rank (error = the dropped singular values)."""
    rng = np.random.default_rng(1)
    U = np.linalg.qr(rng.standard_normal((16, 3)))[0]
    V = np.linalg.qr(rng.standard_normal((9, 3)))[0]
    s = np.array([1.0, 1e-3, 1e-9])
    T = (U * s @ V.T).reshape(4, 4, 3, 3)
    for eps, r in ((1e-6, 2), (1e-10, 3)):
        t = oracle.TT.from_tensor(T, eps=eps)
        assert t.ranks()[1] == r
    t = oracle.TT.from_tensor(T, eps=EPS, max_ranks=[4, 1, 3])
    assert t.ranks()[1] == 1
    assert abs(np.linalg.norm(t.full() - T) - np.linalg.norm(s[1:])) < 1e-9


def test_operator_from_tensor_and_transpose(oracle):
    """An operator tensor (n…, m…) splits into (n_k, m_k) components; identity ⊗ identity has ranks 1; transpose
    swaps rows and columns of the dense operator."""
    I = np.eye(12).reshape(3, 4, 3, 4)
    t = oracle.TT.from_tensor(I, op=True)
    assert t.ranks() == [1]
    assert np.abs(t.full() - I).max() < 1e-14
    rng = np.random.default_rng(2)
    A = rng.standard_normal((2, 3, 4, 2, 3, 4))
    t = oracle.TT.from_tensor(A, op=True)
    assert t.ranks() == [4, 16]
    assert np.abs(t.full() - A).max() < 1e-12
    t.transpose()
    assert np.abs(t.full().reshape(24, 24) - A.reshape(24, 24).T).max() < 1e-12


def test_ones_identity_kronecker_dirac(oracle):
    """Closed forms; ones, identity and kronecker end with canonicalize_left (a QC sweep), so they hold to rounding."""
    dims = [2, 3, 4]
    assert np.abs(oracle.TT.ones(dims).full() - np.ones(dims)).max() < 1e-14
    Iop = oracle.TT.identity([2, 3, 4, 2, 3, 4])
    assert np.abs(Iop.full().reshape(24, 24) - np.eye(24)).max() < 1e-14
    k = oracle.TT.kronecker([3, 4, 5]).full()
    ref = np.zeros((3, 4, 5))
    for i in range(3):
        ref[i, i, i] = 1
    assert np.abs(k - ref).max() < 1e-14 and oracle.TT.kronecker([3, 4, 5]).ranks() == [3, 3]
    e = oracle.TT.dirac([3, 4, 5], [2, 1, 3])
    ref = np.zeros((3, 4, 5)); ref[2, 1, 3] = 1
    assert np.array_equal(e.full(), ref) and e.state()[0] is False
    o = oracle.TT.ones([2, 3, 2, 3], op=True)
    assert np.abs(o.full() - np.ones((2, 3, 2, 3))).max() < 1e-14


# ------------------------------------------------------------------ ALS variants


def quick_start_problem(O):
    """Tutorials:quick_start (tutorials.cxx:6-46): 512² · tridiag(-1,2,-1) as a QTT operator on 9 binary modes,
    b = ones, x random rank 3."""
    n = 512
    L = (2 * np.eye(n) - np.eye(n, k=1) - np.eye(n, k=-1)) * (n * n)
    A = O.TT.from_tensor(L.reshape([2] * 18), op=True)
    b = O.TT.from_tensor(np.ones([2] * 9))
    x = O.TT.random([2] * 9, [3] * 8)
    return L, A, b, x


def test_quick_start_tutorial(oracle):
    """ALS_SPD with xerus's defaults: ‖A·x − b‖ < 3.5e-9 and x equal to the dense solve b/A within 3.5e-9."""
    oracle.seed(SEED)
    L, A, b, x = quick_start_problem(oracle)
    assert b.ranks() == [1] * 8
    oracle.als(A, x, b)
    res = oracle.residual(A, x, b) * np.sqrt(512.0)  # ‖b‖ = √512
    assert res < 3.5e-9, res
    xd = np.linalg.solve(L, np.ones(512))
    assert np.linalg.norm(x.full().ravel() - xd) < 3.5e-9


@pytest.mark.parametrize("sites", [1, 2])
def test_nonsymmetric_full_rank_equals_dense_solve(oracle, sites):
    """ALS / DMRG without assumeSPD solve min ‖A·x − b‖ through AᵀA: at full rank a nonsymmetric, invertible A gives
    x = A⁻¹b. The energy returned is the relative residual by xerus's expansion √(⟨Ax,Ax⟩ − 2⟨Ax,b⟩ + ‖b‖²)/‖b‖
    (als.cpp:507-520); at a zero residual the radicand is a rounding residue that may be negative, so it is ≈ 0 or
    NaN, as in xerus."""
    rng = np.random.default_rng(SEED)
    d, n = 3, 3
    cores = [rng.standard_normal((1 if k == 0 else 2, n, n, 1 if k == d - 1 else 2)) for k in range(d)]
    A = oracle.TT.from_cores(cores, op=True)
    Ad = A.full().reshape(n ** d, n ** d)
    assert np.abs(Ad - Ad.T).max() > 0.1 and np.linalg.cond(Ad) < 1e6
    oracle.seed(SEED)
    b = oracle.TT.random([n] * d, [2] * (d - 1))
    x = oracle.TT.random([n] * d, [3, 3])
    e, _ = oracle.als(A, x, b, sites=sites, spd=False, num_half_sweeps=8, eps=0.0)
    xd = np.linalg.solve(Ad, b.full().ravel())
    assert np.linalg.norm(x.full().ravel() - xd) < 1e-8 * np.linalg.norm(xd)
    assert not e > 1e-6


def test_als_real_reference_test(oracle):
    """ALS:real (als.cxx:70-86): A random TTOperator (ranks 5), b = A·realX with realX random rank 3; ALS from a random
    rank-3 start with convergenceEpsilon 1e-7 returns an energy (the relative residual) below 1e-7 and recovers
    realX to 1e-4."""
    oracle.seed(SEED)
    x = oracle.TT.random([10] * 5, [3] * 4)
    realX = oracle.TT.random([10] * 5, [3] * 4)
    A = oracle.TT.random([10] * 10, [5] * 4, op=True)
    b = A.apply(realX)
    e, st = oracle.als(A, x, b, spd=False, eps=1e-7)
    diff = x.copy()
    diff.scale(-1.0)
    assert e < 1e-7, e
    assert (realX + diff).frob_norm() < 1e-4


def test_projection_als_reference_test(oracle):
    """ALS:projectionALS (als.cxx:88-103): approximating B by ALS (A = None) from its rank-r truncation never does
    worse than the truncation, for r = 7 … 1, and leaves B unchanged."""
    oracle.seed(SEED)
    B = oracle.TT.random([4] * 5, [4, 8, 8, 4])
    nB = B.frob_norm()
    X = B.copy()
    Bf = B.full()
    for r in range(7, 0, -1):
        X.round(r)
        roundNorm = np.linalg.norm(X.full() - Bf)
        oracle.als(None, X, B, eps=1e-4)
        projNorm = np.linalg.norm(X.full() - Bf)
        assert projNorm < roundNorm, (r, projNorm, roundNorm)
    assert B.frob_norm() == nB


def test_dmrg_full_rank_is_exact_solution(oracle):
    """DMRG_SPD with the target ranks at their maximum reaches the closed-form Laplacian solution."""
    d, n = 4, 3
    A = oracle.TT.from_cores(laplace_cores(d, n), op=True)
    oracle.seed(SEED)
    b = oracle.TT.random([n] * d, [2] * (d - 1))
    x = oracle.TT.random([n] * d, [9] * (d - 1))
    assert x.ranks() == [3, 9, 3]
    oracle.als(A, x, b, sites=2, num_half_sweeps=4, eps=0.0)
    xs = exact_laplace_solve(b.full(), d, n)
    assert np.linalg.norm(x.full() - xs) < 1e-12 * np.linalg.norm(xs)
    assert x.ranks() == [3, 9, 3]


def test_dmrg_identity_and_rank_cap(oracle):
    """DMRG_SPD with A = I returns b (rank 2 ≤ target 3); on the Laplacian with target rank 2 the ranks never
    exceed 2 and the energy does not increase over half sweeps."""
    d, n = 5, 3
    I = oracle.TT.from_cores([np.eye(n).reshape(1, n, n, 1)] * d, op=True)
    oracle.seed(SEED)
    b = oracle.TT.random([n] * d, [2] * (d - 1))
    x = oracle.TT.random([n] * d, [3] * (d - 1))
    oracle.als(I, x, b, sites=2, num_half_sweeps=2, eps=0.0)
    assert np.abs(x.full() - b.full()).max() < 1e-12
    assert all(r <= 3 for r in x.ranks())
    A = oracle.TT.from_cores(laplace_cores(d, n), op=True)
    y = oracle.TT.random([n] * d, [2] * (d - 1))
    energies = []
    for _ in range(4):
        e, _ = oracle.als(A, y, b, sites=2, num_half_sweeps=1, eps=0.0)
        energies.append(e)
        assert all(r <= 2 for r in y.ranks())
    # |½⟨x,Ax⟩ − ⟨x,b⟩| = −(½⟨x,Ax⟩ − ⟨x,b⟩) grows as the (negative) quadratic energy falls
    assert all(b2 >= a2 - 1e-12 * abs(a2) for a2, b2 in zip(energies, energies[1:]))


# ------------------------------------------------------------------ ADF


def slice_inner(G1, G2):
    return np.einsum("ajc,ajc->j", G1, G2)


def test_adf_slice_step_is_exact_line_search(oracle):
    """update_x (adf.cpp:468-478) moves slice j of core k by ‖Δ_j‖²/‖P_j A Δ‖²·Δ_j. That is the exact minimiser of
    ½‖P_Ω(x) − b‖² along Δ_j iff the gradient at the new x is orthogonal to Δ_j, slice by slice. A wrong numerator
    or denominator (a missing square, another slice's norm) breaks the orthogonality by O(1)."""
    d, n = 4, 4
    oracle.seed(SEED)
    truth = oracle.TT.random([n] * d, [3] * (d - 1))
    x = oracle.TT.random([n] * d, [3] * (d - 1))
    m = oracle.Measurements.random(150, [n] * d)
    m.measure(truth)
    for k in range(d):
        g_old, _ = oracle.adf_gradient(x, m)
        oracle.adf_core_step(x, m, k)
        g_new, _ = oracle.adf_gradient(x, m)
        old = slice_inner(g_old[k], g_old[k])
        cross = slice_inner(g_new[k], g_old[k])
        assert np.all(old > 0)
        assert np.abs(cross).max() < 1e-10 * old.max(), (k, cross, old)
        # the other cores' gradients changed (the step moved x), this one's component along Δ is gone
        assert np.abs(g_new[k] - g_old[k]).max() > 1e-6 * np.abs(g_old[k]).max()


def test_adf_random_low_rank_reference_test(oracle):
    """Algorithm:adf_random_low_rank (ttCompletion.cxx:107-130), single-point half: d=6, n=5, a random rank-3 truth,
    D·N·CS·R² = 2700 samples, ADFVariant(500, 1e-6, 0.999) from TTTensor::ones with maxRanks 3. The solve starts at
    rank 1, so it reaches rank 3 only through the rank-increase loop (adf.cpp:589-602)."""
    D, N, R, CS = 6, 5, 3, 10
    oracle.seed(SEED)
    truth = oracle.TT.random([N] * D, [R] * (D - 1))
    m = oracle.Measurements.random(D * N * CS * R * R, [N] * D)
    m.measure(truth)
    X = oracle.TT.ones([N] * D)
    assert X.ranks() == [1] * (D - 1)
    res, it = oracle.adf_solve(X, m, max_ranks=[R] * (D - 1), max_iterations=500, target=1e-6, min_decrease=0.999)
    assert X.ranks() == [R] * (D - 1)
    diff = X.copy()
    diff.scale(-1.0)
    err = (truth + diff).frob_norm() / truth.frob_norm()
    assert err < 1e-3, (err, res, it)


def test_adf_stack_plan_unsorted_with_duplicates(oracle):
    """The stack plan (adf.cpp:102-191) of an unsorted position set with repeated samples and long shared prefixes and
    suffixes: per level, each group of equal prefix (suffix) is one entry represented by its smallest sample id,
    listed in ascending order, as brute-force grouping gives."""
    rng = np.random.default_rng(SEED)
    pos = rng.integers(0, 2, size=(300, 7)).astype(np.uint32)
    pos[:, 2:5] = 0
    pos = pos[rng.permutation(len(pos))]
    m = oracle.Measurements.from_arrays(pos, np.zeros(len(pos)))
    fc, bc, fi, bi, nu = oracle.adf_plan(m, 7)
    for k in range(7):
        fwd, bwd = {}, {}
        for i in range(len(pos)):
            fwd.setdefault(tuple(pos[i, : k + 1]), i)
            bwd.setdefault(tuple(pos[i, k:]), i)
        assert fc[k] == len(fwd) and bc[k] == len(bwd)
        assert list(fi[sum(fc[:k]): sum(fc[: k + 1])]) == sorted(fwd.values())
        assert list(bi[sum(bc[:k]): sum(bc[: k + 1])]) == sorted(bwd.values())
    assert nu == sum(fc) + 1
