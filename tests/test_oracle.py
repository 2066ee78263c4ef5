"""Pins of the oracle (CPU): reference KATs, closed forms, invariants, brute force and finite differences.

Each check is chosen so a plausible mistake fails it: a transposed operand or wrong contraction side fails
the product KATs, a wrong permutation convention the reshuffle KATs, a wrong Householder sign the QC
rank-rule cases, a dropped stack term the ALS-vs-exact-solve and ADF finite-difference checks.
"""
import json
import os

import numpy as np
import pytest

from helpers import SEED, apply_dense_laplace, exact_laplace_solve, laplace_cores

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def einsum_via_oracle(O, A, B, lhs, rhs, out):
    """Evaluate out = lhs·rhs (index notation, no traces) with the oracle's contract + reshuffle only."""
    A, B = np.asarray(A, float), np.asarray(B, float)
    common = [c for c in lhs if c in rhs and c not in out]
    la = [c for c in lhs if c not in common] + common      # contracted modes last in A
    rb = common + [c for c in rhs if c not in common]      # … and first in B
    A2 = O.reshuffle(A, [la.index(c) for c in lhs]) if A.ndim > 1 else A
    B2 = O.reshuffle(B, [rb.index(c) for c in rhs]) if B.ndim > 1 else B
    R = O.contract(A2, False, B2, False, len(common))
    cur = la[: len(la) - len(common)] + rb[len(common):]
    if len(cur) <= 1:
        return R.ravel()
    return O.reshuffle(R, [out.index(c) for c in cur]).ravel()


def test_product_kats(oracle):
    kats = json.load(open(os.path.join(GOLDEN, "reference_kats.json")))
    for k in kats["products"]:
        got = einsum_via_oracle(oracle, k["A"], k["B"], k["lhs"], k["rhs"], k["out"])
        assert np.array_equal(got, np.array(k["expected"], float)), k["ref"]


def test_reshuffle_kats(oracle):
    kats = json.load(open(os.path.join(GOLDEN, "reference_kats.json")))
    for k in kats["reshuffles"]:
        A = np.arange(1, 25, dtype=float).reshape(k["dims"])
        # res(lhs) = A(rhs): input mode j (label rhs[j]) moves to the position of that label in lhs
        got = oracle.reshuffle(A, [k["lhs"].index(c) for c in k["rhs"]]).ravel()
        assert np.array_equal(got, np.array(k["expected"], float)), k["ref"]


def test_gemm_brute_force(oracle):
    rng = np.random.default_rng(1)
    for ta in (0, 1):
        for tb in (0, 1):
            A = rng.integers(-3, 4, (5, 7)).astype(float)
            B = rng.integers(-3, 4, (7, 4) if not tb else (4, 7)).astype(float)
            if ta:
                A = A.T.copy()
            got = oracle.gemm(A, B, ta=bool(ta), tb=bool(tb))
            ref = np.zeros((5, 4))
            for i in range(5):
                for j in range(4):
                    for k in range(7):
                        ref[i, j] += (A[k, i] if ta else A[i, k]) * (B[j, k] if tb else B[k, j])
            assert np.array_equal(got, ref)  # small integers: exact


@pytest.mark.parametrize("m,n", [(12, 5), (5, 12), (8, 8), (1, 4), (4, 1)])
def test_qr_rq_invariants(oracle, m, n):
    rng = np.random.default_rng(m * 10 + n)
    A = rng.standard_normal((m, n))
    Q, R = oracle.qr(A)
    k = min(m, n)
    assert np.abs(Q @ R - A).max() < 1e-14 * 10
    assert np.abs(Q.T @ Q - np.eye(k)).max() < 1e-14 * 10
    assert np.abs(np.tril(R, -1)).max() == 0
    # Householder sign convention (dlarfg): R_00 = -sign(a_00)·‖a_0‖, and H = I for a single row
    assert np.sign(R[0, 0]) == (-np.sign(A[0, 0]) if m > 1 else np.sign(A[0, 0]))
    R2, Q2 = oracle.rq(A)
    assert np.abs(R2 @ Q2 - A).max() < 1e-14 * 10
    assert np.abs(Q2 @ Q2.T - np.eye(k)).max() < 1e-14 * 10
    # R of an RQ is upper triangular in its last k rows
    assert np.abs(np.tril(R2[m - k:, :], -1)).max() == 0


def test_qc_rank_rule_sign(oracle):
    """xerus's QC rank rule |R_kk| < 16·ε·R_00 uses R_00 with LAPACK's sign (blasLapackWrapper.cpp:268-272):
    a rank-1 matrix whose pivot column starts positive gets R_00 < 0, so no rank reduction happens."""
    _, _, r_pos = oracle.qc(np.ones((2, 2)))
    _, _, r_neg = oracle.qc(-np.ones((2, 2)))
    assert r_pos == 2 and r_neg == 1
    _, _, r_cq_neg = oracle.cq(-np.ones((2, 2)))
    assert r_cq_neg == 1


@pytest.mark.parametrize("m,n,r", [(20, 8, 3), (8, 20, 3), (30, 30, 30)])
def test_qc_cq_reconstruct(oracle, m, n, r):
    rng = np.random.default_rng(r)
    A = -np.abs(rng.standard_normal((m, r))) @ rng.standard_normal((r, n))
    Q, Cm, rk = oracle.qc(A)
    assert np.abs(Q @ Cm - A).max() < 1e-12 * np.abs(A).max()
    assert np.abs(Q.T @ Q - np.eye(rk)).max() < 1e-13
    C2, Q2, rk2 = oracle.cq(A)
    assert np.abs(C2 @ Q2 - A).max() < 1e-12 * np.abs(A).max()
    assert rk <= min(m, n) and rk2 <= min(m, n)


def test_svd_closed_forms_and_truncation(oracle):
    U, S, Vt = oracle.svd(np.eye(5))
    assert np.allclose(S, 1) and np.allclose((U * S) @ Vt, np.eye(5))
    U, S, Vt = oracle.svd(np.zeros((4, 3)))
    assert np.all(S == 0) and np.allclose(U.T @ U, np.eye(3))
    rng = np.random.default_rng(4)
    Q1, _ = np.linalg.qr(rng.standard_normal((10, 6)))
    Q2, _ = np.linalg.qr(rng.standard_normal((8, 6)))
    s = np.array([3.0, 1.0, 0.5, 1e-9, 1e-13, 1e-15])
    A = (Q1 * s) @ Q2.T
    U, S, Vt = oracle.svd_truncated(A, 1, 0, 1e-12)   # first j with s_j <= eps·s_0 (tensor.cpp:1463-1474)
    assert len(S) == 4
    U, S, Vt = oracle.svd_truncated(A, 1, 2, 0.0)     # maxRank caps first
    assert len(S) == 2 and np.allclose(S, s[:2])


def test_reduce_to_maximal_ranks_and_random(oracle):
    oracle.seed(SEED)
    x = oracle.TT.random([4] * 8, [8] * 7)
    assert x.ranks() == [4, 8, 8, 8, 8, 8, 4]
    y = oracle.TT.random([16] * 16, [64] * 15)
    assert y.ranks() == [16] + [64] * 13 + [16]
    assert x.state() == (True, 0)


def test_tt_sum_ranks_and_difference(oracle):
    """TT:real_difference (ttArithmetic.cxx:74-115)."""
    oracle.seed(SEED)
    a = oracle.TT.random([10] * 5, [4] * 4)
    b = oracle.TT.random([10] * 5, [4] * 4)
    c = a + b
    assert c.ranks() == [8, 8, 8, 8]
    na = a.copy(); na.scale(-1.0)
    assert (a + na).frob_norm() < 1e-11
    assert np.abs(c.full() - a.full() - b.full()).max() < 1e-12


def test_no_rounding(oracle):
    """TT:no_rounding (ttRounding.cxx:104-117)."""
    oracle.seed(SEED)
    a = oracle.TT.random([2] * 7, [2] * 6)
    ref = a.full()
    a.round(2)
    assert np.allclose(a.full(), ref, atol=1e-14)
    c = oracle.TT.random([2] * 7, [2] * 6)
    c.scale(0.0)
    s = a + c
    s.round(2)
    assert np.allclose(s.full(), ref, atol=1e-14)


def test_move_core_preserves_tensor(oracle):
    oracle.seed(SEED)
    x = oracle.TT.random([3, 4, 5, 3], [3, 6, 3])
    ref = x.full()
    for pos, keep in ((3, True), (0, False), (2, True), (1, False)):
        x.move_core(pos, keep)
        assert np.abs(x.full() - ref).max() < 1e-13
        assert x.state() == (True, pos)


def test_residual_matches_dense(oracle):
    d, n = 5, 4
    rng = np.random.default_rng(2)
    pots = [rng.random(n) for _ in range(d)]
    A = oracle.TT.from_cores(laplace_cores(d, n, pots), op=True)
    oracle.seed(SEED)
    x = oracle.TT.random([n] * d, [3] * (d - 1))
    b = oracle.TT.random([n] * d, [2] * (d - 1))
    dense = np.linalg.norm(apply_dense_laplace(x.full(), d, n, pots) - b.full()) / np.linalg.norm(b.full())
    assert abs(oracle.residual(A, x, b) - dense) < 1e-13 * dense


def test_als_full_rank_reaches_exact_solution(oracle):
    """With ranks at their maximum the ALS fixed point is the exact solution x* = A⁻¹b, known in closed form
    for the Laplacian (sine eigenbasis)."""
    d, n = 4, 4
    A = oracle.TT.from_cores(laplace_cores(d, n), op=True)
    oracle.seed(SEED)
    b = oracle.TT.random([n] * d, [2] * (d - 1))
    x = oracle.TT.random([n] * d, [16] * (d - 1))
    oracle.als(A, x, b, num_half_sweeps=4, eps=0.0)
    xs = exact_laplace_solve(b.full(), d, n)
    assert np.linalg.norm(x.full() - xs) < 1e-12 * np.linalg.norm(xs)


def test_als_config1_energy_monotone_and_cg_agrees(oracle):
    d, n = 8, 4
    A = oracle.TT.from_cores(laplace_cores(d, n), op=True)
    oracle.seed(SEED)
    b = oracle.TT.random([n] * d, [2] * (d - 1))
    b.scale(1.0 / b.frob_norm())
    x = oracle.TT.random([n] * d, [8] * (d - 1))
    x2 = x.copy()
    e, st = oracle.als(A, x, b, num_half_sweeps=10, eps=0.0)
    assert st["local_solves"] == 51 and st["half_sweeps"] == 10
    e2, _ = oracle.als(A, x2, b, num_half_sweeps=10, eps=0.0, local_solver="cg", cg_tol=1e-14)
    assert abs(e - e2) < 1e-13 * abs(e)
    res = oracle.residual(A, x, b)
    dense = np.linalg.norm(apply_dense_laplace(x.full(), d, n) - b.full()) / np.linalg.norm(b.full())
    assert abs(res - dense) < 1e-10 * dense
    assert res < 1e-4


def test_als_identity_operator(oracle):
    """ALS:identity (als.cxx:28-67)."""
    n, d = 10, 3
    A = oracle.TT.from_cores([np.eye(n).reshape(1, n, n, 1)] * d, op=True)
    oracle.seed(SEED)
    b = oracle.TT.random([n] * d, [4, 4])
    x = oracle.TT.random([n] * d, [4, 4])
    oracle.als(A, x, b, eps=1e-4, use_residual=True)
    assert np.abs(x.full() - b.full()).max() < 1e-12


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


def test_adf_stack_plan_brute_force(oracle):
    oracle.seed(SEED)
    m = oracle.Measurements.random(400, [5, 4, 6, 5])
    pos = m.positions(4)
    assert len({tuple(p) for p in pos}) == 400
    assert all(tuple(pos[i]) < tuple(pos[i + 1]) for i in range(399))  # sorted lexicographically, unique
    fc, bc, fi, bi, nu = oracle.adf_plan(m, 4)
    fwd, bwd = brute_plan(pos)
    assert list(fc) == [len(g) for g in fwd] and list(bc) == [len(g) for g in bwd]
    assert list(fi) == [i for g in fwd for i in g] and list(bi) == [i for g in bwd for i in g]
    assert nu == sum(fc) + 1


def test_adf_gradient_finite_differences(oracle):
    """Δ_k = −∂/∂G_k ½Σ(x_i − b_i)², checked by central differences (h = 1e-6, FP64) and by brute force."""
    d, n = 3, 3
    oracle.seed(SEED)
    truth = oracle.TT.random([n] * d, [2, 2])
    x = oracle.TT.random([n] * d, [2, 2])
    m = oracle.Measurements.random(15, [n] * d)
    m.measure(truth)
    pos, b = m.positions(d), m.values()
    cores = x.cores()

    def f(cs):
        full = oracle.TT.from_cores(cs).full()
        return 0.5 * sum((full[tuple(p)] - bi) ** 2 for p, bi in zip(pos, b))

    grads, rn = oracle.adf_gradient(x, m)
    full = x.full()
    assert abs(rn - np.sqrt(sum((full[tuple(p)] - bi) ** 2 for p, bi in zip(pos, b)) / (b @ b))) < 1e-13
    h = 1e-6
    for k in range(d):
        for idx in np.ndindex(cores[k].shape):
            cp = [c.copy() for c in cores]
            cm = [c.copy() for c in cores]
            cp[k][idx] += h
            cm[k][idx] -= h
            fd = (f(cp) - f(cm)) / (2 * h)
            assert abs(-grads[k][idx] - fd) < 1e-7 * max(1.0, abs(fd))


def test_adf_solve_reduces_residual(oracle):
    d, n = 4, 5
    oracle.seed(SEED)
    truth = oracle.TT.random([n] * d, [2] * (d - 1))
    x = oracle.TT.random([n] * d, [2] * (d - 1))
    m = oracle.Measurements.random(400, [n] * d)
    m.measure(truth)
    res, it = oracle.adf_solve(x, m, max_iterations=30)
    assert it <= 30 and res < 1e-3


def test_rng_prefix_fixture(oracle):
    """Drift detector for the seeded mt19937_64 + libstdc++ distributions every workload is drawn from."""
    fx = json.load(open(os.path.join(GOLDEN, "rng_prefix.json")))
    oracle.seed(fx["seed"])
    assert np.array_equal(oracle.normal_draws(len(fx["normal"])), np.array(fx["normal"]))
    assert [int(v) for v in oracle.uniform_draws(len(fx["uniform9"]), 9)] == fx["uniform9"]


def _solve_cases(rng):
    """Matrices that take each branch of xerus's dispatch (blasLapackWrapper.cpp:501-651)."""
    n = 12
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    spd = Q @ np.diag(np.linspace(1.0, 5.0, n)) @ Q.T
    indef = Q @ np.diag(np.linspace(-1.0, 5.0, n)) @ Q.T
    indef = indef + np.diag(np.abs(np.diag(indef)) + 0.1 - np.diag(indef))  # positive diagonal, still indefinite
    mixed = spd.copy()
    mixed[3, 3] = -1.0                                                         # diagonal of both signs
    neg = -(np.abs(rng.standard_normal((n, n))) + 0.1)
    neg = neg + neg.T                                                          # symmetric, every entry < 0
    return [("nonsym", rng.standard_normal((n, n)), 0), ("spd", spd, 1), ("indef_posdiag", indef, 2), ("mixed_diag", mixed, 2),
            ("negdef", -spd, 2), ("all_negative", neg, 0)]


@pytest.mark.parametrize("case", range(6))
def test_oracle_solve_dispatch(oracle, case):
    """The oracle's dense solve: xerus's method choice on constructed matrices (the all-negative symmetric matrix
    fails xerus's symmetry test because its maximum entry starts at 0) and the solution against numpy."""
    name, A, method = _solve_cases(np.random.default_rng(55))[case]
    if name == "indef_posdiag":
        assert np.all(np.diag(A) > 0) and np.linalg.eigvalsh(A).min() < 0
    b = np.random.default_rng(56).standard_normal(A.shape[0])
    x, m = oracle.solve(A, b)
    assert m == method, name
    ref = np.linalg.solve(A, b)
    assert np.abs(x - ref).max() < 1e-10 * max(1.0, np.abs(ref).max())
