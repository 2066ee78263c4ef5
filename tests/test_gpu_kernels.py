"""GPU parity of the boundary-1 kernels (xerus::blasWrapper replacement) against the oracle and closed forms.

GEMM: every transpose combination, ragged shapes and large K against float64 numpy (FP64 bit-close, FP32
and BF16 to their stated tolerances). QR/RQ/QC/CQ/SVD: reconstruction, orthogonality, triangularity and the
integer rank decisions of xerus's QC rule (blasLapackWrapper.cpp:268-272), compared with the oracle.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def dev(a, dtype=np.float64):
    from xerus_amd._lib import DeviceArray
    return DeviceArray.from_numpy(np.asarray(a, dtype=dtype))


def empty(shape, dtype=np.float64):
    from xerus_amd._lib import DeviceArray
    return DeviceArray.from_numpy(np.zeros(shape, dtype=dtype))


def ptr(t):
    return t.ptr


def host(t):
    return t.numpy()


@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (17, 33, 5), (64, 64, 64), (192, 1024, 64), (1024, 64, 192), (300, 130, 700),
                                   (64, 192, 1024), (4, 64, 1024), (33, 17, 2001), (36, 108, 576)])
def test_gemm_f64(xe, ta, tb, m, n, k):
    from xerus_amd._lib import call
    print(f"gemm_f64 ta={ta} tb={tb} m={m} n={n} k={k}", flush=True)
    rng = np.random.default_rng(m * 7 + n * 3 + k)
    A = rng.standard_normal((k, m) if ta else (m, k))
    B = rng.standard_normal((n, k) if tb else (k, n))
    ref = 0.5 * (A.T if ta else A) @ (B.T if tb else B)
    dA, dB = dev(A), dev(B)
    dC = empty((m, n))
    call("xt_gemm", ta, tb, m, n, k, 0.5, ptr(dA), ptr(dB), ptr(dC), None)
    err = np.abs(host(dC) - ref).max() / max(1.0, np.abs(ref).max())
    assert err < 1e-14 * max(1, k) ** 0.5 * 10


@pytest.mark.parametrize("prec,tol", [("f32", 1e-5), ("bf16", 2e-2)])
@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0), (0, 1)])
@pytest.mark.parametrize("m,n,k", [(48, 64, 48), (1024, 64, 192), (33, 17, 100), (64, 192, 1024), (5, 36, 577)])
def test_gemm_lowp(xe, prec, tol, ta, tb, m, n, k):
    from xerus_amd._lib import call
    print(f"gemm_{prec} ta={ta} tb={tb} m={m} n={n} k={k}", flush=True)
    rng = np.random.default_rng(k)
    A = rng.standard_normal((k, m) if ta else (m, k)).astype(np.float32)
    B = rng.standard_normal((n, k) if tb else (k, n)).astype(np.float32)
    ref = (A.T if ta else A).astype(np.float64) @ (B.T if tb else B).astype(np.float64)
    dA, dB = dev(A, np.float32), dev(B, np.float32)
    dC = empty((m, n), np.float32)
    call("xt_gemm_f32" if prec == "f32" else "xt_gemm_bf16", ta, tb, m, n, k, 1.0, ptr(dA), ptr(dB), ptr(dC), None)
    # relative to the scale of a length-k dot of unit normals
    err = np.abs(host(dC) - ref).max() / np.sqrt(k)
    assert err < tol


def _qr_checks(Q, R, A, tol=1e-13):
    k = min(A.shape)
    assert np.abs(Q @ R - A).max() < tol * max(1, np.abs(A).max()) * A.shape[1]
    assert np.abs(Q.T @ Q - np.eye(k)).max() < tol * 10


@pytest.mark.parametrize("m,n", [(32, 8), (8, 32), (1024, 64), (64, 1024), (7, 7), (1, 5), (5, 1), (576, 36), (4096, 64), (150, 40)])
def test_qr_rq(xe, oracle, m, n):
    from xerus_amd._lib import call
    print(f"qr/rq m={m} n={n}", flush=True)
    rng = np.random.default_rng(m + n)
    A = rng.standard_normal((m, n))
    k = min(m, n)
    dA = dev(A)
    dQ, dR = empty((m, k)), empty((k, n))
    call("xt_qr", ptr(dA), m, n, ptr(dQ), ptr(dR))
    Q, R = host(dQ), host(dR)
    _qr_checks(Q, R, A)
    assert np.abs(np.tril(R, -1)).max() == 0
    Qo, Ro = oracle.qr(A)
    # same Householder convention ⇒ the same factors up to rounding
    assert np.abs(np.abs(R) - np.abs(Ro)).max() < 1e-10 * np.abs(Ro).max()
    dR2, dQ2 = empty((m, k)), empty((k, n))
    call("xt_rq", ptr(dA), m, n, ptr(dR2), ptr(dQ2))
    R2, Q2 = host(dR2), host(dQ2)
    assert np.abs(R2 @ Q2 - A).max() < 1e-12 * n
    assert np.abs(Q2 @ Q2.T - np.eye(k)).max() < 1e-12
    Ro2, Qo2 = oracle.rq(A)
    assert np.abs(np.abs(R2) - np.abs(Ro2)).max() < 1e-10 * np.abs(Ro2).max()


@pytest.mark.parametrize("m,n,r", [(40, 12, 12), (40, 12, 5), (12, 40, 3), (1024, 64, 64), (1024, 64, 20), (64, 64, 1),
                                   (576, 36, 36), (576, 36, 11), (700, 33, 33), (4096, 64, 30),
                                   (150, 40, 40), (170, 64, 20)])
def test_qc_cq_rank(xe, oracle, m, n, r):
    """Rank decisions must equal the oracle's (xerus's 16·ε·R_00 rule with LAPACK signs)."""
    from xerus_amd._lib import call
    print(f"qc/cq m={m} n={n} r={r}", flush=True)
    rng = np.random.default_rng(m * n + r)
    A = rng.standard_normal((m, r)) @ rng.standard_normal((r, n))
    k = min(m, n)
    dA = dev(A)
    dQ, dC = empty(m * k), empty(k * n)
    rank = ctypes.c_size_t()
    call("xt_qc", ptr(dA), m, n, ptr(dQ), ptr(dC), ctypes.byref(rank))
    rk = rank.value
    Qo, Co, ro = oracle.qc(A)
    assert rk == ro
    Q = host(dQ)[: m * rk].reshape(m, rk)
    Cm = host(dC)[: rk * n].reshape(rk, n)
    assert np.abs(Q @ Cm - A).max() < 1e-11 * np.abs(A).max()
    assert np.abs(Q.T @ Q - np.eye(rk)).max() < 1e-12
    dC2, dQ2 = empty(m * k), empty(k * n)
    call("xt_cq", ptr(dA), m, n, ptr(dC2), ptr(dQ2), ctypes.byref(rank))
    Co2, Qo2, ro2 = oracle.cq(A)
    assert rank.value == ro2
    rk = rank.value
    C2 = host(dC2)[: m * rk].reshape(m, rk)
    Q2 = host(dQ2)[: rk * n].reshape(rk, n)
    assert np.abs(C2 @ Q2 - A).max() < 1e-11 * np.abs(A).max()
    assert np.abs(Q2 @ Q2.T - np.eye(rk)).max() < 1e-12


@pytest.mark.parametrize("case", range(6))
def test_solve_dispatch_matches_oracle(xe, oracle, case):
    """xt_solve (blasWrapper::solve): the same method as the oracle's restatement of xerus's dispatch and the same
    solution to 1e-10 relative."""
    from test_oracle import _solve_cases
    name, A, method = _solve_cases(np.random.default_rng(55))[case]
    print(f"solve {name}", flush=True)
    b = np.random.default_rng(56).standard_normal(A.shape[0])
    x, m = xe.solve(A, b, return_method=True)
    xo, mo = oracle.solve(A, b)
    assert m == mo == method, name
    assert np.abs(x - xo).max() < 1e-10 * max(1.0, np.abs(xo).max())


@pytest.mark.parametrize("n,nrhs,kind", [(300, 3, "nonsym"), (257, 2, "spd"), (1, 1, "spd"), (2, 1, "nonsym")])
def test_solve_sizes(xe, n, nrhs, kind):
    """Several right-hand sides and sizes past one thread per row: the residual of the solution is at rounding level
    (backward stability), and it matches numpy's solve to the conditioning."""
    print(f"solve n={n} nrhs={nrhs} {kind}", flush=True)
    rng = np.random.default_rng(n + nrhs)
    if kind == "spd":
        M = rng.standard_normal((n, n))
        A = M @ M.T + n * np.eye(n)
    else:
        A = rng.standard_normal((n, n)) + n ** 0.5 * np.eye(n)
    B = rng.standard_normal((n, nrhs))
    X, m = xe.solve(A, B, return_method=True)
    assert m == (1 if kind == "spd" else 0)
    assert np.abs(A @ X - B).max() < 1e-12 * np.abs(A).max() * np.abs(X).max() * n
    ref = np.linalg.solve(A, B)
    assert np.abs(X - ref).max() < 1e-9 * np.abs(ref).max()


def test_solve_singular_is_an_error(xe):
    A = np.ones((5, 5))
    A[:, 0] = np.arange(5.0)  # rank 2, not symmetric → LU meets an exactly zero pivot
    A[:, 1] = 2 * A[:, 0]
    with pytest.raises(xe.XttError):
        xe.solve(A, np.ones(5))


@pytest.mark.parametrize("m,n,rank", [(120, 40, 40), (30, 70, 30), (60, 20, 10)])
def test_solve_least_squares(xe, m, n, rank):
    """m ≠ n: dgelsd's minimum-norm least-squares solution with rcond = 8ε, against numpy's pseudo-inverse with the
    same cut-off: full-rank over- and underdetermined systems, and a rank-deficient one whose missing rank is
    exactly zero columns (singular values exactly zero, clear of the cut-off)."""
    print(f"lstsq m={m} n={n} rank={rank}", flush=True)
    rng = np.random.default_rng(m * n + rank)
    A = rng.standard_normal((m, n))
    A[:, rank:] = 0.0
    A = A[:, rng.permutation(n)]
    B = rng.standard_normal((m, 2))
    X, meth = xe.solve(A, B, return_method=True)
    assert meth == 3
    ref = np.linalg.pinv(A, rcond=8 * np.finfo(float).eps) @ B
    assert np.abs(X - ref).max() < 1e-9 * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("m,n", [(64, 64), (100, 30), (30, 100), (1, 1), (9, 4)])
def test_svd(xe, oracle, m, n):
    from xerus_amd._lib import call
    print(f"svd m={m} n={n}", flush=True)
    rng = np.random.default_rng(m * 31 + n)
    # graded spectrum down to 1e-14 so that the small singular values are resolved
    k = min(m, n)
    U0, _ = np.linalg.qr(rng.standard_normal((m, k)))
    V0, _ = np.linalg.qr(rng.standard_normal((n, k)))
    s0 = np.logspace(0, -14, k)
    A = (U0 * s0) @ V0.T
    dA = dev(A)
    dU, dS, dV = (empty(s) for s in ((m, k), (k,), (k, n)))
    call("xt_svd", ptr(dA), m, n, ptr(dU), ptr(dS), ptr(dV))
    U, S, Vt = host(dU), host(dS), host(dV)
    assert np.all(np.diff(S) <= 0)
    _, So, _ = oracle.svd(A)
    # one-sided Jacobi: small singular values to high relative accuracy
    assert np.all(np.abs(S - So) <= 1e-13 * S[0] + 1e-6 * So)
    assert np.abs((U * S) @ Vt - A).max() < 1e-13
    assert np.abs(U.T @ U - np.eye(k)).max() < 1e-10
    assert np.abs(Vt @ Vt.T - np.eye(k)).max() < 1e-10


def test_svd_closed_forms(xe):
    """SVD of the identity and of the zero matrix (the reference's Tensor:SVD_Identity / SVD_zero cases)."""
    from xerus_amd._lib import call
    for A in (np.eye(6), np.zeros((5, 3))):
        m, n = A.shape
        k = min(m, n)
        dA = dev(A)
        dU, dS, dV = (empty(s) for s in ((m, k), (k,), (k, n)))
        call("xt_svd", ptr(dA), m, n, ptr(dU), ptr(dS), ptr(dV))
        S = host(dS)
        assert np.allclose(S, np.linalg.svd(A, compute_uv=False), atol=1e-15)


def test_reshuffle(xe, oracle):
    from xerus_amd._lib import call
    C = ctypes
    rng = np.random.default_rng(3)
    T = rng.standard_normal((2, 3, 4, 5))
    for sh in ([0, 1, 2, 3], [3, 2, 1, 0], [1, 3, 0, 2], [2, 0, 3, 1]):
        dims = np.array(T.shape, dtype=np.uintp)
        shuf = np.array(sh, dtype=np.uintp)
        out = empty(T.size)
        call("xt_reshuffle", 4, dims.ctypes.data_as(C.POINTER(C.c_size_t)), shuf.ctypes.data_as(C.POINTER(C.c_size_t)),
             ptr(dT := dev(T)), ptr(out))
        ref = oracle.reshuffle(T, sh)
        assert np.array_equal(host(out), ref.ravel())  # data movement: bit-exact


def _sampled_step(d, n, r, cores, pos, vals, fused):
    """One ADF gradient step through the raw sampled-entry entry points: the fused call, or the residual call
    followed by the adjoint call seeded with that residual."""
    from xerus_amd._lib import call
    flat = dev(np.concatenate([c.ravel() for c in cores]), np.float32)
    dpos, db = dev(pos, np.uint32), dev(vals, np.float32)
    P = len(vals)
    resid, grad = empty(P, np.float32), empty(flat.shape, np.float32)
    ns, rs = (ctypes.c_int * d)(*n), (ctypes.c_int * (d + 1))(*r)
    sums = (ctypes.c_double * 2)()
    if fused:
        call("xt_tt_sample_residual_adjoint_f32", d, ns, rs, ptr(flat), ptr(dpos), P, ptr(db), ptr(resid), sums, ptr(grad), None)
    else:
        call("xt_tt_sample_residual_f32", d, ns, rs, ptr(flat), ptr(dpos), P, ptr(db), ptr(resid), sums, None)
        call("xt_tt_sample_adjoint_f32", d, ns, rs, ptr(flat), ptr(dpos), P, ptr(resid), ptr(grad), None)
    return host(resid), host(grad), np.array(sums[:])


@pytest.mark.parametrize("d,n,r,P", [(10, 10, 8, 20000), (4, 3, 2, 77), (6, 16, 8, 3001), (3, 4, 3, 1), (5, 6, 12, 2000)])
def test_sampled_residual_adjoint_fused(xe, oracle, d, n, r, P):
    """xt_tt_sample_residual_adjoint_f32 against the two calls it replaces: residuals and gradients bit-identical
    (the fused kernel's forward walk takes the residual kernel's steps in its order), the FP64 sums to rounding;
    the gradient against the oracle's ADF gradient (adf.cpp:314-396, Δ = −∂½Σ(x−b)²/∂G, pinned to FP64 finite
    differences in test_oracle). Ranks above 8 take the two-kernel path."""
    from helpers import SEED
    from test_gpu_tt import twin_random
    print(f"fused sampled step d={d} n={n} r={r} P={P}", flush=True)
    _, xo = twin_random(xe, oracle, [n] * d, [r] * (d - 1))
    _, trutho = twin_random(xe, oracle, [n] * d, [r] * (d - 1), seed=SEED + 2)
    oracle.seed(SEED + 3)
    mo = oracle.Measurements.random(P, [n] * d)
    mo.measure(trutho)
    cores = [c.astype(np.float32) for c in xo.cores()]
    rk = [1] + [c.shape[2] for c in cores]
    pos, vals = mo.positions(d), mo.values()
    rf, gf, sf = _sampled_step(d, [n] * d, rk, cores, pos, vals, True)
    rs, gs, ss = _sampled_step(d, [n] * d, rk, cores, pos, vals, False)
    np.testing.assert_array_equal(rf, rs)
    np.testing.assert_array_equal(gf, gs)
    assert np.abs(sf - ss).max() <= 1e-12 * np.abs(ss).max()
    go, _ = oracle.adf_gradient(oracle.TT.from_cores([c.astype(np.float64) for c in cores]), mo)
    o = 0
    for k in range(d):
        g = gf[o: o + go[k].size].reshape(go[k].shape)
        o += go[k].size
        assert np.abs(g + go[k]).max() <= 3e-5 * max(np.abs(go[k]).max(), 1e-30), k


def test_sampled_residual_adjoint_fused_out_of_range(xe, oracle):
    """A position outside its mode: residual NaN and no gradient contribution, in the fused call as in the two."""
    from helpers import SEED
    from test_gpu_tt import twin_random
    print("fused sampled step, out-of-range position", flush=True)
    d, n, r, P = 5, 6, 4, 500
    _, xo = twin_random(xe, oracle, [n] * d, [r] * (d - 1))
    oracle.seed(SEED + 4)
    mo = oracle.Measurements.random(P, [n] * d)
    pos = mo.positions(d).copy()
    pos[17, 2] = n
    vals = np.random.default_rng(SEED).normal(size=P)
    cores = [c.astype(np.float32) for c in xo.cores()]
    rk = [1] + [c.shape[2] for c in cores]
    rf, gf, sf = _sampled_step(d, [n] * d, rk, cores, pos, vals, True)
    rs, gs, ss = _sampled_step(d, [n] * d, rk, cores, pos, vals, False)
    assert np.isnan(rf[17]) and np.isnan(rs[17]) and np.isnan(sf[0]) and np.isnan(ss[0])
    np.testing.assert_array_equal(rf, rs)
    np.testing.assert_array_equal(gf, gs)
    assert np.all(np.isfinite(gf))


def test_blas_level1_level2(xe):
    """one_norm (dasum), two_norm (dnrm2, scaled: entries of 1e200 and 1e-200 neither overflow nor vanish),
    matrix_vector_product in both orientations and dyadic_vector_product (blasLapackWrapper.h:42-56) against numpy."""
    from xerus_amd._lib import call
    print("blas level 1 and 2", flush=True)
    rng = np.random.default_rng(3)
    for n in (1, 7, 1000, 100003):
        x = rng.standard_normal(n)
        out = ctypes.c_double()
        dx = dev(x)
        call("xt_one_norm", ptr(dx), n, ctypes.byref(out))
        assert abs(out.value - np.abs(x).sum()) <= 1e-12 * np.abs(x).sum()
        call("xt_two_norm", ptr(dx), n, ctypes.byref(out))
        assert abs(out.value - np.linalg.norm(x)) <= 1e-13 * np.linalg.norm(x)
    for scale in (1e200, 1e-200):
        x = rng.standard_normal(513) * scale
        out = ctypes.c_double()
        dx = dev(x)  # device buffers stay referenced for the whole call (a temporary would be freed before it)
        call("xt_two_norm", ptr(dx), 513, ctypes.byref(out))
        ref = scale * np.linalg.norm(x / scale)
        assert np.isfinite(out.value) and abs(out.value - ref) <= 1e-13 * ref
    for m, n in ((1, 5), (5, 1), (37, 200), (300, 64)):
        A = rng.standard_normal((m, n))
        y = rng.standard_normal(n)
        out, dA, dy = empty(m), dev(A), dev(y)
        call("xt_matrix_vector_product", ptr(out), m, 2.5, ptr(dA), n, 0, ptr(dy), None)
        assert np.abs(host(out) - 2.5 * A @ y).max() <= 1e-12 * max(1.0, np.abs(A @ y).max())
        At = rng.standard_normal((n, m))
        out, dAt = empty(m), dev(At)
        call("xt_matrix_vector_product", ptr(out), m, -1.0, ptr(dAt), n, 1, ptr(dy), None)
        assert np.abs(host(out) + At.T @ y).max() <= 1e-12 * max(1.0, np.abs(At.T @ y).max())
        xv, yv = rng.standard_normal(m), rng.standard_normal(n)
        D, dxv, dyv = dev(np.full((m, n), 9.0)), dev(xv), dev(yv)
        call("xt_dyadic_vector_product", ptr(D), m, n, 0.5, ptr(dxv), ptr(dyv), None)
        assert np.abs(host(D) - 0.5 * np.outer(xv, yv)).max() <= 1e-14 * max(1.0, np.abs(np.outer(xv, yv)).max())


@pytest.mark.parametrize("m,n,p,rank", [(8, 8, 1, 8), (8, 8, 3, 5), (40, 12, 2, 12), (12, 40, 2, 12), (30, 20, 1, 7)])
def test_solve_least_squares_entry(xe, m, n, p, rank):
    """solve_least_squares (blasLapackWrapper.cpp:654-663, dgelsd): the minimum-norm least squares solution, for square
    and rank-deficient systems too, equal to numpy's pseudo-inverse solution."""
    from xerus_amd._lib import call
    print(f"least squares {m}x{n} p={p} rank={rank}", flush=True)
    rng = np.random.default_rng(m * 100 + n)
    A = rng.standard_normal((m, rank)) @ rng.standard_normal((rank, n))
    B = rng.standard_normal((m, p))
    X, dA, dB = empty((n, p)), dev(A), dev(B)
    call("xt_solve_least_squares", ptr(X), ptr(dA), m, n, ptr(dB), p)
    ref = np.linalg.pinv(A) @ B
    assert np.abs(host(X) - ref).max() <= 1e-9 * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("m,n,k,ta,tb", [(36, 108, 576, 1, 0), (1024, 64, 192, 0, 1), (17, 33, 130, 0, 0), (300, 130, 64, 1, 1)])
def test_gemm_tuned_choices(xe, prec, m, n, k, ta, tb):
    """Every choice the tile autotuner can install (64-tiles, 32-tiles with 0..32 split-K slices) computes the same
    product (float64 numpy), and xt_gemm_tune installs a choice no slower than the heuristic's."""
    rng = np.random.default_rng(m * 7 + k)
    dt = np.float64 if prec == "f64" else np.float32
    A = rng.standard_normal((k, m) if ta else (m, k))
    B = rng.standard_normal((n, k) if tb else (k, n))
    ref = (A.T if ta else A) @ (B.T if tb else B)
    P = xe.XT_F64 if prec == "f64" else xe.XT_F32
    fn = "xt_gemm" if prec == "f64" else "xt_gemm_f32"
    alpha = 1.0 if prec == "f64" else ctypes.c_float(1.0)
    dA, dB, dC = dev(A, dt), dev(B, dt), empty((m, n), dt)
    tol = 1e-13 if prec == "f64" else 1e-5
    try:
        for tile, slices in [(64, 0), (32, 0), (32, 2), (32, 4), (32, 32)]:
            xe._lib.call("xt_gemm_tune_set", P, m, n, k, 1, ta, tb, tile, slices)
            xe._lib.call(fn, ta, tb, m, n, k, alpha, ctypes.c_void_p(ptr(dA)), ctypes.c_void_p(ptr(dB)), ctypes.c_void_p(ptr(dC)), None)
            err = np.abs(host(dC) - ref).max() / np.abs(ref).max()
            assert err < tol, (tile, slices, err)
        t, s, ht, hs = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        us, hus = ctypes.c_double(), ctypes.c_double()
        xe._lib.call("xt_gemm_tune", P, m, n, k, 1, ta, tb, 20, ctypes.byref(t), ctypes.byref(s), ctypes.byref(us), ctypes.byref(ht),
                     ctypes.byref(hs), ctypes.byref(hus))
        assert us.value <= hus.value and t.value in (32, 64)
    finally:
        # back to the committed table for the tests that follow
        xe._lib.call("xt_gemm_tune_clear")
        from xerus_amd import tuning
        tuning.load()
