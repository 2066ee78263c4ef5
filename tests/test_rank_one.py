"""RankOneMeasurementSet and ADF on rank-one measurements (measurments.cpp:250-545, adf.cpp specialisations).

CPU: the numpy oracle (oracle/rank_one.py) is pinned against a dense brute-force contraction, central finite
differences (h = 1e-6, FP64), the single-point special case, a hand-ordered sort and the exact line search that
the rank-one step size must equal. GPU: libxtt's sets, values, gradients and ADF iterates against the oracle.
"""
import numpy as np
import pytest

from oracle import rank_one as R1


def _tiny(seed=0, dims=(2, 3, 2), ranks=(1, 2, 2, 1)):
    rng = np.random.default_rng(seed)
    cores = [rng.standard_normal((ranks[k], dims[k], ranks[k + 1])) for k in range(len(dims))]
    return cores, list(dims)


def test_oracle_value_brute_force():
    cores, dims = _tiny()
    rng = np.random.default_rng(1)
    U = rng.standard_normal((7, sum(dims)))
    full = R1.full(cores)
    o = R1.offsets(dims)
    for i in range(U.shape[0]):
        u = [U[i, o[k]:o[k + 1]] for k in range(3)]
        ref = np.einsum("abc,a,b,c->", full, *u)
        assert abs(R1.value(cores, U, dims)[i] - ref) < 1e-13 * max(1, abs(ref))


def test_oracle_gradient_finite_differences():
    cores, dims = _tiny(2)
    rng = np.random.default_rng(3)
    U = rng.standard_normal((9, sum(dims)))
    b = rng.standard_normal(9)
    h = 1e-6

    def loss(cs):
        return 0.5 * np.sum((R1.value(cs, U, dims) - b) ** 2)

    for k in range(3):
        D = R1.projected_gradient(cores, U, b, dims, k)
        fd = np.zeros_like(cores[k])
        for idx in np.ndindex(*cores[k].shape):
            cp = [c.copy() for c in cores]
            cm = [c.copy() for c in cores]
            cp[k][idx] += h
            cm[k][idx] -= h
            fd[idx] = (loss(cp) - loss(cm)) / (2 * h)
        # xerus's Δ is the negative gradient of ½Σ(x − b)²
        assert np.abs(D + fd).max() < 1e-7 * max(1.0, np.abs(fd).max())


def test_oracle_single_point_special_case():
    cores, dims = _tiny(4)
    full = R1.full(cores)
    pos = np.array([[0, 2, 1], [1, 0, 0], [1, 1, 1]])
    U, v = R1.from_single_point(pos, np.zeros(3), dims)
    assert np.allclose(R1.value(cores, U, dims), [full[tuple(p)] for p in pos], rtol=1e-14, atol=0)


def test_oracle_sort_order():
    # first differing entry decides, the larger entry first (internal::comp with the sort comparator)
    U = np.array([[0.5, 1.0], [-1.0, 9.0], [0.5, 2.0]])
    v = np.array([1.0, 2.0, 3.0])
    Us, vs = R1.sort_rows(U, v, positions_only=False)
    assert np.array_equal(Us, [[0.5, 2.0], [0.5, 1.0], [-1.0, 9.0]])
    assert np.array_equal(vs, [3.0, 1.0, 2.0])
    Us, vs = R1.sort_rows(U, v, positions_only=True)
    assert np.array_equal(vs, v)


def test_oracle_step_is_exact_line_search():
    """update_x's step ‖Δ‖²/‖AΔ‖² (adf.cpp:481-487) minimises ‖b − x(G_k + αΔ)‖² over α, since x is linear
    in G_k and Δ = Aᵀ(b − Ax)."""
    cores, dims = _tiny(5)
    rng = np.random.default_rng(6)
    U = rng.standard_normal((11, sum(dims)))
    b = rng.standard_normal(11)
    k = 1
    D = R1.projected_gradient(cores, U, b, dims, k)
    step = np.sum(D * D) / R1.step_denominator(cores, U, dims, k, D)

    def f(alpha):
        cs = list(cores)
        cs[k] = cores[k] + alpha * D
        return np.sum((b - R1.value(cs, U, dims)) ** 2)

    h = 1e-4 * step
    assert abs(f(step + h) - f(step - h)) / (2 * h) < 1e-8 * f(0.0) / step
    assert f(step) < f(0.0) and f(step) <= min(f(step * 0.9), f(step * 1.1))


def test_oracle_normalize():
    dims = [3, 2]
    rng = np.random.default_rng(8)
    U = rng.standard_normal((5, 5))
    v = rng.standard_normal(5)
    Un, vn = R1.normalize(U, v, dims)
    o = R1.offsets(dims)
    prod = np.ones(5)
    for k in range(2):
        assert np.allclose(np.linalg.norm(Un[:, o[k]:o[k + 1]], axis=1), 1.0)
        prod *= np.linalg.norm(U[:, o[k]:o[k + 1]], axis=1)
    assert np.allclose(vn * prod, v, rtol=1e-14)


# ---------------------------------------------------------------------------------------------------- GPU


def _tt(xe, dims, ranks, seed):
    xe.seed(seed)
    return xe.TTTensor.random(dims, ranks)


def _cores(t):
    return [t.get_component(k) for k in range(t.order())]


@pytest.mark.gpu
def test_random_set_matches_oracle(xe, oracle):
    dims = [4, 3, 5, 6]  # xerus requires Π n_k >= P for rank-one sets too (measurments.cpp:515)
    P = 300
    xe.seed(77)
    m = xe.RankOneMeasurementSet.random(P, dims)
    oracle.seed(77)
    U, _ = R1.random_set(P, dims, oracle.normal_draws(P * sum(dims)))
    assert np.array_equal(m.vectors(), U)  # same draws, same order: bit-exact
    assert m.size() == P and m.degree() == 4 and np.all(m.measuredValues == 0)
    with pytest.raises(xe.XttError):
        xe.RankOneMeasurementSet.random(361, dims)


@pytest.mark.gpu
def test_measure_test_normalize(xe, oracle):
    dims, ranks = [4, 3, 5, 6], [2, 3, 2]
    x = _tt(xe, dims, ranks, 5)
    m = xe.RankOneMeasurementSet.random(350, dims)
    m.measure(x)
    ref = R1.value(_cores(x), m.vectors(), dims)
    assert np.allclose(m.measuredValues, ref, rtol=1e-12, atol=1e-14)
    y = _tt(xe, dims, ranks, 6)
    assert abs(m.test(y) - R1.test(_cores(y), m.vectors(), m.measuredValues, dims)) < 1e-12
    U0, v0 = m.vectors(), m.measuredValues
    m.normalize()
    Un, vn = R1.normalize(U0, v0, dims)
    assert np.allclose(m.vectors(), Un, rtol=1e-15, atol=0) and np.allclose(m.measuredValues, vn, rtol=1e-15, atol=0)
    # sorting with values keeps (position, value) pairs together
    m.sort(False)
    Us, vs = R1.sort_rows(Un, vn, positions_only=False)
    assert np.array_equal(m.vectors(), Us) and np.array_equal(m.measuredValues, vs)


@pytest.mark.gpu
@pytest.mark.parametrize("precision,tol", [("f64", 1e-11), ("f32", 2e-4)])
def test_gradient_matches_oracle(xe, oracle, precision, tol):
    print(f"rank-one gradient {precision}", flush=True)
    dims, ranks = [6, 5, 4, 7, 3], [2, 4, 3, 2]
    x = _tt(xe, dims, ranks, 9)
    truth = _tt(xe, dims, ranks, 10)
    m = xe.RankOneMeasurementSet.random(2000, truth)
    g, rn = xe.adf_gradient(x, m, xe.XT_F64 if precision == "f64" else xe.XT_F32)
    cores = _cores(x)
    ref_rn = R1.test(cores, m.vectors(), m.measuredValues, dims)
    assert abs(rn - ref_rn) < tol * max(1, ref_rn)
    for k in range(len(dims)):
        ref = R1.projected_gradient(cores, m.vectors(), m.measuredValues, dims, k)
        assert np.abs(g[k] - ref).max() < tol * np.abs(ref).max(), k


@pytest.mark.gpu
def test_single_point_conversion(xe):
    dims, ranks = [4, 3, 5], [2, 3]
    truth = _tt(xe, dims, ranks, 12)
    sp = xe.SinglePointMeasurementSet.random(50, truth)
    r1 = xe.RankOneMeasurementSet(sp, dims)
    assert np.array_equal(r1.measuredValues, sp.measuredValues)
    x = _tt(xe, dims, ranks, 13)
    g1, rn1 = xe.adf_gradient(x, sp)
    g2, rn2 = xe.adf_gradient(x, r1)
    assert abs(rn1 - rn2) < 1e-13
    for a, b in zip(g1, g2):
        assert np.abs(a - b).max() < 1e-12 * max(1, np.abs(a).max())


@pytest.mark.gpu
def test_adf_rank_one_iterates_match_oracle(xe):
    print("rank-one ADF 3 iterations", flush=True)
    dims, ranks = [5, 4, 6, 3], [2, 3, 2]
    truth = _tt(xe, dims, ranks, 20)
    m = xe.RankOneMeasurementSet.random(300, truth)
    x = _tt(xe, dims, ranks, 21)
    x0 = _cores(x)
    adf = xe.ADFVariant(3, 1e-12, 0.999)
    res = adf(x, m, ranks)
    cores, ref_res, it = R1.adf_solve(x0, m.vectors(), m.measuredValues, dims, 3, 1e-12, 0.999)
    assert adf.last_iterations == it == 3
    assert abs(res - ref_res) < 1e-10 * ref_res
    fx, fr = x.full(), R1.full(cores)
    assert np.linalg.norm(fx - fr) < 1e-9 * np.linalg.norm(fr)


@pytest.mark.gpu
def test_adf_rank_one_recovers_low_rank(xe):
    """Algorithm:adf_random_low_rank on a RankOneMeasurementSet built from single points (ttCompletion.cxx:106-137)."""
    dims, ranks = [5] * 5, [2] * 4
    truth = _tt(xe, dims, ranks, 30)
    sp = xe.SinglePointMeasurementSet.random(5 * 5 * 10 * 2 * 2, truth)  # D·N·CS·R² as the reference test
    r1 = xe.RankOneMeasurementSet(sp, dims)
    x = _tt(xe, dims, ranks, 31)
    xe.ADFVariant(500, 1e-6, 0.999)(x, r1, ranks)
    err = np.linalg.norm(x.full() - truth.full()) / np.linalg.norm(truth.full())
    assert err < 1e-3, err
