// Compiled against include/xtt_xerus.hpp and libxtt: config 1 (d=8, n=4 Dirichlet Laplacian, rank-2 b normalised,
// rank-8 start, seed 0xBAADF00D) through the xerus-named C++ API, written as xerus user code would be. Prints the
// values the GPU test compares with the oracle; exits non-zero on an exception.
#include <cmath>
#include <cstdio>
#include <vector>

#include "xtt_xerus.hpp"

using namespace xerus;

static TTOperator laplace(size_t d, size_t n) {
	TTOperator A(std::vector<size_t>(2 * d, n));
	for (size_t k = 0; k < d; ++k) {
		const size_t rl = k == 0 ? 1 : 2, rr = k + 1 == d ? 1 : 2;
		Tensor c({rl, n, n, rr});
		auto at = [&](size_t a, size_t i, size_t j, size_t b) -> value_t& { return c[((a * n + i) * n + j) * rr + b]; };
		for (size_t i = 0; i < n; ++i)
			for (size_t j = 0; j < n; ++j) {
				const value_t L = i == j ? 2.0 : (i + 1 == j || j + 1 == i ? -1.0 : 0.0), I = i == j ? 1.0 : 0.0;
				// cores [L I], [[I,0],[L,I]], [I;L]
				if (d == 1) { at(0, i, j, 0) = L; continue; }
				if (k == 0) { at(0, i, j, 0) = L; at(0, i, j, 1) = I; }
				else if (k + 1 == d) { at(0, i, j, 0) = I; at(1, i, j, 0) = L; }
				else { at(0, i, j, 0) = I; at(1, i, j, 0) = L; at(1, i, j, 1) = I; }
			}
		A.set_component(k, c);
	}
	return A;
}

int main() {
	try {
		const size_t d = 8, n = 4;
		xt_seed(0xBAADF00Dull);
		TTOperator A = laplace(d, n);
		TTTensor b = TTTensor::random(std::vector<size_t>(d, n), std::vector<size_t>(d - 1, 2));
		b /= frob_norm(b);
		TTTensor x = TTTensor::random(std::vector<size_t>(d, n), std::vector<size_t>(d - 1, 8));
		ALSVariant als = ALS_SPD;
		als.convergenceEpsilon = 0.0;
		als.localTolerance = 1e-14;
		PerformanceData perf;
		const value_t energy = als(A, x, b, size_t(10), perf);
		double res = 0;
		if (xt_tt_residual(A.handle(), x.handle(), b.handle(), &res) != XT_OK) throw misc::generic_error(xt_last_error());
		std::printf("energy %.17g\nresidual %.17g\nranks", energy, res);
		for (size_t r : x.ranks()) std::printf(" %zu", r);
		std::printf("\ncore %zu %d\nperf_points %zu\n", x.corePosition, x.canonicalized ? 1 : 0, perf.data.size());
		// round / move_core / sum through the same face
		TTTensor y = x + x;
		y.round(1e-12);
		std::printf("sum_ranks");
		for (size_t r : y.ranks()) std::printf(" %zu", r);
		std::printf("\nsum_norm_ratio %.17g\n", frob_norm(y) / frob_norm(x));
		// TT-SVD of a dense rank-1 tensor and TTTensor::ones
		Tensor ones3({3, 4, 5});
		for (auto& v : ones3.data) v = 1.0;
		TTTensor t(ones3);
		TTTensor o = TTTensor::ones({3, 4, 5});
		std::printf("from_tensor_ranks %zu %zu\nones_diff %.3g\n", t.ranks()[0], t.ranks()[1], frob_norm(t - o));
		// ADF on single points from the same engine
		TTTensor truth = TTTensor::random({5, 5, 5, 5}, {2, 2, 2});
		SinglePointMeasurementSet m = SinglePointMeasurementSet::random(400, {5, 5, 5, 5});
		m.measure(truth);
		TTTensor z = TTTensor::random({5, 5, 5, 5}, {2, 2, 2});
		const value_t adfRes = ADFVariant(30, 1e-10, 0.9999)(z, m);
		std::printf("adf_residual %.17g\n", adfRes);
		// errors come back as generic_error
		try {
			TTTensor bad = TTTensor::random({2, 2}, {1});
			bad.set_component(1, Tensor({3, 2, 1}));
			bad.move_core(1);
			std::printf("format_check missing\n");
			return 2;
		} catch (const misc::generic_error& e) {
			std::printf("format_check ok\n");
		}
		return 0;
	} catch (const std::exception& e) {
		std::fprintf(stderr, "error: %s\n", e.what());
		return 1;
	}
}
