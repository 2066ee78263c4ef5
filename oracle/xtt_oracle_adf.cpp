// TEST INFRASTRUCTURE ONLY — the oracle: SinglePointMeasurementSet (src/xerus/measurments.cpp) and ADF
// (src/xerus/algorithms/adf.cpp) restated in FP64. Stack entries are evaluated per measurement; the
// deduplicated stack plan of adf.cpp:102-191 is restated separately (adf_stack_plan) because its
// representative ids and group counts are integer outputs checked bit-exactly.
#include "xtt_oracle.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <numeric>
#include <set>
#include <stdexcept>

namespace orc {

SinglePointMeasurementSet meas_random(size_t P, const dims_t& dims) {  // measurments.cpp:222-248
	size_t total = 1;
	for (size_t n : dims) total *= n;
	if (total < P) throw std::invalid_argument("It's impossible to perform as many measurements as requested.");
	std::vector<std::uniform_int_distribution<size_t>> dist;
	for (size_t n : dims) dist.emplace_back(0, n - 1);
	std::set<size_t> seen;
	std::vector<std::vector<size_t>> positions;
	std::vector<size_t> idx(dims.size());
	while (positions.size() < P) {
		size_t pos = 0;
		for (size_t i = 0; i < dims.size(); ++i) { idx[i] = dist[i](rng()); pos *= dims[i]; pos += idx[i]; }
		if (seen.insert(pos).second) positions.push_back(idx);
	}
	std::sort(positions.begin(), positions.end());  // lexicographic, measurments.cpp:89-105
	SinglePointMeasurementSet m;
	m.d = dims.size();
	m.pos.resize(P * m.d);
	for (size_t i = 0; i < P; ++i) for (size_t k = 0; k < m.d; ++k) m.pos[i * m.d + k] = (uint32_t)positions[i][k];
	m.values.assign(P, 0.0);
	return m;
}

namespace {
// F_k[i] (length r_k) and B_k[i] (length r_{k-1}) for every measurement.
struct Stacks {
	size_t P, d;
	std::vector<std::vector<double>> F, B;  // F[k][i*r_k + ·], B[k][i*r_{k-1} + ·]
};
void forward_stack(const TT& x, const SinglePointMeasurementSet& m, std::vector<std::vector<double>>& F, size_t k) {
	const Tensor& G = x.cores[k];
	const size_t rl = G.dims[0], n = G.dims[1], rr = G.dims[2], P = m.size(), d = m.d;
	F[k].assign(P * rr, 0.0);
	#pragma omp parallel for schedule(static)
	for (size_t i = 0; i < P; ++i) {
		const size_t j = m.pos[i * d + k];
		const double* f = k == 0 ? nullptr : F[k - 1].data() + i * rl;
		double* o = F[k].data() + i * rr;
		for (size_t a = 0; a < rl; ++a) {
			const double fa = f ? f[a] : 1.0;
			const double* g = G.data() + (a * n + j) * rr;
			for (size_t c = 0; c < rr; ++c) o[c] += fa * g[c];
		}
	}
}
void backward_stack(const TT& x, const SinglePointMeasurementSet& m, std::vector<std::vector<double>>& B, size_t k) {
	const Tensor& G = x.cores[k];
	const size_t rl = G.dims[0], n = G.dims[1], rr = G.dims[2], P = m.size(), d = m.d;
	B[k].assign(P * rl, 0.0);
	#pragma omp parallel for schedule(static)
	for (size_t i = 0; i < P; ++i) {
		const size_t j = m.pos[i * d + k];
		const double* bb = k + 1 == d ? nullptr : B[k + 1].data() + i * rr;
		double* o = B[k].data() + i * rl;
		for (size_t a = 0; a < rl; ++a) {
			const double* g = G.data() + (a * n + j) * rr;
			double s = 0;
			for (size_t c = 0; c < rr; ++c) s += g[c] * (bb ? bb[c] : 1.0);
			o[a] = s;
		}
	}
}
// value_i = F_{k-1}[i]·G_k[:, i_k, :]·B_{k+1}[i]
void evaluate_at(const TT& x, const Tensor& G, const SinglePointMeasurementSet& m, const std::vector<std::vector<double>>& F,
                 const std::vector<std::vector<double>>& B, size_t k, std::vector<double>& out) {
	const size_t rl = G.dims[0], n = G.dims[1], rr = G.dims[2], P = m.size(), d = m.d;
	out.resize(P);
	#pragma omp parallel for schedule(static)
	for (size_t i = 0; i < P; ++i) {
		const size_t j = m.pos[i * d + k];
		const double* f = k == 0 ? nullptr : F[k - 1].data() + i * rl;
		const double* bb = k + 1 == d ? nullptr : B[k + 1].data() + i * rr;
		double s = 0;
		for (size_t a = 0; a < rl; ++a) {
			const double* g = G.data() + (a * n + j) * rr;
			double t = 0;
			for (size_t c = 0; c < rr; ++c) t += g[c] * (bb ? bb[c] : 1.0);
			s += (f ? f[a] : 1.0) * t;
		}
		out[i] = s;
	}
	(void)x;
}
// Δ_k (r_{k-1}, n_k, r_k) = Σ_i res_i·F_{k-1}[i] ⊗ B_{k+1}[i] at slice i_k (adf.cpp:314-396, reshuffled as :395).
Tensor projected_gradient(const TT& x, const SinglePointMeasurementSet& m, const std::vector<std::vector<double>>& F,
                          const std::vector<std::vector<double>>& B, const std::vector<double>& res, size_t k) {
	const Tensor& G = x.cores[k];
	const size_t rl = G.dims[0], n = G.dims[1], rr = G.dims[2], P = m.size(), d = m.d;
	Tensor D({rl, n, rr});
	for (size_t i = 0; i < P; ++i) {
		const size_t j = m.pos[i * d + k];
		const double* f = k == 0 ? nullptr : F[k - 1].data() + i * rl;
		const double* bb = k + 1 == d ? nullptr : B[k + 1].data() + i * rr;
		for (size_t a = 0; a < rl; ++a) {
			const double fa = res[i] * (f ? f[a] : 1.0);
			double* o = D.data() + (a * n + j) * rr;
			for (size_t c = 0; c < rr; ++c) o[c] += fa * (bb ? bb[c] : 1.0);
		}
	}
	return D;
}
double sqnorm(const std::vector<double>& v) { long double s = 0; for (double x : v) s += (long double)x * x; return (double)s; }

// One ADF core step at component k (adf.cpp:360-396 gradient, :413-465 slice norms, :468-478 update_x): with the
// residual res = b − x on the samples and the stacks F_{k-1}, B_{k+1} of x, Δ = projected gradient, and every
// slice j of the core moves by ‖Δ_j‖²/‖P_j A Δ‖² · Δ_j, the exact minimiser of ½‖P_Ω(x) − b‖² along Δ_j (the
// samples with i_k = j are the only ones the slice reaches).
void core_step(TT& x, const SinglePointMeasurementSet& m, const std::vector<std::vector<double>>& F,
               const std::vector<std::vector<double>>& B, const std::vector<double>& res, size_t k) {
	const size_t P = m.size(), d = m.d;
	Tensor D = projected_gradient(x, m, F, B, res, k);
	const size_t n = D.dims[1], rl = D.dims[0], rr = D.dims[2];
	std::vector<double> val;
	evaluate_at(x, D, m, F, B, k, val);
	std::vector<double> normA(n, 0.0);
	for (size_t i = 0; i < P; ++i) normA[m.pos[i * d + k]] += val[i] * val[i];
	Tensor& G = x.cores[k];
	for (size_t j = 0; j < n; ++j) {
		double pyr = 0;
		for (size_t a = 0; a < rl; ++a) for (size_t c = 0; c < rr; ++c) { const double t = D.v[(a * n + j) * rr + c]; pyr += t * t; }
		const double step = pyr / normA[j];
		for (size_t a = 0; a < rl; ++a) for (size_t c = 0; c < rr; ++c) G.v[(a * n + j) * rr + c] += step * D.v[(a * n + j) * rr + c];
	}
}
} // namespace

void adf_core_step(TT& x, const SinglePointMeasurementSet& m, size_t k) {
	const size_t d = x.order();
	if (k >= d || m.d != d) throw std::invalid_argument("adf_core_step: component or degree mismatch");
	std::vector<std::vector<double>> F(d), B(d);
	for (size_t j = 0; j < k; ++j) forward_stack(x, m, F, j);
	for (size_t j = d; j-- > k + 1;) backward_stack(x, m, B, j);
	std::vector<double> val, res(m.size());
	evaluate_at(x, x.cores[k], m, F, B, k, val);
	for (size_t i = 0; i < m.size(); ++i) res[i] = m.values[i] - val[i];
	core_step(x, m, F, B, res, k);
}

void meas_measure(SinglePointMeasurementSet& m, const TT& x) {  // measurments.cpp:125-153
	std::vector<std::vector<double>> F(x.order()), B(x.order());
	for (size_t k = 0; k + 1 < x.order(); ++k) forward_stack(x, m, F, k);
	evaluate_at(x, x.cores.back(), m, F, B, x.order() - 1, m.values);
}

// adf.cpp:102-191: forward groups = equal prefixes (i_0..i_k) in lexicographic order, backward groups = equal
// suffixes (i_k..i_{d-1}); each group is represented by its smallest measurement id, listed in ascending order.
ADFStacks adf_stack_plan(const SinglePointMeasurementSet& m) {
	ADFStacks s;
	const size_t P = m.size(), d = m.d;
	s.fwd.resize(d); s.bwd.resize(d);
	for (int dir = 0; dir < 2; ++dir) {
		const bool forward = dir == 0;
		std::vector<uint32_t> order(P);
		std::iota(order.begin(), order.end(), 0u);
		auto key = [&](uint32_t a, uint32_t b) {
			for (size_t t = 0; t < d; ++t) {
				const size_t j = forward ? t : d - 1 - t;
				if (m.pos[a * d + j] != m.pos[b * d + j]) return m.pos[a * d + j] < m.pos[b * d + j];
			}
			return false;
		};
		std::sort(order.begin(), order.end(), key);
		auto& upd = forward ? s.fwd : s.bwd;
		for (size_t k = 0; k < d; ++k) {
			std::vector<uint32_t> reps;
			uint32_t groupMin = order.empty() ? 0 : order[0];
			for (size_t t = 1; t <= P; ++t) {
				bool same = false;
				if (t < P) {
					same = true;
					// shared prefix positions 0..k (forward) or suffix positions k..d-1 (backward)
					const size_t lo = forward ? 0 : k, hi = forward ? k + 1 : d;
					for (size_t j = lo; j < hi; ++j) if (m.pos[order[t] * d + j] != m.pos[order[t - 1] * d + j]) { same = false; break; }
				}
				if (same) { groupMin = std::min(groupMin, order[t]); }
				else { reps.push_back(groupMin); if (t < P) groupMin = order[t]; }
			}
			std::sort(reps.begin(), reps.end());
			s.numUnique += reps.size();
			upd[k] = std::move(reps);
		}
	}
	s.numUnique = 0;
	for (size_t k = 0; k < d; ++k) s.numUnique += s.fwd[k].size();
	s.numUnique += 1;  // the shared slot for positions -1 and d (adf.cpp:162-166)
	return s;
}

std::vector<Tensor> adf_full_gradient(const TT& x, const SinglePointMeasurementSet& m, double* residualNorm) {
	const size_t d = x.order();
	std::vector<std::vector<double>> F(d), B(d);
	for (size_t k = 0; k + 1 < d; ++k) forward_stack(x, m, F, k);
	for (size_t k = d; k-- > 1;) backward_stack(x, m, B, k);
	std::vector<double> val, res(m.size());
	evaluate_at(x, x.cores[0], m, F, B, 0, val);
	for (size_t i = 0; i < m.size(); ++i) res[i] = m.values[i] - val[i];
	if (residualNorm) *residualNorm = std::sqrt(sqnorm(res) / sqnorm(m.values));
	std::vector<Tensor> g;
	for (size_t k = 0; k < d; ++k) g.push_back(projected_gradient(x, m, F, B, res, k));
	return g;
}

double adf_solve(const ADFOptions& o, TT& x, const SinglePointMeasurementSet& m, const dims_t& maxRanksIn, size_t* iterations) {
	const size_t d = x.order(), P = m.size();
	if (P == 0) throw std::invalid_argument("Need at very least one measurment.");
	if (m.d != d) throw std::invalid_argument("Measurment degree must coincide with x degree.");
	const dims_t maxRanks = reduce_to_maximal_ranks(maxRanksIn, x.dims, false);
	const double normMeasured = std::sqrt(sqnorm(m.values));
	const ADFStacks plan = adf_stack_plan(m);
	std::vector<std::vector<double>> F(d), B(d);
	std::vector<double> res(P), val;
	size_t iteration = 0;
	double residualNorm = DBL_MAX, lastResidualNorm = DBL_MAX;

	auto calculate_residual = [&](size_t k) {  // adf.cpp:290-312 (same stack side choice)
		const bool fwdSide = plan.fwd[k].size() < plan.bwd[k].size();
		if (fwdSide) forward_stack(x, m, F, k); else backward_stack(x, m, B, k);
		evaluate_at(x, x.cores[k], m, F, B, k, val);
		for (size_t i = 0; i < P; ++i) res[i] = m.values[i] - val[i];
	};

	auto solve_with_current_ranks = [&]() {  // adf.cpp:489-542
		double resDec1 = 0.0, resDec2 = 0.0, resDec3 = 0.0;
		for (; o.maxIterations == 0 || iteration < o.maxIterations; ++iteration) {
			move_core(x, 0, true);
			for (size_t k = d - 1; k > 0; --k) backward_stack(x, m, B, k);
			calculate_residual(0);
			lastResidualNorm = residualNorm;
			residualNorm = std::sqrt(sqnorm(res)) / normMeasured;
			const double resDec4 = resDec3; resDec3 = resDec2; resDec2 = resDec1;
			resDec1 = residualNorm / lastResidualNorm;
			if (residualNorm < o.targetResidualNorm || resDec1 * resDec2 * resDec3 * resDec4 > std::pow(o.minimalResidualNormDecrease, 4)) break;
			for (size_t k = 0; k < d; ++k) {
				if (k > 0) calculate_residual(k);
				core_step(x, m, F, B, res, k);
				if (k + 1 < d) { move_core(x, k + 1, true); forward_stack(x, m, F, k); }
			}
		}
	};

	move_core(x, 0, false);  // canonicalize_left
	solve_with_current_ranks();
	while (residualNorm > o.targetResidualNorm && x.ranks() != maxRanks && (o.maxIterations == 0 || iteration < o.maxIterations)) {
		move_core(x, 0, true);  // adf.cpp:590-602
		TT rnd = tt_random(x.dims, dims_t(d - 1, 1), false);
		const double scale = 1e-6 * tt_frob_norm(x) / tt_frob_norm(rnd);
		tt_scale(rnd, scale);
		x = tt_add(x, rnd);
		round(x, maxRanks, 8 * DBL_EPSILON);
		solve_with_current_ranks();
	}
	if (iterations) *iterations = iteration;
	return residualNorm;
}

} // namespace orc
