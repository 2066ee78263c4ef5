// TEST INFRASTRUCTURE ONLY — C entry points of the oracle for ctypes (tests/, smoke(), bench.py's cpu_baseline).
#include "xtt_oracle.hpp"

#include <cstring>
#include <exception>
#include <string>

using namespace orc;

namespace {
thread_local std::string g_err;
template <class F> int guard(F&& f) {
	try { f(); return 0; } catch (const std::exception& e) { g_err = e.what(); return -1; }
}
TT* tt(void* h) { return static_cast<TT*>(h); }
SinglePointMeasurementSet* ms(void* h) { return static_cast<SinglePointMeasurementSet*>(h); }
dims_t mk(const size_t* p, size_t n) { return dims_t(p, p + n); }
}

extern "C" {

const char* or_last_error() { return g_err.c_str(); }
void or_seed(uint64_t s) { seed(s); }
void or_normal_draws(size_t n, double* out) { for (size_t i = 0; i < n; ++i) out[i] = normal()(rng()); }
void or_uniform_draws(size_t n, size_t hi, size_t* out) { std::uniform_int_distribution<size_t> u(0, hi); for (size_t i = 0; i < n; ++i) out[i] = u(rng()); }
void or_raw_draws(size_t n, uint64_t* out) { for (size_t i = 0; i < n; ++i) out[i] = rng()(); }

// ---- dense kernels
int or_gemm(int ta, int tb, size_t m, size_t n, size_t k, double alpha, const double* A, const double* B, double* C) {
	return guard([&] { gemm(ta, tb, m, n, k, alpha, A, B, C); });
}
int or_qr(const double* A, size_t m, size_t n, double* Q, double* R) {
	return guard([&] { std::vector<double> q, r; qr(A, m, n, q, r); std::memcpy(Q, q.data(), q.size() * 8); std::memcpy(R, r.data(), r.size() * 8); });
}
int or_rq(const double* A, size_t m, size_t n, double* R, double* Q) {
	return guard([&] { std::vector<double> q, r; rq(A, m, n, r, q); std::memcpy(Q, q.data(), q.size() * 8); std::memcpy(R, r.data(), r.size() * 8); });
}
int or_qc(const double* A, size_t m, size_t n, double* Q, double* C, size_t* rank) {
	return guard([&] { std::vector<double> q, c; *rank = qc(A, m, n, q, c); std::memcpy(Q, q.data(), q.size() * 8); std::memcpy(C, c.data(), c.size() * 8); });
}
int or_cq(const double* A, size_t m, size_t n, double* C, double* Q, size_t* rank) {
	return guard([&] { std::vector<double> q, c; *rank = cq(A, m, n, c, q); std::memcpy(Q, q.data(), q.size() * 8); std::memcpy(C, c.data(), c.size() * 8); });
}
int or_svd(const double* A, size_t m, size_t n, double* U, double* S, double* Vt) {
	return guard([&] { std::vector<double> u, s, vt; svd(A, m, n, u, s, vt);
		std::memcpy(U, u.data(), u.size() * 8); std::memcpy(S, s.data(), s.size() * 8); std::memcpy(Vt, vt.data(), vt.size() * 8); });
}
int or_solve(const double* A, size_t n, const double* b, double* x, int* method) {
	return guard([&] { *method = solve(A, n, b, x); });
}
int or_reshuffle(size_t deg, const size_t* dims, const size_t* shuffle, const double* in, double* out) {
	return guard([&] { Tensor t(mk(dims, deg)); std::memcpy(t.data(), in, t.size() * 8);
		Tensor r = reshuffle(t, mk(shuffle, deg)); std::memcpy(out, r.data(), r.size() * 8); });
}
int or_contract(size_t ld, const size_t* ldims, const double* l, int lt, size_t rd, const size_t* rdims, const double* r, int rt,
                size_t modes, double* out) {
	return guard([&] { Tensor a(mk(ldims, ld)), b(mk(rdims, rd)); std::memcpy(a.data(), l, a.size() * 8); std::memcpy(b.data(), r, b.size() * 8);
		Tensor c = contract(a, lt, b, rt, modes); std::memcpy(out, c.data(), c.size() * 8); });
}
int or_svd_truncated(size_t deg, const size_t* dims, const double* in, size_t split, size_t maxRank, double eps, size_t* rank, double* U, double* S, double* Vt) {
	return guard([&] { Tensor t(mk(dims, deg)); std::memcpy(t.data(), in, t.size() * 8);
		Tensor u, s, vt; calculate_svd(u, s, vt, t, split, maxRank, eps); *rank = s.dims[0];
		std::memcpy(U, u.data(), u.size() * 8); for (size_t i = 0; i < *rank; ++i) S[i] = s.v[i * *rank + i];
		std::memcpy(Vt, vt.data(), vt.size() * 8); });
}

// ---- TT handles
void* or_tt_new(int op, size_t nd, const size_t* dims) { try { return new TT(tt_zero(mk(dims, nd), op)); } catch (const std::exception& e) { g_err = e.what(); return nullptr; } }
void* or_tt_random(int op, size_t nd, const size_t* dims, const size_t* ranks) {
	try { const size_t d = nd / (op ? 2 : 1); return new TT(tt_random(mk(dims, nd), mk(ranks, d ? d - 1 : 0), op)); }
	catch (const std::exception& e) { g_err = e.what(); return nullptr; }
}
void* or_tt_copy(void* h) { return new TT(*tt(h)); }
void or_tt_free(void* h) { delete tt(h); }
size_t or_tt_order(void* h) { return tt(h)->order(); }
int or_tt_is_operator(void* h) { return tt(h)->op; }
void or_tt_dims(void* h, size_t* out) { std::memcpy(out, tt(h)->dims.data(), tt(h)->dims.size() * sizeof(size_t)); }
void or_tt_ranks(void* h, size_t* out) { auto r = tt(h)->ranks(); if (!r.empty()) std::memcpy(out, r.data(), r.size() * sizeof(size_t)); }
void or_tt_core_dims(void* h, size_t k, size_t* out) { const auto& d = tt(h)->cores[k].dims; std::memcpy(out, d.data(), d.size() * sizeof(size_t)); }
void or_tt_get_core(void* h, size_t k, double* out) { const auto& c = tt(h)->cores[k]; std::memcpy(out, c.data(), c.size() * 8); }
int or_tt_set_core(void* h, size_t k, size_t nd, const size_t* dims, const double* data) {
	return guard([&] { Tensor t(mk(dims, nd)); std::memcpy(t.data(), data, t.size() * 8); tt(h)->set_component(k, std::move(t)); });
}
void or_tt_state(void* h, int* canon, size_t* pos) { *canon = tt(h)->canonicalized; *pos = tt(h)->corePosition; }
void or_tt_set_state(void* h, int canon, size_t pos) { tt(h)->canonicalized = canon; tt(h)->corePosition = pos; }
int or_tt_move_core(void* h, size_t pos, int keepRank) { return guard([&] { move_core(*tt(h), pos, keepRank); }); }
int or_tt_round(void* h, const size_t* maxRanks, double eps) {
	return guard([&] { round(*tt(h), mk(maxRanks, tt(h)->order() - 1), eps); });
}
double or_tt_frob_norm(void* h) { return tt_frob_norm(*tt(h)); }
double or_tt_dot(void* a, void* b) { return tt_dot(*tt(a), *tt(b)); }
void* or_tt_add(void* a, void* b) { try { return new TT(tt_add(*tt(a), *tt(b))); } catch (const std::exception& e) { g_err = e.what(); return nullptr; } }
void or_tt_scale(void* h, double s) { tt_scale(*tt(h), s); }
void* or_tt_apply(void* A, void* x) { try { return new TT(tt_apply(*tt(A), *tt(x))); } catch (const std::exception& e) { g_err = e.what(); return nullptr; } }
int or_tt_full(void* h, double* out) { return guard([&] { Tensor t = tt_full(*tt(h)); std::memcpy(out, t.data(), t.size() * 8); }); }
double or_tt_residual(void* A, void* x, void* b) { return tt_residual(*tt(A), *tt(x), *tt(b)); }
void* or_tt_from_tensor(int op, size_t deg, const size_t* dims, const double* data, double eps, const size_t* maxRanks) {
	try {
		Tensor t(mk(dims, deg)); std::memcpy(t.data(), data, t.size() * 8);
		const size_t d = deg / (op ? 2 : 1);
		return new TT(tt_from_tensor(t, op, eps, maxRanks ? mk(maxRanks, d ? d - 1 : 0) : dims_t{}));
	} catch (const std::exception& e) { g_err = e.what(); return nullptr; }
}
void* or_tt_special(int kind, int op, size_t nd, const size_t* dims, const size_t* position) {
	try {
		switch (kind) {
		case 0: return new TT(tt_ones(mk(dims, nd), op));
		case 1: return new TT(tt_identity(mk(dims, nd)));
		case 2: return new TT(tt_kronecker(mk(dims, nd), op));
		default: return new TT(tt_dirac(mk(dims, nd), mk(position, nd), op));
		}
	} catch (const std::exception& e) { g_err = e.what(); return nullptr; }
}
int or_tt_transpose(void* h) { return guard([&] { tt_transpose(*tt(h)); }); }

int or_als(void* A, void* x, void* b, unsigned sites, size_t numHalfSweeps, double eps, int useResidual, int preserveCore,
           int assumeSPD, int localSolver, double cgTol, size_t cgMaxIter, double* energy, size_t* stats3) {
	return guard([&] {
		ALSOptions o; o.sites = sites; o.numHalfSweeps = numHalfSweeps; o.convergenceEpsilon = eps;
		o.useResidualForEndCriterion = useResidual; o.preserveCorePosition = preserveCore; o.assumeSPD = assumeSPD;
		o.localSolver = localSolver; o.cgTol = cgTol; o.cgMaxIter = cgMaxIter;
		ALSStats st;
		*energy = als_solve(o, A ? tt(A) : nullptr, *tt(x), *tt(b), &st);
		if (stats3) { stats3[0] = st.halfSweeps; stats3[1] = st.localSolves; stats3[2] = st.cgIterations; }
	});
}

// ---- measurements / ADF
void* or_meas_random(size_t P, size_t d, const size_t* dims) {
	try { return new SinglePointMeasurementSet(meas_random(P, mk(dims, d))); } catch (const std::exception& e) { g_err = e.what(); return nullptr; }
}
void* or_meas_new(size_t P, size_t d, const uint32_t* pos, const double* values) {
	auto* m = new SinglePointMeasurementSet; m->d = d; m->pos.assign(pos, pos + P * d); m->values.assign(values, values + P); return m;
}
void or_meas_free(void* h) { delete ms(h); }
size_t or_meas_size(void* h) { return ms(h)->size(); }
void or_meas_positions(void* h, uint32_t* out) { std::memcpy(out, ms(h)->pos.data(), ms(h)->pos.size() * 4); }
void or_meas_values(void* h, double* out) { std::memcpy(out, ms(h)->values.data(), ms(h)->values.size() * 8); }
void or_meas_set_values(void* h, const double* v) { std::memcpy(ms(h)->values.data(), v, ms(h)->values.size() * 8); }
int or_meas_measure(void* h, void* x) { return guard([&] { meas_measure(*ms(h), *tt(x)); }); }
int or_adf_plan(void* h, size_t* fwdCounts, size_t* bwdCounts, uint32_t* fwdIds, uint32_t* bwdIds, size_t* numUnique) {
	return guard([&] {
		ADFStacks s = adf_stack_plan(*ms(h));
		size_t of = 0, ob = 0;
		for (size_t k = 0; k < s.fwd.size(); ++k) {
			fwdCounts[k] = s.fwd[k].size(); bwdCounts[k] = s.bwd[k].size();
			if (fwdIds) { std::memcpy(fwdIds + of, s.fwd[k].data(), s.fwd[k].size() * 4); }
			if (bwdIds) { std::memcpy(bwdIds + ob, s.bwd[k].data(), s.bwd[k].size() * 4); }
			of += s.fwd[k].size(); ob += s.bwd[k].size();
		}
		*numUnique = s.numUnique;
	});
}
int or_adf_gradient(void* x, void* h, double* out, double* residualNorm) {
	return guard([&] {
		auto g = adf_full_gradient(*tt(x), *ms(h), residualNorm);
		size_t o = 0;
		for (auto& t : g) { std::memcpy(out + o, t.data(), t.size() * 8); o += t.size(); }
	});
}
int or_adf_core_step(void* x, void* h, size_t k) { return guard([&] { adf_core_step(*tt(x), *ms(h), k); }); }
int or_adf_solve(void* x, void* h, const size_t* maxRanks, size_t maxIt, double target, double minDec, double* residual, size_t* iterations) {
	return guard([&] {
		ADFOptions o; o.maxIterations = maxIt; o.targetResidualNorm = target; o.minimalResidualNormDecrease = minDec;
		dims_t mr = maxRanks ? mk(maxRanks, tt(x)->order() - 1) : tt(x)->ranks();
		*residual = adf_solve(o, *tt(x), *ms(h), mr, iterations);
	});
}

} // extern "C"
