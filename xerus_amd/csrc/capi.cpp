// extern "C" surface of libxtt (declared in include/xtt.h). Exceptions become status codes here.
#include <cstring>
#include <fstream>
#include <sstream>
#include <map>
#include <mutex>
#include <memory>

#include "../../include/xtt.h"
#include "adf_driver.hpp"
#include "tt.hpp"

struct xt_tt_s { xtt::TT t; };
struct xt_meas_s { xtt::Measurements m; };
struct xt_r1meas_s { xtt::RankOneMeasurements m; };

namespace xtt { const std::string& last_error(); }

namespace {
template <class F>
int guard(F&& f) {
	(void)hipGetLastError();  // a failed call before this one must not be reported by this one's launch checks
	try {
		f();
		return XT_OK;
	} catch (const xtt::Error& e) {
		xtt::set_error(e.what());
		return e.code;
	} catch (const std::bad_alloc&) {
		xtt::set_error("host allocation failed");
		return XT_ERR_NOMEM;
	} catch (const std::exception& e) {
		xtt::set_error(e.what());
		return XT_ERR_INVALID;
	}
}
#define NN(p) XTT_REQUIRE((p) != nullptr, #p " must not be NULL")
// a TT handed to device work must be in xerus's format (TTNetwork::require_correct_format, ttNetwork.cpp:290-340):
// the kernels derive strides from the rank chain, so a component set with an inconsistent rank is refused here
#define FMT(h) do { NN(h); ::xtt::require_correct_format((h)->t); } while (0)
xtt::dims_t mk(const size_t* p, size_t n) { return xtt::dims_t(p, p + n); }
void sync_stream(void* s) { XTT_HIP(hipStreamSynchronize(xtt::pick(s))); }
}  // namespace

namespace {
xtt::GemmDesc tune_desc(int precision, size_t m, size_t n, size_t k, size_t batch, int ta, int tb) {
	XTT_REQUIRE(precision == XT_F64 || precision == XT_F32, "GEMM tuning: precision XT_F64 or XT_F32");
	XTT_REQUIRE(m && n && k && batch, "GEMM tuning: empty shape");
	xtt::GemmDesc g = xtt::gemm_desc(ta ? 1 : 0, tb ? 1 : 0, m, n, k);
	g.batch = batch;
	return g;
}
}  // namespace

extern "C" {

const char* xt_last_error(void) { return xtt::last_error().c_str(); }
const char* xt_version(void) { return "xtt 0.1 (gfx950)"; }
int xt_synchronize(void) { return guard([] { xtt::sync(); }); }
int xt_release_cache(void) { return guard([] { xtt::sync(); xtt::release_cache(); }); }
int xt_cache_bytes(size_t* in_use, size_t* cached) {
	return guard([&] { if (in_use) *in_use = xtt::bytes_in_use(); if (cached) *cached = xtt::bytes_cached(); });
}
int xt_graph_replays(size_t* als, size_t* adf) {
	return guard([&] {
		size_t r[2];
		xtt::graph_replays(r);
		if (als) *als = r[0];
		if (adf) *adf = r[1];
	});
}
int xt_profile_enable(int on) { return guard([&] { xtt::prof_enable(on != 0); }); }
int xt_profile_read(size_t max_records, char* tags, size_t tag_len, size_t* launches, double* ms, double* flops, double* bytes,
                    size_t* count) {
	return guard([&] {
		NN(count);
		auto recs = xtt::prof_collect();
		*count = recs.size();
		for (size_t i = 0; i < recs.size() && i < max_records; ++i) {
			if (tags && tag_len) { std::strncpy(tags + i * tag_len, recs[i].tag.c_str(), tag_len - 1); tags[i * tag_len + tag_len - 1] = 0; }
			if (launches) launches[i] = recs[i].launches;
			if (ms) ms[i] = recs[i].ms;
			if (flops) flops[i] = recs[i].flops;
			if (bytes) bytes[i] = recs[i].bytes;
		}
	});
}
int xt_malloc(size_t bytes, void** ptr) { return guard([&] { NN(ptr); *ptr = nullptr; XTT_HIP(hipMalloc(ptr, bytes ? bytes : 1)); }); }
int xt_free(void* ptr) { return guard([&] { if (ptr) { XTT_HIP(hipDeviceSynchronize()); XTT_HIP(hipFree(ptr)); } }); }
int xt_memcpy_htod(void* dst, const void* src, size_t bytes) {
	return guard([&] { if (bytes) { NN(dst); NN(src); XTT_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, xtt::stream())); xtt::sync(); } });
}
int xt_memcpy_dtoh(void* dst, const void* src, size_t bytes) {
	return guard([&] { if (bytes) { NN(dst); NN(src); XTT_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, xtt::stream())); xtt::sync(); } });
}
int xt_seed(uint64_t s) { return guard([&] { xtt::rng_seed(s); }); }
int xt_normal_draws(size_t n, double* out) { return guard([&] { NN(out); for (size_t i = 0; i < n; ++i) out[i] = xtt::rng_normal(); }); }
int xt_uniform_draws(size_t n, size_t hi, size_t* out) {
	return guard([&] { NN(out); for (size_t i = 0; i < n; ++i) out[i] = xtt::rng_uniform(hi); });
}

// ---------------------------------------------------------------- boundary 1

int xt_gemm(int ta, int tb, size_t m, size_t n, size_t k, double alpha, const double* A, const double* B, double* C, void* stream) {
	return guard([&] {
		NN(C);
		XTT_REQUIRE(k == 0 || (A && B), "A and B must not be NULL");
		xtt::GemmDesc g = xtt::gemm_desc(ta, tb, m, n, k);
		g.alpha = alpha;
		if (k == 0) { XTT_HIP(hipMemsetAsync(C, 0, m * n * 8, xtt::pick(stream))); }
		else xtt::gemm_f64(g, A, B, C, xtt::pick(stream));
		sync_stream(stream);
	});
}
int xt_gemm_tune(int precision, size_t m, size_t n, size_t k, size_t batch, int ta, int tb, int iters, int* tile, int* slices, double* bestUs,
                 int* hTile, int* hSlices, double* hUs) {
	return guard([&] {
		const xtt::GemmTuneResult r = xtt::gemm_tune(precision == XT_F64 ? 0 : 1, tune_desc(precision, m, n, k, batch, ta, tb), iters, xtt::stream());
		if (tile) *tile = r.best.tile;
		if (slices) *slices = r.best.slices;
		if (bestUs) *bestUs = r.bestUs;
		if (hTile) *hTile = r.heuristic.tile;
		if (hSlices) *hSlices = r.heuristic.slices;
		if (hUs) *hUs = r.heuristicUs;
	});
}
int xt_gemm_tune_set(int precision, size_t m, size_t n, size_t k, size_t batch, int ta, int tb, int tile, int slices) {
	return guard([&] { xtt::gemm_tune_set(precision == XT_F64 ? 0 : 1, tune_desc(precision, m, n, k, batch, ta, tb), xtt::GemmChoice{tile, slices}); });
}
int xt_gemm_tune_clear(void) { return guard([&] { xtt::gemm_tune_clear(); }); }
int xt_gemm_tune_load(const char* path, size_t* installed) {
	return guard([&] {
		NN(path);
		std::ifstream in(path);
		XTT_REQUIRE(in.good(), std::string("GEMM tuning: cannot read ") + path);
		std::string line;
		size_t count = 0;
		while (std::getline(in, line)) {
			if (line.empty() || line[0] == '#' || line.compare(0, 9, "precision") == 0) continue;
			std::istringstream ls(line);
			long long v[14];
			double us = 0, hus = 0;
			// precision m n k batch ta tb launches tile slices us heuristic_tile heuristic_slices heuristic_us
			for (int i = 0; i < 10; ++i) ls >> v[i];
			ls >> us >> v[11] >> v[12] >> hus;
			XTT_REQUIRE(!ls.fail(), "GEMM tuning: malformed table line: " + line);
			if (v[8] == v[11] && v[9] == v[12]) continue;  // the heuristic's own choice
			xtt::gemm_tune_set(v[0] == XT_F64 ? 0 : 1, tune_desc((int)v[0], (size_t)v[1], (size_t)v[2], (size_t)v[3], (size_t)v[4], (int)v[5], (int)v[6]),
			                   xtt::GemmChoice{(int)v[8], (int)v[9]});
			++count;
		}
		if (installed) *installed = count;
	});
}
int xt_gemm_f32(int ta, int tb, size_t m, size_t n, size_t k, float alpha, const float* A, const float* B, float* C, void* stream) {
	return guard([&] {
		NN(A); NN(B); NN(C); XTT_REQUIRE(k > 0, "k must be positive");
		xtt::GemmDesc g = xtt::gemm_desc(ta, tb, m, n, k);
		g.alpha = alpha;
		xtt::gemm_f32(g, A, B, C, xtt::pick(stream));
		sync_stream(stream);
	});
}
int xt_gemm_bf16(int ta, int tb, size_t m, size_t n, size_t k, float alpha, const float* A, const float* B, float* C, void* stream) {
	return guard([&] {
		NN(A); NN(B); NN(C); XTT_REQUIRE(k > 0, "k must be positive");
		xtt::GemmDesc g = xtt::gemm_desc(ta, tb, m, n, k);
		g.alpha = alpha;
		xtt::gemm_bf16(g, A, B, C, xtt::pick(stream));
		sync_stream(stream);
	});
}

}  // extern "C"

namespace {
xtt::DTensor wrap_copy(const double* A, size_t m, size_t n) {
	xtt::DTensor t({m, n});
	XTT_HIP(hipMemcpyAsync(t.data(), A, m * n * 8, hipMemcpyDeviceToDevice, xtt::stream()));
	return t;
}
void put(double* dst, const xtt::DTensor& t) {
	if (t.size()) XTT_HIP(hipMemcpyAsync(dst, t.data(), t.size() * 8, hipMemcpyDeviceToDevice, xtt::stream()));
}
}  // namespace

extern "C" {

int xt_qr(const double* A, size_t m, size_t n, double* Q, double* R) {
	return guard([&] { NN(A); NN(Q); NN(R); xtt::DTensor q, r; xtt::calculate_qr(q, r, wrap_copy(A, m, n), 1); put(Q, q); put(R, r); xtt::sync(); });
}
int xt_rq(const double* A, size_t m, size_t n, double* R, double* Q) {
	return guard([&] { NN(A); NN(Q); NN(R); xtt::DTensor q, r; xtt::calculate_rq(r, q, wrap_copy(A, m, n), 1); put(Q, q); put(R, r); xtt::sync(); });
}
int xt_qc(const double* A, size_t m, size_t n, double* Q, double* C, size_t* rank) {
	return guard([&] {
		NN(A); NN(Q); NN(C); NN(rank);
		xtt::DTensor q, c; xtt::calculate_qc(q, c, wrap_copy(A, m, n), 1); put(Q, q); put(C, c); *rank = c.dims[0]; xtt::sync();
	});
}
int xt_cq(const double* A, size_t m, size_t n, double* C, double* Q, size_t* rank) {
	return guard([&] {
		NN(A); NN(Q); NN(C); NN(rank);
		xtt::DTensor q, c; xtt::calculate_cq(c, q, wrap_copy(A, m, n), 1); put(Q, q); put(C, c); *rank = c.dims[1]; xtt::sync();
	});
}
int xt_solve(const double* A, size_t m, size_t n, const double* B, size_t nrhs, double* X, int* method) {
	return guard([&] {
		NN(A); NN(B); NN(X);
		const int mth = xtt::dense_solve(A, m, n, B, nrhs, X, xtt::stream());
		if (method) *method = mth;
	});
}
int xt_svd(const double* A, size_t m, size_t n, double* U, double* S, double* Vt) {
	return guard([&] {
		NN(A); NN(U); NN(S); NN(Vt);
		xtt::DTensor u, vt;
		std::vector<double> s;
		xtt::calculate_svd(u, s, vt, wrap_copy(A, m, n), 1, 0, 0.0);
		// keep all min(m,n) values: a zero singular value is dropped by the eps=0 rule only when exactly zero
		const size_t k = std::min(m, n);
		XTT_HIP(hipMemsetAsync(U, 0, m * k * 8, xtt::stream()));
		XTT_HIP(hipMemsetAsync(S, 0, k * 8, xtt::stream()));
		XTT_HIP(hipMemsetAsync(Vt, 0, k * n * 8, xtt::stream()));
		const size_t r = s.size();
		XTT_HIP(hipMemcpy2DAsync(U, k * 8, u.data(), r * 8, r * 8, m, hipMemcpyDeviceToDevice, xtt::stream()));
		XTT_HIP(hipMemcpyAsync(S, s.data(), r * 8, hipMemcpyHostToDevice, xtt::stream()));
		put(Vt, vt);
		xtt::sync();
	});
}
int xt_reshuffle(size_t degree, const size_t* dims, const size_t* shuffle, const double* in, double* out) {
	return guard([&] {
		NN(dims); NN(shuffle); NN(in); NN(out);
		std::vector<bool> seen(degree, false);
		for (size_t i = 0; i < degree; ++i) { XTT_REQUIRE(shuffle[i] < degree && !seen[shuffle[i]], "reshuffle: not a permutation"); seen[shuffle[i]] = true; }
		xtt::permute_f64(degree, dims, shuffle, in, out, xtt::stream());
		xtt::sync();
	});
}
int xt_dot(const double* x, const double* y, size_t n, double* result) {
	return guard([&] {
		NN(x); NN(y); NN(result);
		xtt::DTensor a({n}), b({n});
		XTT_HIP(hipMemcpyAsync(a.data(), x, n * 8, hipMemcpyDeviceToDevice, xtt::stream()));
		XTT_HIP(hipMemcpyAsync(b.data(), y, n * 8, hipMemcpyDeviceToDevice, xtt::stream()));
		*result = xtt::dt_dot(a, b);
	});
}

int xt_one_norm(const double* x, size_t n, double* result) {
	return guard([&] {
		NN(result);
		XTT_REQUIRE(n == 0 || x, "x must not be NULL");
		if (n == 0) { *result = 0; return; }
		xtt::DBuf<double> out(3), partials(2 * 1024);
		xtt::DBuf<unsigned> counter(1);
		XTT_HIP(hipMemsetAsync(counter.get(), 0, sizeof(unsigned), xtt::stream()));
		xtt::norms_f64(x, n, out.get(), partials.get(), counter.get(), xtt::stream());
		xtt::read_back(result, out.get(), 8);
	});
}
int xt_two_norm(const double* x, size_t n, double* result) {
	return guard([&] {
		NN(result);
		XTT_REQUIRE(n == 0 || x, "x must not be NULL");
		if (n == 0) { *result = 0; return; }
		xtt::DBuf<double> out(3), partials(2 * 1024);
		xtt::DBuf<unsigned> counter(1);
		XTT_HIP(hipMemsetAsync(counter.get(), 0, sizeof(unsigned), xtt::stream()));
		xtt::norms_f64(x, n, out.get(), partials.get(), counter.get(), xtt::stream());
		xtt::read_back(result, out.get() + 2, 8);
	});
}
int xt_matrix_vector_product(double* x, size_t m, double alpha, const double* A, size_t n, int transposed, const double* y, void* stream) {
	return guard([&] {
		NN(x); NN(A); NN(y);
		XTT_REQUIRE(m > 0 && n > 0, "Dimension must be larger than zero");
		// x (m × 1) = alpha·op(A)·y: A m×n row-major, or n×m read transposed
		xtt::GemmDesc g = xtt::gemm_desc(transposed ? 1 : 0, 0, m, 1, n);
		g.alpha = alpha;
		xtt::gemm_f64(g, A, y, x, xtt::pick(stream));
		sync_stream(stream);
	});
}
int xt_dyadic_vector_product(double* A, size_t m, size_t n, double alpha, const double* x, const double* y, void* stream) {
	return guard([&] {
		NN(A); NN(x); NN(y);
		XTT_REQUIRE(m > 0 && n > 0, "Dimension must be larger than zero");
		xtt::GemmDesc g = xtt::gemm_desc(0, 0, m, n, 1);  // (m × 1)·(1 × n)
		g.alpha = alpha;
		xtt::gemm_f64(g, x, y, A, xtt::pick(stream));
		sync_stream(stream);
	});
}
int xt_solve_least_squares(double* X, const double* A, size_t m, size_t n, const double* B, size_t p) {
	return guard([&] { NN(X); NN(A); NN(B); xtt::least_squares_solve(A, m, n, B, p, X, xtt::stream()); xtt::sync(); });
}

// ---------------------------------------------------------------- boundary 2

int xt_tt_new(int op, size_t nd, const size_t* dims, xt_tt* out) {
	return guard([&] { NN(out); XTT_REQUIRE(nd == 0 || dims, "dims must not be NULL"); auto* h = new xt_tt_s{xtt::tt_zero(mk(dims, nd), op != 0)}; *out = h; });
}
int xt_tt_random(int op, size_t nd, const size_t* dims, const size_t* ranks, xt_tt* out) {
	return guard([&] {
		NN(out); NN(dims);
		const size_t d = nd / (op ? 2 : 1);
		XTT_REQUIRE(d > 0, "random TT needs at least one component");
		XTT_REQUIRE(d == 1 || ranks, "ranks must not be NULL");
		*out = new xt_tt_s{xtt::tt_random(mk(dims, nd), d > 1 ? mk(ranks, d - 1) : xtt::dims_t{}, op != 0)};
	});
}
int xt_tt_from_tensor(int op, size_t nd, const size_t* dims, const double* host, double eps, const size_t* maxRanks, xt_tt* out) {
	return guard([&] {
		NN(out); NN(dims); NN(host);
		const size_t d = nd / (op ? 2 : 1);
		*out = new xt_tt_s{xtt::tt_from_tensor(mk(dims, nd), host, op != 0, eps, maxRanks && d > 1 ? mk(maxRanks, d - 1) : xtt::dims_t{})};
	});
}
int xt_tt_ones(int op, size_t nd, const size_t* dims, xt_tt* out) {
	return guard([&] { NN(out); NN(dims); *out = new xt_tt_s{xtt::tt_ones(mk(dims, nd), op != 0)}; });
}
int xt_tt_identity(size_t nd, const size_t* dims, xt_tt* out) {
	return guard([&] { NN(out); NN(dims); *out = new xt_tt_s{xtt::tt_identity(mk(dims, nd))}; });
}
int xt_tt_kronecker(int op, size_t nd, const size_t* dims, xt_tt* out) {
	return guard([&] { NN(out); NN(dims); *out = new xt_tt_s{xtt::tt_kronecker(mk(dims, nd), op != 0)}; });
}
int xt_tt_dirac(int op, size_t nd, const size_t* dims, const size_t* position, xt_tt* out) {
	return guard([&] { NN(out); NN(dims); NN(position); *out = new xt_tt_s{xtt::tt_dirac(mk(dims, nd), mk(position, nd), op != 0)}; });
}
int xt_tt_transpose(xt_tt t) { return guard([&] { FMT(t); xtt::tt_transpose(t->t); xtt::sync(); }); }
int xt_tt_copy(xt_tt src, xt_tt* out) { return guard([&] { NN(src); NN(out); *out = new xt_tt_s{xtt::tt_copy(src->t)}; }); }
int xt_tt_free(xt_tt t) { return guard([&] { delete t; }); }
int xt_tt_order(xt_tt t, size_t* o) { return guard([&] { NN(t); NN(o); *o = t->t.order(); }); }
int xt_tt_is_operator(xt_tt t, int* o) { return guard([&] { NN(t); NN(o); *o = t->t.op; }); }
int xt_tt_dims(xt_tt t, size_t* d) { return guard([&] { NN(t); NN(d); std::memcpy(d, t->t.dims.data(), t->t.dims.size() * sizeof(size_t)); }); }
int xt_tt_ranks(xt_tt t, size_t* r) {
	return guard([&] { NN(t); auto v = t->t.ranks(); if (!v.empty()) { NN(r); std::memcpy(r, v.data(), v.size() * sizeof(size_t)); } });
}
int xt_tt_component_dims(xt_tt t, size_t k, size_t* d) {
	return guard([&] { NN(t); NN(d); XTT_REQUIRE(k < t->t.order(), "Illegal component index"); const auto& c = t->t.cores[k]; std::memcpy(d, c.dims.data(), c.dims.size() * sizeof(size_t)); });
}
int xt_tt_get_component(xt_tt t, size_t k, double* host) {
	return guard([&] { NN(t); NN(host); XTT_REQUIRE(k < t->t.order(), "Illegal component index"); xtt::dt_download(t->t.cores[k], host); });
}
int xt_tt_set_component(xt_tt t, size_t k, size_t nd, const size_t* dims, const double* host) {
	return guard([&] {
		NN(t); NN(dims); NN(host);
		XTT_REQUIRE(k < t->t.order(), "Illegal index in TTNetwork::set_component");
		XTT_REQUIRE(nd == (t->t.op ? 4u : 3u), "Component must have degree N+2");
		xtt::DTensor c = xtt::dt_upload(mk(dims, nd), host);
		xtt::sync();
		t->t.set_component(k, std::move(c));
	});
}
int xt_tt_component_ptr(xt_tt t, size_t k, double** p) {
	return guard([&] { NN(t); NN(p); XTT_REQUIRE(k < t->t.order(), "Illegal component index"); *p = t->t.cores[k].data(); });
}
int xt_tt_canonicalization(xt_tt t, int* c, size_t* pos) {
	return guard([&] { NN(t); NN(c); NN(pos); *c = t->t.canonicalized; *pos = t->t.corePosition; });
}
int xt_tt_assume_core_position(xt_tt t, size_t pos) {
	return guard([&] { NN(t); XTT_REQUIRE(pos < t->t.order(), "Invalid core position."); t->t.corePosition = pos; t->t.canonicalized = true; });
}
int xt_tt_move_core(xt_tt t, size_t pos, int keep) { return guard([&] { FMT(t); xtt::move_core(t->t, pos, keep != 0); xtt::sync(); }); }
int xt_tt_round(xt_tt t, const size_t* maxRanks, double eps) {
	return guard([&] {
		FMT(t);
		const size_t d = t->t.order();
		xtt::dims_t mr = maxRanks ? mk(maxRanks, d - 1) : xtt::dims_t(d - 1, SIZE_MAX);
		xtt::round(t->t, mr, eps);
		xtt::sync();
	});
}
int xt_tt_frob_norm(xt_tt t, double* o) { return guard([&] { FMT(t); NN(o); *o = xtt::tt_frob_norm(t->t); }); }
int xt_tt_dot(xt_tt a, xt_tt b, double* o) { return guard([&] { FMT(a); FMT(b); NN(o); *o = xtt::tt_dot(a->t, b->t); }); }
int xt_tt_add(xt_tt a, xt_tt b, xt_tt* out) { return guard([&] { FMT(a); FMT(b); NN(out); *out = new xt_tt_s{xtt::tt_add(a->t, b->t)}; xtt::sync(); }); }
int xt_tt_scale(xt_tt t, double s) { return guard([&] { NN(t); xtt::tt_scale(t->t, s); xtt::sync(); }); }
int xt_tt_apply(xt_tt A, xt_tt x, xt_tt* out) { return guard([&] { FMT(A); FMT(x); NN(out); *out = new xt_tt_s{xtt::tt_apply(A->t, x->t)}; xtt::sync(); }); }
int xt_tt_full(xt_tt t, double* host) {
	return guard([&] {
		FMT(t); NN(host);
		const xtt::TT& x = t->t;
		xtt::DTensor acc = xtt::dt_copy(x.cores[0]);
		for (size_t k = 1; k < x.order(); ++k) acc = xtt::contract(acc, false, x.cores[k], false, 1);
		if (x.op) {  // (n0,m0,n1,m1,…) → (n…, m…)
			const size_t d = x.order();
			xtt::dims_t inter;
			for (size_t k = 0; k < d; ++k) { inter.push_back(x.dims[k]); inter.push_back(x.dims[d + k]); }
			acc.reshape(inter);
			std::vector<size_t> sh(2 * d);
			for (size_t k = 0; k < d; ++k) { sh[2 * k] = k; sh[2 * k + 1] = d + k; }
			acc = xtt::reshuffle(acc, sh);
		}
		xtt::dt_download(acc, host);
	});
}
int xt_tt_save(xt_tt t, const char* filename, int format) {
	return guard([&] { FMT(t); NN(filename); xtt::tt_save(t->t, filename, format); });
}
int xt_tt_load(const char* filename, xt_tt* out) {
	return guard([&] { NN(filename); NN(out); *out = nullptr; *out = new xt_tt_s{xtt::tt_load(filename)}; });
}
int xt_tt_residual(xt_tt A, xt_tt x, xt_tt b, double* o) {
	return guard([&] { FMT(A); FMT(x); FMT(b); NN(o); *o = xtt::tt_residual(A->t, x->t, b->t); });
}

int xt_als_default_options(xt_als_options* o, unsigned sites, int spd) {
	return guard([&] {
		NN(o);
		XTT_REQUIRE(sites > 0, "sites must be > 0");
		xtt::ALSOptions d;
		o->sites = sites;
		o->numHalfSweeps = 0;
		o->convergenceEpsilon = d.convergenceEpsilon;
		o->useResidualForEndCriterion = 0;
		o->preserveCorePosition = 1;
		o->assumeSPD = spd;
		o->precision = XT_F64;
		o->localTolerance = d.localTol;
		o->maxLocalIterations = d.maxLocalIter;
		o->innerTolerance = 0;
		o->maxRefinements = d.maxRefine;
	});
}
extern "C++" {
namespace {
xtt::ALSOptions als_options(const xt_als_options* o) {
	XTT_REQUIRE(o->precision >= XT_F64 && o->precision <= XT_BF16, "unknown precision");
	xtt::ALSOptions p;
	p.sites = o->sites;
	p.numHalfSweeps = o->numHalfSweeps;
	p.convergenceEpsilon = o->convergenceEpsilon;
	p.useResidualForEndCriterion = o->useResidualForEndCriterion;
	p.preserveCorePosition = o->preserveCorePosition;
	p.assumeSPD = o->assumeSPD;
	p.precision = o->precision;
	p.localTol = o->localTolerance;
	p.maxLocalIter = o->maxLocalIterations;
	p.innerTol = o->innerTolerance;
	p.maxRefine = o->maxRefinements;
	return p;
}
void put_stats(const xtt::ALSStats& s, xt_als_stats* st) {
	if (!st) return;
	st->halfSweeps = s.halfSweeps; st->localSolves = s.localSolves; st->cgIterations = s.cgIterations;
	st->refinements = s.refinements; st->matvecs64 = s.matvecs64; st->matvecsLow = s.matvecsLow;
	st->flopsMatvec = s.flopsMatvec; st->flopsEnv = s.flopsEnv;
}
}  // namespace
}  // extern "C++"

int xt_als_round_batch(const xt_als_options* o, size_t count, const xt_tt* A, const xt_tt* x, const xt_tt* b, size_t sweeps,
                       double eps, int threads, double* residuals, xt_als_stats* st) {
	return guard([&] {
		NN(o); NN(x); NN(b);
		std::vector<const xtt::TT*> av(count), bv(count);
		std::vector<xtt::TT*> xv(count);
		for (size_t p = 0; p < count; ++p) {
			FMT(x[p]); FMT(b[p]);
			if (A && A[p]) FMT(A[p]);
			av[p] = A && A[p] ? &A[p]->t : nullptr;
			xv[p] = &x[p]->t;
			bv[p] = &b[p]->t;
		}
		xtt::ALSStats s;
		xtt::als_round_batch(als_options(o), av, xv, bv, sweeps, eps, threads, residuals, &s);
		put_stats(s, st);
	});
}

int xt_als_env_left(const double* L, const double* X, const double* A, size_t a, size_t n, size_t c, size_t p, size_t q, double* out) {
	return guard([&] { NN(L); NN(X); NN(A); NN(out); xtt::als_env_left(L, X, A, a, n, c, p, q, out); });
}
int xt_als_env_right(const double* R, const double* X, const double* A, size_t a, size_t n, size_t c, size_t p, size_t q, double* out) {
	return guard([&] { NN(R); NN(X); NN(A); NN(out); xtt::als_env_right(R, X, A, a, n, c, p, q, out); });
}
int xt_als_local_apply(const double* L, const double* A, const double* R, size_t a, size_t n, size_t c, size_t p, size_t q,
                       const double* x, double* y) {
	return guard([&] { NN(L); NN(A); NN(R); NN(x); NN(y); xtt::als_local_apply(L, A, R, a, n, c, p, q, x, y); });
}
int xt_als_local_cg(const double* L, const double* A, const double* R, size_t a, size_t n, size_t c, size_t p, size_t q,
                    const double* b, double* x, double tol, int max_iter, size_t* iterations) {
	return guard([&] {
		NN(L); NN(A); NN(R); NN(b); NN(x);
		const size_t it = xtt::als_local_cg(L, A, R, a, n, c, p, q, b, x, tol, max_iter);
		if (iterations) *iterations = it;
	});
}
int xt_als_solve(const xt_als_options* o, xt_tt A, xt_tt x, xt_tt b, double* energy, xt_als_stats* st) {
	return guard([&] {
		NN(o); FMT(x); FMT(b); NN(energy);
		if (A) FMT(A);
		XTT_REQUIRE(o->precision >= XT_F64 && o->precision <= XT_BF16, "unknown precision");
		xtt::ALSOptions p;
		p.sites = o->sites;
		p.numHalfSweeps = o->numHalfSweeps;
		p.convergenceEpsilon = o->convergenceEpsilon;
		p.useResidualForEndCriterion = o->useResidualForEndCriterion;
		p.preserveCorePosition = o->preserveCorePosition;
		p.assumeSPD = o->assumeSPD;
		p.precision = o->precision;
		p.localTol = o->localTolerance;
		p.maxLocalIter = o->maxLocalIterations;
		p.innerTol = o->innerTolerance;
		p.maxRefine = o->maxRefinements;
		xtt::ALSStats s;
		*energy = xtt::als_solve(p, A ? &A->t : nullptr, x->t, b->t, &s);
		xtt::sync();
		if (st) {
			st->halfSweeps = s.halfSweeps; st->localSolves = s.localSolves; st->cgIterations = s.cgIterations;
			st->refinements = s.refinements; st->matvecs64 = s.matvecs64; st->matvecsLow = s.matvecsLow;
			st->flopsMatvec = s.flopsMatvec; st->flopsEnv = s.flopsEnv;
		}
	});
}

// ---------------------------------------------------------------- measurements / ADF

int xt_meas_random(size_t P, size_t d, const size_t* dims, xt_meas* out) {
	return guard([&] { NN(dims); NN(out); *out = new xt_meas_s{xtt::meas_random(P, mk(dims, d))}; });
}
int xt_meas_new(size_t P, size_t d, const uint32_t* pos, const double* values, xt_meas* out) {
	return guard([&] {
		NN(out); XTT_REQUIRE(P == 0 || (pos && values), "positions/values must not be NULL");
		auto* h = new xt_meas_s;
		h->m.d = d;
		h->m.pos.assign(pos, pos + P * d);
		h->m.values.assign(values, values + P);
		h->m.upload();
		*out = h;
	});
}
int xt_meas_free(xt_meas m) { return guard([&] { delete m; }); }
int xt_meas_size(xt_meas m, size_t* P) { return guard([&] { NN(m); NN(P); *P = m->m.size(); }); }
int xt_meas_positions(xt_meas m, uint32_t* o) { return guard([&] { NN(m); NN(o); std::memcpy(o, m->m.pos.data(), m->m.pos.size() * 4); }); }
int xt_meas_positions_ptr(xt_meas m, const uint32_t** p) { return guard([&] { NN(m); NN(p); *p = m->m.dpos.get(); }); }
int xt_meas_values(xt_meas m, double* o) { return guard([&] { NN(m); NN(o); std::memcpy(o, m->m.values.data(), m->m.values.size() * 8); }); }
int xt_meas_set_values(xt_meas m, const double* v) { return guard([&] { NN(m); NN(v); std::memcpy(m->m.values.data(), v, m->m.values.size() * 8); }); }
int xt_meas_measure(xt_meas m, xt_tt x) { return guard([&] { NN(m); FMT(x); xtt::meas_measure(m->m, x->t); }); }
int xt_meas_stack_plan(xt_meas m, size_t* fc, size_t* bc, uint32_t* fi, uint32_t* bi, uint32_t* fs, uint32_t* bs, size_t* nu) {
	return guard([&] {
		NN(m); NN(fc); NN(bc);
		const xtt::AdfPlan& p = m->m.stack_plan();
		const size_t d = (size_t)p.d, P = (size_t)p.P;
		size_t all = 0;
		for (size_t k = 0; k < d; ++k) { fc[k] = (size_t)p.side[0].count[k]; bc[k] = (size_t)p.side[1].count[k]; all += fc[k]; }
		if (nu) *nu = all + 1;  // the slot shared by positions -1 and d (adf.cpp:162-166)
		const size_t nf = (size_t)(p.side[0].first[d - 1] + p.side[0].count[d - 1]);
		const size_t nb = (size_t)(p.side[1].first[d - 1] + p.side[1].count[d - 1]);
		if (fi) xtt::read_back(fi, p.side[0].rep.get(), nf * 4);
		if (bi) xtt::read_back(bi, p.side[1].rep.get(), nb * 4);
		if (fs) xtt::read_back(fs, p.side[0].slot.get(), d * P * 4);
		if (bs) xtt::read_back(bs, p.side[1].slot.get(), d * P * 4);
	});
}

int xt_adf_gradient(xt_tt x, xt_meas m, int precision, double* host, double* rn) {
	return guard([&] {
		FMT(x); NN(m); NN(host);
		XTT_REQUIRE(precision == XT_F64 || precision == XT_F32, "ADF gradient precision must be XT_F64 or XT_F32");
		auto g = precision == XT_F64 ? xtt::adf_full_gradient<double>(x->t, m->m, rn) : xtt::adf_full_gradient<float>(x->t, m->m, rn);
		size_t o = 0;
		for (auto& c : g) { std::memcpy(host + o, c.data(), c.size() * 8); o += c.size(); }
	});
}
int xt_r1meas_random(size_t P, size_t d, const size_t* dims, xt_r1meas* out) {
	return guard([&] { NN(dims); NN(out); *out = new xt_r1meas_s{xtt::r1_random(P, mk(dims, d))}; });
}
int xt_r1meas_from_single_point(xt_meas m, size_t d, const size_t* dims, xt_r1meas* out) {
	return guard([&] { NN(m); NN(dims); NN(out); *out = new xt_r1meas_s{xtt::r1_from_single_point(m->m, mk(dims, d))}; });
}
int xt_r1meas_new(size_t P, size_t d, const size_t* dims, const double* vectors, const double* values, xt_r1meas* out) {
	return guard([&] {
		NN(dims); NN(out);
		auto* h = new xt_r1meas_s;
		h->m.dims = mk(dims, d);
		const size_t S = h->m.sumN();
		XTT_REQUIRE(P == 0 || (vectors && values), "vectors/values must not be NULL");
		h->m.u.assign(vectors, vectors + P * S);
		h->m.values.assign(values, values + P);
		*out = h;
	});
}
int xt_r1meas_free(xt_r1meas m) { return guard([&] { delete m; }); }
int xt_r1meas_size(xt_r1meas m, size_t* P) { return guard([&] { NN(m); NN(P); *P = m->m.size(); }); }
int xt_r1meas_vectors(xt_r1meas m, double* o) { return guard([&] { NN(m); NN(o); std::memcpy(o, m->m.u.data(), m->m.u.size() * 8); }); }
int xt_r1meas_values(xt_r1meas m, double* o) { return guard([&] { NN(m); NN(o); std::memcpy(o, m->m.values.data(), m->m.values.size() * 8); }); }
int xt_r1meas_set_values(xt_r1meas m, const double* v) {
	return guard([&] { NN(m); NN(v); std::memcpy(m->m.values.data(), v, m->m.values.size() * 8); });
}
int xt_r1meas_sort(xt_r1meas m, int positions_only) { return guard([&] { NN(m); xtt::r1_sort(m->m, positions_only != 0); }); }
int xt_r1meas_normalize(xt_r1meas m) { return guard([&] { NN(m); xtt::r1_normalize(m->m); }); }
int xt_r1meas_measure(xt_r1meas m, xt_tt x) { return guard([&] { NN(m); FMT(x); xtt::r1_measure(m->m, x->t); }); }
int xt_r1meas_test(xt_r1meas m, xt_tt x, double* out) { return guard([&] { NN(m); FMT(x); NN(out); *out = xtt::r1_test(m->m, x->t); }); }
int xt_adf_gradient_r1(xt_tt x, xt_r1meas m, int precision, double* host, double* rn) {
	return guard([&] {
		FMT(x); NN(m); NN(host);
		XTT_REQUIRE(precision == XT_F64 || precision == XT_F32, "ADF gradient precision must be XT_F64 or XT_F32");
		auto g = precision == XT_F64 ? xtt::adf_full_gradient<double>(x->t, m->m, rn) : xtt::adf_full_gradient<float>(x->t, m->m, rn);
		size_t o = 0;
		for (auto& c : g) { std::memcpy(host + o, c.data(), c.size() * 8); o += c.size(); }
	});
}
int xt_adf_solve_r1(xt_tt x, xt_r1meas m, const size_t* maxRanks, size_t maxIt, double target, double minDec, int precision,
                    double* res, size_t* its) {
	return guard([&] {
		FMT(x); NN(m); NN(res);
		XTT_REQUIRE(precision == XT_F64 || precision == XT_F32, "ADF precision must be XT_F64 or XT_F32");
		xtt::ADFOptions o;
		o.maxIterations = maxIt; o.targetResidualNorm = target; o.minimalResidualNormDecrease = minDec;
		const size_t d = x->t.order();
		xtt::dims_t mr = maxRanks ? mk(maxRanks, d - 1) : x->t.ranks();
		size_t it = 0;
		*res = precision == XT_F64 ? xtt::adf_solve<double>(o, x->t, m->m, mr, &it) : xtt::adf_solve<float>(o, x->t, m->m, mr, &it);
		if (its) *its = it;
		xtt::sync();
	});
}
int xt_perf_sink(xt_perf_callback cb, void* user) {
	return guard([&] {
		xtt::PerfSink& ps = xtt::perf_sink();
		ps.cb = cb;
		ps.user = user;
	});
}

int xt_adf_solve(xt_tt x, xt_meas m, const size_t* maxRanks, size_t maxIt, double target, double minDec, int precision, double* res,
                 size_t* its) {
	return guard([&] {
		FMT(x); NN(m); NN(res);
		XTT_REQUIRE(precision == XT_F64 || precision == XT_F32, "ADF precision must be XT_F64 or XT_F32");
		xtt::ADFOptions o;
		o.maxIterations = maxIt; o.targetResidualNorm = target; o.minimalResidualNormDecrease = minDec;
		const size_t d = x->t.order();
		xtt::dims_t mr = maxRanks ? mk(maxRanks, d - 1) : x->t.ranks();
		size_t it = 0;
		*res = precision == XT_F64 ? xtt::adf_solve<double>(o, x->t, m->m, mr, &it) : xtt::adf_solve<float>(o, x->t, m->m, mr, &it);
		if (its) *its = it;
		xtt::sync();
	});
}

}  // extern "C"

namespace {
struct RawDesc {
	xtt::DBuf<int> n, r;
	xtt::DBuf<long long> off;
	xtt::AdfDesc desc{};
	RawDesc(int d, const int* hn, const int* hr, int P, hipStream_t s) {
		XTT_REQUIRE(d > 0 && P >= 0, "invalid order / count");
		std::vector<long long> o(d);
		long long acc = 0;
		int maxR = 1;
		XTT_REQUIRE(hr[0] == 1 && hr[d] == 1, "boundary ranks must be 1");
		for (int k = 0; k < d; ++k) {
			XTT_REQUIRE(hn[k] > 0 && hr[k] > 0 && hr[k + 1] > 0, "dims and ranks must be positive");
			o[k] = acc;
			acc += (long long)hr[k] * hn[k] * hr[k + 1];
			maxR = std::max(maxR, std::max(hr[k], hr[k + 1]));
		}
		XTT_REQUIRE(maxR <= 16, "ranks above 16 are not supported");
		n.resize(d); r.resize(d + 1); off.resize(d);
		XTT_HIP(hipMemcpyAsync(n.get(), hn, d * 4, hipMemcpyHostToDevice, s));
		XTT_HIP(hipMemcpyAsync(r.get(), hr, (d + 1) * 4, hipMemcpyHostToDevice, s));
		XTT_HIP(hipMemcpyAsync(off.get(), o.data(), d * 8, hipMemcpyHostToDevice, s));
		XTT_HIP(hipStreamSynchronize(s));
		desc.d = d; desc.P = P; desc.n = n.get(); desc.rank = r.get(); desc.coreOff = off.get(); desc.maxR = maxR; desc.gradSize = (int)acc;
		desc.sumN = 0;
		desc.maxN = 0;
		for (int k = 0; k < d; ++k) { desc.sumN += hn[k]; desc.maxN = std::max(desc.maxN, hn[k]); }
	}
};

// Uploaded descriptors by (d, n, r, P): a repeated call (the per-iteration gradients of a solver) reuses the
// device copy instead of uploading and synchronising again. Bounded; entries are immutable once built and held
// by each call that uses them, so clearing the cache never frees a descriptor in use.
std::shared_ptr<const RawDesc> raw_desc(int d, const int* hn, const int* hr, int P, hipStream_t s) {
	XTT_REQUIRE(d > 0 && P >= 0, "invalid order / count");
	std::vector<int> key;
	key.reserve(2 * d + 3);
	key.push_back(d);
	key.push_back(P);
	key.insert(key.end(), hn, hn + d);
	key.insert(key.end(), hr, hr + d + 1);
	static std::mutex mu;
	static std::map<std::vector<int>, std::shared_ptr<const RawDesc>> cache;
	std::lock_guard<std::mutex> lock(mu);
	auto it = cache.find(key);
	if (it == cache.end()) {
		auto rd = std::make_shared<const RawDesc>(d, hn, hr, P, s);
		if (cache.size() >= 64) cache.clear();
		it = cache.emplace(std::move(key), std::move(rd)).first;
	}
	return it->second;
}
}  // namespace

extern "C" {

int xt_tt_sample_f32(int d, const int* n, const int* r, const float* cores, const uint32_t* pos, int P, float* values, void* stream) {
	return guard([&] {
		NN(n); NN(r); NN(cores); NN(values); XTT_REQUIRE(P == 0 || pos, "pos must not be NULL");
		hipStream_t s = xtt::pick(stream);
		const auto held = raw_desc(d, n, r, P, s);
		const xtt::AdfDesc& desc = held->desc;
		if (P == 0) return;
		xtt::DBuf<double> norms(2), partials(2 * 1024);
		xtt::DBuf<unsigned> counter(1);
		XTT_HIP(hipMemsetAsync(counter.get(), 0, 4, s));
		xtt::adf_residual<float>(desc, cores, pos, nullptr, values, norms.get(), partials.get(), counter.get(), s);
		XTT_HIP(hipStreamSynchronize(s));
	});
}
int xt_tt_sample_residual_f32(int d, const int* n, const int* r, const float* cores, const uint32_t* pos, int P, const float* b,
                               float* resid, double* sums, void* stream) {
	return guard([&] {
		NN(n); NN(r); NN(cores); NN(resid); XTT_REQUIRE(P == 0 || (pos && b), "pos/b must not be NULL");
		hipStream_t s = xtt::pick(stream);
		const auto held = raw_desc(d, n, r, P, s);
		const xtt::AdfDesc& desc = held->desc;
		if (P == 0) { if (sums) sums[0] = sums[1] = 0; return; }
		xtt::DBuf<double> norms(2), partials(2 * 1024);
		xtt::DBuf<unsigned> counter(1);
		XTT_HIP(hipMemsetAsync(counter.get(), 0, 4, s));
		xtt::adf_residual<float>(desc, cores, pos, b, resid, norms.get(), partials.get(), counter.get(), s);
		if (sums) xtt::read_back(sums, norms.get(), 16, s);
		else XTT_HIP(hipStreamSynchronize(s));
	});
}
int xt_tt_sample_adjoint_f32(int d, const int* n, const int* r, const float* cores, const uint32_t* pos, int P, const float* seed,
                             float* grads, void* stream) {
	return guard([&] {
		NN(n); NN(r); NN(cores); NN(grads); XTT_REQUIRE(P == 0 || (pos && seed), "pos/seed must not be NULL");
		hipStream_t s = xtt::pick(stream);
		const auto held = raw_desc(d, n, r, P, s);
		const xtt::AdfDesc& desc = held->desc;
		if (P == 0) { XTT_HIP(hipMemsetAsync(grads, 0, (size_t)desc.gradSize * 4, s)); XTT_HIP(hipStreamSynchronize(s)); return; }
		xtt::DBuf<float> partials((size_t)xtt::adf_grad_blocks(P) * desc.gradSize);
		xtt::adf_gradient<float>(desc, cores, pos, seed, grads, partials.get(), -1, s);
		XTT_HIP(hipStreamSynchronize(s));
	});
}
int xt_tt_sample_residual_adjoint_f32(int d, const int* n, const int* r, const float* cores, const uint32_t* pos, int P,
                                      const float* b, float* resid, double* sums, float* grads, void* stream) {
	return guard([&] {
		NN(n); NN(r); NN(cores); NN(resid); NN(grads); XTT_REQUIRE(P == 0 || (pos && b), "pos/b must not be NULL");
		hipStream_t s = xtt::pick(stream);
		const auto held = raw_desc(d, n, r, P, s);
		const xtt::AdfDesc& desc = held->desc;
		if (P == 0) {
			if (sums) sums[0] = sums[1] = 0;
			XTT_HIP(hipMemsetAsync(grads, 0, (size_t)desc.gradSize * 4, s));
			XTT_HIP(hipStreamSynchronize(s));
			return;
		}
		xtt::DBuf<double> norms(2), normPart(2 * 1024);
		xtt::DBuf<unsigned> counter(1);
		xtt::DBuf<float> partials((size_t)xtt::adf_grad_blocks(P) * desc.gradSize);
		if (!xtt::adf_residual_gradient_fused_ok(desc)) XTT_HIP(hipMemsetAsync(counter.get(), 0, 4, s));  // two-kernel path only
		xtt::adf_residual_gradient_f32(desc, cores, pos, b, resid, norms.get(), grads, partials.get(), normPart.get(), counter.get(), s);
		if (sums) xtt::read_back(sums, norms.get(), 16, s);
		else XTT_HIP(hipStreamSynchronize(s));
	});
}

}  // extern "C"
