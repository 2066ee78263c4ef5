// Tensor-level operations and TTNetwork algorithms (move_core, round, sum, operator application, stable
// residual) on device-resident FP64 cores. Control flow follows xerus (file:line cited per function); the
// arithmetic runs in libxtt's gfx950 kernels.
#include "tt.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <map>
#include <random>

namespace xtt {

namespace {
size_t prodr(const dims_t& d, size_t b, size_t e) { size_t p = 1; for (size_t i = b; i < e; ++i) p *= d[i]; return p; }
dims_t cat(dims_t a, const dims_t& b) { a.insert(a.end(), b.begin(), b.end()); return a; }
dims_t sl(const dims_t& d, size_t b, size_t e) { return dims_t(d.begin() + b, d.begin() + e); }

// Reduction scratch shared by the host-side helpers (one stream, so reuse is ordered).
struct Scratch {
	DBuf<double> partials;
	DBuf<unsigned> counter;
	DBuf<double> scalar;
	Scratch() : Scratch(ArenaScope(nullptr)) {}
	explicit Scratch(const ArenaScope&) : partials(2 * 1024), counter(1), scalar(4) {  // persistent: never from a capture arena
		XTT_HIP(hipMemsetAsync(counter.get(), 0, sizeof(unsigned), stream()));
	}
};
Scratch& scratch() { static thread_local Scratch s; return s; }

std::mt19937_64& engine() { static std::mt19937_64 e(0xBAADF00Dull); return e; }
std::normal_distribution<double>& normal() { static std::normal_distribution<double> d; return d; }
}  // namespace

void rng_seed(uint64_t s) { engine().seed(s); normal().reset(); }
double rng_normal() { return normal()(engine()); }
size_t rng_uniform(size_t hi) { std::uniform_int_distribution<size_t> u(0, hi); return u(engine()); }

// ------------------------------------------------------------------------------- Tensor level

DTensor dt_upload(const dims_t& dims, const double* host) {
	DTensor t(dims);
	if (t.size()) XTT_HIP(hipMemcpyAsync(t.data(), host, t.size() * 8, hipMemcpyHostToDevice, stream()));
	return t;
}
void dt_download(const DTensor& t, double* host) {
	if (t.size()) XTT_HIP(hipMemcpyAsync(host, t.data(), t.size() * 8, hipMemcpyDeviceToHost, stream()));
	sync();
}
DTensor dt_copy(const DTensor& t) {
	DTensor o(t.dims);
	if (t.size()) XTT_HIP(hipMemcpyAsync(o.data(), t.data(), t.size() * 8, hipMemcpyDeviceToDevice, stream()));
	return o;
}
DTensor dt_zeros(const dims_t& dims) {
	DTensor t(dims);
	if (t.size()) XTT_HIP(hipMemsetAsync(t.data(), 0, t.size() * 8, stream()));
	return t;
}

DTensor contract(const DTensor& lhs, bool lt, const DTensor& rhs, bool rt, size_t modes) {
	const size_t ld = lhs.degree(), rd = rhs.degree();
	XTT_REQUIRE(modes <= ld && modes <= rd, "contract: too many modes");
	dims_t lrem = lt ? sl(lhs.dims, modes, ld) : sl(lhs.dims, 0, ld - modes);
	dims_t lcon = lt ? sl(lhs.dims, 0, modes) : sl(lhs.dims, ld - modes, ld);
	dims_t rrem = rt ? sl(rhs.dims, 0, rd - modes) : sl(rhs.dims, modes, rd);
	dims_t rcon = rt ? sl(rhs.dims, rd - modes, rd) : sl(rhs.dims, 0, modes);
	XTT_REQUIRE(lcon == rcon, "contract: dimension mismatch");
	DTensor out(cat(lrem, rrem));
	const size_t M = prodr(lrem, 0, lrem.size()), N = prodr(rrem, 0, rrem.size()), K = prodr(lcon, 0, lcon.size());
	if (M * N == 0) return out;
	if (K == 0) { XTT_HIP(hipMemsetAsync(out.data(), 0, out.size() * 8, stream())); return out; }
	gemm_f64(gemm_desc(lt, rt, M, N, K), lhs.data(), rhs.data(), out.data(), stream());
	return out;
}

DTensor reshuffle(const DTensor& t, const std::vector<size_t>& shuffle) {
	XTT_REQUIRE(shuffle.size() == t.degree(), "reshuffle: wrong permutation size");
	dims_t nd(t.degree());
	std::vector<bool> seen(t.degree(), false);
	for (size_t i = 0; i < t.degree(); ++i) {
		XTT_REQUIRE(shuffle[i] < t.degree() && !seen[shuffle[i]], "reshuffle: not a permutation");
		seen[shuffle[i]] = true;
		nd[shuffle[i]] = t.dims[i];
	}
	DTensor o(nd);
	permute_f64(t.degree(), t.dims.data(), shuffle.data(), t.data(), o.data(), stream());
	return o;
}

DTensor transpose2(const DTensor& t, size_t split) {
	const size_t m = prodr(t.dims, 0, split), n = prodr(t.dims, split, t.degree());
	DTensor o(cat(sl(t.dims, split, t.degree()), sl(t.dims, 0, split)));
	const size_t d2[2] = {m, n}, sh[2] = {1, 0};
	permute_f64(2, d2, sh, t.data(), o.data(), stream());
	return o;
}

namespace {
struct Factor {  // work buffers of one Householder factorisation
	DBuf<double> a, tau;
	DBuf<int> jpvt, rank;
};
// Factor the m×n matrix view of `in`; returns the rank decided on the device (synchronises).
int factor(Factor& f, const DTensor& in, size_t m, size_t n, bool pivot) {
	f.a.resize(m * n);
	f.tau.resize(std::min(m, n));
	f.jpvt.resize(n);
	f.rank.resize(1);
	XTT_HIP(hipMemcpyAsync(f.a.get(), in.data(), m * n * 8, hipMemcpyDeviceToDevice, stream()));
	geqp2(f.a.get(), m, n, f.tau.get(), f.jpvt.get(), pivot ? 1 : 0, nullptr, f.rank.get(), stream());
	if (!pivot) return (int)std::min(m, n);
	int r = 0;
	read_back(&r, f.rank.get(), sizeof(int));
	return r;
}
}  // namespace

// tensor.cpp:1492-1507 + blasLapackWrapper.cpp:374-431
void calculate_qr(DTensor& Q, DTensor& R, const DTensor& in, size_t split) {
	const size_t m = prodr(in.dims, 0, split), n = prodr(in.dims, split, in.degree()), k = std::min(m, n);
	XTT_REQUIRE(m > 0 && n > 0, "Dimension must be larger than zero");
	Q = DTensor(cat(sl(in.dims, 0, split), {k}));
	R = DTensor(cat({k}, sl(in.dims, split, in.degree())));
	if (hh_supported(m, n)) {
		hh_qr(in.data(), m, n, false, Q.data(), R.data(), stream());
		return;
	}
	Factor f;
	factor(f, in, m, n, false);
	org2r(f.a.get(), m, n, f.tau.get(), Q.data(), k, stream());
	extract_r(f.a.get(), m, n, R.data(), stream());
}

// tensor.cpp:1510-1525 + blasLapackWrapper.cpp:442-498, through QR(J·Aᵀ·J) = Q̃R̃ ⇒ A = (J·R̃ᵀ·J)(J·Q̃ᵀ·J).
void calculate_rq(DTensor& R, DTensor& Q, const DTensor& in, size_t split) {
	const size_t m = prodr(in.dims, 0, split), n = prodr(in.dims, split, in.degree()), k = std::min(m, n);
	XTT_REQUIRE(m > 0 && n > 0, "Dimension must be larger than zero");
	DTensor B({n, m});
	revtrans_f64(in.data(), m, n, B.data(), stream());
	DTensor Qt, Rt;
	calculate_qr(Qt, Rt, B, 1);  // Qt n×k, Rt k×m
	R = DTensor(cat(sl(in.dims, 0, split), {k}));
	Q = DTensor(cat({k}, sl(in.dims, split, in.degree())));
	revtrans_f64(Rt.data(), k, m, R.data(), stream());
	revtrans_f64(Qt.data(), n, k, Q.data(), stream());
}

// tensor.cpp:1528-1545 + blasLapackWrapper.cpp:235-305
void calculate_qc(DTensor& Q, DTensor& C, const DTensor& in, size_t split) {
	const size_t m = prodr(in.dims, 0, split), n = prodr(in.dims, split, in.degree());
	XTT_REQUIRE(m > 0 && n > 0, "Dimension must be larger than zero");
	if (hh_supported(m, n)) {
		const size_t k = std::min(m, n);
		DBuf<double> q(m * k), c(k * n);
		DBuf<int> jp(n), rk(1);
		hh_qc(in.data(), m, n, q.get(), c.get(), jp.get(), rk.get(), stream());
		int r = 0;
		read_back(&r, rk.get(), sizeof(int));
		Q = DTensor(cat(sl(in.dims, 0, split), {(size_t)r}), std::move(q));
		C = DTensor(cat({(size_t)r}, sl(in.dims, split, in.degree())), std::move(c));
		return;
	}
	Factor f;
	const size_t r = (size_t)factor(f, in, m, n, true);
	Q = DTensor(cat(sl(in.dims, 0, split), {r}));
	C = DTensor(cat({r}, sl(in.dims, split, in.degree())));
	org2r(f.a.get(), m, n, f.tau.get(), Q.data(), r, stream());
	extract_c(f.a.get(), m, n, f.jpvt.get(), f.rank.get(), C.data(), stream());
}

// tensor.cpp:1548-1565 + blasLapackWrapper.cpp:316-371: pivoted QR of Aᵀ, A = (R·Pᵀ)ᵀ·Qᵀ.
void calculate_cq(DTensor& C, DTensor& Q, const DTensor& in, size_t split) {
	const size_t m = prodr(in.dims, 0, split), n = prodr(in.dims, split, in.degree());
	XTT_REQUIRE(m > 0 && n > 0, "Dimension must be larger than zero");
	DTensor At({n, m});
	{ const size_t d2[2] = {m, n}, sh[2] = {1, 0}; permute_f64(2, d2, sh, in.data(), At.data(), stream()); }
	DTensor Qt, Ct;
	calculate_qc(Qt, Ct, At, 1);  // Qt n×r, Ct r×m
	const size_t r = Ct.dims[0];
	C = DTensor(cat(sl(in.dims, 0, split), {r}));
	Q = DTensor(cat({r}, sl(in.dims, split, in.degree())));
	{ const size_t d2[2] = {r, m}, sh[2] = {1, 0}; permute_f64(2, d2, sh, Ct.data(), C.data(), stream()); }
	{ const size_t d2[2] = {n, r}, sh[2] = {1, 0}; permute_f64(2, d2, sh, Qt.data(), Q.data(), stream()); }
}

namespace {
// U (m×rank) and Vt (rank×n) from the thin factors u (m×k) and vt (k×n): the buffers themselves at full rank,
// else copies of the leading columns / rows.
void take_svd_factors(DTensor& U, DTensor& Vt, DBuf<double>& u, DBuf<double>& vt, dims_t uDims, dims_t vDims, size_t m, size_t n,
                      size_t k, size_t rank) {
	if (rank == k) {
		U = DTensor(std::move(uDims), std::move(u));
		Vt = DTensor(std::move(vDims), std::move(vt));
		return;
	}
	U = DTensor(std::move(uDims));
	Vt = DTensor(std::move(vDims));
	XTT_HIP(hipMemcpy2DAsync(U.data(), rank * 8, u.get(), k * 8, rank * 8, m, hipMemcpyDeviceToDevice, stream()));
	XTT_HIP(hipMemcpyAsync(Vt.data(), vt.get(), rank * n * 8, hipMemcpyDeviceToDevice, stream()));
}
}  // namespace

// tensor.cpp:1424-1489 (dense branch); truncation rank = min(k, maxRank) then first j with s_j <= eps·s_0.
void calculate_svd(DTensor& U, std::vector<double>& S, DTensor& Vt, const DTensor& in, size_t split, size_t maxRank, double eps,
                   DBuf<double>* sDev, bool jacobiOnTranspose, const double* inTransposed) {
	XTT_REQUIRE(eps >= 0 && eps < 1, "Epsilon must be fullfill 0 <= _eps < 1.");
	const size_t m = prodr(in.dims, 0, split), n = prodr(in.dims, split, in.degree()), k = std::min(m, n);
	XTT_REQUIRE(m > 0 && n > 0, "Dimension must be larger than zero");
	DBuf<double> u(m * k), s(k), vt(k * n);
	const size_t M = std::max(m, n), N = k, np = N + (N & 1);
	DBuf<double> ws(M * np + np * np + 2 * np);
	XTT_REQUIRE(!inTransposed || !(m > n || (m == n && !jacobiOnTranspose)), "svd: a transposed input is only used for m <= n");
	if (m > n || (m == n && !jacobiOnTranspose)) {
		jacobi_svd(in.data(), m, n, u.get(), s.get(), vt.get(), ws.get(), stream());
	} else {  // Aᵀ = U'·S·V'ᵀ  ⇒  A = V'·S·U'ᵀ
		DBuf<double> at;
		if (!inTransposed) {
			at.resize(n * m);
			const size_t d2[2] = {m, n}, sh[2] = {1, 0};
			permute_f64(2, d2, sh, in.data(), at.get(), stream());
		}
		// written transposed by the kernel: U = V' (m×m) and Vt = U'ᵀ (m×n)
		jacobi_svd(inTransposed ? inTransposed : at.get(), n, m, vt.get(), s.get(), u.get(), ws.get(), stream(), true);
	}
	std::vector<double> sh(k);
	read_back(sh.data(), s.get(), k * 8);
	if (std::isnan(sh[0])) throw Error(ERR_NUMERIC, "SVD: the one-sided Jacobi iteration did not converge in 60 sweeps");
	size_t rank = k;
	if (maxRank != 0) rank = std::min(rank, maxRank);
	for (size_t j = 1; j < rank; ++j) if (sh[j] <= eps * sh[0]) { rank = j; break; }
	take_svd_factors(U, Vt, u, vt, cat(sl(in.dims, 0, split), {rank}), cat({rank}, sl(in.dims, split, in.degree())), m, n, k, rank);
	S.assign(sh.begin(), sh.begin() + rank);
	if (sDev) *sDev = std::move(s);
}

DTensor qr_r_only(const DTensor& in, size_t split) {
	const size_t m = prodr(in.dims, 0, split), n = prodr(in.dims, split, in.degree()), k = std::min(m, n);
	DTensor R({k, n});
	if (hh_supported(m, n)) hh_qr(in.data(), m, n, false, nullptr, R.data(), stream());
	else hh_r_blocked(in.data(), m, n, R.data(), stream());
	return R;
}

double dt_dot(const DTensor& a, const DTensor& b) {
	XTT_REQUIRE(a.size() == b.size(), "dot: size mismatch");
	Scratch& s = scratch();
	dot_f64(a.data(), b.data(), a.size(), s.scalar.get(), s.partials.get(), s.counter.get(), stream());
	double r = 0;
	read_back(&r, s.scalar.get(), 8);
	return r;
}
std::pair<double, double> dt_dot2(const DTensor& a, const DTensor& b, const DTensor& c, const DTensor& d) {
	XTT_REQUIRE(a.size() == b.size() && c.size() == d.size(), "dot: size mismatch");
	Scratch& s = scratch();
	dot_f64(a.data(), b.data(), a.size(), s.scalar.get(), s.partials.get(), s.counter.get(), stream());
	dot_f64(c.data(), d.data(), c.size(), s.scalar.get() + 1, s.partials.get(), s.counter.get(), stream());
	double r[2] = {0, 0};
	read_back(r, s.scalar.get(), 16);
	return {r[0], r[1]};
}
double dt_norm(const DTensor& t) { return std::sqrt(std::max(0.0, dt_dot(t, t))); }
void dt_scale(DTensor& t, double s) { scale_f64(t.data(), s, t.size(), stream()); }

// ------------------------------------------------------------------------------- TTNetwork

dims_t TT::ranks() const {
	dims_t r;
	for (size_t k = 0; k + 1 < cores.size(); ++k) r.push_back(cores[k].dims.back());
	return r;
}

void TT::set_component(size_t k, DTensor t) {
	XTT_REQUIRE(k < order(), "Illegal index in TTNetwork::set_component");
	XTT_REQUIRE(t.degree() == (op ? 4u : 3u), "Component must have degree N+2");
	dims[k] = t.dims[1];
	if (op) dims[order() + k] = t.dims[2];
	cores[k] = std::move(t);
	canonicalized = canonicalized && (corePosition == k);
}

void require_correct_format(const TT& x) {  // ttNetwork.cpp:290-340 (the checks that apply to a core list)
	using std::to_string;
	const size_t d = x.order(), N = x.op ? 2 : 1;
	XTT_REQUIRE(x.dims.size() == N * d, "TT format: " + to_string(x.dims.size()) + " dimensions for " + to_string(d) + " components");
	XTT_REQUIRE(!x.canonicalized || d == 0 || x.corePosition < d, "TT format: invalid core position " + to_string(x.corePosition));
	for (size_t k = 0; k < d; ++k) {
		const DTensor& c = x.cores[k];
		XTT_REQUIRE(c.degree() == N + 2, "TT format: component " + to_string(k) + " has degree " + to_string(c.degree()));
		XTT_REQUIRE(c.dims[1] == x.dims[k] && (!x.op || c.dims[2] == x.dims[d + k]),
		            "TT format: component " + to_string(k) + " does not match the external dimensions");
		const size_t left = k == 0 ? 1 : x.cores[k - 1].dims.back();
		XTT_REQUIRE(c.dims[0] == left, "TT format: component " + to_string(k) + " has left rank " + to_string(c.dims[0]) +
		                                   (k == 0 ? ", the first component's must be 1" : ", its left neighbour's right rank is " + to_string(left)));
	}
	XTT_REQUIRE(d == 0 || x.cores[d - 1].dims.back() == 1, "TT format: the last component's right rank must be 1");
}

TT tt_zero(const dims_t& dims, bool op) {  // ttNetwork.cpp:57-108
	TT x;
	x.op = op;
	x.dims = dims;
	const size_t d = dims.size() / (op ? 2 : 1);
	XTT_REQUIRE(dims.size() % (op ? 2 : 1) == 0, "Illegal degree for TTOperator.");
	for (size_t n : dims) XTT_REQUIRE(n > 0, "Zero is no valid dimension.");
	for (size_t k = 0; k < d; ++k) x.cores.push_back(dt_zeros(op ? dims_t{1, dims[k], dims[d + k], 1} : dims_t{1, dims[k], 1}));
	x.canonicalized = true;
	x.corePosition = 0;
	return x;
}

TT tt_copy(const TT& x) {
	TT y;
	y.op = x.op; y.dims = x.dims; y.canonicalized = x.canonicalized; y.corePosition = x.corePosition;
	// all cores in one launch per kMultiCopy cores (the per-core copies are host-bound at TT ranks <= 64)
	MultiCopy mc;
	for (const auto& c : x.cores) {
		y.cores.emplace_back(c.dims);
		if (mc.count == kMultiCopy) { multi_copy_f64(mc, stream()); mc.count = 0; }
		mc.src[mc.count] = c.data(); mc.dst[mc.count] = y.cores.back().data(); mc.n[mc.count] = c.count(); mc.count++;
	}
	multi_copy_f64(mc, stream());
	return y;
}

dims_t reduce_to_maximal_ranks(dims_t ranks, const dims_t& dims, bool op) {  // ttNetwork.cpp:369-402
	const size_t d = dims.size() / (op ? 2 : 1);
	XTT_REQUIRE(d == ranks.size() + 1, "Invalid number of ranks or dimensions given.");
	size_t cur = 1;
	for (size_t i = 0; i + 1 < d; ++i) {
		cur *= dims[i]; if (op) cur *= dims[d + i];
		if (cur < ranks[i]) ranks[i] = cur; else cur = ranks[i];
	}
	cur = 1;
	for (size_t i = 1; i < d; ++i) {
		cur *= dims[d - i]; if (op) cur *= dims[2 * d - i];
		if (cur < ranks[d - i - 1]) ranks[d - i - 1] = cur; else cur = ranks[d - i - 1];
	}
	return ranks;
}

TT tt_random(const dims_t& dims, const dims_t& ranks, bool op) {  // ttNetwork.h:128-155
	const size_t d = dims.size() / (op ? 2 : 1);
	XTT_REQUIRE(ranks.size() + 1 == d, "Non-matching amount of ranks given to TTNetwork::random.");
	for (size_t r : ranks) XTT_REQUIRE(r > 0, "Trying to construct random TTTensor with rank 0 is illegal.");
	TT x = tt_zero(dims_t(dims.size(), 1), op);
	const dims_t target = reduce_to_maximal_ranks(ranks, dims, op);
	std::vector<double> h;
	for (size_t i = 0; i < d; ++i) {
		const size_t l = i == 0 ? 1 : target[i - 1], r = i + 1 == d ? 1 : target[i];
		const dims_t cd = op ? dims_t{l, dims[i], dims[d + i], r} : dims_t{l, dims[i], r};
		h.resize(prodr(cd, 0, cd.size()));
		for (double& v : h) v = rng_normal();
		x.set_component(i, dt_upload(cd, h.data()));
		sync();  // h is reused
	}
	move_core(x, 0, false);
	return x;
}

// tensorNetwork.cpp:821-909 for adjacent components.
void transfer_core(TT& x, size_t from, size_t to, bool allowRankReduction) {
	DTensor& F = x.cores[from];
	DTensor& T = x.cores[to];
	DTensor Q, R;
	if (to == from + 1) {
		if (allowRankReduction) calculate_qc(Q, R, F, F.degree() - 1); else calculate_qr(Q, R, F, F.degree() - 1);
		F = std::move(Q);
		T = contract(R, false, T, false, 1);
	} else {
		XTT_REQUIRE(to + 1 == from, "transfer_core: components not adjacent");
		if (allowRankReduction) calculate_cq(R, Q, F, 1); else calculate_rq(R, Q, F, 1);
		F = std::move(Q);
		T = contract(T, false, R, false, 1);
	}
}

namespace {
bool exceeds_maximal_ranks(const TT& x) {  // ttNetwork.cpp:349-359
	for (size_t i = 0; i < x.order(); ++i) {
		const DTensor& c = x.cores[i];
		const size_t e = x.ext(i);
		if (c.dims.front() > e * c.dims.back() || c.dims.back() > e * c.dims.front()) return true;
	}
	return false;
}
}  // namespace

void move_core(TT& x, size_t pos, bool keepRank) {  // ttNetwork.cpp:581-628
	const size_t d = x.order();
	XTT_REQUIRE(pos < d, "Illegal core-position");
	const bool red = !keepRank;
	if (x.canonicalized) {
		for (size_t n = x.corePosition; n < pos; ++n) transfer_core(x, n, n + 1, red);
		for (size_t n = x.corePosition; n > pos; --n) transfer_core(x, n, n - 1, red);
	} else {
		for (size_t n = 0; n < pos; ++n) transfer_core(x, n, n + 1, red);
		for (size_t n = d; n > pos + 1; --n) transfer_core(x, n - 1, n - 2, red);
	}
	while (exceeds_maximal_ranks(x)) {
		for (size_t n = pos; n > 0; --n) transfer_core(x, n, n - 1, red);
		for (size_t n = 0; n + 1 < d; ++n) transfer_core(x, n, n + 1, red);
		for (size_t n = d; n > pos + 1; --n) transfer_core(x, n - 1, n - 2, red);
	}
	x.canonicalized = true;
	x.corePosition = pos;
}

// tensorNetwork.cpp:678-818 for (from, to) = (k, k-1): the shared edge is from's first and to's last mode.
void round_edge(TT& x, size_t from, size_t to, size_t maxRank, double eps, bool toIsLeftOrthogonal) {
	XTT_REQUIRE(to + 1 == from, "round_edge: expects (k, k-1)");
	DTensor& F = x.cores[from];
	DTensor& T = x.cores[to];
	const double r = (double)F.dims[0];
	std::vector<double> S;
	DBuf<double> sd;
	if (5.0 * (double)F.size() * (double)T.size() >= 6.0 * r * r * r * r) {
		DTensor coreA, Fq, U, Vt, X;
		calculate_cq(coreA, Fq, F, 1);
		if (toIsLeftOrthogonal) {
			// QC(T) = (T, I): X = coreAᵀ, and T·Vtᵀ replaces Tq·(coreB·…). X has no more rows than columns, so
			// the Jacobi SVD runs on its transpose, which is coreA itself (X's entries are never formed).
			X.dims = {coreA.dims[1], coreA.dims[0]};
			calculate_svd(U, S, Vt, X, 1, maxRank, eps, &sd, true, coreA.data());
			scale_by_f64(Vt.data(), sd.get(), S.size(), Vt.size() / std::max<size_t>(S.size(), 1), true, stream());
			F = contract(U, true, Fq, false, 1);
			T = contract(T, false, Vt, true, 1);
		} else {
			DTensor coreB, Tq;
			calculate_qc(Tq, coreB, T, T.degree() - 1);
			X = contract(coreA, true, coreB, true, 1);
			calculate_svd(U, S, Vt, X, 1, maxRank, eps, &sd, true);
			scale_by_f64(Vt.data(), sd.get(), S.size(), Vt.size() / std::max<size_t>(S.size(), 1), true, stream());
			F = contract(U, true, Fq, false, 1);
			T = contract(Tq, false, Vt, true, 1);
		}
	} else {
		DTensor X = contract(F, true, T, true, 1);
		DTensor U, Vt;
		calculate_svd(U, S, Vt, X, F.degree() - 1, maxRank, eps, &sd);
		scale_by_f64(Vt.data(), sd.get(), S.size(), Vt.size() / std::max<size_t>(S.size(), 1), true, stream());
		T = transpose2(Vt, 1);
		std::vector<size_t> back(U.degree());
		for (size_t i = 0; i + 1 < U.degree(); ++i) back[i] = i + 1;
		back[U.degree() - 1] = 0;
		F = reshuffle(U, back);
	}
}

// ---- speculative round ---------------------------------------------------------------------------------------
// Every rank decision of round (the QC sweep, each edge's CQ and SVD truncation) normally costs one read-back.
// When the last round of this thread on the same shapes and parameters kept every rank, round first runs with
// each decision assumed (QC and CQ keep full rank, the SVD keeps the current edge rank). Each kernel writes the
// rank it decides to a device slot, and one read-back at the end compares them with the assumptions. A mismatch
// restores x and runs the normal round, so the result is always the normal round's. (A wrong assumption leaves
// unwritten columns in the pass being discarded; every kernel loop is bounded, so that pass still ends.)
namespace {
struct NotSpeculable {};  // a shape outside the register-resident kernels, or a canonicalisation detour
struct RankSpec {
	DBuf<int> slots;
	std::vector<int> expect;
	explicit RankSpec(size_t cap) : slots(cap) {}
	int* slot(int expected) {
		if (expect.size() >= slots.size()) throw NotSpeculable{};
		expect.push_back(expected);
		return slots.get() + (expect.size() - 1);
	}
	bool holds() {
		std::vector<int> h(expect.size());
		if (!h.empty()) read_back(h.data(), slots.get(), h.size() * sizeof(int));
		return h == expect;
	}
};

// calculate_qc with the rank assumed full (min(m, n))
void qc_spec(DTensor& Q, DTensor& C, const DTensor& in, size_t split, RankSpec& rs) {
	const size_t m = prodr(in.dims, 0, split), n = prodr(in.dims, split, in.degree());
	if (!hh_supported(m, n)) throw NotSpeculable{};
	const size_t k = std::min(m, n);
	DBuf<double> q(m * k), c(k * n);
	DBuf<int> jp(n);
	hh_qc(in.data(), m, n, q.get(), c.get(), jp.get(), rs.slot((int)k), stream());
	Q = DTensor(cat(sl(in.dims, 0, split), {k}), std::move(q));
	C = DTensor(cat({k}, sl(in.dims, split, in.degree())), std::move(c));
}
// calculate_cq with the rank assumed full
void cq_spec(DTensor& C, DTensor& Q, const DTensor& in, size_t split, RankSpec& rs) {
	const size_t m = prodr(in.dims, 0, split), n = prodr(in.dims, split, in.degree());
	DTensor At({n, m});
	{ const size_t d2[2] = {m, n}, sh[2] = {1, 0}; permute_f64(2, d2, sh, in.data(), At.data(), stream()); }
	DTensor Qt, Ct;
	qc_spec(Qt, Ct, At, 1, rs);
	const size_t r = Ct.dims[0];
	C = DTensor(cat(sl(in.dims, 0, split), {r}));
	Q = DTensor(cat({r}, sl(in.dims, split, in.degree())));
	{ const size_t d2[2] = {r, m}, sh[2] = {1, 0}; permute_f64(2, d2, sh, Ct.data(), C.data(), stream()); }
	{ const size_t d2[2] = {n, r}, sh[2] = {1, 0}; permute_f64(2, d2, sh, Qt.data(), Q.data(), stream()); }
}
// calculate_svd with the truncation rank assumed (rank), the decided one written to a slot
void svd_spec(DTensor& U, DTensor& Vt, DBuf<double>& sDev, const DTensor& in, size_t split, size_t rank, size_t maxRank, double eps,
              bool jacobiOnTranspose, RankSpec& rs, const double* inTransposed = nullptr) {
	const size_t m = prodr(in.dims, 0, split), n = prodr(in.dims, split, in.degree()), k = std::min(m, n);
	if (rank == 0 || rank > k) throw NotSpeculable{};
	DBuf<double> u(m * k), s(k), vt(k * n);
	const size_t M = std::max(m, n), N = k, np = N + (N & 1);
	DBuf<double> ws(M * np + np * np + 2 * np);
	XTT_REQUIRE(!inTransposed || !(m > n || (m == n && !jacobiOnTranspose)), "svd: a transposed input is only used for m <= n");
	if (m > n || (m == n && !jacobiOnTranspose)) {
		jacobi_svd(in.data(), m, n, u.get(), s.get(), vt.get(), ws.get(), stream());
	} else {
		DBuf<double> at;
		if (!inTransposed) {
			at.resize(n * m);
			const size_t d2[2] = {m, n}, sh[2] = {1, 0};
			permute_f64(2, d2, sh, in.data(), at.get(), stream());
		}
		// written transposed by the kernel: U = V' (m×m) and Vt = U'ᵀ (m×n)
		jacobi_svd(inTransposed ? inTransposed : at.get(), n, m, vt.get(), s.get(), u.get(), ws.get(), stream(), true);
	}
	svd_rank_f64(s.get(), k, maxRank, eps, rs.slot((int)rank), stream());
	take_svd_factors(U, Vt, u, vt, cat(sl(in.dims, 0, split), {rank}), cat({rank}, sl(in.dims, split, in.degree())), m, n, k, rank);
	sDev = std::move(s);
}
// move_core(x, pos, false) from a canonical position, QC / CQ speculative
void move_core_spec(TT& x, size_t pos, RankSpec& rs) {
	if (!x.canonicalized) throw NotSpeculable{};
	for (size_t i = x.corePosition; i < pos; ++i) {  // transfer_core(i, i+1) with rank reduction
		DTensor& F = x.cores[i];
		DTensor Q, R;
		qc_spec(Q, R, F, F.degree() - 1, rs);
		F = std::move(Q);
		x.cores[i + 1] = contract(R, false, x.cores[i + 1], false, 1);
	}
	for (size_t i = x.corePosition; i > pos; --i) {  // transfer_core(i, i-1)
		DTensor& F = x.cores[i];
		DTensor Q, R;
		cq_spec(R, Q, F, 1, rs);
		F = std::move(Q);
		x.cores[i - 1] = contract(x.cores[i - 1], false, R, false, 1);
	}
	if (exceeds_maximal_ranks(x)) throw NotSpeculable{};
	x.corePosition = pos;
}
// round_edge(x, from, from - 1, maxRank, eps, true) with the edge's current rank assumed
void round_edge_spec(TT& x, size_t from, size_t to, size_t maxRank, double eps, RankSpec& rs) {
	DTensor& F = x.cores[from];
	DTensor& T = x.cores[to];
	const size_t r = F.dims[0];
	DBuf<double> sd;
	DTensor U, Vt;
	if (5.0 * (double)F.size() * (double)T.size() >= 6.0 * (double)r * r * r * r) {
		DTensor coreA, Fq;
		cq_spec(coreA, Fq, F, 1, rs);
		DTensor X;  // coreAᵀ, which is never formed: the Jacobi SVD runs on coreA (see round_edge)
		X.dims = {coreA.dims[1], coreA.dims[0]};
		svd_spec(U, Vt, sd, X, 1, r, maxRank, eps, true, rs, coreA.data());
		scale_by_f64(Vt.data(), sd.get(), r, Vt.size() / r, true, stream());
		F = contract(U, true, Fq, false, 1);
		T = contract(T, false, Vt, true, 1);
	} else {
		DTensor X = contract(F, true, T, true, 1);
		svd_spec(U, Vt, sd, X, F.degree() - 1, r, maxRank, eps, false, rs);
		scale_by_f64(Vt.data(), sd.get(), r, Vt.size() / r, true, stream());
		T = transpose2(Vt, 1);
		std::vector<size_t> back(U.degree());
		for (size_t i = 0; i + 1 < U.degree(); ++i) back[i] = i + 1;
		back[U.degree() - 1] = 0;
		F = reshuffle(U, back);
	}
}
}  // namespace

void round(TT& x, const dims_t& maxRanks, double eps) {  // ttNetwork.cpp:643-665
	const size_t d = x.order();
	XTT_REQUIRE(eps < 1, "_eps must be smaller than one.");
	XTT_REQUIRE(maxRanks.size() + 1 == d, "There must be exactly degree/N-1 maxRanks.");
	for (size_t r : maxRanks) XTT_REQUIRE(r > 0, "Trying to round a TTTensor to rank 0 is not possible.");
	const bool initCanon = x.canonicalized;
	const size_t initPos = x.corePosition;
	static thread_local std::map<std::vector<size_t>, bool> keptRanks;
	std::vector<size_t> key = x.dims;
	const dims_t r0 = x.ranks();
	key.insert(key.end(), r0.begin(), r0.end());
	key.insert(key.end(), maxRanks.begin(), maxRanks.end());
	uint64_t eb;
	std::memcpy(&eb, &eps, sizeof(eb));
	key.push_back((size_t)eb);
	key.push_back(x.op);
	key.push_back(initCanon);
	key.push_back(initPos);
	auto kept = keptRanks.find(key);
	if (d > 1 && initCanon && kept != keptRanks.end() && kept->second && !getenv("XTT_NO_SPECULATION")) {
		auto speculate = [&](TT& y, RankSpec& rs) {
			move_core_spec(y, d - 1, rs);
			for (size_t i = 0; i + 1 < d; ++i) round_edge_spec(y, d - 1 - i, d - 2 - i, maxRanks[d - 2 - i], eps, rs);
			y.corePosition = 0;
			y.canonicalized = true;
			move_core_spec(y, initPos, rs);
		};
		TT saved = tt_copy(x);
		bool ok = false;
		try {
			RankSpec rs(4 * d + 8);
			speculate(x, rs);
			ok = rs.holds();
		} catch (const NotSpeculable&) {
		}
		if (ok) return;
		x = std::move(saved);
	}
	move_core(x, d - 1, false);
	// after move_core(d-1) every core left of the edge being rounded is left-orthogonal
	for (size_t i = 0; i + 1 < d; ++i) round_edge(x, d - 1 - i, d - 2 - i, maxRanks[d - 2 - i], eps, true);
	x.corePosition = 0;
	x.canonicalized = true;
	if (initCanon) move_core(x, initPos, false);
	if (keptRanks.size() > 256) keptRanks.clear();
	keptRanks[key] = x.ranks() == r0;
}

double tt_dot(const TT& a, const TT& b) {
	XTT_REQUIRE(a.dims == b.dims, "dot: dimensions differ");
	DTensor E = dt_upload({1, 1}, std::vector<double>{1.0}.data());
	sync();
	for (size_t k = 0; k < a.order(); ++k) {
		const DTensor& A = a.cores[k];
		const DTensor& B = b.cores[k];
		const size_t e = a.ext(k);
		DTensor T({E.dims[0], e * B.dims.back()});
		gemm_f64(gemm_desc(0, 0, E.dims[0], e * B.dims.back(), E.dims[1]), E.data(), B.data(), T.data(), stream());
		DTensor En({A.dims.back(), B.dims.back()});
		gemm_f64(gemm_desc(1, 0, A.dims.back(), B.dims.back(), E.dims[0] * e), A.data(), T.data(), En.data(), stream());
		E = std::move(En);
	}
	double r = 0;
	read_back(&r, E.data(), 8);
	return r;
}

double tt_frob_norm(const TT& x) {  // ttNetwork.cpp:781-789
	if (x.canonicalized) return dt_norm(x.cores[x.corePosition]);
	return std::sqrt(std::max(0.0, tt_dot(x, x)));
}

void tt_scale(TT& x, double s) { dt_scale(x.cores[x.canonicalized ? x.corePosition : 0], s); }

// ttNetwork.cpp:796-847: block-diagonal cores, first core concatenated horizontally, last vertically.
TT tt_add(const TT& a, const TT& b) {
	XTT_REQUIRE(a.dims == b.dims, "The dimensions in TT sum must coincide.");
	TT x = tt_copy(a);
	const size_t d = x.order();
	const bool initCanon = x.canonicalized;
	const size_t initPos = x.corePosition;
	if (d <= 1) { axpy_f64(x.cores[0].data(), 1.0, b.cores[0].data(), x.cores[0].size(), stream()); return x; }
	for (size_t k = 0; k < d; ++k) {
		const DTensor& A = a.cores[k];
		const DTensor& B = b.cores[k];
		const size_t la = A.dims.front(), lb = B.dims.front(), ra = A.dims.back(), rb = B.dims.back(), e = x.ext(k);
		const size_t L = k == 0 ? 1 : la + lb, R = k + 1 == d ? 1 : ra + rb;
		dims_t nd = A.dims; nd.front() = L; nd.back() = R;
		DTensor t = dt_zeros(nd);
		const size_t lo = k == 0 ? 0 : la, ro = k + 1 == d ? 0 : ra;
		XTT_HIP(hipMemcpy2DAsync(t.data(), R * 8, A.data(), ra * 8, ra * 8, la * e, hipMemcpyDeviceToDevice, stream()));
		XTT_HIP(hipMemcpy2DAsync(t.data() + lo * e * R + ro, R * 8, B.data(), rb * 8, rb * 8, lb * e, hipMemcpyDeviceToDevice, stream()));
		x.set_component(k, std::move(t));
	}
	if (initCanon) move_core(x, initPos, false);
	return x;
}

// A·x (ttNetwork.cpp:888-967, ttStack.cpp:197-310): Z[(α,β), n, (α',β')] = Σ_m A[α,n,m,α']·x[β,m,β'].
TT tt_apply(const TT& A, const TT& x) {
	const size_t d = x.order();
	XTT_REQUIRE(A.op && A.order() == d, "tt_apply: operator/tensor mismatch");
	TT z;
	z.op = false; z.dims = sl(A.dims, 0, d); z.canonicalized = false; z.corePosition = 0;
	for (size_t k = 0; k < d; ++k) {
		const DTensor& a = A.cores[k];
		const DTensor& c = x.cores[k];
		const size_t p = a.dims[0], n = a.dims[1], m = a.dims[2], q = a.dims[3], r = c.dims[0], s = c.dims[2];
		XTT_REQUIRE(c.dims[1] == m, "tt_apply: mode mismatch");
		DTensor ap = reshuffle(a, {0, 1, 3, 2});          // (p, n, q, m)
		DTensor cp = reshuffle(c, {1, 0, 2});             // (m, r, s)
		DTensor T = contract(ap, false, cp, false, 1);   // (p, n, q, r, s)
		DTensor Z = reshuffle(T, {0, 2, 3, 1, 4});        // (p, r, n, q, s)
		Z.reshape({p * r, n, q * s});
		z.cores.push_back(std::move(Z));
	}
	return z;
}

// ‖Ax − b‖/‖b‖: TT of Ax − b (rank r_A·r_x + r_b), left-to-right Householder R sweep, norm of the last core
// (the SPD residual_f of als.cpp:483-486 evaluates frob_norm of the canonicalized TT difference).
double tt_residual(const TT& A, const TT& x, const TT& b) {
	TT z = tt_apply(A, x);
	TT nb = tt_copy(b);
	nb.canonicalized = false;
	dt_scale(nb.cores[0], -1.0);
	TT w = tt_add(z, nb);
	DTensor R;
	for (size_t k = 0; k < w.order(); ++k) {
		DTensor c = k == 0 ? dt_copy(w.cores[0]) : contract(R, false, w.cores[k], false, 1);
		if (k + 1 == w.order()) {
			TT bb = tt_copy(b);
			bb.canonicalized = false;
			return dt_norm(c) / std::sqrt(tt_dot(bb, bb));
		}
		R = qr_r_only(c, c.degree() - 1);
	}
	return 0.0;
}

// ------------------------------------------------------------------------------- TT constructors

// TTNetwork(const Tensor&, eps, maxRanks) (ttNetwork.cpp:112-160): operator modes interleaved (n_0, m_0, n_1, …),
// ghost modes of size 1 at both ends, then SVDs from the last component to the first on the device, each keeping
// min(maxRank, #{s_j > eps·s_0}) values and multiplying S into the remaining left part; canonical at 0.
TT tt_from_tensor(const dims_t& fullDims, const double* host, bool op, double eps, const dims_t& maxRanksIn) {
	const size_t N = op ? 2 : 1, deg = fullDims.size();
	XTT_REQUIRE(deg % N == 0, "Number of indicis must be even for TTOperator");
	XTT_REQUIRE(eps >= 0 && eps < 1, "_eps must be positive and smaller than one.");
	const size_t d = deg / N;
	XTT_REQUIRE(d > 0, "a TT of degree 0 is not supported");
	const dims_t maxRanks = maxRanksIn.empty() ? dims_t(d - 1, SIZE_MAX) : maxRanksIn;
	XTT_REQUIRE(maxRanks.size() == d - 1, "We need degree/N - 1 maximal ranks");
	for (size_t r : maxRanks) XTT_REQUIRE(r > 0, "Maximal ranks must be strictly positive.");
	TT x = tt_zero(fullDims, op);
	DTensor remains;
	if (op) {  // (n_0..n_{d-1}, m_0..m_{d-1}) → (n_0, m_0, n_1, m_1, …), on the host (the data arrives from there)
		const size_t total = prodr(fullDims, 0, deg);
		std::vector<double> inter(total);
		dims_t idims(deg);
		for (size_t i = 0; i < d; ++i) { idims[2 * i] = fullDims[i]; idims[2 * i + 1] = fullDims[d + i]; }
		std::vector<size_t> idx(deg, 0);  // multi-index of the source, last mode fastest
		for (size_t e = 0; e < total; ++e) {
			size_t o = 0;
			for (size_t i = 0; i < d; ++i) o = (o * idims[2 * i] + idx[i]) * idims[2 * i + 1] + idx[d + i];
			inter[o] = host[e];
			for (size_t j = deg; j-- > 0;) { if (++idx[j] < fullDims[j]) break; idx[j] = 0; }
		}
		remains = dt_upload(idims, inter.data());
		sync();  // `inter` is a local
	} else {
		remains = dt_upload(fullDims, host);
	}
	remains.reshape(cat(cat({1}, remains.dims), {1}));
	for (size_t pos = d - 1; pos > 0; --pos) {
		DTensor U, Vt;
		std::vector<double> S;
		DBuf<double> sd;
		calculate_svd(U, S, Vt, remains, 1 + pos * N, maxRanks[pos - 1], eps, &sd);
		const size_t rk = S.size();
		scale_by_f64(U.data(), sd.get(), U.size() / rk, rk, false, stream());  // U·S: column j times s_j
		x.set_component(pos, std::move(Vt));
		remains = std::move(U);
	}
	x.set_component(0, std::move(remains));
	x.canonicalized = true;
	x.corePosition = 0;
	sync();
	return x;
}

namespace {
TT from_host_cores(const dims_t& dims, bool op, const std::vector<std::pair<dims_t, std::vector<double>>>& cores) {
	TT x = tt_zero(dims, op);
	for (size_t k = 0; k < cores.size(); ++k) x.set_component(k, dt_upload(cores[k].first, cores[k].second.data()));
	sync();  // the host vectors are locals
	return x;
}
void check_dims(const dims_t& dims, bool op) {
	XTT_REQUIRE(dims.size() % (op ? 2 : 1) == 0, "Illegal number of dimensions for ttOperator");
	XTT_REQUIRE(!dims.empty(), "a TT of degree 0 is not supported");
	for (size_t n : dims) XTT_REQUIRE(n > 0, "Trying to construct a TTTensor with dimension 0 is not possible.");
}
}  // namespace

// TTNetwork::ones (ttNetwork.cpp:170-191): all-ones rank-1 components, then canonicalize_left (move_core(0)).
TT tt_ones(const dims_t& dims, bool op) {
	check_dims(dims, op);
	const size_t d = dims.size() / (op ? 2 : 1);
	std::vector<std::pair<dims_t, std::vector<double>>> cores;
	for (size_t k = 0; k < d; ++k) {
		dims_t cd = op ? dims_t{1, dims[k], dims[d + k], 1} : dims_t{1, dims[k], 1};
		cores.push_back({cd, std::vector<double>(prodr(cd, 0, cd.size()), 1.0)});
	}
	TT x = from_host_cores(dims, op, cores);
	move_core(x, 0, false);
	return x;
}

// TTNetwork<true>::identity (ttNetwork.cpp:194-221): components δ(i_k, j_k), then canonicalize_left.
TT tt_identity(const dims_t& dims) {
	check_dims(dims, true);
	const size_t d = dims.size() / 2;
	std::vector<std::pair<dims_t, std::vector<double>>> cores;
	for (size_t k = 0; k < d; ++k) {
		std::vector<double> h(dims[k] * dims[d + k], 0.0);
		for (size_t i = 0; i < std::min(dims[k], dims[d + k]); ++i) h[i * dims[d + k] + i] = 1.0;
		cores.push_back({{1, dims[k], dims[d + k], 1}, std::move(h)});
	}
	TT x = from_host_cores(dims, true, cores);
	move_core(x, 0, false);
	return x;
}

// TTNetwork::kronecker (ttNetwork.cpp:224-254): each component a Kronecker tensor (ones where every mode index is
// equal) with inner ranks min(dims), then canonicalize_left.
TT tt_kronecker(const dims_t& dims, bool op) {
	check_dims(dims, op);
	const size_t d = dims.size() / (op ? 2 : 1);
	const size_t minN = *std::min_element(dims.begin(), dims.end());
	std::vector<std::pair<dims_t, std::vector<double>>> cores;
	for (size_t k = 0; k < d; ++k) {
		const size_t l = k == 0 ? 1 : minN, r = k + 1 == d ? 1 : minN;
		const dims_t cd = op ? dims_t{l, dims[k], dims[d + k], r} : dims_t{l, dims[k], r};
		std::vector<double> h(prodr(cd, 0, cd.size()), 0.0);
		for (size_t i = 0; i < minN; ++i) {
			const size_t a = k == 0 ? 0 : i, c = k + 1 == d ? 0 : i;
			h[op ? ((a * cd[1] + i) * cd[2] + i) * cd[3] + c : (a * cd[1] + i) * cd[2] + c] = 1.0;
		}
		cores.push_back({cd, std::move(h)});
	}
	TT x = from_host_cores(dims, op, cores);
	move_core(x, 0, false);
	return x;
}

// TTNetwork::dirac (ttNetwork.cpp:257-283): rank-1 components with a single 1 at the position; as in xerus the
// result is not canonicalized (set_component of a non-core component clears the flag) unless it has one component.
TT tt_dirac(const dims_t& dims, const dims_t& position, bool op) {
	check_dims(dims, op);
	XTT_REQUIRE(dims.size() == position.size(), "Inconsitend number of entries in _dimensions and _position.");
	for (size_t i = 0; i < dims.size(); ++i) XTT_REQUIRE(position[i] < dims[i], "dirac position out of range");
	const size_t d = dims.size() / (op ? 2 : 1);
	std::vector<std::pair<dims_t, std::vector<double>>> cores;
	for (size_t k = 0; k < d; ++k) {
		const dims_t cd = op ? dims_t{1, dims[k], dims[d + k], 1} : dims_t{1, dims[k], 1};
		std::vector<double> h(prodr(cd, 0, cd.size()), 0.0);
		h[op ? position[k] * dims[d + k] + position[d + k] : position[k]] = 1.0;
		cores.push_back({cd, std::move(h)});
	}
	TT x = from_host_cores(dims, op, cores);
	if (d == 1) { x.canonicalized = true; x.corePosition = 0; }
	return x;
}

// TTOperator::transpose (ttNetwork.h:443-448): each component's two external modes swapped; the row and column
// dimensions are exchanged with them.
void tt_transpose(TT& x) {
	XTT_REQUIRE(x.op, "transpose: not a TTOperator");
	const size_t d = x.order();
	for (size_t k = 0; k < d; ++k) {
		x.cores[k] = reshuffle(x.cores[k], {0, 2, 1, 3});
		std::swap(x.dims[k], x.dims[d + k]);
	}
}

}  // namespace xtt
