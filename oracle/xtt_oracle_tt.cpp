// TEST INFRASTRUCTURE ONLY — the oracle: Tensor and TTNetwork level of xerus's TT path
// (src/xerus/tensor.cpp, tensorNetwork.cpp, ttNetwork.cpp, ttStack.cpp) restated in FP64.
#include "xtt_oracle.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <stdexcept>

namespace orc {

namespace {
size_t prod(const dims_t& d, size_t b, size_t e) { size_t p = 1; for (size_t i = b; i < e; ++i) p *= d[i]; return p; }
dims_t cat(dims_t a, const dims_t& b) { a.insert(a.end(), b.begin(), b.end()); return a; }
dims_t slice(const dims_t& d, size_t b, size_t e) { return dims_t(d.begin() + b, d.begin() + e); }
}

double frob_norm(const Tensor& t) {
	long double s = 0; for (double x : t.v) s += (long double)x * x; return std::sqrt((double)s);
}

// tensor.cpp:1252-1352: lhsTrans ⇒ lhs's FIRST numModes are contracted; rhsTrans ⇒ rhs's LAST numModes.
Tensor contract(const Tensor& lhs, bool lhsTrans, const Tensor& rhs, bool rhsTrans, size_t numModes) {
	const size_t ld = lhs.degree(), rd = rhs.degree();
	if (numModes > ld || numModes > rd) throw std::invalid_argument("contract: too many modes");
	dims_t lrem = lhsTrans ? slice(lhs.dims, numModes, ld) : slice(lhs.dims, 0, ld - numModes);
	dims_t lcon = lhsTrans ? slice(lhs.dims, 0, numModes) : slice(lhs.dims, ld - numModes, ld);
	dims_t rrem = rhsTrans ? slice(rhs.dims, 0, rd - numModes) : slice(rhs.dims, numModes, rd);
	dims_t rcon = rhsTrans ? slice(rhs.dims, rd - numModes, rd) : slice(rhs.dims, 0, numModes);
	if (lcon != rcon) throw std::invalid_argument("contract: dimension mismatch");
	Tensor out(cat(lrem, rrem));
	const size_t M = prod(lrem, 0, lrem.size()), N = prod(rrem, 0, rrem.size()), K = prod(lcon, 0, lcon.size());
	gemm(lhsTrans, rhsTrans, M, N, K, 1.0, lhs.data(), rhs.data(), out.data());
	return out;
}

// indexedTensor_tensor_evaluate.cpp:55-137: shuffle[i] = new position of old mode i.
Tensor reshuffle(const Tensor& t, const std::vector<size_t>& shuffle) {
	const size_t d = t.degree();
	if (shuffle.size() != d) throw std::invalid_argument("reshuffle: wrong permutation size");
	dims_t nd(d);
	for (size_t i = 0; i < d; ++i) nd[shuffle[i]] = t.dims[i];
	Tensor out(nd);
	std::vector<size_t> newStride(d, 1);
	for (size_t i = d; i-- > 1;) newStride[i - 1] = newStride[i] * nd[i];
	std::vector<size_t> idx(d, 0);
	for (size_t lin = 0; lin < t.size(); ++lin) {
		size_t o = 0;
		for (size_t i = 0; i < d; ++i) o += idx[i] * newStride[shuffle[i]];
		out.v[o] = t.v[lin];
		for (size_t i = d; i-- > 0;) { if (++idx[i] < t.dims[i]) break; idx[i] = 0; }
	}
	return out;
}

void calculate_qr(Tensor& Q, Tensor& R, const Tensor& in, size_t split) {
	const size_t m = prod(in.dims, 0, split), n = prod(in.dims, split, in.degree()), k = std::min(m, n);
	Q = Tensor(cat(slice(in.dims, 0, split), {k}));
	R = Tensor(cat({k}, slice(in.dims, split, in.degree())));
	qr(in.data(), m, n, Q.v, R.v);
}
void calculate_rq(Tensor& R, Tensor& Q, const Tensor& in, size_t split) {
	const size_t m = prod(in.dims, 0, split), n = prod(in.dims, split, in.degree()), k = std::min(m, n);
	R = Tensor(cat(slice(in.dims, 0, split), {k}));
	Q = Tensor(cat({k}, slice(in.dims, split, in.degree())));
	rq(in.data(), m, n, R.v, Q.v);
}
void calculate_qc(Tensor& Q, Tensor& C, const Tensor& in, size_t split) {
	const size_t m = prod(in.dims, 0, split), n = prod(in.dims, split, in.degree());
	std::vector<double> q, c;
	const size_t r = qc(in.data(), m, n, q, c);
	Q.dims = cat(slice(in.dims, 0, split), {r}); Q.v = std::move(q);
	C.dims = cat({r}, slice(in.dims, split, in.degree())); C.v = std::move(c);
}
void calculate_cq(Tensor& C, Tensor& Q, const Tensor& in, size_t split) {
	const size_t m = prod(in.dims, 0, split), n = prod(in.dims, split, in.degree());
	std::vector<double> q, c;
	const size_t r = cq(in.data(), m, n, c, q);
	C.dims = cat(slice(in.dims, 0, split), {r}); C.v = std::move(c);
	Q.dims = cat({r}, slice(in.dims, split, in.degree())); Q.v = std::move(q);
}

// tensor.cpp:1424-1489: rank = min(k, maxRank) (maxRank 0 = no cap), then the first j ≥ 1 with s_j ≤ eps·s_0.
void calculate_svd(Tensor& U, Tensor& S, Tensor& Vt, const Tensor& in, size_t split, size_t maxRank, double eps) {
	if (!(eps >= 0 && eps < 1)) throw std::invalid_argument("Epsilon must be fullfill 0 <= _eps < 1.");
	const size_t m = prod(in.dims, 0, split), n = prod(in.dims, split, in.degree());
	std::vector<double> u, s, vt;
	svd(in.data(), m, n, u, s, vt);
	const size_t k = s.size();
	size_t rank = k;
	if (maxRank != 0) rank = std::min(rank, maxRank);
	for (size_t j = 1; j < rank; ++j) if (s[j] <= eps * s[0]) { rank = j; break; }
	U = Tensor(cat(slice(in.dims, 0, split), {rank}));
	for (size_t i = 0; i < m; ++i) for (size_t j = 0; j < rank; ++j) U.v[i * rank + j] = u[i * k + j];
	S = Tensor({rank, rank});
	for (size_t i = 0; i < rank; ++i) S.v[i * rank + i] = s[i];
	Vt = Tensor(cat({rank}, slice(in.dims, split, in.degree())));
	std::copy(vt.begin(), vt.begin() + rank * n, Vt.v.begin());
}

// ---------------------------------------------------------------- TTNetwork

std::vector<size_t> TT::ranks() const {
	std::vector<size_t> r;
	for (size_t k = 0; k + 1 < cores.size(); ++k) r.push_back(cores[k].dims.back());
	return r;
}

void TT::set_component(size_t k, Tensor t) {  // ttNetwork.cpp:475-492
	if (t.degree() != (op ? 4u : 3u)) throw std::invalid_argument("set_component: wrong degree");
	dims[k] = t.dims[1];
	if (op) dims[order() + k] = t.dims[2];
	cores[k] = std::move(t);
	canonicalized = canonicalized && (corePosition == k);
}

dims_t reduce_to_maximal_ranks(dims_t ranks, const dims_t& dims, bool op) {  // ttNetwork.cpp:369-402
	const size_t N = op ? 2 : 1, d = dims.size() / N;
	if (d != ranks.size() + 1) throw std::invalid_argument("reduce_to_maximal_ranks: rank count");
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

TT tt_zero(const dims_t& dims, bool op) {  // ttNetwork.cpp:57-108: rank-1 zero TT, canonical at 0
	TT x; x.op = op; x.dims = dims;
	const size_t d = dims.size() / (op ? 2 : 1);
	for (size_t k = 0; k < d; ++k) x.cores.emplace_back(op ? dims_t{1, dims[k], dims[d + k], 1} : dims_t{1, dims[k], 1});
	x.canonicalized = true; x.corePosition = 0;
	return x;
}

TT tt_random(const dims_t& dims, const dims_t& ranks, bool op) {  // ttNetwork.h:128-155
	const size_t d = dims.size() / (op ? 2 : 1);
	if (ranks.size() + 1 != d) throw std::invalid_argument("Non-matching amount of ranks given to TTNetwork::random.");
	for (size_t r : ranks) if (r == 0) throw std::invalid_argument("Trying to construct random TTTensor with rank 0 is illegal.");
	TT x = tt_zero(dims_t(dims.size(), 1), op);
	const dims_t target = reduce_to_maximal_ranks(ranks, dims, op);
	for (size_t i = 0; i < d; ++i) {
		const size_t l = i == 0 ? 1 : target[i - 1], r = i + 1 == d ? 1 : target[i];
		Tensor t(op ? dims_t{l, dims[i], dims[d + i], r} : dims_t{l, dims[i], r});
		for (double& v : t.v) v = normal()(rng());
		x.set_component(i, std::move(t));
	}
	move_core(x, 0, false);
	return x;
}

// tensorNetwork.cpp:821-909 for two adjacent TT components.
void transfer_core(TT& x, size_t from, size_t to, bool allowRankReduction) {
	Tensor& F = x.cores[from];
	Tensor& T = x.cores[to];
	Tensor Q, R;
	if (to == from + 1) {
		if (allowRankReduction) calculate_qc(Q, R, F, F.degree() - 1); else calculate_qr(Q, R, F, F.degree() - 1);
		F = std::move(Q);
		T = contract(R, false, T, false, 1);
	} else if (to + 1 == from) {
		if (allowRankReduction) calculate_cq(R, Q, F, 1); else calculate_rq(R, Q, F, 1);
		F = std::move(Q);
		T = contract(T, false, R, false, 1);
	} else {
		throw std::invalid_argument("transfer_core: components not adjacent");
	}
}

namespace {
bool exceeds_maximal_ranks(const TT& x) {  // ttNetwork.cpp:349-359
	for (size_t i = 0; i < x.order(); ++i) {
		const Tensor& c = x.cores[i];
		const size_t e = x.ext(i);
		if (c.dims.front() > e * c.dims.back() || c.dims.back() > e * c.dims.front()) return true;
	}
	return false;
}
}

void move_core(TT& x, size_t pos, bool keepRank) {  // ttNetwork.cpp:581-628
	const size_t d = x.order();
	if (pos >= d) throw std::invalid_argument("Illegal core-position");
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

// tensorNetwork.cpp:678-818 for from = component k, to = component k-1 (the common edge is from's mode 0 and
// to's last mode, so transFrom = transTo = true).
void round_edge(TT& x, size_t from, size_t to, size_t maxRank, double eps) {
	if (to + 1 != from) throw std::invalid_argument("round_edge: expects (k, k-1)");
	Tensor& F = x.cores[from];
	Tensor& T = x.cores[to];
	const size_t r = F.dims[0];
	const double r4 = double(r) * r * r * r;
	Tensor S, X;
	if (5.0 * double(F.size()) * double(T.size()) >= 6.0 * r4) {
		Tensor coreA, coreB, Fq, Tq;
		calculate_cq(coreA, Fq, F, 1);
		calculate_qc(Tq, coreB, T, T.degree() - 1);
		X = contract(coreA, true, coreB, true, 1);
		Tensor U, Vt;
		calculate_svd(U, S, Vt, X, 1, maxRank, eps);
		Vt = contract(S, false, Vt, false, 1);
		F = contract(U, true, Fq, false, 1);
		T = contract(Tq, false, Vt, true, 1);
	} else {
		X = contract(F, true, T, true, 1);
		Tensor U, Vt;
		calculate_svd(U, S, Vt, X, F.degree() - 1, maxRank, eps);
		T = contract(Vt, true, S, true, 1);
		std::vector<size_t> back(U.degree());
		for (size_t i = 0; i + 1 < U.degree(); ++i) back[i] = i + 1;
		back[U.degree() - 1] = 0;
		F = reshuffle(U, back);
	}
}

void round(TT& x, const dims_t& maxRanks, double eps) {  // ttNetwork.cpp:643-665
	const size_t d = x.order();
	if (!(eps < 1)) throw std::invalid_argument("_eps must be smaller than one.");
	if (maxRanks.size() + 1 != d) throw std::invalid_argument("There must be exactly degree/N-1 maxRanks.");
	for (size_t r : maxRanks) if (r == 0) throw std::invalid_argument("Trying to round a TTTensor to rank 0 is not possible.");
	const bool initCanon = x.canonicalized;
	const size_t initPos = x.corePosition;
	move_core(x, d - 1, false);
	for (size_t i = 0; i + 1 < d; ++i) round_edge(x, d - 1 - i, d - 2 - i, maxRanks[d - 2 - i], eps);
	x.corePosition = 0; x.canonicalized = true;
	if (initCanon) move_core(x, initPos, false);
}

double tt_dot(const TT& a, const TT& b) {
	// Left-to-right transfer matrices E(α,β) = Σ a(α',n,α) b(β',n,β) E(α',β').
	std::vector<double> E(1, 1.0);
	size_t ra = 1, rb = 1;
	for (size_t k = 0; k < a.order(); ++k) {
		const Tensor& A = a.cores[k]; const Tensor& B = b.cores[k];
		const size_t n = a.ext(k), na = A.dims.back(), nb = B.dims.back();
		std::vector<double> T(ra * n * nb, 0.0);  // T(α', n, β) = Σ_β' E(α',β') B(β', n, β)
		gemm(false, false, ra, n * nb, rb, 1.0, E.data(), B.data(), T.data());
		std::vector<double> En(na * nb, 0.0);     // En(α, β) = Σ_{α',n} A(α',n,α) T(α',n,β)
		gemm(true, false, na, nb, ra * n, 1.0, A.data(), T.data(), En.data());
		E.swap(En); ra = na; rb = nb;
	}
	return E[0];
}

double tt_frob_norm(const TT& x) {  // ttNetwork.cpp:781-789
	if (x.canonicalized) return frob_norm(x.cores[x.corePosition]);
	return std::sqrt(std::max(0.0, tt_dot(x, x)));
}

TT tt_add(const TT& a, const TT& b) {  // ttNetwork.cpp:796-847 (operator+= on a copy of a)
	if (a.dims != b.dims) throw std::invalid_argument("The dimensions in TT sum must coincide.");
	TT x = a;
	const size_t d = x.order();
	const bool initCanon = x.canonicalized; const size_t initPos = x.corePosition;
	if (d <= 1) { for (size_t i = 0; i < x.cores[0].size(); ++i) x.cores[0].v[i] += b.cores[0].v[i]; return x; }
	for (size_t k = 0; k < d; ++k) {
		const Tensor& A = a.cores[k]; const Tensor& B = b.cores[k];
		const size_t la = A.dims.front(), lb = B.dims.front(), ra = A.dims.back(), rb = B.dims.back(), e = x.ext(k);
		const size_t L = k == 0 ? 1 : la + lb, R = k + 1 == d ? 1 : ra + rb;
		dims_t nd = A.dims; nd.front() = L; nd.back() = R;
		Tensor t(nd);
		const size_t lo = k == 0 ? 0 : la, ro = k + 1 == d ? 0 : ra;
		for (size_t i = 0; i < la; ++i) for (size_t j = 0; j < e; ++j) for (size_t l = 0; l < ra; ++l)
			t.v[(i * e + j) * R + l] += A.v[(i * e + j) * ra + l];
		for (size_t i = 0; i < lb; ++i) for (size_t j = 0; j < e; ++j) for (size_t l = 0; l < rb; ++l)
			t.v[((i + lo) * e + j) * R + l + ro] += B.v[(i * e + j) * rb + l];
		x.set_component(k, std::move(t));
	}
	if (initCanon) move_core(x, initPos, false);
	return x;
}

void tt_scale(TT& x, double alpha) {
	Tensor& c = x.cores[x.canonicalized ? x.corePosition : 0];
	for (double& v : c.v) v *= alpha;
}

// A·x as a TT of rank r_A·r_x (ttNetwork.cpp:888-967, ttStack.cpp:197-310): Z[(α,β),n,(α',β')] = Σ_m A[α,n,m,α']x[β,m,β'].
// The stack is returned uncanonicalized (the caller decides which core move to apply).
TT tt_apply(const TT& A, const TT& x) {
	const size_t d = x.order();
	if (!A.op || A.order() != d) throw std::invalid_argument("tt_apply: operator/tensor mismatch");
	TT z; z.op = false; z.dims = slice(A.dims, 0, d); z.canonicalized = false; z.corePosition = 0;
	for (size_t k = 0; k < d; ++k) {
		const Tensor& a = A.cores[k]; const Tensor& c = x.cores[k];
		const size_t p = a.dims[0], n = a.dims[1], m = a.dims[2], q = a.dims[3], r = c.dims[0], s = c.dims[2];
		if (c.dims[1] != m) throw std::invalid_argument("tt_apply: mode mismatch");
		Tensor t({p * r, n, q * s});
		for (size_t i = 0; i < p; ++i) for (size_t j = 0; j < n; ++j) for (size_t l = 0; l < m; ++l) for (size_t o = 0; o < q; ++o) {
			const double av = a.v[((i * n + j) * m + l) * q + o];
			if (av == 0.0) continue;
			for (size_t b = 0; b < r; ++b) for (size_t e = 0; e < s; ++e)
				t.v[((i * r + b) * n + j) * (q * s) + o * s + e] += av * c.v[(b * m + l) * s + e];
		}
		z.cores.push_back(std::move(t));
	}
	return z;
}

Tensor tt_full(const TT& x) {
	Tensor acc = x.cores[0];
	for (size_t k = 1; k < x.order(); ++k) acc = contract(acc, false, x.cores[k], false, 1);
	// drop the leading/trailing rank-1 modes; for operators reorder (n0,m0,n1,m1,…) → (n…, m…)
	const size_t d = x.order();
	Tensor t; t.v = acc.v;
	if (!x.op) { t.dims = x.dims; return t; }
	dims_t inter; for (size_t k = 0; k < d; ++k) { inter.push_back(x.dims[k]); inter.push_back(x.dims[d + k]); }
	t.dims = inter;
	std::vector<size_t> sh(2 * d);
	for (size_t k = 0; k < d; ++k) { sh[2 * k] = k; sh[2 * k + 1] = d + k; }
	return reshuffle(t, sh);
}

// ‖Ax − b‖/‖b‖ computed stably as the norm of the TT of Ax − b after a left-to-right QR sweep (the SPD
// residual_f of als.cpp:483-486 evaluates frob_norm(A·x − b) on the canonicalized TT sum).
double tt_residual(const TT& A, const TT& x, const TT& b) {
	TT z = tt_apply(A, x);
	TT nb = b; nb.canonicalized = false;
	tt_scale(nb, -1.0);
	z.canonicalized = false;
	TT w = tt_add(z, nb);
	w.canonicalized = false;
	move_core(w, w.order() - 1, true);
	TT bb = b; bb.canonicalized = false;
	return tt_frob_norm(w) / std::sqrt(tt_dot(bb, bb));
}


// ---------------------------------------------------------------- TT constructors

// TTNetwork(const Tensor&, eps, maxRanks) (ttNetwork.cpp:112-160): operator modes interleaved (n_0, m_0, n_1, …),
// ghost modes of size 1 at both ends, then SVDs from the last component to the first, each keeping
// min(maxRank, #{s_j > eps·s_0}) values and multiplying S into the remaining left part; canonical at 0.
TT tt_from_tensor(const Tensor& full, bool op, double eps, const dims_t& maxRanksIn) {
	const size_t N = op ? 2 : 1;
	if (full.degree() % N) throw std::invalid_argument("Number of indicis must be even for TTOperator");
	if (!(eps >= 0 && eps < 1)) throw std::invalid_argument("_eps must be positive and smaller than one.");
	const size_t d = full.degree() / N;
	if (d == 0) throw std::invalid_argument("a TT of degree 0 is not supported");
	const dims_t maxRanks = maxRanksIn.empty() ? dims_t(d - 1, SIZE_MAX) : maxRanksIn;
	if (maxRanks.size() != d - 1) throw std::invalid_argument("wrong number of maximal ranks");
	for (size_t r : maxRanks) if (r == 0) throw std::invalid_argument("Maximal ranks must be strictly positive.");
	TT x = tt_zero(full.dims, op);
	Tensor remains;
	if (op) {
		std::vector<size_t> sh(full.degree());
		for (size_t i = 0; i < d; ++i) { sh[i] = 2 * i; sh[d + i] = 2 * i + 1; }
		remains = reshuffle(full, sh);
	} else {
		remains = full;
	}
	remains.dims = cat(cat({1}, remains.dims), {1});
	for (size_t pos = d - 1; pos > 0; --pos) {
		Tensor U, S, Vt;
		calculate_svd(U, S, Vt, remains, 1 + pos * N, maxRanks[pos - 1], eps);
		x.set_component(pos, std::move(Vt));
		remains = contract(U, false, S, false, 1);
	}
	x.set_component(0, std::move(remains));
	x.canonicalized = true;
	x.corePosition = 0;
	return x;
}

// TTNetwork::ones (ttNetwork.cpp:170-191): all-ones rank-1 components, then canonicalize_left (move_core(0)).
TT tt_ones(const dims_t& dims, bool op) {
	TT x = tt_zero(dims, op);
	for (size_t k = 0; k < x.order(); ++k) {
		Tensor t(x.cores[k].dims);
		std::fill(t.v.begin(), t.v.end(), 1.0);
		x.set_component(k, std::move(t));
	}
	move_core(x, 0, false);
	return x;
}

// TTNetwork<true>::identity (ttNetwork.cpp:194-221): components δ(i_k, j_k), then canonicalize_left.
TT tt_identity(const dims_t& dims) {
	TT x = tt_zero(dims, true);
	const size_t d = x.order();
	for (size_t k = 0; k < d; ++k) {
		Tensor t({1, dims[k], dims[d + k], 1});
		for (size_t i = 0; i < std::min(dims[k], dims[d + k]); ++i) t.v[i * dims[d + k] + i] = 1.0;
		x.set_component(k, std::move(t));
	}
	move_core(x, 0, false);
	return x;
}

// TTNetwork::kronecker (ttNetwork.cpp:224-254): every component a Kronecker tensor (ones on its full diagonal) with
// inner ranks min(dims), then canonicalize_left.
TT tt_kronecker(const dims_t& dims, bool op) {
	TT x = tt_zero(dims, op);
	const size_t d = x.order();
	size_t minN = SIZE_MAX;
	for (size_t v : dims) minN = std::min(minN, v);
	for (size_t k = 0; k < d; ++k) {
		const size_t l = k == 0 ? 1 : minN, r = k + 1 == d ? 1 : minN;
		dims_t cd = op ? dims_t{l, dims[k], dims[d + k], r} : dims_t{l, dims[k], r};
		Tensor t(cd);
		// the diagonal entries of the unreshaped component: every mode index equal (boundary ranks are ghosts)
		for (size_t i = 0; i < minN; ++i) {
			const size_t a = k == 0 ? 0 : i, c = k + 1 == d ? 0 : i;
			const size_t idx = op ? ((a * cd[1] + i) * cd[2] + i) * cd[3] + c : (a * cd[1] + i) * cd[2] + c;
			t.v[idx] = 1.0;
		}
		x.set_component(k, std::move(t));
	}
	move_core(x, 0, false);
	return x;
}

// TTNetwork::dirac (ttNetwork.cpp:257-283): rank-1 components with a single 1 at the position (not canonicalized
// afterwards, as in xerus: set_component of components other than the core clears the flag).
TT tt_dirac(const dims_t& dims, const dims_t& position, bool op) {
	if (dims.size() != position.size()) throw std::invalid_argument("Inconsitend number of entries in _dimensions and _position.");
	for (size_t i = 0; i < dims.size(); ++i) if (position[i] >= dims[i]) throw std::invalid_argument("dirac position out of range");
	TT x = tt_zero(dims, op);
	const size_t d = x.order();
	for (size_t k = 0; k < d; ++k) {
		Tensor t(x.cores[k].dims);
		t.v[op ? position[k] * dims[d + k] + position[d + k] : position[k]] = 1.0;
		x.set_component(k, std::move(t));
	}
	if (d <= 1) { x.canonicalized = true; x.corePosition = 0; }  // the Tensor constructor's result
	return x;
}

// TTOperator::transpose (ttNetwork.h:443-448): every component's two external modes swapped (row and column
// dimensions exchanged with them).
void tt_transpose(TT& x) {
	if (!x.op) throw std::invalid_argument("transpose: not a TTOperator");
	const size_t d = x.order();
	for (size_t k = 0; k < d; ++k) {
		x.cores[k] = reshuffle(x.cores[k], {0, 2, 1, 3});
		std::swap(x.dims[k], x.dims[d + k]);
	}
}

} // namespace orc
