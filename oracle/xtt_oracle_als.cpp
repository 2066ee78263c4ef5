// TEST INFRASTRUCTURE ONLY — the oracle: ALSVariant::solve and its helpers (src/xerus/algorithms/als.cpp)
// restated in FP64 on the oracle's Tensor/TT layer.
//
// The non-SPD variants (ALS, DMRG) keep xerus's (r_x, r_A, r_A, r_x) AᵀA stacks and (r_b, r_A, r_x) Aᵀb stacks
// (als.cpp:409-440, 608-648) by running the SPD machinery on the TT operator B = AᵀA (rank r_A²) and the TT
// right-hand side Aᵀb: the stacks, local operators and local right-hand sides are the same contractions.
#include "xtt_oracle.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <stdexcept>

namespace orc {

namespace {

constexpr double XERUS_EPSILON = 8 * DBL_EPSILON;  // include/xerus/basic.h:50

Tensor ones(dims_t d) { Tensor t(d); std::fill(t.v.begin(), t.v.end(), 1.0); return t; }

// B_k[(p1,p2), n, n', (q1,q2)] = Σ_m A_k[p1, m, n, q1]·A_k[p2, m, n', q2]   (AᵀA, als.cpp:422-428, 617-623)
TT op_gram(const TT& A) {
	TT B = A; B.canonicalized = false;
	for (size_t k = 0; k < A.order(); ++k) {
		const Tensor& a = A.cores[k];
		const size_t p = a.dims[0], m = a.dims[1], n = a.dims[2], q = a.dims[3];
		Tensor t({p * p, n, n, q * q});
		for (size_t p1 = 0; p1 < p; ++p1) for (size_t p2 = 0; p2 < p; ++p2) for (size_t mm = 0; mm < m; ++mm)
			for (size_t i = 0; i < n; ++i) for (size_t q1 = 0; q1 < q; ++q1) {
				const double av = a.v[((p1 * m + mm) * n + i) * q + q1];
				if (av == 0.0) continue;
				for (size_t j = 0; j < n; ++j) for (size_t q2 = 0; q2 < q; ++q2)
					t.v[(((p1 * p + p2) * n + i) * n + j) * (q * q) + q1 * q + q2] += av * a.v[((p2 * m + mm) * n + j) * q + q2];
			}
		B.cores[k] = std::move(t);
		B.dims[k] = n; B.dims[A.order() + k] = n;
	}
	return B;
}

// (Aᵀb)_k[(p,β), n', (q,β')] = Σ_m A_k[p, m, n', q]·b_k[β, m, β']   (als.cpp:432-437, 639-645)
TT op_transpose_apply(const TT& A, const TT& b) {
	TT c; c.op = false; c.canonicalized = false; c.corePosition = 0;
	const size_t d = b.order();
	for (size_t k = 0; k < d; ++k) {
		const Tensor& a = A.cores[k]; const Tensor& bb = b.cores[k];
		const size_t p = a.dims[0], m = a.dims[1], n = a.dims[2], q = a.dims[3], r = bb.dims[0], s = bb.dims[2];
		Tensor t({p * r, n, q * s});
		for (size_t i = 0; i < p; ++i) for (size_t mm = 0; mm < m; ++mm) for (size_t j = 0; j < n; ++j) for (size_t o = 0; o < q; ++o) {
			const double av = a.v[((i * m + mm) * n + j) * q + o];
			if (av == 0.0) continue;
			for (size_t be = 0; be < r; ++be) for (size_t e = 0; e < s; ++e)
				t.v[((i * r + be) * n + j) * (q * s) + o * s + e] += av * bb.v[(be * m + mm) * s + e];
		}
		c.cores.push_back(std::move(t));
		c.dims.push_back(n);
	}
	return c;
}

// L'(c,q,c') = Σ L(a,p,a')·x(a,n,c)·B(p,n,n',q)·x(a',n',c')   (left stack push, als.cpp:580-583)
Tensor op_left(const Tensor& L, const Tensor& x, const Tensor& B) {
	Tensor T = contract(L, true, x, false, 1);           // (p, a', n, c)
	T = reshuffle(T, {2, 0, 3, 1});                      // (a', c, p, n)
	Tensor U = contract(T, false, B, false, 2);          // (a', c, n', q)
	U = reshuffle(U, {2, 0, 3, 1});                      // (c, q, a', n')
	return contract(U, false, x, false, 2);              // (c, q, c')
}
// R'(a,p,a') = Σ x(a,n,c)·B(p,n,n',q)·x(a',n',c')·R(c,q,c')   (right stack push, als.cpp:460-463, 596-597)
Tensor op_right(const Tensor& R, const Tensor& x, const Tensor& B) {
	Tensor T = contract(x, false, R, false, 1);          // (a, n, q, c')
	T = reshuffle(T, {0, 2, 3, 1});                      // (a, c', n, q)
	Tensor B2 = reshuffle(B, {2, 0, 3, 1});              // (n, q, p, n')
	Tensor U = contract(T, false, B2, false, 2);         // (a, c', p, n')
	U = reshuffle(U, {0, 3, 1, 2});                      // (a, p, n', c')
	return contract(U, false, x, true, 2);               // (a, p, a')
}
Tensor rhs_left(const Tensor& Lb, const Tensor& b, const Tensor& x) {   // (β', c)
	Tensor T = contract(Lb, true, b, false, 1);          // (a, n, β')
	return contract(T, true, x, false, 2);
}
Tensor rhs_right(const Tensor& Rb, const Tensor& b, const Tensor& x) {  // (β, a)
	Tensor T = contract(b, false, Rb, false, 1);         // (β, n, c)
	return contract(T, false, x, true, 2);
}

Tensor merge_op(const Tensor& B1, const Tensor& B2) {  // (p, n1 n2, n1' n2', q)
	Tensor t = contract(B1, false, B2, false, 1);        // (p, n1, n1', n2, n2', q)
	t = reshuffle(t, {0, 1, 3, 2, 4, 5});
	t.dims = {B1.dims[0], B1.dims[1] * B2.dims[1], B1.dims[2] * B2.dims[2], B2.dims[3]};
	return t;
}
Tensor merge_vec(const Tensor& x1, const Tensor& x2) {  // (a, n1 n2, c)
	Tensor t = contract(x1, false, x2, false, 1);
	t.dims = {x1.dims[0], x1.dims[1] * x2.dims[1], x2.dims[2]};
	return t;
}

// y(a,N,c) = Σ L(a,p,a')·B(p,N,N',q)·R(c,q,c')·x(a',N',c')  — the matrix-free local operator.
void local_matvec(const Tensor& L, const Tensor& B, const Tensor& R, const double* x, double* y, size_t a, size_t N, size_t c) {
	const size_t P = B.dims[0], Q = B.dims[3];
	std::vector<double> T1(a * P * N * c), T1r(a * c * P * N), Br(P * N * N * Q), T2(a * c * N * Q), T2r(a * N * Q * c);
	gemm(false, false, a * P, N * c, a, 1.0, L.data(), x, T1.data());                     // (a,p,N',c')
	for (size_t i = 0; i < a; ++i) for (size_t p = 0; p < P; ++p) for (size_t n = 0; n < N; ++n) for (size_t j = 0; j < c; ++j)
		T1r[((i * c + j) * P + p) * N + n] = T1[((i * P + p) * N + n) * c + j];                 // (a,c',p,N')
	for (size_t p = 0; p < P; ++p) for (size_t n = 0; n < N; ++n) for (size_t m = 0; m < N; ++m) for (size_t q = 0; q < Q; ++q)
		Br[((p * N + m) * N + n) * Q + q] = B.v[((p * N + n) * N + m) * Q + q];               // (p,N',N,q)
	gemm(false, false, a * c, N * Q, P * N, 1.0, T1r.data(), Br.data(), T2.data());       // (a,c',N,q)
	for (size_t i = 0; i < a; ++i) for (size_t j = 0; j < c; ++j) for (size_t n = 0; n < N; ++n) for (size_t q = 0; q < Q; ++q)
		T2r[((i * N + n) * Q + q) * c + j] = T2[((i * c + j) * N + n) * Q + q];               // (a,N,q,c')
	gemm(false, true, a * N, c, Q * c, 1.0, T2r.data(), R.data(), y);                     // (a,N,c)
}

Tensor dense_local_operator(const Tensor& L, const Tensor& B, const Tensor& R) {  // ((a,N,c),(a',N',c')), als.cpp:608-626
	Tensor L2 = reshuffle(L, {0, 2, 1});                  // (a, a', p)
	Tensor T = contract(L2, false, B, false, 1);          // (a, a', N, N', q)
	Tensor R2 = reshuffle(R, {1, 0, 2});                  // (q, c, c')
	Tensor T2 = contract(T, false, R2, false, 1);         // (a, a', N, N', c, c')
	return reshuffle(T2, {0, 3, 1, 4, 2, 5});
}

double dot(const std::vector<double>& a, const std::vector<double>& b) { long double s = 0; for (size_t i = 0; i < a.size(); ++i) s += (long double)a[i] * b[i]; return (double)s; }

struct ALSData {
	const ALSOptions& o;
	const TT* A;             // the user's operator (nullptr: approximate b)
	TT B;                    // SPD operator used by the stacks: A (SPD) or AᵀA
	TT rhs;                  // b (SPD / no operator) or Aᵀb
	TT& x;
	const TT& b;
	dims_t targetRank;
	double normB;
	bool canonAtEnd; size_t corePosAtEnd;
	std::vector<Tensor> opL, opR, rhsL, rhsR;
	size_t first = 0, second = 0, curr = 0;
	double lastEnergy2 = 1e102, lastEnergy = 1e101, energy = 1e100;
	size_t halfSweepCount = 0;
	bool increasing = true;
	ALSStats* stats;

	ALSData(const ALSOptions& o_, const TT* A_, TT& x_, const TT& b_, ALSStats* s)
		: o(o_), A(A_), x(x_), b(b_), targetRank(x_.ranks()), normB(tt_frob_norm(b_)),
		  canonAtEnd(x_.canonicalized), corePosAtEnd(x_.corePosition), stats(s) {
		if (A) {
			if (o.assumeSPD) { B = *A; rhs = b; } else { B = op_gram(*A); rhs = op_transpose_apply(*A, b); }
		} else {
			rhs = b;
		}
		prepare_x();
		prepare_stacks();
		curr = first;
	}

	// als.cpp:330-407
	void prepare_x() {
		const size_t d = x.order();
		size_t f = 0, dimProd = 1;
		while (f + 1 < d) {
			const size_t localDim = x.dims[f], newProd = dimProd * localDim;
			if (x.rank(f) < newProd) break;
			Tensor cur = x.cores[f];
			cur.dims = {cur.dims[0] * cur.dims[1], cur.dims[2]};
			x.set_component(f + 1, contract(cur, false, x.cores[f + 1], false, 1));
			Tensor id({dimProd, localDim, newProd});
			for (size_t i = 0; i < dimProd; ++i) for (size_t j = 0; j < localDim; ++j) id.v[(i * localDim + j) * newProd + i * localDim + j] = 1.0;
			x.set_component(f, std::move(id));
			f += 1; dimProd = newProd;
		}
		size_t l = d; dimProd = 1;
		while (l > f + o.sites) {
			const size_t localDim = x.dims[l - 1], newProd = dimProd * localDim;
			if (x.rank(l - 2) < newProd) break;
			Tensor cur = x.cores[l - 1];
			cur.dims = {cur.dims[0], cur.dims[1] * cur.dims[2]};
			Tensor prev = contract(x.cores[l - 2], false, cur, false, 1);
			x.set_component(l - 2, std::move(prev));
			Tensor id({newProd, localDim, dimProd});
			for (size_t j = 0; j < localDim; ++j) for (size_t k = 0; k < dimProd; ++k) id.v[((j * dimProd + k) * localDim + j) * dimProd + k] = 1.0;
			x.set_component(l - 1, std::move(id));
			l -= 1; dimProd = newProd;
		}
		if (canonAtEnd && corePosAtEnd < f) {
			x.corePosition = f; x.canonicalized = true;
		} else {
			if (canonAtEnd && corePosAtEnd >= l) { x.corePosition = l - 1; x.canonicalized = true; }
			move_core(x, f, true);
		}
		first = f; second = l;
	}

	// als.cpp:442-478
	void prepare_stacks() {
		const size_t d = x.order();
		opL = {ones({1, 1, 1})}; opR = {ones({1, 1, 1})};
		rhsL = {ones({1, 1})}; rhsR = {ones({1, 1})};
		for (size_t i = d - 1; i > first + o.sites - 1; --i) {
			if (A) opR.push_back(op_right(opR.back(), x.cores[i], B.cores[i]));
			rhsR.push_back(rhs_right(rhsR.back(), rhs.cores[i], x.cores[i]));
		}
		for (size_t i = 0; i < first; ++i) {
			if (A) opL.push_back(op_left(opL.back(), x.cores[i], B.cores[i]));
			rhsL.push_back(rhs_left(rhsL.back(), rhs.cores[i], x.cores[i]));
		}
	}

	Tensor merged_op() const { Tensor m = B.cores[curr]; for (size_t p = 1; p < o.sites; ++p) m = merge_op(m, B.cores[curr + p]); return m; }
	Tensor merged(const TT& t) const { Tensor m = t.cores[curr]; for (size_t p = 1; p < o.sites; ++p) m = merge_vec(m, t.cores[curr + p]); return m; }

	Tensor local_rhs() const {  // als.cpp:629-648
		Tensor bm = merged(rhs);
		Tensor T = contract(rhsL.back(), true, bm, false, 1);  // (a, N, β')
		return contract(T, false, rhsR.back(), false, 1);       // (a, N, c)
	}

	// The energy functional of als.cpp:480-545, evaluated from the local quantities at curr.
	double energy_f() const {
		if (o.useResidualForEndCriterion && A && o.assumeSPD) return tt_residual(*A, x, b);
		Tensor xm = merged(x);
		Tensor bl = local_rhs();
		const double xb = dot(xm.v, bl.v);
		if (!A) return 0.5 * dot(xm.v, xm.v) - xb;
		Tensor Bm = merged_op();
		std::vector<double> y(xm.size());
		local_matvec(opL.back(), Bm, opR.back(), xm.data(), y.data(), xm.dims[0], xm.dims[1], xm.dims[2]);
		const double xAx = dot(xm.v, y);
		if (o.assumeSPD) return std::abs(0.5 * xAx - xb);
		return std::sqrt(xAx - 2 * xb + normB * normB) / normB;
	}

	void local_solve() {
		Tensor xm = merged(x);
		const size_t a = xm.dims[0], N = xm.dims[1], c = xm.dims[2], n = a * N * c;
		Tensor bl = local_rhs();
		Tensor sol(xm.dims);
		if (!A) {
			sol = bl;
		} else if (o.localSolver == 0) {  // lapack_solver, als.cpp:268-273
			Tensor Al = dense_local_operator(opL.back(), merged_op(), opR.back());
			solve(Al.data(), n, bl.data(), sol.data());
		} else {                          // matrix-free CG, warm start from the current core
			Tensor Bm = merged_op();
			std::vector<double> xv = xm.v, r(n), p(n), q(n);
			local_matvec(opL.back(), Bm, opR.back(), xv.data(), q.data(), a, N, c);
			for (size_t i = 0; i < n; ++i) r[i] = bl.v[i] - q[i];
			const double bn = std::sqrt(dot(bl.v, bl.v));
			double rr = dot(r, r);
			// A warm start whose residual exceeds ‖b‖ is worse than zero, and CG's recursive residual cannot resolve
			// the true one below ε·‖A‖·max‖x_k‖ (A. Greenbaum, SIAM J. Matrix Anal. Appl. 18 (1997) 535-551): such a
			// solve starts from x = 0, r = b (the dense solve of lapack_solver has no start value at all).
			if (rr > bn * bn) { std::fill(xv.begin(), xv.end(), 0.0); r = bl.v; rr = bn * bn; }
			p = r;
			size_t it = 0;
			for (; it < o.cgMaxIter && std::sqrt(rr) > o.cgTol * bn; ++it) {
				local_matvec(opL.back(), Bm, opR.back(), p.data(), q.data(), a, N, c);
				const double alpha = rr / dot(p, q);
				for (size_t i = 0; i < n; ++i) { xv[i] += alpha * p[i]; r[i] -= alpha * q[i]; }
				const double rn = dot(r, r);
				const double beta = rn / rr;
				rr = rn;
				for (size_t i = 0; i < n; ++i) p[i] = r[i] + beta * p[i];
			}
			if (stats) stats->cgIterations += it;
			sol.v = xv;
		}
		if (stats) stats->localSolves++;
		if (o.sites == 1) {
			sol.dims = x.cores[curr].dims;
			x.set_component(curr, std::move(sol));
			return;
		}
		// sites == 2: SVD split capped at targetRank (als.cpp:276-295)
		sol.dims = {a, x.cores[curr].dims[1], x.cores[curr + 1].dims[1], c};
		Tensor U, S, Vt;
		calculate_svd(U, S, Vt, sol, 2, targetRank[curr], XERUS_EPSILON);
		if (increasing) {
			Tensor SV = contract(S, false, Vt, false, 1);
			x.set_component(curr, std::move(U));
			x.set_component(curr + 1, std::move(SV));
			x.corePosition = curr + 1;
		} else {
			Tensor US = contract(U, false, S, false, 1);
			x.set_component(curr + 1, std::move(Vt));
			x.set_component(curr, std::move(US));
			x.corePosition = curr;
		}
		x.canonicalized = true;
	}

	// als.cpp:651-700
	bool end_of_sweep() {
		if ((!increasing && curr == first) || (increasing && curr == second - o.sites)) {
			halfSweepCount += 1;
			lastEnergy2 = lastEnergy; lastEnergy = energy; energy = energy_f();
			if (stats) { stats->halfSweeps = halfSweepCount; stats->energies.push_back(energy); }
			if (halfSweepCount == o.numHalfSweeps || std::abs(lastEnergy - energy) < o.convergenceEpsilon ||
			    std::abs(lastEnergy2 - energy) < o.convergenceEpsilon || second - first <= o.sites) {
				if (canonAtEnd && o.preserveCorePosition) move_core(x, corePosAtEnd, true);
				return true;
			}
			increasing = !increasing;
		}
		return false;
	}

	// als.cpp:565-605. For sites ≥ 2 the stack pushed on a decreasing step is that of site curr+sites-1 (the site
	// leaving the window) after the core has been moved into the window; for sites = 1 this is exactly xerus's step.
	void next() {
		if (increasing) {
			move_core(x, curr + 1, true);
			if (A) { opR.pop_back(); opL.push_back(op_left(opL.back(), x.cores[curr], B.cores[curr])); }
			rhsR.pop_back(); rhsL.push_back(rhs_left(rhsL.back(), rhs.cores[curr], x.cores[curr]));
			curr++;
		} else {
			const size_t leaving = curr + o.sites - 1;
			move_core(x, leaving - 1, true);
			if (A) { opL.pop_back(); opR.push_back(op_right(opR.back(), x.cores[leaving], B.cores[leaving])); }
			rhsL.pop_back(); rhsR.push_back(rhs_right(rhsR.back(), rhs.cores[leaving], x.cores[leaving]));
			curr--;
		}
	}
};

} // namespace

double als_solve(const ALSOptions& o, const TT* A, TT& x, const TT& b, ALSStats* stats) {  // als.cpp:708-778
	if (o.sites == 0) throw std::invalid_argument("ALSVariant: sites must be > 0");
	if (x.order() == 0) throw std::invalid_argument("ALS: x must have degree > 0");
	if (x.dims != b.dims) throw std::invalid_argument("ALS: x and b dimensions differ");
	if (A) {
		if (A->dims.size() != 2 * b.dims.size()) throw std::invalid_argument("ALS: operator degree mismatch");
		for (size_t i = 0; i < x.order(); ++i)
			if (A->dims[i] != x.dims[i] || A->dims[i + x.order()] != x.dims[i]) throw std::invalid_argument("ALS: operator dimension mismatch");
	} else if (o.sites != 1) {
		throw std::invalid_argument("approximation dmrg not implemented yet");
	}
	ALSData data(o, A, x, b, stats);
	data.energy = data.energy_f();
	while (true) {
		data.local_solve();
		if (data.end_of_sweep()) return data.energy;
		data.next();
	}
}

} // namespace orc
