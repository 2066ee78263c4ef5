// ALSVariant::solve on the device (src/xerus/algorithms/als.cpp). The sweep control (prepare_x_for_als,
// prepare_stacks, check_for_end_of_sweep, move_to_next_index) follows xerus step by step; the stacks and
// local problems are device tensors and the local solve is a matrix-free CG instead of xerus's dense
// LAPACK solve (the dense local operator is 34 GB at d=16, n=16, r=64).
//
// Local operator of a site (a = left x-rank, n = mode size, c = right x-rank, p/q = operator ranks):
//   y(a,n,c) = Σ L(a,p,a')·A(p,n,n',q)·R(c,q,c')·x(a',n',c')
// evaluated as three MFMA GEMMs without any permutation of the CG vectors:
//   T1[(a,p)][(n',c')] = L[(a,p)][a']·x[a'][(n',c')]                      M=a·p  N=n·c  K=a
//   T2[a][(n,q)][c']   = Am[(n,q)][(p,n')]·T1[a][(p,n')][c']  (batched over a)  M=n·q  N=c  K=p·n
//   y[(a,n)][c]        = T2[(a,n)][(q,c')]·R[c][(q,c')]ᵀ                   M=a·n  N=c    K=q·c
// The left stack push reuses T1/T2 with x = X_k and finishes with L'[c][(q,c')] = X_kᵀ·T2; the right push
// is the left push on reversed cores (X(a,n,c) → (c,n,a), A(p,n,n',q) → (q,n,n',p)).
//
// Precision: FP64 runs CG on FP64 GEMMs. FP32 / BF16 run FP64 iterative refinement: r = b − A·x with the
// FP64 operator, an inner CG on FP32 vectors with FP32 (v_mfma_f32_16x16x4_f32) or BF16
// (v_mfma_f32_16x16x32_bf16, FP32 accumulate) GEMMs solves A·e = r/‖r‖, and x += ‖r‖·e in FP64.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <cstdlib>
#include <atomic>
#include <exception>
#include <map>
#include <mutex>
#include <set>
#include <thread>

#include "tt.hpp"

namespace xtt {

namespace {

constexpr double XERUS_EPSILON = 8 * DBL_EPSILON;  // include/xerus/basic.h:50

DTensor ones(const dims_t& d) {
	DTensor t(d);
	fill_f64(t.data(), 1.0, t.size(), stream());
	return t;
}

// B_k[(p1,p2), n, n', (q1,q2)] = Σ_m A[p1,m,n,q1]·A[p2,m,n',q2]  (the AᵀA stacks of als.cpp:422-428, 617-623)
TT op_gram(const TT& A) {
	TT B;
	B.op = true; B.dims = A.dims; B.canonicalized = false; B.corePosition = 0;
	for (size_t k = 0; k < A.order(); ++k) {
		const DTensor& a = A.cores[k];
		const size_t p = a.dims[0], n = a.dims[2], q = a.dims[3];
		DTensor am = reshuffle(a, {1, 0, 2, 3});                 // (m, p, n, q)
		DTensor g = contract(am, true, am, false, 1);            // (p, n, q, p', n', q')
		DTensor h = reshuffle(g, {0, 2, 4, 1, 3, 5});            // (p, p', n, n', q, q')
		h.reshape({p * p, n, n, q * q});
		B.cores.push_back(std::move(h));
		B.dims[k] = n; B.dims[A.order() + k] = n;
	}
	return B;
}

// (Aᵀb)_k[(p,β), n', (q,β')] = Σ_m A[p,m,n',q]·b[β,m,β']  (als.cpp:432-437, 639-645)
TT op_transpose_apply(const TT& A, const TT& b) {
	TT c;
	c.op = false; c.canonicalized = false; c.corePosition = 0;
	for (size_t k = 0; k < b.order(); ++k) {
		const DTensor& a = A.cores[k];
		const DTensor& bb = b.cores[k];
		const size_t p = a.dims[0], n = a.dims[2], q = a.dims[3], r = bb.dims[0], s = bb.dims[2];
		DTensor am = reshuffle(a, {1, 0, 2, 3});   // (m, p, n', q)
		DTensor bm = reshuffle(bb, {1, 0, 2});     // (m, β, β')
		DTensor g = contract(am, true, bm, false, 1);  // (p, n', q, β, β')
		DTensor h = reshuffle(g, {0, 2, 3, 1, 4});     // (p, β, n', q, β')
		h.reshape({p * r, n, q * s});
		c.cores.push_back(std::move(h));
		c.dims.push_back(n);
	}
	return c;
}

DTensor merge_op(const DTensor& B1, const DTensor& B2) {  // (p, n1·n2, n1'·n2', q)
	DTensor t = contract(B1, false, B2, false, 1);        // (p, n1, n1', n2, n2', q)
	DTensor u = reshuffle(t, {0, 1, 3, 2, 4, 5});
	u.reshape({B1.dims[0], B1.dims[1] * B2.dims[1], B1.dims[2] * B2.dims[2], B2.dims[3]});
	return u;
}
DTensor merge_vec(const DTensor& x1, const DTensor& x2) {  // (a, n1·n2, c)
	DTensor t = contract(x1, false, x2, false, 1);
	t.reshape({x1.dims[0], x1.dims[1] * x2.dims[1], x2.dims[2]});
	return t;
}

// Operator pieces of one site in the GEMM-friendly layouts, FP64 plus optional FP32 copies.
struct SiteOp {
	const DTensor* L = nullptr;    // (a, p, a')
	const DTensor* R = nullptr;    // (c, q, c')
	const DTensor* AmP = nullptr;  // (n, q, p, n'): AmOwn, or the caller's permuted copy of the operator core
	DTensor AmOwn;
	size_t a = 0, n = 0, c = 0, P = 0, Q = 0;
	SiteOp() = default;
	SiteOp(const SiteOp&) = delete;  // AmP may point into the object itself
	size_t N() const { return a * n * c; }
	const DTensor& am() const { return *AmP; }
};

// The operator core (p, n, n', q) in the layout of the local operator's second GEMM, (n, q, p, n').
const std::vector<size_t> kAmPerm{2, 0, 3, 1};
// The same for right stack pushes, which run the left push on the mirrored site: the core is first mirrored
// to (q, n, n', p) and then laid out as above; both permutations together are this one, (n, p, q, n').
const std::vector<size_t> kAmPermMirrored{1, 0, 3, 2};

// amCached: the core already permuted by kAmPerm (reused across the visits of one site), or null
void site_from(SiteOp& s, const DTensor& L, const DTensor& Acore, const DTensor& R, const DTensor* amCached = nullptr) {
	s.L = &L; s.R = &R;
	s.P = Acore.dims[0]; s.n = Acore.dims[1]; s.Q = Acore.dims[3];
	s.a = L.dims[0]; s.c = R.dims[0];
	XTT_REQUIRE(Acore.dims[1] == Acore.dims[2], "ALS: local operator must be square");
	XTT_REQUIRE(L.dims[1] == s.P && R.dims[1] == s.Q, "ALS: stack/operator rank mismatch");
	if (amCached) {
		s.AmP = amCached;
	} else {
		s.AmOwn = reshuffle(Acore, kAmPerm);
		s.AmP = &s.AmOwn;
	}
}
GemmDesc stage2_desc(size_t a, size_t n, size_t c, size_t P, size_t Q, const int* skip) {
	GemmDesc g = gemm_desc(0, 0, n * Q, c, P * n);
	g.batch = a;
	g.sA = 0;
	g.sB = (long long)(P * n * c);
	g.sC = (long long)(n * Q * c);
	g.skip = skip;
	return g;
}

template <class T>
void gemm_t(const GemmDesc& g, const T* A, const T* B, T* C, int prec);
template <>
void gemm_t<double>(const GemmDesc& g, const double* A, const double* B, double* C, int) { gemm_f64(g, A, B, C, stream()); }
template <>
void gemm_t<float>(const GemmDesc& g, const float* A, const float* B, float* C, int prec) {
	if (prec == P_BF16) gemm_bf16(g, A, B, C, stream()); else gemm_f32(g, A, B, C, stream());
}

// y = A_loc·x
template <class T>
void local_matvec(const T* L, const T* Am, const T* R, size_t a, size_t n, size_t c, size_t P, size_t Q, const T* x, T* y,
                  T* T1, T* T2, const int* skip, int prec) {
	GemmDesc g1 = gemm_desc(0, 0, a * P, n * c, a); g1.skip = skip;
	gemm_t<T>(g1, L, x, T1, prec);
	gemm_t<T>(stage2_desc(a, n, c, P, Q, skip), Am, T1, T2, prec);
	GemmDesc g3 = gemm_desc(0, 1, a * n, c, Q * c); g3.skip = skip;
	gemm_t<T>(g3, T2, R, y, prec);
}

// L'(c,q,c') = Σ L(a,p,a')·X(a,n,c)·A(p,n,n',q)·X(a',n',c'), with the operator core given as
// Am = reshuffle(A, kAmPerm) = (n, q, p, n')
DTensor op_left_am(const DTensor& L, const DTensor& X, const DTensor& Am) {
	const size_t a = X.dims[0], n = X.dims[1], c = X.dims[2], P = Am.dims[2], Q = Am.dims[1];
	XTT_REQUIRE(L.dims[0] == a && L.dims[2] == a && L.dims[1] == P && Am.dims[0] == n && Am.dims[3] == n,
	            "stack push: shape mismatch");
	DBuf<double> T1(a * P * n * c), T2(a * n * Q * c);
	gemm_f64(gemm_desc(0, 0, a * P, n * c, a), L.data(), X.data(), T1.get(), stream());
	gemm_f64(stage2_desc(a, n, c, P, Q, nullptr), Am.data(), T1.get(), T2.get(), stream());
	DTensor out({c, Q, c});
	gemm_f64(gemm_desc(1, 0, c, Q * c, a * n), X.data(), T2.get(), out.data(), stream());
	return out;
}
DTensor op_left(const DTensor& L, const DTensor& X, const DTensor& Acore) { return op_left_am(L, X, reshuffle(Acore, kAmPerm)); }
// R'(a,p,a') from R(c,q,c'): the left push on the mirrored site (X reversed to (c, n, a))
DTensor op_right(const DTensor& R, const DTensor& X, const DTensor& Acore) {
	return op_left_am(R, reshuffle(X, {2, 1, 0}), reshuffle(Acore, kAmPermMirrored));
}
// Lb'(β',c) = Σ Lb(β,a)·b(β,n,β')·X(a,n,c)
// The right-hand-side half of a stack push: two products, (g1: A1·B1 → C1) then (g2: A2·B2 → C2).
struct RhsPush {
	GemmDesc g1, g2;
	const double *A1 = nullptr, *B1 = nullptr, *A2 = nullptr, *B2 = nullptr;
	double *C1 = nullptr, *C2 = nullptr;
};
// Left push of the right-hand-side stack, Lb'(β',c) = Σ Lb(β,a)·b(β,n,β')·X(a,n,c): Y = Lb·X (K = a), then
// bᵀ·Y (K = β·n). Both products have short K, where Lb·b first would leave a K = a·n reduction onto a β'×c
// result. `Y` holds β·n·c elements.
RhsPush rhs_push_left(const DTensor& Lb, const DTensor& bc, const DTensor& X, double* Y, DTensor& out) {
	const size_t a = X.dims[0], n = X.dims[1], c = X.dims[2], beta = Lb.dims[0], beta2 = bc.dims[2];
	XTT_REQUIRE(Lb.dims[1] == a && bc.dims[0] == beta && bc.dims[1] == n, "stack push: right-hand side shape mismatch");
	out = DTensor({beta2, c});
	RhsPush r;
	r.g1 = gemm_desc(0, 0, beta, n * c, a); r.A1 = Lb.data(); r.B1 = X.data(); r.C1 = Y;
	r.g2 = gemm_desc(1, 0, beta2, c, beta * n); r.A2 = bc.data(); r.B2 = Y; r.C2 = out.data();
	return r;
}
// Right push, Rb'(β,a) = Σ b(β,n,β')·X(a,n,c)·Rb(β',c), on the unmirrored cores: Z(n,β',a) = Rb·X(:,n,:)ᵀ for
// every n (one batched product), then Rb' = b·Z with b as (β × n·β') and Z as (n·β' × a). `Z` holds n·β'·a.
RhsPush rhs_push_right(const DTensor& Rb, const DTensor& bc, const DTensor& X, double* Z, DTensor& out) {
	const size_t a = X.dims[0], n = X.dims[1], c = X.dims[2], beta = bc.dims[0], beta2 = bc.dims[2];
	XTT_REQUIRE(Rb.dims[0] == beta2 && Rb.dims[1] == c && bc.dims[1] == n, "stack push: right-hand side shape mismatch");
	out = DTensor({beta, a});
	RhsPush r;
	r.g1 = gemm_desc(0, 1, beta2, a, c);
	r.g1.ldb = n * c; r.g1.sA = 0; r.g1.sB = (long long)c; r.g1.sC = (long long)(beta2 * a); r.g1.batch = n;
	r.A1 = Rb.data(); r.B1 = X.data(); r.C1 = Z;
	r.g2 = gemm_desc(0, 0, beta, a, n * beta2); r.A2 = bc.data(); r.B2 = Z; r.C2 = out.data();
	return r;
}
// Both stacks of one site pushed together: the operator stack (op_left_am, when L is given) and the right-hand
// side stack; their independent products share launches (3 instead of 5).
void push_both(const DTensor* L, const DTensor& X, const DTensor* Am, const RhsPush& rp, DTensor* opOut) {
	if (!L) {
		gemm_f64(rp.g1, rp.A1, rp.B1, rp.C1, stream());
		gemm_f64(rp.g2, rp.A2, rp.B2, rp.C2, stream());
		return;
	}
	const size_t a = X.dims[0], n = X.dims[1], c = X.dims[2], P = Am->dims[2], Q = Am->dims[1];
	XTT_REQUIRE(L->dims[0] == a && L->dims[2] == a && L->dims[1] == P && Am->dims[0] == n && Am->dims[3] == n,
	            "stack push: shape mismatch");
	DBuf<double> T1(a * P * n * c), T2(a * n * Q * c);
	gemm_f64_pair(gemm_desc(0, 0, a * P, n * c, a), L->data(), X.data(), T1.get(), rp.g1, rp.A1, rp.B1, rp.C1, stream());
	gemm_f64_pair(stage2_desc(a, n, c, P, Q, nullptr), Am->data(), T1.get(), T2.get(), rp.g2, rp.A2, rp.B2, rp.C2, stream());
	*opOut = DTensor({c, Q, c});
	gemm_f64(gemm_desc(1, 0, c, Q * c, a * n), X.data(), T2.get(), opOut->data(), stream());
}

// A speculative ALS call recorded as a graph (TTGraph): the device slots it leaves its results in (the failed-site
// counter, then (x·b_loc, x·A_loc·x) per energy evaluation), the half-sweep count it was recorded with, and its
// statistics. Replays are accepted when no site failed and the replayed energies stop the sweeps at the same count.
struct ALSRecord {
	TTGraph g;
	void* slots = nullptr;        // arena memory: unsigned fail count, 4 bytes pad, then 2 doubles per energy
	size_t energies = 0;          // energy evaluations per call (the start and each half-sweep end)
	size_t halfSweeps = 0;        // of the eager speculative call the recording follows (0: none yet)
	bool shapeStop = false;       // second − first <= sites: the range alone ends the sweep
	bool failed = false;          // recording failed: stay eager for this key
	ALSStats delta;               // statistics of one call (energies excluded)
};

// Persistent CG workspace of the library (one per process). Every buffer a CG chunk touches has a fixed
// address while the capacities do not grow, so a chunk of CG iterations recorded once per local shape into a
// hipGraph is replayed for every site of that shape (and across ALS calls). Growing a buffer drops the graphs.
struct Workspace {
	size_t capN = 0, capT1 = 0, capT2 = 0, capL = 0, capR = 0, capAm = 0;
	DBuf<double> x, bl, r, p, q, T1, T2, L, R, Am, e64;
	DBuf<float> r32, p32, q32, e32, z32, T1f, T2f, L32, R32, Am32;
	DBuf<CGState> st;
	DBuf<double> partials;
	DBuf<double> fcg;  // the fused CG's seven vectors (FP64, or FP32 in the first half)
	DBuf<double> fsync;            // the fused CG's barrier slots
	unsigned long long fepoch = 0;  // launches on fsync so far
	std::map<std::vector<size_t>, hipGraphExec_t> graphs;
	std::set<std::vector<size_t>> seen;

	// Recorded speculative ALS calls (they read the workspace buffers in place), by call key.
	std::map<std::vector<size_t>, std::unique_ptr<struct ALSRecord>> calls;

	void drop_graphs() {
		for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second);
		graphs.clear();
		seen.clear();
		calls.clear();
	}
	// The executables are destroyed here, while the runtime (and any tool intercepting it) is still up, instead of
	// being left to the runtime's own teardown at process exit.
	~Workspace() {
		if (graphs.empty() && calls.empty()) return;
		(void)hipDeviceSynchronize();
		drop_graphs();
	}
	void ensure(size_t N, size_t t1, size_t t2, size_t nl, size_t nr, size_t nam) {
		if (N <= capN && t1 <= capT1 && t2 <= capT2 && nl <= capL && nr <= capR && nam <= capAm && st.get()) return;
		ArenaScope persistent(nullptr);  // the workspace outlives any recorded call
		sync();
		drop_graphs();
		capN = std::max(capN, N); capT1 = std::max(capT1, t1); capT2 = std::max(capT2, t2);
		capL = std::max(capL, nl); capR = std::max(capR, nr); capAm = std::max(capAm, nam);
		for (DBuf<double>* b : {&x, &bl, &r, &p, &q, &e64}) b->resize(capN);
		for (DBuf<float>* b : {&r32, &p32, &q32, &e32, &z32}) b->resize(capN);
		fcg.resize(7 * capN);
		T1.resize(capT1); T2.resize(capT2); T1f.resize(capT1); T2f.resize(capT2);
		L.resize(capL); L32.resize(capL); R.resize(capR); R32.resize(capR); Am.resize(capAm); Am32.resize(capAm);
		if (!st.get()) { st.resize(1); partials.resize(2 * 1024); }
		XTT_HIP(hipMemsetAsync(z32.get(), 0, capN * sizeof(float), stream()));
	}
	// Runs `body` (which enqueues kernels on the library stream): eagerly the first time a key is seen,
	// from a recorded hipGraph afterwards. No graphs while the launch profiler is on.
	template <class F>
	void run(const std::vector<size_t>& key, F&& body) {
		if (prof_enabled() || getenv("XTT_NO_GRAPHS")) { body(); return; }
		auto it = graphs.find(key);
		if (it == graphs.end()) {
			if (!seen.count(key)) { seen.insert(key); body(); return; }
			hipGraph_t g = nullptr;
			XTT_HIP(hipStreamBeginCapture(stream(), hipStreamCaptureModeRelaxed));
			try {
				body();
			} catch (...) {
				(void)hipStreamEndCapture(stream(), &g);
				if (g) (void)hipGraphDestroy(g);
				throw;
			}
			XTT_HIP(hipStreamEndCapture(stream(), &g));
			hipGraphExec_t e = nullptr;
			const hipError_t ie = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
			(void)hipGraphDestroy(g);
			XTT_HIP(ie);
			it = graphs.emplace(key, e).first;
		}
		XTT_HIP(hipGraphLaunch(it->second, stream()));
	}
};

// Set on the batch driver's worker threads: their streams run concurrently, so a solve must not take a
// device-wide grid (fused CG) there.
thread_local bool t_batch_worker = false;

Workspace& workspace() {
	static thread_local Workspace w;  // per thread: its graphs replay on that thread's stream
	return w;
}
// A batch worker drops its recorded CG graphs before its stream goes (the buffers follow at thread exit, into the
// pool of the library stream).
void release_thread_graphs() { workspace().drop_graphs(); }

struct ALSData {
	const ALSOptions& o;
	const TT* A;
	TT Bown, rhsOwn;             // AᵀA and Aᵀb when not SPD (copies of b when it aliases x)
	const TT* Bp = nullptr;      // operator of the stacks: A (SPD) or AᵀA
	const TT* rhsp = nullptr;    // b or Aᵀb
	const TT& B() const { return *Bp; }
	const TT& rhs() const { return *rhsp; }
	// Per core k, filled on first use: the operator core permuted for left pushes and local operators (kAmPerm)
	// and for right pushes (kAmPermMirrored).
	std::vector<DTensor> amLeft, amRight;
	TT& x;
	const TT& b;
	dims_t targetRank;
	double normB = -1;  // ‖b‖, taken on first use (only the non-SPD energy needs it)
	bool canonAtEnd;
	size_t corePosAtEnd;
	std::vector<DTensor> opL, opR, rhsL, rhsR;
	size_t first = 0, second = 0, curr = 0;
	double lastEnergy2 = 1e102, lastEnergy = 1e101, energy = 1e100;
	size_t halfSweepCount = 0;
	bool increasing = true;
	ALSStats* stats;
	Workspace& ws = workspace();
	// Speculative call (see als_solve): every local solve is taken as converged at its start residual; the
	// residual checks run on the device and count the sites that were not into *specFail, read once at the end.
	bool spec = false;
	unsigned* specFail = nullptr;
	// Recording a speculative call (graphMode): energies go to device slots (x·b_loc, x·A_loc·x at
	// energySlots[2e], [2e + 1]) instead of being read back, and the sweep ends after plannedHalfSweeps.
	bool graphMode = false;
	double* energySlots = nullptr;
	size_t energyCount = 0;
	size_t plannedHalfSweeps = 0;
	size_t cgIters = 0;  // inner CG iterations of this call
	// Fused CG solves (one grid-synchronised launch each) count their iterations and matvec flops on the device;
	// finish_counts() reads them once at the end of the call.
	DBuf<double> itDev;
	bool usedFused = false;
	bool noFused = false;        // the rerun after a fused solve ran out of time uses the chunked CG only
	bool fusedTimedOut = false;

	ALSData(const ALSOptions& o_, const TT* A_, TT& x_, const TT& b_, ALSStats* s)
	    : o(o_), A(A_), x(x_), b(b_), targetRank(x_.ranks()), canonAtEnd(x_.canonicalized),
	      corePosAtEnd(x_.corePosition), stats(s) {
		// A and b are read only (const in xerus's signature); b is copied only when it is x itself
		if (A && !o.assumeSPD) {
			Bown = op_gram(*A);
			rhsOwn = op_transpose_apply(*A, b);
			Bp = &Bown;
			rhsp = &rhsOwn;
		} else {
			Bp = A;
			if (&b == &x) { rhsOwn = tt_copy(b); rhsp = &rhsOwn; } else rhsp = &b;
		}
		amLeft.resize(x.order()); amRight.resize(x.order());
		prepare_x();
		prepare_stacks();
		curr = first;
	}

	void prepare_x() {  // als.cpp:330-407
		const size_t d = x.order();
		size_t f = 0, dimProd = 1;
		// A full-rank boundary core becomes the identity of its unfolding (dimProd·localDim = newProd), absorbed into
		// its neighbour; built on the device (no host data, so a recorded call can contain it).
		auto identity = [](dims_t dims, size_t n) {
			DTensor t(std::move(dims));
			eye_f64(t.data(), n, stream());
			return t;
		};
		while (f + 1 < d) {
			const size_t localDim = x.dims[f], newProd = dimProd * localDim;
			if (x.rank(f) < newProd) break;
			DTensor cur = dt_copy(x.cores[f]);
			cur.reshape({cur.dims[0] * cur.dims[1], cur.dims[2]});
			x.set_component(f + 1, contract(cur, false, x.cores[f + 1], false, 1));
			x.set_component(f, identity({dimProd, localDim, newProd}, newProd));
			f += 1; dimProd = newProd;
		}
		size_t l = d;
		dimProd = 1;
		while (l > f + o.sites) {
			const size_t localDim = x.dims[l - 1], newProd = dimProd * localDim;
			if (x.rank(l - 2) < newProd) break;
			DTensor cur = dt_copy(x.cores[l - 1]);
			cur.reshape({cur.dims[0], cur.dims[1] * cur.dims[2]});
			x.set_component(l - 2, contract(x.cores[l - 2], false, cur, false, 1));
			x.set_component(l - 1, identity({newProd, localDim, dimProd}, newProd));
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

	void prepare_stacks() {  // als.cpp:442-478
		const size_t d = x.order();
		opL.clear(); opR.clear(); rhsL.clear(); rhsR.clear();
		opL.push_back(ones({1, 1, 1})); opR.push_back(ones({1, 1, 1}));
		rhsL.push_back(ones({1, 1})); rhsR.push_back(ones({1, 1}));
		for (size_t i = d - 1; i > first + o.sites - 1; --i) push_right(i);
		for (size_t i = 0; i < first; ++i) push_left(i);
	}
	void count_env(const DTensor& X, const DTensor& Acore) {
		if (!stats) return;
		const double a = (double)X.dims[0], n = (double)X.dims[1], c = (double)X.dims[2], p = (double)Acore.dims[0], q = (double)Acore.dims[3];
		stats->flopsEnv += 2 * a * n * c * (a * p + n * p * q + c * q);
	}
	const DTensor& am_left(size_t k) {
		if (amLeft[k].dims.empty()) amLeft[k] = reshuffle(B().cores[k], kAmPerm);
		return amLeft[k];
	}
	const DTensor& am_right(size_t k) {
		if (amRight[k].dims.empty()) amRight[k] = reshuffle(B().cores[k], kAmPermMirrored);
		return amRight[k];
	}
	// stack pushes of site i (operator and right-hand side); the right push mirrors x's core once for both
	void push_left(size_t i) {
		const DTensor& X = x.cores[i];
		DBuf<double> Y(rhsL.back().dims[0] * X.dims[1] * X.dims[2]);
		DTensor op, rb;
		const RhsPush rp = rhs_push_left(rhsL.back(), rhs().cores[i], X, Y.get(), rb);
		push_both(A ? &opL.back() : nullptr, X, A ? &am_left(i) : nullptr, rp, &op);
		if (A) { opL.push_back(std::move(op)); count_env(X, B().cores[i]); }
		rhsL.push_back(std::move(rb));
	}
	// the operator push runs as the left push on the mirrored site; the right-hand side needs no mirrored cores
	void push_right(size_t i) {
		const DTensor& X = x.cores[i];
		DBuf<double> Z(X.dims[1] * rhsR.back().dims[0] * X.dims[0]);
		DTensor op, rb;
		const RhsPush rp = rhs_push_right(rhsR.back(), rhs().cores[i], X, Z.get(), rb);
		if (A) {
			const DTensor Xr = reshuffle(X, {2, 1, 0});
			push_both(&opR.back(), Xr, &am_right(i), rp, &op);
			opR.push_back(std::move(op));
			count_env(X, B().cores[i]);
		} else {
			push_both(nullptr, X, nullptr, rp, nullptr);
		}
		rhsR.push_back(std::move(rb));
	}

	DTensor merged_op() const { DTensor m = dt_copy(B().cores[curr]); for (size_t p = 1; p < o.sites; ++p) m = merge_op(m, B().cores[curr + p]); return m; }
	DTensor merged(const TT& t) const { DTensor m = dt_copy(t.cores[curr]); for (size_t p = 1; p < o.sites; ++p) m = merge_vec(m, t.cores[curr + p]); return m; }
	// The site's operator pieces; with one site the operator core is the cached permutation (no copy).
	void site_op(SiteOp& s, DTensor& keep) {
		if (o.sites == 1) {
			site_from(s, opL.back(), B().cores[curr], opR.back(), &am_left(curr));
		} else {
			keep = merged_op();
			site_from(s, opL.back(), keep, opR.back());
		}
	}

	// One site: the local right-hand side (as local_rhs) and q = A_loc·x, their independent products paired.
	DTensor rhs_and_matvec(const SiteOp& s, const double* xv, double* q) {
		const DTensor &Lb = rhsL.back(), &Rb = rhsR.back(), &bc = rhs().cores[curr];
		const size_t a = s.a, n = s.n, c = s.c, beta = Lb.dims[0], beta2 = Rb.dims[0];
		XTT_REQUIRE(o.sites == 1 && Lb.dims[1] == a && bc.dims[0] == beta && bc.dims[1] == n && bc.dims[2] == beta2 && Rb.dims[1] == c,
		            "ALS: local right-hand side shape mismatch");
		reserve(s);
		DTensor Tb({a, n, beta2});
		gemm_f64_pair(gemm_desc(0, 0, a * s.P, n * c, a), s.L->data(), xv, ws.T1.get(),          // T1 = L·x
		              gemm_desc(1, 0, a, n * beta2, beta), Lb.data(), bc.data(), Tb.data(), stream());  // Tb = Lbᵀ·b
		DTensor bl({a, n, c});
		gemm_f64_pair(stage2_desc(a, n, c, s.P, s.Q, nullptr), s.am().data(), ws.T1.get(), ws.T2.get(),
		              gemm_desc(0, 0, a * n, c, beta2), Tb.data(), Rb.data(), bl.data(), stream());      // bl = Tb·Rb
		gemm_f64(gemm_desc(0, 1, a * n, c, s.Q * c), ws.T2.get(), s.R->data(), q, stream());
		count_matvec(s);
		return bl;
	}

	DTensor local_rhs() const {  // als.cpp:629-648
		DTensor keep;
		if (o.sites > 1) keep = merged(rhs());
		const DTensor& bm = o.sites > 1 ? keep : rhs().cores[curr];
		DTensor T = contract(rhsL.back(), true, bm, false, 1);  // (a, N, β')
		return contract(T, false, rhsR.back(), false, 1);       // (a, N, c)
	}

	void reserve(const SiteOp& s) {
		ws.ensure(s.N(), s.a * s.P * s.n * s.c, s.a * s.n * s.Q * s.c, s.L->size(), s.R->size(), s.am().size());
	}
	void count_matvec(const SiteOp& s) {
		if (stats) { stats->matvecs64++; stats->flopsMatvec += 2.0 * s.a * s.n * s.c * (s.a * s.P + s.n * s.P * s.Q + s.c * s.Q); }
	}
	// matvec with the site's own (non-workspace) operator pieces; used outside the CG loops
	void matvec64(const SiteOp& s, const double* in, double* out, const int* skip) {
		reserve(s);
		local_matvec<double>(s.L->data(), s.am().data(), s.R->data(), s.a, s.n, s.c, s.P, s.Q, in, out, ws.T1.get(), ws.T2.get(), skip, P_F64);
		if (stats) { stats->matvecs64++; stats->flopsMatvec += 2.0 * s.a * s.n * s.c * (s.a * s.P + s.n * s.P * s.Q + s.c * s.Q); }
	}

	double energy_f() {  // als.cpp:480-545, evaluated from the local quantities at curr
		if (graphMode) {  // one-site SPD: the two dot products into the next pair of slots, nothing read back
			const DTensor& xm = x.cores[curr];
			DTensor bl = local_rhs();
			DTensor keepOp;
			SiteOp s;
			site_op(s, keepOp);
			DTensor y(xm.dims);
			matvec64(s, xm.data(), y.data(), nullptr);
			double* dst = energySlots + 2 * energyCount++;
			dot_f64(xm.data(), bl.data(), xm.size(), dst, scratch.partials.get(), scratch.counter.get(), stream());
			dot_f64(xm.data(), y.data(), xm.size(), dst + 1, scratch.partials.get(), scratch.counter.get(), stream());
			return std::nan("");
		}
		if (o.useResidualForEndCriterion && A && o.assumeSPD) return tt_residual(*A, x, b);
		DTensor keepX;
		if (o.sites > 1) keepX = merged(x);
		const DTensor& xm = o.sites > 1 ? keepX : x.cores[curr];  // read only
		DTensor bl = local_rhs();
		if (!A) {
			const auto [xx, xb] = dt_dot2(xm, xm, xm, bl);
			return 0.5 * xx - xb;
		}
		DTensor keepOp;
		SiteOp s;
		site_op(s, keepOp);
		DTensor y(xm.dims);
		matvec64(s, xm.data(), y.data(), nullptr);
		const auto [xb, xAx] = dt_dot2(xm, bl, xm, y);  // one read-back for both
		if (o.assumeSPD) return std::abs(0.5 * xAx - xb);
		if (normB < 0) normB = tt_frob_norm(b);
		return std::sqrt(xAx - 2 * xb + normB * normB) / normB;
	}

	CGState read_state() {
		CGState h{};
		read_back(&h, ws.st.get(), sizeof(CGState));
		return h;
	}

	// Copies the site's operator pieces into the workspace (fixed addresses for the recorded chunks).
	void stage_site(const SiteOp& s, bool lowp) {
		reserve(s);
		XTT_HIP(hipMemcpyAsync(ws.L.get(), s.L->data(), s.L->size() * 8, hipMemcpyDeviceToDevice, stream()));
		XTT_HIP(hipMemcpyAsync(ws.R.get(), s.R->data(), s.R->size() * 8, hipMemcpyDeviceToDevice, stream()));
		XTT_HIP(hipMemcpyAsync(ws.Am.get(), s.am().data(), s.am().size() * 8, hipMemcpyDeviceToDevice, stream()));
		if (lowp) {
			convert_f64_f32(ws.L.get(), ws.L32.get(), s.L->size(), stream());
			convert_f64_f32(ws.R.get(), ws.R32.get(), s.R->size(), stream());
			convert_f64_f32(ws.Am.get(), ws.Am32.get(), s.am().size(), stream());
		}
	}
	void ws_matvec64(const SiteOp& s, const double* in, double* out, const int* skip) {
		local_matvec<double>(ws.L.get(), ws.Am.get(), ws.R.get(), s.a, s.n, s.c, s.P, s.Q, in, out, ws.T1.get(), ws.T2.get(), skip, P_F64);
	}

	static constexpr int kChunk = 16;

	template <class T>
	bool fused_ok(const SiteOp& s) const {
		return !noFused && !t_batch_worker && !getenv("XTT_NO_FUSED_CG") && fused_cg_fits<T>(s.a, s.n, s.c, s.P, s.Q);
	}
	// false when the launch could not be made (nothing was enqueued)
	template <class T>
	bool run_fused(const SiteOp& s, FusedCG d) {
		if (!usedFused) {
			itDev.resize(2);
			fill_f64(itDev.get(), 0.0, 2, stream());
			usedFused = true;
		}
		d.L = s.L->data(); d.Am = s.am().data(); d.R = s.R->data();
		d.vec = ws.fcg.get();
		d.T2 = sizeof(T) == 8 ? (void*)ws.T2.get() : (void*)ws.T2f.get();
		if (!ws.fsync.get()) {
			ws.fsync.resize(FUSED_CG_SYNC);
			XTT_HIP(hipMemsetAsync(ws.fsync.get(), 0, FUSED_CG_SYNC * sizeof(double), stream()));
		}
		d.sync = ws.fsync.get();
		d.epoch = ++ws.fepoch;
		d.st = ws.st.get();
		d.iterAcc = itDev.get();
		d.flopsPerIter = 2.0 * s.a * s.n * s.c * (s.a * s.P + s.n * s.P * s.Q + s.c * s.Q);
		d.a = s.a; d.n = s.n; d.c = s.c; d.P = s.P; d.Q = s.Q;
		return fused_cg<T>(d, stream());
	}
	// Reads the fused solves' counters once per call. A barrier wait that ran out of time (the device was shared
	// and did not keep the grid resident) leaves that site's core unsolved: the call is then rerun on the chunked
	// CG (als_solve), so fusedTimedOut is all that is recorded here.
	void finish_counts() {
		if (!usedFused) return;
		double h[2];
		read_back(h, itDev.get(), sizeof h);
		if (std::isnan(h[0])) { fusedTimedOut = true; return; }
		const size_t it = (size_t)h[0];
		cgIters += it;
		if (!stats) return;
		stats->cgIterations += it;
		if (o.precision == P_F64) { stats->matvecs64 += it; stats->flopsMatvec += h[1]; } else stats->matvecsLow += it;
	}

	// FP64 CG on A_loc·x = b, warm start from x (device vector, overwritten). The start residual is read back
	// first, so a site that is already converged costs one matvec and one synchronisation.
	void cg64(SiteOp& s, double* xout, const double* blin) {
		const size_t N = s.N();
		reserve(s);
		if (spec) {  // cg_start's test, on the device: converged iff ‖b − Ax‖² <= tol²·‖b‖² (or no iterations allowed)
			matvec64(s, xout, ws.q.get(), nullptr);
			scratch.check(ws.r.get(), blin, ws.q.get(), N, o.localTol * o.localTol, o.maxLocalIter > 0 ? specFail : nullptr);
			return;
		}
		if (fused_ok<double>(s)) {  // no synchronisation: the kernel reads cg_start's verdict itself
			matvec64(s, xout, ws.q.get(), nullptr);
			cg_state_init(ws.st.get(), o.localTol * o.localTol, o.maxLocalIter, stream());
			cg_start<double>(ws.r.get(), ws.p.get(), blin, ws.q.get(), N, ws.st.get(), ws.partials.get(), stream());
			FusedCG d;
			d.r0 = ws.r.get(); d.b = blin; d.x = xout; d.mode = 0;
			if (run_fused<double>(s, d)) return;
		}
		double* x = ws.x.get();
		const double* bl = ws.bl.get();
		XTT_HIP(hipMemcpyAsync(x, xout, N * 8, hipMemcpyDeviceToDevice, stream()));
		XTT_HIP(hipMemcpyAsync(ws.bl.get(), blin, N * 8, hipMemcpyDeviceToDevice, stream()));
		matvec64(s, x, ws.q.get(), nullptr);
		cg_state_init(ws.st.get(), o.localTol * o.localTol, o.maxLocalIter, stream());
		cg_start<double>(ws.r.get(), ws.p.get(), bl, ws.q.get(), N, ws.st.get(), ws.partials.get(), stream());
		const CGState h0 = read_state();
		if (h0.done) return;  // x unchanged
		if (h0.rr > h0.rr0) cg_zero_start(x, ws.r.get(), ws.p.get(), bl, N, ws.st.get(), stream());  // warm start worse than 0
		stage_site(s, false);
		const int* done = &ws.st.get()->done;
		const std::vector<size_t> key{64, s.a, s.n, s.c, s.P, s.Q};
		int iters = 0;
		for (int it = 0; it < o.maxLocalIter; it += kChunk) {
			ws.run(key, [&] {
				for (int c = 0; c < kChunk; ++c) {
					ws_matvec64(s, ws.p.get(), ws.q.get(), done);
					cg_alpha<double>(ws.p.get(), ws.q.get(), N, ws.st.get(), ws.partials.get(), stream());
					cg_update<double>(x, ws.r.get(), ws.p.get(), ws.q.get(), N, ws.st.get(), ws.partials.get(), stream());
					cg_direction<double>(ws.p.get(), ws.r.get(), N, ws.st.get(), stream());
				}
			});
			const CGState h = read_state();
			iters = h.iter;
			if (h.done) break;
		}
		cgIters += iters;
		if (stats) { stats->cgIterations += iters; for (int i = 0; i < iters; ++i) count_matvec(s); }
		XTT_HIP(hipMemcpyAsync(xout, x, N * 8, hipMemcpyDeviceToDevice, stream()));
	}

	// Mixed precision: FP64 iterative refinement around an FP32 / BF16 inner CG. The refinement works on the
	// caller's vectors directly; the low-precision operator pieces are staged only once a correction is needed,
	// and each residual check is one fused kernel and one synchronisation.
	void cg_mixed(SiteOp& s, double* x, const double* bl) {
		const size_t N = s.N();
		reserve(s);
		const double inner = o.innerTol > 0 ? o.innerTol : (o.precision == P_BF16 ? 3e-2 : 1e-5);
		Scratch64& sc = scratch;
		const int* done = &ws.st.get()->done;
		const std::vector<size_t> key{(size_t)o.precision, s.a, s.n, s.c, s.P, s.Q};
		bool staged = false;
		bool fused = o.precision == P_F32 && fused_ok<float>(s);
		if (spec) {  // the first refinement test, on the device
			matvec64(s, x, ws.q.get(), nullptr);
			sc.check(ws.r.get(), bl, ws.q.get(), N, o.localTol * o.localTol, o.maxRefine > 0 ? specFail : nullptr);
			return;
		}
		for (int ref = 0; ref <= o.maxRefine; ++ref) {
			matvec64(s, x, ws.q.get(), nullptr);
			double rb[2];
			sc.residual(ws.r.get(), bl, ws.q.get(), N, rb);
			double rr = rb[0];
			const double tol2 = o.localTol * o.localTol * rb[1];
			if (rr <= tol2 || ref == o.maxRefine) break;
			if (ref == 0 && rr > rb[1]) {  // the warm start is worse than zero: refine from x = 0 (r = b)
				fill_f64(x, 0.0, N, stream());
				XTT_HIP(hipMemcpyAsync(ws.r.get(), bl, N * sizeof(double), hipMemcpyDeviceToDevice, stream()));
				rr = rb[1];
			}
			const double innerEff = inner_tolerance(inner, tol2, rr);
			if (fused) {  // e from 0 for r/‖r‖ and x += ‖r‖·e, all inside one launch
				if (stats) stats->refinements++;
				const double rn = std::sqrt(rr);
				cg_state_init(ws.st.get(), 0.0, 0, stream());
				FusedCG d;
				d.r0 = ws.r.get(); d.r0scale = 1.0 / rn; d.x = x; d.xscale = rn; d.mode = 1;
				d.tol2 = innerEff * innerEff; d.maxIter = o.maxLocalIter;
				if (run_fused<float>(s, d)) continue;
				fused = false;  // launch refused: this and later refinements on the chunked CG
				if (stats) stats->refinements--;
			}
			if (!staged) {
				convert_f64_f32(s.L->data(), ws.L32.get(), s.L->size(), stream());
				convert_f64_f32(s.R->data(), ws.R32.get(), s.R->size(), stream());
				convert_f64_f32(s.am().data(), ws.Am32.get(), s.am().size(), stream());
				staged = true;
			}
			if (stats) stats->refinements++;
			const double rn = std::sqrt(rr);
			scale_f64(ws.r.get(), 1.0 / rn, N, stream());
			convert_f64_f32(ws.r.get(), ws.q32.get(), N, stream());  // right-hand side r/‖r‖
			cg_state_init(ws.st.get(), innerEff * innerEff, o.maxLocalIter, stream());
			// e = 0: r32 = p32 = rhs
			cg_start<float>(ws.r32.get(), ws.p32.get(), ws.q32.get(), ws.z32.get(), N, ws.st.get(), ws.partials.get(), stream());
			XTT_HIP(hipMemcpyAsync(ws.e32.get(), ws.z32.get(), N * 4, hipMemcpyDeviceToDevice, stream()));
			int iters = 0;
			for (int it = 0; it < o.maxLocalIter; it += kChunk) {
				ws.run(key, [&] {
					for (int c = 0; c < kChunk; ++c) {
						local_matvec<float>(ws.L32.get(), ws.Am32.get(), ws.R32.get(), s.a, s.n, s.c, s.P, s.Q, ws.p32.get(), ws.q32.get(),
						                    ws.T1f.get(), ws.T2f.get(), done, o.precision);
						cg_alpha<float>(ws.p32.get(), ws.q32.get(), N, ws.st.get(), ws.partials.get(), stream());
						cg_update<float>(ws.e32.get(), ws.r32.get(), ws.p32.get(), ws.q32.get(), N, ws.st.get(), ws.partials.get(), stream());
						cg_direction<float>(ws.p32.get(), ws.r32.get(), N, ws.st.get(), stream());
					}
				});
				const CGState h = read_state();
				iters = h.iter;
				if (h.done) break;
			}
			cgIters += iters;
			if (stats) { stats->cgIterations += iters; stats->matvecsLow += iters; }
			convert_f32_f64(ws.e32.get(), ws.e64.get(), N, stream());
			axpy_f64(x, rn, ws.e64.get(), N, stream());
		}
	}

	struct Scratch64 {
		DBuf<double> partials{2 * 1024}, out{2};
		DBuf<unsigned> counter{1};
		Scratch64() { zero_u32(counter.get(), 1, stream()); }  // a kernel: the speculative call may be recorded
		// r = b − q; rb = {‖r‖², ‖b‖²}
		void residual(double* r, const double* b, const double* q, size_t n, double* rb) {
			residual_norms_f64(r, b, q, n, out.get(), partials.get(), counter.get(), stream());
			read_back(rb, out.get(), 16);
		}
		// the test ‖r‖² <= tol2·‖b‖² on the device: a failure adds one to *fail (nothing is read back)
		void check(double* r, const double* b, const double* q, size_t n, double tol2, unsigned* fail) {
			residual_norms_f64(r, b, q, n, out.get(), partials.get(), counter.get(), stream(), tol2, fail);
		}
	};
	Scratch64 scratch;  // one per solve: the reduction counter is zeroed once and returned to zero by each use

	void local_solve() {
		// A speculative one-site solve only tests the start residual of x's core: no working copy of it
		if (spec && A && o.sites == 1) {
			SiteOp s;
			DTensor unused;  // only a merged multi-site operator is kept there
			site_op(s, unused);
			// cg_start's test on the device, as cg64 / cg_mixed make it: converged iff ‖b − Ax‖² <= tol²·‖b‖²
			// (or no iterations / refinements allowed); a failure counts into *specFail
			const DTensor bl = rhs_and_matvec(s, x.cores[curr].data(), ws.q.get());
			const bool canIterate = o.precision == P_F64 ? o.maxLocalIter > 0 : o.maxRefine > 0;
			scratch.check(ws.r.get(), bl.data(), ws.q.get(), s.N(), o.localTol * o.localTol, canIterate ? specFail : nullptr);
			if (stats) stats->localSolves++;
			return;
		}
		DTensor xm = merged(x);
		const size_t a = xm.dims[0], c = xm.dims[2];
		DTensor bl = local_rhs();
		if (!A) {
			xm = std::move(bl);
		} else {
			DTensor keepOp;
			SiteOp s;
			site_op(s, keepOp);
			if (o.precision == P_F64) cg64(s, xm.data(), bl.data()); else cg_mixed(s, xm.data(), bl.data());
		}
		if (stats) stats->localSolves++;
		if (o.sites == 1) {
			xm.reshape(x.cores[curr].dims);
			x.set_component(curr, std::move(xm));
			return;
		}
		// sites == 2: SVD split capped at the target rank (als.cpp:276-295)
		xm.reshape({a, x.cores[curr].dims[1], x.cores[curr + 1].dims[1], c});
		DTensor U, Vt;
		std::vector<double> S;
		DBuf<double> sd;
		calculate_svd(U, S, Vt, xm, 2, targetRank[curr], XERUS_EPSILON, &sd);
		const size_t rk = S.size(), cols = Vt.size() / std::max<size_t>(rk, 1), rows = U.size() / std::max<size_t>(rk, 1);
		if (increasing) {
			scale_by_f64(Vt.data(), sd.get(), rk, cols, true, stream());
			x.set_component(curr, std::move(U));
			x.set_component(curr + 1, std::move(Vt));
			x.corePosition = curr + 1;
		} else {
			// U·S: scale column j of U by s_j
			scale_by_f64(U.data(), sd.get(), rows, rk, false, stream());
			x.set_component(curr + 1, std::move(Vt));
			x.set_component(curr, std::move(U));
			x.corePosition = curr;
		}
		x.canonicalized = true;
	}

	// residual_f of als.cpp:240-320: ‖Ax − b‖/‖b‖ (SPD, computed stably as a TT), the energy functional of the
	// non-SPD case (which is its residual), ‖x − b‖ without an operator.
	double residual_f() {
		if (!A) {
			TT nb = tt_copy(b);
			nb.canonicalized = false;
			dt_scale(nb.cores[0], -1.0);
			return tt_frob_norm(tt_add(x, nb));
		}
		if (o.assumeSPD) return tt_residual(*A, x, b);
		return energy_f();
	}

	// PerformanceData::add(residual, x, flags) (performanceData.cpp:50-80): residual with the clock paused
	void perf_add(double residual, size_t flags) {
		PerfSink& ps = perf_sink();
		if (!ps.cb) return;
		clock.pause();
		const dims_t r = x.ranks();
		ps.cb(ps.user, perfIteration++, clock.elapsed(), residual, r.data(), r.size(), flags);
		clock.resume();
	}
	void perf_point(size_t flags, bool residualIsEnergy) {
		if (!perf_sink().cb) return;
		sync();
		clock.pause();
		const double res = residualIsEnergy ? energy : residual_f();
		clock.resume();
		perf_add(res, flags);
	}
	PerfClock clock;
	size_t perfIteration = 0;

	bool end_of_sweep() {  // als.cpp:651-700
		if ((!increasing && curr == first) || (increasing && curr == second - o.sites)) {
			halfSweepCount += 1;
			lastEnergy2 = lastEnergy;
			lastEnergy = energy;
			energy = energy_f();
			if (stats) { stats->halfSweeps = halfSweepCount; stats->energies.push_back(energy); }
			perf_point(increasing ? 1 : 3, o.useResidualForEndCriterion);
			const bool stop = graphMode ? halfSweepCount == plannedHalfSweeps
			                            : (halfSweepCount == o.numHalfSweeps || std::abs(lastEnergy - energy) < o.convergenceEpsilon ||
			                               std::abs(lastEnergy2 - energy) < o.convergenceEpsilon || second - first <= o.sites);
			if (stop) {
				if (canonAtEnd && o.preserveCorePosition) move_core(x, corePosAtEnd, true);
				return true;
			}
			increasing = !increasing;
		} else {
			perf_point(0, false);
		}
		return false;
	}

	// als.cpp:565-605; for sites >= 2 the pushed stack is that of the site leaving the window, after the core has
	// moved into the window (for sites = 1 this is exactly xerus's step).
	void next() {
		if (increasing) {
			move_core(x, curr + 1, true);
			if (A) opR.pop_back();
			rhsR.pop_back();
			push_left(curr);
			curr++;
		} else {
			const size_t leaving = curr + o.sites - 1;
			move_core(x, leaving - 1, true);
			if (A) opL.pop_back();
			rhsL.pop_back();
			push_right(leaving);
			curr--;
		}
	}
};

}  // namespace

// ---- the ALS building blocks on raw device buffers (row-major): stacks L (a, p, a) and R (c, q, c), operator core
// A (p, n, n, q), site vector x (a, n, c) --------------------------------------------------------------------------
namespace {
DTensor wrap_dev(const double* p, std::vector<size_t> dims) {
	DTensor t(std::move(dims));
	XTT_HIP(hipMemcpyAsync(t.data(), p, t.count() * sizeof(double), hipMemcpyDeviceToDevice, stream()));
	return t;
}
void put_dev(double* p, const DTensor& t) {
	XTT_HIP(hipMemcpyAsync(p, t.data(), t.count() * sizeof(double), hipMemcpyDeviceToDevice, stream()));
}
}  // namespace

void als_env_left(const double* L, const double* X, const double* A, size_t a, size_t n, size_t c, size_t p, size_t q, double* out) {
	XTT_REQUIRE(a && n && c && p && q, "env_update: empty dimension");
	put_dev(out, op_left(wrap_dev(L, {a, p, a}), wrap_dev(X, {a, n, c}), wrap_dev(A, {p, n, n, q})));
}
void als_env_right(const double* R, const double* X, const double* A, size_t a, size_t n, size_t c, size_t p, size_t q, double* out) {
	XTT_REQUIRE(a && n && c && p && q, "env_update: empty dimension");
	put_dev(out, op_right(wrap_dev(R, {c, q, c}), wrap_dev(X, {a, n, c}), wrap_dev(A, {p, n, n, q})));
}
void als_local_apply(const double* L, const double* A, const double* R, size_t a, size_t n, size_t c, size_t p, size_t q, const double* x,
                     double* y) {
	XTT_REQUIRE(a && n && c && p && q, "local_apply: empty dimension");
	const DTensor Ld = wrap_dev(L, {a, p, a}), Rd = wrap_dev(R, {c, q, c}), Ad = wrap_dev(A, {p, n, n, q});
	SiteOp s;
	site_from(s, Ld, Ad, Rd);
	DBuf<double> T1(a * p * n * c), T2(a * n * q * c);
	local_matvec<double>(Ld.data(), s.am().data(), Rd.data(), a, n, c, p, q, x, y, T1.get(), T2.get(), nullptr, P_F64);
}
size_t als_local_cg(const double* L, const double* A, const double* R, size_t a, size_t n, size_t c, size_t p, size_t q, const double* b,
                    double* x, double tol, int maxIter) {
	XTT_REQUIRE(a && n && c && p && q, "local_cg: empty dimension");
	XTT_REQUIRE(tol >= 0 && maxIter >= 0, "local_cg: tolerance and iteration limit must be non-negative");
	const DTensor Ld = wrap_dev(L, {a, p, a}), Rd = wrap_dev(R, {c, q, c}), Ad = wrap_dev(A, {p, n, n, q});
	SiteOp s;
	site_from(s, Ld, Ad, Rd);
	const size_t N = s.N();
	DBuf<double> r(N), pv(N), qv(N), T1(a * p * n * c), T2(a * n * q * c), partials(2 * 1024);
	DBuf<CGState> st(1);
	auto mv = [&](const double* in, double* out, const int* skip) {
		local_matvec<double>(Ld.data(), s.am().data(), Rd.data(), a, n, c, p, q, in, out, T1.get(), T2.get(), skip, P_F64);
	};
	mv(x, qv.get(), nullptr);
	cg_state_init(st.get(), tol * tol, maxIter, stream());
	cg_start<double>(r.get(), pv.get(), b, qv.get(), N, st.get(), partials.get(), stream());
	const int* done = &st.get()->done;
	CGState h{};
	read_back(&h, st.get(), sizeof(CGState));
	while (!h.done) {
		for (int i = 0; i < 16; ++i) {  // iterations after convergence are predicated off on the device
			mv(pv.get(), qv.get(), done);
			cg_alpha<double>(pv.get(), qv.get(), N, st.get(), partials.get(), stream());
			cg_update<double>(x, r.get(), pv.get(), qv.get(), N, st.get(), partials.get(), stream());
			cg_direction<double>(pv.get(), r.get(), N, st.get(), stream());
		}
		read_back(&h, st.get(), sizeof(CGState));
	}
	return (size_t)h.iter;
}

double als_solve(const ALSOptions& o, const TT* A, TT& x, const TT& b, ALSStats* stats) {  // als.cpp:708-778
	XTT_REQUIRE(o.sites > 0, "ALSVariant: sites must be > 0");
	XTT_REQUIRE(x.order() > 0, "ALS: x must have degree > 0");
	XTT_REQUIRE(x.dims == b.dims, "ALS: x and b dimensions differ");
	XTT_REQUIRE(!x.op && !b.op, "ALS: x and b must be TTTensors");
	if (A) {
		XTT_REQUIRE(A->op && A->dims.size() == 2 * b.dims.size(), "ALS: operator degree mismatch");
		for (size_t i = 0; i < x.order(); ++i)
			XTT_REQUIRE(A->dims[i] == x.dims[i] && A->dims[i + x.order()] == x.dims[i], "ALS: operator dimension mismatch");
	} else {
		XTT_REQUIRE(o.sites == 1, "approximation dmrg not implemented yet");
	}
	// Speculation: when the last call of this thread on a problem of the same shapes needed no inner CG
	// iteration, this call first runs with every local solve taken as converged at its start residual, which
	// removes the per-site synchronisation. The device counts the sites whose start residual failed the test; if
	// any did, x and the statistics are restored and the call runs again normally. Either way the result is the
	// normal run's. Not with a PerformanceData sink (its points would be reported twice) or for DMRG, and not
	// while XTT_NO_SPECULATION is set (read per call).
	static thread_local std::map<std::vector<size_t>, bool> lastConverged;
	std::vector<size_t> key{(size_t)o.precision, o.sites, o.numHalfSweeps, (size_t)o.assumeSPD};
	for (const TT* t : {A, (const TT*)&x, &b})
		if (t) { for (size_t r : t->ranks()) key.push_back(r); key.push_back(0); }
	key.insert(key.end(), x.dims.begin(), x.dims.end());
	bool timedOut = false;
	size_t lastHalfSweeps = 0;
	bool lastShapeStop = false;
	// graph: record a speculative call on the stand-in y (energies to device slots, plannedHalfSweeps sweeps)
	auto run_on = [&](TT& y, ALSStats* st, bool spec, unsigned* fail, size_t& cgIters, bool noFused, double* energySlots,
	                  size_t plannedHalfSweeps) {
		ALSData data(o, A, y, b, st);
		data.spec = spec;
		data.specFail = fail;
		data.noFused = noFused;
		data.graphMode = energySlots != nullptr;
		data.energySlots = energySlots;
		data.plannedHalfSweeps = plannedHalfSweeps;
		data.clock.begin();
		data.energy = data.energy_f();
		data.perf_point(3, false);
		while (true) {
			data.local_solve();
			if (data.end_of_sweep()) break;
			data.next();
		}
		data.finish_counts();
		cgIters = data.cgIters;
		timedOut = data.fusedTimedOut;
		lastHalfSweeps = data.halfSweepCount;
		lastShapeStop = data.second - data.first <= o.sites;
		return data.energy;
	};
	auto run = [&](bool spec, unsigned* fail, size_t& cgIters, bool noFused) {
		return run_on(x, stats, spec, fail, cgIters, noFused, nullptr, 0);
	};
	auto it = lastConverged.find(key);
	if (A && o.sites == 1 && !perf_sink().cb && it != lastConverged.end() && it->second && !getenv("XTT_NO_SPECULATION")) {
		// From the third speculative call with the same key (shapes, options, the addresses of A's and b's cores, x's
		// core position), the call is a recorded graph (TTGraph). Its energies are checked against the recorded
		// half-sweep count; a failed site or a different count falls back to the normal run, with x untouched.
		ALSRecord* rec = nullptr;
		if (graphs_enabled() && o.assumeSPD && !o.useResidualForEndCriterion && &b != &x) {
			std::vector<size_t> gkey = key;
			for (const TT* t : {A, &b})
				for (const DTensor& c : t->cores) gkey.push_back((size_t)c.data());
			for (double v : {o.convergenceEpsilon, o.localTol, o.innerTol}) { uint64_t u; std::memcpy(&u, &v, 8); gkey.push_back((size_t)u); }
			gkey.insert(gkey.end(), {(size_t)o.maxLocalIter, (size_t)o.maxRefine, (size_t)o.preserveCorePosition, (size_t)x.canonicalized,
			                         x.corePosition});
			auto& calls = workspace().calls;
			if (calls.size() > 32) calls.clear();
			auto& slot = calls[gkey];
			if (!slot) { slot = std::make_unique<ALSRecord>(); slot->g.kind = 0; }
			rec = slot.get();
		}
		// a replay: launch, read the slots once, accept when no site failed and the stop rule agrees
		auto replay = [&](ALSRecord& r, double& energyOut) {
			r.g.launch(x);
			std::vector<double> h(1 + 2 * r.energies);
			read_back(h.data(), r.slots, h.size() * sizeof(double));
			unsigned failedSites;
			std::memcpy(&failedSites, h.data(), sizeof(unsigned));
			if (failedSites != 0) return false;
			std::vector<double> E(r.energies);
			for (size_t e = 0; e < r.energies; ++e) E[e] = std::abs(0.5 * h[2 + 2 * e] - h[1 + 2 * e]);
			double le2 = 1e102, le = 1e101, en = E[0];
			for (size_t hs = 1; hs <= r.halfSweeps; ++hs) {
				le2 = le; le = en; en = E[hs];
				const bool stop = hs == o.numHalfSweeps || std::abs(le - en) < o.convergenceEpsilon || std::abs(le2 - en) < o.convergenceEpsilon ||
				                  r.shapeStop;
				if (stop != (hs == r.halfSweeps)) return false;
			}
			r.g.take(x);
			if (stats) {
				stats->localSolves += r.delta.localSolves;
				stats->matvecs64 += r.delta.matvecs64;
				stats->flopsMatvec += r.delta.flopsMatvec;
				stats->flopsEnv += r.delta.flopsEnv;
				stats->halfSweeps = r.halfSweeps;
				for (size_t hs = 1; hs <= r.halfSweeps; ++hs) stats->energies.push_back(E[hs]);
			}
			energyOut = E[r.halfSweeps];
			return true;
		};
		double e = 0;
		if (rec && rec->g.ready()) {
			if (replay(*rec, e)) return e;
		} else if (rec && rec->halfSweeps > 0 && !rec->failed) {
			ALSStats local;
			size_t cgIters = 0;
			const size_t planned = rec->halfSweeps, nE = planned + 1;
			void* slots = nullptr;
			const bool recorded = rec->g.record(x, [&](TT& y) {
				slots = dmalloc(8 * (1 + 2 * nE));  // the capture arena's
				zero_u32(static_cast<unsigned*>(slots), 1, stream());
				run_on(y, &local, true, static_cast<unsigned*>(slots), cgIters, false, static_cast<double*>(slots) + 1, planned);
			});
			if (recorded && lastHalfSweeps == planned) {
				rec->slots = slots;
				rec->energies = nE;
				rec->shapeStop = lastShapeStop;
				rec->delta = local;
				if (replay(*rec, e)) return e;
			} else {
				if (recorded && getenv("XTT_DEBUG_GRAPHS"))
					fprintf(stderr, "xtt: recorded ALS call ran %zu half-sweeps, planned %zu\n", lastHalfSweeps, planned);
				rec->g.reset();
				rec->failed = true;
			}
		} else {
			TT saved = tt_copy(x);
			const ALSStats savedStats = stats ? *stats : ALSStats{};
			DBuf<unsigned> fail(1);
			XTT_HIP(hipMemsetAsync(fail.get(), 0, sizeof(unsigned), stream()));
			size_t cgIters = 0;
			e = run(true, fail.get(), cgIters, false);
			unsigned failed = 0;
			read_back(&failed, fail.get(), sizeof(unsigned));
			if (failed == 0) {
				if (rec) rec->halfSweeps = lastHalfSweeps;
				return e;
			}
			x = std::move(saved);
			if (stats) *stats = savedStats;
		}
	}
	size_t cgIters = 0;
	// A fused (grid-synchronised) local solve whose grid barrier ran out of time left its site unsolved: the whole call
	// runs again from x's start on the chunked CG, so the result is that of a call without fused solves.
	const bool mayFuse = A && !t_batch_worker && !getenv("XTT_NO_FUSED_CG");
	TT start;
	ALSStats startStats;
	if (mayFuse) { start = tt_copy(x); if (stats) startStats = *stats; }
	double e = run(false, nullptr, cgIters, false);
	if (timedOut) {
		x = std::move(start);
		if (stats) *stats = startStats;
		e = run(false, nullptr, cgIters, true);
	}
	if (lastConverged.size() > 256) lastConverged.clear();
	lastConverged[key] = cgIters == 0;
	return e;
}

// Many independent problems at once: problem p runs `sweeps` × (ALS with options o, then round(eps)) on one of
// `threads` worker threads. Every worker binds a stream of its own (default priority), so its allocations,
// captured graphs and kernels stay on that stream and the latency-bound launches of different problems overlap
// on the device. residuals[p] (optional) = ‖A x − b‖/‖b‖ afterwards.
void als_round_batch(const ALSOptions& o, const std::vector<const TT*>& A, const std::vector<TT*>& x,
                     const std::vector<const TT*>& b, size_t sweeps, double eps, int threads, double* residuals,
                     ALSStats* total) {
	const size_t count = x.size();
	XTT_REQUIRE(A.size() == count && b.size() == count, "batch: A, x and b lists differ in length");
	if (threads <= 0) {  // lockstep: one grouped launch per step for all problems, where the options allow it
		if (lockstep_supported(o, A)) {
			als_round_lockstep(o, A, x, b, sweeps, eps, residuals, total);
			return;
		}
		threads = 16;
	}
	XTT_REQUIRE(threads >= 1 && threads <= 64, "batch: 1..64 worker threads");
	int dev = 0;
	XTT_HIP(hipGetDevice(&dev));
	sync();  // inputs built on the library stream are complete before workers use them
	park_stream_pool(stream());  // … and so the library stream's cached blocks can serve the workers
	std::atomic<size_t> next{0};
	std::mutex mu;
	std::exception_ptr err;
	auto worker = [&]() {
		hipStream_t s = nullptr;
		try {
			XTT_HIP(hipSetDevice(dev));
			XTT_HIP(hipStreamCreateWithFlags(&s, hipStreamDefault));
			bind_thread_stream(s);
			t_batch_worker = true;
			ALSStats st;
			for (size_t p = next++; p < count; p = next++) {
				TT& xp = *x[p];
				const dims_t noCap(xp.order() > 0 ? xp.order() - 1 : 0, SIZE_MAX);
				for (size_t sw = 0; sw < sweeps; ++sw) {
					ALSStats one;  // als_solve sets halfSweeps (the call's count); the others it adds to
					als_solve(o, A[p], xp, *b[p], &one);
					st.halfSweeps += one.halfSweeps; st.localSolves += one.localSolves;
					st.cgIterations += one.cgIterations; st.refinements += one.refinements;
					st.matvecs64 += one.matvecs64; st.matvecsLow += one.matvecsLow;
					st.flopsMatvec += one.flopsMatvec; st.flopsEnv += one.flopsEnv;
					round(xp, noCap, eps);
				}
				if (residuals) residuals[p] = A[p] ? tt_residual(*A[p], xp, *b[p]) : 0.0;
			}
			sync();
			if (total) {
				std::lock_guard<std::mutex> lk(mu);
				total->halfSweeps += st.halfSweeps; total->localSolves += st.localSolves;
				total->cgIterations += st.cgIterations; total->refinements += st.refinements;
				total->matvecs64 += st.matvecs64; total->matvecsLow += st.matvecsLow;
				total->flopsMatvec += st.flopsMatvec; total->flopsEnv += st.flopsEnv;
			}
		} catch (...) {
			std::lock_guard<std::mutex> lk(mu);
			if (!err) err = std::current_exception();
			next = count;  // the other workers stop after their current problem
		}
		if (s) {
			(void)hipStreamSynchronize(s);
			bind_thread_stream(nullptr);
			t_batch_worker = false;
			release_thread_graphs();  // graph workspaces captured on s, and their buffers, go before s does
			park_stream_pool(s);
			(void)hipStreamDestroy(s);
		}
	};
	std::vector<std::thread> pool;
	for (int t = 0; t < threads; ++t) pool.emplace_back(worker);
	for (auto& t : pool) t.join();
	if (err) std::rethrow_exception(err);
}

}  // namespace xtt
