// Device-resident TT tensors / operators and the Tensor-level operations xerus's TT path is built from.
// All FP64; cores stay in HBM, only integer rank decisions and scalar norms cross to the host.
#pragma once
#include <cmath>
#include <functional>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "kernels.hpp"
#include "xtt_common.hpp"

namespace xtt {

using dims_t = std::vector<size_t>;

// ---- Tensor level (tensor.cpp / blasLapackWrapper.cpp equivalents) ------------------------------------
DTensor dt_upload(const dims_t& dims, const double* host);
void dt_download(const DTensor& t, double* host);
DTensor dt_copy(const DTensor& t);
DTensor dt_zeros(const dims_t& dims);
// xerus::contract (tensor.cpp:1252-1352): lhsTrans ⇒ lhs's first `modes` contracted, rhsTrans ⇒ rhs's last.
DTensor contract(const DTensor& lhs, bool lhsTrans, const DTensor& rhs, bool rhsTrans, size_t modes);
DTensor reshuffle(const DTensor& t, const std::vector<size_t>& shuffle);  // shuffle[i] = new position of mode i
DTensor transpose2(const DTensor& t, size_t split);                       // (rows…, cols…) → (cols…, rows…)
void calculate_qr(DTensor& Q, DTensor& R, const DTensor& in, size_t split);
void calculate_rq(DTensor& R, DTensor& Q, const DTensor& in, size_t split);
void calculate_qc(DTensor& Q, DTensor& C, const DTensor& in, size_t split);
// blasWrapper::solve with xerus's dispatch (row-major A m×n, B m×nrhs, X n×nrhs, device); returns the method:
// 0 LU, 1 Cholesky, 2 symmetric indefinite (solved by LU), 3 minimum-norm least squares
int dense_solve(const double* A, size_t m, size_t n, const double* B, size_t nrhs, double* X, hipStream_t s);
// blasWrapper::solve_least_squares: minimum-norm least squares (dgelsd's cut-off) for any m, n
int least_squares_solve(const double* A, size_t m, size_t n, const double* B, size_t nrhs, double* X, hipStream_t s);
void calculate_cq(DTensor& C, DTensor& Q, const DTensor& in, size_t split);
// Thin SVD with xerus's truncation (tensor.cpp:1463-1474); S returned as host singular values (kept ones).
// sDev, when given, receives the kept singular values on the device as well. jacobiOnTranspose: for a
// square matrix whose rows are nearly orthogonal (Rᵀ of a pivoted QR, as in round's precompression), the
// one-sided Jacobi runs on the transpose, where it needs fewer sweeps (Drmač–Veselić preconditioning).
// inTransposed, when given, is the transpose of `in` (n×m row-major) for that case, and in's data is not read.
void calculate_svd(DTensor& U, std::vector<double>& S, DTensor& Vt, const DTensor& in, size_t split, size_t maxRank, double eps,
                   DBuf<double>* sDev = nullptr, bool jacobiOnTranspose = false, const double* inTransposed = nullptr);
// R factor only (k×n) of a Householder QR of the m×n matrix view (used by the stable residual).
DTensor qr_r_only(const DTensor& in, size_t split);
double dt_norm(const DTensor& t);
double dt_dot(const DTensor& a, const DTensor& b);
// (a·b, c·d) with one read-back
std::pair<double, double> dt_dot2(const DTensor& a, const DTensor& b, const DTensor& c, const DTensor& d);
void dt_scale(DTensor& t, double s);

// ---- TTNetwork ---------------------------------------------------------------------------------------
struct TT {
	bool op = false;
	dims_t dims;  // d (tensor) or 2d (operator: row dims then column dims)
	std::vector<DTensor> cores;
	bool canonicalized = true;
	size_t corePosition = 0;
	size_t order() const { return cores.size(); }
	size_t ext(size_t k) const { return op ? dims[k] * dims[order() + k] : dims[k]; }
	size_t rank(size_t i) const { return cores[i].dims.back(); }
	dims_t ranks() const;
	void set_component(size_t k, DTensor t);  // ttNetwork.cpp:475-492
};
// ttNetwork.cpp:290-340: ranks chain (first and last 1, neighbours equal), degrees and external dimensions
void require_correct_format(const TT& x);
TT tt_zero(const dims_t& dims, bool op);
TT tt_copy(const TT& x);
dims_t reduce_to_maximal_ranks(dims_t ranks, const dims_t& dims, bool op);
TT tt_random(const dims_t& dims, const dims_t& ranks, bool op);  // host RNG draws as xerus, then move_core(0)
void transfer_core(TT& x, size_t from, size_t to, bool allowRankReduction);
void move_core(TT& x, size_t pos, bool keepRank);
// toIsLeftOrthogonal: core `to` has orthonormal columns (true inside round), so its QC is skipped (it would
// return the core itself up to a signed permutation and full rank).
void round_edge(TT& x, size_t from, size_t to, size_t maxRank, double eps, bool toIsLeftOrthogonal = false);
void round(TT& x, const dims_t& maxRanks, double eps);
double tt_frob_norm(const TT& x);
double tt_dot(const TT& a, const TT& b);
TT tt_add(const TT& a, const TT& b);
void tt_scale(TT& x, double s);
TT tt_apply(const TT& A, const TT& x);   // rank r_A·r_x, uncanonicalized
double tt_residual(const TT& A, const TT& x, const TT& b);  // ‖Ax − b‖/‖b‖, stable (QR sweep of the TT of Ax − b)
// TTNetwork(const Tensor&, eps, maxRanks) from a host tensor (an operator's as (n…, m…)); empty maxRanks: no cap
TT tt_from_tensor(const dims_t& fullDims, const double* host, bool op, double eps, const dims_t& maxRanks);
TT tt_ones(const dims_t& dims, bool op);
TT tt_identity(const dims_t& dims);
TT tt_kronecker(const dims_t& dims, bool op);
TT tt_dirac(const dims_t& dims, const dims_t& position, bool op);
void tt_transpose(TT& x);

// One call of a TT algorithm recorded as a hipGraph, for calls whose kernels and shapes repeat (the speculative
// steady ALS call, the ADF sweep): the launches of a call are replayed by one graph launch instead of being issued one by
// one. record() runs `body` under stream capture on a stand-in TT whose cores are copies of x's cores in the
// record's capture arena; every allocation the body makes comes from that arena, so the buffers the graph touches
// keep their addresses and are never handed to anyone else. x is not modified. launch() copies x's cores into the
// stand-in's input buffers and launches the graph; take() gives x fresh copies of the output cores. The body must not
// synchronise or read device data back (record() then fails and returns false).
struct TTGraph {
	std::unique_ptr<Arena> arena;
	hipGraphExec_t exec = nullptr;
	std::vector<double*> in, out;
	std::vector<dims_t> inDims, outDims;
	bool canonicalized = true;
	size_t corePosition = 0;
	int kind = 0;  // 0 ALS call, 1 ADF sweep: the counter take() adds to (graph_replays)
	std::string why;  // why the last record() failed (XTT_DEBUG_GRAPHS=1 prints it)
	TTGraph() = default;
	TTGraph(const TTGraph&) = delete;
	TTGraph& operator=(const TTGraph&) = delete;
	~TTGraph() { reset(); }
	void reset();
	bool ready() const { return exec != nullptr; }
	bool record(const TT& x, const std::function<void(TT&)>& body);
	void launch(const TT& x);
	void take(TT& x) const;
};
// Recorded calls are used unless XTT_NO_GRAPHS is set or the launch profiler is on (read per call).
bool graphs_enabled();
// Accepted replays since the process started, per kind (all threads).
void graph_replays(size_t out[2]);
void count_replay(int kind);

// xerus datafiles (misc::save_to_file / load_from_file of a TTNetwork): format 0 = BINARY, 1 = TSV.
void tt_save(const TT& x, const std::string& filename, int format);
TT tt_load(const std::string& filename);  // TTTensor or TTOperator, as the header says

// Shared RNG mirroring xerus::misc::randomEngine / defaultNormalDistribution (misc/random.cpp:29-30).
void rng_seed(uint64_t s);
double rng_normal();
size_t rng_uniform(size_t hi);  // uniform_int_distribution<size_t>(0, hi)

// ---- ALS --------------------------------------------------------------------------------------------
enum Precision { P_F64 = 0, P_F32 = 1, P_BF16 = 2 };
struct ALSOptions {
	unsigned sites = 1;
	size_t numHalfSweeps = 0;
	double convergenceEpsilon = 1e-6;
	bool useResidualForEndCriterion = false;
	bool preserveCorePosition = true;
	bool assumeSPD = true;
	int precision = P_F64;      // arithmetic of the local matvecs
	double localTol = 1e-13;    // relative local residual target of the local solve
	int maxLocalIter = 2000;    // CG iterations per local solve (per refinement step in mixed precision)
	double innerTol = 0;        // mixed precision: inner relative tolerance (0: 1e-5 FP32, 3e-2 BF16)
	int maxRefine = 30;         // mixed precision: FP64 refinement steps
};
struct ALSStats {
	size_t halfSweeps = 0, localSolves = 0, cgIterations = 0, refinements = 0, matvecs64 = 0, matvecsLow = 0;
	double flopsMatvec = 0, flopsEnv = 0;  // algorithmic flop counts
	std::vector<double> energies;
};
double als_solve(const ALSOptions& o, const TT* A, TT& x, const TT& b, ALSStats* stats);
// ALS building blocks on device buffers: stacks L (a,p,a), R (c,q,c), operator core A (p,n,n,q), vectors (a,n,c)
void als_env_left(const double* L, const double* X, const double* A, size_t a, size_t n, size_t c, size_t p, size_t q, double* out);
void als_env_right(const double* R, const double* X, const double* A, size_t a, size_t n, size_t c, size_t p, size_t q, double* out);
void als_local_apply(const double* L, const double* A, const double* R, size_t a, size_t n, size_t c, size_t p, size_t q, const double* x,
                     double* y);
size_t als_local_cg(const double* L, const double* A, const double* R, size_t a, size_t n, size_t c, size_t p, size_t q, const double* b,
                    double* x, double tol, int maxIter);
// problem p: `sweeps` × (als_solve, round(eps)) on one of `threads` worker threads with streams of their own
void als_round_batch(const ALSOptions& o, const std::vector<const TT*>& A, const std::vector<TT*>& x,
                     const std::vector<const TT*>& b, size_t sweeps, double eps, int threads, double* residuals,
                     ALSStats* total);
// The same in lockstep (threads <= 0 in als_round_batch): every step of every problem is one grouped launch for all
// problems. Supported for one-site ALS_SPD with an operator and FP64 or FP32 (FP64-refined) local solves.
bool lockstep_supported(const ALSOptions& o, const std::vector<const TT*>& A);
// Relative tolerance of the low-precision inner solve of one refinement step: no tighter than needed to bring the
// FP64 residual (‖r‖² = rr) under the outer bound (tol2 = tol²·‖b‖²) with a factor two to spare, no looser than one
// half, and never tighter than the configured inner tolerance (the floor). The outer test stays the FP64 one.
inline double inner_tolerance(double inner, double tol2, double rr) {
	return std::max(inner, std::min(0.5, 0.5 * std::sqrt(tol2 / rr)));
}
void als_round_lockstep(const ALSOptions& o, const std::vector<const TT*>& A, const std::vector<TT*>& x,
                        const std::vector<const TT*>& b, size_t sweeps, double eps, double* residuals, ALSStats* total);

}  // namespace xtt
