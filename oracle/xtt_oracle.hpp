// TEST INFRASTRUCTURE ONLY — the oracle.
//
// FP64 CPU restatement of xerus 3.0.1's Tensor-Train hot path (TTTensor/TTOperator, move_core, round,
// ALS/ALS_SPD/DMRG/DMRG_SPD, ADF over SinglePointMeasurementSet). It is the parity checker for the HIP
// library and the timed CPU baseline ("port") in bench.py. Nothing in the product links or calls it.
//
// Every routine cites the reference file:line whose behaviour it restates. Layout is xerus's: row-major,
// last index fastest (include/xerus/tensor.h:84-88); TT component k has dims (r_{k-1}, n_k, r_k), a TT
// operator component (r_{k-1}, n_k, m_k, r_k) (include/xerus/ttNetwork.h:316-321).
#pragma once
#include <cstddef>
#include <cstdint>
#include <random>
#include <vector>

namespace orc {

using dims_t = std::vector<size_t>;

struct Tensor {
	dims_t dims;
	std::vector<double> v;
	Tensor() = default;
	explicit Tensor(dims_t d) : dims(std::move(d)), v(prod(), 0.0) {}
	size_t prod() const { size_t p = 1; for (size_t x : dims) p *= x; return p; }
	size_t size() const { return v.size(); }
	size_t degree() const { return dims.size(); }
	double* data() { return v.data(); }
	const double* data() const { return v.data(); }
};

// Shared random state, mirroring xerus::misc::randomEngine / defaultNormalDistribution
// (src/xerus/misc/random.cpp:29-30): one mt19937_64 and one stateful normal_distribution<double>.
std::mt19937_64& rng();
std::normal_distribution<double>& normal();
void seed(uint64_t s);

// ---- dense kernels (blasLapackWrapper.cpp restatements) ----
// C(m×n) = alpha·op(A)·op(B) with op(A) m×k, beta = 0 (blasLapackWrapper.cpp:149-195).
void gemm(bool ta, bool tb, size_t m, size_t n, size_t k, double alpha, const double* A, const double* B, double* C);
// Householder QR, Q m×min(m,n), R min(m,n)×n (blasLapackWrapper.cpp:374-431, dgeqrf+dorgqr).
void qr(const double* A, size_t m, size_t n, std::vector<double>& Q, std::vector<double>& R);
// RQ, R m×min(m,n) upper triangular, Q min(m,n)×n (blasLapackWrapper.cpp:442-498, dgerqf+dorgrq).
void rq(const double* A, size_t m, size_t n, std::vector<double>& R, std::vector<double>& Q);
// Column-pivoted QR with xerus's rank rule (blasLapackWrapper.cpp:235-305). Returns rank.
size_t qc(const double* A, size_t m, size_t n, std::vector<double>& Q, std::vector<double>& C);
// CQ via QRP of the transpose (blasLapackWrapper.cpp:316-371). Returns rank.
size_t cq(const double* A, size_t m, size_t n, std::vector<double>& C, std::vector<double>& Q);
// Thin SVD, singular values descending (blasLapackWrapper.cpp:201-230, dgesdd 'S').
void svd(const double* A, size_t m, size_t n, std::vector<double>& U, std::vector<double>& S, std::vector<double>& Vt);
// Dense solve with xerus's dispatch (blasLapackWrapper.cpp:501-651). Returns 0 LU, 1 Cholesky, 2 LDL^T(LU fallback).
int solve(const double* A, size_t n, const double* b, double* x);

// ---- Tensor level (tensor.cpp, indexedTensor_tensor_evaluate.cpp) ----
Tensor contract(const Tensor& lhs, bool lhsTrans, const Tensor& rhs, bool rhsTrans, size_t numModes); // tensor.cpp:1252-1352
Tensor reshuffle(const Tensor& t, const std::vector<size_t>& shuffle);                               // evaluate.cpp:55-137
void calculate_qr(Tensor& Q, Tensor& R, const Tensor& in, size_t split);                              // tensor.cpp:1492
void calculate_rq(Tensor& R, Tensor& Q, const Tensor& in, size_t split);                              // tensor.cpp:1510
void calculate_qc(Tensor& Q, Tensor& C, const Tensor& in, size_t split);                              // tensor.cpp:1528
void calculate_cq(Tensor& C, Tensor& Q, const Tensor& in, size_t split);                              // tensor.cpp:1548
void calculate_svd(Tensor& U, Tensor& S, Tensor& Vt, const Tensor& in, size_t split, size_t maxRank, double eps); // tensor.cpp:1424-1489
double frob_norm(const Tensor& t);

// ---- TTNetwork (ttNetwork.cpp) ----
struct TT {
	bool op = false;
	dims_t dims;                 // d (tensor) or 2d (operator: rows then cols), ttNetwork.cpp:70-80
	std::vector<Tensor> cores;
	bool canonicalized = true;
	size_t corePosition = 0;
	size_t order() const { return cores.size(); }
	std::vector<size_t> ranks() const;
	size_t rank(size_t i) const { return cores[i].dims.back(); }
	void set_component(size_t k, Tensor t);                                      // ttNetwork.cpp:475-492
	size_t ext(size_t k) const { return op ? dims[k] * dims[order() + k] : dims[k]; }
};
dims_t reduce_to_maximal_ranks(dims_t ranks, const dims_t& dims, bool op);       // ttNetwork.cpp:369-402
TT tt_random(const dims_t& dims, const dims_t& ranks, bool op);                  // ttNetwork.h:128-155
TT tt_zero(const dims_t& dims, bool op);                                          // ttNetwork.cpp:57-108
void transfer_core(TT& x, size_t from, size_t to, bool allowRankReduction);      // tensorNetwork.cpp:821-909 (component ids)
void move_core(TT& x, size_t pos, bool keepRank);                                // ttNetwork.cpp:581-628
void round_edge(TT& x, size_t from, size_t to, size_t maxRank, double eps);      // tensorNetwork.cpp:678-818
void round(TT& x, const dims_t& maxRanks, double eps);                           // ttNetwork.cpp:643-665
double tt_frob_norm(const TT& x);                                                // ttNetwork.cpp:781-789
double tt_dot(const TT& a, const TT& b);
TT tt_add(const TT& a, const TT& b);                                             // ttNetwork.cpp:796-847
void tt_scale(TT& a, double alpha);
TT tt_apply(const TT& A, const TT& x);                                           // ttNetwork.cpp:888-967 + ttStack.cpp:197-310
Tensor tt_full(const TT& x);
double tt_residual(const TT& A, const TT& x, const TT& b);                       // als.cpp:483-486 (stable form)
TT tt_from_tensor(const Tensor& full, bool op, double eps, const dims_t& maxRanks); // ttNetwork.cpp:112-160
TT tt_ones(const dims_t& dims, bool op);                                         // ttNetwork.cpp:170-191
TT tt_identity(const dims_t& dims);                                              // ttNetwork.cpp:194-221
TT tt_kronecker(const dims_t& dims, bool op);                                    // ttNetwork.cpp:224-254
TT tt_dirac(const dims_t& dims, const dims_t& position, bool op);               // ttNetwork.cpp:257-283
void tt_transpose(TT& x);                                                        // ttNetwork.h:443-448

// ---- ALS (als.cpp) ----
struct ALSOptions {
	unsigned sites = 1;
	size_t numHalfSweeps = 0;
	double convergenceEpsilon = 1e-6;
	bool useResidualForEndCriterion = false;
	bool preserveCorePosition = true;
	bool assumeSPD = true;
	int localSolver = 0;         // 0: dense (xerus lapack_solver), 1: matrix-free CG (SPD only)
	double cgTol = 1e-13;        // relative tolerance of the matrix-free CG
	size_t cgMaxIter = 4000;
};
struct ALSStats { size_t halfSweeps = 0; size_t localSolves = 0; size_t cgIterations = 0; std::vector<double> energies; };
double als_solve(const ALSOptions& o, const TT* A, TT& x, const TT& b, ALSStats* stats);   // als.cpp:708-778

// ---- measurements + ADF (measurments.cpp, adf.cpp) ----
struct SinglePointMeasurementSet {
	size_t d = 0;
	std::vector<uint32_t> pos;   // P×d, row-major
	std::vector<double> values;
	size_t size() const { return values.size(); }
};
SinglePointMeasurementSet meas_random(size_t P, const dims_t& dims);            // measurments.cpp:222-248
void meas_measure(SinglePointMeasurementSet& m, const TT& x);                    // measurments.cpp:125-153
struct ADFStacks {
	// forwardUpdates[k] / backwardUpdates[k]: representative measurement ids (adf.cpp:102-191)
	std::vector<std::vector<uint32_t>> fwd, bwd;
	size_t numUnique = 0;
};
ADFStacks adf_stack_plan(const SinglePointMeasurementSet& m);
// Projected gradient Δ_k (r_{k-1}, n_k, r_k) at x for every core k (adf.cpp:314-396), i.e. -∂/∂G_k ½Σ(x_i-b_i)².
std::vector<Tensor> adf_full_gradient(const TT& x, const SinglePointMeasurementSet& m, double* residualNorm);
struct ADFOptions { size_t maxIterations = 0; double targetResidualNorm = 1e-8; double minimalResidualNormDecrease = 0.999; };
double adf_solve(const ADFOptions& o, TT& x, const SinglePointMeasurementSet& m, const dims_t& maxRanks, size_t* iterations); // adf.cpp:489-604
// One step of the sweep at component k on x as it is (adf.cpp:360-478): gradient, slice norms, update_x.
void adf_core_step(TT& x, const SinglePointMeasurementSet& m, size_t k);

} // namespace orc
