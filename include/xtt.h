/*
 * libxtt — MI355X (gfx950) Tensor-Train engine behind xerus's TT API. C ABI: plain pointers, sizes and
 * opaque handles; every function returns 0 (XT_OK) or an XT_ERR_* code, with a message in xt_last_error().
 *
 * Layout conventions are xerus's: row-major, last index fastest (include/xerus/tensor.h:84-88); a TT
 * component k has dims (r_{k-1}, n_k, r_k), an operator component (r_{k-1}, n_k, m_k, r_k)
 * (include/xerus/ttNetwork.h:316-321); operator dims are [n_0..n_{d-1}, m_0..m_{d-1}] (ttNetwork.cpp:70-80).
 * All device work is ordered on the library's stream of the current device unless a `stream` argument is
 * given; entry points taking a stream return only after that stream has finished the call's work.
 */
#ifndef XTT_H
#define XTT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
	XT_OK = 0,
	XT_ERR_INVALID = 1,     /* bad argument / shape: xerus REQUIRE → generic_error (misc/check.h:60-65) */
	XT_ERR_HIP = 2,         /* HIP runtime failure */
	XT_ERR_NUMERIC = 3,     /* numerical breakdown */
	XT_ERR_NOMEM = 4,       /* device allocation failed */
	XT_ERR_UNSUPPORTED = 5  /* shape outside what the kernels handle (e.g. ADF rank > 16) */
};

enum { XT_F64 = 0, XT_F32 = 1, XT_BF16 = 2 };

typedef struct xt_tt_s* xt_tt;      /* TTTensor / TTOperator resident on the device */
typedef struct xt_meas_s* xt_meas;  /* SinglePointMeasurementSet */
typedef struct xt_r1meas_s* xt_r1meas;  /* RankOneMeasurementSet */

/* ---------------------------------------------------------------- runtime */
const char* xt_last_error(void);
const char* xt_version(void);
int xt_synchronize(void);
/* plain device buffers for callers without their own allocator (ordered on the library stream) */
int xt_malloc(size_t bytes, void** device_ptr);
int xt_free(void* device_ptr);
int xt_memcpy_htod(void* device_dst, const void* host_src, size_t bytes);
int xt_memcpy_dtoh(void* host_dst, const void* device_src, size_t bytes);
int xt_release_cache(void);
/* device bytes held by live library buffers and by the allocator's cache of freed blocks (either nullable) */
int xt_cache_bytes(size_t* in_use, size_t* cached);
/* Diagnostics, no xerus counterpart: replays of recorded calls accepted since the process started (all threads): the
 * speculative steady ALS call (from the third speculative call with the same shapes, options and operands) and the
 * ADF sweep (from the second sweep of a rank phase) run as one hipGraph launch each. XTT_NO_GRAPHS=1 turns recording
 * off. */
int xt_graph_replays(size_t* als, size_t* adf);
/* Launch profiler (no counterpart in xerus; replaces its XERUS_PA_START/END timers, misc/performanceAnalysis.h:33-56):
 * while enabled, GEMM, Householder, Jacobi and ADF launches get a HIP event pair handed to the dispatch on their
 * own stream. xt_profile_read synchronises the device, returns per-kernel (tag, launches, total ms, algorithmic
 * flops, algorithmic bytes) and clears the records. */
int xt_profile_enable(int on);
int xt_profile_read(size_t max_records, char* tags, size_t tag_len, size_t* launches, double* ms, double* flops,
                    double* bytes, size_t* count);
/* misc::randomEngine.seed(s) (src/xerus/misc/random.cpp:29-30); also resets the shared normal distribution */
int xt_seed(uint64_t seed);
int xt_normal_draws(size_t n, double* out);                 /* defaultNormalDistribution(randomEngine) */
int xt_uniform_draws(size_t n, size_t hi, size_t* out);     /* uniform_int_distribution<size_t>(0, hi) */

/* ------------------------------------------- boundary 1: blasWrapper and Tensor-level ops (device FP64)
 * Replaces xerus::blasWrapper (include/xerus/blasLapackWrapper.h:37-146) under xerus::contract /
 * calculate_* (include/xerus/tensor.h:948-1016). Pointers are device pointers. */

/* C = alpha·op(A)·op(B) (β = 0): blasWrapper::matrix_matrix_product, blasLapackWrapper.h:61-85 */
int xt_gemm(int transA, int transB, size_t m, size_t n, size_t k, double alpha, const double* A, const double* B, double* C,
            void* stream);
/* the same with FP32 operands (FP32 MFMA) and with FP32 storage / BF16 products / FP32 accumulation */
int xt_gemm_f32(int transA, int transB, size_t m, size_t n, size_t k, float alpha, const float* A, const float* B, float* C,
                void* stream);
int xt_gemm_bf16(int transA, int transB, size_t m, size_t n, size_t k, float alpha, const float* A, const float* B, float* C,
                 void* stream);
/* Tile autotuner of the GEMMs above and of every contraction inside libxtt (no xerus counterpart: xerus hands its
 * products to a BLAS). A choice per shape (m, n, k, batch, transA, transB; precision XT_F64 or XT_F32) is the tile
 * edge (32 or 64) and, for 32-tiles of a single product, the number of split-K slices (0: none); a shape with a
 * choice no longer uses the built-in heuristic. xt_gemm_tune times the heuristic and each candidate over `iters`
 * launches with HIP events, installs the fastest and reports both (tile, slices, µs per launch; outputs nullable).
 * xt_gemm_tune_set installs a choice (a saved table), xt_gemm_tune_clear drops all. Process-wide. */
int xt_gemm_tune(int precision, size_t m, size_t n, size_t k, size_t batch, int transA, int transB, int iters, int* tile, int* slices,
                 double* best_us, int* heuristic_tile, int* heuristic_slices, double* heuristic_us);
int xt_gemm_tune_set(int precision, size_t m, size_t n, size_t k, size_t batch, int transA, int transB, int tile, int slices);
int xt_gemm_tune_clear(void);
/* installs the choices of a table written by xerus_amd.tune_gemm (xerus_amd/tables/gemm_gfx950.tsv: tab-separated
 * precision m n k batch transA transB launches tile slices µs heuristic_tile heuristic_slices heuristic_µs, '#'
 * comments); rows whose choice is the heuristic's are skipped. installed (nullable): the number installed. */
int xt_gemm_tune_load(const char* path, size_t* installed);
/* Q (m×min(m,n)), R (min×n): blasWrapper::qr, blasLapackWrapper.h:109-112 */
int xt_qr(const double* A, size_t m, size_t n, double* Q, double* R);
/* R (m×min), Q (min×n): blasWrapper::rq, blasLapackWrapper.h:118-121 */
int xt_rq(const double* A, size_t m, size_t n, double* R, double* Q);
/* rank-revealing QR; Q packed m×rank, C packed rank×n (buffers sized for min(m,n)): blasWrapper::qc,
 * blasLapackWrapper.h:97-100, rank rule of blasLapackWrapper.cpp:268-272 */
int xt_qc(const double* A, size_t m, size_t n, double* Q, double* C, size_t* rank);
/* C packed m×rank, Q packed rank×n: blasWrapper::cq, blasLapackWrapper.h:103-106 */
int xt_cq(const double* A, size_t m, size_t n, double* C, double* Q, size_t* rank);
/* blasWrapper::solve (blasLapackWrapper.cpp:501-651, xerus::solve tensor.cpp:1654-1704): X (n×nrhs) solves
 * A·X = B for A m×n, B m×nrhs (device, row-major) with xerus's dispatch: m ≠ n → minimum-norm least squares
 * (dgelsd, rcond = 8ε); A not symmetric by xerus's test → LU with partial pivoting; symmetric with a definite
 * diagonal → Cholesky, falling back on failure; else symmetric indefinite (xerus: LDLᵀ; here LU). *method
 * (nullable) = 0 LU, 1 Cholesky, 2 symmetric indefinite, 3 least squares. A singular system is an error. */
int xt_solve(const double* A, size_t m, size_t n, const double* B, size_t nrhs, double* X, int* method);
/* thin SVD, S descending (device, min(m,n)): blasWrapper::svd, blasLapackWrapper.h:88-91 */
int xt_svd(const double* A, size_t m, size_t n, double* U, double* S, double* Vt);
/* xerus::reshuffle (indexedTensor_tensor_evaluate.cpp:55-137): shuffle[i] = new position of mode i */
int xt_reshuffle(size_t degree, const size_t* dims, const size_t* shuffle, const double* in, double* out);
/* blasWrapper::dot_product (blasLapackWrapper.h:55-58); result to host */
int xt_dot(const double* x, const double* y, size_t n, double* result);
/* blasWrapper::one_norm Σ|x_i| and two_norm ‖x‖₂ (blasLapackWrapper.h:42-46; dasum / dnrm2, the latter scaled by
 * max|x_i| so the squares neither overflow nor underflow); results to host */
int xt_one_norm(const double* x, size_t n, double* result);
int xt_two_norm(const double* x, size_t n, double* result);
/* blasWrapper::matrix_vector_product (blasLapackWrapper.h:52-53): x (m) = alpha·op(A)·y with A row-major m×n
 * (transposed = 0) or n×m (transposed = 1, x = alpha·Aᵀ·y), y of length n */
int xt_matrix_vector_product(double* x, size_t m, double alpha, const double* A, size_t n, int transposed, const double* y,
                             void* stream);
/* blasWrapper::dyadic_vector_product (blasLapackWrapper.h:55-56): A (m×n, row-major) = alpha·x·yᵀ (overwritten) */
int xt_dyadic_vector_product(double* A, size_t m, size_t n, double alpha, const double* x, const double* y, void* stream);
/* blasWrapper::solve_least_squares (blasLapackWrapper.h:136-137, .cpp:654-663): X (n×p) = the minimum-norm solution
 * of min ‖A·X − B‖ (A m×n, B m×p) with dgelsd's cut-off rcond = 8·DBL_EPSILON, for square A as well */
int xt_solve_least_squares(double* X, const double* A, size_t m, size_t n, const double* B, size_t p);

/* ------------------------------------------------------ boundary 2: TTTensor / TTOperator (ttNetwork.h) */
/* TTNetwork(dims): rank-1 zero TT, canonical at 0 (ttNetwork.cpp:57-108) */
int xt_tt_new(int is_operator, size_t ndims, const size_t* dims, xt_tt* out);
/* TTNetwork::random(dims, ranks) with the shared engine (ttNetwork.h:128-155) */
int xt_tt_random(int is_operator, size_t ndims, const size_t* dims, const size_t* ranks, xt_tt* out);
/* TTNetwork(const Tensor&, eps, maxRanks) (ttNetwork.cpp:112-160): TT-SVD of a dense host tensor (row-major; an
 * operator's dims are [n_0..n_{d-1}, m_0..m_{d-1}]); max_ranks NULL = no cap (xerus's default eps is 8·DBL_EPSILON) */
int xt_tt_from_tensor(int is_operator, size_t ndims, const size_t* dims, const double* host_data, double eps,
                      const size_t* max_ranks, xt_tt* out);
int xt_tt_ones(int is_operator, size_t ndims, const size_t* dims, xt_tt* out);       /* TTNetwork::ones, ttNetwork.cpp:170-191 */
int xt_tt_identity(size_t ndims, const size_t* dims, xt_tt* out);                    /* TTOperator::identity, :194-221 */
int xt_tt_kronecker(int is_operator, size_t ndims, const size_t* dims, xt_tt* out);  /* TTNetwork::kronecker, :224-254 */
/* TTNetwork::dirac (:257-283): position has one entry per dimension */
int xt_tt_dirac(int is_operator, size_t ndims, const size_t* dims, const size_t* position, xt_tt* out);
int xt_tt_transpose(xt_tt t);                                 /* TTOperator::transpose, ttNetwork.h:443-448 */
int xt_tt_copy(xt_tt src, xt_tt* out);
int xt_tt_free(xt_tt t);
int xt_tt_order(xt_tt t, size_t* order);
int xt_tt_is_operator(xt_tt t, int* is_operator);
int xt_tt_dims(xt_tt t, size_t* dims);                       /* order (or 2·order) entries */
int xt_tt_ranks(xt_tt t, size_t* ranks);                     /* TTNetwork::ranks, ttNetwork.cpp:716-731 */
int xt_tt_component_dims(xt_tt t, size_t k, size_t* dims);   /* 3 (or 4) entries */
int xt_tt_get_component(xt_tt t, size_t k, double* host_out);  /* TTNetwork::get_component, ttNetwork.cpp:466-470 */
int xt_tt_set_component(xt_tt t, size_t k, size_t ndim, const size_t* dims, const double* host_data); /* :475-492 */
int xt_tt_component_ptr(xt_tt t, size_t k, double** device_ptr);  /* valid until the TT changes shape */
int xt_tt_canonicalization(xt_tt t, int* canonicalized, size_t* core_position);
int xt_tt_assume_core_position(xt_tt t, size_t pos);          /* ttNetwork.cpp:734-738 */
int xt_tt_move_core(xt_tt t, size_t pos, int keep_rank);      /* TTNetwork::move_core, ttNetwork.cpp:581-628 */
int xt_tt_round(xt_tt t, const size_t* max_ranks, double eps);  /* TTNetwork::round, ttNetwork.cpp:643-665 */
int xt_tt_frob_norm(xt_tt t, double* out);                    /* ttNetwork.cpp:781-789 */
int xt_tt_dot(xt_tt a, xt_tt b, double* out);
int xt_tt_add(xt_tt a, xt_tt b, xt_tt* out);                  /* a + b, ttNetwork.cpp:796-847 */
int xt_tt_scale(xt_tt t, double alpha);
int xt_tt_apply(xt_tt A, xt_tt x, xt_tt* out);                /* A·x as TT (TTStack), ttStack.cpp:197-310 */
int xt_tt_full(xt_tt t, double* host_out);                    /* dense tensor (small TTs only) */
/* xerus datafiles: misc::save_to_file(TTNetwork, filename, format) (misc/fileIO.h:105-121, stream writers
   ttNetwork.cpp:1454-1470, tensorNetwork.cpp:1429-1469, tensor.cpp:1781-1806); format 0 = BINARY, 1 = TSV
   (misc::FileFormat). Files are byte compatible with xerus in both directions. */
int xt_tt_save(xt_tt t, const char* filename, int format);
/* misc::load_from_file<TTTensor|TTOperator> (fileIO.h:125-160, readers ttNetwork.cpp:1474-1488,
   tensorNetwork.cpp:1471-1504, tensor.cpp:1810-1843); the header decides tensor vs operator. */
int xt_tt_load(const char* filename, xt_tt* out);
/* ‖A·x − b‖/‖b‖ from the TT of A·x − b (the SPD residual_f, als.cpp:483-486) */
int xt_tt_residual(xt_tt A, xt_tt x, xt_tt b, double* out);

/* ------------------------------------------------------------ ALSVariant (als.h:37-223, als.cpp:708-778) */
typedef struct {
	unsigned sites;                 /* 1: ALS / ALS_SPD, 2: DMRG / DMRG_SPD */
	size_t numHalfSweeps;           /* 0 = until convergence */
	double convergenceEpsilon;      /* on the energy, default 1e-6 */
	int useResidualForEndCriterion;
	int preserveCorePosition;
	int assumeSPD;
	int precision;                  /* XT_F64, XT_F32 or XT_BF16 local matvecs (FP64 refinement for the latter two) */
	double localTolerance;          /* relative residual of each local solve */
	int maxLocalIterations;
	double innerTolerance;          /* mixed precision inner CG tolerance (0 = default) */
	int maxRefinements;
} xt_als_options;
typedef struct {
	size_t halfSweeps, localSolves, cgIterations, refinements, matvecs64, matvecsLow;
	double flopsMatvec, flopsEnv;   /* algorithmic flops: 2·a·n·c·(a·p + n·p·q + c·q) per matvec / stack push */
} xt_als_stats;
/* the presets ALS, ALS_SPD, DMRG, DMRG_SPD (als.cpp:781-788): sites ∈ {1,2}, assumeSPD ∈ {0,1} */
int xt_als_default_options(xt_als_options* o, unsigned sites, int assumeSPD);
/* ALSVariant::operator()(A, x, b, …) (als.h:152-207); A may be NULL (x ≈ b). Returns the energy. */
int xt_als_solve(const xt_als_options* o, xt_tt A, xt_tt x, xt_tt b, double* energy, xt_als_stats* stats);
/* ALS building blocks on device buffers (row-major): stacks L (a,p,a) and R (c,q,c), operator core A (p,n,n,q),
 * site vectors X, x, y, b (a,n,c). Left stack push, localOperatorSlice (als.cpp:184-200) contracted onto the left
 * stack as in prepare_stacks (:247): out (c,q,c) = Σ L(a,p,a')·X(a,n,c)·A(p,n,n',q)·X(a',n',c'). */
int xt_als_env_left(const double* L, const double* X, const double* A, size_t a, size_t n, size_t c, size_t p, size_t q, double* out);
/* Right stack push (als.cpp:239): out (a,p,a) = Σ R(c,q,c')·X(a,n,c)·A(p,n,n',q)·X(a',n',c'). */
int xt_als_env_right(const double* R, const double* X, const double* A, size_t a, size_t n, size_t c, size_t p, size_t q, double* out);
/* The local operator of construct_local_operator (als.cpp:383-402) applied without densifying it:
 * y(a,n,c) = Σ L(a,p,a')·A(p,n,n',q)·R(c,q,c')·x(a',n',c'). */
int xt_als_local_apply(const double* L, const double* A, const double* R, size_t a, size_t n, size_t c, size_t p, size_t q,
                       const double* x, double* y);
/* The local solve (localSolver, als.cpp:537) by CG on that operator (SPD): x holds the start value and receives
 * the solution; stops when ‖b − A_loc·x‖ <= tol·‖b‖ or after max_iter iterations; *iterations (nullable) = the
 * iterations taken. xerus's lapack_solver (als.cpp:43-75) densifies the same operator and solves it directly. */
int xt_als_local_cg(const double* L, const double* A, const double* R, size_t a, size_t n, size_t c, size_t p, size_t q,
                    const double* b, double* x, double tol, int max_iter, size_t* iterations);

/* PerformanceData (performanceData.h:43-157, PerformanceData::add at performanceData.cpp:34-80). While a sink
 * is set on the calling thread, xt_als_solve adds a point after the initial energy (flags 3), after every local
 * step (flags 0) and at the end of every half sweep (1 after a left-to-right, 3 after a right-to-left one), as
 * als.cpp:426-480,515-526 do; xt_adf_solve adds one per iteration (adf.cpp:514). elapsed_us excludes the
 * residual evaluation (stop_timer / continue_timer). Pass NULL to clear. */
typedef void (*xt_perf_callback)(void* user, size_t iteration, double elapsed_us, double residual, const size_t* ranks,
                                 size_t nranks, size_t flags);
int xt_perf_sink(xt_perf_callback cb, void* user);

/* Many independent problems (no single xerus call; a loop over ALSVariant::operator() and TTNetwork::round,
 * als.h:152-207, ttNetwork.cpp:643-684): problem p runs `sweeps` × (ALS with options o, then round(eps)) on one
 * of `threads` worker threads, each with a stream of its own. A[p] may be NULL (x ≈ b). residuals (optional, one
 * per problem) receives ‖A x − b‖/‖b‖; st (optional) the summed counters.
 * threads <= 0 runs the problems in lockstep on the library stream: every step of every problem (stack pushes,
 * local products, CG steps, QR / QC moves, SVDs) is one grouped launch for all problems. It applies to one-site
 * ALS_SPD with an operator and FP64 local solves (XT_F64) or FP32 CG inside FP64 refinement (XT_F32); other
 * options fall back to 16 worker threads. */
int xt_als_round_batch(const xt_als_options* o, size_t count, const xt_tt* A, const xt_tt* x, const xt_tt* b, size_t sweeps,
                       double eps, int threads, double* residuals, xt_als_stats* st);

/* ------------------------------------------- SinglePointMeasurementSet (measurments.h:43-91) and ADF */
int xt_meas_random(size_t P, size_t order, const size_t* dims, xt_meas* out);  /* measurments.cpp:222-248 */
int xt_meas_new(size_t P, size_t order, const uint32_t* positions, const double* values, xt_meas* out);
int xt_meas_free(xt_meas m);
int xt_meas_size(xt_meas m, size_t* P);
int xt_meas_positions(xt_meas m, uint32_t* host_out);
int xt_meas_positions_ptr(xt_meas m, const uint32_t** device_ptr);
int xt_meas_values(xt_meas m, double* host_out);
int xt_meas_set_values(xt_meas m, const double* host_values);
int xt_meas_measure(xt_meas m, xt_tt x);                                       /* measurments.cpp:125-153 */
/* The stack plan ADF builds for these positions (ADFVariant::InternalSolver::construct_stacks, adf.cpp:102-191): per
 * level k, the forward entries (samples sharing positions 0..k) and the backward ones (sharing k..d-1), each entry
 * listed by its representative, the group's smallest sample id, in ascending order (xerus's forwardUpdates[k] /
 * backwardUpdates[k]). fwd_counts / bwd_counts: d entries each; fwd_ids / bwd_ids (optional): the concatenated
 * representative lists; fwd_slots / bwd_slots (optional, P×d, [k·P + i]): sample i's entry within level k;
 * num_unique (optional): xerus's numUniqueStackEntries (forward entries + 1). Built on the device on first use. */
int xt_meas_stack_plan(xt_meas m, size_t* fwd_counts, size_t* bwd_counts, uint32_t* fwd_ids, uint32_t* bwd_ids, uint32_t* fwd_slots,
                       uint32_t* bwd_slots, size_t* num_unique);
/* xerus's projected gradient Δ_k = Σ (b_i − x_i)·∂x_i/∂G_k for all k (adf.cpp:314-396), concatenated cores */
int xt_adf_gradient(xt_tt x, xt_meas m, int precision, double* host_grad, double* residual_norm);
/* ADFVariant::operator() (adf.h:248-266); max_ranks NULL = x's ranks */
int xt_adf_solve(xt_tt x, xt_meas m, const size_t* max_ranks, size_t max_iterations, double target_residual,
                 double min_decrease, int precision, double* residual, size_t* iterations);

/* ------------------------------------------- RankOneMeasurementSet (measurments.h:94-150) and its ADF path.
 * Measurement i is u_{i,0} ⊗ … ⊗ u_{i,d-1}; `vectors` holds P rows of Σ_k n_k doubles (mode k's vector at
 * column offset Σ_{j<k} n_j). */
int xt_r1meas_random(size_t P, size_t order, const size_t* dims, xt_r1meas* out);  /* random(P, dims), measurments.cpp:267-272,513-536 */
int xt_r1meas_from_single_point(xt_meas m, size_t order, const size_t* dims, xt_r1meas* out);  /* measurments.cpp:253-265 */
int xt_r1meas_new(size_t P, size_t order, const size_t* dims, const double* vectors, const double* values, xt_r1meas* out); /* add */
int xt_r1meas_free(xt_r1meas m);
int xt_r1meas_size(xt_r1meas m, size_t* P);
int xt_r1meas_vectors(xt_r1meas m, double* host_out);
int xt_r1meas_values(xt_r1meas m, double* host_out);
int xt_r1meas_set_values(xt_r1meas m, const double* host_values);
int xt_r1meas_sort(xt_r1meas m, int positions_only);   /* measurments.cpp:329-344 (internal::comp order) */
int xt_r1meas_normalize(xt_r1meas m);                  /* measurments.cpp:346-355 */
int xt_r1meas_measure(xt_r1meas m, xt_tt x);           /* measurments.cpp:392-424 */
int xt_r1meas_test(xt_r1meas m, xt_tt x, double* out); /* measurments.cpp:469-505: ‖b − x(u)‖/‖b‖ */
/* Δ_k = Σ (b_i − x(u_i))·∂x(u_i)/∂G_k for all k (adf.cpp:314-396 with the rank-one dyadic product :333-357) */
int xt_adf_gradient_r1(xt_tt x, xt_r1meas m, int precision, double* host_grad, double* residual_norm);
/* ADFVariant::operator() on a RankOneMeasurementSet (adf.h:248-266; update_x adf.cpp:481-487) */
int xt_adf_solve_r1(xt_tt x, xt_r1meas m, const size_t* max_ranks, size_t max_iterations, double target_residual,
                    double min_decrease, int precision, double* residual, size_t* iterations);

/* Raw sampled-entry kernels on caller-owned device buffers (FP32 cores, all cores concatenated in TT order;
 * dims n[d] and ranks r[d+1] given on the host; positions P×d uint32 on the device). The rank list describes
 * every core: core k is (r[k], n[k], r[k+1]) with r[0] = r[d] = 1, so `cores` and `grads` must each hold exactly
 * Σ_k r[k]·n[k]·r[k+1] floats; the kernels take every core offset from that list and cannot see the buffers'
 * real extent. Positions must satisfy pos[i·d + k] < n[k].
 * values[i] = x(position_i); grads = Σ_i seed_i·∂x_i/∂G (the adjoint of the sampling). */
int xt_tt_sample_f32(int d, const int* n, const int* r, const float* cores, const uint32_t* pos, int P, float* values,
                     void* stream);
/* resid[i] = x(position_i) − b[i]; sums (host, nullable) = {Σ resid², Σ b²} accumulated in FP64 */
int xt_tt_sample_residual_f32(int d, const int* n, const int* r, const float* cores, const uint32_t* pos, int P, const float* b,
                              float* resid, double* sums, void* stream);
int xt_tt_sample_adjoint_f32(int d, const int* n, const int* r, const float* cores, const uint32_t* pos, int P,
                             const float* seed, float* grads, void* stream);
/* Both of the above for one gradient step of ADF (calculate_residual adf.cpp:290-312 followed by
 * calculate_projected_gradient adf.cpp:314-396, here for all cores): resid[i] = x(position_i) − b[i], sums (host,
 * nullable) = {Σ resid², Σ b²} in FP64, grads = Σ_i resid_i·∂x_i/∂G. One pass over the samples where the ranks are
 * <= 8 (the same resid and grads as the two calls; sums to FP64 rounding), else the two kernels in turn. */
int xt_tt_sample_residual_adjoint_f32(int d, const int* n, const int* r, const float* cores, const uint32_t* pos, int P,
                                      const float* b, float* resid, double* sums, float* grads, void* stream);

/* No counterpart in xerus. For callers that cannot allocate device memory (code under HIP graph capture, serving
 * loops, harnesses that forbid allocation): xt_tt_sample_residual_adjoint_f32's one-pass kernels with every scratch
 * buffer (block partials, padded cores, descriptors) in a static 16 MiB device array per GPU, held until the process
 * ends. Same inputs and outputs except that `sums` is a DEVICE pointer to 2 doubles (nullable) and the call does not
 * synchronise: it is ordered on `stream` (NULL = the library stream). Limits (XT_ERR_UNSUPPORTED beyond them):
 * order 2..16, ranks <= 8, n_k <= 16, Σ_k n_k <= 180 and Σ_k r_k·n_k·r_{k+1} <= 16314 floats. Every call on a GPU
 * shares that GPU's array, so calls that overlap on one GPU (several streams or threads) race. */
int xt_tt_sample_residual_adjoint_f32_noalloc(int d, const int* n, const int* r, const float* cores, const uint32_t* pos, int P,
                                              const float* b, float* resid, double* sums, float* grads, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* XTT_H */
