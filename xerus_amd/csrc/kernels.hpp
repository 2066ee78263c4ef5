// Launchers of libxtt's gfx950 kernels. Every launcher enqueues on the given stream and checks the launch.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <vector>

#include "xtt_common.hpp"

namespace xtt {

// C[b] = alpha·op(A[b])·op(B[b]) + beta·C[b], row-major, op(A) m×k, op(B) k×n. Batch strides in elements.
// `skip` (device int, may be null): when *skip != 0 the launch does nothing (predicated CG tails).
struct GemmDesc {
	int ta = 0, tb = 0;
	size_t m = 0, n = 0, k = 0;
	size_t lda = 0, ldb = 0, ldc = 0;
	long long sA = 0, sB = 0, sC = 0;
	size_t batch = 1;
	double alpha = 1.0, beta = 0.0;
	const int* skip = nullptr;
};
GemmDesc gemm_desc(int ta, int tb, size_t m, size_t n, size_t k);  // packed leading dimensions, batch 1
void gemm_f64(const GemmDesc& g, const double* A, const double* B, double* C, hipStream_t s);
// Two independent products in one launch when both have small tile grids (never split over K), else two launches.
void gemm_f64_pair(const GemmDesc& g0, const double* A0, const double* B0, double* C0, const GemmDesc& g1, const double* A1,
                   const double* B1, double* C1, hipStream_t s);
void gemm_f32(const GemmDesc& g, const float* A, const float* B, float* C, hipStream_t s);
// BF16 products with FP32 accumulation; operands are stored as FP32 and rounded to bf16 when staged in LDS.
void gemm_bf16(const GemmDesc& g, const float* A, const float* B, float* C, hipStream_t s);
// Tile autotuner of gemm_f64 / gemm_f32 (dtype 0 / 1): a choice is the tile (32 or 64) and, for 32-tiles of a
// single product, the number of split-K slices (0: none). A choice set for a shape overrides the built-in heuristic
// for exactly that shape (m, n, k, batch, ta, tb). gemm_tune times the heuristic and every candidate over `iters`
// launches on stream s (HIP events), installs the fastest and returns both.
struct GemmChoice { int tile = 32, slices = 0; };
struct GemmTuneResult { GemmChoice heuristic, best; double heuristicUs = 0, bestUs = 0; };
GemmTuneResult gemm_tune(int dtype, const GemmDesc& d, int iters, hipStream_t s);
void gemm_tune_set(int dtype, const GemmDesc& d, GemmChoice c);
void gemm_tune_clear();
GemmChoice gemm_heuristic(const GemmDesc& d);

// out = reshuffle(in, shuffle): shuffle[i] = new position of old mode i (xerus::reshuffle semantics).
void permute_f64(size_t degree, const size_t* dims, const size_t* shuffle, const double* in, double* out, hipStream_t s);
void revtrans_f64(const double* in, size_t rows, size_t cols, double* out, hipStream_t s);  // out[i][j] = in[rows-1-j][cols-1-i]
void convert_f64_f32(const double* in, float* out, size_t n, hipStream_t s);
void convert_f32_f64(const float* in, double* out, size_t n, hipStream_t s);
void fill_f64(double* p, double v, size_t n, hipStream_t s);
void zero_u32(unsigned* p, size_t n, hipStream_t s);  // a kernel, so it may be recorded into a graph
void eye_f64(double* p, size_t n, hipStream_t s);     // the n×n identity
// dst[j][0..n[j]) = src[j][0..n[j]) for j < count, one launch (count <= kMultiCopy; empty copies allowed)
constexpr int kMultiCopy = 32;
struct MultiCopy {
	const double* src[kMultiCopy];
	double* dst[kMultiCopy];
	size_t n[kMultiCopy];
	int count = 0;
};
void multi_copy_f64(const MultiCopy& c, hipStream_t s);
// the same with FP32 destinations (dst[j] points to floats) and conversion
void multi_convert_f64_f32(const MultiCopy& c, hipStream_t s);
// *out = the truncation rank of calculate_svd for singular values sv[0..k) (device): min(k, maxRank unless 0), cut
// at the first j >= 1 with sv[j] <= eps·sv[0]
void svd_rank_f64(const double* sv, size_t k, size_t maxRank, double eps, int* out, hipStream_t s);
void scale_f64(double* p, double alpha, size_t n, hipStream_t s);
void axpy_f64(double* y, double alpha, const double* x, size_t n, hipStream_t s);
// m (rows×cols) *= sv[row] (byRow) or sv[col], sv on the device
void scale_by_f64(double* m, const double* sv, size_t rows, size_t cols, bool byRow, hipStream_t s);        // y += alpha·x
// Deterministic FP64 reductions into *out (device). `partials` holds >= reduce_blocks(n) doubles, `counter`
// one zero-initialised unsigned that the kernel returns to zero.
size_t reduce_blocks(size_t n);
void dot_f64(const double* x, const double* y, size_t n, double* out, double* partials, unsigned* counter, hipStream_t s);
// r = b − q; out2[0] = ‖r‖², out2[1] = ‖b‖² (same partials/counter contract as dot_f64). With fail set, the
// last block also adds one to *fail unless ‖r‖² <= tol2·‖b‖² (a convergence test the host reads later).
void residual_norms_f64(double* r, const double* b, const double* q, size_t n, double* out2, double* partials, unsigned* counter,
                        hipStream_t s, double tol2 = 0.0, unsigned* fail = nullptr);
// out3 = {Σ|x_i|, max|x_i|, ‖x‖₂} (the two norm scaled by max|x_i| as dnrm2 does); same partials/counter contract
void norms_f64(const double* x, size_t n, double* out3, double* partials, unsigned* counter, hipStream_t s);
void dot_f32_acc64(const float* x, const float* y, size_t n, double* out, double* partials, unsigned* counter, hipStream_t s);

// ---- local CG on the device; all scalars live in CGState (device) -------------------------------------
struct CGState {
	double rr;        // current ‖r‖²
	double rr0;       // ‖b‖² of the system being solved (tolerance reference)
	double tol2;      // stop when rr <= tol2·rr0
	double alpha, beta, pq;
	int done;         // 1 once converged (every later CG launch is a no-op)
	int iter;         // iterations performed
	int maxIter;
	int pad;
	unsigned counter; // reduction ticket
	unsigned pad2;
	double rrPrev;    // chunked grouped CG: ‖r‖² and the iteration count before the running step
	int iterPrev;
	int pad3;
};
template <class T> void cg_start(T* r, T* p, const T* b, const T* q, size_t n, CGState* st, double* partials, hipStream_t s);
template <class T> void cg_alpha(const T* p, const T* q, size_t n, CGState* st, double* partials, hipStream_t s);
template <class T> void cg_update(T* x, T* r, const T* p, const T* q, size_t n, CGState* st, double* partials, hipStream_t s);
template <class T> void cg_direction(T* p, const T* r, size_t n, CGState* st, hipStream_t s);
void cg_state_init(CGState* st, double tol2, int maxIter, hipStream_t s);
// After cg_start with a warm start worse than zero (st->rr > st->rr0): x = 0, r = p = b, st->rr = st->rr0. CG's
// recursive residual cannot resolve the true one below ε·‖A‖·max‖x_k‖, so a start of large norm caps the accuracy.
void cg_zero_start(double* x, double* r, double* p, const double* b, size_t n, CGState* st, hipStream_t s);

// Whole CG solve of an ALS local system in one launch with grid-wide barriers (cgfused.hip). Operator pieces
// are FP64 in the layouts L (a,P,a), Am (n,Q,P,n), R (c,Q,c); the arithmetic type T is double or float.
//   mode 0: warm start from x (FP64, in/out); r0 = b − A·x and the state (rr0, tol2, maxIter, done, the zeroed
//           fail flag) come from cg_state_init + cg_start. A warm start with ‖r0‖ > ‖b‖ is replaced by x = 0, r = b.
//   mode 1: correction e from 0 for the right-hand side r0·r0scale; x += xscale·e on exit; tol2 and maxIter are
//           taken from here, relative to ‖r0·r0scale‖²; the state only needs cg_state_init (the fail flag).
// Not for concurrent use on several streams: the grid takes every compute unit.
// iterAcc (optional, device): [0] += iterations, [1] += iterations·flopsPerIter; set to NaN if a barrier wait
// ran out of time (then x is left unchanged and the state's `pad` flag is set).
struct FusedCG {
	const double* L = nullptr;
	const double* Am = nullptr;
	const double* R = nullptr;
	const double* r0 = nullptr;
	const double* b = nullptr;  // mode 0: the right-hand side, the start residual when the warm start is discarded
	double* x = nullptr;
	void* vec = nullptr;       // 7·a·n·c elements of T
	void* T2 = nullptr;        // a·n·c·Q elements of T
	double* sync = nullptr;      // FUSED_CG_SYNC doubles, zeroed once when allocated, then never cleared
	unsigned long long epoch = 0;  // 1, 2, … per launch on the same `sync` buffer (below 2^47)
	CGState* st = nullptr;
	double* iterAcc = nullptr;
	double r0scale = 1, xscale = 1, tol2 = 0, flopsPerIter = 0;
	int maxIter = 0, mode = 0;
	size_t a = 0, n = 0, c = 0, P = 0, Q = 0;
};
constexpr size_t FUSED_CG_SYNC = 1025 * 8;
template <class T> bool fused_cg_fits(size_t a, size_t n, size_t c, size_t P, size_t Q);
// The grid is sized to the workgroups the device holds at once (occupancy query). Returns false, with nothing
// enqueued, when the launch cannot be made (the caller then solves on the chunked CG). A barrier wait that runs out
// of time (the grid was not resident because other work shared the device) leaves x untouched and marks the solve
// failed (iterAcc[0] = NaN). XTT_FUSED_CG_WAIT_TICKS (wall-clock ticks, read per launch) overrides the 2 s wait limit;
// 0 makes every wait give up at once, which exercises the failure path with every workgroup exiting normally.
template <class T> bool fused_cg(const FusedCG& d, hipStream_t s);

// ---- dense factorisations (one workgroup per matrix, FP64, matrix in device memory) -------------------
// In-place Householder QR (pivot = 0) or column-pivoted QR (pivot = 1) of a row-major m×n matrix, LAPACK
// dgeqr2 / dlaqp2 conventions: R on and above the diagonal, reflectors below, tau[min(m,n)], jpvt[n]
// (original column index of each position). `ws` needs 2n doubles. *rank receives xerus's QC rank
// (blasLapackWrapper.cpp:268-272) when pivot = 1, min(m,n) otherwise.
void geqp2(double* A, size_t m, size_t n, double* tau, int* jpvt, int pivot, double* ws, int* rank, hipStream_t s);
// Q (m×k, row-major, k <= min(m,n)) = first k columns of H_0…H_{min-1}·[I;0] from the reflectors in A.
void org2r(const double* A, size_t m, size_t n, const double* tau, double* Q, size_t k, hipStream_t s);
// C (rank×n, row-major) from the pivoted R in A: C[row][jpvt[col]] = R[row][col] for row <= col.
void extract_c(const double* A, size_t m, size_t n, const int* jpvt, const int* rank, double* C, hipStream_t s);
// R (k×n) = upper triangle of A (row-major), k = min(m,n).
void extract_r(const double* A, size_t m, size_t n, double* R, hipStream_t s);
// Register-resident Householder QR for n <= 64 columns (any row count, tree-reduced above 256 rows).
bool hh_supported(size_t m, size_t n);
// R factor (min(m,n) × n, zero below the diagonal) of a row-major m × n matrix of any width: 64-column panels of
// register-resident TSQR with the trailing columns updated per tree node.
void hh_r_blocked(const double* A, size_t m, size_t n, double* R, hipStream_t s);
// A (m×n row-major, or n×m row-major read transposed when `trans`) = Q·R: Q m×k (may be null), R k×n upper,
// k = min(m, n).
void hh_qr(const double* A, size_t m, size_t n, bool trans, double* Q, double* R, hipStream_t s);
// xerus's QC: A·P = Q·R with column pivoting, rank by the 16·ε·R_00 rule. Writes *rank (device), Q (m×rank,
// ld rank, capacity m×k), C (rank×n, capacity k×n) and jpvt[n].
void hh_qc(const double* A, size_t m, size_t n, double* Q, double* C, int* jpvt, int* rank, hipStream_t s);
// One-sided Jacobi SVD of a row-major m×n matrix with m >= n: U (m×n), S (n, descending), Vt (n×n).
// Workspace: m·n + n·n + 2n doubles.
// transposedOut: U is written as Uᵀ (n×m) and Vt as V (n×n), row-major.
void jacobi_svd(const double* A, size_t m, size_t n, double* U, double* S, double* Vt, double* ws, hipStream_t s,
                bool transposedOut = false);

// ---- grouped launches (lockstep batch driver): one launch, many independent problems ----------------------
// Workgroup b belongs to the task t with start[t] <= b < start[t + 1] (start ascending, start[0] = 0).
__device__ inline int group_task(const int* start, int count, int b) {
	int lo = 0, hi = count - 1;
	while (lo < hi) {
		const int mid = (lo + hi + 1) >> 1;
		if (start[mid] <= b) lo = mid; else hi = mid - 1;
	}
	return lo;
}
// Copies a task table to device memory on stream s (through a per-thread pinned ring) and returns its device
// address, valid for the launches that follow on s. table_ring_release(s) after s was synchronised frees the ring.
const void* table_upload(const void* src, size_t bytes, hipStream_t s);
void table_ring_release(hipStream_t s);
// Copies a table to dst (device memory the caller owns) through the pinned ring, on stream s.
void table_upload_to(void* dst, const void* src, size_t bytes, hipStream_t s);

enum MapOp { MAP_PERMUTE = 0, MAP_REVTRANS = 1, MAP_SCALE_BY = 2, MAP_EYE = 3, MAP_FILL = 4 };
struct MapTask {
	int op, deg;
	long long n;
	long long dims[4], outStride[4];
	const double* in;
	double* out;
	const double* aux;
	double alpha;
};
MapTask map_permute(size_t degree, const size_t* dims, const size_t* shuffle, const double* in, double* out);  // as permute_f64
MapTask map_copy(const double* in, double* out, size_t n);
MapTask map_revtrans(const double* in, size_t rows, size_t cols, double* out);                              // as revtrans_f64
MapTask map_scale_by(double* m, const double* sv, size_t rows, size_t cols, bool byRow);                      // as scale_by_f64
MapTask map_eye(double* out, size_t n);
MapTask map_fill(double* out, double v, size_t n);
void map_group(const std::vector<MapTask>& tasks, hipStream_t s);

// CG of many local systems, one workgroup each (vectors in T = double or float, reductions in FP64); q = A·p (or
// A·x before the start) comes from grouped GEMMs. q == nullptr at the start means x = 0 (q = 0).
template <class T>
struct CGTaskT {
	int n;
	T* x;
	T* r;
	T* p;
	const T* q;
	const T* b;
	CGState* st;  // tol2 and maxIter set beforehand (cg_state_init semantics)
};
using CGTask = CGTaskT<double>;
template <class T> void cg_group_start(const std::vector<CGTaskT<T>>& tasks, hipStream_t s);  // cg_start + the zero-start rule
template <class T> void cg_group_step(const std::vector<CGTaskT<T>>& tasks, hipStream_t s);   // cg_alpha, cg_update, cg_direction
template <class T> void cg_group_step_run(const CGTaskT<T>* devTasks, int count, hipStream_t s);  // on a table already on the device
// The same step for systems split into chunks of kCGChunk elements, one workgroup per chunk, in three launches
// (p·q partials; α, x and r updates with ‖r‖² partials; the stop test and p = r + β·p). Each workgroup sums its
// system's partials itself in chunk order, so every chunk computes the same α and β; a system of one chunk gets
// the result of cg_group_step_run bit for bit. part: 2·nch doubles per system.
constexpr int kCGChunk = 4096;
template <class T>
struct CGChunk {
	CGTaskT<T> k;
	int c0, len, nch, ci;
	double* part;
};
template <class T> void cg_chunk_step_run(const CGChunk<T>* devTasks, int count, hipStream_t s);
// Mixed-precision plumbing of the refinement loop, many vectors per launch:
//   MIX_CVT32:  out (float) = alpha · a (double);  MIX_AXPY32: out (double) += alpha · a (float);
//   MIX_ZERO:   out (double) = 0, out2 (double) = a (double) (a warm start worse than zero: x = 0, r = b).
enum MixOp { MIX_CVT32 = 0, MIX_AXPY32 = 1, MIX_ZERO = 2 };
struct MixTask { int op; long long n; const void* a; void* out; void* out2; double alpha; };
void mix_group(const std::vector<MixTask>& tasks, hipStream_t s);
// r = b − q with out[2t] = ‖r‖², out[2t + 1] = ‖b‖² (FP64), one workgroup per task
struct ResidTask { int n; double* r; const double* b; const double* q; };
void resid_group(const std::vector<ResidTask>& tasks, double* out, hipStream_t s);
// Core k of w = A·x + (−b) as a TT (tt_apply, then tt_add with b's first core negated): rank index (α, i) of the
// product, then b's ranks; the first core concatenates horizontally, the last vertically, the others block-diagonal.
// A core (rA, n, m, rA2), x core (rx, m, rx2), b core (rb, n, rb2); w (L, n, R) with L = first ? 1 : rA·rx + rb,
// R = last ? 1 : rA2·rx2 + rb2.
struct ResCoreTask { const double* A; const double* x; const double* b; double* w; int rA, rA2, rx, rx2, rb, rb2, n, m, first, last; };
void res_core_group(const std::vector<ResCoreTask>& tasks, hipStream_t s);
struct Dot2Task { int n; const double* a; const double* b; const double* c; const double* d; };
void dot2_group(const std::vector<Dot2Task>& tasks, double* out, hipStream_t s);  // out[2t] = a·b, out[2t+1] = c·d

// C = alpha·op(A)·op(B) + beta·C for many products in one launch (64×64 tiles, no split-K).
struct GemmTask { GemmDesc d; const void* A; const void* B; void* C; };
void gemm_group_f64(const std::vector<GemmTask>& tasks, hipStream_t s);
// A grouped GEMM launch whose tables stay on the device (in buf) and that runs again with gemm_group_run_f64 (the
// launch of a CG step's products, repeated every iteration). plan_f64 also runs it once.
struct GemmGroupPlan {
	DBuf<char> buf;
	const void* tasks = nullptr;
	const int* start = nullptr;
	int count = 0, blocks = 0;
	size_t lds = 0;
	double flops = 0, bytes = 0;
	const char* tag = nullptr;  // launch profiler label of the replays (default: the kernel's name)
};
void gemm_group_plan_f64(const std::vector<GemmTask>& tasks, GemmGroupPlan& plan, hipStream_t s);
void gemm_group_run_f64(const GemmGroupPlan& plan, hipStream_t s);
// the same with FP32 operands (v_mfma_f32_16x16x4_f32)
void gemm_group_f32(const std::vector<GemmTask>& tasks, hipStream_t s);
void gemm_group_plan_f32(const std::vector<GemmTask>& tasks, GemmGroupPlan& plan, hipStream_t s);
void gemm_group_run_f32(const GemmGroupPlan& plan, hipStream_t s);
// Householder QR of many matrices (each as hh_qr): Q (m × k) and R (k × n), k = min(m, n), n <= 64.
struct QRJob { const double* A; int m, n; bool trans; double* Q; double* R; };
void hh_qr_group(const std::vector<QRJob>& jobs, hipStream_t s);
// xerus's QC of many matrices (each as hh_qc): rank, jpvt, Q (m × rank, ld rank) and C (rank × n) on the device.
// R factor (min(m,n) × n) of a row-major m × n matrix of any width; A is overwritten
struct RJob { double* A; long long m, n; double* R; };
void hh_r_group(const std::vector<RJob>& jobs, hipStream_t s);
struct QCJob { const double* A; int m, n; double* Q; double* C; int* jpvt; int* rank; };
void hh_qc_group(const std::vector<QCJob>& jobs, hipStream_t s);
// One-sided Jacobi SVD of many matrices (each as jacobi_svd, m >= n, workspace ws as there).
struct SVDJob { const double* A; int m, n; double* U; double* S; double* Vt; double* ws; bool transposedOut; };
void jacobi_svd_group(const std::vector<SVDJob>& jobs, hipStream_t s);

// ---- ADF sampled-entry kernels (FP32 cores / values, FP64 norms) ---------------------------------------
struct AdfDesc {
	int d;                    // order
	int P;                    // number of measurements
	const int* n;             // device: mode sizes [d]
	const int* rank;          // device: ranks r_0..r_d (r_0 = r_d = 1) [d+1]
	const long long* coreOff; // device: element offset of core k in `cores` [d]
	int maxR;                 // max rank (<= 16)
	int gradSize;             // total elements of all cores
	int sumN;                 // Σ_k n_k
	int maxN;                 // max_k n_k
};
// value_i = x(position_i); resid_i = value_i - b_i; Σ resid² and Σ b² in FP64 into norms[0..1].
template <class T>
void adf_residual(const AdfDesc& a, const T* cores, const uint32_t* pos, const T* values, T* resid, double* norms,
                  double* partials, unsigned* counter, hipStream_t s);
// grad (layout of `cores`) = Σ_i resid_i·∂x_i/∂G, restricted to core kSel when kSel >= 0 (other entries 0).
// `partials` holds adf_grad_blocks(P)·gradSize elements.
int adf_grad_blocks(int P);
template <class T>
void adf_gradient(const AdfDesc& a, const T* cores, const uint32_t* pos, const T* resid, T* grad, T* partials, int kSel,
                  hipStream_t s);
// resid = x(pos) − values and norms = {‖resid‖², ‖values‖²} (as adf_residual) together with grad = the adjoint
// with seed resid (as adf_gradient, every core). `partials` holds adf_grad_blocks(P)·gradSize floats, `normPart`
// 2·1024 doubles, `counter` one zeroed unsigned (used only when the fused kernel does not apply).
// whether adf_residual_gradient_f32 runs as one gradient launch (else the two kernels in turn)
bool adf_residual_gradient_fused_ok(const AdfDesc& a);
void adf_residual_gradient_f32(const AdfDesc& a, const float* cores, const uint32_t* pos, const float* values, float* resid, double* norms,
                               float* grad, float* partials, double* normPart, unsigned* counter, hipStream_t s);
// The one-pass path alone (requires adf_residual_gradient_fused_ok) with caller-provided buffers for the padded cores:
// Gp holds adf_pad_floats(a) floats, padOff ADF_PAD_OFFSETS long longs; partials min(adf_grad_blocks(P), 256)·gradSize
// floats, normPart 2·256 doubles. Allocates nothing.
constexpr int ADF_PAD_OFFSETS = 16;
int adf_pad_floats(const AdfDesc& a);
void adf_residual_gradient_f32_into(const AdfDesc& a, const float* cores, const uint32_t* pos, const float* values, float* resid,
                                    double* norms, float* grad, float* partials, double* normPart, float* Gp, long long* padOff,
                                    hipStream_t s);
// Rank-one measurements (U: P × sumN, measurement i's vectors u_{i,0..d-1} concatenated):
// value_i = x(u_i) with core kSel replaced by D (kSel < 0: none); resid_i = value_i − b_i (values/resid may be
// null: b = 0 / not stored); norms2 = {Σ resid², Σ b²}. `partials` holds 2·r1_blocks(P) doubles.
int r1_blocks(int P);
template <class T>
void r1_residual(const AdfDesc& a, const T* cores, const T* U, const T* values, T* resid, double* norms2, double* partials,
                 int kSel, const T* D, hipStream_t s);
// grad (core k's layout, rl × n × rr) = Σ_i resid_i·∂x(u_i)/∂G_k; `partials` holds r1_blocks(P)·rl·n·rr elements.
template <class T>
void r1_gradient(const AdfDesc& a, const T* cores, const T* U, const T* resid, int k, int rl, int n, int rr, T* grad,
                 T* partials, hipStream_t s);
// out[j] = Σ_{i: i_k = j} (x_i with core k replaced by D)²; `partials` holds adf_slice_blocks(P)·nslices doubles.
int adf_slice_blocks(int P);
template <class T>
void adf_slice_norms(const AdfDesc& a, const T* cores, const uint32_t* pos, int k, const T* D, int nslices, double* out,
                     double* partials, hipStream_t s);
// ---- ADF stacks (adf.cpp:102-312) ----------------------------------------------------------------------------
// Stack plan of one direction. Forward entry (k, i) is the row G_0[i_0]·…·G_k[i_k] (r_{k+1} values), backward
// entry (k, i) the column G_k[i_k]·…·G_{d-1}[i_{d-1}] (r_k values). Samples with equal positions 0..k (forward) or
// k..d-1 (backward) share the entry of level k. Each shared entry is computed once, from its representative (the
// group's smallest sample id); the entries of a level are ordered by representative, as xerus's _updates[k].
struct AdfPlanSide {
	std::vector<long long> count;  // entries per level
	std::vector<long long> first;  // index in rep of each level's first entry (prefix sums of count)
	DBuf<uint32_t> slot;           // [k·P + i]: the entry of sample i at level k
	DBuf<uint32_t> rep;            // [first[k] + g]: the representative of level k's entry g
};
struct AdfPlan {
	int d = 0, P = 0;
	AdfPlanSide side[2];  // 0: forward (shared prefixes), 1: backward (shared suffixes)
	DBuf<uint32_t> posT;  // [k·P + i]: the positions mode by mode (the sweep kernels read the column of their core)
};
// Builds both sides on the device from P×d positions (device, row-major); maxPos[j] = largest position of mode j.
void adf_plan_build(AdfPlan& plan, const uint32_t* pos, int P, int d, const std::vector<uint32_t>& maxPos, hipStream_t s);
// In the sweep kernels below `pos` is the positions' column of mode k (P entries, AdfPlan::posT + k·P).
// Forward level k of the stacks: F_k[g] = F_{k-1}[slot_{k-1}(rep_g)]·G_k[:, i_k(rep_g), :] (F_{-1} = 1). `core` is
// G_k (rl × n × rr), `prevSlot`/`prev` level k−1's slots and values (null at k = 0), out count × rr values.
template <class T>
void adf_forward_stack(const T* core, int rl, int n, int rr, int k, int d, const uint32_t* pos, int P, const uint32_t* rep, long long count,
                       const uint32_t* prevSlot, const T* prev, T* out, hipStream_t s);
// Backward level k: B_k[g] = G_k[:, i_k(rep_g), :]·B_{k+1}[slot_{k+1}(rep_g)] (B_d = 1), out count × rl values.
template <class T>
void adf_backward_stack(const T* core, int rl, int n, int rr, int k, int d, const uint32_t* pos, int P, const uint32_t* rep, long long count,
                        const uint32_t* nextSlot, const T* next, T* out, hipStream_t s);
// The stacks around core k: left = F_{k-1} (slots of level k−1, values; null at k = 0), right = B_{k+1} (null at
// k = d−1).
template <class T>
struct AdfSides {
	const uint32_t* fSlot;
	const T* F;
	const uint32_t* bSlot;
	const T* B;
};
// Blocks of the residual-gradient and slice kernels (at most 512).
int adf_stack_blocks(int P);
// Residual and gradient at core k in one pass: resid_i = F·G_k[:, i_k, :]·B − b_i, norms = {Σ resid², Σ b²} (FP64,
// fixed-order sums; skipped when norms is null) and grad (rl × n × rr) = Σ_i resid_i·F_iᵀ ⊗ e_{i_k} ⊗ B_iᵀ.
// `partials` holds adf_stack_blocks(P)·rl·n·rr elements, `normPart` 2·adf_stack_blocks(P) doubles.
template <class T>
void adf_stack_residual_gradient(const T* core, int rl, int n, int rr, int k, int d, const uint32_t* pos, int P, const AdfSides<T>& sd,
                                 const T* values, T* resid, double* norms, T* grad, T* partials, double* normPart, hipStream_t s);
// update_x of core k (adf.cpp:413-478): normA_j = Σ_{i: i_k = j} (F_i·g[:, j, :]·B_i)², then slice j of the FP64 core
// moves by −(‖g_j‖²/normA_j)·g_j, g the gradient of ½Σ(x − b)² (xerus's Δ = −g). `partials` holds
// adf_stack_blocks(P)·n doubles.
template <class T>
void adf_stack_step(double* core, const T* g, int rl, int n, int rr, int k, int d, const uint32_t* pos, int P, const AdfSides<T>& sd,
                    double* partials, hipStream_t s);
// out[e] = Σ_b partials[b·size + e] in a fixed order (blocks' partial gradients); out[j] = Σ_b partials[b·nslices + j]
// in block order (slice sums).
template <class T>
void adf_sum_partials(const T* partials, int nblocks, int size, T* out, hipStream_t s);
void adf_sum_slices(const double* partials, int nblocks, int nslices, double* out, hipStream_t s);
constexpr int ADF_STACK_MAX_N = 2048;  // largest mode size of the stack kernels (slice sums in LDS)

// update_x (adf.cpp:468-487): core (FP64, rl × n × rr) −= step·g, where g is the gradient of ½Σ(x − b)² for that core
// (xerus's Δ = −g) and step = ‖g_j‖²/normA[j] per slice j (single points) or ‖g‖²/normA[0] (rank-one, perSlice false).
template <class T>
void adf_update_core(double* core, const T* g, const double* normA, int rl, int n, int rr, bool perSlice, hipStream_t s);

}  // namespace xtt
