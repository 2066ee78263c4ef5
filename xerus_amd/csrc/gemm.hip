// MFMA GEMM for gfx950: C[b] = alpha·op(A[b])·op(B[b]) + beta·C[b], row-major, batched with element strides.
//
// One workgroup computes a BM×BN tile with WM×WN waves; each wave owns TM×TN 16×16 MFMA tiles. The K
// extent is staged through LDS in chunks of up to KMAX (our TT shapes have K <= 256, so one chunk: one
// global-load phase per tile). Staging loads are issued 8 at a time before the LDS stores so the wave keeps
// many loads in flight. Ragged edges are predicated (zero fill), so any m, n, k works.
//
// FP64 uses v_mfma_f64_16x16x4_f64, FP32 v_mfma_f32_16x16x4_f32 and the BF16 path
// v_mfma_f32_16x16x32_bf16 with FP32 accumulation (operands are held as FP32 in memory and rounded to bf16
// while being staged). Operand lane layouts (16×16 tile, lane l):
//   x4 shapes : A[i=l%16][k=l/16], B[k=l/16][j=l%16]
//   x32 bf16  : A[i=l%16][k=8(l/16)+e], B[k=8(l/16)+e][j=l%16], e = 0..7
//   result    : FP32 / BF16: D[i=4(l/16)+r][j=l%16]; FP64: D[i=4r+l/16][j=l%16], r = 0..3 (register r)
#include <map>
#include <mutex>
#include <set>
#include <string>
#include <tuple>
#include <vector>

#include "kernels.hpp"
#include "xtt_common.hpp"

namespace xtt {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

struct Args {
	const void* A;
	const void* B;
	void* C;
	int m, n, k;
	int lda, ldb, ldc;
	long long sA, sB, sC;
	int ta, tb;
	double alpha, beta;
	const int* skip;
	int kc;      // K chunk staged in LDS
	int ksplit;  // split-K: blockIdx.z = K slice of this length, C = slice partials (m × n each); 0 = off
};

__device__ inline d4 mfma_x4(double a, double b, d4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
__device__ inline f4 mfma_x4(float a, float b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }

// Row of result register r in lane group lk (= lane/16) of a 16×16 MFMA tile: the FP64 instruction interleaves
// rows across lane groups (row 4r + lk), the FP32 one gives each lane group 4 consecutive rows (row 4lk + r).
template <class T> struct RowOf;
template <> struct RowOf<double> { __device__ static int row(int lk, int r) { return 4 * r + lk; } };
template <> struct RowOf<float> { __device__ static int row(int lk, int r) { return 4 * lk + r; } };

template <class T> struct AccOf;
template <> struct AccOf<double> { using type = d4; };
template <> struct AccOf<float> { using type = f4; };

// Stage op(X) restricted to rows [r0, r0+R) × k-range [k0, k0+kc) into LDS as S[kk*(R+PAD) + i] (k-major).
// trans = 0: X[i*ld + k] (k contiguous); trans = 1: X[k*ld + i] (i contiguous). Out-of-range → 0.
template <class T, int R, int PAD, int NT>
__device__ inline void stage_kmajor(T* S, const T* X, int ld, int trans, int r0, int rows, int k0, int kc, int klim, int tid) {
	const int total = R * kc;
	for (int base = 0; base < total; base += NT * 8) {
		T v[8];
		int li[8], lk[8];
#pragma unroll
		for (int u = 0; u < 8; ++u) {
			const int e = base + u * NT + tid;
			int i, kk;
			if (!trans) { i = e / kc; kk = e - i * kc; } else { kk = e / R; i = e - kk * R; }
			li[u] = i; lk[u] = kk;
			const int gi = r0 + i, gk = k0 + kk;
			v[u] = T(0);
			if (e < total && gi < rows && gk < klim)
				v[u] = trans ? X[(size_t)gk * ld + gi] : X[(size_t)gi * ld + gk];
		}
#pragma unroll
		for (int u = 0; u < 8; ++u) {
			const int e = base + u * NT + tid;
			if (e < total) S[lk[u] * (R + PAD) + li[u]] = v[u];
		}
	}
}

// One BM×BN output tile (bx, by) of product bz (batch entry or K slice). Called by the whole workgroup.
template <class T, int BM, int BN, int WM, int WN>
__device__ inline void gemm_x4_tile(const Args& g, int bx, int by, long long bz) {
	if (g.skip && *g.skip) return;
	constexpr int NT = 64 * WM * WN, TM = BM / WM / 16, TN = BN / WN / 16, PAD = 1;
	using Acc = typename AccOf<T>::type;
	extern __shared__ __align__(16) unsigned char smem_raw[];
	T* As = reinterpret_cast<T*>(smem_raw);
	T* Bs = As + (size_t)g.kc * (BM + PAD);
	const T* A = static_cast<const T*>(g.A);
	const T* B = static_cast<const T*>(g.B);
	T* C = static_cast<T*>(g.C) + bz * g.sC;
	int K = g.k;
	if (g.ksplit) {  // K slice bz of op(A)·op(B)
		const int kb = (int)bz * g.ksplit;
		K = min(g.ksplit, g.k - kb);
		A += g.ta ? (long long)kb * g.lda : (long long)kb;
		B += g.tb ? (long long)kb : (long long)kb * g.ldb;
	} else {
		A += bz * g.sA;
		B += bz * g.sB;
	}
	const int m0 = by * BM, n0 = bx * BN;
	const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
	const int wm = wave / WN, wn = wave % WN;
	const int li = lane & 15, lk = lane >> 4;
	Acc acc[TM][TN];
#pragma unroll
	for (int a = 0; a < TM; ++a)
#pragma unroll
		for (int b = 0; b < TN; ++b) acc[a][b] = Acc{0, 0, 0, 0};

	for (int k0 = 0; k0 < K; k0 += g.kc) {
		const int kc = min(g.kc, K - k0);
		const int kcp = (kc + 3) & ~3;  // MFMA consumes K in steps of 4; staging zero-fills the tail
		stage_kmajor<T, BM, PAD, NT>(As, A, g.lda, g.ta, m0, g.m, k0, kcp, K, tid);
		// B is staged as op(B)ᵀ rows: element (j, kk) = op(B)[kk][j]; tb = 0 means k-major in memory.
		stage_kmajor<T, BN, PAD, NT>(Bs, B, g.ldb, !g.tb, n0, g.n, k0, kcp, K, tid);
		__syncthreads();
		for (int kk = 0; kk < kcp; kk += 4) {
			T av[TM], bv[TN];
#pragma unroll
			for (int a = 0; a < TM; ++a) av[a] = As[(kk + lk) * (BM + PAD) + (wm * TM + a) * 16 + li];
#pragma unroll
			for (int b = 0; b < TN; ++b) bv[b] = Bs[(kk + lk) * (BN + PAD) + (wn * TN + b) * 16 + li];
#pragma unroll
			for (int a = 0; a < TM; ++a)
#pragma unroll
				for (int b = 0; b < TN; ++b) acc[a][b] = mfma_x4(av[a], bv[b], acc[a][b]);
		}
		__syncthreads();
	}
	const T alpha = T(g.alpha), beta = T(g.beta);
#pragma unroll
	for (int a = 0; a < TM; ++a)
#pragma unroll
		for (int b = 0; b < TN; ++b)
#pragma unroll
			for (int r = 0; r < 4; ++r) {
				const int i = m0 + (wm * TM + a) * 16 + RowOf<T>::row(lk, r);
				const int j = n0 + (wn * TN + b) * 16 + li;
				if (i < g.m && j < g.n) {
					T* c = C + (size_t)i * g.ldc + j;
					const T v = alpha * acc[a][b][r];
					*c = (g.beta != 0.0) ? v + beta * *c : v;
				}
			}
}

template <class T, int BM, int BN, int WM, int WN>
__global__ __launch_bounds__(64 * WM * WN) void gemm_x4_kernel(Args g) {
	gemm_x4_tile<T, BM, BN, WM, WN>(g, blockIdx.x, blockIdx.y, blockIdx.z);
}

// Two independent products in one launch. Workgroups [0, nb0) take problem 0's tiles, the rest problem 1's;
// each problem's (gx × gy × batch) grid is flattened with x fastest.
struct PairGrid {
	int gx[2], gy[2];
	int nb0;
};
template <class T, int BM, int BN, int WM, int WN>
__global__ __launch_bounds__(64 * WM * WN) void gemm_x4_pair_kernel(Args g0, Args g1, PairGrid pg) {
	const int b = blockIdx.x;
	const int p = b >= pg.nb0 ? 1 : 0;
	const int l = p ? b - pg.nb0 : b;
	const int gx = pg.gx[p], gy = pg.gy[p];
	const int bx = l % gx, t = l / gx;
	gemm_x4_tile<T, BM, BN, WM, WN>(p ? g1 : g0, bx, t % gy, t / gy);
}

// Many independent products in one launch (the lockstep batch driver): task t owns workgroups
// [start[t], start[t+1]); its (gx × gy × batch) grid is flattened with x fastest.
struct GroupArgs {
	Args a;
	int gx, gy;
};
// Grouped GEMM: 64×64 output tiles (each wave 32×32, four MFMA blocks per K step) and K chunks of at most 32, so a
// launch's LDS size (set by its largest task) stays at 2·65·32 elements and several workgroups share a compute
// unit. The MFMA order over K does not depend on either choice.
constexpr int kGroupTile = 64;
constexpr size_t kGroupKc = 32;
template <class T, int BT>
__global__ __launch_bounds__(256) void gemm_group_kernel(const GroupArgs* tasks, const int* start, int count) {
	const int t = group_task(start, count, (int)blockIdx.x);
	const GroupArgs g = tasks[t];
	const int l = (int)blockIdx.x - start[t];
	const int bx = l % g.gx, r = l / g.gx;
	gemm_x4_tile<T, BT, BT, 2, 2>(g.a, bx, r % g.gy, r / g.gy);
}

// Split-K epilogue: C = alpha·Σ_s P_s + beta·C, slices summed in index order (deterministic).
template <class T>
__global__ __launch_bounds__(256) void splitk_reduce_kernel(const T* P, int slices, int m, int n, T* C, int ldc, double alpha,
                                                           double beta, const int* skip) {
	if (skip && *skip) return;
	const size_t mn = (size_t)m * n;
	for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < mn; e += (size_t)gridDim.x * blockDim.x) {
		T acc = P[e];
		for (int q = 1; q < slices; ++q) acc += P[(size_t)q * mn + e];
		const size_t i = e / n, j = e - i * n;
		T* c = C + i * ldc + j;
		const T v = T(alpha) * acc;
		*c = beta != 0.0 ? v + T(beta) * *c : v;
	}
}

// BF16 compute: LDS holds bf16 row-major tiles As[i][k], Bs[j][k] (k contiguous, 8 elements per lane read).
template <int R, int NT>
__device__ inline void stage_rowk_bf16(__bf16* S, const float* X, int ld, int trans, int r0, int rows, int k0, int kc, int klim,
                                       int tid) {
	constexpr int KP = 8;  // row padding (elements) keeps 16-byte alignment and breaks bank aliasing
	const int total = R * kc;
	for (int base = 0; base < total; base += NT * 8) {
		float v[8];
		int li[8], lk[8];
#pragma unroll
		for (int u = 0; u < 8; ++u) {
			const int e = base + u * NT + tid;
			int i, kk;
			if (!trans) { i = e / kc; kk = e - i * kc; } else { kk = e / R; i = e - kk * R; }
			li[u] = i; lk[u] = kk;
			const int gi = r0 + i, gk = k0 + kk;
			v[u] = 0.f;
			if (e < total && gi < rows && gk < klim) v[u] = trans ? X[(size_t)gk * ld + gi] : X[(size_t)gi * ld + gk];
		}
#pragma unroll
		for (int u = 0; u < 8; ++u) {
			const int e = base + u * NT + tid;
			if (e < total) S[li[u] * (kc + KP) + lk[u]] = static_cast<__bf16>(v[u]);
		}
	}
}

template <int BM, int BN, int WM, int WN>
__global__ __launch_bounds__(64 * WM * WN) void gemm_bf16_kernel(Args g) {
	if (g.skip && *g.skip) return;
	constexpr int NT = 64 * WM * WN, TM = BM / WM / 16, TN = BN / WN / 16, KP = 8;
	extern __shared__ __align__(16) unsigned char smem_raw[];
	const long long bz = blockIdx.z;
	const float* A = static_cast<const float*>(g.A) + bz * g.sA;
	const float* B = static_cast<const float*>(g.B) + bz * g.sB;
	float* C = static_cast<float*>(g.C) + bz * g.sC;
	const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
	const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
	const int wm = wave / WN, wn = wave % WN;
	const int li = lane & 15, lk = lane >> 4;
	f4 acc[TM][TN];
#pragma unroll
	for (int a = 0; a < TM; ++a)
#pragma unroll
		for (int b = 0; b < TN; ++b) acc[a][b] = f4{0, 0, 0, 0};
	for (int k0 = 0; k0 < g.k; k0 += g.kc) {
		const int kc = min(g.kc, g.k - k0);
		const int kcp = (kc + 31) & ~31;
		__bf16* As = reinterpret_cast<__bf16*>(smem_raw);
		__bf16* Bs = As + (size_t)BM * (kcp + KP);
		stage_rowk_bf16<BM, NT>(As, A, g.lda, g.ta, m0, g.m, k0, kcp, g.k, tid);
		stage_rowk_bf16<BN, NT>(Bs, B, g.ldb, !g.tb, n0, g.n, k0, kcp, g.k, tid);
		__syncthreads();
		for (int kk = 0; kk < kcp; kk += 32) {
			bf8 av[TM], bv[TN];
#pragma unroll
			for (int a = 0; a < TM; ++a) av[a] = *reinterpret_cast<const bf8*>(As + ((wm * TM + a) * 16 + li) * (kcp + KP) + kk + 8 * lk);
#pragma unroll
			for (int b = 0; b < TN; ++b) bv[b] = *reinterpret_cast<const bf8*>(Bs + ((wn * TN + b) * 16 + li) * (kcp + KP) + kk + 8 * lk);
#pragma unroll
			for (int a = 0; a < TM; ++a)
#pragma unroll
				for (int b = 0; b < TN; ++b) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[a], bv[b], acc[a][b], 0, 0, 0);
		}
		__syncthreads();
	}
	const float alpha = float(g.alpha), beta = float(g.beta);
#pragma unroll
	for (int a = 0; a < TM; ++a)
#pragma unroll
		for (int b = 0; b < TN; ++b)
#pragma unroll
			for (int r = 0; r < 4; ++r) {
				const int i = m0 + (wm * TM + a) * 16 + 4 * lk + r;
				const int j = n0 + (wn * TN + b) * 16 + li;
				if (i < g.m && j < g.n) {
					float* c = C + (size_t)i * g.ldc + j;
					const float v = alpha * acc[a][b][r];
					*c = (g.beta != 0.0) ? v + beta * *c : v;
				}
			}
}

Args make_args(const GemmDesc& d, const void* A, const void* B, void* C) {
	XTT_REQUIRE(d.m <= INT32_MAX && d.n <= INT32_MAX && d.k <= INT32_MAX, "gemm: dimension too large");
	XTT_REQUIRE(d.batch >= 1 && d.batch <= 65535, "gemm: batch out of range");
	XTT_REQUIRE(d.lda >= (d.ta ? d.m : d.k) && d.ldb >= (d.tb ? d.k : d.n) && d.ldc >= d.n, "gemm: leading dimension too small");
	Args a{};
	a.A = A; a.B = B; a.C = C;
	a.m = (int)d.m; a.n = (int)d.n; a.k = (int)d.k;
	a.lda = (int)d.lda; a.ldb = (int)d.ldb; a.ldc = (int)d.ldc;
	a.sA = d.sA; a.sB = d.sB; a.sC = d.sC;
	a.ta = d.ta; a.tb = d.tb; a.alpha = d.alpha; a.beta = d.beta; a.skip = d.skip;
	return a;
}

// LDS budget per workgroup (gfx950 has 160 KiB per CU); kernels needing more than the 64 KiB default opt in.
constexpr size_t kLdsBudget = 150 * 1024;

// Raised once per instantiation to the full budget, so launches recorded into a hipGraph never need it.
template <class K>
void allow_lds(K kernel, size_t bytes) {
	static std::mutex mu;
	static std::set<const void*> raised;
	if (bytes <= 64 * 1024) return;
	const void* f = reinterpret_cast<const void*>(kernel);
	std::lock_guard<std::mutex> lk(mu);
	if (raised.count(f)) return;
	XTT_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLdsBudget));
	raised.insert(f);
}

// GEMM shape for profiler tags
std::string shape_suffix(const GemmDesc& d) {
	return " m=" + std::to_string(d.m) + " n=" + std::to_string(d.n) + " k=" + std::to_string(d.k) +
	       " batch=" + std::to_string(d.batch) + " t=" + std::to_string(d.ta) + std::to_string(d.tb);
}

// ---- tile autotuner: measured choices per shape override the heuristic below --------------------------------
struct TuneKey {
	int dtype;  // 0 FP64, 1 FP32
	size_t m, n, k, batch;
	int ta, tb;
	bool operator<(const TuneKey& o) const {
		return std::tie(dtype, m, n, k, batch, ta, tb) < std::tie(o.dtype, o.m, o.n, o.k, o.batch, o.ta, o.tb);
	}
};
std::mutex g_tuneMu;
std::map<TuneKey, GemmChoice> g_tune;
TuneKey tune_key(int dtype, const GemmDesc& d) { return TuneKey{dtype, d.m, d.n, d.k, d.batch, d.ta, d.tb}; }
bool tuned(int dtype, const GemmDesc& d, GemmChoice& c) {
	std::lock_guard<std::mutex> lk(g_tuneMu);
	if (g_tune.empty()) return false;
	auto it = g_tune.find(tune_key(dtype, d));
	if (it == g_tune.end()) return false;
	c = it->second;
	return true;
}

// The heuristic: 64×64 tiles once they still give >= 256 workgroups, else 32×32; split-K over
// max(2, 128 / tiles) slices of >= 64 when the 32×32 grid has fewer than 32 tiles and K >= 256 (batch 1).
GemmChoice heuristic_choice(const GemmDesc& d) {
	const size_t tiles64 = ((d.m + 63) / 64) * ((d.n + 63) / 64) * d.batch;
	if (tiles64 >= 256) return GemmChoice{64, 0};
	const size_t tiles32 = ((d.m + 31) / 32) * ((d.n + 31) / 32);
	if (d.batch == 1 && tiles32 < 32 && d.k >= 256) {
		const size_t want = std::max<size_t>(2, 128 / tiles32);
		return GemmChoice{32, (int)std::min<size_t>(d.k / 64, want)};
	}
	return GemmChoice{32, 0};
}

// One launch (or a split-K pair of launches) with the given choice. The K chunk is the largest multiple of 4
// (x4 MFMA) whose two operand tiles fit the LDS budget.
template <class T>
void launch_x4_with(const GemmDesc& d, const T* A, const T* B, T* C, hipStream_t s, GemmChoice ch) {
	if (d.m == 0 || d.n == 0) return;
	Args a = make_args(d, A, B, C);
	const bool big = ch.tile == 64;
	const size_t BMp = big ? 65 : 33;
	size_t kmax = kLdsBudget / (2 * BMp * sizeof(T));
	kmax &= ~size_t(3);
	a.kc = (int)std::max<size_t>(4, std::min<size_t>((d.k + 3) & ~size_t(3), kmax));
	const size_t lds = (size_t)a.kc * BMp * 2 * sizeof(T);
	const double flops = 2.0 * d.m * d.n * d.k * d.batch;
	const Work work(flops, (double)sizeof(T) * d.batch * ((double)d.m * d.k + (double)d.k * d.n + (d.beta != 0.0 ? 2.0 : 1.0) * d.m * d.n));
	std::string tag = big ? (sizeof(T) == 8 ? "gemm_x4_kernel<double,64,64>" : "gemm_x4_kernel<float,64,64>")
	                      : (sizeof(T) == 8 ? "gemm_x4_kernel<double,32,32>" : "gemm_x4_kernel<float,32,32>");
	if (prof_shapes()) tag += shape_suffix(d);
	// Split-K for few output tiles over a long K (the stack contractions Xᵀ·T over a·n rows): K slices
	// run as blockIdx.z, partials go to a stream-ordered scratch buffer and one epilogue launch sums them.
	// While the stream is being captured only with a capture arena (its scratch then belongs to the graph); without
	// one the scratch would be returned to the cache while the graph still uses it.
	const bool split = !big && d.batch == 1 && ch.slices >= 2 && d.k >= 8;
	hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
	if (split && !arena_active()) XTT_HIP(hipStreamIsCapturing(s, &cap));
	if (split && cap == hipStreamCaptureStatusNone) {
		const size_t slices0 = std::min<size_t>((size_t)ch.slices, d.k / 4);
		const size_t ks = (((d.k + slices0 - 1) / slices0) + 3) & ~size_t(3);
		const size_t slices = (d.k + ks - 1) / ks;
		DBuf<T> part(slices * d.m * d.n);
		Args p = a;
		p.C = part.get();
		p.ldc = (int)d.n;
		p.sC = (long long)(d.m * d.n);
		p.alpha = 1.0;
		p.beta = 0.0;
		p.ksplit = (int)ks;
		p.kc = (int)std::min<size_t>(ks, (size_t)a.kc);
		const size_t plds = (size_t)p.kc * BMp * 2 * sizeof(T);
		auto kern = gemm_x4_kernel<T, 32, 32, 2, 2>;
		allow_lds(kern, plds);
		dim3 grid((unsigned)((d.n + 31) / 32), (unsigned)((d.m + 31) / 32), (unsigned)slices);
		launch_timed((tag + " splitk").c_str(), work, kern, grid, dim3(256), plds, s, p);
		const size_t mn = d.m * d.n;
		const unsigned rb = (unsigned)std::min<size_t>(1024, (mn + 255) / 256);
		hipLaunchKernelGGL(splitk_reduce_kernel<T>, dim3(rb), dim3(256), 0, s, (const T*)part.get(), (int)slices, (int)d.m,
		                   (int)d.n, C, (int)d.ldc, d.alpha, d.beta, d.skip);
		XTT_LAUNCH_CHECK();
		return;
	}
	if (big) {
		auto kern = gemm_x4_kernel<T, 64, 64, 2, 2>;
		allow_lds(kern, lds);
		dim3 grid((unsigned)((d.n + 63) / 64), (unsigned)((d.m + 63) / 64), (unsigned)d.batch);
		launch_timed(tag.c_str(), work, kern, grid, dim3(256), lds, s, a);
	} else {
		auto kern = gemm_x4_kernel<T, 32, 32, 2, 2>;
		allow_lds(kern, lds);
		dim3 grid((unsigned)((d.n + 31) / 32), (unsigned)((d.m + 31) / 32), (unsigned)d.batch);
		launch_timed(tag.c_str(), work, kern, grid, dim3(256), lds, s, a);
	}
}

template <class T>
void launch_x4(const GemmDesc& d, const T* A, const T* B, T* C, hipStream_t s) {
	GemmChoice ch;
	if (!tuned(sizeof(T) == 8 ? 0 : 1, d, ch)) ch = heuristic_choice(d);
	launch_x4_with<T>(d, A, B, C, s, ch);
}

// Two products with small tile grids (no split-K) as one launch of 32×32 tiles; anything else as two launches.
template <class T>
void launch_x4_pair(const GemmDesc& d0, const T* A0, const T* B0, T* C0, const GemmDesc& d1, const T* A1, const T* B1, T* C1,
                    hipStream_t s) {
	auto small = [](const GemmDesc& d) { return ((d.m + 63) / 64) * ((d.n + 63) / 64) * d.batch < 256; };
	if (d0.m == 0 || d0.n == 0 || d1.m == 0 || d1.n == 0 || !small(d0) || !small(d1)) {
		launch_x4<T>(d0, A0, B0, C0, s);
		launch_x4<T>(d1, A1, B1, C1, s);
		return;
	}
	constexpr size_t BMp = 33;
	size_t kmax = kLdsBudget / (2 * BMp * sizeof(T));
	kmax &= ~size_t(3);
	Args a[2] = {make_args(d0, A0, B0, C0), make_args(d1, A1, B1, C1)};
	const GemmDesc* d[2] = {&d0, &d1};
	PairGrid pg;
	size_t blocks[2], lds = 0;
	double flops = 0, bytes = 0;
	for (int p = 0; p < 2; ++p) {
		const GemmDesc& g = *d[p];
		a[p].kc = (int)std::max<size_t>(4, std::min<size_t>((g.k + 3) & ~size_t(3), kmax));
		lds = std::max(lds, (size_t)a[p].kc * BMp * 2 * sizeof(T));
		pg.gx[p] = (int)((g.n + 31) / 32);
		pg.gy[p] = (int)((g.m + 31) / 32);
		blocks[p] = (size_t)pg.gx[p] * pg.gy[p] * g.batch;
		flops += 2.0 * g.m * g.n * g.k * g.batch;
		bytes += (double)sizeof(T) * g.batch * ((double)g.m * g.k + (double)g.k * g.n + (g.beta != 0.0 ? 2.0 : 1.0) * g.m * g.n);
	}
	pg.nb0 = (int)blocks[0];
	std::string tag = sizeof(T) == 8 ? "gemm_x4_pair_kernel<double,32,32>" : "gemm_x4_pair_kernel<float,32,32>";
	if (prof_shapes()) tag += shape_suffix(d0) + " +" + shape_suffix(d1);
	auto kern = gemm_x4_pair_kernel<T, 32, 32, 2, 2>;
	allow_lds(kern, lds);
	launch_timed(tag.c_str(), Work(flops, bytes), kern, dim3((unsigned)(blocks[0] + blocks[1])), dim3(256), lds, s, a[0], a[1], pg);
}

template <class T>
void launch_group(const std::vector<GemmTask>& tasks, hipStream_t s, GemmGroupPlan* keep = nullptr) {
	constexpr int BT = kGroupTile;
	constexpr size_t BMp = BT + 1;
	const size_t kmax = std::min<size_t>(kGroupKc, (kLdsBudget / (2 * BMp * sizeof(T))) & ~size_t(3));
	std::vector<GroupArgs> ga;
	std::vector<int> start;
	ga.reserve(tasks.size());
	start.reserve(tasks.size());
	int blocks = 0;
	size_t lds = 0;
	double flops = 0, bytes = 0;
	for (const GemmTask& t : tasks) {
		const GemmDesc& d = t.d;
		if (d.m == 0 || d.n == 0 || d.batch == 0) continue;
		GroupArgs g;
		g.a = make_args(d, t.A, t.B, t.C);
		g.a.kc = (int)std::max<size_t>(4, std::min<size_t>((d.k + 3) & ~size_t(3), kmax));
		g.gx = (int)((d.n + BT - 1) / BT);
		g.gy = (int)((d.m + BT - 1) / BT);
		lds = std::max(lds, (size_t)g.a.kc * BMp * 2 * sizeof(T));
		start.push_back(blocks);
		blocks += g.gx * g.gy * (int)d.batch;
		ga.push_back(g);
		flops += 2.0 * d.m * d.n * d.k * d.batch;
		bytes += (double)sizeof(T) * d.batch * ((double)d.m * d.k + (double)d.k * d.n + (d.beta != 0.0 ? 2.0 : 1.0) * d.m * d.n);
	}
	if (ga.empty()) {
		if (keep) keep->count = 0;
		return;
	}
	const GroupArgs* dt;
	const int* ds;
	if (keep) {  // tables in a buffer of the plan: replayed by gemm_group_run_f64 until the plan goes
		const size_t tb = ga.size() * sizeof(GroupArgs), sb = start.size() * sizeof(int);
		keep->buf.resize((tb + 255) / 256 * 256 + sb);
		char* base = keep->buf.get();
		table_upload_to(base, ga.data(), tb, s);
		table_upload_to(base + (tb + 255) / 256 * 256, start.data(), sb, s);
		dt = reinterpret_cast<const GroupArgs*>(base);
		ds = reinterpret_cast<const int*>(base + (tb + 255) / 256 * 256);
		keep->tasks = dt; keep->start = ds; keep->count = (int)ga.size(); keep->blocks = blocks; keep->lds = lds;
		keep->flops = flops; keep->bytes = bytes;
	} else {
		dt = static_cast<const GroupArgs*>(table_upload(ga.data(), ga.size() * sizeof(GroupArgs), s));
		ds = static_cast<const int*>(table_upload(start.data(), start.size() * sizeof(int), s));
	}
	auto kern = gemm_group_kernel<T, kGroupTile>;
	allow_lds(kern, lds);
	launch_timed(sizeof(T) == 8 ? "gemm_group_kernel<double>" : "gemm_group_kernel<float>", Work(flops, bytes), kern, dim3((unsigned)blocks),
	             dim3(256), lds, s, dt, ds, (int)ga.size());
}

}  // namespace

void gemm_group_f64(const std::vector<GemmTask>& tasks, hipStream_t s) { launch_group<double>(tasks, s); }
void gemm_group_plan_f64(const std::vector<GemmTask>& tasks, GemmGroupPlan& plan, hipStream_t s) { launch_group<double>(tasks, s, &plan); }
void gemm_group_f32(const std::vector<GemmTask>& tasks, hipStream_t s) { launch_group<float>(tasks, s); }
void gemm_group_plan_f32(const std::vector<GemmTask>& tasks, GemmGroupPlan& plan, hipStream_t s) { launch_group<float>(tasks, s, &plan); }
namespace {
template <class T>
void run_group(const GemmGroupPlan& p, hipStream_t s) {
	if (p.count == 0) return;
	auto kern = gemm_group_kernel<T, kGroupTile>;
	const char* tag = p.tag ? p.tag : (sizeof(T) == 8 ? "gemm_group_kernel<double>" : "gemm_group_kernel<float>");
	launch_timed(tag, Work(p.flops, p.bytes), kern, dim3((unsigned)p.blocks), dim3(256), p.lds, s,
	             static_cast<const GroupArgs*>(p.tasks), p.start, p.count);
}
}  // namespace
void gemm_group_run_f64(const GemmGroupPlan& p, hipStream_t s) { run_group<double>(p, s); }
void gemm_group_run_f32(const GemmGroupPlan& p, hipStream_t s) { run_group<float>(p, s); }

void gemm_f64_pair(const GemmDesc& g0, const double* A0, const double* B0, double* C0, const GemmDesc& g1, const double* A1,
                   const double* B1, double* C1, hipStream_t s) {
	launch_x4_pair<double>(g0, A0, B0, C0, g1, A1, B1, C1, s);
}

GemmDesc gemm_desc(int ta, int tb, size_t m, size_t n, size_t k) {
	GemmDesc d;
	d.ta = ta; d.tb = tb; d.m = m; d.n = n; d.k = k;
	d.lda = ta ? m : k; d.ldb = tb ? k : n; d.ldc = n;
	d.sA = (long long)(m * k); d.sB = (long long)(k * n); d.sC = (long long)(m * n);
	return d;
}

void gemm_f64(const GemmDesc& g, const double* A, const double* B, double* C, hipStream_t s) { launch_x4<double>(g, A, B, C, s); }

void gemm_tune_set(int dtype, const GemmDesc& d, GemmChoice c) {
	XTT_REQUIRE(dtype == 0 || dtype == 1, "gemm tuning: FP64 (0) or FP32 (1)");
	XTT_REQUIRE(c.tile == 32 || c.tile == 64, "gemm tuning: tile 32 or 64");
	XTT_REQUIRE(c.slices == 0 || (c.tile == 32 && c.slices >= 2 && c.slices <= 256), "gemm tuning: split-K needs 32-tiles and 2..256 slices");
	std::lock_guard<std::mutex> lk(g_tuneMu);
	g_tune[tune_key(dtype, d)] = c;
}
void gemm_tune_clear() {
	std::lock_guard<std::mutex> lk(g_tuneMu);
	g_tune.clear();
}
GemmChoice gemm_heuristic(const GemmDesc& d) { return heuristic_choice(d); }

namespace {
__global__ void tune_fill_kernel(double* p, size_t n) {  // deterministic operand values in [-1, 1)
	for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
		p[i] = (double)((i * 2654435761u) % 1000) * 0.002 - 1.0;
}
template <class T>
double time_choice(const GemmDesc& d, const T* A, const T* B, T* C, hipStream_t s, GemmChoice ch, int iters) {
	for (int i = 0; i < 3; ++i) launch_x4_with<T>(d, A, B, C, s, ch);
	hipEvent_t e0, e1;
	XTT_HIP(hipEventCreate(&e0));
	XTT_HIP(hipEventCreate(&e1));
	XTT_HIP(hipEventRecord(e0, s));
	for (int i = 0; i < iters; ++i) launch_x4_with<T>(d, A, B, C, s, ch);
	XTT_HIP(hipEventRecord(e1, s));
	XTT_HIP(hipEventSynchronize(e1));
	float ms = 0;
	XTT_HIP(hipEventElapsedTime(&ms, e0, e1));
	(void)hipEventDestroy(e0);
	(void)hipEventDestroy(e1);
	return 1e3 * ms / iters;
}
template <class T>
GemmTuneResult tune_t(const GemmDesc& d, int iters, hipStream_t s) {
	const size_t na = d.batch * (size_t)d.m * d.k, nb = d.batch * (size_t)d.k * d.n, nc = d.batch * (size_t)d.m * d.n;
	// FP64 scratch of the largest operand, reinterpreted for FP32 (the values only feed the timing)
	DBuf<double> a(na), b(nb), c(nc);
	for (DBuf<double>* p : {&a, &b, &c}) {
		hipLaunchKernelGGL(tune_fill_kernel, dim3(256), dim3(256), 0, s, p->get(), p->size());
		XTT_LAUNCH_CHECK();
	}
	const T* A = reinterpret_cast<const T*>(a.get());
	const T* B = reinterpret_cast<const T*>(b.get());
	T* C = reinterpret_cast<T*>(c.get());
	std::vector<GemmChoice> cand;
	if (((d.m + 63) / 64) * ((d.n + 63) / 64) * d.batch >= 4) cand.push_back({64, 0});
	cand.push_back({32, 0});
	if (d.batch == 1 && d.k >= 128)
		for (int sl = 2; sl <= 64 && (size_t)sl * 32 <= d.k; sl *= 2) cand.push_back({32, sl});
	GemmTuneResult r;
	r.heuristic = heuristic_choice(d);
	r.heuristicUs = time_choice<T>(d, A, B, C, s, r.heuristic, iters);
	r.best = r.heuristic;
	r.bestUs = r.heuristicUs;
	for (const GemmChoice& ch : cand) {
		if (ch.tile == r.heuristic.tile && ch.slices == r.heuristic.slices) continue;
		const double us = time_choice<T>(d, A, B, C, s, ch, iters);
		if (us < r.bestUs) { r.bestUs = us; r.best = ch; }
	}
	return r;
}
}  // namespace

GemmTuneResult gemm_tune(int dtype, const GemmDesc& d, int iters, hipStream_t s) {
	XTT_REQUIRE(dtype == 0 || dtype == 1, "gemm tuning: FP64 (0) or FP32 (1)");
	XTT_REQUIRE(d.m > 0 && d.n > 0 && d.k > 0 && d.batch > 0, "gemm tuning: empty shape");
	XTT_REQUIRE(iters >= 1 && iters <= 10000, "gemm tuning: 1..10000 timed launches");
	const GemmTuneResult r = dtype == 0 ? tune_t<double>(d, iters, s) : tune_t<float>(d, iters, s);
	gemm_tune_set(dtype, d, r.best);
	return r;
}
void gemm_f32(const GemmDesc& g, const float* A, const float* B, float* C, hipStream_t s) { launch_x4<float>(g, A, B, C, s); }

void gemm_bf16(const GemmDesc& d, const float* A, const float* B, float* C, hipStream_t s) {
	if (d.m == 0 || d.n == 0) return;
	Args a = make_args(d, A, B, C);
	const size_t tiles64 = ((d.m + 63) / 64) * ((d.n + 63) / 64) * d.batch;
	const bool big = tiles64 >= 256;
	const size_t rows = big ? 128 : 64;
	size_t kmax = kLdsBudget / (rows * sizeof(__bf16)) - 8;
	kmax &= ~size_t(31);
	a.kc = (int)std::max<size_t>(32, std::min<size_t>((d.k + 31) & ~size_t(31), kmax));
	const size_t lds = rows * (size_t)(a.kc + 8) * sizeof(__bf16);
	const Work work(2.0 * d.m * d.n * d.k * d.batch,
	                4.0 * d.batch * ((double)d.m * d.k + (double)d.k * d.n + (d.beta != 0.0 ? 2.0 : 1.0) * d.m * d.n));
	if (big) {
		auto kern = gemm_bf16_kernel<64, 64, 2, 2>;
		allow_lds(kern, lds);
		dim3 grid((unsigned)((d.n + 63) / 64), (unsigned)((d.m + 63) / 64), (unsigned)d.batch);
		launch_timed("gemm_bf16_kernel<64,64>", work, kern, grid, dim3(256), lds, s, a);
	} else {
		auto kern = gemm_bf16_kernel<32, 32, 2, 2>;
		allow_lds(kern, lds);
		dim3 grid((unsigned)((d.n + 31) / 32), (unsigned)((d.m + 31) / 32), (unsigned)d.batch);
		launch_timed("gemm_bf16_kernel<32,32>", work, kern, grid, dim3(256), lds, s, a);
	}
}

}  // namespace xtt
