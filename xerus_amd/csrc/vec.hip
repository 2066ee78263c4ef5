// Element-wise, permutation and reduction kernels of libxtt, plus the device-resident CG step kernels.
//
// Reductions are deterministic: a fixed grid writes one partial per block, and the last block to finish
// (ticket from an agent-scope atomic, after an agent-scope fence) sums the partials in index order. No
// float atomics, so repeated runs are bit-identical.
#include <algorithm>

#include "kernels.hpp"
#include "xtt_common.hpp"

namespace xtt {
namespace {

constexpr int RB = 256;  // threads per reduction block

struct PermArgs {
	int deg;
	long long n;
	long long inStride[8];   // stride of old mode i in the input
	long long outStride[8];  // stride (in the output) of old mode i
	long long dims[8];
};

__global__ void permute_kernel(PermArgs a, const double* __restrict__ in, double* __restrict__ out) {
	for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < a.n; e += (long long)gridDim.x * blockDim.x) {
		long long rem = e, o = 0;
		for (int i = a.deg - 1; i >= 0; --i) {
			const long long idx = rem % a.dims[i];
			rem /= a.dims[i];
			o += idx * a.outStride[i];
		}
		out[o] = in[e];
	}
}

__global__ void convert_kernel_64_32(const double* in, float* out, size_t n) {
	for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) out[i] = (float)in[i];
}
__global__ void convert_kernel_32_64(const float* in, double* out, size_t n) {
	for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) out[i] = (double)in[i];
}
// Up to kMultiCopy device-to-device copies in one launch: blockIdx.y picks the copy, the x-blocks stride over it.
__global__ void multi_copy_kernel(MultiCopy c) {
	const unsigned j = blockIdx.y;
	const double* __restrict__ in = c.src[j];
	double* __restrict__ out = c.dst[j];
	const size_t n = c.n[j];
	for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) out[i] = in[i];
}
// The same for FP64 → FP32 conversions (dst entries are float*)
__global__ void multi_convert_kernel(MultiCopy c) {
	const unsigned j = blockIdx.y;
	const double* __restrict__ in = c.src[j];
	float* __restrict__ out = reinterpret_cast<float*>(c.dst[j]);
	const size_t n = c.n[j];
	for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) out[i] = (float)in[i];
}
__global__ void fill_kernel(double* p, double v, size_t n) {
	for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void scale_kernel(double* p, double a, size_t n) {
	for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] *= a;
}
// M (rows×cols, row-major) scaled by a device vector: along rows (M[r][c] *= s[r]) or columns (*= s[c]).
__global__ void scale_by_kernel(double* m, const double* sv, size_t rows, size_t cols, int byRow) {
	const size_t n = rows * cols;
	for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
		m[e] *= sv[byRow ? e / cols : e % cols];
}
// calculate_svd's truncation rank on the device: min(k, maxRank), cut at the first j with s_j <= eps·s_0
__global__ void svd_rank_kernel(const double* sv, size_t k, size_t maxRank, double eps, int* out) {
	if (threadIdx.x != 0 || blockIdx.x != 0) return;
	size_t rank = k;
	if (maxRank != 0 && maxRank < rank) rank = maxRank;
	for (size_t j = 1; j < rank; ++j)
		if (sv[j] <= eps * sv[0]) { rank = j; break; }
	*out = sv[0] != sv[0] ? 0 : (int)rank;  // an unconverged SVD (NaN values) matches no speculated rank
}
__global__ void axpy_kernel(double* y, double a, const double* x, size_t n) {
	for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) y[i] += a * x[i];
}

// out (cols×rows) = J·inᵀ·J for a rows×cols input: out[i][j] = in[rows-1-j][cols-1-i].
__global__ void revtrans_kernel(const double* in, size_t rows, size_t cols, double* out) {
	const size_t n = rows * cols;
	for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
		const size_t i = e / rows, j = e % rows;
		out[e] = in[(rows - 1 - j) * cols + (cols - 1 - i)];
	}
}

unsigned grid_for(size_t n, unsigned block = 256) {
	size_t g = (n + block - 1) / block;
	if (g > 2048) g = 2048;
	if (g == 0) g = 1;
	return (unsigned)g;
}

// ---- block reductions (two running sums) ---------------------------------------------------------------

__device__ inline double wave_sum(double v) {
#pragma unroll
	for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
	return v;
}

// Sum over the block; result valid in every thread. `ws` needs RB/64 * 2 doubles.
__device__ inline void block_sum2(double& a, double& b, double* ws) {
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	a = wave_sum(a);
	b = wave_sum(b);
	__syncthreads();
	if (lane == 0) { ws[2 * wave] = a; ws[2 * wave + 1] = b; }
	__syncthreads();
	a = 0; b = 0;
	for (int w = 0; w < RB / 64; ++w) { a += ws[2 * w]; b += ws[2 * w + 1]; }
	__syncthreads();
}

// Publishes this block's (a, b) into partials[2*block..] and returns true in every thread of the block that
// finishes last; that block then holds the in-order totals in (a, b).
__device__ inline bool last_block_sum2(double& a, double& b, double* partials, unsigned* counter, double* ws) {
	__shared__ int amLast;
	block_sum2(a, b, ws);
	if (threadIdx.x == 0) {
		partials[2 * blockIdx.x] = a;
		partials[2 * blockIdx.x + 1] = b;
		__threadfence();
		const unsigned t = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
		amLast = (t == gridDim.x - 1);
	}
	__syncthreads();
	if (!amLast) return false;
	__threadfence();
	a = 0; b = 0;
	for (unsigned i = threadIdx.x; i < gridDim.x; i += RB) {
		a += __hip_atomic_load(partials + 2 * i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		b += __hip_atomic_load(partials + 2 * i + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
	}
	block_sum2(a, b, ws);
	if (threadIdx.x == 0) *counter = 0;
	return true;
}

template <class T>
__global__ __launch_bounds__(RB) void dot_kernel(const T* x, const T* y, size_t n, double* out, double* partials, unsigned* counter) {
	__shared__ double ws[RB / 64 * 2];
	double a = 0, b = 0;
	for (size_t i = blockIdx.x * (size_t)RB + threadIdx.x; i < n; i += (size_t)gridDim.x * RB) a += (double)x[i] * (double)y[i];
	if (last_block_sum2(a, b, partials, counter, ws) && threadIdx.x == 0) *out = a;
}

// out[0] = Σ|x_i|, out[1] = max|x_i| (block partials reduced in block order by the last block)
__global__ __launch_bounds__(RB) void asum_amax_kernel(const double* x, size_t n, double* out, double* partials, unsigned* counter) {
	__shared__ double ws[RB / 64 * 2];
	__shared__ int amLast;
	double a = 0, m = 0;
	for (size_t i = blockIdx.x * (size_t)RB + threadIdx.x; i < n; i += (size_t)gridDim.x * RB) {
		const double v = fabs(x[i]);
		a += v;
		m = fmax(m, v);
	}
	for (int o = 32; o > 0; o >>= 1) { a += __shfl_xor(a, o); m = fmax(m, __shfl_xor(m, o)); }
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	if (lane == 0) { ws[2 * wave] = a; ws[2 * wave + 1] = m; }
	__syncthreads();
	if (threadIdx.x == 0) {
		a = 0; m = 0;
		for (int w = 0; w < RB / 64; ++w) { a += ws[2 * w]; m = fmax(m, ws[2 * w + 1]); }
		partials[2 * blockIdx.x] = a;
		partials[2 * blockIdx.x + 1] = m;
		__threadfence();
		const unsigned t = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
		amLast = (t == gridDim.x - 1);
	}
	__syncthreads();
	if (!amLast || threadIdx.x != 0) return;
	__threadfence();
	a = 0; m = 0;
	for (unsigned b = 0; b < gridDim.x; ++b) {
		a += __hip_atomic_load(partials + 2 * b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		m = fmax(m, __hip_atomic_load(partials + 2 * b + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
	}
	out[0] = a;
	out[1] = m;
	*counter = 0;
}
// out[2] = out[1]·√(Σ (x_i/out[1])²): the two norm without overflow or underflow of the squares (dnrm2's scaling)
__global__ __launch_bounds__(RB) void scaled_nrm2_kernel(const double* x, size_t n, double* out, double* partials, unsigned* counter) {
	__shared__ double ws[RB / 64 * 2];
	const double scale = out[1];
	double a = 0, b = 0;
	if (scale > 0)
		for (size_t i = blockIdx.x * (size_t)RB + threadIdx.x; i < n; i += (size_t)gridDim.x * RB) { const double v = x[i] / scale; a = fma(v, v, a); }
	if (last_block_sum2(a, b, partials, counter, ws) && threadIdx.x == 0) out[2] = scale > 0 ? scale * sqrt(a) : 0.0;
}

// r = b − q with out[0] = ‖r‖², out[1] = ‖b‖² (one pass, deterministic)
__global__ __launch_bounds__(RB) void residual_norms_kernel(double* r, const double* b, const double* q, size_t n, double* out,
                                                            double* partials, unsigned* counter, double tol2, unsigned* fail) {
	__shared__ double ws[RB / 64 * 2];
	double rr = 0, bb = 0;
	for (size_t i = blockIdx.x * (size_t)RB + threadIdx.x; i < n; i += (size_t)gridDim.x * RB) {
		const double ri = b[i] - q[i];
		r[i] = ri;
		rr = fma(ri, ri, rr);
		bb = fma(b[i], b[i], bb);
	}
	if (last_block_sum2(rr, bb, partials, counter, ws) && threadIdx.x == 0) {
		out[0] = rr;
		out[1] = bb;
		if (fail && !(rr <= tol2 * bb)) atomicAdd(fail, 1u);  // NaN counts as a failure
	}
}

// ---- CG steps ------------------------------------------------------------------------------------------

template <class T>
__global__ __launch_bounds__(RB) void cg_start_kernel(T* r, T* p, const T* b, const T* q, size_t n, CGState* st, double* partials) {
	__shared__ double ws[RB / 64 * 2];
	double rr = 0, bb = 0;
	for (size_t i = blockIdx.x * (size_t)RB + threadIdx.x; i < n; i += (size_t)gridDim.x * RB) {
		const T ri = b[i] - q[i];
		r[i] = ri;
		p[i] = ri;
		rr += (double)ri * (double)ri;
		bb += (double)b[i] * (double)b[i];
	}
	if (last_block_sum2(rr, bb, partials, &st->counter, ws) && threadIdx.x == 0) {
		st->rr = rr;
		st->rr0 = bb;
		st->iter = 0;
		st->done = (rr <= st->tol2 * bb || st->maxIter <= 0) ? 1 : 0;
	}
}

template <class T>
__global__ __launch_bounds__(RB) void cg_alpha_kernel(const T* p, const T* q, size_t n, CGState* st, double* partials) {
	if (st->done) return;
	__shared__ double ws[RB / 64 * 2];
	double pq = 0, z = 0;
	for (size_t i = blockIdx.x * (size_t)RB + threadIdx.x; i < n; i += (size_t)gridDim.x * RB) pq += (double)p[i] * (double)q[i];
	if (last_block_sum2(pq, z, partials, &st->counter, ws) && threadIdx.x == 0) {
		st->pq = pq;
		if (pq > 0) {
			st->alpha = st->rr / pq;
		} else {  // breakdown (non-positive curvature or exact solution): stop where we are
			st->alpha = 0;
			st->done = 1;
		}
	}
}

template <class T>
__global__ __launch_bounds__(RB) void cg_update_kernel(T* x, T* r, const T* p, const T* q, size_t n, CGState* st, double* partials) {
	if (st->done) return;
	__shared__ double ws[RB / 64 * 2];
	const T alpha = (T)st->alpha;
	double rr = 0, z = 0;
	for (size_t i = blockIdx.x * (size_t)RB + threadIdx.x; i < n; i += (size_t)gridDim.x * RB) {
		x[i] += alpha * p[i];
		const T ri = r[i] - alpha * q[i];
		r[i] = ri;
		rr += (double)ri * (double)ri;
	}
	if (last_block_sum2(rr, z, partials, &st->counter, ws) && threadIdx.x == 0) {
		st->beta = rr / st->rr;
		st->rr = rr;
		st->iter += 1;
		if (rr <= st->tol2 * st->rr0 || st->iter >= st->maxIter) st->done = 1;
	}
}

template <class T>
__global__ __launch_bounds__(RB) void cg_direction_kernel(T* p, const T* r, size_t n, CGState* st) {
	if (st->done) return;
	const T beta = (T)st->beta;
	for (size_t i = blockIdx.x * (size_t)RB + threadIdx.x; i < n; i += (size_t)gridDim.x * RB) p[i] = r[i] + beta * p[i];
}

__global__ void cg_zero_start_kernel(double* x, double* r, double* p, const double* b, size_t n, CGState* st) {
	for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
		x[i] = 0.0;
		r[i] = b[i];
		p[i] = b[i];
	}
	if (blockIdx.x == 0 && threadIdx.x == 0) st->rr = st->rr0;
}
__global__ void cg_state_init_kernel(CGState* st, double tol2, int maxIter) {
	if (threadIdx.x == 0 && blockIdx.x == 0) {
		st->rr = 0; st->rr0 = 0; st->tol2 = tol2; st->alpha = 0; st->beta = 0; st->pq = 0;
		st->done = 0; st->iter = 0; st->maxIter = maxIter; st->pad = 0; st->counter = 0; st->pad2 = 0;
	}
}

}  // namespace

void cg_zero_start(double* x, double* r, double* p, const double* b, size_t n, CGState* st, hipStream_t s) {
	if (!n) return;
	hipLaunchKernelGGL(cg_zero_start_kernel, dim3(grid_for(n)), dim3(256), 0, s, x, r, p, b, n, st);
	XTT_LAUNCH_CHECK();
}

size_t reduce_blocks(size_t n) {
	size_t g = (n + RB * 4 - 1) / (RB * 4);
	if (g > 512) g = 512;
	if (g == 0) g = 1;
	return g;
}

void permute_f64(size_t degree, const size_t* dims, const size_t* shuffle, const double* in, double* out, hipStream_t s) {
	XTT_REQUIRE(degree <= 8, "permute: at most 8 modes");
	PermArgs a{};
	a.deg = (int)degree;
	long long n = 1;
	for (size_t i = 0; i < degree; ++i) { a.dims[i] = (long long)dims[i]; n *= (long long)dims[i]; }
	a.n = n;
	if (n == 0) return;
	long long newDims[8];
	for (size_t i = 0; i < degree; ++i) { XTT_REQUIRE(shuffle[i] < degree, "permute: bad shuffle"); newDims[shuffle[i]] = (long long)dims[i]; }
	long long newStride[8];
	long long st = 1;
	for (size_t j = degree; j-- > 0;) { newStride[j] = st; st *= newDims[j]; }
	for (size_t i = 0; i < degree; ++i) a.outStride[i] = newStride[shuffle[i]];
	hipLaunchKernelGGL(permute_kernel, dim3(grid_for((size_t)n)), dim3(256), 0, s, a, in, out);
	XTT_LAUNCH_CHECK();
}

void revtrans_f64(const double* in, size_t rows, size_t cols, double* out, hipStream_t s) {
	if (!rows || !cols) return;
	hipLaunchKernelGGL(revtrans_kernel, dim3(grid_for(rows * cols)), dim3(256), 0, s, in, rows, cols, out);
	XTT_LAUNCH_CHECK();
}

void convert_f64_f32(const double* in, float* out, size_t n, hipStream_t s) {
	if (!n) return;
	hipLaunchKernelGGL(convert_kernel_64_32, dim3(grid_for(n)), dim3(256), 0, s, in, out, n);
	XTT_LAUNCH_CHECK();
}
void convert_f32_f64(const float* in, double* out, size_t n, hipStream_t s) {
	if (!n) return;
	hipLaunchKernelGGL(convert_kernel_32_64, dim3(grid_for(n)), dim3(256), 0, s, in, out, n);
	XTT_LAUNCH_CHECK();
}
void multi_copy_f64(const MultiCopy& c, hipStream_t s) {
	if (c.count <= 0) return;
	if (c.count > kMultiCopy) throw Error(ERR_INVALID, "multi_copy: too many copies in one launch");
	size_t most = 0;
	for (int j = 0; j < c.count; ++j) most = std::max(most, c.n[j]);
	if (!most) return;
	hipLaunchKernelGGL(multi_copy_kernel, dim3(std::min<size_t>(grid_for(most), 64), c.count), dim3(256), 0, s, c);
	XTT_LAUNCH_CHECK();
}
void multi_convert_f64_f32(const MultiCopy& c, hipStream_t s) {
	if (c.count <= 0) return;
	if (c.count > kMultiCopy) throw Error(ERR_INVALID, "multi_convert: too many conversions in one launch");
	size_t most = 0;
	for (int j = 0; j < c.count; ++j) most = std::max(most, c.n[j]);
	if (!most) return;
	hipLaunchKernelGGL(multi_convert_kernel, dim3(std::min<size_t>(grid_for(most), 64), c.count), dim3(256), 0, s, c);
	XTT_LAUNCH_CHECK();
}
namespace {
__global__ void zero_u32_kernel(unsigned* p, size_t n) {
	for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0u;
}
}  // namespace
namespace {
__global__ void eye_kernel(double* p, size_t n) {
	for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n * n; e += (size_t)gridDim.x * blockDim.x) p[e] = (e / n == e % n) ? 1.0 : 0.0;
}
}  // namespace
void eye_f64(double* p, size_t n, hipStream_t s) {
	if (!n) return;
	hipLaunchKernelGGL(eye_kernel, dim3(grid_for(n * n)), dim3(256), 0, s, p, n);
	XTT_LAUNCH_CHECK();
}
void zero_u32(unsigned* p, size_t n, hipStream_t s) {
	if (!n) return;
	hipLaunchKernelGGL(zero_u32_kernel, dim3(grid_for(n)), dim3(256), 0, s, p, n);
	XTT_LAUNCH_CHECK();
}
void fill_f64(double* p, double v, size_t n, hipStream_t s) {
	if (!n) return;
	hipLaunchKernelGGL(fill_kernel, dim3(grid_for(n)), dim3(256), 0, s, p, v, n);
	XTT_LAUNCH_CHECK();
}
void scale_f64(double* p, double a, size_t n, hipStream_t s) {
	if (!n) return;
	hipLaunchKernelGGL(scale_kernel, dim3(grid_for(n)), dim3(256), 0, s, p, a, n);
	XTT_LAUNCH_CHECK();
}
void scale_by_f64(double* m, const double* sv, size_t rows, size_t cols, bool byRow, hipStream_t s) {
	if (!rows || !cols) return;
	hipLaunchKernelGGL(scale_by_kernel, dim3(grid_for(rows * cols)), dim3(256), 0, s, m, sv, rows, cols, byRow ? 1 : 0);
	XTT_LAUNCH_CHECK();
}
void svd_rank_f64(const double* sv, size_t k, size_t maxRank, double eps, int* out, hipStream_t s) {
	hipLaunchKernelGGL(svd_rank_kernel, dim3(1), dim3(64), 0, s, sv, k, maxRank, eps, out);
	XTT_LAUNCH_CHECK();
}
void axpy_f64(double* y, double a, const double* x, size_t n, hipStream_t s) {
	if (!n) return;
	hipLaunchKernelGGL(axpy_kernel, dim3(grid_for(n)), dim3(256), 0, s, y, a, x, n);
	XTT_LAUNCH_CHECK();
}
void dot_f64(const double* x, const double* y, size_t n, double* out, double* partials, unsigned* counter, hipStream_t s) {
	hipLaunchKernelGGL(dot_kernel<double>, dim3((unsigned)reduce_blocks(n)), dim3(RB), 0, s, x, y, n, out, partials, counter);
	XTT_LAUNCH_CHECK();
}
void norms_f64(const double* x, size_t n, double* out3, double* partials, unsigned* counter, hipStream_t s) {
	const unsigned g = (unsigned)reduce_blocks(n);
	hipLaunchKernelGGL(asum_amax_kernel, dim3(g), dim3(RB), 0, s, x, n, out3, partials, counter);
	XTT_LAUNCH_CHECK();
	hipLaunchKernelGGL(scaled_nrm2_kernel, dim3(g), dim3(RB), 0, s, x, n, out3, partials, counter);
	XTT_LAUNCH_CHECK();
}
void residual_norms_f64(double* r, const double* b, const double* q, size_t n, double* out2, double* partials, unsigned* counter,
                        hipStream_t s, double tol2, unsigned* fail) {
	hipLaunchKernelGGL(residual_norms_kernel, dim3((unsigned)reduce_blocks(n)), dim3(RB), 0, s, r, b, q, n, out2, partials, counter,
	                   tol2, fail);
	XTT_LAUNCH_CHECK();
}
void dot_f32_acc64(const float* x, const float* y, size_t n, double* out, double* partials, unsigned* counter, hipStream_t s) {
	hipLaunchKernelGGL(dot_kernel<float>, dim3((unsigned)reduce_blocks(n)), dim3(RB), 0, s, x, y, n, out, partials, counter);
	XTT_LAUNCH_CHECK();
}

template <class T>
void cg_start(T* r, T* p, const T* b, const T* q, size_t n, CGState* st, double* partials, hipStream_t s) {
	hipLaunchKernelGGL(cg_start_kernel<T>, dim3((unsigned)reduce_blocks(n)), dim3(RB), 0, s, r, p, b, q, n, st, partials);
	XTT_LAUNCH_CHECK();
}
template <class T>
void cg_alpha(const T* p, const T* q, size_t n, CGState* st, double* partials, hipStream_t s) {
	hipLaunchKernelGGL(cg_alpha_kernel<T>, dim3((unsigned)reduce_blocks(n)), dim3(RB), 0, s, p, q, n, st, partials);
	XTT_LAUNCH_CHECK();
}
template <class T>
void cg_update(T* x, T* r, const T* p, const T* q, size_t n, CGState* st, double* partials, hipStream_t s) {
	hipLaunchKernelGGL(cg_update_kernel<T>, dim3((unsigned)reduce_blocks(n)), dim3(RB), 0, s, x, r, p, q, n, st, partials);
	XTT_LAUNCH_CHECK();
}
template <class T>
void cg_direction(T* p, const T* r, size_t n, CGState* st, hipStream_t s) {
	hipLaunchKernelGGL(cg_direction_kernel<T>, dim3((unsigned)reduce_blocks(n)), dim3(RB), 0, s, p, r, n, st);
	XTT_LAUNCH_CHECK();
}
void cg_state_init(CGState* st, double tol2, int maxIter, hipStream_t s) {
	hipLaunchKernelGGL(cg_state_init_kernel, dim3(1), dim3(64), 0, s, st, tol2, maxIter);
	XTT_LAUNCH_CHECK();
}

template void cg_start<double>(double*, double*, const double*, const double*, size_t, CGState*, double*, hipStream_t);
template void cg_start<float>(float*, float*, const float*, const float*, size_t, CGState*, double*, hipStream_t);
template void cg_alpha<double>(const double*, const double*, size_t, CGState*, double*, hipStream_t);
template void cg_alpha<float>(const float*, const float*, size_t, CGState*, double*, hipStream_t);
template void cg_update<double>(double*, double*, const double*, const double*, size_t, CGState*, double*, hipStream_t);
template void cg_update<float>(float*, float*, const float*, const float*, size_t, CGState*, double*, hipStream_t);
template void cg_direction<double>(double*, const double*, size_t, CGState*, hipStream_t);
template void cg_direction<float>(float*, const float*, size_t, CGState*, hipStream_t);

}  // namespace xtt
