// ADF kernels for rank-one measurements (RankOneMeasurementSet, src/xerus/measurments.cpp:250-545, and the
// RankOneMeasurementSet specialisations of src/xerus/algorithms/adf.cpp).
//
// Measurement i is a rank-one tensor u_{i,0} ⊗ … ⊗ u_{i,d-1}; its value is x(u_i) = M_0·M_1·…·M_{d-1} with the
// mixed matrices M_k = Σ_j u_{i,k}[j]·G_k[:, j, :] (adf.cpp:233-288 builds its stacks from the same matrices).
// The vectors of all measurements lie in HBM as one P × Σn_k row-major array (measurement-major, mode k's
// vector at column offset Σ_{j<k} n_j), so a thread reads its own measurement's row.
//
// r1_residual   : one thread per measurement walks the chain left to right in registers (rank <= 16):
//                 cost Σ_k r_{k-1}·n_k·r_k FMAs; writes resid_i = x(u_i) − b_i and per-block (Σ resid², Σ b²).
// r1_gradient   : ∂½Σ resid²/∂G_k[a, j, c] = Σ_i resid_i·F_{k-1}[i][a]·u_{i,k}[j]·B_{k+1}[i][c] (the rank-one
//                 dyadic product of adf.cpp:333-357). Phase 1 (thread per measurement) stages resid·F, u_k and B
//                 of a batch in LDS; phase 2 gives each thread exclusive gradient entries and accumulates over
//                 the batch in measurement order. Block partials are summed in block order by a second kernel.
// r1_slice_norm : Σ_i (x(u_i) with core k replaced by D)² — the rank-one step denominator (adf.cpp:413-465 with
//                 position_or_zero = 0).
// Every reduction is per-block partials summed in a fixed order: results are bit-reproducible.
#include <algorithm>
#include <mutex>

#include "kernels.hpp"
#include "xtt_common.hpp"

namespace xtt {
namespace {

constexpr int R1B = 256;  // threads per block
constexpr int R1M = 16;   // max TT rank in registers

__device__ inline int mode_offset(const AdfDesc& a, int k) {
	int o = 0;
	for (int j = 0; j < k; ++j) o += a.n[j];
	return o;
}

// f (row vector, length rl on entry, a.rank[k1] on exit) ← f · M_k0 · … · M_{k1-1}; core kSel replaced by D.
template <class T>
__device__ inline void left_walk(const AdfDesc& a, const T* cores, const T* u, int k0, int k1, T* f, int kSel, const T* D) {
	int uo = mode_offset(a, k0);
	for (int k = k0; k < k1; ++k) {
		const int n = a.n[k], rl = a.rank[k], rr = a.rank[k + 1];
		const T* G = k == kSel ? D : cores + a.coreOff[k];
		T g[R1M];
#pragma unroll
		for (int c = 0; c < R1M; ++c) g[c] = T(0);
#pragma unroll
		for (int al = 0; al < R1M; ++al) {  // static register index
			if (al < rl) {
				const T fa = f[al];
				for (int j = 0; j < n; ++j) {
					const T w = fa * u[uo + j];
					const T* row = G + ((size_t)al * n + j) * rr;
#pragma unroll
					for (int c = 0; c < R1M; ++c)
						if (c < rr) g[c] = fma(w, row[c], g[c]);
				}
			}
		}
#pragma unroll
		for (int c = 0; c < R1M; ++c) f[c] = g[c];
		uo += n;
	}
}

// b (column vector of length a.rank[k0+1]) ← M_{k0+1} · … · M_{d-1} · 1, walked from the right end down to k0+1.
template <class T>
__device__ inline void right_walk(const AdfDesc& a, const T* cores, const T* u, int k0, T* b) {
	b[0] = T(1);
	int uo = mode_offset(a, a.d);
	for (int k = a.d - 1; k > k0; --k) {
		const int n = a.n[k], rl = a.rank[k], rr = a.rank[k + 1];
		uo -= n;
		const T* G = cores + a.coreOff[k];
		T g[R1M];
#pragma unroll
		for (int al = 0; al < R1M; ++al) {  // static register index
			T s = T(0);
			if (al < rl) {
				for (int j = 0; j < n; ++j) {
					const T* row = G + ((size_t)al * n + j) * rr;
					T t = T(0);
#pragma unroll
					for (int c = 0; c < R1M; ++c)
						if (c < rr) t = fma(row[c], b[c], t);
					s = fma(u[uo + j], t, s);
				}
			}
			g[al] = s;
		}
#pragma unroll
		for (int c = 0; c < R1M; ++c) b[c] = g[c];
	}
}

// Block sum of two doubles (fixed order: waves, then lanes by butterfly) written to partials[2·block].
__device__ inline void block_pair(double x, double y, double* partials) {
	__shared__ double s[2][R1B / 64];
#pragma unroll
	for (int o = 32; o > 0; o >>= 1) {
		x += __shfl_xor(x, o, 64);
		y += __shfl_xor(y, o, 64);
	}
	const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
	if (lane == 0) {
		s[0][w] = x;
		s[1][w] = y;
	}
	__syncthreads();
	if (threadIdx.x == 0) {
		double a = 0, b = 0;
		for (int q = 0; q < R1B / 64; ++q) {
			a += s[0][q];
			b += s[1][q];
		}
		partials[2 * blockIdx.x] = a;
		partials[2 * blockIdx.x + 1] = b;
	}
}

template <class T>
__global__ __launch_bounds__(R1B) void r1_residual_kernel(AdfDesc a, const T* cores, const T* U, const T* values, T* resid,
                                                         double* partials, int kSel, const T* D) {
	double rr = 0, bb = 0;
	for (int i = blockIdx.x * R1B + threadIdx.x; i < a.P; i += gridDim.x * R1B) {
		T f[R1M];
		f[0] = T(1);
		left_walk<T>(a, cores, U + (size_t)i * a.sumN, 0, a.d, f, kSel, D);
		const T b = values ? values[i] : T(0);
		const T r = f[0] - b;
		if (resid) resid[i] = r;
		rr += (double)r * (double)r;
		bb += (double)b * (double)b;
	}
	block_pair(rr, bb, partials);
}

__global__ void r1_pair_reduce_kernel(const double* partials, int nblocks, double* out) {
	if (threadIdx.x != 0) return;
	double a = 0, b = 0;
	for (int q = 0; q < nblocks; ++q) {
		a += partials[2 * q];
		b += partials[2 * q + 1];
	}
	out[0] = a;
	out[1] = b;
}

// Gradient of core k. LDS: acc[gk] | sF[SB][rl] | sU[SB][n] | sB[SB][rr].
template <class T>
__global__ __launch_bounds__(R1B) void r1_grad_kernel(AdfDesc a, const T* cores, const T* U, const T* resid, int k, int SB,
                                                     T* partials) {
	extern __shared__ __align__(16) unsigned char r1_smem[];
	const int n = a.n[k], rl = a.rank[k], rr = a.rank[k + 1];
	const int gk = rl * n * rr;
	T* acc = reinterpret_cast<T*>(r1_smem);
	T* sF = acc + gk;
	T* sU = sF + (size_t)SB * rl;
	T* sB = sU + (size_t)SB * n;
	const int uo = mode_offset(a, k);
	for (int e = threadIdx.x; e < gk; e += R1B) acc[e] = T(0);
	const int batches = (a.P + SB - 1) / SB;
	for (int bt = blockIdx.x; bt < batches; bt += gridDim.x) {
		const int i0 = bt * SB, cnt = min(SB, a.P - i0);
		__syncthreads();  // the previous batch's phase 2 is done with the staging arrays
		if (threadIdx.x < cnt) {
			const int i = i0 + threadIdx.x;
			const T* u = U + (size_t)i * a.sumN;
			T f[R1M], b[R1M];
			f[0] = T(1);
			left_walk<T>(a, cores, u, 0, k, f, -1, nullptr);
			right_walk<T>(a, cores, u, k, b);
			const T r = resid[i];
#pragma unroll
			for (int q = 0; q < R1M; ++q) {
				if (q < rl) sF[threadIdx.x * rl + q] = r * f[q];
				if (q < rr) sB[threadIdx.x * rr + q] = b[q];
			}
			for (int j = 0; j < n; ++j) sU[threadIdx.x * n + j] = u[uo + j];
		}
		__syncthreads();
		for (int e = threadIdx.x; e < gk; e += R1B) {
			const int al = e / (n * rr), j = (e / rr) % n, c = e % rr;
			T s = acc[e];
			for (int t = 0; t < cnt; ++t) s = fma(sF[t * rl + al] * sU[t * n + j], sB[t * rr + c], s);
			acc[e] = s;
		}
	}
	__syncthreads();
	for (int e = threadIdx.x; e < gk; e += R1B) partials[(size_t)blockIdx.x * gk + e] = acc[e];
}

template <class T>
__global__ __launch_bounds__(256) void r1_grad_reduce_kernel(const T* partials, int nblocks, int gk, T* grad) {
	const int e = blockIdx.x * 256 + threadIdx.x;
	if (e >= gk) return;
	T s = T(0);
	for (int q = 0; q < nblocks; ++q) s += partials[(size_t)q * gk + e];
	grad[e] = s;
}

}  // namespace

int r1_blocks(int P) { return std::max(1, std::min(512, (P + R1B - 1) / R1B)); }

template <class T>
void r1_residual(const AdfDesc& a, const T* cores, const T* U, const T* values, T* resid, double* norms2, double* partials,
                 int kSel, const T* D, hipStream_t s) {
	XTT_REQUIRE(a.maxR <= R1M, "ADF (rank-one): ranks above 16 are not supported");
	const int blocks = r1_blocks(a.P);
	launch_timed(sizeof(T) == 8 ? "r1_residual_kernel<double>" : "r1_residual_kernel<float>", 0.0, r1_residual_kernel<T>,
	             dim3(blocks), dim3(R1B), 0, s, a, cores, U, values, resid, partials, kSel, D);
	hipLaunchKernelGGL(r1_pair_reduce_kernel, dim3(1), dim3(64), 0, s, (const double*)partials, blocks, norms2);
	XTT_LAUNCH_CHECK();
}

size_t r1_grad_lds(const AdfDesc& a, int rl, int n, int rr, int SB, size_t E) {
	(void)a;
	return ((size_t)rl * n * rr + (size_t)SB * (rl + n + rr)) * E;
}

template <class T>
void r1_gradient(const AdfDesc& a, const T* cores, const T* U, const T* resid, int k, int rl, int n, int rr, T* grad,
                 T* partials, hipStream_t s) {
	XTT_REQUIRE(a.maxR <= R1M, "ADF (rank-one): ranks above 16 are not supported");
	constexpr size_t kMax = 150 * 1024;
	int SB = R1B;
	while (SB > 16 && r1_grad_lds(a, rl, n, rr, SB, sizeof(T)) > kMax) SB >>= 1;
	const size_t lds = r1_grad_lds(a, rl, n, rr, SB, sizeof(T));
	XTT_REQUIRE(lds <= kMax, "ADF (rank-one): core too large for the LDS-staged gradient");
	static std::once_flag once;
	std::call_once(once, [] {
		XTT_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(r1_grad_kernel<T>), hipFuncAttributeMaxDynamicSharedMemorySize,
		                            (int)kMax));
	});
	const int blocks = std::min(r1_blocks(a.P), (a.P + SB - 1) / SB);
	const int gk = rl * n * rr;
	launch_timed(sizeof(T) == 8 ? "r1_grad_kernel<double>" : "r1_grad_kernel<float>", 0.0, r1_grad_kernel<T>, dim3(blocks),
	             dim3(R1B), lds, s, a, cores, U, resid, k, SB, partials);
	hipLaunchKernelGGL(r1_grad_reduce_kernel<T>, dim3((gk + 255) / 256), dim3(256), 0, s, (const T*)partials, blocks, gk, grad);
	XTT_LAUNCH_CHECK();
}

template void r1_residual<double>(const AdfDesc&, const double*, const double*, const double*, double*, double*, double*, int,
                                  const double*, hipStream_t);
template void r1_residual<float>(const AdfDesc&, const float*, const float*, const float*, float*, double*, double*, int, const float*,
                                 hipStream_t);
template void r1_gradient<double>(const AdfDesc&, const double*, const double*, const double*, int, int, int, int, double*, double*,
                                  hipStream_t);
template void r1_gradient<float>(const AdfDesc&, const float*, const float*, const float*, int, int, int, int, float*, float*,
                                 hipStream_t);

}  // namespace xtt
