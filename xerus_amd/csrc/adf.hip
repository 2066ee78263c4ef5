// ADF sampled-entry kernels (measurements are single points of a TT tensor).
//
// adf_residual : one thread per measurement walks the TT left to right in registers
//                (x_i = G_0[:,i_0,:]·…·G_{d-1}[:,i_{d-1},:]) and writes resid_i = x_i - b_i; Σ resid² and Σ b² are
//                reduced in FP64 (deterministic last-block reduction).
// adf_gradient : the adjoint of that contraction. For a batch of measurements, phase 1 (thread per
//                measurement) stores the left stacks F_k[i] and right stacks B_k[i] (recomputed per batch —
//                nothing is kept between kernels) in LDS; phase 2 gives each thread exclusive ownership of
//                gradient entries (k, a, c) and accumulates ∂½Σres²/∂G_k[a, i_k, c] += res_i·F_{k-1}[i][a]·B_{k+1}[i][c]
//                over the batch in measurement order into LDS. Block partials are summed in block order by a
//                second kernel, so the result is bit-reproducible (no float atomics).
// adf_slice_norms : Σ_{i: i_k = j} (F_{k-1}[i]·D[:, j, :]·B_{k+1}[i])² for a direction D of core k (the
//                denominator of ADF's slice-wise step, adf.cpp:413-465).
#include "kernels.hpp"
#include "xtt_common.hpp"

namespace xtt {
namespace {

constexpr int AB = 256;     // threads per ADF block
constexpr int RMAX = 16;    // max TT rank handled in registers

__device__ inline double wsum64(double v) {
#pragma unroll
	for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
	return v;
}

// Left-to-right walk of one measurement; returns x_i.
template <class T>
__device__ inline T tt_value(const AdfDesc& a, const T* cores, const uint32_t* p, int kSel = -1, const T* D = nullptr) {
	T f[RMAX];
	f[0] = 1.f;
	int rl = 1;
	for (int k = 0; k < a.d; ++k) {
		const int n = a.n[k], rr = a.rank[k + 1];
		if (p[k] >= (uint32_t)n) return T(NAN);  // position outside the tensor: no memory access
		const T* G = (k == kSel ? D : cores + a.coreOff[k]) + (size_t)p[k] * rr;
		T g[RMAX];
#pragma unroll
		for (int c = 0; c < RMAX; ++c) g[c] = T(0);
		for (int al = 0; al < rl; ++al) {
			const T fa = f[al];
			const T* row = G + (size_t)al * n * rr;
#pragma unroll
			for (int c = 0; c < RMAX; ++c) if (c < rr) g[c] += fa * row[c];
		}
#pragma unroll
		for (int c = 0; c < RMAX; ++c) f[c] = g[c];
		rl = rr;
	}
	return f[0];
}

template <class T>
__global__ __launch_bounds__(AB) void adf_residual_kernel(AdfDesc a, const T* cores, const uint32_t* pos, const T* values,
                                                          T* resid, double* norms, double* partials, unsigned* counter) {
	__shared__ double ws[2 * AB / 64];
	__shared__ int amLast;
	double rr = 0, bb = 0;
	for (int i = blockIdx.x * AB + threadIdx.x; i < a.P; i += gridDim.x * AB) {
		const T v = tt_value<T>(a, cores, pos + (size_t)i * a.d);
		const T bi = values ? values[i] : T(0);
		const T r = v - bi;
		resid[i] = r;
		rr += (double)r * r;
		bb += (double)bi * bi;
	}
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	rr = wsum64(rr); bb = wsum64(bb);
	if (lane == 0) { ws[2 * wave] = rr; ws[2 * wave + 1] = bb; }
	__syncthreads();
	if (threadIdx.x == 0) {
		double s0 = 0, s1 = 0;
		for (int w = 0; w < AB / 64; ++w) { s0 += ws[2 * w]; s1 += ws[2 * w + 1]; }
		partials[2 * blockIdx.x] = s0;
		partials[2 * blockIdx.x + 1] = s1;
		__threadfence();
		const unsigned t = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
		amLast = (t == gridDim.x - 1);
	}
	__syncthreads();
	if (amLast && threadIdx.x == 0) {
		__threadfence();
		double s0 = 0, s1 = 0;
		for (unsigned b = 0; b < gridDim.x; ++b) {
			s0 += __hip_atomic_load(partials + 2 * b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
			s1 += __hip_atomic_load(partials + 2 * b + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		}
		norms[0] = s0;
		norms[1] = s1;
		*counter = 0;
	}
}

// LDS layout of the gradient kernel for a batch of SB measurements:
//   F[k][al][s]  left stack after core k (k = 0..d-2), ranks r_{k+1}
//   B[k][c][s]   right stack from core k on (k = 1..d-1), ranks r_k
//   idx[k][s], res[s], acc[gradSize]
struct GradLayout {
	int SB;
	size_t fOff, bOff, iOff, rOff, aOff, total;  // in 4-byte words
};

__device__ inline int triple_count(const AdfDesc& a) {
	int t = 0;
	for (int k = 0; k < a.d; ++k) t += a.rank[k] * a.rank[k + 1];
	return t;
}

template <class T>
__global__ __launch_bounds__(AB) void adf_grad_kernel(AdfDesc a, const T* cores, const uint32_t* pos, const T* resid,
                                                      T* partials, GradLayout L, int kSel) {
	extern __shared__ __align__(16) unsigned char smraw[];
	T* sm = reinterpret_cast<T*>(smraw);
	T* Fs = sm + L.fOff;
	T* Bs = sm + L.bOff;
	int* Is = reinterpret_cast<int*>(sm + L.iOff);
	T* Rs = sm + L.rOff;
	T* acc = sm + L.aOff;
	const int d = a.d, SB = L.SB, tid = threadIdx.x;
	for (int e = tid; e < a.gradSize; e += AB) acc[e] = T(0);
	// per-block measurement range
	const int per = (a.P + gridDim.x - 1) / gridDim.x;
	const int s0 = blockIdx.x * per, s1 = min(a.P, s0 + per);
	const int R = a.maxR;
	const int ntrip = triple_count(a);
	__syncthreads();
	for (int base = s0; base < s1; base += SB) {
		const int cnt = min(SB, s1 - base);
		if (tid < SB) {
			const int s = tid;
			if (s < cnt) {
				const uint32_t* p = pos + (size_t)(base + s) * d;
				bool valid = true;
				for (int k = 0; k < d; ++k) valid = valid && p[k] < (uint32_t)a.n[k];
				for (int k = 0; k < d; ++k) Is[k * SB + s] = valid ? (int)p[k] : 0;
				Rs[s] = valid ? resid[base + s] : T(0);
				if (!valid) p = nullptr;
				// forward: F_k = F_{k-1}·G_k[:, i_k, :] (an invalid position walks index 0 with zero weight)
				static const uint32_t zeros[32] = {0};
				if (!p) p = zeros;
				T f[RMAX];
				f[0] = T(1);
				int rl = 1;
				for (int k = 0; k + 1 < d; ++k) {
					const int n = a.n[k], rr = a.rank[k + 1];
					const T* G = cores + a.coreOff[k] + (size_t)p[k] * rr;
					T g[RMAX];
#pragma unroll
					for (int c = 0; c < RMAX; ++c) g[c] = T(0);
					for (int al = 0; al < rl; ++al) {
						const T fa = f[al];
						const T* row = G + (size_t)al * n * rr;
#pragma unroll
						for (int c = 0; c < RMAX; ++c) if (c < rr) g[c] += fa * row[c];
					}
					for (int c = 0; c < rr; ++c) { f[c] = g[c]; Fs[((size_t)k * R + c) * SB + s] = g[c]; }
					rl = rr;
				}
				// backward: B_k = G_k[:, i_k, :]·B_{k+1}
				T b[RMAX];
				b[0] = T(1);
				int rr = 1;
				for (int k = d - 1; k >= 1; --k) {
					const int n = a.n[k], rlk = a.rank[k];
					const T* G = cores + a.coreOff[k] + (size_t)p[k] * rr;
					T g[RMAX];
#pragma unroll
					for (int al = 0; al < RMAX; ++al) {
						T t = T(0);
						if (al < rlk) {
							const T* row = G + (size_t)al * n * rr;
							for (int c = 0; c < rr; ++c) t += row[c] * b[c];
						}
						g[al] = t;
					}
					for (int al = 0; al < rlk; ++al) { b[al] = g[al]; Bs[((size_t)k * R + al) * SB + s] = g[al]; }
					rr = rlk;
				}
			}
		}
		__syncthreads();
		// phase 2: thread-owned gradient triples (k, al, c)
		for (int t = tid; t < ntrip; t += AB) {
			int k = 0, rem = t;
			while (rem >= a.rank[k] * a.rank[k + 1]) { rem -= a.rank[k] * a.rank[k + 1]; ++k; }
			if (kSel >= 0 && k != kSel) continue;
			const int rl = a.rank[k], rr = a.rank[k + 1], n = a.n[k];
			const int al = rem / rr, c = rem % rr;
			T* out = acc + (a.coreOff[k] + (size_t)al * n * rr + c);
			const T* fcol = k > 0 ? Fs + ((size_t)(k - 1) * R + al) * SB : nullptr;
			const T* bcol = k + 1 < d ? Bs + ((size_t)(k + 1) * R + c) * SB : nullptr;
			const int* icol = Is + k * SB;
			for (int s = 0; s < cnt; ++s) {
				const T v = Rs[s] * (fcol ? fcol[s] : T(1)) * (bcol ? bcol[s] : T(1));
				out[(size_t)icol[s] * rr] += v;
			}
			(void)rl;
		}
		__syncthreads();
	}
	T* dst = partials + (size_t)blockIdx.x * a.gradSize;
	for (int e = tid; e < a.gradSize; e += AB) dst[e] = acc[e];
}

template <class T>
__global__ void adf_grad_reduce_kernel(const T* partials, int nblocks, int gradSize, T* grad) {
	for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < gradSize; e += gridDim.x * blockDim.x) {
		T s = T(0);
		for (int b = 0; b < nblocks; ++b) s += partials[(size_t)b * gradSize + e];
		grad[e] = s;
	}
}

GradLayout grad_layout(int d, int maxR, int gradSize, size_t E) {
	GradLayout L{};
	for (int SB = 256; SB >= 16; SB >>= 1) {
		L.SB = SB;
		L.fOff = 0;
		L.bOff = L.fOff + (size_t)d * maxR * SB;
		L.iOff = L.bOff + (size_t)d * maxR * SB;
		L.rOff = L.iOff + (size_t)d * SB;
		L.aOff = L.rOff + SB;
		L.total = L.aOff + gradSize;
		if (L.total * E <= 150 * 1024) return L;
	}
	throw Error(ERR_UNSUPPORTED, "ADF: problem too large for the LDS-staged gradient kernel");
}


// Σ_{i: i_k = j} value_i² with value_i the TT evaluated with core k replaced by D (slice-wise ‖P_j A D‖²).
template <class T>
__global__ __launch_bounds__(AB) void adf_slice_kernel(AdfDesc a, const T* cores, const uint32_t* pos, int k, const T* D,
                                                       double* partials, int nslices) {
	__shared__ double ws[AB / 64];
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	double* dst = partials + (size_t)blockIdx.x * nslices;
	for (int j = 0; j < nslices; ++j) {
		double v = 0;
		for (int i = blockIdx.x * AB + threadIdx.x; i < a.P; i += gridDim.x * AB) {
			const uint32_t* p = pos + (size_t)i * a.d;
			if ((int)p[k] == j) {
				const double t = (double)tt_value<T>(a, cores, p, k, D);
				v += t * t;
			}
		}
		v = wsum64(v);
		__syncthreads();
		if (lane == 0) ws[wave] = v;
		__syncthreads();
		if (threadIdx.x == 0) {
			double s0 = 0;
			for (int w = 0; w < AB / 64; ++w) s0 += ws[w];
			dst[j] = s0;
		}
	}
}

__global__ void adf_slice_reduce_kernel(const double* partials, int nblocks, int nslices, double* out) {
	for (int j = threadIdx.x; j < nslices; j += blockDim.x) {
		double s = 0;
		for (int b = 0; b < nblocks; ++b) s += partials[(size_t)b * nslices + j];
		out[j] = s;
	}
}

}  // namespace

template <class T>
void adf_residual(const AdfDesc& a, const T* cores, const uint32_t* pos, const T* values, T* resid, double* norms,
                  double* partials, unsigned* counter, hipStream_t s) {
	XTT_REQUIRE(a.maxR <= RMAX, "ADF: ranks above 16 are not supported");
	XTT_REQUIRE(a.d >= 1 && a.d <= 32, "ADF: order must be between 1 and 32");
	const int blocks = std::max(1, std::min(1024, (a.P + AB - 1) / AB));
	const int slot = prof_begin(s);
	hipLaunchKernelGGL(adf_residual_kernel<T>, dim3(blocks), dim3(AB), 0, s, a, cores, pos, values, resid, norms, partials, counter);
	XTT_LAUNCH_CHECK();
	prof_end(slot, s, sizeof(T) == 8 ? "adf_residual_kernel<double>" : "adf_residual_kernel<float>", 0.0);
}

int adf_grad_blocks(int P) { return std::max(1, std::min(512, (P + 255) / 256)); }

template <class T>
void adf_gradient(const AdfDesc& a, const T* cores, const uint32_t* pos, const T* resid, T* grad, T* partials, int kSel,
                  hipStream_t s) {
	XTT_REQUIRE(a.maxR <= RMAX, "ADF: ranks above 16 are not supported");
	XTT_REQUIRE(a.d >= 1 && a.d <= 32, "ADF: order must be between 1 and 32");
	const GradLayout L = grad_layout(a.d, a.maxR, a.gradSize, sizeof(T));
	const size_t lds = L.total * sizeof(T);
	if (lds > 64 * 1024)
		XTT_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(adf_grad_kernel<T>), hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
	const int blocks = adf_grad_blocks(a.P);
	const int slot = prof_begin(s);
	hipLaunchKernelGGL(adf_grad_kernel<T>, dim3(blocks), dim3(AB), lds, s, a, cores, pos, resid, partials, L, kSel);
	XTT_LAUNCH_CHECK();
	prof_end(slot, s, sizeof(T) == 8 ? "adf_grad_kernel<double>" : "adf_grad_kernel<float>", 0.0);
	const int rb = std::max(1, std::min(256, (a.gradSize + 255) / 256));
	hipLaunchKernelGGL(adf_grad_reduce_kernel<T>, dim3(rb), dim3(256), 0, s, partials, blocks, a.gradSize, grad);
	XTT_LAUNCH_CHECK();
}

int adf_slice_blocks(int P) { return std::max(1, std::min(256, (P + 255) / 256)); }

template <class T>
void adf_slice_norms(const AdfDesc& a, const T* cores, const uint32_t* pos, int k, const T* D, int nslices, double* out,
                     double* partials, hipStream_t s) {
	const int blocks = adf_slice_blocks(a.P);
	hipLaunchKernelGGL(adf_slice_kernel<T>, dim3(blocks), dim3(AB), 0, s, a, cores, pos, k, D, partials, nslices);
	XTT_LAUNCH_CHECK();
	hipLaunchKernelGGL(adf_slice_reduce_kernel, dim3(1), dim3(64), 0, s, partials, blocks, nslices, out);
	XTT_LAUNCH_CHECK();
}

template void adf_residual<float>(const AdfDesc&, const float*, const uint32_t*, const float*, float*, double*, double*, unsigned*, hipStream_t);
template void adf_residual<double>(const AdfDesc&, const double*, const uint32_t*, const double*, double*, double*, double*, unsigned*, hipStream_t);
template void adf_gradient<float>(const AdfDesc&, const float*, const uint32_t*, const float*, float*, float*, int, hipStream_t);
template void adf_gradient<double>(const AdfDesc&, const double*, const uint32_t*, const double*, double*, double*, int, hipStream_t);
template void adf_slice_norms<float>(const AdfDesc&, const float*, const uint32_t*, int, const float*, int, double*, double*, hipStream_t);
template void adf_slice_norms<double>(const AdfDesc&, const double*, const uint32_t*, int, const double*, int, double*, double*, hipStream_t);

}  // namespace xtt
