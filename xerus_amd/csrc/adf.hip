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
#include <mutex>
#include <type_traits>

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

// Left-to-right walk of one measurement; returns x_i. RM bounds the ranks (registers), `cores` may point to
// an LDS copy.
template <class T, int RM = RMAX>
__device__ inline T tt_value(const AdfDesc& a, const T* cores, const uint32_t* p, int kSel = -1, const T* D = nullptr,
                             const int* nS = nullptr, const int* rS = nullptr, const long long* offS = nullptr) {
	T f[RM];
	f[0] = 1.f;
	int rl = 1;
	for (int k = 0; k < a.d; ++k) {
		const int n = nS ? nS[k] : a.n[k], rr = rS ? rS[k + 1] : a.rank[k + 1];
		if (p[k] >= (uint32_t)n) return T(NAN);  // position outside the tensor: no memory access
		const T* G = (k == kSel ? D : cores + (offS ? offS[k] : a.coreOff[k])) + (size_t)p[k] * rr;
		T g[RM];
#pragma unroll
		for (int c = 0; c < RM; ++c) g[c] = T(0);
#pragma unroll
		for (int al = 0; al < RM; ++al) {
			if (al < rl) {
				const T fa = f[al];
				const T* row = G + (size_t)al * n * rr;
#pragma unroll
				for (int c = 0; c < RM; ++c) if (c < rr) g[c] = fma(fa, row[c], g[c]);
			}
		}
#pragma unroll
		for (int c = 0; c < RM; ++c) f[c] = g[c];
		rl = rr;
	}
	return f[0];
}

template <class T, int RM, bool LDSCORES>
__global__ __launch_bounds__(AB) void adf_residual_kernel(AdfDesc a, const T* cores, const uint32_t* pos, const T* values,
                                                          T* resid, double* norms, double* partials, unsigned* counter) {
	extern __shared__ __align__(16) unsigned char rsm[];
	__shared__ double ws[2 * AB / 64];
	__shared__ int amLast;
	__shared__ int sN[32], sRk[33];
	__shared__ long long sOff[32];
	if (threadIdx.x < a.d) { sN[threadIdx.x] = a.n[threadIdx.x]; sOff[threadIdx.x] = a.coreOff[threadIdx.x]; }
	if (threadIdx.x <= a.d) sRk[threadIdx.x] = a.rank[threadIdx.x];
	const T* src = cores;
	if (LDSCORES) {
		T* sc = reinterpret_cast<T*>(rsm);
		for (int e = threadIdx.x; e < a.gradSize; e += AB) sc[e] = cores[e];
		src = sc;
	}
	__syncthreads();
	double rr = 0, bb = 0;
	for (int i = blockIdx.x * AB + threadIdx.x; i < a.P; i += gridDim.x * AB) {
		const T v = tt_value<T, RM>(a, src, pos + (size_t)i * a.d, -1, nullptr, sN, sRk, sOff);
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
	if (amLast) {  // every thread sums a fixed stride of the partials, then a fixed-order block sum
		__threadfence();
		double s0 = 0, s1 = 0;
		for (unsigned b = threadIdx.x; b < gridDim.x; b += AB) {
			s0 += __hip_atomic_load(partials + 2 * b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
			s1 += __hip_atomic_load(partials + 2 * b + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		}
		s0 = wsum64(s0); s1 = wsum64(s1);
		__syncthreads();
		if (lane == 0) { ws[2 * wave] = s0; ws[2 * wave + 1] = s1; }
		__syncthreads();
		if (threadIdx.x == 0) {
			double t0 = 0, t1 = 0;
			for (int w = 0; w < AB / 64; ++w) { t0 += ws[2 * w]; t1 += ws[2 * w + 1]; }
			norms[0] = t0;
			norms[1] = t1;
			*counter = 0;
		}
	}
}

// ---- fast FP32 path for ranks <= 8 and order <= 16 ---------------------------------------------------------
// The cores are first copied into a zero-padded layout Gp[k] = (n_k, 8, 8) (slice j of core k is one
// contiguous 8×8 block, blocks PB floats apart), so the per-sample chains are branch-free 8×8 products with
// 16-byte loads; padded entries are zero and contribute nothing.
// Gradient, phase 1 (per wave, 16 samples): lanes 0..15 walk the left stacks F_k, lanes 16..31 the right
// stacks B_k of the same samples, into the wave's LDS slot. Phase 2: lane (a, c) = (lane/8, lane%8) owns entry
// (a, ·, c) of every core and adds res_s·F_{k-1}[s][a]·B_{k+1}[s][c] to its own accumulator (k, i_k(s)) in the
// wave's LDS region: every address has one owner lane, so the adds need no ordering between lanes and the
// result is deterministic (samples in order, waves summed in order, blocks by the reduce kernel).
constexpr int FW = 16;       // waves per gradient block (four per SIMD); the residual kernel uses AB threads
constexpr int FR = 8;        // rank bound
constexpr int PB = FR * FR + 4;  // floats per padded 8×8 block: the extra 4 spread blocks over the LDS banks
constexpr int FD = 16;       // order bound

__global__ void pad_cores_kernel(AdfDesc a, const float* cores, long long* padOff, float* Gp, int total) {
	if (blockIdx.x == 0 && threadIdx.x < a.d) {
		long long o = 0;
		for (int k = 0; k < (int)threadIdx.x; ++k) o += (long long)a.n[k] * PB;
		padOff[threadIdx.x] = o;
	}
	for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < total; e += gridDim.x * blockDim.x) {
		int k = 0, o = 0;
		while (k + 1 < a.d && o + a.n[k] * PB <= e) { o += a.n[k] * PB; ++k; }
		const int rem = e - o;
		const int j = rem / PB, w = rem % PB, al = w / FR, c = w % FR;
		const int rl = a.rank[k], rr = a.rank[k + 1], n = a.n[k];
		Gp[e] = (w < FR * FR && al < rl && c < rr) ? cores[a.coreOff[k] + ((size_t)al * n + j) * rr + c] : 0.f;
	}
}

// v ← v·G (row vector times the 8×8 block)
__device__ inline void mul_row8(float (&v)[FR], const float* G) {
	float g[FR];
#pragma unroll
	for (int c = 0; c < FR; ++c) g[c] = 0.f;
#pragma unroll
	for (int al = 0; al < FR; ++al) {
		const float4 lo = *reinterpret_cast<const float4*>(G + al * FR);
		const float4 hi = *reinterpret_cast<const float4*>(G + al * FR + 4);
		g[0] = fmaf(v[al], lo.x, g[0]); g[1] = fmaf(v[al], lo.y, g[1]); g[2] = fmaf(v[al], lo.z, g[2]); g[3] = fmaf(v[al], lo.w, g[3]);
		g[4] = fmaf(v[al], hi.x, g[4]); g[5] = fmaf(v[al], hi.y, g[5]); g[6] = fmaf(v[al], hi.z, g[6]); g[7] = fmaf(v[al], hi.w, g[7]);
	}
#pragma unroll
	for (int c = 0; c < FR; ++c) v[c] = g[c];
}

// v ← v·G and v ← G·v with two adjacent lanes per sample (h = lane & 1): lane h holds components 4h..4h+3 of v
// before and after, takes the other half from its partner (DPP quad_perm [1,0,3,2]) and computes its half of the
// result; v·G with the operations of mul_row8 in the same order, so its values equal mul_row8's. Every lane of the
// wave must take part.
__device__ inline float pair_swap(float x) {
	return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0xB1, 0xF, 0xF, false));
}
__device__ inline void full8(const float (&v)[4], int h, float (&f)[FR]) {
	float o[4];
#pragma unroll
	for (int c = 0; c < 4; ++c) o[c] = pair_swap(v[c]);
#pragma unroll
	for (int c = 0; c < 4; ++c) { f[c] = h ? o[c] : v[c]; f[4 + c] = h ? v[c] : o[c]; }
}
__device__ inline void mul_row8_half(float (&v)[4], const float* G, int h) {
	float f[FR];
	full8(v, h, f);
	float g[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
	for (int al = 0; al < FR; ++al) {
		const float4 x = *reinterpret_cast<const float4*>(G + al * FR + 4 * h);
		g[0] = fmaf(f[al], x.x, g[0]); g[1] = fmaf(f[al], x.y, g[1]); g[2] = fmaf(f[al], x.z, g[2]); g[3] = fmaf(f[al], x.w, g[3]);
	}
#pragma unroll
	for (int c = 0; c < 4; ++c) v[c] = g[c];
}
__device__ inline void mul_col8_half(float (&v)[4], const float* G, int h) {
	float f[FR];
	full8(v, h, f);
	float g[4];
#pragma unroll
	for (int al = 0; al < 4; ++al) {
		const float4 lo = *reinterpret_cast<const float4*>(G + (4 * h + al) * FR);
		const float4 hi = *reinterpret_cast<const float4*>(G + (4 * h + al) * FR + 4);
		g[al] = lo.x * f[0] + lo.y * f[1] + lo.z * f[2] + lo.w * f[3] + hi.x * f[4] + hi.y * f[5] + hi.z * f[6] + hi.w * f[7];
	}
#pragma unroll
	for (int al = 0; al < 4; ++al) v[al] = g[al];
}

template <class T>
__global__ __launch_bounds__(AB) void adf_residual_fast_kernel(AdfDesc a, const float* Gp, const long long* padOff, const uint32_t* pos,
                                                               const T* values, T* resid, double* norms, double* partials,
                                                               unsigned* counter) {
	extern __shared__ __align__(16) float gsm[];  // padded cores
	__shared__ double ws[2 * AB / 64];
	__shared__ int amLast;
	__shared__ int sN[FD];
	__shared__ long long sPo[FD];
	if (threadIdx.x < a.d) { sN[threadIdx.x] = a.n[threadIdx.x]; sPo[threadIdx.x] = padOff[threadIdx.x]; }
	for (int e = threadIdx.x; e < a.sumN * PB / 4; e += AB)
		reinterpret_cast<float4*>(gsm)[e] = reinterpret_cast<const float4*>(Gp)[e];
	__syncthreads();
	double rr = 0, bb = 0;
	for (int i = blockIdx.x * AB + threadIdx.x; i < a.P; i += gridDim.x * AB) {
		const uint32_t* p = pos + (size_t)i * a.d;
		float v[FR];
#pragma unroll
		for (int c = 0; c < FR; ++c) v[c] = c == 0 ? 1.f : 0.f;
		bool valid = true;
		uint32_t pk[FD];
#pragma unroll
		for (int k = 0; k < FD; ++k) pk[k] = p[k < a.d ? k : a.d - 1];
#pragma unroll
		for (int k = 0; k < FD; ++k) {
			if (k < a.d) {
				valid = valid && pk[k] < (uint32_t)sN[k];
				mul_row8(v, gsm + sPo[k] + (size_t)(valid ? pk[k] : 0) * PB);
			}
		}
		const T val = valid ? T(v[0]) : T(NAN);
		const T bi = values ? values[i] : T(0);
		const T r = val - bi;
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
	if (amLast) {  // every thread sums a fixed stride of the partials, then a fixed-order block sum
		__threadfence();
		double s0 = 0, s1 = 0;
		for (unsigned b = threadIdx.x; b < gridDim.x; b += AB) {
			s0 += __hip_atomic_load(partials + 2 * b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
			s1 += __hip_atomic_load(partials + 2 * b + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		}
		s0 = wsum64(s0); s1 = wsum64(s1);
		__syncthreads();
		if (lane == 0) { ws[2 * wave] = s0; ws[2 * wave + 1] = s1; }
		__syncthreads();
		if (threadIdx.x == 0) {
			double t0 = 0, t1 = 0;
			for (int w = 0; w < AB / 64; ++w) { t0 += ws[2 * w]; t1 += ws[2 * w + 1]; }
			norms[0] = t0;
			norms[1] = t1;
			*counter = 0;
		}
	}
}

// Gradient: one block of FW = 16 waves takes rounds of FS = 128 samples. Phase 1: waves 0-3 walk the left stacks
// F_k of the round's samples, waves 4-7 the right stacks B_k, two lanes per sample (each lane half of the rank
// components), from the LDS copy of the padded cores, into the block's staging area. Phase 2: wave w owns core k = w
// (orders up to 16: one core per wave, so the latency-bound accumulations of different cores interleave on a SIMD)
// and adds the round's contribution to its gradient with MFMA: G_k[a, j, c] += Σ_{s: i_k(s) = j} (res_s·F_{k-1}[s][a])
// · B_{k+1}[s][c]. The wave first counting-sorts the round's samples by i_k (ballot / mbcnt per bin); bins 2t and
// 2t+1 form tile t, padded to a multiple of 4 samples. A v_mfma_f32_16x16x4_f32 step takes 4 samples of the tile and
// is block diagonal: rows (g, a) and columns (g, c) of bin parity g, a sample entering only its own bin's block. The
// summation order is fixed (samples in sorted round order, blocks in block order by the reduce kernel), so the
// result is deterministic.
constexpr int FS = 128;
constexpr int WW = 2 * FS / 32;  // walker waves: FS/32 forward, FS/32 backward (32 samples per wave)
constexpr int FSP = FS + 1;  // staging row stride: lanes reading rows a = 0..7 of one sample hit different banks

constexpr int GNT = 8;       // 16-column MFMA tiles per core (8·n_k <= 128, n_k <= 16)
constexpr int GKC = FD / FW; // cores per wave

constexpr int FPERM = FS + 3 * 16;  // sorted sample slots: a round, the tiles' padding and the prefetch past the last tile

__host__ __device__ inline size_t fast_grad_floats(int d, int sumN) {
	return (size_t)sumN * PB + (size_t)FR * FSP                      // padded cores, e_0 rows
	       + (size_t)2 * (d - 1) * FR * FSP + (size_t)d * FS + FSP   // F, B, positions, residuals (and a zero)
	       + (size_t)2 * FW * FPERM;                                 // per-wave sorted order, one copy per bin parity
}

typedef float gf4 __attribute__((ext_vector_type(4)));

// With RES, the residual is computed here instead of read: the forward walkers take one more step (the last
// core) to the sampled value x_s, the same operations in the same order as adf_residual_fast_kernel, and use
// res_s = x_s − b_s (written to residOut; NaN for positions out of range, whose gradient contribution is 0). The
// block's FP64 sums Σ res² and Σ b² go to normPart[2·block], [2·block + 1].
template <bool RES>
__global__ __launch_bounds__(FW * 64) void adf_grad_fast_kernel(AdfDesc a, const float* Gp, const long long* padOff, const uint32_t* pos,
                                                                const float* resid, float* partials, int kSel, const float* values,
                                                                float* residOut, double* normPart) {
	extern __shared__ __align__(16) float fsm[];
	__shared__ int sN[FD], sRk[FD + 1], sSlot[FD + 1];
	__shared__ long long sOff[FD], sPo[FD];
	const int tid = threadIdx.x, lane = tid & 63;
	const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
	const int d = a.d;
	if (tid < d) { sN[tid] = a.n[tid]; sOff[tid] = a.coreOff[tid]; sPo[tid] = padOff[tid]; }
	if (tid <= d) sRk[tid] = a.rank[tid];
	__syncthreads();
	if (tid == 0) {
		int acc0 = 0;
		for (int k = 0; k < d; ++k) { sSlot[k] = acc0; acc0 += sN[k]; }
		sSlot[d] = acc0;
	}
	__syncthreads();
	const int nslots = sSlot[d];
	float* Gs = fsm;                                                    // padded cores
	float* Es = Gs + (size_t)nslots * PB;                               // rows e_0: F_{-1} and B_d
	float* Fs = Es + (size_t)FR * FSP;                                  // F_k at [(k·FR + c)·FSP + s], k = 0..d-2
	float* Bs = Fs + (size_t)(d - 1) * FR * FSP;                        // B_k at [((k-1)·FR + a)·FSP + s], k = 1..d-1
	int* Is = reinterpret_cast<int*>(Bs + (size_t)(d - 1) * FR * FSP);  // i_k at [k·FS + s]
	float* Rs = reinterpret_cast<float*>(Is + (size_t)d * FS);         // res at [s]; Rs[FS] = 0
	// this wave's sorted sample slots, one copy per bin parity g: byte offset 4·s of the slot's sample when its bin
	// has parity g, else 4·FS, the zero column (Rs, the staging rows and Es are 0 there)
	int* Ps = reinterpret_cast<int*>(Rs + FSP) + w * 2 * FPERM;
	for (int e = tid; e < nslots * PB / 4; e += FW * 64)
		reinterpret_cast<float4*>(Gs)[e] = reinterpret_cast<const float4*>(Gp)[e];
	for (int e = tid; e < FR * FSP; e += FW * 64) Es[e] = e < FS ? 1.f : 0.f;
	for (int e = tid; e < 2 * (d - 1) * FR; e += FW * 64) Fs[(size_t)e * FSP + FS] = 0.f;  // F and B rows, column FS
	if (tid == 0) Rs[FS] = 0.f;
	const int li = lane & 15, lk = lane >> 4;  // MFMA: A row / B column li, K index lk
	gf4 acc[GKC][GNT];
#pragma unroll
	for (int q = 0; q < GKC; ++q)
#pragma unroll
		for (int t = 0; t < GNT; ++t) acc[q][t] = gf4{0.f, 0.f, 0.f, 0.f};
	const int per = (a.P + gridDim.x - 1) / gridDim.x;
	const int s0 = blockIdx.x * per, s1 = min(a.P, s0 + per);
	double rr = 0, bb = 0;  // RES: this lane's residual sums (forward walkers only)
	__syncthreads();
	for (int base = s0; base < s1; base += FS) {
		const int cnt = min(FS, s1 - base);
		if (w < WW) {  // phase 1 walkers (wave-uniform): two lanes per sample, lane half h holds components 4h..4h+3
			const bool fwd = w < WW / 2;
			const int s = (fwd ? w : w - WW / 2) * 32 + (lane >> 1);
			const int h = lane & 1;
			bool valid = s < cnt;
			const uint32_t* p = pos + (size_t)(base + (valid ? s : 0)) * d;
			uint32_t pk[FD];
#pragma unroll
			for (int k = 0; k < FD; ++k) pk[k] = p[k < d ? k : d - 1];
#pragma unroll
			for (int k = 0; k < FD; ++k) valid = valid && (k >= d || pk[k] < (uint32_t)sN[k]);
			if (fwd && h == 0) {
				for (int k = 0; k < d; ++k) Is[k * FS + s] = valid ? (int)pk[k] : 0;
				if (!RES) Rs[s] = valid ? resid[base + s] : 0.f;
			}
			float v[4];
#pragma unroll
			for (int c = 0; c < 4; ++c) v[c] = (h == 0 && c == 0) ? 1.f : 0.f;
			if (fwd) {
#pragma unroll
				for (int k = 0; k + 1 < FD; ++k) {
					if (k + 1 < d) {
						mul_row8_half(v, Gs + sPo[k] + (size_t)(valid ? pk[k] : 0) * PB, h);
#pragma unroll
						for (int c = 0; c < 4; ++c) Fs[(k * FR + 4 * h + c) * FSP + s] = v[c];
					}
				}
				if (RES) {
					mul_row8_half(v, Gs + sPo[d - 1] + (size_t)(valid ? pk[d - 1] : 0) * PB, h);
					if (h == 0) {
						if (s < cnt) {
							const float bi = values[base + s];
							const float r = (valid ? v[0] : NAN) - bi;
							residOut[base + s] = r;
							rr += (double)r * r;
							bb += (double)bi * bi;
							Rs[s] = valid ? r : 0.f;
						} else {
							Rs[s] = 0.f;
						}
					}
				}
			} else {
#pragma unroll
				for (int k = FD - 1; k >= 1; --k) {
					if (k < d) {
						mul_col8_half(v, Gs + sPo[k] + (size_t)(valid ? pk[k] : 0) * PB, h);
#pragma unroll
						for (int al = 0; al < 4; ++al) Bs[((k - 1) * FR + 4 * h + al) * FSP + s] = v[al];
					}
				}
			}
		}
		__syncthreads();
#pragma unroll
		for (int q = 0; q < GKC; ++q) {
			const int k = w + q * FW;
			if (k >= d || (kSel >= 0 && k != kSel)) continue;
			const int nk = sN[k];
			const float* fk = (k > 0 ? Fs + (k - 1) * FR * FSP : Es) + (li & 7) * FSP;
			const float* bk = (k + 1 < d ? Bs + k * FR * FSP : Es) + (li & 7) * FSP;
			const int* ik = Is + k * FS;
			// counting sort of the round's samples by i_k: lane holds samples lane and lane + 64; the bins 2t and 2t+1
			// of tile t are adjacent and the tile is padded to a multiple of 4 slots
			const int ja = lane < cnt ? ik[lane] : -1;
			const int jb = lane + 64 < cnt ? ik[lane + 64] : -1;
			for (int e = lane; e < 2 * FPERM; e += 64) Ps[e] = 4 * FS;
			__builtin_amdgcn_wave_barrier();
			int tstart[GNT], tend[GNT];
			int st = 0;
#pragma unroll
			for (int t = 0; t < GNT; ++t) {
				tstart[t] = st;
				if (2 * t < nk) {
#pragma unroll
					for (int hb = 0; hb < 2; ++hb) {
						const int bin = 2 * t + hb;
						const unsigned long long ma = __ballot(ja == bin), mb = __ballot(jb == bin);
						const int ca = __popcll(ma), cb = __popcll(mb);
						const int below_a = __builtin_amdgcn_mbcnt_hi((unsigned)(ma >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)ma, 0));
						const int below_b = __builtin_amdgcn_mbcnt_hi((unsigned)(mb >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)mb, 0));
						if (ja == bin) Ps[hb * FPERM + st + below_a] = 4 * lane;
						if (jb == bin) Ps[hb * FPERM + st + ca + below_b] = 4 * (lane + 64);
						st += ca + cb;
					}
					st = tstart[t] + ((st - tstart[t] + 3) & ~3);
				}
				tend[t] = st;
			}
			__builtin_amdgcn_wave_barrier();
			// Block-diagonal MFMA tiles: A row i = (g, a) and B column j = (g, c) with g = i >> 3 (j >> 3) the bin parity.
			// The lanes of parity g read the slot copy of parity g, where samples of the other bin (and padding) point at
			// the zero column, so a sample enters the rows and columns of its own bin only: the blocks (0,0) and (1,1)
			// accumulate bins 2t and 2t + 1, and a step takes 4 samples of either bin. Two steps per iteration; the next
			// iteration's slots are loaded with this one's operands (reads past the tile stay inside the wave's slots).
			const int* Pp = Ps + (li >> 3) * FPERM;
			const char* rB = reinterpret_cast<const char*>(Rs);
			const char* fB = reinterpret_cast<const char*>(fk);
			const char* bB = reinterpret_cast<const char*>(bk);
#pragma unroll
			for (int t = 0; t < GNT; ++t) {
				if (2 * t >= nk) continue;
				const int te = tend[t];
				int ks = tstart[t];
				int n0 = Pp[ks + lk], n1 = Pp[ks + 4 + lk];
				for (; ks < te; ks += 8) {
					const int x0 = n0, x1 = n1;
					n0 = Pp[ks + 8 + lk];
					n1 = Pp[ks + 12 + lk];
					const float r0 = *reinterpret_cast<const float*>(rB + x0), f0 = *reinterpret_cast<const float*>(fB + x0);
					const float b0 = *reinterpret_cast<const float*>(bB + x0);
					const float r1 = *reinterpret_cast<const float*>(rB + x1), f1 = *reinterpret_cast<const float*>(fB + x1);
					const float b1 = *reinterpret_cast<const float*>(bB + x1);
					acc[q][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(r0 * f0, b0, acc[q][t], 0, 0, 0);
					if (ks + 4 < te) acc[q][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(r1 * f1, b1, acc[q][t], 0, 0, 0);
				}
			}
		}
		__syncthreads();
	}
	if (RES) {  // the block's residual sums: wave sums, then the waves in order
		__shared__ double sNorm[2 * FW];
		rr = wsum64(rr);
		bb = wsum64(bb);
		if (lane == 0) { sNorm[2 * w] = rr; sNorm[2 * w + 1] = bb; }
		__syncthreads();
		if (tid == 0) {
			double t0 = 0, t1 = 0;
			for (int q = 0; q < FW; ++q) { t0 += sNorm[2 * q]; t1 += sNorm[2 * q + 1]; }
			normPart[2 * blockIdx.x] = t0;
			normPart[2 * blockIdx.x + 1] = t1;
		}
	}
	// the block's accumulators, written in the gradient layout: register r of tile t holds row i = 4·lk + r and column
	// li of the block-diagonal tile, i.e. G_k[a = i & 7, j = 2t + (i >> 3), c = li & 7] when i >> 3 == li >> 3
	float* dst = partials + (size_t)blockIdx.x * a.gradSize;
#pragma unroll
	for (int q = 0; q < GKC; ++q) {
		const int k = w + q * FW;
		if (k >= d) continue;
		const int rl = sRk[k], rr = sRk[k + 1], n = sN[k];
#pragma unroll
		for (int t = 0; t < GNT; ++t) {
#pragma unroll
			for (int r = 0; r < 4; ++r) {
				const int i = 4 * lk + r, al = i & 7, j = 2 * t + (i >> 3), c = li & 7;
				if ((i >> 3) == (li >> 3) && j < n && al < rl && c < rr) dst[sOff[k] + ((size_t)al * n + j) * rr + c] = acc[q][t][r];
			}
		}
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

// grad[e] = Σ_b partials[b][e]: RE entries per block, sixteen threads per entry each summing every sixteenth
// block partial, then the sixteen sums by a fixed pairwise tree (deterministic).
constexpr int RE = 16;
template <class T>
__global__ __launch_bounds__(256) void adf_grad_reduce_kernel(const T* partials, int nblocks, int gradSize, T* grad) {
	__shared__ T part[16][RE + 1];
	const int o = threadIdx.x % RE, q = threadIdx.x / RE;
	const int e = blockIdx.x * RE + o;
	T s = T(0);
	if (e < gradSize)
		for (int b = q; b < nblocks; b += 16) s += partials[(size_t)b * gradSize + e];
	part[q][o] = s;
	__syncthreads();
	if (q == 0 && e < gradSize) {
		T t[8];
#pragma unroll
		for (int i = 0; i < 8; ++i) t[i] = part[2 * i][o] + part[2 * i + 1][o];
		grad[e] = ((t[0] + t[1]) + (t[2] + t[3])) + ((t[4] + t[5]) + (t[6] + t[7]));
	}
}

// norms[0..1] = Σ_b part[2b], Σ_b part[2b + 1]: lane l sums blocks l, l + 64, … in order, then a fixed tree.
__global__ __launch_bounds__(64) void adf_norm_reduce_kernel(const double* part, int nblocks, double* norms) {
	double s0 = 0, s1 = 0;
	for (int b = threadIdx.x; b < nblocks; b += 64) { s0 += part[2 * b]; s1 += part[2 * b + 1]; }
	s0 = wsum64(s0);
	s1 = wsum64(s1);
	if (threadIdx.x == 0) { norms[0] = s0; norms[1] = s1; }
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

// one wave per slice: lane l sums blocks l, l + 64, … in order, then a fixed butterfly (deterministic)
__global__ __launch_bounds__(256) void adf_slice_reduce_kernel(const double* partials, int nblocks, int nslices, double* out) {
	const int lane = threadIdx.x & 63;
	const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
	if (j >= nslices) return;
	double s = 0;
	for (int b = lane; b < nblocks; b += 64) s += partials[(size_t)b * nslices + j];
	s = wsum64(s);
	if (lane == 0) out[j] = s;
}

}  // namespace

namespace {
// ranks <= 8, order <= 16 and the padded cores within 48 KB of LDS
bool fast_ok(const AdfDesc& a) { return a.maxR <= FR && a.d >= 1 && a.d <= FD && (size_t)a.sumN * PB * sizeof(float) <= 48 * 1024; }
struct Padded {
	DBuf<float> Gp;
	DBuf<long long> off;
};
// zero-padded (n_k, 8, 8) copy of the cores (and its core offsets) for the fast kernels
Padded pad_cores(const AdfDesc& a, const float* cores, hipStream_t s) {
	Padded pc;
	const int total = a.sumN * PB;
	pc.Gp.resize((size_t)total);
	pc.off.resize(FD);
	hipLaunchKernelGGL(pad_cores_kernel, dim3(std::max(1, std::min(1024, (total + 255) / 256))), dim3(256), 0, s, a, cores, pc.off.get(),
	                   pc.Gp.get(), total);
	XTT_LAUNCH_CHECK();
	return pc;
}
}  // namespace

template <class T>
void adf_residual(const AdfDesc& a, const T* cores, const uint32_t* pos, const T* values, T* resid, double* norms,
                  double* partials, unsigned* counter, hipStream_t s) {
	XTT_REQUIRE(a.maxR <= RMAX, "ADF: ranks above 16 are not supported");
	XTT_REQUIRE(a.d >= 1 && a.d <= 32, "ADF: order must be between 1 and 32");
	const int blocks = std::max(1, std::min(1024, (a.P + AB - 1) / AB));
	if (std::is_same<T, float>::value && fast_ok(a)) {
		Padded pc = pad_cores(a, reinterpret_cast<const float*>(cores), s);
		// positions, values and residuals of every sample, and the cores once
		launch_timed("adf_residual_fast_kernel<float>", Work(0.0, (double)a.P * (4.0 * a.d + 8.0) + 4.0 * a.gradSize), adf_residual_fast_kernel<T>, dim3(blocks), dim3(AB),
		             (size_t)a.sumN * PB * sizeof(float), s, a, (const float*)pc.Gp.get(), (const long long*)pc.off.get(), pos, values,
		             resid, norms, partials, counter);
		return;
	}
	const size_t coreBytes = (size_t)a.gradSize * sizeof(T);
	const bool lds = coreBytes <= 48 * 1024;
	const char* tag = sizeof(T) == 8 ? "adf_residual_kernel<double>" : "adf_residual_kernel<float>";
	const dim3 g(blocks), t(AB);
	if (a.maxR <= 8) {
		if (lds) launch_timed(tag, 0.0, adf_residual_kernel<T, 8, true>, g, t, coreBytes, s, a, cores, pos, values, resid, norms, partials, counter);
		else launch_timed(tag, 0.0, adf_residual_kernel<T, 8, false>, g, t, 0, s, a, cores, pos, values, resid, norms, partials, counter);
	} else {
		if (lds) launch_timed(tag, 0.0, adf_residual_kernel<T, RMAX, true>, g, t, coreBytes, s, a, cores, pos, values, resid, norms, partials, counter);
		else launch_timed(tag, 0.0, adf_residual_kernel<T, RMAX, false>, g, t, 0, s, a, cores, pos, values, resid, norms, partials, counter);
	}
}

int adf_grad_blocks(int P) { return std::max(1, std::min(512, (P + 255) / 256)); }

namespace {
void allow_fast_grad_lds() {
	static std::once_flag once;
	std::call_once(once, [&] {
		for (const void* f : {reinterpret_cast<const void*>(adf_grad_fast_kernel<false>), reinterpret_cast<const void*>(adf_grad_fast_kernel<true>)})
			XTT_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, 150 * 1024));
	});
}
}  // namespace

bool adf_residual_gradient_fused_ok(const AdfDesc& a) {
	return fast_ok(a) && a.d >= 2 && a.maxN * FR <= GNT * 16 && fast_grad_floats(a.d, a.sumN) * sizeof(float) <= 150 * 1024;
}

static_assert(ADF_PAD_OFFSETS >= FD, "padded core offsets");
int adf_pad_floats(const AdfDesc& a) { return a.sumN * PB; }

// resid = x(pos) − values, norms = {‖resid‖², ‖values‖²} and grad (the adjoint with seed resid) in one gradient
// launch where the fast kernel applies; otherwise adf_residual followed by adf_gradient.
void adf_residual_gradient_f32(const AdfDesc& a, const float* cores, const uint32_t* pos, const float* values, float* resid, double* norms,
                               float* grad, float* partials, double* normPart, unsigned* counter, hipStream_t s) {
	if (!adf_residual_gradient_fused_ok(a)) {
		adf_residual<float>(a, cores, pos, values, resid, norms, normPart, counter, s);
		adf_gradient<float>(a, cores, pos, resid, grad, partials, -1, s);
		return;
	}
	Padded pc;
	pc.Gp.resize((size_t)adf_pad_floats(a));
	pc.off.resize(ADF_PAD_OFFSETS);
	adf_residual_gradient_f32_into(a, cores, pos, values, resid, norms, grad, partials, normPart, pc.Gp.get(), pc.off.get(), s);
}

void adf_residual_gradient_f32_into(const AdfDesc& a, const float* cores, const uint32_t* pos, const float* values, float* resid,
                                    double* norms, float* grad, float* partials, double* normPart, float* Gp, long long* padOff,
                                    hipStream_t s) {
	XTT_REQUIRE(adf_residual_gradient_fused_ok(a), "ADF: shape outside the one-pass residual and gradient kernel");
	const int total = adf_pad_floats(a);
	hipLaunchKernelGGL(pad_cores_kernel, dim3(std::max(1, std::min(1024, (total + 255) / 256))), dim3(256), 0, s, a, cores, padOff, Gp,
	                   total);
	XTT_LAUNCH_CHECK();
	allow_fast_grad_lds();
	const int blocks = std::min(adf_grad_blocks(a.P), 256);
	// positions and values of every sample, residuals out, the cores in and the gradient out
	launch_timed("adf_grad_fast_kernel<float> residual", Work(0.0, (double)a.P * (4.0 * a.d + 8.0) + 8.0 * a.gradSize),
	             adf_grad_fast_kernel<true>, dim3(blocks), dim3(FW * 64), fast_grad_floats(a.d, a.sumN) * sizeof(float), s, a,
	             (const float*)Gp, (const long long*)padOff, pos, (const float*)nullptr, partials, -1, values, resid, normPart);
	hipLaunchKernelGGL(adf_grad_reduce_kernel<float>, dim3((a.gradSize + RE - 1) / RE), dim3(256), 0, s, (const float*)partials, blocks,
	                   a.gradSize, grad);
	XTT_LAUNCH_CHECK();
	hipLaunchKernelGGL(adf_norm_reduce_kernel, dim3(1), dim3(64), 0, s, (const double*)normPart, blocks, norms);
	XTT_LAUNCH_CHECK();
}

template <class T>
void adf_gradient(const AdfDesc& a, const T* cores, const uint32_t* pos, const T* resid, T* grad, T* partials, int kSel,
                  hipStream_t s) {
	XTT_REQUIRE(a.maxR <= RMAX, "ADF: ranks above 16 are not supported");
	XTT_REQUIRE(a.d >= 1 && a.d <= 32, "ADF: order must be between 1 and 32");
	const size_t fastLds = fast_grad_floats(a.d, a.sumN) * sizeof(float);
	if (std::is_same<T, float>::value && fast_ok(a) && a.d >= 2 && a.maxN * FR <= GNT * 16 && fastLds <= 150 * 1024) {
		Padded pc = pad_cores(a, reinterpret_cast<const float*>(cores), s);
		allow_fast_grad_lds();
		const int blocks = std::min(adf_grad_blocks(a.P), 256);
		// positions and residuals of every sample, the cores in and the gradient out
		launch_timed("adf_grad_fast_kernel<float>", Work(0.0, (double)a.P * (4.0 * a.d + 4.0) + 8.0 * a.gradSize), adf_grad_fast_kernel<false>, dim3(blocks), dim3(FW * 64), fastLds, s, a,
		             (const float*)pc.Gp.get(), (const long long*)pc.off.get(), pos, reinterpret_cast<const float*>(resid),
		             reinterpret_cast<float*>(partials), kSel, (const float*)nullptr, (float*)nullptr, (double*)nullptr);
		hipLaunchKernelGGL(adf_grad_reduce_kernel<T>, dim3((a.gradSize + RE - 1) / RE), dim3(256), 0, s, partials, blocks, a.gradSize, grad);
		XTT_LAUNCH_CHECK();
		return;
	}
	const GradLayout L = grad_layout(a.d, a.maxR, a.gradSize, sizeof(T));
	const size_t lds = L.total * sizeof(T);
	if (lds > 64 * 1024)
		XTT_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(adf_grad_kernel<T>), hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
	const int blocks = adf_grad_blocks(a.P);
	launch_timed(sizeof(T) == 8 ? "adf_grad_kernel<double>" : "adf_grad_kernel<float>", 0.0, adf_grad_kernel<T>, dim3(blocks),
	             dim3(AB), lds, s, a, cores, pos, resid, partials, L, kSel);
	hipLaunchKernelGGL(adf_grad_reduce_kernel<T>, dim3((a.gradSize + RE - 1) / RE), dim3(256), 0, s, partials, blocks, a.gradSize, grad);
	XTT_LAUNCH_CHECK();
}

int adf_slice_blocks(int P) { return std::max(1, std::min(256, (P + 255) / 256)); }

template <class T>
void adf_slice_norms(const AdfDesc& a, const T* cores, const uint32_t* pos, int k, const T* D, int nslices, double* out,
                     double* partials, hipStream_t s) {
	const int blocks = adf_slice_blocks(a.P);
	hipLaunchKernelGGL(adf_slice_kernel<T>, dim3(blocks), dim3(AB), 0, s, a, cores, pos, k, D, partials, nslices);
	XTT_LAUNCH_CHECK();
	hipLaunchKernelGGL(adf_slice_reduce_kernel, dim3((nslices + 3) / 4), dim3(256), 0, s, partials, blocks, nslices, out);
	XTT_LAUNCH_CHECK();
}

template <class T>
void adf_sum_partials(const T* partials, int nblocks, int size, T* out, hipStream_t s) {
	hipLaunchKernelGGL(adf_grad_reduce_kernel<T>, dim3((size + RE - 1) / RE), dim3(256), 0, s, partials, nblocks, size, out);
	XTT_LAUNCH_CHECK();
}
void adf_sum_slices(const double* partials, int nblocks, int nslices, double* out, hipStream_t s) {
	hipLaunchKernelGGL(adf_slice_reduce_kernel, dim3((nslices + 3) / 4), dim3(256), 0, s, partials, nblocks, nslices, out);
	XTT_LAUNCH_CHECK();
}
template void adf_sum_partials<float>(const float*, int, int, float*, hipStream_t);
template void adf_sum_partials<double>(const double*, int, int, double*, hipStream_t);

namespace {
// update_x of adf.cpp:468-487 on the device: with the gradient g of ½Σ(x − b)² (xerus's Δ = −g), slice j of the
// core moves by (‖g_j‖²/normA[j])·Δ_j (single points), or the whole core by (‖g‖²/normA[0])·Δ (rank-one).
template <class T>
__global__ __launch_bounds__(256) void adf_update_kernel(double* core, const T* g, const double* normA, int rl, int n, int rr, int perSlice) {
	__shared__ double red[4];
	const int tid = threadIdx.x;
	if (perSlice) {
		for (int j = tid; j < n; j += 256) {
			double pyr = 0;
			for (int al = 0; al < rl; ++al)
				for (int c = 0; c < rr; ++c) { const double t = (double)g[((size_t)al * n + j) * rr + c]; pyr += t * t; }
			const double st = pyr / normA[j];
			for (int al = 0; al < rl; ++al)
				for (int c = 0; c < rr; ++c) {
					const size_t e = ((size_t)al * n + j) * rr + c;
					core[e] -= st * (double)g[e];
				}
		}
		return;
	}
	const int total = rl * n * rr;
	double pyr = 0;
	for (int e = tid; e < total; e += 256) { const double t = (double)g[e]; pyr += t * t; }
	for (int o = 32; o > 0; o >>= 1) pyr += __shfl_xor(pyr, o);
	if ((tid & 63) == 0) red[tid >> 6] = pyr;
	__syncthreads();
	const double st = (red[0] + red[1] + red[2] + red[3]) / normA[0];
	for (int e = tid; e < total; e += 256) core[e] -= st * (double)g[e];
}
}  // namespace

template <class T>
void adf_update_core(double* core, const T* g, const double* normA, int rl, int n, int rr, bool perSlice, hipStream_t s) {
	hipLaunchKernelGGL(adf_update_kernel<T>, dim3(1), dim3(256), 0, s, core, g, normA, rl, n, rr, perSlice ? 1 : 0);
	XTT_LAUNCH_CHECK();
}
template void adf_update_core<float>(double*, const float*, const double*, int, int, int, bool, hipStream_t);
template void adf_update_core<double>(double*, const double*, const double*, int, int, int, bool, hipStream_t);

template void adf_residual<float>(const AdfDesc&, const float*, const uint32_t*, const float*, float*, double*, double*, unsigned*, hipStream_t);
template void adf_residual<double>(const AdfDesc&, const double*, const uint32_t*, const double*, double*, double*, double*, unsigned*, hipStream_t);
template void adf_gradient<float>(const AdfDesc&, const float*, const uint32_t*, const float*, float*, float*, int, hipStream_t);
template void adf_gradient<double>(const AdfDesc&, const double*, const uint32_t*, const double*, double*, double*, int, hipStream_t);
template void adf_slice_norms<float>(const AdfDesc&, const float*, const uint32_t*, int, const float*, int, double*, double*, hipStream_t);
template void adf_slice_norms<double>(const AdfDesc&, const double*, const uint32_t*, int, const double*, int, double*, double*, hipStream_t);

}  // namespace xtt
