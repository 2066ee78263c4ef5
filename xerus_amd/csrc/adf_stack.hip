// ADF with xerus's stacks (adf.cpp:102-312): the plan of shared prefixes and suffixes, built on the device, and the
// sweep kernels that read stack entries instead of walking every sample's chain through all cores.
//
// Plan (adf.cpp:102-191). Samples whose positions agree in modes 0..k share forward entry k; samples agreeing in
// modes k..d-1 share backward entry k. Each direction is built in five passes over the samples:
//   1. a stable LSD radix sort of the sample ids by the direction's key (modes d-1, …, 0 as keys of the passes for
//      the forward order, 0, …, d-1 for the backward one), one hipCUB pass per mode;
//   2. cut masks: in sorted order, sample t opens a group at level k when it differs from sample t−1 in a mode of
//      level k's prefix (suffix); bit k of mask[t];
//   3. a scan per level numbers the runs (groups) of the sorted order;
//   4. an atomic minimum per run gives its representative, the smallest sample id of the group (integer minimum:
//      the result does not depend on the order of the atomics);
//   5. a second scan per level, in id order over the representatives, numbers the entries by ascending
//      representative, the order of xerus's _updates[k]; every sample then reads its entry through its run.
// Stack sweeps: forward level k (F_k = F_{k-1}·G_k[:, i_k, :]) and backward level k (B_k = G_k[:, i_k, :]·B_{k+1})
// touch one entry per group; the residual, gradient and slice norms at core k read F_{k-1} and B_{k+1} of each
// sample through its slots. A sweep over all d cores thus costs O(P·r²) per core instead of a chain of d cores per
// sample per core.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <mutex>

#include "kernels.hpp"
#include "xtt_common.hpp"

namespace xtt {
namespace {

constexpr int PT = 256;          // threads per plan block
constexpr int PQ = 4;            // items per thread in the plan scans
constexpr int PIB = PT * PQ;     // items per scan block
constexpr int SB = 256;          // threads per stack block
constexpr int SR = 16;           // max rank held in registers

__global__ void iota_kernel(uint32_t* v, int P) {
	for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < P; t += gridDim.x * blockDim.x) v[t] = (uint32_t)t;
}

__global__ void gather_mode_kernel(const uint32_t* pos, int d, int j, const uint32_t* perm, uint32_t* keys, int P) {
	for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < P; t += gridDim.x * blockDim.x) keys[t] = pos[(size_t)perm[t] * d + j];
}

// posT[k·P + i] = pos[i·d + k]: the mode-major copy the sweep kernels read (one coalesced column per core)
__global__ void transpose_positions_kernel(const uint32_t* pos, int P, int d, uint32_t* posT) {
	for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < (long long)P * d; e += (long long)gridDim.x * blockDim.x) {
		const long long i = e / d, k = e % d;
		posT[k * P + i] = pos[e];
	}
}

// bit k of mask[t]: sorted sample t opens a group at level k
__global__ void cut_mask_kernel(const uint32_t* pos, int d, const uint32_t* perm, int P, int forward, uint32_t* mask) {
	const uint32_t full = d >= 32 ? ~0u : ((1u << d) - 1u);
	for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < P; t += gridDim.x * blockDim.x) {
		if (t == 0) { mask[0] = full; continue; }
		const uint32_t* a = pos + (size_t)perm[t] * d;
		const uint32_t* b = pos + (size_t)perm[t - 1] * d;
		uint32_t m;
		if (forward) {  // first differing mode c: a new prefix at every level k >= c
			int c = 0;
			while (c < d && a[c] == b[c]) ++c;
			m = c >= d ? 0u : (full & (~0u << c));
		} else {        // last differing mode c: a new suffix at every level k <= c
			int c = d - 1;
			while (c >= 0 && a[c] == b[c]) --c;
			m = c < 0 ? 0u : (full & ((2u << c) - 1u));  // c = 31: 2u << 31 wraps to 0, so the mask is all ones
		}
		mask[t] = m;
	}
}

__device__ inline uint32_t lanes_below(unsigned long long ball) {
	return __builtin_amdgcn_mbcnt_hi((unsigned)(ball >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)ball, 0));
}

// count[k·nb + b] = items of scan block b with bit k set
__global__ __launch_bounds__(PT) void mask_count_kernel(const uint32_t* mask, int P, int d, uint32_t* count) {
	__shared__ uint32_t part[PT / 64][32];
	const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
	uint32_t m[PQ];
#pragma unroll
	for (int q = 0; q < PQ; ++q) {
		const long long i = (long long)blockIdx.x * PIB + q * PT + threadIdx.x;
		m[q] = i < P ? mask[i] : 0u;
	}
	for (int k = 0; k < d; ++k) {
		uint32_t c = 0;
#pragma unroll
		for (int q = 0; q < PQ; ++q) c += (uint32_t)__popcll(__ballot((m[q] >> k) & 1u));
		if (lane == 0) part[w][k] = c;
	}
	__syncthreads();
	if ((int)threadIdx.x < d) {
		uint32_t s = 0;
		for (int q = 0; q < PT / 64; ++q) s += part[q][threadIdx.x];
		count[(size_t)threadIdx.x * gridDim.x + blockIdx.x] = s;
	}
}

// exclusive scan of count[k·nb .. k·nb + nb) in place, one workgroup per level k; total[k] = the level's sum
__global__ __launch_bounds__(PT) void block_scan_kernel(uint32_t* count, int nb, uint32_t* total) {
	__shared__ uint32_t wsum[PT / 64];
	__shared__ uint32_t carry;
	const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
	uint32_t* v = count + (size_t)blockIdx.x * nb;
	if (threadIdx.x == 0) carry = 0;
	__syncthreads();
	for (int base = 0; base < nb; base += PT) {
		const int i = base + threadIdx.x;
		const uint32_t x = i < nb ? v[i] : 0u;
		uint32_t s = x;
#pragma unroll
		for (int o = 1; o < 64; o <<= 1) {
			const uint32_t y = __shfl_up(s, o, 64);
			if (lane >= o) s += y;
		}
		if (lane == 63) wsum[w] = s;
		__syncthreads();
		uint32_t off = carry;
		for (int q = 0; q < w; ++q) off += wsum[q];
		if (i < nb) v[i] = off + s - x;
		__syncthreads();
		if (threadIdx.x == PT - 1) carry = off + s;
		__syncthreads();
	}
	if (threadIdx.x == 0) total[blockIdx.x] = carry;
}

// out[k·P + t] = (items t' <= t with bit k set) − 1: for a cut mask the run of sorted sample t, for a representative
// mask (id order) the entry index of each representative
__global__ __launch_bounds__(PT) void mask_scan_kernel(const uint32_t* mask, int P, int d, const uint32_t* blockOff, uint32_t* out) {
	__shared__ uint32_t part[PQ][PT / 64][32];
	const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nb = gridDim.x;
	uint32_t m[PQ];
#pragma unroll
	for (int q = 0; q < PQ; ++q) {
		const long long i = (long long)blockIdx.x * PIB + q * PT + threadIdx.x;
		m[q] = i < P ? mask[i] : 0u;
	}
	for (int k = 0; k < d; ++k) {
#pragma unroll
		for (int q = 0; q < PQ; ++q) {
			const uint32_t c = (uint32_t)__popcll(__ballot((m[q] >> k) & 1u));
			if (lane == 0) part[q][w][k] = c;
		}
	}
	__syncthreads();
	for (int k = 0; k < d; ++k) {
		uint32_t pre = blockOff[(size_t)k * nb + blockIdx.x];
#pragma unroll
		for (int q = 0; q < PQ; ++q) {
			const uint32_t bit = (m[q] >> k) & 1u;
			const unsigned long long ball = __ballot(bit);
			uint32_t off = pre;
			for (int w2 = 0; w2 < w; ++w2) off += part[q][w2][k];
			const long long i = (long long)blockIdx.x * PIB + q * PT + threadIdx.x;
			if (i < P) out[(size_t)k * P + i] = off + lanes_below(ball) + bit - 1u;
			for (int w2 = 0; w2 < PT / 64; ++w2) pre += part[q][w2][k];
		}
	}
}

// rmin[k·P + run] = smallest sample id of the run; inv[perm[t]] = t. The runs of a level are contiguous in sorted
// order, so a wave first takes the minimum over each of its segments (a segmented scan with shuffles) and only a
// segment's last lane issues the atomic: a level with few long runs does not serialise on a few addresses.
__global__ __launch_bounds__(256) void run_min_kernel(const uint32_t* run, const uint32_t* perm, int P, int d, uint32_t* rmin, uint32_t* inv) {
	const int lane = threadIdx.x & 63;
	const int stride = gridDim.x * blockDim.x;
	for (int base = blockIdx.x * blockDim.x; base < P; base += stride) {  // every lane runs the same trip count
		const int t = base + threadIdx.x;
		const bool in = t < P;
		const uint32_t id = in ? perm[t] : 0xFFFFFFFFu;
		if (in) inv[id] = (uint32_t)t;
		for (int k = 0; k < d; ++k) {
			const uint32_t r = in ? run[(size_t)k * P + t] : 0xFFFFFFFFu;
			uint32_t v = id;
#pragma unroll
			for (int o = 1; o < 64; o <<= 1) {
				const uint32_t ov = __shfl_up(v, o, 64), orun = __shfl_up(r, o, 64);
				if (lane >= o && orun == r) v = min(v, ov);
			}
			const uint32_t nextRun = __shfl_down(r, 1, 64);
			const bool lastOfSegment = lane == 63 || nextRun != r;
			if (in && lastOfSegment) atomicMin(rmin + (size_t)k * P + r, v);
		}
	}
}

// bit k of mask[i] (id order): sample i represents its group of level k
__global__ void rep_mask_kernel(const uint32_t* run, const uint32_t* rmin, const uint32_t* inv, int P, int d, uint32_t* mask) {
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < P; i += gridDim.x * blockDim.x) {
		const uint32_t t = inv[i];
		uint32_t m = 0;
		for (int k = 0; k < d; ++k)
			if (rmin[(size_t)k * P + run[(size_t)k * P + t]] == (uint32_t)i) m |= 1u << k;
		mask[i] = m;
	}
}

// slot[k·P + i] = entry of sample i at level k (its representative's index); rep[first_k + entry] = representative
__global__ void slot_kernel(const uint32_t* run, const uint32_t* rmin, const uint32_t* inv, const uint32_t* entry, const uint32_t* mask,
                            const uint32_t* total, int P, int d, uint32_t* slot, uint32_t* rep) {
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < P; i += gridDim.x * blockDim.x) {
		const uint32_t t = inv[i];
		long long first = 0;
		for (int k = 0; k < d; ++k) {
			const uint32_t r = rmin[(size_t)k * P + run[(size_t)k * P + t]];
			slot[(size_t)k * P + i] = entry[(size_t)k * P + r];
			if ((mask[i] >> k) & 1u) rep[first + entry[(size_t)k * P + i]] = (uint32_t)i;
			first += total[k];
		}
	}
}

int grid_for(long long n, int threads = 256) { return (int)std::max<long long>(1, std::min<long long>(2048, (n + threads - 1) / threads)); }

// out[k·P + t] = (items t' <= t with bit k of mask) − 1 for every level; totals (device, d entries) = per-level counts
void scan_columns(const uint32_t* mask, int P, int d, uint32_t* out, uint32_t* totals, hipStream_t s) {
	const int nb = (P + PIB - 1) / PIB;
	DBuf<uint32_t> count((size_t)d * nb);
	hipLaunchKernelGGL(mask_count_kernel, dim3(nb), dim3(PT), 0, s, mask, P, d, count.get());
	XTT_LAUNCH_CHECK();
	hipLaunchKernelGGL(block_scan_kernel, dim3(d), dim3(PT), 0, s, count.get(), nb, totals);
	XTT_LAUNCH_CHECK();
	hipLaunchKernelGGL(mask_scan_kernel, dim3(nb), dim3(PT), 0, s, mask, P, d, (const uint32_t*)count.get(), out);
	XTT_LAUNCH_CHECK();
}

void build_side(AdfPlanSide& side, bool forward, const uint32_t* pos, int P, int d, const std::vector<uint32_t>& maxPos, hipStream_t s) {
	const size_t dP = (size_t)d * P;
	DBuf<uint32_t> permA(P), permB(P), keysA(P), keysB(P), mask(P), inv(P), run(dP), rmin(dP), entry(dP), totals(d);
	hipLaunchKernelGGL(iota_kernel, dim3(grid_for(P)), dim3(256), 0, s, permA.get(), P);
	XTT_LAUNCH_CHECK();
	// 1. stable LSD radix sort: the least significant mode first
	size_t tmpBytes = 0;
	XTT_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmpBytes, keysA.get(), keysB.get(), permA.get(), permB.get(), P, 0, 32, s));
	DBuf<unsigned char> tmp(std::max<size_t>(tmpBytes, 1));
	uint32_t* perm = permA.get();
	uint32_t* permOut = permB.get();
	for (int t = 0; t < d; ++t) {
		const int j = forward ? d - 1 - t : t;
		if (maxPos[j] == 0) continue;  // one value: nothing to order
		const int bits = 32 - __builtin_clz(maxPos[j]);
		hipLaunchKernelGGL(gather_mode_kernel, dim3(grid_for(P)), dim3(256), 0, s, pos, d, j, (const uint32_t*)perm, keysA.get(), P);
		XTT_LAUNCH_CHECK();
		XTT_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.get(), tmpBytes, keysA.get(), keysB.get(), perm, permOut, P, 0, bits, s));
		std::swap(perm, permOut);
	}
	// 2.-3. cut masks and the run of every sorted sample per level
	hipLaunchKernelGGL(cut_mask_kernel, dim3(grid_for(P)), dim3(256), 0, s, pos, d, (const uint32_t*)perm, P, forward ? 1 : 0, mask.get());
	XTT_LAUNCH_CHECK();
	scan_columns(mask.get(), P, d, run.get(), totals.get(), s);
	// 4. representatives
	XTT_HIP(hipMemsetAsync(rmin.get(), 0xFF, dP * sizeof(uint32_t), s));
	hipLaunchKernelGGL(run_min_kernel, dim3(grid_for(P)), dim3(256), 0, s, (const uint32_t*)run.get(), (const uint32_t*)perm, P, d, rmin.get(),
	                   inv.get());
	XTT_LAUNCH_CHECK();
	// 5. entries by ascending representative, slots and representative lists
	hipLaunchKernelGGL(rep_mask_kernel, dim3(grid_for(P)), dim3(256), 0, s, (const uint32_t*)run.get(), (const uint32_t*)rmin.get(),
	                   (const uint32_t*)inv.get(), P, d, mask.get());
	XTT_LAUNCH_CHECK();
	scan_columns(mask.get(), P, d, entry.get(), totals.get(), s);
	std::vector<uint32_t> cnt(d);
	read_back(cnt.data(), totals.get(), (size_t)d * sizeof(uint32_t), s);
	side.count.assign(d, 0);
	side.first.assign(d, 0);
	long long all = 0;
	for (int k = 0; k < d; ++k) { side.first[k] = all; side.count[k] = cnt[k]; all += cnt[k]; }
	side.slot.resize(dP);
	side.rep.resize((size_t)std::max<long long>(all, 1));
	hipLaunchKernelGGL(slot_kernel, dim3(grid_for(P)), dim3(256), 0, s, (const uint32_t*)run.get(), (const uint32_t*)rmin.get(),
	                   (const uint32_t*)inv.get(), (const uint32_t*)entry.get(), (const uint32_t*)mask.get(), (const uint32_t*)totals.get(), P, d,
	                   side.slot.get(), side.rep.get());
	XTT_LAUNCH_CHECK();
	XTT_HIP(hipStreamSynchronize(s));  // the temporaries above are released on return
}

// ---- stack kernels ---------------------------------------------------------------------------------------------

template <class T>
constexpr int core_cap() { return (32 * 1024) / (int)sizeof(T); }

// LDS copy of core G (rl × n × rr) slice by slice: slice j at j·SS, entry (a, c) at a·rp + c, rp = rr rounded up to 4,
// SS = rl·rp + 4 (the 4 extra elements spread the slices of neighbouring lanes over the banks), zero padded; every
// row starts on a 16-byte boundary for float. Fits when n·SS <= core_cap.
struct SlicePad {
	int rp, SS;
	bool fits;
};
template <class T>
__host__ __device__ inline SlicePad slice_pad(int rl, int n, int rr) {
	SlicePad p;
	p.rp = (rr + 3) & ~3;
	p.SS = rl * p.rp + 4;
	p.fits = (long long)n * p.SS <= core_cap<T>();
	return p;
}
template <class T>
__device__ inline void stage_slices(const T* G, int rl, int n, int rr, const SlicePad& p, T* lds) {
	for (int e = threadIdx.x; e < n * p.SS; e += blockDim.x) {
		const int j = e / p.SS, w = e % p.SS, a = w / p.rp, c = w % p.rp;
		lds[e] = (a < rl && c < rr) ? G[((size_t)a * n + j) * rr + c] : T(0);
	}
	__syncthreads();
}

// f·G[:, j, :]·b: from the padded LDS slices (PAD) or from G in place
template <bool PAD, class T>
__device__ inline T sandwich(const T* G, int rl, int n, int rr, const SlicePad& p, uint32_t j, const T (&f)[SR], const T (&b)[SR]) {
	T v = T(0);
#pragma unroll
	for (int a = 0; a < SR; ++a) {
		if (a < rl) {
			T t = T(0);
			if (PAD) {
				const T* row = G + (size_t)j * p.SS + a * p.rp;
#pragma unroll
				for (int c = 0; c < SR; c += 4) {
					if (c < rr) {
						t = fma(row[c], b[c], t);
						t = fma(row[c + 1], b[c + 1], t);
						t = fma(row[c + 2], b[c + 2], t);
						t = fma(row[c + 3], b[c + 3], t);
					}
				}
			} else {
				const T* row = G + ((size_t)a * n + j) * rr;
#pragma unroll
				for (int c = 0; c < SR; ++c) if (c < rr) t = fma(row[c], b[c], t);
			}
			v = fma(f[a], t, v);
		}
	}
	return v;
}

template <class T>
__global__ __launch_bounds__(SB) void fstack_kernel(const T* core, int rl, int n, int rr, int k, int d, const uint32_t* pos, const uint32_t* rep,
                                                   long long count, const uint32_t* prevSlot, const T* prev, T* out) {
	__shared__ __align__(16) T sg[core_cap<T>()];
	const SlicePad p = slice_pad<T>(rl, n, rr);
	if (p.fits) stage_slices(core, rl, n, rr, p, sg);
	for (long long g = (long long)blockIdx.x * SB + threadIdx.x; g < count; g += (long long)gridDim.x * SB) {
		const uint32_t i = rep[g];
		const uint32_t j = pos[i];
		T f[SR];
		if (prev) {
			const T* pv = prev + (size_t)prevSlot[i] * rl;
#pragma unroll
			for (int a = 0; a < SR; ++a) f[a] = a < rl ? pv[a] : T(0);
		} else {
#pragma unroll
			for (int a = 0; a < SR; ++a) f[a] = a == 0 ? T(1) : T(0);
		}
		T o[SR];
#pragma unroll
		for (int c = 0; c < SR; ++c) o[c] = T(0);
#pragma unroll
		for (int a = 0; a < SR; ++a) {
			if (a < rl) {
				const T* row = p.fits ? sg + (size_t)j * p.SS + a * p.rp : core + ((size_t)a * n + j) * rr;
#pragma unroll
				for (int c = 0; c < SR; ++c) if (c < rr) o[c] = fma(f[a], row[c], o[c]);
			}
		}
		T* dst = out + (size_t)g * rr;
#pragma unroll
		for (int c = 0; c < SR; ++c) if (c < rr) dst[c] = o[c];
	}
}

template <class T>
__global__ __launch_bounds__(SB) void bstack_kernel(const T* core, int rl, int n, int rr, int k, int d, const uint32_t* pos, const uint32_t* rep,
                                                   long long count, const uint32_t* nextSlot, const T* next, T* out) {
	__shared__ __align__(16) T sg[core_cap<T>()];
	const SlicePad p = slice_pad<T>(rl, n, rr);
	if (p.fits) stage_slices(core, rl, n, rr, p, sg);
	for (long long g = (long long)blockIdx.x * SB + threadIdx.x; g < count; g += (long long)gridDim.x * SB) {
		const uint32_t i = rep[g];
		const uint32_t j = pos[i];
		T b[SR];
		if (next) {
			const T* nv = next + (size_t)nextSlot[i] * rr;
#pragma unroll
			for (int c = 0; c < SR; ++c) b[c] = c < rr ? nv[c] : T(0);
		} else {
#pragma unroll
			for (int c = 0; c < SR; ++c) b[c] = c == 0 ? T(1) : T(0);
		}
		T* dst = out + (size_t)g * rl;
#pragma unroll
		for (int a = 0; a < SR; ++a) {
			if (a < rl) {
				const T* row = p.fits ? sg + (size_t)j * p.SS + a * p.rp : core + ((size_t)a * n + j) * rr;
				T v = T(0);
#pragma unroll
				for (int c = 0; c < SR; ++c) if (c < rr) v = fma(row[c], b[c], v);
				dst[a] = v;
			}
		}
	}
}

// The left and right stack entries of sample i (ones outside the chain)
template <class T>
__device__ inline void load_sides(const AdfSides<T>& sd, uint32_t i, int rl, int rr, T (&f)[SR], T (&b)[SR]) {
	if (sd.F) {
		const T* p = sd.F + (size_t)sd.fSlot[i] * rl;
#pragma unroll
		for (int a = 0; a < SR; ++a) f[a] = a < rl ? p[a] : T(0);
	} else {
#pragma unroll
		for (int a = 0; a < SR; ++a) f[a] = a == 0 ? T(1) : T(0);
	}
	if (sd.B) {
		const T* p = sd.B + (size_t)sd.bSlot[i] * rr;
#pragma unroll
		for (int c = 0; c < SR; ++c) b[c] = c < rr ? p[c] : T(0);
	} else {
#pragma unroll
		for (int c = 0; c < SR; ++c) b[c] = c == 0 ? T(1) : T(0);
	}
}

__device__ inline double wave_sum(double v) {
#pragma unroll
	for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
	return v;
}

// Residual and gradient at core k in one pass (calculate_residual + calculate_projected_gradient, adf.cpp:290-396).
// Rounds of RG samples per block: thread t evaluates sample t of the round (x_s = F·G_k[:, i_k, :]·B from its stack
// entries, res_s = x_s − b_s, written to resid) and stages F, B, res and i_k in LDS; then thread (q, a, c) owns
// accumulator acc[q][j][a][c] of every slice j and adds res_s·F_s[a]·B_s[c] for the round's samples s ≡ q in order
// (no two threads write one address: deterministic). At the end the Q copies are summed in q order into the block's
// partial gradient (all slices) and the block's FP64 Σ res², Σ b² go to normPart.
template <class T> constexpr int resgrad_round() { return sizeof(T) == 4 ? 256 : 128; }
template <class T>
__global__ __launch_bounds__(SB) void stack_resgrad_kernel(const T* core, int rl, int n, int rr, int k, int d, const uint32_t* pos, int P,
                                                          AdfSides<T> sd, const T* values, T* resid, T* partials, double* normPart, int j0,
                                                          int j1) {
	constexpr int RG = resgrad_round<T>();
	extern __shared__ __align__(16) unsigned char gsmem[];
	const SlicePad p = slice_pad<T>(rl, n, rr);
	const int E = rl * rr, Q = SB / E;
	T* sg = reinterpret_cast<T*>(gsmem);                   // padded core slices (when they fit)
	const int nj = j1 - j0;                                 // the slices accumulated by this launch
	T* acc = sg + (p.fits ? n * p.SS : 0);                 // [Q][nj][E]
	// staging entry-major with an odd stride (RG + 1): the writes of one entry index go to consecutive addresses and
	// the owners' reads of entries a (or c) of one sample fall in different banks
	T* sF = acc + (size_t)Q * nj * E;                      // [rl][RG + 1]
	T* sB = sF + (RG + 1) * rl;                            // [rr][RG + 1]
	T* sR = sB + (RG + 1) * rr;                            // [RG]
	int* sJ = reinterpret_cast<int*>(sR + RG);             // [RG]
	__shared__ double ws[2 * SB / 64];
	const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
	for (int e = tid; e < Q * nj * E; e += SB) acc[e] = T(0);
	if (p.fits) stage_slices(core, rl, n, rr, p, sg);      // has a barrier
	const T* G = p.fits ? sg : core;
	const int q = tid / E, e = tid % E, a = e / rr, c = e % rr;
	const bool owner = q < Q;
	const int per = (P + gridDim.x - 1) / gridDim.x;
	const int s0 = blockIdx.x * per, s1 = min(P, s0 + per);
	double r2 = 0, b2 = 0;
	for (int base = s0; base < s1; base += RG) {
		const int cnt = min(RG, s1 - base);
		__syncthreads();
		if (tid < RG) {
			if (tid < cnt) {
				const uint32_t i = (uint32_t)(base + tid);
				T f[SR], b[SR];
				load_sides(sd, i, rl, rr, f, b);
				const uint32_t j = pos[i];
				const T v = p.fits ? sandwich<true>(G, rl, n, rr, p, j, f, b) : sandwich<false>(G, rl, n, rr, p, j, f, b);
				const T bi = values[i];
				const T r = v - bi;
				resid[i] = r;
				r2 += (double)r * r;
				b2 += (double)bi * bi;
#pragma unroll
				for (int x = 0; x < SR; ++x) if (x < rl) sF[x * (RG + 1) + tid] = f[x];
#pragma unroll
				for (int x = 0; x < SR; ++x) if (x < rr) sB[x * (RG + 1) + tid] = b[x];
				sR[tid] = r;
				sJ[tid] = (int)j - j0;
			}
		}
		__syncthreads();
		if (owner) {
			T* my = acc + (size_t)q * nj * E + e;
			for (int s = q; s < cnt; s += Q) {
				const int jj = sJ[s];
				if (jj >= 0 && jj < nj) my[(size_t)jj * E] += sR[s] * sF[a * (RG + 1) + s] * sB[c * (RG + 1) + s];
			}
		}
	}
	__syncthreads();
	T* dst = partials + (size_t)blockIdx.x * rl * n * rr;
	for (int x = tid; x < nj * E; x += SB) {
		T v = T(0);
		for (int qq = 0; qq < Q; ++qq) v += acc[(size_t)qq * nj * E + x];
		const int jj = x / E, ee = x % E;
		dst[((size_t)(ee / rr) * n + j0 + jj) * rr + ee % rr] = v;
	}
	r2 = wave_sum(r2);
	b2 = wave_sum(b2);
	if (lane == 0) { ws[2 * wave] = r2; ws[2 * wave + 1] = b2; }
	__syncthreads();
	if (tid == 0) {
		double t0 = 0, t1 = 0;
		for (int w = 0; w < SB / 64; ++w) { t0 += ws[2 * w]; t1 += ws[2 * w + 1]; }
		normPart[2 * blockIdx.x] = t0;
		normPart[2 * blockIdx.x + 1] = t1;
	}
}

// norms = {Σ_b normPart[2b], Σ_b normPart[2b + 1]}: lane l sums blocks l, l + 64, … in order, then a fixed tree
__global__ __launch_bounds__(64) void stack_norm_reduce_kernel(const double* part, int nblocks, double* norms) {
	double s0 = 0, s1 = 0;
	for (int b = threadIdx.x; b < nblocks; b += 64) { s0 += part[2 * b]; s1 += part[2 * b + 1]; }
	s0 = wave_sum(s0);
	s1 = wave_sum(s1);
	if (threadIdx.x == 0) { norms[0] = s0; norms[1] = s1; }
}

// Slice norms of a direction D of core k: lane l of each wave evaluates one sample per round (v = F·D[:, i_k, :]·B),
// then the wave walks the round's 64 samples in order and the lane owning bin i_k (bins j ≡ l mod 64 in the wave's
// LDS row) adds v². Waves are summed in order into the block's partial.
template <class T>
__global__ __launch_bounds__(SB) void stack_slice_kernel(const T* D, int rl, int n, int rr, int k, int d, const uint32_t* pos, int P,
                                                        AdfSides<T> sd, double* partials) {
	extern __shared__ __align__(16) unsigned char ssmem[];
	const SlicePad p = slice_pad<T>(rl, n, rr);
	T* sg = reinterpret_cast<T*>(ssmem);
	double* acc = reinterpret_cast<double*>(sg + (p.fits ? ((n * p.SS + 1) & ~1) : 0));  // [SB / 64][n]
	const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
	double* my = acc + (size_t)wave * n;
	for (int j = lane; j < n; j += 64) my[j] = 0.0;
	if (p.fits) stage_slices(D, rl, n, rr, p, sg);  // has a barrier
	const T* G = p.fits ? sg : D;
	const int per = (P + gridDim.x - 1) / gridDim.x;
	const int s0 = blockIdx.x * per, s1 = min(P, s0 + per);
	for (int base = s0 + wave * 64; base < s1; base += SB) {
		const int cnt = min(64, s1 - base);
		double v2 = 0;
		int j = 0;
		if (lane < cnt) {
			const uint32_t i = (uint32_t)(base + lane);
			T f[SR], b[SR];
			load_sides(sd, i, rl, rr, f, b);
			j = (int)pos[i];
			const double v = (double)(p.fits ? sandwich<true>(G, rl, n, rr, p, (uint32_t)j, f, b) : sandwich<false>(G, rl, n, rr, p, (uint32_t)j, f, b));
			v2 = v * v;
		}
		for (int s = 0; s < cnt; ++s) {
			const int js = __shfl(j, s, 64);
			const double vs = __shfl(v2, s, 64);
			if ((js & 63) == lane) my[js] += vs;
		}
	}
	__syncthreads();
	for (int j = tid; j < n; j += SB) {
		double t = 0;
		for (int w = 0; w < SB / 64; ++w) t += acc[(size_t)w * n + j];
		partials[(size_t)blockIdx.x * n + j] = t;
	}
}

// update_x (adf.cpp:468-478) from the blocks' slice partials: normA_j = Σ_b partials[b][j] in block order, then
// slice j of the core moves by −(‖g_j‖²/normA_j)·g_j (g the gradient of ½Σ(x − b)², xerus's Δ = −g).
// One wave per slice: lane l sums the partials of blocks l, l + 64, … in order and the lanes are added by a fixed
// butterfly (deterministic); the same wave forms ‖g_j‖² and updates the slice's rl·rr entries.
template <class T>
__global__ __launch_bounds__(256) void stack_update_kernel(double* core, const T* g, const double* partials, int nblocks, int rl, int n, int rr) {
	const int lane = threadIdx.x & 63;
	const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
	if (j >= n) return;
	double na = 0;
	for (int b = lane; b < nblocks; b += 64) na += partials[(size_t)b * n + j];
	na = wave_sum(na);
	const int E = rl * rr;
	double pyr = 0;
	for (int e = lane; e < E; e += 64) {
		const double t = (double)g[((size_t)(e / rr) * n + j) * rr + e % rr];
		pyr += t * t;
	}
	pyr = wave_sum(pyr);
	const double st = pyr / na;
	for (int e = lane; e < E; e += 64) {
		const size_t idx = ((size_t)(e / rr) * n + j) * rr + e % rr;
		core[idx] -= st * (double)g[idx];
	}
}

}  // namespace

int adf_stack_blocks(int P) { return std::max(1, std::min(512, (P + SB - 1) / SB)); }

void adf_plan_build(AdfPlan& plan, const uint32_t* pos, int P, int d, const std::vector<uint32_t>& maxPos, hipStream_t s) {
	XTT_REQUIRE(d >= 1 && d <= 32, "ADF stacks: order must be between 1 and 32");
	XTT_REQUIRE(P >= 1, "ADF stacks: empty measurement set");
	XTT_REQUIRE(maxPos.size() == (size_t)d, "ADF stacks: one maximal position per mode");
	plan.d = d;
	plan.P = P;
	plan.posT.resize((size_t)d * P);
	hipLaunchKernelGGL(transpose_positions_kernel, dim3(grid_for((long long)P * d)), dim3(256), 0, s, pos, P, d, plan.posT.get());
	XTT_LAUNCH_CHECK();
	build_side(plan.side[0], true, pos, P, d, maxPos, s);
	build_side(plan.side[1], false, pos, P, d, maxPos, s);
}

template <class T>
void adf_forward_stack(const T* core, int rl, int n, int rr, int k, int d, const uint32_t* pos, int P, const uint32_t* rep, long long count,
                       const uint32_t* prevSlot, const T* prev, T* out, hipStream_t s) {
	XTT_REQUIRE(rl <= SR && rr <= SR, "ADF stacks: ranks above 16 are not supported");
	(void)P;
	// per entry: representative, its position, the parent entry and the result
	launch_timed(sizeof(T) == 8 ? "adf_fstack_kernel<double>" : "adf_fstack_kernel<float>",
	             Work(2.0 * count * rl * rr, (double)count * (8.0 + sizeof(T) * (rl + rr))), fstack_kernel<T>, dim3(grid_for(count, SB)), dim3(SB), 0,
	             s, core, rl, n, rr, k, d, pos, rep, count, prevSlot, prev, out);
}

template <class T>
void adf_backward_stack(const T* core, int rl, int n, int rr, int k, int d, const uint32_t* pos, int P, const uint32_t* rep, long long count,
                        const uint32_t* nextSlot, const T* next, T* out, hipStream_t s) {
	XTT_REQUIRE(rl <= SR && rr <= SR, "ADF stacks: ranks above 16 are not supported");
	(void)P;
	launch_timed(sizeof(T) == 8 ? "adf_bstack_kernel<double>" : "adf_bstack_kernel<float>",
	             Work(2.0 * count * rl * rr, (double)count * (8.0 + sizeof(T) * (rl + rr))), bstack_kernel<T>, dim3(grid_for(count, SB)), dim3(SB), 0,
	             s, core, rl, n, rr, k, d, pos, rep, count, nextSlot, next, out);
}

namespace {
constexpr size_t STACK_LDS = 150 * 1024;  // dynamic LDS limit of the stack kernels (their static LDS comes on top)
// LDS of the residual-gradient kernel for nj slices per launch
template <class T>
size_t resgrad_lds(int rl, int n, int rr, int nj) {
	const int E = rl * rr, Q = SB / E, RG = resgrad_round<T>();
	const SlicePad p = slice_pad<T>(rl, n, rr);
	return ((p.fits ? (size_t)n * p.SS : 0) + (size_t)Q * nj * E + (size_t)(RG + 1) * (rl + rr) + RG) * sizeof(T) + RG * sizeof(int);
}
// the kernels with dynamic LDS may take all of it (once per process; the attribute holds for every device)
void allow_stack_lds() {
	static std::once_flag once;
	std::call_once(once, [] {
		XTT_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(stack_resgrad_kernel<float>), hipFuncAttributeMaxDynamicSharedMemorySize, (int)STACK_LDS));
		XTT_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(stack_resgrad_kernel<double>), hipFuncAttributeMaxDynamicSharedMemorySize, (int)STACK_LDS));
		XTT_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(stack_slice_kernel<float>), hipFuncAttributeMaxDynamicSharedMemorySize, (int)STACK_LDS));
		XTT_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(stack_slice_kernel<double>), hipFuncAttributeMaxDynamicSharedMemorySize, (int)STACK_LDS));
	});
}
}  // namespace

template <class T>
void adf_stack_residual_gradient(const T* core, int rl, int n, int rr, int k, int d, const uint32_t* pos, int P, const AdfSides<T>& sd,
                                 const T* values, T* resid, double* norms, T* grad, T* partials, double* normPart, hipStream_t s) {
	XTT_REQUIRE(rl <= SR && rr <= SR, "ADF stacks: ranks above 16 are not supported");
	XTT_REQUIRE(n <= ADF_STACK_MAX_N, "ADF stacks: mode size above the limit of the stack kernels");
	allow_stack_lds();
	// slices per launch: all of them when the accumulators fit, else chunks (each launch recomputes the residuals)
	int chunk = n;
	while (chunk > 1 && resgrad_lds<T>(rl, n, rr, chunk) > STACK_LDS) chunk = (chunk + 1) / 2;
	XTT_REQUIRE(resgrad_lds<T>(rl, n, rr, chunk) <= STACK_LDS, "ADF stacks: core slices too large for the LDS");
	const int blocks = adf_stack_blocks(P);
	for (int j0 = 0; j0 < n; j0 += chunk) {
		const int j1 = std::min(n, j0 + chunk);
		// per sample: position, two slots, both stack entries, value in and residual out; the blocks' partials out
		launch_timed(sizeof(T) == 8 ? "adf_stack_resgrad_kernel<double>" : "adf_stack_resgrad_kernel<float>",
		             Work(2.0 * P * (rl * rr + rl) + 3.0 * P * rl * rr,
		                  (double)P * (12.0 + sizeof(T) * (rl + rr + 2)) + (double)blocks * rl * (j1 - j0) * rr * sizeof(T)),
		             stack_resgrad_kernel<T>, dim3(blocks), dim3(SB), resgrad_lds<T>(rl, n, rr, j1 - j0), s, core, rl, n, rr, k, d, pos, P, sd, values,
		             resid, partials, normPart, j0, j1);
	}
	adf_sum_partials<T>(partials, blocks, rl * n * rr, grad, s);
	if (norms) {
		hipLaunchKernelGGL(stack_norm_reduce_kernel, dim3(1), dim3(64), 0, s, (const double*)normPart, blocks, norms);
		XTT_LAUNCH_CHECK();
	}
}

template <class T>
void adf_stack_step(double* core, const T* g, int rl, int n, int rr, int k, int d, const uint32_t* pos, int P, const AdfSides<T>& sd,
                    double* partials, hipStream_t s) {
	XTT_REQUIRE(rl <= SR && rr <= SR, "ADF stacks: ranks above 16 are not supported");
	XTT_REQUIRE(n <= ADF_STACK_MAX_N, "ADF stacks: mode size above the limit of the stack kernels");
	allow_stack_lds();
	const int blocks = adf_stack_blocks(P);
	const SlicePad p = slice_pad<T>(rl, n, rr);
	const size_t lds = (p.fits ? (size_t)((n * p.SS + 1) & ~1) * sizeof(T) : 0) + (size_t)(SB / 64) * n * sizeof(double);
	launch_timed(sizeof(T) == 8 ? "adf_stack_slice_kernel<double>" : "adf_stack_slice_kernel<float>",
	             Work(2.0 * P * (rl * rr + rl), (double)P * (12.0 + sizeof(T) * (rl + rr))), stack_slice_kernel<T>, dim3(blocks), dim3(SB), lds, s,
	             g, rl, n, rr, k, d, pos, P, sd, partials);
	hipLaunchKernelGGL(stack_update_kernel<T>, dim3((n + 3) / 4), dim3(256), 0, s, core, g, (const double*)partials, blocks, rl, n, rr);
	XTT_LAUNCH_CHECK();
}

#define XTT_ADF_STACK_INST(T)                                                                                                                   \
	template void adf_forward_stack<T>(const T*, int, int, int, int, int, const uint32_t*, int, const uint32_t*, long long, const uint32_t*,      \
	                                   const T*, T*, hipStream_t);                                                                             \
	template void adf_backward_stack<T>(const T*, int, int, int, int, int, const uint32_t*, int, const uint32_t*, long long, const uint32_t*,     \
	                                    const T*, T*, hipStream_t);                                                                            \
	template void adf_stack_residual_gradient<T>(const T*, int, int, int, int, int, const uint32_t*, int, const AdfSides<T>&, const T*, T*,      \
	                                             double*, T*, T*, double*, hipStream_t);                                                       \
	template void adf_stack_step<T>(double*, const T*, int, int, int, int, int, const uint32_t*, int, const AdfSides<T>&, double*, hipStream_t);
XTT_ADF_STACK_INST(float)
XTT_ADF_STACK_INST(double)

}  // namespace xtt
