// Fused local CG of an ALS site: one launch runs every iteration of the CG solve of
// A_loc·x = b (als.cpp:268-296 solves the same local system densely), with the local operator
//   y(a,n,c) = Σ L(a,p,a')·A(p,n,n',q)·R(c,q,c')·x(a',n',c')
// applied matrix-free. The iteration is Chronopoulos–Gear CG (A. T. Chronopoulos, C. W. Gear, "s-step
// iterative methods for symmetric linear systems", J. Comput. Appl. Math. 25 (1989)): w = A·r, γ = rᵀr and
// δ = wᵀr come out of one reduction, then
//   β = γ/γ_old, α = γ/(δ − β·γ/α_old)   (first step: β = 0, α = γ/δ)
//   p = r + β·p, s = w + β·s, x += α·p, r −= α·s
// which produces the iterates of standard CG (s = A·p) with one global reduction per step.
//
// Work split per step, two grid barriers:
//   stage 1, unit (c', a-block of 16): the vector update of the c' slice (every a', computed redundantly by
//     the slice's units, written by the unit owning the a-block), then
//     T1[(a,p)][n'] = Σ_a' L[(a,p)][a']·r[a'][n'][c']         (MFMA, M = 16·P, N = n, K = a)
//     T2[a][c'][(n,q)] = Σ_(p,n') T1[a][(p,n')]·Am[(n,q)][(p,n')]   (MFMA, M = 16, N = n·Q, K = P·n)
//   grid barrier
//   stage 3, unit (a, c-block of 16): w[c][a][n] = Σ_(c',q) R[c][q][c']·T2[a][c'][(n,q)]  (MFMA, K = c·Q),
//     with the unit's partial γ and δ (FP64) in its slot
//   grid barrier with the sum of the partials: workgroup 0 adds the slots in slot order and broadcasts the
//   totals, so every workgroup takes the same decisions.
// L, Am and R are read once (FP64, converted to the arithmetic type) into LDS and stay there for the whole
// solve. The CG vectors use the layout (c, a, n), which makes a c' slice contiguous; r and s are double
// buffered because the units of a slice read the whole slice while its owners write their part.
//
// Cross-workgroup data (the r and s slices, w, T2, the barrier slots) is written and read only with
// agent-scope relaxed atomic accesses, which gfx950 performs coherently across the XCDs' L2 caches (sc1). A
// workgroup publishes after every thread has waited for its own stores (s_waitcnt vmcnt(0)), so whoever reads
// the tag reads the data after it; no L2 write-back or invalidation (agent-scope fence) is needed, which on
// MI355X costs about 20 µs per barrier at 256 workgroups against 3 µs for the tag exchange itself.
// Barriers use no shared counter: every workgroup publishes a tag (and its partials) in a slot of its own,
// workgroup 0 polls the slots and publishes a broadcast slot, which the others poll. Tags carry a launch epoch,
// so the slots are never cleared. Waits are bounded by the wall clock: a wait that exceeds the limit raises the
// state's fail flag, every workgroup leaves at its next wait, and the host reports an error.
// The grid never exceeds the workgroups the device holds at once (occupancy query) and the workgroups wait only for
// each other; if the device is shared and the grid is not resident after all, a wait runs out of time, every
// workgroup leaves, and the ALS call reruns on the chunked CG.
#include <algorithm>
#include <cstdlib>

#include "kernels.hpp"
#include "xtt_common.hpp"

namespace xtt {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
__device__ inline d4 mma(double a, double b, d4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
__device__ inline f4 mma(float a, float b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
// accumulator type and the row held by result register r of lane group lk (16×16×4 MFMA tiles)
template <class T> struct Mfma;
template <> struct Mfma<double> { using acc = d4; __device__ static int row(int lk, int r) { return 4 * r + lk; } };
template <> struct Mfma<float> { using acc = f4; __device__ static int row(int lk, int r) { return 4 * lk + r; } };

// Device-coherent (sc1) loads and stores for data that other workgroups of the launch read or wrote.
template <class T> __device__ inline T ld_dev(const T* p) {
	return __hip_atomic_load(const_cast<T*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T> __device__ inline void st_dev(T* p, T v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ inline void stores_done() { __builtin_amdgcn_s_waitcnt(0x0F70); }  // s_waitcnt vmcnt(0) (gfx9 encoding)

constexpr int FT = 256, FW = FT / 64;  // threads and waves per workgroup
constexpr int UB = 16;                 // a-block (stage 1) and c-block (stage 3) of a unit

template <class T>
struct FcgArgs {
	const double* L;   // (a, P, a)
	const double* Am;  // (n, Q, P, n)
	const double* R;   // (c, Q, c)
	const double* r0;  // start residual (a, n, c), scaled by r0scale
	const double* b;   // mode 0: right-hand side (a, n, c)
	double* x;         // (a, n, c): warm start and result (mode 0), or x += xscale·e with e from 0 (mode 1)
	T* vec;            // 7·N: r[2], s[2], w, p, x in the layout (c, a, n)
	T* T2g;            // (a, c, n·Q)
	double* sync;      // barrier slots: SLOT doubles per workgroup, the broadcast slot at BCAST
	unsigned long long epoch;  // tag prefix of this launch (distinct per launch on one slot buffer)
	CGState* st;       // mode 0: rr0, tol2, maxIter, done from cg_start; the fail flag (pad) always
	double* iterAcc;   // optional: [0] += iterations, [1] += iterations·flopsPerIter
	double r0scale, xscale, tol2, flopsPerIter;
	int maxIter, mode;
	int a, n, c, P, Q;
	int U1, U3, nab, ncb;
	unsigned long long timeout;  // wall-clock ticks a barrier wait may take
	int ldL, ldAm, ldR, ldT1, ldT2;
	int oL, oAm, oR, oV, oT1, oT2, oS, oRc;  // LDS offsets in elements
};

// C = A·B over LDS operands, A(i,k) = A[i·sai + k·sak], B(k,j) = B[k·sbk + j·sbj]; out(i, j, v) receives every
// entry once. Waves take 16×16 tiles; with fewer tiles than waves, the K range is split over the waves and the
// partial tiles are summed in wave order through `scr` (FW·256 elements). Called by the whole workgroup.
template <class T, class F>
__device__ inline void mm(const T* A, int sai, int sak, const T* B, int sbk, int sbj, int M, int N, int K, T* scr, F&& out) {
	using V = typename Mfma<T>::acc;
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, li = lane & 15, lk = lane >> 4;
	const int tn = (N + 15) >> 4, tiles = ((M + 15) >> 4) * tn, ks = (K + 3) >> 2;
	auto step = [&](int i0, int j0, int k4, V acc) {
		const int kk = 4 * k4 + lk, i = i0 + li, j = j0 + li;
		const T av = (i < M && kk < K) ? A[i * sai + kk * sak] : T(0);
		const T bv = (j < N && kk < K) ? B[kk * sbk + j * sbj] : T(0);
		return mma(av, bv, acc);
	};
	auto run = [&](int i0, int j0, int kb, int ke) {
		V a0 = {0, 0, 0, 0}, a1 = {0, 0, 0, 0};
		int k = kb;
		for (; k + 1 < ke; k += 2) { a0 = step(i0, j0, k, a0); a1 = step(i0, j0, k + 1, a1); }
		if (k < ke) a0 = step(i0, j0, k, a0);
		return a0 + a1;
	};
	if (tiles >= FW) {
		for (int t = wave; t < tiles; t += FW) {
			const int i0 = (t / tn) * 16, j0 = (t % tn) * 16;
			const V acc = run(i0, j0, 0, ks);
#pragma unroll
			for (int r = 0; r < 4; ++r) {
				const int i = i0 + Mfma<T>::row(lk, r), j = j0 + li;
				if (i < M && j < N) out(i, j, acc[r]);
			}
		}
		return;
	}
	const int split = FW / tiles, per = (ks + split - 1) / split;
	const int t = wave / split, part = wave - t * split;
	const int i0 = (t / tn) * 16, j0 = (t % tn) * 16;
	if (t < tiles) {
		const V acc = run(i0, j0, part * per, min(ks, (part + 1) * per));
#pragma unroll
		for (int r = 0; r < 4; ++r) scr[wave * 256 + r * 64 + lane] = acc[r];
	}
	__syncthreads();
	if (t < tiles && part == 0) {
#pragma unroll
		for (int r = 0; r < 4; ++r) {
			T v = scr[wave * 256 + r * 64 + lane];
			for (int q = 1; q < split; ++q) v += scr[(wave + q) * 256 + r * 64 + lane];
			const int i = i0 + Mfma<T>::row(lk, r), j = j0 + li;
			if (i < M && j < N) out(i, j, v);
		}
	}
	__syncthreads();
}

// (x, y) summed over the workgroup in a fixed order; every thread receives the totals.
__device__ inline void wg_sum2(double& x, double& y, double* red) {
	for (int off = 32; off > 0; off >>= 1) { x += __shfl_xor(x, off); y += __shfl_xor(y, off); }
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	if (lane == 0) { red[2 * wave] = x; red[2 * wave + 1] = y; }
	__syncthreads();
	x = 0; y = 0;
	for (int w = 0; w < FW; ++w) { x += red[2 * w]; y += red[2 * w + 1]; }
	__syncthreads();
}

constexpr int SLOT = 8;      // doubles per barrier slot (64 bytes: tag, x, y)
constexpr int BCAST = 1024;  // slot of the broadcast; workgroups use slots 0 .. G-1 (G <= 1024)
using ull = unsigned long long;

// Grid barrier with a sum: every workgroup contributes (x, y) and receives the grid totals, added in slot order
// by workgroup 0 (the same values everywhere). `tag` is unique per barrier and launch. Returns false once any
// wait has run past the time limit (the fail flag is then set).
__device__ inline bool grid_sum(double* S, unsigned* fail, ull tag, double& x, double& y, ull timeout, double* red) {
	const int tid = threadIdx.x, G = gridDim.x;
	stores_done();
	__syncthreads();
	if (tid == 0) {
		st_dev(S + blockIdx.x * SLOT + 1, x);
		st_dev(S + blockIdx.x * SLOT + 2, y);
		stores_done();
		st_dev(reinterpret_cast<ull*>(S + blockIdx.x * SLOT), tag);
	}
	ull* const bc = reinterpret_cast<ull*>(S + BCAST * SLOT);
	if (blockIdx.x == 0 && tid < 64) {  // wave 0 of workgroup 0 gathers
		const ull t0 = wall_clock64();
		int bad = timeout == 0;  // a zero limit gives up at once (the host's way to exercise the failure path)
		for (int u = tid; u < G && !bad; u += 64) {
			while (ld_dev(reinterpret_cast<const ull*>(S + u * SLOT)) != tag) {
				if (wall_clock64() - t0 > timeout) { bad = 1; break; }
				__builtin_amdgcn_s_sleep(1);
			}
		}
		bad = __any(bad);
		double sx = 0, sy = 0;
		for (int u = tid; u < G; u += 64) { sx += ld_dev(S + u * SLOT + 1); sy += ld_dev(S + u * SLOT + 2); }
		for (int off = 32; off > 0; off >>= 1) { sx += __shfl_xor(sx, off); sy += __shfl_xor(sy, off); }
		if (tid == 0) {
			if (bad) st_dev(fail, 1u);
			st_dev(S + BCAST * SLOT + 1, sx);
			st_dev(S + BCAST * SLOT + 2, sy);
			stores_done();
			st_dev(bc, tag);
		}
	}
	if (tid == 0) {
		const ull t0 = wall_clock64();
		int bad = 0;
		if (timeout == 0) st_dev(fail, 1u);
		while (timeout != 0 && ld_dev(bc) != tag) {
			if (wall_clock64() - t0 > timeout) {
				st_dev(fail, 1u);
				bad = 1;
				break;
			}
			__builtin_amdgcn_s_sleep(1);
		}
		if (ld_dev(fail)) bad = 1;
		red[0] = ld_dev(S + BCAST * SLOT + 1);
		red[1] = ld_dev(S + BCAST * SLOT + 2);
		red[2] = bad;
	}
	__syncthreads();
	x = red[0];
	y = red[1];
	const bool ok = red[2] == 0.0;
	__syncthreads();
	return ok;
}

template <class T>
__global__ __launch_bounds__(FT) void fused_cg_kernel(FcgArgs<T> g) {
	extern __shared__ __align__(16) unsigned char lds_raw[];
	T* lds = reinterpret_cast<T*>(lds_raw);
	__shared__ double red[2 * FW];
	const int tid = threadIdx.x, wg = blockIdx.x;
	const int a = g.a, n = g.n, c = g.c, P = g.P, Q = g.Q, an = a * n, nQ = n * Q, Pn = P * n, cQ = c * Q;
	const size_t N = (size_t)an * c;
	T* const rb0 = g.vec;
	T* const sb0 = g.vec + 2 * N;
	T* const w = g.vec + 4 * N;
	T* const p = g.vec + 5 * N;
	T* const xv = g.vec + 6 * N;
	unsigned* const fail = reinterpret_cast<unsigned*>(&g.st->pad);
	const ull tagBase = g.epoch << 16;

	int stopAtStart = 0;
	double rr0 = 0, tol2 = g.tol2;
	int maxIter = g.maxIter;
	bool zeroStart = false;
	if (g.mode == 0) {
		stopAtStart = g.st->done;
		rr0 = g.st->rr0;
		tol2 = g.st->tol2;
		maxIter = g.st->maxIter;
		// A warm start whose residual exceeds ‖b‖ is worse than zero, and CG's recursive residual cannot resolve the
		// true one below ε·‖A‖·‖x‖ (x of the start's size): such a solve starts from x = 0, r = b instead.
		zeroStart = g.st->rr > rr0;
	}
	if (stopAtStart || maxIter <= 0) return;  // uniform: nothing to do, x stays as it is
	const double* const rstart = zeroStart ? g.b : g.r0;

	const bool has1 = wg < g.U1, has3 = wg < g.U3;
	const int cp = has1 ? wg / g.nab : 0, a0 = has1 ? (wg - cp * g.nab) * UB : 0, ab = min(UB, a - a0);
	const int a3 = has3 ? wg / g.ncb : 0, c0 = has3 ? (wg - a3 * g.ncb) * UB : 0, cb = min(UB, c - c0);
	T* const Ls = lds + g.oL;
	T* const As = lds + g.oAm;
	T* const Rs = lds + g.oR;
	T* const Vs = lds + g.oV;
	T* const T1s = lds + g.oT1;
	T* const T2s = lds + g.oT2;
	T* const scr = lds + g.oS;
	T* const Rc = lds + g.oRc;  // stage 3: r of the unit's own elements, (c local, n)

	// operators into LDS (FP64 → T)
	if (has1) {
		for (int e = tid; e < ab * P * a; e += FT) {
			const int i = e / a, k = e - i * a;
			Ls[i * g.ldL + k] = (T)g.L[(size_t)a0 * P * a + e];
		}
		for (int e = tid; e < nQ * Pn; e += FT) {
			const int i = e / Pn, k = e - i * Pn;
			As[i * g.ldAm + k] = (T)g.Am[e];
		}
	}
	if (has3) {
		for (int e = tid; e < cb * cQ; e += FT) {  // Rs[j][(c',q)] = R[c0+j][q][c']
			const int j = e / cQ, k = e - j * cQ, cc = k / Q, q = k - cc * Q;
			Rs[j * g.ldR + k] = (T)g.R[((size_t)(c0 + j) * Q + q) * c + cc];
		}
	}
	// own elements (stage-1 partition): r = r0, x = start, p = s = 0; the slice r0[·][·][c'] for the first product
	if (has1) {
		for (int e = tid; e < ab * n; e += FT) {
			const size_t src = (size_t)(a0 * n + e) * c + cp;
			const size_t dst = (size_t)cp * an + a0 * n + e;
			st_dev(rb0 + dst, (T)(rstart[src] * g.r0scale));
			st_dev(sb0 + dst, T(0));
			xv[dst] = g.mode == 0 && !zeroStart ? (T)g.x[src] : T(0);
			p[dst] = T(0);
		}
		for (int e = tid; e < an; e += FT) Vs[e] = (T)(rstart[(size_t)e * c + cp] * g.r0scale);
	}
	__syncthreads();

	auto stage1 = [&]() {
		mm(Ls, g.ldL, 1, Vs, n, 1, ab * P, n, a, scr, [&](int i, int j, T v) {
			const int al = i / P, pp = i - al * P;
			T1s[al * g.ldT1 + pp * n + j] = v;
		});
		__syncthreads();
		T* const out = g.T2g + ((size_t)a0 * c + cp) * nQ;
		mm(T1s, g.ldT1, 1, As, 1, g.ldAm, ab, nQ, Pn, scr, [&](int i, int j, T v) { st_dev(out + (size_t)i * c * nQ + j, v); });
	};
	auto stage3 = [&](const T* rcur, double& pg, double& pd) {
		const T* src = g.T2g + (size_t)a3 * c * nQ;
		const int tot = c * nQ;
		for (int e0 = 0; e0 < tot; e0 += 8 * FT) {  // T2s[nn][(c',q)] = T2[a3][c'][(nn,q)], 8 loads in flight
			T v[8];
#pragma unroll
			for (int u = 0; u < 8; ++u) {
				const int e = e0 + u * FT + tid;
				v[u] = e < tot ? ld_dev(src + e) : T(0);
			}
#pragma unroll
			for (int u = 0; u < 8; ++u) {
				const int e = e0 + u * FT + tid;
				if (e < tot) {
					const int cc = e / nQ, k = e - cc * nQ, nn = k / Q, q = k - nn * Q;
					T2s[nn * g.ldT2 + cc * Q + q] = v[u];
				}
			}
		}
		for (int e = tid; e < cb * n; e += FT) {
			const int jc = e / n, nn = e - jc * n;
			Rc[e] = ld_dev(rcur + (size_t)(c0 + jc) * an + a3 * n + nn);
		}
		__syncthreads();
		mm(Rs, g.ldR, 1, T2s, 1, g.ldT2, cb, n, cQ, scr, [&](int i, int j, T v) {
			st_dev(w + (size_t)(c0 + i) * an + a3 * n + j, v);
			const double rv = (double)Rc[i * n + j];
			pg = fma(rv, rv, pg);
			pd = fma((double)v, rv, pd);
		});
		wg_sum2(pg, pd, red);
	};
	// one product and its reduction: T2 from Vs (stage 1), barrier, w and (γ, δ) of r (stage 3), barrier
	unsigned kb = 0;
	auto product = [&](const T* rcur, double& gam, double& del) {
		if (has1) stage1();
		double zx = 0, zy = 0;
		const bool good = grid_sum(g.sync, fail, tagBase | ++kb, zx, zy, g.timeout, red);
		gam = 0;
		del = 0;
		if (good && has3) stage3(rcur, gam, del);
		return good && grid_sum(g.sync, fail, tagBase | ++kb, gam, del, g.timeout, red);
	};

	double gam = 0, del = 0;
	bool ok = product(rb0, gam, del);
	double gamOld = 0, alphaOld = 0;
	int it = 0, cur = 0;
	for (int k = 1; ok; ++k) {
		if (k == 1 && g.mode == 1) rr0 = gam;  // e starts at 0: r0 is the right-hand side
		bool stop = !(gam > tol2 * rr0) || k - 1 >= maxIter;
		double alpha = 0, beta = 0;
		if (!stop) {
			double den = del;
			if (k > 1) { beta = gam / gamOld; den = del - beta * gam / alphaOld; }
			if (den > 0) alpha = gam / den; else stop = true;  // breakdown: keep the current iterate
		}
		if (stop) break;
		T* const rIn = rb0 + (size_t)cur * N;
		T* const sIn = sb0 + (size_t)cur * N;
		T* const rOut = rb0 + (size_t)(cur ^ 1) * N;
		T* const sOut = sb0 + (size_t)(cur ^ 1) * N;
		if (has1) {
			const T al = (T)alpha, be = (T)beta;
			const size_t base = (size_t)cp * an;
			for (int e0 = 0; e0 < an; e0 += 4 * FT) {  // 12 loads in flight per thread
				T ro[4], wv[4], si[4];
#pragma unroll
				for (int u = 0; u < 4; ++u) {
					const int e = e0 + u * FT + tid;
					ro[u] = e < an ? ld_dev(rIn + base + e) : T(0);
					wv[u] = e < an ? ld_dev(w + base + e) : T(0);
					si[u] = e < an ? ld_dev(sIn + base + e) : T(0);
				}
#pragma unroll
				for (int u = 0; u < 4; ++u) {
					const int e = e0 + u * FT + tid;
					if (e >= an) continue;
					const size_t idx = base + e;
					const T sn = wv[u] + be * si[u];
					const T rn = ro[u] - al * sn;
					Vs[e] = rn;
					if (e >= a0 * n && e < (a0 + ab) * n) {
						const T pn = ro[u] + be * p[idx];
						p[idx] = pn;
						xv[idx] += al * pn;
						st_dev(sOut + idx, sn);
						st_dev(rOut + idx, rn);
					}
				}
			}
			__syncthreads();
		}
		cur ^= 1;
		gamOld = gam;
		alphaOld = alpha;
		ok = product(rOut, gam, del);
		it = k;
	}
	if (!ok) {
		if (wg == 0 && tid == 0) {
			g.st->iter = -1;
			g.st->done = 1;
			if (g.iterAcc) g.iterAcc[0] = __builtin_nan("");
		}
		return;
	}
	if (has1) {
		for (int e = tid; e < ab * n; e += FT) {
			const size_t src = (size_t)(a0 * n + e) * c + cp;
			const double v = (double)xv[(size_t)cp * an + a0 * n + e];
			if (g.mode == 0) g.x[src] = v; else g.x[src] += g.xscale * v;
		}
	}
	if (wg == 0 && tid == 0) {
		g.st->iter = it;
		g.st->rr = gam;
		g.st->done = 1;
		if (g.iterAcc) { g.iterAcc[0] += (double)it; g.iterAcc[1] += (double)it * g.flopsPerIter; }
	}
}

int pad4(int k) { return k + ((68 - (k & 63)) & 63); }  // smallest ld >= k with ld ≡ 4 (mod 64)

struct FcgPlan {
	int U1, U3, nab, ncb, G;
	int ldL, ldAm, ldR, ldT1, ldT2;
	int oL, oAm, oR, oV, oT1, oT2, oS, oRc;
	size_t lds;
};

template <class T>
bool plan(int a, int n, int c, int P, int Q, FcgPlan& f) {
	f.nab = (a + UB - 1) / UB; f.ncb = (c + UB - 1) / UB;
	f.U1 = c * f.nab; f.U3 = a * f.ncb; f.G = std::max(f.U1, f.U3);
	f.ldL = pad4(a); f.ldAm = pad4(P * n); f.ldR = pad4(c * Q); f.ldT1 = pad4(P * n); f.ldT2 = pad4(c * Q);
	size_t o = 0;
	f.oL = (int)o; o += (size_t)UB * P * f.ldL;
	f.oAm = (int)o; o += (size_t)n * Q * f.ldAm;
	f.oR = (int)o; o += (size_t)UB * f.ldR;
	f.oV = (int)o; o += (size_t)a * n;
	f.oT1 = (int)o; o += (size_t)UB * f.ldT1;
	f.oT2 = (int)o; o += (size_t)n * f.ldT2;
	f.oS = (int)o; o += (size_t)FW * 256;
	f.oRc = (int)o; o += (size_t)UB * n;
	f.lds = o * sizeof(T);
	return f.lds <= 150 * 1024 && f.G <= 1024 && f.U3 <= 1024;
}

template <class T>
int max_resident() {  // workgroups of fused_cg_kernel<T> that fit on the device at once, at full LDS use
	static thread_local int cached = -1;
	if (cached < 0) {
		int dev = 0, cus = 0, per = 0;
		XTT_HIP(hipGetDevice(&dev));
		XTT_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
		XTT_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(fused_cg_kernel<T>), hipFuncAttributeMaxDynamicSharedMemorySize,
		                            150 * 1024));
		XTT_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fused_cg_kernel<T>, FT, 150 * 1024));
		cached = cus * std::max(per, 0);
	}
	return cached;
}

unsigned long long barrier_ticks() {
	static unsigned long long ticks = 0;
	if (!ticks) {
		int dev = 0, khz = 0;
		XTT_HIP(hipGetDevice(&dev));
		if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess || khz <= 0) khz = 100000;
		ticks = (unsigned long long)khz * 2000ull;  // 2 s
	}
	if (const char* e = getenv("XTT_FUSED_CG_WAIT_TICKS")) return strtoull(e, nullptr, 10);
	return ticks;
}

}  // namespace

template <class T>
bool fused_cg_fits(size_t a, size_t n, size_t c, size_t P, size_t Q) {
	if (a > 4096 || n > 4096 || c > 4096 || P > 64 || Q > 64) return false;
	FcgPlan f;
	return plan<T>((int)a, (int)n, (int)c, (int)P, (int)Q, f) && f.G <= max_resident<T>();
}

template <class T>
bool fused_cg(const FusedCG& d, hipStream_t s) {
	FcgPlan f;
	XTT_REQUIRE(plan<T>((int)d.a, (int)d.n, (int)d.c, (int)d.P, (int)d.Q, f) && f.G <= max_resident<T>(),
	            "fused CG: local shape does not fit one grid-synchronised launch");
	FcgArgs<T> g{};
	g.L = d.L; g.Am = d.Am; g.R = d.R; g.r0 = d.r0; g.b = d.b; g.x = d.x;
	g.vec = static_cast<T*>(d.vec); g.T2g = static_cast<T*>(d.T2); g.sync = d.sync; g.epoch = d.epoch; g.st = d.st;
	g.iterAcc = d.iterAcc;
	g.r0scale = d.r0scale; g.xscale = d.xscale; g.tol2 = d.tol2; g.flopsPerIter = d.flopsPerIter; g.maxIter = d.maxIter; g.mode = d.mode;
	g.a = (int)d.a; g.n = (int)d.n; g.c = (int)d.c; g.P = (int)d.P; g.Q = (int)d.Q;
	g.U1 = f.U1; g.U3 = f.U3; g.nab = f.nab; g.ncb = f.ncb; g.timeout = barrier_ticks();
	g.ldL = f.ldL; g.ldAm = f.ldAm; g.ldR = f.ldR; g.ldT1 = f.ldT1; g.ldT2 = f.ldT2;
	g.oL = f.oL; g.oAm = f.oAm; g.oR = f.oR; g.oV = f.oV; g.oT1 = f.oT1; g.oT2 = f.oT2; g.oS = f.oS; g.oRc = f.oRc;
	XTT_REQUIRE(d.maxIter <= 30000 && d.epoch > 0 && d.epoch < (1ull << 47), "fused CG: iteration limit or epoch out of range");
	// A plain launch: hipLaunchCooperativeKernel would refuse a grid the device cannot hold at once, but in this
	// runtime a single cooperative launch slows every later multi-stream workload of the process (the 16-stream
	// batch of 512 problems ran at 111 instead of 187 sweeps/s after one). Co-residency is instead sized from the
	// occupancy query (f.G <= max_resident), and a grid that still is not resident (the device shared with other
	// work) ends at the bounded wait, after which the caller reruns the call on the chunked CG.
	hipLaunchKernelGGL(fused_cg_kernel<T>, dim3(f.G), dim3(FT), f.lds, s, g);
	XTT_HIP(hipGetLastError());
	return true;
}

template bool fused_cg_fits<double>(size_t, size_t, size_t, size_t, size_t);
template bool fused_cg_fits<float>(size_t, size_t, size_t, size_t, size_t);
template bool fused_cg<double>(const FusedCG&, hipStream_t);
template bool fused_cg<float>(const FusedCG&, hipStream_t);

}  // namespace xtt
