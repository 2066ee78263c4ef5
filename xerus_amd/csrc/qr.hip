// Householder QR for the tall-skinny cores of move_core / round (n <= 64 columns), register resident.
//
// One workgroup (8 waves) owns a block of up to 256 rows: lane j of every wave holds column j, wave w holds
// rows w, w+8, w+16, … of the block in registers. In a Householder step every wave reads the pivot column's
// values on its rows from lane i into scalar registers (v_readlane, no LDS traffic); the per-wave partial sums
// of vᵀA (whose pivot lane is the column's ‖x‖²) cross waves through LDS: one barrier per column. The reflector is
// applied unscaled and scaled once at the end (u = scal·a below the diagonal).
//
// Matrices taller than one block are factored as a reduction tree (TSQR): the leaves factor row blocks, each
// level stacks the n×n R factors of up to 256/n children and factors the stack again, until one block is
// left. Q is formed top-down: the root applies its reflectors to the identity (in the same launch as its
// factorisation), every node applies its reflectors to its n×c slice of the parent's result. The result is a Householder QR
// of the whole matrix (Q with orthonormal columns also for rank-deficient input); its diagonal signs follow
// the tree, not LAPACK's single sweep, which only changes the TT gauge.
//
// Column pivoting (xerus's QC, blasLapackWrapper.cpp:244-305 with dgeqp3) runs on one block: directly on
// the matrix when it has at most 256 rows, else as the tree's root on the stacked R factors of the level
// below (same column norms and the same partial-norm downdating, since the levels' Q are orthogonal). The pivot rule and downdating are dlaqp2's: largest
// partial norm, first index on ties, recompute when the downdated norm falls below sqrt(eps/2). The rank rule
// is xerus's: the first r with |R_rr| < 16·ε·R_00, where R_00 carries dlarfg's sign for the original first
// row (blasLapackWrapper.cpp:266-272).
#include <algorithm>
#include <cfloat>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "kernels.hpp"
#include "xtt_common.hpp"

namespace xtt {
namespace {

constexpr int HW = 8;          // waves per workgroup
constexpr int HT = 64 * HW;    // threads per workgroup
constexpr int HN = 64;         // maximum columns (one lane each)
constexpr int HBR = 32 * HW;   // maximum rows per block

// DPP move of a double inside rows of 16 lanes (quad permutations, row rotations).
template <int CTRL>
__device__ inline double dpp_f64(double v) {
	const long long b = __double_as_longlong(v);
	const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, false);
	const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
	return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ inline double lane_value(double v, int l) {  // v of lane l (l wave-uniform)
	const long long b = __double_as_longlong(v);
	const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)b, l);
	const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(b >> 32), l);
	return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
// Maximum over the wave, in every lane: quad swaps and row rotations, then the four row maxima.
__device__ inline double wave_max(double v) {
	v = fmax(v, dpp_f64<0xB1>(v));   // quad_perm [1,0,3,2]
	v = fmax(v, dpp_f64<0x4E>(v));   // quad_perm [2,3,0,1]
	v = fmax(v, dpp_f64<0x124>(v));  // row_ror:4
	v = fmax(v, dpp_f64<0x128>(v));  // row_ror:8
	return fmax(fmax(lane_value(v, 0), lane_value(v, 16)), fmax(lane_value(v, 32), lane_value(v, 48)));
}

struct FactorArgs {
	const double* in;    // level matrix: element (r, c) at in[r·ld + c], or in[c·ld + r] when trans
	long long ld;
	int trans;
	int rows;            // rows of the level matrix
	int rpb;             // rows per workgroup (the last one may have fewer)
	int n;               // columns
	double* V;           // rows × n: R on/above, reflectors below the diagonal of each block (may be null)
	double* tau;         // blocks × n (may be null when V is)
	double* R;           // blocks × (n × n) upper triangles, ld n (may be null); block b writes min(rows_b, n) rows
	// Q of a single-block level (FORMQ): Q = H_0…H_{k-1}·[init; 0], rows × nc (ld nc); init null = identity
	double* Q;
	const double* init;
	int nc;
	// column pivoting (single block); with FORMQ, nc = the rank
	int* jpvt;           // n: original column of each position
	int* rank;           // xerus's QC rank
	double* C;           // rank × n, C[row][jpvt[col]] = R[row][col] (row <= col)
	const double* orig;  // when set, R_00's sign is that of the original matrix (origRows × n, row-major)
	int origRows;
};

template <int RPT, bool PIV, bool FORMQ>
__global__ __launch_bounds__(HT) void hh_factor_kernel(FactorArgs a) {
	__shared__ double sW[2][HW][HN];   // per-wave Σ_{own r > i} a_ri·a_rj (lane i: the norm² partial)
	__shared__ double sRow[2][HN];     // row i
	__shared__ double sR[HN][HN];      // finished rows of R (lanes >= row)
	__shared__ double sScal[HN], sTau[HN];
	__shared__ double sNrm[HW][HN];    // per-wave partial column norms (pivoting)
	__shared__ int sRank;
	const int lane = threadIdx.x & 63;
	const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
	const int b = blockIdx.x;
	const int r0 = b * a.rpb;
	const int rows = min(a.rpb, a.rows - r0);
	const int n = a.n;
	const int k = min(rows, n);

	double x[RPT], v[RPT];
#pragma unroll
	for (int t = 0; t < RPT; ++t) {
		const int r = t * HW + w;
		double z = 0.0;
		if (r < rows && lane < n) z = a.trans ? a.in[(size_t)lane * a.ld + (r0 + r)] : a.in[(size_t)(r0 + r) * a.ld + lane];
		x[t] = z;
	}

	double vn1 = 0.0, vn2 = 0.0;  // partial column norms, identical in every wave
	int jp = lane;
	const double tol3z = sqrt(DBL_EPSILON * 0.5);
	if (PIV) {
		double s0 = 0.0, s1 = 0.0;
#pragma unroll
		for (int t = 0; t < RPT; t += 2) {
			s0 = fma(x[t], x[t], s0);
			if (t + 1 < RPT) s1 = fma(x[t + 1], x[t + 1], s1);
		}
		sNrm[w][lane] = s0 + s1;
		__syncthreads();
		double tot = 0.0;
		for (int q = 0; q < HW; ++q) tot += sNrm[q][lane];
		vn1 = vn2 = sqrt(tot);
	}

	// Step i touches rows > i through v (zero on rows <= i) and lanes > i through g (zero on lanes <= i), so
	// the register loops need no row conditions: rows < i keep (stale) values that are dead — row i's final
	// values go to sR, column i's reflector stays in lane i. Columns run in chunks of HW: in chunk C, row i
	// is register x[C] of wave i - C·HW, and registers t < C hold rows < i only, so they are skipped.
	auto chunk = [&](auto cc) {
		constexpr int C = decltype(cc)::value;
		if (C * HW >= k) return;
		const int iend = min(k, (C + 1) * HW);
		for (int i = C * HW; i < iend; ++i) {
			const int p = i & 1;
			const int wi = i - C * HW;  // wave owning row i
			if (PIV) {
				// dlaqp2's pivot: the largest partial norm over [i, n), the first index on ties
				const double cand = (lane >= i && lane < n) ? vn1 : -1.0;
				const double mx = wave_max(cand);
				const unsigned long long hit = __ballot(cand == mx);
				const int pv = hit ? __builtin_amdgcn_readfirstlane((int)__builtin_ctzll(hit)) : i;  // all-NaN: no swap
				if (pv != i) {
					const int src = lane == i ? pv : (lane == pv ? i : lane);
#pragma unroll
					for (int t = C; t < RPT; ++t) x[t] = __shfl(x[t], src, 64);
					vn1 = __shfl(vn1, src, 64);
					vn2 = __shfl(vn2, src, 64);
					jp = __shfl(jp, src, 64);
					if (w == 0 && lane < i) {  // the finished rows of R swap columns too (wave 0 owns sR)
						const double tmp = sR[lane][i];
						sR[lane][i] = sR[lane][pv];
						sR[lane][pv] = tmp;
					}
				}
			}
			// phase A: d_j = Σ_{own r > i} a_ri·a_rj (unscaled reflector; lane i gets the norm² partial), row i.
			// Column i of the wave's rows is read from lane i into scalar registers (v_readlane).
			{
				const double vc = lane_value(x[C], i);
				v[C] = w > wi ? vc : 0.0;
#pragma unroll
				for (int t = C + 1; t < RPT; ++t) v[t] = lane_value(x[t], i);
			}
			double d0 = 0.0, d1 = 0.0;
#pragma unroll
			for (int t = C; t < RPT; ++t) {
				if ((t - C) & 1) d1 = fma(v[t], x[t], d1);
				else d0 = fma(v[t], x[t], d0);
			}
			sW[p][w][lane] = d0 + d1;
			if (w == wi) sRow[p][lane] = x[C];
			__syncthreads();
			// phase B: reflector (dlarfg) and A(i:, i+1:) -= u·wᵀ with u = [1; scal·a(i+1:, i)]
			double sa = 0.0, sb = 0.0;
#pragma unroll
			for (int q = 0; q < HW; q += 2) {
				sa += sW[p][q][lane];
				sb += sW[p][q + 1][lane];
			}
			const double s = sa + sb;
			const double xn2 = lane_value(s, i);  // lane i's sum is ‖a(i+1:, i)‖²
			const double alpha = sRow[p][i];
			double beta = alpha, tau = 0.0, scal = 1.0;
			if (xn2 != 0.0) {
				const double a2 = fma(alpha, alpha, xn2);
				if (a2 < 1e300 && a2 > 1e-300) {
					// With y = 1/‖(alpha, x)‖: beta = −sign(alpha)/y, tau = (beta − alpha)/beta = 1 + |alpha|·y and
					// scal = 1/(alpha − beta) = sign(alpha)·y/tau. y is v_rsq_f64 refined by two Newton steps
					// y ← y + y·(1 − a2·y²)/2, 1/tau (tau in [1, 2]) is v_rcp_f64 refined by two steps r ← r + r·(1 − tau·r).
					double y = __builtin_amdgcn_rsq(a2);
					y = fma(0.5 * y, fma(-(a2 * y), y, 1.0), y);
					y = fma(0.5 * y, fma(-(a2 * y), y, 1.0), y);
					const double aa = fabs(alpha);
					tau = fma(aa, y, 1.0);
					double rt = __builtin_amdgcn_rcp(tau);
					rt = fma(rt, fma(-tau, rt, 1.0), rt);
					rt = fma(rt, fma(-tau, rt, 1.0), rt);
					beta = -copysign(a2 * y, alpha);
					scal = copysign(y * rt, alpha);
				} else {  // outside the safe range: dlapy2's scaled norm and plain divisions
					const double nrm = hypot(alpha, sqrt(xn2));
					beta = -copysign(nrm, alpha);
					tau = (beta - alpha) / beta;
					scal = 1.0 / (alpha - beta);
				}
			}
			const double wj = (lane > i && lane < n) ? tau * fma(scal, s, sRow[p][lane]) : 0.0;
			const double rij = sRow[p][lane] - wj;  // row i after the update (lanes > i)
			if (w == 0) {
				sR[i][lane] = lane > i ? rij : (lane == i ? beta : 0.0);
				if (lane == 0) {
					sTau[i] = tau;
					sScal[i] = scal;
				}
			}
			if (tau != 0.0) {
				const double g = scal * wj;
#pragma unroll
				for (int t = C; t < RPT; ++t) x[t] = fma(-v[t], g, x[t]);
			}
			if (PIV) {
				// dlaqp2's partial-norm downdating with the updated row i
				bool need = false;
				if (lane > i && lane < n && vn1 != 0.0) {
					double temp = fabs(rij) / vn1;
					temp = fmax(0.0, 1.0 - temp * temp);
					const double ratio = vn1 / vn2;
					const double temp2 = temp * ratio * ratio;
					if (temp2 <= tol3z) need = true;
					else vn1 *= sqrt(temp);
				}
				if (__any(need)) {  // the same decision in every wave (replicated norms)
					double sq = w > wi ? x[C] * x[C] : 0.0;
#pragma unroll
					for (int t = C + 1; t < RPT; ++t) sq = fma(x[t], x[t], sq);
					sNrm[w][lane] = sq;
					__syncthreads();
					double tot = 0.0;
					for (int q = 0; q < HW; ++q) tot += sNrm[q][lane];
					if (need) vn1 = vn2 = (i + 1 < rows) ? sqrt(tot) : 0.0;
				}
			}
		}
	};
	chunk(std::integral_constant<int, 0>{});
	chunk(std::integral_constant<int, 1>{});
	chunk(std::integral_constant<int, 2>{});
	chunk(std::integral_constant<int, 3>{});
	chunk(std::integral_constant<int, 4>{});
	chunk(std::integral_constant<int, 5>{});
	chunk(std::integral_constant<int, 6>{});
	chunk(std::integral_constant<int, 7>{});
	__syncthreads();

	// compact output (R on and above the diagonal, scaled reflectors below); x becomes the reflector
	// columns u (one on the diagonal, zero above)
#pragma unroll
	for (int t = 0; t < RPT; ++t) {
		const int r = t * HW + w;
		const double u = lane < k ? (r > lane ? x[t] * sScal[lane] : (r == lane ? 1.0 : 0.0)) : 0.0;
		if (a.V && r < rows && lane < n) a.V[(size_t)(r0 + r) * n + lane] = (r < k && lane >= r) ? sR[r][lane] : u;
		x[t] = u;
	}
	if (a.tau && threadIdx.x < k) a.tau[(size_t)b * n + threadIdx.x] = sTau[threadIdx.x];
	int nc = a.nc;
	if (!PIV) {
		if (a.R) {
			for (int r = w; r < k; r += HW)
				if (lane < n) a.R[((size_t)b * n + r) * n + lane] = sR[r][lane];
		}
	} else {
		double r00 = sR[0][0];
		if (a.orig) {
			// R_00 as LAPACK's single sweep over the original matrix gives it: beta of the pivot column p0
			const int p0 = __builtin_amdgcn_readfirstlane(__shfl(jp, 0, 64));
			int nz = 0;
			for (int r = 1 + threadIdx.x; r < a.origRows; r += HT) nz |= a.orig[(size_t)r * n + p0] != 0.0;
			nz = __syncthreads_or(nz);
			const double alpha0 = a.orig[p0];
			r00 = nz ? -copysign(fabs(sR[0][0]), alpha0) : alpha0;
		}
		if (threadIdx.x == 0) {
			int rk;
			for (rk = 1; rk <= k; ++rk)
				if (rk == k || fabs(sR[rk][rk]) < 16 * DBL_EPSILON * r00) break;
			sRank = rk;
			*a.rank = rk;
		}
		if (w == 0 && lane < n) a.jpvt[lane] = jp;
		__syncthreads();
		const int rank = sRank;
		for (int r = w; r < rank; r += HW)
			if (lane < n) a.C[(size_t)r * n + jp] = sR[r][lane];
		nc = rank;
	}
	if (!FORMQ) return;

	// Q = H_0 … H_{k-1} · [init; 0] for this block (single-block levels only: r0 = 0)
	double q[RPT];
#pragma unroll
	for (int t = 0; t < RPT; ++t) {
		const int r = t * HW + w;
		double z = 0.0;
		if (r < k && lane < nc) z = a.init ? a.init[(size_t)r * nc + lane] : (r == lane ? 1.0 : 0.0);
		q[t] = z;
	}
	int ph = 0;
	// reflectors in reverse; in chunk C the rows of registers t < C lie above every reflector of the chunk
	auto qchunk = [&](auto cc) {
		constexpr int C = decltype(cc)::value;
		if (C * HW >= k) return;
		for (int i = min(k, (C + 1) * HW) - 1; i >= C * HW; --i) {
			const double ti = sTau[i];
			if (ti == 0.0) continue;
			double c0 = 0.0, c1 = 0.0;
#pragma unroll
			for (int t = C; t < RPT; ++t) {
				v[t] = lane_value(x[t], i);  // reflector i (lane i's column) on the wave's rows
				if ((t - C) & 1) c1 = fma(v[t], q[t], c1);
				else c0 = fma(v[t], q[t], c0);
			}
			sW[ph][w][lane] = c0 + c1;
			__syncthreads();
			double wa = 0.0, wb = 0.0;
#pragma unroll
			for (int s2 = 0; s2 < HW; s2 += 2) {
				wa += sW[ph][s2][lane];
				wb += sW[ph][s2 + 1][lane];
			}
			const double wj = (wa + wb) * ti;
#pragma unroll
			for (int t = C; t < RPT; ++t) q[t] = fma(-v[t], wj, q[t]);
			ph ^= 1;
		}
	};
	qchunk(std::integral_constant<int, 7>{});
	qchunk(std::integral_constant<int, 6>{});
	qchunk(std::integral_constant<int, 5>{});
	qchunk(std::integral_constant<int, 4>{});
	qchunk(std::integral_constant<int, 3>{});
	qchunk(std::integral_constant<int, 2>{});
	qchunk(std::integral_constant<int, 1>{});
	qchunk(std::integral_constant<int, 0>{});
#pragma unroll
	for (int t = 0; t < RPT; ++t) {
		const int r = t * HW + w;
		if (r < rows && lane < nc) a.Q[(size_t)r * nc + lane] = q[t];
	}
}

struct ApplyArgs {
	const double* V;      // a factored level (same geometry as its FactorArgs)
	const double* tau;
	int rows, rpb, n;
	const double* init;   // (blocks·n) × nc, ld nc: block b starts from rows [b·n, b·n+n); null: identity
	double* out;          // rows × nc, ld nc
	int nc;               // columns of the product (<= 64)
	const int* ncDev;     // when set, the column count is read here (<= nc)
};

// out_b = H_0 · … · H_{k-1} · [init_b; 0] for every block b of the level. Lane i holds the block's reflector
// column u_i (zero above row i, one on it) for the wave's rows; the waves read it with v_readlane.
template <int RPT>
__global__ __launch_bounds__(HT) void hh_apply_kernel(ApplyArgs a) {
	__shared__ double sW[2][HW][HN];
	__shared__ double sTau[HN];
	const int lane = threadIdx.x & 63;
	const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
	const int b = blockIdx.x;
	const int r0 = b * a.rpb;
	const int rows = min(a.rpb, a.rows - r0);
	const int n = a.n;
	const int k = min(rows, n);
	const int nc = a.ncDev ? __builtin_amdgcn_readfirstlane(*a.ncDev) : a.nc;
	double vr[RPT];
#pragma unroll
	for (int t = 0; t < RPT; ++t) {
		const int r = t * HW + w;
		double z = 0.0;
		if (r < rows && lane < k) z = r > lane ? a.V[(size_t)(r0 + r) * n + lane] : (r == lane ? 1.0 : 0.0);
		vr[t] = z;
	}
	if (threadIdx.x < k) sTau[threadIdx.x] = a.tau[(size_t)b * n + threadIdx.x];
	double q[RPT], u[RPT];
#pragma unroll
	for (int t = 0; t < RPT; ++t) {
		const int r = t * HW + w;
		double z = 0.0;
		if (r < k && lane < nc) z = a.init ? a.init[((size_t)b * n + r) * nc + lane] : (r == lane ? 1.0 : 0.0);
		q[t] = z;
	}
	__syncthreads();
	int ph = 0;
	auto chunk = [&](auto cc) {
		constexpr int C = decltype(cc)::value;
		if (C * HW >= k) return;
		for (int i = min(k, (C + 1) * HW) - 1; i >= C * HW; --i) {
			const double ti = sTau[i];
			if (ti == 0.0) continue;
			double c0 = 0.0, c1 = 0.0;
#pragma unroll
			for (int t = C; t < RPT; ++t) {
				u[t] = lane_value(vr[t], i);
				if ((t - C) & 1) c1 = fma(u[t], q[t], c1);
				else c0 = fma(u[t], q[t], c0);
			}
			sW[ph][w][lane] = c0 + c1;
			__syncthreads();
			double wa = 0.0, wb = 0.0;
#pragma unroll
			for (int s = 0; s < HW; s += 2) {
				wa += sW[ph][s][lane];
				wb += sW[ph][s + 1][lane];
			}
			const double wj = (wa + wb) * ti;
#pragma unroll
			for (int t = C; t < RPT; ++t) q[t] = fma(-u[t], wj, q[t]);
			ph ^= 1;
		}
	};
	chunk(std::integral_constant<int, 7>{});
	chunk(std::integral_constant<int, 6>{});
	chunk(std::integral_constant<int, 5>{});
	chunk(std::integral_constant<int, 4>{});
	chunk(std::integral_constant<int, 3>{});
	chunk(std::integral_constant<int, 2>{});
	chunk(std::integral_constant<int, 1>{});
	chunk(std::integral_constant<int, 0>{});
#pragma unroll
	for (int t = 0; t < RPT; ++t) {
		const int r = t * HW + w;
		if (r < rows && lane < nc) a.out[(size_t)(r0 + r) * nc + lane] = q[t];
	}
}

int rows_per_thread(int rpb) { return rpb <= 8 * HW ? 8 : (rpb <= 16 * HW ? 16 : 32); }

// Householder flops of one level (dgeqrf's count per block: 2n²(m − n/3) for m >= n, 2m²(n − m/3) otherwise),
// plus, with FORMQ, k reflectors applied to m × nc: 4·nc·(m·k − k²/2).
double factor_flops(const FactorArgs& a, int blocks, bool formq) {
	double f = 0;
	for (int b = 0; b < blocks; ++b) {
		const double m = std::min(a.rpb, a.rows - b * a.rpb), n = a.n, k = std::min(m, n);
		f += m >= n ? 2.0 * n * n * (m - n / 3.0) : 2.0 * m * m * (n - m / 3.0);
		if (formq) f += 4.0 * a.nc * (m * k - k * k / 2.0);
	}
	return f;
}

template <bool PIV, bool FORMQ>
void factor_variant(const FactorArgs& a, int blocks, hipStream_t s) {
	const int rpt = rows_per_thread(a.rpb);
	const dim3 g(blocks), t(HT);
	static const char* tags[3][2] = {{"hh_factor_kernel<8,0,0>", "hh_factor_kernel<8,0,1>"},
	                                 {"hh_factor_kernel<16,0,0>", "hh_factor_kernel<16,0,1>"},
	                                 {"hh_factor_kernel<32,0,0>", "hh_factor_kernel<32,0,1>"}};
	static const char* ptags[3][2] = {{"hh_factor_kernel<8,1,0>", "hh_factor_kernel<8,1,1>"},
	                                  {"hh_factor_kernel<16,1,0>", "hh_factor_kernel<16,1,1>"},
	                                  {"hh_factor_kernel<32,1,0>", "hh_factor_kernel<32,1,1>"}};
	const int ri = rpt == 8 ? 0 : (rpt == 16 ? 1 : 2);
	std::string tag = (PIV ? ptags : tags)[ri][FORMQ ? 1 : 0];
	if (prof_shapes()) tag += " rows=" + std::to_string(a.rows) + " n=" + std::to_string(a.n) + " blocks=" + std::to_string(blocks);
	Work work;
	if (prof_enabled()) {
		// level matrix in; reflectors, τ, the blocks' R rows and (FORMQ) Q out
		const double k = std::min(a.rpb, a.n);
		work.flops = factor_flops(a, blocks, FORMQ);
		work.bytes = 8.0 * ((double)a.rows * a.n + (a.V ? (double)a.rows * a.n + (double)blocks * a.n : 0.0) +
		                    (a.R ? (double)blocks * k * a.n : 0.0) + (FORMQ ? (double)a.rows * a.nc : 0.0));
	}
	if (rpt == 8) launch_timed(tag.c_str(), work, hh_factor_kernel<8, PIV, FORMQ>, g, t, 0, s, a);
	else if (rpt == 16) launch_timed(tag.c_str(), work, hh_factor_kernel<16, PIV, FORMQ>, g, t, 0, s, a);
	else launch_timed(tag.c_str(), work, hh_factor_kernel<32, PIV, FORMQ>, g, t, 0, s, a);
}

void launch_factor(const FactorArgs& a, int blocks, bool piv, hipStream_t s) {
	const bool formq = a.Q != nullptr;
	XTT_REQUIRE(!formq || blocks == 1, "hh_factor: Q formation in the factor launch needs a single block");
	if (piv) {
		if (formq) factor_variant<true, true>(a, blocks, s); else factor_variant<true, false>(a, blocks, s);
	} else {
		if (formq) factor_variant<false, true>(a, blocks, s); else factor_variant<false, false>(a, blocks, s);
	}
	XTT_LAUNCH_CHECK();
}

template <int RPT>
void apply_rpt(const ApplyArgs& a, int blocks, hipStream_t s) {
	const size_t lds = 0;
	double f = 0;
	if (prof_enabled())
		for (int b = 0; b < blocks; ++b) {
			const double m = std::min(a.rpb, a.rows - b * a.rpb), k = std::min<double>(m, a.n);
			f += 4.0 * a.nc * (m * k - k * k / 2.0);
		}
	std::string tag = RPT == 8 ? "hh_apply_kernel<8>" : (RPT == 16 ? "hh_apply_kernel<16>" : "hh_apply_kernel<32>");
	if (prof_shapes()) tag += " rows=" + std::to_string(a.rows) + " n=" + std::to_string(a.n) + " nc=" + std::to_string(a.nc);
	// reflectors, τ and the blocks' start rows in; the level's Q rows out
	const Work work(f, 8.0 * ((double)a.rows * a.n + (double)blocks * a.n + (double)blocks * a.n * a.nc + (double)a.rows * a.nc));
	launch_timed(tag.c_str(), work, hh_apply_kernel<RPT>, dim3(blocks), dim3(HT), lds, s, a);
}

void launch_apply(const ApplyArgs& a, int blocks, hipStream_t s) {
	const int rpt = rows_per_thread(a.rpb);
	if (rpt == 8) apply_rpt<8>(a, blocks, s);
	else if (rpt == 16) apply_rpt<16>(a, blocks, s);
	else apply_rpt<32>(a, blocks, s);
	XTT_LAUNCH_CHECK();
}

// C_b ← Q_bᵀ·C_b = H_{k-1}·…·H_0·C_b for every block b of a factored level, in place: C holds the level's rows
// (ld ldc), 64 columns per grid.y chunk. The trailing-column update of a panel factorisation; reflectors are
// held and read as in hh_apply_kernel and applied in forward order.
struct ApplyTArgs {
	const double* V;
	const double* tau;
	int rows, rpb, n;
	double* C;
	long long ldc;
	int nc;
};

template <int RPT>
__global__ __launch_bounds__(HT) void hh_applyT_kernel(ApplyTArgs a) {
	__shared__ double sW[2][HW][HN];
	__shared__ double sTau[HN];
	const int lane = threadIdx.x & 63;
	const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
	const int b = blockIdx.x;
	const int r0 = b * a.rpb;
	const int rows = min(a.rpb, a.rows - r0);
	const int n = a.n;
	const int k = min(rows, n);
	const int c0 = blockIdx.y * HN;
	const int ncc = min(HN, a.nc - c0);
	double vr[RPT];
#pragma unroll
	for (int t = 0; t < RPT; ++t) {
		const int r = t * HW + w;
		double z = 0.0;
		if (r < rows && lane < k) z = r > lane ? a.V[(size_t)(r0 + r) * n + lane] : (r == lane ? 1.0 : 0.0);
		vr[t] = z;
	}
	if (threadIdx.x < k) sTau[threadIdx.x] = a.tau[(size_t)b * n + threadIdx.x];
	double q[RPT], u[RPT];
#pragma unroll
	for (int t = 0; t < RPT; ++t) {
		const int r = t * HW + w;
		q[t] = (r < rows && lane < ncc) ? a.C[(size_t)(r0 + r) * a.ldc + c0 + lane] : 0.0;
	}
	__syncthreads();
	int ph = 0;
	// in chunk C every reflector i >= C·HW is zero on the rows of registers t < C
	auto chunk = [&](auto cc) {
		constexpr int C = decltype(cc)::value;
		if (C * HW >= k) return;
		const int iend = min(k, (C + 1) * HW);
		for (int i = C * HW; i < iend; ++i) {
			const double ti = sTau[i];
			if (ti == 0.0) continue;
			double c0s = 0.0, c1s = 0.0;
#pragma unroll
			for (int t = C; t < RPT; ++t) {
				u[t] = lane_value(vr[t], i);
				if ((t - C) & 1) c1s = fma(u[t], q[t], c1s);
				else c0s = fma(u[t], q[t], c0s);
			}
			sW[ph][w][lane] = c0s + c1s;
			__syncthreads();
			double wa = 0.0, wb = 0.0;
#pragma unroll
			for (int s2 = 0; s2 < HW; s2 += 2) {
				wa += sW[ph][s2][lane];
				wb += sW[ph][s2 + 1][lane];
			}
			const double wj = (wa + wb) * ti;
#pragma unroll
			for (int t = C; t < RPT; ++t) q[t] = fma(-u[t], wj, q[t]);
			ph ^= 1;
		}
	};
	chunk(std::integral_constant<int, 0>{});
	chunk(std::integral_constant<int, 1>{});
	chunk(std::integral_constant<int, 2>{});
	chunk(std::integral_constant<int, 3>{});
	chunk(std::integral_constant<int, 4>{});
	chunk(std::integral_constant<int, 5>{});
	chunk(std::integral_constant<int, 6>{});
	chunk(std::integral_constant<int, 7>{});
#pragma unroll
	for (int t = 0; t < RPT; ++t) {
		const int r = t * HW + w;
		if (r < rows && lane < ncc) a.C[(size_t)(r0 + r) * a.ldc + c0 + lane] = q[t];
	}
}

// After a level of a panel factorisation: level row r of block b (rb = r − b·rpb, k_b = min(rows_b, w) R rows)
// goes to the next level as [R_b row rb | C row r] when rb < k_b, else its C part joins the next panel's rows.
// Every block but the last is full, so the next-level row is b·w + rb and the remainder row b·(rpb − w) + rb − k_b.
__global__ void panel_split_kernel(const double* Rb, int w, const double* C, long long ldc, int nT, int rows, int rpb,
                                   double* next, long long ldNext, double* rem, long long remRow0) {
	const long long cols = w + nT;
	const long long total = (long long)rows * cols;
	for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long long)gridDim.x * blockDim.x) {
		const int r = (int)(e / cols), c = (int)(e - (long long)r * cols);
		const int b = r / rpb, rb = r - b * rpb;
		const int kb = min(min(rpb, rows - b * rpb), w);
		if (rb < kb) {
			const double v = c < w ? Rb[((size_t)b * w + rb) * w + c] : C[(size_t)r * ldc + (c - w)];
			next[(size_t)(b * w + rb) * ldNext + c] = v;
		} else if (c >= w) {
			rem[(size_t)(remRow0 + (long long)b * (rpb - w) + (rb - kb)) * nT + (c - w)] = C[(size_t)r * ldc + (c - w)];
		}
	}
}

template <int RPT>
void applyT_rpt(const ApplyTArgs& a, int blocks, hipStream_t s) {
	const size_t lds = 0;
	double f = 0;
	if (prof_enabled())
		for (int b = 0; b < blocks; ++b) {
			const double m = std::min(a.rpb, a.rows - b * a.rpb), k = std::min<double>(m, a.n);
			f += 4.0 * a.nc * (m * k - k * k / 2.0);
		}
	std::string tag = RPT == 8 ? "hh_applyT_kernel<8>" : (RPT == 16 ? "hh_applyT_kernel<16>" : "hh_applyT_kernel<32>");
	if (prof_shapes()) tag += " rows=" + std::to_string(a.rows) + " n=" + std::to_string(a.n) + " nc=" + std::to_string(a.nc);
	const Work work(f, 8.0 * ((double)a.rows * a.n + (double)blocks * a.n + 2.0 * a.rows * a.nc));
	launch_timed(tag.c_str(), work, hh_applyT_kernel<RPT>, dim3(blocks, (a.nc + HN - 1) / HN), dim3(HT), lds, s, a);
}

void launch_applyT(const ApplyTArgs& a, int blocks, hipStream_t s) {
	const int rpt = rows_per_thread(a.rpb);
	if (rpt == 8) applyT_rpt<8>(a, blocks, s);
	else if (rpt == 16) applyT_rpt<16>(a, blocks, s);
	else applyT_rpt<32>(a, blocks, s);
}

// Row blocks of one tree level (rows of the level matrix, n columns): a single block up to HBR rows; leaves of
// equal size (>= HBR/2 >= n rows each); inner levels of whole children (n rows each).
void level_geometry(int rows, bool leaves, int n, int& blocks, int& rpb) {
	if (rows <= HBR) {
		blocks = 1;
		rpb = rows;
	} else if (leaves) {
		blocks = (rows + HBR - 1) / HBR;
		rpb = (rows + blocks - 1) / blocks;
	} else {
		rpb = (HBR / n) * n;
		blocks = (rows + rpb - 1) / rpb;
	}
}

struct Level {
	int rows = 0, rpb = 0, blocks = 0;
	DBuf<double> V, tau, R;
};

// Factor the m×n matrix through the tree. The root's R (min(m,n)×n) goes to rootR when given; when rootQ is
// given the root also forms its Q (root rows × min(m,n), identity start) in the same launch. With rootIn set,
// the root is left to the caller: the levels below it are factored and *rootIn / *rootRows receive the root's
// input (the stacked R factors of the last level, ld n, owned by that level).
std::vector<Level> tree_factor(const double* A, long long ld, bool trans, int m, int n, double* rootR, double* rootQ,
                               hipStream_t s, const double** rootIn = nullptr, int* rootRows = nullptr) {
	std::vector<Level> lv;
	const double* in = A;
	int rows = m;
	for (;;) {
		Level L;
		L.rows = rows;
		level_geometry(rows, lv.empty(), n, L.blocks, L.rpb);
		const bool root = L.blocks == 1;
		if (root && rootIn) {
			*rootIn = in;
			*rootRows = rows;
			break;
		}
		FactorArgs f{};
		f.in = in; f.ld = lv.empty() ? ld : n; f.trans = lv.empty() && trans;
		f.rows = rows; f.rpb = L.rpb; f.n = n;
		if (!root || !rootQ) {  // reflectors are read back by a later Q formation
			L.V.resize((size_t)rows * n);
			L.tau.resize((size_t)L.blocks * n);
			f.V = L.V.get(); f.tau = L.tau.get();
		}
		double* R = rootR;
		if (!root || !rootR) {
			L.R.resize((size_t)L.blocks * n * n);
			R = L.R.get();
		}
		f.R = R;
		if (root && rootQ) { f.Q = rootQ; f.nc = std::min(rows, n); }
		launch_factor(f, L.blocks, false, s);
		lv.push_back(std::move(L));
		if (root) break;
		in = R;
		rows = lv.back().blocks * n;
	}
	return lv;
}

// Q (m × nc, ld nc) = (product of the levels [0, count)) · init, init the n × nc start of level count-1.
void tree_apply(std::vector<Level>& lv, int count, int n, const double* init, int nc, const int* ncDev, double* Q,
                hipStream_t s) {
	DBuf<double> tmp[2];
	int which = 0;
	const double* cur = init;
	for (int l = count - 1; l >= 0; --l) {
		Level& L = lv[l];
		double* out = Q;
		if (l > 0) {
			tmp[which].resize((size_t)L.rows * nc);
			out = tmp[which].get();
			which ^= 1;
		}
		ApplyArgs a{};
		a.V = L.V.get(); a.tau = L.tau.get(); a.rows = L.rows; a.rpb = L.rpb; a.n = n;
		a.init = cur; a.out = out; a.nc = nc; a.ncDev = ncDev;
		launch_apply(a, L.blocks, s);
		cur = out;
	}
}

void check_shape(size_t m, size_t n) {
	XTT_REQUIRE(m > 0 && n > 0, "factorisation of an empty matrix");
	XTT_REQUIRE(hh_supported(m, n), "register-resident QR: more than 64 columns");
}

// R factor (min(m,n) × n, upper triangular) of a row-major m × n matrix with more than 64 columns: panels of
// 64 columns, each factored as a TSQR tree whose nodes also apply their reflectors to the panel's trailing
// columns (hh_applyT_kernel). A node's R rows travel up with their trailing parts; all other rows form the next
// panel's matrix. Only R is kept, so the order of those rows is free.
void blocked_r(const double* A, long long m, long long n, double* R, hipStream_t s) {
	fill_f64(R, 0.0, (size_t)std::min(m, n) * n, s);
	DBuf<double> cur((size_t)m * n);
	XTT_HIP(hipMemcpyAsync(cur.get(), A, (size_t)m * n * sizeof(double), hipMemcpyDeviceToDevice, s));
	long long mp = m;
	for (long long c0 = 0; c0 < n && mp > 0; c0 += HN) {
		const int np = (int)(n - c0), w = std::min(np, HN), nT = np - w;
		// remainder rows of this panel: everything but the root's R rows
		const long long remRows = mp - std::min<long long>(mp, w);
		DBuf<double> rem((size_t)std::max<long long>(remRows, 1) * std::max(nT, 1));
		long long remRow = 0;
		DBuf<double> lvl, nextLvl;
		const double* L = cur.get();
		int rows = (int)mp;
		bool leaves = true;
		for (;;) {
			int blocks, rpb;
			level_geometry(rows, leaves, w, blocks, rpb);
			DBuf<double> V((size_t)rows * w), tau((size_t)blocks * w), Rb((size_t)blocks * w * w);
			FactorArgs f{};
			f.in = L; f.ld = np; f.trans = 0;
			f.rows = rows; f.rpb = rpb; f.n = w;
			f.V = V.get(); f.tau = tau.get(); f.R = Rb.get();
			launch_factor(f, blocks, false, s);
			if (nT > 0) {
				ApplyTArgs t{};
				t.V = V.get(); t.tau = tau.get(); t.rows = rows; t.rpb = rpb; t.n = w;
				t.C = const_cast<double*>(L) + w; t.ldc = np; t.nc = nT;
				launch_applyT(t, blocks, s);
			}
			const bool root = blocks == 1;
			const int nextRows = root ? std::min(rows, w) : (blocks - 1) * w + std::min(rows - (blocks - 1) * rpb, w);
			double* next;
			long long ldNext;
			if (root) {
				next = R + (size_t)c0 * n + c0;
				ldNext = n;
			} else {
				nextLvl.resize((size_t)nextRows * np);
				next = nextLvl.get();
				ldNext = np;
			}
			const long long total = (long long)rows * np;
			const unsigned g = (unsigned)std::min<long long>(4096, (total + 255) / 256);
			hipLaunchKernelGGL(panel_split_kernel, dim3(g), dim3(256), 0, s, (const double*)Rb.get(), w, L + w, (long long)np, nT, rows,
			                   rpb, next, ldNext, rem.get(), remRow);
			XTT_LAUNCH_CHECK();
			remRow += rows - nextRows;
			if (root) break;
			std::swap(lvl, nextLvl);  // the level just written becomes the input
			L = lvl.get();
			rows = nextRows;
			leaves = false;
		}
		XTT_REQUIRE(remRow == remRows, "blocked QR: row bookkeeping");
		std::swap(cur, rem);
		mp = remRows;
	}
}

}  // namespace

void hh_r_blocked(const double* A, size_t m, size_t n, double* R, hipStream_t s) {
	XTT_REQUIRE(m > 0 && n > 0 && m <= (size_t)INT32_MAX / HN && n <= (size_t)INT32_MAX / HN, "blocked QR: matrix size");
	blocked_r(A, (long long)m, (long long)n, R, s);
}

bool hh_supported(size_t m, size_t n) { return n >= 1 && n <= (size_t)HN && m >= 1 && m <= (size_t)(INT32_MAX / HN); }

void hh_qr(const double* A, size_t m, size_t n, bool trans, double* Q, double* R, hipStream_t s) {
	check_shape(m, n);
	const int k = (int)std::min(m, n);
	const long long ld = trans ? (long long)m : (long long)n;
	if (!Q) {
		tree_factor(A, ld, trans, (int)m, (int)n, R, nullptr, s);
		return;
	}
	if (m <= (size_t)HBR) {
		tree_factor(A, ld, trans, (int)m, (int)n, R, Q, s);
		return;
	}
	DBuf<double> rootQ;
	// the root has (blocks of the level below)·n rows, at most HBR
	rootQ.resize((size_t)HBR * k);
	std::vector<Level> lv = tree_factor(A, ld, trans, (int)m, (int)n, R, rootQ.get(), s);
	tree_apply(lv, (int)lv.size() - 1, (int)n, rootQ.get(), k, nullptr, Q, s);
}

void hh_qc(const double* A, size_t m, size_t n, double* Q, double* C, int* jpvt, int* rank, hipStream_t s) {
	check_shape(m, n);
	FactorArgs f{};
	f.n = (int)n; f.jpvt = jpvt; f.rank = rank; f.C = C;
	if (m <= (size_t)HBR) {
		f.in = A; f.ld = (long long)n; f.rows = f.rpb = (int)m;
		f.Q = Q;
		launch_factor(f, 1, true, s);
		return;
	}
	// The levels below the root reduce A to the stacked R factors of its last level (at most HBR rows); the
	// pivoted factorisation runs on that stack directly. Its column norms are A's, since every level's Q is
	// orthogonal.
	const double* rootIn = nullptr;
	int rootRows = 0;
	std::vector<Level> lv = tree_factor(A, (long long)n, false, (int)m, (int)n, nullptr, nullptr, s, &rootIn, &rootRows);
	DBuf<double> Q2((size_t)rootRows * n);
	f.in = rootIn; f.ld = (long long)n; f.rows = f.rpb = rootRows;
	f.orig = A; f.origRows = (int)m;
	f.Q = Q2.get();
	launch_factor(f, 1, true, s);
	tree_apply(lv, (int)lv.size(), (int)n, Q2.get(), (int)n, rank, Q, s);
}

}  // namespace xtt
