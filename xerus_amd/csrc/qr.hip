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
// left. Q is formed top-down, each level in one launch of hh_wy_kernel: the root applies its reflectors to the
// identity, every node applies its reflectors to its n×c slice of the parent's result, both in compact WY form
// from the reflector Gram matrix that the factorisation leaves behind. The result is a Householder QR of the
// whole matrix (Q with orthonormal columns also for rank-deficient input); its diagonal signs follow the tree,
// not LAPACK's single sweep, which only changes the TT gauge.
//
// Column pivoting (xerus's QC, blasLapackWrapper.cpp:244-305 with dgeqp3) runs on one block: directly on
// the matrix when it has at most 256 rows, else as the tree's root on the stacked R factors of the level
// below (same column norms and the same partial-norm downdating, since the levels' Q are orthogonal). The pivot rule and downdating are dlaqp2's: largest
// partial norm, first index on ties, recompute when the downdated norm falls below sqrt(eps/2). Columns never
// move between lanes: each lane tracks the position dlaqp2's swaps would give its column (the pivot and the
// column at position i exchange positions), ties go to the smallest position, and the outputs (reflectors, τ,
// Y, R's rows, jpvt, C) are written by position. The rank rule
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
	double* Y;           // blocks × (n × n): Y[l][i] = u_lᵀu_i for l < i (the reflectors' Gram matrix above the
	                     // diagonal, for the compact WY form of their product); may be null, entries l >= i unset
	// column pivoting (single block)
	int* jpvt;           // n: original column of each position
	int* rank;           // xerus's QC rank
	double* C;           // rank × n, C[row][jpvt[col]] = R[row][col] (row <= col)
	const double* orig;  // when set, R_00's sign is that of the original matrix (origRows × n, row-major)
	int origRows;
};

template <int RPT, bool PIV>
__device__ __forceinline__ void hh_factor_body(const FactorArgs& a, const int b) {
	__shared__ double sW[2][HW][HN];   // per-wave Σ_{own r > i} a_ri·a_rj (lane i: the norm² partial)
	__shared__ double sRow[2][HN];     // row i
	__shared__ double sR[HN][HN];      // finished rows of R (lanes >= row)
	__shared__ double sScal[HN], sTau[HN], sDiag[HN];
	__shared__ double sNrm[HW][HN];    // per-wave partial column norms (pivoting)
	__shared__ int sRank;
	const int lane = threadIdx.x & 63;
	const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
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
	double myScal = 1.0;          // the pivot lane of step i: the scale of reflector i (u_i = scal_i·a below the diagonal)
	int pos = lane;               // the column's position (pivoting), identical in every wave
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
			int pl = i;  // the lane holding step i's pivot column
			if (PIV) {
				// dlaqp2's pivot: the largest partial norm over positions [i, n), the first position on ties; the
				// pivot lane takes position i and the lane at position i takes the pivot's
				const double cand = (pos >= i && lane < n) ? vn1 : -1.0;
				const double mx = wave_max(cand);
				const bool hit = cand == mx && cand >= 0.0;
				const unsigned long long hits = __ballot(hit);
				const int at_i = __builtin_amdgcn_readfirstlane((int)__builtin_ctzll(__ballot(pos == i && lane < n)));
				int P;
				if (hits == 0) {  // all-NaN: position i
					P = i;
					pl = at_i;
				} else if ((hits & (hits - 1)) == 0) {  // a single maximum
					pl = __builtin_amdgcn_readfirstlane((int)__builtin_ctzll(hits));
					P = __builtin_amdgcn_readlane(pos, pl);
				} else {  // ties: the smallest position
					P = (int)(-wave_max(hit ? -(double)pos : -1e9));
					pl = __builtin_amdgcn_readfirstlane((int)__builtin_ctzll(__ballot(pos == P && lane < n)));
				}
				if (lane == pl) pos = i;
				else if (lane == at_i) pos = P;
			}
			const bool act = (PIV ? pos > i : lane > i) && lane < n;  // columns the reflector updates
			const bool isPiv = lane == pl;
			// phase A: d_j = Σ_{own r > i} a_ri·a_rj (unscaled reflector; the pivot lane gets the norm² partial), row i.
			// The pivot column on the wave's rows is read from its lane into scalar registers (v_readlane).
			{
				const double vc = lane_value(x[C], pl);
				v[C] = w > wi ? vc : 0.0;
#pragma unroll
				for (int t = C + 1; t < RPT; ++t) v[t] = lane_value(x[t], pl);
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
			const double xn2 = lane_value(s, pl);  // the pivot lane's sum is ‖a(i+1:, i)‖²
			const double alpha = sRow[p][pl];
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
			// f_j = a_ij + scal·d_j. Updated columns: the update's w_j = tau·f_j. The lanes of earlier steps hold
			// reflector l = their position (unscaled) on rows > l, so u_lᵀu_i = scal_l·(a_il + scal_i·d_l) = scal_l·f_l.
			const double f = fma(scal, s, sRow[p][lane]);
			const double wj = act ? tau * f : 0.0;
			const double rij = sRow[p][lane] - wj;  // row i after the update (updated columns)
			if (isPiv) myScal = scal;
			const bool earlier = (PIV ? pos < i : lane < i) && lane < n;
			if (a.Y && w == 0 && earlier) a.Y[((size_t)b * n + (PIV ? pos : lane)) * n + i] = myScal * f;
			if (w == 0) {
				sR[i][lane] = act ? rij : (isPiv ? beta : 0.0);  // by lane: the column's own entry of row i
				if (lane == 0) {
					sTau[i] = tau;
					sScal[i] = scal;
					sDiag[i] = beta;
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
				if (act && vn1 != 0.0) {
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

	// compact output by position: R on and above the diagonal, the scaled reflectors u below
	const int P = PIV ? pos : lane;
#pragma unroll
	for (int t = 0; t < RPT; ++t) {
		const int r = t * HW + w;
		const double u = P < k && r > P ? x[t] * sScal[P] : 0.0;
		if (a.V && r < rows && lane < n) a.V[(size_t)(r0 + r) * n + P] = (r < k && P >= r) ? sR[r][lane] : u;
	}
	if (a.tau && threadIdx.x < k) a.tau[(size_t)b * n + threadIdx.x] = sTau[threadIdx.x];
	if (!PIV) {
		if (a.R) {
			for (int r = w; r < k; r += HW)
				if (lane < n) a.R[((size_t)b * n + r) * n + lane] = sR[r][lane];
		}
	} else {
		double r00 = sDiag[0];
		if (a.orig) {
			// R_00 as LAPACK's single sweep over the original matrix gives it: beta of the pivot column p0
			const int p0 = __builtin_amdgcn_readfirstlane((int)__builtin_ctzll(__ballot(pos == 0 && lane < n)));
			int nz = 0;
			for (int r = 1 + threadIdx.x; r < a.origRows; r += HT) nz |= a.orig[(size_t)r * n + p0] != 0.0;
			nz = __syncthreads_or(nz);
			const double alpha0 = a.orig[p0];
			r00 = nz ? -copysign(fabs(sDiag[0]), alpha0) : alpha0;
		}
		if (threadIdx.x == 0) {
			int rk;
			for (rk = 1; rk <= k; ++rk)
				if (rk == k || fabs(sDiag[rk]) < 16 * DBL_EPSILON * r00) break;
			sRank = rk;
			*a.rank = rk;
		}
		if (w == 0 && lane < n) a.jpvt[pos] = lane;
		__syncthreads();
		const int rank = sRank;
		for (int r = w; r < rank; r += HW)
			if (lane < n) a.C[(size_t)r * n + lane] = sR[r][lane];  // C[row][original column]
	}
}

template <int RPT, bool PIV>
__global__ __launch_bounds__(HT) void hh_factor_kernel(FactorArgs a) { hh_factor_body<RPT, PIV>(a, (int)blockIdx.x); }
// many factorisations in one launch: task t owns workgroups [start[t], start[t+1])
template <int RPT, bool PIV>
__global__ __launch_bounds__(HT) void hh_factor_group_kernel(const FactorArgs* tasks, const int* start, int count) {
	const int t = group_task(start, count, (int)blockIdx.x);
	const FactorArgs a = tasks[t];
	hh_factor_body<RPT, PIV>(a, (int)blockIdx.x - start[t]);
}

struct ApplyArgs {
	const double* V;      // a factored level (same geometry as its FactorArgs)
	const double* tau;
	const double* Y;      // the level's reflector Gram matrices (FactorArgs::Y)
	int rows, rpb, n;
	const double* init;   // (blocks·n) × nc, ld nc: block b starts from rows [b·n, b·n+n); null: identity
	double* out;          // rows × nc, ld nc
	int nc;               // columns of the product (<= 64)
	const int* ncDev;     // when set, the column count is read here (<= nc)
};

typedef double wy_d4 __attribute__((ext_vector_type(4)));
// v_mfma_f64_16x16x4f64: lane (li, lk) = (lane & 15, lane >> 4) supplies A[li][lk] and B[lk][li] and holds
// D[4v + lk][li] in component v.
__device__ inline wy_d4 mfma_f64(double a, double b, wy_d4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }

// out_b = H_0·…·H_{k-1}·[init_b; 0] for every block b of a factored level, in compact WY form (Schreiber–Van
// Loan): H_0·…·H_{k-1} = I − U·T·Uᵀ, U the unit lower trapezoidal reflectors, T upper triangular with dlarft's
// forward recurrence T[0:i, i] = −τ_i·T[0:i, 0:i]·Y[0:i, i], Y = UᵀU above the diagonal.
//   staging: U's top k×k triangle (unit diagonal, zero above), Y, τ and init_b into LDS;
//   phase 1: P = Uᵀ·[init_b; 0] = U_topᵀ·init_b (k × nc), 16×16 MFMA tiles on waves 0..6;
//   phase 2: Z = T·P in chunks of 8 rows from the bottom, Z_C = T_CC·(P_C − Y_{C,>C}·Z_{>C}); the diagonal blocks
//            T_CC are formed during phase 1 by one wave, a lane per row, from T[w][j] = −τ_j·Σ_{l=w}^{j-1} T[w][l]·Y[l][j]
//            (only row w enters);
//   phase 3: out = [init_b; 0] − U·Z, 16×16 MFMA tiles, U read per lane with its triangle masked.
// τ_i = 0 (an identity reflector) zeroes row and column i of T.
struct WyLds {
	double U[HN][HN + 1];   // U's top triangle, zero outside k×k (odd stride: column reads spread over banks)
	double Y[HN][HN];
	double I[HN][HN];       // init_b, zero outside k × nc
	double P[HN + HW][HN];  // P, then Z; zero outside k × nc
	double C[HW][HN];       // the chunk's right-hand sides
	double T[HN][HW];       // the 8×8 diagonal blocks of T: row i holds T[i][8·(i/8) + j]
	double tau[HN];
};

// Staging of one block (rows from U, ld n; k reflectors): U's triangle, Y, τ, P = 0 and, unless keepI, I from init
// (k × nc, ld nc; null: zero). All loads of a thread are in flight together.
__device__ __forceinline__ void wy_stage(WyLds& L, const double* U, const double* Yg, const double* tau, int n, int k, int nc,
                         const double* init, bool keepI) {
	constexpr int E = HN * HN / HT;
	const int tid = threadIdx.x;
	double u[E], y[E], z[E];
#pragma unroll
	for (int q = 0; q < E; ++q) {
		const int e = q * HT + tid, r = e / HN, c = e % HN;
		u[q] = (r < k && c < r) ? U[(size_t)r * n + c] : 0.0;
		y[q] = (r < c && c < k) ? Yg[(size_t)r * n + c] : 0.0;
		z[q] = (!keepI && init && r < k && c < nc) ? init[(size_t)r * nc + c] : 0.0;
	}
#pragma unroll
	for (int q = 0; q < E; ++q) {
		const int e = q * HT + tid, r = e / HN, c = e % HN;
		L.U[r][c] = (c == r && r < k) ? 1.0 : u[q];
		L.Y[r][c] = y[q];
		if (!keepI) L.I[r][c] = z[q];
		L.P[r][c] = 0.0;
	}
	if (tid < HW * HN) L.P[HN + tid / HN][tid % HN] = 0.0;
	if (tid < k) L.tau[tid] = tau[tid];
}

// Z = T·U_topᵀ·I into L.P (ident: I = the identity's first nc columns); after wy_stage and a barrier.
__device__ __forceinline__ void wy_solve(WyLds& L, int k, int nc, bool ident) {
	const int lane = threadIdx.x & 63;
	const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
	const int li = lane & 15, lk = lane >> 4;
	const int kq = (k + 3) >> 2;
	const int cT = (nc + 15) >> 4;
	if (w == HW - 1) {
		// T's diagonal blocks, lane i = 8C + w' forms row w' of block C:
		// T[i][c0 + j] = −τ_{c0+j}·Σ_{l=w'}^{j-1} T[i][c0 + l]·Y[c0 + l][c0 + j], T[i][i] = τ_i
		const int c0 = lane & ~(HW - 1), wr = lane & (HW - 1);
		double yb[HW * (HW - 1) / 2];  // Y[c0 + l][c0 + j], l < j, at j(j-1)/2 + l
#pragma unroll
		for (int j = 1; j < HW; ++j)
#pragma unroll
			for (int l = 0; l < j; ++l) yb[j * (j - 1) / 2 + l] = L.Y[c0 + l][c0 + j];
		double t[HW];
#pragma unroll
		for (int j = 0; j < HW; ++j) {
			double acc = 0.0;
#pragma unroll
			for (int l = 0; l < j; ++l) acc = fma(t[l], yb[j * (j - 1) / 2 + l], acc);  // t[l] = 0 for l < wr
			const double tau = c0 + j < k ? L.tau[c0 + j] : 0.0;
			t[j] = j < wr ? 0.0 : (j == wr ? tau : -tau * acc);
		}
#pragma unroll
		for (int j = 0; j < HW; ++j) L.T[lane][j] = t[j];
	} else if (ident) {  // phase 1: P = U_topᵀ restricted to the first nc columns (waves 0..6)
		for (int e = threadIdx.x; e < HN * HN; e += (HW - 1) * 64) {
			const int i = e >> 6, c = e & 63;
			if (i < k && c < nc) L.P[i][c] = L.U[c][i];
		}
	} else {  // phase 1: P = U_topᵀ·I (waves 0..6)
		const int kT = (k + 15) >> 4;
		for (int tile = w; tile < kT * cT; tile += HW - 1) {
			const int ti = tile / cT, tc = tile - ti * cT;
			wy_d4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
			for (int q = 0; q < HN / 4; ++q) {
				if (q < kq) {
					const int r = 4 * q + lk;
					acc = mfma_f64(L.U[r][ti * 16 + li], L.I[r][tc * 16 + li], acc);
				}
			}
#pragma unroll
			for (int v = 0; v < 4; ++v) L.P[ti * 16 + 4 * v + lk][tc * 16 + li] = acc[v];
		}
	}
	__syncthreads();
	// phase 2: Z_C = T_CC·(P_C − Y_{C,>C}·Z_{>C}) from the last chunk up
	for (int c0 = ((k - 1) / HW) * HW; c0 >= 0; c0 -= HW) {
		const int i = c0 + w;
		if (i < k) {
			double a0 = L.P[i][lane], a1 = 0.0, a2 = 0.0, a3 = 0.0;
			int l = c0 + HW;
			for (; l + 3 < k; l += 4) {
				a0 = fma(-L.Y[i][l], L.P[l][lane], a0);
				a1 = fma(-L.Y[i][l + 1], L.P[l + 1][lane], a1);
				a2 = fma(-L.Y[i][l + 2], L.P[l + 2][lane], a2);
				a3 = fma(-L.Y[i][l + 3], L.P[l + 3][lane], a3);
			}
			for (; l < k; ++l) a0 = fma(-L.Y[i][l], L.P[l][lane], a0);
			L.C[w][lane] = (a0 + a1) + (a2 + a3);
		}
		__syncthreads();
		if (i < k) {
			double z = 0.0;
#pragma unroll
			for (int j = 0; j < HW; ++j)
				if (c0 + j < k) z = fma(L.T[i][j], L.C[j][lane], z);  // T[i][c0 + j] = 0 for j < w
			L.P[i][lane] = z;
		}
		__syncthreads();
	}
}

// out (rows × nc, ld nc) = E − U_rows·Z: U's rows g0 .. g0 + rows - 1 (ld n, k reflectors: entry (g, i) is a
// reflector entry for g > i, one for g == i, zero above), E[r][c] = (g0 + r == c) when ident, else I[r][c] for the
// rows r < k. 16×16 MFMA tiles, two per pass for independent MFMA chains; each lane loads its own U entries.
__device__ __forceinline__ void wy_product(const WyLds& L, const double* U, int n, int g0, int rows, int k, int nc, bool ident, double* out,
                           bool toLds, double (*lds)[HN]) {
	const int lane = threadIdx.x & 63;
	const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
	const int li = lane & 15, lk = lane >> 4;
	const int kq = (k + 3) >> 2;
	const int cT = (nc + 15) >> 4;
	const int rT = (rows + 15) >> 4;
	const int nt = rT * cT;
	for (int t0 = w; t0 < nt; t0 += 2 * HW) {
		const int t1 = t0 + HW;
		const bool two = t1 < nt;
		const int tr0 = t0 / cT, tc0 = t0 - tr0 * cT;
		const int tr1 = two ? t1 / cT : tr0, tc1 = two ? t1 - tr1 * cT : tc0;
		const int ra0 = tr0 * 16 + li, ra1 = tr1 * 16 + li;  // rows of U relative to g0
		const int ga0 = g0 + ra0, ga1 = g0 + ra1;
		const double* u0 = U + (size_t)(g0 + min(ra0, rows - 1)) * n;
		const double* u1 = U + (size_t)(g0 + min(ra1, rows - 1)) * n;
		double x0[HN / 4], x1[HN / 4];
#pragma unroll
		for (int q = 0; q < HN / 4; ++q) {
			const int ic = min(4 * q + lk, k - 1);
			x0[q] = q < kq ? u0[ic] : 0.0;
			x1[q] = q < kq ? u1[ic] : 0.0;
		}
		wy_d4 acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
		for (int q = 0; q < HN / 4; ++q) {
			if (q < kq) {
				const int i = 4 * q + lk;
				const double a0 = (ra0 < rows && i < k) ? (i < ga0 ? x0[q] : (i == ga0 ? 1.0 : 0.0)) : 0.0;
				const double a1 = (ra1 < rows && i < k) ? (i < ga1 ? x1[q] : (i == ga1 ? 1.0 : 0.0)) : 0.0;
				acc0 = mfma_f64(a0, L.P[i][tc0 * 16 + li], acc0);
				acc1 = mfma_f64(a1, L.P[i][tc1 * 16 + li], acc1);
			}
		}
#pragma unroll
		for (int v = 0; v < 4; ++v) {
#pragma unroll
			for (int h = 0; h < 2; ++h) {
				if (h == 1 && !two) continue;
				const int c = (h ? tc1 : tc0) * 16 + li, r = (h ? tr1 : tr0) * 16 + 4 * v + lk;
				if (r < rows && c < nc) {
					const int g = g0 + r;
					const double base = ident ? (g == c ? 1.0 : 0.0) : (r < k ? L.I[r][c] : 0.0);
					const double val = base - (h ? acc1[v] : acc0[v]);
					if (toLds) lds[r][c] = val;
					else out[(size_t)r * nc + c] = val;
				}
			}
		}
	}
}

__device__ __forceinline__ void hh_wy_body(const ApplyArgs& a, const int b) {
	__shared__ WyLds L;
	const int r0 = b * a.rpb;
	const int rows = min(a.rpb, a.rows - r0);
	const int n = a.n;
	const int k = min(rows, n);
	const int nc = a.ncDev ? __builtin_amdgcn_readfirstlane(*a.ncDev) : a.nc;
	const double* U = a.V + (size_t)r0 * n;
	const bool ident = a.init == nullptr;
	wy_stage(L, U, a.Y + (size_t)b * n * n, a.tau + (size_t)b * n, n, k, nc,
	         ident ? nullptr : a.init + (size_t)b * n * nc, false);
	__syncthreads();
	wy_solve(L, k, nc, ident);
	wy_product(L, U, n, 0, rows, k, nc, ident, a.out + (size_t)r0 * nc, false, nullptr);
}

// Two tree levels in one launch, the root started from the identity: leaf block b first forms rows
// [b·n, b·n + n) of the root's Q itself (every leaf block repeats the root's staging and triangular solve), then
// applies its own reflectors to them. Saves the root's launch and its round trip through global memory.
__device__ __forceinline__ void hh_wy2_body(const ApplyArgs& a, const ApplyArgs& root, const int b) {
	__shared__ WyLds L;
	const int r0 = b * a.rpb;
	const int rows = min(a.rpb, a.rows - r0);
	const int n = a.n;
	const int k = min(rows, n);
	const int kr = min(root.rows, n);
	const int nc = a.ncDev ? __builtin_amdgcn_readfirstlane(*a.ncDev) : a.nc;
	wy_stage(L, root.V, root.Y, root.tau, n, kr, nc, nullptr, false);
	__syncthreads();
	wy_solve(L, kr, nc, true);
	wy_product(L, root.V, n, b * n, k, kr, nc, true, nullptr, true, L.I);
	__syncthreads();
	const double* U = a.V + (size_t)r0 * n;
	wy_stage(L, U, a.Y + (size_t)b * n * n, a.tau + (size_t)b * n, n, k, nc, nullptr, true);
	__syncthreads();
	wy_solve(L, k, nc, false);
	wy_product(L, U, n, 0, rows, k, nc, false, a.out + (size_t)r0 * nc, false, nullptr);
}

__global__ __launch_bounds__(HT) void hh_wy_kernel(ApplyArgs a) { hh_wy_body(a, (int)blockIdx.x); }
__global__ __launch_bounds__(HT) void hh_wy2_kernel(ApplyArgs a, ApplyArgs root) { hh_wy2_body(a, root, (int)blockIdx.x); }
__global__ __launch_bounds__(HT) void hh_wy_group_kernel(const ApplyArgs* tasks, const int* start, int count) {
	const int t = group_task(start, count, (int)blockIdx.x);
	const ApplyArgs a = tasks[t];
	hh_wy_body(a, (int)blockIdx.x - start[t]);
}
// tasks[2t] is a leaf level, tasks[2t + 1] its single-block root
__global__ __launch_bounds__(HT) void hh_wy2_group_kernel(const ApplyArgs* tasks, const int* start, int count) {
	const int t = group_task(start, count, (int)blockIdx.x);
	const ApplyArgs a = tasks[2 * t], root = tasks[2 * t + 1];
	hh_wy2_body(a, root, (int)blockIdx.x - start[t]);
}

int rows_per_thread(int rpb) { return rpb <= 8 * HW ? 8 : (rpb <= 16 * HW ? 16 : (rpb <= 24 * HW ? 24 : 32)); }

// Householder flops of one level (dgeqrf's count per block: 2n²(m − n/3) for m >= n, 2m²(n − m/3) otherwise).
double factor_flops(const FactorArgs& a, int blocks) {
	double f = 0;
	for (int b = 0; b < blocks; ++b) {
		const double m = std::min(a.rpb, a.rows - b * a.rpb), n = a.n;
		f += m >= n ? 2.0 * n * n * (m - n / 3.0) : 2.0 * m * m * (n - m / 3.0);
	}
	return f;
}

template <bool PIV>
void factor_variant(const FactorArgs& a, int blocks, hipStream_t s) {
	const int rpt = rows_per_thread(a.rpb);
	const dim3 g(blocks), t(HT);
	static const char* tags[4] = {"hh_factor_kernel<8,0>", "hh_factor_kernel<16,0>", "hh_factor_kernel<24,0>", "hh_factor_kernel<32,0>"};
	static const char* ptags[4] = {"hh_factor_kernel<8,1>", "hh_factor_kernel<16,1>", "hh_factor_kernel<24,1>", "hh_factor_kernel<32,1>"};
	const int ri = rpt == 8 ? 0 : (rpt == 16 ? 1 : (rpt == 24 ? 2 : 3));
	std::string tag = (PIV ? ptags : tags)[ri];
	if (prof_shapes()) tag += " rows=" + std::to_string(a.rows) + " n=" + std::to_string(a.n) + " blocks=" + std::to_string(blocks);
	Work work;
	if (prof_enabled()) {
		// level matrix in; reflectors, τ, Y and the blocks' R rows out
		const double k = std::min(a.rpb, a.n);
		work.flops = factor_flops(a, blocks);
		work.bytes = 8.0 * ((double)a.rows * a.n + (a.V ? (double)a.rows * a.n + (double)blocks * a.n : 0.0) +
		                    (a.R ? (double)blocks * k * a.n : 0.0) + (a.Y ? (double)blocks * k * (k - 1) / 2.0 : 0.0));
	}
	if (rpt == 8) launch_timed(tag.c_str(), work, hh_factor_kernel<8, PIV>, g, t, 0, s, a);
	else if (rpt == 16) launch_timed(tag.c_str(), work, hh_factor_kernel<16, PIV>, g, t, 0, s, a);
	else if (rpt == 24) launch_timed(tag.c_str(), work, hh_factor_kernel<24, PIV>, g, t, 0, s, a);
	else launch_timed(tag.c_str(), work, hh_factor_kernel<32, PIV>, g, t, 0, s, a);
}

void launch_factor(const FactorArgs& a, int blocks, bool piv, hipStream_t s) {
	XTT_REQUIRE(!piv || blocks == 1, "hh_factor: column pivoting needs a single block");
	if (piv) factor_variant<true>(a, blocks, s);
	else factor_variant<false>(a, blocks, s);
	XTT_LAUNCH_CHECK();
}

void launch_wy(const ApplyArgs& a, int blocks, hipStream_t s) {
	double f = 0;
	if (prof_enabled())
		for (int b = 0; b < blocks; ++b) {
			const double m = std::min(a.rpb, a.rows - b * a.rpb), k = std::min<double>(m, a.n);
			f += 2.0 * a.nc * (m * k + k * k);
		}
	std::string tag = "hh_wy_kernel";
	if (prof_shapes()) tag += " rows=" + std::to_string(a.rows) + " n=" + std::to_string(a.n) + " nc=" + std::to_string(a.nc);
	// reflectors, τ, Y and the blocks' start rows in; the level's Q rows out
	const double kk = std::min(a.rpb, a.n);
	const Work work(f, 8.0 * ((double)a.rows * a.n + (double)blocks * (a.n + kk * (kk - 1) / 2.0) +
	                          (a.init ? (double)blocks * a.n * a.nc : 0.0) + (double)a.rows * a.nc));
	launch_timed(tag.c_str(), work, hh_wy_kernel, dim3(blocks), dim3(HT), 0, s, a);
	XTT_LAUNCH_CHECK();
}

// A leaf level (blocks blocks) under a single-block root, the root from the identity, in one launch.
void launch_wy2(const ApplyArgs& a, int blocks, const ApplyArgs& root, hipStream_t s) {
	double f = 0;
	if (prof_enabled()) {
		const double kr = std::min(root.rows, a.n);
		for (int b = 0; b < blocks; ++b) {
			const double m = std::min(a.rpb, a.rows - b * a.rpb), k = std::min<double>(m, a.n);
			f += 2.0 * a.nc * (m * k + k * k) + 2.0 * a.nc * (k * kr + kr * kr);
		}
	}
	std::string tag = "hh_wy2_kernel";
	if (prof_shapes()) tag += " rows=" + std::to_string(a.rows) + " n=" + std::to_string(a.n) + " nc=" + std::to_string(a.nc);
	// both levels' reflectors, τ and Y in; the leaf level's Q rows out
	const double kk = std::min(a.rpb, a.n);
	const Work work(f, 8.0 * ((double)a.rows * a.n + (double)root.rows * a.n + (double)(blocks + 1) * (a.n + kk * (kk - 1) / 2.0) +
	                          (double)a.rows * a.nc));
	launch_timed(tag.c_str(), work, hh_wy2_kernel, dim3(blocks), dim3(HT), 0, s, a, root);
	XTT_LAUNCH_CHECK();
}

// C_b ← Q_bᵀ·C_b = H_{k-1}·…·H_0·C_b for every block b of a factored level, in place: C holds the level's rows
// (ld ldc), 64 columns per grid.y chunk. The trailing-column update of a panel factorisation: lane i holds
// reflector u_i on the wave's rows, the waves read it with v_readlane, reflectors applied in forward order.
struct ApplyTArgs {
	const double* V;
	const double* tau;
	int rows, rpb, n;
	double* C;
	long long ldc;
	int nc;
};

template <int RPT>
__device__ __forceinline__ void hh_applyT_body(const ApplyTArgs& a, const int b, const int cy) {
	__shared__ double sW[2][HW][HN];
	__shared__ double sTau[HN];
	const int lane = threadIdx.x & 63;
	const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
	const int r0 = b * a.rpb;
	const int rows = min(a.rpb, a.rows - r0);
	const int n = a.n;
	const int k = min(rows, n);
	const int c0 = cy * HN;
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
template <int RPT>
__global__ __launch_bounds__(HT) void hh_applyT_kernel(ApplyTArgs a) { hh_applyT_body<RPT>(a, (int)blockIdx.x, (int)blockIdx.y); }
// many levels in one launch: task t owns workgroups [start[t], start[t+1]), row block fastest, then 64-column chunk
template <int RPT>
__global__ __launch_bounds__(HT) void hh_applyT_group_kernel(const ApplyTArgs* tasks, const int* start, int count) {
	const int t = group_task(start, count, (int)blockIdx.x);
	const ApplyTArgs a = tasks[t];
	const int l = (int)blockIdx.x - start[t];
	const int blocks = (a.rows + a.rpb - 1) / a.rpb;
	hh_applyT_body<RPT>(a, l % blocks, l / blocks);
}

// After a level of a panel factorisation: level row r of block b (rb = r − b·rpb, k_b = min(rows_b, w) R rows)
// goes to the next level as [R_b row rb | C row r] when rb < k_b, else its C part joins the next panel's rows.
// Every block but the last is full, so the next-level row is b·w + rb and the remainder row b·(rpb − w) + rb − k_b.
__device__ __forceinline__ void panel_split_elem(const double* Rb, int w, const double* C, long long ldc, int nT, int rows, int rpb,
                                                 double* next, long long ldNext, double* rem, long long remRow0, long long e) {
	const long long cols = w + nT;
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
__global__ void panel_split_kernel(const double* Rb, int w, const double* C, long long ldc, int nT, int rows, int rpb,
                                   double* next, long long ldNext, double* rem, long long remRow0) {
	const long long total = (long long)rows * (w + nT);
	for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long long)gridDim.x * blockDim.x)
		panel_split_elem(Rb, w, C, ldc, nT, rows, rpb, next, ldNext, rem, remRow0, e);
}
struct SplitTask {
	const double* Rb;
	const double* C;
	double* next;
	double* rem;
	long long ldc, ldNext, remRow0;
	int w, nT, rows, rpb;
};
constexpr int kSplitPerBlock = 4096;
__global__ __launch_bounds__(256) void panel_split_group_kernel(const SplitTask* tasks, const int* start, int count) {
	const int t = group_task(start, count, (int)blockIdx.x);
	const SplitTask a = tasks[t];
	const long long total = (long long)a.rows * (a.w + a.nT);
	const long long e0 = (long long)(blockIdx.x - start[t]) * kSplitPerBlock;
	const long long e1 = min(total, e0 + kSplitPerBlock);
	for (long long e = e0 + threadIdx.x; e < e1; e += 256)
		panel_split_elem(a.Rb, a.w, a.C, a.ldc, a.nT, a.rows, a.rpb, a.next, a.ldNext, a.rem, a.remRow0, e);
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
	std::string tag = "hh_applyT_kernel<" + std::to_string(RPT) + ">";
	if (prof_shapes()) tag += " rows=" + std::to_string(a.rows) + " n=" + std::to_string(a.n) + " nc=" + std::to_string(a.nc);
	const Work work(f, 8.0 * ((double)a.rows * a.n + (double)blocks * a.n + 2.0 * a.rows * a.nc));
	launch_timed(tag.c_str(), work, hh_applyT_kernel<RPT>, dim3(blocks, (a.nc + HN - 1) / HN), dim3(HT), lds, s, a);
}

void launch_applyT(const ApplyTArgs& a, int blocks, hipStream_t s) {
	const int rpt = rows_per_thread(a.rpb);
	if (rpt == 8) applyT_rpt<8>(a, blocks, s);
	else if (rpt == 16) applyT_rpt<16>(a, blocks, s);
	else if (rpt == 24) applyT_rpt<24>(a, blocks, s);
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
	DBuf<double> V, tau, Y, R;
};

// Factor the m×n matrix through the tree. The root's R (min(m,n)×n) goes to rootR when given. With wantQ every
// level keeps its reflectors, τ and Y for a later Q formation. With rootIn set, the root is left to the caller:
// the levels below it are factored and *rootIn / *rootRows receive the root's input (the stacked R factors of
// the last level, ld n, owned by that level).
std::vector<Level> tree_factor(const double* A, long long ld, bool trans, int m, int n, double* rootR, bool wantQ,
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
		if (wantQ) {
			L.V.resize((size_t)rows * n);
			L.tau.resize((size_t)L.blocks * n);
			L.Y.resize((size_t)L.blocks * n * n);
			f.V = L.V.get(); f.tau = L.tau.get(); f.Y = L.Y.get();
		}
		double* R = rootR;
		if (!root || !rootR) {
			L.R.resize((size_t)L.blocks * n * n);
			R = L.R.get();
		}
		f.R = R;
		launch_factor(f, L.blocks, false, s);
		lv.push_back(std::move(L));
		if (root) break;
		in = R;
		rows = lv.back().blocks * n;
	}
	return lv;
}

// Q (m × nc, ld nc) = (product of the levels [0, count)) · init, init the (blocks·n) × nc start of level
// count-1 (null: identity, nc = that level's min(rows, n)).
void tree_apply(std::vector<Level>& lv, int count, int n, const double* init, int nc, const int* ncDev, double* Q,
                hipStream_t s) {
	auto args = [&](Level& L, const double* in, double* out) {
		ApplyArgs a{};
		a.V = L.V.get(); a.tau = L.tau.get(); a.Y = L.Y.get(); a.rows = L.rows; a.rpb = L.rpb; a.n = n;
		a.init = in; a.out = out; a.nc = nc; a.ncDev = ncDev;
		return a;
	};
	if (count == 2 && !init && lv[1].blocks == 1) {  // leaves under the root: one launch
		launch_wy2(args(lv[0], nullptr, Q), lv[0].blocks, args(lv[1], nullptr, nullptr), s);
		return;
	}
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
		launch_wy(args(L, cur, out), L.blocks, s);
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
	std::vector<Level> lv = tree_factor(A, ld, trans, (int)m, (int)n, R, Q != nullptr, s);
	if (Q) tree_apply(lv, (int)lv.size(), (int)n, nullptr, k, nullptr, Q, s);
}

void hh_qc(const double* A, size_t m, size_t n, double* Q, double* C, int* jpvt, int* rank, hipStream_t s) {
	check_shape(m, n);
	// The pivoted factorisation runs on one block: A itself, or the stacked R factors that the levels below the
	// root reduce A to (at most HBR rows). Their column norms are A's, since every level's Q is orthogonal.
	const double* rootIn = A;
	int rootRows = (int)m;
	std::vector<Level> lv;
	if (m > (size_t)HBR) lv = tree_factor(A, (long long)n, false, (int)m, (int)n, nullptr, true, s, &rootIn, &rootRows);
	Level root;
	root.rows = root.rpb = rootRows;
	root.blocks = 1;
	root.V.resize((size_t)rootRows * n);
	root.tau.resize(n);
	root.Y.resize(n * n);
	FactorArgs f{};
	f.in = rootIn; f.ld = (long long)n; f.rows = f.rpb = rootRows; f.n = (int)n;
	f.V = root.V.get(); f.tau = root.tau.get(); f.Y = root.Y.get();
	f.jpvt = jpvt; f.rank = rank; f.C = C;
	if (m > (size_t)HBR) { f.orig = A; f.origRows = (int)m; }
	launch_factor(f, 1, true, s);
	lv.push_back(std::move(root));
	// Q's column count is the rank, read on the device by every level
	tree_apply(lv, (int)lv.size(), (int)n, nullptr, std::min(rootRows, (int)n), rank, Q, s);
}


namespace {
// Factor launches of many levels: tasks bucketed by rows per thread, one grouped launch per bucket.
void factor_group(const std::vector<std::pair<FactorArgs, int>>& tasks, bool piv, hipStream_t s) {
	std::vector<FactorArgs> cls[4];
	std::vector<int> st[4];
	int nb[4] = {0, 0, 0, 0};
	for (const auto& tb : tasks) {
		const int rpt = rows_per_thread(tb.first.rpb);
		const int c = rpt == 8 ? 0 : (rpt == 16 ? 1 : (rpt == 24 ? 2 : 3));
		cls[c].push_back(tb.first);
		st[c].push_back(nb[c]);
		nb[c] += tb.second;
	}
	for (int c = 0; c < 4; ++c) {
		if (cls[c].empty()) continue;
		const FactorArgs* dt = static_cast<const FactorArgs*>(table_upload(cls[c].data(), cls[c].size() * sizeof(FactorArgs), s));
		const int* ds = static_cast<const int*>(table_upload(st[c].data(), st[c].size() * sizeof(int), s));
		const int cnt = (int)cls[c].size();
		const dim3 g((unsigned)nb[c]), t(HT);
		if (piv) {
			if (c == 0) launch_timed("hh_factor_group_kernel<8,1>", Work(), hh_factor_group_kernel<8, true>, g, t, 0, s, dt, ds, cnt);
			else if (c == 1) launch_timed("hh_factor_group_kernel<16,1>", Work(), hh_factor_group_kernel<16, true>, g, t, 0, s, dt, ds, cnt);
			else if (c == 2) launch_timed("hh_factor_group_kernel<24,1>", Work(), hh_factor_group_kernel<24, true>, g, t, 0, s, dt, ds, cnt);
			else launch_timed("hh_factor_group_kernel<32,1>", Work(), hh_factor_group_kernel<32, true>, g, t, 0, s, dt, ds, cnt);
		} else {
			if (c == 0) launch_timed("hh_factor_group_kernel<8,0>", Work(), hh_factor_group_kernel<8, false>, g, t, 0, s, dt, ds, cnt);
			else if (c == 1) launch_timed("hh_factor_group_kernel<16,0>", Work(), hh_factor_group_kernel<16, false>, g, t, 0, s, dt, ds, cnt);
			else if (c == 2) launch_timed("hh_factor_group_kernel<24,0>", Work(), hh_factor_group_kernel<24, false>, g, t, 0, s, dt, ds, cnt);
			else launch_timed("hh_factor_group_kernel<32,0>", Work(), hh_factor_group_kernel<32, false>, g, t, 0, s, dt, ds, cnt);
		}
	}
}
void applyT_group(const std::vector<std::pair<ApplyTArgs, int>>& tasks, hipStream_t s) {
	std::vector<ApplyTArgs> cls[4];
	std::vector<int> st[4];
	int nb[4] = {0, 0, 0, 0};
	for (const auto& tb : tasks) {
		const int rpt = rows_per_thread(tb.first.rpb);
		const int c = rpt == 8 ? 0 : (rpt == 16 ? 1 : (rpt == 24 ? 2 : 3));
		cls[c].push_back(tb.first);
		st[c].push_back(nb[c]);
		nb[c] += tb.second * ((tb.first.nc + HN - 1) / HN);
	}
	for (int c = 0; c < 4; ++c) {
		if (cls[c].empty()) continue;
		const ApplyTArgs* dt = static_cast<const ApplyTArgs*>(table_upload(cls[c].data(), cls[c].size() * sizeof(ApplyTArgs), s));
		const int* ds = static_cast<const int*>(table_upload(st[c].data(), st[c].size() * sizeof(int), s));
		const int cnt = (int)cls[c].size();
		const dim3 g((unsigned)nb[c]), t(HT);
		if (c == 0) launch_timed("hh_applyT_group_kernel<8>", Work(), hh_applyT_group_kernel<8>, g, t, 0, s, dt, ds, cnt);
		else if (c == 1) launch_timed("hh_applyT_group_kernel<16>", Work(), hh_applyT_group_kernel<16>, g, t, 0, s, dt, ds, cnt);
		else if (c == 2) launch_timed("hh_applyT_group_kernel<24>", Work(), hh_applyT_group_kernel<24>, g, t, 0, s, dt, ds, cnt);
		else launch_timed("hh_applyT_group_kernel<32>", Work(), hh_applyT_group_kernel<32>, g, t, 0, s, dt, ds, cnt);
	}
}
void split_group(const std::vector<SplitTask>& tasks, hipStream_t s) {
	if (tasks.empty()) return;
	std::vector<int> start;
	int blocks = 0;
	for (const SplitTask& t : tasks) {
		start.push_back(blocks);
		blocks += (int)(((long long)t.rows * (t.w + t.nT) + kSplitPerBlock - 1) / kSplitPerBlock);
	}
	const SplitTask* dt = static_cast<const SplitTask*>(table_upload(tasks.data(), tasks.size() * sizeof(SplitTask), s));
	const int* ds = static_cast<const int*>(table_upload(start.data(), start.size() * sizeof(int), s));
	launch_timed("panel_split_group_kernel", Work(), panel_split_group_kernel, dim3((unsigned)blocks), dim3(256), 0, s, dt, ds,
	             (int)tasks.size());
}
// Q formation of many trees: a single level from the identity (hh_wy), or leaves under a single-block root
// (hh_wy2); anything deeper one tree at a time.
struct QTree {
	std::vector<Level>* lv;
	int n, nc;
	const int* ncDev;
	double* Q;
};
void qform_group(const std::vector<QTree>& trees, hipStream_t s) {
	std::vector<ApplyArgs> one, two;
	std::vector<int> s1, s2;
	int b1 = 0, b2 = 0;
	auto args = [](Level& L, int n, double* out, int nc, const int* ncDev) {
		ApplyArgs a{};
		a.V = L.V.get(); a.tau = L.tau.get(); a.Y = L.Y.get(); a.rows = L.rows; a.rpb = L.rpb; a.n = n;
		a.init = nullptr; a.out = out; a.nc = nc; a.ncDev = ncDev;
		return a;
	};
	for (const QTree& q : trees) {
		std::vector<Level>& lv = *q.lv;
		if (lv.size() == 1) {
			one.push_back(args(lv[0], q.n, q.Q, q.nc, q.ncDev));
			s1.push_back(b1);
			b1 += lv[0].blocks;
		} else if (lv.size() == 2 && lv[1].blocks == 1) {
			two.push_back(args(lv[0], q.n, q.Q, q.nc, q.ncDev));
			two.push_back(args(lv[1], q.n, nullptr, q.nc, q.ncDev));
			s2.push_back(b2);
			b2 += lv[0].blocks;
		} else {
			tree_apply(lv, (int)lv.size(), q.n, nullptr, q.nc, q.ncDev, q.Q, s);
		}
	}
	if (!one.empty()) {
		const ApplyArgs* dt = static_cast<const ApplyArgs*>(table_upload(one.data(), one.size() * sizeof(ApplyArgs), s));
		const int* ds = static_cast<const int*>(table_upload(s1.data(), s1.size() * sizeof(int), s));
		launch_timed("hh_wy_group_kernel", Work(), hh_wy_group_kernel, dim3((unsigned)b1), dim3(HT), 0, s, dt, ds, (int)one.size());
	}
	if (!two.empty()) {
		const ApplyArgs* dt = static_cast<const ApplyArgs*>(table_upload(two.data(), two.size() * sizeof(ApplyArgs), s));
		const int* ds = static_cast<const int*>(table_upload(s2.data(), s2.size() * sizeof(int), s));
		launch_timed("hh_wy2_group_kernel", Work(), hh_wy2_group_kernel, dim3((unsigned)b2), dim3(HT), 0, s, dt, ds, (int)s2.size());
	}
}
}  // namespace

void hh_qr_group(const std::vector<QRJob>& jobs, hipStream_t s) {
	const size_t J = jobs.size();
	ScratchFrame frame;  // the levels' reflectors, τ, Y and inner R factors
	std::vector<std::vector<Level>> lv(J);
	std::vector<const double*> in(J);
	std::vector<int> rows(J);
	std::vector<char> open(J, 1);
	for (size_t j = 0; j < J; ++j) {
		check_shape((size_t)jobs[j].m, (size_t)jobs[j].n);
		in[j] = jobs[j].A;
		rows[j] = jobs[j].m;
	}
	for (bool any = J > 0; any;) {  // one tree level of every open job per round
		std::vector<std::pair<FactorArgs, int>> tasks;
		for (size_t j = 0; j < J; ++j) {
			if (!open[j]) continue;
			const QRJob& q = jobs[j];
			const int n = q.n;
			Level L;
			L.rows = rows[j];
			level_geometry(rows[j], lv[j].empty(), n, L.blocks, L.rpb);
			const bool root = L.blocks == 1;
			FactorArgs f{};
			f.in = in[j]; f.ld = lv[j].empty() ? (q.trans ? q.m : n) : n; f.trans = lv[j].empty() && q.trans;
			f.rows = rows[j]; f.rpb = L.rpb; f.n = n;
			if (q.Q) {
				L.V.scratch((size_t)rows[j] * n);
				L.tau.scratch((size_t)L.blocks * n);
				L.Y.scratch((size_t)L.blocks * n * n);
				f.V = L.V.get(); f.tau = L.tau.get(); f.Y = L.Y.get();
			}
			double* R = q.R;
			if (!root || !q.R) {
				L.R.scratch((size_t)L.blocks * n * n);
				R = L.R.get();
			}
			f.R = R;
			tasks.emplace_back(f, L.blocks);
			lv[j].push_back(std::move(L));
		}
		factor_group(tasks, false, s);
		any = false;
		for (size_t j = 0; j < J; ++j) {
			if (!open[j]) continue;
			Level& L = lv[j].back();
			if (L.blocks == 1) { open[j] = 0; continue; }
			in[j] = L.R.get();
			rows[j] = L.blocks * jobs[j].n;
			any = true;
		}
	}
	std::vector<QTree> trees;
	for (size_t j = 0; j < J; ++j)
		if (jobs[j].Q) trees.push_back(QTree{&lv[j], jobs[j].n, std::min(jobs[j].m, jobs[j].n), nullptr, jobs[j].Q});
	qform_group(trees, s);
}

void hh_qc_group(const std::vector<QCJob>& jobs, hipStream_t s) {
	const size_t J = jobs.size();
	ScratchFrame frame;  // the levels' reflectors, τ, Y and leaf R factors
	std::vector<std::vector<Level>> lv(J);
	std::vector<const double*> rootIn(J);
	std::vector<int> rootRows(J);
	std::vector<char> grouped(J, 1);
	// the levels below the root (matrices taller than one block), one level: leaves whose stacked R factors fit one block
	std::vector<std::pair<FactorArgs, int>> leaves;
	for (size_t j = 0; j < J; ++j) {
		const QCJob& q = jobs[j];
		check_shape((size_t)q.m, (size_t)q.n);
		rootIn[j] = q.A;
		rootRows[j] = q.m;
		if (q.m <= HBR) continue;
		Level L;
		L.rows = q.m;
		level_geometry(q.m, true, q.n, L.blocks, L.rpb);
		if (L.blocks * q.n > HBR) { grouped[j] = 0; continue; }  // a deeper tree: one matrix at a time below
		L.V.scratch((size_t)q.m * q.n);
		L.tau.scratch((size_t)L.blocks * q.n);
		L.Y.scratch((size_t)L.blocks * q.n * q.n);
		L.R.scratch((size_t)L.blocks * q.n * q.n);
		FactorArgs f{};
		f.in = q.A; f.ld = q.n; f.rows = q.m; f.rpb = L.rpb; f.n = q.n;
		f.V = L.V.get(); f.tau = L.tau.get(); f.Y = L.Y.get(); f.R = L.R.get();
		leaves.emplace_back(f, L.blocks);
		rootIn[j] = L.R.get();
		rootRows[j] = L.blocks * q.n;
		lv[j].push_back(std::move(L));
	}
	factor_group(leaves, false, s);
	std::vector<std::pair<FactorArgs, int>> roots;
	for (size_t j = 0; j < J; ++j) {
		const QCJob& q = jobs[j];
		if (!grouped[j]) {
			hh_qc(q.A, (size_t)q.m, (size_t)q.n, q.Q, q.C, q.jpvt, q.rank, s);
			continue;
		}
		Level root;
		root.rows = root.rpb = rootRows[j];
		root.blocks = 1;
		root.V.scratch((size_t)rootRows[j] * q.n);
		root.tau.scratch(q.n);
		root.Y.scratch((size_t)q.n * q.n);
		FactorArgs f{};
		f.in = rootIn[j]; f.ld = q.n; f.rows = f.rpb = rootRows[j]; f.n = q.n;
		f.V = root.V.get(); f.tau = root.tau.get(); f.Y = root.Y.get();
		f.jpvt = q.jpvt; f.rank = q.rank; f.C = q.C;
		if (q.m > HBR) { f.orig = q.A; f.origRows = q.m; }
		roots.emplace_back(f, 1);
		lv[j].push_back(std::move(root));
	}
	factor_group(roots, true, s);
	std::vector<QTree> trees;
	for (size_t j = 0; j < J; ++j)
		if (grouped[j]) trees.push_back(QTree{&lv[j], jobs[j].n, std::min(rootRows[j], jobs[j].n), jobs[j].rank, jobs[j].Q});
	qform_group(trees, s);
}

// R factors of many row-major matrices of any width (blocked_r's panels and trees, all jobs one level per step:
// one grouped factor, trailing-update and split launch per step). The inputs are overwritten.
void hh_r_group(const std::vector<RJob>& jobs, hipStream_t s) {
	struct St {
		RJob j;
		double* cur;
		DBuf<double> curOwn, rem, lvl, nextLvl, V, tau, Rb;
		long long mp = 0, c0 = 0, remRows = 0, remRow = 0;
		int np = 0, w = 0, nT = 0, rows = 0, blocks = 0, rpb = 0, nextRows = 0;
		const double* L = nullptr;
		bool leaves = true, panel = true, done = false;
	};
	std::vector<St> S(jobs.size());
	std::vector<MapTask> zero;
	for (size_t i = 0; i < jobs.size(); ++i) {
		const RJob& j = jobs[i];
		XTT_REQUIRE(j.m > 0 && j.n > 0 && j.m <= INT32_MAX / HN && j.n <= INT32_MAX / HN, "blocked QR: matrix size");
		S[i].j = j;
		S[i].cur = j.A;
		S[i].mp = j.m;
		zero.push_back(map_fill(j.R, 0.0, (size_t)std::min(j.m, j.n) * j.n));
	}
	map_group(zero, s);
	for (;;) {
		ScratchFrame frame;  // the level's reflectors, τ and R blocks (the launches of this level are their last use)
		std::vector<std::pair<FactorArgs, int>> fa;
		std::vector<std::pair<ApplyTArgs, int>> ap;
		std::vector<SplitTask> sp;
		for (St& t : S) {
			if (t.done) continue;
			if (t.panel) {
				t.np = (int)(t.j.n - t.c0);
				t.w = std::min(t.np, HN);
				t.nT = t.np - t.w;
				t.remRows = t.mp - std::min<long long>(t.mp, t.w);
				t.rem.resize((size_t)std::max<long long>(t.remRows, 1) * std::max(t.nT, 1));
				t.remRow = 0;
				t.L = t.cur;
				t.rows = (int)t.mp;
				t.leaves = true;
				t.panel = false;
			}
			level_geometry(t.rows, t.leaves, t.w, t.blocks, t.rpb);
			t.V.scratch((size_t)t.rows * t.w);
			t.tau.scratch((size_t)t.blocks * t.w);
			t.Rb.scratch((size_t)t.blocks * t.w * t.w);
			FactorArgs f{};
			f.in = t.L; f.ld = t.np; f.trans = 0;
			f.rows = t.rows; f.rpb = t.rpb; f.n = t.w;
			f.V = t.V.get(); f.tau = t.tau.get(); f.R = t.Rb.get();
			fa.push_back({f, t.blocks});
			if (t.nT > 0) {
				ApplyTArgs a{};
				a.V = t.V.get(); a.tau = t.tau.get(); a.rows = t.rows; a.rpb = t.rpb; a.n = t.w;
				a.C = const_cast<double*>(t.L) + t.w; a.ldc = t.np; a.nc = t.nT;
				ap.push_back({a, t.blocks});
			}
			const bool root = t.blocks == 1;
			t.nextRows = root ? std::min(t.rows, t.w) : (t.blocks - 1) * t.w + std::min(t.rows - (t.blocks - 1) * t.rpb, t.w);
			SplitTask p{};
			p.Rb = t.Rb.get(); p.C = t.L + t.w; p.ldc = t.np; p.w = t.w; p.nT = t.nT; p.rows = t.rows; p.rpb = t.rpb;
			if (root) {
				p.next = t.j.R + (size_t)t.c0 * t.j.n + t.c0;
				p.ldNext = t.j.n;
			} else {
				t.nextLvl.resize((size_t)t.nextRows * t.np);
				p.next = t.nextLvl.get();
				p.ldNext = t.np;
			}
			p.rem = t.rem.get();
			p.remRow0 = t.remRow;
			sp.push_back(p);
		}
		if (fa.empty()) break;
		factor_group(fa, false, s);
		applyT_group(ap, s);
		split_group(sp, s);
		for (St& t : S) {
			if (t.done) continue;
			t.remRow += t.rows - t.nextRows;
			if (t.blocks == 1) {  // the panel's root: the remainder rows form the next panel
				XTT_REQUIRE(t.remRow == t.remRows, "blocked QR: row bookkeeping");
				std::swap(t.curOwn, t.rem);
				t.cur = t.curOwn.get();
				t.mp = t.remRows;
				t.c0 += HN;
				t.panel = true;
				t.done = t.c0 >= t.j.n || t.mp <= 0;
			} else {
				std::swap(t.lvl, t.nextLvl);
				t.L = t.lvl.get();
				t.rows = t.nextRows;
				t.leaves = false;
			}
		}
	}
	// no table_ring_release here: the stream has not been synchronised, and the tables' host-to-device copies
	// may still be queued behind the launches
}

}  // namespace xtt
