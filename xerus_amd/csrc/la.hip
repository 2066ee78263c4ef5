// FP64 dense factorisations for the orthogonalisation sweeps (move_core / round): Householder QR, column-
// pivoted QR with xerus's rank rule, explicit Q formation and one-sided Jacobi SVD. One workgroup owns one
// matrix; norms are wavefront + LDS tree reductions in a fixed order (deterministic).
//
// The Householder conventions are LAPACK's (dlarfg: beta = -sign(alpha)·‖x‖, tau = (beta-alpha)/beta,
// v = [1; x/(alpha-beta)]); the pivot rule and partial-norm downdating are dlaqp2's (largest partial norm,
// first index on ties; recompute when the downdated norm falls below sqrt(eps)). These fix the integer rank
// decisions that xerus takes from dgeqp3 (blasLapackWrapper.cpp:262-272).
#include <cfloat>
#include <mutex>

#include "kernels.hpp"
#include "xtt_common.hpp"

namespace xtt {
namespace {

constexpr int QT = 1024;  // threads of the factorisation workgroups

__device__ inline double wsum(double v) {
#pragma unroll
	for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
	return v;
}

// Block-wide sum, valid in all threads. red: QT/64 doubles of LDS.
__device__ inline double bsum(double v, double* red) {
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	v = wsum(v);
	__syncthreads();
	if (lane == 0) red[wave] = v;
	__syncthreads();
	double s = 0;
	for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += red[w];
	return s;
}

// Apply H = I - tau·v·vᵀ (v = [1; X[i+1:m, vc]]) from the left to the row-major m×N matrix Y, columns
// [c0, c1), rows [i, m). V is the matrix holding the reflector in column vc (stride ldv); w is LDS scratch
// for N doubles; red2 is 32×33 doubles of LDS.
__device__ void apply_reflector(double* Y, int ldy, const double* V, int ldv, int vc, int i, int m, int c0, int c1, double tau,
                                double* w, double* red2) {
	const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 column lanes × 32 row groups
	for (int jc = c0; jc < c1; jc += 32) {
		const int j = jc + tx;
		double s = 0;
		if (j < c1)
			for (int r = i + ty; r < m; r += 32) s += (r == i ? 1.0 : V[(size_t)r * ldv + vc]) * Y[(size_t)r * ldy + j];
		red2[ty * 33 + tx] = s;
		__syncthreads();
		if (ty == 0 && j < c1) {
			double t = 0;
			for (int g = 0; g < 32; ++g) t += red2[g * 33 + tx];
			w[j - c0] = t * tau;
		}
		__syncthreads();
	}
	for (int jc = c0; jc < c1; jc += 32) {
		const int j = jc + tx;
		if (j < c1) {
			const double wj = w[j - c0];
			for (int r = i + ty; r < m; r += 32) Y[(size_t)r * ldy + j] -= wj * (r == i ? 1.0 : V[(size_t)r * ldv + vc]);
		}
	}
	__syncthreads();
}

__global__ __launch_bounds__(QT) void geqp2_kernel(double* A, int m, int n, double* tau, int* jpvt, int pivot, int* rankOut) {
	extern __shared__ double sm[];
	double* vn1 = sm;            // n
	double* vn2 = sm + n;        // n
	double* w = sm + 2 * n;      // n
	double* red2 = sm + 3 * n;   // 32*33
	__shared__ double red[QT / 64];
	__shared__ double sTau, sScale, sBeta;
	__shared__ int sPvt;
	const int tid = threadIdx.x;
	const int k = min(m, n);
	const double tol3z = sqrt(DBL_EPSILON * 0.5);

	for (int j = tid; j < n; j += QT) {
		jpvt[j] = j;
		if (pivot) {
			double s = 0;
			for (int r = 0; r < m; ++r) { const double a = A[(size_t)r * n + j]; s += a * a; }
			vn1[j] = vn2[j] = sqrt(s);
		}
	}
	__syncthreads();

	for (int i = 0; i < k; ++i) {
		if (pivot) {
			// argmax of vn1 over [i, n), first index on ties
			double best = -1.0;
			int bidx = n;
			for (int j = i + tid; j < n; j += QT) if (vn1[j] > best) { best = vn1[j]; bidx = j; }
#pragma unroll
			for (int o = 32; o > 0; o >>= 1) {
				const double ob = __shfl_xor(best, o, 64);
				const int oi = __shfl_xor(bidx, o, 64);
				if (ob > best || (ob == best && oi < bidx)) { best = ob; bidx = oi; }
			}
			__shared__ double wb[QT / 64];
			__shared__ int wi[QT / 64];
			if ((tid & 63) == 0) { wb[tid >> 6] = best; wi[tid >> 6] = bidx; }
			__syncthreads();
			if (tid == 0) {
				double b = -1.0; int bi = n;
				for (int q = 0; q < (int)(blockDim.x >> 6); ++q)
					if (wb[q] > b || (wb[q] == b && wi[q] < bi)) { b = wb[q]; bi = wi[q]; }
				sPvt = bi;
			}
			__syncthreads();
			const int pvt = sPvt;
			if (pvt != i && pvt < n) {
				for (int r = tid; r < m; r += QT) {
					const double t = A[(size_t)r * n + i];
					A[(size_t)r * n + i] = A[(size_t)r * n + pvt];
					A[(size_t)r * n + pvt] = t;
				}
				__syncthreads();
				if (tid == 0) {
					const int t = jpvt[pvt]; jpvt[pvt] = jpvt[i]; jpvt[i] = t;
					vn1[pvt] = vn1[i];
					vn2[pvt] = vn2[i];
				}
			}
			__syncthreads();
		}
		// Householder reflector for column i (dlarfg)
		double part = 0;
		for (int r = i + 1 + tid; r < m; r += QT) { const double a = A[(size_t)r * n + i]; part += a * a; }
		const double xnorm = sqrt(bsum(part, red));
		if (tid == 0) {
			const double alpha = A[(size_t)i * n + i];
			if (xnorm == 0.0) {
				sTau = 0.0; sScale = 1.0; sBeta = alpha;
			} else {
				const double beta = -copysign(hypot(alpha, xnorm), alpha);
				sTau = (beta - alpha) / beta;
				sScale = 1.0 / (alpha - beta);
				sBeta = beta;
			}
			tau[i] = sTau;
		}
		__syncthreads();
		const double t_i = sTau;
		if (t_i != 0.0) {
			const double sc = sScale;
			for (int r = i + 1 + tid; r < m; r += QT) A[(size_t)r * n + i] *= sc;
		}
		if (tid == 0) A[(size_t)i * n + i] = sBeta;
		__syncthreads();
		if (t_i != 0.0 && i + 1 < n) apply_reflector(A, n, A, n, i, i, m, i + 1, n, t_i, w, red2);
		if (pivot) {
			for (int j = i + 1 + tid; j < n; j += QT) {
				if (vn1[j] == 0.0) continue;
				double temp = fabs(A[(size_t)i * n + j]) / vn1[j];
				temp = fmax(0.0, 1.0 - temp * temp);
				const double ratio = vn1[j] / vn2[j];
				const double temp2 = temp * ratio * ratio;
				if (temp2 <= tol3z) {
					if (i + 1 < m) {
						double s = 0;
						for (int r = i + 1; r < m; ++r) { const double a = A[(size_t)r * n + j]; s += a * a; }
						vn1[j] = sqrt(s);
						vn2[j] = vn1[j];
					} else {
						vn1[j] = 0.0;
						vn2[j] = 0.0;
					}
				} else {
					vn1[j] *= sqrt(temp);
				}
			}
			__syncthreads();
		}
	}
	if (tid == 0) {
		int rank = k;
		if (pivot) {
			// xerus: the first rank with |R_rr| < 16·ε·R_00 (R_00 with its sign), blasLapackWrapper.cpp:268-272
			for (rank = 1; rank <= k; ++rank)
				if (rank == k || fabs(A[(size_t)rank * n + rank]) < 16 * DBL_EPSILON * A[0]) break;
		}
		*rankOut = rank;
	}
}

__global__ __launch_bounds__(QT) void org2r_kernel(const double* A, int m, int n, const double* tau, double* Q, int kq) {
	extern __shared__ double sm[];
	double* w = sm;          // kq
	double* red2 = sm + kq;  // 32*33
	const int tid = threadIdx.x;
	for (size_t e = tid; e < (size_t)m * kq; e += QT) {
		const int r = (int)(e / kq), c = (int)(e % kq);
		Q[e] = (r == c) ? 1.0 : 0.0;
	}
	__syncthreads();
	for (int i = kq - 1; i >= 0; --i) {
		const double t = tau[i];
		if (t != 0.0) apply_reflector(Q, kq, A, n, i, i, m, i, kq, t, w, red2);
	}
}

__global__ void extract_c_kernel(const double* A, int m, int n, const int* jpvt, const int* rankp, double* C) {
	const int rank = *rankp;
	for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < (size_t)rank * n; e += (size_t)gridDim.x * blockDim.x) {
		const int row = (int)(e / n), col = (int)(e % n);
		C[(size_t)row * n + jpvt[col]] = (row <= col) ? A[(size_t)row * n + col] : 0.0;
	}
}

__global__ void extract_r_kernel(const double* A, int m, int n, double* R) {
	const int k = min(m, n);
	for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < (size_t)k * n; e += (size_t)gridDim.x * blockDim.x) {
		const int row = (int)(e / n), col = (int)(e % n);
		R[e] = (col >= row) ? A[(size_t)row * n + col] : 0.0;
	}
}

// One-sided Jacobi (Hestenes) SVD with round-robin pair ordering. W (column-major m×np) starts as A padded
// to an even column count np, V (column-major np×np) as the identity; every round rotates np/2 disjoint
// column pairs, one thread group per pair. Sweeps stop when no pair needed a rotation.
__global__ __launch_bounds__(QT) void jacobi_kernel(const double* A, int m, int n, int np, double* W, double* V, double* sv, int* ord,
                                                    double* U, double* S, double* Vt, int tout) {
	const int tid = threadIdx.x;
	__shared__ int rotated;
	for (size_t e = tid; e < (size_t)m * np; e += QT) {
		const int c = (int)(e / m), r = (int)(e % m);
		W[e] = c < n ? A[(size_t)r * n + c] : 0.0;
	}
	for (size_t e = tid; e < (size_t)np * np; e += QT) V[e] = ((int)(e / np) == (int)(e % np)) ? 1.0 : 0.0;
	__shared__ double red[QT / 64];
	double f2 = 0;
	for (size_t e = tid; e < (size_t)m * n; e += QT) f2 = fma(A[e], A[e], f2);
	for (int o = 32; o > 0; o >>= 1) f2 += __shfl_xor(f2, o);
	if ((tid & 63) == 0) red[tid >> 6] = f2;
	__syncthreads();
	f2 = 0;
	for (int w = 0; w < QT / 64; ++w) f2 += red[w];
	// a column whose squared norm is below ε²·‖A‖²_F/n (norm below ε·s_0) is rounding noise: it takes part in no
	// rotation (rotating noise against noise never meets the relative orthogonality test)
	const double negligible = DBL_EPSILON * DBL_EPSILON * f2 / n;
	const int npairs = np / 2;
	int G = 64;
	while (G > 1 && G * npairs > QT) G >>= 1;
	const int ngroups = QT / G;
	const int grp = tid / G, gl = tid % G;
	for (int sweep = 0; sweep < 60; ++sweep) {
		if (tid == 0) rotated = 0;
		__syncthreads();
		for (int t = 0; t < np - 1; ++t) {
			for (int pi = grp; pi < npairs; pi += ngroups) {
				const int posA = pi, posB = np - 1 - pi;
				const int pa = posA == 0 ? 0 : 1 + (posA - 1 + t) % (np - 1);
				const int pb = posB == 0 ? 0 : 1 + (posB - 1 + t) % (np - 1);
				const int p = min(pa, pb), q = max(pa, pb);
				double* wp = W + (size_t)p * m;
				double* wq = W + (size_t)q * m;
				double al = 0, be = 0, ga = 0;
				for (int r = gl; r < m; r += G) { const double x = wp[r], y = wq[r]; al += x * x; be += y * y; ga += x * y; }
				for (int o = G / 2; o > 0; o >>= 1) {
					al += __shfl_xor(al, o, G);
					be += __shfl_xor(be, o, G);
					ga += __shfl_xor(ga, o, G);
				}
				if (fmin(al, be) > negligible && ga != 0.0 && fabs(ga) > DBL_EPSILON * sqrt(al * be)) {
					if (gl == 0) rotated = 1;
					const double zeta = (be - al) / (2.0 * ga);
					const double tt = copysign(1.0, zeta) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
					const double c = 1.0 / sqrt(1.0 + tt * tt), s = c * tt;
					for (int r = gl; r < m; r += G) { const double x = wp[r], y = wq[r]; wp[r] = c * x - s * y; wq[r] = s * x + c * y; }
					double* vp = V + (size_t)p * np;
					double* vq = V + (size_t)q * np;
					for (int r = gl; r < np; r += G) { const double x = vp[r], y = vq[r]; vp[r] = c * x - s * y; vq[r] = s * x + c * y; }
				}
			}
			__syncthreads();
		}
		if (!rotated) break;
		__syncthreads();
	}
	const bool unconverged = rotated != 0;  // the 60th sweep still rotated: reported as NaN singular values
	for (int j = tid; j < np; j += QT) {
		double s = 0;
		for (int r = 0; r < m; ++r) { const double x = W[(size_t)j * m + r]; s += x * x; }
		sv[j] = sqrt(s);
	}
	__syncthreads();
	for (int j = tid; j < np; j += QT) {
		int rk = 0;
		for (int i = 0; i < np; ++i) rk += (sv[i] > sv[j]) || (sv[i] == sv[j] && i < j);
		ord[j] = rk;
	}
	__syncthreads();
	for (size_t e = tid; e < (size_t)np * m; e += QT) {
		const int j = (int)(e / m), r = (int)(e % m);
		const int rk = ord[j];
		if (rk < n) U[tout ? (size_t)rk * m + r : (size_t)r * n + rk] = sv[j] > 0 ? W[e] / sv[j] : 0.0;
	}
	for (size_t e = tid; e < (size_t)np * np; e += QT) {
		const int j = (int)(e / np), c = (int)(e % np);
		const int rk = ord[j];
		if (rk < n && c < n) Vt[tout ? (size_t)c * n + rk : (size_t)rk * n + c] = V[e];
	}
	for (int j = tid; j < np; j += QT) if (ord[j] < n) S[ord[j]] = unconverged ? __builtin_nan("") : sv[j];
}

// DPP move of a 64-bit value inside 16-lane rows (butterfly partners only: every lane of a row ends with
// the same bits after a full butterfly).
template <int CTRL>
__device__ inline double dpp64(double v) {
	const long long b = __double_as_longlong(v);
	const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, false);
	const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
	return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
// Sum over a 16-lane row through the partners l^1, l^2, l^7, l^15: identical in all 16 lanes.
__device__ inline double row16_sum(double v) {
	v += dpp64<0xB1>(v);   // quad_perm [1,0,3,2]
	v += dpp64<0x4E>(v);   // quad_perm [2,3,0,1]
	v += dpp64<0x141>(v);  // row_half_mirror
	v += dpp64<0x140>(v);  // row_mirror
	return v;
}

constexpr int JG = 16;            // lanes per column pair
constexpr int JT = 512;           // threads: 32 pairs per round
constexpr int JMAXM = 128, JMAXN = 64;

// One-sided Jacobi for m <= 128, n <= 64 with W (m×np, column-major) and V (np×np) in LDS. Same round-robin
// ordering, rotation formula and stopping rule as jacobi_kernel; the pair sums are DPP butterflies.
// tout: U and Vt are written transposed (U as n×m, Vt as n×n holding V).
__device__ __forceinline__ void jacobi_lds_body(const double* A, int m, int n, int np, double* U, double* S, double* Vt, int tout) {
	extern __shared__ double sm[];
	double* W = sm;                       // m × np, column stride m
	double* V = sm + (size_t)m * np;      // np × np
	__shared__ double sv[JMAXN];
	__shared__ int ord[JMAXN];
	__shared__ int rotated;
	const int tid = threadIdx.x;
	for (int e = tid; e < m * np; e += JT) {
		const int c = e / m, r = e % m;
		W[e] = c < n ? A[(size_t)r * n + c] : 0.0;
	}
	for (int e = tid; e < np * np; e += JT) V[e] = (e / np == e % np) ? 1.0 : 0.0;
	if (tid == 0) rotated = 0;
	__shared__ double red[JT / 64];
	double f2 = 0;
	for (int e = tid; e < m * n; e += JT) f2 = fma(A[e], A[e], f2);
	for (int o = 32; o > 0; o >>= 1) f2 += __shfl_xor(f2, o);
	if ((tid & 63) == 0) red[tid >> 6] = f2;
	__syncthreads();
	f2 = 0;
	for (int w = 0; w < JT / 64; ++w) f2 += red[w];
	const double negligible = DBL_EPSILON * DBL_EPSILON * f2 / n;  // as in jacobi_kernel
	const int npairs = np / 2;
	const int grp = tid / JG, gl = tid % JG;
	bool unconverged = true;  // until a sweep without rotations; reported as NaN singular values (xerus: dgesdd info > 0 throws)
	for (int sweep = 0; sweep < 60; ++sweep) {
		for (int t = 0; t < np - 1; ++t) {
			if (grp < npairs) {
				const int posA = grp, posB = np - 1 - grp;
				const int pa = posA == 0 ? 0 : 1 + (posA - 1 + t) % (np - 1);
				const int pb = posB == 0 ? 0 : 1 + (posB - 1 + t) % (np - 1);
				const int p = min(pa, pb), q = max(pa, pb);
				double* wp = W + (size_t)p * m;
				double* wq = W + (size_t)q * m;
				double al = 0, be = 0, ga = 0;
				for (int r = gl; r < m; r += JG) { const double x = wp[r], y = wq[r]; al = fma(x, x, al); be = fma(y, y, be); ga = fma(x, y, ga); }
				al = row16_sum(al);
				be = row16_sum(be);
				ga = row16_sum(ga);
				if (fmin(al, be) > negligible && ga != 0.0 && fabs(ga) > DBL_EPSILON * sqrt(al * be)) {
					if (gl == 0) rotated = 1;
					const double zeta = (be - al) / (2.0 * ga);
					const double tt = copysign(1.0, zeta) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
					const double c = 1.0 / sqrt(1.0 + tt * tt), s = c * tt;
					for (int r = gl; r < m; r += JG) { const double x = wp[r], y = wq[r]; wp[r] = c * x - s * y; wq[r] = s * x + c * y; }
					double* vp = V + (size_t)p * np;
					double* vq = V + (size_t)q * np;
					for (int r = gl; r < np; r += JG) { const double x = vp[r], y = vq[r]; vp[r] = c * x - s * y; vq[r] = s * x + c * y; }
				}
			}
			__syncthreads();
		}
		const int any = rotated;
		__syncthreads();
		if (!any) { unconverged = false; break; }
		if (tid == 0) rotated = 0;
		__syncthreads();
	}
	for (int j = tid; j < np; j += JT) {
		double s = 0;
		for (int r = 0; r < m; ++r) { const double x = W[(size_t)j * m + r]; s += x * x; }
		sv[j] = sqrt(s);
	}
	__syncthreads();
	for (int j = tid; j < np; j += JT) {
		int rk = 0;
		for (int i = 0; i < np; ++i) rk += (sv[i] > sv[j]) || (sv[i] == sv[j] && i < j);
		ord[j] = rk;
	}
	__syncthreads();
	for (int e = tid; e < np * m; e += JT) {
		const int j = e / m, r = e % m;
		const int rk = ord[j];
		if (rk < n) U[tout ? (size_t)rk * m + r : (size_t)r * n + rk] = sv[j] > 0 ? W[e] / sv[j] : 0.0;
	}
	for (int e = tid; e < np * np; e += JT) {
		const int j = e / np, c = e % np;
		const int rk = ord[j];
		if (rk < n && c < n) Vt[tout ? (size_t)c * n + rk : (size_t)rk * n + c] = V[e];
	}
	for (int j = tid; j < np; j += JT) if (ord[j] < n) S[ord[j]] = unconverged ? __builtin_nan("") : sv[j];
}
__global__ __launch_bounds__(JT) void jacobi_lds_kernel(const double* A, int m, int n, int np, double* U, double* S, double* Vt, int tout) {
	jacobi_lds_body(A, m, n, np, U, S, Vt, tout);
}
struct JacobiTask { const double* A; int m, n, np; double* U; double* S; double* Vt; int tout; };
// one matrix per workgroup
__global__ __launch_bounds__(JT) void jacobi_lds_group_kernel(const JacobiTask* tasks) {
	const JacobiTask t = tasks[blockIdx.x];
	jacobi_lds_body(t.A, t.m, t.n, t.np, t.U, t.S, t.Vt, t.tout);
}

}  // namespace

void geqp2(double* A, size_t m, size_t n, double* tau, int* jpvt, int pivot, double* ws, int* rank, hipStream_t s) {
	XTT_REQUIRE(m > 0 && n > 0, "factorisation of an empty matrix");
	XTT_REQUIRE(n <= 4096 && m <= (size_t)INT32_MAX, "geqp2: matrix too wide");
	(void)ws;
	const size_t lds = (3 * n + 32 * 33) * sizeof(double);
	if (lds > 64 * 1024) XTT_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(geqp2_kernel), hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
	hipLaunchKernelGGL(geqp2_kernel, dim3(1), dim3(QT), lds, s, A, (int)m, (int)n, tau, jpvt, pivot, rank);
	XTT_LAUNCH_CHECK();
}

void org2r(const double* A, size_t m, size_t n, const double* tau, double* Q, size_t k, hipStream_t s) {
	XTT_REQUIRE(k <= std::min(m, n), "org2r: too many columns");
	if (k == 0) return;
	const size_t lds = (k + 32 * 33) * sizeof(double);
	hipLaunchKernelGGL(org2r_kernel, dim3(1), dim3(QT), lds, s, A, (int)m, (int)n, tau, Q, (int)k);
	XTT_LAUNCH_CHECK();
}

void extract_c(const double* A, size_t m, size_t n, const int* jpvt, const int* rank, double* C, hipStream_t s) {
	const size_t tot = std::min(m, n) * n;
	unsigned g = (unsigned)std::min<size_t>(1024, (tot + 255) / 256);
	hipLaunchKernelGGL(extract_c_kernel, dim3(g ? g : 1), dim3(256), 0, s, A, (int)m, (int)n, jpvt, rank, C);
	XTT_LAUNCH_CHECK();
}

void extract_r(const double* A, size_t m, size_t n, double* R, hipStream_t s) {
	const size_t tot = std::min(m, n) * n;
	unsigned g = (unsigned)std::min<size_t>(1024, (tot + 255) / 256);
	hipLaunchKernelGGL(extract_r_kernel, dim3(g ? g : 1), dim3(256), 0, s, A, (int)m, (int)n, R);
	XTT_LAUNCH_CHECK();
}

void jacobi_svd(const double* A, size_t m, size_t n, double* U, double* S, double* Vt, double* ws, hipStream_t s, bool transposedOut) {
	XTT_REQUIRE(m >= n && n > 0, "jacobi_svd expects m >= n > 0");
	const size_t np = n + (n & 1);
	if (m <= (size_t)JMAXM && np <= (size_t)JMAXN) {
		static std::once_flag once;
		const size_t lds = ((size_t)JMAXM * JMAXN + (size_t)JMAXN * JMAXN) * sizeof(double);
		std::call_once(once, [&] {
			XTT_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(jacobi_lds_kernel), hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
		});
		// one sweep's rotations, 6m + 6n flops per pair and n(n-1)/2 pairs (the sweep count is data dependent)
		launch_timed("jacobi_lds_kernel", Work(3.0 * (m + np) * np * (np - 1), 8.0 * (2.0 * m * n + n + (double)n * n)), jacobi_lds_kernel, dim3(1), dim3(JT),
		             (m * np + np * np) * sizeof(double), s, A, (int)m, (int)n, (int)np, U, S, Vt, (int)transposedOut);
		return;
	}
	double* W = ws;
	double* V = W + m * np;
	double* sv = V + np * np;
	int* ord = reinterpret_cast<int*>(sv + np);
	hipLaunchKernelGGL(jacobi_kernel, dim3(1), dim3(QT), 0, s, A, (int)m, (int)n, (int)np, W, V, sv, ord, U, S, Vt, (int)transposedOut);
	XTT_LAUNCH_CHECK();
}


void jacobi_svd_group(const std::vector<SVDJob>& jobs, hipStream_t s) {
	std::vector<JacobiTask> tasks;
	size_t lds = 0;
	for (const SVDJob& j : jobs) {
		XTT_REQUIRE(j.m >= j.n && j.n > 0, "jacobi_svd expects m >= n > 0");
		const int np = j.n + (j.n & 1);
		if (j.m <= JMAXM && np <= JMAXN) {
			tasks.push_back(JacobiTask{j.A, j.m, j.n, np, j.U, j.S, j.Vt, j.transposedOut ? 1 : 0});
			lds = std::max(lds, ((size_t)j.m * np + (size_t)np * np) * sizeof(double));
		} else {
			jacobi_svd(j.A, (size_t)j.m, (size_t)j.n, j.U, j.S, j.Vt, j.ws, s, j.transposedOut);
		}
	}
	if (tasks.empty()) return;
	static std::once_flag once;
	std::call_once(once, [&] {
		const size_t full = ((size_t)JMAXM * JMAXN + (size_t)JMAXN * JMAXN) * sizeof(double);
		XTT_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(jacobi_lds_group_kernel), hipFuncAttributeMaxDynamicSharedMemorySize, (int)full));
	});
	const JacobiTask* dt = static_cast<const JacobiTask*>(table_upload(tasks.data(), tasks.size() * sizeof(JacobiTask), s));
	launch_timed("jacobi_lds_group_kernel", Work(), jacobi_lds_group_kernel, dim3((unsigned)tasks.size()), dim3(JT), lds, s, dt);
}

}  // namespace xtt
