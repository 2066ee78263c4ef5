// Dense linear solve with xerus's dispatch (blasWrapper::solve, blasLapackWrapper.cpp:542-651, its tests :501-538):
//   m ≠ n                                  → minimum-norm least squares (dgelsd, rcond = xerus::EPSILON);
//   A not symmetric (|a_ij − a_ji| >= 4·max_entry·ε, max over the signed entries, blasLapackWrapper.cpp:501-515)
//                                          → LU with partial pivoting (dgesv);
//   symmetric, diagonal all > ε (a_00 > 0) or all < −ε (a_00 <= 0) → Cholesky (dpotrf2 'U' + dpotrs),
//                                            falling back to the next case when the factorisation fails;
//   otherwise                              → xerus calls Bunch–Kaufman LDLᵀ (dsysv); this build solves the same
//                                            system by LU with partial pivoting.
// The square solvers run in one workgroup on a copy of A in global memory: right-looking, one column step
// between barriers, the right-hand sides carried as extra columns. They are meant for the local systems of
// dense ALS (a few hundred unknowns); their cost grows as n³ on one CU.
#include <cfloat>
#include <string>
#include <vector>

#include "kernels.hpp"
#include "tt.hpp"
#include "xtt_common.hpp"

namespace xtt {
namespace {

constexpr int ST = 1024;  // threads of the solver workgroups

__device__ inline double block_max(double v, double* red) {
	for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
	__syncthreads();
	if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
	__syncthreads();
	double m = red[0];
	for (int w = 1; w < ST / 64; ++w) m = fmax(m, red[w]);
	return m;
}

// out[0] = symmetric by xerus's test, out[1] = diagonal all positive / all negative by xerus's test.
__global__ __launch_bounds__(ST) void solve_classify_kernel(const double* A, long long n, int* out) {
	__shared__ double red[ST / 64];
	__shared__ int flags[2];
	double mx = -DBL_MAX;
	for (long long e = threadIdx.x; e < n * n; e += ST) mx = fmax(mx, A[e]);
	mx = block_max(mx, red);
	mx = fmax(mx, 0.0);  // xerus starts its maximum at 0
	if (threadIdx.x == 0) { flags[0] = 1; flags[1] = 1; }
	__syncthreads();
	const double tol = 4 * mx * DBL_EPSILON;
	bool asym = false;
	const int ty = threadIdx.x >> 5, tx = threadIdx.x & 31;
	for (long long i = ty; i < n; i += 32)
		for (long long j = i + 1 + tx; j < n; j += 32)
			if (fabs(A[i * n + j] - A[j * n + i]) >= tol) asym = true;
	if (asym) flags[0] = 0;
	const bool positive = A[0] > 0;
	bool bad = false;
	for (long long i = 1 + threadIdx.x; i < n; i += ST) {
		const double d = A[i * (n + 1)];
		if (positive ? (d < DBL_EPSILON) : (d > -DBL_EPSILON)) bad = true;
	}
	if (bad) flags[1] = 0;
	__syncthreads();
	if (threadIdx.x == 0) { out[0] = flags[0]; out[1] = flags[1]; }
}

// Back substitution U·X = Y in place (U upper triangular in a, row-major n×lda; X in columns [n, n + nrhs)):
// column-oriented, x_i = y_i / u_ii, then y_r −= u_ri·x_i for r < i.
__device__ void back_substitute(double* a, long long n, long long lda, long long nrhs) {
	const int ty = threadIdx.x >> 5, tx = threadIdx.x & 31;
	for (long long i = n - 1; i >= 0; --i) {
		const double uii = a[i * lda + i];
		for (long long c = threadIdx.x; c < nrhs; c += ST) a[i * lda + n + c] /= uii;
		__syncthreads();
		for (long long r = ty; r < i; r += 32)
			for (long long c = tx; c < nrhs; c += 32) a[r * lda + n + c] -= a[r * lda + i] * a[i * lda + n + c];
		__syncthreads();
	}
}

// [A | B] (n × (n + nrhs), ld lda) → LU with partial pivoting (dgetrf's pivot rule: the largest |a_ik|, the first
// index on ties), the eliminations applied to B, then back substitution: columns [n, n + nrhs) hold X.
// *info = 0, or k + 1 when the pivot of column k is exactly zero (dgesv's singular case).
__global__ __launch_bounds__(ST) void lu_solve_kernel(double* a, long long n, long long lda, long long nrhs, int* info) {
	__shared__ double red[ST / 64];
	__shared__ long long sPiv;
	const long long w = n + nrhs;
	const int ty = threadIdx.x >> 5, tx = threadIdx.x & 31;
	for (long long k = 0; k < n; ++k) {
		double best = -1.0;
		for (long long i = k + threadIdx.x; i < n; i += ST) best = fmax(best, fabs(a[i * lda + k]));
		best = block_max(best, red);
		if (threadIdx.x == 0) sPiv = n;
		__syncthreads();
		for (long long i = k + threadIdx.x; i < n; i += ST)
			if (fabs(a[i * lda + k]) == best) atomicMin((unsigned long long*)&sPiv, (unsigned long long)i);
		__syncthreads();
		const long long p = sPiv;
		if (best == 0.0 || p >= n) {  // exactly singular (or NaN column)
			if (threadIdx.x == 0) *info = (int)(k + 1);
			return;
		}
		if (p != k)
			for (long long j = threadIdx.x; j < w; j += ST) {
				const double t = a[k * lda + j];
				a[k * lda + j] = a[p * lda + j];
				a[p * lda + j] = t;
			}
		__syncthreads();
		const double pivot = a[k * lda + k];
		for (long long i = k + 1 + threadIdx.x; i < n; i += ST) a[i * lda + k] /= pivot;  // the multipliers l_ik
		__syncthreads();
		for (long long i = k + 1 + ty; i < n; i += 32)
			for (long long j = k + 1 + tx; j < w; j += 32) a[i * lda + j] -= a[i * lda + k] * a[k * lda + j];
		__syncthreads();
	}
	back_substitute(a, n, lda, nrhs);
	if (threadIdx.x == 0) *info = 0;
}

// [A | B] → A = UᵀU (upper Cholesky, dpotrf2 'U'), then Uᵀ·Y = B and U·X = Y; columns [n, n + nrhs) hold X.
// *info = 0, or j + 1 when the j-th pivot is not positive (A left partly overwritten).
__global__ __launch_bounds__(ST) void chol_solve_kernel(double* a, long long n, long long lda, long long nrhs, int* info) {
	const long long w = n + nrhs;
	const int ty = threadIdx.x >> 5, tx = threadIdx.x & 31;
	for (long long j = 0; j < n; ++j) {
		const double d = a[j * lda + j];
		if (!(d > 0.0)) {
			if (threadIdx.x == 0) *info = (int)(j + 1);
			return;
		}
		const double u = sqrt(d);
		__syncthreads();
		// row j of U (and of Y: the forward solve Uᵀ·Y = B runs along with the factorisation)
		for (long long c = j + threadIdx.x; c < w; c += ST) a[j * lda + c] = c == j ? u : a[j * lda + c] / u;
		__syncthreads();
		// trailing update of the upper triangle and of B: a_rc −= u_jr·u_jc for j < r <= c (c >= n: B's columns)
		for (long long r = j + 1 + ty; r < n; r += 32)
			for (long long c = r + tx; c < w; c += 32) a[r * lda + c] -= a[j * lda + r] * a[j * lda + c];
		__syncthreads();
	}
	back_substitute(a, n, lda, nrhs);
	if (threadIdx.x == 0) *info = 0;
}

// dst (rows × (n + nrhs)) = [A | B]
__global__ void augment_kernel(const double* A, const double* B, long long n, long long nrhs, double* dst) {
	const long long w = n + nrhs;
	for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n * w; e += (long long)gridDim.x * blockDim.x) {
		const long long i = e / w, j = e - i * w;
		dst[e] = j < n ? A[i * n + j] : B[i * nrhs + (j - n)];
	}
}
__global__ void extract_kernel(const double* src, long long n, long long nrhs, double* X) {
	const long long w = n + nrhs;
	for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n * nrhs; e += (long long)gridDim.x * blockDim.x) {
		const long long i = e / nrhs, c = e - i * nrhs;
		X[e] = src[i * w + n + c];
	}
}

// T[j][c] /= s[j] for the first r rows
__global__ void inv_scale_rows_kernel(double* T, const double* s, long long r, long long cols) {
	for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < r * cols; e += (long long)gridDim.x * blockDim.x)
		T[e] /= s[e / cols];
}

unsigned grid_for(long long e) { return (unsigned)std::min<long long>(1024, std::max<long long>(1, (e + 255) / 256)); }

// Runs one square solver on [A | B]; returns its info.
int square_solve(bool chol, const double* A, const double* B, size_t n, size_t nrhs, double* X, hipStream_t s) {
	const long long w = (long long)(n + nrhs);
	DBuf<double> aug(n * (size_t)w);
	DBuf<int> info(1);
	hipLaunchKernelGGL(augment_kernel, dim3(grid_for((long long)n * w)), dim3(256), 0, s, A, B, (long long)n, (long long)nrhs, aug.get());
	XTT_LAUNCH_CHECK();
	const Work work(chol ? (double)n * n * n / 3.0 + 2.0 * n * n * nrhs : 2.0 * n * n * n / 3.0 + 2.0 * n * n * nrhs,
	                8.0 * (double)n * w * 2.0);
	if (chol) launch_timed("chol_solve_kernel", work, chol_solve_kernel, dim3(1), dim3(ST), 0, s, aug.get(), (long long)n, w, (long long)nrhs, info.get());
	else launch_timed("lu_solve_kernel", work, lu_solve_kernel, dim3(1), dim3(ST), 0, s, aug.get(), (long long)n, w, (long long)nrhs, info.get());
	int h = 0;
	read_back(&h, info.get(), sizeof(int), s);
	if (h == 0) {
		hipLaunchKernelGGL(extract_kernel, dim3(grid_for((long long)(n * nrhs))), dim3(256), 0, s, (const double*)aug.get(), (long long)n,
		                   (long long)nrhs, X);
		XTT_LAUNCH_CHECK();
	}
	return h;
}

}  // namespace

// blasWrapper::solve_least_squares (blasLapackWrapper.cpp:654-705, dgelsd): X = V·diag(1/s_j for s_j > rcond·s_0)·Uᵀ·B
// with rcond = xerus::EPSILON (include/xerus/basic.h:50), the minimum-norm least squares solution. Returns 3.
int least_squares_solve(const double* A, size_t m, size_t n, const double* B, size_t nrhs, double* X, hipStream_t s) {
	XTT_REQUIRE(m > 0 && n > 0 && nrhs > 0, "solve: empty system");
	XTT_REQUIRE(m <= (size_t)INT32_MAX && n <= (size_t)INT32_MAX && nrhs <= (size_t)INT32_MAX, "Dimension to large for BLAS/Lapack");
	{
		const double rcond = 8 * DBL_EPSILON;
		DTensor Ad({m, n});
		XTT_HIP(hipMemcpyAsync(Ad.data(), A, m * n * 8, hipMemcpyDeviceToDevice, s));
		DTensor U, Vt;
		std::vector<double> S;
		DBuf<double> sd;
		calculate_svd(U, S, Vt, Ad, 1, 0, 0.0, &sd);
		size_t r = 0;
		while (r < S.size() && S[r] > rcond * S[0]) ++r;
		const size_t k = S.size();
		DBuf<double> T(std::max<size_t>(k * nrhs, 1));
		fill_f64(X, 0.0, n * nrhs, s);
		if (r == 0) return 3;
		// T = U[:, :r]ᵀ·B (r × nrhs), scaled row-wise by 1/s_j, X = Vt[:r, :]ᵀ·T
		GemmDesc g1 = gemm_desc(1, 0, r, nrhs, m);
		g1.lda = (long long)k;
		gemm_f64(g1, U.data(), B, T.get(), s);
		hipLaunchKernelGGL(inv_scale_rows_kernel, dim3(grid_for((long long)(r * nrhs))), dim3(256), 0, s, T.get(), (const double*)sd.get(),
		                   (long long)r, (long long)nrhs);
		XTT_LAUNCH_CHECK();
		gemm_f64(gemm_desc(1, 0, n, nrhs, r), Vt.data(), T.get(), X, s);
		return 3;
	}
}

int dense_solve(const double* A, size_t m, size_t n, const double* B, size_t nrhs, double* X, hipStream_t s) {
	XTT_REQUIRE(m > 0 && n > 0 && nrhs > 0, "solve: empty system");
	XTT_REQUIRE(m <= (size_t)INT32_MAX && n <= (size_t)INT32_MAX && nrhs <= (size_t)INT32_MAX, "Dimension to large for BLAS/Lapack");
	if (m != n) return least_squares_solve(A, m, n, B, nrhs, X, s);
	DBuf<int> cls(2);
	hipLaunchKernelGGL(solve_classify_kernel, dim3(1), dim3(ST), 0, s, A, (long long)n, cls.get());
	XTT_LAUNCH_CHECK();
	int h[2] = {0, 0};
	read_back(h, cls.get(), sizeof(h), s);
	if (!h[0]) {
		XTT_REQUIRE(square_solve(false, A, B, n, nrhs, X, s) == 0, "Unable to solve Ax = b (PLU solver).");
		return 0;
	}
	if (h[1] && square_solve(true, A, B, n, nrhs, X, s) == 0) return 1;
	XTT_REQUIRE(square_solve(false, A, B, n, nrhs, X, s) == 0, "Unable to solve Ax = b (symmetric indefinite solver).");
	return 2;
}

}  // namespace xtt
