// TEST INFRASTRUCTURE ONLY — the oracle: dense FP64 kernels restating xerus's blasWrapper
// (src/xerus/blasLapackWrapper.cpp) on top of textbook Householder / Jacobi / Cholesky routines.
#include "xtt_oracle.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <stdexcept>

namespace orc {

std::mt19937_64& rng() { static std::mt19937_64 e(0xBAADF00Dull); return e; }
std::normal_distribution<double>& normal() { static std::normal_distribution<double> d; return d; }
void seed(uint64_t s) { rng().seed(s); normal().reset(); }

// C = alpha·op(A)·op(B). Row-major; op(A) is m×k, op(B) is k×n (blasLapackWrapper.cpp:149-195, beta = 0).
void gemm(bool ta, bool tb, size_t m, size_t n, size_t k, double alpha, const double* A, const double* B, double* C) {
	const size_t lda = ta ? m : k, ldb = tb ? k : n;
	#pragma omp parallel for schedule(static) if (m * n * k > (1u << 18))
	for (size_t i = 0; i < m; ++i) {
		double* c = C + i * n;
		std::fill(c, c + n, 0.0);
		for (size_t l = 0; l < k; ++l) {
			const double a = ta ? A[l * lda + i] : A[i * lda + l];
			if (a == 0.0) continue;
			if (!tb) {
				const double* b = B + l * ldb;
				for (size_t j = 0; j < n; ++j) c[j] += a * b[j];
			} else {
				for (size_t j = 0; j < n; ++j) c[j] += a * B[j * ldb + l];
			}
		}
		if (alpha != 1.0) for (size_t j = 0; j < n; ++j) c[j] *= alpha;
	}
}

namespace {

double nrm2(const double* x, size_t n, size_t inc = 1) {
	// scaled sum of squares (LAPACK dnrm2 semantics: no spurious overflow)
	double scale = 0.0, ssq = 1.0;
	for (size_t i = 0; i < n; ++i) {
		const double a = std::abs(x[i * inc]);
		if (a == 0.0) continue;
		if (scale < a) { ssq = 1.0 + ssq * (scale / a) * (scale / a); scale = a; }
		else { ssq += (a / scale) * (a / scale); }
	}
	return scale * std::sqrt(ssq);
}

// LAPACK dlarfg: H·[alpha; x] = [beta; 0] with H = I - tau·v·vᵀ, v = [1; x/(alpha-beta)], beta = -sign(alpha)·‖[alpha;x]‖.
void larfg(size_t n, double& alpha, double* x, double& tau) {
	if (n <= 1) { tau = 0.0; return; }
	const double xnorm = nrm2(x, n - 1);
	if (xnorm == 0.0) { tau = 0.0; return; }
	const double beta = -std::copysign(std::hypot(alpha, xnorm), alpha);
	tau = (beta - alpha) / beta;
	const double s = 1.0 / (alpha - beta);
	for (size_t i = 0; i + 1 < n; ++i) x[i] *= s;
	alpha = beta;
}

// Apply H = I - tau·v·vᵀ (v[0] = 1 implied by caller having set it) from the left to columns [c0,c1) of the
// column-major block rows [r0, m).
void apply_left(double* a, size_t m, size_t r0, size_t c0, size_t c1, const double* v, double tau) {
	if (tau == 0.0) return;
	for (size_t j = c0; j < c1; ++j) {
		double* col = a + j * m;
		double w = 0.0;
		for (size_t r = r0; r < m; ++r) w += v[r - r0] * col[r];
		w *= tau;
		for (size_t r = r0; r < m; ++r) col[r] -= w * v[r - r0];
	}
}

// Householder QR with optional column pivoting on a column-major m×n matrix (LAPACK dgeqr2 / dlaqp2 as
// called through dgeqp3): pivot = column of largest partial norm (first on ties), partial norms downdated
// with LAPACK's tol3z recomputation rule.
void geqp2(size_t m, size_t n, double* a, std::vector<double>& tau, std::vector<size_t>& jpvt, bool pivot) {
	const size_t k = std::min(m, n);
	tau.assign(k, 0.0);
	jpvt.resize(n);
	std::vector<double> vn1(n), vn2(n), v(m);
	for (size_t j = 0; j < n; ++j) { jpvt[j] = j; vn1[j] = vn2[j] = nrm2(a + j * m, m); }
	const double tol3z = std::sqrt(DBL_EPSILON * 0.5);
	for (size_t i = 0; i < k; ++i) {
		if (pivot) {
			size_t pvt = i;
			for (size_t j = i + 1; j < n; ++j) if (vn1[j] > vn1[pvt]) pvt = j;
			if (pvt != i) {
				for (size_t r = 0; r < m; ++r) std::swap(a[r + pvt * m], a[r + i * m]);
				std::swap(jpvt[pvt], jpvt[i]);
				vn1[pvt] = vn1[i];
				vn2[pvt] = vn2[i];
			}
		}
		double* col = a + i * m;
		larfg(m - i, col[i], col + i + 1, tau[i]);
		if (i + 1 < n) {
			v[0] = 1.0;
			for (size_t r = i + 1; r < m; ++r) v[r - i] = col[r];
			apply_left(a, m, i, i + 1, n, v.data(), tau[i]);
		}
		if (pivot) {
			for (size_t j = i + 1; j < n; ++j) {
				if (vn1[j] == 0.0) continue;
				double temp = std::abs(a[i + j * m]) / vn1[j];
				temp = std::max(0.0, 1.0 - temp * temp);
				const double ratio = vn1[j] / vn2[j];
				const double temp2 = temp * ratio * ratio;
				if (temp2 <= tol3z) {
					if (i + 1 < m) { vn1[j] = nrm2(a + i + 1 + j * m, m - i - 1); vn2[j] = vn1[j]; }
					else { vn1[j] = 0.0; vn2[j] = 0.0; }
				} else {
					vn1[j] *= std::sqrt(temp);
				}
			}
		}
	}
}

// LAPACK dorg2r: overwrite the first `ncol` columns of the reflector storage with Q = H_0…H_{k-1}[I;0].
void org2r(size_t m, size_t ncol, size_t k, double* a, const std::vector<double>& tau) {
	std::vector<double> v(m);
	for (size_t j = k; j < ncol; ++j) { for (size_t r = 0; r < m; ++r) a[r + j * m] = 0.0; a[j + j * m] = 1.0; }
	for (size_t ii = k; ii-- > 0;) {
		double* col = a + ii * m;
		if (ii + 1 < ncol) {
			v[0] = 1.0;
			for (size_t r = ii + 1; r < m; ++r) v[r - ii] = col[r];
			apply_left(a, m, ii, ii + 1, ncol, v.data(), tau[ii]);
		}
		for (size_t r = ii + 1; r < m; ++r) col[r] *= -tau[ii];
		col[ii] = 1.0 - tau[ii];
		for (size_t r = 0; r < ii; ++r) col[r] = 0.0;
	}
}

std::vector<double> to_colmajor(const double* A, size_t m, size_t n) {
	std::vector<double> a(m * n);
	for (size_t i = 0; i < m; ++i) for (size_t j = 0; j < n; ++j) a[i + j * m] = A[i * n + j];
	return a;
}

// xerus's rank rule after dgeqp3 (blasLapackWrapper.cpp:268-272): the first k with |R_kk| < 16·ε·R_00. R_00 is
// taken with its LAPACK sign (dlarfg gives R_00 = -sign(a_00)·‖a_0‖), so a negative R_00 never reduces the rank.
size_t qc_rank(const double* a, size_t m, size_t maxRank) {
	size_t rank;
	for (rank = 1; rank <= maxRank; ++rank) {
		if (rank == maxRank || std::abs(a[rank + rank * m]) < 16 * DBL_EPSILON * a[0]) break;
	}
	return rank;
}

} // namespace

void qr(const double* A, size_t m, size_t n, std::vector<double>& Q, std::vector<double>& R) {
	if (m == 0 || n == 0) throw std::invalid_argument("qr: empty matrix");
	const size_t k = std::min(m, n);
	std::vector<double> a = to_colmajor(A, m, n), tau;
	std::vector<size_t> jpvt;
	geqp2(m, n, a.data(), tau, jpvt, false);
	R.assign(k * n, 0.0);
	for (size_t r = 0; r < k; ++r) for (size_t c = r; c < n; ++c) R[r * n + c] = a[r + c * m];
	org2r(m, k, k, a.data(), tau);
	Q.resize(m * k);
	for (size_t i = 0; i < m; ++i) for (size_t j = 0; j < k; ++j) Q[i * k + j] = a[i + j * m];
}

// RQ through the reversal identity: with J the exchange matrix, QR(J·Aᵀ·J) = Q̃·R̃ gives A = (J·R̃ᵀ·J)(J·Q̃ᵀ·J),
// R upper triangular in its last min(m,n) rows and Q with orthonormal rows, as dgerqf+dorgrq return.
void rq(const double* A, size_t m, size_t n, std::vector<double>& R, std::vector<double>& Q) {
	const size_t k = std::min(m, n);
	std::vector<double> B(n * m);
	for (size_t i = 0; i < n; ++i) for (size_t j = 0; j < m; ++j) B[i * m + j] = A[(m - 1 - j) * n + (n - 1 - i)];
	std::vector<double> Qt, Rt;  // Qt n×k, Rt k×m
	qr(B.data(), n, m, Qt, Rt);
	R.resize(m * k);
	for (size_t i = 0; i < m; ++i) for (size_t j = 0; j < k; ++j) R[i * k + j] = Rt[(k - 1 - j) * m + (m - 1 - i)];
	Q.resize(k * n);
	for (size_t i = 0; i < k; ++i) for (size_t j = 0; j < n; ++j) Q[i * n + j] = Qt[(n - 1 - j) * k + (k - 1 - i)];
}

size_t qc(const double* A, size_t m, size_t n, std::vector<double>& Q, std::vector<double>& C) {
	if (m == 0 || n == 0) throw std::invalid_argument("qc: empty matrix");
	const size_t k = std::min(m, n);
	std::vector<double> a = to_colmajor(A, m, n), tau;
	std::vector<size_t> jpvt;
	geqp2(m, n, a.data(), tau, jpvt, true);
	const size_t rank = qc_rank(a.data(), m, k);
	C.assign(rank * n, 0.0);
	for (size_t col = 0; col < n; ++col)
		for (size_t row = 0; row < rank && row < col + 1; ++row) C[row * n + jpvt[col]] = a[row + col * m];
	org2r(m, k, k, a.data(), tau);
	Q.resize(m * rank);
	for (size_t i = 0; i < m; ++i) for (size_t j = 0; j < rank; ++j) Q[i * rank + j] = a[i + j * m];
	return rank;
}

size_t cq(const double* A, size_t m, size_t n, std::vector<double>& C, std::vector<double>& Q) {
	if (m == 0 || n == 0) throw std::invalid_argument("cq: empty matrix");
	// The column-major n×m view of A's row-major storage is Aᵀ (blasLapackWrapper.cpp:313-314).
	const size_t k = std::min(m, n);
	std::vector<double> a(A, A + m * n), tau;
	std::vector<size_t> jpvt;
	geqp2(n, m, a.data(), tau, jpvt, true);
	const size_t rank = qc_rank(a.data(), n, k);
	C.assign(m * rank, 0.0);
	for (size_t col = 0; col < m; ++col)
		for (size_t row = 0; row < std::min(rank, col + 1); ++row) C[jpvt[col] * rank + row] = a[row + col * n];
	org2r(n, k, k, a.data(), tau);
	Q.resize(rank * n);
	for (size_t r = 0; r < rank; ++r) for (size_t j = 0; j < n; ++j) Q[r * n + j] = a[j + r * n];
	return rank;
}

namespace {
// One-sided (Hestenes) Jacobi SVD of a column-major m×n matrix with m >= n: W ← W·V until the columns are
// mutually orthogonal to working precision. Returns column norms in s (unsorted), W and V overwritten.
void jacobi_svd_tall(size_t m, size_t n, std::vector<double>& W, std::vector<double>& V, std::vector<double>& s) {
	V.assign(n * n, 0.0);
	for (size_t i = 0; i < n; ++i) V[i + i * n] = 1.0;
	for (int sweep = 0; sweep < 80; ++sweep) {
		bool rotated = false;
		for (size_t p = 0; p + 1 < n; ++p) {
			for (size_t q = p + 1; q < n; ++q) {
				double* wp = W.data() + p * m; double* wq = W.data() + q * m;
				double alpha = 0, beta = 0, gamma = 0;
				for (size_t r = 0; r < m; ++r) { alpha += wp[r] * wp[r]; beta += wq[r] * wq[r]; gamma += wp[r] * wq[r]; }
				if (gamma == 0.0 || std::abs(gamma) <= DBL_EPSILON * std::sqrt(alpha * beta)) continue;
				rotated = true;
				const double zeta = (beta - alpha) / (2.0 * gamma);
				const double t = std::copysign(1.0, zeta) / (std::abs(zeta) + std::sqrt(1.0 + zeta * zeta));
				const double c = 1.0 / std::sqrt(1.0 + t * t), sn = c * t;
				for (size_t r = 0; r < m; ++r) { const double x = wp[r], y = wq[r]; wp[r] = c * x - sn * y; wq[r] = sn * x + c * y; }
				double* vp = V.data() + p * n; double* vq = V.data() + q * n;
				for (size_t r = 0; r < n; ++r) { const double x = vp[r], y = vq[r]; vp[r] = c * x - sn * y; vq[r] = sn * x + c * y; }
			}
		}
		if (!rotated) break;
	}
	s.resize(n);
	for (size_t j = 0; j < n; ++j) s[j] = nrm2(W.data() + j * m, m);
}

// Complete zero columns of a column-major m×k matrix with orthonormal vectors (dgesdd returns an orthonormal U
// even for zero singular values).
void complete_orthonormal(size_t m, size_t k, std::vector<double>& U, const std::vector<bool>& zero) {
	size_t e = 0;
	for (size_t j = 0; j < k; ++j) {
		if (!zero[j]) continue;
		for (; e < m; ++e) {
			std::vector<double> u(m, 0.0); u[e] = 1.0;
			for (int pass = 0; pass < 2; ++pass)
				for (size_t i = 0; i < k; ++i) {
					if (i == j || (zero[i] && i > j)) continue;
					const double* ui = U.data() + i * m;
					double d = 0; for (size_t r = 0; r < m; ++r) d += ui[r] * u[r];
					for (size_t r = 0; r < m; ++r) u[r] -= d * ui[r];
				}
			const double nu = nrm2(u.data(), m);
			if (nu > 0.5) { for (size_t r = 0; r < m; ++r) U[r + j * m] = u[r] / nu; ++e; break; }
		}
	}
}
} // namespace

void svd(const double* A, size_t m, size_t n, std::vector<double>& U, std::vector<double>& S, std::vector<double>& Vt) {
	const bool tall = m >= n;
	const size_t M = tall ? m : n, N = tall ? n : m, k = N;
	std::vector<double> W(M * N), V, s;
	for (size_t i = 0; i < m; ++i) for (size_t j = 0; j < n; ++j) {
		if (tall) W[i + j * M] = A[i * n + j]; else W[j + i * M] = A[i * n + j];
	}
	jacobi_svd_tall(M, N, W, V, s);
	std::vector<size_t> order(k);
	for (size_t i = 0; i < k; ++i) order[i] = i;
	std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return s[a] > s[b]; });
	std::vector<double> Us(M * k), Vs(N * k);
	std::vector<bool> zero(k, false);
	S.resize(k);
	for (size_t jj = 0; jj < k; ++jj) {
		const size_t j = order[jj];
		S[jj] = s[j];
		zero[jj] = !(s[j] > 0.0);
		for (size_t r = 0; r < M; ++r) Us[r + jj * M] = zero[jj] ? 0.0 : W[r + j * M] / s[j];
		for (size_t r = 0; r < N; ++r) Vs[r + jj * N] = V[r + j * N];
	}
	complete_orthonormal(M, k, Us, zero);
	// tall: A = Us·S·Vsᵀ.  wide: Aᵀ = Us·S·Vsᵀ → A = Vs·S·Usᵀ.
	const std::vector<double>& Left = tall ? Us : Vs;
	const std::vector<double>& Right = tall ? Vs : Us;
	const size_t LM = tall ? M : N, RM = tall ? N : M;
	U.resize(m * k); Vt.resize(k * n);
	for (size_t i = 0; i < m; ++i) for (size_t j = 0; j < k; ++j) U[i * k + j] = Left[i + j * LM];
	for (size_t i = 0; i < k; ++i) for (size_t j = 0; j < n; ++j) Vt[i * n + j] = Right[j + i * RM];
}

namespace {
bool is_symmetric(const double* A, size_t n) {  // blasLapackWrapper.cpp:501-515
	double mx = 0;
	for (size_t i = 0; i < n * n; ++i) mx = std::max(mx, A[i]);
	for (size_t i = 0; i < n; ++i)
		for (size_t j = i + 1; j < n; ++j)
			if (std::abs(A[i * n + j] - A[i + j * n]) >= 4 * mx * DBL_EPSILON) return false;
	return true;
}
bool pos_neg_definite_diagonal(const double* A, size_t n) {  // blasLapackWrapper.cpp:519-538
	const bool positive = A[0] > 0;
	for (size_t i = 1; i < n; ++i) {
		if (positive ? (A[i * (n + 1)] < DBL_EPSILON) : (A[i * (n + 1)] > -DBL_EPSILON)) return false;
	}
	return true;
}
bool cholesky_upper(std::vector<double>& a, size_t n) {  // A = UᵀU, U stored in the upper triangle
	for (size_t j = 0; j < n; ++j) {
		double d = a[j * n + j];
		for (size_t k = 0; k < j; ++k) d -= a[k * n + j] * a[k * n + j];
		if (!(d > 0.0)) return false;
		d = std::sqrt(d);
		a[j * n + j] = d;
		#pragma omp parallel for schedule(static) if (n > 512)
		for (size_t c = j + 1; c < n; ++c) {
			double s = a[j * n + c];
			for (size_t k = 0; k < j; ++k) s -= a[k * n + j] * a[k * n + c];
			a[j * n + c] = s / d;
		}
	}
	return true;
}
void lu_solve(std::vector<double>& a, size_t n, double* x) {  // partial pivoting, x holds b on entry
	for (size_t k = 0; k < n; ++k) {
		size_t p = k;
		for (size_t i = k + 1; i < n; ++i) if (std::abs(a[i * n + k]) > std::abs(a[p * n + k])) p = i;
		if (a[p * n + k] == 0.0) throw std::runtime_error("solve: singular matrix");
		if (p != k) { for (size_t j = 0; j < n; ++j) std::swap(a[p * n + j], a[k * n + j]); std::swap(x[p], x[k]); }
		for (size_t i = k + 1; i < n; ++i) {
			const double f = a[i * n + k] / a[k * n + k];
			a[i * n + k] = f;
			for (size_t j = k + 1; j < n; ++j) a[i * n + j] -= f * a[k * n + j];
			x[i] -= f * x[k];
		}
	}
	for (size_t i = n; i-- > 0;) {
		double s = x[i];
		for (size_t j = i + 1; j < n; ++j) s -= a[i * n + j] * x[j];
		x[i] = s / a[i * n + i];
	}
}
} // namespace

int solve(const double* A, size_t n, const double* b, double* x) {
	std::vector<double> a(A, A + n * n);
	std::copy(b, b + n, x);
	if (!is_symmetric(A, n)) { lu_solve(a, n, x); return 0; }
	if (pos_neg_definite_diagonal(A, n) && cholesky_upper(a, n)) {
		for (size_t i = 0; i < n; ++i) {  // Uᵀ y = b
			double s = x[i];
			for (size_t k = 0; k < i; ++k) s -= a[k * n + i] * x[k];
			x[i] = s / a[i * n + i];
		}
		for (size_t i = n; i-- > 0;) {    // U x = y
			double s = x[i];
			for (size_t k = i + 1; k < n; ++k) s -= a[i * n + k] * x[k];
			x[i] = s / a[i * n + i];
		}
		return 1;
	}
	// xerus falls back to Bunch–Kaufman LDLᵀ (dsysv); the oracle solves the same system by pivoted LU.
	a.assign(A, A + n * n);
	std::copy(b, b + n, x);
	lu_solve(a, n, x);
	return 2;
}

} // namespace orc
