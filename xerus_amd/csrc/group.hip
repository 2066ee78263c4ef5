// Grouped launches for the lockstep batch driver: one launch does the same step for many independent problems,
// each with its own shapes and buffers. The per-problem arguments travel in a task table that is copied to the
// device on the stream right before the launch; workgroup b finds its task by a binary search over the tasks'
// first-workgroup offsets (group_task). This file holds the table upload, the elementwise operations (mode
// permutations, reversed transposes, row / column scaling, identities, fills), the FP64 CG vector steps and the
// dot products; the grouped GEMM, Householder and Jacobi kernels sit beside their single-problem versions.
#include <algorithm>
#include <cstring>
#include <vector>

#include "kernels.hpp"
#include "xtt_common.hpp"

namespace xtt {

namespace {
// A ring of pinned host memory and device memory per thread. A table is written to the host ring, copied to the
// device ring on the stream and used by the launch that follows. When the ring is full the stream is synchronised
// first, so no table is overwritten while a copy or kernel still reads it.
struct TableRing {
	char* host = nullptr;
	char* dev = nullptr;
	size_t cap = 0, used = 0;
	hipStream_t last = nullptr;
	~TableRing() {
		// the device memory and pinned host memory are kept until the process ends (the runtime may already be
		// shutting down when thread-local storage is destroyed)
	}
};
thread_local TableRing t_ring;
constexpr size_t kRingBytes = 32u << 20;
}  // namespace

namespace {
// The pinned host slot of a table of `bytes` (the ring is created on first use; when full, the stream that used it
// last is synchronised and the ring starts over). `dev` receives the matching device slot.
char* ring_slot(size_t bytes, hipStream_t s, char** dev) {
	TableRing& R = t_ring;
	if (!R.host) {
		XTT_HIP(hipHostMalloc(reinterpret_cast<void**>(&R.host), kRingBytes, hipHostMallocDefault));
		XTT_HIP(hipMalloc(reinterpret_cast<void**>(&R.dev), kRingBytes));
		R.cap = kRingBytes;
	}
	XTT_REQUIRE(bytes <= R.cap, "grouped launch: task table larger than the upload ring");
	const size_t need = (bytes + 255) & ~size_t(255);
	if (R.used + need > R.cap || (R.last && R.last != s)) {
		if (R.last) XTT_HIP(hipStreamSynchronize(R.last));
		R.used = 0;
	}
	R.last = s;
	char* h = R.host + R.used;
	*dev = R.dev + R.used;
	R.used += need;
	return h;
}
}  // namespace

const void* table_upload(const void* src, size_t bytes, hipStream_t s) {
	char* dev = nullptr;
	char* h = ring_slot(bytes, s, &dev);
	std::memcpy(h, src, bytes);
	XTT_HIP(hipMemcpyAsync(dev, h, bytes, hipMemcpyHostToDevice, s));
	return dev;
}

void table_upload_to(void* dst, const void* src, size_t bytes, hipStream_t s) {
	char* dev = nullptr;
	char* h = ring_slot(bytes, s, &dev);
	std::memcpy(h, src, bytes);
	XTT_HIP(hipMemcpyAsync(dst, h, bytes, hipMemcpyHostToDevice, s));
}

void table_ring_release(hipStream_t s) {
	TableRing& R = t_ring;
	if (R.last == s) R.used = 0;  // the caller has synchronised s
}

namespace {

constexpr int GT = 256;        // threads of the elementwise kernels
constexpr int GE = 4 * GT;     // elements per workgroup
constexpr int CT = 512;        // threads of the CG kernels (one workgroup per problem)

__global__ __launch_bounds__(GT) void map_group_kernel(const MapTask* tasks, const int* start, int count) {
	const int t = group_task(start, count, (int)blockIdx.x);
	const MapTask m = tasks[t];
	const long long e0 = (long long)(blockIdx.x - start[t]) * GE;
	const long long e1 = min(m.n, e0 + GE);
	for (long long e = e0 + threadIdx.x; e < e1; e += GT) {
		switch (m.op) {
		case MAP_PERMUTE: {  // e indexes the input (row-major over dims)
			long long rem = e, o = 0;
			for (int i = m.deg - 1; i >= 0; --i) {
				const long long c = rem % m.dims[i];
				rem /= m.dims[i];
				o += c * m.outStride[i];
			}
			m.out[o] = m.in[e];
			break;
		}
		case MAP_REVTRANS: {  // out (cols × rows)[i][j] = in (rows × cols)[rows-1-j][cols-1-i]
			const long long rows = m.dims[0], cols = m.dims[1];
			const long long i = e / rows, j = e - i * rows;
			m.out[e] = m.in[(rows - 1 - j) * cols + (cols - 1 - i)];
			break;
		}
		case MAP_SCALE_BY: {  // out (rows × cols) *= aux[row] (dims[2] != 0) or aux[col]
			const long long cols = m.dims[1];
			m.out[e] *= m.aux[m.dims[2] ? e / cols : e % cols];
			break;
		}
		case MAP_EYE: {  // out = the dims[0] × dims[0] identity
			const long long n = m.dims[0];
			m.out[e] = (e / n == e % n) ? 1.0 : 0.0;
			break;
		}
		default:  // MAP_FILL
			m.out[e] = m.alpha;
			break;
		}
	}
}

// Sum of (x, y) over the workgroup, fixed order; every thread receives the totals.
template <int NT>
__device__ inline void block_sum2(double& x, double& y, double* red) {
	for (int off = 32; off > 0; off >>= 1) { x += __shfl_xor(x, off); y += __shfl_xor(y, off); }
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	__syncthreads();
	if (lane == 0) { red[2 * wave] = x; red[2 * wave + 1] = y; }
	__syncthreads();
	x = 0; y = 0;
	for (int w = 0; w < NT / 64; ++w) { x += red[2 * w]; y += red[2 * w + 1]; }
}

// r = b − q, p = r; rr = ‖r‖², rr0 = ‖b‖²; done when rr <= tol2·rr0 or no iteration is allowed. A warm start
// worse than zero (rr > rr0) restarts from x = 0, r = p = b (CG's recursive residual cannot resolve the true one
// below ε·‖A‖·‖x‖). The state's tol2 and maxIter are set beforehand. q == nullptr: x = 0 (q = 0).
template <class T>
__global__ __launch_bounds__(CT) void cg_group_start_kernel(const CGTaskT<T>* tasks) {
	__shared__ double red[2 * CT / 64];
	const CGTaskT<T> k = tasks[blockIdx.x];
	double rr = 0, bb = 0;
	for (int i = threadIdx.x; i < k.n; i += CT) {
		const T ri = k.q ? k.b[i] - k.q[i] : k.b[i];
		k.r[i] = ri;
		k.p[i] = ri;
		if (!k.q) k.x[i] = T(0);
		rr += (double)ri * (double)ri;
		bb += (double)k.b[i] * (double)k.b[i];
	}
	block_sum2<CT>(rr, bb, red);
	CGState* st = k.st;
	const bool done = rr <= st->tol2 * bb || st->maxIter <= 0;
	if (!done && rr > bb) {
		for (int i = threadIdx.x; i < k.n; i += CT) {
			k.x[i] = T(0);
			k.r[i] = k.b[i];
			k.p[i] = k.b[i];
		}
		rr = bb;
	}
	if (threadIdx.x == 0) {
		st->rr = rr;
		st->rr0 = bb;
		st->iter = 0;
		st->done = done ? 1 : 0;
	}
}

// One CG step after q = A·p: α = rr/(p·q), x += α·p, r −= α·q, then the stop test and p = r + β·p.
template <class T>
__global__ __launch_bounds__(CT) void cg_group_step_kernel(const CGTaskT<T>* tasks) {
	__shared__ double red[2 * CT / 64];
	const CGTaskT<T> k = tasks[blockIdx.x];
	CGState* st = k.st;
	if (st->done) return;
	double pq = 0, z = 0;
	for (int i = threadIdx.x; i < k.n; i += CT) pq += (double)k.p[i] * (double)k.q[i];
	block_sum2<CT>(pq, z, red);
	if (!(pq > 0)) {  // breakdown (non-positive curvature or exact solution): stop where we are
		if (threadIdx.x == 0) { st->alpha = 0; st->done = 1; }
		return;
	}
	const double rrOld = st->rr, alpha = rrOld / pq;
	const T al = (T)alpha;
	double rr = 0;
	z = 0;
	for (int i = threadIdx.x; i < k.n; i += CT) {
		k.x[i] += al * k.p[i];
		const T ri = k.r[i] - al * k.q[i];
		k.r[i] = ri;
		rr += (double)ri * (double)ri;
	}
	block_sum2<CT>(rr, z, red);
	const int iter = st->iter + 1;
	const bool done = rr <= st->tol2 * st->rr0 || iter >= st->maxIter;
	if (!done) {
		const T beta = (T)(rr / rrOld);
		for (int i = threadIdx.x; i < k.n; i += CT) k.p[i] = k.r[i] + beta * k.p[i];
	}
	__syncthreads();  // every thread has read st before it changes
	if (threadIdx.x == 0) {
		st->alpha = alpha;
		st->rr = rr;
		st->iter = iter;
		st->done = done ? 1 : 0;
	}
}

// Chunked CG step (cg_group_step_kernel's arithmetic over chunks of a system; see CGChunk).
template <class T>
__global__ __launch_bounds__(CT) void cg_chunk_dot_kernel(const CGChunk<T>* tasks) {
	__shared__ double red[2 * CT / 64];
	const CGChunk<T> t = tasks[blockIdx.x];
	if (t.k.st->done) return;
	double pq = 0, z = 0;
	for (int i = t.c0 + threadIdx.x; i < t.c0 + t.len; i += CT) pq += (double)t.k.p[i] * (double)t.k.q[i];
	block_sum2<CT>(pq, z, red);
	if (threadIdx.x == 0) t.part[t.ci] = pq;
}
template <class T>
__global__ __launch_bounds__(CT) void cg_chunk_update_kernel(const CGChunk<T>* tasks) {
	__shared__ double red[2 * CT / 64];
	const CGChunk<T> t = tasks[blockIdx.x];
	CGState* st = t.k.st;
	if (st->done) return;
	double pq = 0;
	for (int c = 0; c < t.nch; ++c) pq += t.part[c];
	if (!(pq > 0)) {  // breakdown: every chunk sees the same sum and stops
		if (t.ci == 0 && threadIdx.x == 0) { st->alpha = 0; st->done = 1; }
		return;
	}
	const double rrOld = st->rr, alpha = rrOld / pq;
	const T al = (T)alpha;
	double rr = 0, z = 0;
	for (int i = t.c0 + threadIdx.x; i < t.c0 + t.len; i += CT) {
		t.k.x[i] += al * t.k.p[i];
		const T ri = t.k.r[i] - al * t.k.q[i];
		t.k.r[i] = ri;
		rr += (double)ri * (double)ri;
	}
	block_sum2<CT>(rr, z, red);
	if (threadIdx.x == 0) {
		t.part[t.nch + t.ci] = rr;
		if (t.ci == 0) {  // st->rr and st->iter are rewritten by the next launch, which reads these copies
			st->alpha = alpha;
			st->rrPrev = rrOld;
			st->iterPrev = st->iter;
		}
	}
}
template <class T>
__global__ __launch_bounds__(CT) void cg_chunk_direction_kernel(const CGChunk<T>* tasks) {
	const CGChunk<T> t = tasks[blockIdx.x];
	CGState* st = t.k.st;
	// chunk 0 may set done in this launch: a chunk that reads it already set would have stopped too
	if (st->done) return;
	double rr = 0;
	for (int c = 0; c < t.nch; ++c) rr += t.part[t.nch + c];
	const int iter = st->iterPrev + 1;
	const bool done = rr <= st->tol2 * st->rr0 || iter >= st->maxIter;
	if (!done) {
		const T beta = (T)(rr / st->rrPrev);
		for (int i = t.c0 + threadIdx.x; i < t.c0 + t.len; i += CT) t.k.p[i] = t.k.r[i] + beta * t.k.p[i];
	}
	if (t.ci == 0 && threadIdx.x == 0) {
		st->rr = rr;
		st->iter = iter;
		st->done = done ? 1 : 0;
	}
}

__global__ __launch_bounds__(GT) void mix_group_kernel(const MixTask* tasks, const int* start, int count) {
	const int t = group_task(start, count, (int)blockIdx.x);
	const MixTask m = tasks[t];
	const long long e0 = (long long)(blockIdx.x - start[t]) * GE;
	const long long e1 = min(m.n, e0 + GE);
	for (long long e = e0 + threadIdx.x; e < e1; e += GT) {
		if (m.op == MIX_CVT32) {
			static_cast<float*>(m.out)[e] = (float)(m.alpha * static_cast<const double*>(m.a)[e]);
		} else if (m.op == MIX_AXPY32) {
			static_cast<double*>(m.out)[e] += m.alpha * (double)static_cast<const float*>(m.a)[e];
		} else {  // MIX_ZERO
			static_cast<double*>(m.out)[e] = 0.0;
			static_cast<double*>(m.out2)[e] = static_cast<const double*>(m.a)[e];
		}
	}
}

__global__ __launch_bounds__(CT) void resid_group_kernel(const ResidTask* tasks, double* out) {
	__shared__ double red[2 * CT / 64];
	const ResidTask k = tasks[blockIdx.x];
	double rr = 0, bb = 0;
	for (int i = threadIdx.x; i < k.n; i += CT) {
		const double ri = k.b[i] - k.q[i];
		k.r[i] = ri;
		rr += ri * ri;
		bb += k.b[i] * k.b[i];
	}
	block_sum2<CT>(rr, bb, red);
	if (threadIdx.x == 0) { out[2 * blockIdx.x] = rr; out[2 * blockIdx.x + 1] = bb; }
}

// out[2t] = a·b, out[2t + 1] = c·d (c may be null: 0)
__global__ __launch_bounds__(CT) void dot2_group_kernel(const Dot2Task* tasks, double* out) {
	__shared__ double red[2 * CT / 64];
	const Dot2Task k = tasks[blockIdx.x];
	double s0 = 0, s1 = 0;
	for (int i = threadIdx.x; i < k.n; i += CT) {
		s0 += k.a[i] * k.b[i];
		if (k.c) s1 += k.c[i] * k.d[i];
	}
	block_sum2<CT>(s0, s1, red);
	if (threadIdx.x == 0) { out[2 * blockIdx.x] = s0; out[2 * blockIdx.x + 1] = s1; }
}

__global__ __launch_bounds__(GT) void res_core_group_kernel(const ResCoreTask* tasks, const int* start, int count) {
	const int t = group_task(start, count, (int)blockIdx.x);
	const ResCoreTask k = tasks[t];
	const int pl = k.rA * k.rx, pr = k.rA2 * k.rx2;
	const int L = k.first ? 1 : pl + k.rb, R = k.last ? 1 : pr + k.rb2;
	const long long total = (long long)L * k.n * R;
	const long long e0 = (long long)(blockIdx.x - start[t]) * GE;
	const long long e1 = min(total, e0 + GE);
	for (long long e = e0 + threadIdx.x; e < e1; e += GT) {
		const int r = (int)(e % R);
		const long long lj = e / R;
		const int j = (int)(lj % k.n), l = (int)(lj / k.n);
		double v = 0.0;
		if ((k.first || l < pl) && (k.last || r < pr)) {  // the product block
			const int al = k.first ? 0 : l / k.rx, i = k.first ? 0 : l - (l / k.rx) * k.rx;
			const int ar = k.last ? 0 : r / k.rx2, ir = k.last ? 0 : r - (r / k.rx2) * k.rx2;
			const double* a = k.A + ((size_t)(al * k.n + j) * k.m) * k.rA2 + ar;
			const double* x = k.x + (size_t)i * k.m * k.rx2 + ir;
			for (int mm = 0; mm < k.m; ++mm) v = fma(a[(size_t)mm * k.rA2], x[(size_t)mm * k.rx2], v);
		}
		if ((k.first || l >= pl) && (k.last || r >= pr)) {  // the right-hand side block
			const int bl = k.first ? 0 : l - pl, br = k.last ? 0 : r - pr;
			const double bv = k.b[((size_t)bl * k.n + j) * k.rb2 + br];
			v += k.first ? -bv : bv;
		}
		k.w[e] = v;
	}
}

}  // namespace

void res_core_group(const std::vector<ResCoreTask>& tasks, hipStream_t s) {
	if (tasks.empty()) return;
	std::vector<int> start;
	int blocks = 0;
	for (const ResCoreTask& k : tasks) {
		start.push_back(blocks);
		const long long L = k.first ? 1 : k.rA * k.rx + k.rb, R = k.last ? 1 : k.rA2 * k.rx2 + k.rb2;
		blocks += (int)((L * k.n * R + GE - 1) / GE);
	}
	const ResCoreTask* dt = static_cast<const ResCoreTask*>(table_upload(tasks.data(), tasks.size() * sizeof(ResCoreTask), s));
	const int* ds = static_cast<const int*>(table_upload(start.data(), start.size() * sizeof(int), s));
	launch_timed("res_core_group_kernel", Work(0.0, 0.0), res_core_group_kernel, dim3((unsigned)blocks), dim3(GT), 0, s, dt, ds,
	             (int)tasks.size());
}

void map_group(const std::vector<MapTask>& tasks, hipStream_t s) {
	std::vector<MapTask> live;
	std::vector<int> start;
	int blocks = 0;
	for (const MapTask& m : tasks) {
		if (m.n <= 0) continue;
		live.push_back(m);
		start.push_back(blocks);
		blocks += (int)((m.n + GE - 1) / GE);
	}
	if (live.empty()) return;
	const MapTask* dt = static_cast<const MapTask*>(table_upload(live.data(), live.size() * sizeof(MapTask), s));
	const int* ds = static_cast<const int*>(table_upload(start.data(), start.size() * sizeof(int), s));
	launch_timed("map_group_kernel", Work(0.0, 0.0), map_group_kernel, dim3(blocks), dim3(GT), 0, s, dt, ds, (int)live.size());
}

MapTask map_permute(size_t degree, const size_t* dims, const size_t* shuffle, const double* in, double* out) {
	XTT_REQUIRE(degree >= 1 && degree <= 4, "grouped permute: 1 to 4 modes");
	MapTask m{};
	m.op = MAP_PERMUTE;
	m.deg = (int)degree;
	long long n = 1, newDims[4], newStride[4];
	for (size_t i = 0; i < degree; ++i) { m.dims[i] = (long long)dims[i]; n *= (long long)dims[i]; newDims[shuffle[i]] = (long long)dims[i]; }
	long long st = 1;
	for (size_t j = degree; j-- > 0;) { newStride[j] = st; st *= newDims[j]; }
	for (size_t i = 0; i < degree; ++i) m.outStride[i] = newStride[shuffle[i]];
	m.n = n;
	m.in = in;
	m.out = out;
	return m;
}
MapTask map_copy(const double* in, double* out, size_t n) {
	const size_t dims[1] = {n}, sh[1] = {0};
	return map_permute(1, dims, sh, in, out);
}
MapTask map_revtrans(const double* in, size_t rows, size_t cols, double* out) {
	MapTask m{};
	m.op = MAP_REVTRANS; m.n = (long long)(rows * cols); m.dims[0] = (long long)rows; m.dims[1] = (long long)cols;
	m.in = in; m.out = out;
	return m;
}
MapTask map_scale_by(double* mtx, const double* sv, size_t rows, size_t cols, bool byRow) {
	MapTask m{};
	m.op = MAP_SCALE_BY; m.n = (long long)(rows * cols); m.dims[0] = (long long)rows; m.dims[1] = (long long)cols; m.dims[2] = byRow;
	m.out = mtx; m.aux = sv;
	return m;
}
MapTask map_eye(double* out, size_t n) {
	MapTask m{};
	m.op = MAP_EYE; m.n = (long long)(n * n); m.dims[0] = (long long)n; m.out = out;
	return m;
}
MapTask map_fill(double* out, double v, size_t n) {
	MapTask m{};
	m.op = MAP_FILL; m.n = (long long)n; m.out = out; m.alpha = v;
	return m;
}

template <class T>
void cg_group_start(const std::vector<CGTaskT<T>>& tasks, hipStream_t s) {
	if (tasks.empty()) return;
	const CGTaskT<T>* dt = static_cast<const CGTaskT<T>*>(table_upload(tasks.data(), tasks.size() * sizeof(CGTaskT<T>), s));
	launch_timed(sizeof(T) == 8 ? "cg_group_start_kernel<double>" : "cg_group_start_kernel<float>", Work(0.0, 0.0), cg_group_start_kernel<T>,
	             dim3((unsigned)tasks.size()), dim3(CT), 0, s, dt);
}
template <class T>
void cg_group_step_run(const CGTaskT<T>* devTasks, int count, hipStream_t s) {
	launch_timed(sizeof(T) == 8 ? "cg_group_step_kernel<double>" : "cg_group_step_kernel<float>", Work(0.0, 0.0), cg_group_step_kernel<T>,
	             dim3((unsigned)count), dim3(CT), 0, s, devTasks);
}
template <class T>
void cg_group_step(const std::vector<CGTaskT<T>>& tasks, hipStream_t s) {
	if (tasks.empty()) return;
	const CGTaskT<T>* dt = static_cast<const CGTaskT<T>*>(table_upload(tasks.data(), tasks.size() * sizeof(CGTaskT<T>), s));
	cg_group_step_run<T>(dt, (int)tasks.size(), s);
}
template <class T>
void cg_chunk_step_run(const CGChunk<T>* devTasks, int count, hipStream_t s) {
	const bool f64 = sizeof(T) == 8;
	launch_timed(f64 ? "cg_chunk_dot_kernel<double>" : "cg_chunk_dot_kernel<float>", Work(0.0, 0.0), cg_chunk_dot_kernel<T>,
	             dim3((unsigned)count), dim3(CT), 0, s, devTasks);
	launch_timed(f64 ? "cg_chunk_update_kernel<double>" : "cg_chunk_update_kernel<float>", Work(0.0, 0.0), cg_chunk_update_kernel<T>,
	             dim3((unsigned)count), dim3(CT), 0, s, devTasks);
	launch_timed(f64 ? "cg_chunk_direction_kernel<double>" : "cg_chunk_direction_kernel<float>", Work(0.0, 0.0),
	             cg_chunk_direction_kernel<T>, dim3((unsigned)count), dim3(CT), 0, s, devTasks);
}
template void cg_chunk_step_run<double>(const CGChunk<double>*, int, hipStream_t);
template void cg_chunk_step_run<float>(const CGChunk<float>*, int, hipStream_t);
template void cg_group_start<double>(const std::vector<CGTaskT<double>>&, hipStream_t);
template void cg_group_start<float>(const std::vector<CGTaskT<float>>&, hipStream_t);
template void cg_group_step<double>(const std::vector<CGTaskT<double>>&, hipStream_t);
template void cg_group_step<float>(const std::vector<CGTaskT<float>>&, hipStream_t);
template void cg_group_step_run<double>(const CGTaskT<double>*, int, hipStream_t);
template void cg_group_step_run<float>(const CGTaskT<float>*, int, hipStream_t);

void mix_group(const std::vector<MixTask>& tasks, hipStream_t s) {
	std::vector<MixTask> live;
	std::vector<int> start;
	int blocks = 0;
	for (const MixTask& m : tasks) {
		if (m.n <= 0) continue;
		live.push_back(m);
		start.push_back(blocks);
		blocks += (int)((m.n + GE - 1) / GE);
	}
	if (live.empty()) return;
	const MixTask* dt = static_cast<const MixTask*>(table_upload(live.data(), live.size() * sizeof(MixTask), s));
	const int* ds = static_cast<const int*>(table_upload(start.data(), start.size() * sizeof(int), s));
	launch_timed("mix_group_kernel", Work(0.0, 0.0), mix_group_kernel, dim3(blocks), dim3(GT), 0, s, dt, ds, (int)live.size());
}

void resid_group(const std::vector<ResidTask>& tasks, double* out, hipStream_t s) {
	if (tasks.empty()) return;
	const ResidTask* dt = static_cast<const ResidTask*>(table_upload(tasks.data(), tasks.size() * sizeof(ResidTask), s));
	launch_timed("resid_group_kernel", Work(0.0, 0.0), resid_group_kernel, dim3((unsigned)tasks.size()), dim3(CT), 0, s, dt, out);
}

void dot2_group(const std::vector<Dot2Task>& tasks, double* out, hipStream_t s) {
	if (tasks.empty()) return;
	const Dot2Task* dt = static_cast<const Dot2Task*>(table_upload(tasks.data(), tasks.size() * sizeof(Dot2Task), s));
	launch_timed("dot2_group_kernel", Work(0.0, 0.0), dot2_group_kernel, dim3((unsigned)tasks.size()), dim3(CT), 0, s, dt, out);
}

}  // namespace xtt
