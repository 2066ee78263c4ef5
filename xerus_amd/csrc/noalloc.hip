// An entry point that allocates nothing: the one-pass ADF residual and gradient (the kernels of
// xt_tt_sample_residual_adjoint_f32) with every scratch buffer carved from a file-scope device array.
//
// The array is a per-GPU copy held until the process ends. Host code never passes the symbol itself: its device
// address on the current GPU is looked up once per device with hipGetSymbolAddress and kept. Nothing in this file
// calls hipMalloc/hipFree or the library's caching allocator, and nothing synchronises, so the call can sit in a
// loop that must not allocate or be recorded into a graph. The scratch contents of one call are overwritten by the
// next call on the same GPU: calls that overlap on one GPU (two streams, two threads) race on it.
#include <mutex>

#include "../../include/xtt.h"
#include "kernels.hpp"
#include "xtt_common.hpp"

namespace xtt {
namespace {

// Layout (bytes, each segment 256-aligned): the descriptor arrays, the padded cores, the per-block norm partials,
// a norm pair for callers that pass no `sums`, then the per-block gradient partials.
constexpr size_t kDescBytes = 1024;                     // n[32] int, rank[33] int, coreOff[32] long long
constexpr size_t kPadBytes = 64 * 1024;                 // adf_pad_floats(a) floats; the kernel bounds Σ n_k·68 floats by 48 KiB
constexpr size_t kPadOffBytes = 256;                    // ADF_PAD_OFFSETS long longs
constexpr size_t kNormPartBytes = 2 * 256 * sizeof(double);
constexpr size_t kNormBytes = 256;
constexpr size_t kHead = kDescBytes + kPadBytes + kPadOffBytes + kNormPartBytes + kNormBytes;
constexpr size_t kScratchBytes = 16u << 20;             // 16 MiB per GPU
constexpr size_t kMaxBlocks = 256;                      // gradient blocks of the one-pass kernel
constexpr size_t kMaxGradFloats = (kScratchBytes - kHead) / (kMaxBlocks * sizeof(float));

__device__ __align__(256) unsigned char scratch[kScratchBytes];

struct DescValues {
	int d;
	int n[32];
	int rank[33];
	long long off[32];
};

__global__ void fill_desc_kernel(DescValues v, int* n, int* rank, long long* off) {
	const int t = threadIdx.x;
	if (t < v.d) { n[t] = v.n[t]; off[t] = v.off[t]; }
	if (t <= v.d) rank[t] = v.rank[t];
}

__global__ void zero_outputs_kernel(float* grads, size_t count, double* sums) {
	for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x) grads[i] = 0.f;
	if (sums && blockIdx.x == 0 && threadIdx.x < 2) sums[threadIdx.x] = 0.0;
}

// The array's address on the current device, looked up at the first call on that device.
unsigned char* scratch_on_device() {
	constexpr int kMaxDevices = 64;
	static std::mutex mu;
	static unsigned char* ptr[kMaxDevices] = {};
	int dev = 0;
	XTT_HIP(hipGetDevice(&dev));
	XTT_REQUIRE(dev >= 0 && dev < kMaxDevices, "device ordinal out of range");
	std::lock_guard<std::mutex> lock(mu);
	if (!ptr[dev]) {
		void* p = nullptr;
		XTT_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
		ptr[dev] = static_cast<unsigned char*>(p);
	}
	return ptr[dev];
}

}  // namespace
}  // namespace xtt

extern "C" int xt_tt_sample_residual_adjoint_f32_noalloc(int d, const int* n, const int* r, const float* cores, const uint32_t* pos, int P,
                                                         const float* b, float* resid, double* sums, float* grads, void* stream) {
	using namespace xtt;
	(void)hipGetLastError();
	try {
		XTT_REQUIRE(n && r && cores && resid && grads, "n, r, cores, resid and grads must not be NULL");
		XTT_REQUIRE(P == 0 || (pos && b), "pos/b must not be NULL");
		XTT_REQUIRE(d >= 2 && d <= 16, "the allocation-free call supports orders 2..16");
		XTT_REQUIRE(P >= 0, "invalid count");
		XTT_REQUIRE(r[0] == 1 && r[d] == 1, "boundary ranks must be 1");
		DescValues v{};
		v.d = d;
		long long total = 0;
		int maxR = 1, sumN = 0, maxN = 0;
		for (int k = 0; k < d; ++k) {
			XTT_REQUIRE(n[k] > 0 && r[k] > 0 && r[k + 1] > 0, "dims and ranks must be positive");
			v.n[k] = n[k];
			v.off[k] = total;
			total += (long long)r[k] * n[k] * r[k + 1];
			maxR = std::max(maxR, std::max(r[k], r[k + 1]));
			sumN += n[k];
			maxN = std::max(maxN, n[k]);
		}
		for (int k = 0; k <= d; ++k) v.rank[k] = r[k];
		if (total > (long long)kMaxGradFloats)
			throw Error(ERR_UNSUPPORTED, "allocation-free call: the cores hold " + std::to_string(total) + " floats, at most " +
			                                 std::to_string(kMaxGradFloats) + " fit its scratch");
		hipStream_t s = pick(stream);
		unsigned char* base = scratch_on_device();
		int* dn = reinterpret_cast<int*>(base);
		int* drank = dn + 32;
		long long* doff = reinterpret_cast<long long*>(base + 512);
		float* Gp = reinterpret_cast<float*>(base + kDescBytes);
		long long* padOff = reinterpret_cast<long long*>(base + kDescBytes + kPadBytes);
		double* normPart = reinterpret_cast<double*>(base + kDescBytes + kPadBytes + kPadOffBytes);
		double* normOwn = reinterpret_cast<double*>(base + kDescBytes + kPadBytes + kPadOffBytes + kNormPartBytes);
		float* partials = reinterpret_cast<float*>(base + kHead);
		AdfDesc desc{};
		desc.d = d; desc.P = P; desc.n = dn; desc.rank = drank; desc.coreOff = doff; desc.maxR = maxR; desc.gradSize = (int)total;
		desc.sumN = sumN; desc.maxN = maxN;
		if (!adf_residual_gradient_fused_ok(desc) || (size_t)adf_pad_floats(desc) * sizeof(float) > kPadBytes)
			throw Error(ERR_UNSUPPORTED, "allocation-free call: needs ranks <= 8, n_k <= 16 and the padded cores within 48 KiB "
			                             "(the one-pass kernel); use xt_tt_sample_residual_adjoint_f32 otherwise");
		if (P == 0) {
			hipLaunchKernelGGL(zero_outputs_kernel, dim3(std::max<unsigned>(1, std::min<unsigned>(1024, ((unsigned)total + 255) / 256))),
			                   dim3(256), 0, s, grads, (size_t)total, sums);
			XTT_LAUNCH_CHECK();
			return XT_OK;
		}
		XTT_REQUIRE((size_t)std::min(adf_grad_blocks(P), 256) <= kMaxBlocks, "gradient block count");
		hipLaunchKernelGGL(fill_desc_kernel, dim3(1), dim3(64), 0, s, v, dn, drank, doff);
		XTT_LAUNCH_CHECK();
		adf_residual_gradient_f32_into(desc, cores, pos, b, resid, sums ? sums : normOwn, grads, partials, normPart, Gp, padOff, s);
		return XT_OK;
	} catch (const Error& e) {
		set_error(e.what());
		return e.code;
	} catch (const std::exception& e) {
		set_error(e.what());
		return XT_ERR_INVALID;
	}
}
