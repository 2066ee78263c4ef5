// libxtt runtime: thread-local error text, the per-device library stream and the caching device allocator.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <unordered_map>

#include "xtt_common.hpp"

namespace xtt {

namespace {
thread_local std::string g_error;

struct DeviceState {
	hipStream_t stream = nullptr;
};
std::mutex g_mu;
std::unordered_map<int, DeviceState> g_dev;

// Allocation cache: (device, stream, size class rounded to 256 B below 1 MiB, to 64 KiB above) → free blocks.
// A block freed while a stream is current goes back to that stream's pool, so reuse stays ordered by the
// stream that last used it. Blocks under the null stream are idle (their stream had finished all work when they
// were parked there) and serve any stream.
struct Block { size_t bytes; int device; };
using PoolKey = std::tuple<int, hipStream_t, size_t>;
struct PoolKeyHash {
	size_t operator()(const PoolKey& k) const {
		size_t h = std::hash<size_t>()(std::get<2>(k));
		h ^= std::hash<const void*>()(std::get<1>(k)) + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2);
		h ^= (size_t)(unsigned)std::get<0>(k) * 0x100000001b3ull;
		return h;
	}
};
std::mutex g_amu;
// one free list per key: lookups are a hash and a vector pop, frees a hash and a push (no node allocation)
std::unordered_map<PoolKey, std::vector<void*>, PoolKeyHash> g_free;
std::unordered_map<void*, Block> g_live;
size_t g_inUse = 0;

size_t size_class(size_t b) {
	if (b <= (1u << 20)) return (b + 255) & ~size_t(255);
	return (b + 65535) & ~size_t(65535);
}
}  // namespace

void set_error(const std::string& s) { g_error = s; }
const std::string& last_error() { return g_error; }

namespace {
thread_local hipStream_t t_stream = nullptr;
}
void bind_thread_stream(hipStream_t s) { t_stream = s; }

namespace {
thread_local int t_lastDev = -1;                 // the device of the last stream() lookup on this thread
thread_local hipStream_t t_lastStream = nullptr;  // and its library stream (streams live until the process ends)
hipStream_t device_stream(int dev) {
	if (dev == t_lastDev) return t_lastStream;
	std::lock_guard<std::mutex> lk(g_mu);
	DeviceState& d = g_dev[dev];
	if (!d.stream) XTT_HIP(hipStreamCreateWithFlags(&d.stream, hipStreamDefault));
	t_lastDev = dev;
	t_lastStream = d.stream;
	return d.stream;
}
}  // namespace

hipStream_t stream() {
	if (t_stream) return t_stream;
	int dev = 0;
	XTT_HIP(hipGetDevice(&dev));
	return device_stream(dev);
}

void sync() { XTT_HIP(hipStreamSynchronize(stream())); }

namespace {
// Pinned staging buffers for read_back, shared by all threads; kept until the process ends.
constexpr size_t kPinnedBytes = 64 * 1024;
std::mutex g_pinMu;
std::vector<void*> g_pinFree;
}  // namespace

void read_back(void* host, const void* dev, size_t bytes) { read_back(host, dev, bytes, stream()); }

void read_back(void* host, const void* dev, size_t bytes, hipStream_t s) {
	if (bytes == 0) return;
	if (bytes > kPinnedBytes) {
		XTT_HIP(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, s));
		XTT_HIP(hipStreamSynchronize(s));
		return;
	}
	void* buf = nullptr;
	{
		std::lock_guard<std::mutex> lk(g_pinMu);
		if (!g_pinFree.empty()) {
			buf = g_pinFree.back();
			g_pinFree.pop_back();
		}
	}
	if (!buf) XTT_HIP(hipHostMalloc(&buf, kPinnedBytes, hipHostMallocDefault));
	hipError_t e = hipMemcpyAsync(buf, dev, bytes, hipMemcpyDeviceToHost, s);
	if (e == hipSuccess) e = hipStreamSynchronize(s);
	if (e == hipSuccess) std::memcpy(host, buf, bytes);
	{
		std::lock_guard<std::mutex> lk(g_pinMu);
		g_pinFree.push_back(buf);
	}
	XTT_HIP(e);
}

namespace {
thread_local Arena* t_arena = nullptr;
constexpr size_t kArenaChunk = 64ull << 20;
}  // namespace

void* Arena::alloc(size_t bytes) {
	const size_t b = (std::max<size_t>(bytes, 1) + 255) & ~size_t(255);
	if (chunks.empty() || used + b > chunks.back().second) {
		const size_t sz = std::max(kArenaChunk, b);
		void* p = nullptr;
		const hipError_t e = hipMalloc(&p, sz);
		if (e != hipSuccess) {
			(void)hipGetLastError();
			throw Error(e == hipErrorOutOfMemory ? ERR_NOMEM : ERR_HIP, std::string("hipMalloc (capture arena): ") + hipGetErrorString(e));
		}
		chunks.push_back({static_cast<char*>(p), sz});
		used = 0;
	}
	char* p = chunks.back().first + used;
	used += b;
	total += b;
	return p;
}
Arena::~Arena() {
	if (chunks.empty()) return;
	(void)hipDeviceSynchronize();
	for (auto& c : chunks) (void)hipFree(c.first);
}
namespace {
struct ScratchState {
	std::vector<std::pair<char*, size_t>> chunks;
	size_t chunk = 0, off = 0;
	int depth = 0;
};
constexpr size_t kScratchChunk = 256ull << 20;
thread_local std::unordered_map<int, ScratchState> t_scratch;  // chunks are kept until the process ends
ScratchState& scratch_state(int* devOut) {
	int dev = 0;
	XTT_HIP(hipGetDevice(&dev));
	if (devOut) *devOut = dev;
	return t_scratch[dev];
}
}  // namespace

void* scratch_alloc(size_t bytes) {
	ScratchState& S = scratch_state(nullptr);
	XTT_REQUIRE(S.depth > 0, "scratch allocation outside a ScratchFrame");
	const size_t b = (std::max<size_t>(bytes, 1) + 255) & ~size_t(255);
	for (;;) {
		if (S.chunk < S.chunks.size() && S.off + b <= S.chunks[S.chunk].second) break;
		if (S.chunk < S.chunks.size() && S.off == 0) {
			// an empty chunk too small for this request: a larger one takes its place
			(void)hipStreamSynchronize(stream());  // earlier launches may still use it
			(void)hipFree(S.chunks[S.chunk].first);
			S.chunks.erase(S.chunks.begin() + (long)S.chunk);
		} else if (S.chunk < S.chunks.size()) {
			++S.chunk;
			S.off = 0;
			continue;
		}
		if (S.chunk >= S.chunks.size()) {
			void* p = nullptr;
			const size_t sz = std::max(kScratchChunk, b);
			hipError_t e = hipMalloc(&p, sz);
			if (e == hipErrorOutOfMemory) {
				(void)hipGetLastError();
				release_cache();
				e = hipMalloc(&p, sz);
			}
			if (e != hipSuccess) {
				(void)hipGetLastError();
				throw Error(e == hipErrorOutOfMemory ? ERR_NOMEM : ERR_HIP, std::string("hipMalloc (scratch): ") + hipGetErrorString(e));
			}
			S.chunks.insert(S.chunks.begin() + (long)S.chunk, {static_cast<char*>(p), sz});
			S.off = 0;
		}
	}
	char* p = S.chunks[S.chunk].first + S.off;
	S.off += b;
	return p;
}
ScratchFrame::ScratchFrame() {
	ScratchState& S = scratch_state(&dev);
	chunk = S.chunk;
	off = S.off;
	S.depth++;
}
ScratchFrame::~ScratchFrame() {
	ScratchState& S = t_scratch[dev];
	S.chunk = chunk;
	S.off = off;
	S.depth--;
}

ArenaScope::ArenaScope(Arena* a) : prev(t_arena) { t_arena = a; }
bool arena_active() { return t_arena != nullptr; }
ArenaScope::~ArenaScope() { t_arena = prev; }

void* dmalloc(size_t bytes) {
	if (t_arena) return t_arena->alloc(bytes);
	if (bytes == 0) bytes = 1;
	int dev = 0;
	XTT_HIP(hipGetDevice(&dev));
	const size_t cls = size_class(bytes);
	const hipStream_t st = t_stream ? t_stream : device_stream(dev);
	{
		std::lock_guard<std::mutex> lk(g_amu);
		auto it = g_free.find(PoolKey{dev, st, cls});
		if (it == g_free.end() || it->second.empty()) it = g_free.find(PoolKey{dev, nullptr, cls});  // idle blocks
		if (it != g_free.end() && !it->second.empty()) {
			void* p = it->second.back();
			it->second.pop_back();
			g_live[p] = {cls, dev};
			g_inUse += cls;
			return p;
		}
	}
	void* p = nullptr;
	hipError_t e = hipMalloc(&p, cls);
	if (e == hipErrorOutOfMemory) {  // give cached blocks back and retry once
		(void)hipGetLastError();
		release_cache();
		e = hipMalloc(&p, cls);
	}
	if (e != hipSuccess) {
		(void)hipGetLastError();  // clear the sticky error so later launch checks do not report it
		throw Error(e == hipErrorOutOfMemory ? ERR_NOMEM : ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
	}
	std::lock_guard<std::mutex> lk(g_amu);
	g_live[p] = {cls, dev};
	g_inUse += cls;
	return p;
}

void dfree(void* p) {
	if (!p) return;
	const hipStream_t st = stream();
	std::lock_guard<std::mutex> lk(g_amu);
	auto it = g_live.find(p);
	if (it == g_live.end()) return;
	g_free[PoolKey{it->second.device, st, it->second.bytes}].push_back(p);
	g_inUse -= it->second.bytes;
	g_live.erase(it);
}

void park_stream_pool(hipStream_t from) {
	std::lock_guard<std::mutex> lk(g_amu);
	std::vector<std::pair<PoolKey, std::vector<void*>>> moved;
	for (auto it = g_free.begin(); it != g_free.end();) {
		if (std::get<1>(it->first) == from && from != nullptr) {
			moved.push_back({PoolKey{std::get<0>(it->first), nullptr, std::get<2>(it->first)}, std::move(it->second)});
			it = g_free.erase(it);
		} else {
			++it;
		}
	}
	for (auto& m : moved) {
		std::vector<void*>& dst = g_free[m.first];
		dst.insert(dst.end(), m.second.begin(), m.second.end());
	}
}

size_t bytes_cached() {
	std::lock_guard<std::mutex> lk(g_amu);
	size_t b = 0;
	for (auto& kv : g_free) b += std::get<2>(kv.first) * kv.second.size();
	return b;
}

void release_cache() {
	std::unordered_map<PoolKey, std::vector<void*>, PoolKeyHash> blocks;
	{
		std::lock_guard<std::mutex> lk(g_amu);
		blocks.swap(g_free);
	}
	bool any = false;
	for (auto& kv : blocks) any |= !kv.second.empty();
	if (any) (void)hipDeviceSynchronize();
	for (auto& kv : blocks)
		for (void* p : kv.second) (void)hipFree(p);
}

namespace {
struct Slot { hipEvent_t a, b; std::string tag; double flops, bytes; };
std::mutex g_pmu;
bool g_prof = false;
std::vector<Slot> g_slots;
}  // namespace

bool prof_enabled() { return g_prof; }
bool prof_shapes() {
	static const bool on = getenv("XTT_PROFILE_SHAPES") != nullptr;
	return on;
}
void prof_enable(bool on) { std::lock_guard<std::mutex> lk(g_pmu); g_prof = on; }

ProfEvents prof_events(const char* tag, const Work& w) {
	ProfEvents e;
	if (!g_prof) return e;
	Slot sl;
	XTT_HIP(hipEventCreate(&sl.a));
	XTT_HIP(hipEventCreate(&sl.b));
	sl.tag = tag;
	sl.flops = w.flops;
	sl.bytes = w.bytes;
	e.start = sl.a;
	e.stop = sl.b;
	std::lock_guard<std::mutex> lk(g_pmu);
	g_slots.push_back(sl);
	return e;
}

std::vector<ProfRecord> prof_collect() {
	XTT_HIP(hipDeviceSynchronize());
	std::lock_guard<std::mutex> lk(g_pmu);
	std::map<std::string, ProfRecord> agg;
	for (Slot& sl : g_slots) {
		float ms = 0;
		XTT_HIP(hipEventElapsedTime(&ms, sl.a, sl.b));
		ProfRecord& pr = agg[sl.tag];
		pr.tag = sl.tag;
		pr.launches++;
		pr.ms += ms;
		pr.flops += sl.flops;
		pr.bytes += sl.bytes;
		(void)hipEventDestroy(sl.a);
		(void)hipEventDestroy(sl.b);
	}
	g_slots.clear();
	std::vector<ProfRecord> out;
	for (auto& kv : agg) out.push_back(kv.second);
	return out;
}

size_t bytes_in_use() {
	std::lock_guard<std::mutex> lk(g_amu);
	return g_inUse;
}

void DTensor::reshape(std::vector<size_t> d) {
	size_t p = 1;
	for (size_t x : d) p *= x;
	XTT_REQUIRE(p == count(), "reshape: element count changes");
	dims = std::move(d);
}

PerfSink& perf_sink() {
	static thread_local PerfSink s;
	return s;
}
double PerfClock::now_us() {
	return std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace xtt
