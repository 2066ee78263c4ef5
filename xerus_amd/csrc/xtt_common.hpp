// Shared runtime pieces of libxtt: status handling, the library stream, a caching device allocator and the
// device tensor type used by the TT layer. Everything here is host code compiled by hipcc for gfx950.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <cstddef>
#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace xtt {

enum Status : int { OK = 0, ERR_INVALID = 1, ERR_HIP = 2, ERR_NUMERIC = 3, ERR_NOMEM = 4, ERR_UNSUPPORTED = 5 };

struct Error : std::runtime_error {
	int code;
	Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define XTT_REQUIRE(cond, msg)                                                                          \
	do {                                                                                                \
		if (!(cond)) throw ::xtt::Error(::xtt::ERR_INVALID, std::string(msg));                          \
	} while (0)

#define XTT_HIP(call)                                                                                   \
	do {                                                                                                \
		hipError_t e_ = (call);                                                                         \
		if (e_ != hipSuccess)                                                                           \
			throw ::xtt::Error(e_ == hipErrorOutOfMemory ? ::xtt::ERR_NOMEM : ::xtt::ERR_HIP,           \
			                   std::string(#call) + ": " + hipGetErrorString(e_));                     \
	} while (0)

// Kernel launch check: a launch error surfaces on the next runtime call, so query it right away.
#define XTT_LAUNCH_CHECK() XTT_HIP(hipGetLastError())

void set_error(const std::string& s);

// Device → host copy of a small result (ranks, norms, solver state) on the calling thread's stream (or on s),
// waiting for it: staged through pinned host memory, so the copy is one DMA without the runtime's pageable-memory
// staging.
void read_back(void* host, const void* dev, size_t bytes);
void read_back(void* host, const void* dev, size_t bytes, hipStream_t s);

// Optional launch profiler: while enabled, launches made through launch_timed hand a HIP event pair to the
// dispatch itself (hipExtLaunchKernel), so the pair brackets the kernel's execution and not the host's
// submission gap. Each record carries the kernel instantiation's tag and its algorithmic flop count; read
// back after a synchronisation.
struct ProfRecord { std::string tag; size_t launches = 0; double ms = 0, flops = 0, bytes = 0; };
struct ProfEvents { hipEvent_t start = nullptr, stop = nullptr; };
// Algorithmic work of one launch: flops and the HBM bytes its inputs and outputs occupy (compulsory traffic).
struct Work {
	double flops = 0, bytes = 0;
	Work(double f = 0, double b = 0) : flops(f), bytes(b) {}
};
bool prof_enabled();
bool prof_shapes();                                                // XTT_PROFILE_SHAPES set: tags carry problem shapes
void prof_enable(bool on);
ProfEvents prof_events(const char* tag, const Work& w);           // registers one record; null events when disabled
std::vector<ProfRecord> prof_collect();                            // synchronises, aggregates, clears

template <class F, class... Args>
void launch_timed(const char* tag, const Work& work, F kern, dim3 grid, dim3 block, size_t lds, hipStream_t s, Args... args) {
	if (prof_enabled()) {
		const ProfEvents e = prof_events(tag, work);
		hipExtLaunchKernelGGL(kern, grid, block, (std::uint32_t)lds, s, e.start, e.stop, 0u, args...);
	} else {
		hipLaunchKernelGGL(kern, grid, block, lds, s, args...);
	}
	XTT_LAUNCH_CHECK();
}

// PerformanceData sink (performanceData.h:43-157): while a callback is set on the calling thread, the ALS and
// ADF drivers report data points (iteration, elapsed µs with the residual evaluation excluded, residual,
// ranks, flags) as xerus's PerformanceData::add does.
using PerfCallback = void (*)(void* user, size_t iteration, double elapsedMicros, double residual, const size_t* ranks,
                              size_t nranks, size_t flags);
struct PerfSink {
	PerfCallback cb = nullptr;
	void* user = nullptr;
};
PerfSink& perf_sink();  // thread-local
// Wall clock of one solve, with pauses (xerus's stop_timer / continue_timer).
struct PerfClock {
	double start = 0, pausedAt = -1;
	static double now_us();
	void begin() { start = now_us(); pausedAt = -1; }
	void pause() { pausedAt = now_us(); }
	void resume() { if (pausedAt >= 0) { start += now_us() - pausedAt; pausedAt = -1; } }
	double elapsed() const { return (pausedAt >= 0 ? pausedAt : now_us()) - start; }
};

// The library's stream on the current device (default priority, blocking: it orders with the legacy null
// stream torch uses), created on first use.
hipStream_t stream();
// Binds the calling thread to a stream of its own (nullptr: back to the device's library stream). Work and
// cached allocations of that thread then stay on it; used by the batched drivers, one stream per worker.
void bind_thread_stream(hipStream_t s);
// Stream used by an entry point: the caller's if given, else the library stream.
inline hipStream_t pick(void* s) { return s ? static_cast<hipStream_t>(s) : stream(); }
void sync();

// Caching allocator: freed blocks are kept per size class and reused, so ALS / round / ADF loops do not call
// hipMalloc once warm. Blocks are 256-byte aligned (hipMalloc's alignment).
void* dmalloc(size_t bytes);
void dfree(void* p);
void release_cache();
size_t bytes_in_use();
size_t bytes_cached();  // free blocks held by the cache
// Marks the free blocks of stream `from` idle, so that any stream may reuse them. The caller guarantees that
// `from` has finished all its work (it was synchronised); a stream about to be destroyed parks its pool this way.
void park_stream_pool(hipStream_t from);

// Capture arena: while an ArenaScope is active on a thread, dmalloc on that thread hands out fresh memory from the
// arena (bump allocation, never reused) and dfree ignores it. Work recorded into a hipGraph under the scope thus
// keeps every buffer it touched at a fixed address that nothing else gets, for as long as the arena lives.
struct Arena {
	std::vector<std::pair<char*, size_t>> chunks;  // hipMalloc'ed chunks (the last one being filled)
	size_t used = 0;                               // bytes used in the last chunk
	size_t total = 0;                              // bytes handed out
	void* alloc(size_t bytes);
	Arena() = default;
	Arena(const Arena&) = delete;
	Arena& operator=(const Arena&) = delete;
	~Arena();  // synchronises the device, then frees the chunks
};
struct ArenaScope {
	Arena* prev;
	explicit ArenaScope(Arena* a);
	~ArenaScope();
};
bool arena_active();  // an ArenaScope with an arena is active on this thread

// Stream-ordered scratch of the calling thread (per device): bump allocation from chunks kept until the process
// ends. Memory taken while a ScratchFrame lives is handed out again once the frame ends, so a frame may end only
// after every launch that touches its buffers has been enqueued on the thread's stream: launches enqueued later on
// that stream run after them. Frames nest; scratch_alloc outside any frame is an error.
void* scratch_alloc(size_t bytes);
struct ScratchFrame {
	int dev;
	size_t chunk, off;
	ScratchFrame();
	~ScratchFrame();
	ScratchFrame(const ScratchFrame&) = delete;
	ScratchFrame& operator=(const ScratchFrame&) = delete;
};

template <class T>
struct DBuf {  // device buffer of T: owned (pool) or a view of the thread's scratch (scratch())
	T* p = nullptr;
	size_t n = 0;
	bool own = true;
	DBuf() = default;
	explicit DBuf(size_t count) { resize(count); }
	DBuf(const DBuf&) = delete;
	DBuf& operator=(const DBuf&) = delete;
	DBuf(DBuf&& o) noexcept : p(o.p), n(o.n), own(o.own) { o.p = nullptr; o.n = 0; o.own = true; }
	DBuf& operator=(DBuf&& o) noexcept {
		if (this != &o) { reset(); p = o.p; n = o.n; own = o.own; o.p = nullptr; o.n = 0; o.own = true; }
		return *this;
	}
	~DBuf() { reset(); }
	void reset() { if (p && own) dfree(p); p = nullptr; n = 0; own = true; }
	void resize(size_t count) {
		if (count == n && p && own) return;
		reset();
		if (count) { p = static_cast<T*>(dmalloc(count * sizeof(T))); n = count; }
	}
	// count elements from the scratch of the innermost ScratchFrame (valid until that frame ends)
	void scratch(size_t count) {
		reset();
		if (count) { p = static_cast<T*>(scratch_alloc(count * sizeof(T))); n = count; own = false; }
	}
	T* get() const { return p; }
	size_t size() const { return n; }
};

// Dense row-major FP64 tensor resident on the device (xerus::Tensor's dense representation,
// include/xerus/tensor.h:84-123, without the lazy scalar factor).
struct DTensor {
	std::vector<size_t> dims;
	DBuf<double> buf;
	DTensor() = default;
	explicit DTensor(std::vector<size_t> d) : dims(std::move(d)) { buf.resize(count()); }
	// Adopt an owned device buffer holding at least count() elements (the first count() are the tensor).
	DTensor(std::vector<size_t> d, DBuf<double>&& b) : dims(std::move(d)), buf(std::move(b)) {
		if (count() > buf.n) throw Error(ERR_INVALID, "DTensor: adopted buffer too small");
		if (!buf.own) throw Error(ERR_INVALID, "DTensor: a scratch buffer cannot be adopted");
		buf.n = count();
	}
	// A temporary in the thread's scratch (valid until the innermost ScratchFrame ends).
	static DTensor scratch(std::vector<size_t> d) {
		DTensor t;
		t.dims = std::move(d);
		t.buf.scratch(t.count());
		return t;
	}
	size_t count() const { size_t p = 1; for (size_t x : dims) p *= x; return p; }
	size_t size() const { return buf.size(); }
	size_t degree() const { return dims.size(); }
	double* data() const { return buf.get(); }
	void reshape(std::vector<size_t> d);  // same element count
};

}  // namespace xtt
