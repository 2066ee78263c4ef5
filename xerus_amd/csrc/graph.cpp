// Recorded TT calls (TTGraph): stream capture of one call into a hipGraph with its allocations in a capture arena.
#include <atomic>
#include <cstdio>
#include <cstdlib>

#include "tt.hpp"

namespace xtt {

bool graphs_enabled() { return !prof_enabled() && !getenv("XTT_NO_GRAPHS"); }

namespace {
std::atomic<size_t> g_replays[2];
}
void count_replay(int kind) { if (kind >= 0 && kind < 2) g_replays[kind]++; }
void graph_replays(size_t out[2]) { for (int k = 0; k < 2; ++k) out[k] = g_replays[k].load(); }

void TTGraph::reset() {
	if (exec) (void)hipGraphExecDestroy(exec);
	exec = nullptr;
	in.clear();
	out.clear();
	inDims.clear();
	outDims.clear();
	arena.reset();  // after the graph: the arena's destructor waits for the device
}

bool TTGraph::record(const TT& x, const std::function<void(TT&)>& body) {
	reset();
	arena = std::make_unique<Arena>();
	TT xc;
	xc.op = x.op;
	xc.dims = x.dims;
	xc.canonicalized = x.canonicalized;
	xc.corePosition = x.corePosition;
	{
		ArenaScope scope(arena.get());
		for (const DTensor& c : x.cores) {
			xc.cores.emplace_back(c.dims);
			in.push_back(xc.cores.back().data());
			inDims.push_back(c.dims);
		}
	}
	hipGraph_t g = nullptr;
	XTT_HIP(hipStreamBeginCapture(stream(), hipStreamCaptureModeRelaxed));
	bool ok = true;
	why.clear();
	try {
		ArenaScope scope(arena.get());
		body(xc);
	} catch (const std::exception& ex) {
		ok = false;
		why = ex.what();
	} catch (...) {
		ok = false;
		why = "not recordable";
	}
	const hipError_t e = hipStreamEndCapture(stream(), &g);
	if (ok && e != hipSuccess) why = std::string("capture: ") + hipGetErrorString(e);
	if (ok && e == hipSuccess && g) {
		const hipError_t ie = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
		if (ie != hipSuccess) { exec = nullptr; ok = false; }
	} else {
		ok = false;
	}
	if (g) (void)hipGraphDestroy(g);
	if (!ok) {
		(void)hipGetLastError();  // a capture the body invalidated leaves its error behind
		if (getenv("XTT_DEBUG_GRAPHS")) fprintf(stderr, "xtt: recording (kind %d) failed: %s\n", kind, why.c_str());
		for (DTensor& c : xc.cores) { c.buf.p = nullptr; c.buf.n = 0; }  // arena memory: dropped with the arena
		reset();
		return false;
	}
	for (const DTensor& c : xc.cores) {
		out.push_back(c.data());
		outDims.push_back(c.dims);
	}
	canonicalized = xc.canonicalized;
	corePosition = xc.corePosition;
	return true;  // xc's cores live in the arena: their release is a no-op
}

void TTGraph::launch(const TT& x) {
	XTT_REQUIRE(exec && x.order() == in.size(), "recorded call: no graph for this TT");
	MultiCopy mc;
	for (size_t k = 0; k < x.order(); ++k) {
		XTT_REQUIRE(x.cores[k].dims == inDims[k], "recorded call: core shapes differ from the recording");
		mc.src[mc.count] = x.cores[k].data();
		mc.dst[mc.count] = in[k];
		mc.n[mc.count] = x.cores[k].size();
		if (++mc.count == kMultiCopy) { multi_copy_f64(mc, stream()); mc.count = 0; }
	}
	multi_copy_f64(mc, stream());
	XTT_HIP(hipGraphLaunch(exec, stream()));
}

void TTGraph::take(TT& x) const {
	MultiCopy mc;
	for (size_t k = 0; k < out.size(); ++k) {
		DTensor c(outDims[k]);
		mc.src[mc.count] = out[k];
		mc.dst[mc.count] = c.data();
		mc.n[mc.count] = c.size();
		x.cores[k] = std::move(c);
		if (++mc.count == kMultiCopy) { multi_copy_f64(mc, stream()); mc.count = 0; }
	}
	multi_copy_f64(mc, stream());
	x.canonicalized = canonicalized;
	x.corePosition = corePosition;
	count_replay(kind);
}

}  // namespace xtt
