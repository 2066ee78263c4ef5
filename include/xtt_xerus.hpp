/*
 * C++ face of libxtt with xerus's names (header only, over the C ABI of xtt.h), so that code written against
 * xerus's TT API compiles against libxtt with a changed include:
 *   TTTensor / TTOperator            include/xerus/ttNetwork.h:44-58, 352-448 (random, ones, identity, kronecker, dirac,
 *                                    the Tensor constructor, component access, ranks, move_core, round, transpose)
 *   ALSVariant, ALS, ALS_SPD, DMRG, DMRG_SPD   include/xerus/algorithms/als.h:112-223
 *   SinglePointMeasurementSet, RankOneMeasurementSet   include/xerus/measurments.h:43-150
 *   ADFVariant, ADF                  include/xerus/algorithms/adf.h:230-270
 *   PerformanceData, NoPerfData      include/xerus/performanceData.h:43-157
 *   misc::save_to_file / load_from_file   include/xerus/misc/fileIO.h:105-160
 * Every TT lives on the GPU; dense tensors (xerus::Tensor here: dimensions plus row-major entries) live on the host.
 * Errors throw xerus::misc::generic_error with libxtt's message, as xerus's REQUIRE does.
 * Differences from xerus: ALSVariant has no localSolver member (the local systems are solved on the device by CG,
 * with the extra fields precision, localTolerance, maxLocalIterations, innerTolerance, maxRefinements), and the
 * index notation (x(i&0) = A(i/2, j/2)·y(j&0)) is replaced by operator* for A·x and the named functions below.
 */
#ifndef XTT_XERUS_HPP
#define XTT_XERUS_HPP

#include <cstddef>
#include <cstdint>
#include <functional>
#include <limits>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "xtt.h"

namespace xerus {

using value_t = double;
constexpr value_t EPSILON = 8 * std::numeric_limits<value_t>::epsilon();  // include/xerus/basic.h:50

namespace misc {
struct generic_error : std::runtime_error {
	explicit generic_error(const std::string& what) : std::runtime_error(what) {}
};
}  // namespace misc

namespace xtt_detail {
inline void check(int rc) {
	if (rc != XT_OK) throw misc::generic_error("[libxtt error " + std::to_string(rc) + "] " + xt_last_error());
}
inline size_t product(const std::vector<size_t>& d) {
	size_t p = 1;
	for (size_t v : d) p *= v;
	return p;
}
}  // namespace xtt_detail

/// Dense tensor on the host: the dimensions and the row-major (last index fastest) entries of a xerus::Tensor.
struct Tensor {
	std::vector<size_t> dimensions;
	std::vector<value_t> data;
	Tensor() = default;
	explicit Tensor(std::vector<size_t> dims) : dimensions(std::move(dims)), data(xtt_detail::product(dimensions), 0.0) {}
	Tensor(std::vector<size_t> dims, std::vector<value_t> entries) : dimensions(std::move(dims)), data(std::move(entries)) {
		if (data.size() != xtt_detail::product(dimensions)) throw misc::generic_error("Tensor: entry count does not match the dimensions");
	}
	size_t degree() const { return dimensions.size(); }
	size_t size() const { return data.size(); }
	value_t& operator[](size_t i) { return data[i]; }
	value_t operator[](size_t i) const { return data[i]; }
};

template <bool isOperator>
class TTNetwork;
using TTTensor = TTNetwork<false>;
using TTOperator = TTNetwork<true>;

/// PerformanceData (performanceData.h:43-157): a point per local step / half sweep (ALS) or iteration (ADF).
class PerformanceData {
public:
	struct DataPoint {
		size_t iterationCount;
		size_t elapsedTime;  // µs, residual evaluations excluded
		value_t residual;
		value_t error;
		std::vector<size_t> ranks;
		size_t flags;
	};
	bool active;
	bool printProgress;
	std::vector<DataPoint> data;
	std::function<value_t(const TTTensor&)> errorFunction;

	explicit PerformanceData(bool printProgress_ = false, bool active_ = true) : active(active_), printProgress(printProgress_) {}
	explicit PerformanceData(std::function<value_t(const TTTensor&)> errorFunction_, bool printProgress_ = false, bool active_ = true)
	    : active(active_), printProgress(printProgress_), errorFunction(std::move(errorFunction_)) {}
	explicit operator bool() const { return active; }
	void reset() { data.clear(); }
};
inline PerformanceData NoPerfData(false, false);

/// Common part of TTTensor and TTOperator: owns one libxtt handle.
class TTNetworkBase {
protected:
	xt_tt h_ = nullptr;
	explicit TTNetworkBase(xt_tt h) : h_(h) {}
	TTNetworkBase() = default;
	~TTNetworkBase() { if (h_) xt_tt_free(h_); }
	void reset(xt_tt h) { if (h_) xt_tt_free(h_); h_ = h; }
public:
	TTNetworkBase(const TTNetworkBase&) = delete;
	TTNetworkBase& operator=(const TTNetworkBase&) = delete;
	xt_tt handle() const { return h_; }
	size_t order() const { size_t o = 0; xtt_detail::check(xt_tt_order(h_, &o)); return o; }
};

template <bool isOperator>
class TTNetwork : public TTNetworkBase {
	static constexpr size_t N = isOperator ? 2 : 1;
	explicit TTNetwork(xt_tt h) : TTNetworkBase(h) { refresh(); }

public:
	/// re-reads the dimensions and the canonical state after a libxtt call made directly on handle()
	void refresh() {
		const size_t d = order();
		dimensions.assign(d * N, 0);
		if (d) xtt_detail::check(xt_tt_dims(h_, dimensions.data()));
		int c = 0;
		xtt_detail::check(xt_tt_canonicalization(h_, &c, &corePosition));
		canonicalized = c != 0;
	}

	std::vector<size_t> dimensions;  // ttNetwork.h: operator dims are [n_0..n_{d-1}, m_0..m_{d-1}]
	bool canonicalized = true;
	size_t corePosition = 0;

	/// TTNetwork(dims): rank-1 zero TT (ttNetwork.cpp:57-108)
	explicit TTNetwork(const std::vector<size_t>& dims) {
		xt_tt h = nullptr;
		xtt_detail::check(xt_tt_new(isOperator, dims.size(), dims.data(), &h));
		reset(h);
		refresh();
	}
	/// TTNetwork(const Tensor&, eps, maxRank) (ttNetwork.cpp:50-51, 112-160)
	explicit TTNetwork(const Tensor& t, value_t eps = EPSILON, size_t maxRank = std::numeric_limits<size_t>::max())
	    : TTNetwork(t, eps, std::vector<size_t>(t.degree() / N > 0 ? t.degree() / N - 1 : 0, maxRank)) {}
	TTNetwork(const Tensor& t, value_t eps, const std::vector<size_t>& maxRanks) {
		xt_tt h = nullptr;
		xtt_detail::check(xt_tt_from_tensor(isOperator, t.degree(), t.dimensions.data(), t.data.data(), eps, maxRanks.data(), &h));
		reset(h);
		refresh();
	}
	TTNetwork(const TTNetwork& o) : TTNetworkBase() {
		xt_tt h = nullptr;
		xtt_detail::check(xt_tt_copy(o.h_, &h));
		reset(h);
		refresh();
	}
	TTNetwork(TTNetwork&& o) noexcept : TTNetworkBase(), dimensions(std::move(o.dimensions)), canonicalized(o.canonicalized), corePosition(o.corePosition) {
		h_ = o.h_;
		o.h_ = nullptr;
	}
	TTNetwork& operator=(const TTNetwork& o) {
		if (this != &o) { TTNetwork tmp(o); *this = std::move(tmp); }
		return *this;
	}
	TTNetwork& operator=(TTNetwork&& o) noexcept {
		if (this != &o) {
			reset(o.h_);
			o.h_ = nullptr;
			dimensions = std::move(o.dimensions);
			canonicalized = o.canonicalized;
			corePosition = o.corePosition;
		}
		return *this;
	}

	// ---- construction (ttNetwork.h:128-245)
	static TTNetwork random(const std::vector<size_t>& dims, const std::vector<size_t>& ranks) {
		xt_tt h = nullptr;
		xtt_detail::check(xt_tt_random(isOperator, dims.size(), dims.data(), ranks.data(), &h));
		return TTNetwork(h);
	}
	static TTNetwork random(const std::vector<size_t>& dims, size_t rank) {
		return random(dims, std::vector<size_t>(dims.size() / N > 0 ? dims.size() / N - 1 : 0, rank));
	}
	static TTNetwork ones(const std::vector<size_t>& dims) {
		xt_tt h = nullptr;
		xtt_detail::check(xt_tt_ones(isOperator, dims.size(), dims.data(), &h));
		return TTNetwork(h);
	}
	static TTNetwork kronecker(const std::vector<size_t>& dims) {
		xt_tt h = nullptr;
		xtt_detail::check(xt_tt_kronecker(isOperator, dims.size(), dims.data(), &h));
		return TTNetwork(h);
	}
	static TTNetwork dirac(const std::vector<size_t>& dims, const std::vector<size_t>& position) {
		xt_tt h = nullptr;
		xtt_detail::check(xt_tt_dirac(isOperator, dims.size(), dims.data(), position.data(), &h));
		return TTNetwork(h);
	}
	template <bool B = isOperator, typename std::enable_if<B, int>::type = 0>
	static TTNetwork identity(const std::vector<size_t>& dims) {
		xt_tt h = nullptr;
		xtt_detail::check(xt_tt_identity(dims.size(), dims.data(), &h));
		return TTNetwork(h);
	}

	// ---- structure (ttNetwork.cpp:456-492, 716-731)
	size_t degree() const { return dimensions.size(); }
	std::vector<size_t> ranks() const {
		const size_t d = order();
		std::vector<size_t> r(d > 0 ? d - 1 : 0);
		if (!r.empty()) xtt_detail::check(xt_tt_ranks(h_, r.data()));
		return r;
	}
	size_t rank(size_t i) const { return ranks().at(i); }
	Tensor get_component(size_t k) const {
		std::vector<size_t> dims(N + 2);
		xtt_detail::check(xt_tt_component_dims(h_, k, dims.data()));
		Tensor t(dims);
		xtt_detail::check(xt_tt_get_component(h_, k, t.data.data()));
		return t;
	}
	void set_component(size_t k, const Tensor& t) {
		xtt_detail::check(xt_tt_set_component(h_, k, t.degree(), t.dimensions.data(), t.data.data()));
		refresh();
	}
	void assume_core_position(size_t pos) { xtt_detail::check(xt_tt_assume_core_position(h_, pos)); refresh(); }

	// ---- canonicalisation and rounding (ttNetwork.cpp:581-684)
	void move_core(size_t position, bool keepRank = false) { xtt_detail::check(xt_tt_move_core(h_, position, keepRank)); refresh(); }
	void canonicalize_left() { move_core(0); }
	void canonicalize_right() { move_core(order() - 1); }
	void round(const std::vector<size_t>& maxRanks, value_t eps = EPSILON) {
		xtt_detail::check(xt_tt_round(h_, maxRanks.data(), eps));
		refresh();
	}
	void round(size_t maxRank) { round(std::vector<size_t>(order() > 0 ? order() - 1 : 0, maxRank), EPSILON); }
	void round(int maxRank) {
		if (maxRank <= 0) throw misc::generic_error("MaxRank must be positive");
		round(static_cast<size_t>(maxRank));
	}
	void round(value_t eps) { round(std::vector<size_t>(order() > 0 ? order() - 1 : 0, std::numeric_limits<size_t>::max()), eps); }
	template <bool B = isOperator, typename std::enable_if<B, int>::type = 0>
	void transpose() { xtt_detail::check(xt_tt_transpose(h_)); refresh(); }

	// ---- arithmetic (ttNetwork.cpp:781-847)
	value_t frob_norm() const { value_t v = 0; xtt_detail::check(xt_tt_frob_norm(h_, &v)); return v; }
	TTNetwork& operator+=(const TTNetwork& o) {
		xt_tt h = nullptr;
		xtt_detail::check(xt_tt_add(h_, o.h_, &h));
		reset(h);
		refresh();
		return *this;
	}
	TTNetwork& operator*=(value_t s) { xtt_detail::check(xt_tt_scale(h_, s)); return *this; }
	TTNetwork& operator/=(value_t s) { return *this *= 1.0 / s; }
	TTNetwork& operator-=(const TTNetwork& o) { TTNetwork n(o); n *= -1.0; return *this += n; }
	friend TTNetwork operator+(TTNetwork a, const TTNetwork& b) { a += b; return a; }
	friend TTNetwork operator-(TTNetwork a, const TTNetwork& b) { a -= b; return a; }
	friend TTNetwork operator*(TTNetwork a, value_t s) { a *= s; return a; }
	friend TTNetwork operator*(value_t s, TTNetwork a) { a *= s; return a; }
	friend TTNetwork operator/(TTNetwork a, value_t s) { a /= s; return a; }
	/// the dense tensor (small TTs only); an operator's entries are ordered (n…, m…)
	Tensor full() const {
		Tensor t(dimensions);
		xtt_detail::check(xt_tt_full(h_, t.data.data()));
		return t;
	}

	template <bool> friend class TTNetwork;
	friend TTNetwork<false> operator*(const TTNetwork<true>& A, const TTNetwork<false>& x);
	template <bool B> friend TTNetwork<B> tt_load_from_file(const std::string& filename);
};

/// A·x as a TT of rank r_A·r_x (TTStack contraction, ttStack.cpp:197-310), not canonicalized.
inline TTTensor operator*(const TTOperator& A, const TTTensor& x) {
	xt_tt h = nullptr;
	xtt_detail::check(xt_tt_apply(A.handle(), x.handle(), &h));
	return TTTensor(h);
}
template <bool B>
inline value_t frob_norm(const TTNetwork<B>& x) { return x.frob_norm(); }

namespace misc {
enum class FileFormat { BINARY = 0, TSV = 1 };
template <bool B>
inline void save_to_file(const TTNetwork<B>& x, const std::string& filename, FileFormat format = FileFormat::BINARY) {
	xtt_detail::check(xt_tt_save(x.handle(), filename.c_str(), static_cast<int>(format)));
}
}  // namespace misc
template <bool B>
inline TTNetwork<B> tt_load_from_file(const std::string& filename) {
	xt_tt h = nullptr;
	xtt_detail::check(xt_tt_load(filename.c_str(), &h));
	int op = 0;
	xtt_detail::check(xt_tt_is_operator(h, &op));
	if ((op != 0) != B) { xt_tt_free(h); throw misc::generic_error("load_from_file: the file holds the other TT kind"); }
	return TTNetwork<B>(h);
}
namespace misc {
template <class T>
inline T load_from_file(const std::string& filename) { return tt_load_from_file<std::is_same<T, TTOperator>::value>(filename); }
}  // namespace misc

namespace xtt_detail {
// Records PerformanceData points through libxtt's per-thread sink while `run` executes.
template <class F>
void with_perf(PerformanceData& p, const TTTensor* x, F&& run) {
	if (!p.active) { run(); return; }
	struct Ctx { PerformanceData* p; const TTTensor* x; };
	Ctx ctx{&p, x};
	auto cb = [](void* user, size_t it, double us, double res, const size_t* ranks, size_t nranks, size_t flags) {
		Ctx* c = static_cast<Ctx*>(user);
		const value_t err = c->p->errorFunction && c->x ? c->p->errorFunction(*c->x) : 0.0;
		c->p->data.push_back({it, static_cast<size_t>(us), res, err, std::vector<size_t>(ranks, ranks + nranks), flags});
	};
	check(xt_perf_sink(cb, &ctx));
	try {
		run();
	} catch (...) {
		xt_perf_sink(nullptr, nullptr);
		throw;
	}
	check(xt_perf_sink(nullptr, nullptr));
}
}  // namespace xtt_detail

/// ALSVariant (als.h:112-207). x is updated in place; the return value is the energy functional (als.cpp:773).
class ALSVariant {
	value_t solve(const TTOperator* A, TTTensor& x, const TTTensor& b, size_t halfSweeps, value_t eps, PerformanceData& perf) const {
		xt_als_options o{};
		xtt_detail::check(xt_als_default_options(&o, sites, assumeSPD));
		o.numHalfSweeps = halfSweeps;
		o.convergenceEpsilon = eps;
		o.useResidualForEndCriterion = useResidualForEndCriterion;
		o.preserveCorePosition = preserveCorePosition;
		o.precision = precision;
		o.localTolerance = localTolerance;
		o.maxLocalIterations = maxLocalIterations;
		o.innerTolerance = innerTolerance;
		o.maxRefinements = maxRefinements;
		value_t energy = 0;
		xtt_detail::with_perf(perf, &x, [&] { xtt_detail::check(xt_als_solve(&o, A ? A->handle() : nullptr, x.handle(), b.handle(), &energy, nullptr)); });
		x.refresh();
		return energy;
	}

public:
	static constexpr size_t FLAG_FINISHED_HALFSWEEP = 1;
	static constexpr size_t FLAG_FINISHED_FULLSWEEP = 3;
	unsigned sites;
	size_t numHalfSweeps;
	value_t convergenceEpsilon = 1e-6;
	bool useResidualForEndCriterion;
	bool preserveCorePosition = true;
	bool assumeSPD;
	// device local solve
	int precision = XT_F64;
	value_t localTolerance = 1e-13;
	int maxLocalIterations = 2000;
	value_t innerTolerance = 0;
	int maxRefinements = 30;

	ALSVariant(unsigned sites_, size_t numHalfSweeps_, bool assumeSPD_, bool useResidual = false)
	    : sites(sites_), numHalfSweeps(numHalfSweeps_), useResidualForEndCriterion(useResidual), assumeSPD(assumeSPD_) {
		if (sites_ == 0) throw misc::generic_error("ALSVariant: sites must be > 0");
	}

	value_t operator()(const TTOperator& A, TTTensor& x, const TTTensor& b, value_t eps, PerformanceData& perf = NoPerfData) const {
		return solve(&A, x, b, numHalfSweeps, eps, perf);
	}
	value_t operator()(const TTOperator& A, TTTensor& x, const TTTensor& b, size_t halfSweeps, PerformanceData& perf = NoPerfData) const {
		return solve(&A, x, b, halfSweeps, convergenceEpsilon, perf);
	}
	value_t operator()(const TTOperator& A, TTTensor& x, const TTTensor& b, PerformanceData& perf = NoPerfData) const {
		return solve(&A, x, b, numHalfSweeps, convergenceEpsilon, perf);
	}
	value_t operator()(TTTensor& x, const TTTensor& b, value_t eps, PerformanceData& perf = NoPerfData) const {
		return solve(nullptr, x, b, numHalfSweeps, eps, perf);
	}
	value_t operator()(TTTensor& x, const TTTensor& b, size_t halfSweeps, PerformanceData& perf = NoPerfData) const {
		return solve(nullptr, x, b, halfSweeps, convergenceEpsilon, perf);
	}
	value_t operator()(TTTensor& x, const TTTensor& b, PerformanceData& perf = NoPerfData) const {
		return solve(nullptr, x, b, numHalfSweeps, convergenceEpsilon, perf);
	}
};
// the presets of als.cpp:781-788
inline const ALSVariant ALS(1, 0, false);
inline const ALSVariant ALS_SPD(1, 0, true);
inline const ALSVariant DMRG(2, 0, false);
inline const ALSVariant DMRG_SPD(2, 0, true);

/// SinglePointMeasurementSet (measurments.h:43-91)
class SinglePointMeasurementSet {
	xt_meas h_ = nullptr;
	size_t d_ = 0;

public:
	SinglePointMeasurementSet() = default;
	SinglePointMeasurementSet(const SinglePointMeasurementSet&) = delete;
	SinglePointMeasurementSet& operator=(const SinglePointMeasurementSet&) = delete;
	SinglePointMeasurementSet(SinglePointMeasurementSet&& o) noexcept : h_(o.h_), d_(o.d_) { o.h_ = nullptr; }
	~SinglePointMeasurementSet() { if (h_) xt_meas_free(h_); }
	/// random(P, dims) with the shared engine (measurments.cpp:222-248): sorted, without duplicates
	static SinglePointMeasurementSet random(size_t P, const std::vector<size_t>& dims) {
		SinglePointMeasurementSet m;
		xtt_detail::check(xt_meas_random(P, dims.size(), dims.data(), &m.h_));
		m.d_ = dims.size();
		return m;
	}
	/// positions (P × d, row-major) and values
	SinglePointMeasurementSet(size_t d, const std::vector<uint32_t>& positions, const std::vector<value_t>& values) : d_(d) {
		xtt_detail::check(xt_meas_new(values.size(), d, positions.data(), values.data(), &h_));
	}
	size_t size() const { size_t P = 0; xtt_detail::check(xt_meas_size(h_, &P)); return P; }
	size_t degree() const { return d_; }
	std::vector<uint32_t> positions() const {
		std::vector<uint32_t> p(size() * d_);
		xtt_detail::check(xt_meas_positions(h_, p.data()));
		return p;
	}
	std::vector<value_t> measuredValues() const {
		std::vector<value_t> v(size());
		xtt_detail::check(xt_meas_values(h_, v.data()));
		return v;
	}
	void set_measuredValues(const std::vector<value_t>& v) { xtt_detail::check(xt_meas_set_values(h_, v.data())); }
	void measure(const TTTensor& x) { xtt_detail::check(xt_meas_measure(h_, x.handle())); }
	xt_meas handle() const { return h_; }
};

/// RankOneMeasurementSet (measurments.h:94-150)
class RankOneMeasurementSet {
	xt_r1meas h_ = nullptr;

public:
	RankOneMeasurementSet(const RankOneMeasurementSet&) = delete;
	RankOneMeasurementSet& operator=(const RankOneMeasurementSet&) = delete;
	RankOneMeasurementSet(RankOneMeasurementSet&& o) noexcept : h_(o.h_) { o.h_ = nullptr; }
	~RankOneMeasurementSet() { if (h_) xt_r1meas_free(h_); }
	/// RankOneMeasurementSet(const SinglePointMeasurementSet&, dims) (measurments.cpp:253-265)
	RankOneMeasurementSet(const SinglePointMeasurementSet& sp, const std::vector<size_t>& dims) {
		xtt_detail::check(xt_r1meas_from_single_point(sp.handle(), dims.size(), dims.data(), &h_));
	}
	explicit RankOneMeasurementSet(xt_r1meas h) : h_(h) {}
	static RankOneMeasurementSet random(size_t P, const std::vector<size_t>& dims) {
		xt_r1meas h = nullptr;
		xtt_detail::check(xt_r1meas_random(P, dims.size(), dims.data(), &h));
		return RankOneMeasurementSet(h);
	}
	size_t size() const { size_t P = 0; xtt_detail::check(xt_r1meas_size(h_, &P)); return P; }
	void measure(const TTTensor& x) { xtt_detail::check(xt_r1meas_measure(h_, x.handle())); }
	value_t test(const TTTensor& x) const { value_t v = 0; xtt_detail::check(xt_r1meas_test(h_, x.handle(), &v)); return v; }
	xt_r1meas handle() const { return h_; }
};

/// ADFVariant (adf.h:230-266): returns ‖P_Ω(x) − b‖/‖b‖; x is updated in place.
class ADFVariant {
	value_t run(TTTensor& x, xt_meas sp, xt_r1meas r1, const std::vector<size_t>* maxRanks, PerformanceData& perf) const {
		value_t res = 0;
		size_t it = 0;
		xtt_detail::with_perf(perf, &x, [&] {
			if (sp)
				xtt_detail::check(xt_adf_solve(x.handle(), sp, maxRanks ? maxRanks->data() : nullptr, maxIterations, targetResidualNorm,
				                               minimalResidualNormDecrease, precision, &res, &it));
			else
				xtt_detail::check(xt_adf_solve_r1(x.handle(), r1, maxRanks ? maxRanks->data() : nullptr, maxIterations, targetResidualNorm,
				                                  minimalResidualNormDecrease, precision, &res, &it));
		});
		x.refresh();
		return res;
	}

public:
	size_t maxIterations;
	value_t targetResidualNorm;
	value_t minimalResidualNormDecrease;
	int precision = XT_F64;  // FP64 or FP32 sampled-entry kernels (cores and norms stay FP64)
	ADFVariant(size_t maxIteration, value_t targetResidual, value_t minimalResidualDecrease)
	    : maxIterations(maxIteration), targetResidualNorm(targetResidual), minimalResidualNormDecrease(minimalResidualDecrease) {}
	value_t operator()(TTTensor& x, const SinglePointMeasurementSet& m, PerformanceData& perf = NoPerfData) const {
		return run(x, m.handle(), nullptr, nullptr, perf);
	}
	value_t operator()(TTTensor& x, const SinglePointMeasurementSet& m, const std::vector<size_t>& maxRanks,
	                   PerformanceData& perf = NoPerfData) const {
		return run(x, m.handle(), nullptr, &maxRanks, perf);
	}
	value_t operator()(TTTensor& x, const RankOneMeasurementSet& m, PerformanceData& perf = NoPerfData) const {
		return run(x, nullptr, m.handle(), nullptr, perf);
	}
	value_t operator()(TTTensor& x, const RankOneMeasurementSet& m, const std::vector<size_t>& maxRanks,
	                   PerformanceData& perf = NoPerfData) const {
		return run(x, nullptr, m.handle(), &maxRanks, perf);
	}
};
inline const ADFVariant ADF(0, 1e-8, 0.999);  // adf.cpp:610

}  // namespace xerus

#endif  // XTT_XERUS_HPP
