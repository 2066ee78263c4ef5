// SinglePointMeasurementSet (src/xerus/measurments.cpp) and ADFVariant (src/xerus/algorithms/adf.cpp) on
// the device. Positions live in HBM as uint32 (P×d, lexicographically sorted like xerus); TT cores are
// flattened per step into one buffer of the working precision (FP64 for parity runs, FP32 for the
// throughput configuration) and every sampled-entry contraction runs in the ADF kernels.
#include "adf_driver.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <set>

namespace xtt {

Measurements meas_random(size_t P, const dims_t& dims) {  // measurments.cpp:222-248
	size_t total = 1;
	bool overflow = false;
	for (size_t n : dims) {
		XTT_REQUIRE(n > 0, "dimension 0 is not valid");
		if (total > SIZE_MAX / n) overflow = true; else total *= n;
	}
	XTT_REQUIRE(overflow || total >= P, "It's impossible to perform as many measurements as requested.");
	std::set<size_t> seen;
	std::vector<std::vector<size_t>> positions;
	std::vector<size_t> idx(dims.size());
	while (positions.size() < P) {
		size_t pos = 0;
		for (size_t i = 0; i < dims.size(); ++i) { idx[i] = rng_uniform(dims[i] - 1); pos *= dims[i]; pos += idx[i]; }
		if (seen.insert(pos).second) positions.push_back(idx);
	}
	std::sort(positions.begin(), positions.end());  // measurments.cpp:89-105
	Measurements m;
	m.d = dims.size();
	m.pos.resize(P * m.d);
	for (size_t i = 0; i < P; ++i) for (size_t k = 0; k < m.d; ++k) m.pos[i * m.d + k] = (uint32_t)positions[i][k];
	m.values.assign(P, 0.0);
	m.upload();
	return m;
}

void Measurements::check_against(const TT& x) const {
	XTT_REQUIRE(d == x.order(), "Measurment degree must coincide with x degree.");
	for (size_t i = 0; i < pos.size(); ++i) XTT_REQUIRE(pos[i] < x.dims[i % d], "measurement position outside the tensor");
}

void Measurements::upload() {
	plan.reset();
	dpos.resize(pos.size());
	if (!pos.empty()) XTT_HIP(hipMemcpyAsync(dpos.get(), pos.data(), pos.size() * 4, hipMemcpyHostToDevice, stream()));
	sync();
}

const AdfPlan& Measurements::stack_plan() const {  // adf.cpp:102-191, once per position set
	if (!plan) {
		XTT_REQUIRE(size() > 0, "empty measurement set");
		XTT_REQUIRE(size() <= (size_t)INT32_MAX, "ADF stacks: more than 2^31 - 1 measurements");
		std::vector<uint32_t> maxPos(d, 0);
		for (size_t i = 0; i < pos.size(); ++i) maxPos[i % d] = std::max(maxPos[i % d], pos[i]);
		auto p = std::make_shared<AdfPlan>();
		adf_plan_build(*p, dpos.get(), (int)size(), (int)d, maxPos, stream());
		plan = std::move(p);
	}
	return *plan;
}

template <class T>
void Flat<T>::build(const TT& x) {
	require_correct_format(x);  // strides below come from the rank chain
	const size_t d = x.order();
	std::vector<int> n(d), rank(d + 1);
	std::vector<long long> off(d);
	long long o = 0;
	int maxR = 1;
	rank[0] = 1;
	for (size_t k = 0; k < d; ++k) {
		const DTensor& c = x.cores[k];
		XTT_REQUIRE(c.dims[0] <= 16 && c.dims[2] <= 16, "ADF: ranks above 16 are not supported");
		n[k] = (int)c.dims[1];
		rank[k + 1] = (int)c.dims[2];
		off[k] = o;
		o += (long long)c.size();
		maxR = std::max(maxR, (int)std::max(c.dims[0], c.dims[2]));
	}
	if (dn.size() < d) { dn.resize(d); drank.resize(d + 1); doff.resize(d); }
	if (cores.size() < (size_t)o) cores.resize(o);
	hn = n; hrank = rank; hoff = off;
	XTT_HIP(hipMemcpyAsync(dn.get(), n.data(), d * 4, hipMemcpyHostToDevice, stream()));
	XTT_HIP(hipMemcpyAsync(drank.get(), rank.data(), (d + 1) * 4, hipMemcpyHostToDevice, stream()));
	XTT_HIP(hipMemcpyAsync(doff.get(), off.data(), d * 8, hipMemcpyHostToDevice, stream()));
	desc.d = (int)d;
	desc.n = dn.get();
	desc.rank = drank.get();
	desc.coreOff = doff.get();
	desc.maxR = maxR;
	desc.gradSize = (int)o;
	desc.sumN = 0;
	desc.maxN = 0;
	for (int v : n) { desc.sumN += v; desc.maxN = std::max(desc.maxN, v); }
	refresh(x);
	sync();  // the host arrays above are locals
}

template <class T>
bool Flat<T>::matches(const TT& x) const {
	if (hn.size() != x.order()) return false;
	for (size_t k = 0; k < x.order(); ++k)
		if (hn[k] != (int)x.cores[k].dims[1] || hrank[k] != (int)x.cores[k].dims[0] || hrank[k + 1] != (int)x.cores[k].dims[2]) return false;
	return true;
}

template <class T>
void Flat<T>::refresh(const TT& x, size_t first, size_t last) {  // one launch per kMultiCopy cores
	MultiCopy mc;
	auto flush = [&] {
		if constexpr (sizeof(T) == 8) multi_copy_f64(mc, stream()); else multi_convert_f64_f32(mc, stream());
		mc.count = 0;
	};
	for (size_t k = first; k < x.order() && k <= last; ++k) {
		mc.src[mc.count] = x.cores[k].data();
		mc.dst[mc.count] = reinterpret_cast<double*>(cores.get() + hoff[k]);
		mc.n[mc.count] = x.cores[k].size();
		if (++mc.count == kMultiCopy) flush();
	}
	flush();
}

template struct Flat<double>;
template struct Flat<float>;

namespace {
template <class T>
struct AdfWork {
	DBuf<T> vals, resid, grad, partials;
	DBuf<double> norms, dpart, slices;
	DBuf<unsigned> counter;
	void ensure(size_t P, size_t gradSize, size_t nslice) {
		if (resid.size() < P) { resid.resize(P); vals.resize(P); }
		if (grad.size() < gradSize) grad.resize(gradSize);
		const size_t pb = (size_t)std::max(adf_grad_blocks((int)P), adf_stack_blocks((int)P)) * gradSize;
		if (partials.size() < pb) partials.resize(pb);
		if (!norms.get()) { norms.resize(2); counter.resize(1); XTT_HIP(hipMemsetAsync(counter.get(), 0, 4, stream())); }
		const size_t sp = std::max<size_t>((size_t)std::max(adf_slice_blocks((int)P), adf_stack_blocks((int)P)) * std::max<size_t>(nslice, 1),
		                                   2 * 1024);
		if (dpart.size() < sp) dpart.resize(sp);
		if (slices.size() < nslice) slices.resize(std::max<size_t>(nslice, 1));
	}
};

template <class T>
void upload_values(const std::vector<double>& values, DBuf<T>& v) {
	std::vector<T> h(values.begin(), values.end());
	v.resize(h.size());
	XTT_HIP(hipMemcpyAsync(v.get(), h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, stream()));
	sync();
}
}  // namespace

void meas_measure(Measurements& m, const TT& x) {  // measurments.cpp:125-153 (FP64)
	XTT_REQUIRE(m.d == x.order(), "Degrees of solution and measurements must match!");
	m.check_against(x);
	const size_t P = m.size();
	if (P == 0) return;
	Flat<double> f;
	f.build(x);
	AdfWork<double> w;
	w.ensure(P, f.desc.gradSize, 1);
	XTT_HIP(hipMemsetAsync(w.vals.get(), 0, P * 8, stream()));
	AdfDesc a = f.desc;
	a.P = (int)P;
	adf_residual<double>(a, f.cores.get(), m.dpos.get(), w.vals.get(), w.resid.get(), w.norms.get(), w.dpart.get(), w.counter.get(), stream());
	XTT_HIP(hipMemcpyAsync(m.values.data(), w.resid.get(), P * 8, hipMemcpyDeviceToHost, stream()));
	sync();
}

// ---- RankOneMeasurementSet ------------------------------------------------------------------------------

size_t RankOneMeasurements::sumN() const {
	size_t s = 0;
	for (size_t n : dims) s += n;
	return s;
}

void RankOneMeasurements::check_against(const TT& x) const {
	XTT_REQUIRE(d() == x.order(), "Measurment degree must coincide with x degree.");
	for (size_t k = 0; k < d(); ++k) XTT_REQUIRE(dims[k] == x.dims[k], "rank-one measurement dimensions differ from x's");
	XTT_REQUIRE(u.size() == size() * sumN(), "rank-one measurement set is inconsistent");
}

template <>
const double* RankOneMeasurements::device_u<double>() const {
	if (dirty) {
		du64.resize(std::max<size_t>(u.size(), 1));
		if (!u.empty()) XTT_HIP(hipMemcpyAsync(du64.get(), u.data(), u.size() * 8, hipMemcpyHostToDevice, stream()));
		sync();
		dirty = false;
		dirty32 = true;
	}
	return du64.get();
}
template <>
const float* RankOneMeasurements::device_u<float>() const {
	const double* d64 = device_u<double>();
	if (dirty32) {
		du32.resize(std::max<size_t>(u.size(), 1));
		if (!u.empty()) convert_f64_f32(d64, du32.get(), u.size(), stream());
		dirty32 = false;
	}
	return du32.get();
}

// xerus's internal::comp on dense vectors (measurments.cpp:549-558): the first differing entry decides, and the
// comparator of sort() puts the vector with the LARGER entry first.
static bool r1_before(const double* a, const double* b, size_t len) {
	for (size_t k = 0; k < len; ++k) {
		if (a[k] < b[k]) return false;
		if (a[k] > b[k]) return true;
	}
	return false;
}

void r1_sort(RankOneMeasurements& m, bool positionsOnly) {
	const size_t P = m.size(), S = m.sumN();
	std::vector<size_t> order(P);
	for (size_t i = 0; i < P; ++i) order[i] = i;
	// mode-by-mode comparison equals comparing the concatenated rows (all vectors of a mode have one length)
	std::sort(order.begin(), order.end(), [&](size_t a, size_t b) { return r1_before(&m.u[a * S], &m.u[b * S], S); });
	std::vector<double> u(P * S), v(P);
	for (size_t i = 0; i < P; ++i) {
		std::copy(&m.u[order[i] * S], &m.u[order[i] * S] + S, &u[i * S]);
		v[i] = m.values[positionsOnly ? i : order[i]];
	}
	m.u.swap(u);
	m.values.swap(v);
	m.invalidate();
}

RankOneMeasurements r1_random(size_t P, const dims_t& dims) {  // create_random_positions, measurments.cpp:513-536
	size_t total = 1;
	bool overflow = false;
	for (size_t n : dims) {
		XTT_REQUIRE(n > 0, "dimension 0 is not valid");
		if (total > SIZE_MAX / n) overflow = true; else total *= n;
	}
	XTT_REQUIRE(overflow || total >= P, "It's impossible to perform as many measurements as requested.");
	RankOneMeasurements m;
	m.dims = dims;
	const size_t S = m.sumN();
	m.u.resize(P * S);
	for (size_t i = 0; i < P * S; ++i) m.u[i] = rng_normal();  // Tensor::random({n_k}) per mode, measurement-major
	m.values.assign(P, 0.0);
	r1_sort(m, true);
	return m;
}

RankOneMeasurements r1_from_single_point(const Measurements& sp, const dims_t& dims) {
	XTT_REQUIRE(sp.d == dims.size(), "Inconsistent degrees.");
	RankOneMeasurements m;
	m.dims = dims;
	const size_t S = m.sumN(), P = sp.size();
	m.u.assign(P * S, 0.0);
	for (size_t i = 0; i < P; ++i) {
		size_t o = 0;
		for (size_t k = 0; k < dims.size(); ++k) {
			XTT_REQUIRE(sp.pos[i * sp.d + k] < dims[k], "measurement position outside the dimensions");
			m.u[i * S + o + sp.pos[i * sp.d + k]] = 1.0;
			o += dims[k];
		}
	}
	m.values = sp.values;
	return m;
}

void r1_normalize(RankOneMeasurements& m) {
	const size_t S = m.sumN();
	for (size_t i = 0; i < m.size(); ++i) {
		size_t o = 0;
		for (size_t k = 0; k < m.d(); ++k) {
			double nrm = 0;
			for (size_t j = 0; j < m.dims[k]; ++j) nrm += m.u[i * S + o + j] * m.u[i * S + o + j];
			nrm = std::sqrt(nrm);
			for (size_t j = 0; j < m.dims[k]; ++j) m.u[i * S + o + j] /= nrm;
			m.values[i] /= nrm;
			o += m.dims[k];
		}
	}
	m.invalidate();
}

namespace {
// Σ (x(u_i) − b_i)² and Σ b_i² in FP64; the values x(u_i) go to vals when given.
void r1_eval(const RankOneMeasurements& m, const TT& x, double* norms2, std::vector<double>* vals) {
	m.check_against(x);
	const size_t P = m.size();
	Flat<double> f;
	f.build(x);
	AdfDesc a = f.desc;
	a.P = (int)P;
	AdfWork<double> w;
	w.ensure(P, f.desc.gradSize, 1);
	DBuf<double> b(std::max<size_t>(P, 1));
	XTT_HIP(hipMemcpyAsync(b.get(), m.values.data(), P * 8, hipMemcpyHostToDevice, stream()));
	r1_residual<double>(a, f.cores.get(), m.device_u<double>(), vals ? nullptr : b.get(), w.resid.get(), w.norms.get(), w.dpart.get(), -1,
	                    nullptr, stream());
	XTT_HIP(hipMemcpyAsync(norms2, w.norms.get(), 16, hipMemcpyDeviceToHost, stream()));
	if (vals) {
		vals->resize(P);
		XTT_HIP(hipMemcpyAsync(vals->data(), w.resid.get(), P * 8, hipMemcpyDeviceToHost, stream()));
	}
	sync();
}
}  // namespace

void r1_measure(RankOneMeasurements& m, const TT& x) {
	if (m.size() == 0) return;
	double n2[2];
	std::vector<double> v;
	r1_eval(m, x, n2, &v);
	m.values = v;
}

double r1_test(const RankOneMeasurements& m, const TT& x) {
	XTT_REQUIRE(m.size() > 0, "empty measurement set");
	double n2[2];
	r1_eval(m, x, n2, nullptr);
	return std::sqrt(n2[0] / n2[1]);
}

// ---- gradient and solver, shared by both measurement kinds ----------------------------------------------------

namespace {
// The per-kind device work of one ADF step. Single points: slice-wise steps (adf.cpp:468-478); rank-one: one
// step for the whole core (adf.cpp:481-487).
template <class T>
struct SinglePointOps {
	const Measurements& m;
	void check(const TT& x) const { m.check_against(x); }
	size_t size() const { return m.size(); }
	const std::vector<double>& values() const { return m.values; }
	void residual(const AdfDesc& a, const Flat<T>& f, AdfWork<T>& w) const {
		adf_residual<T>(a, f.cores.get(), m.dpos.get(), w.vals.get(), w.resid.get(), w.norms.get(), w.dpart.get(), w.counter.get(),
		                stream());
	}
	void gradient(const AdfDesc& a, const Flat<T>& f, AdfWork<T>& w, int k) const {
		adf_gradient<T>(a, f.cores.get(), m.dpos.get(), w.resid.get(), w.grad.get(), w.partials.get(), k, stream());
	}
	// update_x (adf.cpp:468-478) on the device: x_k[:, j, :] += (‖Δ_j‖²/‖P_j A Δ‖²)·Δ_j with Δ = −(the gradient)
	void step(const AdfDesc& a, const Flat<T>& f, AdfWork<T>& w, size_t k, DTensor& core) const {
		const int rl = (int)core.dims[0], n = (int)core.dims[1], rr = (int)core.dims[2];
		adf_slice_norms<T>(a, f.cores.get(), m.dpos.get(), (int)k, w.grad.get() + f.hoff[k], n, w.slices.get(), w.dpart.get(), stream());
		adf_update_core<T>(core.data(), w.grad.get() + f.hoff[k], w.slices.get(), rl, n, rr, true, stream());
	}
};

// Single points with xerus's stacks (adf.cpp:102-312): backward stacks rebuilt at the start of an iteration,
// forward level k−1 refreshed after each core move; the residual, gradient and slice norms at core k read F_{k-1}
// and B_{k+1} through the plan's slots.
template <class T>
struct StackedSinglePointOps {
	const Measurements& m;
	const AdfPlan& plan;
	DBuf<T> F, B;                       // stack values, level after level
	std::vector<long long> fOff, bOff;  // first value of each level
	std::vector<int> rk;                // ranks the layout was made for
	void check(const TT& x) const { m.check_against(x); }
	size_t size() const { return m.size(); }
	const std::vector<double>& values() const { return m.values; }
	// forward level k holds count × r_{k+1} values, backward level k count × r_k
	void layout(const Flat<T>& f) {
		if (rk == f.hrank) return;
		rk = f.hrank;
		const size_t d = f.hn.size();
		fOff.assign(d, 0);
		bOff.assign(d, 0);
		long long fo = 0, bo = 0;
		for (size_t k = 0; k < d; ++k) {
			fOff[k] = fo;
			bOff[k] = bo;
			fo += plan.side[0].count[k] * f.hrank[k + 1];
			bo += plan.side[1].count[k] * f.hrank[k];
		}
		if (F.size() < (size_t)fo) F.resize((size_t)fo);
		if (B.size() < (size_t)bo) B.resize((size_t)bo);
	}
	const uint32_t* col(size_t k) const { return plan.posT.get() + k * m.size(); }  // positions of mode k
	AdfSides<T> sides(size_t k) const {
		const size_t d = fOff.size(), P = m.size();
		AdfSides<T> sd{};
		if (k > 0) { sd.fSlot = plan.side[0].slot.get() + (k - 1) * P; sd.F = F.get() + fOff[k - 1]; }
		if (k + 1 < d) { sd.bSlot = plan.side[1].slot.get() + (k + 1) * P; sd.B = B.get() + bOff[k + 1]; }
		return sd;
	}
	void forward(const Flat<T>& f, size_t k) {  // update_forward_stack (adf.cpp:255-270)
		const int P = (int)m.size(), d = (int)fOff.size();
		adf_forward_stack<T>(f.cores.get() + f.hoff[k], f.hrank[k], f.hn[k], f.hrank[k + 1], (int)k, d, col(k), P,
		                     plan.side[0].rep.get() + plan.side[0].first[k], plan.side[0].count[k],
		                     k > 0 ? plan.side[0].slot.get() + (k - 1) * (size_t)P : nullptr, k > 0 ? F.get() + fOff[k - 1] : nullptr,
		                     F.get() + fOff[k], stream());
	}
	void backward(const Flat<T>& f, size_t k) {  // update_backward_stack (adf.cpp:217-232)
		const int P = (int)m.size(), d = (int)fOff.size();
		const bool last = (int)k + 1 == d;
		adf_backward_stack<T>(f.cores.get() + f.hoff[k], f.hrank[k], f.hn[k], f.hrank[k + 1], (int)k, d, col(k), P,
		                      plan.side[1].rep.get() + plan.side[1].first[k], plan.side[1].count[k],
		                      last ? nullptr : plan.side[1].slot.get() + (k + 1) * (size_t)P, last ? nullptr : B.get() + bOff[k + 1],
		                      B.get() + bOff[k], stream());
	}
	// calculate_residual + calculate_projected_gradient (adf.cpp:290-396) in one pass: resid, norms, and the gradient of
	// core k in w.grad
	// (norms only when wanted: the solve reads them at core 0)
	void residual_gradient(const Flat<T>& f, AdfWork<T>& w, size_t k, bool norms) const {
		adf_stack_residual_gradient<T>(f.cores.get() + f.hoff[k], f.hrank[k], f.hn[k], f.hrank[k + 1], (int)k, (int)fOff.size(), col(k),
		                               (int)m.size(), sides(k), w.vals.get(), w.resid.get(), norms ? w.norms.get() : nullptr, w.grad.get(),
		                               w.partials.get(), w.dpart.get(), stream());
	}
	// update_x (adf.cpp:468-478) with the slice-wise norms of adf.cpp:413-465
	void step(AdfWork<T>& w, size_t k, DTensor& core) const {
		adf_stack_step<T>(core.data(), w.grad.get(), (int)core.dims[0], (int)core.dims[1], (int)core.dims[2], (int)k, (int)fOff.size(), col(k),
		                  (int)m.size(), sides(k), w.dpart.get(), stream());
	}
};

template <class T>
struct RankOneOps {
	const RankOneMeasurements& m;
	void check(const TT& x) const { m.check_against(x); }
	size_t size() const { return m.size(); }
	const std::vector<double>& values() const { return m.values; }
	void residual(const AdfDesc& a, const Flat<T>& f, AdfWork<T>& w) const {
		r1_residual<T>(a, f.cores.get(), m.device_u<T>(), w.vals.get(), w.resid.get(), w.norms.get(), w.dpart.get(), -1, nullptr,
		               stream());
	}
	void gradient(const AdfDesc& a, const Flat<T>& f, AdfWork<T>& w, int k) const {
		const int d = a.d;
		for (int q = (k < 0 ? 0 : k); q < (k < 0 ? d : k + 1); ++q)
			r1_gradient<T>(a, f.cores.get(), m.device_u<T>(), w.resid.get(), q, f.hrank[q], f.hn[q], f.hrank[q + 1], w.grad.get() + f.hoff[q],
			               w.partials.get(), stream());
	}
	// update_x (adf.cpp:481-487) on the device: one step ‖Δ‖²/Σ_i (x(u_i) with core k replaced by Δ)² for the
	// whole core (calculate_slicewise_norm_A_projGrad, adf.cpp:413-465, puts every measurement at slot 0)
	void step(const AdfDesc& a, const Flat<T>& f, AdfWork<T>& w, size_t k, DTensor& core) const {
		r1_residual<T>(a, f.cores.get(), m.device_u<T>(), nullptr, nullptr, w.slices.get(), w.dpart.get(), (int)k, w.grad.get() + f.hoff[k],
		               stream());
		adf_update_core<T>(core.data(), w.grad.get() + f.hoff[k], w.slices.get(), (int)core.dims[0], (int)core.dims[1], (int)core.dims[2],
		                   false, stream());
	}
};

template <class T, class Ops>
std::vector<std::vector<double>> full_gradient(const TT& x, const Ops& ops, double* residualNorm) {
	XTT_REQUIRE(ops.size() > 0, "ADF: empty measurement set");
	ops.check(x);
	const size_t P = ops.size();
	Flat<T> f;
	f.build(x);
	AdfWork<T> w;
	w.ensure(P, f.desc.gradSize, 1);
	upload_values<T>(ops.values(), w.vals);
	AdfDesc a = f.desc;
	a.P = (int)P;
	ops.residual(a, f, w);
	ops.gradient(a, f, w, -1);
	std::vector<T> g(f.desc.gradSize);
	double norms[2];
	XTT_HIP(hipMemcpyAsync(g.data(), w.grad.get(), g.size() * sizeof(T), hipMemcpyDeviceToHost, stream()));
	XTT_HIP(hipMemcpyAsync(norms, w.norms.get(), 16, hipMemcpyDeviceToHost, stream()));
	sync();
	if (residualNorm) *residualNorm = std::sqrt(norms[0] / norms[1]);
	std::vector<std::vector<double>> out(x.order());
	for (size_t k = 0; k < x.order(); ++k) {
		const size_t sz = x.cores[k].size();
		out[k].resize(sz);
		// xerus's projected gradient is Σ (b − x)·∂x = −Σ resid·∂x
		for (size_t e = 0; e < sz; ++e) out[k][e] = -(double)g[f.hoff[k] + e];
	}
	return out;
}

// adf.cpp:489-604 with the stack contractions replaced by the sampled-entry kernels.
template <class T, class Ops>
double solve(const ADFOptions& o, TT& x, const Ops& ops, const dims_t& maxRanksIn, size_t* iterations) {
	const size_t d = x.order(), P = ops.size();
	XTT_REQUIRE(P > 0, "Need at very least one measurment.");
	ops.check(x);
	const dims_t maxRanks = reduce_to_maximal_ranks(maxRanksIn, x.dims, false);
	double normMeasured = 0;
	for (double v : ops.values()) normMeasured += v * v;
	normMeasured = std::sqrt(normMeasured);
	AdfWork<T> w;
	upload_values<T>(ops.values(), w.vals);
	Flat<T> f;
	const size_t maxN = std::max<size_t>(2, *std::max_element(x.dims.begin(), x.dims.end()));  // slice norms per mode index
	size_t iteration = 0;
	double residualNorm = DBL_MAX, lastResidualNorm = DBL_MAX;
	PerfClock clock;
	clock.begin();

	// resid = x − b at the current x, after re-flattening cores first..last (the descriptor is rebuilt when a shape
	// changed); nothing is read back
	auto residual = [&](size_t first, size_t last) {
		if (!f.matches(x)) { f.build(x); w.ensure(P, f.desc.gradSize, maxN); } else f.refresh(x, first, last);
		AdfDesc a = f.desc;
		a.P = (int)P;
		ops.residual(a, f, w);
	};

	auto solve_with_current_ranks = [&]() {
		double resDec1 = 0.0, resDec2 = 0.0, resDec3 = 0.0;
		for (; o.maxIterations == 0 || iteration < o.maxIterations; ++iteration) {
			move_core(x, 0, true);
			residual(0, SIZE_MAX);
			double norms[2];
			read_back(norms, w.norms.get(), 16);  // the only synchronisation of an iteration
			lastResidualNorm = residualNorm;
			residualNorm = std::sqrt(norms[0]) / normMeasured;
			if (PerfSink& ps = perf_sink(); ps.cb) {  // adf.cpp:514, perfData.add(iteration, residualNorm, x, 0)
				const dims_t rk = x.ranks();
				ps.cb(ps.user, iteration, clock.elapsed(), residualNorm, rk.data(), rk.size(), 0);
			}
			const double resDec4 = resDec3;
			resDec3 = resDec2;
			resDec2 = resDec1;
			resDec1 = residualNorm / lastResidualNorm;
			if (residualNorm < o.targetResidualNorm || resDec1 * resDec2 * resDec3 * resDec4 > std::pow(o.minimalResidualNormDecrease, 4)) break;
			for (size_t k = 0; k < d; ++k) {
				if (k > 0) residual(k - 1, k);  // move_core(k) changed cores k−1 and k
				AdfDesc a = f.desc;
				a.P = (int)P;
				// gradient of ½Σ(x−b)² w.r.t. core k (xerus's Δ, adf.cpp:359-396, is its negative), then update_x
				ops.gradient(a, f, w, (int)k);
				ops.step(a, f, w, k, x.cores[k]);
				if (k + 1 < d) move_core(x, k + 1, true);
			}
		}
	};

	move_core(x, 0, false);  // canonicalize_left
	solve_with_current_ranks();
	while (residualNorm > o.targetResidualNorm && x.ranks() != maxRanks && (o.maxIterations == 0 || iteration < o.maxIterations)) {
		move_core(x, 0, true);  // adf.cpp:590-602
		TT rnd = tt_random(x.dims, dims_t(d - 1, 1), false);
		const double scale = 1e-6 * tt_frob_norm(x) / tt_frob_norm(rnd);
		tt_scale(rnd, scale);
		x = tt_add(x, rnd);
		round(x, maxRanks, 8 * DBL_EPSILON);
		solve_with_current_ranks();
	}
	if (iterations) *iterations = iteration;
	return residualNorm;
}

// adf.cpp:489-604 with xerus's stacks on the device (single points).
template <class T>
double solve_stacked(const ADFOptions& o, TT& x, const Measurements& m, const dims_t& maxRanksIn, size_t* iterations) {
	const size_t d = x.order(), P = m.size();
	XTT_REQUIRE(P > 0, "Need at very least one measurment.");
	m.check_against(x);
	const dims_t maxRanks = reduce_to_maximal_ranks(maxRanksIn, x.dims, false);
	double normMeasured = 0;
	for (double v : m.values) normMeasured += v * v;
	normMeasured = std::sqrt(normMeasured);
	StackedSinglePointOps<T> ops{m, m.stack_plan()};
	AdfWork<T> w;
	upload_values<T>(m.values, w.vals);
	Flat<T> f;
	const size_t maxN = std::max<size_t>(2, *std::max_element(x.dims.begin(), x.dims.end()));
	size_t iteration = 0;
	double residualNorm = DBL_MAX, lastResidualNorm = DBL_MAX;
	PerfClock clock;
	clock.begin();

	// one sweep of update_x over every core (adf.cpp:525-540)
	auto sweep = [&](TT& y) {
		for (size_t k = 0; k < d; ++k) {
			if (k > 0) {  // move_core(k) changed cores k−1 and k: forward level k−1, then residual and gradient at k
				f.refresh(y, k - 1, k);
				ops.forward(f, k - 1);
				ops.residual_gradient(f, w, k, false);
			}
			ops.step(w, k, y.cores[k]);
			if (k + 1 < d) move_core(y, k + 1, true);
		}
	};
	// the start of an iteration (adf.cpp:495-511): core to 0, backward stacks, residual (and gradient) at core 0
	auto start = [&](TT& y) {
		move_core(y, 0, true);
		f.refresh(y);
		for (size_t k = d - 1; k > 0; --k) ops.backward(f, k);
		ops.residual_gradient(f, w, 0, true);
	};

	// Within a rank phase every sweep has the same kernels and shapes: from the second sweep on, the sweep and the
	// next iteration's start are one recorded graph (TTGraph), launched once per iteration; the residual norms are
	// read after it. The last allowed iteration's sweep runs alone, so x ends as the eager loop leaves it.
	auto solve_with_current_ranks = [&]() {
		double resDec1 = 0.0, resDec2 = 0.0, resDec3 = 0.0;
		if (o.maxIterations != 0 && iteration >= o.maxIterations) return;
		move_core(x, 0, true);
		if (!f.matches(x)) {
			f.build(x);
			w.ensure(P, f.desc.gradSize, maxN);
		} else {
			f.refresh(x);
		}
		ops.layout(f);
		for (size_t k = d - 1; k > 0; --k) ops.backward(f, k);
		ops.residual_gradient(f, w, 0, true);
		TTGraph g;
		g.kind = 1;
		bool recordFailed = !graphs_enabled();
		size_t eagerSweeps = 0;
		for (; o.maxIterations == 0 || iteration < o.maxIterations; ++iteration) {
			double norms[2];
			read_back(norms, w.norms.get(), 16);  // the only synchronisation of an iteration
			lastResidualNorm = residualNorm;
			residualNorm = std::sqrt(norms[0]) / normMeasured;
			if (PerfSink& ps = perf_sink(); ps.cb) {  // adf.cpp:514, perfData.add(iteration, residualNorm, x, 0)
				const dims_t rk = x.ranks();
				ps.cb(ps.user, iteration, clock.elapsed(), residualNorm, rk.data(), rk.size(), 0);
			}
			const double resDec4 = resDec3;
			resDec3 = resDec2;
			resDec2 = resDec1;
			resDec1 = residualNorm / lastResidualNorm;
			if (residualNorm < o.targetResidualNorm || resDec1 * resDec2 * resDec3 * resDec4 > std::pow(o.minimalResidualNormDecrease, 4)) break;
			const bool last = o.maxIterations != 0 && iteration + 1 >= o.maxIterations;
			if (last) {
				sweep(x);
			} else if (g.ready()) {
				g.launch(x);
				g.take(x);
			} else if (eagerSweeps >= 1 && !recordFailed && g.record(x, [&](TT& y) { sweep(y); start(y); })) {
				g.launch(x);
				g.take(x);
			} else {
				if (eagerSweeps >= 1) recordFailed = true;
				sweep(x);
				start(x);
				++eagerSweeps;
			}
		}
	};

	move_core(x, 0, false);  // canonicalize_left
	solve_with_current_ranks();
	while (residualNorm > o.targetResidualNorm && x.ranks() != maxRanks && (o.maxIterations == 0 || iteration < o.maxIterations)) {
		move_core(x, 0, true);  // adf.cpp:590-602
		TT rnd = tt_random(x.dims, dims_t(d - 1, 1), false);
		const double scale = 1e-6 * tt_frob_norm(x) / tt_frob_norm(rnd);
		tt_scale(rnd, scale);
		x = tt_add(x, rnd);
		round(x, maxRanks, 8 * DBL_EPSILON);
		solve_with_current_ranks();
	}
	if (iterations) *iterations = iteration;
	return residualNorm;
}
}  // namespace

template <class T>
std::vector<std::vector<double>> adf_full_gradient(const TT& x, const Measurements& m, double* residualNorm) {
	return full_gradient<T>(x, SinglePointOps<T>{m}, residualNorm);
}
template <class T>
std::vector<std::vector<double>> adf_full_gradient(const TT& x, const RankOneMeasurements& m, double* residualNorm) {
	return full_gradient<T>(x, RankOneOps<T>{m}, residualNorm);
}
template <class T>
double adf_solve(const ADFOptions& o, TT& x, const Measurements& m, const dims_t& maxRanks, size_t* iterations) {
	const bool stacked = !getenv("XTT_ADF_NO_STACKS") &&
	                     std::all_of(x.dims.begin(), x.dims.end(), [](size_t n) { return n <= (size_t)ADF_STACK_MAX_N; });
	if (stacked) return solve_stacked<T>(o, x, m, maxRanks, iterations);
	return solve<T>(o, x, SinglePointOps<T>{m}, maxRanks, iterations);
}
template <class T>
double adf_solve(const ADFOptions& o, TT& x, const RankOneMeasurements& m, const dims_t& maxRanks, size_t* iterations) {
	return solve<T>(o, x, RankOneOps<T>{m}, maxRanks, iterations);
}
template std::vector<std::vector<double>> adf_full_gradient<double>(const TT&, const Measurements&, double*);
template std::vector<std::vector<double>> adf_full_gradient<float>(const TT&, const Measurements&, double*);
template std::vector<std::vector<double>> adf_full_gradient<double>(const TT&, const RankOneMeasurements&, double*);
template std::vector<std::vector<double>> adf_full_gradient<float>(const TT&, const RankOneMeasurements&, double*);
template double adf_solve<double>(const ADFOptions&, TT&, const Measurements&, const dims_t&, size_t*);
template double adf_solve<float>(const ADFOptions&, TT&, const Measurements&, const dims_t&, size_t*);
template double adf_solve<double>(const ADFOptions&, TT&, const RankOneMeasurements&, const dims_t&, size_t*);
template double adf_solve<float>(const ADFOptions&, TT&, const RankOneMeasurements&, const dims_t&, size_t*);

}  // namespace xtt
