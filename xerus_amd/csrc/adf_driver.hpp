// Measurement sets and the ADF solver / gradient on the device.
#pragma once
#include "tt.hpp"

namespace xtt {

struct Measurements {  // SinglePointMeasurementSet (include/xerus/measurments.h:43-91)
	size_t d = 0;
	std::vector<uint32_t> pos;     // P×d, lexicographically sorted
	std::vector<double> values;
	DBuf<uint32_t> dpos;           // device copy of pos
	mutable std::shared_ptr<AdfPlan> plan;  // ADF's stack plan of these positions, built on first use
	size_t size() const { return values.size(); }
	void upload();
	const AdfPlan& stack_plan() const;
	void check_against(const TT& x) const;  // every position inside x's dimensions
};
Measurements meas_random(size_t P, const dims_t& dims);
void meas_measure(Measurements& m, const TT& x);

// RankOneMeasurementSet (include/xerus/measurments.h:94-150): measurement i is u_{i,0} ⊗ … ⊗ u_{i,d-1}.
struct RankOneMeasurements {
	dims_t dims;                 // n_k
	std::vector<double> u;       // P × Σn_k: measurement i's vectors concatenated (mode k at offset Σ_{j<k} n_j)
	std::vector<double> values;
	size_t d() const { return dims.size(); }
	size_t sumN() const;
	size_t size() const { return values.size(); }
	void check_against(const TT& x) const;
	// device copy of u in the working precision (uploaded on first use after a change)
	template <class T>
	const T* device_u() const;
	void invalidate() { dirty = true; }
	mutable DBuf<double> du64;
	mutable DBuf<float> du32;
	mutable bool dirty = true, dirty32 = true;
};
RankOneMeasurements r1_random(size_t P, const dims_t& dims);                     // measurments.cpp:267-272, 513-536
RankOneMeasurements r1_from_single_point(const Measurements& m, const dims_t& dims);  // measurments.cpp:253-265
void r1_sort(RankOneMeasurements& m, bool positionsOnly);                          // measurments.cpp:329-344
void r1_normalize(RankOneMeasurements& m);                                        // measurments.cpp:346-355
void r1_measure(RankOneMeasurements& m, const TT& x);                             // measurments.cpp:392-424
double r1_test(const RankOneMeasurements& m, const TT& x);                        // measurments.cpp:469-505

// Flattened cores of a TT in one device buffer (working precision T) plus the AdfDesc describing them.
template <class T>
struct Flat {
	DBuf<T> cores;
	DBuf<int> dn, drank;
	DBuf<long long> doff;
	std::vector<int> hn, hrank;
	std::vector<long long> hoff;
	AdfDesc desc{};
	void build(const TT& x);
	void refresh(const TT& x, size_t first = 0, size_t last = SIZE_MAX);  // cores first..last (inclusive) only
	bool matches(const TT& x) const;  // the descriptor still describes x's shapes
};

struct ADFOptions { size_t maxIterations = 0; double targetResidualNorm = 1e-8; double minimalResidualNormDecrease = 0.999; };

// xerus's projected gradient Δ_k = Σ_i (b_i − x_i)·∂x_i/∂G_k for every core at x (adf.cpp:314-396).
template <class T>
std::vector<std::vector<double>> adf_full_gradient(const TT& x, const Measurements& m, double* residualNorm);
template <class T>
double adf_solve(const ADFOptions& o, TT& x, const Measurements& m, const dims_t& maxRanks, size_t* iterations);
// The same for rank-one measurements (the RankOneMeasurementSet specialisations of adf.cpp).
template <class T>
std::vector<std::vector<double>> adf_full_gradient(const TT& x, const RankOneMeasurements& m, double* residualNorm);
template <class T>
double adf_solve(const ADFOptions& o, TT& x, const RankOneMeasurements& m, const dims_t& maxRanks, size_t* iterations);

}  // namespace xtt
