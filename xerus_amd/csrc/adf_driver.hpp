// Measurement sets and the ADF solver / gradient on the device.
#pragma once
#include "tt.hpp"

namespace xtt {

struct Measurements {  // SinglePointMeasurementSet (include/xerus/measurments.h:43-91)
	size_t d = 0;
	std::vector<uint32_t> pos;     // P×d, lexicographically sorted
	std::vector<double> values;
	DBuf<uint32_t> dpos;           // device copy of pos
	size_t size() const { return values.size(); }
	void upload();
	void check_against(const TT& x) const;  // every position inside x's dimensions
};
Measurements meas_random(size_t P, const dims_t& dims);
void meas_measure(Measurements& m, const TT& x);

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
	void refresh(const TT& x);
};

struct ADFOptions { size_t maxIterations = 0; double targetResidualNorm = 1e-8; double minimalResidualNormDecrease = 0.999; };

// xerus's projected gradient Δ_k = Σ_i (b_i − x_i)·∂x_i/∂G_k for every core at x (adf.cpp:314-396).
template <class T>
std::vector<std::vector<double>> adf_full_gradient(const TT& x, const Measurements& m, double* residualNorm);
template <class T>
double adf_solve(const ADFOptions& o, TT& x, const Measurements& m, const dims_t& maxRanks, size_t* iterations);

}  // namespace xtt
