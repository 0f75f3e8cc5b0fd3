// oz_nearsum.hpp -- device restatement of the reference NearSum (private header).
//
// Rump-Ogita-Oishi NearSum as written in the reference src/cpu/ozblas_sum.cpp:10-112
// (ozblasTransform, ozblasTransformK, ozblasNearsum), operation for operation, so that the device
// result is bit-identical to the CPU one.  The vector is modified in place, as in the reference.
// Compile with -ffp-contract=off: every + - * below must stay a separately rounded operation.
#pragma once

#include "oz_internal.hpp"

namespace oz {

struct StridedVec {
	static constexpr int kMax = 0;  // length known only at run time
	double *p;
	int64_t ld;
	__device__ double &operator[](int i) const { return p[(int64_t)i * ld]; }
};

template <int N>
struct RegVec {  // register array: loops over it run to N with a guard, so every index is static
	static constexpr int kMax = N;
	double v[N];
	__device__ double &operator[](int i) { return v[i]; }
};

// for (i = i0; i < n; i++) f(i), unrolled to the static bound of V when it has one
template <class V, class F>
__device__ __forceinline__ void ns_for(int i0, int n, F &&f) {
	if constexpr (V::kMax > 0) {
#pragma unroll
		for (int i = 0; i < V::kMax; i++)
			if (i >= i0 && i < n) f(i);
	} else {
		for (int i = i0; i < n; i++) f(i);
	}
}

__device__ inline double mac_max(double a, double b) { return (a > b) ? a : b; }  // MAX macro

// ceil(log2((double)N)) for an integer N >= 1 (exact for every int: log2 of a power of two is
// exact and the other values are at least ~2^-31 away from an integer)
__device__ inline int ceil_log2_int(int N) { return N <= 1 ? 0 : 32 - __builtin_clz((unsigned)(N - 1)); }

template <class V>
__device__ inline void ns_transform(int n, V &vec, double rho, double &tau1, double &tau2,
                                    const uint16_t *J) {
	// ozblasTransform (src/cpu/ozblas_sum.cpp:10-56)
	if (n == 0) {
		tau1 = rho;
		tau2 = 0.;
		return;
	}
	double mu = fabs(vec[0]);
	ns_for<V>(1, n, [&](int i) { mu = mac_max(mu, fabs(vec[i])); });
	if (mu == 0.) {
		tau1 = rho;
		tau2 = 0.;
		return;
	}
	const int m = ceil_log2_int(n + 2);
	// scalbn1(1., m + ceil(log2(mu))): +-inf/NaN exponents convert to INT_MIN -> sigma = 0
	double sigma = (mu <= 1.7976931348623157e308) ? scalbn_clamped(1., (int64_t)m + ceil_log2_tab(mu, J))
	                                              : 0.0;
	double t = rho;
	const double tiny = 2.2250738585072014e-308;  // scalbn(1., emin)
	const double fac = scalbn_clamped(1., 2 * m + 1 - 53);
	const double shrink = scalbn_clamped(1., m - 53);
	while (true) {
		double tau = 0.;
		ns_for<V>(0, n, [&](int i) {
			const double tmp = (sigma + vec[i]) - sigma;
			vec[i] = vec[i] - tmp;
			tau = tau + tmp;
		});
		tau1 = t + tau;
		if ((sigma <= tiny) || (fabs(tau1) >= fac * sigma)) {
			tau2 = tau - (tau1 - t);
			return;
		}
		sigma = shrink * sigma;
		t = tau1;
	}
}

template <class V>
__device__ inline void ns_transform_k(int n, V &vec, double rho, double &res, double &r,
                                      const uint16_t *J) {
	// ozblasTransformK (src/cpu/ozblas_sum.cpp:58-73)
	double tmp = 0., tau1, tau2;
	ns_transform(n, vec, rho, tau1, tau2, J);
	ns_for<V>(0, n, [&](int i) { tmp = tmp + vec[i]; });
	res = tau1 + (tau2 + tmp);
	r = tau2 - (res - tau1);
}

__device__ inline double ns_sign(double v) { return (v < 0) ? -1. : 1.; }

template <class V>
__device__ inline double nearsum(int n, V &vec, const uint16_t *J) {
	// ozblasNearsum (src/cpu/ozblas_sum.cpp:81-112)
	double tmp, res, res2, r, r2, mu, delta, delta2;
	const double eps = 1.1102230246251565e-16;  // scalbn(1., -53)
	ns_transform_k(n, vec, 0., res, r, J);
	ns_transform_k(n, vec, r, delta, r2, J);
	if (delta == 0) return res;
	res2 = res + ns_sign(delta) * eps * fabs(res);
	if (res2 == res) {
		mu = ns_sign(delta) * eps * fabs(res);
		res2 = res + 2. * ns_sign(delta) * eps * fabs(res);
	} else {
		mu = (res2 - res) / 2.;
	}
	if (fabs(delta) < fabs(mu)) return res;
	if (fabs(delta) > fabs(mu)) return res2;
	ns_transform_k(n, vec, r2, delta2, tmp, J);
	if (delta2 == 0) return res + mu;
	if (ns_sign(delta2) == ns_sign(mu)) return res2;
	return res;
}

}  // namespace oz
