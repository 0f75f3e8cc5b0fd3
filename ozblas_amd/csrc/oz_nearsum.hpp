// oz_nearsum.hpp -- device restatement of the reference NearSum (private header).
//
// Rump-Ogita-Oishi NearSum as written in the reference src/cpu/ozblas_sum.cpp:10-112
// (ozblasTransform, ozblasTransformK, ozblasNearsum), operation for operation, so that the device
// result is bit-identical to the CPU one.  The vector is modified in place, as in the reference.
// Compile with -ffp-contract=off: every + - * below must stay a separately rounded operation.
#pragma once

#include "oz_internal.hpp"

namespace oz {

template <class T>
struct StridedVecT {
	static constexpr int kMax = 0;  // length known only at run time
	T *p;
	int64_t ld;
	__device__ T &operator[](int i) const { return p[(int64_t)i * ld]; }
};
using StridedVec = StridedVecT<double>;
using StridedVecF = StridedVecT<float>;

template <int N, class T = double>
struct RegVec {  // register array: loops over it run to N with a guard, so every index is static
	static constexpr int kMax = N;
	T v[N];
	__device__ T &operator[](int i) { return v[i]; }
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

// ---------------------------------------------------------------------------------------------
// The same NearSum over n <= 64 finite values held one per lane of a wave (lane i: vec[i], lanes
// i >= n: 0), every lane computing the same result.  Each pass of the transform extracts
// tmp_i = (sigma + vec[i]) - sigma with sigma >= 2^m max|vec|, 2^m >= n + 2: the tmp_i are
// multiples of ulp(sigma) / 2 and every partial sum of them stays within sigma, so the
// reference's running sum tau is exact and equals a sum over the lanes in any order, as mu equals a
// maximum over the lanes (values finite).  The one inexact sum -- of the remainders after a transform --
// runs in lane order (src/cpu/ozblas_sum.cpp:64-67), each lane's value read in turn.
// ---------------------------------------------------------------------------------------------
__device__ inline double ns_wave_max(double v) { return wave_max_nonneg_f64(v); }  // of |values|: finite, >= 0
__device__ inline double ns_wave_sum(double v) { return wave_sum_f64(v); }

__device__ inline void ns_transform_wave(int n, double &v, double rho, double &tau1, double &tau2,
                                         const uint16_t *J) {
	if (n == 0) {
		tau1 = rho;
		tau2 = 0.;
		return;
	}
	const double mu = ns_wave_max(fabs(v));
	if (mu == 0.) {
		tau1 = rho;
		tau2 = 0.;
		return;
	}
	const int m = ceil_log2_int(n + 2);
	double sigma = scalbn_clamped(1., (int64_t)m + ceil_log2_tab(mu, J));  // mu finite
	double t = rho;
	const double tiny = 2.2250738585072014e-308;
	const double fac = scalbn_clamped(1., 2 * m + 1 - 53);
	const double shrink = scalbn_clamped(1., m - 53);
	while (true) {
		const double tmp = (sigma + v) - sigma;
		v = v - tmp;
		const double tau = ns_wave_sum(tmp);
		tau1 = t + tau;
		if ((sigma <= tiny) || (fabs(tau1) >= fac * sigma)) {
			tau2 = tau - (tau1 - t);
			return;
		}
		sigma = shrink * sigma;
		t = tau1;
	}
}

__device__ inline void ns_transform_k_wave(int n, double &v, double rho, double &res, double &r,
                                           const uint16_t *J) {
	double tmp = 0., tau1, tau2;
	ns_transform_wave(n, v, rho, tau1, tau2, J);
	for (int i = 0; i < n; i++) tmp = tmp + lane_f64(v, i);
	res = tau1 + (tau2 + tmp);
	r = tau2 - (res - tau1);
}

// all 64 lanes of the wave call it; v: this lane's value (0 for lanes >= n)
__device__ inline double nearsum_wave(int n, double v, const uint16_t *J) {
	double tmp, res, res2, r, r2, mu, delta, delta2;
	const double eps = 1.1102230246251565e-16;
	ns_transform_k_wave(n, v, 0., res, r, J);
	ns_transform_k_wave(n, v, r, delta, r2, J);
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
	ns_transform_k_wave(n, v, r2, delta2, tmp, J);
	if (delta2 == 0) return res + mu;
	if (ns_sign(delta2) == ns_sign(mu)) return res2;
	return res;
}

// ---------------------------------------------------------------------------------------------
// binary32 NearSum: the <float> instantiation of the same functions.  There the element type is
// float, but unqualified fabs and scalbn and the double literals 2. and /2. promote part of each
// expression to double before the result is rounded to float on assignment; ceil(std::log2(mu))
// is log2f (table JF2).  Each mixed step below is written as that double expression, then cast.
// ---------------------------------------------------------------------------------------------
__device__ inline float mac_maxf(float a, float b) { return (a > b) ? a : b; }

template <class V>
__device__ inline void ns_transform_f(int n, V &vec, float rho, float &tau1, float &tau2,
                                      const uint16_t *JF2) {
	// ozblasTransform<float> (src/cpu/ozblas_sum.cpp:10-56): emin = -126, epse = 24
	if (n == 0) {
		tau1 = rho;
		tau2 = 0.f;
		return;
	}
	float mu = fabsf(vec[0]);
	ns_for<V>(1, n, [&](int i) { mu = mac_maxf(mu, fabsf(vec[i])); });
	if (mu == 0.f) {
		tau1 = rho;
		tau2 = 0.f;
		return;
	}
	const int m = ceil_log2_int(n + 2);
	// (float)scalbn(1., m + ceil(log2f(mu))): an infinite or NaN mu gives exponent INT_MIN -> 0
	float sigma = (mu <= 3.40282347e38f)
	                  ? (float)scalbn_clamped(1., (int64_t)m + ceil_log2_tab_f(mu, JF2))
	                  : 0.f;
	float t = rho;
	const double fac = scalbn_clamped(1., 2 * m + 1 - 24);
	const double shrink = scalbn_clamped(1., m - 24);
	while (true) {
		float tau = 0.f;
		ns_for<V>(0, n, [&](int i) {
			const float tmp = (sigma + vec[i]) - sigma;
			vec[i] = vec[i] - tmp;
			tau = tau + tmp;
		});
		tau1 = t + tau;
		if (((double)sigma <= 0x1p-126) || (fabs((double)tau1) >= fac * (double)sigma)) {
			tau2 = tau - (tau1 - t);
			return;
		}
		sigma = (float)(shrink * (double)sigma);
		t = tau1;
	}
}

template <class V>
__device__ inline void ns_transform_k_f(int n, V &vec, float rho, float &res, float &r,
                                        const uint16_t *JF2) {
	// ozblasTransformK<float> (src/cpu/ozblas_sum.cpp:58-73)
	float tmp = 0.f, tau1, tau2;
	ns_transform_f(n, vec, rho, tau1, tau2, JF2);
	ns_for<V>(0, n, [&](int i) { tmp = tmp + vec[i]; });
	res = tau1 + (tau2 + tmp);
	r = tau2 - (res - tau1);
}

__device__ inline float ns_sign_f(float v) { return (v < 0) ? -1.f : 1.f; }

template <class V>
__device__ inline float nearsum_f(int n, V &vec, const uint16_t *JF2) {
	// ozblasNearsum<float> (src/cpu/ozblas_sum.cpp:81-112)
	float tmp, res, res2, r, r2, mu, delta, delta2;
	const float eps = 0x1p-24f;
	ns_transform_k_f(n, vec, 0.f, res, r, JF2);
	ns_transform_k_f(n, vec, r, delta, r2, JF2);
	if (delta == 0) return res;
	const float se = ns_sign_f(delta) * eps;  // float product (exact)
	res2 = (float)((double)res + (double)se * fabs((double)res));
	if (res2 == res) {
		mu = (float)((double)se * fabs((double)res));
		res2 = (float)((double)res + 2. * (double)ns_sign_f(delta) * (double)eps * fabs((double)res));
	} else {
		mu = (float)((double)(res2 - res) / 2.);
	}
	if (fabsf(delta) < fabsf(mu)) return res;
	if (fabsf(delta) > fabsf(mu)) return res2;
	ns_transform_k_f(n, vec, r2, delta2, tmp, JF2);
	if (delta2 == 0) return res + mu;
	if (ns_sign_f(delta2) == ns_sign_f(mu)) return res2;
	return res;
}

}  // namespace oz
