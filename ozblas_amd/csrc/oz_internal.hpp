// oz_internal.hpp -- shared host/device helpers of the MI355X Ozaki BLAS (private header).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <vector>

#include "../../include/ozblas_mi355x.h"

namespace oz {

constexpr int kNumSplitDefaultMax = 30;  // reference src/cpu/ozblas_common.h:5
constexpr int kTauBinades = 2098;        // binary64 binades 2^-1074 .. 2^1023
constexpr int kTauBias = 1074;

struct Status {
	int32_t code;
};

#define OZ_HIP(call)                                                                        \
	do {                                                                                    \
		hipError_t e__ = (call);                                                            \
		if (e__ != hipSuccess) {                                                            \
			fprintf(stderr, "ozblas_mi355x: %s failed: %s (%s:%d)\n", #call,                \
			        hipGetErrorString(e__), __FILE__, __LINE__);                            \
			return OZBLAS_STATUS_EXECUTION_FAILED;                                          \
		}                                                                                   \
	} while (0)

#define OZ_TRY(expr)                              \
	do {                                          \
		int32_t s__ = (expr);                     \
		if (s__ != OZBLAS_STATUS_SUCCESS) return s__; \
	} while (0)

// Library-private state behind ozblasHandle_t::backend.
struct Backend {
	hipStream_t stream = nullptr;
	int device = 0;
	char *work = nullptr;      // device workspace (devWork)
	uint64_t cap = 0;
	uint64_t top = 0;          // bump allocator offset (reset per call, like counterInit)
	uint64_t keep = 0;         // bytes preserved across calls (pre-split CSR matrix)
	uint16_t *dTau = nullptr;  // device copy of the ceil(log2) table
	int64_t *hPinned = nullptr;  // pinned host scratch (64 x 8 bytes)
	int32_t *dCounters = nullptr;  // device scratch counters (256 x 4 bytes)
	hipEvent_t ev[8] = {};
	void *batch = nullptr;  // cached plan of ozblasDgemmBatched (oz_batched.hip)
	int nnzA = 0;           // nonzeros of the matrix pre-split for CG (slice pitch rup(nnz, 32))
	uint16_t *dTauF = nullptr;   // device copies of the binary32 tables JF, JF2
	uint16_t *dTauF2 = nullptr;
	int mixed = 0;  // 1 while a mixed (DS: binary64 data, binary32 slices) routine runs
	char *scratch = nullptr;  // this device's copy of the static scratch array (ozblasDdotNoAlloc)
};

// The static scratch array of the allocation-free entry point: kStaticScratchBytes of device
// memory per GPU, held until the process ends.  static_scratch() returns the current device's
// address of it (looked up once per device), or nullptr on failure.
constexpr uint64_t kStaticScratchBytes = 64ull << 20;
char *static_scratch(int device);

void batch_plan_free(void *plan);

inline Backend *be(ozblasHandle_t *oh) { return reinterpret_cast<Backend *>(oh->backend); }
inline const Backend *be(const ozblasHandle_t *oh) {
	return reinterpret_cast<const Backend *>(oh->backend);
}

// bump allocation inside the workspace, 256-byte aligned; returns nullptr when it does not fit
inline void *walloc(ozblasHandle_t *oh, uint64_t bytes) {
	Backend *b = be(oh);
	uint64_t off = (b->top + 255) & ~uint64_t(255);
	if (off + bytes > b->cap) return nullptr;
	b->top = off + bytes;
	oh->memAddr = b->top;
	return b->work + off;
}
inline uint64_t wmark(ozblasHandle_t *oh) { return be(oh)->top; }
inline void wreset(ozblasHandle_t *oh, uint64_t mark) {
	be(oh)->top = mark;
	oh->memAddr = mark;
}

// ---------------------------------------------------------------------------------------------
// ceil(log2(x)) exactly as the host libm computes it.
//
// The reference computes tau = (short)ceil(log2(max)) with the C library's log2
// (src/cpu/ozblas_split.cpp:86,158,808; NearSum's ceil(log2(mu)) at src/cpu/ozblas_sum.cpp:35).
// For x = 2^e exactly the result is e; for x in (2^e, 2^(e+1)) it is e+1 unless log2(x) rounds
// down to the integer e, which happens for the first few representable x above 2^e (how many
// depends on the binade and on the libm's rounding).  J[e + 1074] holds that count, measured on
// the host libm at handle creation; the device evaluates the same rule from the table.
// ---------------------------------------------------------------------------------------------
__host__ __device__ inline uint64_t d2u(double x) { return __builtin_bit_cast(uint64_t, x); }

// x > 0 and finite
__host__ __device__ inline int ceil_log2_tab(double x, const uint16_t *J) {
	const uint64_t b = d2u(x);
	const int E = (int)((b >> 52) & 0x7ff);
	const uint64_t f = b & ((uint64_t(1) << 52) - 1);
	int e;
	uint64_t off;
	if (E == 0) {  // subnormal: x = f * 2^-1074
		const int msb = 63 - __builtin_clzll(f);
		e = -1074 + msb;
		off = f - (uint64_t(1) << msb);
	} else {
		e = E - 1023;
		off = f;
	}
	if (off == 0) return e;
	if (off > 0xFFFFu) return e + 1;  // beyond any table entry (16 bits): no lookup
	return (off <= (uint64_t)J[e + kTauBias]) ? e : e + 1;
}

// ---------------------------------------------------------------------------------------------
// Wave reductions on the DPP network (no LDS round trips): quad permutes, half-row and row mirrors,
// then the row broadcasts of lanes 15 and 31; the total lands in lane 63.  Lanes a step does not
// write take 0, the identity of the sums and of maxima of non-negative values.  Binary64 values
// move as two 32-bit halves.
// ---------------------------------------------------------------------------------------------
template <int CTRL, int ROWMASK = 0xF>
__device__ __forceinline__ double dpp_f64(double v) {
	const uint64_t b = __builtin_bit_cast(uint64_t, v);
	const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)b, CTRL, ROWMASK, 0xF, false);
	const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(b >> 32), CTRL, ROWMASK, 0xF, false);
	return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ double lane_f64(double v, int lane) {
	const uint64_t b = __builtin_bit_cast(uint64_t, v);
	const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, lane);
	const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), lane);
	return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
// the sum over the 64 lanes, valid in lane 63 (exact sums: any order gives the same bits)
__device__ __forceinline__ double wave_sum63_f64(double v) {
	v += dpp_f64<0xB1>(v);         // quad_perm [1,0,3,2]
	v += dpp_f64<0x4E>(v);         // quad_perm [2,3,0,1]
	v += dpp_f64<0x141>(v);        // row_half_mirror
	v += dpp_f64<0x140>(v);        // row_mirror: every lane holds its row's sum
	v += dpp_f64<0x142, 0xA>(v);   // row_bcast:15 into rows 1 and 3
	v += dpp_f64<0x143, 0xC>(v);   // row_bcast:31 into rows 2 and 3
	return v;
}
__device__ __forceinline__ double wave_sum_f64(double v) { return lane_f64(wave_sum63_f64(v), 63); }
// the maximum of non-negative values over the 64 lanes, in every lane
__device__ __forceinline__ double wave_max_nonneg_f64(double v) {
	v = fmax(v, dpp_f64<0xB1>(v));
	v = fmax(v, dpp_f64<0x4E>(v));
	v = fmax(v, dpp_f64<0x141>(v));
	v = fmax(v, dpp_f64<0x140>(v));
	v = fmax(v, dpp_f64<0x142, 0xA>(v));
	v = fmax(v, dpp_f64<0x143, 0xC>(v));
	return lane_f64(v, 63);
}

// (short)ceil(log2(fabs(max))): 0 for zero, NaN and Inf (x86-64 conversion of +-inf/NaN -> 0)
__host__ __device__ inline int tau_of(double max, const uint16_t *J) {
	const double x = fabs(max);
	if (!(x > 0.0) || !(x <= 1.7976931348623157e308)) return 0;
	return ceil_log2_tab(x, J);
}

// ---------------------------------------------------------------------------------------------
// binary32: the reference's <float,float> split takes tau = (short)ceil(log21(max)) with
// log21<float>(v) = (float)log2((double)v) (src/cpu/ozblas_aux.cpp:37-42), and its NearSum takes
// ceil(std::log2(mu)), i.e. log2f, for a float mu (src/cpu/ozblas_sum.cpp:35).  Both round log2 to
// float, so ceil gives e for the first few floats above 2^e; tables JF (split rule) and JF2
// (NearSum rule), indexed by e + 149, hold those counts, measured on the host libm.
// ---------------------------------------------------------------------------------------------
constexpr int kTauBinadesF = 277;  // binary32 binades 2^-149 .. 2^127
constexpr int kTauBiasF = 149;

__host__ __device__ inline uint32_t f2u(float x) { return __builtin_bit_cast(uint32_t, x); }

// x > 0 and finite
__host__ __device__ inline int ceil_log2_tab_f(float x, const uint16_t *J) {
	const uint32_t b = f2u(x);
	const int E = (int)((b >> 23) & 0xff);
	const uint32_t f = b & ((1u << 23) - 1);
	int e;
	uint32_t off;
	if (E == 0) {  // subnormal: x = f * 2^-149
		const int msb = 31 - __builtin_clz(f);
		e = -149 + msb;
		off = f - (1u << msb);
	} else {
		e = E - 127;
		off = f;
	}
	if (off == 0) return e;
	if (off > 0xFFFFu) return e + 1;  // beyond any table entry (16 bits): no lookup
	return (off <= (uint32_t)J[e + kTauBiasF]) ? e : e + 1;
}

// (short)ceil(log21(fabs1(max))) for a float max: 0 for zero, NaN and Inf
__host__ __device__ inline int tau_of_f(float max, const uint16_t *JF) {
	const float x = fabsf(max);
	if (!(x > 0.0f) || !(x <= 3.40282347e38f)) return 0;
	return ceil_log2_tab_f(x, JF);
}

// rho of getRho<float,float> (getEpse = 24, getEpse2 = 53) from integers only, as rho_dev:
// mode 0: ceil(24 - (24 - log2 d) / 2) = ceil((24 + log2 d) / 2); mode 9: ceil(12.5 + log2(d) / 4)
// = 12 + ceil((2 + q) / 4); mode 2: ceil((log2(d) - 5) / 2) = ceil((q - 5) / 2); q = ceil(log2 d)
__host__ __device__ inline int rho_dev_f(int dim, int mode) {
	if (dim <= 0) return INT32_MIN;
	const int p = 31 - __builtin_clz((unsigned)dim);
	const bool pow2 = (dim & (dim - 1)) == 0;
	const int q = pow2 ? p : p + 1;
	if (mode == 2) return (q - 5 >= 0) ? (q - 5 + 1) / 2 : -((5 - q) / 2);
	if (mode == 9) return 12 + (q + 2 + 3) / 4;
	if (pow2) return (24 + p + 1) / 2;
	return (p & 1) ? (25 + p) / 2 : (26 + p) / 2;
}

// std::max semantics (NaN never replaces the running maximum)
__host__ __device__ inline double std_max(double a, double b) { return (a < b) ? b : a; }

__host__ __device__ inline double scalbn_clamped(double v, int64_t e) {
	if (e > 4000) e = 4000;
	if (e < -4000) e = -4000;
	return ldexp(v, (int)e);
}

// sigma = 0.75 * 2^(rho+tau) * splitShift (src/cpu/ozblas_split.cpp:88,159)
__host__ __device__ inline double split_sigma(int rho, int tau, int shift) {
	return 0.75 * scalbn_clamped(1.0, (int64_t)rho + tau) * (double)shift;
}

// rho for a CSR row of `dim` entries: ceil(53 - (53 - log2(dim))/2) without a log:
// for dim = 2^p it is ceil((53+p)/2); otherwise (55+p)/2 for odd p and (54+p)/2 for even p,
// p = floor(log2 dim).  dim = 0 gives INT_MIN as the reference's conversion of -inf does.
__host__ __device__ inline int rho_of_dim(int dim) {
	if (dim <= 0) return INT32_MIN;
	const int p = 31 - __builtin_clz((unsigned)dim);
	if ((dim & (dim - 1)) == 0) return (53 + p + 1) / 2;
	return (p & 1) ? (55 + p) / 2 : (54 + p) / 2;
}

// rho of getRho for every splitEpsMode, from integers only (the per-row split of a CSR matrix runs
// it on the device): mode 2 (Dot2) ceil(53 - (106 - log2 d) / 2) = ceil(log2(d) / 2)
// = ceil(q / 2); mode 9 ceil(53 - (53 - log2(2 sqrt d)) / 2) = 27 + ceil(log2(d) / 4)
// = 27 + ceil(q / 4); q = ceil(log2 d).  d = 0 gives INT_MIN as for mode 0.
__host__ __device__ inline int rho_dev(int dim, int mode) {
	if (mode != 2 && mode != 9) return rho_of_dim(dim);
	if (dim <= 0) return INT32_MIN;
	const int p = 31 - __builtin_clz((unsigned)dim);
	const int q = ((dim & (dim - 1)) == 0) ? p : p + 1;
	return (mode == 2) ? (q + 1) / 2 : 27 + (q + 3) / 4;
}

// host helpers (oz_handle.cpp)
int rho_host(int dim, int splitEpsMode);
int rho_host_f(int dim, int splitEpsMode);  // getRho<float,float> with the host libm

// fma1<float> (src/cpu/ozblas_aux.cpp:4-9): the double fma of the promoted operands, rounded to float
__host__ __device__ inline float fma1_f(float a, float b, float c) {
	return (float)fma((double)a, (double)b, (double)c);
}
__host__ __device__ inline double fma1_t(double a, double b, double c) { return fma(a, b, c); }
__host__ __device__ inline float fma1_t(float a, float b, float c) { return fma1_f(a, b, c); }

// Rules of the split per element type: the reference's <double,double> and <float,float>
// instantiations of src/cpu/ozblas_split.cpp.  Slices are always stored as doubles (a binary32
// slice converts exactly), so one set of FP64 product kernels serves both.
//   kPrec     significand bits (the unit of sigma is 2^(rho + tau - kPrec))
//   kSigEmax  largest rho + tau for which 2^(rho+tau) and sigma stay finite (with margin)
//   kEmin     exponent of the smallest normal number
template <class T>
struct SplitRule;
template <>
struct SplitRule<double> {
	static constexpr int kPrec = 53, kSigEmax = 1022, kEmin = -1022;
	__host__ __device__ static int tau(double mx, const uint16_t *J) { return tau_of(mx, J); }
	// CONST * scalbn1(1, rho + tau) * splitShift (src/cpu/ozblas_split.cpp:88,159)
	__host__ __device__ static double sigma(int rho, int tau, int shift) {
		return split_sigma(rho, tau, shift);
	}
	// CONST * scalbn1(1, rho + tau) / splitShift (the sparse split, src/cpu/ozblas_split.cpp:809)
	__host__ __device__ static double sigma_div(int rho, int tau, int shift) {
		return 0.75 * scalbn_clamped(1.0, (int64_t)rho + tau) / (double)shift;
	}
	__host__ __device__ static int rho_row(int dim, int mode) { return rho_dev(dim, mode); }
};
template <>
struct SplitRule<float> {
	static constexpr int kPrec = 24, kSigEmax = 126, kEmin = -126;
	__host__ __device__ static int tau(double mx, const uint16_t *JF) { return tau_of_f((float)mx, JF); }
	// scalbn1<float> rounds 2^(rho+tau) to float; CONST * p * shift is a double, rounded to float
	__host__ __device__ static float sigma(int rho, int tau, int shift) {
		const float p = (float)scalbn_clamped(1.0, (int64_t)rho + tau);
		return (float)(0.75 * (double)p * (double)shift);
	}
	__host__ __device__ static float sigma_div(int rho, int tau, int shift) {
		const float p = (float)scalbn_clamped(1.0, (int64_t)rho + tau);
		return (float)(0.75 * (double)p / (double)shift);
	}
	__host__ __device__ static int rho_row(int dim, int mode) { return rho_dev_f(dim, mode); }
};

// the reproMode = 0 remainder slot of the split loop: same-type instantiations only
// (`if constexpr (std::is_same<TYPE1, TYPE2>::value)`, src/cpu/ozblas_split.cpp:275-287,916-925)
inline bool repro_slot(const ozblasHandle_t *oh) { return oh->reproMode == 0 && !be(oh)->mixed; }

// cap on slices of the reference split loop (src/cpu/ozblas_split.cpp:275-277)
inline int split_cap(const ozblasHandle_t *oh) {
	int maxS = (oh->nSplitMax > 0) ? oh->nSplitMax : kNumSplitDefaultMax;
	if (repro_slot(oh)) maxS--;
	return maxS;
}

// getRho<double,float> (the mixed split: epse of TYPE1 53, of TYPE2 24, epse2 of TYPE2 53) from
// integers: mode 0 ceil(41 + log2(d) / 2) = 41 + ceil(q / 2); mode 9 ceil(41.5 + log2(d) / 4)
// = 41 + ceil((2 + q) / 4); mode 2 is the binary64 mode-0 value; q = ceil(log2 d)
__host__ __device__ inline int rho_dev_ds(int dim, int mode) {
	if (mode == 2) return rho_of_dim(dim);
	if (dim <= 0) return INT32_MIN;
	const int p = 31 - __builtin_clz((unsigned)dim);
	const int q = ((dim & (dim - 1)) == 0) ? p : p + 1;
	return (mode == 9) ? 41 + (q + 5) / 4 : 41 + (q + 1) / 2;
}
int rho_host_ds(int dim, int splitEpsMode);  // getRho<double,float> with the host libm

// getRho<float,double> (the mixed split of binary32 data into binary64-held slices: epse of TYPE1
// 24, of TYPE2 53, epse2 of TYPE2 106) from integers: mode 0 ceil((log2(d) - 5) / 2)
// = ceil((q - 5) / 2); mode 9 ceil(log2(d) / 4 - 2) = ceil(q / 4) - 2; mode 2
// ceil(log2(d) / 2 - 29) = ceil(q / 2) - 29; q = ceil(log2 d)
__host__ __device__ inline int rho_dev_sd(int dim, int mode) {
	if (dim <= 0) return INT32_MIN;
	const int p = 31 - __builtin_clz((unsigned)dim);
	const int q = ((dim & (dim - 1)) == 0) ? p : p + 1;
	if (mode == 9) return (q + 3) / 4 - 2;
	if (mode == 2) return (q + 1) / 2 - 29;
	return (q - 5 >= 0) ? (q - 5 + 1) / 2 : -((5 - q) / 2);
}
int rho_host_sd(int dim, int splitEpsMode);  // getRho<float,double> with the host libm

// A slice product P (exact in binary64) as the binary32 term of the reference's float sums.  The
// <float,double> Fsum takes c = (float)P of the product of the SCALED slices and adds
// scalbn1<float>(c, seA + seB) = (float)scalbn((double)c, seA + seB): the product rounds to 24
// significant bits first, and only then, if it lies below 2^-126, onto the binary32 subnormal
// grid.  The slices here are unscaled, so the first rounding is done on the bits of P.  Above
// 2^-126 (and for every product that is itself a binary32 value, as in the <float,float>
// instantiation) this is the plain conversion.
__host__ __device__ inline float f32_term(double p) {
	if (!(fabs(p) < 0x1p-126)) return (float)p;
	uint64_t u = __builtin_bit_cast(uint64_t, p);
	const uint64_t low = u & 0x1FFFFFFFull, half = 0x10000000ull;
	u -= low;
	if (low > half || (low == half && ((u >> 29) & 1))) u += 0x20000000ull;
	return (float)__builtin_bit_cast(double, u);
}
// term of a sum in T from a product accumulated in V (V = T, or V = double for T = float)
template <class T>
__host__ __device__ inline T sum_term(double p);
template <>
__host__ __device__ inline double sum_term<double>(double p) { return p; }
template <>
__host__ __device__ inline float sum_term<float>(double p) { return f32_term(p); }
__host__ __device__ inline float sum_term_f(float p) { return p; }

// host image of the digit statistics of one operand (ExpStats of oz_digits.hpp)
struct ExpStatsH {
	int maxS, minLsb, maxE, nonfinite, maxSpan, maxNorm, pad[2];
};

// True when the plain FP64 split of these operands (digit statistics a of A, b of B, inner
// dimension k) is exact, so NearSum returns RN(exact), and no slice product can underflow or
// overflow: the precondition for the fast forms' bit identity with the plain form.
inline bool fast_valid(const ozblasHandle_t *oh, const ExpStatsH &a, const ExpStatsH &b, int k) {
	if (oh->nSplitMax != 0 || oh->sumMode != 1 || oh->splitShift != 1) return false;
	if (oh->splitEpsMode == 2) return false;  // Dot2 slices: the plain form's products are inexact
	if (a.nonfinite || b.nonfinite) return false;
	if (a.maxS == 0 || b.maxS == 0) return true;  // an all-zero operand: exact zeros either way
	const int rho = rho_host(k, oh->splitEpsMode);
	const int bits = 53 - rho - 1;  // every FP64 slice removes at least this many bits
	const int cap = split_cap(oh);
	if (bits <= 0) return false;
	if (a.maxSpan > bits * (cap - 3) || b.maxSpan > bits * (cap - 3)) return false;
	if (a.minLsb + b.minLsb < -1000 || a.maxE + b.maxE > 1000) return false;
	if (a.minLsb < -1000 || b.minLsb < -1000) return false;
	// the exact reconstruction holds 8 Lmax + 41 bits in at most 9 words (k_reconstruct<9>)
	if (8 * (a.maxS + b.maxS - 2) + 41 > 9 * 64) return false;
	return true;
}

const uint16_t *tau_table_host();
const uint16_t *tau_table_host_f();   // JF: the binary32 split rule
const uint16_t *tau_table_host_f2();  // JF2: the binary32 NearSum rule (log2f)
int32_t ensure_init(ozblasHandle_t *oh);

}  // namespace oz
