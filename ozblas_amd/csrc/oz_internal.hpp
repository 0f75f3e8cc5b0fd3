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
};

void batch_plan_free(void *plan);

inline Backend *be(ozblasHandle_t *oh) { return reinterpret_cast<Backend *>(oh->backend); }

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
	return (off <= (uint64_t)J[e + kTauBias]) ? e : e + 1;
}

// (short)ceil(log2(fabs(max))): 0 for zero, NaN and Inf (x86-64 conversion of +-inf/NaN -> 0)
__host__ __device__ inline int tau_of(double max, const uint16_t *J) {
	const double x = fabs(max);
	if (!(x > 0.0) || !(x <= 1.7976931348623157e308)) return 0;
	return ceil_log2_tab(x, J);
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

// cap on slices of the reference split loop (src/cpu/ozblas_split.cpp:275-277)
inline int split_cap(const ozblasHandle_t *oh) {
	int maxS = (oh->nSplitMax > 0) ? oh->nSplitMax : kNumSplitDefaultMax;
	if (oh->reproMode == 0) maxS--;
	return maxS;
}

// host image of the digit statistics of one operand (ExpStats of oz_digits.hpp)
struct ExpStatsH {
	int maxS, minLsb, maxE, nonfinite, maxSpan, pad[3];
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
int32_t ensure_init(ozblasHandle_t *oh);

}  // namespace oz
