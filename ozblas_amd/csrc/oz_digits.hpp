// oz_digits.hpp -- device routines of the fast forms shared by the single-problem kernels
// (oz_split.hip, oz_gemm_digits.hip) and the grouped batched kernels (oz_batched.hip):
//   digit statistics and balanced radix-256 digit extraction of one column,
//   the INT8 tiled-plane writer, the INT8 MFMA tile of one level bin, and the exact
//   multi-word reconstruction of one output element.
#pragma once
#include "oz_kernels.hpp"

namespace oz {

typedef int i16v __attribute__((ext_vector_type(16)));
typedef int i4v __attribute__((ext_vector_type(4)));
typedef long l2v __attribute__((ext_vector_type(2)));

template <int BS>
__device__ inline double block_max(double v, double *red) {
	// wavefront max via cross-lane shuffles, then across the waves through LDS
	for (int o = 32; o > 0; o >>= 1) v = std_max(v, __shfl_xor(v, o, 64));
	const int w = threadIdx.x / 64, l = threadIdx.x % 64;
	__syncthreads();
	if (l == 0) red[w] = v;
	__syncthreads();
	double r = red[0];
	for (int i = 1; i < BS / 64; i++) r = std_max(r, red[i]);
	return r;
}


struct ExpStats {  // global reductions written by k_digit_stats
	int maxS;        // max digit count over columns
	int minLsb;      // min exponent of the lowest set bit over all nonzero elements
	int maxE;        // max column scale exponent
	int nonfinite;   // any Inf/NaN
	int maxSpan;     // max (top exponent - lowest bit) over columns
	int maxNorm;     // max over columns of an integer upper bound on log2 ||x 2^-lowest bit||_2
	int pad[2];
};

// exponent of the lowest set bit of a finite nonzero double
__device__ inline int lsb_exp(double x) {
	const uint64_t b = d2u(x);
	const int E = (int)((b >> 52) & 0x7ff);
	uint64_t f = b & ((uint64_t(1) << 52) - 1);
	if (E == 0) return -1074 + __builtin_ctzll(f);
	f |= (uint64_t(1) << 52);
	return (E - 1075) + __builtin_ctzll(f);
}

template <int BS>
__device__ inline int block_max_i(int v, int *red) {
	for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
	const int w = threadIdx.x / 64, l = threadIdx.x % 64;
	__syncthreads();
	if (l == 0) red[w] = v;
	__syncthreads();
	int r = red[0];
	for (int i = 1; i < BS / 64; i++) r = max(r, red[i]);
	__syncthreads();
	return r;
}

// Global reductions of the column statistics, accumulated by thread 0 of a workgroup over its
// columns and folded into ExpStats once (few atomics on the shared counters).
struct ColAcc {
	int minLsb = INT32_MAX, maxSpan = INT32_MIN, maxS = 0, maxE = INT32_MIN, bad = 0, maxNorm = INT32_MIN;
};

__device__ inline void col_acc_flush(const ColAcc &a, ExpStats *st) {
	if (a.minLsb != INT32_MAX) atomicMin(&st->minLsb, a.minLsb);
	if (a.maxSpan != INT32_MIN) atomicMax(&st->maxSpan, a.maxSpan);
	if (a.maxNorm != INT32_MIN) atomicMax(&st->maxNorm, a.maxNorm);
	atomicMax(&st->maxS, a.maxS);
	if (a.maxE != INT32_MIN) atomicMax(&st->maxE, a.maxE);
	if (a.bad) atomicOr(&st->nonfinite, 1);
}

// Digit statistics of one k-contiguous column x[0..m) (whole workgroup of BS threads):
// the column scale E (|x| <= 127 * 2^(E-8) for every element) and digit count S (every element is
// a multiple of 2^(E - 8S)); thread 0 adds the column to `acc`.  Column j of the operand is
// x = sum_{s<S} d_s 2^(E - 8(s+1)) with d_s in [-128, 127].
template <int BS>
__device__ inline double block_sum(double v, double *red) {
	for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
	const int w = threadIdx.x / 64, l = threadIdx.x % 64;
	__syncthreads();
	if (l == 0) red[w] = v;
	__syncthreads();
	double r = red[0];
	for (int i = 1; i < BS / 64; i++) r += red[i];
	return r;
}

// elementwise part of the column statistics: running max |x|, max of -(lowest set bit), non-finite
__device__ __forceinline__ void col_stat_elem(double v, double &mx, int &nlo, int &bad) {
	if (!isfinite(v)) {
		bad = 1;
		return;
	}
	if (v == 0.0) return;
	mx = std_max(mx, fabs(v));
	nlo = max(nlo, -lsb_exp(v));
}

// the three workgroup reductions of the column statistics behind one barrier pair (red holds
// 2 BS / 64 ints, redd BS / 64 doubles); every thread gets the results
template <int BS>
__device__ inline void block_max3(double &mx, int &nlo, int &bad, int *red, double *redd) {
	for (int o = 32; o > 0; o >>= 1) {
		mx = std_max(mx, __shfl_xor(mx, o, 64));
		nlo = max(nlo, __shfl_xor(nlo, o, 64));
		bad = max(bad, __shfl_xor(bad, o, 64));
	}
	const int w = threadIdx.x / 64, l = threadIdx.x % 64;
	__syncthreads();
	if (l == 0) {
		redd[w] = mx;
		red[w] = nlo;
		red[BS / 64 + w] = bad;
	}
	__syncthreads();
	mx = redd[0];
	nlo = red[0];
	bad = red[BS / 64];
	for (int i = 1; i < BS / 64; i++) {
		mx = std_max(mx, redd[i]);
		nlo = max(nlo, red[i]);
		bad = max(bad, red[BS / 64 + i]);
	}
}

// Columns of up to kColReg * BS elements are loaded once, all loads issued before the first use, and
// kept in registers for the norm pass; longer ones are read in batches of 8 per thread and read
// again (from L2) for the norm.  `red` holds 2 BS / 64 ints.
constexpr int kColReg = 32;
template <int BS>
__device__ inline void digit_col_stats_acc(const double *x, int m, int *colE, int *colS, ColAcc &acc,
                                           int *red, double *redd, int *colLo = nullptr) {
	int nlo = INT32_MIN, bad = 0;  // nlo = -(min lsb)
	double mx = 0.0;
	const bool inreg = m <= kColReg * BS;
	double v[kColReg];
	if (inreg) {
#pragma unroll
		for (int u = 0; u < kColReg; u++) {
			const int i = threadIdx.x + u * BS;
			v[u] = (i < m) ? x[i] : 0.0;
		}
#pragma unroll
		for (int u = 0; u < kColReg; u++) col_stat_elem(v[u], mx, nlo, bad);
	} else {
		for (int i0 = threadIdx.x; i0 < m; i0 += 8 * BS) {
			double b[8];
#pragma unroll
			for (int u = 0; u < 8; u++) b[u] = (i0 + u * BS < m) ? x[i0 + u * BS] : 0.0;
#pragma unroll
			for (int u = 0; u < 8; u++) col_stat_elem(b[u], mx, nlo, bad);
		}
	}
	block_max3<BS>(mx, nlo, bad, red, redd);
	// residue form: an upper bound on log2 of the Euclidean norm of the column's integers x 2^-lo
	// (Cauchy-Schwarz bounds |X| by the product of two such norms).  The squares are summed as
	// (x 2^-hi)^2 <= 1, in a second pass over the column (registers or L2); terms that underflow are
	// covered by the m 2^-1000 added, the rounding of the m-term sum by the factor 1 + m 2^-52.
	double ss = 0.0;
	int hiN = 0;
	if (colLo && mx > 0.0 && !bad) {
		frexp(mx, &hiN);
		if (inreg) {  // the same order as the loop below; the zeros past m add nothing
#pragma unroll
			for (int u = 0; u < kColReg; u++) {
				const double t = ldexp(v[u], -hiN);
				ss += t * t;
			}
		} else {
			for (int i = threadIdx.x; i < m; i += BS) {
				const double t = ldexp(x[i], -hiN);
				ss += t * t;
			}
		}
		ss = block_sum<BS>(ss, redd);
	}
	if (threadIdx.x == 0) {
		if (colLo && mx > 0.0 && !bad) {
			const double up = sqrt((ss * (1.0 + m * 0x1p-52) + m * 0x1p-1000)) * (1.0 + 0x1p-50);
			int eN;
			frexp(up, &eN);  // up < 2^eN
			acc.maxNorm = max(acc.maxNorm, eN + hiN + nlo);  // log2 ||x 2^-lo|| < eN + hi - lo
		}
		int E = 0, S = 0;
		if (mx > 0.0) {
			int hi;
			frexp(mx, &hi);  // mx in [2^(hi-1), 2^hi)
			// smallest E with mx <= 127 * 2^(E-8): hi+1 when mx <= (127/128) 2^hi, else hi+2
			E = (mx <= ldexp(127.0, hi - 7)) ? hi + 1 : hi + 2;
			const int lo = -nlo;
			S = (E - lo + 7) / 8;  // 8S >= E - lo
			acc.minLsb = min(acc.minLsb, lo);
			acc.maxSpan = max(acc.maxSpan, hi - lo);
		}
		*colE = E;
		*colS = S;
		if (colLo) *colLo = (mx > 0.0) ? -nlo : 0;  // exponent of the column's lowest set bit
		acc.maxS = max(acc.maxS, S);
		acc.maxE = max(acc.maxE, E);
		if (bad) acc.bad = 1;
	}
}

// The same statistics of one column computed by one wave, with no barrier (the batched launch's
// columns are at most a few thousand long): max, lowest set bit and the norm's sum of squares reduced
// across the wave; lane 0 adds the column to `acc`.  The results are those of digit_col_stats_acc
// (the sum of squares only in another order, inside the bound's rounding margin).
__device__ inline int wave_max_i32(int v) {
	for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
	return v;
}
__device__ inline void digit_col_stats_wave(const double *x, int m, int *colE, int *colS, ColAcc &acc,
                                            int *colLo) {
	const int l = threadIdx.x % 64;
	int nlo = INT32_MIN, bad = 0;
	double mx = 0.0;
	constexpr int RW = 16;  // columns of up to RW * 64 elements stay in registers for the norm pass
	const bool inreg = m <= RW * 64;
	double v[RW];
	if (inreg) {
#pragma unroll
		for (int u = 0; u < RW; u++) v[u] = (l + u * 64 < m) ? x[l + u * 64] : 0.0;
#pragma unroll
		for (int u = 0; u < RW; u++) col_stat_elem(v[u], mx, nlo, bad);
	} else {
		for (int i0 = l; i0 < m; i0 += 8 * 64) {
			double b[8];
#pragma unroll
			for (int u = 0; u < 8; u++) b[u] = (i0 + u * 64 < m) ? x[i0 + u * 64] : 0.0;
#pragma unroll
			for (int u = 0; u < 8; u++) col_stat_elem(b[u], mx, nlo, bad);
		}
	}
	mx = wave_max_nonneg_f64(mx);
	nlo = wave_max_i32(nlo);
	bad = wave_max_i32(bad);
	double ss = 0.0;
	int hiN = 0;
	if (colLo && mx > 0.0 && !bad) {
		frexp(mx, &hiN);
		if (inreg) {
#pragma unroll
			for (int u = 0; u < RW; u++) {
				const double t = ldexp(v[u], -hiN);
				ss += t * t;
			}
		} else {
			for (int i = l; i < m; i += 64) {
				const double t = ldexp(x[i], -hiN);
				ss += t * t;
			}
		}
		ss = wave_sum_f64(ss);
	}
	if (l == 0) {
		if (colLo && mx > 0.0 && !bad) {
			const double up = sqrt((ss * (1.0 + m * 0x1p-52) + m * 0x1p-1000)) * (1.0 + 0x1p-50);
			int eN;
			frexp(up, &eN);
			acc.maxNorm = max(acc.maxNorm, eN + hiN + nlo);
		}
		int E = 0, S = 0;
		if (mx > 0.0) {
			int hi;
			frexp(mx, &hi);
			E = (mx <= ldexp(127.0, hi - 7)) ? hi + 1 : hi + 2;
			const int lo = -nlo;
			S = (E - lo + 7) / 8;
			acc.minLsb = min(acc.minLsb, lo);
			acc.maxSpan = max(acc.maxSpan, hi - lo);
		}
		*colE = E;
		*colS = S;
		if (colLo) *colLo = (mx > 0.0) ? -nlo : 0;
		acc.maxS = max(acc.maxS, S);
		acc.maxE = max(acc.maxE, E);
		if (bad) acc.bad = 1;
	}
}

// the same for one column with its reductions folded into *st at once
template <int BS>
__device__ inline void digit_col_stats(const double *x, int m, int *colE, int *colS, ExpStats *st,
                                       int *red, double *redd) {
	ColAcc acc;
	digit_col_stats_acc<BS>(x, m, colE, colS, acc, red, redd);
	if (threadIdx.x == 0) col_acc_flush(acc, st);
}

template <typename T>
__device__ inline T digit_cast(int d);
template <>
__device__ inline int8_t digit_cast<int8_t>(int d) { return (int8_t)d; }
template <>
__device__ inline _Float16 digit_cast<_Float16>(int d) { return (_Float16)(float)d; }
template <>
__device__ inline __bf16 digit_cast<__bf16>(int d) { return (__bf16)(float)d; }

// X = M + 0x80808080808080808080 (80-bit, two words) for the element's aligned integer M;
// digit q of M (from the bottom) = byte q of X - 128.  Zero (and a non-finite value, which the
// fast forms never accept) is X = the bias alone: every digit 0.
struct Digits {
	uint64_t lo, hi;
	int base;  // plane-index offset: digit q lands in plane S-1-(q+base)
};

__device__ inline Digits make_digits(double v, int E, int S) {
	Digits d{0x8080808080808080ULL, 0x8080ULL, 0};
	if (v == 0.0 || !isfinite(v)) return d;
	const uint64_t b = d2u(v);
	const int Ex = (int)((b >> 52) & 0x7ff);
	uint64_t f = b & ((uint64_t(1) << 52) - 1);
	int e0;  // v = +-f * 2^e0
	if (Ex == 0) {
		e0 = -1074;
	} else {
		f |= (uint64_t(1) << 52);
		e0 = Ex - 1075;
	}
	const int bottom = E - 8 * S;        // exponent of the unit of the lowest plane
	const int sh = e0 - bottom;          // >= 0 by the choice of S (or f has trailing zeros)
	int q0, bits;
	if (sh >= 0) {
		q0 = sh / 8;
		bits = sh % 8;
	} else {  // e0 below the grid: trailing zeros of f cover it (checked by the stats pass)
		f >>= -sh;
		q0 = 0;
		bits = 0;
	}
	const int64_t M = (int64_t)(f << bits) * ((b >> 63) ? -1 : 1);  // |M| < 2^61
	// 128-bit X = M + K, K = 0x80 in bytes 0..9
	const uint64_t Klo = 0x8080808080808080ULL, Khi = 0x8080ULL;
	const uint64_t mlo = (uint64_t)M, mhi = (M < 0) ? ~0ULL : 0ULL;
	const uint64_t lo = mlo + Klo;
	const uint64_t carry = lo < mlo ? 1 : 0;
	d.lo = lo;
	d.hi = mhi + Khi + carry;
	d.base = q0;
	return d;
}

__device__ inline int digit_at(const Digits &d, int q) {  // q relative to the element's base
	if (q < 0 || q >= 10) return 0;
	const uint64_t w = (q < 8) ? (d.lo >> (8 * q)) : (d.hi >> (8 * (q - 8)));
	return (int)(w & 0xff) - 128;
}

// INT8 planes in the tiled layout of the MFMA kernel: element (column j, k index q) of plane s at
//   s * pstride + (j / 16) * 16 * ldp + (q / 64) * 1024 + (j % 16) * 64 + q % 64.
// One 256-thread workgroup writes the 16 x 64 tile (jb, kt) of every plane (1 KiB per plane, 4
// bytes per thread), with zero digits for columns >= n and k >= m.
__device__ inline void digit_write_tile_i8(int m, int n, const double *in, int64_t ldi,
                                           const int *colE, int S, int8_t *planes, int64_t ldp,
                                           int64_t pstride, int kt, int jb) {
	const int t = threadIdx.x;
	const int j = jb * 16 + t / 16;
	const int q0 = kt * 64 + 4 * (t % 16);
	const int E = (j < n) ? colE[j] : 0;
	const double *x = in + (int64_t)(j < n ? j : 0) * ldi;
	const Digits d0 = make_digits(j < n && q0 + 0 < m ? x[q0 + 0] : 0.0, E, S);
	const Digits d1 = make_digits(j < n && q0 + 1 < m ? x[q0 + 1] : 0.0, E, S);
	const Digits d2 = make_digits(j < n && q0 + 2 < m ? x[q0 + 2] : 0.0, E, S);
	const Digits d3 = make_digits(j < n && q0 + 3 < m ? x[q0 + 3] : 0.0, E, S);
	int8_t *dst = planes + (int64_t)jb * 16 * ldp + (int64_t)kt * 1024 + 4 * t;
	for (int s = 0; s < S; s++) {
		const int q = S - 1 - s;
		const uint32_t word = (uint32_t)(uint8_t)digit_at(d0, q - d0.base) |
		                      (uint32_t)(uint8_t)digit_at(d1, q - d1.base) << 8 |
		                      (uint32_t)(uint8_t)digit_at(d2, q - d2.base) << 16 |
		                      (uint32_t)(uint8_t)digit_at(d3, q - d3.base) << 24;
		*(uint32_t *)(dst + (int64_t)s * pstride) = word;
	}
}

// ---- INT8 MFMA tile ---------------------------------------------------------------------------
constexpr int I8_BK = 128;  // k bytes per stage = one LDS row (unpadded, chunk-swizzled)

// LDS rows hold 8 chunks of 16 bytes; logical chunk c of row r sits at physical chunk
// c ^ i8_swz(r), i8_swz(r) = h[(r >> 1) & 7] ^ 4 (r & 1) with h = {0,5,2,7,1,4,3,6}.  A fragment
// read touches consecutive rows at one chunk, a stage write touches 4 rows x 4 chunks; with this
// swizzle neither has a bank conflict for 32 or 64 banks of 4 bytes (8 or 16 lanes of a b128
// access per cycle), without padding.
__device__ inline int i8_swz(int r) {
	return ((0x63417250 >> (4 * ((r >> 1) & 7))) & 7) ^ (4 * (r & 1));
}

// LDS-only workgroup barrier: orders the LDS traffic of the waves without waiting for the
// global loads that are still in flight for later stages.
__device__ inline void lds_barrier() {
	__builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
	__builtin_amdgcn_s_barrier();
	__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// One level bin of one output tile: T[j * ldt + i] = sum over the bin's pairs p and k in [k0, k1)
// of the digit products, for the (64 TM) x (64 TN) tile at (i0, j0).  Pair p multiplies plane
// s0 + p of A with plane t0 - p of B.  The stage buffers are two distinct __shared__ objects of
// (64 TM + 64 TN) * I8_BK bytes each (so that alias analysis lets the LDS writes of one stage move
// past the fragment reads of the other).  Called by every thread of a 256-thread workgroup.
struct I8Tile {
	const int8_t *A, *B;  // plane 0 of each operand (tiled INT8 layout)
	int64_t apstride, bpstride, ldp;
	int s0, t0, npairs, k0, k1;
	int i0, j0;
	int32_t *T;  // result plane of the bin
	int64_t ldt;
	// residue form: the pairs are moduli s0 .. s0 + npairs - 1 (plane t of A with plane t of B), and
	// after each modulus's k range acc mod p goes to plane t of R (t * rstride) instead of T
	uint8_t *R;
	int64_t rstride;
	const ResConst *rc;
};

// v mod p in [0, p) for |v| <= 2^30 and 3 <= p <= 256, in binary64 without a correction step:
// q = floor(fma(v, RN(1/p), 2^-10)) is floor(v / p) exactly -- the fma's error, below
// |v/p| 2^-53 + 2^-23 < 2^-22, and the offset 2^-10 stay inside the gap of at least 1/p >= 2^-8 that
// separates a non-integer v/p from the integer above it, and the offset lifts an integer v/p off the
// one below -- so r = v - q p, exact in one fma, is the canonical residue.
__device__ __forceinline__ int mod_f64(int v, double p, double invp) {
	const double vd = (double)v;
	const double q = floor(fma(vd, invp, 0x1p-10));
	return (int)fma(-q, p, vd);
}

// Product residues of the residue form: one byte per element, four consecutive columns of a row
// in one 32-bit word -- element (i, j) of a plane at (j / 4) * 4 * ldr + 4 * i + j % 4, so a lane's
// four MFMA result columns are one word and a wave's words are contiguous.
__device__ __forceinline__ int64_t res_word_offset(int i, int j, int64_t ldr) {
	return (int64_t)(j >> 2) * 4 * ldr + 4 * (int64_t)i;
}

template <int TM, int TN, bool RES>
__device__ __forceinline__ void i8_tile(const I8Tile &g, char *stage0, char *stage1) {
	constexpr int TBM = 64 * TM, TBN = 64 * TN;
	constexpr int NA = 2 * TM, NB = 2 * TN;  // 16-byte chunks per thread per stage
	const int i0 = g.i0, j0 = g.j0;
	const int64_t ldp = g.ldp, apstride = g.apstride, bpstride = g.bpstride;
	const int t = threadIdx.x, w = t / 64, l = t % 64;
	const int wm = (w % 2) * 32 * TM, wn = (w / 2) * 32 * TN;
	// Stage loads: thread t moves chunk h (0 <= h < NA) = 16 bytes of row 16 (t/128) + l/4 + 32 h,
	// k bytes 64 ((t/128)%2... = kk64) + 16 (l%4); a wave's instruction reads one contiguous KiB.
	const int kk64 = (t / 64) % 2;
	const int64_t rbs = 16 * ldp;  // bytes per 16-row block of a plane
	const char *Ab = (const char *)g.A + (int64_t)g.s0 * apstride + (int64_t)(g.k0 / I8_BK) * 2048 +
	                 (int64_t)(i0 / 16 + t / 128) * rbs + kk64 * 1024 + 16 * l;
	const char *Bb = (const char *)g.B + (int64_t)g.t0 * bpstride + (int64_t)(g.k0 / I8_BK) * 2048 +
	                 (int64_t)(j0 / 16 + t / 128) * rbs + kk64 * 1024 + 16 * l;
	const int wrow = 16 * (t / 128) + l / 4;  // + 32 h (the swizzle only sees bits 1..3)
	const int lofs = wrow * I8_BK + 16 * ((4 * kk64 + l % 4) ^ i8_swz(wrow));
	auto stage = [&](int buf) -> char * { return buf ? stage1 : stage0; };
	// fragment reads: row wm + 32 mi + l%32 (swizzle depends on l%32 only), chunk 2 kk + l/32
	const int fsw = i8_swz(l % 32);
	const int frow = (l % 32) * I8_BK;
	i16v acc[TM][TN];
#pragma unroll
	for (int a = 0; a < TM; a++)
#pragma unroll
		for (int b = 0; b < TN; b++)
#pragma unroll
			for (int r = 0; r < 16; r++) acc[a][b][r] = 0;
	const int kst = (g.k1 - g.k0) / I8_BK;
	const int total = g.npairs * kst;
	if (total <= 0) return;
	// next stage to load: pair lpair, k stage lkt (pair p uses plane s0 + p of A and plane t0 - p
	// of B); past the end it wraps to stage (0, 0)
	int lkt = 0, lpair = 0;
	auto advance = [&]() {  // 32-bit selects only: a branch here would split the scheduling region
		const int nk = lkt + 1;
		const int np = lpair + (nk == kst ? 1 : 0);
		lkt = (nk == kst) ? 0 : nk;
		lpair = (np == g.npairs) ? 0 : np;
	};
	// two register sets for the global stage loads (ext-vector type: HIP's uint4 struct would
	// keep the arrays out of registers)
	i4v ra0[NA], rb0[NB], ra1[NA], rb1[NB];
	auto gload = [&](i4v(&ra)[NA], i4v(&rb)[NB]) {
		const char *Ap = Ab + (int64_t)lpair * apstride + lkt * 2048;
		const char *Bp = Bb + (int64_t)(RES ? lpair : -lpair) * bpstride + lkt * 2048;
#pragma unroll
		for (int h = 0; h < NA; h++) ra[h] = *(const i4v *)(Ap + 2 * h * rbs);
#pragma unroll
		for (int h = 0; h < NB; h++) rb[h] = *(const i4v *)(Bp + 2 * h * rbs);
		advance();
	};
	auto lstore = [&](const i4v(&ra)[NA], const i4v(&rb)[NB], int buf) {
		char *a = stage(buf) + lofs, *b = stage(buf) + TBM * I8_BK + lofs;
#pragma unroll
		for (int h = 0; h < NA; h++) *(i4v *)(a + 32 * h * I8_BK) = ra[h];
#pragma unroll
		for (int h = 0; h < NB; h++) *(i4v *)(b + 32 * h * I8_BK) = rb[h];
	};
	l2v fa0[TM], fb0[TN], fa1[TM], fb1[TN];
	// fragment read number q of k-step kk (order of first use by the MFMA loop below:
	// A0, B0 .. B(TN-1), A1 .. A(TM-1))
	auto fread1 = [&](l2v(&fa)[TM], l2v(&fb)[TN], int buf, int kk, int q) {
		const int co = 16 * ((2 * kk + l / 32) ^ fsw);
		const char *a = stage(buf) + wm * I8_BK + frow + co;
		const char *b = stage(buf) + TBM * I8_BK + wn * I8_BK + frow + co;
		if (q == 0)
			fa[0] = *(const l2v *)a;
		else if (q <= TN)
			fb[q - 1] = *(const l2v *)(b + 32 * (q - 1) * I8_BK);
		else
			fa[q - TN] = *(const l2v *)(a + 32 * (q - TN) * I8_BK);
	};
	auto fread = [&](l2v(&fa)[TM], l2v(&fb)[TN], int buf, int kk) {
#pragma unroll
		for (int q = 0; q < TM + TN; q++) fread1(fa, fb, buf, kk, q);
	};
	constexpr int MF = TM * TN, RD = TM + TN, NG = NA + NB;  // per k-step: MFMAs, reads; per stage: loads
	// One stage: buffer `cur` holds stage s, registers (raN, rbN) hold stage s + 1 and
	// (raF, rbF) are free.  Each k-step issues its MF MFMAs one slot at a time; after MFMA i come
	// the fragment reads for the next k-step (spread over the slots), a quarter of the global loads
	// of stage s + 2 into the free registers, and in k-steps 1 and 2 half of the LDS writes of
	// stage s + 1 into the other buffer (its loads were issued one stage earlier).  One barrier
	// precedes k-step 3, whose slots read the next stage's first fragments.  sched_barrier keeps
	// the slots in this order.
	auto step = [&](i4v(&raF)[NA], i4v(&rbF)[NB], const i4v(&raN)[NA], const i4v(&rbN)[NB], int cur) {
		const char *Ap = Ab + (int64_t)lpair * apstride + lkt * 2048;
		const char *Bp = Bb + (int64_t)(RES ? lpair : -lpair) * bpstride + lkt * 2048;
		advance();
		char *wa = stage(cur ^ 1) + lofs, *wb = stage(cur ^ 1) + TBM * I8_BK + lofs;
#pragma unroll
		for (int ks = 0; ks < 4; ks++) {
			if (ks == 3) lds_barrier();
			l2v(&ca)[TM] = (ks % 2 == 0) ? fa0 : fa1;
			l2v(&cb)[TN] = (ks % 2 == 0) ? fb0 : fb1;
			l2v(&na)[TM] = (ks % 2 == 0) ? fa1 : fa0;
			l2v(&nb)[TN] = (ks % 2 == 0) ? fb1 : fb0;
#pragma unroll
			for (int i = 0; i < MF; i++) {
				const int mi = i / TN, nj = i % TN;
				acc[mi][nj] = __builtin_amdgcn_mfma_i32_32x32x32_i8(cb[nj], ca[mi], acc[mi][nj], 0, 0, 0);
#pragma unroll
				for (int q = 0; q < RD; q++)
					if ((q * MF) / RD == i) {
						if (ks < 3)
							fread1(na, nb, cur, ks + 1, q);
						else
							fread1(na, nb, cur ^ 1, 0, q);
					}
#pragma unroll
				for (int g = 0; g < NG / 4; g++)
					if ((g * MF) / (NG / 4) == i) {
						const int v = ks * (NG / 4) + g;
						if (v < NA)
							raF[v] = *(const i4v *)(Ap + 2 * v * rbs);
						else
							rbF[v - NA] = *(const i4v *)(Bp + 2 * (v - NA) * rbs);
					}
				if (ks == 1 || ks == 2) {
#pragma unroll
					for (int g = 0; g < NG / 2; g++)
						if ((g * MF) / (NG / 2) == i) {
							const int d = (ks - 1) * (NG / 2) + g;
							if (d < NA)
								*(i4v *)(wa + 32 * d * I8_BK) = raN[d];
							else
								*(i4v *)(wb + 32 * (d - NA) * I8_BK) = rbN[d - NA];
						}
				}
				__builtin_amdgcn_sched_barrier(0);
			}
		}
	};
	// residue form: the product of modulus s0 + q is complete after its kst stages; its residues go
	// to plane s0 + q (four columns per 32-bit word) and the accumulators restart at zero.  With
	// 2 x 2 wave tiles the flush runs inside the flattened (modulus, k) loop, so a workgroup streams
	// several moduli through one pipeline; the 4 x 4 tile has no registers left for that inside the
	// loop (npairs must be 1) and flushes its one modulus after it.
	constexpr bool FLAT = RES && TM * TN <= 4;
	int left = kst, mq = 0;
	auto flush = [&]() {
		if constexpr (RES) {
			if (--left > 0) return;
			left = kst;
			const int t = g.s0 + mq++;
			const double p = (double)g.rc->p[t];
			const double invp = g.rc->invpd[t];
			uint8_t *Rb = g.R + (int64_t)t * g.rstride;
			// |acc| <= 2^14 k <= 2^30 (chunks of at most 65536 k)
#pragma unroll
			for (int mi = 0; mi < TM; mi++)
#pragma unroll
				for (int nj = 0; nj < TN; nj++)
#pragma unroll
					for (int r4 = 0; r4 < 4; r4++) {
						const int i = i0 + wm + 32 * mi + l % 32;
						const int j = j0 + wn + 32 * nj + 8 * r4 + 4 * (l / 32);
						uint32_t w = 0;
#pragma unroll
						for (int c = 0; c < 4; c++) {
							w |= (uint32_t)mod_f64(acc[mi][nj][4 * r4 + c], p, invp) << (8 * c);
							acc[mi][nj][4 * r4 + c] = 0;
						}
						*(uint32_t *)(Rb + res_word_offset(i, j, g.ldt)) = w;
					}
		}
	};
	// prologue: stage 0 -> buffer 0, stage 1 -> registers set 1
	gload(ra0, rb0);
	gload(ra1, rb1);
	lstore(ra0, rb0, 0);
	lds_barrier();
	fread(fa0, fb0, 0, 0);
	int s = 0;
	for (; s + 1 < total; s += 2) {
		step(ra0, rb0, ra1, rb1, 0);
		if constexpr (FLAT) flush();
		step(ra1, rb1, ra0, rb0, 1);
		if constexpr (FLAT) flush();
	}
	if (s < total) {
		step(ra0, rb0, ra1, rb1, 0);
		if constexpr (FLAT) flush();
	}
	if constexpr (RES) {
		if constexpr (!FLAT) {
			left = 1;
			flush();
		}
		return;  // every modulus flushed
	}
	// D[p][q]: p = B-side (column j) = 8*(r/4) + 4*(l/32) + r%4, q = A-side (row i) = l%32
	int32_t *Tb = g.T;
	const int64_t ldt = g.ldt;
#pragma unroll
	for (int mi = 0; mi < TM; mi++)
#pragma unroll
		for (int nj = 0; nj < TN; nj++)
#pragma unroll
			for (int r = 0; r < 16; r++) {
				const int i = i0 + wm + 32 * mi + l % 32;
				const int j = j0 + wn + 32 * nj + 8 * (r / 4) + 4 * (l / 32) + r % 4;
				Tb[(int64_t)j * ldt + i] = acc[mi][nj][r];
			}
}


// ---- exact reconstruction ------------------------------------------------------------------
template <int NW>
struct BigInt {
	uint64_t w[NW];
};

template <int NW>
__host__ __device__ inline void big_add_shifted(BigInt<NW> &a, int64_t v, int sh) {
	// a += v * 2^sh (two's complement, NW 64-bit words)
	const int wi = sh >> 6, b = sh & 63;
	const uint64_t ext = (v < 0) ? ~0ULL : 0ULL;
	const uint64_t lo = (uint64_t)v << b;
	const uint64_t hi = b ? (uint64_t)(v >> (64 - b)) : ext;
	uint64_t carry = 0;
#pragma unroll
	for (int q = 0; q < NW; q++) {
		uint64_t add;
		if (q < wi)
			continue;
		else if (q == wi)
			add = lo;
		else if (q == wi + 1)
			add = hi;
		else
			add = ext;
		const uint64_t s1 = a.w[q] + add;
		const uint64_t c1 = s1 < add ? 1 : 0;
		const uint64_t s2 = s1 + carry;
		const uint64_t c2 = s2 < carry ? 1 : 0;
		a.w[q] = s2;
		carry = c1 | c2;
	}
}

template <int NW>
__host__ __device__ inline uint64_t big_word(const BigInt<NW> &a, int q) {
	uint64_t r = 0;
#pragma unroll
	for (int i = 0; i < NW; i++)
		if (i == q) r = a.w[i];
	return r;
}

// RN-even of N * 2^e0 (N non-negative, NW words)
template <int NW>
__host__ __device__ inline double big_round(const BigInt<NW> &a, int e0) {
	int top = -1;
#pragma unroll
	for (int q = 0; q < NW; q++)
		if (a.w[q] != 0) top = q;
	if (top < 0) return 0.0;
	const uint64_t tw = big_word(a, top);
	const int p = 64 * top + 63 - __builtin_clzll(tw);  // msb index
	int r = p - 52;
	if (r < -1074 - e0) r = -1074 - e0;
	if (r <= 0) {  // every bit kept: N < 2^53
		return ldexp((double)a.w[0], e0);
	}
	// mantissa = bits [r, p]
	const int wq = r >> 6, wb = r & 63;
	uint64_t mant = big_word(a, wq) >> wb;
	if (wb) mant |= big_word(a, wq + 1) << (64 - wb);
	const int nb = p - r + 1;  // <= 53
	if (nb < 64) mant &= ((uint64_t(1) << nb) - 1);
	// round bit r-1 and sticky bits below it
	const int rb = r - 1;
	const uint64_t rword = big_word(a, rb >> 6);
	const int rbit = (int)((rword >> (rb & 63)) & 1);
	bool sticky = (rword & ((uint64_t(1) << (rb & 63)) - 1)) != 0;
#pragma unroll
	for (int q = 0; q < NW; q++)
		if (q < (rb >> 6) && a.w[q] != 0) sticky = true;
	if (rbit && (sticky || (mant & 1))) mant++;
	return ldexp((double)mant, r + e0);
}

// Exact level sums of the fast forms.  N = sum_b T_b 2^(8 (Lmax - L_b)) over int32 planes (element
// at Te[b * tstride], bins in ascending level order) is accumulated in an NW-word two's-complement
// integer.  Levels are first folded four at a time into an int64 (|T| < 2^31, so up to 128 terms
// of weight <= 2^24 stay below 2^63) and each fold is added to N once.
template <int NW>
__host__ __device__ inline void big_zero(BigInt<NW> &a) {
#pragma unroll
	for (int q = 0; q < NW; q++) a.w[q] = 0;
}

template <int NW>
__host__ __device__ inline void recon_accumulate(BigInt<NW> &a, const int32_t *Te, int64_t tstride,
                                                 const int *binLevel, int b0, int b1, int Lmax) {
	int64_t fold = 0;
	int group = -1, top = 0, terms = 0;
	for (int b = b0; b < b1; b++) {
		const int L = binLevel[b];
		if (L / 4 != group || terms == 128) {
			if (fold != 0) big_add_shifted<NW>(a, fold, 8 * (Lmax - top));
			fold = 0;
			terms = 0;
			group = L / 4;
			top = (4 * group + 3 < Lmax) ? 4 * group + 3 : Lmax;  // unit of the fold: 2^(8 (Lmax - top))
		}
		fold += (int64_t)Te[(int64_t)b * tstride] * ((int64_t)1 << (8 * (top - L)));
		terms++;
	}
	if (fold != 0) big_add_shifted<NW>(a, fold, 8 * (Lmax - top));
}

// a += sgn * r (NW-word two's complement)
template <int NW>
__host__ __device__ inline void big_add_signed(BigInt<NW> &a, const uint64_t *r, int sgn) {
	uint64_t carry = (sgn < 0) ? 1 : 0;  // a + ~r + 1 for subtraction
#pragma unroll
	for (int q = 0; q < NW; q++) {
		const uint64_t add = (sgn < 0) ? ~r[q] : r[q];
		const uint64_t s1 = a.w[q] + add;
		const uint64_t c1 = s1 < add ? 1 : 0;
		const uint64_t s2 = s1 + carry;
		const uint64_t c2 = s2 < carry ? 1 : 0;
		a.w[q] = s2;
		carry = c1 | c2;
	}
}

// RN-even of N * 2^e0 for a signed N
template <int NW>
__host__ __device__ inline double big_finish(BigInt<NW> a, int e0) {
	const bool neg = (a.w[NW - 1] >> 63) != 0;
	if (neg) {  // two's complement negate
		uint64_t carry = 1;
#pragma unroll
		for (int q = 0; q < NW; q++) {
			const uint64_t v = ~a.w[q] + carry;
			carry = (carry && v == 0) ? 1 : 0;
			a.w[q] = v;
		}
	}
	const double t = big_round<NW>(a, e0);
	return neg ? -t : t;
}

// Exact value of one output element: fma(alpha, RN(N 2^e0), beta * c) over all the bins.
// The certified rounding in binary64 when it is clear-cut (k_reconstruct_check with the bound R of
// the levels not computed, the full reconstruction with R = 0): the bin terms
// T_b 2^(8 (Lmax - L_b)) (exact binary64 values) are summed in double-double, hi + lo with an error
// below 4 nb^2 u^2 sum|terms| (u = 2^-53: TwoSum steps, lo accumulated with one rounding per step),
// and renormalised to hi2 + lo2.  When |lo2| plus that error plus R stays below half the smaller
// spacing of binary64 next to hi2 (with a 2^-20 margin for the rounding of this test itself), every
// N in [N1 - R, N1 + R] rounds to hi2; hi2 2^e0 is then exact when its exponent is well inside the
// normal range, so *t = hi2 2^e0 is the value the exact certification returns.  Otherwise false,
// and the exact multi-word path decides.
__host__ __device__ inline bool recon_certify_fast(const int32_t *Te, int64_t tstride, const int *binLevel,
                                          int nb1, int Lmax, double rUp, int e0, double *t) {
	double hi = 0.0, lo = 0.0, asum = 0.0;
	for (int b = 0; b < nb1; b++) {
		const double tb = ldexp((double)Te[(int64_t)b * tstride], 8 * (Lmax - binLevel[b]));
		const double s = hi + tb;
		const double bb = s - hi;
		const double err = (hi - (s - bb)) + (tb - bb);
		hi = s;
		lo = lo + err;
		asum = asum + fabs(tb);
	}
	const double h2 = hi + lo;
	const double bb = h2 - hi;
	const double l2 = (hi - (h2 - bb)) + (lo - bb);
	if (h2 == 0.0) return false;
	int ex;
	const double f = frexp(h2, &ex);  // |f| in [1/2, 1)
	if (ex + e0 < -1020 || ex + e0 > 1022) return false;
	const double nb = (double)nb1;
	const double delta = 4.0 * nb * nb * 0x1p-106 * asum * (1.0 + 0x1p-40);
	const double half = ldexp(1.0, ex - ((fabs(f) == 0.5) ? 55 : 54));  // half the smaller spacing
	if (fabs(l2) + delta + rUp >= half * (1.0 - 0x1p-20)) return false;
	*t = ldexp(h2, e0);
	return true;
}

template <int NW>
__host__ __device__ inline double recon_elem(const int32_t *Te, int64_t tstride, const int *binLevel,
                                            int nbins, int Lmax, int e0, double alpha, double beta,
                                            double c) {
	double t;
	if (recon_certify_fast(Te, tstride, binLevel, nbins, Lmax, 0.0, e0, &t)) return fma(alpha, t, beta * c);
	BigInt<NW> a;
	big_zero<NW>(a);
	recon_accumulate<NW>(a, Te, tstride, binLevel, 0, nbins, Lmax);
	return fma(alpha, big_finish<NW>(a, e0), beta * c);
}

// Certified rounding from the leading levels: with N1 the sum over the computed bins and R a
// bound on |N - N1| (the levels not computed), RN(N 2^e0) is known when RN((N1 - R) 2^e0) and
// RN((N1 + R) 2^e0) agree, RN being monotone.  Returns true and sets *t in that case.
template <int NW>
__host__ __device__ inline bool recon_certify(const BigInt<NW> &n1, const uint64_t *R, int e0,
                                             double *t) {
	BigInt<NW> lo = n1, hi = n1;
	big_add_signed<NW>(lo, R, -1);
	big_add_signed<NW>(hi, R, +1);
	const double tl = big_finish<NW>(lo, e0), th = big_finish<NW>(hi, e0);
	*t = tl;
	return __builtin_bit_cast(uint64_t, tl) == __builtin_bit_cast(uint64_t, th);
}

}  // namespace oz
