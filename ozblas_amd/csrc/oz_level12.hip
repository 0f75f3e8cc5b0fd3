// oz_level12.hip -- bandwidth-bound kernels of the plain form: vector split, slice DOT, CSR
// slice products with fused NearSum/Fsum, AXPY.
#include <algorithm>

#include "oz_kernels.hpp"
#include "oz_nearsum.hpp"

namespace oz {

// block reductions: the wave step on the DPP network (values non-negative for the maximum; exact
// sums for the sum, so any order), then the waves through LDS
template <int BS>
__device__ inline double block_max_d(double v, double *red) {
	v = wave_max_nonneg_f64(v);
	const int w = threadIdx.x / 64, l = threadIdx.x % 64;
	__syncthreads();
	if (l == 0) red[w] = v;
	__syncthreads();
	double r = red[0];
	for (int i = 1; i < BS / 64; i++) r = std_max(r, red[i]);
	return r;
}

template <int BS>
__device__ inline double block_sum_d(double v, double *red) {
	v = wave_sum_f64(v);
	const int w = threadIdx.x / 64, l = threadIdx.x % 64;
	__syncthreads();
	if (l == 0) red[w] = v;
	__syncthreads();
	double r = red[0];
	for (int i = 1; i < BS / 64; i++) r += red[i];
	return r;
}

// ---------------------------------------------------------------------------------------------
// Vector split without grid barriers.
//
// With sigma_s = 0.75 * 2^(rho+tau_s) (rho >= 3, sigma_s normal) the reference's slice
// (r + sigma) - sigma is r rounded to the nearest multiple of u_s = ulp(sigma_s) = 2^(rho+tau_s-53),
// ties to even (sigma_s / u_s = 3 * 2^51 is even and |r| <= 2^(rho+tau_s-2) keeps the sum in the
// binade of sigma_s).  The units shrink (u_{s+1} <= u_s 2^(rho-54)) and ties-to-even commutes with
// shifts by even multiples, so the remainder after slice s is the centred residue
//   r_{s+1} = cmod(x, u_s) = x - u_s * rne(x / u_s)
// of the INPUT, whatever the earlier units were.  Hence the remainder maxima the reference reduces
// level by level are values of one function of the unit, F(k) = max_i |cmod(x_i, 2^k)|:
// M_{s+1} = F(k_s), k_s = rho + tau(M_s) - 53, M_0 = max|x|.  k_vsplit_scan tabulates F over all k
// in one pass (per element only k in (lsb, msb+1] are non-trivial: below, the residue is 0; above,
// it is x itself), its last workgroup walks the chain of levels, and k_vsplit_write stores the
// slices slice_s = cmod(x, u_{s-1}) - cmod(x, u_s) (exact).  Inputs outside the model (non-finite,
// |x| >= 2^1022, rho < 3 or above the precision, a sigma outside the normal range, a non
// power-of-two shift, more levels than VsPlan holds) set plan.bad and are split by
// workgroup 0 of k_vsplit_write instead, which runs the reference loop.  The binary32 split
// is the same with 24 for 53 (unit 2^(rho+tau-24), float sigma in its normal range).
// ---------------------------------------------------------------------------------------------
// a load at the device scope (values other workgroups wrote with atomics or before their release)
__device__ inline unsigned long long vs_load(const unsigned long long *p) {
	return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

constexpr int kVsOff = 1100;          // F/G tables cover k in [-1100, 1099]
constexpr int kVsK = 2200;
constexpr int kVsBlocks = 256;
constexpr int kVsBS = 512;

constexpr int kVsW = 64;   // k values a workgroup keeps in registers when its range is that narrow
constexpr int kVsEC = 8;   // elements per thread kept in registers between the scan's two passes

// x = m 2^eb with m < 2^53 an integer; lo/hi: the range of k with a non-trivial residue (below lo
// it is 0, above hi it is x itself); m = 0 for zero and for inputs outside the model (bad = 1)
struct VsElem {
	uint64_t m;
	int eb, lo, hi;
};
__device__ inline VsElem vs_decompose(double v, int &bad) {
	VsElem r{0ull, -1074, INT32_MAX, INT32_MIN};
	const uint64_t b = d2u(v);
	const int E = (int)((b >> 52) & 0x7ff);
	uint64_t m = b & ((uint64_t(1) << 52) - 1);
	if (E == 0x7ff || E >= 1023 + 1022) {
		bad = 1;
		return r;
	}
	if (v == 0.0) return r;
	int eb = -1074;
	if (E != 0) {
		m |= uint64_t(1) << 52;
		eb = E - 1075;
	}
	r.m = m;
	r.eb = eb;
	r.lo = eb + __builtin_ctzll(m) + 1;
	r.hi = eb + (64 - __builtin_clzll(m)) + 1;
	return r;
}

// max-accumulates the keys of x for k = w0 .. w0 + kVsW - 1 into Fr: with y = x / 2^k,
// |y - rint(y)| = |cmod(x, 2^k)| / 2^k (zero below the element's lowest bit, |x| / 2^k above its
// msb + 1).  Every element of a window is within kVsW binades of w0, so y stays a normal binary64
// below 2^(kVsW + 2) and each halving is exact.
__device__ __forceinline__ void vs_window_keys(double v, int w0, double (&Fr)[kVsW]) {
	int dummy = 0;
	const VsElem q = vs_decompose(v, dummy);
	if (q.m == 0) return;
	double y = ldexp(v, -w0);
#pragma unroll
	for (int kk = 0; kk < kVsW; kk++) {
		Fr[kk] = fmax(Fr[kk], fabs(y - rint(y)));
		y *= 0.5;
	}
}

__device__ inline double vs_cmod(double x, int k) {
	if (x == 0.0) return 0.0;
	const uint64_t b = d2u(x);
	const int E = (int)((b >> 52) & 0x7ff);
	const uint64_t f = b & ((uint64_t(1) << 52) - 1);
	int e, q;
	if (E == 0) {
		e = -1074 + (63 - __builtin_clzll(f));
		q = -1074 + __builtin_ctzll(f);
	} else {
		e = E - 1023;
		q = e - 52 + __builtin_ctzll(f | (uint64_t(1) << 52));
	}
	if (k <= q) return 0.0;
	if (k > e + 1) return x;
	const double n = rint(ldexp(x, -k));
	return x - ldexp(n, k);
}

// TI = float: the binary32 split on the same tables (a float converts exactly; its residues are
// floats), with the binary32 tau rule and unit 2^(rho + tau - 24)
template <int BS, class TI>
__device__ void vs_walk(int nwg, unsigned long long *Fg, const int2 *range, int rho, int cap, int reproMode,
                        const uint16_t *J, VsPlan *plan, int *info, int n, const TI *x, double *slices,
                        int64_t lds, int rhoIn, int shift, int fallback);

// the table scan; its last workgroup then walks the levels (vs_walk)
template <class TI>
__device__ int vs_verify(unsigned long long mb, int lof, int bad, const unsigned long long *Fb, int rho,
                         int cap, int reproMode, const uint16_t *J, VsPlan *plan, int *info, bool write);
template <bool ATOMIC>
__device__ inline void vs_pre_sum(const VsPre *pre, unsigned long long &maxBits, int &loOff, int &bad);

// The second vector of a launch that splits two vectors of one length together (the rows
// blockIdx.y == 1 of the grid): its data and its own reductions, scratch, plan and slices.
template <class TI>
struct VsSecond {
	const TI *x;
	VsPre *pre, *preNext;
	unsigned long long *Fg;
	int2 *range;
	unsigned int *done;
	VsPlan *plan;
	int *info;
	double *slices;
};

// Every workgroup first reduces the predicted chain's slots and walks it (the same result in each;
// workgroup 0 writes the plan).  When the chain holds, that is all: no table.  Otherwise the table
// scan, and its last workgroup walks the levels (vs_walk).
template <int BS, class TI>
__global__ __launch_bounds__(BS) void k_vsplit_scan(int n, const TI *x, unsigned long long *Fg,
                                                    int2 *range, unsigned int *done, const int *stopAt,
                                                    int it, const VsPre *pre, int rho, int cap,
                                                    int reproMode, const uint16_t *J, VsPlan *plan,
                                                    int *info, double *slices, int64_t lds, int rhoIn,
                                                    int shift, int fallback, VsSecond<TI> v2) {
	if (stopAt && it >= *stopAt) return;
	if (blockIdx.y == 1) {
		x = v2.x;
		Fg = v2.Fg;
		range = v2.range;
		done = v2.done;
		pre = v2.pre;
		plan = v2.plan;
		info = v2.info;
		slices = v2.slices;
	}
	{
		__shared__ unsigned long long Fp[kVsPred], mbS;
		__shared__ int lofS, badS, needS;
		const int t = threadIdx.x;
		if (t < kVsPred) {
			unsigned long long f = 0ull;
			for (int q = 0; q < kVsSpread; q++) f = std::max(f, pre->slot[q].F[t]);
			Fp[t] = f;
		} else if (t == 64) {
			vs_pre_sum<false>(pre, mbS, lofS, badS);
		}
		__syncthreads();
		if (t == 0) {
			needS = vs_verify<TI>(mbS, lofS, badS, Fp, rho, cap, reproMode, J, plan, info, blockIdx.x == 0);
			if (blockIdx.x == 0 && needS) plan->needScan = 1;
		}
		__syncthreads();
		if (!needS) return;  // the predicted chain held: no table needed
	}
	__shared__ unsigned long long Fb[kVsK];
	__shared__ int kminB, kmaxB, badB;
	__shared__ int red[2][BS / 64];
	for (int k = threadIdx.x; k < kVsK; k += BS) Fb[k] = 0ull;
	if (threadIdx.x == 0) badB = 0;
	__syncthreads();
	const int lane = threadIdx.x % 64;
	// pass 1: the range of k of this workgroup (elements kept in registers for pass 2)
	const int64_t stride = (int64_t)gridDim.x * BS;
	const int64_t i0 = (int64_t)blockIdx.x * BS + threadIdx.x;
	double xc[kVsEC];
	int tlo = INT32_MAX, thi = INT32_MIN, tbad = 0;
#pragma unroll
	for (int e = 0; e < kVsEC; e++) {
		const int64_t i = i0 + e * stride;
		xc[e] = (i < n) ? (double)x[i] : 0.0;
		const VsElem q = vs_decompose(xc[e], tbad);
		tlo = min(tlo, q.lo);
		thi = max(thi, q.hi);
	}
	for (int64_t i = i0 + kVsEC * stride; i < n; i += stride) {
		const VsElem q = vs_decompose((double)x[i], tbad);
		tlo = min(tlo, q.lo);
		thi = max(thi, q.hi);
	}
	if (tbad) badB = 1;
	// workgroup range of k (a shuffle tree per wave, then the waves)
	for (int o = 32; o > 0; o >>= 1) {
		tlo = min(tlo, __shfl_xor(tlo, o, 64));
		thi = max(thi, __shfl_xor(thi, o, 64));
	}
	if (lane == 0) {
		red[0][threadIdx.x / 64] = tlo;
		red[1][threadIdx.x / 64] = thi;
	}
	__syncthreads();
	if (threadIdx.x == 0) {
		int a = red[0][0], c = red[1][0];
		for (int w = 1; w < BS / 64; w++) {
			a = min(a, red[0][w]);
			c = max(c, red[1][w]);
		}
		kminB = a;
		kmaxB = c;
	}
	__syncthreads();
	const int r0 = kminB, r1 = kmaxB;
	// pass 2: key(k) = |cmod(x, 2^k)| / 2^k, a binary64 in (0, 1/2] kept as its bits (positive
	// binary64 values order like their bit patterns), maximised over the workgroup for every k.
	if (r0 <= r1 && r1 - r0 < kVsW && n <= kVsEC * stride) {
		// narrow range, every element in registers: the keys of the window k = r0 .. r0 + kVsW - 1
		// in registers too, then a butterfly max over the wave, 16 window entries per step
		double Fr[kVsW];
#pragma unroll
		for (int kk = 0; kk < kVsW; kk++) Fr[kk] = 0.0;
#pragma unroll
		for (int e = 0; e < kVsEC; e++) vs_window_keys(xc[e], r0, Fr);
		constexpr int KC = 16;
#pragma unroll
		for (int k0 = 0; k0 < kVsW; k0 += KC) {
#pragma unroll
			for (int o = 32; o > 0; o >>= 1) {
				double t[KC];
#pragma unroll
				for (int kk = 0; kk < KC; kk++) t[kk] = __shfl_xor(Fr[k0 + kk], o, 64);
#pragma unroll
				for (int kk = 0; kk < KC; kk++) Fr[k0 + kk] = fmax(Fr[k0 + kk], t[kk]);
			}
			if (lane == 0) {
#pragma unroll
				for (int kk = 0; kk < KC; kk++)
					if (Fr[k0 + kk] > 0.0)
						atomicMax(&Fb[r0 + k0 + kk + kVsOff], (unsigned long long)d2u(Fr[k0 + kk]));
			}
		}
	} else if (r0 <= r1) {
		// wide range: the keys of k in [lo, hi] of each element into the table directly; each lane
		// starts at its own offset so one LDS atomic instruction hits different table entries.
		// y = x / 2^k has |y| in [1/4, 2^53) for those k (exact), and |y - rint(y)| is exactly
		// |cmod(x, 2^k)| / 2^k in (0, 1/2]: its bits (a positive binary64) order like the value.
		auto add = [&](double v) {
			int dummy = 0;
			const VsElem q = vs_decompose(v, dummy);
			if (q.m == 0) return;
			const int K = q.hi - q.lo + 1;
			const int off = lane % K;  // this lane's starting entry
			for (int j = 0; j < K; j++) {
				int t = j + off;
				if (t >= K) t -= K;
				const int k = q.lo + t;
				const double y = ldexp(v, -k);
				const uint64_t key = d2u(y - rint(y)) & ~(uint64_t(1) << 63);
				if (key) atomicMax(&Fb[k + kVsOff], (unsigned long long)key);
			}
		};
#pragma unroll
		for (int e = 0; e < kVsEC; e++) add(xc[e]);
		for (int64_t i = i0 + kVsEC * stride; i < n; i += stride) add((double)x[i]);
	}
	__syncthreads();
	// fold this workgroup's table over its range into the global one (Fg: zero between uses).  A
	// workgroup holding an input outside the model folds nothing: it reports no range, so the walk
	// would not clear what it folded, and the split takes the reference loop anyway.
	if (r0 <= r1 && !badB) {  // an all-zero workgroup has the empty range [INT32_MAX, INT32_MIN]
		for (int k = r0 + (int)threadIdx.x; k <= r1; k += BS) {
			const unsigned long long f = Fb[k + kVsOff];
			if (f) atomicMax(&Fg[k + kVsOff], f);
		}
	}
	// range: [r0, r1] (empty when r0 > r1); x = INT32_MIN marks an input outside the model
	if (threadIdx.x == 0) range[blockIdx.x] = badB ? make_int2(INT32_MIN, INT32_MIN) : make_int2(r0, r1);
	// the last workgroup to finish walks the levels
	__shared__ int lastWG;
	__syncthreads();
	if (threadIdx.x == 0) {
		__threadfence();
		lastWG = atomicAdd(done, 1u) == gridDim.x - 1;
	}
	__syncthreads();
	if (!lastWG) return;
	__threadfence();
	if (threadIdx.x == 0) *done = 0u;
	vs_walk<BS, TI>(gridDim.x, Fg, range, rho, cap, reproMode, J, plan, info, n, x, slices, lds, rhoIn, shift,
	                fallback);
}

template <int BS, class TI>
__device__ void vsplit_fallback(int n, const TI *x, double *slices, int64_t lds, int rho, int shift,
                                int cap, int reproMode, const uint16_t *J, int *info);

// The level walk of the split, by the last workgroup of k_vsplit_scan (the nwg workgroups' ranges
// and the global table): the union of the ranges, F as the prefix max of the table, then the levels
// M_{s+1} = F(rho + tau(M_s) - kPrec) into the plan.
template <int BS, class TI>
__device__ void vs_walk(int nwg, unsigned long long *Fg, const int2 *range, int rho, int cap, int reproMode,
                        const uint16_t *J, VsPlan *plan, int *info, int n, const TI *x, double *slices,
                        int64_t lds, int rhoIn, int shift, int fallback) {
	__shared__ unsigned long long Gb[kVsK];
	__shared__ int kminB, kmaxB, badB;
	if (threadIdx.x == 0) {
		kminB = INT32_MAX;
		kmaxB = INT32_MIN;
		badB = 0;
	}
	__syncthreads();
	for (int b = threadIdx.x; b < nwg; b += BS) {
		// the other workgroups' ranges and table entries: read at the device scope
		const int2 rg = make_int2(__hip_atomic_load(&range[b].x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
		                          __hip_atomic_load(&range[b].y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
		if (rg.x == INT32_MIN) {
			badB = 1;
		} else if (rg.x <= rg.y) {
			atomicMin(&kminB, rg.x);
			atomicMax(&kmaxB, rg.y);
		}
	}
	__syncthreads();
	const int K0 = kminB, K1 = kmaxB;  // empty ([INT32_MAX, INT32_MIN]) for an all-zero vector
	for (int k = (K0 <= K1 ? K0 + (int)threadIdx.x : 1); k <= K1; k += BS) {
		// F(k) = key 2^k (a residue of a binary64: exact)
		Gb[k + kVsOff] = d2u(ldexp(__builtin_bit_cast(double, (uint64_t)vs_load(&Fg[k + kVsOff])), k));
		Fg[k + kVsOff] = 0ull;  // clear for the next split
	}
	__syncthreads();  // the prefix segments below span entries other threads converted
	// the tau table of the level walk into LDS
	constexpr int nJ = std::is_same<TI, float>::value ? kTauBinadesF : kTauBinades;
	__shared__ uint16_t Js[kTauBinades];
	for (int q = threadIdx.x; q < nJ; q += BS) Js[q] = J[q];
	// |cmod(x, 2^k)| does not decrease with k (a residue modulo 2^k' is one modulo 2^k, and the
	// centred one is the smallest), so F is the prefix max of the entries; beyond K1 it is max |x|.
	// Prefix max in parallel: each thread's segment, then a scan of the segment maxima.
	__shared__ unsigned long long segMax[BS];
	const int Lk = (K0 <= K1) ? K1 - K0 + 1 : 0;
	const int per = (Lk + BS - 1) / BS;
	const int s0 = (Lk > 0) ? K0 + (int)threadIdx.x * per : 0;
	const int s1 = (Lk > 0) ? min(s0 + per, K1 + 1) : 0;
	unsigned long long run = 0ull;
	for (int k = s0; k < s1; k++) {
		run = Gb[k + kVsOff] > run ? Gb[k + kVsOff] : run;
		Gb[k + kVsOff] = run;
	}
	segMax[threadIdx.x] = run;
	__syncthreads();
	for (int o = 1; o < BS; o <<= 1) {
		const unsigned long long v = (threadIdx.x >= (unsigned)o) ? segMax[threadIdx.x - o] : 0ull;
		__syncthreads();
		if (v > segMax[threadIdx.x]) segMax[threadIdx.x] = v;
		__syncthreads();
	}
	const unsigned long long pre = (threadIdx.x > 0) ? segMax[threadIdx.x - 1] : 0ull;
	for (int k = s0; k < s1; k++)
		if (pre > Gb[k + kVsOff]) Gb[k + kVsOff] = pre;
	__syncthreads();
	__shared__ int planBad;
	if (threadIdx.x == 0) {
		const uint16_t *Jl = Js;
		int bad = badB;
		const unsigned long long gmax = segMax[BS - 1];  // max |x|
		auto Fk = [&](int k) -> double {
			if (K0 > K1 || k < K0) return 0.0;
			if (k > K1) return __builtin_bit_cast(double, gmax);
			return __builtin_bit_cast(double, Gb[k + kVsOff]);
		};
		// rho <= kPrec keeps every unit an even divisor of the one before
		// (u_{s+1} <= u_s 2^(rho-1-kPrec)), which the residue chain needs; a binary32 vector longer
		// than 2^24 exceeds it
		if (rho < 3 || rho > SplitRule<TI>::kPrec || cap > kVsMaxLevels) bad = 1;
		double M = __builtin_bit_cast(double, gmax);
		int s = 0;
		bool zero = false;
		for (s = 0; s < cap && !bad; s++) {
			const int tau = SplitRule<TI>::tau(M, Jl);
			if (rho + tau > SplitRule<TI>::kSigEmax || rho + tau - 1 < SplitRule<TI>::kEmin) {
				bad = 1;
				break;
			}
			const int k = rho + tau - SplitRule<TI>::kPrec;
			plan->k[s] = k;
			M = Fk(k);
			if (M == 0.0) {
				zero = true;
				break;
			}
		}
		plan->bad = bad;
		if (!bad) {
			int count;
			if (cap <= 1)
				count = 1;
			else if (zero)
				count = (s < 1 ? 1 : s) + 1;
			else
				count = cap;
			if (reproMode == 0) count++;
			plan->count = count;
			plan->last = zero ? s : cap - 1;
			plan->zero = zero ? 1 : 0;
			info[0] = count;
			info[1] = zero ? s : cap;
		}
		planBad = bad;
	}
	__syncthreads();
	// a plan outside the model: the reference loop stores the slices and the count, here when the
	// consumers read no stored slices otherwise (k_vsplit_write does it when it runs)
	if (fallback && planBad) vsplit_fallback<BS, TI>(n, x, slices, lds, rhoIn, shift, cap, reproMode, J, info);
}

// ---- predicted chain: F at the few units the level walk will visit -------------------------------
// Level s+1's unit exponent is k_{s+1} = rho + tau(F(k_s)) - prec with F(k_s) <= 2^(k_s - 1), so it is
// at most k_s - d, d = prec - rho + 1, and exactly that whenever F(k_s) > 2^(k_s - 2) (tau rounding
// aside) -- the case for all but the last level of vectors of varied mantissas.  The reductions
// max |x| and the lowest non-trivial unit exponent min lo (k_vsplit_pre, or the kernel that produced
// x) fix k_0 and the last level that can be non-zero; k_vsplit_levels evaluates F at
// k_s = k_0 - d s for those levels (plus the zero one), and k_vsplit_verify walks the chain with
// the exact tau rule, keeping the plan when every k_s it meets is the predicted one.  Otherwise it
// sets plan->needScan and the table scan (gated on it) and its walk make the plan.  The F values are
// exact maxima either way, so the plan -- and every slice -- is the same.

// category clock of the device CG: the first workgroup of a category's first kernel charges the
// time since the previous stamp to the category then open; an iteration past the stop closes it
__device__ inline void cg_clock(CgDev *S, int it, int cat) {
	if (S == nullptr || blockIdx.x != 0 || threadIdx.x != 0) return;
	if (it >= S->stopAt) cat = kClkNone;
	if (cat == kClkNone && S->clkCat == kClkNone) return;
	const unsigned long long t = (unsigned long long)wall_clock64();
	if (S->clkCat != kClkNone) S->clkTot[S->clkCat] += t - S->clkLast;
	S->clkLast = t;
	S->clkCat = cat;
}

__global__ void k_cg_clock_close(CgDev *S) { cg_clock(S, S->stopAt, kClkNone); }

void launch_cg_clock_close(hipStream_t st, CgDev *S) {
	hipLaunchKernelGGL(k_cg_clock_close, dim3(1), dim3(64), 0, st, S);
}

// the reductions of the predicted chain for one element, then for the workgroup (every thread of
// the workgroup calls vs_stats_flush)
__device__ inline void vs_stats_acc(double v, double &mx, int &lo, int &bad) {
	const VsElem q = vs_decompose(v, bad);
	if (q.m == 0) return;
	mx = std_max(mx, fabs(v));
	lo = max(lo, kVsLoOff - q.lo);
}

// atomicMax that skips the atomic when the slot already holds at least v (the slot only grows
// while the reductions run), which keeps most workgroups off the shared cache line
__device__ inline void atomic_max_lazy(unsigned long long *p, unsigned long long v) {
	if (__hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < v) atomicMax(p, v);
}
__device__ inline void atomic_max_lazy(int *p, int v) {
	if (__hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < v) atomicMax(p, v);
}

// the reductions of all slots (every workgroup of the level kernel reads them; with atomic loads in
// the verify step, after the other workgroups' atomics)
template <bool ATOMIC>
__device__ inline void vs_pre_sum(const VsPre *pre, unsigned long long &maxBits, int &loOff, int &bad) {
	maxBits = 0ull;
	loOff = 0;
	bad = 0;
#pragma unroll
	for (int q = 0; q < kVsSpread; q++) {
		const VsSlot &sl = pre->slot[q];
		if constexpr (ATOMIC) {
			maxBits = std::max(maxBits, __hip_atomic_load(&sl.maxBits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
			loOff = max(loOff, __hip_atomic_load(&sl.loOff, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
			bad |= __hip_atomic_load(&sl.bad, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		} else {
			maxBits = std::max(maxBits, sl.maxBits);
			loOff = max(loOff, sl.loOff);
			bad |= sl.bad;
		}
	}
}

template <int BS>
__device__ inline void vs_stats_flush(double mx, int lo, int bad, VsPre *pre) {
	__shared__ double redd[BS / 64];
	__shared__ int redi[BS / 64];
	mx = block_max_d<BS>(mx, redd);
	for (int o = 32; o > 0; o >>= 1) {
		lo = max(lo, __shfl_xor(lo, o, 64));
		bad = max(bad, __shfl_xor(bad, o, 64));
	}
	if (threadIdx.x % 64 == 0) redi[threadIdx.x / 64] = lo | (bad << 30);
	__syncthreads();
	if (threadIdx.x == 0) {
		int l = 0, bd = 0;
		for (int w = 0; w < BS / 64; w++) {
			l = max(l, redi[w] & ((1 << 30) - 1));
			bd |= redi[w] >> 30;
		}
		VsSlot &sl = pre->slot[blockIdx.x % kVsSpread];
		if (mx > 0.0) atomic_max_lazy(&sl.maxBits, (unsigned long long)d2u(mx));
		if (l > 0) atomic_max_lazy(&sl.loOff, l);
		if (bd) atomicOr(&sl.bad, 1);
	}
}

template <class TI>
__device__ inline void vs_predict(double M0, int minLo, int rho, int cap, const uint16_t *J, int &k0,
                                  int &d, int &L) {
	k0 = rho + SplitRule<TI>::tau(M0, J) - SplitRule<TI>::kPrec;
	d = SplitRule<TI>::kPrec - rho + 1;
	L = (d > 0 && k0 >= minLo) ? (k0 - minLo) / d + 2 : 1;
	if (L > cap) L = cap;
	if (L > kVsPred) L = kVsPred;
}

constexpr int kVsLevBlocks = 1024;  // workgroups of the reduction kernels at most

template <int BS, class TI>
__global__ __launch_bounds__(BS) void k_vsplit_pre(int n, const TI *x, VsPre *pre, const int *stopAt, int it,
                                                   VsSecond<TI> v2) {
	if (stopAt && it >= *stopAt) return;
	if (blockIdx.y == 1) {
		x = v2.x;
		pre = v2.pre;
	}
	double mx = 0.0;
	int lo = 0, bad = 0;
	const int64_t stride = (int64_t)gridDim.x * BS;
	for (int64_t i = blockIdx.x * (int64_t)BS + threadIdx.x; i < n; i += 4 * stride) {
		double v[4];
#pragma unroll
		for (int e = 0; e < 4; e++) {
			const int64_t ie = i + e * stride;
			const double xl = (double)x[ie < n ? ie : (int64_t)n - 1];
			v[e] = ie < n ? xl : 0.0;
		}
#pragma unroll
		for (int e = 0; e < 4; e++) vs_stats_acc(v[e], mx, lo, bad);
	}
	vs_stats_flush<BS>(mx, lo, bad, pre);
}

// The walk over the predicted F values (the steps of vs_walk), by one thread of k_vsplit_verify.  Anything the prediction does not settle (an input outside the
// model, a zero vector, a level off the chain) goes to the table path.
template <class TI>
__device__ int vs_verify(unsigned long long mb, int lof, int bad, const unsigned long long *Fb, int rho,
                         int cap, int reproMode, const uint16_t *J, VsPlan *plan, int *info, bool write) {
	const double M0 = __builtin_bit_cast(double, (uint64_t)mb);
	if (bad || M0 == 0.0 || rho < 3 || rho > SplitRule<TI>::kPrec || cap > kVsMaxLevels) return 1;
	int k0, d, L;
	vs_predict<TI>(M0, kVsLoOff - lof, rho, cap, J, k0, d, L);
	int ks[kVsPred];
	double M = M0;
	int s = 0;
	bool zero = false;
	for (s = 0; s < cap; s++) {
		const int tau = SplitRule<TI>::tau(M, J);
		const int k = rho + tau - SplitRule<TI>::kPrec;
		if (rho + tau > SplitRule<TI>::kSigEmax || rho + tau - 1 < SplitRule<TI>::kEmin || s >= L ||
		    k != k0 - d * s)  // outside the model, or off the predicted chain
			return 1;
		ks[s] = k;
		M = __builtin_bit_cast(double, (uint64_t)Fb[s]);
		if (M == 0.0) {
			zero = true;
			break;
		}
	}
	if (!write) return 0;
	int count;
	if (cap <= 1)
		count = 1;
	else if (zero)
		count = (s < 1 ? 1 : s) + 1;
	else
		count = cap;
	if (reproMode == 0) count++;
	const int last = zero ? s : cap - 1;
	for (int q = 0; q <= last && q < kVsPred; q++) plan->k[q] = ks[q];
	plan->needScan = 0;
	plan->bad = 0;
	plan->count = count;
	plan->last = last;
	plan->zero = zero ? 1 : 0;
	info[0] = count;
	info[1] = zero ? s : cap;
	return 0;
}

constexpr int kVsSlotWords = (int)(sizeof(VsSlot) / 8) * kVsSpread;

// pre: this split's reductions; preNext: the other buffer of the pair, cleared here by workgroup 0
// for the split after this one (nothing reads or writes it meanwhile)
template <int BS, class TI>
__global__ __launch_bounds__(BS) void k_vsplit_levels(int n, const TI *x, VsPre *pre, VsPre *preNext, int rho,
                                                      int cap, const uint16_t *J, CgDev *S, int it,
                                                      int clkCat, VsSecond<TI> v2) {
	cg_clock(S, it, clkCat);
	if (S && it >= S->stopAt) return;
	if (blockIdx.y == 1) {
		x = v2.x;
		pre = v2.pre;
		preNext = v2.preNext;
	}
	if (blockIdx.x == 0)
		for (int q = threadIdx.x; q < kVsSlotWords; q += BS) ((unsigned long long *)preNext->slot)[q] = 0ull;
	__shared__ double red[BS / 64];
	unsigned long long mb;
	int lof, pbad;
	vs_pre_sum<false>(pre, mb, lof, pbad);
	if (!pbad && mb != 0ull) {
		const double M0 = __builtin_bit_cast(double, (uint64_t)mb);
		int k0, d, L;
		vs_predict<TI>(M0, kVsLoOff - lof, rho, cap, J, k0, d, L);
		double Fm[kVsPred];
#pragma unroll
		for (int q = 0; q < kVsPred; q++) Fm[q] = 0.0;
		// four elements per thread and step, their loads in flight together
		const int64_t stride = (int64_t)gridDim.x * BS;
		for (int64_t i = blockIdx.x * (int64_t)BS + threadIdx.x; i < n; i += 4 * stride) {
			double v[4];
#pragma unroll
			for (int e = 0; e < 4; e++) {
				const int64_t ie = i + e * stride;
				const double xl = (double)x[ie < n ? ie : (int64_t)n - 1];
				v[e] = ie < n ? xl : 0.0;
			}
#pragma unroll
			for (int e = 0; e < 4; e++)
#pragma unroll
				for (int q = 0; q < kVsPred; q++)
					if (q < L) Fm[q] = std_max(Fm[q], fabs(vs_cmod(v[e], k0 - d * q)));
		}
#pragma unroll
		for (int q = 0; q < kVsPred; q++) {
			if (q < L) {  // L is uniform: every thread takes the reductions it needs
				const double f = block_max_d<BS>(Fm[q], red);
				if (threadIdx.x == 0 && f > 0.0)
					atomic_max_lazy(&pre->slot[blockIdx.x % kVsSpread].F[q], (unsigned long long)d2u(f));
			}
		}
	}
}

// Slices of one element from its plan: the reference's loop (src/cpu/ozblas_split.cpp:75-104)
// with the plan's units, sigma_s = 0.75 * 2^(rho + tau_s) = 1.5 * 2^(k_s + prec - 1), in the input's
// precision.  Slot last + 1 holds the remainder (0 when the chain ended in zero) and later slots 0:
// the slots k_vsplit_write stores.  Slots at or above the slice count are for the caller to mask.
// sigma = 1.5 * 2^(k + prec - 1) from its bits (integer work on the uniform k: scalar registers)
template <class TI>
__device__ inline TI vs_sigma(int k);
template <>
__device__ inline double vs_sigma<double>(int k) {
	return __builtin_bit_cast(double, ((uint64_t)(k + 52 + 1023) << 52) | (uint64_t(1) << 51));
}
template <>
__device__ inline float vs_sigma<float>(int k) {
	return __builtin_bit_cast(float, ((uint32_t)(k + 23 + 127) << 23) | (uint32_t(1) << 22));
}

template <class TI, int NS>
struct VsGen {
	TI sig[NS];
	int last, zero;
	__device__ inline void init(const VsPlan *p) {
		last = p->last;
		zero = p->zero;
#pragma unroll
		for (int q = 0; q < NS; q++) {
			const int kq = p->k[q];
			sig[q] = (q <= last) ? vs_sigma<TI>(kq) : (TI)0;
		}
	}
	// Branch-free: with sigma = 0 beyond the last level the step returns the remainder itself at
	// slot last + 1 (exactly 0 when the chain ended in zero) and 0 after it.
	__device__ inline void gen(TI v, double (&out)[NS]) const {
		TI r = v;
#pragma unroll
		for (int q = 0; q < NS; q++) {
			const TI a = (r + sig[q]) - sig[q];
			out[q] = (double)a;
			r = r - a;
		}
	}
};

// slot q alone (the levels before it replayed): the general paths with more slots than registers
template <class TI>
__device__ inline double vs_slot(TI v, const VsPlan *p, int q) {
	TI r = v;
	const int last = p->last;
	for (int t = 0; t <= last; t++) {
		const TI sg = vs_sigma<TI>(p->k[t]);
		const TI a = (r + sg) - sg;
		if (t == q) return (double)a;
		r = r - a;
	}
	return (q == last + 1 && !p->zero) ? (double)r : 0.0;
}

// the reference loop in one workgroup of BS threads (inputs outside the model above): slice s at slot s, its
// remainder at slot s + 1 (read back as the next level's input), as src/cpu/ozblas_split.cpp:96
template <int BS, class TI>
__device__ void vsplit_fallback(int n, const TI *x, double *slices, int64_t lds, int rho, int shift,
                                int cap, int reproMode, const uint16_t *J, int *info) {
	__shared__ double red[BS / 64];
	double mx = 0.0;
	for (int64_t i = threadIdx.x; i < n; i += BS) mx = std_max(mx, fabs((double)x[i]));
	mx = block_max_d<BS>(mx, red);
	int s;
	bool zero = false;
	for (s = 0; s < cap; s++) {
		const TI sigma = SplitRule<TI>::sigma(rho, SplitRule<TI>::tau(mx, J), shift);
		const double *src = slices + (int64_t)s * lds;
		double lm = 0.0;
		for (int64_t i = threadIdx.x; i < n; i += BS) {
			const TI v = (s == 0) ? x[i] : (TI)src[i];  // a stored remainder converts back exactly
			const TI a = (v + sigma) - sigma;
			const TI rr = v - a;
			slices[(int64_t)s * lds + i] = (double)a;
			slices[(int64_t)(s + 1) * lds + i] = (double)rr;
			lm = std_max(lm, fabs((double)rr));
		}
		mx = block_max_d<BS>(lm, red);
		if (mx == 0.0) {
			zero = true;
			break;
		}
	}
	const int last = zero ? s : cap - 1;
	int count;
	if (cap <= 1)
		count = 1;
	else if (zero)
		count = (s < 1 ? 1 : s) + 1;
	else
		count = cap;
	if (reproMode == 0) count++;
	for (int64_t i = threadIdx.x; i < n; i += BS) {
		if (zero) slices[(int64_t)(last + 1) * lds + i] = 0.0;
		for (int q = last + 2; q < count; q++) slices[(int64_t)q * lds + i] = 0.0;
	}
	if (threadIdx.x == 0) {
		info[0] = count;
		info[1] = zero ? s : cap;
	}
}


// slices from the plan: slice_s = cmod(x, u_{s-1}) - cmod(x, u_s), remainder slot, zero slots;
// a plan marked bad is split by workgroup 0 with the reference loop instead
template <class TI>
__global__ __launch_bounds__(256) void k_vsplit_write(int n, const TI *x, double *slices,
                                                      int64_t lds, const VsPlan *plan, int rho,
                                                      int shift, int cap, int reproMode,
                                                      const uint16_t *J, int *info) {
	if (plan->bad) {
		if (blockIdx.x == 0) vsplit_fallback<256, TI>(n, x, slices, lds, rho, shift, cap, reproMode, J, info);
		return;
	}
	const int last = plan->last, count = plan->count, zero = plan->zero;
	for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
	     i += (int64_t)gridDim.x * blockDim.x) {
		const double v = (double)x[i];
		double r = v;
		for (int s = 0; s <= last; s++) {
			const double c = vs_cmod(v, plan->k[s]);
			slices[(int64_t)s * lds + i] = r - c;
			r = c;
		}
		slices[(int64_t)(last + 1) * lds + i] = zero ? 0.0 : r;
		for (int q = last + 2; q < count; q++) slices[(int64_t)q * lds + i] = 0.0;
	}
}

size_t vsplit_scratch_bytes() { return 8ull * kVsK + 8ull * kVsBlocks + 64 + 2 * sizeof(VsPre) + 64; }

// The plan of x (length n): the predicted chain, then the table scan and walk only when it does not
// hold; with fallback the walk's workgroup also stores the slices of a plan marked bad (the
// reference loop).  scratch: vsplit_scratch_bytes(), zeroed once (the kernels leave it zero again).
// the pointers of a vector's split in its scratch (Fg, range, the done counter, the VsPre pair)
template <class TI>
static VsSecond<TI> vs_second(const TI *x, void *scratch, VsPlan *plan, int *info, double *slices) {
	VsSecond<TI> v{};
	v.x = x;
	v.Fg = (unsigned long long *)scratch;
	v.range = (int2 *)(v.Fg + kVsK);
	v.done = (unsigned int *)(v.range + kVsBlocks);
	v.pre = (VsPre *)((char *)scratch + 8ull * kVsK + 8ull * kVsBlocks + 64);
	v.preNext = v.pre + 1;
	v.plan = plan;
	v.info = info;
	v.slices = slices;
	return v;
}

// With y (the DOT's second vector, its own scratch, plan, info and slices in y2) both vectors are
// split by the same three launches (grid rows 0 and 1).
template <class TI>
static hipError_t vsplit_plan(hipStream_t st, int n, const TI *x, double *slices, int64_t lds, int rho,
                              int shift, int cap, int reproMode, const uint16_t *J, int *info,
                              void *scratch, VsPre *pre, VsPlan *plan, bool statsDone, CgDev *S, int it,
                              int clkCat, int fallback, int parity, const VsSecond<TI> *y2 = nullptr) {
	unsigned long long *Fg = (unsigned long long *)scratch;
	int2 *range = (int2 *)(Fg + kVsK);
	if (pre == nullptr) pre = (VsPre *)((char *)scratch + 8ull * kVsK + 8ull * kVsBlocks + 64);
	VsPre *cur = pre + (parity & 1), *next = pre + ((parity & 1) ^ 1);
	const VsSecond<TI> v2 = y2 ? *y2 : VsSecond<TI>{};
	const unsigned rows = y2 ? 2u : 1u;
	const int *stopAt = S ? &S->stopAt : nullptr;
	int rhoEff = rho;
	if (shift > 1) {  // sigma * 2^j acts as rho + j
		if (shift & (shift - 1)) rhoEff = -1000;  // not a power of two: the fallback splits
		else rhoEff = rho + __builtin_ctz((unsigned)shift);
	}
	int grid = (int)std::min<int64_t>(kVsBlocks, ((int64_t)n + kVsBS - 1) / kVsBS);
	if (grid < 1) grid = 1;
	int gridL = (int)std::min<int64_t>(kVsLevBlocks, ((int64_t)n + 4 * kVsBS - 1) / (4 * kVsBS));
	if (gridL < 1) gridL = 1;
	if (!statsDone)
		hipLaunchKernelGGL((k_vsplit_pre<kVsBS, TI>), dim3(gridL, rows), dim3(kVsBS), 0, st, n, x, cur, stopAt, it,
		                   v2);
	hipLaunchKernelGGL((k_vsplit_levels<kVsBS, TI>), dim3(gridL, rows), dim3(kVsBS), 0, st, n, x, cur, next,
	                   rhoEff, cap, J, S, it, clkCat, v2);
	unsigned int *done = (unsigned int *)(range + kVsBlocks);
	hipLaunchKernelGGL((k_vsplit_scan<kVsBS, TI>), dim3(grid, rows), dim3(kVsBS), 0, st, n, x, Fg, range, done,
	                   stopAt, it, (const VsPre *)cur, rhoEff, cap, reproMode, J, plan, info, slices, lds, rho,
	                   shift, fallback, v2);
	return hipGetLastError();
}

// the plan, then the slices (slice_s = cmod(x, u_{s-1}) - cmod(x, u_s), remainder, zero slots)
template <class TI>
static hipError_t vsplit(hipStream_t st, int n, const TI *x, double *slices, int64_t lds, int rho,
                         int shift, int cap, int reproMode, const uint16_t *J, int *info, void *scratch,
                         VsPlan *plan) {
	const hipError_t e = vsplit_plan(st, n, x, slices, lds, rho, shift, cap, reproMode, J, info, scratch,
	                                 nullptr, plan, false, nullptr, 0, kClkNone, 0, 0);
	if (e != hipSuccess) return e;
	int gw = (int)std::min<int64_t>(2048, ((int64_t)n + 255) / 256);
	if (gw < 1) gw = 1;
	hipLaunchKernelGGL(k_vsplit_write<TI>, dim3(gw), dim3(256), 0, st, n, x, slices, lds,
	                   (const VsPlan *)plan, rho, shift, cap, reproMode, J, info);
	return hipGetLastError();
}

hipError_t launch_vsplit(hipStream_t st, int n, const double *x, double *slices, int64_t lds, int rho,
                         int shift, int cap, int reproMode, const uint16_t *J, int *info, void *scratch,
                         VsPlan *plan) {
	return vsplit(st, n, x, slices, lds, rho, shift, cap, reproMode, J, info, scratch, plan);
}

hipError_t launch_vsplit(hipStream_t st, int n, const float *x, double *slices, int64_t lds, int rho,
                         int shift, int cap, int reproMode, const uint16_t *JF, int *info, void *scratch,
                         VsPlan *plan) {
	return vsplit(st, n, x, slices, lds, rho, shift, cap, reproMode, JF, info, scratch, plan);
}

template <class TI>
static hipError_t vsplit_plan2(hipStream_t st, int n, const TI *x, const TI *y, int rho, int shift, int cap,
                               int reproMode, const uint16_t *J, double *const slices[2], int64_t lds,
                               int *const info[2], void *const scratch[2], VsPlan *const plan[2]) {
	const VsSecond<TI> v2 = vs_second<TI>(y, scratch[1], plan[1], info[1], slices[1]);
	return vsplit_plan(st, n, x, slices[0], lds, rho, shift, cap, reproMode, J, info[0], scratch[0], nullptr,
	                   plan[0], false, nullptr, 0, kClkNone, 1, 0, &v2);
}

hipError_t launch_vsplit_plan2(hipStream_t st, int n, const double *x, const double *y, int rho, int shift,
                               int cap, int reproMode, const uint16_t *J, double *const slices[2], int64_t lds,
                               int *const info[2], void *const scratch[2], VsPlan *const plan[2]) {
	return vsplit_plan2(st, n, x, y, rho, shift, cap, reproMode, J, slices, lds, info, scratch, plan);
}

hipError_t launch_vsplit_plan2(hipStream_t st, int n, const float *x, const float *y, int rho, int shift,
                               int cap, int reproMode, const uint16_t *JF, double *const slices[2], int64_t lds,
                               int *const info[2], void *const scratch[2], VsPlan *const plan[2]) {
	return vsplit_plan2(st, n, x, y, rho, shift, cap, reproMode, JF, slices, lds, info, scratch, plan);
}

hipError_t launch_vsplit_plan(hipStream_t st, int n, const double *x, double *slices, int64_t lds,
                              int rho, int shift, int cap, int reproMode, const uint16_t *J, int *info,
                              void *scratch, VsPre *pre, VsPlan *plan, bool statsDone, CgDev *S, int it,
                              int clkCat, int parity) {
	return vsplit_plan(st, n, x, slices, lds, rho, shift, cap, reproMode, J, info, scratch, pre, plan,
	                   statsDone, S, it, clkCat, 1, parity);
}

hipError_t launch_vsplit_plan(hipStream_t st, int n, const float *x, double *slices, int64_t lds,
                              int rho, int shift, int cap, int reproMode, const uint16_t *JF, int *info,
                              void *scratch, VsPre *pre, VsPlan *plan, bool statsDone, CgDev *S, int it,
                              int clkCat, int parity) {
	return vsplit_plan(st, n, x, slices, lds, rho, shift, cap, reproMode, JF, info, scratch, pre, plan,
	                   statsDone, S, it, clkCat, 1, parity);
}

// ---------------------------------------------------------------------------------------------
// slice DOT: out[ia + ib*nA] = sum_i xs_ia[i] * ys_ib[i] (the reference computes these with one
// cblas_dgemm, src/cpu/ozblas_gemm.cpp:209-215).  Every partial sum is an integer multiple of the
// product unit below 2^53, so block partials added with FP64 atomics are exact in any order.
// ---------------------------------------------------------------------------------------------
constexpr int kDotChunk = 16;
constexpr int kDotPairs = 64;    // pair slots per workgroup partial of k_dots_reg
constexpr int kDotMaxWG = 512;  // workgroups of k_dots_reg at most (two waves per SIMD)
constexpr int kDotFinOne = 36;  // pair counts (6 x 6 slices) whose partials k_dots_fin loads in one round

template <int BS>
__global__ __launch_bounds__(BS) void k_slice_dots(int n, const double *xs, int64_t ldx, int nA,
                                                   const double *ys, int64_t ldy, int nB,
                                                   double *out) {
	__shared__ double red[BS / 64];
	const int npairs = nA * nB;
	const int p0 = blockIdx.y * kDotChunk;
	double acc[kDotChunk];
#pragma unroll
	for (int q = 0; q < kDotChunk; q++) acc[q] = 0.0;
	for (int64_t i = blockIdx.x * (int64_t)BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * BS) {
#pragma unroll
		for (int q = 0; q < kDotChunk; q++) {
			const int p = p0 + q;
			if (p < npairs) {
				const int ia = p % nA, ib = p / nA;
				acc[q] = fma(xs[(int64_t)ia * ldx + i], ys[(int64_t)ib * ldy + i], acc[q]);
			}
		}
	}
#pragma unroll
	for (int q = 0; q < kDotChunk; q++) {
		const double v = block_sum_d<BS>(acc[q], red);
		if (threadIdx.x == 0 && p0 + q < npairs && v != 0.0) atomicAdd(&out[p0 + q], v);
	}
}

void launch_slice_dots(hipStream_t st, int n, const double *xs, int64_t ldx, int nA,
                       const double *ys, int64_t ldy, int nB, double *out) {
	int gx = (n + 255) / 256;
	if (gx > 512) gx = 512;
	if (gx < 1) gx = 1;
	const int gy = (nA * nB + kDotChunk - 1) / kDotChunk;
	hipLaunchKernelGGL(k_slice_dots<256>, dim3(gx, gy), dim3(256), 0, st, n, xs, ldx, nA, ys, ldy,
	                   nB, out);
}

// GEMM with k = 0: every slice product is zero, so the reference's sum is +0 and the epilogue
// leaves C = fma(alpha, +0, beta * C)
template <class T>
__global__ void k_gemm_k0(int m, int n, T alpha, T beta, T *C, int64_t ldc) {
	const int64_t total = (int64_t)m * n;
	for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
	     e += (int64_t)gridDim.x * blockDim.x) {
		T *c = C + (e / m) * ldc + e % m;
		*c = fma1_t(alpha, (T)0, beta * *c);
	}
}

static unsigned grid_k0(int m, int n) {
	int64_t g = ((int64_t)m * n + 255) / 256;
	if (g > 4096) g = 4096;
	if (g < 1) g = 1;
	return (unsigned)g;
}

void launch_gemm_k0(hipStream_t st, int m, int n, double alpha, double beta, double *C, int64_t ldc) {
	hipLaunchKernelGGL(k_gemm_k0<double>, dim3(grid_k0(m, n)), dim3(256), 0, st, m, n, alpha, beta, C,
	                   ldc);
}

void launch_gemm_k0(hipStream_t st, int m, int n, float alpha, float beta, float *C, int64_t ldc) {
	hipLaunchKernelGGL(k_gemm_k0<float>, dim3(grid_k0(m, n)), dim3(256), 0, st, m, n, alpha, beta, C,
	                   ldc);
}

// ---------------------------------------------------------------------------------------------
// Dot2 mode (splitEpsMode = 2): error-free transformations of src/cpu/eft.h (scalar forms) and the
// double-double slice dots.  Compiled without contraction, so every expression rounds as written.
// ---------------------------------------------------------------------------------------------
struct dd {
	double x, y;
};
__device__ inline void two_sum(double a, double b, double &s, double &e) {  // TwoSum
	s = a + b;
	const double v = s - a;
	e = (a - (s - v)) + (b - v);
}
__device__ inline void quick_two_sum(double a, double b, double &s, double &e) {  // QuickTwoSum
	s = a + b;
	e = b - (s - a);
}
__device__ inline void dot2i(double a, double b, dd &c) {  // Dot2i: TwoProdFMA + TwoSum
	const double h = a * b;
	const double r = fma(a, b, -h);
	double q;
	two_sum(c.x, h, c.x, q);
	c.y = c.y + (q + r);
}
__device__ inline dd quad_add(dd a, dd b) {  // QuadAdd
	dd t, c;
	two_sum(a.x, b.x, t.x, t.y);
	t.y = t.y + (a.y + b.y);
	quick_two_sum(t.x, t.y, c.x, c.y);
	return c;
}
__device__ inline dd quad_mul(double a, dd b) {  // QuadMul(double, double2)
	dd t, c;
	t.x = a * b.x;
	t.y = fma(a, b.x, -t.x);
	t.y = fma(a, b.y, t.y);
	quick_two_sum(t.x, t.y, c.x, c.y);
	return c;
}

// Dot2 DOT (dgemm_tn_skinny_x2, src/cpu/ozblas_blas_wrapper.cpp:109-152): the reference sums each
// slice dot in OpenMP per-thread partials over a static schedule of k, combined by QuadAdd in a
// critical section.  This is that computation for kDot2Chunks threads entering the section in
// thread order: partial t covers libgomp's static chunk t of [0, n).
constexpr int kDot2Chunks = 1024;

__global__ __launch_bounds__(256) void k_dot2_partials(int n, const double *xs, int64_t ldx, int nA,
                                                       const double *ys, int64_t ldy, dd *part) {
	const int t = blockIdx.x * blockDim.x + threadIdx.x;  // chunk
	const int pr = blockIdx.y;                            // pair: ia + nA * ib
	if (t >= kDot2Chunks) return;
	const int ia = pr % nA, ib = pr / nA;
	const int64_t q = n / kDot2Chunks, r = n % kDot2Chunks;
	const int64_t lo = t * q + (t < r ? t : r), hi = lo + q + (t < r ? 1 : 0);
	const double *x = xs + (int64_t)ia * ldx, *y = ys + (int64_t)ib * ldy;
	dd L{0.0, 0.0};
	// the chunk in order, 16 elements of each operand loaded together ahead of their steps
	constexpr int B = 16;
	int64_t k = lo;
	for (; k + B <= hi; k += B) {
		double xv[B], yv[B];
#pragma unroll
		for (int u = 0; u < B; u++) {
			xv[u] = x[k + u];
			yv[u] = y[k + u];
		}
#pragma unroll
		for (int u = 0; u < B; u++) dot2i(xv[u], yv[u], L);
	}
	for (; k < hi; k++) dot2i(x[k], y[k], L);
	part[(int64_t)pr * kDot2Chunks + t] = L;
}

// one workgroup per pair: the partials staged in LDS, then thread 0 QuadAdds them in chunk order,
// QuadMul(1, .), hi to vals[nA * ib + ia] and lo to vals[nA * (nB + ib) + ia] (C and C + n*ldc of
// the reference)
__global__ __launch_bounds__(256) void k_dot2_combine(int nA, int nB, const dd *part, double *vals) {
	__shared__ dd sp[kDot2Chunks];
	const int pr = blockIdx.x;
	if (pr >= nA * nB) return;
	for (int t = threadIdx.x; t < kDot2Chunks; t += blockDim.x) sp[t] = part[(int64_t)pr * kDot2Chunks + t];
	__syncthreads();
	if (threadIdx.x != 0) return;
	const int ia = pr % nA, ib = pr / nA;
	dd G{0.0, 0.0};
	for (int t = 0; t < kDot2Chunks; t++) G = quad_add(G, sp[t]);
	G = quad_mul(1.0, G);
	vals[nA * ib + ia] = G.x;
	vals[nA * (nB + ib) + ia] = G.y;
}

size_t dot2_partials_bytes(int nA, int nB) { return sizeof(dd) * (size_t)nA * nB * kDot2Chunks; }

void launch_dot2_slices(hipStream_t st, int n, const double *xs, int64_t ldx, int nA,
                        const double *ys, int64_t ldy, int nB, void *part, double *vals) {
	hipLaunchKernelGGL(k_dot2_partials, dim3((kDot2Chunks + 255) / 256, nA * nB), dim3(256), 0, st, n,
	                   xs, ldx, nA, ys, ldy, (dd *)part);
	hipLaunchKernelGGL(k_dot2_combine, dim3(nA * nB), dim3(256), 0, st, nA, nB, (const dd *)part, vals);
}

// DOT / NRM2 summation of the nA x nB slice dots (sumOrder 2, src/cpu/ozblas_gemm.cpp:281-288):
// NearSum of all of them, or Fsum in (level, ia) order; vals is consumed (NearSum works in place)
__device__ inline double dot_pairs_sum(int nvals, double *vals, int nA, int nB, int maxlevel,
                                       int fsum, const uint16_t *J) {
	if (fsum) {
		double t = 0.;
		for (int ik = 0; ik <= maxlevel; ik++)
			for (int ia = 0; ia < nA; ia++) {
				const int ib = ik - ia;
				if (ib < 0 || ib >= nB) continue;
				t = t + vals[nA * ib + ia];
			}
		return t;
	}
	StridedVec v{vals, 1};
	return nearsum(nvals, v, J);
}

// binary32: the same over the pair dots rounded to float (exact: 24 - rho bits per slice keep
// every pair dot within 24 bits), NearSum<float> with the log2f table / Fsum in float
__device__ inline float dot_pairs_sum_f(int nvals, float *vals, int nA, int nB, int maxlevel,
                                        int fsum, const uint16_t *JF2) {
	if (fsum) {
		float t = 0.f;
		for (int ik = 0; ik <= maxlevel; ik++)
			for (int ia = 0; ia < nA; ia++) {
				const int ib = ik - ia;
				if (ib < 0 || ib >= nB) continue;
				t = t + vals[nA * ib + ia];
			}
		return t;
	}
	StridedVecF v{vals, 1};
	return nearsum_f(nvals, v, JF2);
}

// rounds vals[0, nvals) to float in place (f32_term: the term of the <float,float> and the
// <float,double> Fsum alike): float p occupies the bytes of double p / 2, which ascending p has
// already read
__device__ inline float *vals_to_float(int nvals, double *vals) {
	float *f = (float *)vals;
	for (int p = 0; p < nvals; p++) {
		const double v = vals[p];
		f[p] = f32_term(v);
	}
	return f;
}

// one thread; C[0] = fma(alpha, t, beta * C[0])
__global__ void k_dot_sum(int nvals, double *vals, double alpha, double beta, double *C, int nA,
                          int nB, int maxlevel, int fsum, const uint16_t *J) {
	if (threadIdx.x != 0 || blockIdx.x != 0) return;
	const double t = dot_pairs_sum(nvals, vals, nA, nB, maxlevel, fsum, J);
	C[0] = fma(alpha, t, beta * C[0]);
}

void launch_nearsum_small(hipStream_t st, int nvals, double *vals, double alpha, double beta,
                          double *C, int sumOrderFsum, int nA, int nB, int maxlevel, int fsum,
                          const uint16_t *J) {
	(void)sumOrderFsum;
	hipLaunchKernelGGL(k_dot_sum, dim3(1), dim3(64), 0, st, nvals, vals, alpha, beta, C, nA, nB,
	                   maxlevel, fsum, J);
}

// ---------------------------------------------------------------------------------------------
// CSR slice products + global sum fused per row (ozblasRcsrmv, src/cpu/ozblas_csrmv.cpp:113-166):
// v[ia][ib] = sum_j As_ia[j] * Xs_ib[col_j] (fma in row order as csrmm_n, exact), then NearSum
// over the pairs with ia + ib <= maxlevel (or Fsum in (level, ia, ib) order), then
// y = fma(alpha, t, beta * y).
// ---------------------------------------------------------------------------------------------
// T = float: the binary32 instantiations -- with V = float each pair sum is the fma1<float> chain
// (float)fma(a, x, (double)v) of csrmm_n<float> (<float,float>), with V = double the binary64 chain
// of csrmm_n<double> (<float,double>), whose sum then enters the float sum as f32_term; then
// NearSum<float> (J = JF2) or Fsum in float.
// With OTF the slices of x are generated from x and its plan (gen) instead of read from Xs.
template <int NAM, int NBM, class T = double, class TI = double, bool OTF = false, class V = T>
__device__ __forceinline__ T csrmm_row_sum(int r, const double *As, int64_t ldas, int nA,
                                           const int *colind, const int *rowptr, const double *Xs,
                                           int64_t ldxs, int nB, int maxlevel, int fsum,
                                           const uint16_t *J, const TI *X = nullptr,
                                           const VsGen<TI, NBM> *gen = nullptr) {
	const int b = rowptr[r], e = rowptr[r + 1];
	RegVec<NAM * NBM, V> v;
#pragma unroll
	for (int q = 0; q < NAM * NBM; q++) v[q] = (V)0;
	// nonzeros in chunks of CH: the chunk's column indices, then its A and x loads, all in flight
	// together, then the fma chains in row order (the products of one pair stay in sequence).  With
	// OTF every load is unconditional, at an index clamped into the row and the slice counts, and the
	// unused values are selected away afterwards: a conditional load becomes a branch with a full
	// memory wait behind it, which serialises the chunk.
	constexpr int CH = (NAM * NBM <= 16) ? 4 : 2;
	for (int j0 = b; j0 < e; j0 += CH) {
		int c[CH];
		double a[CH][NAM], x[CH][NBM];
		if constexpr (OTF) {
#pragma unroll
			for (int q = 0; q < CH; q++) c[q] = colind[min(j0 + q, e - 1)];
#pragma unroll
			for (int q = 0; q < CH; q++) {
				const bool in = j0 + q < e;
				const int jj = min(j0 + q, e - 1);
#pragma unroll
				for (int ia = 0; ia < NAM; ia++) {
					const double av = As[(int64_t)min(ia, nA - 1) * ldas + jj];
					a[q][ia] = (in && ia < nA) ? av : 0.0;
				}
				const TI xv = X[c[q]];
				gen->gen(in ? xv : (TI)0, x[q]);
#pragma unroll
				for (int ib = 0; ib < NBM; ib++)
					if (ib >= nB) x[q][ib] = 0.0;
			}
		} else {
			// (the stored slices: the compiler sinks such loads into their branches whatever the
			// source form, and this one measured fastest)
#pragma unroll
			for (int q = 0; q < CH; q++) c[q] = (j0 + q < e) ? colind[j0 + q] : 0;
#pragma unroll
			for (int q = 0; q < CH; q++) {
				const bool in = j0 + q < e;
#pragma unroll
				for (int ia = 0; ia < NAM; ia++)
					a[q][ia] = (in && ia < nA) ? As[(int64_t)ia * ldas + j0 + q] : 0.0;
#pragma unroll
				for (int ib = 0; ib < NBM; ib++)
					x[q][ib] = (in && ib < nB) ? Xs[(int64_t)ib * ldxs + c[q]] : 0.0;
			}
		}
#pragma unroll
		for (int q = 0; q < CH; q++) {
			if (j0 + q >= e) break;
#pragma unroll
			for (int ia = 0; ia < NAM; ia++)
#pragma unroll
				for (int ib = 0; ib < NBM; ib++)
					v[ia * NBM + ib] = (V)fma(a[q][ia], x[q][ib], (double)v[ia * NBM + ib]);
		}
	}
	// a pair sum as a term of the sum in T (the binary64 chain of <float,double> rounds as the
	// reference's (float) of the scaled product)
	auto term = [](V p) -> T {
		if constexpr (sizeof(V) == sizeof(T))
			return p;
		else
			return sum_term<T>(p);
	};
	T t;
	if (fsum) {
		t = (T)0;
#pragma unroll
		for (int ik = 0; ik < NAM + NBM - 1; ik++)
#pragma unroll
			for (int ia = 0; ia < NAM; ia++) {
				const int ib = ik - ia;
				if (ib >= 0 && ib < NBM && ia < nA && ib < nB && ik <= maxlevel) t = t + term(v[ia * NBM + ib]);
			}
	} else {
		// NearSum over the products (its result does not depend on the order of the terms):
		// compact the valid ones to the front of a register array
		RegVec<NAM * NBM, T> u;
		int cnt = 0;
#pragma unroll
		for (int ia = 0; ia < NAM; ia++)
#pragma unroll
			for (int ib = 0; ib < NBM; ib++) u[ia * NBM + ib] = (T)0;
#pragma unroll
		for (int ia = 0; ia < NAM; ia++)
#pragma unroll
			for (int ib = 0; ib < NBM; ib++) {
				const bool ok = ia < nA && ib < nB && ia + ib <= maxlevel;
#pragma unroll
				for (int q = 0; q < NAM * NBM; q++)
					if (ok && q == cnt) u[q] = term(v[ia * NBM + ib]);
				cnt += ok ? 1 : 0;
			}
		if constexpr (sizeof(T) == 4)
			t = nearsum_f(cnt, u, J);
		else
			t = nearsum(cnt, u, J);
	}
	return t;
}

template <int NAM, int NBM, class T, class V = T>
__global__ __launch_bounds__(256) void k_csrmm_sum(int m, const double *As, int64_t ldas, int nA,
                                                   const int *colind, const int *rowptr,
                                                   const double *Xs, int64_t ldxs, int nB,
                                                   int maxlevel, int fsum, T alpha, T beta, T *y,
                                                   const uint16_t *J) {
	const int r = blockIdx.x * blockDim.x + threadIdx.x;
	if (r >= m) return;
	const T t = csrmm_row_sum<NAM, NBM, T, double, false, V>(r, As, ldas, nA, colind, rowptr, Xs, ldxs,
	                                                         nB, maxlevel, fsum, J);
	y[r] = fma1_t(alpha, t, beta * y[r]);
}

template <class T, class V = T>
static void csrmm_sum(hipStream_t st, int m, const double *As, int64_t ldas, int nA,
                      const int *colind, const int *rowptr, const double *Xs, int64_t ldxs, int nB,
                      int maxlevel, int fsum, T alpha, T beta, T *y, const uint16_t *J) {
	// register bounds as tight as the counts allow (the per-row sum works on NAM x NBM values)
	const dim3 g((m + 255) / 256), b(256);
	if (nA <= 2 && nB <= 4)
		hipLaunchKernelGGL((k_csrmm_sum<2, 4, T, V>), g, b, 0, st, m, As, ldas, nA, colind, rowptr, Xs,
		                   ldxs, nB, maxlevel, fsum, alpha, beta, y, J);
	else if (nA <= 2)
		hipLaunchKernelGGL((k_csrmm_sum<2, 8, T, V>), g, b, 0, st, m, As, ldas, nA, colind, rowptr, Xs,
		                   ldxs, nB, maxlevel, fsum, alpha, beta, y, J);
	else if (nB <= 4)
		hipLaunchKernelGGL((k_csrmm_sum<4, 4, T, V>), g, b, 0, st, m, As, ldas, nA, colind, rowptr, Xs,
		                   ldxs, nB, maxlevel, fsum, alpha, beta, y, J);
	else
		hipLaunchKernelGGL((k_csrmm_sum<4, 8, T, V>), g, b, 0, st, m, As, ldas, nA, colind, rowptr, Xs,
		                   ldxs, nB, maxlevel, fsum, alpha, beta, y, J);
}

void launch_csrmm_sum(hipStream_t st, int m, int nnz, const double *As, int64_t ldas, int nA,
                      const int *colind, const int *rowptr, const double *Xs, int64_t ldxs, int nB,
                      int maxlevel, int fsum, double alpha, double beta, double *y,
                      const uint16_t *J) {
	(void)nnz;
	csrmm_sum(st, m, As, ldas, nA, colind, rowptr, Xs, ldxs, nB, maxlevel, fsum, alpha, beta, y, J);
}

void launch_csrmm_sum(hipStream_t st, int m, int nnz, const double *As, int64_t ldas, int nA,
                      const int *colind, const int *rowptr, const double *Xs, int64_t ldxs, int nB,
                      int maxlevel, int fsum, float alpha, float beta, float *y,
                      const uint16_t *JF2, int mixed) {
	(void)nnz;
	if (mixed)
		csrmm_sum<float, double>(st, m, As, ldas, nA, colind, rowptr, Xs, ldxs, nB, maxlevel, fsum, alpha,
		                         beta, y, JF2);
	else
		csrmm_sum<float>(st, m, As, ldas, nA, colind, rowptr, Xs, ldxs, nB, maxlevel, fsum, alpha, beta, y,
		                 JF2);
}

// Dot2 CSRMV (csrmm_n_x2, src/cpu/ozblas_blas_wrapper.cpp:262-295, then the global sum over
// 2 nB "slices" of src/cpu/ozblas_csrmv.cpp:159-162): per row and slice pair, Dot2 over the row's
// nonzeros in order and QuadMul(1, .); w[ia][ib'] = hi of pair (ia, ib') for ib' < nB and lo of
// pair (ia, ib' - nB) otherwise; NearSum over all w, or Fsum in (level ia + ib', ia) order.
template <int NAM, int NBM>
__global__ __launch_bounds__(256) void k_csrmm_sum_x2(int m, const double *As, int64_t ldas, int nA,
                                                      const int *colind, const int *rowptr,
                                                      const double *Xs, int64_t ldxs, int nB,
                                                      int fsum, double alpha, double beta, double *y,
                                                      const uint16_t *J) {
	const int r = blockIdx.x * blockDim.x + threadIdx.x;
	if (r >= m) return;
	const int b = rowptr[r], e = rowptr[r + 1];
	dd v[NAM][NBM];
#pragma unroll
	for (int ia = 0; ia < NAM; ia++)
#pragma unroll
		for (int ib = 0; ib < NBM; ib++) v[ia][ib] = dd{0.0, 0.0};
	for (int jj = b; jj < e; jj++) {
		const int c = colind[jj];
		double a[NAM], x[NBM];
#pragma unroll
		for (int ia = 0; ia < NAM; ia++) a[ia] = (ia < nA) ? As[(int64_t)ia * ldas + jj] : 0.0;
#pragma unroll
		for (int ib = 0; ib < NBM; ib++) x[ib] = (ib < nB) ? Xs[(int64_t)ib * ldxs + c] : 0.0;
#pragma unroll
		for (int ia = 0; ia < NAM; ia++)
#pragma unroll
			for (int ib = 0; ib < NBM; ib++) dot2i(a[ia], x[ib], v[ia][ib]);
	}
	double fx[NAM][NBM], fy[NAM][NBM];  // renormalised (hi, lo) of every pair: QuadMul(1, .)
#pragma unroll
	for (int ia = 0; ia < NAM; ia++)
#pragma unroll
		for (int ib = 0; ib < NBM; ib++) {
			const dd f = quad_mul(1.0, v[ia][ib]);
			fx[ia][ib] = f.x;
			fy[ia][ib] = f.y;
		}
	double t;
	if (fsum) {
		t = 0.;
		for (int ik = 0; ik <= (nA - 1) + (2 * nB - 1); ik++)
			for (int ia = 0; ia < nA; ia++) {
				const int ibp = ik - ia;
				if (ibp < 0 || ibp >= 2 * nB) continue;
				const bool hiPart = ibp < nB;
				const int ib = hiPart ? ibp : ibp - nB;
				double w = 0.0;
#pragma unroll
				for (int qa = 0; qa < NAM; qa++)
#pragma unroll
					for (int qb = 0; qb < NBM; qb++)
						if (qa == ia && qb == ib) w = hiPart ? fx[qa][qb] : fy[qa][qb];
				t = t + w;
			}
	} else {
		RegVec<NAM * 2 * NBM> u;  // the 2 nA nB values, compacted to the front
#pragma unroll
		for (int q = 0; q < NAM * 2 * NBM; q++) u[q] = 0.0;
		int cnt = 0;
#pragma unroll
		for (int ia = 0; ia < NAM; ia++)
#pragma unroll
			for (int ib = 0; ib < NBM; ib++) {
				const bool ok = ia < nA && ib < nB;
#pragma unroll
				for (int q = 0; q < NAM * 2 * NBM; q++) {
					if (ok && q == cnt) u[q] = fx[ia][ib];
					if (ok && q == cnt + 1) u[q] = fy[ia][ib];
				}
				cnt += ok ? 2 : 0;
			}
		t = nearsum(cnt, u, J);
	}
	y[r] = fma(alpha, t, beta * y[r]);
}

void launch_csrmm_sum_x2(hipStream_t st, int m, const double *As, int64_t ldas, int nA,
                         const int *colind, const int *rowptr, const double *Xs, int64_t ldxs, int nB,
                         int fsum, double alpha, double beta, double *y, const uint16_t *J) {
	hipLaunchKernelGGL((k_csrmm_sum_x2<4, 4>), dim3((m + 255) / 256), dim3(256), 0, st, m, As, ldas,
	                   nA, colind, rowptr, Xs, ldxs, nB, fsum, alpha, beta, y, J);
}

// ---------------------------------------------------------------------------------------------
// GEMV of the plain form (the reference's GEMV branch, src/cpu/ozblas_gemm.cpp:216-225, summed
// with sumOrder 3, src/cpu/ozblas_sum.cpp:331-380, or NearSum :231-271), bandwidth-bound: one wave
// per output element i reads column i of every A slice once (k-contiguous, 16-byte loads) and the
// x slices through the cache, v[ia][ib] = sum_k As_ia[k, i] Xs_ib[k] (exact: any order), then
// lane 0 sums the nA x nB values -- Fsum in (level, ia) order, the order of the reference's index
// map nB*ia + ib, or NearSum -- and y_i = fma1(alpha, t, beta * y_i).  Slices are padded with zeros
// to a multiple of 16 in k, so the paired loads never leave a slice.  T = float: the binary32 sum.
// ---------------------------------------------------------------------------------------------
template <class T, int NAM, int NBM, int R>
__global__ __launch_bounds__(256) void k_gemv_slices(int M, int K, const double *As, int64_t ldk,
                                                     int64_t sA, int nA, const double *Xs,
                                                     int64_t sX, int nB, int fsum, T alpha, T beta,
                                                     T *y, const uint16_t *J) {
	// R rows per wave share each load of the x slices
	const int lane = threadIdx.x % 64;
	const int row0 = (blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64) * R;
	if (row0 >= M) return;  // the whole wave
	double acc[R][NAM][NBM];
#pragma unroll
	for (int r = 0; r < R; r++)
#pragma unroll
		for (int a = 0; a < NAM; a++)
#pragma unroll
			for (int b = 0; b < NBM; b++) acc[r][a][b] = 0.0;
	const int rows = (M - row0 < R) ? M - row0 : R;
	for (int k = 2 * lane; k < K; k += 128) {
		// every load of the step first, so that they are all in flight together
		double2 xv[NBM], av[R][NAM];
#pragma unroll
		for (int b = 0; b < NBM; b++)
			xv[b] = (b < nB) ? *(const double2 *)(Xs + (int64_t)b * sX + k) : make_double2(0.0, 0.0);
#pragma unroll
		for (int r = 0; r < R; r++) {
			const double *arow = As + (int64_t)(row0 + r) * ldk + k;
#pragma unroll
			for (int a = 0; a < NAM; a++)
				av[r][a] = (r < rows && a < nA) ? *(const double2 *)(arow + (int64_t)a * sA)
				                                : make_double2(0.0, 0.0);
		}
#pragma unroll
		for (int r = 0; r < R; r++)
#pragma unroll
			for (int a = 0; a < NAM; a++)
#pragma unroll
				for (int b = 0; b < NBM; b++) {
					acc[r][a][b] = fma(av[r][a].x, xv[b].x, acc[r][a][b]);
					acc[r][a][b] = fma(av[r][a].y, xv[b].y, acc[r][a][b]);
				}
	}
#pragma unroll
	for (int r = 0; r < R; r++)
#pragma unroll
		for (int a = 0; a < NAM; a++)
#pragma unroll
			for (int b = 0; b < NBM; b++)
				for (int o = 32; o > 0; o >>= 1) acc[r][a][b] += __shfl_xor(acc[r][a][b], o, 64);
	// lane r sums row0 + r
#pragma unroll
	for (int r = 0; r < R; r++) {
		if (lane != r || r >= rows) continue;
		T t;
		if (fsum) {
			t = (T)0;
#pragma unroll
			for (int ik = 0; ik < NAM + NBM - 1; ik++)
#pragma unroll
				for (int a = 0; a < NAM; a++) {
					const int b = ik - a;
					if (b >= 0 && b < NBM && a < nA && b < nB) t = t + sum_term<T>(acc[r][a][b]);
				}
		} else {
			RegVec<NAM * NBM, T> u;
			int cnt = 0;
#pragma unroll
			for (int q = 0; q < NAM * NBM; q++) u[q] = (T)0;
#pragma unroll
			for (int a = 0; a < NAM; a++)
#pragma unroll
				for (int b = 0; b < NBM; b++) {
					const bool ok = a < nA && b < nB;
#pragma unroll
					for (int q = 0; q < NAM * NBM; q++)
						if (ok && q == cnt) u[q] = sum_term<T>(acc[r][a][b]);
					cnt += ok ? 1 : 0;
				}
			if constexpr (sizeof(T) == 4)
				t = nearsum_f(cnt, u, J);
			else
				t = nearsum(cnt, u, J);
		}
		y[row0 + r] = fma1_t(alpha, t, beta * y[row0 + r]);
	}
}

template <class T>
static bool gemv_slices(hipStream_t st, int M, int K, const double *As, int64_t ldk, int64_t sA,
                        int nA, const double *Xs, int64_t sX, int nB, int fsum, T alpha, T beta, T *y,
                        const uint16_t *J) {
	const dim3 b(256);
	auto grid = [&](int R) { return dim3((unsigned)((M + 4 * R - 1) / (4 * R))); };
	if (nA <= 4 && nB <= 4)
		hipLaunchKernelGGL((k_gemv_slices<T, 4, 4, 1>), grid(1), b, 0, st, M, K, As, ldk, sA, nA, Xs, sX, nB,
		                   fsum, alpha, beta, y, J);
	else if (nA <= 8 && nB <= 4)
		hipLaunchKernelGGL((k_gemv_slices<T, 8, 4, 1>), grid(1), b, 0, st, M, K, As, ldk, sA, nA, Xs, sX, nB,
		                   fsum, alpha, beta, y, J);
	else if (nA <= 4 && nB <= 8)
		hipLaunchKernelGGL((k_gemv_slices<T, 4, 8, 1>), grid(1), b, 0, st, M, K, As, ldk, sA, nA, Xs, sX, nB,
		                   fsum, alpha, beta, y, J);
	else if (nA <= 8 && nB <= 8)
		hipLaunchKernelGGL((k_gemv_slices<T, 8, 8, 1>), grid(1), b, 0, st, M, K, As, ldk, sA, nA, Xs, sX, nB,
		                   fsum, alpha, beta, y, J);
	else
		return false;
	return true;
}

bool launch_gemv_slices(hipStream_t st, int M, int K, const double *As, int64_t ldk, int64_t sA, int nA,
                        const double *Xs, int64_t sX, int nB, int fsum, double alpha, double beta,
                        double *y, const uint16_t *J) {
	return gemv_slices(st, M, K, As, ldk, sA, nA, Xs, sX, nB, fsum, alpha, beta, y, J);
}

bool launch_gemv_slices(hipStream_t st, int M, int K, const double *As, int64_t ldk, int64_t sA, int nA,
                        const double *Xs, int64_t sX, int nB, int fsum, float alpha, float beta,
                        float *y, const uint16_t *JF2) {
	return gemv_slices(st, M, K, As, ldk, sA, nA, Xs, sX, nB, fsum, alpha, beta, y, JF2);
}

// ---------------------------------------------------------------------------------------------
// GEMV with the split of op(A) fused (binary64): one 512-thread workgroup per row of op(A), which
// is a k-contiguous column of `in` (K <= 512 R; up to 8 slices of A and 4 of x).  The row is split in registers level by level
// with the rule of k_split_cols_reg (slot s: the part of level s; then the remainder slot and zero
// slots up to nA), and each slot is multiplied at once by the x slices, which every thread keeps
// in registers for its R indices: the pair dots sum_k As_s[k] Xs_b[k] are exact, so the workgroup
// sums them in any order (one LDS round per level carries the nB pair sums and the next level's
// remainder maximum).  The nA x nB pair values of row i go to P[i * kSgPairs + a * NBM + b]; the
// slices of A are never stored, the matrix is read once.  k_pairs_sum then sums each row's values
// as k_gemv_slices does (Fsum in (level, ia) order, or NearSum) and writes y_i.
// ---------------------------------------------------------------------------------------------
constexpr int kSgBS = 512, kSgNAM = 8, kSgPairs = 64;
template <int R, int NBM>
__global__ __launch_bounds__(kSgBS) void k_split_gemv(int K, int M, const double *in, int64_t ldi,
                                                      int rho, int shift, int cap, const uint16_t *Js,
                                                      int nA, const double *Xs, int64_t sX, int nB,
                                                      double *P) {
	constexpr int BS = kSgBS, NW = BS / 64, NV = NBM + 1;  // per round: NBM sums and a maximum
	__shared__ double red[2][NW][NV];
	const int lane = threadIdx.x % 64, w = threadIdx.x / 64;
	double xr[NBM][R];
#pragma unroll
	for (int b = 0; b < NBM; b++)
#pragma unroll
		for (int q = 0; q < R; q++) {
			const int i = threadIdx.x + BS * q;
			xr[b][q] = (b < nB && i < K) ? Xs[(int64_t)b * sX + i] : 0.0;
		}
	int round = 0;
	// one reduction round over the workgroup: val[0..NBM) summed, val[NBM] maximised; every
	// thread gets the totals (double-buffered by round parity: one barrier per round)
	auto reduce = [&](double (&val)[NV]) {
#pragma unroll
		for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
			for (int b = 0; b < NBM; b++) val[b] += __shfl_xor(val[b], o, 64);
			val[NBM] = std_max(val[NBM], __shfl_xor(val[NBM], o, 64));
		}
		const int buf = round & 1;
		if (lane == 0) {
#pragma unroll
			for (int c = 0; c < NV; c++) red[buf][w][c] = val[c];
		}
		__syncthreads();
#pragma unroll
		for (int c = 0; c < NV; c++) val[c] = red[buf][0][c];
		for (int u = 1; u < NW; u++) {
#pragma unroll
			for (int b = 0; b < NBM; b++) val[b] += red[buf][u][b];
			val[NBM] = std_max(val[NBM], red[buf][u][NBM]);
		}
		round++;
	};
	for (int j = blockIdx.x; j < M; j += gridDim.x) {
		const double *x = in + (int64_t)j * ldi;
		double *Pj = P + (int64_t)j * kSgPairs;
		double v[R];
		double val[NV];
#pragma unroll
		for (int c = 0; c < NV; c++) val[c] = 0.0;
#pragma unroll
		for (int q = 0; q < R; q++) {
			const int i = threadIdx.x + BS * q;
			v[q] = (i < K) ? x[i] : 0.0;
			val[NBM] = std_max(val[NBM], fabs(v[q]));
		}
		reduce(val);
		double mx = val[NBM];
		int s = 0;
		bool zero = false;
		for (s = 0; s < cap; s++) {
			const int tau = SplitRule<double>::tau(mx, Js);
			const double sigma = SplitRule<double>::sigma(rho, tau, shift);
#pragma unroll
			for (int c = 0; c < NV; c++) val[c] = 0.0;
#pragma unroll
			for (int q = 0; q < R; q++) {
				const double part = (v[q] + sigma) - sigma;
				const double rr = v[q] - part;
				v[q] = rr;
#pragma unroll
				for (int b = 0; b < NBM; b++) val[b] = fma(part, xr[b][q], val[b]);
				val[NBM] = std_max(val[NBM], fabs(rr));
			}
			reduce(val);
			if (threadIdx.x < NBM && s < kSgNAM) Pj[s * NBM + threadIdx.x] = val[threadIdx.x];
			mx = val[NBM];
			if (mx == 0.0) {
				zero = true;
				break;
			}
		}
		const int last = zero ? s : cap - 1;
		if (!zero && last + 1 < nA) {  // the remainder slot
#pragma unroll
			for (int c = 0; c < NV; c++) val[c] = 0.0;
#pragma unroll
			for (int q = 0; q < R; q++)
#pragma unroll
				for (int b = 0; b < NBM; b++) val[b] = fma(v[q], xr[b][q], val[b]);
			reduce(val);
			if (threadIdx.x < NBM && last + 1 < kSgNAM) Pj[(last + 1) * NBM + threadIdx.x] = val[threadIdx.x];
		}
		// zero slots
		const int z0 = zero ? last + 1 : last + 2;
		for (int t = z0 + (int)threadIdx.x / NBM; t < nA && t < kSgNAM; t += BS / NBM)
			Pj[t * NBM + threadIdx.x % NBM] = 0.0;
	}
}

// y_i = fma(alpha, sum of row i's pair values, beta * y_i); one thread per row
template <int NBM>
__global__ __launch_bounds__(256) void k_pairs_sum(int M, double *P, int nA, int nB, int fsum,
                                                   double alpha, double beta, double *y,
                                                   const uint16_t *J) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= M) return;
	const double *Pi = P + (int64_t)i * kSgPairs;
	double t;
	if (fsum) {
		t = 0.0;
		for (int ik = 0; ik < kSgNAM + NBM - 1; ik++)
			for (int a = 0; a < kSgNAM; a++) {
				const int b = ik - a;
				if (b >= 0 && b < NBM && a < nA && b < nB) t = t + Pi[a * NBM + b];
			}
	} else {
		// the values in (a, b) order, compacted in place (scratch), then NearSum over them
		double *Pw = P + (int64_t)i * kSgPairs;
		int cnt = 0;
		for (int a = 0; a < nA; a++)
			for (int b = 0; b < nB; b++) Pw[cnt++] = Pw[a * NBM + b];
		StridedVec u{Pw, 1};
		t = nearsum(cnt, u, J);
	}
	y[i] = fma(alpha, t, beta * y[i]);
}

size_t split_gemv_pairs_bytes(int M) { return sizeof(double) * kSgPairs * (size_t)(M > 0 ? M : 1); }

bool split_gemv_fits(int K, int nA, int nB) {
	return nA >= 1 && nA <= kSgNAM && nB >= 1 && nB <= 4 && K >= 1 && K <= 16 * kSgBS;
}

bool launch_split_gemv(hipStream_t st, int K, int M, const double *in, int64_t ldi, int rho, int shift,
                       int cap, const uint16_t *Js, int nA, const double *Xs, int64_t sX, int nB,
                       int fsum, double alpha, double beta, double *y, const uint16_t *J, double *P) {
	if (!split_gemv_fits(K, nA, nB) || M < 1) return false;
	const int grid = M < 1024 ? M : 1024;
	const int gs = (M + 255) / 256;
	hipLaunchKernelGGL((k_split_gemv<16, 4>), dim3(grid), dim3(kSgBS), 0, st, K, M, in, ldi, rho, shift, cap,
	                   Js, nA, Xs, sX, nB, P);
	hipLaunchKernelGGL((k_pairs_sum<4>), dim3(gs), dim3(256), 0, st, M, P, nA, nB, fsum, alpha, beta, y, J);
	return true;
}

// general case: store the products where blasRcsrmm puts them (column ic of CSplit, ld ldcs);
// T = float: the fma1<float> chain of csrmm_n<float> (V = float), or the binary64 chain of
// csrmm_n<double> stored as its float term (V = double, <float,double>)
template <class T, class V = T>
__global__ void k_csrmm_store(int m, const double *As, int64_t ldas, int nA, const int *colind,
                              const int *rowptr, const double *Xs, int64_t ldxs, int nB,
                              int maxlevel, T *Cs, int64_t ldcs) {
	const int r = blockIdx.x * blockDim.x + threadIdx.x;
	if (r >= m) return;
	const int b = rowptr[r], e = rowptr[r + 1];
	int col = 0;
	const int na = (maxlevel + 1 < nA) ? maxlevel + 1 : nA;
	for (int ia = 0; ia < na; ia++) {
		const int numB = (nB < maxlevel + 1 - ia) ? nB : maxlevel + 1 - ia;
		for (int ib = 0; ib < numB; ib++) {
			V tr = (V)0;
			for (int jj = b; jj < e; jj++)
				tr = (V)fma(As[(int64_t)ia * ldas + jj], Xs[(int64_t)ib * ldxs + colind[jj]], (double)tr);
			if constexpr (sizeof(V) == sizeof(T))
				Cs[(int64_t)(col + ib) * ldcs + r] = tr;
			else
				Cs[(int64_t)(col + ib) * ldcs + r] = sum_term<T>(tr);
		}
		col += numB;
	}
}

void launch_csrmm_store(hipStream_t st, int m, const double *As, int64_t ldas, int nA,
                        const int *colind, const int *rowptr, const double *Xs, int64_t ldxs,
                        int nB, int maxlevel, double *Cs, int64_t ldcs) {
	hipLaunchKernelGGL(k_csrmm_store<double>, dim3((m + 255) / 256), dim3(256), 0, st, m, As, ldas, nA,
	                   colind, rowptr, Xs, ldxs, nB, maxlevel, Cs, ldcs);
}

void launch_csrmm_store(hipStream_t st, int m, const double *As, int64_t ldas, int nA,
                        const int *colind, const int *rowptr, const double *Xs, int64_t ldxs,
                        int nB, int maxlevel, float *Cs, int64_t ldcs, int mixed) {
	if (mixed)
		hipLaunchKernelGGL((k_csrmm_store<float, double>), dim3((m + 255) / 256), dim3(256), 0, st, m, As,
		                   ldas, nA, colind, rowptr, Xs, ldxs, nB, maxlevel, Cs, ldcs);
	else
		hipLaunchKernelGGL(k_csrmm_store<float>, dim3((m + 255) / 256), dim3(256), 0, st, m, As, ldas, nA,
		                   colind, rowptr, Xs, ldxs, nB, maxlevel, Cs, ldcs);
}

// ozblasRaxpy with reproMode != 0 (src/cpu/ozblas_axpy.cpp:142-144): y = fma1(alpha, x, y)
template <class T>
__global__ void k_axpy(int n, T alpha, const T *x, T *y) {
	for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
	     i += (int64_t)gridDim.x * blockDim.x)
		y[i] = fma1_t(alpha, x[i], y[i]);
}
// y = y + alpha * x without contraction (the cblas_?axpy path of reproMode == 0)
template <class T>
__global__ void k_axpy_plain(int n, T alpha, const T *x, T *y) {
	for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
	     i += (int64_t)gridDim.x * blockDim.x)
		y[i] = y[i] + alpha * x[i];
}
// CG update p = beta * p (cblas_?scal, src/cpu/ozblas_cg.cpp:97), then ozblasRaxpy(1, r, p) (:98):
// fma1(1, r, p), or p + r on the reproMode == 0 path (the same for binary64, not for binary32,
// where fma1 rounds twice)
template <class T>
__global__ void k_scal_axpy(int n, T beta, const T *r, T *p, int fused) {
	for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
	     i += (int64_t)gridDim.x * blockDim.x) {
		const T q = beta * p[i];
		p[i] = fused ? fma1_t((T)1, r[i], q) : q + (T)1 * r[i];
	}
}

// t = b - A x with the reference's plain CSRMV (csrmv, src/cpu/ozblas_blas_wrapper.cpp:755-778:
// an fma1 chain per row, then fma1(-1, Tr, 1 * t) as beta = 1), the first step of the CG's true
// residual (getTrueResidual, src/cpu/ozblas_cg.cpp:4-21)
template <class T>
__global__ void k_true_residual(int n, const T *val, const int *colind, const int *rowptr, const T *x,
                                const T *b, T *t) {
	for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n;
	     r += (int64_t)gridDim.x * blockDim.x) {
		T tr = (T)0;
		for (int i = rowptr[r]; i < rowptr[r + 1]; i++) tr = fma1_t(val[i], x[colind[i]], tr);
		t[r] = fma1_t((T)-1, tr, (T)1 * b[r]);
	}
}

static int grid1(int64_t n) {
	int64_t g = (n + 255) / 256;
	if (g > 2048) g = 2048;
	return g < 1 ? 1 : (int)g;
}

void launch_axpy_host(hipStream_t st, int n, double alpha, const double *x, double *y,
                      int fused) {
	if (fused)
		hipLaunchKernelGGL(k_axpy<double>, dim3(grid1(n)), dim3(256), 0, st, n, alpha, x, y);
	else
		hipLaunchKernelGGL(k_axpy_plain<double>, dim3(grid1(n)), dim3(256), 0, st, n, alpha, x, y);
}

void launch_axpy_host(hipStream_t st, int n, float alpha, const float *x, float *y, int fused) {
	if (fused)
		hipLaunchKernelGGL(k_axpy<float>, dim3(grid1(n)), dim3(256), 0, st, n, alpha, x, y);
	else
		hipLaunchKernelGGL(k_axpy_plain<float>, dim3(grid1(n)), dim3(256), 0, st, n, alpha, x, y);
}

void launch_true_residual(hipStream_t st, int n, const double *val, const int *colind,
                          const int *rowptr, const double *x, const double *b, double *t) {
	hipLaunchKernelGGL(k_true_residual<double>, dim3(grid1(n)), dim3(256), 0, st, n, val, colind, rowptr,
	                   x, b, t);
}

void launch_true_residual(hipStream_t st, int n, const float *val, const int *colind,
                          const int *rowptr, const float *x, const float *b, float *t) {
	hipLaunchKernelGGL(k_true_residual<float>, dim3(grid1(n)), dim3(256), 0, st, n, val, colind, rowptr,
	                   x, b, t);
}

void launch_scal_axpy(hipStream_t st, int n, double beta, const double *r, double *p, int fused) {
	hipLaunchKernelGGL(k_scal_axpy<double>, dim3(grid1(n)), dim3(256), 0, st, n, beta, r, p, fused);
}

void launch_scal_axpy(hipStream_t st, int n, float beta, const float *r, float *p, int fused) {
	hipLaunchKernelGGL(k_scal_axpy<float>, dim3(grid1(n)), dim3(256), 0, st, n, beta, r, p, fused);
}

// ---------------------------------------------------------------------------------------------
// Device-resident CG (ozblasRcg, src/cpu/ozblas_cg.cpp:24-128).  Slice counts and the scalar
// recurrences stay in device memory (CgDev), so the host enqueues every iteration without waiting.
// A kernel of iteration `it` does nothing when it >= S->stopAt: the residual test of iteration it
// sets stopAt = it + 1 (the p update of that iteration still runs: the reference breaks after it),
// and a CSR product whose vector has more slices than its register bound sets stopAt = it and
// overflow = 1, after which the host repeats the solve on the synchronous path.
// ---------------------------------------------------------------------------------------------
__device__ inline bool cg_skip(const CgDev *S, int it) { return S != nullptr && it >= S->stopAt; }

// the row sum of the product when p has more slices than the register kernel holds: each pair
// sum runs over the row's nonzeros in order (exact) into Cs[q * ldc + r], then NearSum over the
// row's pairs, or Fsum in (level, ia) order computed on the fly
__device__ inline double csrmm_big_row(int r, const double *As, int64_t ldas, int nA, const int *colind,
                                       const int *rowptr, const VsRef &X, int nB, int fsum, double *Cs,
                                       int64_t ldc, const uint16_t *J) {
	const bool stored = X.plan->bad;
	const double *Xv = (const double *)X.x;
	const int b = rowptr[r], e = rowptr[r + 1];
	auto pair = [&](int ia, int ib) {
		double acc = 0.0;
		for (int j = b; j < e; j++) {
			const double xv = stored ? X.slices[(int64_t)ib * X.lds + colind[j]]
			                         : vs_slot<double>(Xv[colind[j]], X.plan, ib);
			acc = fma(As[(int64_t)ia * ldas + j], xv, acc);
		}
		return acc;
	};
	const int maxlevel = (nA - 1) + (nB - 1);
	double t;
	if (fsum) {
		t = 0.;
		for (int ik = 0; ik <= maxlevel; ik++)
			for (int ia = 0; ia < nA; ia++) {
				const int ib = ik - ia;
				if (ib >= 0 && ib < nB) t = t + pair(ia, ib);
			}
	} else {
		int cnt = 0;
		for (int ia = 0; ia < nA; ia++)
			for (int ib = 0; ib < nB; ib++) Cs[(int64_t)(cnt++) * ldc + r] = pair(ia, ib);
		StridedVec v{Cs + r, ldc};
		t = nearsum(cnt, v, J);
	}
	return t;
}

// q = A p with the pre-split matrix (nA <= NAM) and the device slice count of p; alpha = 1, beta = 0
// as ozblasRcsrmv(..., 1., ..., 0., q) at src/cpu/ozblas_cg.cpp:82.  p's slices come from p and its
// plan (or its stored slices when the plan is marked bad); more than NBM of them take the
// per-pair row loop (csrmm_big_row, pair sums through Cs: m x nA x maxVecSlices doubles at pitch
// ldc).  The reductions of q's split accumulate into preY.
// T = float: the <float,float> CG -- p, q binary32, the pair sums and NearSum in float (J = JF2); the
// host caps p's slices at NBM for it (more overflow to the synchronous path).
template <int NAM, int NBM, class T, class V = T>
__global__ __launch_bounds__(256) void k_csrmm_sum_dev(int m, const double *As, int64_t ldas, int nA,
                                                       const int *colind, const int *rowptr, VsRef X,
                                                       int fsum, T *y, double *Cs, int64_t ldc,
                                                       const uint16_t *J, CgDev *S, int it, VsPre *preY) {
	if (cg_skip(S, it)) return;
	const int nB = X.info[0];
	if (nB > S->maxVecSlices) {
		if (blockIdx.x == 0 && threadIdx.x == 0) {
			S->overflow = 1;
			S->stopAt = it;
		}
		return;
	}
	const int r = blockIdx.x * blockDim.x + threadIdx.x;
	double mx = 0.0;
	int lo = 0, bad = 0;
	if (r < m) {
		const int maxlevel = (nA - 1) + (nB - 1);
		T t = 0;
		if (nB > NBM) {
			if constexpr (sizeof(T) == 8) t = csrmm_big_row(r, As, ldas, nA, colind, rowptr, X, nB, fsum, Cs, ldc, J);
		} else if (X.plan->bad) {
			t = csrmm_row_sum<NAM, NBM, T, double, false, V>(r, As, ldas, nA, colind, rowptr, X.slices, X.lds, nB,
			                                                 maxlevel, fsum, J);
		} else {
			VsGen<T, NBM> g;
			g.init(X.plan);
			t = csrmm_row_sum<NAM, NBM, T, T, true, V>(r, As, ldas, nA, colind, rowptr, nullptr, 0, nB, maxlevel,
			                                           fsum, J, (const T *)X.x, &g);
		}
		const T v = fma1_t((T)1, t, (T)0 * y[r]);
		y[r] = v;
		vs_stats_acc((double)v, mx, lo, bad);
	}
	vs_stats_flush<256>(mx, lo, bad, preY);
}

// Completion of a device DOT in one workgroup, after the product kernel: sums the nA x nB pair
// dots (sumOrder 2) and applies `mode`
//   0: C[0] = fma(1, t, 0 * C[0])                        (ozblasRdot's GEMM epilogue)
//   1: tmp = t, alpha = dold / tmp                       (src/cpu/ozblas_cg.cpp:83-84)
//   2: dnew = t, resi = sqrt(dnew), beta = dnew / dold,  (src/cpu/ozblas_cg.cpp:91-94,100-112)
//      dold = dnew, history and the residual test
// then clears vals for the next use.
// With f32 (binary32 DOT) the pair dots round to float first, the sum is the float one (J is then
// JF2), C is a float (C[0] = fma1<float>(1, t, 0 * C[0])) and modes 1 and 2 run in binary32.
// The pair dots are the per-workgroup partials part[p * kDotMaxWG + w] of k_dots_reg's nwg
// workgroups, summed here by the whole workgroup (exact in any order), or with part == nullptr
// the sums dots_general accumulated in vals.
template <int BS>
__device__ void dots_finish(int nA, int nB, double *vals, const double *part, int nwg, const uint16_t *J,
                            int fsum, CgDev *S, int it, int mode, void *C, int f32) {
	static_assert(kDotMaxWG <= 2 * BS, "two partials per thread and pair");
	__shared__ double psum[kDotPairs];
	__shared__ double red[BS / 64][16];
	const int npairs = nA * nB;
	double *src = vals;
	if (part) {
		// pairs in chunks of 16: every thread loads its (at most two) partials of each pair of the
		// chunk, all loads in flight together (pair p's partials are contiguous: pair-major), then a
		// sum over each wave (DPP) and the waves' sums in LDS
		const int t = threadIdx.x, w = t / 64, l = t % 64;
		if (npairs <= kDotFinOne) {
			// one round: every load of every pair in flight together (the partials come from
			// memory, so a round is a full memory latency)
			double v[kDotFinOne];
#pragma unroll
			for (int q = 0; q < kDotFinOne; q++) {
				// unconditional loads inside the buffer (kDotPairs x kDotMaxWG); the rows q >= npairs
				// are summed too and never read
				const double *row = part + (int64_t)q * kDotMaxWG;
				const double a = row[t], b = row[t + BS];
				v[q] = (t < nwg ? a : 0.0) + (t + BS < nwg ? b : 0.0);
			}
#pragma unroll
			for (int q = 0; q < kDotFinOne; q++) v[q] = wave_sum63_f64(v[q]);
			__shared__ double redOne[BS / 64][kDotFinOne];
			if (l == 63) {
#pragma unroll
				for (int q = 0; q < kDotFinOne; q++) redOne[w][q] = v[q];
			}
			__syncthreads();
			if (t < npairs) {
				double r = redOne[0][t];
				for (int k = 1; k < BS / 64; k++) r += redOne[k][t];
				psum[t] = r;
			}
			__syncthreads();
		}
		for (int p0 = (npairs <= kDotFinOne) ? npairs : 0; p0 < npairs; p0 += 16) {
			double v[16];
#pragma unroll
			for (int q = 0; q < 16; q++) {
				const double *row = part + (int64_t)min(p0 + q, kDotPairs - 1) * kDotMaxWG;
				const double a = row[t], b = row[t + BS];
				v[q] = (t < nwg ? a : 0.0) + (t + BS < nwg ? b : 0.0);
			}
#pragma unroll
			for (int q = 0; q < 16; q++) v[q] = wave_sum63_f64(v[q]);
			if (l == 63) {
#pragma unroll
				for (int q = 0; q < 16; q++) red[w][q] = v[q];
			}
			__syncthreads();
			if (t < 16 && p0 + t < npairs) {
				double r = red[0][t];
				for (int k = 1; k < BS / 64; k++) r += red[k][t];
				psum[p0 + t] = r;
			}
			__syncthreads();
		}
		src = psum;
	}
	const int maxlevel = (nA - 1) + (nB - 1);  // DOT forces fastMode = 0 (src/cpu/ozblas_gemm.cpp:25-28)
	double tw = 0.0;
	bool waveDone = false;
	if (!f32 && !fsum && npairs <= 64) {
		// NearSum across the lanes of wave 0 when every pair dot is finite
		if (threadIdx.x >= 64) return;
		const double pv = ((int)threadIdx.x < npairs) ? src[threadIdx.x] : 0.0;
		const bool finite = __all(!(fabs(pv) > 1.7976931348623157e308) && pv == pv);
		if (finite) {
			tw = nearsum_wave(npairs, pv, J);
			waveDone = true;
		}
	}
	if (threadIdx.x != 0) return;
	if (f32) {
		float *fv = vals_to_float(npairs, src);
		const float t = dot_pairs_sum_f(npairs, fv, nA, nB, maxlevel, fsum, J);
		if (mode == 0) {
			float *Cf = (float *)C;
			Cf[0] = fma1_f(1.f, t, 0.f * Cf[0]);
		} else {
			// the <float,float> CG's recurrences in binary32 (the CgDev fields hold binary32 values)
			const float r = fma1_f(1.f, t, 0.f * 0.f);
			if (mode == 1) {
				S->tmp = r;
				S->alpha = (float)S->dold / r;
			} else {
				const float dnew = r;
				const float resi = (float)sqrt((double)dnew);
				const float nrmb = (float)S->nrmb, tol = (float)S->tol;
				S->dnew = dnew;
				S->resi = resi;
				S->beta = dnew / (float)S->dold;
				S->dold = dnew;
				S->numiter = it + 1;
				if (S->verbose > 0) {
					if (resi / nrmb < tol || S->numiter % S->verbose == 0) {
						if (S->hist) S->hist[S->vcnt] = resi / nrmb;
						S->vcnt = S->vcnt + 1;
					}
				}
				if (resi / nrmb < tol) S->stopAt = it + 1;
			}
		}
		for (int q = 0; q < npairs; q++) vals[q] = 0.0;
		return;
	}
	const double t = waveDone ? tw : dot_pairs_sum(npairs, src, nA, nB, maxlevel, fsum, J);
	const double r = fma(1.0, t, 0.0 * 0.0);
	if (mode == 0) {
		double *Cd = (double *)C;
		Cd[0] = fma(1.0, t, 0.0 * Cd[0]);
	} else if (mode == 1) {
		S->tmp = r;
		S->alpha = S->dold / r;
	} else {
		const double dnew = r;
		const double resi = sqrt(dnew);
		S->dnew = dnew;
		S->resi = resi;
		S->beta = dnew / S->dold;
		S->dold = dnew;
		S->numiter = it + 1;
		if (S->verbose > 0) {
			if (resi / S->nrmb < S->tol || S->numiter % S->verbose == 0) {
				if (S->hist) S->hist[S->vcnt] = resi / S->nrmb;
				S->vcnt = S->vcnt + 1;
			}
		}
		if (resi / S->nrmb < S->tol) S->stopAt = it + 1;
	}
	for (int q = 0; q < npairs; q++) vals[q] = 0.0;
}

// slot q of element i of a split vector: generated from the plan, or read from the stored slices
template <class TI>
__device__ inline double vs_ref_slot(const VsRef &V, bool stored, int64_t i, int q) {
	return stored ? V.slices[(int64_t)q * V.lds + i] : vs_slot<TI>(((const TI *)V.x)[i], V.plan, q);
}

// The same for any slice counts (pairs in chunks of kDotChunk, exact partials added atomically to
// vals); k_dots_reg runs it when the counts exceed its registers.
template <int BS, class TI>
__device__ void dots_general(int n, int nA, int nB, const VsRef &X, const VsRef &Y, double *vals) {
	__shared__ double red[BS / 64];
	const int npairs = nA * nB;
	const bool sx = X.plan->bad, sy = Y.plan->bad;
	for (int p0 = 0; p0 < npairs; p0 += kDotChunk) {
		double acc[kDotChunk];
#pragma unroll
		for (int q = 0; q < kDotChunk; q++) acc[q] = 0.0;
		for (int64_t i = blockIdx.x * (int64_t)BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * BS) {
#pragma unroll
			for (int q = 0; q < kDotChunk; q++) {
				const int p = p0 + q;
				if (p < npairs) {
					const int ia = p % nA, ib = p / nA;
					acc[q] = fma(vs_ref_slot<TI>(X, sx, i, ia), vs_ref_slot<TI>(Y, sy, i, ib), acc[q]);
				}
			}
		}
#pragma unroll
		for (int q = 0; q < kDotChunk; q++) {
			const double v = block_sum_d<BS>(acc[q], red);
			if (threadIdx.x == 0 && p0 + q < npairs && v != 0.0) atomicAdd(&vals[p0 + q], v);
		}
	}
}

// the slots of element i (0 beyond n) of a split vector into registers, masked to the count
template <class TI, int NS>
__device__ __forceinline__ void vs_ref_load(const VsRef &V, bool stored, const VsGen<TI, NS> &g, int cnt,
                                            int64_t i, bool in, double (&out)[NS]) {
	if (stored) {
#pragma unroll
		for (int q = 0; q < NS; q++) {
			const double sv = V.slices[(int64_t)min(q, cnt - 1) * V.lds + (in ? i : 0)];
			out[q] = (in && q < cnt) ? sv : 0.0;
		}
	} else {
		g.gen(in ? ((const TI *)V.x)[i] : (TI)0, out);
#pragma unroll
		for (int q = 0; q < NS; q++)
			if (q >= cnt) out[q] = 0.0;
	}
}

template <class TI, int NS>
__device__ __forceinline__ void vs_gen_masked(const VsGen<TI, NS> &g, TI v, int cnt, double (&out)[NS]) {
	g.gen(v, out);
#pragma unroll
	for (int q = 0; q < NS; q++)
		if (q >= cnt) out[q] = 0.0;
}

// DOT of two split vectors with device slice counts, slice counts up to NAM x NBM: each element is
// read once (two elements per thread and step, their loads issued together), its slices generated
// in registers from the plans, and all pair products accumulate in registers.  Each workgroup
// stores its pair partials to part[p * kDotMaxWG + blockIdx.x] -- no atomics on a few shared
// addresses -- and k_dots_fin sums them (integer multiples of the product unit below 2^53: any
// order is exact).  Larger counts take dots_general.
// the pair accumulation over this thread's elements; SX / SY: the vector's slices are stored (its
// plan is marked bad) rather than generated
template <int BS, int NAM, int NBM, class TI, bool SX, bool SY>
__device__ __forceinline__ void dots_reg_acc(int n, const VsRef &X, const VsRef &Y, int nA, int nB,
                                             double (&acc)[NAM * NBM]) {
	VsGen<TI, NAM> gx;
	VsGen<TI, NBM> gy;
	if constexpr (!SX) gx.init(X.plan);
	if constexpr (!SY) gy.init(Y.plan);
	const TI *xv = (const TI *)X.x, *yv = (const TI *)Y.x;
	const int64_t stride = (int64_t)gridDim.x * BS;
	constexpr int E = (SX || SY || sizeof(TI) == 4) ? 2 : 4;  // elements per thread and step
	for (int64_t i = blockIdx.x * (int64_t)BS + threadIdx.x; i < n; i += E * stride) {
		// every element's loads in flight, then one element's slices at a time
		TI xe[E], ye[E];
#pragma unroll
		for (int e = 0; e < E; e++) {
			// unconditional loads at an index clamped into the vector, then the select
			const int64_t ie = i + e * stride, ic = ie < n ? ie : (int64_t)n - 1;
			if constexpr (!SX) {
				const TI xl = xv[ic];
				xe[e] = ie < n ? xl : (TI)0;
			} else {
				xe[e] = 0;
			}
			if constexpr (!SY) {
				const TI yl = yv[ic];
				ye[e] = ie < n ? yl : (TI)0;
			} else {
				ye[e] = 0;
			}
		}
#pragma unroll
		for (int e = 0; e < E; e++) {
			const int64_t ie = i + e * stride;
			const bool in = ie < n;
			double xa[NAM], yb[NBM];
			if constexpr (SX) vs_ref_load<TI, NAM>(X, true, gx, nA, ie, in, xa);
			else vs_gen_masked<TI, NAM>(gx, xe[e], nA, xa);
			if constexpr (SY) vs_ref_load<TI, NBM>(Y, true, gy, nB, ie, in, yb);
			else vs_gen_masked<TI, NBM>(gy, ye[e], nB, yb);
#pragma unroll
			for (int a = 0; a < NAM; a++)
#pragma unroll
				for (int b = 0; b < NBM; b++) acc[a * NBM + b] = fma(xa[a], yb[b], acc[a * NBM + b]);
		}
	}
}

template <int BS, int NAM, int NBM, class TI>
__global__ __launch_bounds__(BS) void k_dots_reg(int n, VsRef X, VsRef Y, double *vals, double *part,
                                                 const CgDev *S, int it) {
	static_assert(NAM * NBM <= kDotPairs && BS % 64 == 0 && BS <= 256, "pair partial layout");
	if (cg_skip(S, it)) return;
	const int nA = X.info[0], nB = Y.info[0];
	if (nA > NAM || nB > NBM) {
		dots_general<BS, TI>(n, nA, nB, X, Y, vals);
		return;
	}
	const bool sx = X.plan->bad, sy = Y.plan->bad;
	__shared__ double red[BS / 64][NAM * NBM];
	double acc[NAM * NBM];
#pragma unroll
	for (int q = 0; q < NAM * NBM; q++) acc[q] = 0.0;
	if (!sx && !sy) dots_reg_acc<BS, NAM, NBM, TI, false, false>(n, X, Y, nA, nB, acc);
	else if (!sx) dots_reg_acc<BS, NAM, NBM, TI, false, true>(n, X, Y, nA, nB, acc);
	else if (!sy) dots_reg_acc<BS, NAM, NBM, TI, true, false>(n, X, Y, nA, nB, acc);
	else dots_reg_acc<BS, NAM, NBM, TI, true, true>(n, X, Y, nA, nB, acc);
	const int w = threadIdx.x / 64, l = threadIdx.x % 64;
	// wave sums on the DPP network, 16 pairs at a time (skipping the rows a >= nA; absent pairs hold
	// zeros and are not read); lane 63 holds the sums
	constexpr int QC = 16;
	static_assert((NAM * NBM) % QC == 0, "pair chunks");
#pragma unroll
	for (int q0 = 0; q0 < NAM * NBM; q0 += QC) {
		if (q0 / NBM >= nA) continue;
#pragma unroll
		for (int q = 0; q < QC; q++) acc[q0 + q] = wave_sum63_f64(acc[q0 + q]);
		if (l == 63) {
#pragma unroll
			for (int q = 0; q < QC; q++) red[w][q0 + q] = acc[q0 + q];
		}
	}
	__syncthreads();
	for (int q = threadIdx.x; q < NAM * NBM; q += BS) {
		const int a = q / NBM, b = q % NBM;
		if (a < nA && b < nB) {
			double v = red[0][q];
			for (int k = 1; k < BS / 64; k++) v += red[k][q];
			part[(int64_t)(a + nA * b) * kDotMaxWG + blockIdx.x] = v;  // pair-major
		}
	}
}

// The completion of k_dots_reg in one workgroup (its nwg workgroups' partials, or the pair sums
// the general path accumulated in vals)
template <int BS, int NAM, int NBM>
__global__ __launch_bounds__(BS) void k_dots_fin(int nwg, const int *infoX, const int *infoY,
                                                 double *vals, const double *part, const uint16_t *J,
                                                 int fsum, CgDev *S, int it, int mode, void *C, int f32) {
	if (cg_skip(S, it)) return;
	const int nA = infoX[0], nB = infoY[0];
	const bool reg = nA <= NAM && nB <= NBM;
	dots_finish<BS>(nA, nB, vals, reg ? part : nullptr, nwg, J, fsum, S, it, mode, C, f32);
}

// x = fma(alpha, p, x); r = fma(-alpha, q, r)   (src/cpu/ozblas_cg.cpp:87-88, ozblas_axpy.cpp),
// the first kernel of the AXPY category; r's split reductions into preR
// (T = float: fma1<float> with the binary32 alpha)
template <class T>
__global__ __launch_bounds__(256) void k_axpy2_dev(int n, CgDev *S, int it, const T *P, T *X, const T *Q, T *R,
                                                   VsPre *preR) {
	cg_clock(S, it, kClkAxpy);
	if (cg_skip(S, it)) return;
	const T a = (T)S->alpha;
	double mx = 0.0;
	int lo = 0, bad = 0;
	for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
	     i += (int64_t)gridDim.x * blockDim.x) {
		X[i] = fma1_t(a, P[i], X[i]);
		const T r = fma1_t(-a, Q[i], R[i]);
		R[i] = r;
		vs_stats_acc((double)r, mx, lo, bad);
	}
	vs_stats_flush<256>(mx, lo, bad, preR);
}

// p = beta * p, then p = fma(1, r, p)   (src/cpu/ozblas_cg.cpp:97-98); p's split reductions into preP
template <class T>
__global__ __launch_bounds__(256) void k_scal_axpy_dev(int n, CgDev *S, int it, const T *R, T *P, VsPre *preP) {
	cg_clock(S, it, kClkAxpy);
	if (cg_skip(S, it)) return;
	const T b = (T)S->beta;
	double mx = 0.0;
	int lo = 0, bad = 0;
	for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
	     i += (int64_t)gridDim.x * blockDim.x) {
		const T p = fma1_t((T)1, R[i], b * P[i]);
		P[i] = p;
		vs_stats_acc((double)p, mx, lo, bad);
	}
	vs_stats_flush<256>(mx, lo, bad, preP);
}

template <class T, class V = T>
static void csrmm_sum_dev(hipStream_t st, int m, const double *As, int64_t ldas, int nA, const int *colind,
                          const int *rowptr, VsRef X, int fsum, T *y, double *Cs, int64_t ldc, const uint16_t *J,
                          CgDev *S, int it, VsPre *preY) {
	constexpr int kNbReg = 8;
	if (nA <= 2)  // the matrix's count is known here; p's count only on the device
		hipLaunchKernelGGL((k_csrmm_sum_dev<2, kNbReg, T, V>), dim3((m + 255) / 256), dim3(256), 0, st, m, As,
		                   ldas, nA, colind, rowptr, X, fsum, y, Cs, ldc, J, S, it, preY);
	else
		hipLaunchKernelGGL((k_csrmm_sum_dev<4, kNbReg, T, V>), dim3((m + 255) / 256), dim3(256), 0, st, m, As,
		                   ldas, nA, colind, rowptr, X, fsum, y, Cs, ldc, J, S, it, preY);
}

void launch_csrmm_sum_dev(hipStream_t st, int m, const double *As, int64_t ldas, int nA,
                          const int *colind, const int *rowptr, VsRef X, int fsum, double *y,
                          double *Cs, int64_t ldc, const uint16_t *J, CgDev *S, int it, VsPre *preY) {
	csrmm_sum_dev(st, m, As, ldas, nA, colind, rowptr, X, fsum, y, Cs, ldc, J, S, it, preY);
}

void launch_csrmm_sum_dev(hipStream_t st, int m, const double *As, int64_t ldas, int nA,
                          const int *colind, const int *rowptr, VsRef X, int fsum, float *y,
                          double *Cs, int64_t ldc, const uint16_t *JF2, CgDev *S, int it, VsPre *preY,
                          int mixed) {
	if (mixed)
		csrmm_sum_dev<float, double>(st, m, As, ldas, nA, colind, rowptr, X, fsum, y, Cs, ldc, JF2, S, it, preY);
	else
		csrmm_sum_dev<float>(st, m, As, ldas, nA, colind, rowptr, X, fsum, y, Cs, ldc, JF2, S, it, preY);
}

size_t dots_part_bytes() { return sizeof(double) * kDotPairs * kDotMaxWG; }

void launch_dots_dev(hipStream_t st, int n, VsRef X, VsRef Y, double *vals, double *part,
                     const uint16_t *J, int fsum, CgDev *S, int it, int mode, void *C, int f32) {
	constexpr int kRA = 8, kRB = 8;
	int gx = (n + 1023) / 1024;  // four elements per thread and step
	if (gx > kDotMaxWG) gx = kDotMaxWG;
	if (gx < 1) gx = 1;
	if (f32)
		hipLaunchKernelGGL((k_dots_reg<256, kRA, kRB, float>), dim3(gx), dim3(256), 0, st, n, X, Y, vals,
		                   part, (const CgDev *)S, it);
	else
		hipLaunchKernelGGL((k_dots_reg<256, kRA, kRB, double>), dim3(gx), dim3(256), 0, st, n, X, Y, vals,
		                   part, (const CgDev *)S, it);
	hipLaunchKernelGGL((k_dots_fin<256, kRA, kRB>), dim3(1), dim3(256), 0, st, gx, X.info, Y.info, vals,
	                   (const double *)part, J, fsum, S, it, mode, C, f32);
}

// the AXPYs that also reduce for the next split: fewer workgroups (each reduces once)
static int grid_red(int64_t n) {
	int64_t g = (n + 255) / 256;
	if (g > 512) g = 512;
	return g < 1 ? 1 : (int)g;
}

void launch_axpy2_dev(hipStream_t st, int n, CgDev *S, int it, const double *P, double *X,
                      const double *Q, double *R, VsPre *preR) {
	hipLaunchKernelGGL(k_axpy2_dev<double>, dim3(grid_red(n)), dim3(256), 0, st, n, S, it, P, X, Q, R, preR);
}

void launch_axpy2_dev(hipStream_t st, int n, CgDev *S, int it, const float *P, float *X, const float *Q,
                      float *R, VsPre *preR) {
	hipLaunchKernelGGL(k_axpy2_dev<float>, dim3(grid_red(n)), dim3(256), 0, st, n, S, it, P, X, Q, R, preR);
}

void launch_scal_axpy_dev(hipStream_t st, int n, CgDev *S, int it, const double *R, double *P,
                          VsPre *preP) {
	hipLaunchKernelGGL(k_scal_axpy_dev<double>, dim3(grid_red(n)), dim3(256), 0, st, n, S, it, R, P, preP);
}

void launch_scal_axpy_dev(hipStream_t st, int n, CgDev *S, int it, const float *R, float *P, VsPre *preP) {
	hipLaunchKernelGGL(k_scal_axpy_dev<float>, dim3(grid_red(n)), dim3(256), 0, st, n, S, it, R, P, preP);
}

}  // namespace oz
