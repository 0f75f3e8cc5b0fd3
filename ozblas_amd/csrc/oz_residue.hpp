// oz_residue.hpp -- device routines of the residue form shared by the single-problem kernels
// (oz_residue.hip) and the grouped batched kernels (oz_batched.hip): the balanced residues of one
// element modulo each modulus, the tiled INT8 residue-plane writer, and the Chinese-remainder
// reconstruction of one output element.  The scheme itself is described in oz_residue.hip.
#pragma once
#include "oz_digits.hpp"

namespace oz {

// ---- device: residues of one element ------------------------------------------------------
// x = +-f 2^e (f odd, or 0) relative to the column unit 2^lo: x 2^-lo = +-f 2^sh with sh >= 0
struct ResElem {
	double fd;  // f as a binary64 (exact: f < 2^53)
	int sh;
	int neg;
};

__device__ inline ResElem res_elem(double x, int lo) {
	ResElem r{0.0, 0, 0};
	if (x == 0.0) return r;
	const uint64_t b = d2u(x);
	const int E = (int)((b >> 52) & 0x7ff);
	uint64_t f = b & ((uint64_t(1) << 52) - 1);
	int e0;
	if (E == 0) {
		e0 = -1074;
	} else {
		f |= (uint64_t(1) << 52);
		e0 = E - 1075;
	}
	const int tz = __builtin_ctzll(f);
	f >>= tz;
	r.fd = (double)f;
	r.sh = e0 + tz - lo;  // in [0, span) by the choice of lo
	r.neg = (int)(b >> 63);
	return r;
}

// Residues by groups of three moduli.  First r = (+-f mod P) (2^sh mod P) mod P in binary64, P the
// group's product (< 2^24): f < 2^53 and the product of two values below 2^24 stay exact, the
// quotient estimates are off by at most one (|error| <= 2 f 2^-53 / P < 1).  Then each member's
// balanced residue from r < 2^24 in binary32: in [-(p-1)/2, (p-1)/2] for odd p, [-128, 127] for 256.
__device__ __forceinline__ int res_group(const ResElem &e, double P, double invP, const uint32_t *pw2) {
	if (e.fd == 0.0) return 0;
	double q = floor(e.fd * invP);
	double r = fma(-q, P, e.fd);
	r += (r < 0.0) ? P : 0.0;
	r -= (r >= P) ? P : 0.0;
	const double x = r * (double)pw2[e.sh];  // < 2^48: exact
	q = floor(x * invP);
	r = fma(-q, P, x);
	r += (r < 0.0) ? P : 0.0;
	r -= (r >= P) ? P : 0.0;
	// the sign: (-v) mod P, so that every member's residue comes out of res_member already signed
	if (e.neg && r != 0.0) r = P - r;
	return (int)r;
}

// rx - p rint(rx / p) is the balanced residue up to one correction either way (the binary32 quotient
// estimate is off by at most one near a half-integer); q p with a 24-bit multiply (q < 2^20, p < 2^9)
__device__ __forceinline__ int res_member(int rx, int p, float invpf) {
	const int q = (int)rintf((float)rx * invpf);  // rx < 2^24 is exact
	int r = rx - (int)__umul24((uint32_t)q, (uint32_t)p);
	const int hi = (p & 1) ? (p >> 1) : (p >> 1) - 1;  // [-(p-1)/2, (p-1)/2] odd p, [-128, 127] for 256
	const int lo = -(p >> 1);
	r -= (r > hi) ? p : 0;
	r += (r < lo) ? p : 0;
	return r;
}

// the per-group and per-modulus tables of the residue writers, staged in LDS by the whole workgroup
// (2^sh mod P_g only for sh below the operand's span)
struct ResLds {
	uint32_t pw2[kResMaxGroups][kResMaxShift];
	double Pg[kResMaxGroups], invPg[kResMaxGroups];
	int p[kResMaxMod];
	float ipf[kResMaxMod];
};

__device__ inline void res_lds_load(ResLds &T, const ResConst *rc, int nmod, int shLimit) {
	const int ng = (nmod + 2) / 3;
	for (int q = threadIdx.x; q < ng * shLimit; q += blockDim.x) T.pw2[q / shLimit][q % shLimit] = rc->pow2g[q / shLimit][q % shLimit];
	for (int q = threadIdx.x; q < ng; q += blockDim.x) {
		T.Pg[q] = rc->Pg[q];
		T.invPg[q] = rc->invPg[q];
	}
	for (int q = threadIdx.x; q < nmod; q += blockDim.x) {
		T.p[q] = rc->p[q];
		T.ipf[q] = rc->invpf[q];
	}
}

// One 256-thread workgroup: residues of the 16 columns x 64 k-bytes tile (kt, jb) of every modulus
// plane in the INT8 tiled layout (4 bytes per thread).
__device__ inline void res_write_tile_i8(int m, int n, const double *in, int64_t ldi, const int *colLo,
                                         int nmod, const ResLds &T, int8_t *planes, int64_t ldp,
                                         int64_t pstride, int kt, int jb) {
	const int t = threadIdx.x;
	const int j = jb * 16 + t / 16;
	const int lo = (j < n) ? colLo[j] : 0;
	const double *x = in + (int64_t)(j < n ? j : 0) * ldi;
	const int q0 = kt * 64 + 4 * (t % 16);
	ResElem e[4];
#pragma unroll
	for (int u = 0; u < 4; u++) e[u] = res_elem((j < n && q0 + u < m) ? x[q0 + u] : 0.0, lo);
	int8_t *dst = planes + (int64_t)jb * 16 * ldp + (int64_t)kt * 1024 + 4 * t;
	const int ng = (nmod + 2) / 3;
	for (int gi = 0; gi < ng; gi++) {
		int rx[4];
#pragma unroll
		for (int u = 0; u < 4; u++) rx[u] = res_group(e[u], T.Pg[gi], T.invPg[gi], T.pw2[gi]);
		const int s1 = min(3 * gi + 3, nmod);
		for (int s = 3 * gi; s < s1; s++) {
			uint32_t word = 0;
#pragma unroll
			for (int u = 0; u < 4; u++)
				word |= (uint32_t)(uint8_t)(int8_t)res_member(rx[u], T.p[s], T.ipf[s]) << (8 * u);
			*(uint32_t *)(dst + (int64_t)s * pstride) = word;
		}
	}
}

// v mod p for 0 <= v < 2^24 (binary32 quotient estimate off by at most one)
__device__ __forceinline__ int mod_f(int v, int p, float invpf) {
	const float q = floorf((float)v * invpf);
	int r = v - (int)q * p;
	r += (r < 0) ? p : 0;
	r -= (r >= p) ? p : 0;
	return r;
}

// The four elements (i, 4g .. 4g + 3) of one product-residue word: X of each from its residues
// (nmod moduli, nchunk k-chunk planes per modulus summed modulo p), then t[e] = RN(X 2^e0[e]).
// Rw addresses the word in plane 0; plane b is rstride bytes further.
template <int NW>
__device__ inline void res_group_values(const uint8_t *Rw, int64_t rstride, int nchunk,
                                        const ResConst *rc, int nmod, const int *e0, double *t) {
	constexpr int NL = 2 * NW;
	int64_t acc[4][NL];
	double S[4];
#pragma unroll
	for (int e = 0; e < 4; e++) {
		S[e] = 0.0;
#pragma unroll
		for (int l = 0; l < NL; l++) acc[e][l] = 0;
	}
	for (int q = 0; q < nmod; q++) {
		const int p = rc->p[q];
		const float ipf = rc->invpf[q];
		const int u = rc->u[q];
		const double ipd = rc->invpd[q];
		int rs[4] = {0, 0, 0, 0};
		for (int c = 0; c < nchunk; c++) {
			const uint32_t w = *(const uint32_t *)(Rw + (int64_t)(c * nmod + q) * rstride);
#pragma unroll
			for (int e = 0; e < 4; e++) rs[e] += (int)((w >> (8 * e)) & 255);
		}
		uint32_t mp[NL];
#pragma unroll
		for (int l = 0; l < NL; l++) mp[l] = rc->Mp[q][l];
#pragma unroll
		for (int e = 0; e < 4; e++) {
			const int r = (nchunk > 1) ? mod_f(rs[e], p, ipf) : rs[e];
			const int ct = mod_f(r * u, p, ipf);  // < 2^16 before the reduction
			S[e] += (double)ct * ipd;
#pragma unroll
			for (int l = 0; l < NL; l++) acc[e][l] += (int64_t)((uint64_t)(uint32_t)ct * mp[l]);
		}
	}
#pragma unroll
	for (int e = 0; e < 4; e++) {
		const int64_t K = (int64_t)rint(S[e]);
		BigInt<NW> X;
		int64_t carry = 0;
#pragma unroll
		for (int q = 0; q < NW; q++) {
			const int64_t v0 = acc[e][2 * q] - K * (int64_t)rc->M[2 * q] + carry;
			const uint64_t w0 = (uint32_t)v0;
			carry = v0 >> 32;
			const int64_t v1 = acc[e][2 * q + 1] - K * (int64_t)rc->M[2 * q + 1] + carry;
			const uint64_t w1 = (uint32_t)v1;
			carry = v1 >> 32;
			X.w[q] = w0 | (w1 << 32);
		}
		t[e] = big_finish<NW>(X, e0[e]);
	}
}

// one thread per (row i, column group g) of an M x N product: C(i, 4g + e) = fma(alpha, RN(X 2^(lo_i +
// lo_j)), beta C) for the columns of the group inside N
template <int NW>
__device__ inline void res_recon_group(int M, int N, const uint8_t *R, int64_t ldr, int64_t rstride,
                                       int nchunk, const ResConst *rc, int nmod, const int *rowLo,
                                       const int *colLo, double alpha, double beta, double *C,
                                       int64_t ldc, int i, int g) {
	int e0[4];
	const int li = rowLo[i];
#pragma unroll
	for (int e = 0; e < 4; e++) e0[e] = li + ((4 * g + e < N) ? colLo[4 * g + e] : 0);
	double t[4];
	res_group_values<NW>(R + res_word_offset(i, 4 * g, ldr), rstride, nchunk, rc, nmod, e0, t);
#pragma unroll
	for (int e = 0; e < 4; e++) {
		const int j = 4 * g + e;
		if (j < N) {
			double *c = C + (int64_t)j * ldc + i;
			*c = fma(alpha, t[e], beta * *c);
		}
	}
}

}  // namespace oz
