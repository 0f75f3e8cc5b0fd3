// oz_residue.hpp -- device routines of the residue form shared by the single-problem kernels
// (oz_residue.hip) and the grouped batched kernels (oz_batched.hip): the balanced residues of one
// element modulo each modulus, the tiled INT8 residue-plane writer, and the Chinese-remainder
// reconstruction of one output element.  The scheme itself is described in oz_residue.hip.
#pragma once
#include "oz_digits.hpp"

namespace oz {

// ---- device: residues of one element ------------------------------------------------------
// x 2^-lo = sf 2^sh with sf a signed integer below 2^53 in magnitude and sh >= 0: sh is the exponent
// of x's last place relative to lo (clamped at 0), sf = x 2^-(lo + sh) exactly.  lo is at most the
// lowest set bit of every element of the row or column, so sf is an integer either way, and sh is at
// most the shift of x's lowest set bit (below the operand's span).  A zero x gives sf = 0, sh = 0.
struct ResElem {
	double sf;
	int sh;
};

__device__ inline ResElem res_elem(double x, int lo) {
	const int Eb = (int)((d2u(x) >> 52) & 0x7ff);  // biased exponent; 0 for zero and subnormals
	int sh = (Eb > 1 ? Eb : 1) - 1075 - lo;         // the unit of x's last place is 2^(max(Eb,1) - 1075)
	sh = sh > 0 ? sh : 0;
	return ResElem{ldexp(x, -(lo + sh)), sh};
}

// Residues by groups of three moduli: (sf 2^sh) mod P in binary64, P the group's product (< 2^24),
// left in [-P, 2P) (an exact integer congruent to the element modulo P): each floor estimate is off
// by at most one (|error| <= 2 |v| 2^-53 / P < 1), the first remainder times 2^sh mod P stays exact
// (|x| < 2^25 2^24), and res_member takes any such representative.  The sign rides on sf.
__device__ __forceinline__ double res_group(const ResElem &e, double P, double invP, const uint32_t *pw2) {
	double q = floor(e.sf * invP);
	const double r0 = fma(-q, P, e.sf);              // exact, in [-P, 2P)
	const double x = r0 * (double)pw2[e.sh];         // exact, |x| < 2^49
	q = floor(x * invP);
	return fma(-q, P, x);                            // exact, in [-P, 2P)
}

// The balanced residue of an integer rx (|rx| < 2^25) modulo p: q = floor(rx / p + 1/2) exactly, as
// floor(fma(rx, RN(1/p), 1/2)) -- its error, below 2^-28, cannot cross an integer: rx / p + 1/2 is
// an integer only for p = 256, where 1/p is exact, and otherwise keeps 1/(2p) >= 2^-9 from one --
// so r = rx - q p (exact in one fma) lies in [-p/2, p/2): [-(p-1)/2, (p-1)/2] for odd p, [-128, 127]
// for 256.
__device__ __forceinline__ int res_member(double rx, double p, double invp) {
	const double q = floor(fma(rx, invp, 0.5));
	return (int)fma(-q, p, rx);
}

// the per-group and per-modulus tables of the residue writers, staged in LDS by the whole workgroup
// (2^sh mod P_g only for sh below the operand's span)
struct ResLds {
	double Pg[kResMaxGroups], invPg[kResMaxGroups];
	double pd[kResMaxMod], ipd[kResMaxMod];  // p and RN(1/p)
	int shs;  // row length of the shift table
};
// the shift table 2^sh mod P_g (sh below the operand's span, `shs` per group row) in the kernel's
// dynamic LDS: res_lds_bytes(nmod, shLimit) of it, so that a narrow span leaves room for more
// workgroups per CU
extern __shared__ uint32_t resPw2[];
__host__ __device__ inline size_t res_lds_bytes(int nmod, int shLimit) {
	return sizeof(uint32_t) * (size_t)((nmod + 2) / 3) * (size_t)shLimit;
}

// scaleU: the tables of op(A), whose residues carry the factors u_t (ResConst::pow2gU)
__device__ inline void res_lds_load(ResLds &T, const ResConst *rc, int nmod, int shLimit, int scaleU) {
	const int ng = (nmod + 2) / 3;
	const uint32_t(*src)[kResMaxShift] = scaleU ? rc->pow2gU : rc->pow2g;
	for (int q = threadIdx.x; q < ng * shLimit; q += blockDim.x) resPw2[q] = src[q / shLimit][q % shLimit];
	if (threadIdx.x == 0) T.shs = shLimit;
	for (int q = threadIdx.x; q < ng; q += blockDim.x) {
		T.Pg[q] = rc->Pg[q];
		T.invPg[q] = rc->invPg[q];
	}
	for (int q = threadIdx.x; q < nmod; q += blockDim.x) {
		T.pd[q] = (double)rc->p[q];
		T.ipd[q] = rc->invpd[q];
	}
}

// One 256-thread workgroup: residues of the 16 columns x 64 k-bytes tile (kt, jb) of every modulus
// plane in the INT8 tiled layout (4 bytes per thread).
__device__ inline void res_write_tile_i8(int m, int n, const double *in, int64_t ldi, const int *colLo,
                                         int nmod, const ResLds &T, int8_t *planes, int64_t ldp,
                                         int64_t pstride, int kt, int jb) {
	const int t = threadIdx.x;
	if (jb * 16 >= n || kt * 64 >= m) {  // a tile of padding only: zero residues, nothing to compute
		int8_t *dst = planes + (int64_t)jb * 16 * ldp + (int64_t)kt * 1024 + 4 * t;
		for (int s = 0; s < nmod; s++) *(uint32_t *)(dst + (int64_t)s * pstride) = 0u;
		return;
	}
	const int j = jb * 16 + t / 16;
	const int lo = (j < n) ? colLo[j] : 0;
	const double *x = in + (int64_t)(j < n ? j : 0) * ldi;
	const int q0 = kt * 64 + 4 * (t % 16);
	ResElem e[4];
#pragma unroll
	for (int u = 0; u < 4; u++) e[u] = res_elem((j < n && q0 + u < m) ? x[q0 + u] : 0.0, lo);
	int8_t *dst = planes + (int64_t)jb * 16 * ldp + (int64_t)kt * 1024 + 4 * t;
	const int ng = (nmod + 2) / 3;
	const int shs = T.shs;
	for (int gi = 0; gi < ng; gi++) {
		double rx[4];
#pragma unroll
		for (int u = 0; u < 4; u++) rx[u] = res_group(e[u], T.Pg[gi], T.invPg[gi], resPw2 + gi * shs);
		const int s1 = min(3 * gi + 3, nmod);
		for (int s = 3 * gi; s < s1; s++) {
			uint32_t word = 0;
#pragma unroll
			for (int u = 0; u < 4; u++)
				word |= (uint32_t)(uint8_t)(int8_t)res_member(rx[u], T.pd[s], T.ipd[s]) << (8 * u);
			*(uint32_t *)(dst + (int64_t)s * pstride) = word;
		}
	}
}

// v mod p for 0 <= v < 2^24 (binary32 quotient estimate off by at most one; q p with a 24-bit
// multiply: q < 2^24 / p, p <= 256)
__device__ __forceinline__ int mod_f(int v, int p, float invpf) {
	const float q = floorf((float)v * invpf);
	int r = v - (int)__umul24((uint32_t)(int)q, (uint32_t)p);
	r += (r < 0) ? p : 0;
	r -= (r >= p) ? p : 0;
	return r;
}

// an integer-valued binary64 with |d| < 2^51 as int64: d + 1.5 2^52 is exact and its fraction
// field holds d + 2^51
__device__ __forceinline__ int64_t f64_to_i64_exact(double d) {
	return (int64_t)(d2u(d + 0x1.8p52) - 0x4338000000000000ULL);
}

// a += v 2^sh (two's complement) when the bits of a at and above sh are all zero: no carries
template <int NW>
__device__ __forceinline__ void big_or_shifted(BigInt<NW> &a, int64_t v, int sh) {
	const int wi = sh >> 6, b = sh & 63;
	const uint64_t ext = (v < 0) ? ~0ULL : 0ULL;
	const uint64_t lo = (uint64_t)v << b;
	const uint64_t hi = b ? (uint64_t)(v >> (64 - b)) : ext;
#pragma unroll
	for (int q = 0; q < NW; q++) {
		if (q == wi)
			a.w[q] |= lo;
		else if (q == wi + 1)
			a.w[q] |= hi;
		else if (q > wi + 1)
			a.w[q] |= ext;
	}
}

// The four elements (i, 4g .. 4g + 3) of one product-residue word: X of each from its residues
// (nmod moduli, nchunk k-chunk planes per modulus summed modulo p), then t[e] = RN(X 2^e0[e]).
// The planes hold c_t = X u_t mod p_t < 2^8 (op(A) was written with the factors u_t);
// X = sum_t c_t M/p_t - K M is summed in binary64 chunks of 39 bits
// (every product c_t chunk < 2^47, every chunk sum of up to 49 moduli < 2^53: exact in any order),
// K = rint(sum_t c_t / p_t), and the chunks minus K M's are carried into NW 64-bit words.  The words
// of a block of moduli are loaded together (clamped indices; the surplus ones are not used).
// Rw addresses the word in plane 0; plane b is rstride bytes further.
template <int NW>
__device__ __forceinline__ void res_group_values(const uint8_t *Rw, int64_t rstride, int nchunk,
                                        const ResConst *rc, int nmod, const int *e0, double *t) {
	constexpr int NC = (64 * NW + kResChunkBits - 1) / kResChunkBits;
	double A[4][NC];
	double S[4];
#pragma unroll
	for (int e = 0; e < 4; e++) {
		S[e] = 0.0;
#pragma unroll
		for (int c = 0; c < NC; c++) A[e][c] = 0.0;
	}
	// modulus q's residues ct[e] of the four elements enter the sums
	auto absorb = [&](int q, const double (&ct)[4]) {
		const double ipd = rc->invpd[q];
		double mp[NC];
#pragma unroll
		for (int c = 0; c < NC; c++) mp[c] = rc->Mpd[q][c];
#pragma unroll
		for (int e = 0; e < 4; e++) {
			S[e] = fma(ct[e], ipd, S[e]);
#pragma unroll
			for (int c = 0; c < NC; c++) A[e][c] = fma(ct[e], mp[c], A[e][c]);
		}
	};
	if (nchunk == 1) {
		// one k chunk: the words of QB moduli are loaded while the previous QB are absorbed, so
		// 2 QB loads per thread are in flight
		constexpr int QB = 8;
		uint32_t w[QB];
		auto load = [&](uint32_t(&d)[QB], int q0) {
#pragma unroll
			for (int j = 0; j < QB; j++) d[j] = *(const uint32_t *)(Rw + (int64_t)min(q0 + j, nmod - 1) * rstride);
		};
		load(w, 0);
		for (int q0 = 0; q0 < nmod; q0 += QB) {
			uint32_t wn[QB];
			load(wn, min(q0 + QB, nmod - 1));
#pragma unroll
			for (int j = 0; j < QB; j++) {
				if (q0 + j >= nmod) break;
				double ct[4];
#pragma unroll
				for (int e = 0; e < 4; e++) ct[e] = (double)((w[j] >> (8 * e)) & 255);
				absorb(q0 + j, ct);
			}
#pragma unroll
			for (int j = 0; j < QB; j++) w[j] = wn[j];
		}
	} else {
		// several k chunks: the chunks' residues of a modulus are summed (below 2^24) and reduced
		constexpr int QB = 4;
		for (int q0 = 0; q0 < nmod; q0 += QB) {
			int rs[QB][4];
#pragma unroll
			for (int j = 0; j < QB; j++)
#pragma unroll
				for (int e = 0; e < 4; e++) rs[j][e] = 0;
			for (int c = 0; c < nchunk; c++) {
				uint32_t w[QB];
#pragma unroll
				for (int j = 0; j < QB; j++) {
					const int q = min(q0 + j, nmod - 1);
					w[j] = *(const uint32_t *)(Rw + (int64_t)(c * nmod + q) * rstride);
				}
#pragma unroll
				for (int j = 0; j < QB; j++)
#pragma unroll
					for (int e = 0; e < 4; e++) rs[j][e] += (int)((w[j] >> (8 * e)) & 255);
			}
#pragma unroll
			for (int j = 0; j < QB; j++) {
				const int q = q0 + j;
				if (q >= nmod) break;
				const int p = rc->p[q];
				const float ipf = rc->invpf[q];
				double ct[4];
#pragma unroll
				for (int e = 0; e < 4; e++) ct[e] = (double)mod_f(rs[j][e], p, ipf);  // rs < 2^24
				absorb(q, ct);
			}
		}
	}
#pragma unroll
	for (int e = 0; e < 4; e++) {
		const double K = rint(S[e]);
		// chunk c of X is d_c = A_c - K M_c, exact: 0 <= A_c < 49 * 255 * 2^39 < 2^52.61 and
		// |K M_c| < 2^45.  The carries run in binary64: hi = floor(d_c 2^-39) (at most 2^13.61 in
		// magnitude) moves to d_(c+1), which stays below 2^53, and leaves d_c in [0, 2^39); every step
		// is exact.  The top chunk carries the sign, |d_top| < 2^51 because |X| < 2^(64 NW - 1).
		double d[NC];
#pragma unroll
		for (int c = 0; c < NC; c++) d[c] = fma(-K, rc->Md[c], A[e][c]);
#pragma unroll
		for (int c = 0; c + 1 < NC; c++) {
			const double hi = floor(d[c] * 0x1p-39);
			d[c] = fma(-hi, 0x1p39, d[c]);
			d[c + 1] += hi;
		}
		// the chunks occupy disjoint bit ranges: X is their OR
		BigInt<NW> X;
#pragma unroll
		for (int q = 0; q < NW; q++) X.w[q] = 0;
#pragma unroll
		for (int c = 0; c < NC; c++) big_or_shifted<NW>(X, f64_to_i64_exact(d[c]), kResChunkBits * c);
		t[e] = big_finish<NW>(X, e0[e]);
	}
}

// one thread per (row i, column group g) of an M x N product: C(i, 4g + e) = fma(alpha, RN(X 2^(lo_i +
// lo_j)), beta C) for the columns of the group inside N
template <int NW>
__device__ __forceinline__ void res_recon_group(int M, int N, const uint8_t *R, int64_t ldr, int64_t rstride,
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
