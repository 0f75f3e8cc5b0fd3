// oz_gemm_digits.hip -- fast forms: exact digit-plane products on MFMA + one exact rounding.
//
// Operands are balanced radix-256 digit planes (oz_split.hip): row i of A and column j of B are
// x = sum_s d_s 2^(E - 8(s+1)) with d_s in [-128, 127].  All products of plane s of A and plane t
// of B with the same level L = s + t carry the same weight 2^(E_i + F_j - 16 - 8L), so one integer
// accumulator per level is exact as long as |sum| < 2^31; a "bin" groups the pairs (and a k range)
// of one level that fit.  k_digit_gemm_* computes one int32 plane per bin on MFMA:
//   INT8 : v_mfma_i32_32x32x32_i8 (exact int32 accumulate)
//   FP16/BF16: v_mfma_f32_16x16x32_{f16,bf16}; products of 8-bit digits are exact and partial sums
//          stay integers below 2^24 for 1024 k, after which they are moved to an int32 register.
// k_reconstruct then forms N = sum_bins T_b 2^(8(Lmax - L_b)) exactly in a multi-word integer and
// rounds N * 2^(E_i + F_j - 16 - 8 Lmax) once to nearest-even binary64 -- RN of the exact dot
// product, which is what NearSum returns for the exact FP64 split of the plain form.
#include "oz_kernels.hpp"

namespace oz {

typedef int i16v __attribute__((ext_vector_type(16)));
typedef int i4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));
typedef long l2v __attribute__((ext_vector_type(2)));
typedef _Float16 h8v __attribute__((ext_vector_type(8)));
typedef __bf16 b8v __attribute__((ext_vector_type(8)));

constexpr int BM = 128, BN = 128;  // output tile per workgroup (4 waves, 64 x 64 each)
constexpr int BKB = 64;            // k bytes per stage
constexpr int LDSROW = BKB + 16;   // padded LDS row (bytes)

// global -> register staging: 128 rows x 64 bytes per operand = 512 x 16B, 2 per thread
struct Stage {
	uint4 a[2], b[2];
};

__device__ inline void load_stage(Stage &s, const char *Ap, int64_t ldab, const char *Bp,
                                  int64_t ldbb, int i0, int j0, int M, int N, int kb) {
	const int t = threadIdx.x;
	for (int h = 0; h < 2; h++) {
		const int row = t / 4 + 64 * h, ch = t % 4;
		const int ia = i0 + row, jb = j0 + row;
		s.a[h] = (ia < M) ? *(const uint4 *)(Ap + (int64_t)ia * ldab + kb + 16 * ch) : make_uint4(0, 0, 0, 0);
		s.b[h] = (jb < N) ? *(const uint4 *)(Bp + (int64_t)jb * ldbb + kb + 16 * ch) : make_uint4(0, 0, 0, 0);
	}
}

__device__ inline void store_stage(const Stage &s, char *sA, char *sB) {
	const int t = threadIdx.x;
	for (int h = 0; h < 2; h++) {
		const int row = t / 4 + 64 * h, ch = t % 4;
		*(uint4 *)(sA + row * LDSROW + 16 * ch) = s.a[h];
		*(uint4 *)(sB + row * LDSROW + 16 * ch) = s.b[h];
	}
}

// ---- INT8 ----------------------------------------------------------------------------------
// Workgroup = 2 x 2 waves; wave tile = TM x TN MFMA tiles of 32 x 32 (block tile 64TM x 64TN).
// Operands stream through a double-buffered LDS stage of 64 k-bytes; one barrier per stage.  The
// (pair, k) loop of a bin is flattened so the software pipeline does not drain between pairs.
// The MFMA is issued as B-fragment x A-fragment so that lane l%32 indexes the A row: stores of the
// column-major result planes are contiguous across lanes.
template <int TM, int TN>
__global__ __launch_bounds__(256) void k_digit_gemm_i8(int M, int N, const int8_t *A, int64_t ldap,
                                                       int64_t apstride, const int8_t *B,
                                                       int64_t ldbp, int64_t bpstride,
                                                       const int2 *pairs, const LevelBin *bins,
                                                       int32_t *T, int64_t ldt, int64_t tstride,
                                                       int tilesM) {
	constexpr int TBM = 64 * TM, TBN = 64 * TN;
	extern __shared__ __attribute__((aligned(16))) char smem[];
	char *sA = smem;                              // [2][TBM][LDSROW]
	char *sB = smem + 2 * TBM * LDSROW;           // [2][TBN][LDSROW]
	const LevelBin bin = bins[blockIdx.y];
	const int i0 = (blockIdx.x % tilesM) * TBM, j0 = (blockIdx.x / tilesM) * TBN;
	const int t = threadIdx.x, w = t / 64, l = t % 64;
	const int wm = (w % 2) * 32 * TM, wn = (w / 2) * 32 * TN;
	i16v acc[TM][TN];
#pragma unroll
	for (int a = 0; a < TM; a++)
#pragma unroll
		for (int b = 0; b < TN; b++)
#pragma unroll
			for (int r = 0; r < 16; r++) acc[a][b][r] = 0;
	const int kst = (bin.k1 - bin.k0) / BKB;
	const int total = bin.npairs * kst;
	uint4 ra[TM], rb[TN];
	auto gload = [&](int step) {
		const int p = step / kst, kb = bin.k0 + (step % kst) * BKB;
		const int2 st = pairs[bin.pair0 + p];
		const char *Ap = (const char *)(A + (int64_t)st.x * apstride);
		const char *Bp = (const char *)(B + (int64_t)st.y * bpstride);
#pragma unroll
		for (int h = 0; h < TM; h++) {
			const int c = t + 256 * h, row = c / 4, ch = c % 4;
			const int ia = i0 + row;
			ra[h] = (ia < M) ? *(const uint4 *)(Ap + (int64_t)ia * ldap + kb + 16 * ch)
			                 : make_uint4(0, 0, 0, 0);
		}
#pragma unroll
		for (int h = 0; h < TN; h++) {
			const int c = t + 256 * h, row = c / 4, ch = c % 4;
			const int jb = j0 + row;
			rb[h] = (jb < N) ? *(const uint4 *)(Bp + (int64_t)jb * ldbp + kb + 16 * ch)
			                 : make_uint4(0, 0, 0, 0);
		}
	};
	auto lstore = [&](int buf) {
		char *a = sA + buf * TBM * LDSROW, *b = sB + buf * TBN * LDSROW;
#pragma unroll
		for (int h = 0; h < TM; h++) {
			const int c = t + 256 * h;
			*(uint4 *)(a + (c / 4) * LDSROW + 16 * (c % 4)) = ra[h];
		}
#pragma unroll
		for (int h = 0; h < TN; h++) {
			const int c = t + 256 * h;
			*(uint4 *)(b + (c / 4) * LDSROW + 16 * (c % 4)) = rb[h];
		}
	};
	if (total > 0) {
		gload(0);
		lstore(0);
	}
	__syncthreads();
	for (int step = 0; step < total; step++) {
		const int cur = step & 1;
		if (step + 1 < total) gload(step + 1);
		const char *a = sA + cur * TBM * LDSROW, *b = sB + cur * TBN * LDSROW;
#pragma unroll
		for (int kk = 0; kk < 2; kk++) {
			l2v af[TM], bf[TN];
#pragma unroll
			for (int mi = 0; mi < TM; mi++)
				af[mi] = *(const l2v *)(a + (wm + 32 * mi + l % 32) * LDSROW + 32 * kk + 16 * (l / 32));
#pragma unroll
			for (int nj = 0; nj < TN; nj++)
				bf[nj] = *(const l2v *)(b + (wn + 32 * nj + l % 32) * LDSROW + 32 * kk + 16 * (l / 32));
#pragma unroll
			for (int mi = 0; mi < TM; mi++)
#pragma unroll
				for (int nj = 0; nj < TN; nj++)
					acc[mi][nj] = __builtin_amdgcn_mfma_i32_32x32x32_i8(bf[nj], af[mi], acc[mi][nj], 0, 0, 0);
		}
		if (step + 1 < total) lstore(cur ^ 1);
		__syncthreads();
	}
	// D[p][q]: p = B-side (column j) = 8*(r/4) + 4*(l/32) + r%4, q = A-side (row i) = l%32
	int32_t *Tb = T + (int64_t)blockIdx.y * tstride;
#pragma unroll
	for (int mi = 0; mi < TM; mi++)
#pragma unroll
		for (int nj = 0; nj < TN; nj++)
#pragma unroll
			for (int r = 0; r < 16; r++) {
				const int i = i0 + wm + 32 * mi + l % 32;
				const int j = j0 + wn + 32 * nj + 8 * (r / 4) + 4 * (l / 32) + r % 4;
				if (i < M && j < N) Tb[(int64_t)j * ldt + i] = acc[mi][nj][r];
			}
}

// ---- FP16 / BF16 ---------------------------------------------------------------------------
template <typename V>
__device__ inline f4v mfma16(V a, V b, f4v c);
template <>
__device__ inline f4v mfma16<h8v>(h8v a, h8v b, f4v c) {
	return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}
template <>
__device__ inline f4v mfma16<b8v>(b8v a, b8v b, f4v c) {
	return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// 2-byte digits: a stage is 32 k (64 bytes); a wave's 64x64 region is 4x4 MFMA 16x16 tiles.
template <typename V>
__global__ __launch_bounds__(256) void k_digit_gemm_h(int M, int N, const char *A, int64_t ldab,
                                                      int64_t apstrideb, const char *B,
                                                      int64_t ldbb, int64_t bpstrideb,
                                                      const int2 *pairs, const LevelBin *bins,
                                                      int32_t *T, int64_t ldt, int64_t tstride,
                                                      int tilesM) {
	__shared__ __attribute__((aligned(16))) char sA[BM * LDSROW];
	__shared__ __attribute__((aligned(16))) char sB[BN * LDSROW];
	const LevelBin bin = bins[blockIdx.y];
	const int i0 = (blockIdx.x % tilesM) * BM, j0 = (blockIdx.x / tilesM) * BN;
	const int w = threadIdx.x / 64, l = threadIdx.x % 64;
	const int wm = (w % 2) * 64, wn = (w / 2) * 64;
	int32_t iacc[4][4][4];
	f4v facc[4][4];
	for (int a = 0; a < 4; a++)
		for (int b = 0; b < 4; b++)
			for (int r = 0; r < 4; r++) {
				iacc[a][b][r] = 0;
				facc[a][b][r] = 0.f;
			}
	const int kb0 = 2 * bin.k0, kb1 = 2 * bin.k1;  // byte offsets
	int since = 0;                                  // k accumulated in facc since the last flush
	for (int p = 0; p < bin.npairs; p++) {
		const int2 st = pairs[bin.pair0 + p];
		const char *Ap = A + (int64_t)st.x * apstrideb;
		const char *Bp = B + (int64_t)st.y * bpstrideb;
		Stage stg;
		load_stage(stg, Ap, ldab, Bp, ldbb, i0, j0, M, N, kb0);
		for (int kb = kb0; kb < kb1; kb += BKB) {
			__syncthreads();
			store_stage(stg, sA, sB);
			__syncthreads();
			if (kb + BKB < kb1) load_stage(stg, Ap, ldab, Bp, ldbb, i0, j0, M, N, kb + BKB);
			V af[4], bf[4];
			for (int mi = 0; mi < 4; mi++)
				af[mi] = *(const V *)(sA + (wm + 16 * mi + l % 16) * LDSROW + 16 * (l / 16));
			for (int nj = 0; nj < 4; nj++)
				bf[nj] = *(const V *)(sB + (wn + 16 * nj + l % 16) * LDSROW + 16 * (l / 16));
			for (int mi = 0; mi < 4; mi++)
				for (int nj = 0; nj < 4; nj++) facc[mi][nj] = mfma16<V>(bf[nj], af[mi], facc[mi][nj]);
			since += 32;
			if (since >= 1024) {  // |partial| <= 1024 * 2^14 = 2^24: still exact in fp32
				for (int mi = 0; mi < 4; mi++)
					for (int nj = 0; nj < 4; nj++)
						for (int r = 0; r < 4; r++) {
							iacc[mi][nj][r] += (int32_t)facc[mi][nj][r];
							facc[mi][nj][r] = 0.f;
						}
				since = 0;
			}
		}
	}
	// D[p][q]: p = B side (column j) = 4*(l/16) + r, q = A side (row i) = l%16
	int32_t *Tb = T + (int64_t)blockIdx.y * tstride;
	for (int mi = 0; mi < 4; mi++)
		for (int nj = 0; nj < 4; nj++)
			for (int r = 0; r < 4; r++) {
				const int i = i0 + wm + 16 * mi + l % 16;
				const int j = j0 + wn + 16 * nj + 4 * (l / 16) + r;
				const int32_t v = iacc[mi][nj][r] + (int32_t)facc[mi][nj][r];
				if (i < M && j < N) Tb[(int64_t)j * ldt + i] = v;
			}
}

void launch_digit_gemm(hipStream_t st, int elem, int M, int N, int Kpad, const void *Ap,
                       int64_t ldap, int64_t apstride, const void *Bp, int64_t ldbp,
                       int64_t bpstride, const int2 *pairs, const LevelBin *bins, int nbins,
                       int32_t *T, int64_t ldt, int64_t tstride) {
	(void)Kpad;
	const int tilesM = (M + BM - 1) / BM, tilesN = (N + BN - 1) / BN;
	dim3 grid(tilesM * tilesN, nbins);
	if (elem == 1) {
		if (M >= 512 && N >= 512) {
			constexpr int TM = 4, TN = 4;
			const int tm = (M + 64 * TM - 1) / (64 * TM), tn = (N + 64 * TN - 1) / (64 * TN);
			const size_t lds = 2 * (64 * TM + 64 * TN) * LDSROW;
			static bool attr = false;
			if (!attr) {  // more than 64 KB of dynamic LDS (gfx950 has 160 KB per CU)
				(void)hipFuncSetAttribute((const void *)k_digit_gemm_i8<TM, TN>,
				                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
				attr = true;
			}
			hipLaunchKernelGGL((k_digit_gemm_i8<TM, TN>), dim3(tm * tn, nbins), dim3(256), lds, st, M, N,
			                   (const int8_t *)Ap, ldap, apstride, (const int8_t *)Bp, ldbp, bpstride,
			                   pairs, bins, T, ldt, tstride, tm);
		} else {
			constexpr int TM = 2, TN = 2;
			const int tm = (M + 64 * TM - 1) / (64 * TM), tn = (N + 64 * TN - 1) / (64 * TN);
			const size_t lds = 2 * (64 * TM + 64 * TN) * LDSROW;
			hipLaunchKernelGGL((k_digit_gemm_i8<TM, TN>), dim3(tm * tn, nbins), dim3(256), lds, st, M, N,
			                   (const int8_t *)Ap, ldap, apstride, (const int8_t *)Bp, ldbp, bpstride,
			                   pairs, bins, T, ldt, tstride, tm);
		}
	} else if (elem == 2) {
		hipLaunchKernelGGL(k_digit_gemm_h<h8v>, grid, dim3(256), 0, st, M, N, (const char *)Ap,
		                   2 * ldap, 2 * apstride, (const char *)Bp, 2 * ldbp, 2 * bpstride, pairs,
		                   bins, T, ldt, tstride, tilesM);
	} else {
		hipLaunchKernelGGL(k_digit_gemm_h<b8v>, grid, dim3(256), 0, st, M, N, (const char *)Ap,
		                   2 * ldap, 2 * apstride, (const char *)Bp, 2 * ldbp, 2 * bpstride, pairs,
		                   bins, T, ldt, tstride, tilesM);
	}
}

// ---- exact reconstruction ------------------------------------------------------------------
template <int NW>
struct BigInt {
	uint64_t w[NW];
};

template <int NW>
__device__ inline void big_add_shifted(BigInt<NW> &a, int64_t v, int sh) {
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
__device__ inline uint64_t big_word(const BigInt<NW> &a, int q) {
	uint64_t r = 0;
#pragma unroll
	for (int i = 0; i < NW; i++)
		if (i == q) r = a.w[i];
	return r;
}

// RN-even of N * 2^e0 (N non-negative, NW words)
template <int NW>
__device__ inline double big_round(const BigInt<NW> &a, int e0) {
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

template <int NW>
__global__ __launch_bounds__(256) void k_reconstruct(int M, int N, const int32_t *T, int64_t ldt,
                                                     int64_t tstride, const int *binLevel,
                                                     int nbins, int Lmax, const int *rowE,
                                                     const int *colE, double alpha, double beta,
                                                     double *C, int64_t ldc) {
	const int64_t total = (int64_t)M * N;
	for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
	     e += (int64_t)gridDim.x * blockDim.x) {
		const int i = (int)(e % M), j = (int)(e / M);
		BigInt<NW> a;
#pragma unroll
		for (int q = 0; q < NW; q++) a.w[q] = 0;
		for (int b = 0; b < nbins; b++) {
			const int32_t v = T[(int64_t)b * tstride + (int64_t)j * ldt + i];
			if (v != 0) big_add_shifted<NW>(a, (int64_t)v, 8 * (Lmax - binLevel[b]));
		}
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
		const int e0 = rowE[i] + colE[j] - 16 - 8 * Lmax;
		double t = big_round<NW>(a, e0);
		if (neg) t = -t;
		double *c = C + (int64_t)j * ldc + i;
		*c = fma(alpha, t, beta * *c);
	}
}

void launch_reconstruct(hipStream_t st, int M, int N, const int32_t *T, int64_t ldt,
                        int64_t tstride, const int *binLevel, int nbins, int Lmax, const int *rowE,
                        const int *colE, double alpha, double beta, double *C, int64_t ldc) {
	int64_t g = ((int64_t)M * N + 255) / 256;
	if (g > 16384) g = 16384;
	// bits needed: 8*Lmax + 31 (|T|) + log2(nbins) + sign
	const int bits = 8 * Lmax + 32 + 8 + 1;
	const int nw = (bits + 63) / 64;
	if (nw <= 3)
		hipLaunchKernelGGL(k_reconstruct<3>, dim3((unsigned)g), dim3(256), 0, st, M, N, T, ldt,
		                   tstride, binLevel, nbins, Lmax, rowE, colE, alpha, beta, C, ldc);
	else if (nw <= 4)
		hipLaunchKernelGGL(k_reconstruct<4>, dim3((unsigned)g), dim3(256), 0, st, M, N, T, ldt,
		                   tstride, binLevel, nbins, Lmax, rowE, colE, alpha, beta, C, ldc);
	else if (nw <= 6)
		hipLaunchKernelGGL(k_reconstruct<6>, dim3((unsigned)g), dim3(256), 0, st, M, N, T, ldt,
		                   tstride, binLevel, nbins, Lmax, rowE, colE, alpha, beta, C, ldc);
	else
		hipLaunchKernelGGL(k_reconstruct<9>, dim3((unsigned)g), dim3(256), 0, st, M, N, T, ldt,
		                   tstride, binLevel, nbins, Lmax, rowE, colE, alpha, beta, C, ldc);
}

}  // namespace oz
