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
#include "oz_digits.hpp"

namespace oz {

typedef float f4v __attribute__((ext_vector_type(4)));
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
// INT8 planes use a tiled layout: row r, k-byte q of a plane sits at
//   (r / 16) * 16 * ldp + (q / 64) * 1024 + (r % 16) * 64 + q % 64,
// so the 16 rows x 64 bytes a wave loads with one instruction are one contiguous KiB.  Rows are
// padded to a multiple of 256 and k to a multiple of 128 with zero digits: the kernel has no
// bounds checks.
//
// Workgroup = 2 x 2 waves; wave tile = TM x TN MFMA tiles of 32 x 32 (block tile 64TM x 64TN).
// A stage is 128 k-bytes (4 MFMA k-steps); two LDS stages.  Per stage: the next stage's global
// loads are issued first, the fragments of k-step kk+1 are read while the MFMAs of kk run, the
// next stage is written to LDS after k-step 2, one barrier follows, and k-step 3 (already in
// registers) overlaps the read of the next stage's first fragments.  The (pair, k) loop of a bin
// is flattened, so the pipeline does not drain between pairs.  Pairs of a bin are consecutive:
// pair p uses plane s0 + p of A and plane t0 - p of B.
// The MFMA is issued as B-fragment x A-fragment so that lane l%32 indexes the A row: stores of the
// column-major result planes are contiguous across lanes.
template <int TM, int TN, bool RES>
__global__ __launch_bounds__(256, 1) void k_digit_gemm_i8(const int8_t *A, int64_t apstride,
                                                          const int8_t *B, int64_t bpstride,
                                                          int64_t ldp, const int2 *pairs,
                                                          const LevelBin *bins, const int *binOrder,
                                                          int32_t *T, int64_t ldt, int64_t tstride,
                                                          int tilesM, const int *tileList,
                                                          const ResConst *rc, uint8_t *R) {
	constexpr int TBM = 64 * TM, TBN = 64 * TN;
	__shared__ __attribute__((aligned(16))) char stage0[(TBM + TBN) * I8_BK];
	__shared__ __attribute__((aligned(16))) char stage1[(TBM + TBN) * I8_BK];
	const int binIdx = binOrder[blockIdx.y];
	const LevelBin bin = bins[binIdx];
	const int2 st0 = pairs[bin.pair0];
	I8Tile g;
	g.A = A;
	g.B = B;
	g.apstride = apstride;
	g.bpstride = bpstride;
	g.ldp = ldp;
	g.s0 = st0.x;
	g.t0 = st0.y;
	g.npairs = bin.npairs;
	g.k0 = bin.k0;
	g.k1 = bin.k1;
	const int tile = tileList ? tileList[blockIdx.x] : (int)blockIdx.x;  // a tile subset: fallback pass
	g.i0 = (tile % tilesM) * TBM;
	g.j0 = (tile / tilesM) * TBN;
	g.T = T + (int64_t)binIdx * tstride;
	g.ldt = ldt;
	g.R = nullptr;
	if constexpr (RES) {  // residue form: bin = moduli level .. level + npairs - 1 over one k chunk
		g.R = R + (int64_t)(bin.k0 / kResChunk) * rc->nmod * tstride;
		g.rstride = tstride;
		g.rc = rc;
	}
	i8_tile<TM, TN, RES>(g, stage0, stage1);
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
template <typename V, bool RES>
__global__ __launch_bounds__(256) void k_digit_gemm_h(int M, int N, const char *A, int64_t ldab,
                                                      int64_t apstrideb, const char *B,
                                                      int64_t ldbb, int64_t bpstrideb,
                                                      const int2 *pairs, const LevelBin *bins,
                                                      const int *binOrder, int32_t *T, int64_t ldt,
                                                      int64_t tstride, int tilesM, const int *tileList,
                                                      const ResConst *rc, uint8_t *R) {
	__shared__ __attribute__((aligned(16))) char sA[2][BM * LDSROW];  // two stage buffers
	__shared__ __attribute__((aligned(16))) char sB[2][BN * LDSROW];
	const int binIdx = binOrder[blockIdx.y];
	const LevelBin bin = bins[binIdx];
	const int tile = tileList ? tileList[blockIdx.x] : (int)blockIdx.x;  // a tile subset: fallback pass
	const int i0 = (tile % tilesM) * BM, j0 = (tile / tilesM) * BN;
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
		// stage kb in buffer cur while the next one is stored into the other: one barrier a stage
		Stage stg;
		load_stage(stg, Ap, ldab, Bp, ldbb, i0, j0, M, N, kb0);
		store_stage(stg, sA[0], sB[0]);
		if (kb0 + BKB < kb1) load_stage(stg, Ap, ldab, Bp, ldbb, i0, j0, M, N, kb0 + BKB);
		__syncthreads();
		int cur = 0;
		for (int kb = kb0; kb < kb1; kb += BKB) {
			V af[4], bf[4];
			for (int mi = 0; mi < 4; mi++)
				af[mi] = *(const V *)(sA[cur] + (wm + 16 * mi + l % 16) * LDSROW + 16 * (l / 16));
			for (int nj = 0; nj < 4; nj++)
				bf[nj] = *(const V *)(sB[cur] + (wn + 16 * nj + l % 16) * LDSROW + 16 * (l / 16));
			if (kb + BKB < kb1) {
				store_stage(stg, sA[cur ^ 1], sB[cur ^ 1]);
				if (kb + 2 * BKB < kb1) load_stage(stg, Ap, ldab, Bp, ldbb, i0, j0, M, N, kb + 2 * BKB);
			}
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
			__syncthreads();
			cur ^= 1;
		}
	}
	// D[p][q]: p = B side (column j) = 4*(l/16) + r, q = A side (row i) = l%16
	if constexpr (RES) {  // residue form: bin = (modulus bin.level, k chunk), acc mod p, 4 columns a word
		uint8_t *Rb = R + (int64_t)binIdx * tstride;
		const double p = (double)rc->p[bin.level];
		const double invp = rc->invpd[bin.level];
		for (int mi = 0; mi < 4; mi++)
			for (int nj = 0; nj < 4; nj++) {
				const int i = i0 + wm + 16 * mi + l % 16;
				const int j = j0 + wn + 16 * nj + 4 * (l / 16);
				uint32_t w = 0;
				for (int r = 0; r < 4; r++) {
					const int32_t v = iacc[mi][nj][r] + (int32_t)facc[mi][nj][r];
					w |= (uint32_t)mod_f64(v, p, invp) << (8 * r);
				}
				if (i < M && j < N) *(uint32_t *)(Rb + res_word_offset(i, j, ldt)) = w;
			}
		return;
	}
	int32_t *Tb = T + (int64_t)binIdx * tstride;
	for (int mi = 0; mi < 4; mi++)
		for (int nj = 0; nj < 4; nj++)
			for (int r = 0; r < 4; r++) {
				const int i = i0 + wm + 16 * mi + l % 16;
				const int j = j0 + wn + 16 * nj + 4 * (l / 16) + r;
				const int32_t v = iacc[mi][nj][r] + (int32_t)facc[mi][nj][r];
				if (i < M && j < N) Tb[(int64_t)j * ldt + i] = v;
			}
}

int digit_gemm_tile(int elem, int M, int N) {
	if (elem != 1) return BM;
	return (M >= 512 && N >= 512) ? 256 : 128;
}

void launch_digit_gemm(hipStream_t st, int elem, int M, int N, int Kpad, const void *Ap,
                       int64_t ldap, int64_t apstride, const void *Bp, int64_t ldbp,
                       int64_t bpstride, const int2 *pairs, const LevelBin *bins, int nbins,
                       const int *binOrder, int32_t *T, int64_t ldt, int64_t tstride,
                       const int *tileList, int ntiles, const ResConst *rc, uint8_t *R) {
	(void)Kpad;
	const int tilesM = (M + BM - 1) / BM, tilesN = (N + BN - 1) / BN;
	if (elem == 1) {
		(void)ldbp;  // both operands share the padded k extent ldap
#define OZ_LAUNCH_I8(TMV, RESV)                                                                           \
	do {                                                                                              \
		const int tm = (M + 64 * TMV - 1) / (64 * TMV), tn = (N + 64 * TMV - 1) / (64 * TMV);          \
		hipLaunchKernelGGL((k_digit_gemm_i8<TMV, TMV, RESV>), dim3(tileList ? ntiles : tm * tn, nbins), \
		                   dim3(256), 0, st, (const int8_t *)Ap, apstride, (const int8_t *)Bp, bpstride, \
		                   ldap, pairs, bins, binOrder, T, ldt, tstride, tm, tileList, rc, R);         \
	} while (0)
		if (M >= 512 && N >= 512) {
			if (R) OZ_LAUNCH_I8(4, true);
			else OZ_LAUNCH_I8(4, false);
		} else {
			if (R) OZ_LAUNCH_I8(2, true);
			else OZ_LAUNCH_I8(2, false);
		}
#undef OZ_LAUNCH_I8
	} else {
		const dim3 hgrid(tileList ? ntiles : tilesM * tilesN, nbins);
#define OZ_LAUNCH_H(V, RESV)                                                                       \
	hipLaunchKernelGGL((k_digit_gemm_h<V, RESV>), hgrid, dim3(256), 0, st, M, N, (const char *)Ap, \
	                   2 * ldap, 2 * apstride, (const char *)Bp, 2 * ldbp, 2 * bpstride, pairs, bins,  \
	                   binOrder, T, ldt, tstride, tilesM, tileList, rc, R)
		if (elem == 2) {
			if (R) OZ_LAUNCH_H(h8v, true);
			else OZ_LAUNCH_H(h8v, false);
		} else {
			if (R) OZ_LAUNCH_H(b8v, true);
			else OZ_LAUNCH_H(b8v, false);
		}
#undef OZ_LAUNCH_H
	}
}

// ---- exact reconstruction ------------------------------------------------------------------
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
		double *c = C + (int64_t)j * ldc + i;
		*c = recon_elem<NW>(T + (int64_t)j * ldt + i, tstride, binLevel, nbins, Lmax,
		                    rowE[i] + colE[j] - 16 - 8 * Lmax, alpha, beta, *c);
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

// Certified early termination (pass 1): N1 over the computed bins [0, nb1); where the rounding
// of N1 +- R agrees the element is final, otherwise its tile is marked for the remaining levels
// and C is left untouched.
template <int NW>
__global__ __launch_bounds__(256) void k_reconstruct_check(int M, int N, const int32_t *T, int64_t ldt,
                                                           int64_t tstride, const int *binLevel,
                                                           int nb1, int Lmax, const int *rowE,
                                                           const int *colE, double alpha, double beta,
                                                           double *C, int64_t ldc, const uint64_t *R,
                                                           int tile, int tilesM, int *tileMark) {
	uint64_t r[NW];
	double rUp = 0.0;  // an upper bound on R as a binary64
#pragma unroll
	for (int q = 0; q < NW; q++) {
		r[q] = R[q];
		rUp += ldexp((double)r[q], 64 * q);
	}
	rUp *= 1.0 + 0x1p-40;
	const int64_t total = (int64_t)M * N;
	for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
	     e += (int64_t)gridDim.x * blockDim.x) {
		const int i = (int)(e % M), j = (int)(e / M);
		const int32_t *Te = T + (int64_t)j * ldt + i;
		const int e0 = rowE[i] + colE[j] - 16 - 8 * Lmax;
		double t;
		if (recon_certify_fast(Te, tstride, binLevel, nb1, Lmax, rUp, e0, &t)) {
			double *c = C + (int64_t)j * ldc + i;
			*c = fma(alpha, t, beta * *c);
			continue;
		}
		BigInt<NW> a;
		big_zero<NW>(a);
		recon_accumulate<NW>(a, Te, tstride, binLevel, 0, nb1, Lmax);
		if (recon_certify<NW>(a, r, e0, &t)) {
			double *c = C + (int64_t)j * ldc + i;
			*c = fma(alpha, t, beta * *c);
		} else {
			tileMark[(i / tile) + (j / tile) * tilesM] = 1;
		}
	}
}

// Pass 2: every element of the listed tiles whose pass-1 check failed gets the exact sum over all
// the bins [0, nbins) (the remaining ones were computed for these tiles only).
template <int NW>
__global__ __launch_bounds__(256) void k_reconstruct_fix(int M, int N, const int32_t *T, int64_t ldt,
                                                         int64_t tstride, const int *binLevel,
                                                         int nb1, int nbins, int Lmax,
                                                         const int *rowE, const int *colE,
                                                         double alpha, double beta, double *C,
                                                         int64_t ldc, const uint64_t *R, int tile,
                                                         int tilesM, const int *tileList) {
	const int perTile = tile * tile / 256;  // workgroups per tile
	const int tl = tileList[blockIdx.x / perTile];
	const int local = (blockIdx.x % perTile) * 256 + threadIdx.x;
	const int i = (tl % tilesM) * tile + local % tile, j = (tl / tilesM) * tile + local / tile;
	if (i >= M || j >= N) return;
	uint64_t r[NW];
#pragma unroll
	for (int q = 0; q < NW; q++) r[q] = R[q];
	const int32_t *Te = T + (int64_t)j * ldt + i;
	const int e0 = rowE[i] + colE[j] - 16 - 8 * Lmax;
	BigInt<NW> a;
	big_zero<NW>(a);
	recon_accumulate<NW>(a, Te, tstride, binLevel, 0, nb1, Lmax);
	double t;
	if (recon_certify<NW>(a, r, e0, &t)) return;  // final since pass 1
	recon_accumulate<NW>(a, Te, tstride, binLevel, nb1, nbins, Lmax);
	double *c = C + (int64_t)j * ldc + i;
	*c = fma(alpha, big_finish<NW>(a, e0), beta * *c);
}

int recon_words(int Lmax) { return (8 * Lmax + 32 + 8 + 1 + 63) / 64; }

#define OZ_RECON_NW(NWV, ...)                                     \
	do {                                                          \
		const int nw_ = (NWV);                                    \
		if (nw_ <= 3) { constexpr int NW = 3; __VA_ARGS__; }      \
		else if (nw_ <= 4) { constexpr int NW = 4; __VA_ARGS__; } \
		else if (nw_ <= 6) { constexpr int NW = 6; __VA_ARGS__; } \
		else { constexpr int NW = 9; __VA_ARGS__; }               \
	} while (0)

void launch_reconstruct_check(hipStream_t st, int M, int N, const int32_t *T, int64_t ldt,
                              int64_t tstride, const int *binLevel, int nb1, int Lmax,
                              const int *rowE, const int *colE, double alpha, double beta, double *C,
                              int64_t ldc, const uint64_t *R, int tile, int tilesM, int *tileMark) {
	int64_t g = ((int64_t)M * N + 255) / 256;
	if (g > 16384) g = 16384;
	OZ_RECON_NW(recon_words(Lmax),
	            hipLaunchKernelGGL(k_reconstruct_check<NW>, dim3((unsigned)g), dim3(256), 0, st, M, N, T,
	                               ldt, tstride, binLevel, nb1, Lmax, rowE, colE, alpha, beta, C, ldc, R,
	                               tile, tilesM, tileMark));
}

void launch_reconstruct_fix(hipStream_t st, int M, int N, const int32_t *T, int64_t ldt,
                            int64_t tstride, const int *binLevel, int nb1, int nbins, int Lmax,
                            const int *rowE, const int *colE, double alpha, double beta, double *C,
                            int64_t ldc, const uint64_t *R, int tile, int tilesM, const int *tileList,
                            int ntiles) {
	const int perTile = tile * tile / 256;
	OZ_RECON_NW(recon_words(Lmax),
	            hipLaunchKernelGGL(k_reconstruct_fix<NW>, dim3((unsigned)(ntiles * perTile)), dim3(256), 0,
	                               st, M, N, T, ldt, tstride, binLevel, nb1, nbins, Lmax, rowE, colE,
	                               alpha, beta, C, ldc, R, tile, tilesM, tileList));
}

}  // namespace oz
