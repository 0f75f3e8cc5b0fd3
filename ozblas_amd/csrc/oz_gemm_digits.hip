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

template <int TM, int TN>
__global__ __launch_bounds__(256, 1) void k_digit_gemm_i8(const int8_t *A, int64_t apstride,
                                                          const int8_t *B, int64_t bpstride,
                                                          int64_t ldp, const int2 *pairs,
                                                          const LevelBin *bins, const int *binOrder,
                                                          int32_t *T, int64_t ldt, int64_t tstride,
                                                          int tilesM) {
	constexpr int TBM = 64 * TM, TBN = 64 * TN;
	constexpr int NA = 2 * TM, NB = 2 * TN;  // 16-byte chunks per thread per stage
	// two stage buffers as separate objects, so that alias analysis lets the LDS writes of one
	// stage move past the fragment reads of the other ([A rows | B rows] x I8_BK bytes each)
	__shared__ __attribute__((aligned(16))) char stage0[(TBM + TBN) * I8_BK];
	__shared__ __attribute__((aligned(16))) char stage1[(TBM + TBN) * I8_BK];
	const int binIdx = binOrder[blockIdx.y];
	const LevelBin bin = bins[binIdx];
	const int2 st0 = pairs[bin.pair0];
	const int i0 = (blockIdx.x % tilesM) * TBM, j0 = (blockIdx.x / tilesM) * TBN;
	const int t = threadIdx.x, w = t / 64, l = t % 64;
	const int wm = (w % 2) * 32 * TM, wn = (w / 2) * 32 * TN;
	// Stage loads: thread t moves chunk h (0 <= h < NA) = 16 bytes of row 16 (t/128) + l/4 + 32 h,
	// k bytes 64 ((t/128)%2... = kk64) + 16 (l%4); a wave's instruction reads one contiguous KiB.
	const int kk64 = (t / 64) % 2;
	const int64_t rbs = 16 * ldp;  // bytes per 16-row block of a plane
	const char *Ab = (const char *)A + (int64_t)st0.x * apstride + (int64_t)(bin.k0 / I8_BK) * 2048 +
	                 (int64_t)(i0 / 16 + t / 128) * rbs + kk64 * 1024 + 16 * l;
	const char *Bb = (const char *)B + (int64_t)st0.y * bpstride + (int64_t)(bin.k0 / I8_BK) * 2048 +
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
	const int kst = (bin.k1 - bin.k0) / I8_BK;
	const int total = bin.npairs * kst;
	if (total <= 0) return;
	// next stage to load: pair lpair, k stage lkt (pair p uses plane s0 + p of A and plane t0 - p
	// of B); past the end it wraps to stage (0, 0)
	int lkt = 0, lpair = 0;
	auto advance = [&]() {  // 32-bit selects only: a branch here would split the scheduling region
		const int nk = lkt + 1;
		const int np = lpair + (nk == kst ? 1 : 0);
		lkt = (nk == kst) ? 0 : nk;
		lpair = (np == bin.npairs) ? 0 : np;
	};
	// two register sets for the global stage loads (ext-vector type: HIP's uint4 struct would
	// keep the arrays out of registers)
	i4v ra0[NA], rb0[NB], ra1[NA], rb1[NB];
	auto gload = [&](i4v(&ra)[NA], i4v(&rb)[NB]) {
		const char *Ap = Ab + (int64_t)lpair * apstride + lkt * 2048;
		const char *Bp = Bb - (int64_t)lpair * bpstride + lkt * 2048;
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
		const char *Bp = Bb - (int64_t)lpair * bpstride + lkt * 2048;
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
	// prologue: stage 0 -> buffer 0, stage 1 -> registers set 1
	gload(ra0, rb0);
	gload(ra1, rb1);
	lstore(ra0, rb0, 0);
	lds_barrier();
	fread(fa0, fb0, 0, 0);
	int s = 0;
	for (; s + 1 < total; s += 2) {
		step(ra0, rb0, ra1, rb1, 0);
		step(ra1, rb1, ra0, rb0, 1);
	}
	if (s < total) step(ra0, rb0, ra1, rb1, 0);
	// D[p][q]: p = B-side (column j) = 8*(r/4) + 4*(l/32) + r%4, q = A-side (row i) = l%32
	int32_t *Tb = T + (int64_t)binIdx * tstride;
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
                                                      const int *binOrder, int32_t *T, int64_t ldt,
                                                      int64_t tstride, int tilesM) {
	__shared__ __attribute__((aligned(16))) char sA[BM * LDSROW];
	__shared__ __attribute__((aligned(16))) char sB[BN * LDSROW];
	const int binIdx = binOrder[blockIdx.y];
	const LevelBin bin = bins[binIdx];
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

void launch_digit_gemm(hipStream_t st, int elem, int M, int N, int Kpad, const void *Ap,
                       int64_t ldap, int64_t apstride, const void *Bp, int64_t ldbp,
                       int64_t bpstride, const int2 *pairs, const LevelBin *bins, int nbins,
                       const int *binOrder, int32_t *T, int64_t ldt, int64_t tstride) {
	(void)Kpad;
	const int tilesM = (M + BM - 1) / BM, tilesN = (N + BN - 1) / BN;
	dim3 grid(tilesM * tilesN, nbins);
	if (elem == 1) {
		(void)ldbp;  // both operands share the padded k extent ldap
		if (M >= 512 && N >= 512) {
			constexpr int TM = 4, TN = 4;
			const int tm = (M + 64 * TM - 1) / (64 * TM), tn = (N + 64 * TN - 1) / (64 * TN);
			hipLaunchKernelGGL((k_digit_gemm_i8<TM, TN>), dim3(tm * tn, nbins), dim3(256), 0, st,
			                   (const int8_t *)Ap, apstride, (const int8_t *)Bp, bpstride, ldap,
			                   pairs, bins, binOrder, T, ldt, tstride, tm);
		} else {
			constexpr int TM = 2, TN = 2;
			const int tm = (M + 64 * TM - 1) / (64 * TM), tn = (N + 64 * TN - 1) / (64 * TN);
			hipLaunchKernelGGL((k_digit_gemm_i8<TM, TN>), dim3(tm * tn, nbins), dim3(256), 0, st,
			                   (const int8_t *)Ap, apstride, (const int8_t *)Bp, bpstride, ldap,
			                   pairs, bins, binOrder, T, ldt, tstride, tm);
		}
	} else if (elem == 2) {
		hipLaunchKernelGGL(k_digit_gemm_h<h8v>, grid, dim3(256), 0, st, M, N, (const char *)Ap,
		                   2 * ldap, 2 * apstride, (const char *)Bp, 2 * ldbp, 2 * bpstride, pairs,
		                   bins, binOrder, T, ldt, tstride, tilesM);
	} else {
		hipLaunchKernelGGL(k_digit_gemm_h<b8v>, grid, dim3(256), 0, st, M, N, (const char *)Ap,
		                   2 * ldap, 2 * apstride, (const char *)Bp, 2 * ldbp, 2 * bpstride, pairs,
		                   bins, binOrder, T, ldt, tstride, tilesM);
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
