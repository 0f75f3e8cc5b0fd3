/*
 * ozoracle_sd.c -- CPU oracle, mixed binary32/binary64 (TEST INFRASTRUCTURE ONLY; see ozoracle.h).
 *
 * Restatement of OzBLAS 1.6a src/cpu for the <float,double,double> / <float,double> instantiations
 * (src/cpu/ozblas_gemm.cpp:321, ozblas_dot.cpp:45, ozblas_nrm2.cpp:37, ozblas_gemv.cpp:35,
 * ozblas_csrmv.cpp:187,277, ozblas_cg.cpp:134, ozblas_split.cpp:295,934): binary32 data split in
 * binary32 arithmetic into slices scaled by 2^-tau and stored as doubles, with the exponents tau
 * kept as shorts.
 *  - rho is getRho<float,double>: epse of TYPE1 = 24, of TYPE2 = 53, epse2 of TYPE2 = 106
 *    (src/cpu/ozblas_split.cpp:20-34, src/cpu/ozblas_common.h:93-113);
 *  - tau = ceil(log21<float>(max)) with log21<float>(v) = (float)log2((double)v), sigma the float
 *    CONST * scalbn1<float>(1, rho + tau) * splitShift, the split (x + sigma) - sigma in float, and
 *    the stored slice scalbn1<float>(split, -tau) = (float)scalbn((double)split, -tau), widened
 *    (src/cpu/ozblas_split.cpp:106-135,182-222,871-894);
 *  - no remainder slot and no reproMode adjustment of the split loop (mixed types,
 *    src/cpu/ozblas_split.cpp:275-287,916-925);
 *  - slice products are binary64 GEMM / the csrmm_n<double> fma chain of the scaled slices;
 *  - the sum is the mixed Fsum in float: c = (float)Cs, t += scalbn1<float>(c, seA + seB)
 *    (src/cpu/ozblas_sum.cpp:383-432); C = fma1<float>(alpha, t, beta * C);
 *  - the mixed NearSum exits in the reference (:275-330), so sumMode = 1 returns -3 here.  Dot2
 *    (splitEpsMode = 2) is not restated (returns -3): its rho is negative for every dim below
 *    2^58.
 * The binary64 slice GEMM is exact whenever rho >= 2 (k > 128 in splitEpsMode 0): every scaled
 * slice is then a multiple of 2^(rho-24) of magnitude at most 1, so the k products are integers
 * times 2^(2 rho - 48) whose sum stays below 2^53 units.  Below that the reference's dgemm rounds
 * in its BLAS's order, which no restatement can fix; gemm_tn_sd sums in index order.
 */
#define _GNU_SOURCE
#include "ozoracle_int.h"

static inline float std_maxf(float a, float b) { return (a < b) ? b : a; } /* std::max<float> */

int32_t ozo_sd_rho(int32_t dim, int32_t splitEpsMode) {
	/* getRho<float,double> (src/cpu/ozblas_split.cpp:20-34) */
	double r;
	switch (splitEpsMode) {
	case 2: r = ceil((24 - (106 - log2(1. * dim)) / 2)); break;
	case 9: r = ceil((24 - (53 - log2(2. * sqrt(dim))) / 2)); break;
	default: r = ceil((24 - (53 - log2(1. * dim)) / 2)); break;
	}
	return d2i(r);
}

/* CONST * scalbn1<float>(one, rho + tau) * splitShift (:116,200); the sparse split divides */
static float sigma_sd(int32_t rho, int32_t tau, int32_t shift, int div) {
	const float p = (float)scalbn_l(1.0, (int64_t)rho + tau);
	return div ? (float)(SPLIT_CONST * (double)p / shift) : (float)(SPLIT_CONST * (double)p * shift);
}

/* scalbn1<float>(split, -tau) stored into a double slice (:124,210,884) */
static inline double scaled_sd(float sp, int32_t tau) { return (double)(float)scalbn_l((double)sp, -(int64_t)tau); }

/* ozblasFindMaxDevice :55-73 (binary32 input) */
static void find_max_sd(int32_t m, int32_t n, const float *in, int32_t ldi, float *max) {
	if (n == 1) { /* cblas_isamax: first index of the largest |x| */
		int32_t idx = 0;
		float best = m > 0 ? fabsf(in[0]) : 0.f;
		for (int32_t i = 1; i < m; i++)
			if (fabsf(in[i]) > best) {
				best = fabsf(in[i]);
				idx = i;
			}
		max[0] = fabsf(in[idx]);
	} else {
		for (int32_t j = 0; j < n; j++) {
			float mx = 0.f;
			for (int32_t i = 0; i < m; i++) mx = std_maxf(mx, fabsf(in[(int64_t)j * ldi + i]));
			max[j] = mx;
		}
	}
}

/* ozblasSplitVecKernel<float,double> :106-135 and ozblasSplitCKernel<float,double> :182-222 */
static float split_c_sd(int32_t m, int32_t n, int32_t rho, const float *in, int32_t ldi, float *out,
                        int32_t ldo, double *split, int32_t lds, int16_t *exps, float *max,
                        int32_t shift) {
	float check = 0.f;
	for (int32_t j = 0; j < n; j++) {
		const int32_t tau = exps[j] = (int16_t)ozo_s_tau(max[j]);
		const float sigma = sigma_sd(rho, tau, shift, 0);
		float mx = 0.f;
		for (int32_t i = 0; i < m; i++) {
			float input = in[(int64_t)j * ldi + i];
			const float sp = (input + sigma) - sigma;
			input = input - sp;
			split[(int64_t)j * lds + i] = scaled_sd(sp, tau);
			out[(int64_t)j * ldo + i] = input;
			mx = std_maxf(mx, fabsf(input));
		}
		max[j] = mx;
		check += mx; /* only compared with 0 */
	}
	return check;
}

/* ozblasSplit<float,double> :254-289: no reproMode adjustment for mixed types */
int32_t ozo_sd_split(ozo_handle *h, int32_t m, int32_t n, const float *in, int32_t ldi, float *out,
                     int32_t ldo, double *split, int32_t lds, int16_t *exps, int32_t ldse, float *max) {
	int32_t s;
	const int32_t rho = ozo_sd_rho(m, h->splitEpsMode);
	find_max_sd(m, n, in, ldi, max);
	split_c_sd(m, n, rho, in, ldi, out, ldo, split, lds, exps, max, h->splitShift);
	const int32_t maxS = (h->nSplitMax > 0) ? h->nSplitMax : NUM_SPLIT_DEFAULT_MAX;
	for (s = 1; s < maxS; s++) {
		if (split_c_sd(m, n, rho, out, ldo, out, ldo, &split[(int64_t)lds * n * s], lds,
		               &exps[(int64_t)ldse * s], max, h->splitShift) == 0.f) {
			s++;
			break;
		}
	}
	return s;
}

/* sparse: ozblasSplitSparseNKernel<float,double> :871-894, ozblasSplitSparse :896-928 */
static float split_sparse_sd_kernel(int32_t m, const float *in, const int32_t *rowptr, float *out,
                                    double *split, int16_t *exps, float *max, int32_t epsMode,
                                    int32_t shift) {
	float check = 0.f;
	for (int32_t r = 0; r < m; r++) {
		const int32_t rho = ozo_sd_rho(rowptr[r + 1] - rowptr[r], epsMode);
		const int32_t tau = exps[r] = (int16_t)ozo_s_tau(max[r]);
		const float sigma = sigma_sd(rho, tau, shift, 1);
		float mx = 0.f;
		for (int32_t i = rowptr[r]; i < rowptr[r + 1]; i++) {
			float input = in[i];
			const float sp = (input + sigma) - sigma;
			input = input - sp;
			split[i] = scaled_sd(sp, tau);
			out[i] = input;
			mx = std_maxf(mx, fabsf(input));
		}
		max[r] = mx;
		check += mx;
	}
	return check;
}

static int32_t split_sparse_sd(ozo_handle *h, int32_t m, const float *in, const int32_t *rowptr,
                               float *out, double *split, int32_t lds, int16_t *exps, int32_t ldse,
                               float *max) {
	int32_t s;
	for (int32_t r = 0; r < m; r++) { /* ozblasFindMaxSparseDevice */
		float mx = 0.f;
		for (int32_t i = rowptr[r]; i < rowptr[r + 1]; i++) mx = std_maxf(mx, fabsf(in[i]));
		max[r] = mx;
	}
	split_sparse_sd_kernel(m, in, rowptr, out, split, exps, max, h->splitEpsMode, h->splitShift);
	const int32_t maxS = (h->nSplitMax > 0) ? h->nSplitMax : NUM_SPLIT_DEFAULT_MAX;
	for (s = 1; s < maxS; s++)
		if (split_sparse_sd_kernel(m, out, rowptr, out, &split[(int64_t)lds * s], &exps[(int64_t)ldse * s],
		                           max, h->splitEpsMode, h->splitShift) == 0.f) {
			s++;
			return s;
		}
	return s;
}

/* mixed Fsum (ozblasGlobalFsumKernel<float,double> :383-432) */
static void global_fsum_sd(int32_t m, int32_t n, const int16_t *expA, int32_t ldase, int32_t nA,
                           const int16_t *expB, int32_t ldbse, int32_t nB, const double *Cs, int64_t llsc,
                           int32_t ldsc, float *C, int32_t ldc, float alpha, float beta,
                           int32_t maxlevel, int32_t sumOrder) {
	for (int32_t y = 0; y < n; y++)
		for (int32_t x = 0; x < m; x++) {
			float t = 0.f;
			int32_t ic = 0;
			for (int32_t ik = 0; ik <= maxlevel; ik++)
				for (int32_t ia = 0; ia < nA; ia++) {
					const int32_t seA = expA[(int64_t)ldase * ia + x];
					for (int32_t ib = 0; ib < nB; ib++)
						if (ik == ia + ib) {
							int32_t it;
							switch (sumOrder) {
							case 2: it = nA * ib + ia; break;
							case 3: it = nB * ia + ib; break;
							default: it = ic; break;
							}
							const float c = (float)Cs[llsc * it + (int64_t)y * ldsc + x];
							const int32_t seB = expB[(int64_t)ldbse * ib + y];
							t += (float)scalbn((double)c, seA + seB);
							ic++;
						}
				}
			float *cp = &C[(int64_t)y * ldc + x];
			*cp = (float)fma((double)alpha, (double)t, (double)(beta * *cp));
		}
}

/* binary64 GEMM of the scaled slices (cblas_dgemm, src/cpu/ozblas_blas_wrapper.cpp:158-167), summed
   in index order: exact, and so equal to any order, whenever rho >= 2 (see the header) */
static void gemm_tn_sd(int32_t M, int32_t N, int32_t K, const double *A, int32_t lda, const double *B,
                       int32_t ldb, double *C, int32_t ldc) {
#pragma omp parallel for schedule(static)
	for (int32_t j = 0; j < N; j++)
		for (int32_t i = 0; i < M; i++) {
			const double *a = A + (int64_t)i * lda;
			const double *b = B + (int64_t)j * ldb;
			double s = 0.;
			for (int32_t kk = 0; kk < K; kk++) s += a[kk] * b[kk];
			C[(int64_t)j * ldc + i] = s;
		}
}

static void omatcopy_t_sd(int32_t rows, int32_t cols, const float *src, int32_t lds, float *dst,
                          int32_t ldd) {
	for (int32_t j = 0; j < cols; j++)
		for (int32_t i = 0; i < rows; i++) dst[(int64_t)i * ldd + j] = src[(int64_t)j * lds + i];
}

/* csrmm_n<double> (src/cpu/ozblas_blas_wrapper.cpp:231-260): fma1<double> chain per row */
static void csrmm_n_sd(int32_t m, int32_t n, const double *val, const int32_t *colind,
                       const int32_t *rowptr, const double *B, int32_t ldb, double *C, int32_t ldc) {
	for (int32_t r = 0; r < m; r++)
		for (int32_t j = 0; j < n; j++) {
			double tr = 0.;
			for (int32_t i = rowptr[r]; i < rowptr[r + 1]; i++)
				tr = fma(val[i], B[(int64_t)j * ldb + colind[i]], tr);
			C[(int64_t)j * ldc + r] = 1. * tr; /* alpha = 1, beta = 0 */
		}
}

int32_t ozo_sd_gemm(ozo_handle *h, char transA, char transB, int32_t m, int32_t n, int32_t k,
                    float alpha, const float *A, int32_t lda, const float *B, int32_t ldb, float beta,
                    float *C, int32_t ldc) {
	/* src/cpu/ozblas_gemm.cpp:3-315 with TYPE1 = float, TYPE2 = TYPE3 = double */
	if (h->reproMode == 0 && h->nSplitMax == 1) return -3; /* plain BLAS bypass :14-17 */
	if (h->splitEpsMode == 2) return -3;                   /* Dot2 not restated */
	if (h->sumMode != 0) return -3;                        /* the mixed NearSum exits */
	counter_init(h);
	const int32_t saveBatched = h->useBatchedGemmFlag, saveFast = h->fastMode;
	if (m == 1 && n == 1) {
		h->useBatchedGemmFlag = 0;
		h->fastMode = 0;
	}
	int32_t mbk = m, nbk = n;
	int32_t L = (h->nSplitMax > 1) ? h->nSplitMax + 1 : NUM_SPLIT_DEFAULT_MAX;
	if (h->reproMode == 1) L++;
	if (h->useBatchedGemmFlag) {
		vec_alloc(h, L * L, 8);
		vec_alloc(h, L * L, 8);
		vec_alloc(h, L * L, 8);
	}
	int32_t ldas = 0, ldbs = 0, ldcs = 0, ldt = 0, ldct = 0, ldase = 0, ldbse = 0;
	uint64_t oAS = 0, oBS = 0, oCS = 0, oT = 0, oMax = 0, oAE = 0, oBE = 0;
	const uint64_t memAddrTmp = h->memAddr;
	while (mbk > 0 && nbk > 0) { /* fit loop :59-97 */
		const int32_t sizeCn = nbk * L * L; /* globalSum (sumMode 0) */
		oAS = mat_alloc(h, k, mbk * L, 8, &ldas);
		oBS = mat_alloc(h, k, nbk * L, 8, &ldbs);
		oCS = mat_alloc(h, mbk, sizeCn, 8, &ldcs);
		oT = mat_alloc(h, k, imax(mbk, nbk), 4, &ldt);
		mat_alloc(h, mbk, nbk, 4, &ldct);
		oAE = mat_alloc(h, mbk, L, 2, &ldase);
		oBE = mat_alloc(h, nbk, L, 2, &ldbse);
		oMax = vec_alloc(h, imax(k, imax(mbk, nbk)), 4);
		ldt = ozo_pitch(k);
		if (!mem_check(h)) break;
		h->memAddr = memAddrTmp;
		if (mbk == 1 && nbk == 1 && !mem_check(h)) {
			fprintf(stderr, "ozoracle: out of memory\n");
			return -1;
		}
		if (mbk != 1 || nbk != 1) {
			if (mbk >= nbk)
				mbk = (int32_t)ceil(mbk / 2.);
			else
				nbk = (int32_t)ceil(nbk / 2.);
		}
	}
	arena_check(h);
	h->mbk = mbk;
	h->nbk = nbk;
	double *AS = WP(h, oAS, double), *BS = WP(h, oBS, double), *CS = WP(h, oCS, double);
	float *Tmp = WP(h, oT, float), *Max1 = WP(h, oMax, float);
	int16_t *AE = WP(h, oAE, int16_t), *BE = WP(h, oBE, int16_t);
	(void)ldct;

	int32_t block_count = 0;
	const int32_t nbm = (int32_t)ceilf((float)m / mbk), nbn = (int32_t)ceilf((float)n / nbk);
	for (int32_t im = 0; im < nbm; im++) {
		const int32_t mbk_ = (m - mbk * im >= mbk) ? mbk : m - mbk * im;
		int32_t nSplitA;
		if (is_n(transA)) {
			omatcopy_t_sd(mbk_, k, A + (int64_t)im * mbk, lda, Tmp, ldt);
			nSplitA = ozo_sd_split(h, k, mbk_, Tmp, ldt, Tmp, ldt, AS, ldas, AE, ldase, Max1);
		} else {
			nSplitA = ozo_sd_split(h, k, mbk_, A + (int64_t)im * mbk * lda, lda, Tmp, ldt, AS, ldas, AE,
			                       ldase, Max1);
		}
		for (int32_t in = 0; in < nbn; in++) {
			const int32_t nbk_ = (n - nbk * in >= nbk) ? nbk : n - nbk * in;
			int32_t nSplitB;
			if (is_n(transB)) {
				nSplitB = ozo_sd_split(h, k, nbk_, B + (int64_t)in * nbk * ldb, ldb, Tmp, ldt, BS, ldbs, BE,
				                       ldbse, Max1);
			} else {
				omatcopy_t_sd(nbk_, k, B + (int64_t)in * nbk, ldb, Tmp, ldt);
				nSplitB = ozo_sd_split(h, k, nbk_, Tmp, ldt, Tmp, ldt, BS, ldbs, BE, ldbse, Max1);
			}
			int32_t ic = 0;
			int32_t maxlevel = (nSplitA - 1) + (nSplitB - 1);
			maxlevel = imax(maxlevel - h->fastMode, imin(nSplitA - 1, nSplitB - 1));
			if (h->fastModeRule == 1 && h->fastMode) maxlevel = imin(nSplitA - 1, nSplitB - 1);
			if (n == 1 && m == 1 && h->fastMode == 0) { /* DOT :209-215 */
				gemm_tn_sd(nSplitA, nSplitB, k, AS, ldas, BS, ldbs, CS, nSplitA);
				ic++;
			} else if (n == 1 && h->fastMode == 0) { /* GEMV :216-225 */
				for (int32_t ia = 0; ia < imin(maxlevel + 1, nSplitA); ia++) {
					const int32_t numB = imin(nSplitB, maxlevel + 1 - ia);
					gemm_tn_sd(mbk_, numB, k, AS + (int64_t)ldas * mbk_ * ia, ldas, BS, ldbs,
					           CS + (int64_t)ldcs * numB * ic, ldcs);
					ic++;
				}
			} else { /* GEMM :226-264 */
				for (int32_t ik = 0; ik <= maxlevel; ik++)
					for (int32_t ia = 0; ia < nSplitA; ia++)
						for (int32_t ib = 0; ib < nSplitB; ib++)
							if (ik == ia + ib) {
								gemm_tn_sd(mbk_, nbk_, k, AS + (int64_t)ldas * mbk_ * ia, ldas,
								           BS + (int64_t)ldbs * nbk_ * ib, ldbs,
								           CS + (int64_t)ldcs * nbk_ * ic, ldcs);
								ic++;
							}
			}
			float *Cblk = &C[(int64_t)ldc * (in * nbk) + (int64_t)im * mbk];
			if (m == 1 && n == 1) { /* :277-298 */
				const int32_t sumorder = (h->fastMode == 0) ? 2 : 1;
				global_fsum_sd(1, 1, AE, ldase, nSplitA, BE, ldbse, nSplitB, CS, 1, 1, Cblk, ldc, alpha,
				               beta, maxlevel, sumorder);
			} else {
				const int32_t sumorder = (h->fastMode == 0 && nbk_ == 1) ? 3 : 1;
				global_fsum_sd(mbk_, nbk_, AE, ldase, nSplitA, BE, ldbse, nSplitB, CS, (int64_t)ldcs * nbk_,
				               ldcs, Cblk, ldc, alpha, beta, maxlevel, sumorder);
			}
			block_count++;
			h->nSplitA += nSplitA;
			h->nSplitB += nSplitB;
			h->nSplitC += ic;
		}
	}
	h->nSplitA /= (float)block_count;
	h->nSplitB /= (float)block_count;
	h->nSplitC /= (float)block_count;
	h->useBatchedGemmFlag = saveBatched;
	h->fastMode = saveFast;
	return 0;
}

float ozo_sd_dot(ozo_handle *h, int32_t n, const float *x, const float *y) { /* ozblas_dot.cpp */
	float ret = 0.f; /* the reference leaves ret uninitialised (:25); it is the beta*C operand */
	if (ozo_sd_gemm(h, 't', 'n', 1, 1, n, 1.f, x, n, y, n, 0.f, &ret, 1) != 0) return NAN;
	return ret;
}

float ozo_sd_nrm2(ozo_handle *h, int32_t n, const float *x) { /* ozblas_nrm2.cpp:4-31 */
	counter_init(h);
	return (float)sqrt(ozo_sd_dot(h, n, x, x));
}

int32_t ozo_sd_gemv(ozo_handle *h, char trans, int32_t m, int32_t n, float alpha, const float *A,
                    int32_t lda, const float *x, float beta, float *y) { /* ozblas_gemv.cpp */
	if (is_n(trans)) return ozo_sd_gemm(h, 'n', 'n', m, 1, n, alpha, A, lda, x, n, beta, y, m);
	return ozo_sd_gemm(h, 't', 'n', n, 1, m, alpha, A, lda, x, m, beta, y, n);
}

int32_t ozo_sd_csrmv(ozo_handle *h, int32_t m, int32_t n, int32_t nnz, float alpha, const float *val,
                     const int32_t *colind, const int32_t *rowptr, const float *x, float beta,
                     float *y) {
	/* src/cpu/ozblas_csrmv.cpp:4-181 with TYPE1 = float, TYPE2 = double; with a pre-split matrix
	   (memMaskSplitA != 0) val holds the double slices of ozo_sd_csrmv_splitA */
	if (h->reproMode == 0 && h->nSplitMax == 1) return -3;
	if (h->splitEpsMode == 2 || h->sumMode != 0) return -3;
	if (h->fastMode) return -3; /* the Fsum then reads products the SpMM did not write */
	counter_init(h);
	int32_t L = (h->nSplitMax > 0) ? h->nSplitMax : NUM_SPLIT_DEFAULT_MAX;
	if (h->reproMode == 1) L++;
	int32_t ldas, ldase, ldbs, ldcs;
	if (h->memMaskSplitA != 0) h->memAddr = 0;
	const uint64_t oAS = mat_alloc(h, nnz, L, 8, &ldas);
	const uint64_t oAT = vec_alloc(h, nnz, 4);
	const uint64_t oAmax = vec_alloc(h, m, 4);
	const uint64_t oAE = mat_alloc(h, m, L, 2, &ldase);
	if (h->memMaskSplitA != 0) h->memMaskSplitA = h->memAddr;
	const uint64_t oBS = mat_alloc(h, n, L, 8, &ldbs);
	const uint64_t oCS = mat_alloc(h, m, L * L, 8, &ldcs);
	const uint64_t oBT = vec_alloc(h, n, 4);
	vec_alloc(h, m, 4);
	const uint64_t oBE = vec_alloc(h, L, 2);
	const uint64_t oBmax = vec_alloc(h, 1, 4);
	if (mem_check(h)) {
		fprintf(stderr, "ozoracle: memory shortage\n");
		return -1;
	}
	arena_check(h);
	double *AS;
	int16_t *AE = WP(h, oAE, int16_t), *BE = WP(h, oBE, int16_t);
	int32_t nSplitA;
	if (h->memMaskSplitA == 0) {
		AS = WP(h, oAS, double);
		nSplitA = split_sparse_sd(h, m, val, rowptr, WP(h, oAT, float), AS, ldas, AE, ldase,
		                          WP(h, oAmax, float));
	} else {
		AS = (double *)val;
		nSplitA = h->nSplitA_;
	}
	double *BS = WP(h, oBS, double), *CS = WP(h, oCS, double);
	const int32_t nSplitB = ozo_sd_split(h, n, 1, x, n, WP(h, oBT, float), n, BS, ldbs, BE, 1,
	                                     WP(h, oBmax, float));
	int32_t maxlevel = (nSplitA - 1) + (nSplitB - 1);
	double *pC = CS;
	int32_t ic = 0;
	for (int32_t ia = 0; ia < imin(maxlevel + 1, nSplitA); ia++) {
		const int32_t numB = imin(nSplitB, maxlevel + 1 - ia);
		csrmm_n_sd(m, numB, AS + (int64_t)ldas * ia, colind, rowptr, BS, ldbs, pC, ldcs);
		pC += (int64_t)ldcs * numB;
		ic += numB;
	}
	h->nSplitA += nSplitA;
	h->nSplitB += nSplitB;
	h->nSplitC += ic;
	global_fsum_sd(m, 1, AE, ldase, nSplitA, BE, 1, nSplitB, CS, ldcs, ldcs, y, 1, alpha, beta,
	               maxlevel, 3);
	return 0;
}

double *ozo_sd_csrmv_splitA(ozo_handle *h, int32_t m, int32_t n, int32_t nnz, const float *val,
                            const int32_t *rowptr) {
	/* src/cpu/ozblas_csrmv.cpp:191-271 with TYPE1 = float, TYPE2 = double */
	(void)n;
	if (h->reproMode == 0 && h->nSplitMax == 1) return NULL;
	counter_init(h);
	int32_t L = (h->nSplitMax > 0) ? h->nSplitMax : NUM_SPLIT_DEFAULT_MAX;
	if (h->reproMode == 1) L++;
	int32_t ldas, ldase;
	const uint64_t oAS = mat_alloc(h, nnz, L, 8, &ldas);
	const uint64_t oAT = vec_alloc(h, nnz, 4);
	const uint64_t oAmax = vec_alloc(h, m, 4);
	const uint64_t oAE = mat_alloc(h, m, L, 2, &ldase);
	if (mem_check(h)) {
		fprintf(stderr, "ozoracle: memory shortage\n");
		return NULL;
	}
	arena_check(h);
	h->memMaskSplitA = h->memAddr;
	double *AS = WP(h, oAS, double);
	float *AT = WP(h, oAT, float), *Amax = WP(h, oAmax, float);
	int16_t *AE = WP(h, oAE, int16_t);
	int32_t nA = split_sparse_sd(h, m, val, rowptr, AT, AS, ldas, AE, ldase, Amax);
	if (h->nSplitMax == 0) {
		int32_t nAOld;
		h->splitShift = 1;
		do {
			nAOld = nA;
			h->splitShift *= 2;
			nA = split_sparse_sd(h, m, val, rowptr, AT, AS, ldas, AE, ldase, Amax);
		} while (nAOld == nA && h->splitShift < 512);
		if (nAOld == nA) h->splitShift = 1;
		nA = split_sparse_sd(h, m, val, rowptr, AT, AS, ldas, AE, ldase, Amax);
	}
	h->nSplitA_ = nA;
	h->nSplitA = (float)nA;
	return AS;
}

int32_t ozo_sd_cg(ozo_handle *h, int32_t n, int32_t nnz, const float *val, const int32_t *colind,
                  const int32_t *rowptr, const float *b, float *x, int32_t maxiter, float tol) {
	/* src/cpu/ozblas_cg.cpp:24-128 with TYPE1 = float, TYPE2 = double (trueres diagnostics not
	   restated); the history (resi / nrmb, a float) is stored widened into cg_verbose1 */
	float alpha, beta, resi, nrmb, tmp, dnew, dold;
	float *R = (float *)malloc(sizeof(float) * (size_t)n);
	float *P = (float *)malloc(sizeof(float) * (size_t)n);
	float *Q = (float *)calloc((size_t)n, sizeof(float)); /* read as 0 * q by the first product */
	h->cg_numiter = 0;
	double *AS = ozo_sd_csrmv_splitA(h, n, n, nnz, val, rowptr);
	if (!AS) {
		free(R);
		free(P);
		free(Q);
		return -3;
	}
	nrmb = ozo_sd_nrm2(h, n, b);
	memcpy(R, b, sizeof(float) * (size_t)n);
	int32_t rc = ozo_sd_csrmv(h, n, n, nnz, -1.f, (const float *)AS, colind, rowptr, x, 1.f, R);
	dold = ozo_sd_dot(h, n, R, R);
	resi = (float)sqrt(dold);
	if (h->verbose > 0 && h->cg_verbose1) h->cg_verbose1[0] = resi / nrmb;
	memcpy(P, R, sizeof(float) * (size_t)n);
	int32_t vcnt = 1;
	while (rc == 0 && h->cg_numiter < maxiter) {
		h->cg_numiter++;
		rc = ozo_sd_csrmv(h, n, n, nnz, 1.f, (const float *)AS, colind, rowptr, P, 0.f, Q);
		tmp = ozo_sd_dot(h, n, P, Q);
		alpha = dold / tmp;
		ozo_s_axpy(h, n, alpha, P, x);
		ozo_s_axpy(h, n, -alpha, Q, R);
		dnew = ozo_sd_dot(h, n, R, R);
		resi = (float)sqrt(dnew);
		beta = dnew / dold;
		dold = dnew;
		for (int32_t i = 0; i < n; i++) P[i] = beta * P[i]; /* cblas_sscal */
		ozo_s_axpy(h, n, 1.f, R, P);
		if (h->verbose > 0) {
			if (resi / nrmb < tol || h->cg_numiter % h->verbose == 0) {
				if (h->cg_verbose1) h->cg_verbose1[vcnt] = resi / nrmb;
				vcnt++;
			}
		}
		if (resi / nrmb < tol) break;
	}
	free(R);
	free(P);
	free(Q);
	h->memMaskSplitA = 0;
	h->splitShift = 1;
	return rc;
}
