/*
 * ozoracle_ds.c -- CPU oracle, mixed binary64/binary32 (TEST INFRASTRUCTURE ONLY; see ozoracle.h).
 *
 * Restatement of OzBLAS 1.6a src/cpu for the <double,float,float> / <double,float> instantiations
 * (src/cpu/ozblas_gemm.cpp:319, ozblas_dot.cpp:43, ozblas_nrm2.cpp:35, ozblas_gemv.cpp:33,
 * ozblas_csrmv.cpp:185,275, ozblas_cg.cpp:132, ozblas_split.cpp:293,932): binary64 data split in
 * binary64 arithmetic into binary32 slices scaled by 2^-tau, with the exponents tau kept as shorts.
 *  - rho is getRho<double,float>: ceil(53 - (24 - log2 dim) / 2) (epse of TYPE1 = 53, of TYPE2 = 24);
 *  - the mixed split kernels store scalbn(slice, -tau) as a float and no remainder slot, and the
 *    split loop does not adjust its cap or count for reproMode = 0 (the adjustments apply only to
 *    same-type instantiations, src/cpu/ozblas_split.cpp:275-287,916-925);
 *  - slice products are float GEMM / CSRMM of the scaled slices (24 bits at most, so exact; here
 *    the exact double sums, rounded once to float);
 *  - the sum is the mixed Fsum t += scalbn((double)c, seA + seB) (src/cpu/ozblas_sum.cpp:383-432);
 *    the mixed NearSum is not available in the reference (:275-330 prints an error and exits), so
 *    sumMode = 1 returns -3 here, as does Dot2.
 */
#define _GNU_SOURCE
#include "ozoracle_int.h"

int32_t ozo_ds_rho(int32_t dim, int32_t splitEpsMode) {
	/* getRho<double,float> (src/cpu/ozblas_split.cpp:20-34) with getEpse<double> = 53,
	   getEpse<float> = 24, getEpse2<float> = 53 */
	double r;
	switch (splitEpsMode) {
	case 2: r = ceil((53 - (53 - log2(1. * dim)) / 2)); break;
	case 9: r = ceil((53 - (24 - log2(2. * sqrt(dim))) / 2)); break;
	default: r = ceil((53 - (24 - log2(1. * dim)) / 2)); break;
	}
	return d2i(r);
}

/* ozblasFindMaxDevice :55-73 (binary64 input) */
static void find_max_ds(int32_t m, int32_t n, const double *in, int32_t ldi, double *max) {
	if (n == 1) { /* cblas_idamax: first index of the largest |x| */
		int32_t idx = 0;
		double best = m > 0 ? fabs(in[0]) : 0.;
		for (int32_t i = 1; i < m; i++)
			if (fabs(in[i]) > best) {
				best = fabs(in[i]);
				idx = i;
			}
		max[0] = fabs(in[idx]);
	} else {
		for (int32_t j = 0; j < n; j++) {
			double mx = 0.;
			for (int32_t i = 0; i < m; i++) mx = std_max(mx, fabs(in[(int64_t)j * ldi + i]));
			max[j] = mx;
		}
	}
}

/* ozblasSplitVecKernel<double,float> :106-135 */
static double split_vec_ds(int32_t n, int32_t rho, const double *in, double *out, float *split,
                           int16_t *exps, double *max, int32_t shift) {
	const int32_t tau = exps[0] = (int16_t)ozo_tau(max[0]);
	const double sigma = SPLIT_CONST * scalbn_l(1., (int64_t)rho + tau) * shift;
	double mx = 0.;
	for (int32_t i = 0; i < n; i++) {
		double input = in[i];
		const double sp = (input + sigma) - sigma;
		input = input - sp;
		split[i] = (float)scalbn(sp, -tau);
		out[i] = input;
		mx = std_max(mx, fabs(input));
	}
	max[0] = mx;
	return mx;
}

/* ozblasSplitCKernel<double,float> :182-222 */
static double split_c_ds(int32_t m, int32_t n, int32_t rho, const double *in, int32_t ldi, double *out,
                         int32_t ldo, float *split, int32_t lds, int16_t *exps, double *max,
                         int32_t shift) {
	double check = 0.;
	for (int32_t j = 0; j < n; j++) {
		const int32_t tau = exps[j] = (int16_t)ozo_tau(max[j]);
		const double sigma = SPLIT_CONST * scalbn_l(1., (int64_t)rho + tau) * shift;
		double mx = 0.;
		for (int32_t i = 0; i < m; i++) {
			double input = in[(int64_t)j * ldi + i];
			const double sp = (input + sigma) - sigma;
			input = input - sp;
			split[(int64_t)j * lds + i] = (float)scalbn(sp, -tau);
			out[(int64_t)j * ldo + i] = input;
			mx = std_max(mx, fabs(input));
		}
		max[j] = mx;
		check += mx;
	}
	return check;
}

static double split_device_ds(int32_t m, int32_t n, const double *in, int32_t ldi, double *out,
                              int32_t ldo, float *split, int32_t lds, int16_t *exps, double *max,
                              int32_t epsMode, int32_t shift) {
	const int32_t rho = ozo_ds_rho(m, epsMode);
	if (n == 1) return split_vec_ds(m, rho, in, out, split, exps, max, shift);
	return split_c_ds(m, n, rho, in, ldi, out, ldo, split, lds, exps, max, shift);
}

/* ozblasSplit<double,float> :254-289: no reproMode adjustment for mixed types */
int32_t ozo_ds_split(ozo_handle *h, int32_t m, int32_t n, const double *in, int32_t ldi, double *out,
                     int32_t ldo, float *split, int32_t lds, int16_t *exps, int32_t ldse, double *max) {
	int32_t s;
	find_max_ds(m, n, in, ldi, max);
	split_device_ds(m, n, in, ldi, out, ldo, split, lds, exps, max, h->splitEpsMode, h->splitShift);
	const int32_t maxS = (h->nSplitMax > 0) ? h->nSplitMax : NUM_SPLIT_DEFAULT_MAX;
	for (s = 1; s < maxS; s++) {
		if (split_device_ds(m, n, out, ldo, out, ldo, &split[(int64_t)lds * n * s], lds,
		                    &exps[(int64_t)ldse * s], max, h->splitEpsMode, h->splitShift) == 0.) {
			s++;
			break;
		}
	}
	return s;
}

/* sparse: ozblasSplitSparseNKernel<double,float> :830-869, ozblasSplitSparse :896-928 */
static double split_sparse_ds_kernel(int32_t m, const double *in, const int32_t *rowptr, double *out,
                                     float *split, int16_t *exps, double *max, int32_t epsMode,
                                     int32_t shift) {
	double check = 0.;
	for (int32_t r = 0; r < m; r++) {
		const int32_t rho = ozo_ds_rho(rowptr[r + 1] - rowptr[r], epsMode);
		const int32_t tau = exps[r] = (int16_t)ozo_tau(max[r]);
		const double sigma = SPLIT_CONST * scalbn_l(1., (int64_t)rho + tau) / shift;
		double mx = 0.;
		for (int32_t i = rowptr[r]; i < rowptr[r + 1]; i++) {
			double input = in[i];
			const double sp = (input + sigma) - sigma;
			input = input - sp;
			split[i] = (float)scalbn(sp, -tau);
			out[i] = input;
			mx = std_max(mx, fabs(input));
		}
		max[r] = mx;
		check += mx;
	}
	return check;
}

static int32_t split_sparse_ds(ozo_handle *h, int32_t m, const double *in, const int32_t *rowptr,
                               double *out, float *split, int32_t lds, int16_t *exps, int32_t ldse,
                               double *max) {
	int32_t s;
	for (int32_t r = 0; r < m; r++) { /* ozblasFindMaxSparseDevice */
		double mx = 0.;
		for (int32_t i = rowptr[r]; i < rowptr[r + 1]; i++) mx = std_max(mx, fabs(in[i]));
		max[r] = mx;
	}
	split_sparse_ds_kernel(m, in, rowptr, out, split, exps, max, h->splitEpsMode, h->splitShift);
	const int32_t maxS = (h->nSplitMax > 0) ? h->nSplitMax : NUM_SPLIT_DEFAULT_MAX;
	for (s = 1; s < maxS; s++)
		if (split_sparse_ds_kernel(m, out, rowptr, out, &split[(int64_t)lds * s], &exps[(int64_t)ldse * s],
		                           max, h->splitEpsMode, h->splitShift) == 0.) {
			s++;
			return s;
		}
	return s;
}

/* mixed Fsum (ozblasGlobalFsumKernel<double,float> :383-432) */
static void global_fsum_ds(int32_t m, int32_t n, const int16_t *expA, int32_t ldase, int32_t nA,
                           const int16_t *expB, int32_t ldbse, int32_t nB, const float *Cs, int64_t llsc,
                           int32_t ldsc, double *C, int32_t ldc, double alpha, double beta,
                           int32_t maxlevel, int32_t sumOrder) {
	for (int32_t y = 0; y < n; y++)
		for (int32_t x = 0; x < m; x++) {
			double t = 0.;
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
							const double c = (double)Cs[llsc * it + (int64_t)y * ldsc + x];
							const int32_t seB = expB[(int64_t)ldbse * ib + y];
							t += scalbn(c, seA + seB);
							ic++;
						}
				}
			double *cp = &C[(int64_t)y * ldc + x];
			*cp = fma(alpha, t, (beta * *cp));
		}
}

/* float GEMM of scaled slices: every slice product sum is exact (at most 24 bits), so the exact
   double sum rounded to float is what any float GEMM returns */
static void gemm_tn_ds(int32_t M, int32_t N, int32_t K, const float *A, int32_t lda, const float *B,
                       int32_t ldb, float *C, int32_t ldc) {
#pragma omp parallel for schedule(static)
	for (int32_t j = 0; j < N; j++)
		for (int32_t i = 0; i < M; i++) {
			const float *a = A + (int64_t)i * lda;
			const float *b = B + (int64_t)j * ldb;
			double s = 0.;
			for (int32_t kk = 0; kk < K; kk++) s += (double)a[kk] * (double)b[kk];
			C[(int64_t)j * ldc + i] = (float)s;
		}
}

static void omatcopy_t_ds(int32_t rows, int32_t cols, const double *src, int32_t lds, double *dst,
                          int32_t ldd) {
	for (int32_t j = 0; j < cols; j++)
		for (int32_t i = 0; i < rows; i++) dst[(int64_t)i * ldd + j] = src[(int64_t)j * lds + i];
}

/* csrmm_n<float> (src/cpu/ozblas_blas_wrapper.cpp:231-260): fma1<float> chain per row */
static void csrmm_n_ds(int32_t m, int32_t n, const float *val, const int32_t *colind,
                       const int32_t *rowptr, const float *B, int32_t ldb, float *C, int32_t ldc) {
	for (int32_t r = 0; r < m; r++)
		for (int32_t j = 0; j < n; j++) {
			float tr = 0.f;
			for (int32_t i = rowptr[r]; i < rowptr[r + 1]; i++)
				tr = (float)fma((double)val[i], (double)B[(int64_t)j * ldb + colind[i]], (double)tr);
			C[(int64_t)j * ldc + r] = 1.f * tr; /* alpha = 1, beta = 0 */
		}
}

int32_t ozo_ds_gemm(ozo_handle *h, char transA, char transB, int32_t m, int32_t n, int32_t k,
                    double alpha, const double *A, int32_t lda, const double *B, int32_t ldb,
                    double beta, double *C, int32_t ldc) {
	/* src/cpu/ozblas_gemm.cpp:3-315 with TYPE1 = double, TYPE2 = float */
	if (h->reproMode == 0 && h->nSplitMax == 1) return -3; /* plain BLAS bypass :14-17 */
	if (h->splitEpsMode == 2) return -3; /* the float Dot2 GEMM is a stub */
	if (h->sumMode != 0) return -3;      /* the mixed NearSum exits in the reference */
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
		oAS = mat_alloc(h, k, mbk * L, 4, &ldas);
		oBS = mat_alloc(h, k, nbk * L, 4, &ldbs);
		oCS = mat_alloc(h, mbk, sizeCn, 4, &ldcs);
		oT = mat_alloc(h, k, imax(mbk, nbk), 8, &ldt);
		mat_alloc(h, mbk, nbk, 8, &ldct);
		oAE = mat_alloc(h, mbk, L, 2, &ldase);
		oBE = mat_alloc(h, nbk, L, 2, &ldbse);
		oMax = vec_alloc(h, imax(k, imax(mbk, nbk)), 8);
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
	float *AS = WP(h, oAS, float), *BS = WP(h, oBS, float), *CS = WP(h, oCS, float);
	double *Tmp = WP(h, oT, double), *Max1 = WP(h, oMax, double);
	int16_t *AE = WP(h, oAE, int16_t), *BE = WP(h, oBE, int16_t);
	(void)ldct;

	int32_t block_count = 0;
	const int32_t nbm = (int32_t)ceilf((float)m / mbk), nbn = (int32_t)ceilf((float)n / nbk);
	for (int32_t im = 0; im < nbm; im++) {
		const int32_t mbk_ = (m - mbk * im >= mbk) ? mbk : m - mbk * im;
		int32_t nSplitA;
		if (is_n(transA)) {
			omatcopy_t_ds(mbk_, k, A + (int64_t)im * mbk, lda, Tmp, ldt);
			nSplitA = ozo_ds_split(h, k, mbk_, Tmp, ldt, Tmp, ldt, AS, ldas, AE, ldase, Max1);
		} else {
			nSplitA = ozo_ds_split(h, k, mbk_, A + (int64_t)im * mbk * lda, lda, Tmp, ldt, AS, ldas, AE,
			                       ldase, Max1);
		}
		for (int32_t in = 0; in < nbn; in++) {
			const int32_t nbk_ = (n - nbk * in >= nbk) ? nbk : n - nbk * in;
			int32_t nSplitB;
			if (is_n(transB)) {
				nSplitB = ozo_ds_split(h, k, nbk_, B + (int64_t)in * nbk * ldb, ldb, Tmp, ldt, BS, ldbs, BE,
				                       ldbse, Max1);
			} else {
				omatcopy_t_ds(nbk_, k, B + (int64_t)in * nbk, ldb, Tmp, ldt);
				nSplitB = ozo_ds_split(h, k, nbk_, Tmp, ldt, Tmp, ldt, BS, ldbs, BE, ldbse, Max1);
			}
			int32_t ic = 0;
			int32_t maxlevel = (nSplitA - 1) + (nSplitB - 1);
			maxlevel = imax(maxlevel - h->fastMode, imin(nSplitA - 1, nSplitB - 1));
			if (h->fastModeRule == 1 && h->fastMode) maxlevel = imin(nSplitA - 1, nSplitB - 1);
			if (n == 1 && m == 1 && h->fastMode == 0) { /* DOT :209-215 */
				gemm_tn_ds(nSplitA, nSplitB, k, AS, ldas, BS, ldbs, CS, nSplitA);
				ic++;
			} else if (n == 1 && h->fastMode == 0) { /* GEMV :216-225 */
				for (int32_t ia = 0; ia < imin(maxlevel + 1, nSplitA); ia++) {
					const int32_t numB = imin(nSplitB, maxlevel + 1 - ia);
					gemm_tn_ds(mbk_, numB, k, AS + (int64_t)ldas * mbk_ * ia, ldas, BS, ldbs,
					           CS + (int64_t)ldcs * numB * ic, ldcs);
					ic++;
				}
			} else { /* GEMM :226-264 */
				for (int32_t ik = 0; ik <= maxlevel; ik++)
					for (int32_t ia = 0; ia < nSplitA; ia++)
						for (int32_t ib = 0; ib < nSplitB; ib++)
							if (ik == ia + ib) {
								gemm_tn_ds(mbk_, nbk_, k, AS + (int64_t)ldas * mbk_ * ia, ldas,
								           BS + (int64_t)ldbs * nbk_ * ib, ldbs,
								           CS + (int64_t)ldcs * nbk_ * ic, ldcs);
								ic++;
							}
			}
			double *Cblk = &C[(int64_t)ldc * (in * nbk) + (int64_t)im * mbk];
			if (m == 1 && n == 1) { /* :277-298 */
				const int32_t sumorder = (h->fastMode == 0) ? 2 : 1;
				global_fsum_ds(1, 1, AE, ldase, nSplitA, BE, ldbse, nSplitB, CS, 1, 1, Cblk, ldc, alpha,
				               beta, maxlevel, sumorder);
			} else {
				const int32_t sumorder = (h->fastMode == 0 && nbk_ == 1) ? 3 : 1;
				global_fsum_ds(mbk_, nbk_, AE, ldase, nSplitA, BE, ldbse, nSplitB, CS, (int64_t)ldcs * nbk_,
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

double ozo_ds_dot(ozo_handle *h, int32_t n, const double *x, const double *y) { /* ozblas_dot.cpp */
	double ret = 0.;
	if (ozo_ds_gemm(h, 't', 'n', 1, 1, n, 1., x, n, y, n, 0., &ret, 1) != 0) return NAN;
	return ret;
}

double ozo_ds_nrm2(ozo_handle *h, int32_t n, const double *x) { /* ozblas_nrm2.cpp:4-31 */
	counter_init(h);
	return sqrt(ozo_ds_dot(h, n, x, x));
}

int32_t ozo_ds_gemv(ozo_handle *h, char trans, int32_t m, int32_t n, double alpha, const double *A,
                    int32_t lda, const double *x, double beta, double *y) { /* ozblas_gemv.cpp */
	if (is_n(trans)) return ozo_ds_gemm(h, 'n', 'n', m, 1, n, alpha, A, lda, x, n, beta, y, m);
	return ozo_ds_gemm(h, 't', 'n', n, 1, m, alpha, A, lda, x, m, beta, y, n);
}

int32_t ozo_ds_csrmv(ozo_handle *h, int32_t m, int32_t n, int32_t nnz, double alpha, const double *val,
                     const int32_t *colind, const int32_t *rowptr, const double *x, double beta,
                     double *y) {
	/* src/cpu/ozblas_csrmv.cpp:4-181 with TYPE1 = double, TYPE2 = float */
	if (h->reproMode == 0 && h->nSplitMax == 1) return -3;
	if (h->splitEpsMode == 2 || h->sumMode != 0) return -3;
	if (h->fastMode) return -3; /* the Fsum then reads products the SpMM did not write */
	counter_init(h);
	int32_t L = (h->nSplitMax > 0) ? h->nSplitMax : NUM_SPLIT_DEFAULT_MAX;
	if (h->reproMode == 1) L++;
	int32_t ldas, ldase, ldbs, ldcs;
	if (h->memMaskSplitA != 0) h->memAddr = 0;
	const uint64_t oAS = mat_alloc(h, nnz, L, 4, &ldas);
	const uint64_t oAT = vec_alloc(h, nnz, 8);
	const uint64_t oAmax = vec_alloc(h, m, 8);
	const uint64_t oAE = mat_alloc(h, m, L, 2, &ldase);
	if (h->memMaskSplitA != 0) h->memMaskSplitA = h->memAddr;
	const uint64_t oBS = mat_alloc(h, n, L, 4, &ldbs);
	const uint64_t oCS = mat_alloc(h, m, L * L, 4, &ldcs);
	const uint64_t oBT = vec_alloc(h, n, 8);
	vec_alloc(h, m, 8);
	const uint64_t oBE = vec_alloc(h, L, 2);
	const uint64_t oBmax = vec_alloc(h, 1, 8);
	if (mem_check(h)) {
		fprintf(stderr, "ozoracle: memory shortage\n");
		return -1;
	}
	arena_check(h);
	float *AS;
	int16_t *AE = WP(h, oAE, int16_t), *BE = WP(h, oBE, int16_t);
	int32_t nSplitA;
	if (h->memMaskSplitA == 0) {
		AS = WP(h, oAS, float);
		nSplitA = split_sparse_ds(h, m, val, rowptr, WP(h, oAT, double), AS, ldas, AE, ldase,
		                          WP(h, oAmax, double));
	} else {
		AS = (float *)val;
		nSplitA = h->nSplitA_;
	}
	float *BS = WP(h, oBS, float), *CS = WP(h, oCS, float);
	const int32_t nSplitB = ozo_ds_split(h, n, 1, x, n, WP(h, oBT, double), n, BS, ldbs, BE, 1,
	                                     WP(h, oBmax, double));
	int32_t maxlevel = (nSplitA - 1) + (nSplitB - 1);
	float *pC = CS;
	int32_t ic = 0;
	for (int32_t ia = 0; ia < imin(maxlevel + 1, nSplitA); ia++) {
		const int32_t numB = imin(nSplitB, maxlevel + 1 - ia);
		csrmm_n_ds(m, numB, AS + (int64_t)ldas * ia, colind, rowptr, BS, ldbs, pC, ldcs);
		pC += (int64_t)ldcs * numB;
		ic += numB;
	}
	h->nSplitA += nSplitA;
	h->nSplitB += nSplitB;
	h->nSplitC += ic;
	global_fsum_ds(m, 1, AE, ldase, nSplitA, BE, 1, nSplitB, CS, ldcs, ldcs, y, 1, alpha, beta,
	               maxlevel, 3);
	return 0;
}

float *ozo_ds_csrmv_splitA(ozo_handle *h, int32_t m, int32_t n, int32_t nnz, const double *val,
                           const int32_t *rowptr) {
	/* src/cpu/ozblas_csrmv.cpp:191-271 with TYPE2 = float */
	(void)n;
	if (h->reproMode == 0 && h->nSplitMax == 1) return NULL;
	counter_init(h);
	int32_t L = (h->nSplitMax > 0) ? h->nSplitMax : NUM_SPLIT_DEFAULT_MAX;
	if (h->reproMode == 1) L++;
	int32_t ldas, ldase;
	const uint64_t oAS = mat_alloc(h, nnz, L, 4, &ldas);
	const uint64_t oAT = vec_alloc(h, nnz, 8);
	const uint64_t oAmax = vec_alloc(h, m, 8);
	const uint64_t oAE = mat_alloc(h, m, L, 2, &ldase);
	if (mem_check(h)) {
		fprintf(stderr, "ozoracle: memory shortage\n");
		return NULL;
	}
	arena_check(h);
	h->memMaskSplitA = h->memAddr;
	float *AS = WP(h, oAS, float);
	double *AT = WP(h, oAT, double), *Amax = WP(h, oAmax, double);
	int16_t *AE = WP(h, oAE, int16_t);
	int32_t nA = split_sparse_ds(h, m, val, rowptr, AT, AS, ldas, AE, ldase, Amax);
	if (h->nSplitMax == 0) {
		int32_t nAOld;
		h->splitShift = 1;
		do {
			nAOld = nA;
			h->splitShift *= 2;
			nA = split_sparse_ds(h, m, val, rowptr, AT, AS, ldas, AE, ldase, Amax);
		} while (nAOld == nA && h->splitShift < 512);
		if (nAOld == nA) h->splitShift = 1;
		nA = split_sparse_ds(h, m, val, rowptr, AT, AS, ldas, AE, ldase, Amax);
	}
	h->nSplitA_ = nA;
	h->nSplitA = (float)nA;
	return AS;
}

int32_t ozo_ds_cg(ozo_handle *h, int32_t n, int32_t nnz, const double *val, const int32_t *colind,
                  const int32_t *rowptr, const double *b, double *x, int32_t maxiter, double tol) {
	/* src/cpu/ozblas_cg.cpp:24-128 with TYPE1 = double, TYPE2 = float */
	double alpha, beta, resi, nrmb, tmp, dnew, dold;
	double *R = (double *)malloc(sizeof(double) * (size_t)n);
	double *P = (double *)malloc(sizeof(double) * (size_t)n);
	double *Q = (double *)calloc((size_t)n, sizeof(double));
	h->cg_numiter = 0;
	float *AS = ozo_ds_csrmv_splitA(h, n, n, nnz, val, rowptr);
	if (!AS) {
		free(R);
		free(P);
		free(Q);
		return -3;
	}
	nrmb = ozo_ds_nrm2(h, n, b);
	memcpy(R, b, sizeof(double) * (size_t)n);
	int32_t rc = ozo_ds_csrmv(h, n, n, nnz, -1., (const double *)AS, colind, rowptr, x, 1., R);
	dold = ozo_ds_dot(h, n, R, R);
	resi = sqrt(dold);
	if (h->verbose > 0 && h->cg_verbose1) h->cg_verbose1[0] = resi / nrmb;
	memcpy(P, R, sizeof(double) * (size_t)n);
	int32_t vcnt = 1;
	while (rc == 0 && h->cg_numiter < maxiter) {
		h->cg_numiter++;
		rc = ozo_ds_csrmv(h, n, n, nnz, 1., (const double *)AS, colind, rowptr, P, 0., Q);
		tmp = ozo_ds_dot(h, n, P, Q);
		alpha = dold / tmp;
		ozo_axpy(h, n, alpha, P, x);
		ozo_axpy(h, n, -alpha, Q, R);
		dnew = ozo_ds_dot(h, n, R, R);
		resi = sqrt(dnew);
		beta = dnew / dold;
		dold = dnew;
		for (int32_t i = 0; i < n; i++) P[i] = beta * P[i];
		ozo_axpy(h, n, 1., R, P);
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
