/*
 * ozoracle_s.c -- CPU oracle, binary32 (TEST INFRASTRUCTURE ONLY; see ozoracle.h for scope).
 *
 * Restatement of OzBLAS 1.6a src/cpu for the <float,float,float> instantiations
 * (src/cpu/ozblas_gemm.cpp:320, ozblas_dot.cpp:44, ozblas_nrm2.cpp:36, ozblas_gemv.cpp:34,
 * ozblas_csrmv.cpp:186, ozblas_cg.cpp:133, ozblas_axpy.cpp:30, ozblas_split.cpp:294,933).
 *
 * The reference compiles these templates with <cmath> and no using-directive, so inside them an
 * unqualified log2/fma/scalbn/fabs/sqrt of a float argument is the C library's double function:
 *   - log21<float>(v)   = (float)log2((double)v)          (src/cpu/ozblas_aux.cpp:37-42)
 *   - fma1<float>(a,b,c) = (float)fma(a, b, c) in double   (src/cpu/ozblas_aux.cpp:4-9)
 *   - scalbn1<float>(v,n) = (float)scalbn((double)v, n)    (src/cpu/ozblas_aux.cpp:26-31)
 *   - sqrt(ret) in Rnrm2/Rcg = (float)sqrt((double)ret)
 * while NearSum calls std::log2 explicitly (src/cpu/ozblas_sum.cpp:35), which is log2f for a float.
 * Mixed float/double expressions are evaluated as C does (float operands promoted), then rounded
 * once to float on assignment, exactly as in the reference.  Compiled with -ffp-contract=off.
 *
 * Dot2 (splitEpsMode = 2) is not restated: the reference's binary32 Dot2 GEMM is an empty stub
 * that prints "sgemm_x2 under construction." (src/cpu/ozblas_blas_wrapper.cpp:179-181).
 */
#define _GNU_SOURCE
#include "ozoracle_int.h"

static inline float std_maxf(float a, float b) { return (a < b) ? b : a; } /* std::max<float> */

/* ---------------------------------------------------------------------------------------------
 * split (src/cpu/ozblas_split.cpp) for TYPE1 = TYPE2 = float
 * ------------------------------------------------------------------------------------------- */
int32_t ozo_s_rho(int32_t dim, int32_t splitEpsMode) {
	/* getRho<float,float> :20-34: getEpse<float> = 24, getEpse2<float> = 53
	   (src/cpu/ozblas_common.h:93-95,111-113) */
	double r;
	switch (splitEpsMode) {
	case 2: r = ceil((24 - (53 - log2(1. * dim)) / 2)); break;
	case 9: r = ceil((24 - (24 - log2(2. * sqrt(dim))) / 2)); break;
	default: r = ceil((24 - (24 - log2(1. * dim)) / 2)); break;
	}
	return d2i(r);
}

int32_t ozo_s_tau(float max) { /* (short)ceil(log21(fabs1(max))) :86,158,808 */
	const float l = (float)log2((double)(float)fabs((double)max));
	return (int32_t)(int16_t)d2i(ceil((double)l));
}

/* CONST * scalbn1(one, rho + tau) * splitShift (:88,159), one = (float)1 */
static float sigma_s(int32_t rho, int32_t tau, int32_t shift) {
	const float p = (float)scalbn_l(1.0, (int64_t)rho + tau);
	return (float)(SPLIT_CONST * (double)p * shift);
}

/* ozblasFindMaxDevice :55-73 */
static void find_max_s(int32_t m, int32_t n, const float *in, int32_t ldi, float *max) {
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
		for (int32_t j = 0; j < n; j++) { /* ozblasFindMaxCKernel :37-52 */
			float mx = 0.f;
			for (int32_t i = 0; i < m; i++) mx = std_maxf(mx, fabsf(in[(int64_t)j * ldi + i]));
			max[j] = mx;
		}
	}
}

/* ozblasSplitVecKernel (same type) :75-104 */
static float split_vec_kernel_s(int32_t n, int32_t rho, const float *in, float *out, float *split,
                                float *max, int32_t shift) {
	const int32_t tau = ozo_s_tau(max[0]);
	const float sigma = sigma_s(rho, tau, shift);
	float mx = 0.f;
	for (int32_t i = 0; i < n; i++) {
		float input = in[i];
		const float sp = (input + sigma) - sigma;
		input = input - sp;
		split[i + n] = input; /* remainder slot at +n (:96) */
		split[i] = sp;
		out[i] = input;
		mx = std_maxf(mx, fabsf(input));
	}
	max[0] = mx;
	return mx;
}

/* ozblasSplitCKernel (same type) :137-179 */
static float split_c_kernel_s(int32_t m, int32_t n, int32_t rho, const float *in, int32_t ldi,
                              float *out, int32_t ldo, float *split, int32_t lds, float *max,
                              int32_t shift) {
	float check = 0.f;
	for (int32_t j = 0; j < n; j++) {
		const int32_t tau = ozo_s_tau(max[j]);
		const float sigma = sigma_s(rho, tau, shift);
		float mx = 0.f;
		for (int32_t i = 0; i < m; i++) {
			float input = in[(int64_t)j * ldi + i];
			const float sp = (input + sigma) - sigma;
			input = input - sp;
			split[(int64_t)j * lds + i + (int64_t)lds * n] = input; /* :166 */
			split[(int64_t)j * lds + i] = sp;
			out[(int64_t)j * ldo + i] = input;
			mx = std_maxf(mx, fabsf(input));
		}
		max[j] = mx;
		check += mx; /* only compared with 0 */
	}
	return check;
}

static float split_device_s(int32_t m, int32_t n, const float *in, int32_t ldi, float *out,
                            int32_t ldo, float *split, int32_t lds, float *max, int32_t epsMode,
                            int32_t shift) {
	const int32_t rho = ozo_s_rho(m, epsMode);
	if (n == 1) return split_vec_kernel_s(m, rho, in, out, split, max, shift);
	return split_c_kernel_s(m, n, rho, in, ldi, out, ldo, split, lds, max, shift);
}

/* ozblasSplit :254-289 */
int32_t ozo_s_split(ozo_handle *h, int32_t m, int32_t n, const float *in, int32_t ldi, float *out,
                    int32_t ldo, float *split, int32_t lds, float *max) {
	int32_t s;
	find_max_s(m, n, in, ldi, max);
	split_device_s(m, n, in, ldi, out, ldo, split, lds, max, h->splitEpsMode, h->splitShift);
	int32_t maxS = (h->nSplitMax > 0) ? h->nSplitMax : NUM_SPLIT_DEFAULT_MAX;
	if (h->reproMode == 0) maxS--;
	for (s = 1; s < maxS; s++) {
		if (split_device_s(m, n, out, ldo, out, ldo, &split[(int64_t)lds * n * s], lds, max,
		                   h->splitEpsMode, h->splitShift) == 0.f) {
			s++;
			break;
		}
	}
	if (h->reproMode == 0) s++;
	return s;
}

/* sparse row split :756-928 */
static void find_max_sparse_s(int32_t m, const float *in, const int32_t *rowptr, float *max) {
	for (int32_t r = 0; r < m; r++) {
		float mx = 0.f;
		for (int32_t i = rowptr[r]; i < rowptr[r + 1]; i++) mx = std_maxf(mx, fabsf(in[i]));
		max[r] = mx;
	}
}

static float split_sparse_kernel_s(int32_t m, const float *in, const int32_t *rowptr, float *out,
                                   float *split, int32_t lds, float *max, int32_t epsMode,
                                   int32_t shift) {
	float check = 0.f;
	for (int32_t r = 0; r < m; r++) {
		const int32_t rho = ozo_s_rho(rowptr[r + 1] - rowptr[r], epsMode);
		const int32_t tau = ozo_s_tau(max[r]);
		const float p = (float)scalbn_l(1.0, (int64_t)rho + tau);
		const float sigma = (float)(SPLIT_CONST * (double)p / shift); /* :809 */
		float mx = 0.f;
		for (int32_t i = rowptr[r]; i < rowptr[r + 1]; i++) {
			float input = in[i];
			const float sp = (input + sigma) - sigma;
			input = input - sp;
			split[i + lds] = input;
			split[i] = sp;
			out[i] = input;
			mx = std_maxf(mx, fabsf(input));
		}
		max[r] = mx;
		check += mx;
	}
	return check;
}

static int32_t split_sparse_s(ozo_handle *h, int32_t m, const float *in, const int32_t *rowptr,
                              float *out, float *split, int32_t lds, float *max) {
	int32_t s;
	find_max_sparse_s(m, in, rowptr, max);
	split_sparse_kernel_s(m, in, rowptr, out, split, lds, max, h->splitEpsMode, h->splitShift);
	int32_t maxS = (h->nSplitMax > 0) ? h->nSplitMax : NUM_SPLIT_DEFAULT_MAX;
	if (h->reproMode == 0) maxS--;
	for (s = 1; s < maxS; s++)
		if (split_sparse_kernel_s(m, out, rowptr, out, &split[(int64_t)lds * s], lds, max,
		                          h->splitEpsMode, h->splitShift) == 0.f) {
			s++;
			return s;
		}
	if (h->reproMode == 0) s++;
	return s;
}

/* ---------------------------------------------------------------------------------------------
 * summation (src/cpu/ozblas_sum.cpp) for TYPE = float
 * ------------------------------------------------------------------------------------------- */
static void transform_s(int32_t n, float *vec, int32_t ld, float rho, float *tau1, float *tau2) {
	/* :10-56 with emin = -126, epse = 24 */
	float mu = (float)fabs(vec[0]);
	for (int32_t i = 1; i < n; i++) mu = (float)mac_max(mu, fabs(vec[(int64_t)i * ld]));
	if ((n == 0) || (mu == 0.f)) {
		*tau1 = rho;
		*tau2 = 0.f;
		return;
	}
	const int32_t m = d2i(ceil(log2((double)(n + 2))));
	/* scalbn1(1., m + ceil(std::log2(mu))): std::log2 of a float is log2f */
	float sigma = (float)scalbn_l(1., (int64_t)d2i(m + ceil((double)log2f(mu))));
	float t = rho;
	while (1) {
		float tau = 0.f;
		for (int32_t i = 0; i < n; i++) {
			const float tmp = (sigma + vec[(int64_t)i * ld]) - sigma;
			vec[(int64_t)i * ld] -= tmp;
			tau += tmp;
		}
		*tau1 = t + tau;
		if ((sigma <= scalbn(1., -126)) || (fabs(*tau1) >= scalbn_l(1., 2 * m + 1 - 24) * sigma)) {
			*tau2 = tau - (*tau1 - t);
			return;
		}
		sigma = (float)(scalbn_l(1., m - 24) * sigma);
		t = *tau1;
	}
}

static void transform_k_s(int32_t n, float *vec, int32_t ld, float rho, float *res, float *r) {
	/* :58-73 */
	float tmp = 0.f, tau1, tau2;
	transform_s(n, vec, ld, rho, &tau1, &tau2);
	for (int32_t i = 0; i < n; i++) tmp += vec[(int64_t)i * ld];
	*res = tau1 + (tau2 + tmp);
	*r = tau2 - (*res - tau1);
}

static inline float get_sign_s(float v) { return (v < 0) ? -1.f : 1.f; }

float ozo_s_nearsum(int32_t n, float *vec, int32_t ld) { /* :81-112 */
	float tmp, res, res2, r, r2, mu, delta, delta2;
	const float eps = (float)scalbn(1., -24);
	transform_k_s(n, vec, ld, 0.f, &res, &r);
	transform_k_s(n, vec, ld, r, &delta, &r2);
	if (delta == 0) return res;
	res2 = (float)(res + get_sign_s(delta) * eps * fabs(res));
	if (res2 == res) {
		mu = (float)(get_sign_s(delta) * eps * fabs(res));
		res2 = (float)(res + 2. * get_sign_s(delta) * eps * fabs(res));
	} else {
		mu = (float)((res2 - res) / 2.);
	}
	if (fabs(delta) < fabs(mu)) return res;
	if (fabs(delta) > fabs(mu)) return res2;
	transform_k_s(n, vec, ld, r2, &delta2, &tmp);
	if (delta2 == 0) return res + mu;
	if (get_sign_s(delta2) == get_sign_s(mu)) return res2;
	return res;
}

/* fma1<float> (src/cpu/ozblas_aux.cpp:4-9): the double fma, rounded to float */
static inline float fma1_s(float a, float b, float c) { return (float)fma(a, b, c); }

static void global_nearsum_s(int32_t m, int32_t n, int32_t nA, int32_t nB, float *Cs, int64_t llsc,
                             int32_t ldsc, float *C, int32_t ldc, float alpha, float beta,
                             int32_t maxlevel) { /* :231-271 */
	int32_t ic = 0;
	for (int32_t ik = 0; ik <= maxlevel; ik++)
		for (int32_t ia = 0; ia < nA; ia++)
			for (int32_t ib = 0; ib < nB; ib++)
				if (ik == ia + ib) ic++;
	for (int32_t y = 0; y < n; y++)
		for (int32_t x = 0; x < m; x++) {
			float t = ozo_s_nearsum(ic, &Cs[(int64_t)y * ldsc + x], (int32_t)llsc);
			float *c = &C[(int64_t)y * ldc + x];
			*c = fma1_s(alpha, t, (beta * *c));
		}
}

static void global_fsum_s(int32_t m, int32_t n, int32_t nA, int32_t nB, const float *Cs, int64_t llsc,
                          int32_t ldsc, float *C, int32_t ldc, float alpha, float beta,
                          int32_t maxlevel, int32_t sumOrder) { /* :331-380 */
	for (int32_t y = 0; y < n; y++)
		for (int32_t x = 0; x < m; x++) {
			float t = 0.f;
			int32_t ic = 0;
			for (int32_t ik = 0; ik <= maxlevel; ik++)
				for (int32_t ia = 0; ia < nA; ia++)
					for (int32_t ib = 0; ib < nB; ib++)
						if (ik == ia + ib) {
							int32_t it;
							switch (sumOrder) {
							case 2: it = nA * ib + ia; break;
							case 3: it = nB * ia + ib; break;
							default: it = ic; break;
							}
							t += Cs[llsc * it + (int64_t)y * ldsc + x];
							ic++;
						}
			float *c = &C[(int64_t)y * ldc + x];
			*c = fma1_s(alpha, t, (beta * *c));
		}
}

static int32_t global_sum_s(ozo_handle *h, int32_t m, int32_t n, int32_t nA, int32_t nB, float *Cs,
                            int64_t llsc, int32_t ldsc, float *C, int32_t ldc, float alpha,
                            float beta, int32_t maxlevel, int32_t sumOrder) { /* :513-552 */
	if (h->sumMode == 1)
		global_nearsum_s(m, n, nA, nB, Cs, llsc, ldsc, C, ldc, alpha, beta, maxlevel);
	else if (h->sumMode == 30)
		return -1;
	else
		global_fsum_s(m, n, nA, nB, Cs, llsc, ldsc, C, ldc, alpha, beta, maxlevel, sumOrder);
	return 0;
}

/* slice products, exact by construction (24 - rho bits per slice, rho from the inner dimension):
   any evaluation order gives the same float */
static void gemm_tn_s(int32_t M, int32_t N, int32_t K, const float *A, int32_t lda, const float *B,
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

static void omatcopy_t_s(int32_t rows, int32_t cols, const float *src, int32_t lds, float *dst,
                         int32_t ldd) {
	for (int32_t j = 0; j < cols; j++)
		for (int32_t i = 0; i < rows; i++) dst[(int64_t)i * ldd + j] = src[(int64_t)j * lds + i];
}

/* csrmm_n<float> src/cpu/ozblas_blas_wrapper.cpp:231-260 */
static void csrmm_n_s(int32_t m, int32_t n, float alpha, const float *val, const int32_t *colind,
                      const int32_t *rowptr, const float *B, int32_t ldb, float beta, float *C,
                      int32_t ldc) {
#pragma omp parallel for schedule(static)
	for (int32_t r = 0; r < m; r++)
		for (int32_t j = 0; j < n; j++) {
			float tr = 0.f;
			for (int32_t i = rowptr[r]; i < rowptr[r + 1]; i++)
				tr = fma1_s(val[i], B[(int64_t)j * ldb + colind[i]], tr);
			if (beta == 0.)
				C[(int64_t)j * ldc + r] = alpha * tr;
			else
				C[(int64_t)j * ldc + r] = fma1_s(alpha, tr, (beta * C[(int64_t)j * ldc + r]));
		}
}

/* ---------------------------------------------------------------------------------------------
 * drivers
 * ------------------------------------------------------------------------------------------- */
int32_t ozo_s_gemm(ozo_handle *h, char transA, char transB, int32_t m, int32_t n, int32_t k,
                   float alpha, const float *A, int32_t lda, const float *B, int32_t ldb,
                   float beta, float *C, int32_t ldc) {
	/* src/cpu/ozblas_gemm.cpp:3-315 with sizeof(TYPE1) = sizeof(TYPE2) = 4 */
	if (h->reproMode == 0 && h->nSplitMax == 1) { /* bypass :14-17 (cblas_sgemm; order-dependent) */
		for (int32_t j = 0; j < n; j++)
			for (int32_t i = 0; i < m; i++) {
				float s = 0.f;
				for (int32_t kk = 0; kk < k; kk++) {
					float a = is_n(transA) ? A[i + (int64_t)kk * lda] : A[kk + (int64_t)i * lda];
					float b = is_n(transB) ? B[kk + (int64_t)j * ldb] : B[j + (int64_t)kk * ldb];
					s += a * b;
				}
				C[i + (int64_t)j * ldc] = alpha * s + beta * C[i + (int64_t)j * ldc];
			}
		return 0;
	}
	if (h->splitEpsMode == 2) return -3; /* binary32 Dot2 is a stub in the reference */
	counter_init(h);
	const int32_t saveBatched = h->useBatchedGemmFlag, saveFast = h->fastMode;
	if (m == 1 && n == 1) {
		h->useBatchedGemmFlag = 0;
		h->fastMode = 0;
	}
	int32_t mbk = m, nbk = n;
	int32_t L = (h->nSplitMax > 1) ? h->nSplitMax + 1 : NUM_SPLIT_DEFAULT_MAX;
	if (h->reproMode == 1) L++;
	const int globalSum =
	    (h->useBatchedGemmFlag == 1 || h->sumMode == 0 || h->sumMode == 1 || h->sumMode == 30);
	if (h->useBatchedGemmFlag) {
		vec_alloc(h, L * L, 8);
		vec_alloc(h, L * L, 8);
		vec_alloc(h, L * L, 8);
	}
	int32_t ldas = 0, ldbs = 0, ldcs = 0, ldt = 0, ldct = 0, ldase = 0, ldbse = 0;
	uint64_t oAS = 0, oBS = 0, oCS = 0, oT = 0, oMax = 0;
	const uint64_t memAddrTmp = h->memAddr;
	while (mbk > 0 && nbk > 0) { /* fit loop :59-97 */
		int32_t sizeCn = globalSum ? (nbk * L * L) : nbk;
		oAS = mat_alloc(h, k, mbk * L, 4, &ldas);
		oBS = mat_alloc(h, k, nbk * L, 4, &ldbs);
		oCS = mat_alloc(h, mbk, sizeCn, 4, &ldcs);
		oT = mat_alloc(h, k, imax(mbk, nbk), 4, &ldt);
		mat_alloc(h, mbk, nbk, 4, &ldct);
		if (h->sumMode == 3 && h->useBatchedGemmFlag == 0) {
			int32_t dummy;
			mat_alloc(h, mbk, nbk, 8, &dummy);
			mat_alloc(h, mbk, nbk, 8, &dummy);
			mat_alloc(h, mbk, nbk, 8, &dummy);
		}
		mat_alloc(h, mbk, L, 2, &ldase);
		mat_alloc(h, nbk, L, 2, &ldbse);
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
	float *AS = WP(h, oAS, float), *BS = WP(h, oBS, float), *CS = WP(h, oCS, float);
	float *Tmp = WP(h, oT, float), *Max1 = WP(h, oMax, float);
	(void)ldct;
	(void)ldase;
	(void)ldbse;

	int32_t block_count = 0;
	const int32_t nbm = (int32_t)ceilf((float)m / mbk), nbn = (int32_t)ceilf((float)n / nbk);
	for (int32_t im = 0; im < nbm; im++) {
		const int32_t mbk_ = (m - mbk * im >= mbk) ? mbk : m - mbk * im;
		int32_t nSplitA;
		if (is_n(transA)) {
			omatcopy_t_s(mbk_, k, A + (int64_t)im * mbk, lda, Tmp, ldt);
			nSplitA = ozo_s_split(h, k, mbk_, Tmp, ldt, Tmp, ldt, AS, ldas, Max1);
		} else {
			nSplitA = ozo_s_split(h, k, mbk_, A + (int64_t)im * mbk * lda, lda, Tmp, ldt, AS, ldas, Max1);
		}
		for (int32_t in = 0; in < nbn; in++) {
			const int32_t nbk_ = (n - nbk * in >= nbk) ? nbk : n - nbk * in;
			int32_t nSplitB;
			if (is_n(transB)) {
				nSplitB = ozo_s_split(h, k, nbk_, B + (int64_t)in * nbk * ldb, ldb, Tmp, ldt, BS, ldbs, Max1);
			} else {
				omatcopy_t_s(nbk_, k, B + (int64_t)in * nbk, ldb, Tmp, ldt);
				nSplitB = ozo_s_split(h, k, nbk_, Tmp, ldt, Tmp, ldt, BS, ldbs, Max1);
			}
			int32_t ic = 0;
			int32_t maxlevel = (nSplitA - 1) + (nSplitB - 1);
			maxlevel = imax(maxlevel - h->fastMode, imin(nSplitA - 1, nSplitB - 1));
			if (h->fastModeRule == 1 && h->fastMode)
				maxlevel = imin(nSplitA - 1 - (h->reproMode == 0), nSplitB - 1 - (h->reproMode == 0));
			const int sumOk = (h->sumMode < 2 || h->sumMode == 30);
			if (n == 1 && m == 1 && h->fastMode == 0 && sumOk) { /* DOT :209-215 */
				gemm_tn_s(nSplitA, nSplitB, k, AS, ldas, BS, ldbs, CS, nSplitA);
				ic++;
			} else if (n == 1 && h->fastMode == 0 && sumOk) { /* GEMV :216-225 */
				for (int32_t ia = 0; ia < imin(maxlevel + 1, nSplitA); ia++) {
					const int32_t numB = imin(nSplitB, maxlevel + 1 - ia);
					gemm_tn_s(mbk_, numB, k, AS + (int64_t)ldas * mbk_ * ia, ldas, BS, ldbs,
					          CS + (int64_t)ldcs * numB * ic, ldcs);
					ic++;
				}
			} else { /* GEMM :226-264 */
				if (!sumOk) return -3;
				for (int32_t ik = 0; ik <= maxlevel; ik++)
					for (int32_t ia = 0; ia < nSplitA; ia++)
						for (int32_t ib = 0; ib < nSplitB; ib++)
							if (ik == ia + ib) {
								gemm_tn_s(mbk_, nbk_, k, AS + (int64_t)ldas * mbk_ * ia, ldas,
								          BS + (int64_t)ldbs * nbk_ * ib, ldbs,
								          CS + (int64_t)ldcs * nbk_ * ic, ldcs);
								ic++;
							}
			}
			if (globalSum) { /* :277-298 */
				float *Cblk = &C[(int64_t)ldc * (in * nbk) + (int64_t)im * mbk];
				if (m == 1 && n == 1) {
					const int32_t sumorder = (h->fastMode == 0) ? 2 : 1;
					global_sum_s(h, 1, 1, nSplitA, nSplitB, CS, 1, 1, Cblk, ldc, alpha, beta, maxlevel,
					             sumorder);
				} else {
					const int32_t sumorder = (h->fastMode == 0 && nbk_ == 1) ? 3 : 1;
					global_sum_s(h, mbk_, nbk_, nSplitA, nSplitB, CS, (int64_t)ldcs * nbk_, ldcs, Cblk,
					             ldc, alpha, beta, maxlevel, sumorder);
				}
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

float ozo_s_dot(ozo_handle *h, int32_t n, const float *x, const float *y) { /* ozblas_dot.cpp */
	if (h->reproMode == 0 && h->nSplitMax == 1) {
		float s = 0.f;
		for (int32_t i = 0; i < n; i++) s += x[i] * y[i];
		return s;
	}
	float ret = 0.f; /* the reference leaves ret uninitialised (:25); it is the beta*C operand */
	if (ozo_s_gemm(h, 't', 'n', 1, 1, n, 1.f, x, n, y, n, 0.f, &ret, 1) != 0) return NAN;
	return ret;
}

float ozo_s_nrm2(ozo_handle *h, int32_t n, const float *x) { /* ozblas_nrm2.cpp:4-31 */
	if (h->reproMode == 0 && h->nSplitMax == 1) return (float)sqrt(ozo_s_dot(h, n, x, x));
	counter_init(h);
	const float ret = ozo_s_dot(h, n, x, x);
	return (float)sqrt(ret);
}

int32_t ozo_s_gemv(ozo_handle *h, char trans, int32_t m, int32_t n, float alpha, const float *A,
                   int32_t lda, const float *x, float beta, float *y) { /* ozblas_gemv.cpp */
	if (is_n(trans)) return ozo_s_gemm(h, 'n', 'n', m, 1, n, alpha, A, lda, x, n, beta, y, m);
	return ozo_s_gemm(h, 't', 'n', n, 1, m, alpha, A, lda, x, m, beta, y, n);
}

int32_t ozo_s_axpy(ozo_handle *h, int32_t n, float alpha, const float *x, float *y) {
	/* ozblas_axpy.cpp:4-27 (the reproMode == 0 cblas_saxpy path as an unfused update) */
	if (h->reproMode == 0) {
		for (int32_t i = 0; i < n; i++) y[i] = y[i] + alpha * x[i];
		return 0;
	}
	for (int32_t i = 0; i < n; i++) y[i] = fma1_s(alpha, x[i], y[i]);
	return 0;
}

static void plain_csrmv_s(int32_t m, float alpha, const float *val, const int32_t *colind,
                          const int32_t *rowptr, const float *x, float beta, float *y) {
	for (int32_t r = 0; r < m; r++) {
		float tr = 0.f;
		for (int32_t i = rowptr[r]; i < rowptr[r + 1]; i++) tr = fma1_s(val[i], x[colind[i]], tr);
		if (beta == 0.)
			y[r] = alpha * tr;
		else
			y[r] = fma1_s(alpha, tr, (beta * y[r]));
	}
}

int32_t ozo_s_csrmv(ozo_handle *h, int32_t m, int32_t n, int32_t nnz, float alpha, const float *val,
                    const int32_t *colind, const int32_t *rowptr, const float *x, float beta,
                    float *y) {
	/* src/cpu/ozblas_csrmv.cpp:4-181 */
	if (h->reproMode == 0 && h->nSplitMax == 1) {
		plain_csrmv_s(m, alpha, val, colind, rowptr, x, beta, y);
		return 0;
	}
	if (h->splitEpsMode == 2) return -3;
	counter_init(h);
	int32_t L = (h->nSplitMax > 0) ? h->nSplitMax : NUM_SPLIT_DEFAULT_MAX;
	if (h->reproMode == 1) L++;
	int32_t ldas, ldase, ldbs, ldcs;
	if (h->memMaskSplitA != 0) h->memAddr = 0;
	const uint64_t oAS = mat_alloc(h, nnz, L, 4, &ldas);
	const uint64_t oAT = vec_alloc(h, nnz, 4);
	const uint64_t oAmax = vec_alloc(h, m, 4);
	mat_alloc(h, m, L, 2, &ldase);
	if (h->memMaskSplitA != 0) h->memMaskSplitA = h->memAddr;
	const uint64_t oBS = mat_alloc(h, n, L, 4, &ldbs);
	const uint64_t oCS = mat_alloc(h, m, L * L, 4, &ldcs);
	const uint64_t oBT = vec_alloc(h, n, 4);
	vec_alloc(h, m, 4);
	if (h->sumMode == 3) {
		vec_alloc(h, m, 8);
		vec_alloc(h, m, 8);
		vec_alloc(h, m, 8);
	}
	vec_alloc(h, L, 2);
	const uint64_t oBmax = vec_alloc(h, 1, 4);
	if (mem_check(h)) {
		fprintf(stderr, "ozoracle: memory shortage\n");
		return -1;
	}
	arena_check(h);
	float *AS;
	int32_t nSplitA;
	if (h->memMaskSplitA == 0) {
		AS = WP(h, oAS, float);
		nSplitA = split_sparse_s(h, m, val, rowptr, WP(h, oAT, float), AS, ldas, WP(h, oAmax, float));
	} else {
		AS = (float *)val;
		nSplitA = h->nSplitA_;
	}
	float *BS = WP(h, oBS, float), *CS = WP(h, oCS, float);
	const int32_t nSplitB = ozo_s_split(h, n, 1, x, n, WP(h, oBT, float), n, BS, ldbs, WP(h, oBmax, float));
	int32_t maxlevel = (nSplitA - 1) + (nSplitB - 1);
	maxlevel = imax(maxlevel - h->fastMode, imin(nSplitA - 1, nSplitB - 1));
	float *pC = CS;
	int32_t ic = 0;
	for (int32_t ia = 0; ia < imin(maxlevel + 1, nSplitA); ia++) {
		const int32_t numB = imin(nSplitB, maxlevel + 1 - ia);
		csrmm_n_s(m, numB, 1.f, AS + (int64_t)ldas * ia, colind, rowptr, BS, ldbs, 0.f, pC, ldcs);
		pC += (int64_t)ldcs * numB;
		ic += numB;
	}
	h->nSplitA += nSplitA;
	h->nSplitB += nSplitB;
	h->nSplitC += ic;
	if (h->sumMode == 3) return -3;
	return global_sum_s(h, m, 1, nSplitA, nSplitB, CS, ldcs, ldcs, y, 1, alpha, beta, maxlevel, 3);
}

float *ozo_s_csrmv_splitA(ozo_handle *h, int32_t m, int32_t n, int32_t nnz, const float *val,
                          const int32_t *rowptr) {
	/* src/cpu/ozblas_csrmv.cpp:191-271 */
	(void)n;
	if (h->reproMode == 0 && h->nSplitMax == 1) return (float *)val;
	counter_init(h);
	int32_t L = (h->nSplitMax > 0) ? h->nSplitMax : NUM_SPLIT_DEFAULT_MAX;
	if (h->reproMode == 1) L++;
	int32_t ldas, ldase;
	const uint64_t oAS = mat_alloc(h, nnz, L, 4, &ldas);
	const uint64_t oAT = vec_alloc(h, nnz, 4);
	const uint64_t oAmax = vec_alloc(h, m, 4);
	mat_alloc(h, m, L, 2, &ldase);
	if (mem_check(h)) {
		fprintf(stderr, "ozoracle: memory shortage\n");
		return NULL;
	}
	arena_check(h);
	h->memMaskSplitA = h->memAddr;
	float *AS = WP(h, oAS, float), *AT = WP(h, oAT, float), *Amax = WP(h, oAmax, float);
	int32_t nA = split_sparse_s(h, m, val, rowptr, AT, AS, ldas, Amax);
	if (h->nSplitMax == 0) {
		int32_t nAOld;
		h->splitShift = 1;
		do {
			nAOld = nA;
			h->splitShift *= 2;
			nA = split_sparse_s(h, m, val, rowptr, AT, AS, ldas, Amax);
		} while (nAOld == nA && h->splitShift < 512);
		if (nAOld == nA) h->splitShift = 1;
		nA = split_sparse_s(h, m, val, rowptr, AT, AS, ldas, Amax);
	}
	h->nSplitA_ = nA;
	h->nSplitA = (float)nA;
	return AS;
}

int32_t ozo_s_cg(ozo_handle *h, int32_t n, int32_t nnz, const float *val, const int32_t *colind,
                 const int32_t *rowptr, const float *b, float *x, int32_t maxiter, float tol) {
	/* src/cpu/ozblas_cg.cpp:24-128 for TYPE1 = TYPE2 = float (trueres diagnostics not restated);
	   the history (resi / nrmb, a float) is stored widened into cg_verbose1 */
	float alpha, beta, resi, nrmb, tmp, dnew, dold;
	float *R = (float *)malloc(sizeof(float) * (size_t)n);
	float *P = (float *)malloc(sizeof(float) * (size_t)n);
	float *Q = (float *)calloc((size_t)n, sizeof(float)); /* read as 0 * q by the first product */
	h->cg_numiter = 0;
	float *AS = ozo_s_csrmv_splitA(h, n, n, nnz, val, rowptr);
	nrmb = ozo_s_nrm2(h, n, b);
	memcpy(R, b, sizeof(float) * (size_t)n);
	ozo_s_csrmv(h, n, n, nnz, -1.f, AS, colind, rowptr, x, 1.f, R);
	dold = ozo_s_dot(h, n, R, R);
	resi = (float)sqrt(dold);
	if (h->verbose > 0 && h->cg_verbose1) h->cg_verbose1[0] = resi / nrmb;
	memcpy(P, R, sizeof(float) * (size_t)n);
	int32_t vcnt = 1;
	while (h->cg_numiter < maxiter) {
		h->cg_numiter++;
		ozo_s_csrmv(h, n, n, nnz, 1.f, AS, colind, rowptr, P, 0.f, Q);
		tmp = ozo_s_dot(h, n, P, Q);
		alpha = dold / tmp;
		ozo_s_axpy(h, n, alpha, P, x);
		ozo_s_axpy(h, n, -alpha, Q, R);
		dnew = ozo_s_dot(h, n, R, R);
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
	return 0;
}

/* mublasInitMat of the binary32 harness (testing/testing_common.cpp:664-763 with FP_TYPE float):
   the expressions are evaluated in double and rounded to float on assignment; in erange mode
   f_pi = F_PI = M_PI is a float, so the sign factor carries (double)M_PI / (float)M_PI */
void ozo_s_init_mat(int32_t m, int32_t n, int32_t ld, float *mat, double phi, int32_t mode,
                    int32_t seed) {
	if (ld < m) ld = m;
	ozo_srand48(seed);
	switch (mode) {
	case 0:
		for (int32_t j = 0; j < n; j++)
			for (int32_t i = 0; i < m; i++) mat[(int64_t)j * ld + i] = (float)phi;
		break;
	case 1:
		for (int32_t j = 0; j < n; j++)
			for (int32_t i = 0; i < m; i++) mat[(int64_t)j * ld + i] = (float)ozo_drand48();
		break;
	case 2:
		for (int32_t j = 0; j < n; j++)
			for (int32_t i = 0; i < m; i++) {
				const double r1 = ozo_drand48();
				const double r2 = ozo_drand48();
				const double x1 = 0.0 + (sqrt(-2. * log(r1)) * sin(2. * M_PI * r2)) * 1.0;
				mat[(int64_t)j * ld + i] = (float)((r1 - 0.5) * exp(phi * x1));
			}
		break;
	case 3: {
		const float f_pi = (float)M_PI;
		for (int32_t j = 0; j < n; j++)
			for (int32_t i = 0; i < m; i++) {
				const double d = ozo_drand48();
				const int32_t e = ozo_rand() % (int32_t)phi;
				float valf = (float)((d * 9 + 1) * pow(10, e));
				valf = (float)(((ozo_rand() % 2) ? 1. : -1.) * ((double)M_PI / f_pi) * valf);
				mat[(int64_t)j * ld + i] = valf;
			}
		break;
	}
	}
}
