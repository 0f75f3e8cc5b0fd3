/*
 * ozoracle.c -- CPU oracle (TEST INFRASTRUCTURE ONLY; see ozoracle.h for scope and pinning).
 *
 * Restatement of OzBLAS 1.6a src/cpu for binary64.  Compiled with gcc -O2 -ffp-contract=off and
 * no -march (as src/cpu/Makefile:7-36), so a*b+c is never contracted; fma() is called only where
 * the reference calls fma1().
 */
#define _GNU_SOURCE
#include "ozoracle_int.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/mman.h>

/* ---------------------------------------------------------------------------------------------
 * handle + workspace arena (src/cpu/ozblas.cpp:3-71, src/cpu/ozblas_aux.cpp:228-267)
 * ------------------------------------------------------------------------------------------- */
int32_t ozo_pitch(int32_t n) { return (int32_t)(ceilf((float)n / 8.0f) * 8.0f); }

void ozo_create(ozo_handle *h, uint64_t workSizeBytes) {
	memset(h, 0, sizeof(*h));
	h->workSizeBytes = workSizeBytes;
	h->nSplitMax = 0;
	h->memMaskSplitA = 0;
	h->splitShift = 1;
	h->splitMode = 0;
	h->fastMode = 0;
	h->reproMode = 1;
	h->sumMode = 0;
	h->useBatchedGemmFlag = 0;
	h->splitEpsMode = 0;
	h->precxFlag = 0;
	h->dot2Chunks = 1024;
	/* lazily-zeroed pages: the reference's aligned_alloc of a multi-GB workspace is served by
	   fresh (zero) pages too, so untouched slots read as +0 in both */
	void *p = mmap(NULL, workSizeBytes + 4096, PROT_READ | PROT_WRITE,
	               MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE, -1, 0);
	if (p == MAP_FAILED) {
		fprintf(stderr, "ozoracle: workspace mmap failed\n");
		exit(1);
	}
	h->work = (char *)p;
	h->workCap = workSizeBytes;
}

void ozo_destroy(ozo_handle *h) {
	if (h->work) munmap(h->work, h->workCap + 4096);
	h->work = NULL;
}

ozo_handle *ozo_handle_new(uint64_t workSizeBytes) {
	ozo_handle *h = (ozo_handle *)malloc(sizeof(ozo_handle));
	ozo_create(h, workSizeBytes);
	return h;
}
void ozo_handle_free(ozo_handle *h) {
	ozo_destroy(h);
	free(h);
}


/* ---------------------------------------------------------------------------------------------
 * split (src/cpu/ozblas_split.cpp)
 * ------------------------------------------------------------------------------------------- */
int32_t ozo_rho(int32_t dim, int32_t splitEpsMode) { /* :20-34 with TYPE1=TYPE2=double */
	double r;
	switch (splitEpsMode) {
	case 2: r = ceil((53 - (106 - log2(1. * dim)) / 2)); break;
	case 9: r = ceil((53 - (53 - log2(2. * sqrt(dim))) / 2)); break;
	default: r = ceil((53 - (53 - log2(1. * dim)) / 2)); break;
	}
	return d2i(r);
}

int32_t ozo_tau(double max) { /* (short)ceil(log2(fabs(max))) :86,158,808 */
	double c = ceil(log2(fabs(max)));
	return (int32_t)(int16_t)d2i(c);
}

void ozo_tau_batch(int64_t n, const double *x, int32_t *out) {
	for (int64_t i = 0; i < n; i++) out[i] = ozo_tau(x[i]);
}

static double sigma_of(int32_t rho, int32_t tau, int32_t shift) { /* :88,159 */
	return SPLIT_CONST * scalbn_l(1., (int64_t)rho + tau) * shift;
}

/* ozblasFindMaxDevice :55-73 (column-major only) */
static void find_max(int32_t m, int32_t n, const double *in, int32_t ldi, double *max) {
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
		for (int32_t j = 0; j < n; j++) { /* ozblasFindMaxCKernel :37-52 */
			double mx = 0.;
			for (int32_t i = 0; i < m; i++) mx = std_max(mx, fabs(in[(int64_t)j * ldi + i]));
			max[j] = mx;
		}
	}
}

/* ozblasSplitVecKernel (same type) :75-104 */
static double split_vec_kernel(int32_t n, int32_t rho, const double *in, double *out, double *split,
                               double *max, int32_t shift) {
	const int32_t tau = ozo_tau(max[0]);
	const double sigma = sigma_of(rho, tau, shift);
	double mx = 0.;
	for (int32_t i = 0; i < n; i++) {
		double input = in[i];
		const double sp = (input + sigma) - sigma;
		input = input - sp;
		split[i + n] = input; /* remainder slot at +n, not +lds (:96) */
		split[i] = sp;
		out[i] = input;
		mx = std_max(mx, fabs(input));
	}
	max[0] = mx;
	return mx;
}

/* ozblasSplitCKernel (same type) :137-179 */
static double split_c_kernel(int32_t m, int32_t n, int32_t rho, const double *in, int32_t ldi,
                             double *out, int32_t ldo, double *split, int32_t lds, double *max,
                             int32_t shift) {
	double check = 0.;
	/* OpenMP over the columns as the reference (:156); check is only compared with 0 */
#pragma omp parallel for schedule(static) reduction(+ : check)
	for (int32_t j = 0; j < n; j++) {
		const int32_t tau = ozo_tau(max[j]);
		const double sigma = sigma_of(rho, tau, shift);
		double mx = 0.;
		for (int32_t i = 0; i < m; i++) {
			double input = in[(int64_t)j * ldi + i];
			const double sp = (input + sigma) - sigma;
			input = input - sp;
			split[(int64_t)j * lds + i + (int64_t)lds * n] = input; /* :166 */
			split[(int64_t)j * lds + i] = sp;
			out[(int64_t)j * ldo + i] = input;
			mx = std_max(mx, fabs(input));
		}
		max[j] = mx;
		check += mx; /* omp critical sum, only compared with 0 */
	}
	return check;
}

/* ozblasSplitDevice :224-252 */
static double split_device(int32_t m, int32_t n, const double *in, int32_t ldi, double *out,
                           int32_t ldo, double *split, int32_t lds, double *max, int32_t epsMode,
                           int32_t shift) {
	const int32_t rho = ozo_rho(m, epsMode); /* dim = m for 'c' */
	if (n == 1) return split_vec_kernel(m, rho, in, out, split, max, shift);
	return split_c_kernel(m, n, rho, in, ldi, out, ldo, split, lds, max, shift);
}

/* ozblasSplit :254-289 */
int32_t ozo_split(ozo_handle *h, int32_t m, int32_t n, const double *in, int32_t ldi, double *out,
                  int32_t ldo, double *split, int32_t lds, double *max) {
	int32_t s;
	find_max(m, n, in, ldi, max);
	split_device(m, n, in, ldi, out, ldo, split, lds, max, h->splitEpsMode, h->splitShift);
	int32_t maxS = (h->nSplitMax > 0) ? h->nSplitMax : NUM_SPLIT_DEFAULT_MAX;
	if (h->reproMode == 0) maxS--;
	for (s = 1; s < maxS; s++) {
		if (split_device(m, n, out, ldo, out, ldo, &split[(int64_t)lds * n * s], lds, max,
		                 h->splitEpsMode, h->splitShift) == 0.) {
			s++;
			break;
		}
	}
	if (h->reproMode == 0) s++;
	return s;
}

/* sparse row split :756-928 */
static void find_max_sparse(int32_t m, const double *in, const int32_t *rowptr, double *max) {
	for (int32_t r = 0; r < m; r++) {
		double mx = 0.;
		for (int32_t i = rowptr[r]; i < rowptr[r + 1]; i++) mx = std_max(mx, fabs(in[i]));
		max[r] = mx;
	}
}

static double split_sparse_kernel(int32_t m, const double *in, const int32_t *rowptr, double *out,
                                  double *split, int32_t lds, double *max, int32_t epsMode,
                                  int32_t shift) {
	double check = 0.;
	for (int32_t r = 0; r < m; r++) {
		const int32_t dim = rowptr[r + 1] - rowptr[r];
		const int32_t rho = ozo_rho(dim, epsMode);
		const int32_t tau = ozo_tau(max[r]);
		const double sigma = SPLIT_CONST * scalbn_l(1., (int64_t)rho + tau) / shift; /* :809 */
		double mx = 0.;
		for (int32_t i = rowptr[r]; i < rowptr[r + 1]; i++) {
			double input = in[i];
			const double sp = (input + sigma) - sigma;
			input = input - sp;
			split[i + lds] = input;
			split[i] = sp;
			out[i] = input;
			mx = std_max(mx, fabs(input));
		}
		max[r] = mx;
		check += mx;
	}
	return check;
}

static int32_t split_sparse(ozo_handle *h, int32_t m, const double *in, const int32_t *rowptr,
                            double *out, double *split, int32_t lds, double *max) {
	int32_t s;
	find_max_sparse(m, in, rowptr, max);
	split_sparse_kernel(m, in, rowptr, out, split, lds, max, h->splitEpsMode, h->splitShift);
	int32_t maxS = (h->nSplitMax > 0) ? h->nSplitMax : NUM_SPLIT_DEFAULT_MAX;
	if (h->reproMode == 0) maxS--;
	for (s = 1; s < maxS; s++)
		if (split_sparse_kernel(m, out, rowptr, out, &split[(int64_t)lds * s], lds, max,
		                        h->splitEpsMode, h->splitShift) == 0.) {
			s++;
			return s; /* early return skips the reproMode==0 increment (:920-922) */
		}
	if (h->reproMode == 0) s++;
	return s;
}

/* ---------------------------------------------------------------------------------------------
 * summation (src/cpu/ozblas_sum.cpp)
 * ------------------------------------------------------------------------------------------- */
static void transform(int32_t n, double *vec, int32_t ld, double rho, double *tau1, double *tau2) {
	/* :10-56 */
	double mu = fabs(vec[0]);
	for (int32_t i = 1; i < n; i++) mu = mac_max(mu, fabs(vec[(int64_t)i * ld]));
	if ((n == 0) || (mu == 0.)) {
		*tau1 = rho;
		*tau2 = 0.;
		return;
	}
	const int32_t m = d2i(ceil(log2((double)(n + 2))));
	double sigma = scalbn_l(1., (int64_t)d2i(m + ceil(log2(mu))));
	double t = rho;
	while (1) {
		double tau = 0.;
		for (int32_t i = 0; i < n; i++) {
			const double tmp = (sigma + vec[(int64_t)i * ld]) - sigma;
			vec[(int64_t)i * ld] -= tmp;
			tau += tmp;
		}
		*tau1 = t + tau;
		if ((sigma <= scalbn(1., -1022)) || (fabs(*tau1) >= scalbn_l(1., 2 * m + 1 - 53) * sigma)) {
			*tau2 = tau - (*tau1 - t);
			return;
		}
		sigma = scalbn_l(1., m - 53) * sigma;
		t = *tau1;
	}
}

static void transform_k(int32_t n, double *vec, int32_t ld, double rho, double *res, double *r) {
	/* :58-73 */
	double tmp = 0., tau1, tau2;
	transform(n, vec, ld, rho, &tau1, &tau2);
	for (int32_t i = 0; i < n; i++) tmp += vec[(int64_t)i * ld];
	*res = tau1 + (tau2 + tmp);
	*r = tau2 - (*res - tau1);
}

static inline double get_sign(double v) { return (v < 0) ? -1. : 1.; }

double ozo_nearsum(int32_t n, double *vec, int32_t ld) { /* :81-112 */
	double tmp, res, res2, r, r2, mu, delta, delta2;
	const double eps = scalbn(1., -53);
	transform_k(n, vec, ld, 0., &res, &r);
	transform_k(n, vec, ld, r, &delta, &r2);
	if (delta == 0) return res;
	res2 = res + get_sign(delta) * eps * fabs(res);
	if (res2 == res) {
		mu = get_sign(delta) * eps * fabs(res);
		res2 = res + 2. * get_sign(delta) * eps * fabs(res);
	} else {
		mu = (res2 - res) / 2.;
	}
	if (fabs(delta) < fabs(mu)) return res;
	if (fabs(delta) > fabs(mu)) return res2;
	transform_k(n, vec, ld, r2, &delta2, &tmp);
	if (delta2 == 0) return res + mu;
	if (get_sign(delta2) == get_sign(mu)) return res2;
	return res;
}

/* ozblasGlobalNearsumKernel (1-type) :231-271 */
static void global_nearsum(int32_t m, int32_t n, int32_t nA, int32_t nB, double *Cs, int64_t llsc,
                           int32_t ldsc, double *C, int32_t ldc, double alpha, double beta,
                           int32_t maxlevel) {
	int32_t ic = 0;
	for (int32_t ik = 0; ik <= maxlevel; ik++)
		for (int32_t ia = 0; ia < nA; ia++)
			for (int32_t ib = 0; ib < nB; ib++)
				if (ik == ia + ib) ic++;
#pragma omp parallel for schedule(static)
	for (int32_t y = 0; y < n; y++)
		for (int32_t x = 0; x < m; x++) {
			double t = ozo_nearsum(ic, &Cs[(int64_t)y * ldsc + x], (int32_t)llsc);
			double *c = &C[(int64_t)y * ldc + x];
			*c = fma(alpha, t, (beta * *c));
		}
}

/* ozblasGlobalFsumKernel (1-type) :331-380 */
static void global_fsum(int32_t m, int32_t n, int32_t nA, int32_t nB, const double *Cs, int64_t llsc,
                        int32_t ldsc, double *C, int32_t ldc, double alpha, double beta,
                        int32_t maxlevel, int32_t sumOrder) {
#pragma omp parallel for schedule(static)
	for (int32_t y = 0; y < n; y++)
		for (int32_t x = 0; x < m; x++) {
			double t = 0.;
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
			double *c = &C[(int64_t)y * ldc + x];
			*c = fma(alpha, t, (beta * *c));
		}
}

/* ozblasGlobalSum :513-552 (sumMode 1 NearSum, otherwise Fsum; Sum3 out of scope) */
static int32_t global_sum(ozo_handle *h, int32_t m, int32_t n, int32_t nA, int32_t nB, double *Cs,
                          int64_t llsc, int32_t ldsc, double *C, int32_t ldc, double alpha,
                          double beta, int32_t maxlevel, int32_t sumOrder) {
	if (h->sumMode == 1)
		global_nearsum(m, n, nA, nB, Cs, llsc, ldsc, C, ldc, alpha, beta, maxlevel);
	else if (h->sumMode == 30)
		return -1;
	else
		global_fsum(m, n, nA, nB, Cs, llsc, ldsc, C, ldc, alpha, beta, maxlevel, sumOrder);
	return 0;
}

/* ---------------------------------------------------------------------------------------------
 * slice-product backends (exact by construction, so any summation order gives the same bits)
 * ------------------------------------------------------------------------------------------- */
/* An optional CBLAS dgemm (cblas_dgemm's signature) for the slice products, as the reference
 * calls one (src/cpu/ozblas_blas_wrapper.cpp:158-168); NULL selects the loop below.  The products
 * are exact, so the bits are the same either way: only the time differs. */
typedef void (*ozo_cblas_dgemm_fn)(int order, int transA, int transB, int M, int N, int K,
                                   double alpha, const double *A, int lda, const double *B, int ldb,
                                   double beta, double *C, int ldc);
static ozo_cblas_dgemm_fn g_cblas_dgemm = NULL;
void ozo_set_cblas_dgemm(void *fn) { g_cblas_dgemm = (ozo_cblas_dgemm_fn)fn; }

/* C(M x N, ldc) = A^T B with A (K x M, lda), B (K x N, ldb); beta = 0 (blasRgemm 't','n') */
static void gemm_tn(int32_t M, int32_t N, int32_t K, const double *A, int32_t lda, const double *B,
                    int32_t ldb, double *C, int32_t ldc) {
	if (g_cblas_dgemm && M > 0 && N > 0 && K > 0) {
		g_cblas_dgemm(102 /* CblasColMajor */, 112 /* CblasTrans */, 111 /* CblasNoTrans */, M, N, K,
		              1.0, A, lda, B, ldb, 0.0, C, ldc);
		return;
	}
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

/* cblas_domatcopy(ColMajor, Trans): dst (cols x rows, ldd) = src^T (rows x cols, lds) */
static void omatcopy_t(int32_t rows, int32_t cols, const double *src, int32_t lds, double *dst,
                       int32_t ldd) {
	for (int32_t j = 0; j < cols; j++)
		for (int32_t i = 0; i < rows; i++) dst[(int64_t)i * ldd + j] = src[(int64_t)j * lds + i];
}

/* csrmm_n src/cpu/ozblas_blas_wrapper.cpp:231-260 (the l2..l5 variants compute the same) */
static void csrmm_n(int32_t m, int32_t n, double alpha, const double *val, const int32_t *colind,
                    const int32_t *rowptr, const double *B, int32_t ldb, double beta, double *C,
                    int32_t ldc) {
#pragma omp parallel for schedule(static)
	for (int32_t r = 0; r < m; r++)
		for (int32_t j = 0; j < n; j++) {
			double tr = 0.;
			for (int32_t i = rowptr[r]; i < rowptr[r + 1]; i++)
				tr = fma(val[i], B[(int64_t)j * ldb + colind[i]], tr);
			if (beta == 0.)
				C[(int64_t)j * ldc + r] = alpha * tr;
			else
				C[(int64_t)j * ldc + r] = fma(alpha, tr, (beta * C[(int64_t)j * ldc + r]));
		}
}

/* ---- Dot2: error-free transformations, scalar forms of src/cpu/eft.h:8-182 ---------------- */
typedef struct {
	double x, y;
} dd_t;

static void two_sum(double a, double b, double *s, double *e) { /* eft.h TwoSum */
	double v;
	*s = a + b;
	v = *s - a;
	*e = (a - (*s - v)) + (b - v);
}
static void quick_two_sum(double a, double b, double *s, double *e) { /* eft.h QuickTwoSum */
	*s = a + b;
	*e = b - (*s - a);
}
static void two_prod(double a, double b, double *p, double *e) { /* eft.h TwoProdFMA */
	*p = a * b;
	*e = fma(a, b, -*p);
}
static void dot2i(double a, double b, dd_t *c) { /* eft.h Dot2i */
	double h, r, q;
	two_prod(a, b, &h, &r);
	two_sum(c->x, h, &c->x, &q);
	c->y = c->y + (q + r);
}
static dd_t quad_add(dd_t a, dd_t b) { /* eft.h QuadAdd */
	dd_t t, c;
	two_sum(a.x, b.x, &t.x, &t.y);
	t.y = t.y + (a.y + b.y);
	quick_two_sum(t.x, t.y, &c.x, &c.y);
	return c;
}
static dd_t quad_mul_dd(double a, dd_t b) { /* eft.h QuadMul(double, double2) */
	dd_t t, c;
	two_prod(a, b.x, &t.x, &t.y);
	t.y = fma(a, b.y, t.y);
	quick_two_sum(t.x, t.y, &c.x, &c.y);
	return c;
}

/* libgomp's static partition of [0, n) over T threads: thread t gets q (+1 for t < r) iterations */
static void static_chunk(int64_t n, int32_t T, int32_t t, int64_t *lo, int64_t *hi) {
	const int64_t q = n / T, r = n % T;
	*lo = t * q + (t < r ? t : r);
	*hi = *lo + q + (t < r ? 1 : 0);
}

/* dgemm_tn_skinny_x2 (src/cpu/ozblas_blas_wrapper.cpp:109-152), alpha = 1, beta = 0, flag = 1:
   C[nn*ldc + mm] = hi, C[nn*ldc + mm + n*ldc] = lo of the Dot2 sum of column mm of A and column nn
   of B; T per-thread partials combined in thread order (see ozo_handle.dot2Chunks) */
static void gemm_tn_x2(int32_t m, int32_t n, int32_t k, const double *A, int32_t lda, const double *B,
                       int32_t ldb, double *C, int32_t ldc, int32_t T) {
	if (T < 1) T = 1;
	for (int32_t nn = 0; nn < n; nn++)
		for (int32_t mm = 0; mm < m; mm++) {
			dd_t G = {0., 0.};
			for (int32_t t = 0; t < T; t++) {
				int64_t lo, hi;
				static_chunk(k, T, t, &lo, &hi);
				dd_t L = {0., 0.};
				for (int64_t kk = lo; kk < hi; kk++)
					dot2i(A[(int64_t)mm * lda + kk], B[(int64_t)nn * ldb + kk], &L);
				G = quad_add(G, L);
			}
			G = quad_mul_dd(1., G);
			C[(int64_t)nn * ldc + mm] = G.x;
			C[(int64_t)nn * ldc + mm + (int64_t)n * ldc] = G.y;
		}
}

/* csrmm_n_x2 (src/cpu/ozblas_blas_wrapper.cpp:262-295; the _l2.._l5 forms compute the same
   sequence per row), alpha = 1, beta = 0, flag = 1: per row and x slice j, Dot2 over the row's
   nonzeros in order; C[j*ldc + row] = hi, C[j*ldc + row + ldc*n] = lo */
static void csrmm_n_x2(int32_t m, int32_t n, const double *val, const int32_t *colind,
                       const int32_t *rowptr, const double *B, int32_t ldb, double *C, int32_t ldc) {
#pragma omp parallel for schedule(static)
	for (int32_t r = 0; r < m; r++)
		for (int32_t j = 0; j < n; j++) {
			dd_t Tr = {0., 0.};
			for (int32_t i = rowptr[r]; i < rowptr[r + 1]; i++)
				dot2i(val[i], B[(int64_t)j * ldb + colind[i]], &Tr);
			Tr = quad_mul_dd(1., Tr);
			C[(int64_t)j * ldc + r] = Tr.x;
			C[(int64_t)j * ldc + r + (int64_t)ldc * n] = Tr.y;
		}
}

/* test hooks for the EFT primitives */
void ozo_dd_dot2i(double a, double b, double *cx, double *cy) {
	dd_t c = {*cx, *cy};
	dot2i(a, b, &c);
	*cx = c.x;
	*cy = c.y;
}
void ozo_dd_quad_add(double ax, double ay, double bx, double by, double *cx, double *cy) {
	dd_t a = {ax, ay}, b = {bx, by};
	const dd_t c = quad_add(a, b);
	*cx = c.x;
	*cy = c.y;
}
void ozo_dd_quad_mul(double a, double bx, double by, double *cx, double *cy) {
	dd_t b = {bx, by};
	const dd_t c = quad_mul_dd(a, b);
	*cx = c.x;
	*cy = c.y;
}
void ozo_dot2_chunked(int32_t n, const double *x, const double *y, int32_t T, double *hi, double *lo) {
	double C2[2];
	gemm_tn_x2(1, 1, n, x, n, y, n, C2, 1, T);
	*hi = C2[0];
	*lo = C2[1];
}

/* ---------------------------------------------------------------------------------------------
 * drivers
 * ------------------------------------------------------------------------------------------- */

int32_t ozo_gemm(ozo_handle *h, char transA, char transB, int32_t m, int32_t n, int32_t k,
                 double alpha, const double *A, int32_t lda, const double *B, int32_t ldb,
                 double beta, double *C, int32_t ldc) {
	/* src/cpu/ozblas_gemm.cpp:3-315 */
	if (h->reproMode == 0 && h->nSplitMax == 1) { /* bypass :14-17: plain (non-reproducible) */
		for (int32_t j = 0; j < n; j++)
			for (int32_t i = 0; i < m; i++) {
				double s = 0.;
				for (int32_t kk = 0; kk < k; kk++) {
					double a = is_n(transA) ? A[i + (int64_t)kk * lda] : A[kk + (int64_t)i * lda];
					double b = is_n(transB) ? B[kk + (int64_t)j * ldb] : B[j + (int64_t)kk * ldb];
					s += a * b;
				}
				C[i + (int64_t)j * ldc] = alpha * s + beta * C[i + (int64_t)j * ldc];
			}
		return 0;
	}
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
	int32_t ldas = 0, ldbs = 0, ldcs = 0, ldt = 0, ldct = 0, ldase = 0, ldbse = 0, dummy;
	uint64_t oAS = 0, oBS = 0, oCS = 0, oT = 0, oMax = 0;
	const uint64_t memAddrTmp = h->memAddr;
	while (mbk > 0 && nbk > 0) { /* fit loop :59-97 */
		int32_t sizeCn = globalSum ? (nbk * L * L) : nbk;
		if (h->splitEpsMode == 2) sizeCn *= 2;
		oAS = mat_alloc(h, k, mbk * L, 8, &ldas);
		oBS = mat_alloc(h, k, nbk * L, 8, &ldbs);
		oCS = mat_alloc(h, mbk, sizeCn, 8, &ldcs);
		oT = mat_alloc(h, k, imax(mbk, nbk), 8, &ldt);
		mat_alloc(h, mbk, nbk, 8, &ldct);
		if (h->sumMode == 3 && h->useBatchedGemmFlag == 0) {
			mat_alloc(h, mbk, nbk, 8, &dummy);
			mat_alloc(h, mbk, nbk, 8, &dummy);
			mat_alloc(h, mbk, nbk, 8, &dummy);
		}
		mat_alloc(h, mbk, L, 2, &ldase);
		mat_alloc(h, nbk, L, 2, &ldbse);
		oMax = vec_alloc(h, imax(k, imax(mbk, nbk)), 8);
		if (h->splitMode == 3) {
			vec_alloc(h, imax(k, imax(mbk, nbk)), 8);
			mat_alloc(h, k, imax(mbk, nbk), 8, &ldt);
			mat_alloc(h, k, imax(mbk, nbk), 8, &ldt);
			mat_alloc(h, k, imax(mbk, nbk), 8, &ldt);
		}
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
	double *Tmp = WP(h, oT, double), *Max1 = WP(h, oMax, double);
	(void)ldct;
	(void)ldase;
	(void)ldbse;

	int32_t block_count = 0;
	const int32_t nbm = (int32_t)ceilf((float)m / mbk), nbn = (int32_t)ceilf((float)n / nbk);
	for (int32_t im = 0; im < nbm; im++) {
		const int32_t mbk_ = (m - mbk * im >= mbk) ? mbk : m - mbk * im;
		int32_t nSplitA;
		if (is_n(transA)) {
			omatcopy_t(mbk_, k, A + (int64_t)im * mbk, lda, Tmp, ldt);
			nSplitA = ozo_split(h, k, mbk_, Tmp, ldt, Tmp, ldt, AS, ldas, Max1);
		} else {
			nSplitA = ozo_split(h, k, mbk_, A + (int64_t)im * mbk * lda, lda, Tmp, ldt, AS, ldas, Max1);
		}
		for (int32_t in = 0; in < nbn; in++) {
			const int32_t nbk_ = (n - nbk * in >= nbk) ? nbk : n - nbk * in;
			int32_t nSplitB;
			if (is_n(transB)) {
				nSplitB = ozo_split(h, k, nbk_, B + (int64_t)in * nbk * ldb, ldb, Tmp, ldt, BS, ldbs, Max1);
			} else {
				omatcopy_t(nbk_, k, B + (int64_t)in * nbk, ldb, Tmp, ldt);
				nSplitB = ozo_split(h, k, nbk_, Tmp, ldt, Tmp, ldt, BS, ldbs, Max1);
			}
			int32_t ic = 0;
			int32_t maxlevel = (nSplitA - 1) + (nSplitB - 1);
			maxlevel = imax(maxlevel - h->fastMode, imin(nSplitA - 1, nSplitB - 1));
			if (h->fastModeRule == 1 && h->fastMode) /* legacy rule, see ozo_handle.fastModeRule */
				maxlevel = imin(nSplitA - 1 - (h->reproMode == 0), nSplitB - 1 - (h->reproMode == 0));
			const int sumOk = (h->sumMode < 2 || h->sumMode == 30);
			const int dot2 = (n == 1 && m == 1 && h->splitEpsMode == 2 && h->fastMode == 0 && sumOk);
			if (dot2) { /* Dot2 DOT :166-173: (hi, lo) of every slice pair */
				gemm_tn_x2(nSplitA, nSplitB, k, AS, ldas, BS, ldbs, CS, nSplitA, h->dot2Chunks);
				ic++;
			} else if (n == 1 && m == 1 && h->fastMode == 0 && sumOk) { /* DOT :209-215 */
				gemm_tn(nSplitA, nSplitB, k, AS, ldas, BS, ldbs, CS, nSplitA);
				ic++;
			} else if (n == 1 && h->fastMode == 0 && sumOk) { /* GEMV :216-225 */
				for (int32_t ia = 0; ia < imin(maxlevel + 1, nSplitA); ia++) {
					const int32_t numB = imin(nSplitB, maxlevel + 1 - ia);
					gemm_tn(mbk_, numB, k, AS + (int64_t)ldas * mbk_ * ia, ldas, BS, ldbs,
					        CS + (int64_t)ldcs * numB * ic, ldcs);
					ic++;
				}
			} else { /* GEMM :226-264 (batched enumerates the same products) */
				if (!sumOk) return -3;
				for (int32_t ik = 0; ik <= maxlevel; ik++)
					for (int32_t ia = 0; ia < nSplitA; ia++)
						for (int32_t ib = 0; ib < nSplitB; ib++)
							if (ik == ia + ib) {
								gemm_tn(mbk_, nbk_, k, AS + (int64_t)ldas * mbk_ * ia, ldas,
								        BS + (int64_t)ldbs * nbk_ * ib, ldbs,
								        CS + (int64_t)ldcs * nbk_ * ic, ldcs);
								ic++;
							}
			}
			if (globalSum) { /* :277-298 */
				double *Cblk = &C[(int64_t)ldc * (in * nbk) + (int64_t)im * mbk];
				if (m == 1 && n == 1) {
					const int32_t sumorder = (h->fastMode == 0) ? 2 : 1;
					/* Dot2: the lo parts count as nSplitB more B slices (:281-285) */
					if (h->splitEpsMode == 2) maxlevel = (nSplitA - 1) + (nSplitB * 2 - 1);
					global_sum(h, 1, 1, nSplitA, nSplitB * ((h->splitEpsMode == 2) ? 2 : 1), CS, 1, 1,
					           Cblk, ldc, alpha, beta, maxlevel, sumorder);
				} else {
					const int32_t sumorder = (h->fastMode == 0 && nbk_ == 1) ? 3 : 1;
					global_sum(h, mbk_, nbk_, nSplitA, nSplitB, CS, (int64_t)ldcs * nbk_, ldcs, Cblk,
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

double ozo_dot(ozo_handle *h, int32_t n, const double *x, const double *y) { /* ozblas_dot.cpp */
	if (h->reproMode == 0 && h->nSplitMax == 1) {
		double s = 0.;
		for (int32_t i = 0; i < n; i++) s += x[i] * y[i];
		return s;
	}
	double ret = 0.; /* the reference leaves ret uninitialised (:25); it is the beta*C operand */
	ozo_gemm(h, 't', 'n', 1, 1, n, 1., x, n, y, n, 0., &ret, 1);
	return ret;
}

double ozo_nrm2(ozo_handle *h, int32_t n, const double *x) { /* ozblas_nrm2.cpp:4-31 */
	if (h->reproMode == 0 && h->nSplitMax == 1) return sqrt(ozo_dot(h, n, x, x));
	counter_init(h);
	double ret = ozo_dot(h, n, x, x);
	return sqrt(ret);
}

int32_t ozo_gemv(ozo_handle *h, char trans, int32_t m, int32_t n, double alpha, const double *A,
                 int32_t lda, const double *x, double beta, double *y) { /* ozblas_gemv.cpp */
	if (is_n(trans)) return ozo_gemm(h, 'n', 'n', m, 1, n, alpha, A, lda, x, n, beta, y, m);
	return ozo_gemm(h, 't', 'n', n, 1, m, alpha, A, lda, x, m, beta, y, n);
}

int32_t ozo_axpy(ozo_handle *h, int32_t n, double alpha, const double *x, double *y) {
	/* ozblas_axpy.cpp:4-27; the reproMode==0 cblas_daxpy path is restated as an unfused update */
	if (h->reproMode == 0) {
		for (int32_t i = 0; i < n; i++) y[i] = y[i] + alpha * x[i];
		return 0;
	}
	for (int32_t i = 0; i < n; i++) y[i] = fma(alpha, x[i], y[i]);
	return 0;
}

static void plain_csrmv(int32_t m, double alpha, const double *val, const int32_t *colind,
                        const int32_t *rowptr, const double *x, double beta, double *y) {
	/* csrmv src/cpu/ozblas_blas_wrapper.cpp:755-779 */
	for (int32_t r = 0; r < m; r++) {
		double tr = 0.;
		for (int32_t i = rowptr[r]; i < rowptr[r + 1]; i++) tr = fma(val[i], x[colind[i]], tr);
		if (beta == 0.)
			y[r] = alpha * tr;
		else
			y[r] = fma(alpha, tr, (beta * y[r]));
	}
}

int32_t ozo_csrmv(ozo_handle *h, int32_t m, int32_t n, int32_t nnz, double alpha, const double *val,
                  const int32_t *colind, const int32_t *rowptr, const double *x, double beta,
                  double *y) {
	/* src/cpu/ozblas_csrmv.cpp:4-181 */
	if (h->reproMode == 0 && h->nSplitMax == 1) {
		plain_csrmv(m, alpha, val, colind, rowptr, x, beta, y);
		return 0;
	}
	counter_init(h);
	int32_t L = (h->nSplitMax > 0) ? h->nSplitMax : NUM_SPLIT_DEFAULT_MAX;
	if (h->reproMode == 1) L++;
	int32_t ldas, ldase, ldbs, ldcs;
	if (h->memMaskSplitA != 0) h->memAddr = 0;
	const uint64_t oAS = mat_alloc(h, nnz, L, 8, &ldas);
	const uint64_t oAT = vec_alloc(h, nnz, 8);
	const uint64_t oAmax = vec_alloc(h, m, 8);
	mat_alloc(h, m, L, 2, &ldase);
	if (h->memMaskSplitA != 0) h->memMaskSplitA = h->memAddr;
	const uint64_t oBS = mat_alloc(h, n, L, 8, &ldbs);
	const uint64_t oCS = mat_alloc(h, m, L * L * ((h->splitEpsMode == 2) ? 2 : 1), 8, &ldcs);
	const uint64_t oBT = vec_alloc(h, n, 8);
	vec_alloc(h, m, 8);
	if (h->sumMode == 3) {
		vec_alloc(h, m, 8);
		vec_alloc(h, m, 8);
		vec_alloc(h, m, 8);
	}
	vec_alloc(h, L, 2);
	const uint64_t oBmax = vec_alloc(h, 1, 8);
	if (mem_check(h)) {
		fprintf(stderr, "ozoracle: memory shortage\n");
		return -1;
	}
	arena_check(h);
	double *AS;
	int32_t nSplitA;
	if (h->memMaskSplitA == 0) {
		AS = WP(h, oAS, double);
		nSplitA = split_sparse(h, m, val, rowptr, WP(h, oAT, double), AS, ldas, WP(h, oAmax, double));
	} else {
		AS = (double *)val;
		nSplitA = h->nSplitA_;
	}
	double *BS = WP(h, oBS, double), *CS = WP(h, oCS, double);
	const int32_t nSplitB = ozo_split(h, n, 1, x, n, WP(h, oBT, double), n, BS, ldbs, WP(h, oBmax, double));
	int32_t maxlevel = (nSplitA - 1) + (nSplitB - 1);
	maxlevel = imax(maxlevel - h->fastMode, imin(nSplitA - 1, nSplitB - 1));
	double *pC = CS;
	int32_t ic = 0;
	for (int32_t ia = 0; ia < imin(maxlevel + 1, nSplitA); ia++) {
		const int32_t numB = imin(nSplitB, maxlevel + 1 - ia);
		if (h->splitEpsMode == 2) { /* :139-142 */
			csrmm_n_x2(m, numB, AS + (int64_t)ldas * ia, colind, rowptr, BS, ldbs, pC, ldcs);
			pC += (int64_t)ldcs * numB * 2;
		} else {
			csrmm_n(m, numB, 1., AS + (int64_t)ldas * ia, colind, rowptr, BS, ldbs, 0., pC, ldcs);
			pC += (int64_t)ldcs * numB;
		}
		ic += numB;
	}
	h->nSplitA += nSplitA;
	h->nSplitB += nSplitB;
	h->nSplitC += ic;
	if (h->sumMode == 3) return -3;
	if (h->splitEpsMode == 2) maxlevel = (nSplitA - 1) + (nSplitB * 2 - 1); /* :159-162 */
	return global_sum(h, m, 1, nSplitA, nSplitB * ((h->splitEpsMode == 2) ? 2 : 1), CS, ldcs, ldcs,
	                  y, 1, alpha, beta, maxlevel, 3);
}

double *ozo_csrmv_splitA(ozo_handle *h, int32_t m, int32_t n, int32_t nnz, const double *val,
                         const int32_t *rowptr) {
	/* src/cpu/ozblas_csrmv.cpp:191-271 */
	(void)n;
	if (h->reproMode == 0 && h->nSplitMax == 1) return (double *)val;
	counter_init(h);
	int32_t L = (h->nSplitMax > 0) ? h->nSplitMax : NUM_SPLIT_DEFAULT_MAX;
	if (h->reproMode == 1) L++;
	int32_t ldas, ldase;
	const uint64_t oAS = mat_alloc(h, nnz, L, 8, &ldas);
	const uint64_t oAT = vec_alloc(h, nnz, 8);
	const uint64_t oAmax = vec_alloc(h, m, 8);
	mat_alloc(h, m, L, 2, &ldase);
	if (mem_check(h)) {
		fprintf(stderr, "ozoracle: memory shortage\n");
		return NULL;
	}
	arena_check(h);
	h->memMaskSplitA = h->memAddr;
	double *AS = WP(h, oAS, double), *AT = WP(h, oAT, double), *Amax = WP(h, oAmax, double);
	int32_t nA = split_sparse(h, m, val, rowptr, AT, AS, ldas, Amax);
	if (h->nSplitMax == 0) {
		int32_t nAOld;
		h->splitShift = 1;
		do {
			nAOld = nA;
			h->splitShift *= 2;
			nA = split_sparse(h, m, val, rowptr, AT, AS, ldas, Amax);
		} while (nAOld == nA && h->splitShift < 512);
		if (nAOld == nA) h->splitShift = 1;
		nA = split_sparse(h, m, val, rowptr, AT, AS, ldas, Amax);
	}
	h->nSplitA_ = nA;
	h->nSplitA = (float)nA;
	return AS;
}

int32_t ozo_cg(ozo_handle *h, int32_t n, int32_t nnz, const double *val, const int32_t *colind,
               const int32_t *rowptr, const double *b, double *x, int32_t maxiter, double tol) {
	/* src/cpu/ozblas_cg.cpp:24-128 (trueres diagnostics not restated) */
	double alpha, beta, resi, nrmb, tmp, dnew, dold;
	double *R = (double *)malloc(sizeof(double) * (size_t)n);
	double *P = (double *)malloc(sizeof(double) * (size_t)n);
	/* q is read through 0 * q by the first product (beta = 0): start it at +0, as the fresh
	   zero pages of the reference's new TYPE1[dimN] of any sizeable n */
	double *Q = (double *)calloc((size_t)n, sizeof(double));
	h->cg_numiter = 0;
	double *AS = ozo_csrmv_splitA(h, n, n, nnz, val, rowptr);
	nrmb = ozo_nrm2(h, n, b);
	memcpy(R, b, sizeof(double) * (size_t)n);
	ozo_csrmv(h, n, n, nnz, -1., AS, colind, rowptr, x, 1., R);
	dold = ozo_dot(h, n, R, R);
	resi = sqrt(dold);
	if (h->verbose > 0 && h->cg_verbose1) h->cg_verbose1[0] = resi / nrmb;
	memcpy(P, R, sizeof(double) * (size_t)n);
	int32_t vcnt = 1;
	while (h->cg_numiter < maxiter) {
		h->cg_numiter++;
		ozo_csrmv(h, n, n, nnz, 1., AS, colind, rowptr, P, 0., Q);
		tmp = ozo_dot(h, n, P, Q);
		alpha = dold / tmp;
		ozo_axpy(h, n, alpha, P, x);
		ozo_axpy(h, n, -alpha, Q, R);
		dnew = ozo_dot(h, n, R, R);
		resi = sqrt(dnew);
		beta = dnew / dold;
		dold = dnew;
		for (int32_t i = 0; i < n; i++) P[i] = beta * P[i]; /* cblas_dscal */
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
	return 0;
}

/* ---------------------------------------------------------------------------------------------
 * exact checker: fixed-point superaccumulator with one round-to-nearest-even
 * (independent of the reference; covers every product of two binary64 numbers)
 * ------------------------------------------------------------------------------------------- */
#define XW 140
#define XLO (-2240)
typedef struct {
	int64_t w[XW];
	int64_t nadd;
	int nan, pinf, ninf;
} xacc;

static void xacc_init(xacc *a) { memset(a, 0, sizeof(*a)); }

static void xacc_norm(xacc *a) {
	for (int i = 0; i < XW - 1; i++) {
		int64_t c = a->w[i] >> 32;
		a->w[i] -= c * ((int64_t)1 << 32);
		a->w[i + 1] += c;
	}
	a->nadd = 0;
}

static void xacc_add_mag(xacc *a, int neg, unsigned __int128 mag, int exp) {
	if (mag == 0) return;
	const int pos = exp - XLO;
	const int wi = pos >> 5, sh = pos & 31;
	for (int c = 0; c < 4; c++) {
		uint64_t chunk = (uint64_t)(mag >> (32 * c)) & 0xffffffffULL;
		if (!chunk) continue;
		uint64_t v = chunk << sh; /* < 2^63 */
		int64_t lo = (int64_t)(v & 0xffffffffULL), hi = (int64_t)(v >> 32);
		if (neg) {
			a->w[wi + c] -= lo;
			a->w[wi + c + 1] -= hi;
		} else {
			a->w[wi + c] += lo;
			a->w[wi + c + 1] += hi;
		}
	}
	if (++a->nadd >= ((int64_t)1 << 29)) xacc_norm(a);
}

static int decompose(double x, uint64_t *m, int *e) {
	uint64_t b;
	memcpy(&b, &x, 8);
	const int E = (int)((b >> 52) & 0x7ff);
	const uint64_t f = b & ((1ULL << 52) - 1);
	if (E == 0) {
		*m = f;
		*e = -1074;
	} else {
		*m = f | (1ULL << 52);
		*e = E - 1075;
	}
	return (int)(b >> 63);
}

static void xacc_add_double(xacc *a, double x) {
	if (isnan(x)) { a->nan = 1; return; }
	if (isinf(x)) { if (x > 0) a->pinf = 1; else a->ninf = 1; return; }
	uint64_t m;
	int e;
	int s = decompose(x, &m, &e);
	xacc_add_mag(a, s, (unsigned __int128)m, e);
}

static void xacc_add_prod(xacc *a, double x, double y) {
	if (isnan(x) || isnan(y)) { a->nan = 1; return; }
	if (isinf(x) || isinf(y)) {
		if (x == 0. || y == 0.) { a->nan = 1; return; }
		if ((x > 0) == (y > 0)) a->pinf = 1; else a->ninf = 1;
		return;
	}
	uint64_t mx, my;
	int ex, ey;
	int sx = decompose(x, &mx, &ex), sy = decompose(y, &my, &ey);
	xacc_add_mag(a, sx ^ sy, (unsigned __int128)mx * my, ex + ey);
}

static inline int xacc_bit(const int64_t *w, int p) {
	if (p < XLO) return 0;
	const int q = p - XLO;
	return (int)((w[q >> 5] >> (q & 31)) & 1);
}

static double xacc_round(xacc *a) {
	if (a->nan || (a->pinf && a->ninf)) return NAN;
	if (a->pinf) return INFINITY;
	if (a->ninf) return -INFINITY;
	xacc_norm(a);
	int neg = a->w[XW - 1] < 0;
	int64_t w[XW];
	memcpy(w, a->w, sizeof(w));
	if (neg) {
		for (int i = 0; i < XW; i++) w[i] = -w[i];
		for (int i = 0; i < XW - 1; i++) {
			int64_t c = w[i] >> 32;
			w[i] -= c * ((int64_t)1 << 32);
			w[i + 1] += c;
		}
	}
	int hi = -1;
	for (int i = XW - 1; i >= 0; i--)
		if (w[i] != 0) { hi = i; break; }
	if (hi < 0) return 0.0;
	int b = 63 - __builtin_clzll((unsigned long long)w[hi]);
	const int msb = XLO + 32 * hi + b;
	if (msb > 1024) return neg ? -INFINITY : INFINITY;
	int rp = msb - 52;
	if (rp < -1074) rp = -1074;
	uint64_t mant = 0;
	for (int p = msb; p >= rp; p--) mant = (mant << 1) | (uint64_t)xacc_bit(w, p);
	const int rbit = xacc_bit(w, rp - 1);
	int sticky = 0;
	for (int p = rp - 2; p >= XLO && !sticky; p--) {
		const int q = p - XLO;
		if ((q & 31) == 31) { /* whole word below can be tested at once */
			for (int i = (q >> 5); i >= 0; i--)
				if (w[i] & 0xffffffffLL) { sticky = 1; break; }
			break;
		}
		sticky |= xacc_bit(w, p);
	}
	if (rbit && (sticky || (mant & 1))) mant++;
	double r = ldexp((double)mant, rp);
	return neg ? -r : r;
}

double ozo_exact_sum_rn(int64_t n, const double *v) {
	xacc a;
	xacc_init(&a);
	for (int64_t i = 0; i < n; i++) xacc_add_double(&a, v[i]);
	return xacc_round(&a);
}

double ozo_exact_dot_rn(int64_t n, const double *x, int64_t incx, const double *y, int64_t incy) {
	xacc a;
	xacc_init(&a);
	for (int64_t i = 0; i < n; i++) xacc_add_prod(&a, x[i * incx], y[i * incy]);
	return xacc_round(&a);
}

static inline double a_elem(char t, const double *A, int32_t lda, int32_t i, int32_t kk) {
	return is_n(t) ? A[i + (int64_t)kk * lda] : A[kk + (int64_t)i * lda];
}
static inline double b_elem(char t, const double *B, int32_t ldb, int32_t kk, int32_t j) {
	return is_n(t) ? B[kk + (int64_t)j * ldb] : B[j + (int64_t)kk * ldb];
}

void ozo_exact_gemm_rn(char transA, char transB, int32_t m, int32_t n, int32_t k, const double *A,
                       int32_t lda, const double *B, int32_t ldb, double *C, int32_t ldc,
                       int32_t nthreads) {
	(void)nthreads;
#pragma omp parallel for schedule(dynamic)
	for (int32_t j = 0; j < n; j++)
		for (int32_t i = 0; i < m; i++) {
			xacc a;
			xacc_init(&a);
			for (int32_t kk = 0; kk < k; kk++)
				xacc_add_prod(&a, a_elem(transA, A, lda, i, kk), b_elem(transB, B, ldb, kk, j));
			C[i + (int64_t)j * ldc] = xacc_round(&a);
		}
}

void ozo_exact_gemm_rn_entries(char transA, char transB, int32_t k, const double *A, int32_t lda,
                               const double *B, int32_t ldb, int64_t nent, const int32_t *rows,
                               const int32_t *cols, double *out) {
#pragma omp parallel for schedule(dynamic)
	for (int64_t e = 0; e < nent; e++) {
		xacc a;
		xacc_init(&a);
		for (int32_t kk = 0; kk < k; kk++)
			xacc_add_prod(&a, a_elem(transA, A, lda, rows[e], kk), b_elem(transB, B, ldb, kk, cols[e]));
		out[e] = xacc_round(&a);
	}
}

/* double-double dot (Ogita-Rump-Oishi Dot2): an independent ~106-bit estimate */
void ozo_dd_dot(int64_t n, const double *x, int64_t incx, const double *y, int64_t incy, double *hi,
                double *lo) {
	double p = 0., s = 0.;
	for (int64_t i = 0; i < n; i++) {
		const double a = x[i * incx], b = y[i * incy];
		const double h = a * b;
		const double r = fma(a, b, -h);
		const double t = p + h;
		const double z = t - p;
		const double q = (p - (t - z)) + (h - z);
		p = t;
		s = s + (q + r);
	}
	const double H = p + s;
	*hi = H;
	*lo = s - (H - p);
}

/* ---------------------------------------------------------------------------------------------
 * generators (testing/testing_common.cpp:664-806), restated without calling libc's PRNGs
 * ------------------------------------------------------------------------------------------- */
static uint64_t g_lcg = 0x1234ABCD330EULL;

void ozo_srand48(int64_t seed) { g_lcg = (((uint64_t)(uint32_t)seed) << 16) | 0x330EULL; }

double ozo_drand48(void) {
	g_lcg = (0x5DEECE66DULL * g_lcg + 0xBULL) & 0xFFFFFFFFFFFFULL;
	return ldexp((double)g_lcg, -48);
}

/* glibc random(): additive lagged-Fibonacci generator r[i] = r[i-3] + r[i-31], seeded by a
   Lehmer generator (16807 mod 2^31-1) with 310 warm-up outputs; rand() never seeded == srand(1) */
static int32_t g_rstate[31];
static int g_rf = 3, g_rr = 0, g_rinit = 0;

void ozo_srand(uint32_t seed) {
	if (seed == 0) seed = 1;
	g_rstate[0] = (int32_t)seed;
	int32_t word = (int32_t)seed;
	for (int i = 1; i < 31; i++) {
		const int32_t hi = word / 127773, lo = word % 127773;
		word = 16807 * lo - 2836 * hi;
		if (word < 0) word += 2147483647;
		g_rstate[i] = word;
	}
	g_rf = 3;
	g_rr = 0;
	g_rinit = 1;
	for (int i = 0; i < 310; i++) (void)ozo_rand();
}

int32_t ozo_rand(void) {
	if (!g_rinit) ozo_srand(1);
	uint32_t val = (uint32_t)g_rstate[g_rf] + (uint32_t)g_rstate[g_rr];
	g_rstate[g_rf] = (int32_t)val;
	const int32_t result = (int32_t)(val >> 1);
	if (++g_rf >= 31) {
		g_rf = 0;
		++g_rr;
	} else if (++g_rr >= 31) {
		g_rr = 0;
	}
	return result;
}

void ozo_init_mat(int32_t m, int32_t n, int32_t ld, double *mat, double phi, int32_t mode,
                  int32_t seed) {
	if (ld < m) ld = m;
	ozo_srand48(seed);
	switch (mode) {
	case 0:
		for (int32_t j = 0; j < n; j++)
			for (int32_t i = 0; i < m; i++) mat[(int64_t)j * ld + i] = phi;
		break;
	case 1:
		for (int32_t j = 0; j < n; j++)
			for (int32_t i = 0; i < m; i++) mat[(int64_t)j * ld + i] = ozo_drand48();
		break;
	case 2: /* phi-mode :703-717 (r1 reused) */
		for (int32_t j = 0; j < n; j++)
			for (int32_t i = 0; i < m; i++) {
				const double mu = 0.0, sigma = 1.0;
				const double r1 = ozo_drand48();
				const double r2 = ozo_drand48();
				const double x1 = mu + (sqrt(-2. * log(r1)) * sin(2. * M_PI * r2)) * sigma;
				mat[(int64_t)j * ld + i] = (r1 - 0.5) * exp(phi * x1);
			}
		break;
	case 3: /* erange-mode :718-763 (binary64: M_PI/f_pi == 1) */
		for (int32_t j = 0; j < n; j++)
			for (int32_t i = 0; i < m; i++) {
				const double d = ozo_drand48();
				const int32_t e = ozo_rand() % (int32_t)phi;
				double valf = (d * 9 + 1) * pow(10, e);
				valf = ((ozo_rand() % 2) ? 1. : -1.) * ((double)M_PI / M_PI) * valf;
				mat[(int64_t)j * ld + i] = valf;
			}
		break;
	}
}

int64_t ozo_laplace3d_nnz(int32_t g) {
	const int64_t n = (int64_t)g * g * g;
	return 7 * n - 6 * (int64_t)g * g;
}

void ozo_laplace3d(int32_t g, double shift, int32_t *rowptr, int32_t *colind, double *val) {
	int64_t p = 0;
	const int64_t g2 = (int64_t)g * g;
	const double diag = 6.0 + shift;
	for (int32_t l = 0; l < g; l++)
		for (int32_t j = 0; j < g; j++)
			for (int32_t i = 0; i < g; i++) {
				const int64_t r = i + (int64_t)g * j + g2 * l;
				rowptr[r] = (int32_t)p;
				if (l > 0) { colind[p] = (int32_t)(r - g2); val[p++] = -1.0; }
				if (j > 0) { colind[p] = (int32_t)(r - g); val[p++] = -1.0; }
				if (i > 0) { colind[p] = (int32_t)(r - 1); val[p++] = -1.0; }
				colind[p] = (int32_t)r;
				val[p++] = diag;
				if (i < g - 1) { colind[p] = (int32_t)(r + 1); val[p++] = -1.0; }
				if (j < g - 1) { colind[p] = (int32_t)(r + g); val[p++] = -1.0; }
				if (l < g - 1) { colind[p] = (int32_t)(r + g2); val[p++] = -1.0; }
			}
	rowptr[g2 * g] = (int32_t)p;
}
