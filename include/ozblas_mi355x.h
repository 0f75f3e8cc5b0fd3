/*
 * ozblas_mi355x.h -- C ABI of the MI355X (gfx950) Ozaki-scheme BLAS.
 *
 * Drop-in for the OzBLAS GPU path: the handle keeps the fields of cuozblasHandle_t
 * (reference include/cuozblas.h:7-63) with the same names and meanings; the cuBLAS/cuSPARSE handles
 * are replaced by a HIP stream and a library-private backend pointer.  Entry points are the
 * binary64 (D) and binary32 (S) instantiations of the reference templates, given C names:
 *
 *   reference (include/cuozblas.h)                          here
 *   cuozblasCreate / cuozblasDestroy              :67-68    ozblasCreate / ozblasDestroy
 *   cuozblasRdot<double,double,double>            :75       ozblasDdot
 *   cuozblasRnrm2<double,double,double>           :73       ozblasDnrm2
 *   cuozblasRaxpy<double>                         :74       ozblasDaxpy
 *   cuozblasRgemv<double,double,double>           :76       ozblasDgemv
 *   cuozblasRgemm<double,double,double>           :77       ozblasDgemm
 *   cuozblasRcsrmv<double,double>                 :78       ozblasDcsrmv
 *   cuozblasRcsrmvSplitA<double,double>           :79       ozblasDcsrmvSplitA
 *   cuozblasRcg<double,double>                    :80       ozblasDcg
 *   the same templates at <float,float,float> / <float,float> / <float>
 *   (instantiated for the CPU library in src/cpu/ozblas_dot.cpp:44, ozblas_nrm2.cpp:36,
 *   ozblas_axpy.cpp:30, ozblas_gemv.cpp:34, ozblas_gemm.cpp:320, ozblas_csrmv.cpp:186,276,
 *   ozblas_cg.cpp:133)                                      ozblasS{dot,nrm2,axpy,gemv,gemm,
 *                                                                  csrmv,csrmvSplitA,cg}
 *
 * Semantics follow cuOzBLAS: matrices and vectors are device pointers (column-major, 0-based
 * CSR), alpha/beta are host scalars, DOT/NRM2 write their result to a host pointer, every call is
 * synchronous on the handle's stream.  Numerical semantics (split, slice counts, summation order,
 * the epilogue fma(alpha, t, beta*C) that reads C even for beta == 0) follow the reference CPU
 * OzBLAS (src/cpu), so with sliceType = OZBLAS_SLICE_FP64 the output is bitwise identical to it.
 * Errors are returned as OZBLAS_STATUS_* codes instead of exit(1).
 *
 * Binary32: the split, NearSum/Fsum and epilogue follow the reference's float instantiations
 * (float arithmetic, with its double-evaluated log2, fma and sqrt); slices are held as doubles and
 * multiplied on the FP64 path, which is exact, then rounded once to float as the reference's float
 * slice products.  sliceType is not consulted (the digit forms are binary64 only), and Dot2
 * (splitEpsMode = 2) returns NOT_SUPPORTED, as the reference's float Dot2 GEMM is a stub
 * (src/cpu/ozblas_blas_wrapper.cpp:179-181).  ozblasScg iterates device-resident like ozblasDcg,
 * for search directions of at most 8 slices (more redo the iterations on the synchronous path).
 */
#ifndef OZBLAS_MI355X_H
#define OZBLAS_MI355X_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
	OZBLAS_STATUS_SUCCESS = 0,
	OZBLAS_STATUS_NOT_INITIALIZED = 1,
	OZBLAS_STATUS_ALLOC_FAILED = 2,
	OZBLAS_STATUS_INVALID_VALUE = 3,
	OZBLAS_STATUS_NOT_SUPPORTED = 4,
	OZBLAS_STATUS_WORKSPACE = 5,   /* workspace too small even after m/n blocking */
	OZBLAS_STATUS_EXECUTION_FAILED = 6
};

/* slice formats of the products (no reference counterpart).  FP64 is the reference's plain form.
 * INT8/FP16/BF16 cut every operand into exact 8-bit slices whose products are exact on INT8
 * (INT32 accumulate) or FP16/BF16 (FP32 accumulate) MFMA, and round the exact sum once: they give
 * the bits of sumMode = 1 (NearSum) whenever the FP64 split is exact (nSplitMax = 0).  The slices
 * are either residues of the row/column-scaled integers modulo small coprime moduli (one product
 * per modulus, exact Chinese-remainder reconstruction) or radix-256 digit planes (one product per
 * digit pair, certified early termination); digitScheme selects. */
enum {
	OZBLAS_SLICE_FP64 = 0,
	OZBLAS_SLICE_INT8 = 1,
	OZBLAS_SLICE_FP16 = 2,
	OZBLAS_SLICE_BF16 = 3
};

typedef struct ozblasHandle {
	/* backend: replaces cublasHandle_t ch / cusparseHandle_t csh (include/cuozblas.h:8-9) */
	void *stream;  /* hipStream_t all work is issued on (NULL = default stream) */
	void *backend; /* library-private state */

	/* core (include/cuozblas.h:11-18) */
	char *devWork;
	char *devWorkCommon;
	uint64_t workSizeBytes;
	uint64_t memAddr;
	int32_t nSplitMax;
	char *hstBatchAddr;
	char *devBatchAddr;

	int32_t initialized;

	/* option flags (include/cuozblas.h:22-28) */
	int32_t splitMode;
	int32_t fastMode;
	int32_t reproMode;
	int32_t sumMode;
	int32_t useBatchedGemmFlag;
	int32_t splitEpsMode;
	int32_t precxFlag;

	/* exec info (include/cuozblas.h:31-42) */
	float nSplitA;
	float nSplitB;
	float nSplitC;
	float t_SplitA;
	float t_SplitB;
	float t_comp;
	float t_sum;
	float t_total;
	float n_comp;
	int32_t mbk;
	int32_t nbk;

	/* SpMV in iterative solvers (include/cuozblas.h:45-47) */
	int32_t nSplitA_;
	int32_t splitShift;
	uint64_t memMaskSplitA;

	/* CG (include/cuozblas.h:50-61, src/cpu/ozblas_cg.cpp:40-112).  With verbose > 0 the solver
	 * writes resi/nrmb of iteration 0 and of every verbose-th iteration (and of a tolerance stop)
	 * into cg_verbose1, a host array of >= maxiter/verbose + 2 elements of the routine's type
	 * (double for Dcg, float for Scg).  When cg_verbose2 is set, the same entries of it get the true
	 * relative residual |b - A x| / |b| (trueresFlag = 1; routine's type; iterates synchronously)
	 * or the seconds since the start of the solve (trueresFlag = 0; double).  The t_*_total fields
	 * get the solve's time split as the reference's categories (the synchronous iterations only;
	 * t_SplitVec_total and t_Sum_total are counted in the SpMV and DOT totals, where they run
	 * fused). */
	int32_t trueresFlag;
	int32_t verbose;
	int32_t cg_numiter;
	void *cg_verbose1;
	void *cg_verbose2;
	double t_SplitMat_total;
	double t_SplitVec_total;
	double t_Sum_total;
	double t_AXPY_SCAL_total;
	double t_DOT_NRM2_total;
	double t_SpMV_SpMM_total;

	/* MI355X additions (no reference counterpart) */
	int32_t sliceType;    /* OZBLAS_SLICE_*, default OZBLAS_SLICE_FP64 */
	int32_t lastStatus;   /* status of the last call (also returned by it) */
	int32_t nLevelBins;   /* fast forms: number of exact level planes of the last GEMM */
	int32_t batchTile;    /* ozblasDgemmBatched: output tile edge picked by the autotuner (128, 256) */
	int32_t batchReplay;  /* ozblasDgemmBatched: 1 when the call replayed the cached hipGraph */
	int32_t batchGrouped; /* ozblasDgemmBatched: problems run in the grouped launch */
	int32_t digitLevels;  /* digit forms: digit levels computed before the certified rounding check;
	                         0 = automatic (default), -1 = all levels, n > 0 = the first n levels */
	int32_t fallbackTiles; /* digit forms: output tiles of the last GEMM that needed the levels after
	                          the check */
	int32_t cgDeviceSlices; /* CG: 0 = iterate device-resident (default); -1 = iterate on the
	                           synchronous path; n > 0 = give up the device path (and redo the
	                           iterations synchronously) once a search direction has more than n
	                           slices */
	int32_t cgDevicePath;   /* last CG: 1 iterated device-resident, 2 gave up and redid the iterations
	                           synchronously, 0 iterated synchronously */
	int32_t digitScheme;    /* INT8/FP16/BF16 forms: 0 = automatic (default: residues, unless
	                           digitLevels != 0 or the moduli cannot cover the product, then digits),
	                           1 = radix-256 digit planes, 2 = residues modulo small coprime moduli
	                           (digits when they cannot cover the product) */
	int32_t schemeUsed;     /* last DGEMM: 0 = plain FP64 slices, 1 = digit planes, 2 = residues */
} ozblasHandle_t;

/* lifecycle: cuozblasCreate/cuozblasDestroy (include/cuozblas.h:67-68, src/cuda/cuozblas.cpp:23-75).
 * Defaults as src/cpu/ozblas.cpp:31-49 (reproMode=1, sumMode=0, nSplitMax=0, ...). */
int32_t ozblasCreate(ozblasHandle_t *oh, uint64_t workSizeBytes);
int32_t ozblasDestroy(ozblasHandle_t *oh);
int32_t ozblasSetStream(ozblasHandle_t *oh, void *hipStream);

/* level 1: cuozblasRdot / cuozblasRnrm2 / cuozblasRaxpy (include/cuozblas.h:73-75) */
int32_t ozblasDdot(ozblasHandle_t *oh, int32_t n, const double *x, int32_t incx, const double *y,
                   int32_t incy, double *ret);
/* ozblasDdot without allocation -- no counterpart in the reference.  For callers that cannot
 * allocate device memory at call time (tight serving loops, harnesses that forbid hipMalloc): the
 * same result bits as ozblasDdot, computed in a static 64 MiB scratch array that the library keeps
 * on each GPU until the process ends (looked up per device at ozblasCreate), instead of the handle's
 * workspace.  It calls no hipMalloc / hipFree.  Limit: both vectors' slot buffers plus the sums must
 * fit the 64 MiB, about 16 * (cap + 2) * n bytes with cap the split cap (30 for nSplitMax = 0 and
 * reproMode = 1), so n <= 130,432 for two distinct vectors at the defaults; beyond it the call
 * returns OZBLAS_STATUS_WORKSPACE before any work.  Every call on one GPU shares
 * that GPU's array: calls that overlap in time on one GPU (other host threads, other streams) race
 * and must be serialised by the caller. */
int32_t ozblasDdotNoAlloc(ozblasHandle_t *oh, int32_t n, const double *x, int32_t incx,
                          const double *y, int32_t incy, double *ret);
int32_t ozblasDnrm2(ozblasHandle_t *oh, int32_t n, const double *x, int32_t incx, double *ret);
int32_t ozblasDaxpy(ozblasHandle_t *oh, int32_t n, double alpha, const double *x, int32_t incx,
                    double *y, int32_t incy);
int32_t ozblasSdot(ozblasHandle_t *oh, int32_t n, const float *x, int32_t incx, const float *y,
                   int32_t incy, float *ret);
int32_t ozblasSnrm2(ozblasHandle_t *oh, int32_t n, const float *x, int32_t incx, float *ret);
int32_t ozblasSaxpy(ozblasHandle_t *oh, int32_t n, float alpha, const float *x, int32_t incx,
                    float *y, int32_t incy);

/* level 2/3: cuozblasRgemv / cuozblasRgemm (include/cuozblas.h:76-77) */
int32_t ozblasDgemv(ozblasHandle_t *oh, char transA, int32_t m, int32_t n, double alpha,
                    const double *A, int32_t lda, const double *x, int32_t incx, double beta,
                    double *y, int32_t incy);
int32_t ozblasDgemm(ozblasHandle_t *oh, char transA, char transB, int32_t m, int32_t n, int32_t k,
                    double alpha, const double *A, int32_t lda, const double *B, int32_t ldb,
                    double beta, double *C, int32_t ldc);
int32_t ozblasSgemv(ozblasHandle_t *oh, char transA, int32_t m, int32_t n, float alpha,
                    const float *A, int32_t lda, const float *x, int32_t incx, float beta, float *y,
                    int32_t incy);
int32_t ozblasSgemm(ozblasHandle_t *oh, char transA, char transB, int32_t m, int32_t n, int32_t k,
                    float alpha, const float *A, int32_t lda, const float *B, int32_t ldb, float beta,
                    float *C, int32_t ldc);

/* sparse: cuozblasRcsrmv / cuozblasRcsrmvSplitA / cuozblasRcg (include/cuozblas.h:78-80).
 * descrA is accepted and ignored (general matrix, 0-based indices). */
int32_t ozblasDcsrmv(ozblasHandle_t *oh, char transA, int32_t m, int32_t n, int32_t nnz,
                     double alpha, const void *descrA, const double *A, const int32_t *colind,
                     const int32_t *rowptr, const double *x, double beta, double *y);
double *ozblasDcsrmvSplitA(ozblasHandle_t *oh, char transA, int32_t m, int32_t n, int32_t nnz,
                           const void *descrA, const double *A, const int32_t *rowptr);
int32_t ozblasDcg(ozblasHandle_t *oh, char transA, int32_t n, int32_t nnz, const void *descrA,
                  const double *A, const int32_t *colind, const int32_t *rowptr, const double *b,
                  double *x, int32_t maxiter, double tol);
/* ozblasScsrmvSplitA returns the library's slice buffer (FP64 slices in the workspace), to be
 * passed back as A of ozblasScsrmv; its layout is private, the pointer is typed as the reference's */
int32_t ozblasScsrmv(ozblasHandle_t *oh, char transA, int32_t m, int32_t n, int32_t nnz,
                     float alpha, const void *descrA, const float *A, const int32_t *colind,
                     const int32_t *rowptr, const float *x, float beta, float *y);
float *ozblasScsrmvSplitA(ozblasHandle_t *oh, char transA, int32_t m, int32_t n, int32_t nnz,
                          const void *descrA, const float *A, const int32_t *rowptr);
int32_t ozblasScg(ozblasHandle_t *oh, char transA, int32_t n, int32_t nnz, const void *descrA,
                  const float *A, const int32_t *colind, const int32_t *rowptr, const float *b,
                  float *x, int32_t maxiter, float tol);

/* mixed: binary64 data split into binary32 slices -- the reference's <double,float,float> /
 * <double,float> instantiations (src/cpu/ozblas_gemm.cpp:319, ozblas_gemv.cpp:33, ozblas_dot.cpp:43,
 * ozblas_nrm2.cpp:35, ozblas_csrmv.cpp:185,275, ozblas_cg.cpp:132), reached through the same
 * templates of include/cuozblas.h:73-80.  Bitwise equal to the reference CPU build; sumMode = 0
 * (Fsum) only, as the reference's mixed NearSum is not available (src/cpu/ozblas_sum.cpp:275-330):
 * other sumModes and Dot2 return NOT_SUPPORTED.  ozblasDScsrmvSplitA returns the library's slice
 * buffer typed as the reference's float *, to be passed back as A of ozblasDScsrmv. */
int32_t ozblasDSgemm(ozblasHandle_t *oh, char transA, char transB, int32_t m, int32_t n, int32_t k,
                     double alpha, const double *A, int32_t lda, const double *B, int32_t ldb,
                     double beta, double *C, int32_t ldc);
int32_t ozblasDSgemv(ozblasHandle_t *oh, char transA, int32_t m, int32_t n, double alpha,
                     const double *A, int32_t lda, const double *x, int32_t incx, double beta,
                     double *y, int32_t incy);
int32_t ozblasDSdot(ozblasHandle_t *oh, int32_t n, const double *x, int32_t incx, const double *y,
                    int32_t incy, double *ret);
int32_t ozblasDSnrm2(ozblasHandle_t *oh, int32_t n, const double *x, int32_t incx, double *ret);
int32_t ozblasDScsrmv(ozblasHandle_t *oh, char transA, int32_t m, int32_t n, int32_t nnz,
                      double alpha, const void *descrA, const double *A, const int32_t *colind,
                      const int32_t *rowptr, const double *x, double beta, double *y);
float *ozblasDScsrmvSplitA(ozblasHandle_t *oh, char transA, int32_t m, int32_t n, int32_t nnz,
                           const void *descrA, const double *A, const int32_t *rowptr);
int32_t ozblasDScg(ozblasHandle_t *oh, char transA, int32_t n, int32_t nnz, const void *descrA,
                   const double *A, const int32_t *colind, const int32_t *rowptr, const double *b,
                   double *x, int32_t maxiter, double tol);

/* mixed: binary32 data split into binary64-held slices -- the reference's <float,double,double> /
 * <float,double> instantiations (src/cpu/ozblas_gemm.cpp:321, ozblas_gemv.cpp:35, ozblas_dot.cpp:45,
 * ozblas_nrm2.cpp:37, ozblas_csrmv.cpp:187,277, ozblas_cg.cpp:134; templates of
 * include/cuozblas.h:73-80).  The split runs in binary32 with getRho<float,double>, the slice
 * products in binary64, the Fsum in binary32.  Bitwise equal to the reference CPU build wherever
 * its binary64 slice products are exact (inner dimension above 128 in splitEpsMode 0; every CSR
 * product, whose fma chain the reference fixes); sumMode = 0 only (the mixed NearSum exits in the
 * reference, src/cpu/ozblas_sum.cpp:275-330); Dot2 returns NOT_SUPPORTED.  ozblasSDcsrmvSplitA
 * returns the library's slice buffer typed as the reference's double *, to be passed back as A of
 * ozblasSDcsrmv. */
int32_t ozblasSDgemm(ozblasHandle_t *oh, char transA, char transB, int32_t m, int32_t n, int32_t k,
                     float alpha, const float *A, int32_t lda, const float *B, int32_t ldb,
                     float beta, float *C, int32_t ldc);
int32_t ozblasSDgemv(ozblasHandle_t *oh, char transA, int32_t m, int32_t n, float alpha,
                     const float *A, int32_t lda, const float *x, int32_t incx, float beta, float *y,
                     int32_t incy);
int32_t ozblasSDdot(ozblasHandle_t *oh, int32_t n, const float *x, int32_t incx, const float *y,
                    int32_t incy, float *ret);
int32_t ozblasSDnrm2(ozblasHandle_t *oh, int32_t n, const float *x, int32_t incx, float *ret);
int32_t ozblasSDcsrmv(ozblasHandle_t *oh, char transA, int32_t m, int32_t n, int32_t nnz,
                      float alpha, const void *descrA, const float *A, const int32_t *colind,
                      const int32_t *rowptr, const float *x, float beta, float *y);
double *ozblasSDcsrmvSplitA(ozblasHandle_t *oh, char transA, int32_t m, int32_t n, int32_t nnz,
                            const void *descrA, const float *A, const int32_t *rowptr);
int32_t ozblasSDcg(ozblasHandle_t *oh, char transA, int32_t n, int32_t nnz, const void *descrA,
                   const float *A, const int32_t *colind, const int32_t *rowptr, const float *b,
                   float *x, int32_t maxiter, float tol);

/* batched uneven GEMM (no reference counterpart; the reference batches only the slice pairs of one
 * GEMM, src/cuda/cuozblas_gemm.cu:150-190).  Arrays of per-problem sizes and device pointers live
 * in host memory.  Every C[p] is bitwise equal to ozblasDgemm on problem p alone. */
int32_t ozblasDgemmBatched(ozblasHandle_t *oh, int32_t batch, const char *transA,
                           const char *transB, const int32_t *m, const int32_t *n,
                           const int32_t *k, const double *alpha, const double *const *A,
                           const int32_t *lda, const double *const *B, const int32_t *ldb,
                           const double *beta, double *const *C, const int32_t *ldc);

/* input generator of the reference test harness, phi-mode (testing/testing_common.cpp:703-717):
 * (r1 - 0.5) * exp(phi * sqrt(-2 ln r1) sin(2 pi r2)) from drand48 seeded with `seed`, written
 * column-major into the HOST buffer A (ld >= m) -- no reference counterpart in the library API */
void ozblasGenPhi(int32_t m, int32_t n, int32_t ld, double *A, double phi, int64_t seed);

/* library info */
const char *ozblasGetVersionString(void);

#ifdef __cplusplus
}
#endif
#endif
