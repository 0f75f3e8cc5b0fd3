/*
 * cuozblas_mi355x.hpp -- the cuOzBLAS C++ templates over the C ABI of ozblas_mi355x.h.
 *
 * Callers written against the reference's GPU interface (include/cuozblas.h:67-80, used by the test
 * drivers through the macros of testing/testing_common.h:223-233,355-365) compile against this
 * header with no edits to their calls.  Each specialization forwards to the C entry point of the
 * same instantiation:
 *
 *   cuozblasCreate / cuozblasDestroy                     (include/cuozblas.h:67-68)
 *   cuozblasRnrm2 / Raxpy / Rdot / Rgemv / Rgemm         (include/cuozblas.h:73-77)
 *   cuozblasRcsrmv / RcsrmvSplitA / Rcg                  (include/cuozblas.h:78-80)
 * at <double,...> (ozblasD*), <float,...> (ozblasS*) and the mixed <double,float,...> (ozblasDS*)
 * and <float,double,...> (ozblasSD*).
 *
 * The reference returns 0 and calls exit(1) on an error; these return the OZBLAS_STATUS_* code
 * (0 on success), or exit(1) as the reference does when OZBLAS_SHIM_EXIT_ON_ERROR is defined.
 * Instantiations without a C entry point (the binary128 ones) are declared by the primary templates
 * and left undefined, so using one fails at link time.
 */
#ifndef CUOZBLAS_MI355X_HPP
#define CUOZBLAS_MI355X_HPP

#include <cstdint>
#include <cstdio>
#include <cstdlib>

#include "ozblas_mi355x.h"

typedef ozblasHandle_t cuozblasHandle_t;
#ifndef OZBLAS_SHIM_NO_CUSPARSE_DESCR
typedef const void *cusparseMatDescr_t;  // accepted and ignored: general matrix, 0-based CSR
#endif

inline int32_t cuozblas_status(int32_t s, const char *where) {
#ifdef OZBLAS_SHIM_EXIT_ON_ERROR
	if (s != OZBLAS_STATUS_SUCCESS) {
		fprintf(stderr, "OzBLAS error: %s (status %d)\n", where, (int)s);
		exit(1);
	}
#else
	(void)where;
#endif
	return s;
}

inline void cuozblasCreate(cuozblasHandle_t *oh, uint64_t workSizeBytes) {
	cuozblas_status(ozblasCreate(oh, workSizeBytes), "cuozblasCreate");
}
inline void cuozblasDestroy(cuozblasHandle_t *oh) { ozblasDestroy(oh); }

// ---- primary templates (the reference's signatures) ----------------------------------------------
template <typename T1, typename T2, typename T3>
int32_t cuozblasRnrm2(cuozblasHandle_t *oh, const int32_t n, const T1 *devX, const int32_t incx, T1 *ret);
template <typename T>
int32_t cuozblasRaxpy(cuozblasHandle_t *oh, const int32_t n, const T alpha, const T *devX,
                      const int32_t incx, T *devY, const int32_t incy);
template <typename T1, typename T2, typename T3>
int32_t cuozblasRdot(cuozblasHandle_t *oh, const int32_t n, const T1 *devA, const int32_t incx,
                     const T1 *devB, const int32_t incy, T1 *ret);
template <typename T1, typename T2, typename T3>
int32_t cuozblasRgemv(cuozblasHandle_t *oh, const char tranA, const int32_t m, const int32_t n,
                      const T1 alpha, const T1 *devA, const int32_t lda, const T1 *devB,
                      const int32_t incx, const T1 beta, T1 *devC, const int32_t incy);
template <typename T1, typename T2, typename T3>
int32_t cuozblasRgemm(cuozblasHandle_t *oh, const char tranA, const char tranB, const int32_t m,
                      const int32_t n, const int32_t k, const T1 alpha, const T1 *devA,
                      const int32_t lda, const T1 *devB, const int32_t ldb, const T1 beta, T1 *devC,
                      const int32_t ldc);
template <typename T1, typename T2>
int32_t cuozblasRcsrmv(cuozblasHandle_t *oh, const char tranA, const int32_t m, const int32_t n,
                       const int32_t nnz, const T1 alpha, const cusparseMatDescr_t descrA,
                       const T1 *devA, const int32_t *devAcolind, const int32_t *devArowptr,
                       const T1 *devB, const T1 beta, T1 *devC);
template <typename T1, typename T2>
T2 *cuozblasRcsrmvSplitA(cuozblasHandle_t *oh, const char tranA, const int32_t m, const int32_t n,
                         const int32_t nnz, const cusparseMatDescr_t descrA, const T1 *devA,
                         const int32_t *devArowptr);
template <typename T1, typename T2>
int32_t cuozblasRcg(cuozblasHandle_t *oh, const char tranA, const int32_t dimN, const int32_t dimNNZ,
                    const cusparseMatDescr_t descrA, const T1 *matA, const int32_t *matAcolind,
                    const int32_t *matArowptr, const T1 *vecB, T1 *vecX, int32_t maxiter, T1 tol);

// ---- specializations: one per C entry point -------------------------------------------------------
#define CUOZBLAS_MI355X_LEVEL123(T, P)                                                                   \
	template <>                                                                                         \
	inline int32_t cuozblasRnrm2<T, T, T>(cuozblasHandle_t * oh, const int32_t n, const T *devX,          \
	                                      const int32_t incx, T *ret) {                                 \
		return cuozblas_status(ozblas##P##nrm2(oh, n, devX, incx, ret), "cuozblasRnrm2");               \
	}                                                                                                   \
	template <>                                                                                         \
	inline int32_t cuozblasRaxpy<T>(cuozblasHandle_t * oh, const int32_t n, const T alpha, const T *devX, \
	                                const int32_t incx, T *devY, const int32_t incy) {                   \
		return cuozblas_status(ozblas##P##axpy(oh, n, alpha, devX, incx, devY, incy), "cuozblasRaxpy"); \
	}                                                                                                   \
	template <>                                                                                         \
	inline int32_t cuozblasRdot<T, T, T>(cuozblasHandle_t * oh, const int32_t n, const T *devA,           \
	                                     const int32_t incx, const T *devB, const int32_t incy, T *ret) { \
		return cuozblas_status(ozblas##P##dot(oh, n, devA, incx, devB, incy, ret), "cuozblasRdot");     \
	}                                                                                                   \
	template <>                                                                                         \
	inline int32_t cuozblasRgemv<T, T, T>(cuozblasHandle_t * oh, const char tranA, const int32_t m,       \
	                                      const int32_t n, const T alpha, const T *devA,                \
	                                      const int32_t lda, const T *devB, const int32_t incx,          \
	                                      const T beta, T *devC, const int32_t incy) {                  \
		return cuozblas_status(                                                                         \
		    ozblas##P##gemv(oh, tranA, m, n, alpha, devA, lda, devB, incx, beta, devC, incy),           \
		    "cuozblasRgemv");                                                                           \
	}                                                                                                   \
	template <>                                                                                         \
	inline int32_t cuozblasRgemm<T, T, T>(cuozblasHandle_t * oh, const char tranA, const char tranB,      \
	                                      const int32_t m, const int32_t n, const int32_t k,            \
	                                      const T alpha, const T *devA, const int32_t lda,              \
	                                      const T *devB, const int32_t ldb, const T beta, T *devC,      \
	                                      const int32_t ldc) {                                          \
		return cuozblas_status(                                                                         \
		    ozblas##P##gemm(oh, tranA, tranB, m, n, k, alpha, devA, lda, devB, ldb, beta, devC, ldc),   \
		    "cuozblasRgemm");                                                                           \
	}                                                                                                   \
	template <>                                                                                         \
	inline int32_t cuozblasRcsrmv<T, T>(cuozblasHandle_t * oh, const char tranA, const int32_t m,         \
	                                    const int32_t n, const int32_t nnz, const T alpha,              \
	                                    const cusparseMatDescr_t descrA, const T *devA,                 \
	                                    const int32_t *devAcolind, const int32_t *devArowptr,           \
	                                    const T *devB, const T beta, T *devC) {                         \
		return cuozblas_status(ozblas##P##csrmv(oh, tranA, m, n, nnz, alpha, descrA, devA, devAcolind,   \
		                                        devArowptr, devB, beta, devC),                          \
		                       "cuozblasRcsrmv");                                                       \
	}                                                                                                   \
	template <>                                                                                         \
	inline T *cuozblasRcsrmvSplitA<T, T>(cuozblasHandle_t * oh, const char tranA, const int32_t m,       \
	                                     const int32_t n, const int32_t nnz,                            \
	                                     const cusparseMatDescr_t descrA, const T *devA,                \
	                                     const int32_t *devArowptr) {                                   \
		T *p = ozblas##P##csrmvSplitA(oh, tranA, m, n, nnz, descrA, devA, devArowptr);                  \
		if (!p) cuozblas_status(oh->lastStatus, "cuozblasRcsrmvSplitA");                                \
		return p;                                                                                       \
	}                                                                                                   \
	template <>                                                                                         \
	inline int32_t cuozblasRcg<T, T>(cuozblasHandle_t * oh, const char tranA, const int32_t dimN,        \
	                                 const int32_t dimNNZ, const cusparseMatDescr_t descrA,             \
	                                 const T *matA, const int32_t *matAcolind,                          \
	                                 const int32_t *matArowptr, const T *vecB, T *vecX,                 \
	                                 int32_t maxiter, T tol) {                                          \
		return cuozblas_status(ozblas##P##cg(oh, tranA, dimN, dimNNZ, descrA, matA, matAcolind,          \
		                                     matArowptr, vecB, vecX, maxiter, tol),                      \
		                       "cuozblasRcg");                                                          \
	}

CUOZBLAS_MI355X_LEVEL123(double, D)
CUOZBLAS_MI355X_LEVEL123(float, S)
#undef CUOZBLAS_MI355X_LEVEL123

// ---- the mixed instantiations (binary64 data, binary32 slices) -----------------------------------
template <>
inline int32_t cuozblasRnrm2<double, float, float>(cuozblasHandle_t *oh, const int32_t n, const double *devX,
                                                   const int32_t incx, double *ret) {
	return cuozblas_status(ozblasDSnrm2(oh, n, devX, incx, ret), "cuozblasRnrm2");
}
template <>
inline int32_t cuozblasRdot<double, float, float>(cuozblasHandle_t *oh, const int32_t n, const double *devA,
                                                  const int32_t incx, const double *devB,
                                                  const int32_t incy, double *ret) {
	return cuozblas_status(ozblasDSdot(oh, n, devA, incx, devB, incy, ret), "cuozblasRdot");
}
template <>
inline int32_t cuozblasRgemv<double, float, float>(cuozblasHandle_t *oh, const char tranA, const int32_t m,
                                                   const int32_t n, const double alpha, const double *devA,
                                                   const int32_t lda, const double *devB, const int32_t incx,
                                                   const double beta, double *devC, const int32_t incy) {
	return cuozblas_status(ozblasDSgemv(oh, tranA, m, n, alpha, devA, lda, devB, incx, beta, devC, incy),
	                       "cuozblasRgemv");
}
template <>
inline int32_t cuozblasRgemm<double, float, float>(cuozblasHandle_t *oh, const char tranA, const char tranB,
                                                   const int32_t m, const int32_t n, const int32_t k,
                                                   const double alpha, const double *devA, const int32_t lda,
                                                   const double *devB, const int32_t ldb, const double beta,
                                                   double *devC, const int32_t ldc) {
	return cuozblas_status(
	    ozblasDSgemm(oh, tranA, tranB, m, n, k, alpha, devA, lda, devB, ldb, beta, devC, ldc),
	    "cuozblasRgemm");
}
template <>
inline int32_t cuozblasRcsrmv<double, float>(cuozblasHandle_t *oh, const char tranA, const int32_t m,
                                             const int32_t n, const int32_t nnz, const double alpha,
                                             const cusparseMatDescr_t descrA, const double *devA,
                                             const int32_t *devAcolind, const int32_t *devArowptr,
                                             const double *devB, const double beta, double *devC) {
	return cuozblas_status(ozblasDScsrmv(oh, tranA, m, n, nnz, alpha, descrA, devA, devAcolind, devArowptr,
	                                     devB, beta, devC),
	                       "cuozblasRcsrmv");
}
template <>
inline float *cuozblasRcsrmvSplitA<double, float>(cuozblasHandle_t *oh, const char tranA, const int32_t m,
                                                  const int32_t n, const int32_t nnz,
                                                  const cusparseMatDescr_t descrA, const double *devA,
                                                  const int32_t *devArowptr) {
	float *p = ozblasDScsrmvSplitA(oh, tranA, m, n, nnz, descrA, devA, devArowptr);
	if (!p) cuozblas_status(oh->lastStatus, "cuozblasRcsrmvSplitA");
	return p;
}
template <>
inline int32_t cuozblasRcg<double, float>(cuozblasHandle_t *oh, const char tranA, const int32_t dimN,
                                          const int32_t dimNNZ, const cusparseMatDescr_t descrA,
                                          const double *matA, const int32_t *matAcolind,
                                          const int32_t *matArowptr, const double *vecB, double *vecX,
                                          int32_t maxiter, double tol) {
	return cuozblas_status(ozblasDScg(oh, tranA, dimN, dimNNZ, descrA, matA, matAcolind, matArowptr, vecB,
	                                  vecX, maxiter, tol),
	                       "cuozblasRcg");
}

// ---- the mixed instantiations (binary32 data, binary64-held slices) ------------------------------
template <>
inline int32_t cuozblasRnrm2<float, double, double>(cuozblasHandle_t *oh, const int32_t n, const float *devX,
                                                    const int32_t incx, float *ret) {
	return cuozblas_status(ozblasSDnrm2(oh, n, devX, incx, ret), "cuozblasRnrm2");
}
template <>
inline int32_t cuozblasRdot<float, double, double>(cuozblasHandle_t *oh, const int32_t n, const float *devA,
                                                   const int32_t incx, const float *devB,
                                                   const int32_t incy, float *ret) {
	return cuozblas_status(ozblasSDdot(oh, n, devA, incx, devB, incy, ret), "cuozblasRdot");
}
template <>
inline int32_t cuozblasRgemv<float, double, double>(cuozblasHandle_t *oh, const char tranA, const int32_t m,
                                                    const int32_t n, const float alpha, const float *devA,
                                                    const int32_t lda, const float *devB, const int32_t incx,
                                                    const float beta, float *devC, const int32_t incy) {
	return cuozblas_status(ozblasSDgemv(oh, tranA, m, n, alpha, devA, lda, devB, incx, beta, devC, incy),
	                       "cuozblasRgemv");
}
template <>
inline int32_t cuozblasRgemm<float, double, double>(cuozblasHandle_t *oh, const char tranA, const char tranB,
                                                    const int32_t m, const int32_t n, const int32_t k,
                                                    const float alpha, const float *devA, const int32_t lda,
                                                    const float *devB, const int32_t ldb, const float beta,
                                                    float *devC, const int32_t ldc) {
	return cuozblas_status(
	    ozblasSDgemm(oh, tranA, tranB, m, n, k, alpha, devA, lda, devB, ldb, beta, devC, ldc),
	    "cuozblasRgemm");
}
template <>
inline int32_t cuozblasRcsrmv<float, double>(cuozblasHandle_t *oh, const char tranA, const int32_t m,
                                             const int32_t n, const int32_t nnz, const float alpha,
                                             const cusparseMatDescr_t descrA, const float *devA,
                                             const int32_t *devAcolind, const int32_t *devArowptr,
                                             const float *devB, const float beta, float *devC) {
	return cuozblas_status(ozblasSDcsrmv(oh, tranA, m, n, nnz, alpha, descrA, devA, devAcolind, devArowptr,
	                                     devB, beta, devC),
	                       "cuozblasRcsrmv");
}
template <>
inline double *cuozblasRcsrmvSplitA<float, double>(cuozblasHandle_t *oh, const char tranA, const int32_t m,
                                                   const int32_t n, const int32_t nnz,
                                                   const cusparseMatDescr_t descrA, const float *devA,
                                                   const int32_t *devArowptr) {
	double *p = ozblasSDcsrmvSplitA(oh, tranA, m, n, nnz, descrA, devA, devArowptr);
	if (!p) cuozblas_status(oh->lastStatus, "cuozblasRcsrmvSplitA");
	return p;
}
template <>
inline int32_t cuozblasRcg<float, double>(cuozblasHandle_t *oh, const char tranA, const int32_t dimN,
                                          const int32_t dimNNZ, const cusparseMatDescr_t descrA,
                                          const float *matA, const int32_t *matAcolind,
                                          const int32_t *matArowptr, const float *vecB, float *vecX,
                                          int32_t maxiter, float tol) {
	return cuozblas_status(ozblasSDcg(oh, tranA, dimN, dimNNZ, descrA, matA, matAcolind, matArowptr, vecB,
	                                  vecX, maxiter, tol),
	                       "cuozblasRcg");
}

#endif
