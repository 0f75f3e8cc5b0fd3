// shim_check.cpp -- calls every cuOzBLAS template of include/cuozblas_mi355x.hpp (binary64,
// binary32 and mixed) and the C entry point it forwards to on the same inputs, and compares the outputs byte
// for byte.  Built by ozblas_amd/build.py into ozblas_amd/_build/shim_check; run by
// tests/test_gpu_shim.py.  Prints one line per routine and "shim check: all ok" on success.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "cuozblas_mi355x.hpp"

#define CHECK_HIP(x)                                                          \
	do {                                                                      \
		if ((x) != hipSuccess) {                                              \
			fprintf(stderr, "HIP error %s at %s:%d\n", #x, __FILE__, __LINE__); \
			return 2;                                                         \
		}                                                                     \
	} while (0)

static uint64_t g_state = 88172645463325252ull;
static double next_uniform() {  // xorshift64 in [-1, 1)
	g_state ^= g_state << 13;
	g_state ^= g_state >> 7;
	g_state ^= g_state << 17;
	return (double)(g_state >> 11) * 0x1p-52 - 1.0;
}

template <class T>
struct Dev {
	T *p = nullptr;
	size_t n = 0;
	explicit Dev(const std::vector<T> &h) : n(h.size()) {
		if (hipMalloc(&p, sizeof(T) * (n ? n : 1)) == hipSuccess && n)
			(void)hipMemcpy(p, h.data(), sizeof(T) * n, hipMemcpyHostToDevice);
	}
	~Dev() { (void)hipFree(p); }
	std::vector<T> get() const {
		std::vector<T> h(n);
		if (n) (void)hipMemcpy(h.data(), p, sizeof(T) * n, hipMemcpyDeviceToHost);
		return h;
	}
};

template <class T>
static std::vector<T> rand_vec(size_t n, double scale) {
	std::vector<T> v(n);
	for (auto &x : v) x = (T)(next_uniform() * scale);
	return v;
}

template <class T>
static bool same(const std::vector<T> &a, const std::vector<T> &b) {
	return a.size() == b.size() && std::memcmp(a.data(), b.data(), sizeof(T) * a.size()) == 0;
}

static int g_fail = 0;
static void report(const char *prec, const char *what, bool ok, int32_t s1, int32_t s2) {
	const bool good = ok && s1 == 0 && s2 == 0;
	printf("shim %s %-12s %s (status %d %d)\n", prec, what, good ? "ok" : "MISMATCH", (int)s1, (int)s2);
	fflush(stdout);
	if (!good) g_fail = 1;
}

// direct C entry points per precision
struct CD {
	using T = double;
	static constexpr const char *name = "D";
	static int32_t gemm(ozblasHandle_t *h, char a, char b, int m, int n, int k, T al, const T *A, int la,
	                    const T *B, int lb, T be, T *C, int lc) {
		return ozblasDgemm(h, a, b, m, n, k, al, A, la, B, lb, be, C, lc);
	}
	static int32_t gemv(ozblasHandle_t *h, char a, int m, int n, T al, const T *A, int la, const T *x, T be,
	                    T *y) {
		return ozblasDgemv(h, a, m, n, al, A, la, x, 1, be, y, 1);
	}
	static int32_t dot(ozblasHandle_t *h, int n, const T *x, const T *y, T *r) {
		return ozblasDdot(h, n, x, 1, y, 1, r);
	}
	static int32_t nrm2(ozblasHandle_t *h, int n, const T *x, T *r) { return ozblasDnrm2(h, n, x, 1, r); }
	static int32_t axpy(ozblasHandle_t *h, int n, T a, const T *x, T *y) {
		return ozblasDaxpy(h, n, a, x, 1, y, 1);
	}
	static int32_t csrmv(ozblasHandle_t *h, int m, int n, int nnz, T al, const T *A, const int *ci,
	                     const int *rp, const T *x, T be, T *y) {
		return ozblasDcsrmv(h, 'n', m, n, nnz, al, nullptr, A, ci, rp, x, be, y);
	}
	static T *split(ozblasHandle_t *h, int m, int n, int nnz, const T *A, const int *rp) {
		return ozblasDcsrmvSplitA(h, 'n', m, n, nnz, nullptr, A, rp);
	}
	static int32_t cg(ozblasHandle_t *h, int n, int nnz, const T *A, const int *ci, const int *rp, const T *b,
	                  T *x, int it, T tol) {
		return ozblasDcg(h, 'n', n, nnz, nullptr, A, ci, rp, b, x, it, tol);
	}
};
struct CS {
	using T = float;
	static constexpr const char *name = "S";
	static int32_t gemm(ozblasHandle_t *h, char a, char b, int m, int n, int k, T al, const T *A, int la,
	                    const T *B, int lb, T be, T *C, int lc) {
		return ozblasSgemm(h, a, b, m, n, k, al, A, la, B, lb, be, C, lc);
	}
	static int32_t gemv(ozblasHandle_t *h, char a, int m, int n, T al, const T *A, int la, const T *x, T be,
	                    T *y) {
		return ozblasSgemv(h, a, m, n, al, A, la, x, 1, be, y, 1);
	}
	static int32_t dot(ozblasHandle_t *h, int n, const T *x, const T *y, T *r) {
		return ozblasSdot(h, n, x, 1, y, 1, r);
	}
	static int32_t nrm2(ozblasHandle_t *h, int n, const T *x, T *r) { return ozblasSnrm2(h, n, x, 1, r); }
	static int32_t axpy(ozblasHandle_t *h, int n, T a, const T *x, T *y) {
		return ozblasSaxpy(h, n, a, x, 1, y, 1);
	}
	static int32_t csrmv(ozblasHandle_t *h, int m, int n, int nnz, T al, const T *A, const int *ci,
	                     const int *rp, const T *x, T be, T *y) {
		return ozblasScsrmv(h, 'n', m, n, nnz, al, nullptr, A, ci, rp, x, be, y);
	}
	static T *split(ozblasHandle_t *h, int m, int n, int nnz, const T *A, const int *rp) {
		return ozblasScsrmvSplitA(h, 'n', m, n, nnz, nullptr, A, rp);
	}
	static int32_t cg(ozblasHandle_t *h, int n, int nnz, const T *A, const int *ci, const int *rp, const T *b,
	                  T *x, int it, T tol) {
		return ozblasScg(h, 'n', n, nnz, nullptr, A, ci, rp, b, x, it, tol);
	}
};

// the mixed routines: binary64 data, binary32 slices (templates at <double,float,...>)
struct CDS {
	static int32_t gemm(ozblasHandle_t *h, char a, char b, int m, int n, int k, double al, const double *A,
	                    int la, const double *B, int lb, double be, double *C, int lc) {
		return ozblasDSgemm(h, a, b, m, n, k, al, A, la, B, lb, be, C, lc);
	}
};

static int run_mixed() {
	cuozblasHandle_t h1, h2;
	cuozblasCreate(&h1, 256ull << 20);
	if (ozblasCreate(&h2, 256ull << 20) != 0) return 2;
	h1.sumMode = h2.sumMode = 0;  // Fsum: the only sum of the mixed routines
	const int m = 37, n = 29, k = 53;
	{
		const int lda = m + 3, ldb = n + 1, ldc = m + 2;
		Dev<double> A(rand_vec<double>((size_t)lda * k, 4.0)), B(rand_vec<double>((size_t)ldb * k, 0.5));
		const std::vector<double> C0 = rand_vec<double>((size_t)ldc * n, 1.0);
		Dev<double> C1(C0), C2(C0);
		const int32_t s1 = cuozblasRgemm<double, float, float>(&h1, 'n', 't', m, n, k, 1.5, A.p, lda, B.p,
		                                                       ldb, -0.25, C1.p, ldc);
		const int32_t s2 = CDS::gemm(&h2, 'n', 't', m, n, k, 1.5, A.p, lda, B.p, ldb, -0.25, C2.p, ldc);
		report("DS", "gemm", same(C1.get(), C2.get()) && !same(C1.get(), C0), s1, s2);
	}
	{
		Dev<double> A(rand_vec<double>((size_t)m * n, 2.0)), x(rand_vec<double>(m, 1.0));
		const std::vector<double> y0 = rand_vec<double>(n, 1.0);
		Dev<double> y1(y0), y2(y0);
		const int32_t s1 = cuozblasRgemv<double, float, float>(&h1, 't', m, n, 0.75, A.p, m, x.p, 1, 2.0, y1.p, 1);
		const int32_t s2 = ozblasDSgemv(&h2, 't', m, n, 0.75, A.p, m, x.p, 1, 2.0, y2.p, 1);
		report("DS", "gemv", same(y1.get(), y2.get()), s1, s2);
	}
	{
		const int len = 1001;
		Dev<double> x(rand_vec<double>(len, 3.0)), y(rand_vec<double>(len, 1.0));
		double r1 = 0, r2 = 0;
		int32_t s1 = cuozblasRdot<double, float, float>(&h1, len, x.p, 1, y.p, 1, &r1);
		int32_t s2 = ozblasDSdot(&h2, len, x.p, 1, y.p, 1, &r2);
		report("DS", "dot", std::memcmp(&r1, &r2, 8) == 0, s1, s2);
		s1 = cuozblasRnrm2<double, float, float>(&h1, len, x.p, 1, &r1);
		s2 = ozblasDSnrm2(&h2, len, x.p, 1, &r2);
		report("DS", "nrm2", std::memcmp(&r1, &r2, 8) == 0 && r1 > 0, s1, s2);
	}
	{
		const int nn = 300;
		std::vector<int> rp(nn + 1), ci;
		std::vector<double> val;
		for (int r = 0; r < nn; r++) {
			rp[r] = (int)ci.size();
			for (int c = r - 1; c <= r + 1; c++) {
				if (c < 0 || c >= nn) continue;
				ci.push_back(c);
				val.push_back(c == r ? 2.001 : -1.0);
			}
		}
		rp[nn] = (int)ci.size();
		const int nnz = (int)val.size();
		Dev<double> A(val), x(rand_vec<double>(nn, 1.0)), b(rand_vec<double>(nn, 1.0));
		Dev<int> CI(ci), RP(rp);
		const std::vector<double> y0 = rand_vec<double>(nn, 1.0);
		Dev<double> y1(y0), y2(y0);
		int32_t s1 = cuozblasRcsrmv<double, float>(&h1, 'n', nn, nn, nnz, 1.5, nullptr, A.p, CI.p, RP.p, x.p,
		                                           -0.5, y1.p);
		int32_t s2 = ozblasDScsrmv(&h2, 'n', nn, nn, nnz, 1.5, nullptr, A.p, CI.p, RP.p, x.p, -0.5, y2.p);
		report("DS", "csrmv", same(y1.get(), y2.get()), s1, s2);
		float *sp1 = cuozblasRcsrmvSplitA<double, float>(&h1, 'n', nn, nn, nnz, nullptr, A.p, RP.p);
		float *sp2 = ozblasDScsrmvSplitA(&h2, 'n', nn, nn, nnz, nullptr, A.p, RP.p);
		Dev<double> z1(y0), z2(y0);
		s1 = sp1 ? cuozblasRcsrmv<double, float>(&h1, 'n', nn, nn, nnz, 1.0, nullptr, (const double *)sp1,
		                                         CI.p, RP.p, x.p, 0.0, z1.p)
		         : h1.lastStatus;
		s2 = sp2 ? ozblasDScsrmv(&h2, 'n', nn, nn, nnz, 1.0, nullptr, (const double *)sp2, CI.p, RP.p, x.p,
		                         0.0, z2.p)
		         : h2.lastStatus;
		h1.memMaskSplitA = h2.memMaskSplitA = 0;
		report("DS", "csrmvSplitA", sp1 && sp2 && same(z1.get(), z2.get()), s1, s2);
		const std::vector<double> x0(nn, 0.0);
		Dev<double> xc1(x0), xc2(x0);
		s1 = cuozblasRcg<double, float>(&h1, 'n', nn, nnz, nullptr, A.p, CI.p, RP.p, b.p, xc1.p, 25, 0.0);
		s2 = ozblasDScg(&h2, 'n', nn, nnz, nullptr, A.p, CI.p, RP.p, b.p, xc2.p, 25, 0.0);
		report("DS", "cg", same(xc1.get(), xc2.get()) && h1.cg_numiter == 25, s1, s2);
	}
	cuozblasDestroy(&h1);
	ozblasDestroy(&h2);
	return 0;
}

// the mixed routines: binary32 data, binary64-held slices (templates at <float,double,...>)
static int run_mixed_sd() {
	cuozblasHandle_t h1, h2;
	cuozblasCreate(&h1, 256ull << 20);
	if (ozblasCreate(&h2, 256ull << 20) != 0) return 2;
	h1.sumMode = h2.sumMode = 0;  // Fsum: the only sum of the mixed routines
	const int m = 37, n = 29, k = 153;
	{
		const int lda = m + 3, ldb = n + 1, ldc = m + 2;
		Dev<float> A(rand_vec<float>((size_t)lda * k, 4.0)), B(rand_vec<float>((size_t)ldb * k, 0.5));
		const std::vector<float> C0 = rand_vec<float>((size_t)ldc * n, 1.0);
		Dev<float> C1(C0), C2(C0);
		const int32_t s1 = cuozblasRgemm<float, double, double>(&h1, 'n', 't', m, n, k, 1.5f, A.p, lda, B.p,
		                                                        ldb, -0.25f, C1.p, ldc);
		const int32_t s2 = ozblasSDgemm(&h2, 'n', 't', m, n, k, 1.5f, A.p, lda, B.p, ldb, -0.25f, C2.p, ldc);
		report("SD", "gemm", same(C1.get(), C2.get()) && !same(C1.get(), C0), s1, s2);
	}
	{
		Dev<float> A(rand_vec<float>((size_t)k * n, 2.0)), x(rand_vec<float>(k, 1.0));
		const std::vector<float> y0 = rand_vec<float>(n, 1.0);
		Dev<float> y1(y0), y2(y0);
		const int32_t s1 =
		    cuozblasRgemv<float, double, double>(&h1, 't', k, n, 0.75f, A.p, k, x.p, 1, 2.0f, y1.p, 1);
		const int32_t s2 = ozblasSDgemv(&h2, 't', k, n, 0.75f, A.p, k, x.p, 1, 2.0f, y2.p, 1);
		report("SD", "gemv", same(y1.get(), y2.get()), s1, s2);
	}
	{
		const int len = 1001;
		Dev<float> x(rand_vec<float>(len, 3.0)), y(rand_vec<float>(len, 1.0));
		float r1 = 0, r2 = 0;
		int32_t s1 = cuozblasRdot<float, double, double>(&h1, len, x.p, 1, y.p, 1, &r1);
		int32_t s2 = ozblasSDdot(&h2, len, x.p, 1, y.p, 1, &r2);
		report("SD", "dot", std::memcmp(&r1, &r2, 4) == 0, s1, s2);
		s1 = cuozblasRnrm2<float, double, double>(&h1, len, x.p, 1, &r1);
		s2 = ozblasSDnrm2(&h2, len, x.p, 1, &r2);
		report("SD", "nrm2", std::memcmp(&r1, &r2, 4) == 0 && r1 > 0, s1, s2);
	}
	{
		const int nn = 300;
		std::vector<int> rp(nn + 1), ci;
		std::vector<float> val;
		for (int r = 0; r < nn; r++) {
			rp[r] = (int)ci.size();
			for (int c = r - 1; c <= r + 1; c++) {
				if (c < 0 || c >= nn) continue;
				ci.push_back(c);
				val.push_back(c == r ? 2.001f : -1.0f);
			}
		}
		rp[nn] = (int)ci.size();
		const int nnz = (int)val.size();
		Dev<float> A(val), x(rand_vec<float>(nn, 1.0)), b(rand_vec<float>(nn, 1.0));
		Dev<int> CI(ci), RP(rp);
		const std::vector<float> y0 = rand_vec<float>(nn, 1.0);
		Dev<float> y1(y0), y2(y0);
		int32_t s1 = cuozblasRcsrmv<float, double>(&h1, 'n', nn, nn, nnz, 1.5f, nullptr, A.p, CI.p, RP.p, x.p,
		                                           -0.5f, y1.p);
		int32_t s2 = ozblasSDcsrmv(&h2, 'n', nn, nn, nnz, 1.5f, nullptr, A.p, CI.p, RP.p, x.p, -0.5f, y2.p);
		report("SD", "csrmv", same(y1.get(), y2.get()), s1, s2);
		double *sp1 = cuozblasRcsrmvSplitA<float, double>(&h1, 'n', nn, nn, nnz, nullptr, A.p, RP.p);
		double *sp2 = ozblasSDcsrmvSplitA(&h2, 'n', nn, nn, nnz, nullptr, A.p, RP.p);
		Dev<float> z1(y0), z2(y0);
		s1 = sp1 ? cuozblasRcsrmv<float, double>(&h1, 'n', nn, nn, nnz, 1.0f, nullptr, (const float *)sp1,
		                                         CI.p, RP.p, x.p, 0.0f, z1.p)
		         : h1.lastStatus;
		s2 = sp2 ? ozblasSDcsrmv(&h2, 'n', nn, nn, nnz, 1.0f, nullptr, (const float *)sp2, CI.p, RP.p, x.p,
		                         0.0f, z2.p)
		         : h2.lastStatus;
		h1.memMaskSplitA = h2.memMaskSplitA = 0;
		report("SD", "csrmvSplitA", sp1 && sp2 && same(z1.get(), z2.get()), s1, s2);
		const std::vector<float> x0(nn, 0.0f);
		Dev<float> xc1(x0), xc2(x0);
		s1 = cuozblasRcg<float, double>(&h1, 'n', nn, nnz, nullptr, A.p, CI.p, RP.p, b.p, xc1.p, 25, 0.0f);
		s2 = ozblasSDcg(&h2, 'n', nn, nnz, nullptr, A.p, CI.p, RP.p, b.p, xc2.p, 25, 0.0f);
		report("SD", "cg", same(xc1.get(), xc2.get()) && h1.cg_numiter == 25, s1, s2);
	}
	cuozblasDestroy(&h1);
	ozblasDestroy(&h2);
	return 0;
}

template <class P>
static int run() {
	using T = typename P::T;
	const char *pn = P::name;
	cuozblasHandle_t h1, h2;
	cuozblasCreate(&h1, 256ull << 20);
	if (ozblasCreate(&h2, 256ull << 20) != 0) return 2;
	h1.sumMode = h2.sumMode = 1;
	const int m = 37, n = 29, k = 53;
	{  // GEMM 'n','t': A m x k (lda m + 3), B n x k (ldb n + 1), C m x n (ldc m + 2)
		const int lda = m + 3, ldb = n + 1, ldc = m + 2;
		Dev<T> A(rand_vec<T>((size_t)lda * k, 4.0)), B(rand_vec<T>((size_t)ldb * k, 0.5));
		const std::vector<T> C0 = rand_vec<T>((size_t)ldc * n, 1.0);
		Dev<T> C1(C0), C2(C0);
		const int32_t s1 = cuozblasRgemm<T, T, T>(&h1, 'n', 't', m, n, k, (T)1.5, A.p, lda, B.p, ldb, (T)-0.25,
		                                          C1.p, ldc);
		const int32_t s2 = P::gemm(&h2, 'n', 't', m, n, k, (T)1.5, A.p, lda, B.p, ldb, (T)-0.25, C2.p, ldc);
		report(pn, "gemm", same(C1.get(), C2.get()) && !same(C1.get(), C0), s1, s2);
	}
	{  // GEMV 't': A m x n (lda m), x of m, y of n
		Dev<T> A(rand_vec<T>((size_t)m * n, 2.0)), x(rand_vec<T>(m, 1.0));
		const std::vector<T> y0 = rand_vec<T>(n, 1.0);
		Dev<T> y1(y0), y2(y0);
		const int32_t s1 = cuozblasRgemv<T, T, T>(&h1, 't', m, n, (T)0.75, A.p, m, x.p, 1, (T)2, y1.p, 1);
		const int32_t s2 = P::gemv(&h2, 't', m, n, (T)0.75, A.p, m, x.p, (T)2, y2.p);
		report(pn, "gemv", same(y1.get(), y2.get()), s1, s2);
	}
	{  // DOT, NRM2, AXPY
		const int len = 1001;
		Dev<T> x(rand_vec<T>(len, 3.0)), y(rand_vec<T>(len, 1.0));
		T r1 = 0, r2 = 0;
		int32_t s1 = cuozblasRdot<T, T, T>(&h1, len, x.p, 1, y.p, 1, &r1);
		int32_t s2 = P::dot(&h2, len, x.p, y.p, &r2);
		report(pn, "dot", std::memcmp(&r1, &r2, sizeof(T)) == 0, s1, s2);
		s1 = cuozblasRnrm2<T, T, T>(&h1, len, x.p, 1, &r1);
		s2 = P::nrm2(&h2, len, x.p, &r2);
		report(pn, "nrm2", std::memcmp(&r1, &r2, sizeof(T)) == 0 && r1 > 0, s1, s2);
		const std::vector<T> y0 = y.get();
		Dev<T> ya(y0), yb(y0);
		s1 = cuozblasRaxpy<T>(&h1, len, (T)-1.25, x.p, 1, ya.p, 1);
		s2 = P::axpy(&h2, len, (T)-1.25, x.p, yb.p);
		report(pn, "axpy", same(ya.get(), yb.get()), s1, s2);
	}
	{  // CSRMV, CSRMV with the pre-split matrix, CG: 1-D Laplacian (2.001 on the diagonal, -1 beside)
		const int nn = 300;
		std::vector<int> rp(nn + 1), ci;
		std::vector<T> val;
		for (int r = 0; r < nn; r++) {
			rp[r] = (int)ci.size();
			for (int c = r - 1; c <= r + 1; c++) {
				if (c < 0 || c >= nn) continue;
				ci.push_back(c);
				val.push_back(c == r ? (T)2.001 : (T)-1);
			}
		}
		rp[nn] = (int)ci.size();
		const int nnz = (int)val.size();
		Dev<T> A(val), x(rand_vec<T>(nn, 1.0)), b(rand_vec<T>(nn, 1.0));
		Dev<int> CI(ci), RP(rp);
		const std::vector<T> y0 = rand_vec<T>(nn, 1.0);
		Dev<T> y1(y0), y2(y0);
		int32_t s1 = cuozblasRcsrmv<T, T>(&h1, 'n', nn, nn, nnz, (T)1.5, nullptr, A.p, CI.p, RP.p, x.p, (T)-0.5,
		                                  y1.p);
		int32_t s2 = P::csrmv(&h2, nn, nn, nnz, (T)1.5, A.p, CI.p, RP.p, x.p, (T)-0.5, y2.p);
		report(pn, "csrmv", same(y1.get(), y2.get()), s1, s2);
		T *sp1 = cuozblasRcsrmvSplitA<T, T>(&h1, 'n', nn, nn, nnz, nullptr, A.p, RP.p);
		T *sp2 = P::split(&h2, nn, nn, nnz, A.p, RP.p);
		Dev<T> z1(y0), z2(y0);
		s1 = sp1 ? cuozblasRcsrmv<T, T>(&h1, 'n', nn, nn, nnz, (T)1, nullptr, sp1, CI.p, RP.p, x.p, (T)0, z1.p)
		         : h1.lastStatus;
		s2 = sp2 ? P::csrmv(&h2, nn, nn, nnz, (T)1, sp2, CI.p, RP.p, x.p, (T)0, z2.p) : h2.lastStatus;
		h1.memMaskSplitA = h2.memMaskSplitA = 0;
		report(pn, "csrmvSplitA", sp1 && sp2 && same(z1.get(), z2.get()), s1, s2);
		const std::vector<T> x0(nn, (T)0);
		Dev<T> xc1(x0), xc2(x0);
		s1 = cuozblasRcg<T, T>(&h1, 'n', nn, nnz, nullptr, A.p, CI.p, RP.p, b.p, xc1.p, 25, (T)0);
		s2 = P::cg(&h2, nn, nnz, A.p, CI.p, RP.p, b.p, xc2.p, 25, (T)0);
		report(pn, "cg", same(xc1.get(), xc2.get()) && h1.cg_numiter == 25, s1, s2);
	}
	cuozblasDestroy(&h1);
	ozblasDestroy(&h2);
	return 0;
}

int main() {
	if (run<CD>() != 0 || run<CS>() != 0 || run_mixed() != 0 || run_mixed_sd() != 0) return 2;
	if (g_fail) {
		printf("shim check: FAILED\n");
		return 1;
	}
	printf("shim check: all ok\n");
	return 0;
}
