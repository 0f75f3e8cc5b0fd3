// oz_level12.hip -- bandwidth-bound kernels of the plain form: vector split, slice DOT, CSR
// slice products with fused NearSum/Fsum, AXPY.
#include <hip/hip_cooperative_groups.h>

#include "oz_kernels.hpp"
#include "oz_nearsum.hpp"

namespace cg = cooperative_groups;

namespace oz {

template <int BS>
__device__ inline double block_max_d(double v, double *red) {
	for (int o = 32; o > 0; o >>= 1) v = std_max(v, __shfl_xor(v, o, 64));
	const int w = threadIdx.x / 64, l = threadIdx.x % 64;
	__syncthreads();
	if (l == 0) red[w] = v;
	__syncthreads();
	double r = red[0];
	for (int i = 1; i < BS / 64; i++) r = std_max(r, red[i]);
	return r;
}

template <int BS>
__device__ inline double block_sum_d(double v, double *red) {
	for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
	const int w = threadIdx.x / 64, l = threadIdx.x % 64;
	__syncthreads();
	if (l == 0) red[w] = v;
	__syncthreads();
	double r = red[0];
	for (int i = 1; i < BS / 64; i++) r += red[i];
	return r;
}

// ---------------------------------------------------------------------------------------------
// Vector split in one cooperative launch (ozblasSplit with n == 1, reference
// src/cpu/ozblas_split.cpp:75-104,254-289): the tau of slice s comes from the grid-wide max of the
// previous remainder, so every slice ends with a grid barrier.  maxv[] (zeroed by the caller)
// receives max|remainder| after each slice; info[0] the slice count of the reference loop.
// ---------------------------------------------------------------------------------------------
template <int BS>
__global__ __launch_bounds__(BS) void k_split_vec_coop(int n, const double *x, double *rem,
                                                       double *slices, int64_t lds, int rho,
                                                       int shift, int cap, int reproMode,
                                                       const uint16_t *J,
                                                       unsigned long long *maxv, int *info) {
	cg::grid_group grid = cg::this_grid();
	__shared__ double red[BS / 64];
	const int64_t stride = (int64_t)gridDim.x * BS;
	const int64_t i0 = (int64_t)blockIdx.x * BS + threadIdx.x;
	double mx = 0.0;
	for (int64_t i = i0; i < n; i += stride) mx = std_max(mx, fabs(x[i]));
	mx = block_max_d<BS>(mx, red);
	if (threadIdx.x == 0) atomicMax(&maxv[0], (unsigned long long)d2u(mx));
	grid.sync();
	int s;
	bool zero = false;
	for (s = 0; s < cap; s++) {
		const double m = __builtin_bit_cast(double, maxv[s]);
		const int tau = tau_of(m, J);
		const double sigma = split_sigma(rho, tau, shift);
		const double *src = (s == 0) ? x : rem;
		double *sl = slices + (int64_t)s * lds;
		double lm = 0.0;
		for (int64_t i = i0; i < n; i += stride) {
			const double v = src[i];
			const double a = (v + sigma) - sigma;
			const double rr = v - a;
			sl[i] = a;
			rem[i] = rr;
			lm = std_max(lm, fabs(rr));
		}
		lm = block_max_d<BS>(lm, red);
		if (threadIdx.x == 0) atomicMax(&maxv[s + 1], (unsigned long long)d2u(lm));
		grid.sync();
		if (__builtin_bit_cast(double, maxv[s + 1]) == 0.0) {
			zero = true;
			break;
		}
	}
	const int last = zero ? s : cap - 1;
	double *slot = slices + (int64_t)(last + 1) * lds;
	for (int64_t i = i0; i < n; i += stride) slot[i] = zero ? 0.0 : rem[i];
	if (blockIdx.x == 0 && threadIdx.x == 0) {
		// reference count: first s >= 1 with all-zero remainder, +1 (slice 1 is computed even when
		// slice 0 is already exact); cap slices when never exact; +1 remainder slot for reproMode 0
		int count;
		if (cap <= 1)
			count = 1;
		else if (zero)
			count = (s < 1 ? 1 : s) + 1;
		else
			count = cap;
		if (reproMode == 0) count++;
		info[0] = count;
		info[1] = zero ? s : cap;
	}
}

// zero slices [z+2, count) of a vector (the reference splits the zero remainder into them)
__global__ void k_zero_vec_slots(int n, double *slices, int64_t lds, const int *info) {
	const int count = info[0], z = info[1];
	for (int s = z + 2 + blockIdx.y; s < count; s += gridDim.y)
		for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
		     i += (int64_t)gridDim.x * blockDim.x)
			slices[(int64_t)s * lds + i] = 0.0;
}

int vec_split_grid() {
	static int g = 0;
	if (g == 0) {
		int dev = 0, ncu = 0, per = 0;
		(void)hipGetDevice(&dev);
		(void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
		(void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_split_vec_coop<256>, 256, 0);
		if (per > 2) per = 2;
		g = ncu * (per > 0 ? per : 1);
		if (g < 1) g = 1;
	}
	return g;
}

// returns hipSuccess or the launch error; grid is clamped to the vector size
hipError_t launch_split_vec_coop(hipStream_t st, int n, const double *x, double *rem,
                                 double *slices, int64_t lds, int rho, int shift, int cap,
                                 int reproMode, const uint16_t *J, unsigned long long *maxv,
                                 int *info) {
	int grid = vec_split_grid();
	const int need = (n + 255) / 256;
	if (need < grid) grid = need < 1 ? 1 : need;
	void *args[] = {&n, (void *)&x, &rem, &slices, &lds, &rho, &shift, &cap, &reproMode,
	                (void *)&J, &maxv, &info};
	hipError_t e = hipLaunchCooperativeKernel((const void *)k_split_vec_coop<256>, dim3(grid),
	                                          dim3(256), args, 0, st);
	if (e != hipSuccess) return e;
	hipLaunchKernelGGL(k_zero_vec_slots, dim3((n + 255) / 256 < 256 ? (n + 255) / 256 : 256, 4),
	                   dim3(256), 0, st, n, slices, lds, (const int *)info);
	return hipGetLastError();
}

// ---------------------------------------------------------------------------------------------
// slice DOT: out[ia + ib*nA] = sum_i xs_ia[i] * ys_ib[i] (the reference computes these with one
// cblas_dgemm, src/cpu/ozblas_gemm.cpp:209-215).  Every partial sum is an integer multiple of the
// product unit below 2^53, so block partials added with FP64 atomics are exact in any order.
// ---------------------------------------------------------------------------------------------
constexpr int kDotChunk = 16;

template <int BS>
__global__ __launch_bounds__(BS) void k_slice_dots(int n, const double *xs, int64_t ldx, int nA,
                                                   const double *ys, int64_t ldy, int nB,
                                                   double *out) {
	__shared__ double red[BS / 64];
	const int npairs = nA * nB;
	const int p0 = blockIdx.y * kDotChunk;
	double acc[kDotChunk];
#pragma unroll
	for (int q = 0; q < kDotChunk; q++) acc[q] = 0.0;
	for (int64_t i = blockIdx.x * (int64_t)BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * BS) {
#pragma unroll
		for (int q = 0; q < kDotChunk; q++) {
			const int p = p0 + q;
			if (p < npairs) {
				const int ia = p % nA, ib = p / nA;
				acc[q] = fma(xs[(int64_t)ia * ldx + i], ys[(int64_t)ib * ldy + i], acc[q]);
			}
		}
	}
#pragma unroll
	for (int q = 0; q < kDotChunk; q++) {
		const double v = block_sum_d<BS>(acc[q], red);
		if (threadIdx.x == 0 && p0 + q < npairs && v != 0.0) atomicAdd(&out[p0 + q], v);
	}
}

void launch_slice_dots(hipStream_t st, int n, const double *xs, int64_t ldx, int nA,
                       const double *ys, int64_t ldy, int nB, double *out) {
	int gx = (n + 255) / 256;
	if (gx > 512) gx = 512;
	if (gx < 1) gx = 1;
	const int gy = (nA * nB + kDotChunk - 1) / kDotChunk;
	hipLaunchKernelGGL(k_slice_dots<256>, dim3(gx, gy), dim3(256), 0, st, n, xs, ldx, nA, ys, ldy,
	                   nB, out);
}

// GEMM with k = 0: every slice product is zero, so the reference's sum is +0 and the epilogue
// leaves C = fma(alpha, +0, beta * C)
__global__ void k_gemm_k0(int m, int n, double alpha, double beta, double *C, int64_t ldc) {
	const int64_t total = (int64_t)m * n;
	for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
	     e += (int64_t)gridDim.x * blockDim.x) {
		double *c = C + (e / m) * ldc + e % m;
		*c = fma(alpha, 0.0, beta * *c);
	}
}

void launch_gemm_k0(hipStream_t st, int m, int n, double alpha, double beta, double *C, int64_t ldc) {
	int64_t g = ((int64_t)m * n + 255) / 256;
	if (g > 4096) g = 4096;
	if (g < 1) g = 1;
	hipLaunchKernelGGL(k_gemm_k0, dim3((unsigned)g), dim3(256), 0, st, m, n, alpha, beta, C, ldc);
}

// ---------------------------------------------------------------------------------------------
// Dot2 mode (splitEpsMode = 2): error-free transformations of src/cpu/eft.h (scalar forms) and the
// double-double slice dots.  Compiled without contraction, so every expression rounds as written.
// ---------------------------------------------------------------------------------------------
struct dd {
	double x, y;
};
__device__ inline void two_sum(double a, double b, double &s, double &e) {  // TwoSum
	s = a + b;
	const double v = s - a;
	e = (a - (s - v)) + (b - v);
}
__device__ inline void quick_two_sum(double a, double b, double &s, double &e) {  // QuickTwoSum
	s = a + b;
	e = b - (s - a);
}
__device__ inline void dot2i(double a, double b, dd &c) {  // Dot2i: TwoProdFMA + TwoSum
	const double h = a * b;
	const double r = fma(a, b, -h);
	double q;
	two_sum(c.x, h, c.x, q);
	c.y = c.y + (q + r);
}
__device__ inline dd quad_add(dd a, dd b) {  // QuadAdd
	dd t, c;
	two_sum(a.x, b.x, t.x, t.y);
	t.y = t.y + (a.y + b.y);
	quick_two_sum(t.x, t.y, c.x, c.y);
	return c;
}
__device__ inline dd quad_mul(double a, dd b) {  // QuadMul(double, double2)
	dd t, c;
	t.x = a * b.x;
	t.y = fma(a, b.x, -t.x);
	t.y = fma(a, b.y, t.y);
	quick_two_sum(t.x, t.y, c.x, c.y);
	return c;
}

// Dot2 DOT (dgemm_tn_skinny_x2, src/cpu/ozblas_blas_wrapper.cpp:109-152): the reference sums each
// slice dot in OpenMP per-thread partials over a static schedule of k, combined by QuadAdd in a
// critical section.  This is that computation for kDot2Chunks threads entering the section in
// thread order: partial t covers libgomp's static chunk t of [0, n).
constexpr int kDot2Chunks = 1024;

__global__ __launch_bounds__(256) void k_dot2_partials(int n, const double *xs, int64_t ldx, int nA,
                                                       const double *ys, int64_t ldy, dd *part) {
	const int t = blockIdx.x * blockDim.x + threadIdx.x;  // chunk
	const int pr = blockIdx.y;                            // pair: ia + nA * ib
	if (t >= kDot2Chunks) return;
	const int ia = pr % nA, ib = pr / nA;
	const int64_t q = n / kDot2Chunks, r = n % kDot2Chunks;
	const int64_t lo = t * q + (t < r ? t : r), hi = lo + q + (t < r ? 1 : 0);
	const double *x = xs + (int64_t)ia * ldx, *y = ys + (int64_t)ib * ldy;
	dd L{0.0, 0.0};
	for (int64_t k = lo; k < hi; k++) dot2i(x[k], y[k], L);
	part[(int64_t)pr * kDot2Chunks + t] = L;
}

// one thread per pair: QuadAdd of the partials in chunk order, QuadMul(1, .), then hi to
// vals[nA * ib + ia] and lo to vals[nA * (nB + ib) + ia] (C and C + n*ldc of the reference)
__global__ void k_dot2_combine(int nA, int nB, const dd *part, double *vals) {
	const int pr = blockIdx.x * blockDim.x + threadIdx.x;
	if (pr >= nA * nB) return;
	const int ia = pr % nA, ib = pr / nA;
	dd G{0.0, 0.0};
	for (int t = 0; t < kDot2Chunks; t++) G = quad_add(G, part[(int64_t)pr * kDot2Chunks + t]);
	G = quad_mul(1.0, G);
	vals[nA * ib + ia] = G.x;
	vals[nA * (nB + ib) + ia] = G.y;
}

size_t dot2_partials_bytes(int nA, int nB) { return sizeof(dd) * (size_t)nA * nB * kDot2Chunks; }

void launch_dot2_slices(hipStream_t st, int n, const double *xs, int64_t ldx, int nA,
                        const double *ys, int64_t ldy, int nB, void *part, double *vals) {
	hipLaunchKernelGGL(k_dot2_partials, dim3((kDot2Chunks + 255) / 256, nA * nB), dim3(256), 0, st, n,
	                   xs, ldx, nA, ys, ldy, (dd *)part);
	hipLaunchKernelGGL(k_dot2_combine, dim3((nA * nB + 63) / 64), dim3(64), 0, st, nA, nB,
	                   (const dd *)part, vals);
}

// DOT / NRM2 summation of the nA x nB slice dots (sumOrder 2, src/cpu/ozblas_gemm.cpp:281-288),
// one thread; C[0] = fma(alpha, t, beta * C[0])
__global__ void k_dot_sum(int nvals, double *vals, double alpha, double beta, double *C, int nA,
                          int nB, int maxlevel, int fsum, const uint16_t *J) {
	if (threadIdx.x != 0 || blockIdx.x != 0) return;
	double t;
	if (fsum) {
		t = 0.;
		for (int ik = 0; ik <= maxlevel; ik++)
			for (int ia = 0; ia < nA; ia++) {
				const int ib = ik - ia;
				if (ib < 0 || ib >= nB) continue;
				t = t + vals[nA * ib + ia];
			}
	} else {
		StridedVec v{vals, 1};
		t = nearsum(nvals, v, J);
	}
	C[0] = fma(alpha, t, beta * C[0]);
}

void launch_nearsum_small(hipStream_t st, int nvals, double *vals, double alpha, double beta,
                          double *C, int sumOrderFsum, int nA, int nB, int maxlevel, int fsum,
                          const uint16_t *J) {
	(void)sumOrderFsum;
	hipLaunchKernelGGL(k_dot_sum, dim3(1), dim3(64), 0, st, nvals, vals, alpha, beta, C, nA, nB,
	                   maxlevel, fsum, J);
}

// ---------------------------------------------------------------------------------------------
// CSR slice products + global sum fused per row (ozblasRcsrmv, src/cpu/ozblas_csrmv.cpp:113-166):
// v[ia][ib] = sum_j As_ia[j] * Xs_ib[col_j] (fma in row order as csrmm_n, exact), then NearSum
// over the pairs with ia + ib <= maxlevel (or Fsum in (level, ia, ib) order), then
// y = fma(alpha, t, beta * y).
// ---------------------------------------------------------------------------------------------
template <int NAM, int NBM>
__global__ __launch_bounds__(256) void k_csrmm_sum(int m, const double *As, int64_t ldas, int nA,
                                                   const int *colind, const int *rowptr,
                                                   const double *Xs, int64_t ldxs, int nB,
                                                   int maxlevel, int fsum, double alpha,
                                                   double beta, double *y, const uint16_t *J) {
	const int r = blockIdx.x * blockDim.x + threadIdx.x;
	if (r >= m) return;
	const int b = rowptr[r], e = rowptr[r + 1];
	RegVec<NAM * NBM> v;
#pragma unroll
	for (int q = 0; q < NAM * NBM; q++) v[q] = 0.0;
	for (int jj = b; jj < e; jj++) {
		const int c = colind[jj];
		double a[NAM], x[NBM];
#pragma unroll
		for (int ia = 0; ia < NAM; ia++) a[ia] = (ia < nA) ? As[(int64_t)ia * ldas + jj] : 0.0;
#pragma unroll
		for (int ib = 0; ib < NBM; ib++) x[ib] = (ib < nB) ? Xs[(int64_t)ib * ldxs + c] : 0.0;
#pragma unroll
		for (int ia = 0; ia < NAM; ia++)
#pragma unroll
			for (int ib = 0; ib < NBM; ib++) v[ia * NBM + ib] = fma(a[ia], x[ib], v[ia * NBM + ib]);
	}
	double t;
	if (fsum) {
		t = 0.;
#pragma unroll
		for (int ik = 0; ik < NAM + NBM - 1; ik++)
#pragma unroll
			for (int ia = 0; ia < NAM; ia++) {
				const int ib = ik - ia;
				if (ib >= 0 && ib < NBM && ia < nA && ib < nB && ik <= maxlevel) t = t + v[ia * NBM + ib];
			}
	} else {
		// NearSum over the products (its result does not depend on the order of the terms):
		// compact the valid ones to the front of a register array
		RegVec<NAM * NBM> u;
		int cnt = 0;
#pragma unroll
		for (int ia = 0; ia < NAM; ia++)
#pragma unroll
			for (int ib = 0; ib < NBM; ib++) u[ia * NBM + ib] = 0.0;
#pragma unroll
		for (int ia = 0; ia < NAM; ia++)
#pragma unroll
			for (int ib = 0; ib < NBM; ib++) {
				const bool ok = ia < nA && ib < nB && ia + ib <= maxlevel;
#pragma unroll
				for (int q = 0; q < NAM * NBM; q++)
					if (ok && q == cnt) u[q] = v[ia * NBM + ib];
				cnt += ok ? 1 : 0;
			}
		t = nearsum(cnt, u, J);
	}
	y[r] = fma(alpha, t, beta * y[r]);
}

void launch_csrmm_sum(hipStream_t st, int m, int nnz, const double *As, int64_t ldas, int nA,
                      const int *colind, const int *rowptr, const double *Xs, int64_t ldxs, int nB,
                      int maxlevel, int fsum, double alpha, double beta, double *y,
                      const uint16_t *J) {
	(void)nnz;
	const dim3 g((m + 255) / 256), b(256);
	if (nA <= 4 && nB <= 4)
		hipLaunchKernelGGL((k_csrmm_sum<4, 4>), g, b, 0, st, m, As, ldas, nA, colind, rowptr, Xs,
		                   ldxs, nB, maxlevel, fsum, alpha, beta, y, J);
	else
		hipLaunchKernelGGL((k_csrmm_sum<4, 8>), g, b, 0, st, m, As, ldas, nA, colind, rowptr, Xs,
		                   ldxs, nB, maxlevel, fsum, alpha, beta, y, J);
}

// Dot2 CSRMV (csrmm_n_x2, src/cpu/ozblas_blas_wrapper.cpp:262-295, then the global sum over
// 2 nB "slices" of src/cpu/ozblas_csrmv.cpp:159-162): per row and slice pair, Dot2 over the row's
// nonzeros in order and QuadMul(1, .); w[ia][ib'] = hi of pair (ia, ib') for ib' < nB and lo of
// pair (ia, ib' - nB) otherwise; NearSum over all w, or Fsum in (level ia + ib', ia) order.
template <int NAM, int NBM>
__global__ __launch_bounds__(256) void k_csrmm_sum_x2(int m, const double *As, int64_t ldas, int nA,
                                                      const int *colind, const int *rowptr,
                                                      const double *Xs, int64_t ldxs, int nB,
                                                      int fsum, double alpha, double beta, double *y,
                                                      const uint16_t *J) {
	const int r = blockIdx.x * blockDim.x + threadIdx.x;
	if (r >= m) return;
	const int b = rowptr[r], e = rowptr[r + 1];
	dd v[NAM][NBM];
#pragma unroll
	for (int ia = 0; ia < NAM; ia++)
#pragma unroll
		for (int ib = 0; ib < NBM; ib++) v[ia][ib] = dd{0.0, 0.0};
	for (int jj = b; jj < e; jj++) {
		const int c = colind[jj];
		double a[NAM], x[NBM];
#pragma unroll
		for (int ia = 0; ia < NAM; ia++) a[ia] = (ia < nA) ? As[(int64_t)ia * ldas + jj] : 0.0;
#pragma unroll
		for (int ib = 0; ib < NBM; ib++) x[ib] = (ib < nB) ? Xs[(int64_t)ib * ldxs + c] : 0.0;
#pragma unroll
		for (int ia = 0; ia < NAM; ia++)
#pragma unroll
			for (int ib = 0; ib < NBM; ib++) dot2i(a[ia], x[ib], v[ia][ib]);
	}
	double fx[NAM][NBM], fy[NAM][NBM];  // renormalised (hi, lo) of every pair: QuadMul(1, .)
#pragma unroll
	for (int ia = 0; ia < NAM; ia++)
#pragma unroll
		for (int ib = 0; ib < NBM; ib++) {
			const dd f = quad_mul(1.0, v[ia][ib]);
			fx[ia][ib] = f.x;
			fy[ia][ib] = f.y;
		}
	double t;
	if (fsum) {
		t = 0.;
		for (int ik = 0; ik <= (nA - 1) + (2 * nB - 1); ik++)
			for (int ia = 0; ia < nA; ia++) {
				const int ibp = ik - ia;
				if (ibp < 0 || ibp >= 2 * nB) continue;
				const bool hiPart = ibp < nB;
				const int ib = hiPart ? ibp : ibp - nB;
				double w = 0.0;
#pragma unroll
				for (int qa = 0; qa < NAM; qa++)
#pragma unroll
					for (int qb = 0; qb < NBM; qb++)
						if (qa == ia && qb == ib) w = hiPart ? fx[qa][qb] : fy[qa][qb];
				t = t + w;
			}
	} else {
		RegVec<NAM * 2 * NBM> u;  // the 2 nA nB values, compacted to the front
#pragma unroll
		for (int q = 0; q < NAM * 2 * NBM; q++) u[q] = 0.0;
		int cnt = 0;
#pragma unroll
		for (int ia = 0; ia < NAM; ia++)
#pragma unroll
			for (int ib = 0; ib < NBM; ib++) {
				const bool ok = ia < nA && ib < nB;
#pragma unroll
				for (int q = 0; q < NAM * 2 * NBM; q++) {
					if (ok && q == cnt) u[q] = fx[ia][ib];
					if (ok && q == cnt + 1) u[q] = fy[ia][ib];
				}
				cnt += ok ? 2 : 0;
			}
		t = nearsum(cnt, u, J);
	}
	y[r] = fma(alpha, t, beta * y[r]);
}

void launch_csrmm_sum_x2(hipStream_t st, int m, const double *As, int64_t ldas, int nA,
                         const int *colind, const int *rowptr, const double *Xs, int64_t ldxs, int nB,
                         int fsum, double alpha, double beta, double *y, const uint16_t *J) {
	hipLaunchKernelGGL((k_csrmm_sum_x2<4, 4>), dim3((m + 255) / 256), dim3(256), 0, st, m, As, ldas,
	                   nA, colind, rowptr, Xs, ldxs, nB, fsum, alpha, beta, y, J);
}

// general case: store the products where blasRcsrmm puts them (column ic of CSplit, ld ldcs)
__global__ void k_csrmm_store(int m, const double *As, int64_t ldas, int nA, const int *colind,
                              const int *rowptr, const double *Xs, int64_t ldxs, int nB,
                              int maxlevel, double *Cs, int64_t ldcs) {
	const int r = blockIdx.x * blockDim.x + threadIdx.x;
	if (r >= m) return;
	const int b = rowptr[r], e = rowptr[r + 1];
	int col = 0;
	const int na = (maxlevel + 1 < nA) ? maxlevel + 1 : nA;
	for (int ia = 0; ia < na; ia++) {
		const int numB = (nB < maxlevel + 1 - ia) ? nB : maxlevel + 1 - ia;
		for (int ib = 0; ib < numB; ib++) {
			double tr = 0.;
			for (int jj = b; jj < e; jj++)
				tr = fma(As[(int64_t)ia * ldas + jj], Xs[(int64_t)ib * ldxs + colind[jj]], tr);
			Cs[(int64_t)(col + ib) * ldcs + r] = tr;
		}
		col += numB;
	}
}

void launch_csrmm_store(hipStream_t st, int m, const double *As, int64_t ldas, int nA,
                        const int *colind, const int *rowptr, const double *Xs, int64_t ldxs,
                        int nB, int maxlevel, double *Cs, int64_t ldcs) {
	hipLaunchKernelGGL(k_csrmm_store, dim3((m + 255) / 256), dim3(256), 0, st, m, As, ldas, nA,
	                   colind, rowptr, Xs, ldxs, nB, maxlevel, Cs, ldcs);
}

// ozblasRaxpy with reproMode != 0 (src/cpu/ozblas_axpy.cpp:142-144): y = fma(alpha, x, y)
__global__ void k_axpy(int n, double alpha, const double *x, double *y) {
	for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
	     i += (int64_t)gridDim.x * blockDim.x)
		y[i] = fma(alpha, x[i], y[i]);
}
// y = y + alpha * x without contraction (the cblas_daxpy path of reproMode == 0)
__global__ void k_axpy_plain(int n, double alpha, const double *x, double *y) {
	for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
	     i += (int64_t)gridDim.x * blockDim.x)
		y[i] = y[i] + alpha * x[i];
}
// CG update p = beta * p (cblas_dscal, src/cpu/ozblas_cg.cpp:97), then p = fma(1, r, p) (:98)
__global__ void k_scal_axpy(int n, double beta, const double *r, double *p) {
	for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
	     i += (int64_t)gridDim.x * blockDim.x)
		p[i] = fma(1.0, r[i], beta * p[i]);
}

static int grid1(int64_t n) {
	int64_t g = (n + 255) / 256;
	if (g > 2048) g = 2048;
	return g < 1 ? 1 : (int)g;
}

void launch_axpy_host(hipStream_t st, int n, double alpha, const double *x, double *y,
                      int fused) {
	if (fused)
		hipLaunchKernelGGL(k_axpy, dim3(grid1(n)), dim3(256), 0, st, n, alpha, x, y);
	else
		hipLaunchKernelGGL(k_axpy_plain, dim3(grid1(n)), dim3(256), 0, st, n, alpha, x, y);
}

void launch_scal_axpy(hipStream_t st, int n, double beta, const double *r, double *p) {
	hipLaunchKernelGGL(k_scal_axpy, dim3(grid1(n)), dim3(256), 0, st, n, beta, r, p);
}

}  // namespace oz
