// oz_sum.hip -- global summation of materialised slice products (plain form).
//
// One thread per output element (consecutive threads -> consecutive rows, so every product plane
// is read coalesced).  NearSum restates ozblasGlobalNearsumKernel (reference
// src/cpu/ozblas_sum.cpp:231-271), Fsum restates ozblasGlobalFsumKernel (:331-380) including the
// sumOrder index maps; both end with C = fma(alpha, t, beta * C), which reads C even for beta = 0.
// The binary32 kernels are the <float> instantiations of the same functions.
#include "oz_kernels.hpp"
#include "oz_nearsum.hpp"

namespace oz {

template <int NR>
__global__ __launch_bounds__(256) void k_global_nearsum_reg(int m, int n, int ic, double *Cs,
                                                            int64_t llsc, int64_t ldsc, double *C,
                                                            int64_t ldc, double alpha,
                                                            double beta, const uint16_t *J) {
	const int64_t total = (int64_t)m * n;
	for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
	     e += (int64_t)gridDim.x * blockDim.x) {
		const int x = (int)(e % m), y = (int)(e / m);
		const double *src = Cs + (int64_t)y * ldsc + x;
		RegVec<NR> v;
#pragma unroll
		for (int i = 0; i < NR; i++) v[i] = (i < ic) ? src[(int64_t)i * llsc] : 0.0;
		const double t = nearsum(ic, v, J);
		double *c = C + (int64_t)y * ldc + x;
		*c = fma(alpha, t, beta * *c);
	}
}

__global__ __launch_bounds__(256) void k_global_nearsum_mem(int m, int n, int ic, double *Cs,
                                                            int64_t llsc, int64_t ldsc, double *C,
                                                            int64_t ldc, double alpha,
                                                            double beta, const uint16_t *J) {
	const int64_t total = (int64_t)m * n;
	for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
	     e += (int64_t)gridDim.x * blockDim.x) {
		const int x = (int)(e % m), y = (int)(e / m);
		StridedVec v{Cs + (int64_t)y * ldsc + x, llsc};
		const double t = nearsum(ic, v, J);
		double *c = C + (int64_t)y * ldc + x;
		*c = fma(alpha, t, beta * *c);
	}
}

__global__ __launch_bounds__(256) void k_global_fsum(int m, int n, int nA, int nB, int maxlevel,
                                                     int sumOrder, const double *Cs, int64_t llsc,
                                                     int64_t ldsc, double *C, int64_t ldc,
                                                     double alpha, double beta) {
	const int64_t total = (int64_t)m * n;
	for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
	     e += (int64_t)gridDim.x * blockDim.x) {
		const int x = (int)(e % m), y = (int)(e / m);
		double t = 0.;
		int ic = 0;
		for (int ik = 0; ik <= maxlevel; ik++)
			for (int ia = 0; ia < nA; ia++) {
				const int ib = ik - ia;
				if (ib < 0 || ib >= nB) continue;
				int it;
				switch (sumOrder) {
				case 2: it = nA * ib + ia; break;
				case 3: it = nB * ia + ib; break;
				default: it = ic; break;
				}
				t = t + Cs[llsc * it + (int64_t)y * ldsc + x];
				ic++;
			}
		double *c = C + (int64_t)y * ldc + x;
		*c = fma(alpha, t, beta * *c);
	}
}

// ---- binary32 (<float> instantiations): NearSum<float> with the log2f table JF2, Fsum in float,
// epilogue fma1<float>(alpha, t, beta * C) with beta * C a float product
template <int NR>
__global__ __launch_bounds__(256) void k_global_nearsum_reg_f(int m, int n, int ic, float *Cs,
                                                              int64_t llsc, int64_t ldsc, float *C,
                                                              int64_t ldc, float alpha, float beta,
                                                              const uint16_t *JF2) {
	const int64_t total = (int64_t)m * n;
	for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
	     e += (int64_t)gridDim.x * blockDim.x) {
		const int x = (int)(e % m), y = (int)(e / m);
		const float *src = Cs + (int64_t)y * ldsc + x;
		RegVec<NR, float> v;
#pragma unroll
		for (int i = 0; i < NR; i++) v[i] = (i < ic) ? src[(int64_t)i * llsc] : 0.f;
		const float t = nearsum_f(ic, v, JF2);
		float *c = C + (int64_t)y * ldc + x;
		*c = fma1_f(alpha, t, beta * *c);
	}
}

__global__ __launch_bounds__(256) void k_global_nearsum_mem_f(int m, int n, int ic, float *Cs,
                                                              int64_t llsc, int64_t ldsc, float *C,
                                                              int64_t ldc, float alpha, float beta,
                                                              const uint16_t *JF2) {
	const int64_t total = (int64_t)m * n;
	for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
	     e += (int64_t)gridDim.x * blockDim.x) {
		const int x = (int)(e % m), y = (int)(e / m);
		StridedVecF v{Cs + (int64_t)y * ldsc + x, llsc};
		const float t = nearsum_f(ic, v, JF2);
		float *c = C + (int64_t)y * ldc + x;
		*c = fma1_f(alpha, t, beta * *c);
	}
}

__global__ __launch_bounds__(256) void k_global_fsum_f(int m, int n, int nA, int nB, int maxlevel,
                                                       int sumOrder, const float *Cs, int64_t llsc,
                                                       int64_t ldsc, float *C, int64_t ldc,
                                                       float alpha, float beta) {
	const int64_t total = (int64_t)m * n;
	for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
	     e += (int64_t)gridDim.x * blockDim.x) {
		const int x = (int)(e % m), y = (int)(e / m);
		float t = 0.f;
		int ic = 0;
		for (int ik = 0; ik <= maxlevel; ik++)
			for (int ia = 0; ia < nA; ia++) {
				const int ib = ik - ia;
				if (ib < 0 || ib >= nB) continue;
				int it;
				switch (sumOrder) {
				case 2: it = nA * ib + ia; break;
				case 3: it = nB * ia + ib; break;
				default: it = ic; break;
				}
				t = t + Cs[llsc * it + (int64_t)y * ldsc + x];
				ic++;
			}
		float *c = C + (int64_t)y * ldc + x;
		*c = fma1_f(alpha, t, beta * *c);
	}
}

static inline int grid_for(int64_t total) {
	int64_t g = (total + 255) / 256;
	if (g > 8192) g = 8192;
	if (g < 1) g = 1;
	return (int)g;
}

void launch_global_nearsum(hipStream_t st, int m, int n, int ic, double *Cs, int64_t llsc,
                           int64_t ldsc, double *C, int64_t ldc, double alpha, double beta,
                           const uint16_t *J) {
	const int g = grid_for((int64_t)m * n);
	if (ic <= 16)
		hipLaunchKernelGGL(k_global_nearsum_reg<16>, dim3(g), dim3(256), 0, st, m, n, ic, Cs, llsc,
		                   ldsc, C, ldc, alpha, beta, J);
	else if (ic <= 32)
		hipLaunchKernelGGL(k_global_nearsum_reg<32>, dim3(g), dim3(256), 0, st, m, n, ic, Cs, llsc,
		                   ldsc, C, ldc, alpha, beta, J);
	else if (ic <= 64)
		hipLaunchKernelGGL(k_global_nearsum_reg<64>, dim3(g), dim3(256), 0, st, m, n, ic, Cs, llsc,
		                   ldsc, C, ldc, alpha, beta, J);
	else
		hipLaunchKernelGGL(k_global_nearsum_mem, dim3(g), dim3(256), 0, st, m, n, ic, Cs, llsc,
		                   ldsc, C, ldc, alpha, beta, J);
}

void launch_global_fsum(hipStream_t st, int m, int n, int nA, int nB, int maxlevel, int sumOrder,
                        const double *Cs, int64_t llsc, int64_t ldsc, double *C, int64_t ldc,
                        double alpha, double beta) {
	const int g = grid_for((int64_t)m * n);
	hipLaunchKernelGGL(k_global_fsum, dim3(g), dim3(256), 0, st, m, n, nA, nB, maxlevel, sumOrder,
	                   Cs, llsc, ldsc, C, ldc, alpha, beta);
}

void launch_global_nearsum(hipStream_t st, int m, int n, int ic, float *Cs, int64_t llsc,
                           int64_t ldsc, float *C, int64_t ldc, float alpha, float beta,
                           const uint16_t *JF2) {
	const int g = grid_for((int64_t)m * n);
	if (ic <= 16)
		hipLaunchKernelGGL(k_global_nearsum_reg_f<16>, dim3(g), dim3(256), 0, st, m, n, ic, Cs, llsc,
		                   ldsc, C, ldc, alpha, beta, JF2);
	else if (ic <= 32)
		hipLaunchKernelGGL(k_global_nearsum_reg_f<32>, dim3(g), dim3(256), 0, st, m, n, ic, Cs, llsc,
		                   ldsc, C, ldc, alpha, beta, JF2);
	else
		hipLaunchKernelGGL(k_global_nearsum_mem_f, dim3(g), dim3(256), 0, st, m, n, ic, Cs, llsc,
		                   ldsc, C, ldc, alpha, beta, JF2);
}

void launch_global_fsum(hipStream_t st, int m, int n, int nA, int nB, int maxlevel, int sumOrder,
                        const float *Cs, int64_t llsc, int64_t ldsc, float *C, int64_t ldc,
                        float alpha, float beta) {
	const int g = grid_for((int64_t)m * n);
	hipLaunchKernelGGL(k_global_fsum_f, dim3(g), dim3(256), 0, st, m, n, nA, nB, maxlevel, sumOrder,
	                   Cs, llsc, ldsc, C, ldc, alpha, beta);
}

}  // namespace oz
