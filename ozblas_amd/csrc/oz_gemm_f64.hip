// oz_gemm_f64.hip -- FP64 slice-pair products on v_mfma_f64_16x16x4f64 (plain form).
//
// Every slice product of the plain form is exact (the split keeps 53-rho bits per slice, reference
// src/cpu/ozblas_split.cpp:20-34), so the order of the k reduction does not change a bit: the MFMA
// result equals the reference's cblas_dgemm of the same slices (src/cpu/ozblas_blas_wrapper.cpp:158).
//
// Layout: both operands are k-contiguous ("TN"): A_b is K x M with column i = row i of op(A),
// B_b is K x N.  Workgroup = 4 waves, output tile 128 x 128, each wave 64 x 64 = 4 x 4 MFMA tiles.
// The MFMA is issued as (B fragment) x (A fragment) so that the lane index runs along the rows of
// the column-major C and stores are contiguous.
//
// k_gemm_f64<false>: one launch computes every (pair, tile) product and stores it (NearSum input).
// k_gemm_f64<true> : fused Fsum -- the pairs are walked in the reference's (level, ia, ib) order
//                    and t += product is accumulated in registers, then C = fma(alpha, t, beta*C)
//                    (src/cpu/ozblas_sum.cpp:354-378), so nothing is materialised.
// TC = float serves the binary32 instantiations: their slices (stored as doubles) give exact FP64
// products that round to float as the reference's float slice GEMM (<float,float>) or its
// (float) of the binary64 slice GEMM followed by the exponent scaling (<float,double>: f32_term);
// the Fsum then adds in float and the epilogue is fma1<float> (the double fma, rounded to float).
#include "oz_kernels.hpp"

namespace oz {

typedef double d4v __attribute__((ext_vector_type(4)));

// result element (lane l, register r) of mfma_f64_16x16x4(X, Y): X-side index P = 4r + l/16,
// Y-side index Q = l%16 (operands: lane l holds X[l%16][k = l/16] and Y[k = l/16][l%16]); measured
// on gfx950 by matching all lane/register hypotheses against a host product
#define F64_P(l, r) (4 * (r) + (l) / 16)
#define F64_Q(l, r) ((l) % 16)

constexpr int FBM = 128, FBN = 128, FBK = 16;
constexpr int FLD = FBK + 1;  // padded LDS row (doubles)

__device__ inline void f64_load_rows(double (&reg)[8], const double *P, int64_t ld, int row0,
                                     int nrows, int kb, int K) {
	const int t = threadIdx.x;
	const int row = row0 + t / 2, k = kb + 8 * (t % 2);
	if (row < nrows && k + 8 <= K) {
		const double *src = P + (int64_t)row * ld + k;
#pragma unroll
		for (int q = 0; q < 4; q++) {
			const double2 v = *(const double2 *)(src + 2 * q);
			reg[2 * q] = v.x;
			reg[2 * q + 1] = v.y;
		}
	} else {
#pragma unroll
		for (int q = 0; q < 8; q++)
			reg[q] = (row < nrows && k + q < K) ? P[(int64_t)row * ld + k + q] : 0.0;
	}
}

__device__ inline void f64_store_lds(const double (&reg)[8], double *s) {
	const int t = threadIdx.x;
	double *dst = s + (t / 2) * FLD + 8 * (t % 2);
#pragma unroll
	for (int q = 0; q < 8; q++) dst[q] = reg[q];
}

template <bool FSUM, class TC>
__global__ __launch_bounds__(256) void k_gemm_f64(int M, int N, int K, const double *A, int64_t lda,
                                                  const double *B, int64_t ldb, TC *C,
                                                  int64_t ldc, const int64_t *aOff,
                                                  const int64_t *bOff, const int64_t *cOff,
                                                  int npairs, TC alpha, TC beta, int tilesM) {
	__shared__ double sA[FBM * FLD];
	__shared__ double sB[FBN * FLD];
	const int i0 = (blockIdx.x % tilesM) * FBM, j0 = (blockIdx.x / tilesM) * FBN;
	const int w = threadIdx.x / 64, l = threadIdx.x % 64;
	const int wm = (w % 2) * 64, wn = (w / 2) * 64;
	d4v acc[4][4];
	TC tot[4][4][4];
	const int pbeg = FSUM ? 0 : blockIdx.y, pend = FSUM ? npairs : blockIdx.y + 1;
	for (int a = 0; a < 4; a++)
		for (int b = 0; b < 4; b++)
			for (int r = 0; r < 4; r++) tot[a][b][r] = (TC)0;
	for (int p = pbeg; p < pend; p++) {
		const double *Ap = A + aOff[p];
		const double *Bp = B + bOff[p];
		for (int a = 0; a < 4; a++)
			for (int b = 0; b < 4; b++) acc[a][b] = d4v{0.0, 0.0, 0.0, 0.0};
		double ra[8], rb[8];
		// a stage holds 128 rows x 16 k of each operand: thread t loads half (t%2) of row t/2
		f64_load_rows(ra, Ap, lda, i0, M, 0, K);
		f64_load_rows(rb, Bp, ldb, j0, N, 0, K);
		for (int kb = 0; kb < K; kb += FBK) {
			__syncthreads();
			f64_store_lds(ra, sA);
			f64_store_lds(rb, sB);
			__syncthreads();
			if (kb + FBK < K) {
				f64_load_rows(ra, Ap, lda, i0, M, kb + FBK, K);
				f64_load_rows(rb, Bp, ldb, j0, N, kb + FBK, K);
			}
#pragma unroll
			for (int kk = 0; kk < FBK / 4; kk++) {
				double af[4], bf[4];
#pragma unroll
				for (int mi = 0; mi < 4; mi++) af[mi] = sA[(wm + 16 * mi + l % 16) * FLD + 4 * kk + l / 16];
#pragma unroll
				for (int nj = 0; nj < 4; nj++) bf[nj] = sB[(wn + 16 * nj + l % 16) * FLD + 4 * kk + l / 16];
#pragma unroll
				for (int mi = 0; mi < 4; mi++)
#pragma unroll
					for (int nj = 0; nj < 4; nj++)
						acc[mi][nj] = __builtin_amdgcn_mfma_f64_16x16x4f64(bf[nj], af[mi], acc[mi][nj], 0, 0, 0);
			}
		}
		if (FSUM) {
			for (int a = 0; a < 4; a++)
				for (int b = 0; b < 4; b++)
					for (int r = 0; r < 4; r++) tot[a][b][r] = tot[a][b][r] + sum_term<TC>(acc[a][b][r]);
		}
	}
	TC *Cp = FSUM ? C : C + cOff[blockIdx.y];
	for (int mi = 0; mi < 4; mi++)
		for (int nj = 0; nj < 4; nj++)
			for (int r = 0; r < 4; r++) {
				const int i = i0 + wm + 16 * mi + F64_Q(l, r);
				const int j = j0 + wn + 16 * nj + F64_P(l, r);
				if (i < M && j < N) {
					TC *c = Cp + (int64_t)j * ldc + i;
					if (FSUM)
						*c = fma1_t(alpha, tot[mi][nj][r], beta * *c);
					else
						*c = sum_term<TC>(acc[mi][nj][r]);
				}
			}
}

template <class TC>
static void gemm_f64_batched(hipStream_t st, int M, int N, int K, const double *A, int64_t lda,
                             const double *B, int64_t ldb, TC *C, int64_t ldc, const int64_t *aOff,
                             const int64_t *bOff, const int64_t *cOff, int batch) {
	const int tilesM = (M + FBM - 1) / FBM, tilesN = (N + FBN - 1) / FBN;
	hipLaunchKernelGGL((k_gemm_f64<false, TC>), dim3(tilesM * tilesN, batch), dim3(256), 0, st, M, N,
	                   K, A, lda, B, ldb, C, ldc, aOff, bOff, cOff, batch, (TC)1, (TC)0, tilesM);
}

template <class TC>
static void gemm_f64_fsum(hipStream_t st, int M, int N, int K, const double *A, int64_t lda,
                          const double *B, int64_t ldb, const int64_t *aOff, const int64_t *bOff,
                          int npairs, TC alpha, TC beta, TC *C, int64_t ldc) {
	const int tilesM = (M + FBM - 1) / FBM, tilesN = (N + FBN - 1) / FBN;
	hipLaunchKernelGGL((k_gemm_f64<true, TC>), dim3(tilesM * tilesN, 1), dim3(256), 0, st, M, N, K, A,
	                   lda, B, ldb, C, ldc, aOff, bOff, (const int64_t *)nullptr, npairs, alpha,
	                   beta, tilesM);
}

void launch_gemm_f64_batched(hipStream_t st, int M, int N, int K, const double *A, int64_t lda,
                             const double *B, int64_t ldb, double *C, int64_t ldc,
                             const int64_t *aOff, const int64_t *bOff, const int64_t *cOff,
                             int batch) {
	gemm_f64_batched(st, M, N, K, A, lda, B, ldb, C, ldc, aOff, bOff, cOff, batch);
}
void launch_gemm_f64_batched(hipStream_t st, int M, int N, int K, const double *A, int64_t lda,
                             const double *B, int64_t ldb, float *C, int64_t ldc,
                             const int64_t *aOff, const int64_t *bOff, const int64_t *cOff,
                             int batch) {
	gemm_f64_batched(st, M, N, K, A, lda, B, ldb, C, ldc, aOff, bOff, cOff, batch);
}

void launch_gemm_f64_fsum(hipStream_t st, int M, int N, int K, const double *A, int64_t lda,
                          const double *B, int64_t ldb, const int64_t *aOff, const int64_t *bOff,
                          int npairs, double alpha, double beta, double *C, int64_t ldc) {
	gemm_f64_fsum(st, M, N, K, A, lda, B, ldb, aOff, bOff, npairs, alpha, beta, C, ldc);
}
void launch_gemm_f64_fsum(hipStream_t st, int M, int N, int K, const double *A, int64_t lda,
                          const double *B, int64_t ldb, const int64_t *aOff, const int64_t *bOff,
                          int npairs, float alpha, float beta, float *C, int64_t ldc) {
	gemm_f64_fsum(st, M, N, K, A, lda, B, ldb, aOff, bOff, npairs, alpha, beta, C, ldc);
}

}  // namespace oz
