// oz_split.hip -- error-free splitting kernels.
//
// Plain form (FP64 slices): restates ozblasSplit / ozblasSplitCKernel / ozblasSplitVecKernel /
// ozblasSplitSparse (reference src/cpu/ozblas_split.cpp:75-289,756-928) bit for bit, for the
// binary64 and the binary32 instantiations (binary32 slices are stored as doubles):
//   tau = (short)ceil(log2(max)), sigma = 0.75 * 2^(rho+tau) * splitShift,
//   slice = (x + sigma) - sigma, remainder = x - slice, next max = max |remainder|.
// Because tau is per column (dense) or per row (CSR), one workgroup (dense) or one thread (CSR)
// runs every slice of its column/row in a single launch; the global slice count (the reference
// loops until the sum of all column maxima is zero) is max over columns, read back once.
// Vectors share one tau across all elements, so their split is one grid-wide pass per slice.
//
// Fast forms (INT8/FP16/BF16): exact balanced radix-256 digit planes, see k_digit_*.
#include "oz_digits.hpp"

namespace oz {

// ---------------------------------------------------------------------------------------------
// dense: one workgroup per column of a k-contiguous operand (column j = in + j*ldi, length m)
// ---------------------------------------------------------------------------------------------
// TI = float: the binary32 split (float arithmetic, float remainders), slices stored as doubles;
// maxima are reduced as doubles (exact for floats)
template <int BS, class TI>
__global__ __launch_bounds__(BS) void k_split_cols(int m, int n, const TI *in, int64_t ldi,
                                                   TI *out, int64_t ldo, double *split,
                                                   int64_t lds, int64_t sstride, int rho, int shift,
                                                   int cap, const uint16_t *J, int *zcol, int *zmax) {
	__shared__ double red[BS / 64];
	for (int j = blockIdx.x; j < n; j += gridDim.x) {
		const TI *x = in + (int64_t)j * ldi;
		TI *r = out + (int64_t)j * ldo;
		double *sp = split + (int64_t)j * lds;
		double mx = 0.0;  // ozblasFindMaxCKernel (src/cpu/ozblas_split.cpp:37-52)
		for (int i = threadIdx.x; i < m; i += BS) mx = std_max(mx, fabs((double)x[i]));
		mx = block_max<BS>(mx, red);
		int s = 0;
		bool zero = false;
		for (s = 0; s < cap; s++) {
			const int tau = SplitRule<TI>::tau(mx, J);
			const TI sigma = SplitRule<TI>::sigma(rho, tau, shift);
			const TI *src = (s == 0) ? x : r;
			double lm = 0.0;
			for (int i = threadIdx.x; i < m; i += BS) {
				const TI v = src[i];
				const TI a = (v + sigma) - sigma;
				const TI rr = v - a;
				sp[(int64_t)s * sstride + i] = (double)a;
				r[i] = rr;
				lm = std_max(lm, fabs((double)rr));
			}
			mx = block_max<BS>(lm, red);
			if (mx == 0.0) {
				zero = true;
				break;
			}
		}
		const int last = zero ? s : cap - 1;
		// the reference stores each remainder into the next slot (:166); the last one survives
		for (int i = threadIdx.x; i < m; i += BS)
			sp[(int64_t)(last + 1) * sstride + i] = zero ? 0.0 : (double)r[i];
		if (threadIdx.x == 0) {
			const int z = zero ? s : cap;  // cap == "never reached zero"
			zcol[j] = z;
			atomicMax(zmax, z);
		}
	}
}

// The same split with the column held in registers (R elements per thread, m <= BS * R): the
// remainder never goes back to memory, so a column costs one read and the slice writes (none in
// the counting pass, store = 0).  Same operations, same order per element, same bits.  Slots
// [last + 1, fill) -- the reproMode = 0 remainder slot and the all-zero slices up to the operand's
// slice count -- are written here too; slots the products never read are not.
template <int BS, int R, class TI>
__global__ __launch_bounds__(BS) void k_split_cols_reg(int m, int n, const TI *in, int64_t ldi,
                                                       double *split, int64_t lds, int64_t sstride,
                                                       int rho, int shift, int cap,
                                                       const uint16_t *J, int *zcol, int *zmax,
                                                       int store, int fill) {
	__shared__ double red[BS / 64];
	for (int j = blockIdx.x; j < n; j += gridDim.x) {
		const TI *x = in + (int64_t)j * ldi;
		double *sp = split + (int64_t)j * lds;
		TI v[R];
		double mx = 0.0;
#pragma unroll
		for (int q = 0; q < R; q++) {
			const int i = threadIdx.x + BS * q;
			v[q] = (i < m) ? x[i] : (TI)0;
			mx = std_max(mx, fabs((double)v[q]));
		}
		mx = block_max<BS>(mx, red);
		int s = 0;
		bool zero = false;
		for (s = 0; s < cap; s++) {
			const int tau = SplitRule<TI>::tau(mx, J);
			const TI sigma = SplitRule<TI>::sigma(rho, tau, shift);
			double lm = 0.0;
#pragma unroll
			for (int q = 0; q < R; q++) {
				const int i = threadIdx.x + BS * q;
				if (i < m) {
					const TI a = (v[q] + sigma) - sigma;
					const TI rr = v[q] - a;
					if (store) sp[(int64_t)s * sstride + i] = (double)a;
					v[q] = rr;
					lm = std_max(lm, fabs((double)rr));
				}
			}
			mx = block_max<BS>(lm, red);
			if (mx == 0.0) {
				zero = true;
				break;
			}
		}
		const int last = zero ? s : cap - 1;
		if (store) {
			for (int t = last + 1; t < fill; t++) {
#pragma unroll
				for (int q = 0; q < R; q++) {
					const int i = threadIdx.x + BS * q;
					if (i < m) sp[(int64_t)t * sstride + i] = (zero || t > last + 1) ? 0.0 : (double)v[q];
				}
			}
		}
		if (threadIdx.x == 0) {
			const int z = zero ? s : cap;
			zcol[j] = z;
			atomicMax(zmax, z);
		}
	}
}

// zero slots [zcol[j]+2, count) of every column (what the reference computes by splitting zeros)
__global__ void k_zero_slots(int m, int n, double *split, int64_t lds, int64_t sstride,
                             const int *zcol, int count) {
	const int j = blockIdx.y;
	const int s0 = zcol[j] + 2;
	for (int s = s0 + blockIdx.z; s < count; s += gridDim.z)
		for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x)
			split[(int64_t)s * sstride + (int64_t)j * lds + i] = 0.0;
}

// out (cols x rows, ldo) = in^T (in: rows x cols, ldi), 32x32 tiles through LDS
template <class T>
__global__ __launch_bounds__(256) void k_transpose(int rows, int cols, const T *in, int64_t ldi,
                                                   T *out, int64_t ldo) {
	__shared__ T t[32][33];
	const int bx = blockIdx.x * 32, by = blockIdx.y * 32;  // bx over rows, by over cols
	const int tx = threadIdx.x % 32, ty = threadIdx.x / 32;  // 32 x 8
	for (int k = ty; k < 32; k += 8) {
		const int r = bx + tx, c = by + k;
		if (r < rows && c < cols) t[k][tx] = in[(int64_t)c * ldi + r];
	}
	__syncthreads();
	for (int k = ty; k < 32; k += 8) {
		const int c = by + tx, r = bx + k;
		if (r < rows && c < cols) out[(int64_t)r * ldo + c] = t[tx][k];
	}
}

// ---------------------------------------------------------------------------------------------
// vectors: one tau for the whole vector (ozblasSplitVecKernel, src/cpu/ozblas_split.cpp:75-104)
// ---------------------------------------------------------------------------------------------
__device__ inline double u2d(unsigned long long u) { return __builtin_bit_cast(double, u); }

template <int BS>
__global__ __launch_bounds__(BS) void k_vec_max(int m, const double *x, unsigned long long *slot) {
	__shared__ double red[BS / 64];
	double mx = 0.0;
	for (int i = blockIdx.x * BS + threadIdx.x; i < m; i += gridDim.x * BS)
		mx = std_max(mx, fabs(x[i]));
	mx = block_max<BS>(mx, red);
	if (threadIdx.x == 0) atomicMax(slot, (unsigned long long)d2u(mx));
}

template <int BS>
__global__ __launch_bounds__(BS) void k_split_vec_pass(int m, const double *src, double *rem,
                                                       double *slice, int rho, int shift,
                                                       const uint16_t *J,
                                                       const unsigned long long *maxIn,
                                                       unsigned long long *maxOut) {
	__shared__ double red[BS / 64];
	const int tau = tau_of(u2d(*maxIn), J);
	const double sigma = split_sigma(rho, tau, shift);
	double lm = 0.0;
	for (int i = blockIdx.x * BS + threadIdx.x; i < m; i += gridDim.x * BS) {
		const double v = src[i];
		const double a = (v + sigma) - sigma;
		const double rr = v - a;
		slice[i] = a;
		rem[i] = rr;
		lm = std_max(lm, fabs(rr));
	}
	lm = block_max<BS>(lm, red);
	if (threadIdx.x == 0) atomicMax(maxOut, (unsigned long long)d2u(lm));
}

template <class T>
__global__ void k_fill(T *p, int64_t n, T v) {
	for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
	     i += (int64_t)gridDim.x * blockDim.x)
		p[i] = v;
}

__global__ void k_copy(const double *src, double *dst, int64_t n) {
	for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
	     i += (int64_t)gridDim.x * blockDim.x)
		dst[i] = src[i];
}

// ---------------------------------------------------------------------------------------------
// CSR rows: one thread per row, every slice in one launch (ozblasSplitSparseNKernel :788-829,
// ozblasSplitSparse :896-928); slices at split + s*lds, rho from the row length.  TI = float is
// the binary32 split (slices stored as doubles).
// ---------------------------------------------------------------------------------------------
template <class TI>
__global__ void k_split_csr(int m, const TI *val, const int *rowptr, TI *work, double *split,
                            int64_t lds, int shift, int cap, const uint16_t *J, int *zrow,
                            int *zmax, int epsMode, int mixed) {
	const int r = blockIdx.x * blockDim.x + threadIdx.x;
	if (r >= m) return;
	const int b = rowptr[r], e = rowptr[r + 1];
	// the mixed splits (binary64 data into binary32 slices, binary32 data into binary64 slices)
	// have their own rho, the rest is the same
	const int rho = !mixed ? SplitRule<TI>::rho_row(e - b, epsMode)
	                       : (sizeof(TI) == 8 ? rho_dev_ds(e - b, epsMode) : rho_dev_sd(e - b, epsMode));
	double mx = 0.0;
	for (int i = b; i < e; i++) mx = std_max(mx, fabs((double)val[i]));
	int s = 0;
	bool zero = false;
	for (s = 0; s < cap; s++) {
		const int tau = SplitRule<TI>::tau(mx, J);
		const TI sigma = SplitRule<TI>::sigma_div(rho, tau, shift);
		double lm = 0.0;
		for (int i = b; i < e; i++) {
			const TI v = (s == 0) ? val[i] : work[i];
			const TI a = (v + sigma) - sigma;
			const TI rr = v - a;
			split[(int64_t)s * lds + i] = (double)a;
			work[i] = rr;
			lm = std_max(lm, fabs((double)rr));
		}
		mx = lm;
		if (mx == 0.0) {
			zero = true;
			break;
		}
	}
	const int last = zero ? s : cap - 1;
	for (int i = b; i < e; i++) split[(int64_t)(last + 1) * lds + i] = zero ? 0.0 : (double)work[i];
	const int z = zero ? s : cap;
	zrow[r] = z;
	atomicMax(zmax, z);
}

__global__ void k_zero_slots_csr(int m, const int *rowptr, double *split, int64_t lds,
                                 const int *zrow, int count) {
	const int r = blockIdx.x * blockDim.x + threadIdx.x;
	if (r >= m) return;
	const int b = rowptr[r], e = rowptr[r + 1];
	for (int s = zrow[r] + 2; s < count; s++)
		for (int i = b; i < e; i++) split[(int64_t)s * lds + i] = 0.0;
}

// ---------------------------------------------------------------------------------------------
// fast forms: balanced radix-256 digit planes.
// Column j (k-contiguous, length m) is scaled by 2^(E_j - 8): |x| <= 127 * 2^(E_j - 8) for every
// element, and x = sum_{s<S} d_s 2^(E_j - 8(s+1)) with digits d_s in [-128, 127] -- the unique
// balanced base-256 representation of the integer x / 2^(E_j - 8S).  S is the smallest count for
// which every element is an integer multiple of 2^(E_j - 8S); the operand uses S = max_j S_j.
// ---------------------------------------------------------------------------------------------
template <int BS>
__global__ __launch_bounds__(BS) void k_digit_stats(int m, int n, const double *in, int64_t ldi,
                                                    int *colE, int *colS, ExpStats *st, int *colLo) {
	__shared__ int red[2 * BS / 64];
	__shared__ double redd[BS / 64];
	ColAcc acc;
	for (int j = blockIdx.x; j < n; j += gridDim.x)
		digit_col_stats_acc<BS>(in + (int64_t)j * ldi, m, colE + j, colS + j, acc, red, redd,
		                        colLo ? colLo + j : nullptr);
	if (threadIdx.x == 0) col_acc_flush(acc, st);
}

// digit plane writer for 2-byte digits: plane s of column j at planes + s*pstride + j*ldp
// (k-contiguous).
template <typename T>
__global__ __launch_bounds__(256) void k_digit_write(int m, int n, const double *in, int64_t ldi,
                                                     const int *colE, int S, T *planes,
                                                     int64_t ldp, int64_t pstride) {
	const int j = blockIdx.y;
	const double *x = in + (int64_t)j * ldi;
	const int E = colE[j];
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) {
		const Digits d = make_digits(x[i], E, S);
		for (int s = 0; s < S; s++) {
			const int q = (S - 1 - s) - d.base;
			planes[(int64_t)s * pstride + (int64_t)j * ldp + i] = digit_cast<T>(digit_at(d, q));
		}
	}
}

// INT8 planes in the tiled layout (digit_write_tile_i8).  Grid: (ldp / 64, padded columns / 16).
__global__ __launch_bounds__(256) void k_digit_write_i8t(int m, int n, const double *in, int64_t ldi,
                                                         const int *colE, int S, int8_t *planes,
                                                         int64_t ldp, int64_t pstride) {
	digit_write_tile_i8(m, n, in, ldi, colE, S, planes, ldp, pstride, blockIdx.x, blockIdx.y);
}

// zero the padding rows [m, ldp) of every plane/column so MFMA tiles can read them
template <typename T>
__global__ void k_digit_pad(int m, int n, int S, T *planes, int64_t ldp, int64_t pstride) {
	const int j = blockIdx.x;
	for (int s = 0; s < S; s++)
		for (int i = m + threadIdx.x; i < ldp; i += blockDim.x)
			planes[(int64_t)s * pstride + (int64_t)j * ldp + i] = digit_cast<T>(0);
}

// ---------------------------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------------------------
// columns of up to 256 * 32 elements split in registers; longer ones through the remainder buffer.
// sstride = 0 marks the counting pass (nothing stored); the store pass leaves `fill` valid slots.
template <class TI>
static void split_cols(hipStream_t st, int m, int n, const TI *in, int64_t ldi, TI *out, int64_t ldo,
                       double *split, int64_t lds, int64_t sstride, int rho, int shift, int cap,
                       const uint16_t *J, int *zcol, int *zmax, int fill) {
	const int grid = n < 65535 ? n : 65535;
	const int store = sstride != 0 ? 1 : 0;
	if (m <= 256 * 4)
		hipLaunchKernelGGL((k_split_cols_reg<256, 4, TI>), dim3(grid), dim3(256), 0, st, m, n, in, ldi,
		                   split, lds, sstride, rho, shift, cap, J, zcol, zmax, store, fill);
	else if (m <= 256 * 8)
		hipLaunchKernelGGL((k_split_cols_reg<256, 8, TI>), dim3(grid), dim3(256), 0, st, m, n, in, ldi,
		                   split, lds, sstride, rho, shift, cap, J, zcol, zmax, store, fill);
	else if (m <= 256 * 16)
		hipLaunchKernelGGL((k_split_cols_reg<256, 16, TI>), dim3(grid), dim3(256), 0, st, m, n, in, ldi,
		                   split, lds, sstride, rho, shift, cap, J, zcol, zmax, store, fill);
	else if (m <= 256 * 32)
		hipLaunchKernelGGL((k_split_cols_reg<256, 32, TI>), dim3(grid), dim3(256), 0, st, m, n, in, ldi,
		                   split, lds, sstride, rho, shift, cap, J, zcol, zmax, store, fill);
	else {
		hipLaunchKernelGGL((k_split_cols<256, TI>), dim3(grid), dim3(256), 0, st, m, n, in, ldi, out, ldo,
		                   split, lds, sstride, rho, shift, cap, J, zcol, zmax);
		if (store) launch_zero_slots(st, m, n, split, lds, sstride, zcol, fill);
	}
}

void launch_split_cols(hipStream_t st, int m, int n, const double *in, int64_t ldi, double *out,
                       int64_t ldo, double *split, int64_t lds, int64_t sstride, int rho,
                       int shift, int cap, const uint16_t *J, int *zcol, int *zmax, int fill) {
	split_cols(st, m, n, in, ldi, out, ldo, split, lds, sstride, rho, shift, cap, J, zcol, zmax, fill);
}

void launch_split_cols(hipStream_t st, int m, int n, const float *in, int64_t ldi, float *out,
                       int64_t ldo, double *split, int64_t lds, int64_t sstride, int rho,
                       int shift, int cap, const uint16_t *JF, int *zcol, int *zmax, int fill) {
	split_cols(st, m, n, in, ldi, out, ldo, split, lds, sstride, rho, shift, cap, JF, zcol, zmax, fill);
}

void launch_zero_slots(hipStream_t st, int m, int n, double *split, int64_t lds, int64_t sstride,
                       const int *zcol, int count) {
	if (count <= 2) return;
	const int gx = (m + 255) / 256 < 64 ? (m + 255) / 256 : 64;
	hipLaunchKernelGGL(k_zero_slots, dim3(gx, n, 4), dim3(256), 0, st, m, n, split, lds, sstride,
	                   zcol, count);
}

void launch_transpose(hipStream_t st, int rows, int cols, const double *in, int64_t ldi,
                      double *out, int64_t ldo) {
	hipLaunchKernelGGL(k_transpose<double>, dim3((rows + 31) / 32, (cols + 31) / 32), dim3(256), 0, st,
	                   rows, cols, in, ldi, out, ldo);
}

void launch_transpose(hipStream_t st, int rows, int cols, const float *in, int64_t ldi, float *out,
                      int64_t ldo) {
	hipLaunchKernelGGL(k_transpose<float>, dim3((rows + 31) / 32, (cols + 31) / 32), dim3(256), 0, st,
	                   rows, cols, in, ldi, out, ldo);
}

void launch_vec_max(hipStream_t st, int m, const double *x, unsigned long long *slot) {
	int grid = (m + 255) / 256;
	if (grid > 1024) grid = 1024;
	if (grid < 1) grid = 1;
	hipLaunchKernelGGL(k_vec_max<256>, dim3(grid), dim3(256), 0, st, m, x, slot);
}

void launch_split_vec_pass(hipStream_t st, int m, const double *src, double *rem, double *slice,
                           int rho, int shift, const uint16_t *J, const unsigned long long *maxIn,
                           unsigned long long *maxOut) {
	int grid = (m + 255) / 256;
	if (grid > 1024) grid = 1024;
	if (grid < 1) grid = 1;
	hipLaunchKernelGGL(k_split_vec_pass<256>, dim3(grid), dim3(256), 0, st, m, src, rem, slice,
	                   rho, shift, J, maxIn, maxOut);
}

void launch_fill(hipStream_t st, double *p, int64_t n, double v) {
	if (n <= 0) return;
	int64_t g = (n + 255) / 256;
	if (g > 4096) g = 4096;
	hipLaunchKernelGGL(k_fill<double>, dim3((unsigned)g), dim3(256), 0, st, p, n, v);
}

void launch_fill(hipStream_t st, float *p, int64_t n, float v) {
	if (n <= 0) return;
	int64_t g = (n + 255) / 256;
	if (g > 4096) g = 4096;
	hipLaunchKernelGGL(k_fill<float>, dim3((unsigned)g), dim3(256), 0, st, p, n, v);
}

void launch_copy(hipStream_t st, const double *src, double *dst, int64_t n) {
	if (n <= 0) return;
	int64_t g = (n + 255) / 256;
	if (g > 4096) g = 4096;
	hipLaunchKernelGGL(k_copy, dim3((unsigned)g), dim3(256), 0, st, src, dst, n);
}

void launch_split_csr(hipStream_t st, int m, const double *val, const int *rowptr, double *work,
                      double *split, int64_t lds, int shift, int cap, const uint16_t *J, int *zrow,
                      int *zmax, int epsMode, int mixed) {
	hipLaunchKernelGGL(k_split_csr<double>, dim3((m + 255) / 256), dim3(256), 0, st, m, val, rowptr,
	                   work, split, lds, shift, cap, J, zrow, zmax, epsMode, mixed);
}

void launch_split_csr(hipStream_t st, int m, const float *val, const int *rowptr, float *work,
                      double *split, int64_t lds, int shift, int cap, const uint16_t *JF, int *zrow,
                      int *zmax, int epsMode, int mixed) {
	hipLaunchKernelGGL(k_split_csr<float>, dim3((m + 255) / 256), dim3(256), 0, st, m, val, rowptr,
	                   work, split, lds, shift, cap, JF, zrow, zmax, epsMode, mixed);
}

void launch_zero_slots_csr(hipStream_t st, int m, const int *rowptr, double *split, int64_t lds,
                           const int *zrow, int count) {
	if (count <= 2) return;
	hipLaunchKernelGGL(k_zero_slots_csr, dim3((m + 255) / 256), dim3(256), 0, st, m, rowptr, split,
	                   lds, zrow, count);
}

void launch_digit_stats(hipStream_t st, int m, int n, const double *in, int64_t ldi, int *colE,
                        int *colS, void *stats, int *colLo) {
	const int grid = n < 2048 ? n : 2048;  // columns per workgroup share its few counter atomics
	hipLaunchKernelGGL(k_digit_stats<256>, dim3(grid), dim3(256), 0, st, m, n, in, ldi, colE, colS,
	                   (ExpStats *)stats, colLo);
}

template <typename T>
void launch_digit_write_t(hipStream_t st, int m, int n, const double *in, int64_t ldi,
                          const int *colE, int S, void *planes, int64_t ldp, int64_t pstride) {
	int gx = (m + 255) / 256;
	if (gx > 16) gx = 16;
	hipLaunchKernelGGL(k_digit_write<T>, dim3(gx, n), dim3(256), 0, st, m, n, in, ldi, colE, S,
	                   (T *)planes, ldp, pstride);
	if (ldp > m)
		hipLaunchKernelGGL(k_digit_pad<T>, dim3(n), dim3(64), 0, st, m, n, S, (T *)planes, ldp,
		                   pstride);
}

void launch_digit_write(hipStream_t st, int elem, int m, int n, const double *in, int64_t ldi,
                        const int *colE, int S, void *planes, int64_t ldp, int64_t pstride) {
	if (elem == 1) {  // tiled layout; ldp % 128 == 0 and pstride = ldp * (n rounded up to 256)
		hipLaunchKernelGGL(k_digit_write_i8t, dim3((unsigned)(ldp / 64), (unsigned)(pstride / ldp / 16)),
		                   dim3(256), 0, st, m, n, in, ldi, colE, S, (int8_t *)planes, ldp, pstride);
	}
	else if (elem == 2)
		launch_digit_write_t<_Float16>(st, m, n, in, ldi, colE, S, planes, ldp, pstride);
	else
		launch_digit_write_t<__bf16>(st, m, n, in, ldi, colE, S, planes, ldp, pstride);
}

}  // namespace oz
