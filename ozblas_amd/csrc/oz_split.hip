// oz_split.hip -- error-free splitting kernels.
//
// Plain form (FP64 slices): restates ozblasSplit / ozblasSplitCKernel / ozblasSplitVecKernel /
// ozblasSplitSparse (reference src/cpu/ozblas_split.cpp:75-289,756-928) bit for bit:
//   tau = (short)ceil(log2(max)), sigma = 0.75 * 2^(rho+tau) * splitShift,
//   slice = (x + sigma) - sigma, remainder = x - slice, next max = max |remainder|.
// Because tau is per column (dense) or per row (CSR), one workgroup (dense) or one thread (CSR)
// runs every slice of its column/row in a single launch; the global slice count (the reference
// loops until the sum of all column maxima is zero) is max over columns, read back once.
// Vectors share one tau across all elements, so their split is one grid-wide pass per slice.
//
// Fast forms (INT8/FP16/BF16): exact balanced radix-256 digit planes, see k_digit_*.
#include "oz_kernels.hpp"

namespace oz {

template <int BS>
__device__ inline double block_max(double v, double *red) {
	// wavefront max via cross-lane shuffles, then across the waves through LDS
	for (int o = 32; o > 0; o >>= 1) v = std_max(v, __shfl_xor(v, o, 64));
	const int w = threadIdx.x / 64, l = threadIdx.x % 64;
	__syncthreads();
	if (l == 0) red[w] = v;
	__syncthreads();
	double r = red[0];
	for (int i = 1; i < BS / 64; i++) r = std_max(r, red[i]);
	return r;
}

// ---------------------------------------------------------------------------------------------
// dense: one workgroup per column of a k-contiguous operand (column j = in + j*ldi, length m)
// ---------------------------------------------------------------------------------------------
template <int BS>
__global__ __launch_bounds__(BS) void k_split_cols(int m, int n, const double *in, int64_t ldi,
                                                   double *out, int64_t ldo, double *split,
                                                   int64_t lds, int64_t sstride, int rho, int shift,
                                                   int cap, const uint16_t *J, int *zcol, int *zmax) {
	__shared__ double red[BS / 64];
	for (int j = blockIdx.x; j < n; j += gridDim.x) {
		const double *x = in + (int64_t)j * ldi;
		double *r = out + (int64_t)j * ldo;
		double *sp = split + (int64_t)j * lds;
		double mx = 0.0;  // ozblasFindMaxCKernel (src/cpu/ozblas_split.cpp:37-52)
		for (int i = threadIdx.x; i < m; i += BS) mx = std_max(mx, fabs(x[i]));
		mx = block_max<BS>(mx, red);
		int s = 0;
		bool zero = false;
		for (s = 0; s < cap; s++) {
			const int tau = tau_of(mx, J);
			const double sigma = split_sigma(rho, tau, shift);
			const double *src = (s == 0) ? x : r;
			double lm = 0.0;
			for (int i = threadIdx.x; i < m; i += BS) {
				const double v = src[i];
				const double a = (v + sigma) - sigma;
				const double rr = v - a;
				sp[(int64_t)s * sstride + i] = a;
				r[i] = rr;
				lm = std_max(lm, fabs(rr));
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
			sp[(int64_t)(last + 1) * sstride + i] = zero ? 0.0 : r[i];
		if (threadIdx.x == 0) {
			const int z = zero ? s : cap;  // cap == "never reached zero"
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
__global__ __launch_bounds__(256) void k_transpose(int rows, int cols, const double *in,
                                                   int64_t ldi, double *out, int64_t ldo) {
	__shared__ double t[32][33];
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

__global__ void k_fill(double *p, int64_t n, double v) {
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
// ozblasSplitSparse :896-928); slices at split + s*lds, rho from the row length
// ---------------------------------------------------------------------------------------------
__global__ void k_split_csr(int m, const double *val, const int *rowptr, double *work,
                            double *split, int64_t lds, int shift, int cap, const uint16_t *J,
                            int *zrow, int *zmax) {
	const int r = blockIdx.x * blockDim.x + threadIdx.x;
	if (r >= m) return;
	const int b = rowptr[r], e = rowptr[r + 1];
	const int rho = rho_of_dim(e - b);
	double mx = 0.0;
	for (int i = b; i < e; i++) mx = std_max(mx, fabs(val[i]));
	int s = 0;
	bool zero = false;
	for (s = 0; s < cap; s++) {
		const int tau = tau_of(mx, J);
		const double sigma = 0.75 * scalbn_clamped(1.0, (int64_t)rho + tau) / (double)shift;
		double lm = 0.0;
		for (int i = b; i < e; i++) {
			const double v = (s == 0) ? val[i] : work[i];
			const double a = (v + sigma) - sigma;
			const double rr = v - a;
			split[(int64_t)s * lds + i] = a;
			work[i] = rr;
			lm = std_max(lm, fabs(rr));
		}
		mx = lm;
		if (mx == 0.0) {
			zero = true;
			break;
		}
	}
	const int last = zero ? s : cap - 1;
	for (int i = b; i < e; i++) split[(int64_t)(last + 1) * lds + i] = zero ? 0.0 : work[i];
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
struct ExpStats {  // global reductions written by k_digit_stats
	int maxS;        // max digit count over columns
	int minLsb;      // min exponent of the lowest set bit over all nonzero elements
	int maxE;        // max column scale exponent
	int nonfinite;   // any Inf/NaN
	int maxSpan;     // max (top exponent - lowest bit) over columns
	int pad[3];
};

// exponent of the lowest set bit of a finite nonzero double
__device__ inline int lsb_exp(double x) {
	const uint64_t b = d2u(x);
	const int E = (int)((b >> 52) & 0x7ff);
	uint64_t f = b & ((uint64_t(1) << 52) - 1);
	if (E == 0) return -1074 + __builtin_ctzll(f);
	f |= (uint64_t(1) << 52);
	return (E - 1075) + __builtin_ctzll(f);
}

template <int BS>
__device__ inline int block_max_i(int v, int *red) {
	for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
	const int w = threadIdx.x / 64, l = threadIdx.x % 64;
	__syncthreads();
	if (l == 0) red[w] = v;
	__syncthreads();
	int r = red[0];
	for (int i = 1; i < BS / 64; i++) r = max(r, red[i]);
	__syncthreads();
	return r;
}

template <int BS>
__global__ __launch_bounds__(BS) void k_digit_stats(int m, int n, const double *in, int64_t ldi,
                                                    int *colE, int *colS, ExpStats *st) {
	__shared__ int red[BS / 64];
	__shared__ double redd[BS / 64];
	for (int j = blockIdx.x; j < n; j += gridDim.x) {
		const double *x = in + (int64_t)j * ldi;
		int nlo = INT32_MIN, bad = 0;  // nlo = -(min lsb)
		double mx = 0.0;
		for (int i = threadIdx.x; i < m; i += BS) {
			const double v = x[i];
			if (!isfinite(v)) {
				bad = 1;
				continue;
			}
			if (v == 0.0) continue;
			mx = std_max(mx, fabs(v));
			nlo = max(nlo, -lsb_exp(v));
		}
		mx = block_max<BS>(mx, redd);
		nlo = block_max_i<BS>(nlo, red);
		bad = block_max_i<BS>(bad, red);
		if (threadIdx.x == 0) {
			int E = 0, S = 0;
			if (mx > 0.0) {
				int hi;
				frexp(mx, &hi);  // mx in [2^(hi-1), 2^hi)
				// smallest E with mx <= 127 * 2^(E-8): hi+1 when mx <= (127/128) 2^hi, else hi+2
				E = (mx <= ldexp(127.0, hi - 7)) ? hi + 1 : hi + 2;
				const int lo = -nlo;
				S = (E - lo + 7) / 8;  // 8S >= E - lo
				atomicMin(&st->minLsb, lo);
				atomicMax(&st->maxSpan, hi - lo);
			}
			colE[j] = E;
			colS[j] = S;
			atomicMax(&st->maxS, S);
			atomicMax(&st->maxE, E);
			if (bad) atomicOr(&st->nonfinite, 1);
		}
	}
}

// digit plane writer: plane s of column j at planes + s*pstride + j*ldp (k-contiguous, bytes or
// halfs).  Each thread handles 4 consecutive elements per step.
template <typename T>
__device__ inline T digit_cast(int d);
template <>
__device__ inline int8_t digit_cast<int8_t>(int d) { return (int8_t)d; }
template <>
__device__ inline _Float16 digit_cast<_Float16>(int d) { return (_Float16)(float)d; }
template <>
__device__ inline __bf16 digit_cast<__bf16>(int d) { return (__bf16)(float)d; }

// X = M + 0x80808080808080808080 (80-bit, two words) for the element's aligned integer M;
// digit q of M (from the bottom) = byte q of X - 128
struct Digits {
	uint64_t lo, hi;
	int base;  // plane-index offset: digit q lands in plane S-1-(q+base)
	bool any;
};

__device__ inline Digits make_digits(double v, int E, int S) {
	Digits d{0, 0, 0, false};
	if (v == 0.0 || !isfinite(v)) return d;
	const uint64_t b = d2u(v);
	const int Ex = (int)((b >> 52) & 0x7ff);
	uint64_t f = b & ((uint64_t(1) << 52) - 1);
	int e0;  // v = +-f * 2^e0
	if (Ex == 0) {
		e0 = -1074;
	} else {
		f |= (uint64_t(1) << 52);
		e0 = Ex - 1075;
	}
	const int bottom = E - 8 * S;        // exponent of the unit of the lowest plane
	const int sh = e0 - bottom;          // >= 0 by the choice of S (or f has trailing zeros)
	int q0, bits;
	if (sh >= 0) {
		q0 = sh / 8;
		bits = sh % 8;
	} else {  // e0 below the grid: trailing zeros of f cover it (checked by the stats pass)
		f >>= -sh;
		q0 = 0;
		bits = 0;
	}
	const int64_t M = (int64_t)(f << bits) * ((b >> 63) ? -1 : 1);  // |M| < 2^61
	// 128-bit X = M + K, K = 0x80 in bytes 0..9
	const uint64_t Klo = 0x8080808080808080ULL, Khi = 0x8080ULL;
	const uint64_t mlo = (uint64_t)M, mhi = (M < 0) ? ~0ULL : 0ULL;
	const uint64_t lo = mlo + Klo;
	const uint64_t carry = lo < mlo ? 1 : 0;
	d.lo = lo;
	d.hi = mhi + Khi + carry;
	d.base = q0;
	d.any = true;
	return d;
}

__device__ inline int digit_at(const Digits &d, int q) {  // q relative to the element's base
	if (!d.any || q < 0 || q >= 10) return 0;
	const uint64_t w = (q < 8) ? (d.lo >> (8 * q)) : (d.hi >> (8 * (q - 8)));
	return (int)(w & 0xff) - 128;
}

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

// INT8 planes in the tiled layout of the MFMA kernel: element (column j, k index q) of plane s at
//   s * pstride + (j / 16) * 16 * ldp + (q / 64) * 1024 + (j % 16) * 64 + q % 64.
// One block writes one 16 x 64 tile of every plane (1 KiB per plane, 4 bytes per thread), padding
// columns >= n and k >= m with zero digits.  Grid: (ldp / 64, padded columns / 16).
__global__ __launch_bounds__(256) void k_digit_write_i8t(int m, int n, const double *in, int64_t ldi,
                                                         const int *colE, int S, int8_t *planes,
                                                         int64_t ldp, int64_t pstride) {
	const int t = threadIdx.x;
	const int j = blockIdx.y * 16 + t / 16;
	const int q0 = blockIdx.x * 64 + 4 * (t % 16);
	Digits d[4];
	if (j < n) {
		const int E = colE[j];
		const double *x = in + (int64_t)j * ldi;
#pragma unroll
		for (int e = 0; e < 4; e++) d[e] = make_digits(q0 + e < m ? x[q0 + e] : 0.0, E, S);
	} else {
#pragma unroll
		for (int e = 0; e < 4; e++) d[e] = make_digits(0.0, 0, S);
	}
	int8_t *dst = planes + (int64_t)blockIdx.y * 16 * ldp + (int64_t)blockIdx.x * 1024 + 4 * t;
	for (int s = 0; s < S; s++) {
		uint32_t word = 0;
#pragma unroll
		for (int e = 0; e < 4; e++)
			word |= (uint32_t)(uint8_t)digit_at(d[e], (S - 1 - s) - d[e].base) << (8 * e);
		*(uint32_t *)(dst + (int64_t)s * pstride) = word;
	}
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
void launch_split_cols(hipStream_t st, int m, int n, const double *in, int64_t ldi, double *out,
                       int64_t ldo, double *split, int64_t lds, int64_t sstride, int rho,
                       int shift, int cap, const uint16_t *J, int *zcol, int *zmax) {
	const int grid = n < 65535 ? n : 65535;
	hipLaunchKernelGGL(k_split_cols<256>, dim3(grid), dim3(256), 0, st, m, n, in, ldi, out, ldo,
	                   split, lds, sstride, rho, shift, cap, J, zcol, zmax);
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
	hipLaunchKernelGGL(k_transpose, dim3((rows + 31) / 32, (cols + 31) / 32), dim3(256), 0, st,
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
	hipLaunchKernelGGL(k_fill, dim3((unsigned)g), dim3(256), 0, st, p, n, v);
}

void launch_copy(hipStream_t st, const double *src, double *dst, int64_t n) {
	if (n <= 0) return;
	int64_t g = (n + 255) / 256;
	if (g > 4096) g = 4096;
	hipLaunchKernelGGL(k_copy, dim3((unsigned)g), dim3(256), 0, st, src, dst, n);
}

void launch_split_csr(hipStream_t st, int m, const double *val, const int *rowptr, double *work,
                      double *split, int64_t lds, int shift, int cap, const uint16_t *J, int *zrow,
                      int *zmax) {
	hipLaunchKernelGGL(k_split_csr, dim3((m + 255) / 256), dim3(256), 0, st, m, val, rowptr, work,
	                   split, lds, shift, cap, J, zrow, zmax);
}

void launch_zero_slots_csr(hipStream_t st, int m, const int *rowptr, double *split, int64_t lds,
                           const int *zrow, int count) {
	if (count <= 2) return;
	hipLaunchKernelGGL(k_zero_slots_csr, dim3((m + 255) / 256), dim3(256), 0, st, m, rowptr, split,
	                   lds, zrow, count);
}

void launch_digit_stats(hipStream_t st, int m, int n, const double *in, int64_t ldi, int *colE,
                        int *colS, void *stats) {
	const int grid = n < 65535 ? n : 65535;
	hipLaunchKernelGGL(k_digit_stats<256>, dim3(grid), dim3(256), 0, st, m, n, in, ldi, colE, colS,
	                   (ExpStats *)stats);
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
