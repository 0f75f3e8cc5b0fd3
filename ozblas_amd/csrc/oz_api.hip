// oz_api.hip -- the C ABI drivers: ozblasD{gemm,gemv,dot,nrm2,axpy,csrmv,csrmvSplitA,cg}.
//
// Control flow restates the reference drivers (src/cpu/ozblas_gemm.cpp, ozblas_dot.cpp,
// ozblas_nrm2.cpp, ozblas_gemv.cpp, ozblas_csrmv.cpp, ozblas_cg.cpp, ozblas_axpy.cpp); the work is
// done by the gfx950 kernels of this library.  Differences that cannot change a result bit:
//  - slice counts are measured before the slices are stored, so only the slots in use are
//    allocated (the reference reserves nSplitMax+1 slots per operand and L^2 product planes);
//  - m/n blocking happens only when the workspace cannot hold the whole problem;
//  - the Fsum summation is fused into the FP64 product kernel when the reference's summation
//    index map is the identity (always, except a GEMM whose column block has width 1).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <type_traits>
#include <vector>

#include "oz_digits.hpp"

namespace oz {

static inline bool is_n(char t) { return t == 'N' || t == 'n'; }
static inline int64_t rup(int64_t v, int64_t a) { return (v + a - 1) / a * a; }
static inline int imin(int a, int b) { return a < b ? a : b; }
static inline int imax(int a, int b) { return a > b ? a : b; }

struct WallTimer {
	std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
	float sec() const {
		return std::chrono::duration<float>(std::chrono::steady_clock::now() - t0).count();
	}
};

// elapsed seconds between two events recorded on the handle's stream (after a sync)
static float ev_sec(hipEvent_t a, hipEvent_t b) {
	float ms = 0.f;
	if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.f;
	return ms * 1e-3f;
}

static int32_t sync(ozblasHandle_t *oh) {
	OZ_HIP(hipStreamSynchronize(be(oh)->stream));
	return OZBLAS_STATUS_SUCCESS;
}

// Per-precision rules of the drivers: rho of getRho and the device tau tables (the binary32 split
// takes tau from (float)log2((double)max), its NearSum from log2f).  Both precisions store slices
// as doubles and share the FP64 product kernels.  The mixed routines (Backend::mixed) differ from
// their data type's routines in rho and in the split loop's slot rule: binary64 data split into
// binary32 slices (<double,float>) give unscaled FP64 products that are exactly the reference's
// scalbn(float product, seA + seB); binary32 data split into binary64-held slices (<float,double>)
// give products whose binary32 term (f32_term) is the reference's scalbn1<float>((float)P, seA + seB),
// and their CSR chains run in binary64.
template <class T>
struct Prec;
template <>
struct Prec<double> {
	static int rho(const ozblasHandle_t *oh, int dim) {
		return be(oh)->mixed ? rho_host_ds(dim, oh->splitEpsMode) : rho_host(dim, oh->splitEpsMode);
	}
	static const uint16_t *Jsplit(Backend *b) { return b->dTau; }
	static const uint16_t *Jsum(Backend *b) { return b->dTau; }
};
template <>
struct Prec<float> {
	static int rho(const ozblasHandle_t *oh, int dim) {
		return be(oh)->mixed ? rho_host_sd(dim, oh->splitEpsMode) : rho_host_f(dim, oh->splitEpsMode);
	}
	static const uint16_t *Jsplit(Backend *b) { return b->dTauF; }
	static const uint16_t *Jsum(Backend *b) { return b->dTauF2; }
};

// reference slice count from the per-column first-zero index (max over columns)
static inline int split_count(const ozblasHandle_t *oh, int zmax) {
	const int cap = split_cap(oh);
	int count;
	if (cap <= 1)
		count = 1;
	else if (zmax < cap)
		count = imax(zmax, 1) + 1;
	else
		count = cap;
	if (repro_slot(oh)) count++;
	return count;
}

// ---------------------------------------------------------------------------------------------
// dense column split of a k-contiguous operand (k x cols, ld) into slices (ldk x cols per slot)
// ---------------------------------------------------------------------------------------------
struct DenseSplit {
	double *slices = nullptr;
	int64_t ldk = 0;      // slice column pitch
	int64_t sstride = 0;  // slice stride
	int count = 0;
};

// knownZ >= 0: the first-zero level of these very columns, measured by count_cols (skips the
// counting pass)
template <class T>
static int32_t dense_split(ozblasHandle_t *oh, int k, int cols, const T *in, int64_t ldi, T *rem,
                           int64_t ldr, DenseSplit &ds, int knownZ = -1) {
	Backend *b = be(oh);
	const int cap = split_cap(oh);
	const int rho = Prec<T>::rho(oh, k);
	const uint16_t *J = Prec<T>::Jsplit(b);
	ds.ldk = rup(k, 16);
	ds.sstride = ds.ldk * cols;
	int *zcol = (int *)walloc(oh, sizeof(int) * (size_t)cols);
	if (!zcol) return OZBLAS_STATUS_WORKSPACE;
	int zmax = knownZ;
	if (zmax < 0) {
		// the store pass needs the count; measure it first without storing (slices -> rem scratch)
		OZ_HIP(hipMemsetAsync(b->dCounters, 0, sizeof(int), b->stream));
		// counting pass: slices land in a scratch of one slot that is overwritten (sstride = 0)
		double *scratch = (double *)walloc(oh, sizeof(double) * (size_t)ds.ldk * cols * 2);
		if (!scratch) return OZBLAS_STATUS_WORKSPACE;
		launch_split_cols(b->stream, k, cols, in, ldi, rem, ldr, scratch, ds.ldk, 0, rho,
		                  oh->splitShift, cap, J, zcol, b->dCounters);
		OZ_HIP(hipMemcpyAsync(b->hPinned, b->dCounters, sizeof(int), hipMemcpyDeviceToHost, b->stream));
		OZ_TRY(sync(oh));
		zmax = (int)((int32_t *)b->hPinned)[0];
	}
	ds.count = split_count(oh, zmax);
	// store pass: count slots (+1 for the remainder slot the kernel writes after the last slice)
	const int slots = imax(ds.count, imin(zmax, cap) + 2);
	ds.slices = (double *)walloc(oh, sizeof(double) * (size_t)ds.sstride * slots);
	if (!ds.slices) return OZBLAS_STATUS_WORKSPACE;
	launch_split_cols(b->stream, k, cols, in, ldi, rem, ldr, ds.slices, ds.ldk, ds.sstride, rho,
	                  oh->splitShift, cap, J, zcol, b->dCounters + 1, ds.count);
	// zero the pad rows [k, ldk) of every slot the product kernels read
	if (ds.ldk > k) {
		for (int s = 0; s < ds.count; s++)
			OZ_HIP(hipMemset2DAsync(ds.slices + (size_t)s * ds.sstride + k, ds.ldk * sizeof(double), 0,
			                        (ds.ldk - k) * sizeof(double), cols, b->stream));
	}
	return OZBLAS_STATUS_SUCCESS;
}

// ---------------------------------------------------------------------------------------------
// slice count of a k-contiguous operand without storing slices (for the blocking decision)
// ---------------------------------------------------------------------------------------------
template <class T>
static int32_t count_cols(ozblasHandle_t *oh, int k, int cols, const T *in, int64_t ldi,
                          int *count, int *zmaxOut = nullptr) {
	Backend *b = be(oh);
	const uint64_t mark = wmark(oh);
	const int64_t ldk = rup(k, 16);
	T *rem = (T *)walloc(oh, sizeof(T) * ldk * cols);
	double *scratch = (double *)walloc(oh, 8ull * ldk * cols);
	int *zcol = (int *)walloc(oh, sizeof(int) * (size_t)cols);
	if (!rem || !scratch || !zcol) return OZBLAS_STATUS_WORKSPACE;
	OZ_HIP(hipMemsetAsync(b->dCounters, 0, sizeof(int), b->stream));
	launch_split_cols(b->stream, k, cols, in, ldi, rem, ldk, scratch, ldk, 0,
	                  Prec<T>::rho(oh, k), oh->splitShift, split_cap(oh),
	                  Prec<T>::Jsplit(b), zcol, b->dCounters);
	OZ_HIP(hipMemcpyAsync(b->hPinned, b->dCounters, sizeof(int), hipMemcpyDeviceToHost, b->stream));
	OZ_TRY(sync(oh));
	*count = split_count(oh, ((int32_t *)b->hPinned)[0]);
	if (zmaxOut) *zmaxOut = ((int32_t *)b->hPinned)[0];
	wreset(oh, mark);
	return OZBLAS_STATUS_SUCCESS;
}

// ---------------------------------------------------------------------------------------------
// GEMM, plain form (src/cpu/ozblas_gemm.cpp:3-315)
// ---------------------------------------------------------------------------------------------
template <class T>
static int32_t gemm_plain(ozblasHandle_t *oh, char transA, char transB, int m, int n, int k,
                          T alpha, const T *A, int lda, const T *B, int ldb, T beta, T *C,
                          int ldc) {
	Backend *b = be(oh);
	hipStream_t st = b->stream;
	const bool fsum = (oh->sumMode == 0);
	if (oh->sumMode != 0 && oh->sumMode != 1) return OZBLAS_STATUS_NOT_SUPPORTED;
	if (oh->splitEpsMode != 0 && oh->splitEpsMode != 9) return OZBLAS_STATUS_NOT_SUPPORTED;
	// the reference's mixed NearSum prints an error and exits (src/cpu/ozblas_sum.cpp:275-330)
	if (b->mixed && !fsum) return OZBLAS_STATUS_NOT_SUPPORTED;
	const bool dotPath = (m == 1 && n == 1);  // :25-28: DOT ignores fastMode
	const int fastMode = dotPath ? 0 : oh->fastMode;
	const uint64_t base = wmark(oh);
	const int64_t ldk = rup(k, 16);
	WallTimer tt;
	// whole-operand copies of op(A) / op(B) in k-contiguous form (transposed when needed)
	const T *ain = A, *bin = B;
	int64_t ldai = lda, ldbi = ldb;
	if (is_n(transA)) {
		T *tA = (T *)walloc(oh, sizeof(T) * ldk * m);
		if (!tA) return OZBLAS_STATUS_WORKSPACE;
		launch_transpose(st, m, k, A, lda, tA, ldk);
		ain = tA;
		ldai = ldk;
	}
	if (!is_n(transB)) {
		T *tB = (T *)walloc(oh, sizeof(T) * ldk * n);
		if (!tB) return OZBLAS_STATUS_WORKSPACE;
		launch_transpose(st, n, k, B, ldb, tB, ldk);
		bin = tB;
		ldbi = ldk;
	}
	const uint64_t base2 = wmark(oh);
	// blocking: slice counts of the whole operands bound those of any block, so halving the
	// larger side (as src/cpu/ozblas_gemm.cpp:91-96) until the exact need fits is safe
	int nAf = 0, nBf = 0, zAf = -1, zBf = -1;
	OZ_TRY(count_cols(oh, k, m, ain, ldai, &nAf, &zAf));
	OZ_TRY(count_cols(oh, k, n, bin, ldbi, &nBf, &zBf));
	auto need = [&](int64_t mb, int64_t nb) -> uint64_t {
		uint64_t bytes = 8ull * ldk * std::max(mb, nb) * 2;            // remainder + count scratch
		bytes += 8ull * ldk * (mb * (nAf + 1) + nb * (nBf + 1));        // slices (+ remainder slot)
		const int64_t pairs = (int64_t)nAf * nBf;
		if (!fsum || (oh->fastMode == 0 && nb == 1)) bytes += 8ull * rup(mb, 8) * nb * pairs;
		bytes += 8ull * (mb + nb) + 24ull * pairs + 64 * 1024;         // counters, offsets, align
		return bytes;
	};
	int mbk = m, nbk = n;
	while (need(mbk, nbk) > b->cap - base2) {
		if (mbk == 1 && nbk == 1) return OZBLAS_STATUS_WORKSPACE;
		if (mbk >= nbk)
			mbk = (mbk + 1) / 2;
		else
			nbk = (nbk + 1) / 2;
	}
	oh->mbk = mbk;
	oh->nbk = nbk;
	double sA = 0, sB = 0, sC = 0, tcomp = 0, tsum = 0;
	int blocks = 0;
	for (int im = 0; im * mbk < m; im++) {
		const int mbk_ = imin(mbk, m - im * mbk);
		for (int in = 0; in * nbk < n; in++) {
			const int nbk_ = imin(nbk, n - in * nbk);
			wreset(oh, base2);
			T *rem = (T *)walloc(oh, sizeof(T) * ldk * std::max(mbk_, nbk_));
			if (!rem) return OZBLAS_STATUS_WORKSPACE;
			// binary64 GEMV over the whole matrix: the split of op(A) is fused into the slice
			// products (the matrix is read once, its slices never stored); x is split as usual
			if constexpr (std::is_same<T, double>::value) {
				if (!dotPath && n == 1 && fastMode == 0 && nbk_ == 1 && mbk == m && !b->mixed &&
				    split_gemv_fits(k, nAf, nBf)) {
					DenseSplit dx;
					OZ_TRY(dense_split(oh, k, nbk_, bin + (int64_t)in * nbk * ldbi, ldbi, rem, ldk, dx,
					                   nbk == n ? zBf : -1));
					T *Cblk = C + (int64_t)ldc * ((int64_t)in * nbk) + (int64_t)im * mbk;
					double *P = (double *)walloc(oh, split_gemv_pairs_bytes(mbk_));
					if (!P) return OZBLAS_STATUS_WORKSPACE;
					OZ_HIP(hipEventRecord(b->ev[3], st));
					if (launch_split_gemv(st, k, mbk_, ain + (int64_t)im * mbk * ldai, ldai, Prec<T>::rho(oh, k),
					                      oh->splitShift, split_cap(oh), Prec<T>::Jsplit(b), nAf, dx.slices,
					                      dx.sstride, dx.count, fsum ? 1 : 0, alpha, beta, Cblk,
					                      Prec<T>::Jsum(b), P)) {
						OZ_HIP(hipEventRecord(b->ev[4], st));
						OZ_HIP(hipGetLastError());
						OZ_TRY(sync(oh));
						tcomp += ev_sec(b->ev[3], b->ev[4]);
						sA += nAf;
						sB += dx.count;
						sC += (double)nAf * dx.count;
						blocks++;
						continue;
					}
				}
			}
			DenseSplit da, db;
			// a block that is the whole operand reuses the first-zero level of count_cols
			OZ_TRY(dense_split(oh, k, mbk_, ain + (int64_t)im * mbk * ldai, ldai, rem, ldk, da,
			                   mbk == m ? zAf : -1));
			OZ_TRY(dense_split(oh, k, nbk_, bin + (int64_t)in * nbk * ldbi, ldbi, rem, ldk, db,
			                   nbk == n ? zBf : -1));
			const int nA = da.count, nB = db.count;
			// ---- pairs (src/cpu/ozblas_gemm.cpp:163-164,237-264)
			int maxlevel = (nA - 1) + (nB - 1);
			maxlevel = imax(maxlevel - fastMode, imin(nA - 1, nB - 1));
			const bool gemvPath = !dotPath && (n == 1 && fastMode == 0);  // products at nB*ia+ib
			const int sumorder = dotPath ? 2 : ((fastMode == 0 && nbk_ == 1) ? 3 : 1);
			std::vector<int> pia, pib;
			for (int ik = 0; ik <= maxlevel; ik++)
				for (int ia = 0; ia < nA; ia++) {
					const int ib = ik - ia;
					if (ib < 0 || ib >= nB) continue;
					pia.push_back(ia);
					pib.push_back(ib);
				}
			const int ic = (int)pia.size();
			const int64_t ldcs = rup(mbk_, 8);
			const int64_t llsc = ldcs * nbk_;
			std::vector<int64_t> hoff(3 * (size_t)ic);
			for (int p = 0; p < ic; p++) {
				hoff[p] = (int64_t)pia[p] * da.sstride;
				hoff[ic + p] = (int64_t)pib[p] * db.sstride;
				int64_t it = p;                                   // GEMM: level order (:243)
				if (gemvPath) it = (int64_t)nB * pia[p] + pib[p];  // GEMV (:221)
				if (dotPath) it = (int64_t)nA * pib[p] + pia[p];   // DOT: C[ia + ib*nA] (:214)
				hoff[2 * ic + p] = it * llsc;
			}
			int64_t *dOff = (int64_t *)walloc(oh, sizeof(int64_t) * 3 * (size_t)ic);
			if (!dOff) return OZBLAS_STATUS_WORKSPACE;
			OZ_HIP(hipMemcpyAsync(dOff, hoff.data(), sizeof(int64_t) * 3 * ic, hipMemcpyHostToDevice, st));
			T *Cblk = C + (int64_t)ldc * ((int64_t)in * nbk) + (int64_t)im * mbk;
			OZ_HIP(hipEventRecord(b->ev[3], st));
			// fused Fsum is valid when the summation visits the products in pair order; the
			// reference's GEMM with a width-1 column block reads them through the GEMV map instead
			const bool identityMap = gemvPath || dotPath || sumorder == 1;
			bool summed = false;
			// GEMV: the bandwidth-bound slice kernel with the sum fused (all nA x nB pairs: fastMode 0)
			if (gemvPath && nbk_ == 1 &&
			    launch_gemv_slices(st, mbk_, k, da.slices, da.ldk, da.sstride, nA, db.slices, db.sstride, nB,
			                       fsum ? 1 : 0, alpha, beta, Cblk, Prec<T>::Jsum(b))) {
				OZ_HIP(hipEventRecord(b->ev[4], st));
				summed = true;
			} else if (fsum && identityMap) {
				launch_gemm_f64_fsum(st, mbk_, nbk_, k, da.slices, da.ldk, db.slices, db.ldk, dOff,
				                     dOff + ic, ic, alpha, beta, Cblk, ldc);
				OZ_HIP(hipEventRecord(b->ev[4], st));
				summed = true;
			} else {
				const int64_t nplanes = (gemvPath || dotPath || !identityMap) ? std::max<int64_t>((int64_t)nA * nB, ic) : ic;
				T *Cs = (T *)walloc(oh, sizeof(T) * llsc * nplanes);
				if (!Cs) return OZBLAS_STATUS_WORKSPACE;
				if (nplanes > ic) launch_fill(st, Cs, llsc * nplanes, (T)0);
				launch_gemm_f64_batched(st, mbk_, nbk_, k, da.slices, da.ldk, db.slices, db.ldk, Cs, ldcs,
				                        dOff, dOff + ic, dOff + 2 * ic, ic);
				OZ_HIP(hipEventRecord(b->ev[4], st));
				if (fsum)
					launch_global_fsum(st, mbk_, nbk_, nA, nB, maxlevel, sumorder, Cs, llsc, ldcs, Cblk,
					                   ldc, alpha, beta);
				else
					launch_global_nearsum(st, mbk_, nbk_, ic, Cs, llsc, ldcs, Cblk, ldc, alpha, beta,
					                      Prec<T>::Jsum(b));
			}
			OZ_HIP(hipEventRecord(b->ev[5], st));
			OZ_HIP(hipGetLastError());
			OZ_TRY(sync(oh));
			tcomp += ev_sec(b->ev[3], b->ev[4]);
			if (!summed) tsum += ev_sec(b->ev[4], b->ev[5]);
			sA += nA;
			sB += nB;
			sC += ic;
			blocks++;
		}
	}
	oh->nSplitA = (float)(sA / blocks);
	oh->nSplitB = (float)(sB / blocks);
	oh->nSplitC = (float)(sC / blocks);
	oh->t_comp = (float)tcomp;  // the FP64 MFMA product kernel(s)
	oh->t_sum = (float)tsum;
	oh->t_total = tt.sec();
	oh->n_comp = (float)(2.0 * m * n * (double)k * oh->nSplitC);
	wreset(oh, base);
	return OZBLAS_STATUS_SUCCESS;
}

// ---------------------------------------------------------------------------------------------
// GEMM, fast forms: exact digit planes (no reference counterpart; same bits as NearSum)
// ---------------------------------------------------------------------------------------------
struct DigitOperand {
	void *planes = nullptr;
	int64_t ldp = 0, pstride = 0;  // in elements
	int *colE = nullptr;
	int S = 0;
	ExpStatsH st{};
};

static int32_t digit_split(ozblasHandle_t *oh, int elem, int k, int cols, const double *in,
                           int64_t ldi, DigitOperand &d) {
	Backend *b = be(oh);
	d.colE = (int *)walloc(oh, sizeof(int) * (size_t)cols * 2);
	ExpStatsH *dst = (ExpStatsH *)walloc(oh, sizeof(ExpStatsH));
	if (!d.colE || !dst) return OZBLAS_STATUS_WORKSPACE;
	ExpStatsH init{0, INT32_MAX, INT32_MIN, 0, INT32_MIN, INT32_MIN, {0, 0}};
	OZ_HIP(hipMemcpyAsync(dst, &init, sizeof(init), hipMemcpyHostToDevice, b->stream));
	launch_digit_stats(b->stream, k, cols, in, ldi, d.colE, d.colE + cols, dst);
	OZ_HIP(hipMemcpyAsync(b->hPinned, dst, sizeof(ExpStatsH), hipMemcpyDeviceToHost, b->stream));
	OZ_TRY(sync(oh));
	std::memcpy(&d.st, b->hPinned, sizeof(ExpStatsH));
	d.S = d.st.maxS;
	if (elem == 1) {  // tiled INT8 layout: k padded to 128 bytes, columns to 256 (zero digits)
		d.ldp = rup(k, 128);
		d.pstride = d.ldp * rup(cols, 256);
	} else {
		d.ldp = rup(k, 64);
		d.pstride = d.ldp * cols;
	}
	const size_t esz = (elem == 1) ? 1 : 2;
	if (d.S > 0) {
		d.planes = walloc(oh, esz * (size_t)d.pstride * d.S);
		if (!d.planes) return OZBLAS_STATUS_WORKSPACE;
		launch_digit_write(b->stream, elem, k, cols, in, ldi, d.colE, d.S, d.planes, d.ldp, d.pstride);
	}
	return OZBLAS_STATUS_SUCCESS;
}

static bool fast_form_valid(const ozblasHandle_t *oh, const DigitOperand &a, const DigitOperand &b,
                            int k) {
	return fast_valid(oh, a.st, b.st, k);
}

static int32_t gemm_fast(ozblasHandle_t *oh, char transA, char transB, int m, int n, int k,
                         double alpha, const double *A, int lda, const double *B, int ldb,
                         double beta, double *C, int ldc, bool *handled) {
	Backend *b = be(oh);
	hipStream_t st = b->stream;
	*handled = false;
	const int elem = (oh->sliceType == OZBLAS_SLICE_INT8) ? 1 : (oh->sliceType == OZBLAS_SLICE_FP16 ? 2 : 3);
	const uint64_t base = wmark(oh);
	WallTimer tt;
	const int64_t ldk = rup(k, 16);
	const double *ain = A, *bin = B;
	int64_t ldai = lda, ldbi = ldb;
	if (is_n(transA)) {
		double *tA = (double *)walloc(oh, 8ull * ldk * m);
		if (!tA) return OZBLAS_STATUS_WORKSPACE;
		launch_transpose(st, m, k, A, lda, tA, ldk);
		ain = tA;
		ldai = ldk;
	}
	if (!is_n(transB)) {
		double *tB = (double *)walloc(oh, 8ull * ldk * n);
		if (!tB) return OZBLAS_STATUS_WORKSPACE;
		launch_transpose(st, n, k, B, ldb, tB, ldk);
		bin = tB;
		ldbi = ldk;
	}
	DigitOperand da, db;
	OZ_HIP(hipEventRecord(b->ev[0], st));
	OZ_TRY(digit_split(oh, elem, k, m, ain, ldai, da));
	OZ_HIP(hipEventRecord(b->ev[1], st));
	OZ_TRY(digit_split(oh, elem, k, n, bin, ldbi, db));
	OZ_HIP(hipEventRecord(b->ev[2], st));
	if (!fast_form_valid(oh, da, db, k)) {
		wreset(oh, base);
		return OZBLAS_STATUS_SUCCESS;  // caller falls back to the plain form
	}
	*handled = true;
	const int SA = da.S, SB = db.S;
	if (SA == 0 || SB == 0) {  // exact zero product
		double *zero = (double *)walloc(oh, 8);
		(void)zero;
		// C = fma(alpha, +0, beta*C) through the reconstruction kernel with no bins
		int *lv = (int *)walloc(oh, sizeof(int));
		launch_reconstruct(st, m, n, nullptr, 0, 0, lv, 0, 0, da.colE, db.colE, alpha, beta, C, ldc);
		OZ_HIP(hipGetLastError());
		OZ_TRY(sync(oh));
		wreset(oh, base);
		return OZBLAS_STATUS_SUCCESS;
	}
	// bins: pairs of one level whose exact int32 sum cannot overflow: npairs * kc * 2^14 < 2^31
	const int Kpad = (int)da.ldp;
	const int kc = Kpad <= 65536 ? Kpad : 65536;
	const int G = std::max(1, 131071 / kc);
	const int Lmax = SA + SB - 2;
	std::vector<int2> pairs;
	std::vector<LevelBin> bins;
	std::vector<int> binLevel;
	for (int L = 0; L <= Lmax; L++) {
		std::vector<int2> lp;
		for (int s = 0; s < SA; s++) {
			const int t = L - s;
			if (t >= 0 && t < SB) lp.push_back(make_int2(s, t));
		}
		for (size_t g0 = 0; g0 < lp.size(); g0 += G)
			for (int k0 = 0; k0 < Kpad; k0 += kc) {
				LevelBin bn;
				bn.level = L;
				bn.pair0 = (int)pairs.size();
				bn.npairs = (int)std::min<size_t>(G, lp.size() - g0);
				bn.k0 = k0;
				bn.k1 = std::min(Kpad, k0 + kc);
				for (int q = 0; q < bn.npairs; q++) pairs.push_back(lp[g0 + q]);
				bins.push_back(bn);
				binLevel.push_back(L);
			}
	}
	const int nbins = (int)bins.size();
	oh->nLevelBins = nbins;
	// Certified early termination (INT8): pass 1 computes the levels <= L1; the reconstruction
	// bounds the rest by R = sum_{L > L1} npairs_L * k * 2^14 * 2^(8 (Lmax - L)) (every digit
	// product is at most 2^14 in magnitude) and keeps the elements whose rounding of N1 +- R agrees;
	// pass 2 computes the remaining levels only for the tiles holding the other elements.
	int L1 = Lmax;
	if (oh->digitLevels > 0) L1 = imin(Lmax, oh->digitLevels - 1);
	if (oh->digitLevels == 0) {
		// automatic: the first level with R / 2^(8 Lmax) <= 2^-62, i.e. 10 bits below the binary64
		// precision of an element of size 2^(8 Lmax) units (the product of the two scales)
		std::vector<double> np(Lmax + 1, 0.0);
		for (const LevelBin &bn : bins) np[bn.level] += bn.npairs * (double)(bn.k1 - bn.k0) / Kpad;
		for (L1 = 0; L1 < Lmax; L1++) {
			double r = 0.0;
			for (int Lv = L1 + 1; Lv <= Lmax; Lv++) r += np[Lv] * (double)k * 16384.0 * std::ldexp(1.0, -8 * Lv);
			if (r <= std::ldexp(1.0, -62)) break;
		}
		int64_t p1 = 0;
		for (const LevelBin &bn : bins)
			if (bn.level <= L1) p1 += bn.npairs;
		if (L1 >= Lmax - 1 || p1 > 0.9 * (double)pairs.size()) L1 = Lmax;  // too little to skip
	}
	int nb1 = 0;  // bins are in level order: pass 1 is a prefix
	while (nb1 < nbins && bins[nb1].level <= L1) nb1++;
	// launch order within each pass: bins with the most pairs first, so the short ones fill the tail
	std::vector<int> order(nbins);
	for (int q = 0; q < nbins; q++) order[q] = q;
	auto byPairs = [&](int x, int y) { return bins[x].npairs > bins[y].npairs; };
	std::stable_sort(order.begin(), order.begin() + nb1, byPairs);
	std::stable_sort(order.begin() + nb1, order.end(), byPairs);
	int2 *dPairs = (int2 *)walloc(oh, sizeof(int2) * pairs.size());
	LevelBin *dBins = (LevelBin *)walloc(oh, sizeof(LevelBin) * bins.size());
	int *dLevel = (int *)walloc(oh, sizeof(int) * binLevel.size());
	int *dOrder = (int *)walloc(oh, sizeof(int) * order.size());
	// INT8 result planes cover whole 256 x 256 tiles (the kernel stores without bounds checks)
	const int64_t ldt = (elem == 1) ? rup(m, 256) : rup(m, 64);
	const int64_t tstride = ldt * ((elem == 1) ? rup(n, 256) : n);
	int32_t *T = (int32_t *)walloc(oh, sizeof(int32_t) * (size_t)tstride * nbins);
	const int tile = digit_gemm_tile(elem, m, n);
	const int tilesM = (m + tile - 1) / tile, tilesN = (n + tile - 1) / tile;
	uint64_t *dR = (uint64_t *)walloc(oh, 8 * 16);
	int *dMark = (int *)walloc(oh, sizeof(int) * (size_t)tilesM * tilesN * 2);
	if (!dPairs || !dBins || !dLevel || !dOrder || !T || !dR || !dMark) {
		wreset(oh, base);
		*handled = false;
		return OZBLAS_STATUS_WORKSPACE;
	}
	OZ_HIP(hipMemcpyAsync(dPairs, pairs.data(), sizeof(int2) * pairs.size(), hipMemcpyHostToDevice, st));
	OZ_HIP(hipMemcpyAsync(dBins, bins.data(), sizeof(LevelBin) * bins.size(), hipMemcpyHostToDevice, st));
	OZ_HIP(hipMemcpyAsync(dLevel, binLevel.data(), sizeof(int) * binLevel.size(), hipMemcpyHostToDevice, st));
	OZ_HIP(hipMemcpyAsync(dOrder, order.data(), sizeof(int) * order.size(), hipMemcpyHostToDevice, st));
	int64_t pairs1 = 0;
	for (int q = 0; q < nb1; q++) pairs1 += bins[q].npairs;
	OZ_HIP(hipEventRecord(b->ev[3], st));
	launch_digit_gemm(st, elem, m, n, Kpad, da.planes, da.ldp, da.pstride, db.planes, db.ldp,
	                  db.pstride, dPairs, dBins, nb1, dOrder, T, ldt, tstride);
	OZ_HIP(hipEventRecord(b->ev[4], st));
	int fallback = 0;
	if (nb1 == nbins) {
		launch_reconstruct(st, m, n, T, ldt, tstride, dLevel, nbins, Lmax, da.colE, db.colE, alpha,
		                   beta, C, ldc);
	} else {
		BigInt<9> R;
		big_zero<9>(R);
		for (const LevelBin &bn : bins)
			if (bn.level > L1)
				big_add_shifted<9>(R, (int64_t)bn.npairs * (int64_t)imax(0, imin(k, bn.k1) - bn.k0) * 16384,
				                   8 * (Lmax - bn.level));
		OZ_HIP(hipMemcpyAsync(dR, R.w, sizeof(R.w), hipMemcpyHostToDevice, st));
		int *dList = dMark + tilesM * tilesN;
		OZ_HIP(hipMemsetAsync(dMark, 0, sizeof(int) * (size_t)tilesM * tilesN, st));
		launch_reconstruct_check(st, m, n, T, ldt, tstride, dLevel, nb1, Lmax, da.colE, db.colE, alpha,
		                         beta, C, ldc, dR, tile, tilesM, dMark);
		std::vector<int> mark((size_t)tilesM * tilesN);
		OZ_HIP(hipMemcpyAsync(mark.data(), dMark, sizeof(int) * mark.size(), hipMemcpyDeviceToHost, st));
		OZ_TRY(sync(oh));
		std::vector<int> list;
		for (int q = 0; q < (int)mark.size(); q++)
			if (mark[q]) list.push_back(q);
		fallback = (int)list.size();
		if (fallback > 0) {
			OZ_HIP(hipMemcpyAsync(dList, list.data(), sizeof(int) * list.size(), hipMemcpyHostToDevice, st));
			launch_digit_gemm(st, elem, m, n, Kpad, da.planes, da.ldp, da.pstride, db.planes, db.ldp,
			                  db.pstride, dPairs, dBins, nbins - nb1, dOrder + nb1, T, ldt, tstride, dList,
			                  fallback);
			launch_reconstruct_fix(st, m, n, T, ldt, tstride, dLevel, nb1, nbins, Lmax, da.colE, db.colE,
			                       alpha, beta, C, ldc, dR, tile, tilesM, dList, fallback);
		}
	}
	OZ_HIP(hipEventRecord(b->ev[5], st));
	OZ_HIP(hipGetLastError());
	OZ_TRY(sync(oh));
	oh->fallbackTiles = fallback;
	oh->t_SplitA = ev_sec(b->ev[0], b->ev[1]);
	oh->t_SplitB = ev_sec(b->ev[1], b->ev[2]);
	oh->t_comp = ev_sec(b->ev[3], b->ev[4]);  // the digit-plane MFMA kernel alone
	oh->t_sum = ev_sec(b->ev[4], b->ev[5]);
	oh->nSplitA = (float)SA;
	oh->nSplitB = (float)SB;
	oh->nSplitC = (float)pairs1;  // pairs of the first pass (the kernel timed in t_comp)
	oh->mbk = m;
	oh->nbk = n;
	oh->t_total = tt.sec();
	oh->n_comp = (float)(2.0 * m * n * (double)k * pairs1 / ((double)Kpad / kc));
	wreset(oh, base);
	return OZBLAS_STATUS_SUCCESS;
}

// ---------------------------------------------------------------------------------------------
// GEMM, residue form (no reference counterpart; same bits as NearSum): INT8 (or FP16/BF16)
// products of residues modulo small coprime moduli, exact Chinese-remainder reconstruction
// (oz_residue.hip).  Returns with *handled = false when the moduli cannot cover the product (the
// caller then takes the digit form or the plain form).
// ---------------------------------------------------------------------------------------------
struct ResOperand {
	int *colE = nullptr, *colLo = nullptr;
	ExpStatsH *dst = nullptr;
	ExpStatsH st{};
	void *planes = nullptr;
	int64_t ldp = 0, pstride = 0;
};

static int32_t res_stats(ozblasHandle_t *oh, int k, int cols, const double *in, int64_t ldi,
                         ResOperand &r) {
	Backend *b = be(oh);
	r.colE = (int *)walloc(oh, sizeof(int) * (size_t)cols * 3);
	r.dst = (ExpStatsH *)walloc(oh, sizeof(ExpStatsH));
	if (!r.colE || !r.dst) return OZBLAS_STATUS_WORKSPACE;
	r.colLo = r.colE + 2 * (size_t)cols;
	static const ExpStatsH init{0, INT32_MAX, INT32_MIN, 0, INT32_MIN, INT32_MIN, {0, 0}};
	OZ_HIP(hipMemcpyAsync(r.dst, &init, sizeof(init), hipMemcpyHostToDevice, b->stream));
	launch_digit_stats(b->stream, k, cols, in, ldi, r.colE, r.colE + cols, r.dst, r.colLo);
	return OZBLAS_STATUS_SUCCESS;
}

static int32_t gemm_residue(ozblasHandle_t *oh, char transA, char transB, int m, int n, int k,
                            double alpha, const double *A, int lda, const double *B, int ldb,
                            double beta, double *C, int ldc, bool *handled) {
	Backend *b = be(oh);
	hipStream_t st = b->stream;
	*handled = false;
	const int elem = (oh->sliceType == OZBLAS_SLICE_INT8) ? 1 : (oh->sliceType == OZBLAS_SLICE_FP16 ? 2 : 3);
	const uint64_t base = wmark(oh);
	WallTimer tt;
	const int64_t ldk = rup(k, 16);
	const double *ain = A, *bin = B;
	int64_t ldai = lda, ldbi = ldb;
	OZ_HIP(hipEventRecord(b->ev[0], st));
	if (is_n(transA)) {
		double *tA = (double *)walloc(oh, 8ull * ldk * m);
		if (!tA) return OZBLAS_STATUS_WORKSPACE;
		launch_transpose(st, m, k, A, lda, tA, ldk);
		ain = tA;
		ldai = ldk;
	}
	if (!is_n(transB)) {
		double *tB = (double *)walloc(oh, 8ull * ldk * n);
		if (!tB) return OZBLAS_STATUS_WORKSPACE;
		launch_transpose(st, n, k, B, ldb, tB, ldk);
		bin = tB;
		ldbi = ldk;
	}
	ResOperand ra, rb;
	OZ_TRY(res_stats(oh, k, m, ain, ldai, ra));
	OZ_TRY(res_stats(oh, k, n, bin, ldbi, rb));
	OZ_HIP(hipMemcpyAsync(b->hPinned, ra.dst, sizeof(ExpStatsH), hipMemcpyDeviceToHost, st));
	OZ_HIP(hipMemcpyAsync((char *)b->hPinned + sizeof(ExpStatsH), rb.dst, sizeof(ExpStatsH),
	                      hipMemcpyDeviceToHost, st));
	OZ_TRY(sync(oh));  // the one host round trip: the statistics decide the number of moduli
	std::memcpy(&ra.st, b->hPinned, sizeof(ExpStatsH));
	std::memcpy(&rb.st, (char *)b->hPinned + sizeof(ExpStatsH), sizeof(ExpStatsH));
	auto give_up = [&]() {
		wreset(oh, base);
		return OZBLAS_STATUS_SUCCESS;
	};
	if (!fast_valid(oh, ra.st, rb.st, k)) return give_up();
	if (ra.st.maxS == 0 || rb.st.maxS == 0) {  // an all-zero operand: C = fma(alpha, +0, beta C)
		*handled = true;
		int *lv = (int *)walloc(oh, sizeof(int));
		if (!lv) return OZBLAS_STATUS_WORKSPACE;
		launch_reconstruct(st, m, n, nullptr, 0, 0, lv, 0, 0, ra.colE, rb.colE, alpha, beta, C, ldc);
		OZ_HIP(hipGetLastError());
		OZ_TRY(sync(oh));
		oh->nSplitA = oh->nSplitB = oh->nSplitC = 0;
		oh->schemeUsed = 2;
		wreset(oh, base);
		return OZBLAS_STATUS_SUCCESS;
	}
	// |X| must stay within M / 4; two bounds hold: |X| < k 2^(spanA + spanB) and, by Cauchy-Schwarz,
	// |X| <= ||a'_i|| ||b'_j|| < 2^(normA + normB); the smaller one decides
	int lk = 0;
	while ((int64_t(1) << lk) < (int64_t)k) lk++;
	const double need = (double)std::min(ra.st.maxSpan + rb.st.maxSpan + lk, ra.st.maxNorm + rb.st.maxNorm) + 2;
	if (ra.st.maxSpan >= kResMaxShift || rb.st.maxSpan >= kResMaxShift) return give_up();
	int nmod = 1;
	while (nmod <= kResMaxMod && res_log2_product(nmod) < need + 1e-6) nmod++;
	if (nmod > kResMaxMod) return give_up();
	const ResConst &rc = res_const_host(nmod);
	ResConst *dRc = (ResConst *)walloc(oh, sizeof(ResConst));
	if (elem == 1) {
		ra.ldp = rb.ldp = rup(k, 128);
		ra.pstride = ra.ldp * rup(m, 256);
		rb.pstride = rb.ldp * rup(n, 256);
	} else {
		ra.ldp = rb.ldp = rup(k, 64);
		ra.pstride = ra.ldp * m;
		rb.pstride = rb.ldp * n;
	}
	const size_t esz = (elem == 1) ? 1 : 2;
	ra.planes = walloc(oh, esz * (size_t)ra.pstride * nmod);
	rb.planes = walloc(oh, esz * (size_t)rb.pstride * nmod);
	const int Kpad = (int)ra.ldp;
	const int kc = kResChunk;
	const int nchunk = (Kpad + kc - 1) / kc;
	// bins: groups of consecutive moduli over one k chunk, one workgroup running its group's products
	// back to back; groups small enough to give the grid about 1024 (tile, bin) items
	const int tile = digit_gemm_tile(elem, m, n);
	const int64_t tiles = (int64_t)((m + tile - 1) / tile) * ((n + tile - 1) / tile);
	int ngroups = (int)std::min<int64_t>(nmod, std::max<int64_t>(1, (1024 + tiles * nchunk - 1) / (tiles * nchunk)));
	// one modulus per bin for the half-precision kernel and the 256 x 256 INT8 tile, whose
	// registers leave no room to flush inside the loop
	if (elem != 1 || tile == 256) ngroups = nmod;
	const int G = (nmod + ngroups - 1) / ngroups;
	std::vector<int2> pairs(nmod);
	std::vector<LevelBin> bins;
	for (int t = 0; t < nmod; t++) pairs[t] = make_int2(t, t);
	for (int c = 0; c < nchunk; c++)
		for (int t0 = 0; t0 < nmod; t0 += G) {
			LevelBin bn;
			bn.level = t0;
			bn.pair0 = t0;
			bn.npairs = std::min(G, nmod - t0);
			bn.k0 = c * kc;
			bn.k1 = std::min(Kpad, (c + 1) * kc);
			bins.push_back(bn);
		}
	const int nbins = (int)bins.size();
	std::vector<int> order(nbins);
	for (int q = 0; q < nbins; q++) order[q] = q;
	// product residues: whole INT8 tiles, or (FP16/BF16) every 4-column word of the m x n product
	const int64_t ldt = (elem == 1) ? rup(m, 256) : rup(m, 64);
	const int64_t tstride = ldt * ((elem == 1) ? rup(n, 256) : rup(n, 4));
	int2 *dPairs = (int2 *)walloc(oh, sizeof(int2) * pairs.size());
	LevelBin *dBins = (LevelBin *)walloc(oh, sizeof(LevelBin) * bins.size());
	int *dOrder = (int *)walloc(oh, sizeof(int) * order.size());
	uint8_t *Rp = (uint8_t *)walloc(oh, (size_t)tstride * nmod * nchunk);
	if (!dRc || !ra.planes || !rb.planes || !dPairs || !dBins || !dOrder || !Rp) {
		wreset(oh, base);
		return OZBLAS_STATUS_WORKSPACE;
	}
	*handled = true;
	OZ_HIP(hipMemcpyAsync(dRc, &rc, sizeof(ResConst), hipMemcpyHostToDevice, st));
	OZ_HIP(hipMemcpyAsync(dPairs, pairs.data(), sizeof(int2) * pairs.size(), hipMemcpyHostToDevice, st));
	OZ_HIP(hipMemcpyAsync(dBins, bins.data(), sizeof(LevelBin) * bins.size(), hipMemcpyHostToDevice, st));
	OZ_HIP(hipMemcpyAsync(dOrder, order.data(), sizeof(int) * order.size(), hipMemcpyHostToDevice, st));
	OZ_HIP(hipEventRecord(b->ev[1], st));
	launch_res_write(st, elem, k, m, ain, ldai, ra.colLo, dRc, nmod, ra.st.maxSpan, 1, ra.planes, ra.ldp, ra.pstride);
	OZ_HIP(hipEventRecord(b->ev[2], st));
	launch_res_write(st, elem, k, n, bin, ldbi, rb.colLo, dRc, nmod, rb.st.maxSpan, 0, rb.planes, rb.ldp, rb.pstride);
	OZ_HIP(hipEventRecord(b->ev[3], st));
	launch_digit_gemm(st, elem, m, n, Kpad, ra.planes, ra.ldp, ra.pstride, rb.planes, rb.ldp,
	                  rb.pstride, dPairs, dBins, nbins, dOrder, nullptr, ldt, tstride, nullptr, 0, dRc, Rp);
	OZ_HIP(hipEventRecord(b->ev[4], st));
	launch_res_recon(st, m, n, Rp, ldt, tstride, nchunk, dRc, nmod, rc.nl, ra.colLo, rb.colLo, alpha,
	                 beta, C, ldc);
	OZ_HIP(hipEventRecord(b->ev[5], st));
	OZ_HIP(hipGetLastError());
	OZ_TRY(sync(oh));
	oh->fallbackTiles = 0;
	oh->nLevelBins = nbins;
	oh->schemeUsed = 2;
	oh->t_SplitA = ev_sec(b->ev[0], b->ev[1]) + ev_sec(b->ev[1], b->ev[2]);  // transposes + stats + A
	oh->t_SplitB = ev_sec(b->ev[2], b->ev[3]);
	oh->t_comp = ev_sec(b->ev[3], b->ev[4]);  // the residue MFMA kernel alone
	oh->t_sum = ev_sec(b->ev[4], b->ev[5]);
	oh->nSplitA = (float)nmod;
	oh->nSplitB = (float)nmod;
	oh->nSplitC = (float)nmod;  // products: one per modulus
	oh->mbk = m;
	oh->nbk = n;
	oh->t_total = tt.sec();
	oh->n_comp = (float)(2.0 * m * n * (double)k * nmod);
	wreset(oh, base);
	return OZBLAS_STATUS_SUCCESS;
}

// ---------------------------------------------------------------------------------------------
// vector split (ozblasSplit with one column) -> slots at stride lds
// ---------------------------------------------------------------------------------------------
struct VecSplit {
	const void *x = nullptr;  // the vector itself
	double *slices = nullptr;
	int64_t lds = 0;
	int count = 0;     // host copy of the slice count (valid when read back)
	int *info = nullptr;  // device: info[0] slice count, info[1] first all-zero remainder
	VsPlan *plan = nullptr;
	VsRef ref() const { return VsRef{x, plan, slices, lds, info}; }
};

// Splits x; with readCount the slice count is read back to the host (one synchronisation),
// otherwise it stays on the device in vs.info for kernels that read it there.  With planOnly the
// slices are stored only for a plan outside the residue model: the consumer generates them from
// (x, plan) (VsRef).
// zScratch / zInfo: a zeroed scratch (vsplit_scratch_bytes()) and info (16 bytes) the caller
// cleared together with its other buffers (one fill in place of several)
template <class T>
static int32_t vec_split(ozblasHandle_t *oh, int n, const T *x, VecSplit &vs, bool readCount = true,
                         bool planOnly = false, void *zScratch = nullptr, int *zInfo = nullptr) {
	Backend *b = be(oh);
	const int cap = split_cap(oh);
	const int rho = Prec<T>::rho(oh, n);
	vs.x = x;
	vs.lds = rup(n, 32);
	vs.slices = (double *)walloc(oh, 8ull * vs.lds * (cap + 2));
	vs.info = zInfo ? zInfo : (int *)walloc(oh, 16);
	void *scratch = zScratch ? zScratch : walloc(oh, vsplit_scratch_bytes());
	vs.plan = (VsPlan *)walloc(oh, sizeof(VsPlan));
	if (!vs.slices || !vs.info || !scratch || !vs.plan) return OZBLAS_STATUS_WORKSPACE;
	if (!zScratch) OZ_HIP(hipMemsetAsync(scratch, 0, vsplit_scratch_bytes(), b->stream));
	if (planOnly)
		OZ_HIP(launch_vsplit_plan(b->stream, n, x, vs.slices, vs.lds, rho, oh->splitShift, cap,
		                          repro_slot(oh) ? 0 : 1, Prec<T>::Jsplit(b), vs.info, scratch, nullptr,
		                          vs.plan, false));
	else
		OZ_HIP(launch_vsplit(b->stream, n, x, vs.slices, vs.lds, rho, oh->splitShift, cap,
		                     repro_slot(oh) ? 0 : 1, Prec<T>::Jsplit(b), vs.info, scratch, vs.plan));
	if (!readCount) return OZBLAS_STATUS_SUCCESS;
	OZ_HIP(hipMemcpyAsync(b->hPinned, vs.info, sizeof(int), hipMemcpyDeviceToHost, b->stream));
	OZ_TRY(sync(oh));
	vs.count = ((int32_t *)b->hPinned)[0];
	return OZBLAS_STATUS_SUCCESS;
}

// DOT core: xs/ys already split; result to device scalar dres (fma(1, t, 0*dres))
static int32_t dot_core(ozblasHandle_t *oh, int n, const VecSplit &xs, const VecSplit &ys,
                        double *dres) {
	Backend *b = be(oh);
	const int nA = xs.count, nB = ys.count;
	if (oh->splitEpsMode == 2) {  // Dot2: (hi, lo) of every slice pair, as 2 nB B slices (:166-173)
		double *vals = (double *)walloc(oh, 16ull * nA * nB);
		void *part = walloc(oh, dot2_partials_bytes(nA, nB));
		if (!vals || !part) return OZBLAS_STATUS_WORKSPACE;
		launch_dot2_slices(b->stream, n, xs.slices, xs.lds, nA, ys.slices, ys.lds, nB, part, vals);
		const int maxlevel = (nA - 1) + (2 * nB - 1);  // :281-285
		launch_nearsum_small(b->stream, 2 * nA * nB, vals, 1.0, 0.0, dres, 2, nA, 2 * nB, maxlevel,
		                     oh->sumMode == 0 ? 1 : 0, b->dTau);
		return OZBLAS_STATUS_SUCCESS;
	}
	double *vals = (double *)walloc(oh, 8ull * nA * nB);
	if (!vals) return OZBLAS_STATUS_WORKSPACE;
	OZ_HIP(hipMemsetAsync(vals, 0, 8ull * nA * nB, b->stream));
	launch_slice_dots(b->stream, n, xs.slices, xs.lds, nA, ys.slices, ys.lds, nB, vals);
	const int maxlevel = (nA - 1) + (nB - 1);  // fastMode is forced off for DOT (:25-28)
	launch_nearsum_small(b->stream, nA * nB, vals, 1.0, 0.0, dres, 2, nA, nB, maxlevel,
	                     oh->sumMode == 0 ? 1 : 0, b->dTau);
	return OZBLAS_STATUS_SUCCESS;
}

template <class T>
static int32_t dot_impl(ozblasHandle_t *oh, int n, const T *x, const T *y, T *ret) {
	Backend *b = be(oh);
	constexpr bool f32 = sizeof(T) == 4;
	if (oh->sumMode != 0 && oh->sumMode != 1) return OZBLAS_STATUS_NOT_SUPPORTED;
	if (n <= 0) {  // empty vectors: the sum is +0 (fma(1, +0, 0 * ret) with ret = +0)
		*ret = 0;
		return OZBLAS_STATUS_SUCCESS;
	}
	// the reference's binary32 Dot2 GEMM is a stub (src/cpu/ozblas_blas_wrapper.cpp:179-181); its
	// mixed NearSum exits (src/cpu/ozblas_sum.cpp:275-330)
	if ((f32 || b->mixed) && oh->splitEpsMode == 2) return OZBLAS_STATUS_NOT_SUPPORTED;
	if (b->mixed && oh->sumMode != 0) return OZBLAS_STATUS_NOT_SUPPORTED;
	const uint64_t base = wmark(oh);
	oh->splitShift = 1;  // counterInit (src/cpu/ozblas.cpp:8)
	const bool dot2 = (oh->splitEpsMode == 2);
	VecSplit xs, ys;
	if (dot2) {
		OZ_TRY(vec_split(oh, n, x, xs, true, false));
		if (y == x)
			ys = xs;
		else
			OZ_TRY(vec_split(oh, n, y, ys, true, false));
		T *dres = (T *)walloc(oh, 8);
		if (!dres) return OZBLAS_STATUS_WORKSPACE;
		OZ_HIP(hipMemsetAsync(dres, 0, 8, b->stream));
		if constexpr (!f32) OZ_TRY(dot_core(oh, n, xs, ys, dres));
		OZ_HIP(hipMemcpyAsync(b->hPinned, dres, sizeof(T), hipMemcpyDeviceToHost, b->stream));
		OZ_TRY(sync(oh));
		std::memcpy(ret, b->hPinned, sizeof(T));
	} else {
		// one zeroed block [scratch x | scratch y | pair sums | info x | info y | result] (one fill),
		// the slice counts stay on the device, one launch for the pair dots and their sum, and the
		// counts and the result come back in one copy
		const int cap = split_cap(oh);
		const size_t sb = (size_t)rup((int64_t)vsplit_scratch_bytes(), 256);
		const size_t vb = (size_t)rup(8ll * (cap + 1) * (cap + 1), 256);
		char *Z = (char *)walloc(oh, 2 * sb + vb + 48);
		double *part = (double *)walloc(oh, dots_part_bytes());
		if (!Z || !part) return OZBLAS_STATUS_WORKSPACE;
		int *R = (int *)(Z + 2 * sb + vb);  // [info x (4 ints) | info y (4 ints) | result (8 bytes)]
		T *dres = (T *)(R + 8);
		OZ_HIP(hipMemsetAsync(Z, 0, 2 * sb + vb + 48, b->stream));
		if (y == x) {
			OZ_TRY(vec_split(oh, n, x, xs, false, true, Z, R));
			ys = xs;
		} else {  // both plans by the same three launches
			VecSplit *v[2] = {&xs, &ys};
			const T *src[2] = {x, y};
			double *sl[2];
			int *inf[2];
			void *scr[2];
			VsPlan *pl[2];
			for (int q = 0; q < 2; q++) {
				v[q]->x = src[q];
				v[q]->lds = rup(n, 32);
				v[q]->slices = sl[q] = (double *)walloc(oh, 8ull * v[q]->lds * (cap + 2));
				v[q]->plan = pl[q] = (VsPlan *)walloc(oh, sizeof(VsPlan));
				v[q]->info = inf[q] = R + 4 * q;
				scr[q] = Z + q * sb;
				if (!sl[q] || !pl[q]) return OZBLAS_STATUS_WORKSPACE;
			}
			OZ_HIP(launch_vsplit_plan2(b->stream, n, x, y, Prec<T>::rho(oh, n), oh->splitShift, cap,
			                           repro_slot(oh) ? 0 : 1, Prec<T>::Jsplit(b), sl, xs.lds, inf, scr, pl));
		}
		launch_dots_dev(b->stream, n, xs.ref(), ys.ref(), (double *)(Z + 2 * sb), part, Prec<T>::Jsum(b),
		                oh->sumMode == 0 ? 1 : 0, nullptr, 0, 0, dres, f32 ? 1 : 0);
		OZ_HIP(hipMemcpyAsync(b->hPinned, R, 48, hipMemcpyDeviceToHost, b->stream));
		OZ_TRY(sync(oh));
		const int32_t *hr = (const int32_t *)b->hPinned;
		std::memcpy(ret, hr + 8, sizeof(T));
		xs.count = hr[0];
		ys.count = (y == x) ? hr[0] : hr[4];
	}
	oh->nSplitA = (float)xs.count;
	oh->nSplitB = (float)ys.count;
	oh->nSplitC = 1;
	wreset(oh, base);
	return OZBLAS_STATUS_SUCCESS;
}

// ---------------------------------------------------------------------------------------------
// CSR (src/cpu/ozblas_csrmv.cpp)
// ---------------------------------------------------------------------------------------------
struct CsrSplit {
	double *slices = nullptr;
	int64_t lds = 0;
	int count = 0;
};

template <class T>
static int32_t csr_split(ozblasHandle_t *oh, int m, int nnz, const T *val, const int *rowptr,
                         CsrSplit &cs) {
	Backend *b = be(oh);
	const int cap = split_cap(oh);
	cs.lds = rup(nnz, 32);
	T *work = (T *)walloc(oh, sizeof(T) * cs.lds);
	int *zrow = (int *)walloc(oh, sizeof(int) * (size_t)m);
	if (!work || !zrow) return OZBLAS_STATUS_WORKSPACE;
	if (!cs.slices) {
		cs.slices = (double *)walloc(oh, 8ull * cs.lds * (cap + 2));
		if (!cs.slices) return OZBLAS_STATUS_WORKSPACE;
	}
	OZ_HIP(hipMemsetAsync(b->dCounters, 0, sizeof(int), b->stream));
	launch_split_csr(b->stream, m, val, rowptr, work, cs.slices, cs.lds, oh->splitShift, cap,
	                 Prec<T>::Jsplit(b), zrow, b->dCounters, oh->splitEpsMode, b->mixed);
	OZ_HIP(hipMemcpyAsync(b->hPinned, b->dCounters, sizeof(int), hipMemcpyDeviceToHost, b->stream));
	OZ_TRY(sync(oh));
	const int zmax = ((int32_t *)b->hPinned)[0];
	// ozblasSplitSparse returns s+1 at the first zero without the reproMode==0 extra slot
	// (src/cpu/ozblas_split.cpp:919-927)
	int count;
	if (cap <= 1)
		count = 1 + (repro_slot(oh) ? 1 : 0);
	else if (zmax < cap)
		count = imax(zmax, 1) + 1;
	else
		count = cap + (repro_slot(oh) ? 1 : 0);
	cs.count = count;
	launch_zero_slots_csr(b->stream, m, rowptr, cs.slices, cs.lds, zrow, count);
	return OZBLAS_STATUS_SUCCESS;
}

template <class T>
static int32_t csrmv_core(ozblasHandle_t *oh, int m, int n, int nnz, T alpha, const CsrSplit &as,
                          const int *colind, const int *rowptr, const VecSplit &xs, T beta, T *y) {
	Backend *b = be(oh);
	const int nA = as.count, nB = xs.count;
	int maxlevel = (nA - 1) + (nB - 1);
	maxlevel = imax(maxlevel - oh->fastMode, imin(nA - 1, nB - 1));
	const bool fsum = oh->sumMode == 0;
	if (fsum && oh->fastMode) return OZBLAS_STATUS_NOT_SUPPORTED;  // reference reads unwritten slots
	if (b->mixed && (!fsum || oh->splitEpsMode == 2)) return OZBLAS_STATUS_NOT_SUPPORTED;
	if constexpr (sizeof(T) == 4) {  // no binary32 Dot2 (src/cpu/ozblas_blas_wrapper.cpp:179-181)
		if (oh->splitEpsMode == 2) return OZBLAS_STATUS_NOT_SUPPORTED;
	} else if (oh->splitEpsMode == 2) {  // Dot2 slice products (:139-142, 159-162)
		// with fastMode the reference's 2 numB-column blocks no longer match its summation index
		if (oh->fastMode || nA > 4 || nB > 4) return OZBLAS_STATUS_NOT_SUPPORTED;
		launch_csrmm_sum_x2(b->stream, m, as.slices, as.lds, nA, colind, rowptr, xs.slices, xs.lds, nB,
		                    fsum ? 1 : 0, alpha, beta, y, b->dTau);
		(void)n;
		(void)maxlevel;
		return OZBLAS_STATUS_SUCCESS;
	}
	if (nA <= 4 && nB <= 8) {
		if constexpr (sizeof(T) == 4)
			launch_csrmm_sum(b->stream, m, nnz, as.slices, as.lds, nA, colind, rowptr, xs.slices, xs.lds, nB,
			                 maxlevel, fsum ? 1 : 0, alpha, beta, y, Prec<T>::Jsum(b), b->mixed);
		else  // <double,float>: the binary32 chain of the reference is exact, so the binary64 one
			launch_csrmm_sum(b->stream, m, nnz, as.slices, as.lds, nA, colind, rowptr, xs.slices, xs.lds, nB,
			                 maxlevel, fsum ? 1 : 0, alpha, beta, y, Prec<T>::Jsum(b));
	} else {
		int ic = 0;
		for (int ia = 0; ia < imin(maxlevel + 1, nA); ia++) ic += imin(nB, maxlevel + 1 - ia);
		const int64_t ldcs = rup(m, 8);
		T *Cs = (T *)walloc(oh, sizeof(T) * ldcs * ((int64_t)nA * nB > ic ? (int64_t)nA * nB : ic));
		if (!Cs) return OZBLAS_STATUS_WORKSPACE;
		if constexpr (sizeof(T) == 4)
			launch_csrmm_store(b->stream, m, as.slices, as.lds, nA, colind, rowptr, xs.slices, xs.lds, nB,
			                   maxlevel, Cs, ldcs, b->mixed);
		else
			launch_csrmm_store(b->stream, m, as.slices, as.lds, nA, colind, rowptr, xs.slices, xs.lds, nB,
			                   maxlevel, Cs, ldcs);
		if (fsum)
			launch_global_fsum(b->stream, m, 1, nA, nB, maxlevel, 3, Cs, ldcs, ldcs, y, 1, alpha, beta);
		else
			launch_global_nearsum(b->stream, m, 1, ic, Cs, ldcs, ldcs, y, 1, alpha, beta,
			                      Prec<T>::Jsum(b));
	}
	(void)n;
	return OZBLAS_STATUS_SUCCESS;
}

}  // namespace oz

using namespace oz;

// =============================================================================================
// C ABI
// =============================================================================================
namespace oz {

template <class T>
static int32_t gemm_k0(ozblasHandle_t *oh, int m, int n, T alpha, T beta, T *C, int ldc) {
	launch_gemm_k0(be(oh)->stream, m, n, alpha, beta, C, ldc);
	OZ_HIP(hipGetLastError());
	return sync(oh);
}

// ozblasRdot (src/cpu/ozblas_dot.cpp): the 1 x 1 x n GEMM of the split vectors
template <class T>
static int32_t dot_entry(ozblasHandle_t *oh, int32_t n, const T *x, int32_t incx, const T *y,
                         int32_t incy, T *ret) {
	OZ_TRY(ensure_init(oh));
	if (incx != 1 || incy != 1) return oh->lastStatus = OZBLAS_STATUS_NOT_SUPPORTED;
	if (n <= 0) {
		*ret = 0;
		return oh->lastStatus = OZBLAS_STATUS_SUCCESS;
	}
	if (oh->reproMode == 0 && oh->nSplitMax == 1) return oh->lastStatus = OZBLAS_STATUS_NOT_SUPPORTED;
	return oh->lastStatus = dot_impl(oh, n, x, y, ret);
}

template <class T>
static int32_t axpy_entry(ozblasHandle_t *oh, int32_t n, T alpha, const T *x, int32_t incx, T *y,
                          int32_t incy) {
	OZ_TRY(ensure_init(oh));
	if (incx != 1 || incy != 1) return oh->lastStatus = OZBLAS_STATUS_NOT_SUPPORTED;
	if (n <= 0) return OZBLAS_STATUS_SUCCESS;
	launch_axpy_host(be(oh)->stream, n, alpha, x, y, oh->reproMode != 0);
	OZ_HIP(hipGetLastError());
	return oh->lastStatus = sync(oh);
}

}  // namespace oz

extern "C" int32_t ozblasDgemm(ozblasHandle_t *oh, char transA, char transB, int32_t m, int32_t n,
                               int32_t k, double alpha, const double *A, int32_t lda,
                               const double *B, int32_t ldb, double beta, double *C, int32_t ldc) {
	OZ_TRY(ensure_init(oh));
	if (m < 0 || n < 0 || k < 0) return oh->lastStatus = OZBLAS_STATUS_INVALID_VALUE;
	if (m == 0 || n == 0) return oh->lastStatus = OZBLAS_STATUS_SUCCESS;
	oh->splitShift = 1;  // counterInit
	if (oh->reproMode == 0 && oh->nSplitMax == 1) return oh->lastStatus = OZBLAS_STATUS_NOT_SUPPORTED;
	if (k == 0)  // no slices to multiply: the reference's sum is +0
		return oh->lastStatus = gemm_k0(oh, m, n, alpha, beta, C, ldc);
	oh->schemeUsed = 0;
	if (oh->sliceType != OZBLAS_SLICE_FP64) {
		if (oh->sumMode == 1 && oh->nSplitMax == 0) {
			bool handled = false;
			// residues first (automatic unless digit levels are asked for, or forced), then digits
			const bool residues = oh->digitScheme == 2 || (oh->digitScheme == 0 && oh->digitLevels == 0);
			if (residues) {
				int32_t s = gemm_residue(oh, transA, transB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc,
				                         &handled);
				if (s != OZBLAS_STATUS_SUCCESS || handled) return oh->lastStatus = s;
			}
			int32_t s = gemm_fast(oh, transA, transB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, &handled);
			if (handled && s == OZBLAS_STATUS_SUCCESS) oh->schemeUsed = 1;
			if (s != OZBLAS_STATUS_SUCCESS || handled) return oh->lastStatus = s;
		} else if (oh->sumMode != 1) {
			return oh->lastStatus = OZBLAS_STATUS_NOT_SUPPORTED;
		}
	}
	return oh->lastStatus = gemm_plain(oh, transA, transB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

// binary32: the plain form only (sliceType is not consulted: the digit forms are binary64)
extern "C" int32_t ozblasSgemm(ozblasHandle_t *oh, char transA, char transB, int32_t m, int32_t n,
                               int32_t k, float alpha, const float *A, int32_t lda, const float *B,
                               int32_t ldb, float beta, float *C, int32_t ldc) {
	OZ_TRY(ensure_init(oh));
	if (m < 0 || n < 0 || k < 0) return oh->lastStatus = OZBLAS_STATUS_INVALID_VALUE;
	if (m == 0 || n == 0) return oh->lastStatus = OZBLAS_STATUS_SUCCESS;
	oh->splitShift = 1;  // counterInit
	if (oh->reproMode == 0 && oh->nSplitMax == 1) return oh->lastStatus = OZBLAS_STATUS_NOT_SUPPORTED;
	// the reference's binary32 Dot2 GEMM is a stub (src/cpu/ozblas_blas_wrapper.cpp:179-181)
	if (oh->splitEpsMode == 2) return oh->lastStatus = OZBLAS_STATUS_NOT_SUPPORTED;
	if (k == 0) return oh->lastStatus = gemm_k0(oh, m, n, alpha, beta, C, ldc);
	return oh->lastStatus = gemm_plain(oh, transA, transB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

extern "C" int32_t ozblasDgemv(ozblasHandle_t *oh, char transA, int32_t m, int32_t n, double alpha,
                               const double *A, int32_t lda, const double *x, int32_t incx,
                               double beta, double *y, int32_t incy) {
	OZ_TRY(ensure_init(oh));
	if (incx != 1 || incy != 1) return oh->lastStatus = OZBLAS_STATUS_NOT_SUPPORTED;
	const int32_t saveSlice = oh->sliceType;
	oh->sliceType = OZBLAS_SLICE_FP64;
	int32_t s;
	if (is_n(transA))  // src/cpu/ozblas_gemv.cpp:23-26
		s = ozblasDgemm(oh, 'n', 'n', m, 1, n, alpha, A, lda, x, n, beta, y, m);
	else
		s = ozblasDgemm(oh, 't', 'n', n, 1, m, alpha, A, lda, x, m, beta, y, n);
	oh->sliceType = saveSlice;
	return s;
}

extern "C" int32_t ozblasSgemv(ozblasHandle_t *oh, char transA, int32_t m, int32_t n, float alpha,
                               const float *A, int32_t lda, const float *x, int32_t incx, float beta,
                               float *y, int32_t incy) {
	OZ_TRY(ensure_init(oh));
	if (incx != 1 || incy != 1) return oh->lastStatus = OZBLAS_STATUS_NOT_SUPPORTED;
	if (is_n(transA))  // src/cpu/ozblas_gemv.cpp:23-26
		return ozblasSgemm(oh, 'n', 'n', m, 1, n, alpha, A, lda, x, n, beta, y, m);
	return ozblasSgemm(oh, 't', 'n', n, 1, m, alpha, A, lda, x, m, beta, y, n);
}

extern "C" int32_t ozblasDdot(ozblasHandle_t *oh, int32_t n, const double *x, int32_t incx,
                              const double *y, int32_t incy, double *ret) {
	return dot_entry(oh, n, x, incx, y, incy, ret);
}

// Bytes the binary64 DOT takes from the workspace: the bump allocations of dot_impl / vec_split in
// their order (256-byte aligned), with the Dot2 buffers at the largest slice counts.
static uint64_t dot_workspace_bytes(const ozblasHandle_t *oh, int n, bool same) {
	auto bump = [](uint64_t top, uint64_t bytes) { return ((top + 255) & ~uint64_t(255)) + bytes; };
	const int cap = split_cap(oh);
	const uint64_t lds = (uint64_t)rup(n, 32);
	uint64_t t = 0;
	if (oh->splitEpsMode == 2) {  // dot_impl's Dot2 order: per vector, then result and products
		for (int v = 0; v < (same ? 1 : 2); v++) {
			t = bump(t, 8ull * lds * (cap + 2));  // slices
			t = bump(t, 16);                      // count
			t = bump(t, vsplit_scratch_bytes());
			t = bump(t, sizeof(VsPlan));
		}
		t = bump(t, 8);  // result
		t = bump(t, 16ull * (cap + 2) * (cap + 2));
		t = bump(t, dot2_partials_bytes(cap + 2, cap + 2));
		return t;
	}
	// the zeroed block (two scratches, pair sums, counts and result), the partials, per vector
	const uint64_t sb = (uint64_t)rup((int64_t)vsplit_scratch_bytes(), 256);
	const uint64_t vb = (uint64_t)rup(8ll * (cap + 1) * (cap + 1), 256);
	t = bump(t, 2 * sb + vb + 48);
	t = bump(t, dots_part_bytes());
	for (int v = 0; v < (same ? 1 : 2); v++) {
		t = bump(t, 8ull * lds * (cap + 2));  // slices (stored for a plan outside the model)
		t = bump(t, sizeof(VsPlan));
	}
	return t;
}

// ozblasDdot with the library's static scratch array as its workspace: no allocation on any path.
// The handle's own workspace is swapped out for the call and restored after it.
extern "C" int32_t ozblasDdotNoAlloc(ozblasHandle_t *oh, int32_t n, const double *x, int32_t incx,
                                     const double *y, int32_t incy, double *ret) {
	OZ_TRY(ensure_init(oh));
	Backend *b = be(oh);
	if (b->scratch == nullptr) return oh->lastStatus = OZBLAS_STATUS_NOT_INITIALIZED;
	if (n > 0 && dot_workspace_bytes(oh, n, x == y) > kStaticScratchBytes)
		return oh->lastStatus = OZBLAS_STATUS_WORKSPACE;  // beyond the static scratch array
	char *work = b->work;
	const uint64_t cap = b->cap, top = b->top, memAddr = oh->memAddr;
	b->work = b->scratch;
	b->cap = kStaticScratchBytes;
	b->top = 0;
	const int32_t st = dot_entry(oh, n, x, incx, y, incy, ret);
	b->work = work;
	b->cap = cap;
	b->top = top;
	oh->memAddr = memAddr;
	return st;
}

extern "C" int32_t ozblasSdot(ozblasHandle_t *oh, int32_t n, const float *x, int32_t incx,
                              const float *y, int32_t incy, float *ret) {
	return dot_entry(oh, n, x, incx, y, incy, ret);
}

extern "C" int32_t ozblasDnrm2(ozblasHandle_t *oh, int32_t n, const double *x, int32_t incx,
                               double *ret) {
	OZ_TRY(ensure_init(oh));
	double d = 0.0;
	int32_t s = ozblasDdot(oh, n, x, incx, x, incx, &d);
	if (s != OZBLAS_STATUS_SUCCESS) return s;
	*ret = sqrt(d);  // host sqrt as the reference (src/cpu/ozblas_nrm2.cpp:27)
	return OZBLAS_STATUS_SUCCESS;
}

extern "C" int32_t ozblasSnrm2(ozblasHandle_t *oh, int32_t n, const float *x, int32_t incx,
                               float *ret) {
	OZ_TRY(ensure_init(oh));
	float d = 0.f;
	int32_t s = ozblasSdot(oh, n, x, incx, x, incx, &d);
	if (s != OZBLAS_STATUS_SUCCESS) return s;
	*ret = (float)sqrt((double)d);  // the double sqrt, rounded (src/cpu/ozblas_nrm2.cpp:27)
	return OZBLAS_STATUS_SUCCESS;
}

extern "C" int32_t ozblasDaxpy(ozblasHandle_t *oh, int32_t n, double alpha, const double *x,
                               int32_t incx, double *y, int32_t incy) {
	return axpy_entry(oh, n, alpha, x, incx, y, incy);
}

extern "C" int32_t ozblasSaxpy(ozblasHandle_t *oh, int32_t n, float alpha, const float *x,
                               int32_t incx, float *y, int32_t incy) {
	return axpy_entry(oh, n, alpha, x, incx, y, incy);
}

namespace oz {
// ozblasRcsrmvSplitA (src/cpu/ozblas_csrmv.cpp:191-271): the slices stay in the workspace (kept
// across calls until the next CG or SplitA); returns them, or nullptr with lastStatus set
template <class T>
static double *csrmv_split_a(ozblasHandle_t *oh, char transA, int32_t m, int32_t nnz, const T *A,
                             const int32_t *rowptr) {
	if (ensure_init(oh) != OZBLAS_STATUS_SUCCESS) return nullptr;
	if (!is_n(transA)) {
		oh->lastStatus = OZBLAS_STATUS_NOT_SUPPORTED;
		return nullptr;
	}
	Backend *b = be(oh);
	b->keep = 0;
	wreset(oh, 0);
	oh->splitShift = 1;
	CsrSplit cs;
	cs.lds = rup(nnz, 32);
	cs.slices = (double *)walloc(oh, 8ull * cs.lds * (split_cap(oh) + 2));
	if (!cs.slices) {
		oh->lastStatus = OZBLAS_STATUS_WORKSPACE;
		return nullptr;
	}
	const uint64_t keepTop = wmark(oh);
	int32_t s = csr_split(oh, m, nnz, A, rowptr, cs);
	int nA = cs.count;
	// splitShift tuning (src/cpu/ozblas_csrmv.cpp:240-255)
	if (s == OZBLAS_STATUS_SUCCESS && oh->nSplitMax == 0) {
		int nAOld;
		oh->splitShift = 1;
		do {
			nAOld = nA;
			oh->splitShift *= 2;
			wreset(oh, keepTop);
			s = csr_split(oh, m, nnz, A, rowptr, cs);
			nA = cs.count;
		} while (s == OZBLAS_STATUS_SUCCESS && nAOld == nA && oh->splitShift < 512);
		if (nAOld == nA) oh->splitShift = 1;
		wreset(oh, keepTop);
		if (s == OZBLAS_STATUS_SUCCESS) s = csr_split(oh, m, nnz, A, rowptr, cs);
		nA = cs.count;
	}
	oh->lastStatus = s;
	if (s != OZBLAS_STATUS_SUCCESS) return nullptr;
	oh->nSplitA_ = nA;
	oh->nSplitA = (float)nA;
	b->keep = keepTop;
	oh->memMaskSplitA = keepTop;
	wreset(oh, keepTop);
	return cs.slices;
}

// A: the matrix values, or (memMaskSplitA != 0) the slices returned by csrmv_split_a
template <class T>
static int32_t csrmv_impl(ozblasHandle_t *oh, int m, int n, int nnz, T alpha, const void *A,
                          const int *colind, const int *rowptr, const T *x, T beta, T *y) {
	Backend *b = be(oh);
	if (oh->sumMode != 0 && oh->sumMode != 1) return OZBLAS_STATUS_NOT_SUPPORTED;
	const uint64_t base = (oh->memMaskSplitA != 0) ? b->keep : 0;
	wreset(oh, base);
	CsrSplit as;
	if (oh->memMaskSplitA != 0) {  // pre-split by ozblasDcsrmvSplitA (src/cpu/ozblas_csrmv.cpp:95-97)
		as.slices = (double *)A;
		as.lds = rup(nnz, 32);
		as.count = oh->nSplitA_;
	} else {
		oh->splitShift = 1;
		OZ_TRY(csr_split(oh, m, nnz, (const T *)A, rowptr, as));
	}
	oh->splitShift = 1;  // the vector split always uses splitShift = 1 (counterInit)
	VecSplit xs;
	OZ_TRY(vec_split(oh, n, x, xs));
	OZ_TRY(csrmv_core(oh, m, n, nnz, alpha, as, colind, rowptr, xs, beta, y));
	OZ_HIP(hipGetLastError());
	OZ_TRY(sync(oh));
	oh->nSplitA = (float)as.count;
	oh->nSplitB = (float)xs.count;
	wreset(oh, base);
	return OZBLAS_STATUS_SUCCESS;
}

template <class T>
static int32_t csrmv_entry(ozblasHandle_t *oh, char transA, int32_t m, int32_t n, int32_t nnz,
                           T alpha, const T *A, const int32_t *colind, const int32_t *rowptr,
                           const T *x, T beta, T *y) {
	OZ_TRY(ensure_init(oh));
	if (!is_n(transA)) return oh->lastStatus = OZBLAS_STATUS_NOT_SUPPORTED;
	if (oh->reproMode == 0 && oh->nSplitMax == 1) return oh->lastStatus = OZBLAS_STATUS_NOT_SUPPORTED;
	return oh->lastStatus = csrmv_impl(oh, m, n, nnz, alpha, (const void *)A, colind, rowptr, x, beta, y);
}
}  // namespace oz

extern "C" double *ozblasDcsrmvSplitA(ozblasHandle_t *oh, char transA, int32_t m, int32_t n,
                                      int32_t nnz, const void *descrA, const double *A,
                                      const int32_t *rowptr) {
	(void)descrA;
	(void)n;
	return csrmv_split_a(oh, transA, m, nnz, A, rowptr);
}

extern "C" float *ozblasScsrmvSplitA(ozblasHandle_t *oh, char transA, int32_t m, int32_t n,
                                     int32_t nnz, const void *descrA, const float *A,
                                     const int32_t *rowptr) {
	(void)descrA;
	(void)n;
	return (float *)csrmv_split_a(oh, transA, m, nnz, A, rowptr);
}

extern "C" int32_t ozblasDcsrmv(ozblasHandle_t *oh, char transA, int32_t m, int32_t n, int32_t nnz,
                                double alpha, const void *descrA, const double *A,
                                const int32_t *colind, const int32_t *rowptr, const double *x,
                                double beta, double *y) {
	(void)descrA;
	return csrmv_entry(oh, transA, m, n, nnz, alpha, A, colind, rowptr, x, beta, y);
}

extern "C" int32_t ozblasScsrmv(ozblasHandle_t *oh, char transA, int32_t m, int32_t n, int32_t nnz,
                                float alpha, const void *descrA, const float *A,
                                const int32_t *colind, const int32_t *rowptr, const float *x,
                                float beta, float *y) {
	(void)descrA;
	return csrmv_entry(oh, transA, m, n, nnz, alpha, A, colind, rowptr, x, beta, y);
}

// CG (src/cpu/ozblas_cg.cpp:24-128).  The set-up (pre-split of A, |b|, r0 = b - A x0, <r0,r0>)
// runs through the synchronous entry points; the iterations then run device-resident: every
// iteration is enqueued without waiting (cg_device_loop).  When a vector has more slices than the
// device CSR product holds, or the options leave the device path, the iterations run on the
// synchronous path with the scalar recurrences on the host (cg_host_loop); both give the same bits.
namespace oz {

template <class T>
struct CgVecs {
	T *R = nullptr, *P = nullptr, *Q = nullptr, *X0 = nullptr, *T_ = nullptr;
};

// The second history of ozblasRcg (src/cpu/ozblas_cg.cpp:47-49,64-69,101-109), written next to
// each entry of cg_verbose1: with trueresFlag the true relative residual |b - A x| / |b| (a T array),
// otherwise the seconds since the start of the solve (doubles).  The true residual takes t = b - A x
// from the reference's plain CSRMV and |t| as sqrt(RN(sum t_i^2)) (an Ozaki DOT under NearSum); the
// reference's cblas_?nrm2 rounds in an order of its BLAS's own choosing.
template <class T>
struct CgDiag {
	void *v2 = nullptr;
	bool trueres = false;
	WallTimer t1;
	const T *A = nullptr, *b = nullptr;
	const int *colind = nullptr, *rowptr = nullptr;
	T *t = nullptr;
	int n = 0;
	T nrmb = 0;
};

template <class T>
static int32_t cg_record2(ozblasHandle_t *oh, const CgDiag<T> &d, int idx, const T *x) {
	if (!d.v2) return OZBLAS_STATUS_SUCCESS;
	if (!d.trueres) {
		((double *)d.v2)[idx] = d.t1.sec();
		return OZBLAS_STATUS_SUCCESS;
	}
	launch_true_residual(be(oh)->stream, d.n, d.A, d.colind, d.rowptr, x, d.b, d.t);
	// a correctly rounded sum of squares: NearSum with the exact binary64 (or binary32) split,
	// whatever the handle holds
	const int32_t sv[5] = {oh->sumMode, oh->reproMode, oh->nSplitMax, oh->splitEpsMode, oh->splitShift};
	const float st[3] = {oh->nSplitA, oh->nSplitB, oh->nSplitC};
	const int mixed = be(oh)->mixed;
	be(oh)->mixed = 0;
	oh->sumMode = 1;
	oh->reproMode = 1;
	oh->nSplitMax = 0;
	oh->splitEpsMode = 0;
	T dd = 0;
	const int32_t s = dot_impl(oh, d.n, (const T *)d.t, (const T *)d.t, &dd);
	oh->sumMode = sv[0];
	oh->reproMode = sv[1];
	oh->nSplitMax = sv[2];
	oh->splitEpsMode = sv[3];
	oh->splitShift = sv[4];
	oh->nSplitA = st[0];
	oh->nSplitB = st[1];
	oh->nSplitC = st[2];
	be(oh)->mixed = mixed;
	OZ_TRY(s);
	((T *)d.v2)[idx] = (T)sqrt((double)dd) / d.nrmb;
	return OZBLAS_STATUS_SUCCESS;
}

// T = float: the <float,float> instantiation (float scalars, the double sqrt rounded to float).
// The time totals of the handle follow the reference's categories (src/cpu/ozblas_cg.cpp:89,99,
// ozblas_csrmv.cpp:172-177, ozblas_dot.cpp:30-35) with the vector split and the sum of a CSRMV or DOT
// counted in its SpMV or DOT total, where they run fused.
template <class T>
static int32_t cg_host_loop(ozblasHandle_t *oh, int n, int nnz, const double *AS, const int *colind,
                            const int *rowptr, T *x, int maxiter, T tol, T nrmb, T dold,
                            const CgVecs<T> &v, T *hist, const CgDiag<T> &diag) {
	Backend *b = be(oh);
	hipStream_t st = b->stream;
	T tmp, alpha, dnew, resi, beta;
	int vcnt = 1;
	while (oh->cg_numiter < maxiter) {
		oh->cg_numiter++;
		WallTimer tc;
		OZ_TRY(csrmv_impl(oh, n, n, nnz, (T)1, (const void *)AS, colind, rowptr, v.P, (T)0, v.Q));
		oh->t_SpMV_SpMM_total += tc.sec();
		WallTimer td;
		OZ_TRY(dot_entry(oh, n, v.P, 1, v.Q, 1, &tmp));
		double tdot = td.sec();
		alpha = dold / tmp;
		OZ_HIP(hipEventRecord(b->ev[6], st));
		launch_axpy_host(st, n, alpha, v.P, x, oh->reproMode != 0);
		launch_axpy_host(st, n, -alpha, v.Q, v.R, oh->reproMode != 0);
		OZ_HIP(hipEventRecord(b->ev[7], st));
		WallTimer td2;
		OZ_TRY(dot_entry(oh, n, v.R, 1, v.R, 1, &dnew));
		tdot += td2.sec();
		float axpyMs = 0.f;
		(void)hipEventElapsedTime(&axpyMs, b->ev[6], b->ev[7]);
		resi = (T)sqrt((double)dnew);
		beta = dnew / dold;
		dold = dnew;
		OZ_HIP(hipEventRecord(b->ev[6], st));
		launch_scal_axpy(st, n, beta, v.R, v.P, oh->reproMode != 0);
		OZ_HIP(hipEventRecord(b->ev[7], st));
		OZ_HIP(hipEventSynchronize(b->ev[7]));
		float scalMs = 0.f;
		(void)hipEventElapsedTime(&scalMs, b->ev[6], b->ev[7]);
		oh->t_AXPY_SCAL_total += 1e-3 * ((double)axpyMs + scalMs);
		oh->t_DOT_NRM2_total += tdot;
		if (oh->verbose > 0) {
			if (resi / nrmb < tol || oh->cg_numiter % oh->verbose == 0) {
				if (hist) hist[vcnt] = resi / nrmb;
				OZ_TRY(cg_record2(oh, diag, vcnt, (const T *)x));
				vcnt++;
			}
		}
		if (resi / nrmb < tol) break;
	}
	OZ_HIP(hipStreamSynchronize(st));
	return OZBLAS_STATUS_SUCCESS;
}

// Device-resident iterations; *overflow = 1 when the host path must redo them.
// T = float: the <float,float> solve with binary32 vectors and recurrences; p's slices are capped
// at the CSR kernel's register bound (more overflow to the synchronous path).
template <class T>
static int32_t cg_device_loop(ozblasHandle_t *oh, int n, int nA, const double *AS, const int *colind,
                              const int *rowptr, T *x, int maxiter, T tol, T nrmb, T dold,
                              const CgVecs<T> &v, T *hist, int *overflow, const CgDiag<T> &diag) {
	constexpr bool f32 = sizeof(T) == 4;
	Backend *b = be(oh);
	hipStream_t st = b->stream;
	// times of the second history: an event after every iteration whose residual may be recorded
	// (every verbose-th one, and the last enqueued one for a tolerance stop)
	const bool stamps = oh->verbose > 0 && diag.v2 != nullptr && !diag.trueres;
	std::vector<hipEvent_t> evs;
	std::vector<int> evIt;
	hipEvent_t ev0 = nullptr;
	double off = 0.0;
	auto free_events = [&]() {
		for (hipEvent_t e : evs) (void)hipEventDestroy(e);
		if (ev0) (void)hipEventDestroy(ev0);
	};
	const int cap = split_cap(oh);
	const int rho = Prec<T>::rho(oh, n);
	const uint16_t *Js = Prec<T>::Jsplit(b), *Jn = Prec<T>::Jsum(b);
	const int reproCount = repro_slot(oh) ? 0 : 1;
	const int64_t lds = rup(n, 32);
	const int fsum = oh->sumMode == 0 ? 1 : 0;
	const int nhist = (oh->verbose > 0) ? maxiter / oh->verbose + 2 : 1;
	// slices are stored only for a vector whose plan falls outside the residue model
	double *slP = (double *)walloc(oh, 8ull * lds * (cap + 2));
	double *slQ = (double *)walloc(oh, 8ull * lds * (cap + 2));
	double *slR = (double *)walloc(oh, 8ull * lds * (cap + 2));
	const size_t small = 8ull * (cap + 1) * (cap + 1) + 64 + 8ull * nhist;
	char *sm = (char *)walloc(oh, small);
	CgDev *S = (CgDev *)walloc(oh, sizeof(CgDev));
	void *vscr = walloc(oh, vsplit_scratch_bytes());
	VsPlan *plans = (VsPlan *)walloc(oh, 3 * sizeof(VsPlan));  // P, Q, R
	VsPre *pres = (VsPre *)walloc(oh, 6 * sizeof(VsPre));  // a pair per vector (split parity)
	double *part = (double *)walloc(oh, dots_part_bytes());
	if (!vscr || !plans || !pres || !part) return OZBLAS_STATUS_WORKSPACE;
	int maxVec = (oh->cgDeviceSlices > 0 && oh->cgDeviceSlices < cap) ? oh->cgDeviceSlices : cap;
	if (f32 && maxVec > 8) maxVec = 8;  // the binary32 product has no per-pair row loop
	const int64_t ldc = rup(n, 32);
	double *Cs = (double *)walloc(oh, 8ull * ldc * nA * maxVec);  // pair sums of the general product
	if (!slP || !slQ || !slR || !sm || !S || !Cs) return OZBLAS_STATUS_WORKSPACE;
	double *vals = (double *)sm;
	int *info = (int *)(vals + (cap + 1) * (cap + 1));  // P, Q, R: two ints each
	double *dhist = (double *)(info + 16);
	OZ_HIP(hipMemsetAsync(sm, 0, small, st));
	OZ_HIP(hipMemsetAsync(vscr, 0, vsplit_scratch_bytes(), st));
	OZ_HIP(hipMemsetAsync(pres, 0, 6 * sizeof(VsPre), st));
	CgDev h{};
	h.nrmb = nrmb;
	h.dold = dold;
	h.tol = tol;
	h.hist = (oh->verbose > 0) ? dhist : nullptr;
	h.stopAt = INT32_MAX;
	h.vcnt = 1;
	h.verbose = oh->verbose;
	h.maxVecSlices = maxVec;
	h.clkCat = kClkNone;
	OZ_HIP(hipMemcpyAsync(S, &h, sizeof(CgDev), hipMemcpyHostToDevice, st));
	int *infoP = info, *infoQ = info + 2, *infoR = info + 4;
	VsPlan *planP = plans, *planQ = plans + 1, *planR = plans + 2;
	VsPre *preP = pres, *preQ = pres + 2, *preR = pres + 4;
	const VsRef refP{v.P, planP, slP, lds, infoP}, refQ{v.Q, planQ, slQ, lds, infoQ},
	    refR{v.R, planR, slR, lds, infoR};
	if (stamps) {
		OZ_HIP(hipEventCreate(&ev0));
		OZ_HIP(hipEventRecord(ev0, st));
		off = diag.t1.sec();
	}
	// Each vector's split reductions come from the kernel that wrote it (the CSR product for q, the
	// AXPYs for r and p); p of the first iteration is the copy of r0, reduced here.  The first
	// kernel of each category opens its clock: SpMV (split of p, q = A p), DOT (split of q or r and
	// the dot), AXPY/SCAL.
	for (int it = 0; it < maxiter; it++) {
		const int par = it & 1, parNext = (it + 1) & 1;
		OZ_HIP(launch_vsplit_plan(st, n, v.P, slP, lds, rho, 1, cap, reproCount, Js, infoP, vscr, preP, planP,
		                          it > 0, S, it, kClkSpmv, par));
		if constexpr (sizeof(T) == 4)
			launch_csrmm_sum_dev(st, n, AS, rup(b->nnzA, 32), nA, colind, rowptr, refP, fsum, v.Q, Cs, ldc, Jn, S,
			                     it, preQ + par, b->mixed);
		else
			launch_csrmm_sum_dev(st, n, AS, rup(b->nnzA, 32), nA, colind, rowptr, refP, fsum, v.Q, Cs, ldc, Jn, S,
			                     it, preQ + par);
		OZ_HIP(launch_vsplit_plan(st, n, v.Q, slQ, lds, rho, 1, cap, reproCount, Js, infoQ, vscr, preQ, planQ,
		                          true, S, it, kClkDot, par));
		launch_dots_dev(st, n, refP, refQ, vals, part, Jn, fsum, S, it, 1, nullptr, f32 ? 1 : 0);
		launch_axpy2_dev(st, n, S, it, v.P, x, v.Q, v.R, preR + par);
		OZ_HIP(launch_vsplit_plan(st, n, v.R, slR, lds, rho, 1, cap, reproCount, Js, infoR, vscr, preR, planR,
		                          true, S, it, kClkDot, par));
		launch_dots_dev(st, n, refR, refR, vals, part, Jn, fsum, S, it, 2, nullptr, f32 ? 1 : 0);
		launch_scal_axpy_dev(st, n, S, it, v.R, v.P, preP + parNext);
		if (stamps && ((it + 1) % oh->verbose == 0 || it + 1 == maxiter || it % 16 == 15)) {
			hipEvent_t e;
			OZ_HIP(hipEventCreate(&e));
			OZ_HIP(hipEventRecord(e, st));
			evs.push_back(e);
			evIt.push_back(it);
		}
	}
	launch_cg_clock_close(st, S);
	OZ_HIP(hipGetLastError());
	OZ_HIP(hipMemcpyAsync(&h, S, sizeof(CgDev), hipMemcpyDeviceToHost, st));
	OZ_HIP(hipStreamSynchronize(st));
	*overflow = h.overflow;
	if (h.overflow) {
		free_events();
		return OZBLAS_STATUS_SUCCESS;
	}
	{  // category totals (src/cpu/ozblas_cg.cpp:89,99 and the per-call totals of DOT and CSRMV)
		int khz = 0;
		OZ_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, b->device));
		const double tick = khz > 0 ? 1e-3 / khz : 0.0;
		oh->t_SpMV_SpMM_total += tick * (double)h.clkTot[kClkSpmv];
		oh->t_DOT_NRM2_total += tick * (double)h.clkTot[kClkDot];
		oh->t_AXPY_SCAL_total += tick * (double)h.clkTot[kClkAxpy];
	}
	oh->cg_numiter = (h.stopAt < maxiter) ? h.stopAt : maxiter;
	if (oh->verbose > 0 && hist && h.vcnt > 1) {
		// the device history holds T values as binary64
		std::vector<double> hh(h.vcnt - 1);
		OZ_HIP(hipMemcpy(hh.data(), dhist + 1, 8ull * (h.vcnt - 1), hipMemcpyDeviceToHost));
		for (int j = 1; j < h.vcnt; j++) hist[j] = (T)hh[j - 1];
	}
	if (stamps) {
		// entry j of the history is the j-th iteration (1-based count) with count % verbose == 0,
		// or the last one when the tolerance stopped the solve; an iteration past the stop did no
		// work, so the first event at or after it carries its time
		double *t2 = (double *)diag.v2;
		const int last = oh->cg_numiter - 1;
		size_t q = 0;
		int j = 1;
		for (int it = 0; it <= last && j < h.vcnt; it++) {
			if ((it + 1) % oh->verbose != 0 && it != last) continue;
			while (q + 1 < evIt.size() && evIt[q] < it) q++;
			float ms = 0.f;
			(void)hipEventElapsedTime(&ms, ev0, evs[q]);
			t2[j++] = off + 1e-3 * ms;
		}
	}
	free_events();
	return OZBLAS_STATUS_SUCCESS;
}

// binary64 and binary32 iterate device-resident where they can (cg_device_loop), else synchronously
template <class T>
static int32_t cg_entry(ozblasHandle_t *oh, char transA, int32_t n, int32_t nnz, const T *A,
                        const int32_t *colind, const int32_t *rowptr, const T *b_, T *x,
                        int32_t maxiter, T tol) {
	constexpr size_t E = sizeof(T);
	OZ_TRY(ensure_init(oh));
	if (!is_n(transA)) return oh->lastStatus = OZBLAS_STATUS_NOT_SUPPORTED;
	if (oh->sumMode != 0 && oh->sumMode != 1) return oh->lastStatus = OZBLAS_STATUS_NOT_SUPPORTED;
	Backend *b = be(oh);
	hipStream_t st = b->stream;
	CgDiag<T> diag;  // starts the clock of the second history (t1 of src/cpu/ozblas_cg.cpp:58)
	T *hist = (T *)oh->cg_verbose1;  // TYPE1 *ptr_cg_verbose1 (src/cpu/ozblas_cg.cpp:47)
	oh->cg_numiter = 0;
	oh->cgDevicePath = 0;
	oh->t_SplitMat_total = oh->t_SplitVec_total = oh->t_Sum_total = 0.;
	oh->t_AXPY_SCAL_total = oh->t_DOT_NRM2_total = oh->t_SpMV_SpMM_total = 0.;
	CgVecs<T> v;
	OZ_HIP(hipMalloc(&v.R, E * n));
	OZ_HIP(hipMalloc(&v.P, E * n));
	OZ_HIP(hipMalloc(&v.Q, E * n));
	OZ_HIP(hipMalloc(&v.X0, E * n));
	if (oh->verbose > 0 && oh->cg_verbose2) {
		diag.v2 = oh->cg_verbose2;
		diag.trueres = oh->trueresFlag != 0;
		diag.A = A;
		diag.b = b_;
		diag.colind = colind;
		diag.rowptr = rowptr;
		diag.n = n;
		if (diag.trueres) OZ_HIP(hipMalloc(&v.T_, E * n));
		diag.t = v.T_;
	}
	int32_t s = OZBLAS_STATUS_SUCCESS;
	T nrmb = 0, dold = 0, resi = 0;
	int overflow = 1;
	double *AS = nullptr;
	// q starts finite: the first product reads it through beta * q with beta = 0
	if (hipMemsetAsync(v.Q, 0, E * n, st) != hipSuccess ||
	    hipMemcpyAsync(v.X0, x, E * n, hipMemcpyDeviceToDevice, st) != hipSuccess) {
		s = OZBLAS_STATUS_EXECUTION_FAILED;
		goto done;
	}
	{
		WallTimer ts;
		AS = csrmv_split_a(oh, transA, n, nnz, A, rowptr);
		oh->t_SplitMat_total = ts.sec();
	}
	if (!AS) {
		s = oh->lastStatus;
		goto done;
	}
	b->nnzA = nnz;
	{  // ozblasRnrm2 (src/cpu/ozblas_nrm2.cpp:27: the double sqrt, rounded to T)
		T d = 0;
		s = dot_entry(oh, n, b_, 1, b_, 1, &d);
		if (s) goto done;
		nrmb = (T)sqrt((double)d);
	}
	if (hipMemcpyAsync(v.R, b_, E * n, hipMemcpyDeviceToDevice, st) != hipSuccess) {
		s = OZBLAS_STATUS_EXECUTION_FAILED;
		goto done;
	}
	s = csrmv_impl(oh, n, n, nnz, (T)-1, (const void *)AS, colind, rowptr, x, (T)1, v.R);
	if (s) goto done;
	s = dot_entry(oh, n, v.R, 1, v.R, 1, &dold);
	if (s) goto done;
	resi = (T)sqrt((double)dold);
	diag.nrmb = nrmb;
	if (oh->verbose > 0) {
		if (hist) hist[0] = resi / nrmb;
		s = cg_record2(oh, diag, 0, (const T *)x);
		if (s) goto done;
	}
	if (hipMemcpyAsync(v.P, v.R, E * n, hipMemcpyDeviceToDevice, st) != hipSuccess) {
		s = OZBLAS_STATUS_EXECUTION_FAILED;
		goto done;
	}
	{
		// the true residual of a recorded iteration needs its x: that history runs synchronously
		bool devicePath = oh->cgDeviceSlices >= 0 && !repro_slot(oh) &&
		                  oh->fastMode == 0 && oh->splitEpsMode != 2 && oh->nSplitA_ <= 4 && maxiter > 0 &&
		                  !diag.trueres;
		if (devicePath) {
			wreset(oh, b->keep);
			s = cg_device_loop<T>(oh, n, oh->nSplitA_, AS, colind, rowptr, x, maxiter, tol, nrmb, dold, v, hist,
			                      &overflow, diag);
			wreset(oh, b->keep);
			if (s == OZBLAS_STATUS_WORKSPACE) {  // nothing ran: iterate on the synchronous path
				s = OZBLAS_STATUS_SUCCESS;
				overflow = 1;
				devicePath = false;
			}
			if (s) goto done;
		}
		oh->cgDevicePath = !devicePath ? 0 : (overflow ? 2 : 1);
		if (overflow) {  // synchronous path (from the saved set-up state when the device path gave up)
			if (devicePath) {
				if (hipMemcpyAsync(x, v.X0, E * n, hipMemcpyDeviceToDevice, st) != hipSuccess ||
				    hipMemcpyAsync(v.R, b_, E * n, hipMemcpyDeviceToDevice, st) != hipSuccess) {
					s = OZBLAS_STATUS_EXECUTION_FAILED;
					goto done;
				}
				s = csrmv_impl(oh, n, n, nnz, (T)-1, (const void *)AS, colind, rowptr, x, (T)1, v.R);
				if (s) goto done;
				if (hipMemcpyAsync(v.P, v.R, E * n, hipMemcpyDeviceToDevice, st) != hipSuccess) {
					s = OZBLAS_STATUS_EXECUTION_FAILED;
					goto done;
				}
			}
			oh->cg_numiter = 0;
			s = cg_host_loop(oh, n, nnz, AS, colind, rowptr, x, maxiter, tol, nrmb, dold, v, hist, diag);
		}
	}
done:
	(void)hipStreamSynchronize(st);
	(void)hipFree(v.R);
	(void)hipFree(v.P);
	(void)hipFree(v.Q);
	(void)hipFree(v.X0);
	(void)hipFree(v.T_);
	oh->memMaskSplitA = 0;
	b->keep = 0;
	oh->splitShift = 1;
	wreset(oh, 0);
	return oh->lastStatus = s;
}

}  // namespace oz

extern "C" int32_t ozblasDcg(ozblasHandle_t *oh, char transA, int32_t n, int32_t nnz,
                             const void *descrA, const double *A, const int32_t *colind,
                             const int32_t *rowptr, const double *b, double *x, int32_t maxiter,
                             double tol) {
	(void)descrA;
	return cg_entry(oh, transA, n, nnz, A, colind, rowptr, b, x, maxiter, tol);
}

extern "C" int32_t ozblasScg(ozblasHandle_t *oh, char transA, int32_t n, int32_t nnz,
                             const void *descrA, const float *A, const int32_t *colind,
                             const int32_t *rowptr, const float *b, float *x, int32_t maxiter,
                             float tol) {
	(void)descrA;
	return cg_entry(oh, transA, n, nnz, A, colind, rowptr, b, x, maxiter, tol);
}

// =============================================================================================
// mixed routines: binary64 data, binary32 slices (the reference's <double,float,float> and
// <double,float> instantiations).  Fsum only: the reference's mixed NearSum and its float Dot2 GEMM
// are not available.
// =============================================================================================
namespace oz {
struct MixedScope {  // marks the handle's backend as running a mixed routine for one call
	Backend *b;
	int save;
	explicit MixedScope(ozblasHandle_t *oh) : b(be(oh)), save(b->mixed) { b->mixed = 1; }
	~MixedScope() { b->mixed = save; }
};
}  // namespace oz

extern "C" int32_t ozblasDSgemm(ozblasHandle_t *oh, char transA, char transB, int32_t m, int32_t n,
                                int32_t k, double alpha, const double *A, int32_t lda,
                                const double *B, int32_t ldb, double beta, double *C, int32_t ldc) {
	OZ_TRY(ensure_init(oh));
	if (m < 0 || n < 0 || k < 0) return oh->lastStatus = OZBLAS_STATUS_INVALID_VALUE;
	if (m == 0 || n == 0) return oh->lastStatus = OZBLAS_STATUS_SUCCESS;
	oh->splitShift = 1;  // counterInit
	if (oh->reproMode == 0 && oh->nSplitMax == 1) return oh->lastStatus = OZBLAS_STATUS_NOT_SUPPORTED;
	if (oh->splitEpsMode == 2 || oh->sumMode != 0) return oh->lastStatus = OZBLAS_STATUS_NOT_SUPPORTED;
	if (k == 0) return oh->lastStatus = gemm_k0(oh, m, n, alpha, beta, C, ldc);
	MixedScope mx(oh);
	return oh->lastStatus = gemm_plain(oh, transA, transB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

extern "C" int32_t ozblasDSgemv(ozblasHandle_t *oh, char transA, int32_t m, int32_t n, double alpha,
                                const double *A, int32_t lda, const double *x, int32_t incx,
                                double beta, double *y, int32_t incy) {
	OZ_TRY(ensure_init(oh));
	if (incx != 1 || incy != 1) return oh->lastStatus = OZBLAS_STATUS_NOT_SUPPORTED;
	if (is_n(transA))  // src/cpu/ozblas_gemv.cpp:23-26
		return ozblasDSgemm(oh, 'n', 'n', m, 1, n, alpha, A, lda, x, n, beta, y, m);
	return ozblasDSgemm(oh, 't', 'n', n, 1, m, alpha, A, lda, x, m, beta, y, n);
}

extern "C" int32_t ozblasDSdot(ozblasHandle_t *oh, int32_t n, const double *x, int32_t incx,
                               const double *y, int32_t incy, double *ret) {
	OZ_TRY(ensure_init(oh));
	MixedScope mx(oh);
	return dot_entry(oh, n, x, incx, y, incy, ret);
}

extern "C" int32_t ozblasDSnrm2(ozblasHandle_t *oh, int32_t n, const double *x, int32_t incx,
                                double *ret) {
	OZ_TRY(ensure_init(oh));
	double d = 0.0;
	const int32_t s = ozblasDSdot(oh, n, x, incx, x, incx, &d);
	if (s != OZBLAS_STATUS_SUCCESS) return s;
	*ret = sqrt(d);
	return OZBLAS_STATUS_SUCCESS;
}

extern "C" float *ozblasDScsrmvSplitA(ozblasHandle_t *oh, char transA, int32_t m, int32_t n,
                                      int32_t nnz, const void *descrA, const double *A,
                                      const int32_t *rowptr) {
	(void)descrA;
	(void)n;
	if (ensure_init(oh) != OZBLAS_STATUS_SUCCESS) return nullptr;
	MixedScope mx(oh);
	return (float *)csrmv_split_a(oh, transA, m, nnz, A, rowptr);
}

extern "C" int32_t ozblasDScsrmv(ozblasHandle_t *oh, char transA, int32_t m, int32_t n, int32_t nnz,
                                 double alpha, const void *descrA, const double *A,
                                 const int32_t *colind, const int32_t *rowptr, const double *x,
                                 double beta, double *y) {
	(void)descrA;
	OZ_TRY(ensure_init(oh));
	MixedScope mx(oh);
	return csrmv_entry(oh, transA, m, n, nnz, alpha, A, colind, rowptr, x, beta, y);
}

extern "C" int32_t ozblasDScg(ozblasHandle_t *oh, char transA, int32_t n, int32_t nnz,
                              const void *descrA, const double *A, const int32_t *colind,
                              const int32_t *rowptr, const double *b, double *x, int32_t maxiter,
                              double tol) {
	(void)descrA;
	OZ_TRY(ensure_init(oh));
	MixedScope mx(oh);
	return cg_entry(oh, transA, n, nnz, A, colind, rowptr, b, x, maxiter, tol);
}

// =============================================================================================
// mixed routines: binary32 data, binary64-held slices (the reference's <float,double,double> and
// <float,double> instantiations).  The split runs in binary32 with getRho<float,double>, the slice
// products in binary64, and the Fsum in binary32 over the products rounded as the reference's
// (float) of the scaled product (f32_term).  Fsum only: the reference's mixed NearSum exits; Dot2
// (splitEpsMode 2) is not provided.
// =============================================================================================
extern "C" int32_t ozblasSDgemm(ozblasHandle_t *oh, char transA, char transB, int32_t m, int32_t n,
                                int32_t k, float alpha, const float *A, int32_t lda, const float *B,
                                int32_t ldb, float beta, float *C, int32_t ldc) {
	OZ_TRY(ensure_init(oh));
	if (m < 0 || n < 0 || k < 0) return oh->lastStatus = OZBLAS_STATUS_INVALID_VALUE;
	if (m == 0 || n == 0) return oh->lastStatus = OZBLAS_STATUS_SUCCESS;
	oh->splitShift = 1;  // counterInit
	if (oh->reproMode == 0 && oh->nSplitMax == 1) return oh->lastStatus = OZBLAS_STATUS_NOT_SUPPORTED;
	if (oh->splitEpsMode == 2 || oh->sumMode != 0) return oh->lastStatus = OZBLAS_STATUS_NOT_SUPPORTED;
	if (k == 0) return oh->lastStatus = gemm_k0(oh, m, n, alpha, beta, C, ldc);
	MixedScope mx(oh);
	return oh->lastStatus = gemm_plain(oh, transA, transB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

extern "C" int32_t ozblasSDgemv(ozblasHandle_t *oh, char transA, int32_t m, int32_t n, float alpha,
                                const float *A, int32_t lda, const float *x, int32_t incx, float beta,
                                float *y, int32_t incy) {
	OZ_TRY(ensure_init(oh));
	if (incx != 1 || incy != 1) return oh->lastStatus = OZBLAS_STATUS_NOT_SUPPORTED;
	if (is_n(transA))  // src/cpu/ozblas_gemv.cpp:23-26
		return ozblasSDgemm(oh, 'n', 'n', m, 1, n, alpha, A, lda, x, n, beta, y, m);
	return ozblasSDgemm(oh, 't', 'n', n, 1, m, alpha, A, lda, x, m, beta, y, n);
}

extern "C" int32_t ozblasSDdot(ozblasHandle_t *oh, int32_t n, const float *x, int32_t incx,
                               const float *y, int32_t incy, float *ret) {
	OZ_TRY(ensure_init(oh));
	MixedScope mx(oh);
	return dot_entry(oh, n, x, incx, y, incy, ret);
}

extern "C" int32_t ozblasSDnrm2(ozblasHandle_t *oh, int32_t n, const float *x, int32_t incx,
                                float *ret) {
	OZ_TRY(ensure_init(oh));
	float d = 0.f;
	const int32_t s = ozblasSDdot(oh, n, x, incx, x, incx, &d);
	if (s != OZBLAS_STATUS_SUCCESS) return s;
	*ret = (float)sqrt((double)d);  // the double sqrt, rounded (src/cpu/ozblas_nrm2.cpp:27)
	return OZBLAS_STATUS_SUCCESS;
}

extern "C" double *ozblasSDcsrmvSplitA(ozblasHandle_t *oh, char transA, int32_t m, int32_t n,
                                       int32_t nnz, const void *descrA, const float *A,
                                       const int32_t *rowptr) {
	(void)descrA;
	(void)n;
	if (ensure_init(oh) != OZBLAS_STATUS_SUCCESS) return nullptr;
	MixedScope mx(oh);
	return csrmv_split_a(oh, transA, m, nnz, A, rowptr);
}

extern "C" int32_t ozblasSDcsrmv(ozblasHandle_t *oh, char transA, int32_t m, int32_t n, int32_t nnz,
                                 float alpha, const void *descrA, const float *A,
                                 const int32_t *colind, const int32_t *rowptr, const float *x,
                                 float beta, float *y) {
	(void)descrA;
	OZ_TRY(ensure_init(oh));
	MixedScope mx(oh);
	return csrmv_entry(oh, transA, m, n, nnz, alpha, A, colind, rowptr, x, beta, y);
}

extern "C" int32_t ozblasSDcg(ozblasHandle_t *oh, char transA, int32_t n, int32_t nnz,
                              const void *descrA, const float *A, const int32_t *colind,
                              const int32_t *rowptr, const float *b, float *x, int32_t maxiter,
                              float tol) {
	(void)descrA;
	OZ_TRY(ensure_init(oh));
	MixedScope mx(oh);
	return cg_entry(oh, transA, n, nnz, A, colind, rowptr, b, x, maxiter, tol);
}
