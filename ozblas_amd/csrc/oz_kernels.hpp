// oz_kernels.hpp -- kernel launchers shared between the translation units (private header).
#pragma once

#include "oz_internal.hpp"

namespace oz {

// ---- split (oz_split.hip)
// sstride = 0: the counting pass (zcol / zmax only); otherwise slots [0, fill) of every column are
// valid afterwards (slices, the remainder slot, zeros), fill being the operand's slice count
void launch_split_cols(hipStream_t st, int m, int n, const double *in, int64_t ldi, double *out,
                       int64_t ldo, double *split, int64_t lds, int64_t sstride, int rho,
                       int shift, int cap, const uint16_t *J, int *zcol, int *zmax, int fill = 0);
void launch_split_cols(hipStream_t st, int m, int n, const float *in, int64_t ldi, float *out,
                       int64_t ldo, double *split, int64_t lds, int64_t sstride, int rho,
                       int shift, int cap, const uint16_t *JF, int *zcol, int *zmax, int fill = 0);
void launch_zero_slots(hipStream_t st, int m, int n, double *split, int64_t lds, int64_t sstride,
                       const int *zcol, int count);
void launch_transpose(hipStream_t st, int rows, int cols, const double *in, int64_t ldi,
                      double *out, int64_t ldo);
void launch_transpose(hipStream_t st, int rows, int cols, const float *in, int64_t ldi, float *out,
                      int64_t ldo);
void launch_vec_max(hipStream_t st, int m, const double *x, unsigned long long *slot);
void launch_split_vec_pass(hipStream_t st, int m, const double *src, double *rem, double *slice,
                           int rho, int shift, const uint16_t *J, const unsigned long long *maxIn,
                           unsigned long long *maxOut);
void launch_fill(hipStream_t st, double *p, int64_t n, double v);
void launch_fill(hipStream_t st, float *p, int64_t n, float v);
void launch_copy(hipStream_t st, const double *src, double *dst, int64_t n);
// mixed = 1: the rho of the mixed split (binary64 data into binary32 slices, or binary32 data
// into binary64 slices)
void launch_split_csr(hipStream_t st, int m, const double *val, const int *rowptr, double *work,
                      double *split, int64_t lds, int shift, int cap, const uint16_t *J, int *zrow,
                      int *zmax, int epsMode, int mixed = 0);
void launch_split_csr(hipStream_t st, int m, const float *val, const int *rowptr, float *work,
                      double *split, int64_t lds, int shift, int cap, const uint16_t *JF, int *zrow,
                      int *zmax, int epsMode, int mixed = 0);
void launch_zero_slots_csr(hipStream_t st, int m, const int *rowptr, double *split, int64_t lds,
                           const int *zrow, int count);
void launch_digit_stats(hipStream_t st, int m, int n, const double *in, int64_t ldi, int *colE,
                        int *colS, void *stats, int *colLo = nullptr);
void launch_digit_write(hipStream_t st, int elem, int m, int n, const double *in, int64_t ldi,
                        const int *colE, int S, void *planes, int64_t ldp, int64_t pstride);

// ---- FP64 slice products (oz_gemm_f64.hip)
// C_b (M x N, ldc) = A_b^T B_b with A_b (K x M, lda), B_b (K x N, ldb) for batch entries b:
// A_b = A + aOff[b], B_b = B + bOff[b], C_b = C + cOff[b] (element offsets, device arrays)
void launch_gemm_f64_batched(hipStream_t st, int M, int N, int K, const double *A, int64_t lda,
                             const double *B, int64_t ldb, double *C, int64_t ldc,
                             const int64_t *aOff, const int64_t *bOff, const int64_t *cOff,
                             int batch);
// fused Fsum: t = sum over pairs p (in order) of A_p^T B_p; C = fma(alpha, t, beta*C)
void launch_gemm_f64_fsum(hipStream_t st, int M, int N, int K, const double *A, int64_t lda,
                          const double *B, int64_t ldb, const int64_t *aOff, const int64_t *bOff,
                          int npairs, double alpha, double beta, double *C, int64_t ldc);
// binary32 outputs: products rounded once to float; the Fsum adds in float, epilogue fma1<float>
void launch_gemm_f64_batched(hipStream_t st, int M, int N, int K, const double *A, int64_t lda,
                             const double *B, int64_t ldb, float *C, int64_t ldc,
                             const int64_t *aOff, const int64_t *bOff, const int64_t *cOff,
                             int batch);
void launch_gemm_f64_fsum(hipStream_t st, int M, int N, int K, const double *A, int64_t lda,
                          const double *B, int64_t ldb, const int64_t *aOff, const int64_t *bOff,
                          int npairs, float alpha, float beta, float *C, int64_t ldc);

// ---- sums (oz_sum.hip)
void launch_global_nearsum(hipStream_t st, int m, int n, int ic, double *Cs, int64_t llsc,
                           int64_t ldsc, double *C, int64_t ldc, double alpha, double beta,
                           const uint16_t *J);
void launch_global_fsum(hipStream_t st, int m, int n, int nA, int nB, int maxlevel, int sumOrder,
                        const double *Cs, int64_t llsc, int64_t ldsc, double *C, int64_t ldc,
                        double alpha, double beta);
// binary32: NearSum<float> (JF2: the log2f table) / Fsum in float, epilogue fma1<float>
void launch_global_nearsum(hipStream_t st, int m, int n, int ic, float *Cs, int64_t llsc,
                           int64_t ldsc, float *C, int64_t ldc, float alpha, float beta,
                           const uint16_t *JF2);
void launch_global_fsum(hipStream_t st, int m, int n, int nA, int nB, int maxlevel, int sumOrder,
                        const float *Cs, int64_t llsc, int64_t ldsc, float *C, int64_t ldc,
                        float alpha, float beta);

// ---- fast forms (oz_gemm_digits.hip)
struct LevelBin {      // one exact int32 plane: sum over `npairs` digit pairs, k in [k0, k1)
	int level;
	int npairs;
	int pair0;         // first entry in the pair list
	int k0, k1;
};
// Residue form (oz_residue.hip): the moduli and the constants of the Chinese-remainder
// reconstruction for a product of nmod moduli M = p_0 ... p_{nmod-1}.
constexpr int kResMaxMod = 49;   // 256 and the 48 odd pairwise coprime moduli up to 255
constexpr int kResMaxLimb = 18;  // 32-bit limbs of M (<= 392 bits) plus headroom
constexpr int kResMaxShift = 512;  // element shifts below 2^512 (256^a mod p for a < 64)
constexpr int kResMaxGroups = (kResMaxMod + 2) / 3;
constexpr int kResChunkBits = 39;  // binary64 chunks of the reconstruction: 49 x 255 x 2^39 < 2^53
constexpr int kResMaxChunk = (64 * 9 + kResChunkBits - 1) / kResChunkBits;  // up to 9 64-bit words
constexpr int kResChunk = 65536;   // k per exact int32 product chunk (|acc| <= 2^16 2^14 = 2^30)
// product residues: plane c * nmod + t holds modulus t over k chunk c
struct ResConst {
	int nmod;          // moduli in use
	int nl;            // 32-bit limbs carried by the reconstruction
	int p[kResMaxMod];
	float invpf[kResMaxMod];   // 1 / p rounded to binary32
	double invpd[kResMaxMod];  // 1 / p rounded to binary64
	int u[kResMaxMod];         // (M / p)^-1 mod p
	uint32_t Mp[kResMaxMod][kResMaxLimb];  // M / p, little-endian limbs
	uint32_t M[kResMaxLimb];
	uint8_t pow256[kResMaxMod][kResMaxShift / 8];  // 256^a mod p
	// moduli in groups of three (3g, 3g + 1, 3g + 2): P_g = their product < 2^24, and 2^sh mod P_g
	double Pg[kResMaxGroups];
	double invPg[kResMaxGroups];  // 1 / P_g rounded to binary64
	uint32_t pow2g[kResMaxGroups][kResMaxShift];
	// (2^sh U_g) mod P_g with U_g = u_t mod p_t for the group's moduli t < nmod (1 beyond): the
	// residues of op(A) written with it come out multiplied by u_t, so the product modulo p_t is
	// already c_t = X u_t mod p_t
	uint32_t pow2gU[kResMaxGroups][kResMaxShift];
	// M / p and M in kResChunkBits-bit chunks (chunk c holds bits [39 c, 39 c + 39)), as binary64
	double Mpd[kResMaxMod][kResMaxChunk];
	double Md[kResMaxChunk];
};
// the moduli list and per-count constants (host, built once); log2 of the product of the first n
const ResConst &res_const_host(int nmod);
double res_log2_product(int nmod);

void launch_digit_gemm(hipStream_t st, int elem, int M, int N, int Kpad, const void *Ap,
                       int64_t ldap, int64_t apstride, const void *Bp, int64_t ldbp,
                       int64_t bpstride, const int2 *pairs, const LevelBin *bins, int nbins,
                       const int *binOrder, int32_t *T, int64_t ldt, int64_t tstride,
                       const int *tileList = nullptr, int ntiles = 0, const ResConst *rc = nullptr,
                       uint8_t *R = nullptr);
// residue planes of one operand (k-contiguous columns; INT8 tiled layout for elem 1, else
// 2-byte k-contiguous planes), element x of column j as the integer x 2^-colLo[j] modulo each p,
// times u_t = (M/p_t)^-1 mod p_t with scaleU (op(A): the products are then the c_t of the
// reconstruction)
void launch_res_write(hipStream_t st, int elem, int m, int n, const double *in, int64_t ldi,
                      const int *colLo, const ResConst *rc, int nmod, int shLimit, int scaleU, void *planes,
                      int64_t ldp, int64_t pstride);
// C = fma(alpha, RN(X 2^(rowLo[i] + colLo[j])), beta C), X from the residue planes (nchunk k-chunk
// planes per modulus, summed modulo p)
void launch_res_recon(hipStream_t st, int M, int N, const uint8_t *R, int64_t ldr, int64_t rstride,
                      int nchunk, const ResConst *rc, int nmod, int nl, const int *rowLo,
                      const int *colLo, double alpha, double beta, double *C, int64_t ldc);
int digit_gemm_tile(int elem, int M, int N);
int recon_words(int Lmax);
void launch_reconstruct_check(hipStream_t st, int M, int N, const int32_t *T, int64_t ldt,
                              int64_t tstride, const int *binLevel, int nb1, int Lmax,
                              const int *rowE, const int *colE, double alpha, double beta, double *C,
                              int64_t ldc, const uint64_t *R, int tile, int tilesM, int *tileMark);
void launch_reconstruct_fix(hipStream_t st, int M, int N, const int32_t *T, int64_t ldt,
                            int64_t tstride, const int *binLevel, int nb1, int nbins, int Lmax,
                            const int *rowE, const int *colE, double alpha, double beta, double *C,
                            int64_t ldc, const uint64_t *R, int tile, int tilesM, const int *tileList,
                            int ntiles);
void launch_reconstruct(hipStream_t st, int M, int N, const int32_t *T, int64_t ldt,
                        int64_t tstride, const int *binLevel, int nbins, int Lmax, const int *rowE,
                        const int *colE, double alpha, double beta, double *C, int64_t ldc);

// ---- vectors / sparse (oz_level12.hip)
void launch_slice_dots(hipStream_t st, int n, const double *xs, int64_t ldx, int nA,
                       const double *ys, int64_t ldy, int nB, double *out);
void launch_gemm_k0(hipStream_t st, int m, int n, double alpha, double beta, double *C, int64_t ldc);
void launch_gemm_k0(hipStream_t st, int m, int n, float alpha, float beta, float *C, int64_t ldc);
size_t dot2_partials_bytes(int nA, int nB);
void launch_dot2_slices(hipStream_t st, int n, const double *xs, int64_t ldx, int nA,
                        const double *ys, int64_t ldy, int nB, void *part, double *vals);
void launch_csrmm_sum_x2(hipStream_t st, int m, const double *As, int64_t ldas, int nA,
                         const int *colind, const int *rowptr, const double *Xs, int64_t ldxs, int nB,
                         int fsum, double alpha, double beta, double *y, const uint16_t *J);
void launch_nearsum_small(hipStream_t st, int nvals, double *vals, double alpha, double beta,
                          double *C, int sumOrderFsum, int nA, int nB, int maxlevel, int fsum,
                          const uint16_t *J);
void launch_csrmm_sum(hipStream_t st, int m, int nnz, const double *As, int64_t ldas, int nA,
                      const int *colind, const int *rowptr, const double *Xs, int64_t ldxs, int nB,
                      int maxlevel, int fsum, double alpha, double beta, double *y,
                      const uint16_t *J);
void launch_csrmm_store(hipStream_t st, int m, const double *As, int64_t ldas, int nA,
                        const int *colind, const int *rowptr, const double *Xs, int64_t ldxs,
                        int nB, int maxlevel, double *Cs, int64_t ldcs);
// binary32 (slices stored as doubles; sums, y and Cs in float; JF2 for NearSum<float>); mixed = 1:
// the pair products are binary64 chains (the <float,double> instantiation), not binary32 ones
void launch_csrmm_sum(hipStream_t st, int m, int nnz, const double *As, int64_t ldas, int nA,
                      const int *colind, const int *rowptr, const double *Xs, int64_t ldxs, int nB,
                      int maxlevel, int fsum, float alpha, float beta, float *y,
                      const uint16_t *JF2, int mixed);
void launch_csrmm_store(hipStream_t st, int m, const double *As, int64_t ldas, int nA,
                        const int *colind, const int *rowptr, const double *Xs, int64_t ldxs,
                        int nB, int maxlevel, float *Cs, int64_t ldcs, int mixed);
// GEMV of the plain form on k-contiguous slices (slice s of A column i at As + s*sA + i*ldk, of x
// at Xs + s*sX), fused with the sum (Fsum in level order or NearSum) and the epilogue into y;
// false when the slice counts exceed its register bound (8 x 8)
// GEMV with the split of op(A) fused (binary64; rows of op(A) k-contiguous in `in`): false when the
// sizes or slice counts are outside the kernel's (then nothing was launched; split_gemv_fits tells)
// P: split_gemv_pairs_bytes(M) of scratch for the pair values of every row
bool split_gemv_fits(int K, int nA, int nB);
size_t split_gemv_pairs_bytes(int M);
bool launch_split_gemv(hipStream_t st, int K, int M, const double *in, int64_t ldi, int rho, int shift,
                       int cap, const uint16_t *Js, int nA, const double *Xs, int64_t sX, int nB,
                       int fsum, double alpha, double beta, double *y, const uint16_t *J, double *P);
bool launch_gemv_slices(hipStream_t st, int M, int K, const double *As, int64_t ldk, int64_t sA, int nA,
                        const double *Xs, int64_t sX, int nB, int fsum, double alpha, double beta,
                        double *y, const uint16_t *J);
bool launch_gemv_slices(hipStream_t st, int M, int K, const double *As, int64_t ldk, int64_t sA, int nA,
                        const double *Xs, int64_t sX, int nB, int fsum, float alpha, float beta,
                        float *y, const uint16_t *JF2);
// barrier-free vector split (plan of levels, then slices); VsPlan lives in device memory
constexpr int kVsMaxLevels = 64;
struct VsPlan {
	int count, last, zero, bad;
	int needScan;  // the predicted chain did not hold: the table scan and its walk make the plan
	int pad[3];
	int k[kVsMaxLevels];  // log2 of the slice unit of level s
};
// Per-vector reductions of the predicted level chain (zero between uses: the kernel that consumes
// them clears them).  A kernel that produces the vector may accumulate maxBits / loOff / bad itself
// (vs_stats_* in oz_level12.hip), so that the split does not read the vector for them again.
constexpr int kVsPred = 12;     // predicted levels evaluated at most
constexpr int kVsLoOff = 4096;  // min lo is kept as the maximum of kVsLoOff - lo
constexpr int kVsSpread = 16;   // workgroup b reduces into slot b % kVsSpread (one cache line each)
struct alignas(128) VsSlot {
	unsigned long long maxBits;  // max |x| (bits of a non-negative binary64)
	int loOff;                   // max over non-zero x of kVsLoOff - lo (lo: the lowest unit exponent
	int bad;                     //   with a non-zero residue); bad: an input outside the model
	unsigned long long F[kVsPred];
};
struct VsPre {
	VsSlot slot[kVsSpread];
};
size_t vsplit_scratch_bytes();
// Splits x into slices at stride lds (cap + 2 slots) and its plan; the count goes to info[0].
hipError_t launch_vsplit(hipStream_t st, int n, const double *x, double *slices, int64_t lds,
                         int rho, int shift, int cap, int reproMode, const uint16_t *J, int *info,
                         void *scratch, VsPlan *plan);
// binary32 input (JF: the binary32 tau table); slices are stored as doubles
hipError_t launch_vsplit(hipStream_t st, int n, const float *x, double *slices, int64_t lds,
                         int rho, int shift, int cap, int reproMode, const uint16_t *JF, int *info,
                         void *scratch, VsPlan *plan);
// The plan only: consumers that take (x, plan) generate the slices themselves; the slices are
// stored only for a plan marked bad (the reference loop in one workgroup).  pre: a pair of VsPre
// for the vector (zero at first); this split reads pre[parity] (holding the producer's reductions
// when statsDone) and clears pre[parity ^ 1] for the split after it, so consecutive splits of one
// vector alternate the parity.  In a device CG, S gives the iteration bound (its kernels of
// iteration it >= S->stopAt do nothing) and the category clock that the split's first kernel opens
// (clkCat).
struct CgDev;
hipError_t launch_vsplit_plan(hipStream_t st, int n, const double *x, double *slices, int64_t lds,
                              int rho, int shift, int cap, int reproMode, const uint16_t *J, int *info,
                              void *scratch, VsPre *pre, VsPlan *plan, bool statsDone,
                              CgDev *S = nullptr, int it = 0, int clkCat = -1, int parity = 0);
hipError_t launch_vsplit_plan(hipStream_t st, int n, const float *x, double *slices, int64_t lds,
                              int rho, int shift, int cap, int reproMode, const uint16_t *JF, int *info,
                              void *scratch, VsPre *pre, VsPlan *plan, bool statsDone,
                              CgDev *S = nullptr, int it = 0, int clkCat = -1, int parity = 0);
// the plans of two vectors of one length (the DOT's x and y; each with its own slices, info,
// zeroed scratch and plan), by the same three launches
hipError_t launch_vsplit_plan2(hipStream_t st, int n, const double *x, const double *y, int rho, int shift,
                               int cap, int reproMode, const uint16_t *J, double *const slices[2], int64_t lds,
                               int *const info[2], void *const scratch[2], VsPlan *const plan[2]);
hipError_t launch_vsplit_plan2(hipStream_t st, int n, const float *x, const float *y, int rho, int shift,
                               int cap, int reproMode, const uint16_t *JF, double *const slices[2], int64_t lds,
                               int *const info[2], void *const scratch[2], VsPlan *const plan[2]);
void launch_axpy_host(hipStream_t st, int n, double alpha, const double *x, double *y, int fused);
void launch_axpy_host(hipStream_t st, int n, float alpha, const float *x, float *y, int fused);
// t = b - A x by the reference's plain CSRMV (the CG's true residual)
void launch_true_residual(hipStream_t st, int n, const double *val, const int *colind,
                          const int *rowptr, const double *x, const double *b, double *t);
void launch_true_residual(hipStream_t st, int n, const float *val, const int *colind,
                          const int *rowptr, const float *x, const float *b, float *t);
void launch_scal_axpy(hipStream_t st, int n, double beta, const double *r, double *p, int fused);
void launch_scal_axpy(hipStream_t st, int n, float beta, const float *r, float *p, int fused);

// ---- device-resident CG (oz_level12.hip)
// Scalars of ozblasRcg kept on the device between the kernels of one solve; a kernel of
// iteration `it` does nothing once it >= stopAt.
struct CgDev {
	double nrmb, dold, dnew, tmp, alpha, beta, resi, tol;
	double *hist;    // device history (resi / nrmb), index vcnt
	int stopAt;      // first iteration not to run (INT32_MAX while iterating)
	int overflow;    // a vector had more slices than the CSR product kernel holds
	int numiter;
	int vcnt;
	int verbose;
	int maxVecSlices;  // the CSR product gives up (overflow) when p has more slices than this
	int clkCat;        // category clock: the category open since clkLast (-1: none)
	int pad;
	unsigned long long clkLast;
	unsigned long long clkTot[3];  // wall-clock ticks of SpMV, DOT and AXPY/SCAL
};
enum CgClock { kClkSpmv = 0, kClkDot = 1, kClkAxpy = 2, kClkNone = -1 };
// closes the category clock after the last iteration (one thread)
void launch_cg_clock_close(hipStream_t st, CgDev *S);

// A split vector as its consumers read it: the vector and its plan, from which they generate the
// slices, or the stored slices when the plan is marked bad.
struct VsRef {
	const void *x;  // binary64 (or binary32 where the consumer says so)
	const VsPlan *plan;
	const double *slices;
	int64_t lds;
	const int *info;  // info[0]: the slice count
};
// q = A p on the device path: registers for p with <= 8 slices, a per-pair row loop for more (pair
// sums through Cs, m x nA x maxVecSlices doubles at pitch ldc); the reductions of the predicted
// split chain of q accumulate into preY
void launch_csrmm_sum_dev(hipStream_t st, int m, const double *As, int64_t ldas, int nA,
                          const int *colind, const int *rowptr, VsRef X, int fsum, double *y,
                          double *Cs, int64_t ldc, const uint16_t *J, CgDev *S, int it, VsPre *preY);
// binary32 p and q (JF2: the binary32 NearSum table); p's slices capped at 8 by maxVecSlices;
// mixed = 1: binary64 pair chains (<float,double>)
void launch_csrmm_sum_dev(hipStream_t st, int m, const double *As, int64_t ldas, int nA,
                          const int *colind, const int *rowptr, VsRef X, int fsum, float *y,
                          double *Cs, int64_t ldc, const uint16_t *JF2, CgDev *S, int it, VsPre *preY,
                          int mixed = 0);
// part: dots_part_bytes() of scratch for the per-workgroup pair partials (no initial value)
size_t dots_part_bytes();
// DOT of two split vectors (binary32 inputs with f32), completed by `mode` (0: C[0] = the DOT)
void launch_dots_dev(hipStream_t st, int n, VsRef X, VsRef Y, double *vals, double *part,
                     const uint16_t *J, int fsum, CgDev *S, int it, int mode, void *C, int f32 = 0);
// x += alpha p, r -= alpha q (reductions of r's split into preR); p = r + beta p (into preP)
void launch_axpy2_dev(hipStream_t st, int n, CgDev *S, int it, const double *P, double *X,
                      const double *Q, double *R, VsPre *preR);
void launch_scal_axpy_dev(hipStream_t st, int n, CgDev *S, int it, const double *R, double *P,
                          VsPre *preP);
void launch_axpy2_dev(hipStream_t st, int n, CgDev *S, int it, const float *P, float *X, const float *Q,
                      float *R, VsPre *preR);
void launch_scal_axpy_dev(hipStream_t st, int n, CgDev *S, int it, const float *R, float *P, VsPre *preP);

}  // namespace oz
