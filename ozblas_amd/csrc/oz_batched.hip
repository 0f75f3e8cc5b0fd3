// oz_batched.hip -- grouped batched Ozaki DGEMM in the INT8 digit form (ozblasDgemmBatched).
//
// No reference counterpart: the reference batches the slice pairs of ONE GEMM through
// cublasDgemmBatched (src/cuda/cuozblas_gemm.cu:150-190).  Here a whole batch of uneven GEMMs
// runs as one plan:
//   plan (once per batch signature, one host synchronisation):
//     grouped transposes + digit statistics -> digit count S of every operand and the fast-form
//     validity of every problem; level bins of every problem; work items (problem, bin, tile)
//     for both tile shapes; the autotuner times one grouped MFMA launch per tile shape on the
//     real operands and keeps the faster (exact integer products: the tile shape cannot change
//     a bit of the result).
//   execute (captured once into a hipGraph, replayed by every later call with the same
//     signature): reset -> grouped transposes -> grouped digit statistics -> plan check ->
//     grouped digit planes -> ONE persistent grouped MFMA launch over all (problem, bin, tile)
//     items, pulled from a work counter -> grouped exact reconstruction with the alpha/beta
//     epilogue.
// Every C[p] is bitwise equal to ozblasDgemm on problem p alone (both are RN of the exact
// product).  Problems whose operands fail the fast-form precondition, and whole batches that do
// not use the INT8 form with NearSum, go through ozblasDgemm one by one.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <vector>

#include "oz_residue.hpp"
#include "oz_internal.hpp"

extern "C" int32_t ozblasDgemm(ozblasHandle_t *oh, char transA, char transB, int32_t m, int32_t n,
                               int32_t k, double alpha, const double *A, int32_t lda,
                               const double *B, int32_t ldb, double beta, double *C, int32_t ldc);

namespace oz {

// ---- device tables ---------------------------------------------------------------------------
struct GTrans {  // out[r * ldo + c] = in[c * ldi + r], r < rows, c < cols
	const double *in;
	double *out;
	int64_t ldi, ldo;
	int rows, cols, tilesR, pad;
};
struct GStat {  // statistics of `cols` k-contiguous columns of length m
	const double *in;
	int64_t ldi;
	int m, pad;
	int *colE, *colS, *colLo;
	ExpStats *st;
};
struct GWrite {  // tiled INT8 residue planes of one operand
	const double *in;
	int64_t ldi;
	int m, n;
	const int *colLo;
	int8_t *planes;
	int64_t ldp, pstride;
	int nmod, kts;
	int shLimit;  // 2^sh mod P tables staged for sh below the planned span
	int scaleU;   // op(A): residues times u_t, from the constants of the problem's moduli count
	const ResConst *rcw;
};
struct GProb {  // one problem of the grouped MFMA launch
	const int8_t *A, *B;
	int64_t apstride, bpstride, ldp;
	uint8_t *R;  // residue planes of the products, one per (modulus, k chunk)
	int64_t ldt, tstride;
	const LevelBin *bins;
	const int2 *pairs;
	int tilesM, nmod;
};
struct GItem {
	int prob, bin, tile, pad;
};
struct GRec {  // Chinese-remainder reconstruction + epilogue of one problem
	const uint8_t *R;
	const ResConst *rc;  // constants of the problem's moduli count
	const int *rowLo, *colLo;
	double *C;
	int64_t ldt, tstride, ldc;
	double alpha, beta;
	int M, N, nmod, nchunk;
};
struct GCheck {  // plan assumptions of one problem, re-checked on every execution
	int opA, opB;    // operand indices
	int spanBudget;  // largest bound bits (min of spanA + spanB + log2 k and normA + normB) the moduli cover
	                 // (-1: planned as a zero product)
	int lk;          // ceil(log2 k)
	int spanLimit;   // largest span for which the plain split is exact (the fast-form precondition)
	int shA, shB;    // spans the residue writer's tables were staged for
};

// start[] staged in LDS by the whole workgroup (one round of independent loads, instead of the
// chain of dependent global loads a search would issue), for the statistics, which look up every
// column; tables longer than the stage are searched where they are
constexpr int kOwnerStage = 1024;
__device__ inline const int *stage_starts(const int *start, int n, int *sh) {
	if (n > kOwnerStage) return start;
	for (int i = threadIdx.x; i < n; i += blockDim.x) sh[i] = start[i];
	__syncthreads();
	return sh;
}

// largest o < n with start[o] <= b (start[] ascending, start[0] = 0)
__device__ inline int owner(const int *start, int n, int b) {
	int lo = 0, hi = n - 1;
	while (lo < hi) {
		const int mid = (lo + hi + 1) / 2;
		if (start[mid] <= b)
			lo = mid;
		else
			hi = mid - 1;
	}
	return lo;
}

__global__ void k_batch_reset(ExpStats *st, int nops, int *flag, int *counter) {
	for (int o = blockIdx.x * blockDim.x + threadIdx.x; o < nops; o += gridDim.x * blockDim.x) {
		st[o].maxS = 0;
		st[o].minLsb = INT32_MAX;
		st[o].maxE = INT32_MIN;
		st[o].nonfinite = 0;
		st[o].maxSpan = INT32_MIN;
		st[o].maxNorm = INT32_MIN;
	}
	if (blockIdx.x == 0 && threadIdx.x == 0) {
		*flag = 0;
		*counter = 0;
	}
}

__global__ void k_counter_reset(int *counter) {
	if (threadIdx.x == 0) *counter = 0;
}

__global__ __launch_bounds__(256) void k_transpose_grouped(const GTrans *g, const int *start, int n) {
	__shared__ double t[32][33];
	const int o = owner(start, n, blockIdx.x);
	const GTrans tr = g[o];
	const int lb = blockIdx.x - start[o];
	const int bx = (lb % tr.tilesR) * 32, by = (lb / tr.tilesR) * 32;
	const int tx = threadIdx.x % 32, ty = threadIdx.x / 32;
	for (int k = ty; k < 32; k += 8) {
		const int r = bx + tx, c = by + k;
		if (r < tr.rows && c < tr.cols) t[k][tx] = tr.in[(int64_t)c * tr.ldi + r];
	}
	__syncthreads();
	for (int k = ty; k < 32; k += 8) {
		const int c = by + tx, r = bx + k;
		if (r < tr.rows && c < tr.cols) tr.out[(int64_t)r * tr.ldo + c] = t[tx][k];
	}
}

// kStatCols consecutive columns (of all operands, in order) per workgroup, kStatCols / 4 per wave;
// each wave folds the statistics of each operand's columns into its counters once
constexpr int kStatCols = 16;
__global__ __launch_bounds__(256) void k_stats_grouped(const GStat *g, const int *start, int n,
                                                       int total) {
	constexpr int per = kStatCols / 4;
	__shared__ int sStart[kOwnerStage];
	start = stage_starts(start, n, sStart);
	const int w = threadIdx.x / 64, l = threadIdx.x % 64;
	const int c0 = blockIdx.x * kStatCols + w * per, c1 = min(c0 + per, total);
	int cur = -1;
	ColAcc acc;
	GStat s{};
	for (int c = c0; c < c1; c++) {
		const int o = owner(start, n, c);
		if (o != cur) {
			if (cur >= 0 && l == 0) col_acc_flush(acc, s.st);
			acc = ColAcc();
			cur = o;
			s = g[o];
		}
		const int j = c - start[o];
		digit_col_stats_wave(s.in + (int64_t)j * s.ldi, s.m, s.colE + j, s.colS + j, acc, s.colLo + j);
	}
	if (cur >= 0 && l == 0) col_acc_flush(acc, s.st);
}

// flags a problem whose operands no longer satisfy what its plan was built for
__global__ void k_check_grouped(const GCheck *c, int n, const ExpStats *st, int *flag) {
	for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
		const GCheck q = c[p];
		const ExpStats a = st[q.opA], b = st[q.opB];
		bool bad = a.nonfinite || b.nonfinite;
		if (a.maxS > 0 && b.maxS > 0) {
			bad = bad || q.spanBudget < 0 ||
			      min(a.maxSpan + b.maxSpan + q.lk, a.maxNorm + b.maxNorm) > q.spanBudget;
			bad = bad || a.maxSpan >= kResMaxShift || b.maxSpan >= kResMaxShift;
			bad = bad || a.maxSpan > q.shA || b.maxSpan > q.shB;
			bad = bad || a.maxSpan > q.spanLimit || b.maxSpan > q.spanLimit;
			bad = bad || a.minLsb + b.minLsb < -1000 || a.maxE + b.maxE > 1000;
			bad = bad || a.minLsb < -1000 || b.minLsb < -1000;
		}
		if (bad) atomicOr(flag, 1);
	}
}

// one workgroup per (operand, 16-column block), over every k tile of the block: the tables of the
// operand's moduli, staged up to its span, serve the whole block
__global__ __launch_bounds__(256) void k_write_grouped(const GWrite *g, const int *start, int n) {
	__shared__ ResLds T;
	const int o = owner(start, n, blockIdx.x);
	const GWrite w = g[o];
	res_lds_load(T, w.rcw, w.nmod, w.shLimit, w.scaleU);
	__syncthreads();
	const int jb = blockIdx.x - start[o];
	for (int kt = 0; kt < w.kts; kt++)
		res_write_tile_i8(w.m, w.n, w.in, w.ldi, w.colLo, w.nmod, T, w.planes, w.ldp, w.pstride, kt, jb);
}

// Persistent grouped MFMA launch: every workgroup takes the next (problem, bin, tile) item from
// the counter until the list is exhausted.  Items are sorted by decreasing cost.
template <int TM, int TN>
__global__ __launch_bounds__(256, 1) void k_gemm_grouped(const GProb *P, const GItem *items,
                                                         int nitems, int *counter, const int *flag,
                                                         const ResConst *rc) {
	constexpr int TBM = 64 * TM, TBN = 64 * TN;
	__shared__ __attribute__((aligned(16))) char stage0[(TBM + TBN) * I8_BK];
	__shared__ __attribute__((aligned(16))) char stage1[(TBM + TBN) * I8_BK];
	__shared__ int next;
	if (*flag) return;  // the plan check failed: the call will be planned again
	for (;;) {
		if (threadIdx.x == 0) next = atomicAdd(counter, 1);
		__syncthreads();  // also: the previous item's LDS traffic is complete
		// uniform across the workgroup: the item and problem tables are then read with scalar loads
		const int it = __builtin_amdgcn_readfirstlane(next);
		__syncthreads();
		if (it >= nitems) break;
		const GItem w = items[it];
		const GProb p = P[w.prob];
		const LevelBin bin = p.bins[w.bin];
		const int2 st0 = p.pairs[bin.pair0];
		I8Tile g;
		g.A = p.A;
		g.B = p.B;
		g.apstride = p.apstride;
		g.bpstride = p.bpstride;
		g.ldp = p.ldp;
		g.s0 = st0.x;
		g.t0 = st0.y;
		g.npairs = bin.npairs;
		g.k0 = bin.k0;
		g.k1 = bin.k1;
		g.i0 = (w.tile % p.tilesM) * TBM;
		g.j0 = (w.tile / p.tilesM) * TBN;
		g.T = nullptr;
		g.ldt = p.ldt;
		g.R = p.R + (int64_t)(bin.k0 / kResChunk) * p.nmod * p.tstride;  // bin: moduli over one k chunk
		g.rstride = p.tstride;
		g.rc = rc;
		i8_tile<TM, TN, true>(g, stage0, stage1);
	}
}

// kRecGroups column groups of 4 per thread (the problem lookup and its table read once for them)
constexpr int kRecGroups = 4;
template <int NW>
__global__ __launch_bounds__(256) void k_recon_grouped(const GRec *g, const int *start, int n,
                                                       const int *flag) {
	if (*flag) return;  // C stays untouched when the plan check failed
	const int o = owner(start, n, blockIdx.x);
	const GRec r = g[o];
	// item e = (row e % M, column group e / M): one division, then rows carried (the plan keeps
	// M * G below 2^31 - 2^20)
	const int total = r.M * ((r.N + 3) / 4);
	const int e0 = (blockIdx.x - start[o]) * 256 * kRecGroups + (int)threadIdx.x;
	if (e0 >= total) return;
	int i = e0 % r.M, gq = e0 / r.M;
	for (int e = e0; e < total && e < e0 + 256 * kRecGroups; e += 256) {
		res_recon_group<NW>(r.M, r.N, r.R, r.ldt, r.tstride, r.nchunk, r.rc, r.nmod, r.rowLo, r.colLo,
		                    r.alpha, r.beta, r.C, r.ldc, i, gq);
		i += 256;
		while (i >= r.M) {  // at most 256 / M + 1 steps
			i -= r.M;
			gq++;
		}
	}
}

// ---- host plan -------------------------------------------------------------------------------
struct BatchKey {
	std::vector<int32_t> m, n, k, lda, ldb, ldc;
	std::vector<char> ta, tb;
	std::vector<const double *> A, B;
	std::vector<double *> C;
	std::vector<uint64_t> alpha, beta;  // bit patterns
	int32_t flags[6];
	bool operator==(const BatchKey &o) const {
		return m == o.m && n == o.n && k == o.k && lda == o.lda && ldb == o.ldb && ldc == o.ldc &&
		       ta == o.ta && tb == o.tb && A == o.A && B == o.B && C == o.C && alpha == o.alpha &&
		       beta == o.beta && std::memcmp(flags, o.flags, sizeof(flags)) == 0;
	}
};

struct BatchPlan {
	BatchKey key;
	char *mem1 = nullptr, *mem2 = nullptr;  // device memory owned by the plan
	ExpStatsH *hSt = nullptr;               // pinned host images of the statistics / flag
	int *hFlag = nullptr;
	hipStream_t cap = nullptr;              // capture stream (non-blocking, default priority)
	hipGraph_t graph = nullptr;
	hipGraphExec_t exec = nullptr;
	// phase 1
	int nops = 0, nTrans = 0, nTransBlocks = 0, nStatBlocks = 0;
	GTrans *dTrans = nullptr;
	int *dTransStart = nullptr;
	GStat *dStat = nullptr;
	int *dStatStart = nullptr;
	ExpStats *dSt = nullptr;
	int *dFlag = nullptr, *dCounter = nullptr;
	std::vector<const double *> opIn;  // k-contiguous source of every operand
	std::vector<int64_t> opLd;
	std::vector<int> opCols, opK;
	std::vector<int *> opColE, opColLo;
	// phase 2
	std::vector<int> fast;      // problems in the grouped launch
	std::vector<int> fallback;  // problems computed by ozblasDgemm one by one
	int nWrite = 0, nWriteBlocks = 0, nRec = 0, nRecBlocks = 0, nCheck = 0;
	size_t writeLds = 0;  // dynamic LDS of k_write_grouped: the largest operand's shift table
	GWrite *dWrite = nullptr;
	int *dWriteStart = nullptr;
	GProb *dProb[2] = {nullptr, nullptr};
	GItem *dItems[2] = {nullptr, nullptr};
	int nItems[2] = {0, 0};
	GRec *dRec = nullptr;
	int *dRecStart = nullptr;
	GCheck *dCheck = nullptr;
	// the reconstruction by word count (NW 3, 4, 6, 9): the problems of class c are GRec entries
	// [recFirst[c], recFirst[c] + recN[c]), recBlocks[c] workgroups, block starts relative to the class
	int recFirst[4] = {0, 0, 0, 0}, recN[4] = {0, 0, 0, 0}, recBlocks[4] = {0, 0, 0, 0};
	int tcfg = 0;  // 0: 128 x 128 tiles, 1: 256 x 256 tiles
	int grid[2] = {0, 0};
	float tuneMs[2] = {0.f, 0.f};
	double ops = 0.0;
	int64_t totalPairs = 0;  // slice products (moduli) over the grouped problems
	const ResConst *dRcAll = nullptr;  // constants of all the moduli (residue writer, GEMM epilogue)
};

static void plan_free(BatchPlan *p) {
	if (!p) return;
	if (p->exec) (void)hipGraphExecDestroy(p->exec);
	if (p->graph) (void)hipGraphDestroy(p->graph);
	if (p->cap) (void)hipStreamDestroy(p->cap);
	if (p->mem1) (void)hipFree(p->mem1);
	if (p->mem2) (void)hipFree(p->mem2);
	if (p->hSt) (void)hipHostFree(p->hSt);
	if (p->hFlag) (void)hipHostFree(p->hFlag);
	delete p;
}

void batch_plan_free(void *p) { plan_free(reinterpret_cast<BatchPlan *>(p)); }

// bump layout of one device allocation
struct Layout {
	uint64_t top = 0;
	uint64_t take(uint64_t bytes) {
		const uint64_t off = (top + 255) & ~uint64_t(255);
		top = off + bytes;
		return off;
	}
};

static inline int64_t rup64(int64_t v, int64_t a) { return (v + a - 1) / a * a; }

static int32_t launch_execute(BatchPlan *p, hipStream_t st) {
	hipLaunchKernelGGL(k_batch_reset, dim3((p->nops + 255) / 256), dim3(256), 0, st, p->dSt, p->nops,
	                   p->dFlag, p->dCounter);
	if (p->nTransBlocks > 0)
		hipLaunchKernelGGL(k_transpose_grouped, dim3(p->nTransBlocks), dim3(256), 0, st, p->dTrans,
		                   p->dTransStart, p->nTrans);
	if (p->nStatBlocks > 0)
		hipLaunchKernelGGL(k_stats_grouped, dim3((p->nStatBlocks + kStatCols - 1) / kStatCols), dim3(256),
		                   0, st, p->dStat, p->dStatStart, p->nops, p->nStatBlocks);
	if (p->nCheck > 0)
		hipLaunchKernelGGL(k_check_grouped, dim3((p->nCheck + 255) / 256), dim3(256), 0, st, p->dCheck,
		                   p->nCheck, p->dSt, p->dFlag);
	if (p->nWriteBlocks > 0)
		hipLaunchKernelGGL(k_write_grouped, dim3(p->nWriteBlocks), dim3(256), p->writeLds, st, p->dWrite,
		                   p->dWriteStart, p->nWrite);
	const int c = p->tcfg;
	if (p->nItems[c] > 0) {
		if (c == 0)
			hipLaunchKernelGGL((k_gemm_grouped<2, 2>), dim3(p->grid[0]), dim3(256), 0, st, p->dProb[0],
			                   p->dItems[0], p->nItems[0], p->dCounter, p->dFlag, p->dRcAll);
		else
			hipLaunchKernelGGL((k_gemm_grouped<4, 4>), dim3(p->grid[1]), dim3(256), 0, st, p->dProb[1],
			                   p->dItems[1], p->nItems[1], p->dCounter, p->dFlag, p->dRcAll);
	}
	// one launch per word-count class: the registers of the widest class do not bound the others
	for (int cl = 0; cl < 4; cl++) {
		if (p->recBlocks[cl] == 0) continue;
		const GRec *g = p->dRec + p->recFirst[cl];
		const int *gs = p->dRecStart + p->recFirst[cl];
		const dim3 grid((unsigned)p->recBlocks[cl]);
		if (cl == 0)
			hipLaunchKernelGGL(k_recon_grouped<3>, grid, dim3(256), 0, st, g, gs, p->recN[cl], p->dFlag);
		else if (cl == 1)
			hipLaunchKernelGGL(k_recon_grouped<4>, grid, dim3(256), 0, st, g, gs, p->recN[cl], p->dFlag);
		else if (cl == 2)
			hipLaunchKernelGGL(k_recon_grouped<6>, grid, dim3(256), 0, st, g, gs, p->recN[cl], p->dFlag);
		else
			hipLaunchKernelGGL(k_recon_grouped<9>, grid, dim3(256), 0, st, g, gs, p->recN[cl], p->dFlag);
	}
	OZ_HIP(hipGetLastError());
	return OZBLAS_STATUS_SUCCESS;
}

// Builds the plan: phase 1 (operands, statistics, one synchronisation), phase 2 (planes, bins,
// items, reconstruction tables), autotuning and graph capture.
static int32_t plan_build(ozblasHandle_t *oh, BatchPlan *p, int batch, const char *transA,
                          const char *transB, const int32_t *m, const int32_t *n, const int32_t *k,
                          const double *alpha, const double *const *A, const int32_t *lda,
                          const double *const *B, const int32_t *ldb, const double *beta,
                          double *const *C, const int32_t *ldc) {
	Backend *b = be(oh);
	hipStream_t st = b->stream;
	// ---- phase 1: operands (2p = A of problem p, 2p + 1 = B) --------------------------------
	std::vector<int> live;  // problems with m, n > 0
	for (int q = 0; q < batch; q++)
		if (m[q] > 0 && n[q] > 0) live.push_back(q);
	p->nops = 2 * batch;
	p->opIn.assign(p->nops, nullptr);
	p->opLd.assign(p->nops, 0);
	p->opCols.assign(p->nops, 0);
	p->opK.assign(p->nops, 0);
	p->opColE.assign(p->nops, nullptr);
	p->opColLo.assign(p->nops, nullptr);
	Layout L1;
	std::vector<uint64_t> transOff(p->nops, UINT64_MAX), colEOff(p->nops), colSOff(p->nops), colLoOff(p->nops);
	std::vector<GTrans> hTrans;
	std::vector<int> transStart;
	for (int q : live) {
		const int kk = k[q];
		const int64_t ldk = rup64(kk > 0 ? kk : 1, 16);
		// A: rows of op(A) must be k-contiguous -> transpose when transA = 'N'
		const bool tA = (transA[q] == 'N' || transA[q] == 'n');
		const bool tB = !(transB[q] == 'N' || transB[q] == 'n');
		if (tA && kk > 0) transOff[2 * q] = L1.take(8ull * ldk * m[q]);
		if (tB && kk > 0) transOff[2 * q + 1] = L1.take(8ull * ldk * n[q]);
		p->opCols[2 * q] = m[q];
		p->opCols[2 * q + 1] = n[q];
		p->opK[2 * q] = p->opK[2 * q + 1] = kk;
		colEOff[2 * q] = L1.take(4ull * m[q]);
		colSOff[2 * q] = L1.take(4ull * m[q]);
		colEOff[2 * q + 1] = L1.take(4ull * n[q]);
		colSOff[2 * q + 1] = L1.take(4ull * n[q]);
		colLoOff[2 * q] = L1.take(4ull * m[q]);
		colLoOff[2 * q + 1] = L1.take(4ull * n[q]);
	}
	const uint64_t stOff = L1.take(sizeof(ExpStats) * (uint64_t)p->nops);
	const uint64_t flagOff = L1.take(256);
	const uint64_t ntr = 2ull * live.size();
	const uint64_t trOff = L1.take(sizeof(GTrans) * (ntr + 1));
	const uint64_t trStartOff = L1.take(4 * (ntr + 1));
	const uint64_t statOff = L1.take(sizeof(GStat) * (uint64_t)p->nops);
	const uint64_t statStartOff = L1.take(4ull * (p->nops + 1));
	OZ_HIP(hipMalloc(&p->mem1, L1.top));
	char *M1 = p->mem1;
	p->dSt = (ExpStats *)(M1 + stOff);
	p->dFlag = (int *)(M1 + flagOff);
	p->dCounter = p->dFlag + 1;
	p->dTrans = (GTrans *)(M1 + trOff);
	p->dTransStart = (int *)(M1 + trStartOff);
	p->dStat = (GStat *)(M1 + statOff);
	p->dStatStart = (int *)(M1 + statStartOff);
	std::vector<GStat> hStat(p->nops);
	std::vector<int> statStart(p->nops + 1, 0);
	{
		int tb = 0;
		for (int q : live) {
			for (int side = 0; side < 2; side++) {
				const int o = 2 * q + side;
				const int kk = k[q];
				const int rows = side == 0 ? m[q] : n[q];  // rows of op(A) / columns of op(B)
				const double *src = side == 0 ? A[q] : B[q];
				const int64_t ld = side == 0 ? lda[q] : ldb[q];
				if (transOff[o] != UINT64_MAX) {
					GTrans t;
					t.in = src;
					t.ldi = ld;
					t.out = (double *)(M1 + transOff[o]);
					t.ldo = rup64(kk, 16);
					t.rows = rows;
					t.cols = kk;
					t.tilesR = (rows + 31) / 32;
					t.pad = 0;
					transStart.push_back(tb);
					tb += t.tilesR * ((kk + 31) / 32);
					hTrans.push_back(t);
					p->opIn[o] = t.out;
					p->opLd[o] = t.ldo;
				} else {
					p->opIn[o] = src;
					p->opLd[o] = ld;
				}
				p->opColE[o] = (int *)(M1 + colEOff[o]);
				p->opColLo[o] = (int *)(M1 + colLoOff[o]);
			}
		}
		p->nTrans = (int)hTrans.size();
		p->nTransBlocks = tb;
		int sb = 0;
		for (int o = 0; o < p->nops; o++) {
			GStat &g = hStat[o];
			g.in = p->opIn[o];
			g.ldi = p->opLd[o];
			g.m = p->opK[o];
			g.pad = 0;
			g.colE = p->opColE[o];
			g.colS = p->opColE[o] ? (int *)(M1 + colSOff[o]) : nullptr;
			g.colLo = p->opColLo[o];
			g.st = p->dSt + o;
			statStart[o] = sb;
			sb += p->opCols[o];  // columns, in operand order (0 for absent problems)
		}
		statStart[p->nops] = sb;
		p->nStatBlocks = sb;
	}
	if (p->nTrans > 0) {
		OZ_HIP(hipMemcpyAsync(p->dTrans, hTrans.data(), sizeof(GTrans) * hTrans.size(), hipMemcpyHostToDevice, st));
		OZ_HIP(hipMemcpyAsync(p->dTransStart, transStart.data(), 4 * transStart.size(), hipMemcpyHostToDevice, st));
	}
	OZ_HIP(hipMemcpyAsync(p->dStat, hStat.data(), sizeof(GStat) * hStat.size(), hipMemcpyHostToDevice, st));
	OZ_HIP(hipMemcpyAsync(p->dStatStart, statStart.data(), 4 * statStart.size(), hipMemcpyHostToDevice, st));
	OZ_HIP(hipHostMalloc((void **)&p->hSt, sizeof(ExpStatsH) * p->nops, hipHostMallocDefault));
	OZ_HIP(hipHostMalloc((void **)&p->hFlag, 64, hipHostMallocDefault));
	// statistics pass (the same kernels as the execute sequence, run eagerly)
	p->nCheck = 0;
	p->nWriteBlocks = 0;
	p->nRecBlocks = 0;
	p->nItems[0] = p->nItems[1] = 0;
	hipLaunchKernelGGL(k_batch_reset, dim3((p->nops + 255) / 256), dim3(256), 0, st, p->dSt, p->nops,
	                   p->dFlag, p->dCounter);
	if (p->nTransBlocks > 0)
		hipLaunchKernelGGL(k_transpose_grouped, dim3(p->nTransBlocks), dim3(256), 0, st, p->dTrans,
		                   p->dTransStart, p->nTrans);
	if (p->nStatBlocks > 0)
		hipLaunchKernelGGL(k_stats_grouped, dim3((p->nStatBlocks + kStatCols - 1) / kStatCols), dim3(256),
		                   0, st, p->dStat, p->dStatStart, p->nops, p->nStatBlocks);
	OZ_HIP(hipGetLastError());
	OZ_HIP(hipMemcpyAsync(p->hSt, p->dSt, sizeof(ExpStatsH) * p->nops, hipMemcpyDeviceToHost, st));
	OZ_HIP(hipStreamSynchronize(st));
	// ---- phase 2: residue planes, bins, items ---------------------------------------------------
	// moduli per problem: |X| < k 2^(spanA + spanB) must stay within M / 4 (oz_residue.hip)
	auto ceil_log2 = [](int v) {
		int l = 0;
		while ((int64_t(1) << l) < (int64_t)v) l++;
		return l;
	};
	struct ProbH {
		int q, nmod, nchunk, budget;
		int64_t ldp, apstride, bpstride, ldt, tstride;
		uint64_t planesA, planesB, R;
		int bin0[2], nbins[2];  // per tile shape: all moduli of a chunk per bin (128), one modulus (256)
	};
	std::vector<ProbH> probs;
	for (int q : live) {
		const ExpStatsH &sa = p->hSt[2 * q], &sb = p->hSt[2 * q + 1];
		// (the grouped reconstruction indexes a problem's outputs in 32 bits)
		if (!(k[q] > 0 && fast_valid(oh, sa, sb, k[q])) || (int64_t)m[q] * ((n[q] + 3) / 4) >= (1LL << 31) - (1 << 20)) {
			p->fallback.push_back(q);
			continue;
		}
		ProbH h{};
		h.q = q;
		if (sa.maxS > 0 && sb.maxS > 0) {
			if (sa.maxSpan >= kResMaxShift || sb.maxSpan >= kResMaxShift) {
				p->fallback.push_back(q);
				continue;
			}
			const double need = (double)std::min(sa.maxSpan + sb.maxSpan + ceil_log2(k[q]), sa.maxNorm + sb.maxNorm) + 2;
			int nm = 1;
			while (nm <= kResMaxMod && res_log2_product(nm) < need + 1e-6) nm++;
			if (nm > kResMaxMod) {  // beyond the moduli: ozblasDgemm takes the problem alone
				p->fallback.push_back(q);
				continue;
			}
			h.nmod = nm;
			h.budget = (int)std::floor(res_log2_product(nm) - 1e-6) - 2;
		} else {
			h.nmod = 0;  // an all-zero operand: C = fma(alpha, +0, beta C)
			h.budget = -1;
		}
		p->fast.push_back(q);
		probs.push_back(h);
	}
	const int cap = split_cap(oh);
	Layout L2;
	std::vector<LevelBin> hBins;  // the bins of both tile shapes
	std::vector<int2> hPairs;
	for (int t = 0; t < kResMaxMod; t++) hPairs.push_back(make_int2(t, t));  // product t: plane t x plane t
	std::vector<int> usedNmod;
	for (ProbH &h : probs) {
		const int q = h.q;
		h.ldp = rup64(k[q], 128);
		h.apstride = h.ldp * rup64(m[q], 256);
		h.bpstride = h.ldp * rup64(n[q], 256);
		h.ldt = rup64(m[q], 256);
		h.tstride = h.ldt * rup64(n[q], 256);
		h.nbins[0] = h.nbins[1] = 0;
		h.bin0[0] = h.bin0[1] = (int)hBins.size();
		h.nchunk = 1;
		if (h.nmod > 0) {
			h.planesA = L2.take((uint64_t)h.apstride * h.nmod);
			h.planesB = L2.take((uint64_t)h.bpstride * h.nmod);
			const int Kpad = (int)h.ldp;
			const int kc = kResChunk;
			h.nchunk = (Kpad + kc - 1) / kc;
			// 128 x 128 tiles: one item streams every modulus of a tile and chunk through one
			// pipeline; 256 x 256 tiles: one modulus per item
			for (int c = 0; c < h.nchunk; c++) {
				LevelBin bn;
				bn.level = 0;
				bn.pair0 = 0;
				bn.npairs = h.nmod;
				bn.k0 = c * kc;
				bn.k1 = std::min(Kpad, (c + 1) * kc);
				hBins.push_back(bn);
			}
			h.nbins[0] = (int)hBins.size() - h.bin0[0];
			h.bin0[1] = (int)hBins.size();
			for (int c = 0; c < h.nchunk; c++)
				for (int t = 0; t < h.nmod; t++) {
					LevelBin bn;
					bn.level = t;
					bn.pair0 = t;
					bn.npairs = 1;
					bn.k0 = c * kc;
					bn.k1 = std::min(Kpad, (c + 1) * kc);
					hBins.push_back(bn);
				}
			h.nbins[1] = (int)hBins.size() - h.bin0[1];
			h.R = L2.take((uint64_t)h.tstride * h.nmod * h.nchunk);
			p->totalPairs += h.nmod;
			p->ops += 2.0 * m[q] * (double)n[q] * k[q] * h.nmod;
			if (std::find(usedNmod.begin(), usedNmod.end(), h.nmod) == usedNmod.end()) usedNmod.push_back(h.nmod);
		}
	}
	// constants: the full moduli table, then one per moduli count in use
	const uint64_t rcAllOff = L2.take(sizeof(ResConst));
	std::vector<uint64_t> rcOff(kResMaxMod + 1, 0);
	for (int nm : usedNmod) rcOff[nm] = L2.take(sizeof(ResConst));
	// tables
	std::vector<GWrite> hWrite;
	size_t writeLds = 0;
	std::vector<int> writeStart;
	std::vector<GRec> hRec;
	std::vector<int> recStart, recClass;
	std::vector<GCheck> hCheck;
	std::vector<GProb> hProb[2];
	std::vector<GItem> hItems[2];
	std::vector<double> itemCost[2];
	const uint64_t binsOff = L2.take(sizeof(LevelBin) * (hBins.size() + 1));
	const uint64_t pairsOff = L2.take(sizeof(int2) * (hPairs.size() + 1));
	// sizes of the remaining tables are known from probs
	size_t nItemsMax[2] = {0, 0};
	for (const ProbH &h : probs)
		for (int c = 0; c < 2; c++) {
			const int TB = c == 0 ? 128 : 256;
			nItemsMax[c] += (size_t)h.nbins[c] * ((m[h.q] + TB - 1) / TB) * ((n[h.q] + TB - 1) / TB);
		}
	const uint64_t writeOff = L2.take(sizeof(GWrite) * (2 * probs.size() + 1));
	const uint64_t writeStartOff = L2.take(4 * (2 * probs.size() + 1));
	const uint64_t recOff = L2.take(sizeof(GRec) * (probs.size() + 1));
	const uint64_t recStartOff = L2.take(4 * (probs.size() + 1));
	const uint64_t checkOff = L2.take(sizeof(GCheck) * (probs.size() + 1));
	uint64_t probOff[2], itemOff[2];
	for (int c = 0; c < 2; c++) {
		probOff[c] = L2.take(sizeof(GProb) * (probs.size() + 1));
		itemOff[c] = L2.take(sizeof(GItem) * (nItemsMax[c] + 1));
	}
	OZ_HIP(hipMalloc(&p->mem2, L2.top));
	char *M2 = p->mem2;
	const LevelBin *dBins = (const LevelBin *)(M2 + binsOff);
	const int2 *dPairs = (const int2 *)(M2 + pairsOff);
	p->dRcAll = (const ResConst *)(M2 + rcAllOff);
	OZ_HIP(hipMemcpyAsync(M2 + rcAllOff, &res_const_host(kResMaxMod), sizeof(ResConst), hipMemcpyHostToDevice, st));
	for (int nm : usedNmod)
		OZ_HIP(hipMemcpyAsync(M2 + rcOff[nm], &res_const_host(nm), sizeof(ResConst), hipMemcpyHostToDevice, st));
	int wb = 0, rb = 0;
	for (const ProbH &h : probs) {
		const int q = h.q;
		if (h.nmod > 0) {
			for (int side = 0; side < 2; side++) {
				const int o = 2 * q + side;
				GWrite w;
				w.in = p->opIn[o];
				w.ldi = p->opLd[o];
				w.m = k[q];
				w.n = p->opCols[o];
				w.colLo = p->opColLo[o];
				w.planes = (int8_t *)(M2 + (side == 0 ? h.planesA : h.planesB));
				w.ldp = h.ldp;
				w.pstride = side == 0 ? h.apstride : h.bpstride;
				w.nmod = h.nmod;
				w.kts = (int)(h.ldp / 64);
				w.shLimit = std::max(1, p->hSt[o].maxSpan);
				writeLds = std::max(writeLds, res_lds_bytes(w.nmod, w.shLimit));
				w.scaleU = side == 0 ? 1 : 0;
				w.rcw = (const ResConst *)(M2 + rcOff[h.nmod]);
				writeStart.push_back(wb);
				wb += (int)(rup64(w.n, 256) / 16);
				hWrite.push_back(w);
			}
			for (int c = 0; c < 2; c++) {
				const int TM = c == 0 ? 2 : 4, TB = 64 * TM;
				GProb g;
				g.A = (const int8_t *)(M2 + h.planesA);
				g.B = (const int8_t *)(M2 + h.planesB);
				g.apstride = h.apstride;
				g.bpstride = h.bpstride;
				g.ldp = h.ldp;
				g.R = (uint8_t *)(M2 + h.R);
				g.ldt = h.ldt;
				g.tstride = h.tstride;
				g.bins = dBins + h.bin0[c];
				g.pairs = dPairs;
				g.tilesM = (m[q] + TB - 1) / TB;
				g.nmod = h.nmod;
				const int pi = (int)hProb[c].size();
				hProb[c].push_back(g);
				const int tiles = g.tilesM * ((n[q] + TB - 1) / TB);
				for (int bi = 0; bi < h.nbins[c]; bi++) {
					const LevelBin &bn = hBins[h.bin0[c] + bi];
					const double cost = (double)bn.npairs * (bn.k1 - bn.k0);
					for (int t = 0; t < tiles; t++) {
						hItems[c].push_back(GItem{pi, bi, t, 0});
						itemCost[c].push_back(cost);
					}
				}
			}
		}
		GRec r;
		r.R = (h.nmod > 0) ? (const uint8_t *)(M2 + h.R) : nullptr;
		r.rc = (h.nmod > 0) ? (const ResConst *)(M2 + rcOff[h.nmod]) : p->dRcAll;
		r.rowLo = p->opColLo[2 * q];
		r.colLo = p->opColLo[2 * q + 1];
		r.C = C[q];
		r.ldt = h.ldt;
		r.tstride = h.tstride;
		r.ldc = ldc[q];
		r.alpha = alpha[q];
		r.beta = beta[q];
		r.M = m[q];
		r.N = n[q];
		r.nmod = h.nmod;
		r.nchunk = h.nchunk;
		const int nw = (h.nmod > 0) ? (res_const_host(h.nmod).nl + 1) / 2 : 3;
		recClass.push_back(nw <= 3 ? 0 : nw <= 4 ? 1 : nw <= 6 ? 2 : 3);
		hRec.push_back(r);
		GCheck ck;
		ck.opA = 2 * q;
		ck.opB = 2 * q + 1;
		ck.spanBudget = h.budget;
		ck.lk = ceil_log2(std::max(1, k[q]));
		const int rho = rho_host(k[q], oh->splitEpsMode);
		ck.spanLimit = (53 - rho - 1) * (cap - 3);
		ck.shA = std::max(1, p->hSt[2 * q].maxSpan);
		ck.shB = std::max(1, p->hSt[2 * q + 1].maxSpan);
		hCheck.push_back(ck);
	}
	// longest items first (the work counter then balances the tail)
	for (int c = 0; c < 2; c++) {
		std::vector<int> ord(hItems[c].size());
		for (size_t i = 0; i < ord.size(); i++) ord[i] = (int)i;
		std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return itemCost[c][x] > itemCost[c][y]; });
		std::vector<GItem> sorted(ord.size());
		for (size_t i = 0; i < ord.size(); i++) sorted[i] = hItems[c][ord[i]];
		hItems[c].swap(sorted);
		p->nItems[c] = (int)hItems[c].size();
		p->dProb[c] = (GProb *)(M2 + probOff[c]);
		p->dItems[c] = (GItem *)(M2 + itemOff[c]);
		if (!hProb[c].empty())
			OZ_HIP(hipMemcpyAsync(p->dProb[c], hProb[c].data(), sizeof(GProb) * hProb[c].size(), hipMemcpyHostToDevice, st));
		if (!hItems[c].empty())
			OZ_HIP(hipMemcpyAsync(p->dItems[c], hItems[c].data(), sizeof(GItem) * hItems[c].size(), hipMemcpyHostToDevice, st));
	}
	p->nWrite = (int)hWrite.size();
	p->nWriteBlocks = wb;
	p->writeLds = writeLds;
	p->dWrite = (GWrite *)(M2 + writeOff);
	p->dWriteStart = (int *)(M2 + writeStartOff);
	{  // the reconstruction tables in class order, block starts relative to each class
		std::vector<GRec> byClass;
		for (int cl = 0; cl < 4; cl++) {
			p->recFirst[cl] = (int)byClass.size();
			int blocks = 0;
			for (size_t i = 0; i < hRec.size(); i++) {
				if (recClass[i] != cl) continue;
				recStart.push_back(blocks);
				blocks += (int)(((int64_t)hRec[i].M * ((hRec[i].N + 3) / 4) + 256 * kRecGroups - 1) /
				                (256 * kRecGroups));
				byClass.push_back(hRec[i]);
			}
			p->recN[cl] = (int)byClass.size() - p->recFirst[cl];
			p->recBlocks[cl] = blocks;
			rb += blocks;
		}
		hRec.swap(byClass);
	}
	p->nRec = (int)hRec.size();
	p->nRecBlocks = rb;
	p->dRec = (GRec *)(M2 + recOff);
	p->dRecStart = (int *)(M2 + recStartOff);
	p->nCheck = (int)hCheck.size();
	p->dCheck = (GCheck *)(M2 + checkOff);
	if (!hBins.empty())
		OZ_HIP(hipMemcpyAsync(M2 + binsOff, hBins.data(), sizeof(LevelBin) * hBins.size(), hipMemcpyHostToDevice, st));
	OZ_HIP(hipMemcpyAsync(M2 + pairsOff, hPairs.data(), sizeof(int2) * hPairs.size(), hipMemcpyHostToDevice, st));
	if (p->nWrite > 0) {
		OZ_HIP(hipMemcpyAsync(p->dWrite, hWrite.data(), sizeof(GWrite) * hWrite.size(), hipMemcpyHostToDevice, st));
		OZ_HIP(hipMemcpyAsync(p->dWriteStart, writeStart.data(), 4 * writeStart.size(), hipMemcpyHostToDevice, st));
	}
	if (p->nRec > 0) {
		OZ_HIP(hipMemcpyAsync(p->dRec, hRec.data(), sizeof(GRec) * hRec.size(), hipMemcpyHostToDevice, st));
		OZ_HIP(hipMemcpyAsync(p->dRecStart, recStart.data(), 4 * recStart.size(), hipMemcpyHostToDevice, st));
		OZ_HIP(hipMemcpyAsync(p->dCheck, hCheck.data(), sizeof(GCheck) * hCheck.size(), hipMemcpyHostToDevice, st));
	}
	// persistent grids: resident workgroups of each tile shape on the whole device
	int cus = 0;
	OZ_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device));
	for (int c = 0; c < 2; c++) {
		int per = 0;
		if (c == 0)
			OZ_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, (const void *)k_gemm_grouped<2, 2>, 256, 0));
		else
			OZ_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, (const void *)k_gemm_grouped<4, 4>, 256, 0));
		p->grid[c] = std::max(1, per) * cus;
	}
	// ---- autotune: digit planes once, then one grouped launch per tile shape -------------------
	if (p->nWriteBlocks > 0)
		hipLaunchKernelGGL(k_write_grouped, dim3(p->nWriteBlocks), dim3(256), p->writeLds, st, p->dWrite,
		                   p->dWriteStart, p->nWrite);
	p->tcfg = 0;
	if (p->nItems[0] > 0) {
		for (int c = 0; c < 2; c++) {
			hipLaunchKernelGGL(k_counter_reset, dim3(1), dim3(64), 0, st, p->dCounter);
			OZ_HIP(hipEventRecord(b->ev[6], st));
			if (c == 0)
				hipLaunchKernelGGL((k_gemm_grouped<2, 2>), dim3(p->grid[0]), dim3(256), 0, st, p->dProb[0],
				                   p->dItems[0], p->nItems[0], p->dCounter, p->dFlag, p->dRcAll);
			else
				hipLaunchKernelGGL((k_gemm_grouped<4, 4>), dim3(p->grid[1]), dim3(256), 0, st, p->dProb[1],
				                   p->dItems[1], p->nItems[1], p->dCounter, p->dFlag, p->dRcAll);
			OZ_HIP(hipEventRecord(b->ev[7], st));
			OZ_HIP(hipEventSynchronize(b->ev[7]));
			OZ_HIP(hipEventElapsedTime(&p->tuneMs[c], b->ev[6], b->ev[7]));
		}
		p->tcfg = (p->tuneMs[1] < p->tuneMs[0]) ? 1 : 0;
	}
	OZ_HIP(hipGetLastError());
	// ---- capture the execute sequence ----------------------------------------------------------
	OZ_HIP(hipStreamCreateWithFlags(&p->cap, hipStreamNonBlocking));
	OZ_HIP(hipStreamBeginCapture(p->cap, hipStreamCaptureModeThreadLocal));
	const int32_t ls = launch_execute(p, p->cap);
	hipGraph_t g = nullptr;
	const hipError_t ec = hipStreamEndCapture(p->cap, &g);
	if (ls != OZBLAS_STATUS_SUCCESS) return ls;
	OZ_HIP(ec);
	p->graph = g;
	OZ_HIP(hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0));
	return OZBLAS_STATUS_SUCCESS;
}

static BatchKey make_key(const ozblasHandle_t *oh, int batch, const char *transA, const char *transB,
                         const int32_t *m, const int32_t *n, const int32_t *k, const double *alpha,
                         const double *const *A, const int32_t *lda, const double *const *B,
                         const int32_t *ldb, const double *beta, double *const *C, const int32_t *ldc) {
	BatchKey key;
	key.m.assign(m, m + batch);
	key.n.assign(n, n + batch);
	key.k.assign(k, k + batch);
	key.lda.assign(lda, lda + batch);
	key.ldb.assign(ldb, ldb + batch);
	key.ldc.assign(ldc, ldc + batch);
	key.ta.assign(transA, transA + batch);
	key.tb.assign(transB, transB + batch);
	key.A.assign(A, A + batch);
	key.B.assign(B, B + batch);
	key.C.assign(C, C + batch);
	for (int q = 0; q < batch; q++) {
		key.alpha.push_back(d2u(alpha[q]));
		key.beta.push_back(d2u(beta[q]));
	}
	const int32_t f[6] = {oh->nSplitMax, oh->sumMode, oh->reproMode, oh->fastMode, oh->splitEpsMode,
	                      oh->sliceType};
	std::memcpy(key.flags, f, sizeof(f));
	return key;
}

static bool valid_args(int batch, const int32_t *m, const int32_t *n, const int32_t *k,
                       const int32_t *lda, const int32_t *ldb, const int32_t *ldc, const char *transA,
                       const char *transB) {
	for (int q = 0; q < batch; q++) {
		if (m[q] < 0 || n[q] < 0 || k[q] < 0) return false;
		const bool nA = transA[q] == 'N' || transA[q] == 'n';
		const bool nB = transB[q] == 'N' || transB[q] == 'n';
		if (!nA && !(transA[q] == 'T' || transA[q] == 't')) return false;
		if (!nB && !(transB[q] == 'T' || transB[q] == 't')) return false;
		if (lda[q] < std::max(1, nA ? m[q] : k[q])) return false;
		if (ldb[q] < std::max(1, nB ? k[q] : n[q])) return false;
		if (ldc[q] < std::max(1, m[q])) return false;
	}
	return true;
}

}  // namespace oz

using namespace oz;

extern "C" int32_t ozblasDgemmBatched(ozblasHandle_t *oh, int32_t batch, const char *transA,
                                      const char *transB, const int32_t *m, const int32_t *n,
                                      const int32_t *k, const double *alpha,
                                      const double *const *A, const int32_t *lda,
                                      const double *const *B, const int32_t *ldb,
                                      const double *beta, double *const *C, const int32_t *ldc) {
	OZ_TRY(ensure_init(oh));
	if (batch < 0) return oh->lastStatus = OZBLAS_STATUS_INVALID_VALUE;
	if (batch == 0) return oh->lastStatus = OZBLAS_STATUS_SUCCESS;
	if (!valid_args(batch, m, n, k, lda, ldb, ldc, transA, transB))
		return oh->lastStatus = OZBLAS_STATUS_INVALID_VALUE;
	const bool grouped = oh->sliceType == OZBLAS_SLICE_INT8 && oh->sumMode == 1 && oh->nSplitMax == 0 &&
	                     !(oh->reproMode == 0 && oh->nSplitMax == 1);
	if (!grouped) {  // every problem on its own, as ozblasDgemm
		for (int q = 0; q < batch; q++) {
			const int32_t s = ozblasDgemm(oh, transA[q], transB[q], m[q], n[q], k[q], alpha[q], A[q], lda[q],
			                              B[q], ldb[q], beta[q], C[q], ldc[q]);
			if (s != OZBLAS_STATUS_SUCCESS) return oh->lastStatus = s;
		}
		return oh->lastStatus = OZBLAS_STATUS_SUCCESS;
	}
	Backend *b = be(oh);
	const auto t0 = std::chrono::steady_clock::now();
	oh->splitShift = 1;
	BatchKey key = make_key(oh, batch, transA, transB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
	BatchPlan *p = reinterpret_cast<BatchPlan *>(b->batch);
	int replay = (p != nullptr && p->key == key) ? 1 : 0;
	for (int attempt = 0; attempt < 2; attempt++) {
		if (!replay) {
			plan_free(p);
			p = new BatchPlan();
			b->batch = p;
			p->key = key;
			const int32_t s = plan_build(oh, p, batch, transA, transB, m, n, k, alpha, A, lda, B, ldb,
			                             beta, C, ldc);
			if (s != OZBLAS_STATUS_SUCCESS) {
				plan_free(p);
				b->batch = nullptr;
				return oh->lastStatus = s;
			}
		}
		OZ_HIP(hipGraphLaunch(p->exec, b->stream));
		OZ_HIP(hipMemcpyAsync(p->hFlag, p->dFlag, sizeof(int), hipMemcpyDeviceToHost, b->stream));
		OZ_HIP(hipStreamSynchronize(b->stream));
		if (*p->hFlag == 0) break;
		// the operands no longer fit the cached plan (more digits, or the fast-form precondition
		// fails): build a new plan from these operands and execute again
		if (attempt == 1) return oh->lastStatus = OZBLAS_STATUS_EXECUTION_FAILED;
		replay = 0;
	}
	for (int q : p->fallback) {
		const int32_t s = ozblasDgemm(oh, transA[q], transB[q], m[q], n[q], k[q], alpha[q], A[q], lda[q],
		                              B[q], ldb[q], beta[q], C[q], ldc[q]);
		if (s != OZBLAS_STATUS_SUCCESS) return oh->lastStatus = s;
	}
	oh->batchTile = p->tcfg ? 256 : 128;
	oh->batchReplay = replay;
	oh->batchGrouped = (int32_t)p->fast.size();
	oh->t_comp = p->tuneMs[p->tcfg] * 1e-3f;
	oh->n_comp = (float)p->ops;
	oh->nSplitC = (float)p->totalPairs;
	oh->t_total = std::chrono::duration<float>(std::chrono::steady_clock::now() - t0).count();
	return oh->lastStatus = OZBLAS_STATUS_SUCCESS;
}
