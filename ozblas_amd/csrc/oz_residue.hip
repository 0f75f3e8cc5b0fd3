// oz_residue.hip -- the residue form of the fast GEMM (no reference counterpart; same bits as
// NearSum): the exact product from INT8 (or FP16/BF16) products of residues modulo small moduli.
//
// Scheme (Chinese remainder theorem, as in Ozaki, Uchino and Imamura's "Ozaki scheme II", used
// here for the EXACT product).  Row i of op(A) is the integer vector a'_i = a_i 2^-lo_i, lo_i the
// exponent of the lowest set bit in the row (every element is a multiple of 2^lo_i); column j of
// op(B) likewise with lo_j.  X_ij = sum_k a'_ik b'_kj is then an integer with
// |X| < k 2^(spanA + spanB) (span = top exponent - lo), and C_ij = X_ij 2^(lo_i + lo_j) exactly.
// With pairwise coprime moduli p_t <= 256 whose product M >= 4 k 2^(spanA + spanB):
//   - each operand element is reduced modulo every p_t to a balanced residue in [-128, 127]
//     (exact in int8, fp16 and bf16);
//   - one INT8 GEMM per modulus (int32 accumulation is exact below 2^31: k <= 2^17 per chunk)
//     gives X mod p_t -- or directly c_t = X u_t mod p_t, u_t = (M/p_t)^-1 mod p_t, since op(A)'s
//     residues are written multiplied by u_t (through the group tables, at no cost);
//   - the reconstruction: X = sum_t c_t M/p_t - K M with
//     K = round(sum_t c_t / p_t) (exact, since |X| <= M/4 keeps the sum a quarter away from the
//     rounding boundary), summed exactly in binary64 chunks of 39 bits (c_t < 2^8, so every chunk
//     sum of up to 49 moduli stays below 2^53), carried into a multi-word integer, then one RN-even
//     of X 2^(lo_i + lo_j).
// The number of GEMMs is the number of moduli, about (spanA + spanB + log2 k + 2) / 7.9, against
// about (spanA / 8) (spanB / 8) / 2 digit-pair GEMMs of the radix-256 digit form.
#include "oz_residue.hpp"

#include <cstring>
#include <mutex>
#include <vector>

namespace oz {

// ---- host: moduli and reconstruction constants --------------------------------------------
static std::vector<int> res_moduli() {
	// 256, then the odd numbers below it that are coprime with every one taken so far
	std::vector<int> ms{256};
	auto gcd = [](int a, int b) {
		while (b) {
			const int t = a % b;
			a = b;
			b = t;
		}
		return a;
	};
	for (int p = 255; p >= 3 && (int)ms.size() < kResMaxMod; p -= 2) {
		bool ok = true;
		for (int q : ms) ok = ok && gcd(p, q) == 1;
		if (ok) ms.push_back(p);
	}
	return ms;
}

static ResConst g_rc[kResMaxMod + 1];
static double g_log2M[kResMaxMod + 1];
static std::once_flag g_rc_once;

// little-endian 32-bit limb integers
static void limb_mul_small(std::vector<uint32_t> &a, uint32_t s) {
	uint64_t carry = 0;
	for (uint32_t &w : a) {
		const uint64_t v = (uint64_t)w * s + carry;
		w = (uint32_t)v;
		carry = v >> 32;
	}
	if (carry) a.push_back((uint32_t)carry);
}
// bits [39 c, 39 c + 39) of a limb integer, as a binary64 (exact)
static double limb_chunk(const std::vector<uint32_t> &a, int c) {
	uint64_t v = 0;
	for (int b = 0; b < kResChunkBits; b++) {
		const int bit = kResChunkBits * c + b;
		const size_t q = (size_t)(bit / 32);
		if (q < a.size() && ((a[q] >> (bit % 32)) & 1u)) v |= uint64_t(1) << b;
	}
	return (double)v;
}
static uint32_t limb_mod_small(const std::vector<uint32_t> &a, uint32_t p) {
	uint64_t r = 0;
	for (size_t q = a.size(); q-- > 0;) r = ((r << 32) | a[q]) % p;
	return (uint32_t)r;
}

static void build_res_consts() {
	const std::vector<int> ms = res_moduli();
	const int avail = (int)ms.size();
	for (int nmod = 1; nmod <= kResMaxMod; nmod++) {
		ResConst &c = g_rc[nmod];
		std::memset(&c, 0, sizeof(c));
		if (nmod > avail) continue;
		c.nmod = nmod;
		double lg = 0.0;
		std::vector<uint32_t> M{1};
		for (int t = 0; t < nmod; t++) {
			limb_mul_small(M, (uint32_t)ms[t]);
			lg += std::log2((double)ms[t]);
		}
		g_log2M[nmod] = lg;
		// limbs carried: sum_t c_t M/p_t < nmod M and |X| <= M/4, so bits(nmod M) plus a sign bit
		std::vector<uint32_t> nM = M;
		limb_mul_small(nM, (uint32_t)nmod);
		while (nM.size() > 1 && nM.back() == 0) nM.pop_back();
		const int bitsNM = 32 * ((int)nM.size() - 1) + (32 - __builtin_clz(nM.back()));
		c.nl = (bitsNM + 1 + 31) / 32;
		if (c.nl > kResMaxLimb) c.nl = kResMaxLimb;
		for (size_t q = 0; q < M.size() && q < (size_t)kResMaxLimb; q++) c.M[q] = M[q];
		for (int ch = 0; ch < kResMaxChunk; ch++) c.Md[ch] = limb_chunk(M, ch);
		for (int t = 0; t < nmod; t++) {
			const int p = ms[t];
			c.p[t] = p;
			c.invpf[t] = 1.0f / (float)p;
			c.invpd[t] = 1.0 / (double)p;
			std::vector<uint32_t> Mp{1};
			for (int s = 0; s < nmod; s++)
				if (s != t) limb_mul_small(Mp, (uint32_t)ms[s]);
			for (size_t q = 0; q < Mp.size() && q < (size_t)kResMaxLimb; q++) c.Mp[t][q] = Mp[q];
			for (int ch = 0; ch < kResMaxChunk; ch++) c.Mpd[t][ch] = limb_chunk(Mp, ch);
			const uint32_t w = limb_mod_small(Mp, (uint32_t)p);
			int u = 0;
			for (int v = 1; v < p; v++)
				if ((uint32_t)((w * (uint32_t)v) % (uint32_t)p) == 1u) {
					u = v;
					break;
				}
			c.u[t] = u;
			uint32_t pw = 1 % (uint32_t)p;
			for (int a = 0; a < kResMaxShift / 8; a++) {
				c.pow256[t][a] = (uint8_t)pw;
				pw = (pw * 256u) % (uint32_t)p;
			}
		}
		// groups of three consecutive moduli of the full list (the last one shorter), independent of
		// nmod: a group's members beyond nmod are simply not used
		for (int gi = 0; gi < kResMaxGroups; gi++) {
			uint64_t P = 1;
			for (int t = 3 * gi; t < 3 * gi + 3 && t < avail; t++) P *= (uint64_t)ms[t];
			c.Pg[gi] = (double)P;
			c.invPg[gi] = 1.0 / (double)P;
			// U = u_t mod p_t for the members t < nmod and 1 mod the others: the Chinese remainder
			// sum over the members of v_t (P/p_t) ((P/p_t)^-1 mod p_t)
			uint64_t U = 0;
			for (int t = 3 * gi; t < 3 * gi + 3 && t < avail; t++) {
				const uint64_t pt = (uint64_t)ms[t], Q = P / pt;
				const uint64_t vt = (t < nmod) ? (uint64_t)c.u[t] : 1u;
				uint64_t inv = 0;
				for (uint64_t v = 1; v < pt; v++)
					if ((Q % pt) * v % pt == 1) {
						inv = v;
						break;
					}
				U = (U + vt * Q % P * inv) % P;  // each factor below 2^24: no overflow
			}
			uint64_t pw2 = 1 % P;
			for (int sh = 0; sh < kResMaxShift; sh++) {
				c.pow2g[gi][sh] = (uint32_t)pw2;
				c.pow2gU[gi][sh] = (uint32_t)(pw2 * U % P);
				pw2 = (pw2 * 2) % P;
			}
		}
	}
}

const ResConst &res_const_host(int nmod) {
	std::call_once(g_rc_once, build_res_consts);
	return g_rc[nmod < 1 ? 1 : (nmod > kResMaxMod ? kResMaxMod : nmod)];
}

double res_log2_product(int nmod) {
	std::call_once(g_rc_once, build_res_consts);
	if (nmod < 1) return 0.0;
	if (nmod > kResMaxMod) return 1e300;
	return g_rc[nmod].nmod == nmod ? g_log2M[nmod] : 0.0;
}

// INT8 tiled planes (the layout of the digit planes): one 256-thread workgroup writes 16 columns x
// 64 k-bytes of every modulus plane per step (4 bytes per thread), looping over k tiles.
__global__ __launch_bounds__(256) void k_res_write_i8t(int m, int n, const double *in, int64_t ldi,
                                                       const int *colLo, const ResConst *rc, int nmod,
                                                       int shLimit, int scaleU, int8_t *planes,
                                                       int64_t ldp, int64_t pstride) {
	__shared__ ResLds T;
	res_lds_load(T, rc, nmod, shLimit, scaleU);
	__syncthreads();
	const int ktiles = (int)(ldp / 64);
	for (int kt = blockIdx.x; kt < ktiles; kt += gridDim.x)
		res_write_tile_i8(m, n, in, ldi, colLo, nmod, T, planes, ldp, pstride, kt, blockIdx.y);
}

// 2-byte planes (FP16 / BF16): plane s of column j at planes + s*pstride + j*ldp, rows [m, ldp) zero;
// a workgroup takes columns blockIdx.x, blockIdx.x + gridDim.x, ... over their whole length
template <typename T>
__global__ __launch_bounds__(256) void k_res_write_h(int m, int n, const double *in, int64_t ldi,
                                                     const int *colLo, const ResConst *rc, int nmod,
                                                     int shLimit, int scaleU, T *planes, int64_t ldp,
                                                     int64_t pstride) {
	__shared__ ResLds L;
	res_lds_load(L, rc, nmod, shLimit, scaleU);
	__syncthreads();
	const int ng = (nmod + 2) / 3;
	for (int j = blockIdx.x; j < n; j += gridDim.x) {
		const double *x = in + (int64_t)j * ldi;
		const int lo = colLo[j];
		for (int i = threadIdx.x; i < ldp; i += blockDim.x) {
			const ResElem e = res_elem(i < m ? x[i] : 0.0, lo);
			for (int gi = 0; gi < ng; gi++) {
				const double rx = res_group(e, L.Pg[gi], L.invPg[gi], resPw2 + gi * shLimit);
				const int s1 = min(3 * gi + 3, nmod);
				for (int s = 3 * gi; s < s1; s++)
					planes[(int64_t)s * pstride + (int64_t)j * ldp + i] = digit_cast<T>(res_member(rx, L.pd[s], L.ipd[s]));
			}
		}
	}
}

void launch_res_write(hipStream_t st, int elem, int m, int n, const double *in, int64_t ldi,
                      const int *colLo, const ResConst *rc, int nmod, int shLimit, int scaleU, void *planes,
                      int64_t ldp, int64_t pstride) {
	if (shLimit < 1) shLimit = 1;
	if (elem == 1) {  // tiled layout; ldp % 128 == 0 and pstride = ldp * (n rounded up to 256)
		const int ktiles = (int)(ldp / 64);
		const int gx = ktiles < 16 ? ktiles : 16;
		hipLaunchKernelGGL(k_res_write_i8t, dim3((unsigned)gx, (unsigned)(pstride / ldp / 16)), dim3(256),
		                   res_lds_bytes(nmod, shLimit), st, m, n, in, ldi, colLo, rc, nmod, shLimit, scaleU,
		                   (int8_t *)planes, ldp, pstride);
		return;
	}
	const int g = n < 2048 ? n : 2048;
	if (elem == 2)
		hipLaunchKernelGGL(k_res_write_h<_Float16>, dim3(g), dim3(256), res_lds_bytes(nmod, shLimit), st, m, n,
		                   in, ldi, colLo, rc, nmod, shLimit, scaleU, (_Float16 *)planes, ldp, pstride);
	else
		hipLaunchKernelGGL(k_res_write_h<__bf16>, dim3(g), dim3(256), res_lds_bytes(nmod, shLimit), st, m, n,
		                   in, ldi, colLo, rc, nmod, shLimit, scaleU, (__bf16 *)planes, ldp, pstride);
}

template <int NW>
// (C written only here, the inputs only read: __restrict__ lets the constants come through the
// scalar cache)
__global__ __launch_bounds__(256) void k_res_recon(int M, int N, const uint8_t *__restrict__ R, int64_t ldr,
                                                   int64_t rstride, int nchunk,
                                                   const ResConst *__restrict__ rc, int nmod,
                                                   const int *__restrict__ rowLo, const int *__restrict__ colLo,
                                                   double alpha, double beta, double *__restrict__ C,
                                                   int64_t ldc) {
	// item e = (row e % M, column group e / M): one division for the first item, then the grid
	// stride split the same way once, and the row carried by compare-and-subtract (M * G < 2^31)
	const int G = (N + 3) / 4;
	if ((int64_t)M * G > (int64_t)INT32_MAX - (int64_t)gridDim.x * blockDim.x) {  // 64-bit items
		const int64_t tot = (int64_t)M * G;
		for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
		     e += (int64_t)gridDim.x * blockDim.x)
			res_recon_group<NW>(M, N, R, ldr, rstride, nchunk, rc, nmod, rowLo, colLo, alpha, beta, C, ldc,
			                    (int)(e % M), (int)(e / M));
		return;
	}
	const int total = M * G;
	const int stride = (int)(gridDim.x * blockDim.x);
	const int e0 = (int)(blockIdx.x * blockDim.x + threadIdx.x);
	if (e0 >= total) return;
	int i = e0 % M, g = e0 / M;
	const int si = stride % M, sg = stride / M;
	for (int e = e0; e < total; e += stride) {
		res_recon_group<NW>(M, N, R, ldr, rstride, nchunk, rc, nmod, rowLo, colLo, alpha, beta, C, ldc, i, g);
		i += si;
		g += sg;
		if (i >= M) {
			i -= M;
			g++;
		}
	}
}

void launch_res_recon(hipStream_t st, int M, int N, const uint8_t *R, int64_t ldr, int64_t rstride,
                      int nchunk, const ResConst *rc, int nmod, int nl, const int *rowLo,
                      const int *colLo, double alpha, double beta, double *C, int64_t ldc) {
	int64_t g = ((int64_t)M * ((N + 3) / 4) + 255) / 256;
	if (g > 16384) g = 16384;
	const int nw = (nl + 1) / 2;
#define OZ_RES_RECON(NWV)                                                                              \
	hipLaunchKernelGGL(k_res_recon<NWV>, dim3((unsigned)g), dim3(256), 0, st, M, N, R, ldr, rstride, nchunk, \
	                   rc, nmod, rowLo, colLo, alpha, beta, C, ldc)
	if (nw <= 3)
		OZ_RES_RECON(3);
	else if (nw <= 4)
		OZ_RES_RECON(4);
	else if (nw <= 6)
		OZ_RES_RECON(6);
	else
		OZ_RES_RECON(9);
#undef OZ_RES_RECON
}

}  // namespace oz

// test hooks: the host constants for nmod moduli (copied into out when it holds sizeof(ResConst)
// bytes; returns sizeof(ResConst)) and log2 of their product
extern "C" int32_t ozblas_internal_res_const(int32_t nmod, void *out, uint64_t bytes) {
	const oz::ResConst &c = oz::res_const_host(nmod);
	if (out && bytes >= sizeof(oz::ResConst)) std::memcpy(out, &c, sizeof(oz::ResConst));
	return (int32_t)sizeof(oz::ResConst);
}

extern "C" double ozblas_internal_res_log2(int32_t nmod) { return oz::res_log2_product(nmod); }
