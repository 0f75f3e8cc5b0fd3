// oz_handle.cpp -- handle lifecycle, workspace and the libm ceil(log2) table.
//
// Mirrors cuozblasCreate/cuozblasDestroy (reference src/cuda/cuozblas.cpp:23-75) and the defaults
// of src/cpu/ozblas.cpp:31-49; the workspace is one hipMalloc'd region with a bump allocator reset
// on every call (counterInit, src/cpu/ozblas.cpp:3-20).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>

#include "oz_digits.hpp"
#include "oz_internal.hpp"

namespace oz {

static uint16_t g_tau[kTauBinades];
static std::once_flag g_tau_once;

static void build_tau_table() {
	for (int e = -1074; e <= 1023; e++) {
		const double base = ldexp(1.0, e);
		const double u = ldexp(1.0, e - 52 > -1074 ? e - 52 : -1074);
		// number of representable values strictly above 2^e inside the binade
		const uint64_t nvals = (e - 52 >= -1074) ? ((uint64_t(1) << 52) - 1)
		                                         : ((uint64_t(1) << (e + 1074)) - 1);
		uint64_t i = 1;
		for (; i <= nvals && i < 65535; i++) {
			const double x = base + (double)i * u;
			if (ceil(log2(x)) != (double)e) break;
		}
		g_tau[e + kTauBias] = (uint16_t)(i - 1);
	}
}

const uint16_t *tau_table_host() {
	std::call_once(g_tau_once, build_tau_table);
	return g_tau;
}

// binary32 tables: for each binade 2^e the number of floats above 2^e for which the rule gives e
static uint16_t g_tauf[kTauBinadesF], g_tauf2[kTauBinadesF];
static std::once_flag g_tauf_once;

static void build_tau_tables_f() {
	for (int e = -149; e <= 127; e++) {
		const float base = ldexpf(1.0f, e);
		const float u = ldexpf(1.0f, e - 23 > -149 ? e - 23 : -149);
		const uint32_t nvals = (e - 23 >= -149) ? ((1u << 23) - 1) : ((1u << (e + 149)) - 1);
		uint32_t i = 1;
		for (; i <= nvals && i < 65535; i++) {  // split rule: (float)log2((double)x)
			const float x = base + (float)i * u;
			if (ceil((double)(float)log2((double)x)) != (double)e) break;
		}
		g_tauf[e + kTauBiasF] = (uint16_t)(i - 1);
		for (i = 1; i <= nvals && i < 65535; i++) {  // NearSum rule: log2f(x)
			const float x = base + (float)i * u;
			if (ceil((double)log2f(x)) != (double)e) break;
		}
		g_tauf2[e + kTauBiasF] = (uint16_t)(i - 1);
	}
}

const uint16_t *tau_table_host_f() {
	std::call_once(g_tauf_once, build_tau_tables_f);
	return g_tauf;
}
const uint16_t *tau_table_host_f2() {
	std::call_once(g_tauf_once, build_tau_tables_f);
	return g_tauf2;
}

// getRho<double,float> (src/cpu/ozblas_split.cpp:20-34 with getEpse<double> = 53,
// getEpse<float> = 24, getEpse2<float> = 53)
int rho_host_ds(int dim, int splitEpsMode) {
	double r;
	switch (splitEpsMode) {
	case 2: r = ceil((53 - (53 - log2(1. * dim)) / 2)); break;
	case 9: r = ceil((53 - (24 - log2(2. * sqrt((double)dim))) / 2)); break;
	default: r = ceil((53 - (24 - log2(1. * dim)) / 2)); break;
	}
	if (!(r >= -2147483648.0 && r < 2147483648.0)) return INT32_MIN;
	return (int)r;
}

// getRho<float,double> (src/cpu/ozblas_split.cpp:20-34 with getEpse<float> = 24,
// getEpse<double> = 53, getEpse2<double> = 106)
int rho_host_sd(int dim, int splitEpsMode) {
	double r;
	switch (splitEpsMode) {
	case 2: r = ceil((24 - (106 - log2(1. * dim)) / 2)); break;
	case 9: r = ceil((24 - (53 - log2(2. * sqrt((double)dim))) / 2)); break;
	default: r = ceil((24 - (53 - log2(1. * dim)) / 2)); break;
	}
	if (!(r >= -2147483648.0 && r < 2147483648.0)) return INT32_MIN;
	return (int)r;
}

// getRho<float,float> (src/cpu/ozblas_split.cpp:20-34 with getEpse = 24, getEpse2 = 53)
int rho_host_f(int dim, int splitEpsMode) {
	double r;
	switch (splitEpsMode) {
	case 2: r = ceil((24 - (53 - log2(1. * dim)) / 2)); break;
	case 9: r = ceil((24 - (24 - log2(2. * sqrt((double)dim))) / 2)); break;
	default: r = ceil((24 - (24 - log2(1. * dim)) / 2)); break;
	}
	if (!(r >= -2147483648.0 && r < 2147483648.0)) return INT32_MIN;
	return (int)r;
}

// getRho<double,double> (src/cpu/ozblas_split.cpp:20-34), evaluated with the host libm
int rho_host(int dim, int splitEpsMode) {
	double r;
	switch (splitEpsMode) {
	case 2: r = ceil((53 - (106 - log2(1. * dim)) / 2)); break;
	case 9: r = ceil((53 - (53 - log2(2. * sqrt((double)dim))) / 2)); break;
	default: r = ceil((53 - (53 - log2(1. * dim)) / 2)); break;
	}
	if (!(r >= -2147483648.0 && r < 2147483648.0)) return INT32_MIN;
	return (int)r;
}

// Device memory of ozblasDdotNoAlloc: one copy per GPU, at its own address on each.  Host code
// never passes `scratch` itself to a kernel (in host code the name is a host-side stand-in, not a
// device address); it uses the address hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) char scratch[kStaticScratchBytes];
static std::mutex g_scratch_mu;
static char *g_scratch_dev[64];

char *static_scratch(int device) {
	if (device < 0 || device >= 64) return nullptr;
	std::lock_guard<std::mutex> lk(g_scratch_mu);
	if (g_scratch_dev[device] == nullptr) {
		void *ptr = nullptr;
		const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
		if (e != hipSuccess || ptr == nullptr) {
			fprintf(stderr, "ozblas_mi355x: hipGetSymbolAddress(scratch) failed on device %d: %s\n",
			        device, hipGetErrorString(e));
			return nullptr;
		}
		g_scratch_dev[device] = (char *)ptr;
	}
	return g_scratch_dev[device];
}

int32_t ensure_init(ozblasHandle_t *oh) {
	if (oh == nullptr || oh->initialized != 1 || oh->backend == nullptr)
		return OZBLAS_STATUS_NOT_INITIALIZED;
	return OZBLAS_STATUS_SUCCESS;
}

}  // namespace oz

using namespace oz;

extern "C" int32_t ozblasCreate(ozblasHandle_t *oh, uint64_t workSizeBytes) {
	if (oh == nullptr) return OZBLAS_STATUS_INVALID_VALUE;
	std::memset(oh, 0, sizeof(*oh));
	Backend *b = new Backend();
	OZ_HIP(hipGetDevice(&b->device));
	if (workSizeBytes < (1u << 20)) workSizeBytes = 1u << 20;
	if (hipMalloc(&b->work, workSizeBytes) != hipSuccess) {
		delete b;
		fprintf(stderr, "OzBLAS error: work memory allocation error (%1.3e Bytes requested).\n",
		        (double)workSizeBytes);
		return OZBLAS_STATUS_ALLOC_FAILED;
	}
	b->cap = workSizeBytes;
	OZ_HIP(hipMalloc(&b->dTau, sizeof(uint16_t) * kTauBinades));
	OZ_HIP(hipMemcpy(b->dTau, tau_table_host(), sizeof(uint16_t) * kTauBinades,
	                 hipMemcpyHostToDevice));
	OZ_HIP(hipMalloc(&b->dTauF, sizeof(uint16_t) * kTauBinadesF));
	OZ_HIP(hipMemcpy(b->dTauF, tau_table_host_f(), sizeof(uint16_t) * kTauBinadesF,
	                 hipMemcpyHostToDevice));
	OZ_HIP(hipMalloc(&b->dTauF2, sizeof(uint16_t) * kTauBinadesF));
	OZ_HIP(hipMemcpy(b->dTauF2, tau_table_host_f2(), sizeof(uint16_t) * kTauBinadesF,
	                 hipMemcpyHostToDevice));
	OZ_HIP(hipHostMalloc(&b->hPinned, 64 * sizeof(int64_t), hipHostMallocDefault));
	OZ_HIP(hipMalloc(&b->dCounters, 256 * sizeof(int32_t)));
	for (auto &e : b->ev) OZ_HIP(hipEventCreate(&e));
	b->scratch = static_scratch(b->device);  // with the handle's device current
	oh->backend = b;
	oh->stream = nullptr;
	oh->devWork = b->work;
	oh->workSizeBytes = workSizeBytes;
	// defaults: src/cpu/ozblas.cpp:31-49
	oh->nSplitMax = 0;
	oh->memMaskSplitA = 0;
	oh->splitShift = 1;
	oh->splitMode = 0;
	oh->fastMode = 0;
	oh->reproMode = 1;
	oh->sumMode = 0;
	oh->useBatchedGemmFlag = 0;
	oh->splitEpsMode = 0;
	oh->precxFlag = 0;
	oh->sliceType = OZBLAS_SLICE_FP64;
	oh->initialized = 1;
	return OZBLAS_STATUS_SUCCESS;
}

extern "C" int32_t ozblasDestroy(ozblasHandle_t *oh) {
	if (oh == nullptr || oh->backend == nullptr) return OZBLAS_STATUS_NOT_INITIALIZED;
	Backend *b = be(oh);
	if (b->stream) (void)hipStreamSynchronize(b->stream);
	else (void)hipDeviceSynchronize();
	batch_plan_free(b->batch);
	b->batch = nullptr;
	for (auto &e : b->ev)
		if (e) (void)hipEventDestroy(e);
	(void)hipFree(b->work);
	(void)hipFree(b->dTau);
	(void)hipFree(b->dTauF);
	(void)hipFree(b->dTauF2);
	(void)hipFree(b->dCounters);
	(void)hipHostFree(b->hPinned);
	delete b;
	oh->backend = nullptr;
	oh->devWork = nullptr;
	oh->initialized = 0;
	return OZBLAS_STATUS_SUCCESS;
}

extern "C" int32_t ozblasSetStream(ozblasHandle_t *oh, void *stream) {
	OZ_TRY(ensure_init(oh));
	be(oh)->stream = (hipStream_t)stream;
	oh->stream = stream;
	return OZBLAS_STATUS_SUCCESS;
}

extern "C" const char *ozblasGetVersionString(void) {
	return "ozblas_mi355x 0.1 (Ozaki-scheme BLAS for gfx950; OzBLAS 1.6a semantics)";
}

// Harness input generator (reference testing/testing_common.cpp:703-717, phi-mode): srand48(seed),
// then per element r1, r2 = drand48(), x1 = sqrt(-2 log r1) sin(2 pi r2), a = (r1 - 0.5) exp(phi x1),
// column by column.  drand48 is the POSIX 48-bit LCG (a = 0x5DEECE66D, c = 11, seed in bits 16..47
// over 0x330E); log/sin/exp/sqrt are the host libm's, as in the harness.
extern "C" void ozblasGenPhi(int32_t m, int32_t n, int32_t ld, double *A, double phi, int64_t seed) {
	uint64_t x = (((uint64_t)(uint32_t)seed) << 16) | 0x330EULL;
	auto drand = [&x]() {
		x = (0x5DEECE66DULL * x + 0xBULL) & 0xFFFFFFFFFFFFULL;
		return ldexp((double)x, -48);
	};
	if (ld < m) ld = m;
	for (int32_t j = 0; j < n; j++)
		for (int32_t i = 0; i < m; i++) {
			const double r1 = drand();
			const double r2 = drand();
			const double x1 = 0.0 + (sqrt(-2. * log(r1)) * sin(2. * M_PI * r2)) * 1.0;
			A[(int64_t)j * ld + i] = (r1 - 0.5) * exp(phi * x1);
		}
}

// test hooks (not part of the reference API): the tau rule and rho evaluated on the host with the
// same code the kernels run
extern "C" int32_t ozblas_internal_tau(double max) { return tau_of(max, tau_table_host()); }
extern "C" int32_t ozblas_internal_rho(int32_t dim, int32_t splitEpsMode) {
	return rho_host(dim, splitEpsMode);
}
extern "C" int32_t ozblas_internal_rho_row(int32_t dim) { return rho_of_dim(dim); }
// binary32 rules: tau of the split, the NearSum exponent rule, rho (libm) and its integer form
extern "C" int32_t ozblas_internal_tau_f(float max) { return tau_of_f(max, tau_table_host_f()); }
extern "C" int32_t ozblas_internal_nearsum_exp_f(float mu) {
	return ceil_log2_tab_f(mu, tau_table_host_f2());
}
extern "C" int32_t ozblas_internal_rho_f(int32_t dim, int32_t mode) { return rho_host_f(dim, mode); }
extern "C" int32_t ozblas_internal_rho_row_f(int32_t dim, int32_t mode) { return rho_dev_f(dim, mode); }
extern "C" int32_t ozblas_internal_rho_row_mode(int32_t dim, int32_t mode) { return rho_dev(dim, mode); }
extern "C" int32_t ozblas_internal_rho_ds(int32_t dim, int32_t mode) { return rho_host_ds(dim, mode); }
extern "C" int32_t ozblas_internal_rho_row_ds(int32_t dim, int32_t mode) { return rho_dev_ds(dim, mode); }
extern "C" int32_t ozblas_internal_rho_sd(int32_t dim, int32_t mode) { return rho_host_sd(dim, mode); }
extern "C" float ozblas_internal_f32_term(double p) { return f32_term(p); }
extern "C" int32_t ozblas_internal_rho_row_sd(int32_t dim, int32_t mode) { return rho_dev_sd(dim, mode); }
extern "C" void ozblas_internal_tau_batch(int64_t n, const double *x, int32_t *out) {
	const uint16_t *J = tau_table_host();
	for (int64_t i = 0; i < n; i++) out[i] = tau_of(x[i], J);
}

// test hook: the exact reconstruction of one element (host build of recon_elem)
extern "C" double ozblas_internal_recon(int32_t nbins, const int32_t *T, const int32_t *levels,
                                        int32_t Lmax, int32_t e0, double alpha, double beta, double c) {
	const int nw = (8 * Lmax + 32 + 8 + 1 + 63) / 64;
	if (nw <= 3) return oz::recon_elem<3>(T, 1, levels, nbins, Lmax, e0, alpha, beta, c);
	if (nw <= 4) return oz::recon_elem<4>(T, 1, levels, nbins, Lmax, e0, alpha, beta, c);
	if (nw <= 6) return oz::recon_elem<6>(T, 1, levels, nbins, Lmax, e0, alpha, beta, c);
	return oz::recon_elem<9>(T, 1, levels, nbins, Lmax, e0, alpha, beta, c);
}

// test hook: the certified rounding of one element from its leading levels: returns 1 and *t when
// RN((N1 +- R) 2^e0) agree, else 0 (N1 from nb1 bins, R given as 9 little-endian words)
extern "C" int32_t ozblas_internal_recon_certify(int32_t nb1, const int32_t *T, const int32_t *levels,
                                                 int32_t Lmax, int32_t e0, const uint64_t *R, double *t) {
	oz::BigInt<9> a;
	oz::big_zero<9>(a);
	oz::recon_accumulate<9>(a, T, 1, levels, 0, nb1, Lmax);
	return oz::recon_certify<9>(a, R, e0, t) ? 1 : 0;
}
