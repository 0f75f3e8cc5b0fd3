/*
 * ozoracle_int.h -- helpers shared by the binary64 (ozoracle.c) and binary32 (ozoracle_s.c) oracle
 * restatements (TEST INFRASTRUCTURE ONLY).
 */
#ifndef OZORACLE_INT_H
#define OZORACLE_INT_H

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ozoracle.h"

#define NUM_SPLIT_DEFAULT_MAX 30 /* src/cpu/ozblas_common.h:5 */
#define SPLIT_CONST 0.75         /* src/cpu/ozblas_split.cpp:9 */

static inline double std_max(double a, double b) { return (a < b) ? b : a; } /* std::max */
static inline double mac_max(double a, double b) { return (a > b) ? a : b; } /* MAX macro */
static inline int32_t imin(int32_t a, int32_t b) { return a < b ? a : b; }
static inline int32_t imax(int32_t a, int32_t b) { return a > b ? a : b; }

/* x86-64 cvttsd2si semantics of the reference's implicit double->int conversions */
static inline int32_t d2i(double x) {
	if (!(x >= -2147483648.0 && x < 2147483648.0)) return INT32_MIN;
	return (int32_t)x;
}
static inline double scalbn_l(double v, int64_t e) {
	if (e > 100000) e = 100000;
	if (e < -100000) e = -100000;
	return scalbn(v, (int)e);
}


static inline void counter_init(ozo_handle *h) {
	h->splitShift = 1;
	h->nSplitA = h->nSplitB = h->nSplitC = 0;
	h->mbk = h->nbk = 0;
	h->memAddr = h->memMaskSplitA;
}

static inline uint64_t mat_alloc(ozo_handle *h, int32_t m, int32_t n, int32_t size, int32_t *ld) {
	*ld = ozo_pitch(m);
	uint64_t off = h->memAddr;
	h->memAddr += (uint64_t)size * (uint64_t)(*ld) * (uint64_t)n;
	return off;
}
static inline uint64_t vec_alloc(ozo_handle *h, int32_t n, int32_t size) {
	int32_t ld = ozo_pitch(n);
	uint64_t off = h->memAddr;
	h->memAddr += (uint64_t)size * (uint64_t)ld;
	return off;
}
static inline int mem_check(ozo_handle *h) { return h->memAddr > h->workSizeBytes; }
static inline void arena_check(ozo_handle *h) {
	if (h->memAddr > h->workCap) {
		fprintf(stderr, "ozoracle: workspace too small (%llu > %llu)\n",
		        (unsigned long long)h->memAddr, (unsigned long long)h->workCap);
		exit(1);
	}
}
#define WP(h, off, T) ((T *)((h)->work + (off)))

static inline int is_n(char t) { return t == 'N' || t == 'n'; }

#endif
