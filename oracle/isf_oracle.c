/*
 * isf_oracle.c -- CPU restatement of the SchroedingersSmoke.jl ISF step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the oracle the HIP path is checked against and the
 * CPU baseline that bench.py times beside it.  The product library never links, loads
 * or calls it, and it is never the thing measured as the GPU result.
 *
 * Built as oracle/_build/libisf_oracle.so by oracle/Makefile.  Exports every function
 * twice: suffix _f32 (Float32 arithmetic, the reference's Array{...Float32} path of
 * test/runtests.jl:14-16) and _f64 (Float64 authority).  The function bodies live in
 * isf_oracle_impl.h.
 */
#define _GNU_SOURCE
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>
#include "isf_philox.h"

/* Optional replacement of the 3-D FFT (isfo_fft3_*): the timed CPU baseline installs a
 * library FFT here (MKL through torch.fft); z, n and inverse as for isfo_fft3, dbl = 1
 * for the Float64 instance.  NULL: the radix-2 restatement below. */
typedef void (*isfo_fft_hook_t)(void* z, const int* n, int inverse, int dbl);
static isfo_fft_hook_t isfo_fft_hook = NULL;
void isfo_set_fft_hook(isfo_fft_hook_t f) { isfo_fft_hook = f; }
/* OpenMP threads of the loops (and of the radix-2 FFT) */
void isfo_set_threads(int t) { omp_set_num_threads(t > 0 ? t : 1); }
int isfo_get_threads(void) { return omp_get_max_threads(); }

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* Slot order of the nozzle fill: draw indices sorted by the 24-bit key
 * (x: top 5 bits of word 0, theta: top 7 bits of word 2, r: top 12 bits of word 1) of
 * their raw Philox words, ties by index. */
typedef struct { unsigned key; long long i; } isf_keyed;
static int isf_keyed_cmp(const void* a, const void* b) {
    const isf_keyed* x = (const isf_keyed*)a;
    const isf_keyed* y = (const isf_keyed*)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return x->i < y->i ? -1 : (x->i > y->i);
}
static void isf_fill_order(long long n, unsigned long long seed, long long* order) {
    isf_keyed* k = (isf_keyed*)malloc(sizeof(isf_keyed) * (size_t)(n > 0 ? n : 1));
    for (long long i = 0; i < n; i++) {
        unsigned ctr[4] = {(unsigned)i, 0xFFFFFFFFu, 1u, 0u};
        unsigned key[2] = {(unsigned)seed, (unsigned)(seed >> 32)};
        unsigned o[4];
        isf_philox4x32_10(ctr, key, o);
        k[i].key = ((o[0] >> 27) << 19) | ((o[2] >> 25) << 12) | (o[1] >> 20);
        k[i].i = i;
    }
    qsort(k, (size_t)n, sizeof(isf_keyed), isf_keyed_cmp);
    for (long long i = 0; i < n; i++) order[i] = k[i].i;
    free(k);
}

/* Rank order of a spawn chunk: draws c0 .. c0+m-1 (m <= 1024) sorted by
 * (top 22 bits of raw Philox word 0) << 10 | (index within the chunk). */
static int isf_u32_cmp(const void* a, const void* b) {
    const unsigned x = *(const unsigned*)a, y = *(const unsigned*)b;
    return x < y ? -1 : (x > y);
}
static void isf_spawn_order(unsigned c0, int m, unsigned step, unsigned long long seed, int* rank_to_draw) {
    unsigned keys[1024];
    for (int t = 0; t < m; t++) {
        unsigned ctr[4] = {c0 + (unsigned)t, step, 0u, 0u};
        unsigned key[2] = {(unsigned)seed, (unsigned)(seed >> 32)};
        unsigned o[4];
        isf_philox4x32_10(ctr, key, o);
        keys[t] = ((o[0] >> 10) << 10) | (unsigned)t;
    }
    qsort(keys, (size_t)m, sizeof(unsigned), isf_u32_cmp);
    for (int k = 0; k < m; k++) rank_to_draw[k] = (int)(keys[k] & 1023u);
}

/* ---- Float32 instance */
#define R float
#define SUF _f32
static inline float r_fmod_f32(float x, float y) { return fmodf(x, y); }
static inline float r_hypot_f32(float x, float y) { return hypotf(x, y); }
static inline float r_atan2_f32(float y, float x) { return atan2f(y, x); }
static inline float r_exp_f32(float x) { return expf(x); }
static inline float r_cos_f32(float x) { return cosf(x); }
static inline float r_sin2_f32(float x) { return sinf(x); }
static inline float r_floor_f32(float x) { return floorf(x); }
static inline float r_fabs_f32(float x) { return fabsf(x); }
#include "isf_oracle_impl.h"
#undef R
#undef SUF

/* ---- Float64 instance */
#define R double
#define SUF _f64
static inline double r_fmod_f64(double x, double y) { return fmod(x, y); }
static inline double r_hypot_f64(double x, double y) { return hypot(x, y); }
static inline double r_atan2_f64(double y, double x) { return atan2(y, x); }
static inline double r_exp_f64(double x) { return exp(x); }
static inline double r_cos_f64(double x) { return cos(x); }
static inline double r_sin2_f64(double x) { return sin(x); }
static inline double r_floor_f64(double x) { return floor(x); }
static inline double r_fabs_f64(double x) { return fabs(x); }
#include "isf_oracle_impl.h"
#undef R
#undef SUF

/* Philox4x32-10 block, exported for known-answer tests. */
void isfo_philox4x32_10(const unsigned* ctr, const unsigned* key, unsigned* out) {
    isf_philox4x32_10(ctr, key, out);
}

void isfo_philox_uniform4(unsigned a, unsigned b, unsigned c, unsigned long long seed, float* u) {
    isf_philox_uniform4(a, b, c, seed, u);
}
