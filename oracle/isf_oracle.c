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
#include <stdlib.h>
#include <string.h>
#include "isf_philox.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

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
