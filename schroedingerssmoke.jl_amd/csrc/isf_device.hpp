// isf_device.hpp -- device helpers shared by the step kernels (isf_kernels.hip) and
// their adjoints (isf_adjoint.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "isf_internal.hpp"

namespace isf {

// angle(re + i im) = atan2(im, re) (Julia's angle, src/parallel.jl:108).  Range reduction
// to a = min/max in [0, 1], atan(a) = a P(a^2) with a degree-8 polynomial fitted for this
// project (weighted least squares driven to equal ripple; |relative error| <= 1.5e-7 on
// [0, 1] in FP32 Horner evaluation), then the octant is restored.  angle(0) = 0 and the
// sign of a zero imaginary part is kept, as atan2 does.
__device__ __forceinline__ float angle_of(float re, float im) {
    const float ax = fabsf(re), ay = fabsf(im);
    // a = mn / mx with the divisor clamped to the smallest normal float: the same value for
    // every normal mx, and 0 (not 0 * inf) when both parts are zero
    const float mx = fmaxf(fmaxf(ax, ay), 1.17549435e-38f), mn = fminf(ax, ay);
    const float a = mn * __builtin_amdgcn_rcpf(mx);
    const float s = a * a;
    float p = 2.903499175e-03f;
    p = fmaf(p, s, -1.628280245e-02f);
    p = fmaf(p, s, 4.303904623e-02f);
    p = fmaf(p, s, -7.533650100e-02f);
    p = fmaf(p, s, 1.065466627e-01f);
    p = fmaf(p, s, -1.420713067e-01f);
    p = fmaf(p, s, 1.999305338e-01f);
    p = fmaf(p, s, -3.333309293e-01f);
    p = fmaf(p, s, 1.0f);
    float r = p * a;
    if (ay > ax) r = 1.57079632679489661923f - r;
    if (__float_as_uint(re) >> 31) r = 3.14159265358979323846f - r;  // re < 0 or re = -0
    return copysignf(r, im);
}

// angle(conj(a1) b1 + conj(a2) b2) for psi cells a, b (src/parallel.jl:108-111)
__device__ __forceinline__ float edge_angle(float4 a, float4 b) {
    const float re = (a.x * b.x + a.y * b.y) + (a.z * b.z + a.w * b.w);
    const float im = (a.x * b.y - a.y * b.x) + (a.z * b.w - a.w * b.z);
    return angle_of(re, im);
}

// isjet, src/particles.jl:6-10, at position (i-1)*d (src/parallel.jl:75-77).  Rounded
// operations (no contraction) so the predicate is the one the Float32 reference evaluates.
__device__ __forceinline__ bool is_jet_cell(int x, int y, int z, const GridInfo& g, const JetParams& j) {
    const float px = __fmul_rn((float)x, g.dx), py = __fmul_rn((float)y, g.dy), pz = __fmul_rn((float)z, g.dz);
    const float ax = fabsf(__fsub_rn(px, j.cx));
    const float dy = __fsub_rn(py, j.cy), dz = __fsub_rn(pz, j.cz);
    const float r2 = __fadd_rn(__fmul_rn(dy, dy), __fmul_rn(dz, dz));
    return (ax <= j.len * 0.5f) & (r2 <= __fmul_rn(j.rad, j.rad));
}

// restrict_kernel inside the jet, src/particles.jl:12-20: psi_a <- |psi_a| exp(i phase)
__device__ __forceinline__ float4 jet_phase(float4 p, int x, int y, int z, const GridInfo& g, const JetParams& j) {
    const float px = __fmul_rn((float)x, g.dx), py = __fmul_rn((float)y, g.dy), pz = __fmul_rn((float)z, g.dz);
    float ph;
    if (j.phase_mode == 1)
        ph = __fadd_rn(__fadd_rn(__fadd_rn(j.kx, px), __fadd_rn(j.ky, py)), __fadd_rn(j.kz, pz));
    else
        ph = __fadd_rn(__fadd_rn(__fmul_rn(j.kx, px), __fmul_rn(j.ky, py)), __fmul_rn(j.kz, pz));
    ph = __fsub_rn(ph, j.omgterm);
    float s, c;
    sincosf(ph, &s, &c);
    const float a1 = hypotf(p.x, p.y), a2 = hypotf(p.z, p.w);
    return make_float4(a1 * c, a1 * s, a2 * c, a2 * s);
}

// normalize_psi (src/parallel.jl:162-165): psi_a / hypot(|psi1|, |psi2|).  |psi| is O(1)
// here, so the norm is formed as sqrt(sum of squares) (the reference uses @fastmath
// hypot) and applied as one reciprocal.
__device__ __forceinline__ float4 normalized(float4 p) {
    // rounded operations: the same value wherever the compiler inlines this (the fused
    // inverse x pass and the standalone kernels of the differentiable rollout)
    const float ss = __fadd_rn(__fadd_rn(__fmul_rn(p.x, p.x), __fmul_rn(p.y, p.y)),
                               __fadd_rn(__fmul_rn(p.z, p.z), __fmul_rn(p.w, p.w)));
    const float inv = 1.0f / sqrtf(ss);
    return make_float4(p.x * inv, p.y * inv, p.z * inv, p.w * inv);
}

// block-wide sum of a double (NT a power of two); result valid in thread 0.
// Fixed tree order, so the sum is deterministic.
template <int NT>
__device__ __forceinline__ double block_sum(double v) {
    __shared__ double red[NT];
    red[threadIdx.x] = v;
    __syncthreads();
#pragma unroll
    for (int s = NT / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    return red[0];
}

__device__ __forceinline__ float2* poisson_P(const Inst& m) {
    return reinterpret_cast<float2*>(reinterpret_cast<char*>(m.S) + 4 * (size_t)m.g.nx * m.g.ny * m.g.nz);
}

__device__ __forceinline__ void cell_xyz(size_t c, const GridInfo& g, int& x, int& y, int& z) {
    x = (int)(c % g.nx);
    const size_t r = c / g.nx;
    y = (int)(r % g.ny);
    z = (int)(r / g.ny);
}

// Julia mod(x, y) for floats (result has the sign of y), src/parallel.jl:220:
// r = rem(x, y); r != 0 && sign(r) != sign(y) ? r + y : r.  For y > 0 and the ranges
// a particle position takes, rem is exact and simple: x in [0, y) -> x;
// x in (-y, 0) -> x + y; x in [y, 2y) -> x - y (exact by Sterbenz); otherwise fmodf.
__device__ __forceinline__ float jl_mod(float x, float y) {
    if (x >= 0.0f && x < y) return x;
    if (x < 0.0f && x > -y) return x + y;
    if (x >= y && x < 2.0f * y) return x - y;
    float r = fmodf(x, y);
    if (r != 0.0f && ((r < 0.0f) != (y < 0.0f))) r += y;
    return r;
}

struct F3 {
    float x, y, z;
};
// three consecutive floats (the compiler merges them into one dwordx3 load)
__device__ __forceinline__ F3 ld3(const float* __restrict__ p) { return F3{p[0], p[1], p[2]}; }

// Cell indices and weights of staggered_velocity at a particle position (src/parallel.jl:
// 219-238): p = mod(p_in, L), i = floor(p/d) (0-based; an index that mod() rounds onto
// the upper face wraps to the first cell), ip = i + 1 (periodic), w = p - i d (not
// divided by d, :238).
__device__ __forceinline__ void stag_index(const float p_in[3], const GridInfo& g, int i[3], int ip[3], float w[3]) {
    const float d[3] = {g.dx, g.dy, g.dz};
    const float id[3] = {g.idx, g.idy, g.idz};
    const float L[3] = {g.Lx, g.Ly, g.Lz};
    const int n[3] = {g.nx, g.ny, g.nz};
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const float p = jl_mod(p_in[a], L[a]);
        int ia = (int)floorf(g.d_pow2 ? p * id[a] : p / d[a]);
        if (ia >= n[a]) ia -= n[a];
        i[a] = ia;
        ip[a] = ia + 1 == n[a] ? 0 : ia + 1;
        w[a] = p - (float)ia * d[a];
    }
}

// staggered_velocity, src/parallel.jl:219-246.  Weights are not divided by d (:238).
// An index that mod() rounds onto the upper face wraps to the first cell.
__device__ __forceinline__ void stag_vel(const float* __restrict__ vel, const float p_in[3], const GridInfo& g, float u[3]) {
    int i[3], ip[3];
    float w[3];
    stag_index(p_in, g, i, ip, w);
    // the 7 cells as 4 rows: (j,k) and (j+1,k), (j,k+1) at x = i and i+1, (j+1,k+1) at i
    const unsigned sy = (unsigned)g.nx, sz = (unsigned)g.nx * (unsigned)g.ny;
    const unsigned rjk = sy * i[1] + sz * i[2], rjpk = sy * ip[1] + sz * i[2];
    const unsigned rjkp = sy * i[1] + sz * ip[2], rjpkp = sy * ip[1] + sz * ip[2];
    const F3 a0 = ld3(vel + 3u * (i[0] + rjk)), a1 = ld3(vel + 3u * (ip[0] + rjk));
    const F3 b0 = ld3(vel + 3u * (i[0] + rjpk)), b1 = ld3(vel + 3u * (ip[0] + rjpk));
    const F3 c0 = ld3(vel + 3u * (i[0] + rjkp)), c1 = ld3(vel + 3u * (ip[0] + rjkp));
    const float d0x = vel[3u * (i[0] + rjpkp)];
    const float v0[3] = {a0.x, a0.y, a0.z};   // velocity[i, j, k]
    const float vn[3] = {d0x, c1.y, b1.z};    // (v[i,j+1,k+1].x, v[i+1,j,k+1].y, v[i+1,j+1,k].z)
    const float pp[3] = {b0.x, a1.y, a1.z};   // (pyp.x, pxp.y, pxp.z)
    const float pp2[3] = {c0.x, c0.y, b0.z};  // (pzp.x, pzp.y, pyp.z)
    const float w1[3] = {w[2], w[2], w[1]}, w2[3] = {w[1], w[0], w[0]};
#pragma unroll
    for (int a = 0; a < 3; a++)
        u[a] = (1.0f - w1[a]) * ((1.0f - w2[a]) * v0[a] + w2[a] * pp[a]) + w1[a] * ((1.0f - w2[a]) * pp2[a] + w2[a] * vn[a]);
}

// StaggeredVelocity on separate component arrays vx, vy, vz (src/MatlabPort.jl:237-273):
// the same cells, weights and arithmetic as stag_vel, so both layouts give equal results.
__device__ __forceinline__ void stag_vel_split(const float* __restrict__ vx, const float* __restrict__ vy,
                                               const float* __restrict__ vz, const float p_in[3], const GridInfo& g,
                                               float u[3]) {
    int i[3], ip[3];
    float w[3];
    stag_index(p_in, g, i, ip, w);
    const unsigned sy = (unsigned)g.nx, sz = (unsigned)g.nx * (unsigned)g.ny;
    const unsigned rjk = sy * i[1] + sz * i[2], rjpk = sy * ip[1] + sz * i[2];
    const unsigned rjkp = sy * i[1] + sz * ip[2], rjpkp = sy * ip[1] + sz * ip[2];
    const float v0[3] = {vx[i[0] + rjk], vy[i[0] + rjk], vz[i[0] + rjk]};
    const float vn[3] = {vx[i[0] + rjpkp], vy[ip[0] + rjkp], vz[ip[0] + rjpk]};
    const float pp[3] = {vx[i[0] + rjpk], vy[ip[0] + rjk], vz[ip[0] + rjk]};
    const float pp2[3] = {vx[i[0] + rjkp], vy[i[0] + rjkp], vz[i[0] + rjpk]};
    const float w1[3] = {w[2], w[2], w[1]}, w2[3] = {w[1], w[0], w[0]};
#pragma unroll
    for (int a = 0; a < 3; a++)
        u[a] = (1.0f - w1[a]) * ((1.0f - w2[a]) * v0[a] + w2[a] * pp[a]) + w1[a] * ((1.0f - w2[a]) * pp2[a] + w2[a] * vn[a]);
}

}  // namespace isf
