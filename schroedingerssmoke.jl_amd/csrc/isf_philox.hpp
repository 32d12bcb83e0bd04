// isf_philox.hpp -- Philox4x32-10 counter-based generator (host + device).
//
// Published method: J. K. Salmon, M. A. Moraes, R. O. Dror, D. E. Shaw, "Parallel random
// numbers: as easy as 1, 2, 3", SC'11.  It replaces the reference's unseeded
// rand(Float32) / gpu_rand draws of the particle spawner (src/particles.jl:37-42,78-83)
// so that every implementation sees the same seeded spawn positions.
// Uniform float: u = (x >> 8) * 2^-24 in [0, 1).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace isf {

__host__ __device__ inline void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; r++) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
        const uint32_t n1 = (uint32_t)p1;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
        const uint32_t n3 = (uint32_t)p0;
        c[0] = n0;
        c[1] = n1;
        c[2] = n2;
        c[3] = n3;
    }
}

// Four uniforms from counter (a, b, c, 0) and the 64-bit seed as key.
__host__ __device__ inline void philox_uniform4(uint32_t a, uint32_t b, uint32_t c, unsigned long long seed, float u[4]) {
    uint32_t x[4] = {a, b, c, 0u};
    philox4x32_10(x, (uint32_t)seed, (uint32_t)(seed >> 32));
#pragma unroll
    for (int i = 0; i < 4; i++) u[i] = (float)(x[i] >> 8) * (1.0f / 16777216.0f);
}

}  // namespace isf
