/*
 * isf_philox.h -- Philox4x32-10 counter-based generator for the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY (see isf_oracle_impl.h).
 *
 * Published method: J. K. Salmon, M. A. Moraes, R. O. Dror, D. E. Shaw, "Parallel random
 * numbers: as easy as 1, 2, 3", SC'11 (the Random123 Philox4x32 round with 10 rounds).
 * The reference draws spawn positions from the unseeded rand(Float32) / gpu_rand
 * (src/particles.jl:37-42,78-83); the seeded counter-based draw replaces it so that CPU
 * and GPU paths see identical inputs.
 *
 * Uniform float: u = (x >> 8) * 2^-24, in [0, 1).
 */
#ifndef ISF_PHILOX_H
#define ISF_PHILOX_H
#include <stdint.h>

static inline void isf_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; r++) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static inline void isf_philox_uniform4(uint32_t a, uint32_t b, uint32_t c, unsigned long long seed, float u[4]) {
    uint32_t ctr[4] = {a, b, c, 0u};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t o[4];
    isf_philox4x32_10(ctr, key, o);
    for (int i = 0; i < 4; i++) u[i] = (float)(o[i] >> 8) * (1.0f / 16777216.0f);
}
#endif
