// isf_sort.hip -- spatial processing order of the particle slots for the advection.
//
// Slots are sorted by bucket = (ix / 8) + (nx / 8) * (iy + ny * iz) of the cell holding
// the particle (slots outside the box last), with rocPRIM's stable LSD radix sort of
// (bucket, slot) pairs.  A wavefront then advects particles from the same few cells and
// its velocity gathers share cache lines.  Each slot is advanced independently, so the
// order never changes a result.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>

#include "isf_internal.hpp"

namespace isf {

static int bucket_count(const GridInfo& g) { return (g.nx / 8) * g.ny * g.nz; }

static unsigned key_bits(const GridInfo& g) {
    unsigned b = 1;
    while ((1ull << b) <= (unsigned long long)bucket_count(g)) b++;
    return b;
}

__global__ void __launch_bounds__(256)
k_bucket_keys(const float* __restrict__ parts, long long nslots, GridInfo g, int nb, unsigned* __restrict__ keys,
              int* __restrict__ slots) {
    const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslots) return;
    const float x = parts[3 * s], y = parts[3 * s + 1], z = parts[3 * s + 2];
    unsigned key = (unsigned)nb;
    if ((x > 0.0f) & (x < g.Lx) & (y > 0.0f) & (y < g.Ly) & (z > 0.0f) & (z < g.Lz)) {
        const int ix = min((int)(x * g.idx), g.nx - 1), iy = min((int)(y * g.idy), g.ny - 1);
        const int iz = min((int)(z * g.idz), g.nz - 1);
        key = (unsigned)((ix >> 3) + (g.nx >> 3) * (iy + g.ny * iz));
    }
    keys[s] = key;
    slots[s] = (int)s;
}

static size_t align256(size_t b) { return (b + 255) / 256 * 256; }

static size_t rocprim_bytes(const GridInfo& g, long long n) {
    size_t bytes = 0;
    (void)rocprim::radix_sort_pairs(nullptr, bytes, (unsigned*)nullptr, (unsigned*)nullptr, (int*)nullptr,
                                    (int*)nullptr, (size_t)n, 0u, key_bits(g));
    return bytes;
}

size_t order_scratch_bytes(const GridInfo& g, long long nslots) {
    return 3 * align256(sizeof(int) * (size_t)nslots) + align256(rocprim_bytes(g, nslots));
}

int launch_particle_order(const GridInfo& g, const float* parts, long long nslots, int* order, void* scratch,
                          size_t scratch_bytes, hipStream_t s) {
    if (nslots <= 0) return 0;
    if (scratch_bytes < order_scratch_bytes(g, nslots)) return (int)hipErrorInvalidValue;
    char* p = static_cast<char*>(scratch);
    const size_t a = align256(sizeof(int) * (size_t)nslots);
    unsigned* keys_in = reinterpret_cast<unsigned*>(p);
    unsigned* keys_out = reinterpret_cast<unsigned*>(p + a);
    int* slots = reinterpret_cast<int*>(p + 2 * a);
    void* tmp = p + 3 * a;
    size_t tmp_bytes = scratch_bytes - 3 * a;
    const int nb = bucket_count(g);
    hipLaunchKernelGGL(k_bucket_keys, dim3((unsigned)((nslots + 255) / 256)), dim3(256), 0, s, parts, nslots, g, nb,
                       keys_in, slots);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return (int)e;
    e = rocprim::radix_sort_pairs(tmp, tmp_bytes, keys_in, keys_out, slots, order, (size_t)nslots, 0u, key_bits(g), s);
    return (int)e;
}

}  // namespace isf
