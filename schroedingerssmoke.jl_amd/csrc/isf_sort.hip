// isf_sort.hip -- spatial processing orders of the particle slots.
//
// Advection: slots are sorted by bucket = (ix / 8) + (nx / 8) * (iy + ny * iz) of the
// cell holding the particle (slots outside the box last), with rocPRIM's stable LSD radix
// sort of (bucket, slot) pairs.  A wavefront then advects particles from the same few
// cells and its velocity gathers share cache lines.  Each slot is advanced independently,
// so the order never changes a result.
//
// Advection adjoint: slots grouped by their exact cell with a counting sort (count per
// cell, exclusive scan, scatter; order inside a cell arbitrary), so the velocity scatter
// of a wavefront and of a workgroup hits few cells.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "isf_internal.hpp"
#include "isf_philox.hpp"

namespace isf {

// xs: log2 of the x cells per bucket (3: buckets of 8 cells; 0: single cells)
static int bucket_count(const GridInfo& g, int xs) { return (g.nx >> xs) * g.ny * g.nz; }

static unsigned key_bits(const GridInfo& g, int xs) {
    unsigned b = 1;
    while ((1ull << b) <= (unsigned long long)bucket_count(g, xs)) b++;
    return b;
}

// bucket of a particle: cell (x / 2^xs, y, z) strictly inside the box, else nb
__device__ __forceinline__ unsigned particle_bucket(float x, float y, float z, const GridInfo& g, int nb, int xs) {
    unsigned key = (unsigned)nb;
    if ((x > 0.0f) & (x < g.Lx) & (y > 0.0f) & (y < g.Ly) & (z > 0.0f) & (z < g.Lz)) {
        const int ix = min((int)(x * g.idx), g.nx - 1), iy = min((int)(y * g.idy), g.ny - 1);
        const int iz = min((int)(z * g.idz), g.nz - 1);
        key = (unsigned)((ix >> xs) + (g.nx >> xs) * (iy + g.ny * iz));
    }
    return key;
}

__global__ void __launch_bounds__(256)
k_bucket_keys(const float* __restrict__ parts, long long nslots, GridInfo g, int nb, int xs, unsigned* __restrict__ keys,
              int* __restrict__ slots) {
    const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslots) return;
    keys[s] = particle_bucket(parts[3 * s], parts[3 * s + 1], parts[3 * s + 2], g, nb, xs);
    slots[s] = (int)s;
}

// counting sort by cell: rank of each slot inside its cell, then the scatter
// zero n counters (p 16-byte aligned): uint4 i of the n / 4, the last n % 4 by block 0
__global__ void __launch_bounds__(256) k_fill_u32(unsigned* __restrict__ p, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n / 4) reinterpret_cast<uint4*>(p)[i] = make_uint4(0u, 0u, 0u, 0u);
    if (blockIdx.x == 0 && threadIdx.x < n % 4) p[n / 4 * 4 + threadIdx.x] = 0u;
}
// Rank of each slot inside its cell.  Neighbouring slots mostly share a cell, so the lanes
// of a wavefront take their ranks with one atomic per distinct cell: the lowest pending
// lane's cell is added once with the count of lanes in it, and each of those lanes takes
// base + the number of lower lanes in the same cell.
__global__ void __launch_bounds__(256)
k_cell_rank(const float* __restrict__ parts, long long nslots, GridInfo g, int nb, unsigned* __restrict__ count,
            unsigned* __restrict__ key, unsigned* __restrict__ rank) {
    const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const bool act = s < nslots;
    const unsigned k = act ? particle_bucket(parts[3 * s], parts[3 * s + 1], parts[3 * s + 2], g, nb, 0) : 0u;
    if (act) key[s] = k;
    const int lane = (int)(threadIdx.x & 63u);
    const unsigned long long below = (1ull << lane) - 1ull;
    unsigned long long pending = __ballot(act);
    unsigned r = 0u;
    while (pending) {
        const int leader = __ffsll((long long)pending) - 1;
        const unsigned kl = __shfl(k, leader, 64);
        const unsigned long long same = __ballot(act && k == kl) & pending;
        unsigned base = 0u;
        if (lane == leader) base = atomicAdd(count + kl, (unsigned)__popcll(same));
        base = __shfl(base, leader, 64);
        if ((same >> lane) & 1ull) r = base + (unsigned)__popcll(same & below);
        pending &= ~same;
    }
    if (act) rank[s] = r;
}
__global__ void __launch_bounds__(256)
k_cell_scatter(long long nslots, const unsigned* __restrict__ key, const unsigned* __restrict__ rank,
               const unsigned* __restrict__ start, int* __restrict__ order) {
    const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslots) return;
    order[start[key[s]] + rank[s]] = (int)s;
}

static size_t align256(size_t b) { return (b + 255) / 256 * 256; }

static size_t rocprim_bytes(const GridInfo& g, long long n, int xs) {
    size_t bytes = 0;
    (void)rocprim::radix_sort_pairs(nullptr, bytes, (unsigned*)nullptr, (unsigned*)nullptr, (int*)nullptr,
                                    (int*)nullptr, (size_t)n, 0u, key_bits(g, xs));
    return bytes;
}

size_t order_scratch_bytes(const GridInfo& g, long long nslots) {
    return 3 * align256(sizeof(int) * (size_t)nslots) + align256(rocprim_bytes(g, nslots, 3));
}

static size_t scan_bytes(size_t n) {
    size_t bytes = 0;
    (void)rocprim::exclusive_scan(nullptr, bytes, (unsigned*)nullptr, (unsigned*)nullptr, 0u, n,
                                  rocprim::plus<unsigned>());
    return bytes;
}

// counts and starts (cells + 1 each), keys and ranks (nslots each), scan temporary
size_t cell_order_scratch_bytes(const GridInfo& g, long long nslots) {
    const size_t nb = (size_t)bucket_count(g, 0) + 1;
    return 2 * align256(sizeof(unsigned) * nb) + 2 * align256(sizeof(unsigned) * (size_t)nslots) +
           align256(scan_bytes(nb));
}

int launch_cell_order(const GridInfo& g, const float* parts, long long nslots, int* order, void* scratch,
                      size_t scratch_bytes, hipStream_t s) {
    if (nslots <= 0) return 0;
    if (scratch_bytes < cell_order_scratch_bytes(g, nslots)) return (int)hipErrorInvalidValue;
    const int nbk = bucket_count(g, 0);
    const size_t nb = (size_t)nbk + 1, ab = align256(sizeof(unsigned) * nb),
                 an = align256(sizeof(unsigned) * (size_t)nslots);
    char* p = static_cast<char*>(scratch);
    unsigned* count = reinterpret_cast<unsigned*>(p);
    unsigned* start = reinterpret_cast<unsigned*>(p + ab);
    unsigned* key = reinterpret_cast<unsigned*>(p + 2 * ab);
    unsigned* rank = reinterpret_cast<unsigned*>(p + 2 * ab + an);
    void* tmp = p + 2 * ab + 2 * an;
    size_t tmp_bytes = scratch_bytes - (2 * ab + 2 * an);
    const dim3 gs((unsigned)((nslots + 255) / 256));
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((nb / 4 + 255) / 256 + 1)), dim3(256), 0, s, count, nb);
    hipLaunchKernelGGL(k_cell_rank, gs, dim3(256), 0, s, parts, nslots, g, nbk, count, key, rank);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return (int)e;
    e = rocprim::exclusive_scan(tmp, tmp_bytes, count, start, 0u, nb, rocprim::plus<unsigned>(), s);
    if (e != hipSuccess) return (int)e;
    hipLaunchKernelGGL(k_cell_scatter, gs, dim3(256), 0, s, nslots, key, rank, start, order);
    return (int)hipGetLastError();
}

static int particle_order(const GridInfo& g, const float* parts, long long nslots, int* order, void* scratch,
                          size_t scratch_bytes, hipStream_t s, int xs) {
    if (nslots <= 0) return 0;
    if (scratch_bytes < order_scratch_bytes(g, nslots)) return (int)hipErrorInvalidValue;
    char* p = static_cast<char*>(scratch);
    const size_t a = align256(sizeof(int) * (size_t)nslots);
    unsigned* keys_in = reinterpret_cast<unsigned*>(p);
    unsigned* keys_out = reinterpret_cast<unsigned*>(p + a);
    int* slots = reinterpret_cast<int*>(p + 2 * a);
    void* tmp = p + 3 * a;
    size_t tmp_bytes = scratch_bytes - 3 * a;
    const int nb = bucket_count(g, xs);
    hipLaunchKernelGGL(k_bucket_keys, dim3((unsigned)((nslots + 255) / 256)), dim3(256), 0, s, parts, nslots, g, nb, xs,
                       keys_in, slots);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return (int)e;
    e = rocprim::radix_sort_pairs(tmp, tmp_bytes, keys_in, keys_out, slots, order, (size_t)nslots, 0u, key_bits(g, xs),
                                  s);
    return (int)e;
}

int launch_particle_order(const GridInfo& g, const float* parts, long long nslots, int* order, void* scratch,
                          size_t scratch_bytes, hipStream_t s) {
    return particle_order(g, parts, nslots, order, scratch, scratch_bytes, s, 3);
}

// ---- initial particle cloud uniform in the nozzle cylinder, stored in a spatially
// coherent slot order.  Draw i uses Philox4x32-10 with counter (i, 0xFFFFFFFF, 1, 0):
// x = c_x + len (u0 - 1/2), r = rad sqrt(u1), theta = 2 pi u2.  Draws are assigned to
// slots in increasing order of the 24-bit key (x slab: top 5 bits of word 0, theta wedge:
// top 7 bits of word 2, r: top 12 bits of word 1; ties by draw index), so a wavefront of
// consecutive slots holds particles from a few neighbouring cells.  No reference
// counterpart.
__global__ void __launch_bounds__(256)
k_fill_keys(long long n, unsigned long long seed, unsigned* __restrict__ keys, int* __restrict__ idx) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t x[4] = {(uint32_t)i, 0xFFFFFFFFu, 1u, 0u};
    philox4x32_10(x, (uint32_t)seed, (uint32_t)(seed >> 32));
    keys[i] = ((x[0] >> 27) << 19) | ((x[2] >> 25) << 12) | (x[1] >> 20);
    idx[i] = (int)i;
}

__global__ void __launch_bounds__(256)
k_fill_ordered(float* __restrict__ parts, long long nslots, long long first, long long n, const int* __restrict__ draw,
               JetParams j, unsigned long long seed) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const long long s = (first + k) % nslots;
    float u[4];
    philox_uniform4((uint32_t)draw[k], 0xFFFFFFFFu, 1u, seed, u);
    const float r = j.rad * sqrtf(u[1]);
    const float th = u[2] * 2.0f * 3.14159265358979323846f;
    float sn, cs;
    sincosf(th, &sn, &cs);
    parts[3 * s + 0] = j.cx + j.len * (u[0] - 0.5f);
    parts[3 * s + 1] = j.cy + r * cs;
    parts[3 * s + 2] = j.cz + r * sn;
}

int launch_fill_cylinder(float* parts, long long nslots, long long first, long long n, const JetParams& jet,
                         unsigned long long seed, hipStream_t s) {
    if (n <= 0) return 0;
    size_t tmp_bytes = 0;
    (void)rocprim::radix_sort_pairs(nullptr, tmp_bytes, (unsigned*)nullptr, (unsigned*)nullptr, (int*)nullptr,
                                    (int*)nullptr, (size_t)n, 0u, 24u);
    const size_t a = align256(sizeof(int) * (size_t)n);
    void* mem = nullptr;
    hipError_t e = hipMalloc(&mem, 4 * a + align256(tmp_bytes));
    if (e != hipSuccess) return (int)e;
    char* p = static_cast<char*>(mem);
    unsigned* keys_in = reinterpret_cast<unsigned*>(p);
    unsigned* keys_out = reinterpret_cast<unsigned*>(p + a);
    int* idx = reinterpret_cast<int*>(p + 2 * a);
    int* draw = reinterpret_cast<int*>(p + 3 * a);
    const dim3 grid((unsigned)((n + 255) / 256));
    hipLaunchKernelGGL(k_fill_keys, grid, dim3(256), 0, s, n, seed, keys_in, idx);
    e = hipGetLastError();
    if (e == hipSuccess)
        e = rocprim::radix_sort_pairs(p + 4 * a, tmp_bytes, keys_in, keys_out, idx, draw, (size_t)n, 0u, 24u, s);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_fill_ordered, grid, dim3(256), 0, s, parts, nslots, first, n, draw, jet, seed);
        e = hipGetLastError();
    }
    hipError_t e2 = hipStreamSynchronize(s);
    (void)hipFree(mem);
    return (int)(e != hipSuccess ? e : e2);
}

// Order measure of a particle buffer: number of consecutive slot pairs, both inside the
// box in x, more than two cells apart on some axis; one partial count per block.
__global__ void __launch_bounds__(256)
k_order_jumps(const float* __restrict__ parts, long long nslots, GridInfo g, unsigned* __restrict__ partial) {
    const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned jump = 0;
    if (s + 1 < nslots) {
        const float* a = parts + 3 * s;
        const float* b = a + 3;
        const bool in = (a[0] > 0.0f) & (a[0] < g.Lx) & (b[0] > 0.0f) & (b[0] < g.Lx);
        jump = in && (fabsf(a[0] - b[0]) > 2.0f * g.dx || fabsf(a[1] - b[1]) > 2.0f * g.dy ||
                      fabsf(a[2] - b[2]) > 2.0f * g.dz);
    }
    __shared__ unsigned sh[256];
    sh[threadIdx.x] = jump;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

size_t order_measure_ints(long long nslots) { return (size_t)((nslots + 255) / 256); }

int measure_order_jumps(const GridInfo& g, const float* parts, long long nslots, unsigned* dev_partial,
                        unsigned long long* host_count, hipStream_t s) {
    *host_count = 0;
    if (nslots < 2) return 0;
    const size_t nb = order_measure_ints(nslots);
    hipLaunchKernelGGL(k_order_jumps, dim3((unsigned)nb), dim3(256), 0, s, parts, nslots, g, dev_partial);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return (int)e;
    std::vector<unsigned> h(nb);
    if ((e = hipMemcpyAsync(h.data(), dev_partial, sizeof(unsigned) * nb, hipMemcpyDeviceToHost, s)) != hipSuccess)
        return (int)e;
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return (int)e;
    for (unsigned v : h) *host_count += v;
    return 0;
}

}  // namespace isf
