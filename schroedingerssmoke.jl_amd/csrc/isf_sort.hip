// isf_sort.hip -- spatial processing orders of the particle slots, hand-written counting
// sorts and a bitonic sort.
//
// Advection: slots are grouped by bucket = (ix / 8) + (nx / 8) * (iy + ny * iz) of the
// cell holding the particle (slots outside the box last) with a counting sort (count per
// bucket, exclusive scan, scatter).  A wavefront then advects particles from the same few
// cells and its velocity gathers share cache lines.  Each slot is advanced independently,
// so the order never changes a result, and the order inside a bucket is free.
//
// Advection adjoint: the same counting sort by exact cell, so the velocity scatter of a
// wavefront and of a workgroup hits few cells.
//
// Nozzle fill: the draws are stored in the order of a 24-bit spatial key, ties by draw
// index, by a bitonic sort of (key << 32 | draw) words.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "isf_internal.hpp"
#include "isf_philox.hpp"

namespace isf {

// xs: log2 of the x cells per bucket (3: buckets of 8 cells; 0: single cells)
static int bucket_count(const GridInfo& g, int xs) { return (g.nx >> xs) * g.ny * g.nz; }

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
k_cell_rank(const float* __restrict__ parts, long long nslots, GridInfo g, int nb, int xs, unsigned* __restrict__ count,
            unsigned* __restrict__ key, unsigned* __restrict__ rank) {
    const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const bool act = s < nslots;
    const unsigned k = act ? particle_bucket(parts[3 * s], parts[3 * s + 1], parts[3 * s + 2], g, nb, xs) : 0u;
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

// ---- exclusive scan of unsigned counts: block t scans tile t (SCAN_TILE elements, SCAN_E
// per thread) into out and writes the tile's total; one block scans the totals in place
// (chunks of SCAN_TILE with a running carry); every block then adds its tile's offset.
constexpr int SCAN_T = 1024, SCAN_E = 4, SCAN_TILE = SCAN_T * SCAN_E;

// exclusive scan of one value per thread over the block (SCAN_T threads); *total = sum
__device__ __forceinline__ unsigned block_exclusive_scan(unsigned v, unsigned* total) {
    __shared__ unsigned wsum[SCAN_T / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    unsigned inc = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const unsigned t = __shfl_up(inc, o, 64);
        if (lane >= o) inc += t;
    }
    if (lane == 63) wsum[wave] = inc;
    __syncthreads();
    if (wave == 0) {
        unsigned w = lane < SCAN_T / 64 ? wsum[lane] : 0u;
#pragma unroll
        for (int o = 1; o < SCAN_T / 64; o <<= 1) {
            const unsigned t = __shfl_up(w, o, 64);
            if (lane >= o) w += t;
        }
        if (lane < SCAN_T / 64) wsum[lane] = w;  // inclusive wave totals
    }
    __syncthreads();
    const unsigned before = wave > 0 ? wsum[wave - 1] : 0u;
    *total = wsum[SCAN_T / 64 - 1];
    __syncthreads();  // wsum is reused by the next call
    return before + inc - v;
}

__global__ void __launch_bounds__(SCAN_T)
k_scan_tiles(const unsigned* __restrict__ in, unsigned* __restrict__ out, size_t n, unsigned* __restrict__ sums) {
    const size_t i0 = (size_t)blockIdx.x * SCAN_TILE + (size_t)threadIdx.x * SCAN_E;
    unsigned v[SCAN_E], t = 0u;
#pragma unroll
    for (int k = 0; k < SCAN_E; k++) {
        v[k] = i0 + k < n ? in[i0 + k] : 0u;
        t += v[k];
    }
    unsigned total;
    unsigned run = block_exclusive_scan(t, &total);
#pragma unroll
    for (int k = 0; k < SCAN_E; k++) {
        if (i0 + k < n) out[i0 + k] = run;
        run += v[k];
    }
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}
__global__ void __launch_bounds__(SCAN_T) k_scan_sums(unsigned* __restrict__ sums, size_t ntiles) {
    unsigned carry = 0u;
    for (size_t c0 = 0; c0 < ntiles; c0 += SCAN_TILE) {
        const size_t i0 = c0 + (size_t)threadIdx.x * SCAN_E;
        unsigned v[SCAN_E], t = 0u;
#pragma unroll
        for (int k = 0; k < SCAN_E; k++) {
            v[k] = i0 + k < ntiles ? sums[i0 + k] : 0u;
            t += v[k];
        }
        unsigned total;
        unsigned run = carry + block_exclusive_scan(t, &total);
#pragma unroll
        for (int k = 0; k < SCAN_E; k++) {
            if (i0 + k < ntiles) sums[i0 + k] = run;
            run += v[k];
        }
        carry += total;
    }
}
__global__ void __launch_bounds__(256) k_scan_add(unsigned* __restrict__ out, size_t n, const unsigned* __restrict__ sums) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) out[i] += sums[i / SCAN_TILE];
}
static size_t scan_tiles(size_t n) { return (n + SCAN_TILE - 1) / SCAN_TILE; }
static int exclusive_scan(const unsigned* in, unsigned* out, size_t n, unsigned* sums, hipStream_t s) {
    const size_t nt = scan_tiles(n);
    hipLaunchKernelGGL(k_scan_tiles, dim3((unsigned)nt), dim3(SCAN_T), 0, s, in, out, n, sums);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(SCAN_T), 0, s, sums, nt);
    hipLaunchKernelGGL(k_scan_add, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, out, n, sums);
    return (int)hipGetLastError();
}

// ---- counting sort of the slots by bucket (xs: log2 of the x cells per bucket)
// scratch: counts and starts (buckets + 1 each), keys and ranks (nslots each), tile sums
static size_t counting_scratch_bytes(const GridInfo& g, long long nslots, int xs) {
    const size_t nb = (size_t)bucket_count(g, xs) + 1;
    return 2 * align256(sizeof(unsigned) * nb) + 2 * align256(sizeof(unsigned) * (size_t)nslots) +
           align256(sizeof(unsigned) * scan_tiles(nb));
}
static int counting_order(const GridInfo& g, const float* parts, long long nslots, int* order, void* scratch,
                          size_t scratch_bytes, hipStream_t s, int xs) {
    if (nslots <= 0) return 0;
    if (scratch_bytes < counting_scratch_bytes(g, nslots, xs)) return (int)hipErrorInvalidValue;
    const int nbk = bucket_count(g, xs);
    const size_t nb = (size_t)nbk + 1, ab = align256(sizeof(unsigned) * nb),
                 an = align256(sizeof(unsigned) * (size_t)nslots);
    char* p = static_cast<char*>(scratch);
    unsigned* count = reinterpret_cast<unsigned*>(p);
    unsigned* start = reinterpret_cast<unsigned*>(p + ab);
    unsigned* key = reinterpret_cast<unsigned*>(p + 2 * ab);
    unsigned* rank = reinterpret_cast<unsigned*>(p + 2 * ab + an);
    unsigned* sums = reinterpret_cast<unsigned*>(p + 2 * ab + 2 * an);
    const dim3 gs((unsigned)((nslots + 255) / 256));
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((nb / 4 + 255) / 256 + 1)), dim3(256), 0, s, count, nb);
    hipLaunchKernelGGL(k_cell_rank, gs, dim3(256), 0, s, parts, nslots, g, nbk, xs, count, key, rank);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return (int)e;
    const int r = exclusive_scan(count, start, nb, sums, s);
    if (r) return r;
    hipLaunchKernelGGL(k_cell_scatter, gs, dim3(256), 0, s, nslots, key, rank, start, order);
    return (int)hipGetLastError();
}

size_t order_scratch_bytes(const GridInfo& g, long long nslots) {
    return nslots > 0 ? counting_scratch_bytes(g, nslots, 3) : 0;
}
size_t cell_order_scratch_bytes(const GridInfo& g, long long nslots) { return counting_scratch_bytes(g, nslots, 0); }

int launch_cell_order(const GridInfo& g, const float* parts, long long nslots, int* order, void* scratch,
                      size_t scratch_bytes, hipStream_t s) {
    return counting_order(g, parts, nslots, order, scratch, scratch_bytes, s, 0);
}
int launch_particle_order(const GridInfo& g, const float* parts, long long nslots, int* order, void* scratch,
                          size_t scratch_bytes, hipStream_t s) {
    return counting_order(g, parts, nslots, order, scratch, scratch_bytes, s, 3);
}

// ---- bitonic sort of 64-bit words (ascending), n a power of two >= BITONIC_TILE: tiles
// of BITONIC_TILE words are sorted in LDS; each later stage k runs its strides j >= the
// tile in global memory (one pass each) and the strides below the tile in LDS.
constexpr int BITONIC_T = 1024, BITONIC_TILE = 2 * BITONIC_T;

__device__ __forceinline__ void bitonic_cas(unsigned long long* a, size_t lo, size_t hi, bool asc) {
    const unsigned long long x = a[lo], y = a[hi];
    if ((x > y) == asc) {
        a[lo] = y;
        a[hi] = x;
    }
}
// k == 0: sort the tile (all stages up to BITONIC_TILE); else the strides j < the tile of stage k
__global__ void __launch_bounds__(BITONIC_T) k_bitonic_tile(unsigned long long* __restrict__ keys, size_t k) {
    __shared__ unsigned long long t[BITONIC_TILE];
    const size_t base = (size_t)blockIdx.x * BITONIC_TILE;
    const int tid = threadIdx.x;
    t[tid] = keys[base + tid];
    t[tid + BITONIC_T] = keys[base + tid + BITONIC_T];
    __syncthreads();
    const size_t k0 = k ? k : 2, k1 = k ? k : BITONIC_TILE;
    for (size_t kk = k0; kk <= k1; kk <<= 1) {
        for (size_t j = (kk < BITONIC_TILE ? kk : BITONIC_TILE) >> 1; j > 0; j >>= 1) {
            const size_t lo = (tid / j) * 2 * j + tid % j;
            bitonic_cas(t, lo, lo + j, ((base + lo) & kk) == 0);
            __syncthreads();
        }
    }
    keys[base + tid] = t[tid];
    keys[base + tid + BITONIC_T] = t[tid + BITONIC_T];
}
// one stride j >= BITONIC_TILE of stage k over the whole array (one pair per thread)
__global__ void __launch_bounds__(256) k_bitonic_global(unsigned long long* __restrict__ keys, size_t n, size_t k, size_t j) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n / 2) return;
    const size_t lo = (i / j) * 2 * j + i % j;
    bitonic_cas(keys, lo, lo + j, (lo & k) == 0);
}
static int bitonic_sort(unsigned long long* keys, size_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_bitonic_tile, dim3((unsigned)(n / BITONIC_TILE)), dim3(BITONIC_T), 0, s, keys, (size_t)0);
    for (size_t k = 2 * (size_t)BITONIC_TILE; k <= n; k <<= 1) {
        for (size_t j = k / 2; j >= (size_t)BITONIC_TILE; j >>= 1)
            hipLaunchKernelGGL(k_bitonic_global, dim3((unsigned)((n / 2 + 255) / 256)), dim3(256), 0, s, keys, n, k, j);
        hipLaunchKernelGGL(k_bitonic_tile, dim3((unsigned)(n / BITONIC_TILE)), dim3(BITONIC_T), 0, s, keys, k);
    }
    return (int)hipGetLastError();
}

// ---- initial particle cloud uniform in the nozzle cylinder, stored in a spatially
// coherent slot order.  Draw i uses Philox4x32-10 with counter (i, 0xFFFFFFFF, 1, 0):
// x = c_x + len (u0 - 1/2), r = rad sqrt(u1), theta = 2 pi u2.  Draws are assigned to
// slots in increasing order of the 24-bit key (x slab: top 5 bits of word 0, theta wedge:
// top 7 bits of word 2, r: top 12 bits of word 1; ties by draw index), so a wavefront of
// consecutive slots holds particles from a few neighbouring cells.  No reference
// counterpart.
__global__ void __launch_bounds__(256)
k_fill_keys(long long n, long long npad, unsigned long long seed, unsigned long long* __restrict__ keys) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npad) return;
    unsigned long long key = ~0ull;  // padding sorts last
    if (i < n) {
        uint32_t x[4] = {(uint32_t)i, 0xFFFFFFFFu, 1u, 0u};
        philox4x32_10(x, (uint32_t)seed, (uint32_t)(seed >> 32));
        key = ((unsigned long long)(((x[0] >> 27) << 19) | ((x[2] >> 25) << 12) | (x[1] >> 20)) << 32) | (uint32_t)i;
    }
    keys[i] = key;
}

__global__ void __launch_bounds__(256)
k_fill_ordered(float* __restrict__ parts, long long nslots, long long first, long long n,
               const unsigned long long* __restrict__ keys, JetParams j, unsigned long long seed) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const long long s = (first + k) % nslots;
    const uint32_t draw = (uint32_t)keys[k];  // the draw stored k-th
    float u[4];
    philox_uniform4(draw, 0xFFFFFFFFu, 1u, seed, u);
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
    size_t npad = BITONIC_TILE;
    while (npad < (size_t)n) npad <<= 1;
    void* mem = nullptr;
    hipError_t e = hipMalloc(&mem, sizeof(unsigned long long) * npad);
    if (e != hipSuccess) return (int)e;
    unsigned long long* keys = static_cast<unsigned long long*>(mem);
    hipLaunchKernelGGL(k_fill_keys, dim3((unsigned)((npad + 255) / 256)), dim3(256), 0, s, n, (long long)npad, seed, keys);
    e = hipGetLastError();
    if (e == hipSuccess) e = (hipError_t)bitonic_sort(keys, npad, s);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_fill_ordered, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, parts, nslots, first, n,
                           keys, jet, seed);
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
