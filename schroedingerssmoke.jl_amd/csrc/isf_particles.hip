// isf_particles.hip -- the MatlabPort particle store (src/MatlabPort.jl:178-230) on the
// device: structure-of-arrays positions x, y, z (Float32, `capacity` slots), the number
// of slots ever filled (`length`, grows on append! only) and the list of active slot
// indices, compacted in order by removethedead! (examples/matlabport/simple.jl:63-70).
// The active count lives on the device, so births, compaction and advection run without
// host synchronisation; it is read back only when asked for.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "isf_device.hpp"
#include "isf_handle.hpp"
#include "isf_internal.hpp"
#include "isf_philox.hpp"

using namespace isf;

namespace {

constexpr uint32_t kPartMagic = 0x5A0A5A0Au;
constexpr int NT = 256;

// isinside(point, (0,0,0), vol_size): strict on both sides (simple.jl:9-14)
__device__ __forceinline__ bool inside_box(float x, float y, float z, const GridInfo& g) {
    return (x > 0.0f) & (x < g.Lx) & (y > 0.0f) & (y < g.Ly) & (z > 0.0f) & (z < g.Lz);
}

// ---- append!: active[count + k] = first + k, then count += n
__global__ void __launch_bounds__(NT) k_active_append(int* __restrict__ active, const int* __restrict__ count, int first, int n) {
    const int k = blockIdx.x * NT + threadIdx.x;
    if (k < n) active[*count + k] = first + k;
}
__global__ void k_count_add(int* count, int n) { count[0] += n; }

__global__ void __launch_bounds__(NT)
k_copy_soa(float* __restrict__ x, float* __restrict__ y, float* __restrict__ z, const float* __restrict__ sx,
           const float* __restrict__ sy, const float* __restrict__ sz, int first, int n) {
    const int k = blockIdx.x * NT + threadIdx.x;
    if (k >= n) return;
    x[first + k] = sx[k];
    y[first + k] = sy[k];
    z[first + k] = sz[k];
}

// Births of the MatlabPort driver (simple.jl:113-117): rt = 2 pi u, x = c_x,
// y = c_y + 0.9 rad cos(rt), z = c_z + 0.9 rad sin(rt); u from Philox4x32-10 counter
// (j, step, 2, 0) (the reference draws rand(Float32), unseeded).
__global__ void __launch_bounds__(NT)
k_ring_births(float* __restrict__ x, float* __restrict__ y, float* __restrict__ z, int first, int n, JetParams jt,
              unsigned long long seed, unsigned step) {
    const int k = blockIdx.x * NT + threadIdx.x;
    if (k >= n) return;
    uint32_t raw[4] = {(uint32_t)k, step, 2u, 0u};
    philox4x32_10(raw, (uint32_t)seed, (uint32_t)(seed >> 32));
    const float u = (float)(raw[0] >> 8) * (1.0f / 16777216.0f);
    const float rt = u * 2.0f * 3.14159265358979323846f;
    float sn, cs;
    sincosf(rt, &sn, &cs);
    x[first + k] = jt.cx;
    y[first + k] = jt.cy + 0.9f * jt.rad * cs;
    z[first + k] = jt.cz + 0.9f * jt.rad * sn;
}

// ---- removethedead!: order-preserving filter of the active list
// pass 1: kept entries per block
__global__ void __launch_bounds__(NT)
k_keep_count(const int* __restrict__ active, const int* __restrict__ count, const float* __restrict__ x,
             const float* __restrict__ y, const float* __restrict__ z, GridInfo g, int* __restrict__ block_kept) {
    __shared__ int s[NT];
    const int i = blockIdx.x * NT + threadIdx.x;
    int keep = 0;
    if (i < *count) {
        const int a = active[i];
        keep = inside_box(x[a], y[a], z[a], g) ? 1 : 0;
    }
    s[threadIdx.x] = keep;
    __syncthreads();
    for (int st = NT / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) s[threadIdx.x] += s[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) block_kept[blockIdx.x] = s[0];
}
// pass 2: exclusive scan of the block counts (one block), new count
__global__ void __launch_bounds__(NT) k_keep_scan(int* __restrict__ block_kept, int nblocks, int* __restrict__ count) {
    __shared__ int s[NT];
    int carry = 0;
    for (int base = 0; base < nblocks; base += NT) {
        const int i = base + threadIdx.x;
        const int v = i < nblocks ? block_kept[i] : 0;
        s[threadIdx.x] = v;
        __syncthreads();
        for (int off = 1; off < NT; off <<= 1) {  // inclusive Hillis-Steele scan
            const int t = (int)threadIdx.x >= off ? s[threadIdx.x - off] : 0;
            __syncthreads();
            s[threadIdx.x] += t;
            __syncthreads();
        }
        if (i < nblocks) block_kept[i] = carry + s[threadIdx.x] - v;
        carry += s[NT - 1];
        __syncthreads();
    }
    if (threadIdx.x == 0) count[0] = carry;
}
// pass 3: scatter the kept entries in order
__global__ void __launch_bounds__(NT)
k_keep_scatter(const int* __restrict__ active, const float* __restrict__ x, const float* __restrict__ y,
               const float* __restrict__ z, GridInfo g, const int* __restrict__ block_off, int* __restrict__ out,
               const int* __restrict__ count_before) {
    __shared__ int s[NT];
    const int i = blockIdx.x * NT + threadIdx.x;
    int keep = 0, a = 0;
    if (i < *count_before) {
        a = active[i];
        keep = inside_box(x[a], y[a], z[a], g) ? 1 : 0;
    }
    s[threadIdx.x] = keep;
    __syncthreads();
    for (int off = 1; off < NT; off <<= 1) {
        const int t = (int)threadIdx.x >= off ? s[threadIdx.x - off] : 0;
        __syncthreads();
        s[threadIdx.x] += t;
        __syncthreads();
    }
    if (keep) out[block_off[blockIdx.x] + s[threadIdx.x] - 1] = a;
}
__global__ void k_copy_count(const int* __restrict__ src, int* __restrict__ dst) { dst[0] = src[0]; }

// ---- StaggeredAdvect over the active slots (src/MatlabPort.jl:201-222): RK4, no
// in-grid test (the driver filters first)
__global__ void __launch_bounds__(NT)
k_advect_active(float* __restrict__ x, float* __restrict__ y, float* __restrict__ z, const int* __restrict__ active,
                const int* __restrict__ count, const float* __restrict__ vel, const float* __restrict__ vx,
                const float* __restrict__ vy, const float* __restrict__ vz, GridInfo g, float dt) {
    const int i = blockIdx.x * NT + threadIdx.x;
    if (i >= *count) return;
    const int a = active[i];
    const float p[3] = {x[a], y[a], z[a]};
    // interleaved velocity (vel) or separate component arrays (vx, vy, vz)
    auto U = [&](const float q[3], float u[3]) {
        if (vel)
            stag_vel(vel, q, g, u);
        else
            stag_vel_split(vx, vy, vz, q, g, u);
    };
    float k1[3], k2[3], k3[3], k4[3], t[3];
    U(p, k1);
#pragma unroll
    for (int c = 0; c < 3; c++) t[c] = p[c] + k1[c] * dt / 2.0f;
    U(t, k2);
#pragma unroll
    for (int c = 0; c < 3; c++) t[c] = p[c] + k2[c] * dt / 2.0f;
    U(t, k3);
#pragma unroll
    for (int c = 0; c < 3; c++) t[c] = p[c] + k3[c] * dt;
    U(t, k4);
    x[a] = p[0] + dt / 6.0f * (((k1[0] + 2.0f * k2[0]) + 2.0f * k3[0]) + k4[0]);
    y[a] = p[1] + dt / 6.0f * (((k1[1] + 2.0f * k2[1]) + 2.0f * k3[1]) + k4[1]);
    z[a] = p[2] + dt / 6.0f * (((k1[2] + 2.0f * k2[2]) + 2.0f * k3[2]) + k4[2]);
}

// ---- Keep(particle, ind): x = x[ind], y = y[ind], z = z[ind] (src/MatlabPort.jl:226-230).
// The position arrays become n long: slots 0..n-1 receive the gathered positions and the
// slots beyond (which the reference's shorter arrays no longer have) the dead marker -1.
// length and the active list are left as they are, as in the reference.
__global__ void __launch_bounds__(NT)
k_gather_soa(float* __restrict__ nx_, float* __restrict__ ny_, float* __restrict__ nz_, const float* __restrict__ x,
             const float* __restrict__ y, const float* __restrict__ z, const int* __restrict__ ind, int n, int cap,
             int* __restrict__ bad) {
    const int k = blockIdx.x * NT + threadIdx.x;
    if (k >= cap) return;
    const int s = k < n ? ind[k] : -1;
    if (k < n && (s < 0 || s >= cap)) atomicOr(bad, 1);  // the reference raises a BoundsError
    if (s < 0 || s >= cap) {
        nx_[k] = ny_[k] = nz_[k] = -1.0f;
    } else {
        nx_[k] = x[s];
        ny_[k] = y[s];
        nz_[k] = z[s];
    }
}
__global__ void k_set_int(int* p, int v) { p[0] = v; }

}  // namespace

struct isf_particles {
    uint32_t magic;
    int capacity;
    int length;          // slots ever filled (host copy, append! only)
    int size;            // length(p.x): capacity, or n after Keep(p, ind) with n indices
    float *x, *y, *z;    // capacity each
    float *x2, *y2, *z2; // gather targets of Keep
    int *active, *active2;
    int* count;          // device: [0] active count, [1] scratch count
    int* block_kept;     // per-block scratch of the compaction
};

static int check_parts(isf_particles_t p) {
    if (!p || p->magic != kPartMagic) return fail(ISF_EHANDLE, "null or destroyed particle store");
    return ISF_OK;
}
static int nblocks(int n) { return (n + NT - 1) / NT; }

extern "C" {

int isf_particles_create(isf_particles_t* out, long long capacity) {
    CHECK_PTR(out, "out");
    *out = nullptr;
    if (capacity <= 0 || capacity > (1LL << 30)) return fail(ISF_EINVAL, "capacity must be in 1..2^30");
    isf_particles* p = new isf_particles();
    p->magic = kPartMagic;
    p->capacity = (int)capacity;
    p->size = (int)capacity;
    const size_t fb = sizeof(float) * (size_t)capacity, ib = sizeof(int) * (size_t)capacity;
    hipError_t e = hipSuccess;
    float** fbufs[6] = {&p->x, &p->y, &p->z, &p->x2, &p->y2, &p->z2};
    for (float** b : fbufs)
        if (e == hipSuccess) e = hipMalloc(b, fb);
    if (e == hipSuccess) e = hipMalloc(&p->active, ib);
    if (e == hipSuccess) e = hipMalloc(&p->active2, ib);
    if (e == hipSuccess) e = hipMalloc(&p->count, 2 * sizeof(int));
    if (e == hipSuccess) e = hipMalloc(&p->block_kept, sizeof(int) * (size_t)nblocks((int)capacity));
    if (e == hipSuccess) e = hipMemset(p->count, 0, 2 * sizeof(int));
    // zeros(Float32, max_particles) (simple.jl:74-78)
    for (float** b : fbufs)
        if (e == hipSuccess) e = hipMemset(*b, 0, fb);
    if (e != hipSuccess) {
        isf_particles_destroy(p);
        return fail(ISF_ENOMEM, "particle store: %s", hipGetErrorString(e));
    }
    *out = p;
    return ISF_OK;
}

int isf_particles_destroy(isf_particles_t p) {
    int r = check_parts(p);
    if (r) return r;
    (void)hipDeviceSynchronize();
    for (void* b : {(void*)p->x, (void*)p->y, (void*)p->z, (void*)p->x2, (void*)p->y2, (void*)p->z2, (void*)p->active,
                    (void*)p->active2, (void*)p->count, (void*)p->block_kept})
        if (b) (void)hipFree(b);
    p->magic = 0;
    delete p;
    return ISF_OK;
}

static int stream_of(isf_t h, isf_particles_t p, hipStream_t* s) {
    int r = check_parts(p);
    if (!r) r = check_handle(h);
    if (!r) r = join(h);
    if (r) return r;
    *s = h->stream;
    return ISF_OK;
}

// append!: the reference asserts length(x) + p.length < length(p.x) (src/MatlabPort.jl:187)
static int room(isf_particles_t p, long long n) {
    if (n < 0) return fail(ISF_EINVAL, "negative count");
    if ((long long)p->length + n >= p->size)
        return fail(ISF_EINVAL, "append!: %lld + %d particles need more than %d slots", n, p->length, p->size);
    return ISF_OK;
}
static int append_active(isf_particles_t p, int first, int n, hipStream_t s) {
    if (n == 0) return ISF_OK;
    hipLaunchKernelGGL(k_active_append, dim3(nblocks(n)), dim3(NT), 0, s, p->active, p->count, first, n);
    hipLaunchKernelGGL(k_count_add, dim3(1), dim3(1), 0, s, p->count, n);
    HIP_CHECK(hipGetLastError(), "append!");
    return ISF_OK;
}

int isf_particles_append(isf_t h, isf_particles_t p, const float* x, const float* y, const float* z, long long n) {
    hipStream_t s;
    int r = stream_of(h, p, &s);
    if (!r) r = room(p, n);
    if (r) return r;
    if (n > 0 && (!x || !y || !z)) return fail(ISF_EINVAL, "x, y, z must not be NULL");
    if (n > 0) {
        hipLaunchKernelGGL(k_copy_soa, dim3(nblocks((int)n)), dim3(NT), 0, s, p->x, p->y, p->z, x, y, z, p->length, (int)n);
        HIP_CHECK(hipGetLastError(), "append!");
    }
    if ((r = append_active(p, p->length, (int)n, s))) return r;
    p->length += (int)n;
    return ISF_OK;
}

int isf_particles_spawn_ring(isf_t h, isf_particles_t p, long long n, unsigned long long seed, unsigned step) {
    hipStream_t s;
    int r = stream_of(h, p, &s);
    if (!r) r = room(p, n);
    if (r) return r;
    if (!h->jet_set) return fail(ISF_EINVAL, "isf_jet_configure must be called first (nozzle)");
    if (n > 0) {
        hipLaunchKernelGGL(k_ring_births, dim3(nblocks((int)n)), dim3(NT), 0, s, p->x, p->y, p->z, p->length, (int)n,
                           h->jet, seed, step);
        HIP_CHECK(hipGetLastError(), "particle births");
    }
    if ((r = append_active(p, p->length, (int)n, s))) return r;
    p->length += (int)n;
    return ISF_OK;
}

int isf_particles_remove_dead(isf_t h, isf_particles_t p) {
    hipStream_t s;
    int r = stream_of(h, p, &s);
    if (r) return r;
    const int nb = nblocks(p->capacity);
    hipLaunchKernelGGL(k_copy_count, dim3(1), dim3(1), 0, s, p->count, p->count + 1);  // count before
    hipLaunchKernelGGL(k_keep_count, dim3(nb), dim3(NT), 0, s, p->active, p->count + 1, p->x, p->y, p->z, h->g,
                       p->block_kept);
    hipLaunchKernelGGL(k_keep_scan, dim3(1), dim3(NT), 0, s, p->block_kept, nb, p->count);
    hipLaunchKernelGGL(k_keep_scatter, dim3(nb), dim3(NT), 0, s, p->active, p->x, p->y, p->z, h->g, p->block_kept,
                       p->active2, p->count + 1);
    HIP_CHECK(hipGetLastError(), "removethedead!");
    std::swap(p->active, p->active2);
    return ISF_OK;
}

int isf_particles_advect(isf_t h, isf_particles_t p, const float* vel) {
    hipStream_t s;
    int r = stream_of(h, p, &s);
    if (r) return r;
    const int n = std::max(p->length, 1);
    hipLaunchKernelGGL(k_advect_active, dim3(nblocks(n)), dim3(NT), 0, s, p->x, p->y, p->z, p->active, p->count,
                       vel ? vel : h->vel, nullptr, nullptr, nullptr, h->g, h->dt);
    HIP_CHECK(hipGetLastError(), "StaggeredAdvect");
    return ISF_OK;
}

int isf_mp_staggered_advect(isf_t h, isf_particles_t p, const float* vx, const float* vy, const float* vz, float dt) {
    hipStream_t s;
    int r = stream_of(h, p, &s);
    if (r) return r;
    CHECK_PTR(vx, "vx");
    CHECK_PTR(vy, "vy");
    CHECK_PTR(vz, "vz");
    const int n = std::max(p->length, 1);
    hipLaunchKernelGGL(k_advect_active, dim3(nblocks(n)), dim3(NT), 0, s, p->x, p->y, p->z, p->active, p->count,
                       nullptr, vx, vy, vz, h->g, dt);
    HIP_CHECK(hipGetLastError(), "StaggeredAdvect");
    return ISF_OK;
}

int isf_particles_keep(isf_t h, isf_particles_t p, const int* ind, long long n) {
    hipStream_t s;
    int r = stream_of(h, p, &s);
    if (r) return r;
    if (n < 0 || n > p->capacity) return fail(ISF_EINVAL, "Keep: bad count");
    if (n > 0 && !ind) return fail(ISF_EINVAL, "ind must not be NULL");
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, s, p->count + 1, 0);
    hipLaunchKernelGGL(k_gather_soa, dim3(nblocks(p->capacity)), dim3(NT), 0, s, p->x2, p->y2, p->z2, p->x, p->y, p->z,
                       ind, (int)n, p->capacity, p->count + 1);
    HIP_CHECK(hipGetLastError(), "Keep");
    int bad = 0;
    HIP_CHECK(hipMemcpyAsync(&bad, p->count + 1, sizeof(int), hipMemcpyDeviceToHost, s), "Keep");
    HIP_CHECK(hipStreamSynchronize(s), "hipStreamSynchronize");
    std::swap(p->x, p->x2);
    std::swap(p->y, p->y2);
    std::swap(p->z, p->z2);
    p->size = (int)n;
    if (bad) return fail(ISF_EINVAL, "Keep: an index is outside 0..%d (those slots now hold -1)", p->capacity - 1);
    return ISF_OK;
}

int isf_particles_set_active(isf_t h, isf_particles_t p, const int* active, long long n) {
    hipStream_t s;
    int r = stream_of(h, p, &s);
    if (r) return r;
    if (n < 0 || n > p->capacity) return fail(ISF_EINVAL, "set_active: %lld entries for %d slots", n, p->capacity);
    if (n > 0 && !active) return fail(ISF_EINVAL, "active must not be NULL");
    for (long long k = 0; k < n; k++)
        if (active[k] < 0 || active[k] >= p->size)
            return fail(ISF_EINVAL, "set_active: entry %lld = %d outside 0..%d", k, active[k], p->size - 1);
    if (n > 0) HIP_CHECK(hipMemcpyAsync(p->active, active, sizeof(int) * (size_t)n, hipMemcpyHostToDevice, s), "set_active");
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, s, p->count, (int)n);
    HIP_CHECK(hipGetLastError(), "set_active");
    HIP_CHECK(hipStreamSynchronize(s), "hipStreamSynchronize");
    return ISF_OK;
}

int isf_particles_state(isf_t h, isf_particles_t p, long long* length, long long* active_count, const float** x,
                        const float** y, const float** z, const int** active) {
    hipStream_t s;
    int r = stream_of(h, p, &s);
    if (r) return r;
    if (length) *length = p->length;
    if (active_count) {
        int c = 0;
        HIP_CHECK(hipMemcpyAsync(&c, p->count, sizeof(int), hipMemcpyDeviceToHost, s), "count copy");
        HIP_CHECK(hipStreamSynchronize(s), "hipStreamSynchronize");
        *active_count = c;
    }
    if (x) *x = p->x;
    if (y) *y = p->y;
    if (z) *z = p->z;
    if (active) *active = p->active;
    return ISF_OK;
}

}  // extern "C"
