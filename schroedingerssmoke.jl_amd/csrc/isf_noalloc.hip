// isf_noalloc.hip -- isf_norm2_noalloc: the FP64 norm reduction of isf_norm2 for callers
// that cannot allocate device memory (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc).  It needs no handle and calls no allocating routine:
// the per-block partial sums live in `scratch`, a file-scope device array of fixed size
// (kNorm2Blocks doubles, 8 KiB) that each GPU the library runs on holds until the process
// ends, and the sum is written to a device double the caller passes in.
//
// Every call on one GPU shares that GPU's copy of `scratch`: calls that overlap on one GPU
// (on different streams, or from different host threads without ordering) race.  Each
// call overwrites all kNorm2Blocks partials before the reduction reads them, so nothing
// has to be cleared between calls.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/isf.h"
#include "isf_handle.hpp"
#include "isf_internal.hpp"

using namespace isf;

namespace {

// k_norm2's partial sums: exactly kNorm2Blocks doubles per call
static __device__ __align__(256) double scratch[kNorm2Blocks];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices];  // device address of `scratch` on each GPU (null: not resolved yet)

// The device address of this GPU's `scratch`.  In host code the name `scratch` is a
// host-side stand-in, not a device address, so it is resolved once per device with
// hipGetSymbolAddress (with that device current) and only the resolved pointer is used.
int scratch_ptr(double** out) {
    int dev = 0;
    HIP_CHECK(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) return fail(ISF_EINVAL, "device %d out of range (0..%d)", dev, kMaxDevices - 1);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    *out = g_scratch_ptr[dev];
    return ISF_OK;
}

}  // namespace

extern "C" {

int isf_norm2_noalloc(const float* psi, long long ncells, double* out, void* stream) {
    CHECK_PTR(psi, "psi");
    CHECK_PTR(out, "out");
    if (ncells <= 0 || ncells > ISF_NOALLOC_MAX_CELLS)
        return fail(ISF_EINVAL, "isf_norm2_noalloc: ncells = %lld outside 1..%lld", ncells,
                    (long long)ISF_NOALLOC_MAX_CELLS);
    double* part = nullptr;
    int r = scratch_ptr(&part);
    if (r) return r;
    GridInfo g{};
    g.nx = (int)ncells;  // k_norm2 walks nx * ny * nz cells as one range
    g.ny = 1;
    g.nz = 1;
    const hipStream_t s = static_cast<hipStream_t>(stream);
    size_t np = 0;
    LAUNCH_CHECK(launch_norm2(g, reinterpret_cast<const float4*>(psi), part, &np, s), "norm2 (no-alloc)");
    LAUNCH_CHECK(launch_reduce_sum(part, np, out, s), "norm2 reduce (no-alloc)");
    return ISF_OK;
}

}  // extern "C"
