// isf_export.hip -- asynchronous, double-buffered export of particle positions to host
// memory, the device-side form of the visualisation driver's per-frame copy
// `Array(particles)` (examples/jet.jl:83).
//
// push(k): on the handle's side stream (where isf_step advects particles), after all work
// queued on the handle so far, the positions are copied device-to-device into staging
// buffer k mod 2 (the stream first waits until the host copy of push k-2 has drained that
// buffer); a second stream then copies the staging buffer into pinned host buffer k mod 2.
// The handle's stream, which carries the grid part of later steps, waits for neither copy;
// work that writes particles after the push is ordered after the staging copy like after
// an advect (join).  previous() waits for the host copy
// of push k-1 and returns its buffer, valid until the next push.
#include <hip/hip_runtime.h>

#include "isf_handle.hpp"
#include "isf_internal.hpp"

using namespace isf;

static constexpr uint32_t kExportMagic = 0xE8B0A7EDu;

struct isf_export {
    uint32_t magic;
    isf_t h;
    long long nslots;
    size_t bytes;
    float* stage[2];      // device staging buffers
    float* host[2];       // pinned host buffers
    hipStream_t copy;     // host copies
    hipEvent_t staged[2], drained[2];
    long long pushed;     // number of pushes so far
};

static int check_export(isf_export_t e) {
    if (!e || e->magic != kExportMagic) return fail(ISF_EHANDLE, "null or destroyed export");
    return ISF_OK;
}

static void export_free(isf_export_t e) {
    if (e->copy) (void)hipStreamSynchronize(e->copy);
    for (int i = 0; i < 2; i++) {
        if (e->stage[i]) (void)hipFree(e->stage[i]);
        if (e->host[i]) (void)hipHostFree(e->host[i]);
        if (e->staged[i]) (void)hipEventDestroy(e->staged[i]);
        if (e->drained[i]) (void)hipEventDestroy(e->drained[i]);
    }
    if (e->copy) (void)hipStreamDestroy(e->copy);
    e->magic = 0;
    delete e;
}

extern "C" {

int isf_export_create(isf_export_t* out, isf_t h, long long nslots) {
    CHECK_PTR(out, "out");
    *out = nullptr;
    int r = check_handle(h);
    if (r) return r;
    if (nslots <= 0) return fail(ISF_EINVAL, "nslots must be > 0");
    isf_export* e = new isf_export();
    e->magic = kExportMagic;
    e->h = h;
    e->nslots = nslots;
    e->bytes = sizeof(float) * 3 * (size_t)nslots;
    hipError_t err = hipStreamCreateWithFlags(&e->copy, hipStreamNonBlocking);
    for (int i = 0; i < 2 && err == hipSuccess; i++) {
        err = hipMalloc(&e->stage[i], e->bytes);
        if (err == hipSuccess) err = hipHostMalloc(&e->host[i], e->bytes, hipHostMallocDefault);
        if (err == hipSuccess) err = hipEventCreateWithFlags(&e->staged[i], hipEventDisableTiming);
        if (err == hipSuccess) err = hipEventCreateWithFlags(&e->drained[i], hipEventDisableTiming);
        // a drained event that is complete before the first push of its buffer
        if (err == hipSuccess) err = hipEventRecord(e->drained[i], e->copy);
    }
    if (err != hipSuccess) {
        export_free(e);
        return fail(ISF_ENOMEM, "particle export buffers: %s", hipGetErrorString(err));
    }
    *out = e;
    return ISF_OK;
}

int isf_export_destroy(isf_export_t e) {
    int r = check_export(e);
    if (r) return r;
    export_free(e);
    return ISF_OK;
}

int isf_export_push(isf_export_t e, const float* parts) {
    int r = check_export(e);
    if (r) return r;
    CHECK_PTR(parts, "parts");
    isf_t h = e->h;
    if ((r = side_after_main(h))) return r;  // the side stream already follows the last advect
    const int b = (int)(e->pushed & 1);
    HIP_CHECK(hipStreamWaitEvent(h->pstream, e->drained[b], 0), "hipStreamWaitEvent");
    HIP_CHECK(hipMemcpyAsync(e->stage[b], parts, e->bytes, hipMemcpyDeviceToDevice, h->pstream), "export stage copy");
    HIP_CHECK(hipEventRecord(e->staged[b], h->pstream), "hipEventRecord");
    HIP_CHECK(hipEventRecord(h->ev_adv, h->pstream), "hipEventRecord");
    h->side_pending = true;
    HIP_CHECK(hipStreamWaitEvent(e->copy, e->staged[b], 0), "hipStreamWaitEvent");
    HIP_CHECK(hipMemcpyAsync(e->host[b], e->stage[b], e->bytes, hipMemcpyDeviceToHost, e->copy), "export host copy");
    HIP_CHECK(hipEventRecord(e->drained[b], e->copy), "hipEventRecord");
    e->pushed++;
    return ISF_OK;
}

int isf_export_previous(isf_export_t e, const float** host, long long* frame) {
    int r = check_export(e);
    if (r) return r;
    CHECK_PTR(host, "host");
    if (e->pushed < 2) return fail(ISF_EINVAL, "no completed previous frame yet (push at least twice)");
    const long long k = e->pushed - 2;
    const int b = (int)(k & 1);
    HIP_CHECK(hipEventSynchronize(e->drained[b]), "hipEventSynchronize");
    *host = e->host[b];
    if (frame) *frame = k;
    return ISF_OK;
}

int isf_export_latest(isf_export_t e, const float** host, long long* frame) {
    int r = check_export(e);
    if (r) return r;
    CHECK_PTR(host, "host");
    if (e->pushed < 1) return fail(ISF_EINVAL, "nothing pushed yet");
    const long long k = e->pushed - 1;
    const int b = (int)(k & 1);
    HIP_CHECK(hipEventSynchronize(e->drained[b]), "hipEventSynchronize");
    *host = e->host[b];
    if (frame) *frame = k;
    return ISF_OK;
}

}  // extern "C"
