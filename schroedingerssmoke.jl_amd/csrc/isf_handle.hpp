// isf_handle.hpp -- the ISF handle (struct isf_handle behind isf_t) and the error
// helpers shared by the C-ABI sources (isf_capi.hip, isf_rollout.hip).  Not part of
// the public interface.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/isf.h"
#include "isf_internal.hpp"

namespace isf {
// set the thread-local message returned by isf_last_error and return code
int fail(int code, const char* fmt, ...);
int hip_fail(hipError_t e, const char* what);
const char* last_error();
}  // namespace isf

#define HIP_CHECK(expr, what)                                         \
    do {                                                              \
        hipError_t e_ = (expr);                                       \
        if (e_ != hipSuccess) return isf::hip_fail(e_, what);         \
    } while (0)
#define LAUNCH_CHECK(expr, what)                                      \
    do {                                                              \
        int e_ = (expr);                                              \
        if (e_ != 0) return isf::hip_fail((hipError_t)e_, what);      \
    } while (0)
#define CHECK_PTR(p, name)                                                         \
    do {                                                                           \
        if (!(p)) return isf::fail(ISF_EINVAL, "%s must not be NULL", name);       \
    } while (0)

using isf::DevTables;
using isf::GridInfo;
using isf::JetParams;

constexpr uint32_t kMagic = 0x15F0CAFEu;

struct isf_handle {
    uint32_t magic;
    int device;
    GridInfo g;
    float hbar, dt;
    size_t N;
    hipStream_t stream;   // the handle's stream: every operator, the grid part of isf_step
    hipStream_t pstream;  // side of isf_step (velocity, spawn, order, advect), joined by events
    hipEvent_t ev0, ev1;
    hipEvent_t ev_vel, ev_adv, ev_vread;
    bool side_pending;    // side work of isf_step may still read psi / vel or write particles
    bool vread_pending;   // the velocity pass on pstream may still read psi (ev_vread)
    hipEvent_t ext_ev;    // recorded on the handle's stream when a batched step must wait for it
    hipEvent_t batch_ev;  // (borrowed) end of the last batched step that used this handle's data
    bool ext_pending;     // the handle's streams must wait for batch_ev before new work
    unsigned long long seq;  // bumped by every call that enqueues work on the handle's streams
    const void* owner;       // the batch that owns this handle (isf_destroy refuses it)
    void* table_mem;
    DevTables tb;
    float2* S;          // Schroedinger spectrum, 2N complex; f and P alias it in projections
    float* vel;         // velocity field, 3N floats
    double* norm_parts; // per-block FP64 partial sums of the last step's gauge pass
    size_t norm_nparts; // room in norm_parts
    size_t norm_count;  // partials the last step's gauge pass wrote (0: no step yet)
    double* norm_out;
    double* norm2_parts;  // partials of isf_norm2 (kNorm2Blocks), kept apart from the step's
    double* norm2_out;
    // jet (isf_jet_configure)
    bool jet_set;
    JetParams jet;
    float omega;
    long long n_spawn;
    unsigned long long seed;
    // spatial processing order of particle slots for the advection
    int* order;
    void* order_scratch;
    size_t order_scratch_bytes;
    long long order_cap;
    const float* order_parts;
    long long order_n;
    int order_iter;
    int sort_interval;
    int order_check_iter;
    bool order_sorted;               // slots are advected through `order` (else slot order)
    unsigned* jump_partial;          // per-block counts of the order measure
    // isf_simloop: second velocity buffer (allocated on first use), the event after each
    // fused forward x pass, and per velocity buffer the event after its last reader
    float* vel_alt;
    hipEvent_t ev_v;
    hipEvent_t ev_buf[2];
    bool buf_rec[2];
    void* vjp_mem;                   // scratch of the operator VJPs (allocated on first use)
    float4* mp_psi;                  // interleaved staging of the MatlabPort PressureProject
    int* vjp_order;                  // cell-sorted slot order of the advection VJP
    void* vjp_order_scratch;
    size_t vjp_order_bytes;
    long long vjp_order_cap;
    long long jump_cap;
};


namespace isf {
int check_handle(isf_t h);
// the spectrum buffer's regions: the Poisson right-hand side (first 4N bytes, also the
// real q of a Poisson solve) and the Poisson half spectrum (from byte 4N)
float* f_of(isf_t h);
float2* P_of(isf_t h);
double poisson_scale(isf_t h);
// pressure projection of interleaved psi: one-form (hbar 1) + div + Poisson + gauge
// (c2r_mode 1), or + jet restriction (mode 2)
int project(isf_t h, float* psi, int c2r_mode, const JetParams* jet, double* parts);
// order the handle's streams after its pending side-stream and batched work
int join(isf_t h);
// order the side stream (pstream) after all work queued on the handle's stream
int side_after_main(isf_t h);
}  // namespace isf
