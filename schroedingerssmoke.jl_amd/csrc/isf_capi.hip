// isf_capi.hip -- host side of libisf: the C ABI declared in include/isf.h.
//
// The handle plays the role of the reference's ISF struct (src/parallel.jl:4-35): grid
// constants, the spectral tables (here per-axis, derived in FP64), the scratch arrays
// and the velocity field, plus the HIP stream every call is enqueued on.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/isf.h"
#include "isf_handle.hpp"
#include "isf_internal.hpp"

using namespace isf;

namespace isf {
static thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}
int hip_fail(hipError_t e, const char* what) {
    return fail(ISF_EHIP, "%s: %s", what, hipGetErrorString(e));
}
const char* last_error() { return g_err.c_str(); }
}  // namespace isf

float* isf::f_of(isf_t h) { return reinterpret_cast<float*>(h->S); }
float2* isf::P_of(isf_t h) { return reinterpret_cast<float2*>(reinterpret_cast<char*>(h->S) + 4 * h->N); }

int isf::check_handle(isf_t h) {
    if (!h || h->magic != kMagic) return fail(ISF_EHANDLE, "null or destroyed isf handle");
    return ISF_OK;
}
#define CHECK_H(h)                              \
    do {                                        \
        int r_ = check_handle(h);               \
        if (r_) return r_;                      \
    } while (0)
// Order the handle's streams after a batched step that used this handle's buffers.
static int wait_ext(isf_t h) {
    h->seq++;
    if (h->ext_pending) {
        hipError_t e = hipStreamWaitEvent(h->stream, h->batch_ev, 0);
        if (e == hipSuccess) e = hipStreamWaitEvent(h->pstream, h->batch_ev, 0);
        if (e != hipSuccess) return fail(ISF_EHIP, "hipStreamWaitEvent: %s", hipGetErrorString(e));
        h->ext_pending = false;
    }
    return ISF_OK;
}
// Order the side stream after all work queued on the handle's stream (ev_vel is recorded
// afresh wherever the side stream must follow the handle's stream).
int isf::side_after_main(isf_t h) {
    int r = wait_ext(h);
    if (r) return r;
    HIP_CHECK(hipEventRecord(h->ev_vel, h->stream), "hipEventRecord");
    HIP_CHECK(hipStreamWaitEvent(h->pstream, h->ev_vel, 0), "hipStreamWaitEvent");
    return ISF_OK;
}
// Order the handle's stream after a pending particle-side advect of isf_step.
int isf::join(isf_t h) {
    int r = wait_ext(h);
    if (r) return r;
    if (h->side_pending) {
        hipError_t e = hipStreamWaitEvent(h->stream, h->ev_adv, 0);
        if (e != hipSuccess) return fail(ISF_EHIP, "hipStreamWaitEvent: %s", hipGetErrorString(e));
        h->side_pending = false;
        h->vread_pending = false;  // ev_adv is recorded after the velocity pass
    }
    return ISF_OK;
}
#define CHECK_HJ(h)                             \
    do {                                        \
        int r_ = check_handle(h);               \
        if (!r_) r_ = join(h);                  \
        if (r_) return r_;                      \
    } while (0)

extern "C" {

const char* isf_last_error(void) { return isf::last_error(); }
const char* isf_version(void) { return "libisf 0.1 (gfx950)"; }

int isf_device_count(int* out) {
    CHECK_PTR(out, "out");
    HIP_CHECK(hipGetDeviceCount(out), "hipGetDeviceCount");
    return ISF_OK;
}
int isf_malloc(void** ptr, size_t bytes) {
    CHECK_PTR(ptr, "ptr");
    *ptr = nullptr;
    if (bytes == 0) return ISF_OK;
    hipError_t e = hipMalloc(ptr, bytes);
    if (e != hipSuccess) return fail(ISF_ENOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    return ISF_OK;
}
int isf_free(void* ptr) {
    if (ptr) HIP_CHECK(hipFree(ptr), "hipFree");
    return ISF_OK;
}
// The copy and set helpers are synchronous: they first wait for all work queued on the
// device (including every handle's stream), so results of earlier calls are complete.
int isf_memcpy_h2d(void* dst, const void* src, size_t bytes) {
    HIP_CHECK(hipDeviceSynchronize(), "hipDeviceSynchronize");
    HIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "hipMemcpy H2D");
    return ISF_OK;
}
int isf_memcpy_d2h(void* dst, const void* src, size_t bytes) {
    HIP_CHECK(hipDeviceSynchronize(), "hipDeviceSynchronize");
    HIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "hipMemcpy D2H");
    return ISF_OK;
}
int isf_memcpy_d2d(void* dst, const void* src, size_t bytes) {
    HIP_CHECK(hipDeviceSynchronize(), "hipDeviceSynchronize");
    HIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice), "hipMemcpy D2D");
    return ISF_OK;
}
int isf_memset(void* dst, int value, size_t bytes) {
    HIP_CHECK(hipDeviceSynchronize(), "hipDeviceSynchronize");
    HIP_CHECK(hipMemset(dst, value, bytes), "hipMemset");
    return ISF_OK;
}

static bool pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }

int isf_create(isf_t* out, const int res[3], const int box[3], float hbar, float dt) {
    CHECK_PTR(out, "out");
    CHECK_PTR(res, "res");
    CHECK_PTR(box, "box");
    *out = nullptr;
    for (int a = 0; a < 3; a++) {
        if (!pow2(res[a])) return fail(ISF_EINVAL, "grid_res[%d] = %d is not a power of two", a, res[a]);
        if (box[a] <= 0) return fail(ISF_EINVAL, "physical_size[%d] = %d must be positive", a, box[a]);
    }
    if (!x_len_supported(res[0])) return fail(ISF_EINVAL, "grid_res[0] = %d unsupported (16..512)", res[0]);
    if (!plane_supported(res[1], res[2]))
        return fail(ISF_EINVAL, "grid_res[1:2] = (%d, %d) unsupported (16..512 each)", res[1], res[2]);
    if (!(hbar > 0.0f) || !(dt > 0.0f)) return fail(ISF_EINVAL, "hbar and dt must be positive");

    isf_handle* h = new isf_handle();
    std::memset(h, 0, sizeof *h);
    h->magic = kMagic;
    hipError_t e = hipGetDevice(&h->device);
    if (e != hipSuccess) {
        delete h;
        return hip_fail(e, "hipGetDevice");
    }
    GridInfo& g = h->g;
    g.nx = res[0];
    g.ny = res[1];
    g.nz = res[2];
    // d = VF(physical_size ./ grid_res), src/parallel.jl:46
    g.dx = (float)((double)box[0] / res[0]);
    g.dy = (float)((double)box[1] / res[1]);
    g.dz = (float)((double)box[2] / res[2]);
    g.Lx = (float)box[0];
    g.Ly = (float)box[1];
    g.Lz = (float)box[2];
    g.idx = 1.0f / g.dx;
    g.idy = 1.0f / g.dy;
    g.idz = 1.0f / g.dz;
    {
        int ex;
        const float dd3[3] = {g.dx, g.dy, g.dz};
        g.d_pow2 = 1;
        for (int a = 0; a < 3; a++) g.d_pow2 &= (std::frexp(dd3[a], &ex) == 0.5f);
    }
    h->hbar = hbar;
    h->dt = dt;
    h->N = (size_t)g.nx * g.ny * g.nz;
    h->sort_interval = 32;

    // ---- per-axis tables (host, FP64)
    const int nmax = 512;
    std::vector<float2> tw(2 * nmax);
    for (int n = 2; n <= nmax; n *= 2)
        for (int m = 0; m < n; m++) {
            const double a = -2.0 * M_PI * (double)m / (double)n;
            tw[n + m] = make_float2((float)std::cos(a), (float)std::sin(a));
        }
    // Schroedinger mask, src/parallel.jl:49,58-62 in FFT order: exp(i c k^2) per axis,
    // c = -4 pi^2 hbar dt / 2, k = j/L (j < n/2) or (j-n)/L; 1/N folded into the x factor.
    const double c = -4.0 * M_PI * M_PI * (double)hbar * (double)dt / 2.0;
    const int n3[3] = {g.nx, g.ny, g.nz};
    const double dd[3] = {(double)g.dx, (double)g.dy, (double)g.dz};
    std::vector<float2> mask[3];
    std::vector<double> s2[3];
    for (int a = 0; a < 3; a++) {
        mask[a].resize(n3[a]);
        s2[a].resize(n3[a]);
        const double scale = a == 0 ? 1.0 / (double)h->N : 1.0;
        for (int j = 0; j < n3[a]; j++) {
            const double k = (j < n3[a] / 2 ? j : j - n3[a]) / (double)box[a];
            const double ph = c * k * k;
            mask[a][j] = make_float2((float)(std::cos(ph) * scale), (float)(std::sin(ph) * scale));
            const double s = std::sin(M_PI * j / n3[a]) / dd[a];  // src/parallel.jl:53
            s2[a][j] = s * s;
        }
    }
    size_t bytes = tw.size() * sizeof(float2);
    size_t off_mask[3], off_s2[3];
    for (int a = 0; a < 3; a++) {
        off_mask[a] = bytes;
        bytes += ((n3[a] * sizeof(float2) + 255) / 256) * 256;
    }
    for (int a = 0; a < 3; a++) {
        off_s2[a] = bytes;
        bytes += ((n3[a] * sizeof(double) + 255) / 256) * 256;
    }
    auto cleanup = [&](int code) {
        isf_destroy(h);
        return code;
    };
    e = hipMalloc(&h->table_mem, bytes);
    if (e != hipSuccess) return cleanup(fail(ISF_ENOMEM, "tables: %s", hipGetErrorString(e)));
    char* base = (char*)h->table_mem;
    std::vector<char> host(bytes, 0);
    std::memcpy(host.data(), tw.data(), tw.size() * sizeof(float2));
    for (int a = 0; a < 3; a++) {
        std::memcpy(host.data() + off_mask[a], mask[a].data(), n3[a] * sizeof(float2));
        std::memcpy(host.data() + off_s2[a], s2[a].data(), n3[a] * sizeof(double));
    }
    e = hipMemcpy(base, host.data(), bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) return cleanup(hip_fail(e, "tables upload"));
    h->tb.tw = (float2*)base;
    h->tb.mask_x = (float2*)(base + off_mask[0]);
    h->tb.mask_y = (float2*)(base + off_mask[1]);
    h->tb.mask_z = (float2*)(base + off_mask[2]);
    h->tb.s2x = (double*)(base + off_s2[0]);
    h->tb.s2y = (double*)(base + off_s2[1]);
    h->tb.s2z = (double*)(base + off_s2[2]);

    // ---- scratch
    if ((e = hipMalloc(&h->S, 16 * h->N)) != hipSuccess) return cleanup(fail(ISF_ENOMEM, "scratch: %s", hipGetErrorString(e)));
    if ((e = hipMalloc(&h->vel, 12 * h->N)) != hipSuccess) return cleanup(fail(ISF_ENOMEM, "velocity: %s", hipGetErrorString(e)));
    {
        // partials of the gauge pass for the smallest x tile (8 lines per block)
        const size_t most = (size_t)(g.ny / (g.ny < 8 ? g.ny : 8)) * g.nz;
        h->norm_nparts = most > 1024 ? most : 1024;
    }
    // [step partials | step sum | isf_norm2 partials | isf_norm2 sum]
    const size_t norm_doubles = h->norm_nparts + 1 + kNorm2Blocks + 1;
    if ((e = hipMalloc(&h->norm_parts, sizeof(double) * norm_doubles)) != hipSuccess)
        return cleanup(fail(ISF_ENOMEM, "norm partials: %s", hipGetErrorString(e)));
    h->norm_out = h->norm_parts + h->norm_nparts;
    h->norm2_parts = h->norm_out + 1;
    h->norm2_out = h->norm2_parts + kNorm2Blocks;
    if ((e = hipMemset(h->vel, 0, 12 * h->N)) != hipSuccess) return cleanup(hip_fail(e, "hipMemset"));
    if ((e = hipMemset(h->norm_parts, 0, sizeof(double) * norm_doubles)) != hipSuccess)
        return cleanup(hip_fail(e, "hipMemset"));
    // a blocking stream: it also orders against the null stream used by plain hipMemcpy
    if ((e = hipStreamCreate(&h->stream)) != hipSuccess)
        return cleanup(hip_fail(e, "hipStreamCreate"));
    if ((e = hipStreamCreate(&h->pstream)) != hipSuccess) return cleanup(hip_fail(e, "hipStreamCreate"));
    if ((e = hipEventCreateWithFlags(&h->ev_vel, hipEventDisableTiming)) != hipSuccess)
        return cleanup(hip_fail(e, "hipEventCreate"));
    if ((e = hipEventCreateWithFlags(&h->ev_vread, hipEventDisableTiming)) != hipSuccess)
        return cleanup(hip_fail(e, "hipEventCreate"));
    if ((e = hipEventCreateWithFlags(&h->ev_adv, hipEventDisableTiming)) != hipSuccess)
        return cleanup(hip_fail(e, "hipEventCreate"));
    if ((e = hipEventCreateWithFlags(&h->ext_ev, hipEventDisableTiming)) != hipSuccess)
        return cleanup(hip_fail(e, "hipEventCreate"));
    if ((e = hipEventCreate(&h->ev0)) != hipSuccess) return cleanup(hip_fail(e, "hipEventCreate"));
    if ((e = hipEventCreate(&h->ev1)) != hipSuccess) return cleanup(hip_fail(e, "hipEventCreate"));
    *out = h;
    return ISF_OK;
}

int isf_destroy(isf_t h) {
    CHECK_H(h);
    if (h->owner) return fail(ISF_EINVAL, "this handle belongs to a batch; destroy the batch instead");
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    if (h->pstream) (void)hipStreamSynchronize(h->pstream);
    if (h->ev_vel) (void)hipEventDestroy(h->ev_vel);
    if (h->ev_adv) (void)hipEventDestroy(h->ev_adv);
    if (h->ev_vread) (void)hipEventDestroy(h->ev_vread);
    if (h->ext_ev) (void)hipEventDestroy(h->ext_ev);
    if (h->pstream) (void)hipStreamDestroy(h->pstream);
    if (h->ev0) (void)hipEventDestroy(h->ev0);
    if (h->ev1) (void)hipEventDestroy(h->ev1);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    if (h->S) (void)hipFree(h->S);
    if (h->vel) (void)hipFree(h->vel);
    if (h->norm_parts) (void)hipFree(h->norm_parts);
    if (h->order) (void)hipFree(h->order);
    if (h->order_scratch) (void)hipFree(h->order_scratch);
    if (h->jump_partial) (void)hipFree(h->jump_partial);
    if (h->vjp_mem) (void)hipFree(h->vjp_mem);
    if (h->vel_alt) (void)hipFree(h->vel_alt);
    if (h->ev_v) (void)hipEventDestroy(h->ev_v);
    for (hipEvent_t e : h->ev_buf)
        if (e) (void)hipEventDestroy(e);
    if (h->mp_psi) (void)hipFree(h->mp_psi);
    if (h->vjp_order) (void)hipFree(h->vjp_order);
    if (h->vjp_order_scratch) (void)hipFree(h->vjp_order_scratch);
    if (h->table_mem) (void)hipFree(h->table_mem);
    h->magic = 0;
    delete h;
    return ISF_OK;
}

int isf_synchronize(isf_t h) {
    CHECK_HJ(h);
    HIP_CHECK(hipStreamSynchronize(h->stream), "hipStreamSynchronize");
    return ISF_OK;
}

int isf_get_grid(isf_t h, int res[3], float d[3], float box[3], float* hbar, float* dt) {
    CHECK_H(h);
    if (res) {
        res[0] = h->g.nx;
        res[1] = h->g.ny;
        res[2] = h->g.nz;
    }
    if (d) {
        d[0] = h->g.dx;
        d[1] = h->g.dy;
        d[2] = h->g.dz;
    }
    if (box) {
        box[0] = h->g.Lx;
        box[1] = h->g.Ly;
        box[2] = h->g.Lz;
    }
    if (hbar) *hbar = h->hbar;
    if (dt) *dt = h->dt;
    return ISF_OK;
}

int isf_velocity_ptr(isf_t h, float** out) {
    CHECK_HJ(h);
    CHECK_PTR(out, "out");
    *out = h->vel;
    return ISF_OK;
}

// ------------------------------------------------------------------ operators

static int flow(isf_t h, float* psi, bool normalize) {
    float4* p = reinterpret_cast<float4*>(psi);
    LAUNCH_CHECK(launch_flow_x_fwd(h->g, p, h->S, h->tb, h->stream), "flow x forward");
    LAUNCH_CHECK(launch_flow_plane(h->g, h->S, h->tb, h->stream), "flow plane");
    LAUNCH_CHECK(launch_flow_x_inv(h->g, h->S, p, normalize, h->tb, h->stream), "flow x inverse");
    return ISF_OK;
}

}  // extern "C"

double isf::poisson_scale(isf_t h) { return -0.25 * 2.0 / (double)h->N; }

// velocity one-form (hbar=1) + div + Poisson + gauge; mode C2R_GAUGE(_RESTRICT)
int isf::project(isf_t h, float* psi, int c2r_mode, const JetParams* jet, double* parts) {
    float4* p = reinterpret_cast<float4*>(psi);
    LAUNCH_CHECK(launch_div_r2c(h->g, p, f_of(h), P_of(h), h->tb, h->stream), "div(one_form(psi)) + poisson x forward");
    LAUNCH_CHECK(launch_poisson_plane(h->g, P_of(h), poisson_scale(h), h->tb, h->stream), "poisson plane");
    LAUNCH_CHECK(launch_c2r_x(h->g, P_of(h), c2r_mode, nullptr, p, jet, parts, h->tb, h->stream), "poisson x inverse + gauge");
    return ISF_OK;
}

extern "C" {

int isf_schroedinger_flow(isf_t h, float* psi) {
    CHECK_HJ(h);
    CHECK_PTR(psi, "psi");
    return flow(h, psi, false);
}

int isf_normalize(isf_t h, float* psi) {
    CHECK_HJ(h);
    CHECK_PTR(psi, "psi");
    LAUNCH_CHECK(launch_normalize(h->g, reinterpret_cast<float4*>(psi), h->stream), "normalize");
    return ISF_OK;
}

int isf_velocity_one_form(isf_t h, const float* psi, float hbar, int sharp, float* vel) {
    CHECK_HJ(h);
    CHECK_PTR(psi, "psi");
    LAUNCH_CHECK(launch_velocity(h->g, reinterpret_cast<const float4*>(psi), hbar, sharp, vel ? vel : h->vel, h->stream),
                 "velocity_one_form");
    return ISF_OK;
}

int isf_staggered_sharp(isf_t h, float* vel) {
    CHECK_HJ(h);
    LAUNCH_CHECK(launch_sharp(h->g, vel ? vel : h->vel, h->stream), "StaggeredSharp");
    return ISF_OK;
}

int isf_div(isf_t h, const float* vel, float* f) {
    CHECK_HJ(h);
    CHECK_PTR(f, "f");
    LAUNCH_CHECK(launch_div_vel(h->g, vel ? vel : h->vel, f, h->stream), "div");
    return ISF_OK;
}

int isf_poisson_solve(isf_t h, const float* f, float* q) {
    CHECK_HJ(h);
    CHECK_PTR(f, "f");
    CHECK_PTR(q, "q");
    LAUNCH_CHECK(launch_r2c_x(h->g, f, P_of(h), h->tb, h->stream), "poisson x forward");
    LAUNCH_CHECK(launch_poisson_plane(h->g, P_of(h), poisson_scale(h), h->tb, h->stream), "poisson plane");
    LAUNCH_CHECK(launch_c2r_x(h->g, P_of(h), 0, q, nullptr, nullptr, nullptr, h->tb, h->stream), "poisson x inverse");
    return ISF_OK;
}

int isf_gauge_transform(isf_t h, float* psi, const float* q) {
    CHECK_HJ(h);
    CHECK_PTR(psi, "psi");
    CHECK_PTR(q, "q");
    LAUNCH_CHECK(launch_gauge(h->g, reinterpret_cast<float4*>(psi), q, h->stream), "gauge_transform");
    return ISF_OK;
}

int isf_pressure_project(isf_t h, float* psi) {
    CHECK_HJ(h);
    CHECK_PTR(psi, "psi");
    return project(h, psi, 1, nullptr, nullptr);
}

static JetParams jet_from(const float cen[3], float len, float rad) {
    JetParams j{};
    j.cx = cen[0];
    j.cy = cen[1];
    j.cz = cen[2];
    j.len = len;
    j.rad = rad;
    return j;
}

int isf_isjet(isf_t h, const float cen[3], float len, float rad, unsigned char* out) {
    CHECK_HJ(h);
    CHECK_PTR(cen, "cen");
    CHECK_PTR(out, "out");
    LAUNCH_CHECK(launch_isjet(h->g, jet_from(cen, len, rad), out, h->stream), "isjet");
    return ISF_OK;
}

int isf_restrict_velocity(isf_t h, float* psi, const unsigned char* isjet, const float kvec[3], float omgterm,
                          int phase_mode) {
    CHECK_HJ(h);
    CHECK_PTR(psi, "psi");
    CHECK_PTR(kvec, "kvec");
    if (phase_mode != ISF_PHASE_DOT && phase_mode != ISF_PHASE_LEGACY) return fail(ISF_EINVAL, "bad phase_mode %d", phase_mode);
    if (!isjet && !h->jet_set) return fail(ISF_EINVAL, "isjet is NULL and no nozzle was configured");
    JetParams j = h->jet_set ? h->jet : JetParams{};
    j.kx = kvec[0];
    j.ky = kvec[1];
    j.kz = kvec[2];
    j.omgterm = omgterm;
    j.phase_mode = phase_mode;
    LAUNCH_CHECK(launch_restrict(h->g, reinterpret_cast<float4*>(psi), isjet, j, h->stream), "restrict_velocity");
    return ISF_OK;
}

int isf_add_particles(isf_t h, float* parts, long long nslots, long long first, long long n, const float cen[3],
                      float rad, unsigned long long seed, unsigned step) {
    CHECK_HJ(h);
    CHECK_PTR(parts, "parts");
    CHECK_PTR(cen, "cen");
    if (nslots <= 0 || first < 0 || n < 0) return fail(ISF_EINVAL, "bad particle range");
    LAUNCH_CHECK(launch_spawn(parts, nslots, first % nslots, n, jet_from(cen, 0.0f, rad), seed, step, h->stream),
                 "add_particles");
    return ISF_OK;
}

int isf_staggered_advect(isf_t h, float* parts, long long nslots, const float* vel) {
    CHECK_HJ(h);
    CHECK_PTR(parts, "parts");
    if (nslots < 0) return fail(ISF_EINVAL, "bad nslots");
    LAUNCH_CHECK(launch_advect(h->g, parts, nslots, nullptr, vel ? vel : h->vel, h->dt, h->stream), "staggered_advect");
    return ISF_OK;
}

int isf_jet_configure(isf_t h, const float kvec[3], float omega, const float cen[3], float len, float rad,
                      long long n_spawn, unsigned long long seed, int phase_mode) {
    CHECK_H(h);
    CHECK_PTR(kvec, "kvec");
    CHECK_PTR(cen, "cen");
    if (phase_mode != ISF_PHASE_DOT && phase_mode != ISF_PHASE_LEGACY) return fail(ISF_EINVAL, "bad phase_mode %d", phase_mode);
    if (n_spawn < 0) return fail(ISF_EINVAL, "n_spawn must be >= 0");
    JetParams j = jet_from(cen, len, rad);
    j.kx = kvec[0];
    j.ky = kvec[1];
    j.kz = kvec[2];
    j.phase_mode = phase_mode;
    h->jet = j;
    h->omega = omega;
    h->n_spawn = n_spawn;
    h->seed = seed;
    h->jet_set = true;
    return ISF_OK;
}

int isf_fill_nozzle(isf_t h, float* parts, long long nslots, long long first, long long n, unsigned long long seed) {
    CHECK_HJ(h);
    CHECK_PTR(parts, "parts");
    if (!h->jet_set) return fail(ISF_EINVAL, "isf_jet_configure must be called first");
    if (nslots <= 0 || first < 0 || n < 0) return fail(ISF_EINVAL, "bad particle range");
    LAUNCH_CHECK(launch_fill_cylinder(parts, nslots, first % nslots, n, h->jet, seed, h->stream), "fill_nozzle");
    return ISF_OK;
}

int isf_constrain(isf_t h, float* psi, int iterations) {
    CHECK_HJ(h);
    CHECK_PTR(psi, "psi");
    if (!h->jet_set) return fail(ISF_EINVAL, "isf_jet_configure must be called first");
    JetParams j = h->jet;
    j.omgterm = 0.0f;  // test/runtests.jl:28
    for (int it = 0; it < iterations; it++) {
        LAUNCH_CHECK(launch_restrict(h->g, reinterpret_cast<float4*>(psi), nullptr, j, h->stream), "restrict_velocity");
        int r = project(h, psi, 1, nullptr, nullptr);
        if (r) return r;
    }
    return ISF_OK;
}

static int ensure_order(isf_t h, const float* parts, long long nslots, int iter, const int** order,
                        hipStream_t stream);

// Stage names of one step, in launch order (isf_step_profile reports one time each).
static const char* kStageNames[] = {
    "flow_x_fwd", "flow_plane", "flow_x_inv_normalize",                          // :48-49
    "p1_div_r2c_x", "p1_plane", "p1_c2r_x_gauge_restrict",                       // :50, :57
    "spawn",                                                                     // :52-54
    "p2_div_r2c_x", "p2_plane", "p2_c2r_x_gauge",                                // :58
    "velocity_sharp", "particle_order", "advect"};                               // :59-63
static constexpr int kNumStages = (int)(sizeof(kStageNames) / sizeof(kStageNames[0]));

// One simloop body (test/runtests.jl:45-64).  The transforms and projections run on the
// handle's stream; the velocity pass and the particle part (spawn, processing order,
// advect) run on pstream, in the order velocity(k-1), order, advect(k-1), spawn(k): the
// velocity of step k after the last projection of step k (event), and the inverse flow
// pass of step k+1, the next write of psi, after that velocity pass (event).  The
// bandwidth-bound velocity pass and the latency-bound particle work so overlap the next
// step's forward flow transforms.  If ev is non-null (profiling, kNumStages + 1 events)
// everything runs on the handle's stream in stage order, ev[0] recorded before the first
// stage and ev[k + 1] after stage k.
static int step_impl(isf_t h, float* psi, float* parts, long long nslots, int iter, hipEvent_t* ev) {
    const float t = (float)iter * h->dt;  // test/runtests.jl:46
    JetParams j = h->jet;
    j.omgterm = h->omega * t;             // test/runtests.jl:57
    float4* p = reinterpret_cast<float4*>(psi);
    hipStream_t s = h->stream;
    // Outside profiling the particle work and the velocity pass run on the side stream:
    // the velocity field is read only by the advect, and the next step's forward flow pass
    // only reads psi, so both overlap it; its inverse pass, the next write of psi, waits
    // for the velocity pass (ev_vread).
    const bool side = ev == nullptr;
    hipStream_t ps = side ? h->pstream : s;
    {
        int r = wait_ext(h);
        if (r) return r;
    }
    int k = 0;
    auto mark = [&]() -> int {
        if (ev) {
            hipError_t e = hipEventRecord(ev[++k], s);
            if (e != hipSuccess) return hip_fail(e, "hipEventRecord");
        }
        return ISF_OK;
    };
#define STAGE(expr, what)              \
    do {                               \
        LAUNCH_CHECK(expr, what);      \
        int m_ = mark();               \
        if (m_) return m_;             \
    } while (0)
    if (ev) {
        int r = join(h);
        if (r) return r;
        HIP_CHECK(hipEventRecord(ev[0], s), "hipEventRecord");
    }
    // :52-54 add_particles! into the ring buffer (first 0-based slot of the batch)
    long long first0 = 0, nsp = 0;
    if (nslots > 0 && h->n_spawn > 0) {
        const long long start1 = ((long long)(iter - 1) * h->n_spawn + 1) % nslots;
        first0 = start1 == 0 ? nslots - 1 : start1 - 1;
        nsp = h->n_spawn;
    }
    if (side && nslots > 0)
        LAUNCH_CHECK(launch_spawn(parts, nslots, first0, nsp, h->jet, h->seed, (unsigned)iter, ps), "add_particles");
    // :48-49 schroedinger_flow! + normalize
    STAGE(launch_flow_x_fwd(h->g, p, h->S, h->tb, s), "flow x forward");
    STAGE(launch_flow_plane(h->g, h->S, h->tb, s), "flow plane");
    if (h->vread_pending) {
        HIP_CHECK(hipStreamWaitEvent(s, h->ev_vread, 0), "hipStreamWaitEvent");
        h->vread_pending = false;
    }
    STAGE(launch_flow_x_inv(h->g, h->S, p, true, h->tb, s), "flow x inverse + normalize");
    // :50 pressure_project! with the restrict of :57 fused into its gauge pass
    const double sc = poisson_scale(h);
    STAGE(launch_div_r2c(h->g, p, f_of(h), P_of(h), h->tb, s), "div(one_form(psi)) + poisson x forward");
    STAGE(launch_poisson_plane(h->g, P_of(h), sc, h->tb, s), "poisson plane");
    STAGE(launch_c2r_x(h->g, P_of(h), 2, nullptr, p, &j, nullptr, h->tb, s), "poisson x inverse + gauge + restrict");
    if (!side) STAGE(launch_spawn(parts, nslots, first0, nsp, h->jet, h->seed, (unsigned)iter, s), "add_particles");
    // :58 pressure_project! (FP64 norm partials accumulated in its gauge pass)
    STAGE(launch_div_r2c(h->g, p, f_of(h), P_of(h), h->tb, s), "div(one_form(psi)) + poisson x forward");
    STAGE(launch_poisson_plane(h->g, P_of(h), sc, h->tb, s), "poisson plane");
    STAGE(launch_c2r_x(h->g, P_of(h), 1, nullptr, p, nullptr, h->norm_parts, h->tb, s), "poisson x inverse + gauge");
    h->norm_count = c2r_x_blocks(h->g);
    // :59-62 velocity_one_form!(isf, psi, hbar) and the in-place StaggeredSharp, after the
    // previous step's advect is done reading vel (same side stream, or joined above)
    if (side) {
        HIP_CHECK(hipEventRecord(h->ev_vel, s), "hipEventRecord");
        HIP_CHECK(hipStreamWaitEvent(ps, h->ev_vel, 0), "hipStreamWaitEvent");
        LAUNCH_CHECK(launch_velocity(h->g, p, h->hbar, 1, h->vel, ps), "velocity_one_form");
        HIP_CHECK(hipEventRecord(h->ev_vread, ps), "hipEventRecord");
        h->vread_pending = true;
    } else {
        STAGE(launch_velocity(h->g, p, h->hbar, 1, h->vel, s), "velocity_one_form");
    }
    // :63 staggered_advect!, slots processed in a spatial order when the slot order is
    // scattered (particles move less than a cell per step)
    if (nslots > 0 || !side) {
        const int* order = nullptr;
        int r = ensure_order(h, parts, nslots, iter, &order, ps);
        if (r) return r;
        if ((r = mark())) return r;
        STAGE(launch_advect(h->g, parts, nslots, order, h->vel, h->dt, ps, Batch{}, 0, nullptr, side),
              "staggered_advect");
    }
    if (side) {
        HIP_CHECK(hipEventRecord(h->ev_adv, ps), "hipEventRecord");
        h->side_pending = true;
    }
#undef STAGE
    return ISF_OK;
}

// Particle processing order.  When the buffer changes, and every 256 steps after, the
// slot order is measured (one host synchronisation): if more than a quarter of the
// consecutive slot pairs are more than two cells apart, slots are advected through a
// spatial sort refreshed every sort_interval steps; otherwise in slot order.
static int ensure_order(isf_t h, const float* parts, long long nslots, int iter, const int** order,
                        hipStream_t stream) {
    *order = nullptr;
    if (nslots <= 0 || h->sort_interval <= 0) return ISF_OK;
    const bool changed = parts != h->order_parts || nslots != h->order_n;
    if (changed || iter < h->order_check_iter || iter - h->order_check_iter >= 256) {
        if (nslots > h->jump_cap) {
            if (h->jump_partial) (void)hipFree(h->jump_partial);
            h->jump_partial = nullptr;
            h->jump_cap = 0;
            hipError_t e = hipMalloc(&h->jump_partial, sizeof(unsigned) * order_measure_ints(nslots));
            if (e != hipSuccess) return fail(ISF_ENOMEM, "order measure: %s", hipGetErrorString(e));
            h->jump_cap = nslots;
        }
        unsigned long long jumps = 0;
        LAUNCH_CHECK(measure_order_jumps(h->g, parts, nslots, h->jump_partial, &jumps, stream), "particle order measure");
        h->order_sorted = jumps * 4 > (unsigned long long)nslots;
        h->order_check_iter = iter;
        h->order_parts = parts;
        h->order_n = nslots;
        h->order_iter = iter - h->sort_interval;  // force a refresh below if sorting
    }
    if (!h->order_sorted) return ISF_OK;
    if (nslots > h->order_cap) {
        if (h->order) (void)hipFree(h->order);
        if (h->order_scratch) (void)hipFree(h->order_scratch);
        h->order = nullptr;
        h->order_scratch = nullptr;
        h->order_cap = 0;
        const size_t sb = order_scratch_bytes(h->g, nslots);
        hipError_t e = hipMalloc(&h->order, sizeof(int) * nslots);
        if (e == hipSuccess) e = hipMalloc(&h->order_scratch, sb);
        if (e != hipSuccess) return fail(ISF_ENOMEM, "particle order: %s", hipGetErrorString(e));
        h->order_scratch_bytes = sb;
        h->order_cap = nslots;
        h->order_iter = iter - h->sort_interval;
    }
    if (iter < h->order_iter || iter - h->order_iter >= h->sort_interval) {
        LAUNCH_CHECK(launch_particle_order(h->g, parts, nslots, h->order, h->order_scratch, h->order_scratch_bytes,
                                           stream),
                     "particle order");
        h->order_iter = iter;
    }
    *order = h->order;
    return ISF_OK;
}

int isf_set_sort_interval(isf_t h, int steps) {
    CHECK_H(h);
    if (steps < 0) return fail(ISF_EINVAL, "sort interval must be >= 0");
    h->sort_interval = steps;
    h->order_parts = nullptr;
    h->order_n = 0;
    return ISF_OK;
}

static int step_args_ok(isf_t h, float* psi, float* parts, long long nslots, int iter) {
    CHECK_H(h);
    CHECK_PTR(psi, "psi");
    if (!h->jet_set) return fail(ISF_EINVAL, "isf_jet_configure must be called first");
    if (iter < 1) return fail(ISF_EINVAL, "iter must be >= 1");
    if (nslots < 0 || (nslots > 0 && !parts)) return fail(ISF_EINVAL, "bad particle buffer");
    return ISF_OK;
}

int isf_step(isf_t h, float* psi, float* parts, long long nslots, int iter) {
    int r = step_args_ok(h, psi, parts, nslots, iter);
    if (r) return r;
    return step_impl(h, psi, parts, nslots, iter, nullptr);
}

// The driver's simloop (test/runtests.jl:40-66): steps iter0 .. iter0 + count - 1.  The
// same kernels as isf_step, pipelined across steps: the forward flow x pass of step k + 1
// also forms the velocity field of step k from the psi it reads (the one-form needs
// psi(c + e_y) and psi(c + e_z), which no pass of step k has at hand), into one of two
// alternating velocity buffers, and step k's advection runs on the side stream after
// that pass (one event), overlapping step k + 1's remaining transforms.  psi is not read
// on the side stream except by the last step's velocity pass, so the inverse flow pass
// waits for nothing.  Before a fused pass rewrites a velocity buffer, the host waits for
// the event after that buffer's last reader (the advection two steps back, long done by
// then), so no stream waits on it.  Results are bit-identical to count calls of isf_step.
static int simloop_impl(isf_t h, float* psi, float* parts, long long nslots, int iter0, int count) {
    // the second velocity buffer (a batched step loop may have allocated it already) and
    // the loop's events
    if (!h->vel_alt) {
        const hipError_t e = hipMalloc(&h->vel_alt, 12 * h->N);
        if (e != hipSuccess) return fail(ISF_ENOMEM, "simloop velocity buffer: %s", hipGetErrorString(e));
    }
    if (!h->ev_v) {
        hipError_t e;
        // completed by the fused forward pass's own launch (a stop event of hipExtLaunchKernel);
        // only the side stream waits on it, so it needs no system-scope release (measured
        // 0.2711-0.2713 against 0.2723-0.2732 ms per step with the default release)
        if ((e = hipEventCreateWithFlags(&h->ev_v, hipEventDisableSystemFence)) != hipSuccess)
            return hip_fail(e, "hipEventCreate");
        for (int b = 0; b < 2; b++)
            if ((e = hipEventCreateWithFlags(&h->ev_buf[b], hipEventDisableTiming)) != hipSuccess)
                return hip_fail(e, "hipEventCreate");
    }
    {
        int r = join(h);  // earlier side work (readers of h->vel and psi) completes first
        if (r) return r;
    }
    hipStream_t s = h->stream, ps = h->pstream;
    float4* p = reinterpret_cast<float4*>(psi);
    float* bufs[2] = {h->vel, h->vel_alt};
    h->buf_rec[0] = h->buf_rec[1] = false;
    const double sc = poisson_scale(h);
    const VelOut vo0{nullptr, h->hbar, 1.0f / h->g.dx, 1.0f / h->g.dy, 1.0f / h->g.dz, 1};
    for (int q = 0; q < count; q++) {
        const int iter = iter0 + q;
        const float t = (float)iter * h->dt;  // test/runtests.jl:46
        JetParams j = h->jet;
        j.omgterm = h->omega * t;             // test/runtests.jl:57
        long long first0 = 0, nsp = 0;
        if (nslots > 0 && h->n_spawn > 0) {
            const long long start1 = ((long long)(iter - 1) * h->n_spawn + 1) % nslots;  // :52-54
            first0 = start1 == 0 ? nslots - 1 : start1 - 1;
            nsp = h->n_spawn;
        }
        // the side stream's part of this step: after an event on the handle's stream, the
        // advection of step iter - 1 (velocity in bufs[side_b]) and this step's spawn
        int side_b = -1;
        auto side = [&]() -> int {
            HIP_CHECK(hipStreamWaitEvent(ps, h->ev_v, 0), "hipStreamWaitEvent");  // the fused pass has ended
            // :63 staggered_advect! of step iter - 1, then :52-54 add_particles! of this step
            const int* order = nullptr;
            int r = ensure_order(h, parts, nslots, iter - 1, &order, ps);
            if (r) return r;
            LAUNCH_CHECK(launch_advect(h->g, parts, nslots, order, bufs[side_b], h->dt, ps, Batch{}, 0, nullptr, true),
                         "staggered_advect");
            HIP_CHECK(hipEventRecord(h->ev_buf[side_b], ps), "hipEventRecord");
            h->buf_rec[side_b] = true;
            LAUNCH_CHECK(launch_spawn(parts, nslots, first0, nsp, h->jet, h->seed, (unsigned)iter, ps), "add_particles");
            return ISF_OK;
        };
        if (q == 0 || nslots == 0) {
            // first step of the call (the previous step's advection is already done), or no
            // particles (the velocity field is only needed after the last step)
            if (nslots > 0)
                LAUNCH_CHECK(launch_spawn(parts, nslots, first0, nsp, h->jet, h->seed, (unsigned)iter, ps), "add_particles");
            LAUNCH_CHECK(launch_flow_x_fwd(h->g, p, h->S, h->tb, s), "flow x forward");
        } else {
            // :59-62 of step iter - 1 inside this step's forward x pass, into bufs[q & 1]
            const int b = q & 1;
            if (h->buf_rec[b]) {  // its last reader, the advection of step iter - 3
                if (hipEventQuery(h->ev_buf[b]) != hipSuccess)
                    HIP_CHECK(hipEventSynchronize(h->ev_buf[b]), "hipEventSynchronize");
            }
            VelOut vo = vo0;
            vo.vel = bufs[b];
            LAUNCH_CHECK(launch_flow_x_fwd(h->g, p, h->S, h->tb, s, Batch{}, &vo, h->ev_v),
                         "flow x forward + velocity_one_form");
            side_b = b;
        }
        // the advection right after the fused pass overlaps the flow plane pass: measured
        // 0.272 ms per step, against 0.284 (after the flow plane pass), 0.290 (after the
        // inverse flow pass) and 0.329 (after the first divergence pass)
        if (side_b >= 0) {
            int r = side();
            if (r) return r;
        }
        // :48-49 schroedinger_flow! + normalize, :50 + :57 projection with the restrict fused
        // into its gauge pass, :58 projection with the FP64 norm partials
        LAUNCH_CHECK(launch_flow_plane(h->g, h->S, h->tb, s), "flow plane");
        LAUNCH_CHECK(launch_flow_x_inv(h->g, h->S, p, true, h->tb, s), "flow x inverse + normalize");
        LAUNCH_CHECK(launch_div_r2c(h->g, p, f_of(h), P_of(h), h->tb, s), "div(one_form(psi)) + poisson x forward");
        LAUNCH_CHECK(launch_poisson_plane(h->g, P_of(h), sc, h->tb, s), "poisson plane");
        LAUNCH_CHECK(launch_c2r_x(h->g, P_of(h), 2, nullptr, p, &j, nullptr, h->tb, s), "poisson x inverse + gauge + restrict");
        LAUNCH_CHECK(launch_div_r2c(h->g, p, f_of(h), P_of(h), h->tb, s), "div(one_form(psi)) + poisson x forward");
        LAUNCH_CHECK(launch_poisson_plane(h->g, P_of(h), sc, h->tb, s), "poisson plane");
        LAUNCH_CHECK(launch_c2r_x(h->g, P_of(h), 1, nullptr, p, nullptr, h->norm_parts, h->tb, s), "poisson x inverse + gauge");
        h->norm_count = c2r_x_blocks(h->g);
    }
    // the last step's velocity and advection on the side stream, as isf_step leaves them
    HIP_CHECK(hipEventRecord(h->ev_vel, s), "hipEventRecord");
    HIP_CHECK(hipStreamWaitEvent(ps, h->ev_vel, 0), "hipStreamWaitEvent");
    LAUNCH_CHECK(launch_velocity(h->g, p, h->hbar, 1, h->vel, ps), "velocity_one_form");
    HIP_CHECK(hipEventRecord(h->ev_vread, ps), "hipEventRecord");
    h->vread_pending = true;
    if (nslots > 0) {
        const int* order = nullptr;
        int r = ensure_order(h, parts, nslots, iter0 + count - 1, &order, ps);
        if (r) return r;
        LAUNCH_CHECK(launch_advect(h->g, parts, nslots, order, h->vel, h->dt, ps, Batch{}, 0, nullptr, true),
                     "staggered_advect");
    }
    HIP_CHECK(hipEventRecord(h->ev_adv, ps), "hipEventRecord");
    h->side_pending = true;
    return ISF_OK;
}

int isf_simloop(isf_t h, float* psi, float* parts, long long nslots, int iter0, int count) {
    int r = step_args_ok(h, psi, parts, nslots, iter0);
    if (r) return r;
    if (count < 1) return fail(ISF_EINVAL, "count must be >= 1");
    if (count > INT32_MAX - iter0) return fail(ISF_EINVAL, "iter0 + count overflows");
    r = wait_ext(h);
    if (r) return r;
    return simloop_impl(h, psi, parts, nslots, iter0, count);
}

int isf_step_stage_count(void) { return kNumStages; }
const char* isf_step_stage_name(int i) { return (i >= 0 && i < kNumStages) ? kStageNames[i] : ""; }

int isf_step_profile(isf_t h, float* psi, float* parts, long long nslots, int iter, float* stage_ms) {
    int r = step_args_ok(h, psi, parts, nslots, iter);
    if (r) return r;
    CHECK_PTR(stage_ms, "stage_ms");
    hipEvent_t ev[kNumStages + 1] = {};
    for (int i = 0; i <= kNumStages; i++) {
        hipError_t e = hipEventCreate(&ev[i]);
        if (e != hipSuccess) {
            for (int k = 0; k < i; k++) (void)hipEventDestroy(ev[k]);
            return hip_fail(e, "hipEventCreate");
        }
    }
    r = step_impl(h, psi, parts, nslots, iter, ev);
    if (!r) {
        hipError_t e = hipEventSynchronize(ev[kNumStages]);
        if (e != hipSuccess) r = hip_fail(e, "hipEventSynchronize");
        for (int i = 0; !r && i < kNumStages; i++) {
            e = hipEventElapsedTime(&stage_ms[i], ev[i], ev[i + 1]);
            if (e != hipSuccess) r = hip_fail(e, "hipEventElapsedTime");
        }
    }
    for (int i = 0; i <= kNumStages; i++) (void)hipEventDestroy(ev[i]);
    return r;
}

int isf_norm2(isf_t h, const float* psi, double* out) {
    CHECK_HJ(h);
    CHECK_PTR(psi, "psi");
    CHECK_PTR(out, "out");
    size_t np = 0;
    LAUNCH_CHECK(launch_norm2(h->g, reinterpret_cast<const float4*>(psi), h->norm2_parts, &np, h->stream), "norm2");
    LAUNCH_CHECK(launch_reduce_sum(h->norm2_parts, np, h->norm2_out, h->stream), "norm2 reduce");
    HIP_CHECK(hipMemcpyAsync(out, h->norm2_out, sizeof(double), hipMemcpyDeviceToHost, h->stream), "norm2 copy");
    HIP_CHECK(hipStreamSynchronize(h->stream), "hipStreamSynchronize");
    return ISF_OK;
}

int isf_step_norm2(isf_t h, double* out) {
    CHECK_HJ(h);
    CHECK_PTR(out, "out");
    if (h->norm_count == 0) return fail(ISF_EINVAL, "no step has run on this handle");
    LAUNCH_CHECK(launch_reduce_sum(h->norm_parts, h->norm_count, h->norm_out, h->stream), "norm2 reduce");
    HIP_CHECK(hipMemcpyAsync(out, h->norm_out, sizeof(double), hipMemcpyDeviceToHost, h->stream), "norm2 copy");
    HIP_CHECK(hipStreamSynchronize(h->stream), "hipStreamSynchronize");
    return ISF_OK;
}

// Both timer calls first join the side stream (CHECK_HJ: the handle's stream waits for
// the event recorded after the last isf_step's velocity pass and advect), so the stop
// event completes only after the side-stream work of the timed steps.
int isf_timer_start(isf_t h) {
    CHECK_HJ(h);
    HIP_CHECK(hipEventRecord(h->ev0, h->stream), "hipEventRecord");
    return ISF_OK;
}
int isf_timer_stop(isf_t h, float* ms) {
    CHECK_HJ(h);
    CHECK_PTR(ms, "ms");
    HIP_CHECK(hipEventRecord(h->ev1, h->stream), "hipEventRecord");
    HIP_CHECK(hipEventSynchronize(h->ev1), "hipEventSynchronize");
    HIP_CHECK(hipEventElapsedTime(ms, h->ev0, h->ev1), "hipEventElapsedTime");
    return ISF_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ batched ensembles
//
// A batch owns one ISF handle per member (tables, scratch, velocity) and groups the
// members by grid shape.  A batched step launches every kernel stage once per group
// with the group's members in a grid dimension (each block reads its member's
// descriptor); the groups run as parallel branches forked from and joined back into the
// batch stream, captured once into a hipGraph and replayed per step.  The step number
// lives in a device counter set before each replay, so the captured graph stays valid.

static constexpr uint32_t kBatchMagic = 0xBA7C15F0u;

struct BatchGroup {
    GridInfo g;              // shape shared by the members; g.ty = x tile (0: default)
    std::vector<int> idx;    // member indices
    Inst* inst;              // device array of the members' descriptors
    long long max_spawn;     // largest per-step spawn count of the members with slots
    long long max_slots;     // largest particle buffer
    hipStream_t stream;
    hipEvent_t done;
};

struct isf_batch {
    uint32_t magic;
    std::vector<isf_t> mem;
    std::vector<float*> psi, parts;
    std::vector<long long> nslots;
    std::vector<unsigned long long> seen_seq;  // member seq already ordered before the batch stream
    std::vector<BatchGroup> groups;
    hipStream_t stream;
    hipEvent_t fork, done;
    int* counter;            // device step counter read by the batched kernels
    bool use_graph;
    bool dirty;              // descriptors must be rebuilt (bind, tune)
    hipGraphExec_t exec;     // isf_batch_step
    hipGraphExec_t exec_first, exec_steady;  // isf_batch_simloop: first step, later steps
    bool alt_uploaded;       // every uploaded descriptor carries its member's second velocity buffer
};

static int check_batch(isf_batch_t b) {
    if (!b || b->magic != kBatchMagic) return fail(ISF_EHANDLE, "null or destroyed isf batch");
    return ISF_OK;
}
#define CHECK_B(b)                     \
    do {                               \
        int r_ = check_batch(b);       \
        if (r_) return r_;             \
    } while (0)

static void batch_free(isf_batch_t b) {
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->exec_first) (void)hipGraphExecDestroy(b->exec_first);
    if (b->exec_steady) (void)hipGraphExecDestroy(b->exec_steady);
    for (BatchGroup& gr : b->groups) {
        if (gr.stream) (void)hipStreamSynchronize(gr.stream);
        if (gr.inst) (void)hipFree(gr.inst);
        if (gr.done) (void)hipEventDestroy(gr.done);
        if (gr.stream) (void)hipStreamDestroy(gr.stream);
    }
    for (isf_t h : b->mem) {
        h->owner = nullptr;
        h->ext_pending = false;
        h->batch_ev = nullptr;
        (void)isf_destroy(h);
    }
    if (b->counter) (void)hipFree(b->counter);
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->done) (void)hipEventDestroy(b->done);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    b->magic = 0;
    delete b;
}

static Inst inst_of(isf_t h, float* psi, float* parts, long long nslots, int ty) {
    Inst m{};
    m.g = h->g;
    m.g.ty = ty;
    m.jet = h->jet;
    m.tb = h->tb;
    m.psi = reinterpret_cast<float4*>(psi);
    m.S = h->S;
    m.vel = h->vel;
    m.vel_alt = h->vel_alt;
    m.parts = parts;
    m.nslots = parts ? nslots : 0;
    m.n_spawn = h->n_spawn;
    m.seed = h->seed;
    m.norm_parts = h->norm_parts;
    m.hbar = h->hbar;
    m.dt = h->dt;
    m.omega = h->omega;
    m.pscale = poisson_scale(h);
    return m;
}

static void drop_graph(isf_batch_t b) {
    for (hipGraphExec_t* e : {&b->exec, &b->exec_first, &b->exec_steady}) {
        if (*e) (void)hipGraphExecDestroy(*e);
        *e = nullptr;
    }
}

// Upload the members' descriptors (synchronous: the batch stream is idle afterwards).
static int batch_prepare(isf_batch_t b) {
    for (size_t i = 0; i < b->mem.size(); i++)
        if (!b->psi[i]) return fail(ISF_EINVAL, "batch member %zu has no psi bound (isf_batch_bind)", i);
    HIP_CHECK(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
    drop_graph(b);
    for (BatchGroup& gr : b->groups) {
        std::vector<Inst> host;
        gr.max_spawn = 0;
        gr.max_slots = 0;
        for (int i : gr.idx) {
            host.push_back(inst_of(b->mem[i], b->psi[i], b->parts[i], b->nslots[i], gr.g.ty));
            const Inst& m = host.back();
            if (m.nslots > 0) {
                gr.max_slots = std::max(gr.max_slots, m.nslots);
                gr.max_spawn = std::max(gr.max_spawn, m.n_spawn);
            }
        }
        HIP_CHECK(hipMemcpy(gr.inst, host.data(), host.size() * sizeof(Inst), hipMemcpyHostToDevice),
                  "batch descriptors upload");
    }
    b->alt_uploaded = true;
    for (isf_t h : b->mem) b->alt_uploaded = b->alt_uploaded && h->vel_alt != nullptr;
    b->dirty = false;
    return ISF_OK;
}

// What one batched enqueue covers: a whole step (isf_batch_step), or the parts of the
// batched step loop (isf_batch_simloop): its first step without the velocity pass and the
// advection; a later step, whose forward flow pass also writes the previous step's
// velocity, followed by that step's advection and this step's spawn; the tail, the last
// step's velocity pass and advection.
enum { BSTEP_FULL = 0, BSTEP_FIRST = 1, BSTEP_STEADY = 2, BSTEP_TAIL = 3 };

// The kernel stages of one simloop body (test/runtests.jl:45-64) for one shape group,
// in the order of isf_step.  The pointer arguments are ignored in batch mode.
static int group_stages(const BatchGroup& gr, const Inst* inst, const int* counter, const DevTables& t0,
                        double* norm_flag, hipStream_t s, int kind) {
    const Batch bt{inst, counter, (int)gr.idx.size()};
    const GridInfo& g = gr.g;
    if (kind == BSTEP_TAIL) {
        LAUNCH_CHECK(launch_velocity(g, nullptr, 0.0f, 1, nullptr, s, bt), "batch velocity_one_form");
        if (gr.max_slots > 0)
            LAUNCH_CHECK(launch_advect(g, nullptr, 0, nullptr, nullptr, 0.0f, s, bt, gr.max_slots), "batch staggered_advect");
        return ISF_OK;
    }
    if (kind != BSTEP_STEADY && gr.max_spawn > 0)
        LAUNCH_CHECK(launch_spawn(nullptr, 0, 0, 0, JetParams{}, 0, 0, s, bt, gr.max_spawn), "batch add_particles");
    if (kind == BSTEP_STEADY) {
        // the group stays one chain: a fork inside each group's branch made the runtime's
        // graph executor run the groups one after another (measured 0.81 against 0.65 ms)
        const VelOut vo{};  // per-member values come from the descriptors
        LAUNCH_CHECK(launch_flow_x_fwd(g, nullptr, nullptr, t0, s, bt, &vo), "batch flow x forward + velocity_one_form");
        if (gr.max_slots > 0) {
            Batch bp = bt;
            bp.vel_by_step = 1;
            LAUNCH_CHECK(launch_advect(g, nullptr, 0, nullptr, nullptr, 0.0f, s, bp, gr.max_slots), "batch staggered_advect");
        }
        if (gr.max_spawn > 0)
            LAUNCH_CHECK(launch_spawn(nullptr, 0, 0, 0, JetParams{}, 0, 0, s, bt, gr.max_spawn), "batch add_particles");
    } else {
        LAUNCH_CHECK(launch_flow_x_fwd(g, nullptr, nullptr, t0, s, bt), "batch flow x forward");
    }
    LAUNCH_CHECK(launch_flow_plane(g, nullptr, t0, s, bt), "batch flow plane");
    LAUNCH_CHECK(launch_flow_x_inv(g, nullptr, nullptr, true, t0, s, bt), "batch flow x inverse + normalize");
    for (int pass = 0; pass < 2; pass++) {
        LAUNCH_CHECK(launch_div_r2c(g, nullptr, nullptr, nullptr, t0, s, bt),
                     "batch div(one_form(psi)) + poisson x forward");
        LAUNCH_CHECK(launch_poisson_plane(g, nullptr, 0.0, t0, s, bt), "batch poisson plane");
        LAUNCH_CHECK(launch_c2r_x(g, nullptr, pass == 0 ? 2 : 1, nullptr, nullptr, nullptr,
                                  pass == 0 ? nullptr : norm_flag, t0, s, bt),
                     "batch poisson x inverse + gauge");
    }
    if (kind == BSTEP_FULL) {
        LAUNCH_CHECK(launch_velocity(g, nullptr, 0.0f, 1, nullptr, s, bt), "batch velocity_one_form");
        if (gr.max_slots > 0)
            LAUNCH_CHECK(launch_advect(g, nullptr, 0, nullptr, nullptr, 0.0f, s, bt, gr.max_slots), "batch staggered_advect");
    }
    return ISF_OK;
}

// Fork the groups off the batch stream, enqueue their stages, join them back.
static int batch_enqueue(isf_batch_t b, int kind) {
    HIP_CHECK(hipEventRecord(b->fork, b->stream), "hipEventRecord");
    for (BatchGroup& gr : b->groups) {
        HIP_CHECK(hipStreamWaitEvent(gr.stream, b->fork, 0), "hipStreamWaitEvent");
        isf_t h0 = b->mem[gr.idx[0]];
        int r = group_stages(gr, gr.inst, b->counter, h0->tb, h0->norm_parts, gr.stream, kind);
        if (r) return r;
        HIP_CHECK(hipEventRecord(gr.done, gr.stream), "hipEventRecord");
        HIP_CHECK(hipStreamWaitEvent(b->stream, gr.done, 0), "hipStreamWaitEvent");
    }
    return ISF_OK;
}

// One enqueue of `kind` on the batch stream: a replay of its captured graph (captured on
// first use), or direct launches.
static int batch_run(isf_batch_t b, int kind, hipGraphExec_t* exec) {
    if (!b->use_graph || !exec) return batch_enqueue(b, kind);
    if (!*exec) {
        HIP_CHECK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        const int r = batch_enqueue(b, kind);
        hipGraph_t graph = nullptr;
        const hipError_t e = hipStreamEndCapture(b->stream, &graph);
        if (r) {
            if (graph) (void)hipGraphDestroy(graph);
            return r;
        }
        if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
        const hipError_t ei = hipGraphInstantiate(exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ei != hipSuccess) {
            *exec = nullptr;
            return hip_fail(ei, "hipGraphInstantiate");
        }
    }
    HIP_CHECK(hipGraphLaunch(*exec, b->stream), "hipGraphLaunch");
    return ISF_OK;
}

// Order the batch stream after work the members' own handles enqueued since the last
// batched call.
static int batch_join_members(isf_batch_t b) {
    for (size_t i = 0; i < b->mem.size(); i++) {
        isf_t h = b->mem[i];
        if (h->seq == b->seen_seq[i] && !h->side_pending) continue;
        int r = join(h);
        if (r) return r;
        HIP_CHECK(hipEventRecord(h->ext_ev, h->stream), "hipEventRecord");
        HIP_CHECK(hipStreamWaitEvent(b->stream, h->ext_ev, 0), "hipStreamWaitEvent");
        b->seen_seq[i] = h->seq;
    }
    return ISF_OK;
}

// After a batched call: the members' handles wait for it before their own next work.
static int batch_finish(isf_batch_t b) {
    HIP_CHECK(hipEventRecord(b->done, b->stream), "hipEventRecord");
    for (isf_t h : b->mem) {
        h->batch_ev = b->done;
        h->ext_pending = true;
    }
    for (const BatchGroup& gr : b->groups)
        for (int i : gr.idx) b->mem[i]->norm_count = c2r_x_blocks(gr.g);
    return ISF_OK;
}

extern "C" {

int isf_batch_create(isf_batch_t* out, int n, const isf_member_desc* members) {
    CHECK_PTR(out, "out");
    CHECK_PTR(members, "members");
    *out = nullptr;
    if (n <= 0 || n > 65535) return fail(ISF_EINVAL, "batch size %d out of range (1..65535)", n);
    isf_batch* b = new isf_batch();
    b->magic = kBatchMagic;
    b->use_graph = true;
    b->dirty = true;
    auto bail = [&](int code) {
        batch_free(b);
        return code;
    };
    for (int i = 0; i < n; i++) {
        const isf_member_desc& d = members[i];
        if (!plane_in_lds(d.res[1], d.res[2]))
            return bail(fail(ISF_EINVAL, "batch member %d: grid_res[1:2] = (%d, %d); batches take y and z up to 128", i,
                             d.res[1], d.res[2]));
        isf_t h = nullptr;
        int r = isf_create(&h, d.res, d.box, d.hbar, d.dt);
        if (r) return bail(fail(r, "batch member %d: %s", i, std::string(isf::last_error()).c_str()));
        h->owner = b;
        b->mem.push_back(h);
        r = isf_jet_configure(h, d.kvec, d.omega, d.cen, d.len, d.rad, d.n_spawn, d.seed, d.phase_mode);
        if (r) return bail(fail(r, "batch member %d: %s", i, std::string(isf::last_error()).c_str()));
        b->seen_seq.push_back(h->seq);
    }
    b->psi.assign(n, nullptr);
    b->parts.assign(n, nullptr);
    b->nslots.assign(n, 0);
    for (int i = 0; i < n; i++) {
        const GridInfo& g = b->mem[i]->g;
        BatchGroup* gr = nullptr;
        for (BatchGroup& x : b->groups)
            if (x.g.nx == g.nx && x.g.ny == g.ny && x.g.nz == g.nz) gr = &x;
        if (!gr) {
            b->groups.push_back(BatchGroup{});
            gr = &b->groups.back();
            gr->g = g;
            gr->g.ty = 0;
        }
        gr->idx.push_back(i);
    }
    hipError_t e;
    for (BatchGroup& gr : b->groups) {
        if ((e = hipMalloc(&gr.inst, gr.idx.size() * sizeof(Inst))) != hipSuccess)
            return bail(fail(ISF_ENOMEM, "batch descriptors: %s", hipGetErrorString(e)));
        if ((e = hipStreamCreateWithFlags(&gr.stream, hipStreamNonBlocking)) != hipSuccess)
            return bail(hip_fail(e, "hipStreamCreate"));
        if ((e = hipEventCreateWithFlags(&gr.done, hipEventDisableTiming)) != hipSuccess)
            return bail(hip_fail(e, "hipEventCreate"));
    }
    if ((e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking)) != hipSuccess)
        return bail(hip_fail(e, "hipStreamCreate"));
    if ((e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming)) != hipSuccess)
        return bail(hip_fail(e, "hipEventCreate"));
    if ((e = hipEventCreateWithFlags(&b->done, hipEventDisableTiming)) != hipSuccess)
        return bail(hip_fail(e, "hipEventCreate"));
    if ((e = hipMalloc(&b->counter, sizeof(int))) != hipSuccess)
        return bail(fail(ISF_ENOMEM, "batch counter: %s", hipGetErrorString(e)));
    *out = b;
    return ISF_OK;
}

int isf_batch_destroy(isf_batch_t b) {
    CHECK_B(b);
    batch_free(b);
    return ISF_OK;
}

int isf_batch_member(isf_batch_t b, int i, isf_t* out) {
    CHECK_B(b);
    CHECK_PTR(out, "out");
    if (i < 0 || i >= (int)b->mem.size()) return fail(ISF_EINVAL, "member %d out of range", i);
    *out = b->mem[i];
    return ISF_OK;
}

int isf_batch_bind(isf_batch_t b, int i, float* psi, float* parts, long long nslots) {
    CHECK_B(b);
    CHECK_PTR(psi, "psi");
    if (i < 0 || i >= (int)b->mem.size()) return fail(ISF_EINVAL, "member %d out of range", i);
    if (nslots < 0 || (nslots > 0 && !parts)) return fail(ISF_EINVAL, "bad particle buffer");
    b->psi[i] = psi;
    b->parts[i] = nslots > 0 ? parts : nullptr;
    b->nslots[i] = nslots;
    b->dirty = true;
    return ISF_OK;
}

int isf_batch_set_graph(isf_batch_t b, int use_graph) {
    CHECK_B(b);
    b->use_graph = use_graph != 0;
    if (!b->use_graph) {
        HIP_CHECK(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
        drop_graph(b);
    }
    return ISF_OK;
}

int isf_batch_step(isf_batch_t b, int iter) {
    CHECK_B(b);
    if (iter < 1) return fail(ISF_EINVAL, "iter must be >= 1");
    if (b->dirty) {
        int r = batch_prepare(b);
        if (r) return r;
    }
    int r = batch_join_members(b);
    if (r) return r;
    LAUNCH_CHECK(launch_set_counter(b->counter, iter, b->stream), "batch step counter");
    if ((r = batch_run(b, BSTEP_FULL, &b->exec))) return r;
    return batch_finish(b);
}

int isf_batch_simloop(isf_batch_t b, int iter0, int count) {
    CHECK_B(b);
    if (iter0 < 1) return fail(ISF_EINVAL, "iter0 must be >= 1");
    if (count < 1) return fail(ISF_EINVAL, "count must be >= 1");
    if (count == 1) return isf_batch_step(b, iter0);
    // the second velocity buffer of every member (a member's own isf_simloop may have
    // allocated it after the descriptors were uploaded: then they are uploaded again)
    for (isf_t h : b->mem)
        if (!h->vel_alt) {
            const hipError_t e = hipMalloc(&h->vel_alt, 12 * h->N);
            if (e != hipSuccess) return fail(ISF_ENOMEM, "second velocity buffer: %s", hipGetErrorString(e));
        }
    if (!b->alt_uploaded) b->dirty = true;
    if (b->dirty) {
        int r = batch_prepare(b);
        if (r) return r;
    }
    int r = batch_join_members(b);
    if (r) return r;
    for (int q = 0; q < count; q++) {
        LAUNCH_CHECK(launch_set_counter(b->counter, iter0 + q, b->stream), "batch step counter");
        if ((r = batch_run(b, q == 0 ? BSTEP_FIRST : BSTEP_STEADY, q == 0 ? &b->exec_first : &b->exec_steady)))
            return r;
    }
    if ((r = batch_run(b, BSTEP_TAIL, nullptr))) return r;
    return batch_finish(b);
}

int isf_batch_step_sequential(isf_batch_t b, int iter) {
    CHECK_B(b);
    for (size_t i = 0; i < b->mem.size(); i++) {
        if (!b->psi[i]) return fail(ISF_EINVAL, "batch member %zu has no psi bound (isf_batch_bind)", i);
        int r = isf_step(b->mem[i], b->psi[i], b->parts[i], b->nslots[i], iter);
        if (r) return r;
    }
    return ISF_OK;
}

int isf_batch_synchronize(isf_batch_t b) {
    CHECK_B(b);
    HIP_CHECK(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
    for (size_t i = 0; i < b->mem.size(); i++) {
        isf_t h = b->mem[i];
        if (h->seq == b->seen_seq[i] && !h->side_pending) continue;
        int r = isf_synchronize(h);
        if (r) return r;
        b->seen_seq[i] = h->seq;
    }
    return ISF_OK;
}

// Time the four x passes of one step (forward, inverse + normalize, R2C, C2R + gauge)
// for one group with x tile ty, on scratch psi; returns milliseconds of `reps` rounds.
static int time_x_tile(isf_batch_t b, const BatchGroup& gr, int ty, const Inst* inst, float* ms) {
    GridInfo g = gr.g;
    g.ty = ty;
    const Batch bt{inst, b->counter, (int)gr.idx.size()};
    isf_t h0 = b->mem[gr.idx[0]];
    const DevTables& t0 = h0->tb;
    auto once = [&]() -> int {
        LAUNCH_CHECK(launch_flow_x_fwd(g, nullptr, nullptr, t0, b->stream, bt), "tune x forward");
        LAUNCH_CHECK(launch_flow_x_inv(g, nullptr, nullptr, true, t0, b->stream, bt), "tune x inverse");
        LAUNCH_CHECK(launch_div_r2c(g, nullptr, nullptr, nullptr, t0, b->stream, bt), "tune div + r2c");
        LAUNCH_CHECK(launch_c2r_x(g, nullptr, 1, nullptr, nullptr, nullptr, nullptr, t0, b->stream, bt), "tune c2r");
        return ISF_OK;
    };
    int r = once();
    if (r) return r;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_CHECK(hipEventCreate(&e0), "hipEventCreate");
    hipError_t e = hipEventCreate(&e1);
    if (e != hipSuccess) {
        (void)hipEventDestroy(e0);
        return hip_fail(e, "hipEventCreate");
    }
    const int reps = 20;
    e = hipEventRecord(e0, b->stream);
    for (int k = 0; !r && e == hipSuccess && k < reps; k++) r = once();
    if (!r && e == hipSuccess) e = hipEventRecord(e1, b->stream);
    if (!r && e == hipSuccess) e = hipEventSynchronize(e1);
    if (!r && e == hipSuccess) e = hipEventElapsedTime(ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (r) return r;
    if (e != hipSuccess) return hip_fail(e, "tune timing");
    return ISF_OK;
}

// Time the flow and Poisson plane passes of one step for one group with plane elements
// per thread pe (0: default), on the members' spectra; milliseconds of `reps` rounds.
static int time_planes(isf_batch_t b, const BatchGroup& gr, int pe, const Inst* inst, float* ms) {
    GridInfo g = gr.g;
    g.pe = pe;
    const Batch bt{inst, b->counter, (int)gr.idx.size()};
    const DevTables& t0 = b->mem[gr.idx[0]]->tb;
    auto once = [&]() -> int {
        LAUNCH_CHECK(launch_flow_plane(g, nullptr, t0, b->stream, bt), "tune flow plane");
        LAUNCH_CHECK(launch_poisson_plane(g, nullptr, 1.0, t0, b->stream, bt), "tune poisson plane");
        return ISF_OK;
    };
    int r = once();
    if (r) return r;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_CHECK(hipEventCreate(&e0), "hipEventCreate");
    hipError_t e = hipEventCreate(&e1);
    if (e != hipSuccess) {
        (void)hipEventDestroy(e0);
        return hip_fail(e, "hipEventCreate");
    }
    const int reps = 20;
    e = hipEventRecord(e0, b->stream);
    for (int k = 0; !r && e == hipSuccess && k < reps; k++) r = once();
    if (!r && e == hipSuccess) e = hipEventRecord(e1, b->stream);
    if (!r && e == hipSuccess) e = hipEventSynchronize(e1);
    if (!r && e == hipSuccess) e = hipEventElapsedTime(ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (r) return r;
    if (e != hipSuccess) return hip_fail(e, "tune timing");
    return ISF_OK;
}

// planes whose pass can run with 8 (default) or 16 elements per thread
static bool plane_pe_choice(const GridInfo& g) { return g.ny <= 64 && g.nz <= 64; }

int isf_batch_tune(isf_batch_t b, const char* cache_path) {
    CHECK_B(b);
    HIP_CHECK(hipDeviceSynchronize(), "hipDeviceSynchronize");
    // cache lines: nx ny nz members x_tile [plane_elements]; a line without the last
    // column keeps the default plane schedule
    struct Entry {
        int nx, ny, nz, members, ty, pe;
    };
    std::vector<Entry> cache;
    if (cache_path) {
        if (FILE* f = std::fopen(cache_path, "r")) {
            char line[256];
            while (std::fgets(line, sizeof line, f)) {
                Entry en{};
                const int k = std::sscanf(line, "%d %d %d %d %d %d", &en.nx, &en.ny, &en.nz, &en.members, &en.ty,
                                          &en.pe);
                if (k == 5) en.pe = 0;
                if (k == 5 || k == 6) cache.push_back(en);
            }
            std::fclose(f);
        }
    }
    std::vector<Entry> fresh;
    int rc = ISF_OK;
    for (BatchGroup& gr : b->groups) {
        const int members = (int)gr.idx.size();
        const bool pe_choice = plane_pe_choice(gr.g);
        int best = 0, best_pe = 0;
        for (const Entry& en : cache)
            if (en.nx == gr.g.nx && en.ny == gr.g.ny && en.nz == gr.g.nz && en.members == members &&
                (en.ty == 8 || en.ty == 16 || en.ty == 32)) {
                best = en.ty;
                best_pe = (pe_choice && en.pe == 16) ? 16 : 0;
            }
        if (!best) {
            std::vector<int> cand;
            for (int ty : {8, 16, 32})
                if (ty <= gr.g.ny && (ty == 8 || gr.g.nx <= 256) && gr.g.nx < 512) cand.push_back(ty);
            if (cand.empty()) cand.push_back(8);
            if (cand.size() == 1 && !pe_choice) {
                best = cand[0];
            } else {
                // scratch psi per member; the members' own spectra are scratch already
                const size_t bytes = (size_t)gr.g.nx * gr.g.ny * gr.g.nz * 16;
                float* scratch = nullptr;
                Inst* dinst = nullptr;
                hipError_t e = hipMalloc(&scratch, bytes * members);
                if (e == hipSuccess) e = hipMalloc(&dinst, sizeof(Inst) * members);
                if (e == hipSuccess) e = hipMemset(scratch, 0x3c, bytes * members);  // 0.0115f everywhere
                std::vector<Inst> host;
                for (int k = 0; k < members; k++)
                    host.push_back(inst_of(b->mem[gr.idx[k]], scratch + (bytes / 4) * k, nullptr, 0, 0));
                if (e == hipSuccess) e = hipMemcpy(dinst, host.data(), sizeof(Inst) * members, hipMemcpyHostToDevice);
                if (e == hipSuccess && launch_set_counter(b->counter, 1, b->stream) != 0) e = hipErrorLaunchFailure;
                if (e != hipSuccess) rc = hip_fail(e, "tune scratch");
                float best_ms = 0.0f;
                for (int ty : cand) {
                    if (rc) break;
                    float ms = 0.0f;
                    rc = cand.size() == 1 ? ISF_OK : time_x_tile(b, gr, ty, dinst, &ms);
                    if (!rc && (best == 0 || ms < best_ms)) {
                        best = ty;
                        best_ms = ms;
                    }
                }
                // plane schedule: radix 8 (8 elements per thread) or radix 16 (16)
                float pe_ms = 0.0f;
                for (int pe : {0, 16}) {
                    if (rc || !pe_choice) break;
                    float ms = 0.0f;
                    rc = time_planes(b, gr, pe, dinst, &ms);
                    if (!rc && (pe == 0 || ms < pe_ms)) {
                        best_pe = pe;
                        pe_ms = ms;
                    }
                }
                (void)hipStreamSynchronize(b->stream);
                if (scratch) (void)hipFree(scratch);
                if (dinst) (void)hipFree(dinst);
                if (rc) return rc;
            }
            fresh.push_back(Entry{gr.g.nx, gr.g.ny, gr.g.nz, members, best, best_pe});
        }
        gr.g.ty = best;
        gr.g.pe = best_pe;
        for (int i : gr.idx) {
            b->mem[i]->g.ty = best;
            b->mem[i]->g.pe = best_pe;
        }
    }
    if (cache_path && !fresh.empty()) {
        if (FILE* f = std::fopen(cache_path, "a")) {
            for (const Entry& en : fresh)
                std::fprintf(f, "%d %d %d %d %d %d\n", en.nx, en.ny, en.nz, en.members, en.ty, en.pe ? en.pe : 8);
            std::fclose(f);
        }
    }
    b->dirty = true;
    return ISF_OK;
}

int isf_batch_plane_elems(isf_batch_t b, const int res[3], int* pe) {
    CHECK_B(b);
    CHECK_PTR(res, "res");
    CHECK_PTR(pe, "pe");
    for (const BatchGroup& gr : b->groups)
        if (gr.g.nx == res[0] && gr.g.ny == res[1] && gr.g.nz == res[2]) {
            *pe = gr.g.pe ? gr.g.pe : (plane_pe_choice(gr.g) ? 8 : 16);
            return ISF_OK;
        }
    return fail(ISF_EINVAL, "no member of this batch has grid (%d, %d, %d)", res[0], res[1], res[2]);
}

int isf_batch_tile(isf_batch_t b, const int res[3], int* ty) {
    CHECK_B(b);
    CHECK_PTR(res, "res");
    CHECK_PTR(ty, "ty");
    for (const BatchGroup& gr : b->groups)
        if (gr.g.nx == res[0] && gr.g.ny == res[1] && gr.g.nz == res[2]) {
            *ty = gr.g.ty;
            return ISF_OK;
        }
    return fail(ISF_EINVAL, "no member of this batch has grid (%d, %d, %d)", res[0], res[1], res[2]);
}

}  // extern "C"
