// isf_matlabport.hip -- the MatlabPort operators (src/MatlabPort.jl:87-273) on the layout
// that module uses: psi1 and psi2 as two separate complex arrays (2 floats per cell), the
// velocity as three separate arrays vx, vy, vz, the divergence f real, the Poisson
// solution q complex, particle positions as separate x, y, z arrays.  All arrays are x
// fastest.  Arithmetic is Float32 with the same cell formulas as the interleaved kernels;
// the FFT operators reuse the x passes with split loads and stores, so SchroedingerFlow on
// split arrays equals schroedinger_flow! on the interleaved ones bit for bit.
#include <hip/hip_runtime.h>

#include "isf_device.hpp"
#include "isf_handle.hpp"
#include "isf_internal.hpp"

using namespace isf;

namespace {

constexpr int NT = 256;

inline dim3 cells_grid(size_t n) { return dim3((unsigned)((n + NT - 1) / NT)); }
__device__ __forceinline__ size_t ncells(const GridInfo& g) { return (size_t)g.nx * g.ny * g.nz; }

__device__ __forceinline__ float4 cell(const float2* __restrict__ a, const float2* __restrict__ b, size_t c) {
    const float2 u = a[c], v = b[c];
    return make_float4(u.x, u.y, v.x, v.y);
}

// Normalize(psi1, psi2), src/MatlabPort.jl:155-160
__global__ void __launch_bounds__(NT)
k_mp_normalize(const float2* a, const float2* b, float2* oa, float2* ob, size_t N) {
    const size_t c = (size_t)blockIdx.x * NT + threadIdx.x;
    if (c >= N) return;
    const float4 p = normalized(cell(a, b, c));
    oa[c] = make_float2(p.x, p.y);
    ob[c] = make_float2(p.z, p.w);
}

// VelocityOneForm(obj, psi1, psi2, hbar), src/MatlabPort.jl:106-126: forward edges, periodic
__global__ void __launch_bounds__(NT)
k_mp_velocity(const float2* __restrict__ a, const float2* __restrict__ b, GridInfo g, float hbar,
              float* __restrict__ vx, float* __restrict__ vy, float* __restrict__ vz) {
    const size_t c = (size_t)blockIdx.x * NT + threadIdx.x;
    const size_t N = ncells(g);
    if (c >= N) return;
    int x, y, z;
    cell_xyz(c, g, x, y, z);
    const size_t sy = g.nx, sz = (size_t)g.nx * g.ny;
    const size_t xp = x + 1 < g.nx ? c + 1 : c - (size_t)(g.nx - 1);
    const size_t yp = y + 1 < g.ny ? c + sy : c - (size_t)(g.ny - 1) * sy;
    const size_t zp = z + 1 < g.nz ? c + sz : c - (size_t)(g.nz - 1) * sz;
    const float4 p = cell(a, b, c);
    vx[c] = edge_angle(p, cell(a, b, xp)) * hbar;
    vy[c] = edge_angle(p, cell(a, b, yp)) * hbar;
    vz[c] = edge_angle(p, cell(a, b, zp)) * hbar;
}

// StaggeredSharp(obj, vx, vy, vz), src/MatlabPort.jl:99-104: v / d
__global__ void __launch_bounds__(NT)
k_mp_sharp(const float* vx, const float* vy, const float* vz, float* ux, float* uy, float* uz, GridInfo g, size_t N) {
    const size_t c = (size_t)blockIdx.x * NT + threadIdx.x;
    if (c >= N) return;
    ux[c] = vx[c] / g.dx;
    uy[c] = vy[c] / g.dy;
    uz[c] = vz[c] / g.dz;
}

// Div(obj, vx, vy, vz), src/MatlabPort.jl:162-170: backward differences over d^2, periodic
__global__ void __launch_bounds__(NT)
k_mp_div(const float* __restrict__ vx, const float* __restrict__ vy, const float* __restrict__ vz, float* __restrict__ f,
         GridInfo g) {
    const size_t c = (size_t)blockIdx.x * NT + threadIdx.x;
    const size_t N = ncells(g);
    if (c >= N) return;
    int x, y, z;
    cell_xyz(c, g, x, y, z);
    const size_t sy = g.nx, sz = (size_t)g.nx * g.ny;
    const size_t xm = x > 0 ? c - 1 : c + (size_t)(g.nx - 1);
    const size_t ym = y > 0 ? c - sy : c + (size_t)(g.ny - 1) * sy;
    const size_t zm = z > 0 ? c - sz : c + (size_t)(g.nz - 1) * sz;
    float r = (vx[c] - vx[xm]) / (g.dx * g.dx);
    r = r + (vy[c] - vy[ym]) / (g.dy * g.dy);
    r = r + (vz[c] - vz[zm]) / (g.dz * g.dz);
    f[c] = r;
}

// PoissonSolve's result as a complex array: the real q of the solve, imaginary part 0
__global__ void __launch_bounds__(NT) k_mp_complex(const float* __restrict__ qr, float2* __restrict__ q, size_t N) {
    const size_t c = (size_t)blockIdx.x * NT + threadIdx.x;
    if (c >= N) return;
    q[c] = make_float2(qr[c], 0.0f);
}

// GaugeTransform(psi1, psi2, q), src/MatlabPort.jl:140-145: psi * exp(i q), q complex:
// exp(i q) = exp(-Im q) (cos Re q + i sin Re q)
__global__ void __launch_bounds__(NT)
k_mp_gauge(const float2* a, const float2* b, const float2* __restrict__ q, float2* oa, float2* ob, size_t N) {
    const size_t c = (size_t)blockIdx.x * NT + threadIdx.x;
    if (c >= N) return;
    const float2 qc = q[c];
    float s, co;
    sincosf(qc.x, &s, &co);
    const float m = expf(-qc.y);
    const float er = m * co, ei = m * s;
    const float2 u = a[c], v = b[c];
    oa[c] = make_float2(u.x * er - u.y * ei, u.x * ei + u.y * er);
    ob[c] = make_float2(v.x * er - v.y * ei, v.x * ei + v.y * er);
}

__global__ void __launch_bounds__(NT) k_mp_pack(const float2* a, const float2* b, float4* __restrict__ w, size_t N) {
    const size_t c = (size_t)blockIdx.x * NT + threadIdx.x;
    if (c >= N) return;
    w[c] = cell(a, b, c);
}

__global__ void __launch_bounds__(NT) k_mp_unpack(const float4* __restrict__ w, float2* oa, float2* ob, size_t N) {
    const size_t c = (size_t)blockIdx.x * NT + threadIdx.x;
    if (c >= N) return;
    const float4 p = w[c];
    oa[c] = make_float2(p.x, p.y);
    ob[c] = make_float2(p.z, p.w);
}

// StaggeredVelocity(px, py, pz, torus, vx, vy, vz), src/MatlabPort.jl:237-273
__global__ void __launch_bounds__(NT)
k_mp_stag_vel(const float* __restrict__ px, const float* __restrict__ py, const float* __restrict__ pz, long long n,
              const float* __restrict__ vx, const float* __restrict__ vy, const float* __restrict__ vz,
              float* __restrict__ ux, float* __restrict__ uy, float* __restrict__ uz, GridInfo g) {
    const long long k = (long long)blockIdx.x * NT + threadIdx.x;
    if (k >= n) return;
    const float p[3] = {px[k], py[k], pz[k]};
    float u[3];
    stag_vel_split(vx, vy, vz, p, g, u);
    ux[k] = u[0];
    uy[k] = u[1];
    uz[k] = u[2];
}

// the driver's jet constraint, examples/matlabport/simple.jl:104-110: inside the jet
// psi_a <- |psi_a| exp(i (k.x - omega t)); isjet null: the nozzle predicate of j
__global__ void __launch_bounds__(NT)
k_mp_restrict(float2* a, float2* b, const unsigned char* __restrict__ isjet, GridInfo g, JetParams j) {
    const size_t c = (size_t)blockIdx.x * NT + threadIdx.x;
    if (c >= ncells(g)) return;
    int x, y, z;
    cell_xyz(c, g, x, y, z);
    if (!(isjet ? isjet[c] != 0 : is_jet_cell(x, y, z, g, j))) return;
    const float4 p = jet_phase(cell(a, b, c), x, y, z, g, j);
    a[c] = make_float2(p.x, p.y);
    b[c] = make_float2(p.z, p.w);
}

int ready(isf_t h) {
    int e = check_handle(h);
    if (!e) e = join(h);
    return e;
}

}  // namespace

#define MP_LAUNCH(kernel, n, what, ...)                                                         \
    do {                                                                                        \
        hipLaunchKernelGGL(kernel, cells_grid(n), dim3(NT), 0, h->stream, __VA_ARGS__);        \
        HIP_CHECK(hipGetLastError(), what);                                                     \
    } while (0)

extern "C" {

int isf_mp_schroedinger_flow(isf_t h, const float* psi1, const float* psi2, float* out1, float* out2) {
    int e = ready(h);
    if (e) return e;
    CHECK_PTR(psi1, "psi1");
    CHECK_PTR(psi2, "psi2");
    CHECK_PTR(out1, "out1");
    CHECK_PTR(out2, "out2");
    LAUNCH_CHECK(launch_flow_x_fwd_split(h->g, reinterpret_cast<const float2*>(psi1),
                                         reinterpret_cast<const float2*>(psi2), h->S, h->tb, h->stream),
                 "SchroedingerFlow x forward");
    LAUNCH_CHECK(launch_flow_plane(h->g, h->S, h->tb, h->stream), "SchroedingerFlow plane");
    LAUNCH_CHECK(launch_flow_x_inv_split(h->g, h->S, reinterpret_cast<float2*>(out1), reinterpret_cast<float2*>(out2),
                                         h->tb, h->stream),
                 "SchroedingerFlow x inverse");
    return ISF_OK;
}

int isf_mp_normalize(isf_t h, const float* psi1, const float* psi2, float* out1, float* out2) {
    int e = ready(h);
    if (e) return e;
    CHECK_PTR(psi1, "psi1");
    CHECK_PTR(psi2, "psi2");
    CHECK_PTR(out1, "out1");
    CHECK_PTR(out2, "out2");
    MP_LAUNCH(k_mp_normalize, h->N, "Normalize", reinterpret_cast<const float2*>(psi1),
              reinterpret_cast<const float2*>(psi2), reinterpret_cast<float2*>(out1), reinterpret_cast<float2*>(out2),
              h->N);
    return ISF_OK;
}

int isf_mp_velocity_one_form(isf_t h, const float* psi1, const float* psi2, float hbar, float* vx, float* vy,
                             float* vz) {
    int e = ready(h);
    if (e) return e;
    CHECK_PTR(psi1, "psi1");
    CHECK_PTR(psi2, "psi2");
    CHECK_PTR(vx, "vx");
    CHECK_PTR(vy, "vy");
    CHECK_PTR(vz, "vz");
    MP_LAUNCH(k_mp_velocity, h->N, "VelocityOneForm", reinterpret_cast<const float2*>(psi1),
              reinterpret_cast<const float2*>(psi2), h->g, hbar, vx, vy, vz);
    return ISF_OK;
}

int isf_mp_staggered_sharp(isf_t h, const float* vx, const float* vy, const float* vz, float* ux, float* uy,
                           float* uz) {
    int e = ready(h);
    if (e) return e;
    CHECK_PTR(vx, "vx");
    CHECK_PTR(vy, "vy");
    CHECK_PTR(vz, "vz");
    CHECK_PTR(ux, "ux");
    CHECK_PTR(uy, "uy");
    CHECK_PTR(uz, "uz");
    MP_LAUNCH(k_mp_sharp, h->N, "StaggeredSharp", vx, vy, vz, ux, uy, uz, h->g, h->N);
    return ISF_OK;
}

int isf_mp_div(isf_t h, const float* vx, const float* vy, const float* vz, float* f) {
    int e = ready(h);
    if (e) return e;
    CHECK_PTR(vx, "vx");
    CHECK_PTR(vy, "vy");
    CHECK_PTR(vz, "vz");
    CHECK_PTR(f, "f");
    MP_LAUNCH(k_mp_div, h->N, "Div", vx, vy, vz, f, h->g);
    return ISF_OK;
}

int isf_mp_poisson_solve(isf_t h, const float* f, float* q) {
    int e = ready(h);
    if (e) return e;
    CHECK_PTR(f, "f");
    CHECK_PTR(q, "q");
    LAUNCH_CHECK(launch_r2c_x(h->g, f, P_of(h), h->tb, h->stream), "PoissonSolve x forward");
    LAUNCH_CHECK(launch_poisson_plane(h->g, P_of(h), poisson_scale(h), h->tb, h->stream), "PoissonSolve plane");
    LAUNCH_CHECK(launch_c2r_x(h->g, P_of(h), 0, f_of(h), nullptr, nullptr, nullptr, h->tb, h->stream),
                 "PoissonSolve x inverse");
    MP_LAUNCH(k_mp_complex, h->N, "PoissonSolve", f_of(h), reinterpret_cast<float2*>(q), h->N);
    return ISF_OK;
}

int isf_mp_gauge_transform(isf_t h, const float* psi1, const float* psi2, const float* q, float* out1, float* out2) {
    int e = ready(h);
    if (e) return e;
    CHECK_PTR(psi1, "psi1");
    CHECK_PTR(psi2, "psi2");
    CHECK_PTR(q, "q");
    CHECK_PTR(out1, "out1");
    CHECK_PTR(out2, "out2");
    MP_LAUNCH(k_mp_gauge, h->N, "GaugeTransform", reinterpret_cast<const float2*>(psi1),
              reinterpret_cast<const float2*>(psi2), reinterpret_cast<const float2*>(q),
              reinterpret_cast<float2*>(out1), reinterpret_cast<float2*>(out2), h->N);
    return ISF_OK;
}

int isf_mp_pressure_project(isf_t h, const float* psi1, const float* psi2, float* out1, float* out2) {
    int e = ready(h);
    if (e) return e;
    CHECK_PTR(psi1, "psi1");
    CHECK_PTR(psi2, "psi2");
    CHECK_PTR(out1, "out1");
    CHECK_PTR(out2, "out2");
    if (!h->mp_psi) {
        const hipError_t he = hipMalloc(&h->mp_psi, 16 * h->N);
        if (he != hipSuccess) {
            h->mp_psi = nullptr;
            return fail(ISF_ENOMEM, "PressureProject staging: %s", hipGetErrorString(he));
        }
    }
    MP_LAUNCH(k_mp_pack, h->N, "PressureProject", reinterpret_cast<const float2*>(psi1),
              reinterpret_cast<const float2*>(psi2), h->mp_psi, h->N);
    if ((e = project(h, reinterpret_cast<float*>(h->mp_psi), 1, nullptr, nullptr))) return e;
    MP_LAUNCH(k_mp_unpack, h->N, "PressureProject", h->mp_psi, reinterpret_cast<float2*>(out1),
              reinterpret_cast<float2*>(out2), h->N);
    return ISF_OK;
}

int isf_mp_restrict_velocity(isf_t h, float* psi1, float* psi2, const unsigned char* isjet, const float kvec[3],
                              float omgterm, int phase_mode) {
    int e = ready(h);
    if (e) return e;
    CHECK_PTR(psi1, "psi1");
    CHECK_PTR(psi2, "psi2");
    CHECK_PTR(kvec, "kvec");
    if (phase_mode != ISF_PHASE_DOT && phase_mode != ISF_PHASE_LEGACY) return fail(ISF_EINVAL, "bad phase_mode %d", phase_mode);
    if (!isjet && !h->jet_set) return fail(ISF_EINVAL, "isjet is NULL and no nozzle was configured");
    JetParams j = h->jet_set ? h->jet : JetParams{};
    j.kx = kvec[0];
    j.ky = kvec[1];
    j.kz = kvec[2];
    j.omgterm = omgterm;
    j.phase_mode = phase_mode;
    MP_LAUNCH(k_mp_restrict, h->N, "jet constraint", reinterpret_cast<float2*>(psi1), reinterpret_cast<float2*>(psi2),
              isjet, h->g, j);
    return ISF_OK;
}

int isf_mp_staggered_velocity(isf_t h, const float* px, const float* py, const float* pz, long long n, const float* vx,
                              const float* vy, const float* vz, float* ux, float* uy, float* uz) {
    int e = ready(h);
    if (e) return e;
    if (n < 0) return fail(ISF_EINVAL, "negative count");
    if (n == 0) return ISF_OK;
    CHECK_PTR(px, "px");
    CHECK_PTR(py, "py");
    CHECK_PTR(pz, "pz");
    CHECK_PTR(vx, "vx");
    CHECK_PTR(vy, "vy");
    CHECK_PTR(vz, "vz");
    CHECK_PTR(ux, "ux");
    CHECK_PTR(uy, "uy");
    CHECK_PTR(uz, "uz");
    MP_LAUNCH(k_mp_stag_vel, (size_t)n, "StaggeredVelocity", px, py, pz, n, vx, vy, vz, ux, uy, uz, h->g);
    return ISF_OK;
}

}  // extern "C"
