// isf_internal.hpp -- types shared by the kernel launchers (isf_kernels.hip) and the
// C-ABI front end (isf_capi.cpp).  Not part of the public interface.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace isf {

// Grid geometry of one ISF instance (src/parallel.jl:11-13,44-46).
struct GridInfo {
    int nx, ny, nz;     // grid_res
    float dx, dy, dz;   // d = physical_size ./ grid_res, rounded to Float32
    float Lx, Ly, Lz;   // physical_size (integers in the reference, stored as float here)
    float idx, idy, idz;  // 1/d, used instead of a division only when d is a power of two
    int d_pow2;           // all three d are powers of two (then p * (1/d) == p / d exactly)
    int ty;               // y lines per block of the x passes (0: default choice)
    int pe;               // elements per thread of the plane pass (0: default choice; 16 also
                          // for planes up to 64 x 64, whose default is 8)
};

// Jet constraint parameters (src/particles.jl:12-27, test/runtests.jl:7-12,23-25).
struct JetParams {
    float kx, ky, kz;   // kvec = jet_velocity / hbar
    float omgterm;      // omega * t
    float cx, cy, cz;   // nozzle_cen
    float len, rad;     // nozzle_len, nozzle_rad
    int phase_mode;     // 0: k.x - omega t ; 1: reference tuple_dot sum(k .+ x) - omega t
};

// Device-resident per-instance tables (all tiny: O(nx + ny + nz) entries).
struct DevTables {
    float2* tw;          // combined twiddle table, length-N roots at tw + N (N <= 512)
    float2* mask_x;      // exp(i c kx^2) / N  (Schroedinger mask, x factor incl. 1/N)
    float2* mask_y;      // exp(i c ky^2)
    float2* mask_z;      // exp(i c kz^2)
    double* s2x;         // (sin(pi k / nx) / dx)^2, FP64 Poisson eigenvalue table
    double* s2y;
    double* s2z;
};

// One member of a batched launch: everything a step's kernels read for that instance.
// The per-step values (omega t, spawn slots, step number) are derived on the device from
// the batch's step counter.
struct Inst {
    GridInfo g;
    JetParams jet;           // nozzle and kvec; omgterm is formed per step
    DevTables tb;
    float4* psi;
    float2* S;               // spectrum; f = (float*)S and P = S + 4N bytes alias it
    float* vel;
    float* vel_alt;          // second velocity buffer (batched step loop)
    float* parts;
    long long nslots;
    long long n_spawn;
    unsigned long long seed;
    double* norm_parts;      // FP64 partial sums of the last gauge pass
    float hbar, dt, omega;
    double pscale;           // Poisson multiplier scale, -0.5/N
};

// Batch argument of every step kernel: inst == nullptr means a single instance whose
// data come from the kernel's other arguments; otherwise the kernel serves member m of
// inst (m taken from a grid dimension) and iter points to the batch's step counter.
struct Batch {
    const Inst* inst;
    const int* iter;
    int members;
    // batched step loop: the velocity of step *iter - 1 lives in vel_alt when that step is
    // odd, in vel when even (the fused forward pass writes it, the advection reads it)
    int vel_by_step;
};

// Launch-time choices of the x-direction passes (g.ty, when non-zero, overrides).
int x_tile_lines(int nx, int ny);

// ---- transform passes (return hipError_t as int)
// Every launcher takes an optional Batch (default: single instance).  In batch mode the
// pointer arguments are ignored and g gives the shape shared by all members.
// Velocity output of the fused forward x pass: vel = velocity_one_form(psi, hbar), times
// (s0, s1, s2) = 1/d when sharp (the driver's StaggeredSharp).
struct VelOut {
    float* vel;
    float hbar, s0, s1, s2;
    int sharp;
};
// vo != null: the pass also writes the velocity field of psi (in batch mode each member's
// hbar and 1/d, into the buffer of step *iter - 1; vo's fields are then not read).
// done != null (with vo): the launch completes that event when the kernel ends.
int launch_flow_x_fwd(const GridInfo& g, const float4* psi, float2* S, const DevTables& t, hipStream_t s,
                      Batch b = Batch{}, const VelOut* vo = nullptr, hipEvent_t done = nullptr);
int launch_flow_plane(const GridInfo& g, float2* S, const DevTables& t, hipStream_t s, Batch b = Batch{});
int launch_flow_x_inv(const GridInfo& g, const float2* S, float4* psi, bool normalize, const DevTables& t,
                      hipStream_t s, Batch b = Batch{});
// the inverse x pass writing the raw psi and, beside it, normalize_psi of it
int launch_flow_x_inv_raw_normalized(const GridInfo& g, const float2* S, float4* psi, float4* psin,
                                     const DevTables& t, hipStream_t s);
// the same x passes on the split-component layout (psi1 and psi2 as separate complex arrays)
int launch_flow_x_fwd_split(const GridInfo& g, const float2* sa, const float2* sb, float2* S, const DevTables& t,
                            hipStream_t s);
int launch_flow_x_inv_split(const GridInfo& g, const float2* S, float2* oa, float2* ob, const DevTables& t,
                            hipStream_t s);
int launch_r2c_x(const GridInfo& g, const float* f, float2* P, const DevTables& t, hipStream_t s, Batch b = Batch{});
// div(one_form(psi), hbar = 1) and the R2C x pass, in one kernel (f not stored) while a z
// plane of psi is at most 1 MB, else as k_div_psi into f (4N bytes of scratch) and R2C
int launch_div_r2c(const GridInfo& g, const float4* psi, float* f, float2* P, const DevTables& t, hipStream_t s,
                   Batch b = Batch{});
int launch_poisson_plane(const GridInfo& g, float2* P, double scale, const DevTables& t, hipStream_t s,
                         Batch b = Batch{});
// mode 0: write q (float per cell) ; mode 1: gauge psi ; mode 2: gauge psi + restrict ;
// mode 3: gauge psi, and its restriction into rout (not in batch mode).
// In batch mode norm_partials != nullptr requests the members' own partial sums.
int launch_c2r_x(const GridInfo& g, const float2* P, int mode, float* q, float4* psi, const JetParams* jet,
                 double* norm_partials, const DevTables& t, hipStream_t s, Batch b = Batch{},
                 float4* rout = nullptr);
size_t c2r_x_blocks(const GridInfo& g);

// ---- stencil / pointwise
int launch_div_psi(const GridInfo& g, const float4* psi, float* f, hipStream_t s, Batch b = Batch{});
int launch_velocity(const GridInfo& g, const float4* psi, float hbar, int sharp, float* vel, hipStream_t s,
                    Batch b = Batch{});
int launch_div_vel(const GridInfo& g, const float* vel, float* f, hipStream_t s);
int launch_normalize(const GridInfo& g, float4* psi, hipStream_t s);
int launch_sharp(const GridInfo& g, float* vel, hipStream_t s);
int launch_gauge(const GridInfo& g, float4* psi, const float* q, hipStream_t s);
int launch_isjet(const GridInfo& g, const JetParams& jet, unsigned char* out, hipStream_t s);
int launch_restrict(const GridInfo& g, float4* psi, const unsigned char* isjet, const JetParams& jet, hipStream_t s);
// sum of |psi1|^2 + |psi2|^2 as kNorm2Blocks FP64 partials (*nparts = kNorm2Blocks)
constexpr size_t kNorm2Blocks = 1024;
int launch_norm2(const GridInfo& g, const float4* psi, double* partials, size_t* nparts, hipStream_t s);
int launch_reduce_sum(const double* partials, size_t n, double* out, hipStream_t s);

// ---- particles
int launch_spawn(float* parts, long long nslots, long long first, long long n, const JetParams& jet,
                 unsigned long long seed, unsigned step, hipStream_t s, Batch b = Batch{}, long long max_n = 0);
// Uniform cloud in the nozzle cylinder in spatially coherent slot order (synchronous).
int launch_fill_cylinder(float* parts, long long nslots, long long first, long long n, const JetParams& jet,
                         unsigned long long seed, hipStream_t s);
// Number of consecutive in-box slot pairs more than two cells apart (synchronous);
// dev_partial holds order_measure_ints(nslots) unsigned ints.
size_t order_measure_ints(long long nslots);
int measure_order_jumps(const GridInfo& g, const float* parts, long long nslots, unsigned* dev_partial,
                        unsigned long long* host_count, hipStream_t s);
// src (optional, not in batch mode): advect the positions in src into parts (out of place)
// beside: the advection runs beside other kernels (the step's side stream): one block per
// CU, so it slows the concurrent plane passes evenly
int launch_advect(const GridInfo& g, float* parts, long long nslots, const int* order, const float* vel, float dt,
                  hipStream_t s, Batch b = Batch{}, long long max_slots = 0, const float* src = nullptr,
                  bool beside = false);
// Batch step counter: counter[0] = value / counter[0] += 1.
int launch_set_counter(int* counter, int value, hipStream_t s);
int launch_bump_counter(int* counter, hipStream_t s);
// Slot processing order for the advection: slots grouped by bucket of 8 x-cells
// (counting sort; order inside a bucket arbitrary).  order_scratch_bytes(g, n) bytes of
// scratch.
size_t order_scratch_bytes(const GridInfo& g, long long nslots);
int launch_particle_order(const GridInfo& g, const float* parts, long long nslots, int* order, void* scratch,
                          size_t scratch_bytes, hipStream_t s);
// slots grouped by exact cell (counting sort; order inside a cell arbitrary), for the
// scatter of the advection adjoint
size_t cell_order_scratch_bytes(const GridInfo& g, long long nslots);
int launch_cell_order(const GridInfo& g, const float* parts, long long nslots, int* order, void* scratch,
                      size_t scratch_bytes, hipStream_t s);

// ---- adjoint of the step and the differentiable rollout (isf_adjoint.hip)
// Particle-density grid: cells of size h over the box L, Gaussian width sigma truncated
// at radius R, density scale inv_n (1 / number of slots).
struct SplatCfg {
    int res[3];
    float h[3];
    float L[3];
    float sigma, R, inv_n;
};
int launch_flow_plane_adj(const GridInfo& g, float2* S, const DevTables& t, hipStream_t s);
int launch_fill_psi(float4* psi, float4 v, size_t n, hipStream_t s);
int launch_fill_f32(float* p, float v, size_t n, hipStream_t s);
int launch_fill_f64(double* p, double v, size_t n, hipStream_t s);
int launch_apply_phase(const GridInfo& g, float4* psi, const float* theta, const int pr[3], hipStream_t s);
int launch_apply_phase_adj(const GridInfo& g, float4* psibar, const float4* psi_out, const float* theta, float* phibar,
                           const int pr[3], hipStream_t s);
int launch_phase_gather(const GridInfo& g, const float* phibar, float* thbar, const int pr[3], hipStream_t s);
int launch_normalize_to(const GridInfo& g, const float4* in, float4* out, hipStream_t s);
int launch_restrict_to(const GridInfo& g, const float4* in, float4* out, const JetParams& j, hipStream_t s);
int launch_gauge_qbar(const GridInfo& g, const float4* psibar, const float4* psi_out, float* qbar, hipStream_t s);
// pre: how the one-form's psi is formed from the stored one (as stored, normalized, or
// jet-restricted with *j); with PRE_NORMALIZE psibar comes out as the cotangent of the
// stored psi (the normalize adjoint included)
enum { PRE_NONE = 0, PRE_NORMALIZE = 1, PRE_RESTRICT = 2 };
int launch_oneform_adj_div(const GridInfo& g, float4* psibar, const float4* psi, const float* q, const float* fbar,
                           hipStream_t s, int pre = 0, const JetParams* j = nullptr);
// qbar_out (optional): the gauge cotangent of the new psibar against psi
int launch_oneform_adj_vel(const GridInfo& g, float4* psibar, const float4* psi, const float* velbar, float hbar,
                           int sharp, hipStream_t s, float* qbar_out = nullptr);
// partials: restrict_adj_partials(g) doubles (restrict_adj_blocks(g) rows of 4 per-block
// sums, then the first reduction level's rows); out4[k] += the sums (sum phibar,
// sum phibar x, sum phibar y, sum phibar z)
size_t restrict_adj_blocks(const GridInfo& g);
size_t restrict_adj_partials(const GridInfo& g);
// qbar_out (optional): the gauge cotangent of the new psibar against psi_in
int launch_restrict_adj(const GridInfo& g, float4* psibar, const float4* psi_in, const JetParams& j, double* partials,
                        double* out4, hipStream_t s, float* qbar_out = nullptr);
// launch_oneform_adj_div with PRE_RESTRICT followed by launch_restrict_adj at the same psi,
// in one pass over the cells (the same results)
int launch_oneform_adj_div_restrict(const GridInfo& g, float4* psibar, const float4* psi, const float* q,
                                    const float* fbar, const JetParams& j, double* partials, double* out4,
                                    float* qbar_out, hipStream_t s);
int launch_normalize_adj(const GridInfo& g, float4* psibar, const float4* psi_in, hipStream_t s);
// order: slot processing order (null: slot order); cell-sorted slots make the velocity
// scatter of a wavefront hit few cells
int launch_advect_adj(const GridInfo& g, const float* parts0, long long nslots, const int* order, const float* vel,
                      float dt, float* pbar, float* velbar, hipStream_t s);
int launch_zero_slots(float* pbar, long long nslots, long long first, long long n, hipStream_t s);
// density: with partial (splat_partial_doubles(sc) doubles, 0 = unavailable) per-block LDS
// grids reduced in block order, rho overwritten; without, global atomics into a zeroed rho
size_t splat_partial_doubles(const SplatCfg& sc);
int launch_splat(const float* parts, long long nslots, const SplatCfg& sc, double* rho, double* partial,
                 hipStream_t s);
size_t splat_loss_blocks(int M);
int launch_splat_loss(const double* rho, const float* target, int M, float* resid, double* partials, double* out,
                      hipStream_t s);
int launch_splat_adj(const float* parts, long long nslots, const SplatCfg& sc, const float* resid, float* pbar,
                     hipStream_t s);

// Supported axis lengths.
bool x_len_supported(int n);
bool plane_supported(int ny, int nz);
bool plane_in_lds(int ny, int nz);

}  // namespace isf
