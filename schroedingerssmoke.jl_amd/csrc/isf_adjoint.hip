// isf_adjoint.hip -- hand-written adjoints (transposed Jacobians) of the ISF step
// operators, the set-up operators of a differentiable rollout and its particle-density
// loss.  Complex cotangents follow the real inner product <a, b> = Re(conj(a) b) summed
// over cells and components, so for an operator y = A(x) the adjoint maps ybar to
// xbar = J^T ybar with <J dx, ybar> = <dx, xbar>.
//
// Forward operator -> adjoint used here:
//   schroedinger_flow  F^-1 M F        -> F^-1 conj(M) F (plane pass PLANE_FLOW_ADJ)
//   normalize  n = u / |u| (u in R^4)  -> ubar = (nbar - n (n . nbar)) / |u|
//   gauge      psi' = psi exp(-i q)    -> psibar = psi'bar exp(+i q),
//                                         qbar = sum_comp Im(conj(psi'bar) psi')
//   Poisson    q = P f (P real, even)  -> fbar = P qbar (self-adjoint)
//   div        f = sum_a (v_a(c) - v_a(c - e_a)) / d_a^2
//                                      -> vbar_a(c) = (fbar(c) - fbar(c + e_a)) / d_a^2
//   one-form   v_a(c) = s angle(z_a(c)), z_a(c) = sum_comp conj(psi(c)) psi(c + e_a)
//                                      -> zbar = vbar s i z / |z|^2,
//                                         psibar(c) += conj(zbar) psi(c + e_a),
//                                         psibar(c + e_a) += zbar psi(c)
//   restrict   psi_a <- |psi_a| e^{i phi} in the jet
//                                      -> psibar_a = Re(e^{-i phi} psi'bar_a) psi_a / |psi_a|,
//                                         phibar = sum_a |psi_a| Im(e^{-i phi} psi'bar_a)
//   RK4 staggered advection            -> reverse sweep over the four stages with the
//                                         trilinear weights scattered into velbar (atomics)
#include <hip/hip_runtime.h>

#include <cmath>

#include "isf_device.hpp"
#include "isf_internal.hpp"

namespace isf {

#define ISF_LAUNCH(kernel, grid, block, stream, ...)                      \
    do {                                                                  \
        hipLaunchKernelGGL(kernel, grid, block, 0, stream, __VA_ARGS__);  \
        return (int)hipGetLastError();                                    \
    } while (0)

static inline dim3 grid1d(size_t n, int nt) { return dim3((unsigned)((n + nt - 1) / nt)); }
__host__ __device__ static inline size_t ncells(const GridInfo& g) { return (size_t)g.nx * g.ny * g.nz; }

__device__ __forceinline__ size_t cidx(int x, int y, int z, const GridInfo& g) {
    return (size_t)x + (size_t)g.nx * ((size_t)y + (size_t)g.ny * z);
}

// ======================================================================== set-up

__global__ void __launch_bounds__(256) k_fill_psi(float4* __restrict__ psi, float4 v, size_t N) {
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c < N) psi[c] = v;
}

__global__ void __launch_bounds__(256) k_fill_f32(float* __restrict__ p, float v, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
// the same fill with 16-byte stores (p 16-byte aligned): float4 i of the n / 4, and the
// last n % 4 floats by the first threads of block 0
__global__ void __launch_bounds__(256) k_fill_f32x4(float* __restrict__ p, float v, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n / 4) reinterpret_cast<float4*>(p)[i] = make_float4(v, v, v, v);
    if (blockIdx.x == 0 && threadIdx.x < n % 4) p[n / 4 * 4 + threadIdx.x] = v;
}

__global__ void __launch_bounds__(256) k_fill_f64(double* __restrict__ p, double v, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Periodic trilinear upsampling of a coarse field theta (pr[0] x pr[1] x pr[2], x
// fastest) at grid cell (x, y, z): theta coordinate s_a = x_a pr_a / n_a.
struct PhaseGrid {
    int pr[3];
};
__device__ __forceinline__ float phase_at(const float* __restrict__ th, int x, int y, int z, const GridInfo& g,
                                          const PhaseGrid& P) {
    const int c3[3] = {x, y, z}, n3[3] = {g.nx, g.ny, g.nz};
    int i0[3], i1[3];
    float f[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const int r = n3[a] / P.pr[a];  // cells per coarse interval (powers of two)
        i0[a] = c3[a] / r;
        i1[a] = i0[a] + 1 == P.pr[a] ? 0 : i0[a] + 1;
        f[a] = (float)(c3[a] - i0[a] * r) / (float)r;
    }
    auto T = [&](int a, int b, int c) { return th[a + P.pr[0] * (b + P.pr[1] * c)]; };
    const float z0 = (1.0f - f[0]) * ((1.0f - f[1]) * T(i0[0], i0[1], i0[2]) + f[1] * T(i0[0], i1[1], i0[2])) +
                     f[0] * ((1.0f - f[1]) * T(i1[0], i0[1], i0[2]) + f[1] * T(i1[0], i1[1], i0[2]));
    const float z1 = (1.0f - f[0]) * ((1.0f - f[1]) * T(i0[0], i0[1], i1[2]) + f[1] * T(i0[0], i1[1], i1[2])) +
                     f[0] * ((1.0f - f[1]) * T(i1[0], i0[1], i1[2]) + f[1] * T(i1[0], i1[1], i1[2]));
    return (1.0f - f[2]) * z0 + f[2] * z1;
}

// psi <- psi exp(i Phi), Phi = upsampled theta
__global__ void __launch_bounds__(256)
k_apply_phase(float4* __restrict__ psi, const float* __restrict__ th, GridInfo g, PhaseGrid P) {
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncells(g)) return;
    int x, y, z;
    cell_xyz(c, g, x, y, z);
    float s, co;
    sincosf(phase_at(th, x, y, z, g, P), &s, &co);
    const float4 p = psi[c];
    psi[c] = make_float4(p.x * co - p.y * s, p.x * s + p.y * co, p.z * co - p.w * s, p.z * s + p.w * co);
}

// Adjoint of k_apply_phase at psi' (the phased field): phibar = -sum Im(conj(psi'bar) psi'),
// psibar <- psi'bar exp(-i Phi).
__global__ void __launch_bounds__(256)
k_apply_phase_adj(float4* __restrict__ psibar, const float4* __restrict__ psi_out, const float* __restrict__ th,
                  float* __restrict__ phibar, GridInfo g, PhaseGrid P) {
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncells(g)) return;
    int x, y, z;
    cell_xyz(c, g, x, y, z);
    const float4 b = psibar[c], p = psi_out[c];
    phibar[c] = -((b.x * p.y - b.y * p.x) + (b.z * p.w - b.w * p.z));
    float s, co;
    sincosf(phase_at(th, x, y, z, g, P), &s, &co);
    psibar[c] = make_float4(b.x * co + b.y * s, b.y * co - b.x * s, b.z * co + b.w * s, b.w * co - b.z * s);
}

// thetabar = U^T phibar: each coarse node gathers the cells of its support (weight
// (1 - |s - node|) per axis), one workgroup per node: the threads take the support's cells
// x fastest (coalesced rows), sum in FP64 and reduce in a fixed tree order.
__global__ void __launch_bounds__(256)
k_phase_gather(const float* __restrict__ phibar, float* __restrict__ thbar, GridInfo g, PhaseGrid P) {
    const int node = blockIdx.x;
    const int m3[3] = {node % P.pr[0], (node / P.pr[0]) % P.pr[1], node / (P.pr[0] * P.pr[1])};
    const int n3[3] = {g.nx, g.ny, g.nz};
    int r[3], w[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        r[a] = n3[a] / P.pr[a];
        w[a] = 2 * r[a] - 1;
    }
    const int cnt = w[0] * w[1] * w[2];
    double acc = 0.0;
    for (int i = threadIdx.x; i < cnt; i += 256) {
        const int dx = i % w[0] - (r[0] - 1), dy = (i / w[0]) % w[1] - (r[1] - 1), dz = i / (w[0] * w[1]) - (r[2] - 1);
        const int x = (m3[0] * r[0] + dx + n3[0]) % n3[0];
        const int y = (m3[1] * r[1] + dy + n3[1]) % n3[1];
        const int z = (m3[2] * r[2] + dz + n3[2]) % n3[2];
        const float wx = 1.0f - (float)abs(dx) / (float)r[0];
        const float wy = 1.0f - (float)abs(dy) / (float)r[1];
        const float wz = 1.0f - (float)abs(dz) / (float)r[2];
        acc += (double)(wx * wy * wz) * (double)phibar[cidx(x, y, z, g)];
    }
    const double sum = block_sum<256>(acc);
    if (threadIdx.x == 0) thbar[node] = (float)sum;
}

// ======================================================================== out-of-place forms

__global__ void __launch_bounds__(256) k_normalize_to(const float4* __restrict__ in, float4* __restrict__ out, size_t N) {
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c < N) out[c] = normalized(in[c]);
}

__global__ void __launch_bounds__(256)
k_restrict_to(const float4* __restrict__ in, float4* __restrict__ out, GridInfo g, JetParams j) {
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncells(g)) return;
    int x, y, z;
    cell_xyz(c, g, x, y, z);
    const float4 p = in[c];
    out[c] = is_jet_cell(x, y, z, g, j) ? jet_phase(p, x, y, z, g, j) : p;
}

// ======================================================================== adjoints

// qbar(c) = sum_comp Im(conj(psibar) psi_out): the cotangent of q of a gauge transform
// psi_out = psi exp(-i q).
__global__ void __launch_bounds__(256)
k_gauge_qbar(const float4* __restrict__ psibar, const float4* __restrict__ psi_out, float* __restrict__ qbar, size_t N) {
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    const float4 b = psibar[c], p = psi_out[c];
    qbar[c] = (b.x * p.y - b.y * p.x) + (b.z * p.w - b.w * p.z);
}

enum { OF_DIV = 0, OF_VEL = 1 };

// zbar = w i z / |z|^2 for edge weight w (the cotangent of angle(z) times its scale); the
// hardware reciprocal (1 ulp) instead of a correctly rounded division, which takes about
// ten instructions, six times per cell
__device__ __forceinline__ float2 angle_zbar(float re, float im, float w) {
    const float m2 = re * re + im * im;
    const float s = m2 > 0.0f ? w * __builtin_amdgcn_rcpf(m2) : 0.0f;
    return make_float2(-im * s, re * s);
}
__device__ __forceinline__ void edge_z(float4 a, float4 b, float& re, float& im) {
    re = (a.x * b.x + a.y * b.y) + (a.z * b.z + a.w * b.w);
    im = (a.x * b.y - a.y * b.x) + (a.z * b.w - a.w * b.z);
}
// conj(zb) * b added to acc (componentwise over the two wavefunctions)
__device__ __forceinline__ void add_conj_mul(float4& acc, float2 zb, float4 b) {
    acc.x += zb.x * b.x + zb.y * b.y;
    acc.y += zb.x * b.y - zb.y * b.x;
    acc.z += zb.x * b.z + zb.y * b.w;
    acc.w += zb.x * b.w - zb.y * b.z;
}
// zb * a added to acc
__device__ __forceinline__ void add_mul(float4& acc, float2 zb, float4 a) {
    acc.x += zb.x * a.x - zb.y * a.y;
    acc.y += zb.x * a.y + zb.y * a.x;
    acc.z += zb.x * a.z - zb.y * a.w;
    acc.w += zb.x * a.w + zb.y * a.z;
}

// Cotangent of normalize's input u from the cotangent b of its output:
// (b - n (n . b)) / |u| with n = u / |u|.
__device__ __forceinline__ float4 normalize_adj_at(float4 u, float4 b) {
    const float inv = 1.0f / sqrtf((u.x * u.x + u.y * u.y) + (u.z * u.z + u.w * u.w));
    const float4 n = make_float4(u.x * inv, u.y * inv, u.z * inv, u.w * inv);
    const float dn = (n.x * b.x + n.y * b.y) + (n.z * b.z + n.w * b.w);
    return make_float4((b.x - n.x * dn) * inv, (b.y - n.y * dn) * inv, (b.z - n.z * dn) * inv, (b.w - n.w * dn) * inv);
}

// The psi a one-form was evaluated at, formed from the stored psi on the fly: as stored
// (PRE_NONE), normalized (PRE_NORMALIZE: the first projection's input normalize(psi1)) or
// jet-restricted (PRE_RESTRICT: the second projection's input restrict(psi3)).  Same
// arithmetic as k_normalize_to and k_restrict_to.
template <int PRE>
__device__ __forceinline__ float4 psi_pre(const float4* __restrict__ psi, size_t i, int x, int y, int z,
                                          const GridInfo& g, const JetParams& j) {
    const float4 p = psi[i];
    if (PRE == PRE_NORMALIZE) return normalized(p);
    if (PRE == PRE_RESTRICT) return is_jet_cell(x, y, z, g, j) ? jet_phase(p, x, y, z, g, j) : p;
    return p;
}

// The restrict adjoint of one jet cell at its input p (psibar b of the output): the new
// psibar, and phibar with its products by the cell position in s[0..3].
__device__ __forceinline__ float4 restrict_adj_cell(float4 b, float4 p, int x, int y, int z, const GridInfo& g,
                                                    const JetParams& j, double s[4]) {
    const float4 o = jet_phase(p, x, y, z, g, j);
    const float a1 = hypotf(p.x, p.y), a2 = hypotf(p.z, p.w);
    // e^{i phi} recovered from the output (a_k > 0 in practice; guarded)
    const float e1r = a1 > 0.0f ? o.x / a1 : 0.0f, e1i = a1 > 0.0f ? o.y / a1 : 0.0f;
    const float e2r = a2 > 0.0f ? o.z / a2 : 0.0f, e2i = a2 > 0.0f ? o.w / a2 : 0.0f;
    const float r1 = b.x * e1r + b.y * e1i, r2 = b.z * e2r + b.w * e2i;  // Re(conj(e) bar)
    const float phib = a1 * (b.y * e1r - b.x * e1i) + a2 * (b.w * e2r - b.z * e2i);
    const float i1 = a1 > 0.0f ? 1.0f / a1 : 0.0f, i2 = a2 > 0.0f ? 1.0f / a2 : 0.0f;
    const float px = __fmul_rn((float)x, g.dx), py = __fmul_rn((float)y, g.dy), pz = __fmul_rn((float)z, g.dz);
    s[0] = phib;
    s[1] = (double)phib * px;
    s[2] = (double)phib * py;
    s[3] = (double)phib * pz;
    return make_float4(r1 * p.x * i1, r1 * p.y * i1, r2 * p.z * i2, r2 * p.w * i2);
}
// the gauge cotangent of psibar b against the gauge output p (k_gauge_qbar)
__device__ __forceinline__ float gauge_qbar(float4 b, float4 p) {
    return (b.x * p.y - b.y * p.x) + (b.z * p.w - b.w * p.z);
}
// the four per-block sums of the restrict adjoint (every thread of the 256 calls this)
__device__ __forceinline__ void restrict_adj_block_sums(double s[4], double* __restrict__ partials) {
    s[0] = block_sum<256>(s[0]);
    s[1] = block_sum<256>(s[1]);
    s[2] = block_sum<256>(s[2]);
    s[3] = block_sum<256>(s[3]);
    if (threadIdx.x == 0) {
        double* o = partials + 4 * (size_t)blockIdx.x;
        o[0] = s[0];
        o[1] = s[1];
        o[2] = s[2];
        o[3] = s[3];
    }
}

// Adjoint of a one-form evaluated at psi (formed per PRE), fused with the adjoint of what
// consumed it:
//   OF_DIV: div then a Poisson solve produced q; w = fbar (Poisson-solved qbar),
//           edge weight (fbar(c) - fbar(c + e_a)) / d_a^2, hbar = 1 (src/parallel.jl:156)
//   OF_VEL: velocity_one_form(hbar), with 1/d when sharp (test/runtests.jl:59-62); w =
//           velbar (3 per cell), edge weight velbar_a(c) hbar (/ d_a)
// psibar(c) <- rot(psibar(c), +q(c)) (the gauge's own adjoint, when q != nullptr) plus the
// one-form contributions of the six edges that touch c.  With PRE_NORMALIZE the result is
// the cotangent of the stored psi (the normalize adjoint applied last).  With qbar_out, also the gauge
// cotangent of the new psibar against psi (k_gauge_qbar with psi_out = psi), for a gauge
// transform whose output psi is.
// The psibar of cell c (x, y, z) before the normalize adjoint: the gauge adjoint and the
// one-form contributions of the six edges at c; *pc = the psi the one-form saw at c.
template <int MODE, int PRE>
__device__ __forceinline__ float4 oneform_adj_cell(const float4* __restrict__ psibar, const float4* __restrict__ psi,
                                                   const float* __restrict__ q, const float* __restrict__ w,
                                                   const GridInfo& g, float hbar, int sharp, const JetParams& j,
                                                   size_t c, int x, int y, int z, float4* pc_out) {
    const int n3[3] = {g.nx, g.ny, g.nz};
    const float d[3] = {g.dx, g.dy, g.dz};
    float4 acc = psibar[c];
    if (q) {
        float s, co;
        sincosf(q[c], &s, &co);
        acc = make_float4(acc.x * co - acc.y * s, acc.x * s + acc.y * co, acc.z * co - acc.w * s, acc.z * s + acc.w * co);
    }
    const float4 pc = psi_pre<PRE>(psi, c, x, y, z, g, j);
    const float wc = MODE == OF_DIV ? w[c] : 0.0f;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        int cp[3] = {x, y, z}, cm[3] = {x, y, z};
        cp[a] = cp[a] + 1 == n3[a] ? 0 : cp[a] + 1;
        cm[a] = cm[a] == 0 ? n3[a] - 1 : cm[a] - 1;
        const size_t ip = cidx(cp[0], cp[1], cp[2], g), im = cidx(cm[0], cm[1], cm[2], g);
        const float4 pp = psi_pre<PRE>(psi, ip, cp[0], cp[1], cp[2], g, j);
        const float4 pm = psi_pre<PRE>(psi, im, cm[0], cm[1], cm[2], g, j);
        float wf, wb;  // weights of the forward edge (c, c+e_a) and the backward edge (c-e_a, c)
        if (MODE == OF_DIV) {
            const float id2 = 1.0f / (d[a] * d[a]);
            wf = (wc - w[ip]) * id2;
            wb = (w[im] - wc) * id2;
        } else {
            const float sc = sharp ? hbar / d[a] : hbar;
            wf = w[3 * c + a] * sc;
            wb = w[3 * im + a] * sc;
        }
        float re, imz;
        edge_z(pc, pp, re, imz);
        add_conj_mul(acc, angle_zbar(re, imz, wf), pp);  // c is the first end of (c, c+e_a)
        edge_z(pm, pc, re, imz);
        add_mul(acc, angle_zbar(re, imz, wb), pm);       // c is the second end of (c-e_a, c)
    }
    *pc_out = pc;
    return acc;
}

// RADJ (with PRE_RESTRICT): the restrict adjoint at psi follows in the same pass (the
// arithmetic of k_restrict_adj, with its per-block sums in partials), and qbar_out then
// takes the gauge cotangent against psi, the restrict input: one pass over psibar instead
// of two.
template <int MODE, int PRE, bool RADJ = false>
__global__ void __launch_bounds__(256)
k_oneform_adj(float4* __restrict__ psibar, const float4* __restrict__ psi, const float* __restrict__ q,
              const float* __restrict__ w, GridInfo g, float hbar, int sharp, JetParams j,
              float* __restrict__ qbar_out, double* __restrict__ partials = nullptr) {
    static_assert(!RADJ || PRE == PRE_RESTRICT, "the fused restrict adjoint follows a one-form at restrict(psi)");
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if constexpr (RADJ) {
        // every thread reaches the block sums
        double sums[4] = {0.0, 0.0, 0.0, 0.0};
        if (c < ncells(g)) {
            int x, y, z;
            cell_xyz(c, g, x, y, z);
            float4 pc;
            const float4 acc = oneform_adj_cell<MODE, PRE>(psibar, psi, q, w, g, hbar, sharp, j, c, x, y, z, &pc);
            const float4 p = psi[c];
            float4 nb = acc;
            if (is_jet_cell(x, y, z, g, j)) nb = restrict_adj_cell(acc, p, x, y, z, g, j, sums);
            psibar[c] = nb;
            if (qbar_out) qbar_out[c] = gauge_qbar(nb, p);
        }
        restrict_adj_block_sums(sums, partials);
    } else {
        if (c >= ncells(g)) return;
        int x, y, z;
        cell_xyz(c, g, x, y, z);
        float4 pc;
        float4 acc = oneform_adj_cell<MODE, PRE>(psibar, psi, q, w, g, hbar, sharp, j, c, x, y, z, &pc);
        if (PRE == PRE_NORMALIZE) acc = normalize_adj_at(psi[c], acc);
        psibar[c] = acc;
        if (qbar_out) qbar_out[c] = (acc.x * pc.y - acc.y * pc.x) + (acc.z * pc.w - acc.w * pc.z);
    }
}

// The same adjoint, tiled for 64 <= nx <= 256 and computed per edge instead of per cell.
// A block of kOfaThreads threads holds RY = kOfaThreads / nx whole x rows and walks tz z
// planes; a thread owns one (x, y) column.  Each thread forms the one-form input
// (psi_pre) of its own cell once, and evaluates the three forward edges (c, c + e_a) of
// its cell once: it adds the edge's term for c itself and hands the edge's cotangent zbar
// to the cell c + e_a, which adds zbar * psi_pre(c): through LDS for x and y (z-plane
// parity double buffers), in a register for z (the same thread's next plane).  The two
// edges whose first end lies outside the block are evaluated again where they are
// needed: the backward y edge by the block's first row (from a halo row) and the
// backward z edge at the slab's first plane.  So a cell costs about 3 + 1/RY + 1/tz edges
// and 1 + 2/RY + 1/tz psi_pre instead of 6 and 7.  Same terms as k_oneform_adj, summed in
// another order.
constexpr int kOfaThreads = 512;
template <int PRE>
__device__ __forceinline__ float4 pre_of(float4 p, int x, int y, int z, const GridInfo& g, const JetParams& j) {
    if (PRE == PRE_NORMALIZE) return normalized(p);
    if (PRE == PRE_RESTRICT) return is_jet_cell(x, y, z, g, j) ? jet_phase(p, x, y, z, g, j) : p;
    return p;
}
template <int MODE, int PRE, bool RADJ>
__global__ void __launch_bounds__(kOfaThreads)
k_oneform_adj_tile(float4* __restrict__ psibar, const float4* __restrict__ psi, const float* __restrict__ q,
                   const float* __restrict__ w, GridInfo g, float hbar, int sharp, JetParams j, int tz,
                   float* __restrict__ qbar_out, double* __restrict__ partials) {
    static_assert(!RADJ || PRE == PRE_RESTRICT, "the fused restrict adjoint follows a one-form at restrict(psi)");
    __shared__ float4 ps[2][kOfaThreads];  // psi_pre of the block's cells (z-plane parity)
    __shared__ float wsh[2][MODE == OF_DIV ? kOfaThreads : 1];
    __shared__ float2 zx_s[2][kOfaThreads], zy_s[2][kOfaThreads];  // zbar of the cell's backward x / y edge
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const int RY = kOfaThreads / nx;
    const int tid = threadIdx.x, x = tid % nx, ty = tid / nx;
    const int y = blockIdx.x * RY + ty, z0 = blockIdx.y * tz;
    const size_t sy = nx, sz = (size_t)nx * ny;
    const int xp = x + 1 < nx ? x + 1 : 0, xm = x > 0 ? x - 1 : nx - 1;
    const int yp = y + 1 < ny ? y + 1 : 0, ym = y > 0 ? y - 1 : ny - 1;
    const bool top = ty == RY - 1, bottom = ty == 0;
    const float id2[3] = {1.0f / (g.dx * g.dx), 1.0f / (g.dy * g.dy), 1.0f / (g.dz * g.dz)};
    const float sc[3] = {sharp ? hbar / g.dx : hbar, sharp ? hbar / g.dy : hbar, sharp ? hbar / g.dz : hbar};
    const size_t col = x + sy * y;
    double sums[4] = {0.0, 0.0, 0.0, 0.0};
    // this plane's cell (raw and pre), its w (OF_DIV), and the previous plane's pre and
    // forward z cotangent (the backward z edge of this cell)
    float4 praw = psi[col + sz * z0];
    float4 pc = pre_of<PRE>(praw, x, y, z0, g, j);
    float wc = MODE == OF_DIV ? w[col + sz * z0] : 0.0f;
    float4 pprev;
    float2 zzprev;
    {
        const int zm = z0 > 0 ? z0 - 1 : nz - 1;
        const size_t cm = col + sz * zm;
        pprev = pre_of<PRE>(psi[cm], x, y, zm, g, j);
        const float wb = MODE == OF_DIV ? (w[cm] - wc) * id2[2] : w[3 * cm + 2] * sc[2];
        float re, im;
        edge_z(pprev, pc, re, im);
        zzprev = angle_zbar(re, im, wb);
    }
#pragma unroll 1
    for (int k = 0; k < tz; k++) {
        const int z = z0 + k, par = k & 1;
        const size_t c = col + sz * z;
        const int zn = z + 1 < nz ? z + 1 : 0;
        const size_t cn = col + sz * zn;
        // loads of this step: the next plane's cell, the halo rows, this cell's cotangents
        const float4 prawn = psi[cn];
        const float wn = MODE == OF_DIV ? w[cn] : 0.0f;
        float4 phr = make_float4(0.0f, 0.0f, 0.0f, 0.0f), pbr = phr;
        float wh = 0.0f, wbm = 0.0f;
        if (top) {
            phr = psi[x + sy * yp + sz * z];
            if (MODE == OF_DIV) wh = w[x + sy * yp + sz * z];
        }
        if (bottom) {
            pbr = psi[x + sy * ym + sz * z];
            wbm = MODE == OF_DIV ? w[x + sy * ym + sz * z] : w[3 * (x + sy * ym + sz * z) + 1];
        }
        float4 acc = psibar[c];
        if (q) {
            float sn, co;
            sincosf(q[c], &sn, &co);
            acc = make_float4(acc.x * co - acc.y * sn, acc.x * sn + acc.y * co, acc.z * co - acc.w * sn,
                              acc.z * sn + acc.w * co);
        }
        float wv[3] = {0.0f, 0.0f, 0.0f};
        if (MODE == OF_VEL) {
            wv[0] = w[3 * c + 0];
            wv[1] = w[3 * c + 1];
            wv[2] = w[3 * c + 2];
        }
        ps[par][tid] = pc;
        if (MODE == OF_DIV) wsh[par][tid] = wc;
        const float4 pn = pre_of<PRE>(prawn, x, y, zn, g, j);
        const float4 ph = top ? pre_of<PRE>(phr, x, yp, z, g, j) : phr;
        const float4 pb = bottom ? pre_of<PRE>(pbr, x, ym, z, g, j) : pbr;
        __syncthreads();
        // the three forward edges (c, c + e_a)
        const int rx = ty * nx, ru = top ? rx + x : rx + nx + x;  // the row above (in the block)
        const float4 pa[3] = {ps[par][rx + xp], top ? ph : ps[par][ru], pn};
        float wf[3];
        if (MODE == OF_DIV) {
            wf[0] = (wc - wsh[par][rx + xp]) * id2[0];
            wf[1] = (wc - (top ? wh : wsh[par][ru])) * id2[1];
            wf[2] = (wc - wn) * id2[2];
        } else {
            wf[0] = wv[0] * sc[0];
            wf[1] = wv[1] * sc[1];
            wf[2] = wv[2] * sc[2];
        }
        float2 zf[3];
#pragma unroll
        for (int a = 0; a < 3; a++) {
            float re, im;
            edge_z(pc, pa[a], re, im);
            zf[a] = angle_zbar(re, im, wf[a]);
            add_conj_mul(acc, zf[a], pa[a]);  // c is the first end of (c, c + e_a)
        }
        zx_s[par][rx + xp] = zf[0];
        if (!top) zy_s[par][ru] = zf[1];
        __syncthreads();
        // the three backward edges (c - e_a, c): c is their second end
        add_mul(acc, zx_s[par][tid], ps[par][rx + xm]);
        if (!bottom) {
            add_mul(acc, zy_s[par][tid], ps[par][rx - nx + x]);
        } else {
            const float wb = MODE == OF_DIV ? (wbm - wc) * id2[1] : wbm * sc[1];
            float re, im;
            edge_z(pb, pc, re, im);
            add_mul(acc, angle_zbar(re, im, wb), pb);
        }
        add_mul(acc, zzprev, pprev);
        // finish the cell as k_oneform_adj does
        if constexpr (RADJ) {
            float4 nb = acc;
            if (is_jet_cell(x, y, z, g, j)) {
                double cs[4];
                nb = restrict_adj_cell(acc, praw, x, y, z, g, j, cs);
#pragma unroll
                for (int i = 0; i < 4; i++) sums[i] += cs[i];
            }
            psibar[c] = nb;
            if (qbar_out) qbar_out[c] = gauge_qbar(nb, praw);
        } else {
            if (PRE == PRE_NORMALIZE) acc = normalize_adj_at(praw, acc);
            psibar[c] = acc;
            if (qbar_out) qbar_out[c] = (acc.x * pc.y - acc.y * pc.x) + (acc.z * pc.w - acc.w * pc.z);
        }
        pprev = pc;
        zzprev = zf[2];
        praw = prawn;
        pc = pn;
        wc = wn;
    }
    if constexpr (RADJ) {
#pragma unroll
        for (int i = 0; i < 4; i++) sums[i] = block_sum<kOfaThreads>(sums[i]);
        if (tid == 0) {
            double* o = partials + 4 * ((size_t)blockIdx.y * gridDim.x + blockIdx.x);
#pragma unroll
            for (int i = 0; i < 4; i++) o[i] = sums[i];
        }
    }
}

// Adjoint of restrict_velocity (jet cells) at its input psi_in, in place on psibar; per
// block FP64 partial sums of phibar and phibar * pos_a (4 per block) for the gradient with
// respect to kvec and omega t.  With qbar_out, also the gauge cotangent of the new psibar
// against psi_in (k_gauge_qbar with psi_out = psi_in), for a gauge transform whose output
// psi_in is.
__global__ void __launch_bounds__(256)
k_restrict_adj(float4* __restrict__ psibar, const float4* __restrict__ psi_in, GridInfo g, JetParams j,
               double* __restrict__ partials, float* __restrict__ qbar_out) {
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    double sums[4] = {0.0, 0.0, 0.0, 0.0};
    if (c < ncells(g)) {
        int x, y, z;
        cell_xyz(c, g, x, y, z);
        const bool jet = is_jet_cell(x, y, z, g, j);
        if (qbar_out && !jet) qbar_out[c] = gauge_qbar(psibar[c], psi_in[c]);
        if (jet) {
            const float4 p = psi_in[c];
            const float4 nb = restrict_adj_cell(psibar[c], p, x, y, z, g, j, sums);
            psibar[c] = nb;
            if (qbar_out) qbar_out[c] = gauge_qbar(nb, p);
        }
    }
    restrict_adj_block_sums(sums, partials);
}

// out[k] += sum over blocks of partials[4 b + k], k < 4 (one block, fixed order)
// First level of the four sums of the restrict adjoint: block b adds the partial rows
// [b chunk, (b + 1) chunk) (fixed order) into out[4 b .. 4 b + 3].
__global__ void __launch_bounds__(256)
k_reduce4_part(const double* __restrict__ partials, size_t nrows, size_t chunk, double* __restrict__ out) {
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    const size_t lo = (size_t)blockIdx.x * chunk, hi = lo + chunk < nrows ? lo + chunk : nrows;
    for (size_t b = lo + threadIdx.x; b < hi; b += 256)
#pragma unroll
        for (int k = 0; k < 4; k++) a[k] += partials[4 * b + k];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const double s = block_sum<256>(a[k]);
        if (threadIdx.x == 0) out[4 * blockIdx.x + k] = s;
    }
}

__global__ void __launch_bounds__(256) k_reduce4(const double* __restrict__ partials, size_t nblocks, double* __restrict__ out) {
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    for (size_t b = threadIdx.x; b < nblocks; b += 256)
#pragma unroll
        for (int k = 0; k < 4; k++) a[k] += partials[4 * b + k];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const double s = block_sum<256>(a[k]);
        if (threadIdx.x == 0) out[k] += s;
    }
}

// Adjoint of normalize at its input u: ubar = (nbar - n (n . nbar)) / |u|, in place.
__global__ void __launch_bounds__(256)
k_normalize_adj(float4* __restrict__ psibar, const float4* __restrict__ psi_in, size_t N) {
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    psibar[c] = normalize_adj_at(psi_in[c], psibar[c]);
}

// ---- particles

// Staggered interpolation at x: u (3) and, for the adjoint, the position Jacobian terms and
// the 12 (cell, component) weights.
struct StagEval {
    int i[3], ip[3];
    float w[3];
    float v0[3], vn[3], pp[3], pp2[3];
};
__device__ __forceinline__ void stag_eval(const float* __restrict__ vel, const float x[3], const GridInfo& g, StagEval& e,
                                          float u[3]) {
    stag_index(x, g, e.i, e.ip, e.w);
    const int *i = e.i, *ip = e.ip;
    auto V = [&](int ix, int iy, int iz, int comp) { return vel[3 * cidx(ix, iy, iz, g) + comp]; };
    e.v0[0] = V(i[0], i[1], i[2], 0);
    e.v0[1] = V(i[0], i[1], i[2], 1);
    e.v0[2] = V(i[0], i[1], i[2], 2);
    e.vn[0] = V(i[0], ip[1], ip[2], 0);
    e.vn[1] = V(ip[0], i[1], ip[2], 1);
    e.vn[2] = V(ip[0], ip[1], i[2], 2);
    e.pp[0] = V(i[0], ip[1], i[2], 0);
    e.pp[1] = V(ip[0], i[1], i[2], 1);
    e.pp[2] = V(ip[0], i[1], i[2], 2);
    e.pp2[0] = V(i[0], i[1], ip[2], 0);
    e.pp2[1] = V(i[0], i[1], ip[2], 1);
    e.pp2[2] = V(i[0], ip[1], i[2], 2);
    const float w1[3] = {e.w[2], e.w[2], e.w[1]}, w2[3] = {e.w[1], e.w[0], e.w[0]};
#pragma unroll
    for (int a = 0; a < 3; a++)
        u[a] = (1.0f - w1[a]) * ((1.0f - w2[a]) * e.v0[a] + w2[a] * e.pp[a]) +
               w1[a] * ((1.0f - w2[a]) * e.pp2[a] + w2[a] * e.vn[a]);
}

// Workgroup-local accumulation of the velocity cotangent: an LDS box of ADJ_WX x ADJ_WY x
// ADJ_WZ cells (3 components each) whose corner is one cell below the lowest x, y and z of
// the workgroup's particle cells.  Cell-sorted slots put a workgroup's particles in a few
// neighbouring cells; the RK stages move at most about a cell and every target is a base
// cell or its +1 neighbour, so nearly all targets fall in the box.  The rest (and targets
// wrapped by the periodic boundary) go straight to global atomics.  The box is added to
// velbar once per workgroup, one atomic per touched (cell, component).
constexpr int ADJ_WX = 32, ADJ_WY = 8, ADJ_WZ = 8, ADJ_BOX = ADJ_WX * ADJ_WY * ADJ_WZ;
typedef __attribute__((address_space(3))) float lds_float;  // an LDS address (ds_* instructions)
struct AdjAcc {
    lds_float* lds;      // [ADJ_WZ][ADJ_WY][ADJ_WX][3]
    int ox, oy, oz;      // box corner (cell coordinates, workgroup-uniform)
    float* velbar;
    __device__ __forceinline__ void add(int x, int y, int z, size_t t, int a, float v) const {
        const unsigned lx = (unsigned)(x - ox), ly = (unsigned)(y - oy), lz = (unsigned)(z - oz);
        if ((lx < (unsigned)ADJ_WX) & (ly < (unsigned)ADJ_WY) & (lz < (unsigned)ADJ_WZ))
            __hip_atomic_fetch_add(lds + ((lz * ADJ_WY + ly) * ADJ_WX + lx) * 3 + a, v, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
        else
            unsafeAtomicAdd(velbar + 3 * t + a, v);
    }
};

// Cross-lane moves inside 16-lane rows and across rows by DPP (no LDS round trip):
// row_shr:n (ctrl 0x110 + n) reads lane - n of the same row, 0 where that is outside the
// row; row_bcast:15 (0x142, rows 1 and 3) reads the last lane of the previous row,
// row_bcast:31 (0x143, rows 2 and 3) reads lane 31.  Rows a broadcast does not write keep
// the given 0.
template <int CTRL>
__device__ __forceinline__ float dpp_shr(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), CTRL, 0xF, 0xF, true));
}
template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp_bcast(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, ROWS, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ int dpp_shr_i(int v) {
    return __builtin_amdgcn_mov_dpp(v, CTRL, 0xF, 0xF, true);
}
template <int CTRL, int ROWS>
__device__ __forceinline__ int dpp_bcast_i(int v) {
    return __builtin_amdgcn_update_dpp(0, v, CTRL, ROWS, 0xF, false);
}

// The 12 (cell, component) targets of one staggered evaluation are fixed by its base cell
// e.i (the others are its +1 neighbours, periodic).  The wavefront's lanes are regrouped by
// base cell: a uniform loop over the distinct cells of the wavefront (a few for
// cell-sorted slots) gives each lane its rank in the cell-major order, and ds_permute
// moves the key and the 12 values there.  Lanes in any order are grouped, which matters
// for the later RK stages, whose points have crossed cell faces independently of the
// slot order.  A segmented inclusive scan by DPP (row shifts 1, 2, 4, 8, then the row
// broadcasts) sums each group into its last lane, which adds the 12 sums to the
// workgroup's accumulator.  Every lane of the wavefront calls this (act = false: nothing
// to add); all 64 lanes must be active.
__device__ __forceinline__ void wave_scatter12(const AdjAcc& acc, const StagEval& e, const GridInfo& g, float val[12],
                                               bool act) {
    const int lane = (int)(threadIdx.x & 63u);
    constexpr unsigned kNone = 0xFFFFFFFFu;  // no cell (N < 2^32 - 1)
    unsigned key = act ? (unsigned)cidx(e.i[0], e.i[1], e.i[2], g) : kNone;
    unsigned long long rem = __ballot(true);
    int dst = 0, base = 0;
    while (rem) {
        const unsigned kl = (unsigned)__builtin_amdgcn_readlane((int)key, __builtin_ctzll(rem));
        const bool mine = key == kl;
        const unsigned long long m = __ballot(mine);
        const int below = (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
        if (mine) dst = base + below;
        base += __popcll(m);
        rem &= ~m;
    }
    const int addr = dst * 4;
    key = (unsigned)__builtin_amdgcn_ds_permute(addr, (int)key);
#pragma unroll
    for (int j = 0; j < 12; j++)
        val[j] = __builtin_bit_cast(float, __builtin_amdgcn_ds_permute(addr, __builtin_bit_cast(int, val[j])));
    // first lane of each group, propagated by a max-scan
    const unsigned kprev = __shfl_up(key, 1u, 64);
    int head = (lane == 0 || kprev != key) ? lane : 0;
    head = max(head, dpp_shr_i<0x111>(head));
    head = max(head, dpp_shr_i<0x112>(head));
    head = max(head, dpp_shr_i<0x114>(head));
    head = max(head, dpp_shr_i<0x118>(head));
    head = max(head, dpp_bcast_i<0x142, 0xA>(head));
    head = max(head, dpp_bcast_i<0x143, 0xC>(head));
    // 1 where the lane read by each step lies in the lane's group
    const float f1 = lane - 1 >= head ? 1.0f : 0.0f, f2 = lane - 2 >= head ? 1.0f : 0.0f;
    const float f4 = lane - 4 >= head ? 1.0f : 0.0f, f8 = lane - 8 >= head ? 1.0f : 0.0f;
    const float fb15 = (((lane >> 4) & 1) && head < (lane & ~15)) ? 1.0f : 0.0f;
    const float fb31 = (lane >= 32 && head < 32) ? 1.0f : 0.0f;
#pragma unroll
    for (int j = 0; j < 12; j++) {
        float v = val[j];
        v = fmaf(dpp_shr<0x111>(v), f1, v);
        v = fmaf(dpp_shr<0x112>(v), f2, v);
        v = fmaf(dpp_shr<0x114>(v), f4, v);
        v = fmaf(dpp_shr<0x118>(v), f8, v);
        v = fmaf(dpp_bcast<0x142, 0xA>(v), fb15, v);
        v = fmaf(dpp_bcast<0x143, 0xC>(v), fb31, v);
        val[j] = v;
    }
    const unsigned knext = __shfl_down(key, 1u, 64);
    if (key == kNone || (lane != 63 && knext == key)) return;
    // the group's base cell from its key
    int i[3], ip[3];
    const unsigned sxy = (unsigned)g.nx * (unsigned)g.ny;
    i[2] = (int)(key / sxy);
    const unsigned r = key - (unsigned)i[2] * sxy;
    i[1] = (int)(r / (unsigned)g.nx);
    i[0] = (int)(r - (unsigned)i[1] * (unsigned)g.nx);
    ip[0] = i[0] + 1 == g.nx ? 0 : i[0] + 1;
    ip[1] = i[1] + 1 == g.ny ? 0 : i[1] + 1;
    ip[2] = i[2] + 1 == g.nz ? 0 : i[2] + 1;
    // (pp, pp2, vn) cells of each component after the base cell, as (x, y, z)
    const int cv[3][3][3] = {{{i[0], ip[1], i[2]}, {i[0], i[1], ip[2]}, {i[0], ip[1], ip[2]}},
                             {{ip[0], i[1], i[2]}, {i[0], i[1], ip[2]}, {ip[0], i[1], ip[2]}},
                             {{ip[0], i[1], i[2]}, {i[0], ip[1], i[2]}, {ip[0], ip[1], i[2]}}};
#pragma unroll
    for (int a = 0; a < 3; a++) {
        acc.add(i[0], i[1], i[2], cidx(i[0], i[1], i[2], g), a, val[4 * a]);
#pragma unroll
        for (int q = 0; q < 3; q++) {
            const int* c = cv[a][q];
            acc.add(c[0], c[1], c[2], cidx(c[0], c[1], c[2], g), a, val[4 * a + 1 + q]);
        }
    }
}

// ubar at the evaluation point: position cotangent xbar (3) and scatter into the velocity
// cotangent (every lane of the wavefront calls this, act marks real work).
__device__ __forceinline__ void stag_adj(const StagEval& e, const float ub[3], const GridInfo& g, const AdjAcc& acc,
                                         float xb[3], bool act) {
    const float w1[3] = {e.w[2], e.w[2], e.w[1]}, w2[3] = {e.w[1], e.w[0], e.w[0]};
    float dw1[3], dw2[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        dw1[a] = -((1.0f - w2[a]) * e.v0[a] + w2[a] * e.pp[a]) + ((1.0f - w2[a]) * e.pp2[a] + w2[a] * e.vn[a]);
        dw2[a] = (1.0f - w1[a]) * (e.pp[a] - e.v0[a]) + w1[a] * (e.vn[a] - e.pp2[a]);
    }
    // w1 = (w_z, w_z, w_y), w2 = (w_y, w_x, w_x)
    xb[0] = ub[1] * dw2[1] + ub[2] * dw2[2];
    xb[1] = ub[0] * dw2[0] + ub[2] * dw1[2];
    xb[2] = ub[0] * dw1[0] + ub[1] * dw1[1];
    float val[12];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const float u = ub[a];
        val[4 * a] = u * (1.0f - w1[a]) * (1.0f - w2[a]);
        val[4 * a + 1] = u * (1.0f - w1[a]) * w2[a];
        val[4 * a + 2] = u * w1[a] * (1.0f - w2[a]);
        val[4 * a + 3] = u * w1[a] * w2[a];
    }
    wave_scatter12(acc, e, g, val, act);
}

// Adjoint of staggered_advect at the slot positions before the step (parts0): pbar (the
// cotangent of the new positions) becomes the cotangent of parts0; velbar accumulates.
// Thread k handles slot order[k] (order null: slot k).
__global__ void __launch_bounds__(256)
k_advect_adj(const float* __restrict__ parts0, long long nslots, const int* __restrict__ order,
             const float* __restrict__ vel, GridInfo g, float dt, float* __restrict__ pbar, float* __restrict__ velbar) {
    // every lane runs to the end (the scatter works across the wavefront and the
    // workgroup); lanes past the last slot or on a dead slot do the arithmetic on a
    // stand-in position and add nothing
    __shared__ float win[ADJ_BOX * 3];
    __shared__ int cmin_s[3];
    for (int i = threadIdx.x; i < ADJ_BOX * 3; i += blockDim.x) win[i] = 0.0f;
    if (threadIdx.x < 3) cmin_s[threadIdx.x] = 0x7FFFFFFF;
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const bool slot = k < nslots;
    const long long s = slot ? (order ? (long long)order[k] : k) : 0;
    float p[3] = {0.5f * g.Lx, 0.5f * g.Ly, 0.5f * g.Lz};
    if (slot) {
        p[0] = parts0[3 * s];
        p[1] = parts0[3 * s + 1];
        p[2] = parts0[3 * s + 2];
    }
    const bool in = slot & (p[0] > 0.0f) & (p[0] < g.Lx) & (p[1] > 0.0f) & (p[1] < g.Ly) & (p[2] > 0.0f) & (p[2] < g.Lz);
    if (!in) {
        p[0] = 0.5f * g.Lx;
        p[1] = 0.5f * g.Ly;
        p[2] = 0.5f * g.Lz;
    }
    // forward RK4 evaluation points P[0..3] (staggered_advect, src/parallel.jl:201-212);
    // the reverse sweep re-gathers each stage's cells, so only the points stay live
    float P[4][3];
    {
        StagEval e;
        float kk[3];
#pragma unroll
        for (int a = 0; a < 3; a++) P[0][a] = p[a];
        stag_eval(vel, P[0], g, e, kk);
#pragma unroll
        for (int a = 0; a < 3; a++) P[1][a] = p[a] + kk[a] * dt * 0.5f;
        stag_eval(vel, P[1], g, e, kk);
#pragma unroll
        for (int a = 0; a < 3; a++) P[2][a] = p[a] + kk[a] * dt * 0.5f;
        stag_eval(vel, P[2], g, e, kk);
#pragma unroll
        for (int a = 0; a < 3; a++) P[3][a] = p[a] + kk[a] * dt;
    }
    // lowest x, y and z of the workgroup's particle cells
    int m[3] = {0x7FFFFFFF, 0x7FFFFFFF, 0x7FFFFFFF};
    if (in) {
        int i[3], ip[3];
        float w[3];
        stag_index(P[0], g, i, ip, w);
        m[0] = i[0];
        m[1] = i[1];
        m[2] = i[2];
    }
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) m[a] = min(m[a], __shfl_xor(m[a], off, 64));
    __syncthreads();  // zeroed box and cmin_s
    if ((threadIdx.x & 63u) == 0) {
        atomicMin(&cmin_s[0], m[0]);
        atomicMin(&cmin_s[1], m[1]);
        atomicMin(&cmin_s[2], m[2]);
    }
    __syncthreads();
    AdjAcc acc;
    acc.lds = (lds_float*)win;
    acc.velbar = velbar;
    acc.ox = __builtin_amdgcn_readfirstlane(cmin_s[0]) - 1;
    acc.oy = __builtin_amdgcn_readfirstlane(cmin_s[1]) - 1;
    acc.oz = __builtin_amdgcn_readfirstlane(cmin_s[2]) - 1;
    float pb[3] = {0.0f, 0.0f, 0.0f};
    if (in) {
        pb[0] = pbar[3 * s];
        pb[1] = pbar[3 * s + 1];
        pb[2] = pbar[3 * s + 2];
    }
    // reverse sweep: kb[q] = cotangent of k_{q+1}; xb = cotangent of the point P[q]
    const float c6 = dt / 6.0f;
    const float wk[4] = {c6, 2.0f * c6, 2.0f * c6, c6};  // weights of k1..k4 in the update
    const float wp[4] = {0.0f, 0.5f * dt, 0.5f * dt, dt}; // P[q] = p + wp[q] k_q
    float pacc[3], kb[3], xb[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int a = 0; a < 3; a++) pacc[a] = pb[a];
#pragma unroll
    for (int q = 3; q >= 0; q--) {
        // cotangent of k_{q+1}: its weight in the update plus what P[q+1] passed back
#pragma unroll
        for (int a = 0; a < 3; a++) kb[a] = wk[q] * pb[a] + (q < 3 ? wp[q + 1] * xb[a] : 0.0f);
        StagEval e;
        float u[3];
        stag_eval(vel, P[q], g, e, u);
        stag_adj(e, kb, g, acc, xb, in);
#pragma unroll
        for (int a = 0; a < 3; a++) pacc[a] += xb[a];
    }
    if (slot) {
        // the dead marker does not depend on anything
#pragma unroll
        for (int a = 0; a < 3; a++) pbar[3 * s + a] = in ? pacc[a] : 0.0f;
    }
    // add the box to velbar (entries never touched are zero and skipped; a touched entry
    // is a valid cell, since only targets inside the grid were added)
    __syncthreads();
    for (int i = threadIdx.x; i < ADJ_BOX * 3; i += blockDim.x) {
        const float v = win[i];
        if (v != 0.0f) {
            const int c = i / 3, lx = c % ADJ_WX, ly = (c / ADJ_WX) % ADJ_WY, lz = c / (ADJ_WX * ADJ_WY);
            unsafeAtomicAdd(velbar + 3 * cidx(acc.ox + lx, acc.oy + ly, acc.oz + lz, g) + i % 3, v);
        }
    }
}

// Slots overwritten by add_particles! (test/runtests.jl:52-54) carry no cotangent back.
__global__ void __launch_bounds__(256)
k_zero_slots(float* __restrict__ pbar, long long nslots, long long first, long long n) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const long long s = (first + k) % nslots;
    pbar[3 * s] = 0.0f;
    pbar[3 * s + 1] = 0.0f;
    pbar[3 * s + 2] = 0.0f;
}

// ======================================================================== splat loss

// Particle density on a grid of cell centres y = (j + 1/2) h over the box:
// rho(y) = inv_n * sum over live slots of exp(-|y - p|^2 / (2 sigma^2)) for |y - p| <= R.
__device__ __forceinline__ bool splat_live(const float p[3], const SplatCfg& s) {
    return (p[0] > 0.0f) & (p[0] < s.L[0]) & (p[1] > 0.0f) & (p[1] < s.L[1]) & (p[2] > 0.0f) & (p[2] < s.L[2]);
}
__device__ __forceinline__ void splat_range(float p, float h, float R, int n, int& lo, int& hi) {
    lo = (int)ceilf((p - R) / h - 0.5f);
    hi = (int)floorf((p + R) / h - 0.5f);
    lo = lo < 0 ? 0 : lo;
    hi = hi > n - 1 ? n - 1 : hi;
}

__global__ void __launch_bounds__(256)
k_splat(const float* __restrict__ parts, long long nslots, SplatCfg sc, double* __restrict__ rho) {
    const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslots) return;
    const float p[3] = {parts[3 * s], parts[3 * s + 1], parts[3 * s + 2]};
    if (!splat_live(p, sc)) return;
    int lo[3], hi[3];
#pragma unroll
    for (int a = 0; a < 3; a++) splat_range(p[a], sc.h[a], sc.R, sc.res[a], lo[a], hi[a]);
    const float R2 = sc.R * sc.R, k = -0.5f / (sc.sigma * sc.sigma);
    for (int jz = lo[2]; jz <= hi[2]; jz++) {
        const float dz = ((float)jz + 0.5f) * sc.h[2] - p[2];
        for (int jy = lo[1]; jy <= hi[1]; jy++) {
            const float dy = ((float)jy + 0.5f) * sc.h[1] - p[1];
            const float r2yz = dy * dy + dz * dz;
            if (r2yz > R2) continue;
            for (int jx = lo[0]; jx <= hi[0]; jx++) {
                const float dx = ((float)jx + 0.5f) * sc.h[0] - p[0];
                const float r2 = dx * dx + r2yz;
                if (r2 > R2) continue;
                const size_t j = (size_t)jx + (size_t)sc.res[0] * ((size_t)jy + (size_t)sc.res[1] * jz);
                unsafeAtomicAdd(rho + j, (double)(__expf(r2 * k) * sc.inv_n));
            }
        }
    }
}

// Same density, per block in LDS: each block accumulates its share of the slots into its
// own grid and writes it to partial[b * M ...]; k_splat_reduce sums the block grids.
//
// The Gaussian factors over the axes, exp(k |y - p|^2) = exp(k dx^2) exp(k (dy^2 + dz^2)):
// a slot takes one exp per x offset of its support (at most XW, held in registers) and
// one per (y, z) row, instead of one per cell; the cut-off |y - p| <= R is still decided
// per cell on dx^2 + (dy^2 + dz^2).  A slot wider than XW in x (possible only through
// rounding at the support's ends) takes the per-cell path.
//
// LDS traffic: a cloud of slots a few cells across puts many lanes of a wavefront on the
// same cells.  Lanes take slots nslots / 64 apart (neighbouring slots hold neighbouring
// particles: the nozzle fill and the spawn batches are stored in spatial order), each lane
// starts its walk over the support rows at its own row (13 lane mod rows), and the grid
// rows in LDS are padded (x stride 1, y stride res0 + 1, z stride (res0 + 1) res1 + 1) so
// that equal x offsets of different rows fall in different banks.
//
// Exactness (cut-off R <= 4 sigma, fewer than 2^25 slots): every term is a float between
// exp(-8) inv_n > 2^-12 inv_n and inv_n, hence a multiple of 2^-35 of inv_n's binade, and a
// block adds at most 2^17 terms per cell, so its FP64 sums are exact and do not depend on
// the order of the atomics (each halving of the smallest term costs one bit of margin).
constexpr int SPLAT_THREADS = 1024;
constexpr int SPLAT_BLOCKS = 256;
constexpr int SPLAT_LDS_DOUBLES = 16384;  // 128 KiB of the CU's 160
__host__ __device__ inline int splat_sy(const SplatCfg& sc) { return sc.res[0] + 1; }
__host__ __device__ inline int splat_sz(const SplatCfg& sc) { return (sc.res[0] + 1) * sc.res[1] + 1; }
template <int XW>
__global__ void __launch_bounds__(SPLAT_THREADS)
k_splat_lds(const float* __restrict__ parts, long long nslots, SplatCfg sc, double* __restrict__ partial) {
    extern __shared__ double acc[];
    const int M = sc.res[0] * sc.res[1] * sc.res[2];
    const int sy = splat_sy(sc), sz = splat_sz(sc), MP = sz * sc.res[2];
    for (int j = threadIdx.x; j < MP; j += SPLAT_THREADS) acc[j] = 0.0;
    __syncthreads();
    const float R2 = sc.R * sc.R, k = -0.5f / (sc.sigma * sc.sigma);
    // slot lane * W * J + j * W + wave: a bijection of (lane, j, wave) onto [0, 64 W J),
    // which covers [0, nslots) (W waves in the grid, J rounds)
    const long long W = (long long)gridDim.x * (SPLAT_THREADS / 64), J = (nslots + 64 * W - 1) / (64 * W);
    const long long wave = ((long long)blockIdx.x * SPLAT_THREADS + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    for (long long j = 0; j < J; j++) {
        const long long s = lane * W * J + j * W + wave;
        if (s >= nslots) continue;
        const float p[3] = {parts[3 * s], parts[3 * s + 1], parts[3 * s + 2]};
        if (!splat_live(p, sc)) continue;
        int lo[3], hi[3];
#pragma unroll
        for (int a = 0; a < 3; a++) splat_range(p[a], sc.h[a], sc.R, sc.res[a], lo[a], hi[a]);
        const int nx = hi[0] - lo[0] + 1, ny = hi[1] - lo[1] + 1, nz = hi[2] - lo[2] + 1, nr = ny * nz;
        if (nx <= 0 || nr <= 0) continue;
        const int r0 = (lane * 13) % nr;
        int iy = r0 % ny, iz = r0 / ny;
        if (nx <= XW) {
            float dx2[XW], ex[XW];
#pragma unroll
            for (int i = 0; i < XW; i++) {
                const float dx = ((float)(lo[0] + i) + 0.5f) * sc.h[0] - p[0];
                dx2[i] = dx * dx;
                ex[i] = __expf(dx2[i] * k) * sc.inv_n;
            }
            for (int r = 0; r < nr; r++) {
                const int jy = lo[1] + iy, jz = lo[2] + iz;
                const float dz = ((float)jz + 0.5f) * sc.h[2] - p[2];
                const float dy = ((float)jy + 0.5f) * sc.h[1] - p[1];
                const float r2yz = dy * dy + dz * dz;
                if (r2yz <= R2) {
                    const float eyz = __expf(r2yz * k);
                    double* row = acc + lo[0] + sy * jy + sz * jz;
#pragma unroll
                    for (int i = 0; i < XW; i++)
                        if (i < nx && dx2[i] + r2yz <= R2) atomicAdd(row + i, (double)(eyz * ex[i]));
                }
                if (++iy == ny) {
                    iy = 0;
                    if (++iz == nz) iz = 0;
                }
            }
        } else {
            for (int r = 0; r < nr; r++) {
                const int jy = lo[1] + iy, jz = lo[2] + iz;
                const float dz = ((float)jz + 0.5f) * sc.h[2] - p[2];
                const float dy = ((float)jy + 0.5f) * sc.h[1] - p[1];
                const float r2yz = dy * dy + dz * dz;
                if (r2yz <= R2) {
                    const float eyz = __expf(r2yz * k);
                    double* row = acc + sy * jy + sz * jz;
                    for (int jx = lo[0]; jx <= hi[0]; jx++) {
                        const float dx = ((float)jx + 0.5f) * sc.h[0] - p[0];
                        const float dx2 = dx * dx;
                        if (dx2 + r2yz <= R2) atomicAdd(row + jx, (double)(eyz * (__expf(dx2 * k) * sc.inv_n)));
                    }
                }
                if (++iy == ny) {
                    iy = 0;
                    if (++iz == nz) iz = 0;
                }
            }
        }
    }
    __syncthreads();
    double* out = partial + (size_t)blockIdx.x * M;
    for (int j = threadIdx.x; j < M; j += SPLAT_THREADS) {
        const int x = j % sc.res[0], y = (j / sc.res[0]) % sc.res[1], z = j / (sc.res[0] * sc.res[1]);
        out[j] = acc[x + sy * y + sz * z];
    }
}
// rho[j] = sum over the block grids: 32 cells per block, 8 groups of threads each summing
// every 8th block grid, the 8 group sums added in group order (a fixed order)
__global__ void __launch_bounds__(256)
k_splat_reduce(const double* __restrict__ partial, int nblocks, int M, double* __restrict__ rho) {
    __shared__ double part[8][32];
    const int c = threadIdx.x & 31, q = threadIdx.x >> 5, j = blockIdx.x * 32 + c;
    double s = 0.0;
    if (j < M)
        for (int b = q; b < nblocks; b += 8) s += partial[(size_t)b * M + j];
    part[q][c] = s;
    __syncthreads();
    if (q == 0 && j < M) {
        double t = part[0][c];
#pragma unroll
        for (int g = 1; g < 8; g++) t += part[g][c];
        rho[j] = t;
    }
}

// resid = 2/M (rho - target) (the loss cotangent of rho); per-block FP64 sums of (rho - target)^2
__global__ void __launch_bounds__(256)
k_splat_loss(const double* __restrict__ rho, const float* __restrict__ target, int M, float* __restrict__ resid,
             double* __restrict__ partials) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    double e2 = 0.0;
    if (j < M) {
        const double e = rho[j] - (double)target[j];
        resid[j] = (float)(2.0 * e / (double)M);
        e2 = e * e;
    }
    const double s = block_sum<256>(e2);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// pbar = d loss / d p = sum_y resid(y) inv_n G(y - p) (y - p) / sigma^2 over the splat support,
// with the Gaussian factored over the axes as in k_splat_lds: per (y, z) row,
// sum_x resid G_x dx and sum_x resid G_x, then times G_yz (and dy, dz).
template <int XW>
__global__ void __launch_bounds__(256)
k_splat_adj(const float* __restrict__ parts, long long nslots, SplatCfg sc, const float* __restrict__ resid,
            float* __restrict__ pbar) {
    const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslots) return;
    const float p[3] = {parts[3 * s], parts[3 * s + 1], parts[3 * s + 2]};
    float g[3] = {0.0f, 0.0f, 0.0f};
    if (splat_live(p, sc)) {
        int lo[3], hi[3];
#pragma unroll
        for (int a = 0; a < 3; a++) splat_range(p[a], sc.h[a], sc.R, sc.res[a], lo[a], hi[a]);
        const float R2 = sc.R * sc.R, k = -0.5f / (sc.sigma * sc.sigma), is2 = 1.0f / (sc.sigma * sc.sigma);
        const int nx = hi[0] - lo[0] + 1;
        float dxv[XW], ex[XW];
#pragma unroll
        for (int i = 0; i < XW; i++) {
            dxv[i] = ((float)(lo[0] + i) + 0.5f) * sc.h[0] - p[0];
            ex[i] = __expf(dxv[i] * dxv[i] * k) * (sc.inv_n * is2);
        }
        for (int jz = lo[2]; jz <= hi[2]; jz++) {
            const float dz = ((float)jz + 0.5f) * sc.h[2] - p[2];
            for (int jy = lo[1]; jy <= hi[1]; jy++) {
                const float dy = ((float)jy + 0.5f) * sc.h[1] - p[1];
                const float r2yz = dy * dy + dz * dz;
                if (r2yz > R2) continue;
                const float* row = resid + (size_t)sc.res[0] * ((size_t)jy + (size_t)sc.res[1] * jz);
                float sx = 0.0f, s1 = 0.0f;
                if (nx <= XW) {
#pragma unroll
                    for (int i = 0; i < XW; i++)
                        if (i < nx && dxv[i] * dxv[i] + r2yz <= R2) {
                            const float w = row[lo[0] + i] * ex[i];
                            sx += w * dxv[i];
                            s1 += w;
                        }
                } else {
                    for (int jx = lo[0]; jx <= hi[0]; jx++) {
                        const float dx = ((float)jx + 0.5f) * sc.h[0] - p[0];
                        if (dx * dx + r2yz <= R2) {
                            const float w = row[jx] * (__expf(dx * dx * k) * (sc.inv_n * is2));
                            sx += w * dx;
                            s1 += w;
                        }
                    }
                }
                const float eyz = __expf(r2yz * k);
                g[0] += eyz * sx;
                g[1] += eyz * s1 * dy;
                g[2] += eyz * s1 * dz;
            }
        }
    }
    pbar[3 * s] = g[0];
    pbar[3 * s + 1] = g[1];
    pbar[3 * s + 2] = g[2];
}

// ======================================================================== launchers

int launch_fill_psi(float4* psi, float4 v, size_t n, hipStream_t s) {
    ISF_LAUNCH(k_fill_psi, grid1d(n, 256), dim3(256), s, psi, v, n);
}
int launch_fill_f32(float* p, float v, size_t n, hipStream_t s) {
    if (n == 0) return 0;
    if ((reinterpret_cast<uintptr_t>(p) & 15u) == 0)
        ISF_LAUNCH(k_fill_f32x4, grid1d(n / 4 > 0 ? n / 4 : 1, 256), dim3(256), s, p, v, n);
    ISF_LAUNCH(k_fill_f32, grid1d(n, 256), dim3(256), s, p, v, n);
}
int launch_fill_f64(double* p, double v, size_t n, hipStream_t s) {
    if (n == 0) return 0;
    ISF_LAUNCH(k_fill_f64, grid1d(n, 256), dim3(256), s, p, v, n);
}
int launch_apply_phase(const GridInfo& g, float4* psi, const float* theta, const int pr[3], hipStream_t s) {
    const PhaseGrid P{{pr[0], pr[1], pr[2]}};
    ISF_LAUNCH(k_apply_phase, grid1d(ncells(g), 256), dim3(256), s, psi, theta, g, P);
}
int launch_apply_phase_adj(const GridInfo& g, float4* psibar, const float4* psi_out, const float* theta, float* phibar,
                           const int pr[3], hipStream_t s) {
    const PhaseGrid P{{pr[0], pr[1], pr[2]}};
    ISF_LAUNCH(k_apply_phase_adj, grid1d(ncells(g), 256), dim3(256), s, psibar, psi_out, theta, phibar, g, P);
}
int launch_phase_gather(const GridInfo& g, const float* phibar, float* thbar, const int pr[3], hipStream_t s) {
    const PhaseGrid P{{pr[0], pr[1], pr[2]}};
    ISF_LAUNCH(k_phase_gather, dim3((unsigned)(pr[0] * pr[1] * pr[2])), dim3(256), s, phibar, thbar, g, P);
}
int launch_normalize_to(const GridInfo& g, const float4* in, float4* out, hipStream_t s) {
    ISF_LAUNCH(k_normalize_to, grid1d(ncells(g), 256), dim3(256), s, in, out, ncells(g));
}
int launch_restrict_to(const GridInfo& g, const float4* in, float4* out, const JetParams& j, hipStream_t s) {
    ISF_LAUNCH(k_restrict_to, grid1d(ncells(g), 256), dim3(256), s, in, out, g, j);
}
int launch_gauge_qbar(const GridInfo& g, const float4* psibar, const float4* psi_out, float* qbar, hipStream_t s) {
    ISF_LAUNCH(k_gauge_qbar, grid1d(ncells(g), 256), dim3(256), s, psibar, psi_out, qbar, ncells(g));
}
// the tiled one-form adjoint: rows of 64 to 256 cells; z planes per block up to 16, halved
// (down to 2) while the launch would have fewer than 1024 blocks
static bool ofa_tiled(const GridInfo& g) { return g.nx >= 64 && g.nx <= 256; }
static int ofa_tz(const GridInfo& g) {
    const long rows = g.ny / (kOfaThreads / g.nx);
    int tz = 16;
    while (tz > 2 && rows * (g.nz / tz) < 1024) tz >>= 1;
    return tz;
}
static dim3 ofa_grid(const GridInfo& g, int tz) { return dim3(g.ny / (kOfaThreads / g.nx), g.nz / tz); }
int launch_oneform_adj_div(const GridInfo& g, float4* psibar, const float4* psi, const float* q, const float* fbar,
                           hipStream_t s, int pre, const JetParams* j) {
    const JetParams jp = j ? *j : JetParams{};
    if (pre == PRE_RESTRICT && !j) return (int)hipErrorInvalidValue;
    if (ofa_tiled(g)) {
        const int tz = ofa_tz(g);
        const dim3 grid = ofa_grid(g, tz), block(kOfaThreads);
        if (pre == PRE_NORMALIZE)
            ISF_LAUNCH((k_oneform_adj_tile<OF_DIV, PRE_NORMALIZE, false>), grid, block, s, psibar, psi, q, fbar, g,
                       1.0f, 0, jp, tz, nullptr, nullptr);
        if (pre == PRE_RESTRICT)
            ISF_LAUNCH((k_oneform_adj_tile<OF_DIV, PRE_RESTRICT, false>), grid, block, s, psibar, psi, q, fbar, g,
                       1.0f, 0, jp, tz, nullptr, nullptr);
        ISF_LAUNCH((k_oneform_adj_tile<OF_DIV, PRE_NONE, false>), grid, block, s, psibar, psi, q, fbar, g, 1.0f, 0,
                   jp, tz, nullptr, nullptr);
    }
    if (pre == PRE_NORMALIZE)
        ISF_LAUNCH((k_oneform_adj<OF_DIV, PRE_NORMALIZE>), grid1d(ncells(g), 256), dim3(256), s, psibar, psi, q, fbar,
                   g, 1.0f, 0, jp, nullptr);
    if (pre == PRE_RESTRICT) {
        if (!j) return (int)hipErrorInvalidValue;
        ISF_LAUNCH((k_oneform_adj<OF_DIV, PRE_RESTRICT>), grid1d(ncells(g), 256), dim3(256), s, psibar, psi, q, fbar,
                   g, 1.0f, 0, jp, nullptr);
    }
    ISF_LAUNCH((k_oneform_adj<OF_DIV, PRE_NONE>), grid1d(ncells(g), 256), dim3(256), s, psibar, psi, q, fbar, g, 1.0f,
               0, jp, nullptr);
}
int launch_oneform_adj_vel(const GridInfo& g, float4* psibar, const float4* psi, const float* velbar, float hbar,
                           int sharp, hipStream_t s, float* qbar_out) {
    if (ofa_tiled(g)) {
        const int tz = ofa_tz(g);
        ISF_LAUNCH((k_oneform_adj_tile<OF_VEL, PRE_NONE, false>), ofa_grid(g, tz), dim3(kOfaThreads), s, psibar, psi,
                   nullptr, velbar, g, hbar, sharp, JetParams{}, tz, qbar_out, nullptr);
    }
    ISF_LAUNCH((k_oneform_adj<OF_VEL, PRE_NONE>), grid1d(ncells(g), 256), dim3(256), s, psibar, psi, nullptr, velbar,
               g, hbar, sharp, JetParams{}, qbar_out);
}
size_t restrict_adj_blocks(const GridInfo& g) { return (ncells(g) + 255) / 256; }
// the first reduction level's blocks (their sums follow the per-block rows in partials)
constexpr size_t kReduce4Blocks = 128;
size_t restrict_adj_partials(const GridInfo& g) { return 4 * restrict_adj_blocks(g) + 4 * kReduce4Blocks; }
// out4 += the four sums of the nb per-block rows in partials (two levels, fixed order)
static int restrict_adj_reduce(size_t nb, double* partials, double* out4, hipStream_t s) {
    const size_t G = nb < kReduce4Blocks ? nb : kReduce4Blocks, chunk = (nb + G - 1) / G;
    double* part2 = partials + 4 * nb;
    hipLaunchKernelGGL(k_reduce4_part, dim3((unsigned)G), dim3(256), 0, s, partials, nb, chunk, part2);
    const int e = (int)hipGetLastError();
    if (e) return e;
    ISF_LAUNCH(k_reduce4, dim3(1), dim3(256), s, part2, G, out4);
}
int launch_restrict_adj(const GridInfo& g, float4* psibar, const float4* psi_in, const JetParams& j, double* partials,
                        double* out4, hipStream_t s, float* qbar_out) {
    hipLaunchKernelGGL(k_restrict_adj, grid1d(ncells(g), 256), dim3(256), 0, s, psibar, psi_in, g, j, partials,
                       qbar_out);
    const int e = (int)hipGetLastError();
    if (e) return e;
    return restrict_adj_reduce(restrict_adj_blocks(g), partials, out4, s);
}
int launch_oneform_adj_div_restrict(const GridInfo& g, float4* psibar, const float4* psi, const float* q,
                                    const float* fbar, const JetParams& j, double* partials, double* out4,
                                    float* qbar_out, hipStream_t s) {
    size_t nb = restrict_adj_blocks(g);
    if (ofa_tiled(g)) {
        // at most restrict_adj_blocks(g) blocks (each holds 512 cells or more)
        const int tz = ofa_tz(g);
        const dim3 grid = ofa_grid(g, tz);
        nb = (size_t)grid.x * grid.y;
        hipLaunchKernelGGL((k_oneform_adj_tile<OF_DIV, PRE_RESTRICT, true>), grid, dim3(kOfaThreads), 0, s, psibar,
                           psi, q, fbar, g, 1.0f, 0, j, tz, qbar_out, partials);
    } else {
        hipLaunchKernelGGL((k_oneform_adj<OF_DIV, PRE_RESTRICT, true>), grid1d(ncells(g), 256), dim3(256), 0, s,
                           psibar, psi, q, fbar, g, 1.0f, 0, j, qbar_out, partials);
    }
    const int e = (int)hipGetLastError();
    if (e) return e;
    return restrict_adj_reduce(nb, partials, out4, s);
}
int launch_normalize_adj(const GridInfo& g, float4* psibar, const float4* psi_in, hipStream_t s) {
    ISF_LAUNCH(k_normalize_adj, grid1d(ncells(g), 256), dim3(256), s, psibar, psi_in, ncells(g));
}
int launch_advect_adj(const GridInfo& g, const float* parts0, long long nslots, const int* order, const float* vel,
                      float dt, float* pbar, float* velbar, hipStream_t s) {
    if (nslots <= 0) return 0;
    ISF_LAUNCH(k_advect_adj, grid1d((size_t)nslots, 256), dim3(256), s, parts0, nslots, order, vel, g, dt, pbar, velbar);
}
int launch_zero_slots(float* pbar, long long nslots, long long first, long long n, hipStream_t s) {
    if (n <= 0 || nslots <= 0) return 0;
    ISF_LAUNCH(k_zero_slots, grid1d((size_t)n, 256), dim3(256), s, pbar, nslots, first, n);
}
// register width of the x support: a slot's support holds floor(2R / h) + 1 cell centres
// in x (one more only through rounding at its ends, which the kernels send to their
// per-cell path)
static int splat_x_width(const SplatCfg& sc) {
    const double w = std::floor(2.0 * (double)sc.R / (double)sc.h[0]) + 1.0;
    return w <= 9.0 ? 9 : 17;
}
// 0 when the padded grid does not fit in LDS or the support is wider than 17 (global atomics)
static int splat_xw(const SplatCfg& sc) {
    const long long mp = (long long)splat_sz(sc) * sc.res[2];
    if (mp > SPLAT_LDS_DOUBLES) return 0;
    return std::floor(2.0 * (double)sc.R / (double)sc.h[0]) + 1.0 <= 17.0 ? splat_x_width(sc) : 0;
}
size_t splat_partial_doubles(const SplatCfg& sc) {
    const size_t M = (size_t)sc.res[0] * sc.res[1] * sc.res[2];
    return splat_xw(sc) ? M * SPLAT_BLOCKS : 0;
}
// rho must hold zeros when partial is null (global atomics); otherwise it is overwritten
int launch_splat(const float* parts, long long nslots, const SplatCfg& sc, double* rho, double* partial,
                 hipStream_t s) {
    const int M = sc.res[0] * sc.res[1] * sc.res[2];
    const int xw = splat_xw(sc);
    if (partial && xw) {
        const size_t lds = sizeof(double) * (size_t)splat_sz(sc) * sc.res[2];
        if (xw == 9)
            hipLaunchKernelGGL(k_splat_lds<9>, dim3(SPLAT_BLOCKS), dim3(SPLAT_THREADS), lds, s, parts, nslots, sc,
                               partial);
        else
            hipLaunchKernelGGL(k_splat_lds<17>, dim3(SPLAT_BLOCKS), dim3(SPLAT_THREADS), lds, s, parts, nslots, sc,
                               partial);
        int e = (int)hipGetLastError();
        if (e) return e;
        ISF_LAUNCH(k_splat_reduce, dim3((M + 31) / 32), dim3(256), s, partial, SPLAT_BLOCKS, M, rho);
    }
    if (nslots <= 0) return 0;
    ISF_LAUNCH(k_splat, grid1d((size_t)nslots, 256), dim3(256), s, parts, nslots, sc, rho);
}
size_t splat_loss_blocks(int M) { return ((size_t)M + 255) / 256; }
int launch_splat_loss(const double* rho, const float* target, int M, float* resid, double* partials, double* out,
                      hipStream_t s) {
    hipLaunchKernelGGL(k_splat_loss, grid1d((size_t)M, 256), dim3(256), 0, s, rho, target, M, resid, partials);
    int e = (int)hipGetLastError();
    if (e) return e;
    return launch_reduce_sum(partials, splat_loss_blocks(M), out, s);
}
int launch_splat_adj(const float* parts, long long nslots, const SplatCfg& sc, const float* resid, float* pbar,
                     hipStream_t s) {
    if (nslots <= 0) return 0;
    if (splat_x_width(sc) == 9)
        ISF_LAUNCH(k_splat_adj<9>, grid1d((size_t)nslots, 256), dim3(256), s, parts, nslots, sc, resid, pbar);
    ISF_LAUNCH(k_splat_adj<17>, grid1d((size_t)nslots, 256), dim3(256), s, parts, nslots, sc, resid, pbar);
}

}  // namespace isf
