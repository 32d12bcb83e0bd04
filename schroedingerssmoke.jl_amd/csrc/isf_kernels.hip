// isf_kernels.hip -- hand-written gfx950 kernels of the ISF step and their launchers.
//
// Data layouts in HBM (x fastest, the reference's column-major order):
//   psi  : float4 per cell {re1, im1, re2, im2}   (test/runtests.jl:19)
//   vel  : 3 floats per cell {vx, vy, vz}          (src/parallel.jl:14,66)
//   f, q : 1 float per cell
//   S    : Schroedinger spectrum, planar complex, plane (c, kx) = S + (c*nx + kx)*ny*nz,
//          element (ky, kz) at kz*ny + ky
//   P    : Poisson half spectrum, plane kx in [0, nx/2] = P + kx*ny*nz, same plane layout
//   parts: 3 floats per slot                       (test/runtests.jl:33)
//
// One 3-D transform = an x pass (contiguous lines, written transposed into kx planes) and
// a plane pass (y and z transforms of a whole kx plane held in LDS, with the spectral
// multiplier applied between the forward and the inverse transforms).
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "isf_device.hpp"
#include "isf_fft.hpp"
#include "isf_internal.hpp"
#include "isf_philox.hpp"

namespace isf {

template <int NX> struct XCfg {
    static constexpr int E = NX == 512 ? 16 : DefaultE<NX>::value;
    static constexpr int T = NX / E;
    static constexpr int LS = padded_len(NX);
};
// The C2R gauge pass's half-length transform: 8 elements per thread for M = 128 (twice
// the threads of XCfg; that pass also streams psi in and out, and more wavefronts hide
// those loads: 33 -> 29.5 us at 256x128x128).
template <int M> struct C2rCfg {
    static constexpr int E = M == 128 ? 8 : DefaultE<M>::value;
    static constexpr int T = M / E;
};

// Tile of this block: x tile, y tile and z slab (member-major in batch mode).  Blocks are
// mapped XCD-aware: the hardware deals consecutive block ids to the 8 XCDs in turn, so
// tile t is given to id (t mod C) * 8 + t / C, C = blocks / 8; each XCD then walks one
// contiguous range of tiles, and rows or cache lines shared by neighbouring tiles stay in
// that XCD's L2.
struct StTile {
    int bx, by, bz;
};
__device__ __forceinline__ StTile stencil_tile() {
    const unsigned T = gridDim.x * gridDim.y * gridDim.z;
    unsigned lin = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    if ((T & 7u) == 0u) {
        const unsigned C = T >> 3;
        lin = (lin & 7u) * C + (lin >> 3);
    }
    StTile t;
    t.bx = (int)(lin % gridDim.x);
    lin /= gridDim.x;
    t.by = (int)(lin % gridDim.y);
    t.bz = (int)(lin / gridDim.y);
    return t;
}

// ======================================================================== Schroedinger x passes

// Forward x transform of TY psi lines (fixed z), both components, written into the
// (c, kx) planes of S with y fastest (src/parallel.jl:179-181, x part of fft!).  With
// sa != null the components come from two separate complex arrays sa, sb (the MatlabPort
// layout, src/MatlabPort.jl:87-93) instead of psi.
// VEL (the step loop of isf_simloop): the pass also writes velocity_one_form!(isf, psi,
// hbar) of the psi it reads, with the driver's StaggeredSharp (src/parallel.jl:115-122,
// test/runtests.jl:59-62) -- the field the previous step's advection needs.  Each edge
// angle has the operands and arithmetic of k_velocity (bit-identical field): psi(c+e_x)
// and psi(c+e_y) come from the block's lines in LDS, the row above the tile and the lines
// of the next z plane from global memory, loaded with the block's own lines.
template <int NX, int TY, bool VEL>
__global__ void __launch_bounds__(2 * TY * XCfg<NX>::T)
k_flow_x_fwd(const float4* __restrict__ psi, const float2* __restrict__ sa, const float2* __restrict__ sb,
             float2* __restrict__ S, int ny, int nz, const float2* __restrict__ tw, Batch B, VelOut vo) {
    constexpr int T = XCfg<NX>::T, E = XCfg<NX>::E, LS = XCfg<NX>::LS;
    const StTile tl = stencil_tile();  // x: y tile, y: z, z: member
    if (B.inst) {
        const Inst& m = B.inst[tl.bz];
        psi = m.psi;
        S = m.S;
        tw = m.tb.tw;
        if constexpr (VEL) {
            vo.vel = ((*B.iter - 1) & 1) ? m.vel_alt : m.vel;
            vo.hbar = m.hbar;
            vo.s0 = 1.0f / m.g.dx;
            vo.s1 = 1.0f / m.g.dy;
            vo.s2 = 1.0f / m.g.dz;
            vo.sharp = 1;
        }
    }
    constexpr int NT = 2 * TY * T;
    __shared__ float2 lds[2 * TY * LS];
    __shared__ float2 twl[FftTw<NX, E>::SIZE];
    const int tid = threadIdx.x;
    const int y0 = tl.bx * TY, z = tl.by;
    const size_t line0 = ((size_t)z * ny + y0) * NX;
    // the block's lines: all loads issued before the twiddle-table loads and the LDS stores
    constexpr int K = TY * NX / NT;
    static_assert(!VEL || NT >= NX, "one cell per line and thread");
    float2 va[K], vb[K];
    float4 pz[VEL ? K : 1], ph = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (sa) {
#pragma unroll
        for (int k = 0; k < K; k++) {
            va[k] = sa[line0 + tid + k * NT];
            vb[k] = sb[line0 + tid + k * NT];
        }
    } else {
#pragma unroll
        for (int k = 0; k < K; k++) {
            const float4 a = psi[line0 + tid + k * NT];
            va[k] = make_float2(a.x, a.y);
            vb[k] = make_float2(a.z, a.w);
        }
        if constexpr (VEL) {
            // psi(c + e_z) of the block's cells, psi(c + e_y) of its last line
            const size_t znext0 = ((size_t)(z + 1 < nz ? z + 1 : 0) * ny + y0) * NX;
#pragma unroll
            for (int k = 0; k < K; k++) pz[k] = psi[znext0 + tid + k * NT];
            const int i_last = tid + (K - 1) * NT;  // the thread's cell on the last line, if any
            if (i_last / NX == TY - 1) {
                const int yn = y0 + TY < ny ? y0 + TY : 0;
                ph = psi[((size_t)z * ny + yn) * NX + i_last % NX];
            }
        }
    }
    FftTw<NX, E>::template fill<NT>(twl, tw, tid);
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int i = tid + k * NT, l = i / NX, x = i % NX;
        lds[l * LS + pad16(x)] = va[k];
        lds[(TY + l) * LS + pad16(x)] = vb[k];
    }
    __syncthreads();
    if constexpr (VEL) {
        if (!sa) {
#pragma unroll
            for (int k = 0; k < K; k++) {
                const int i = tid + k * NT, l = i / NX, x = i % NX, xp = x + 1 < NX ? x + 1 : 0;
                const float4 pc = make_float4(va[k].x, va[k].y, vb[k].x, vb[k].y);
                const float2 ax = lds[l * LS + pad16(xp)], bx = lds[(TY + l) * LS + pad16(xp)];
                const float4 px = make_float4(ax.x, ax.y, bx.x, bx.y);
                float4 py = ph;
                if (l + 1 < TY) {
                    const float2 ay = lds[(l + 1) * LS + pad16(x)], by = lds[(TY + l + 1) * LS + pad16(x)];
                    py = make_float4(ay.x, ay.y, by.x, by.y);
                }
                float v0 = edge_angle(pc, px) * vo.hbar, v1 = edge_angle(pc, py) * vo.hbar,
                      v2 = edge_angle(pc, pz[k]) * vo.hbar;
                if (vo.sharp) {
                    v0 = v0 * vo.s0;
                    v1 = v1 * vo.s1;
                    v2 = v2 * vo.s2;
                }
                float* v = vo.vel + 3 * (line0 + i);
                v[0] = v0;
                v[1] = v1;
                v[2] = v2;
            }
        }
        __syncthreads();  // the lines' neighbours are read before the transform rewrites them
    }
    const int line = tid / T;
    fft_line<NX, E, -1>(LayContig{lds + line * LS}, tid % T, true, twl);
    for (int i = tid; i < 2 * NX * TY; i += NT) {
        const int ly = i % TY, kx = (i / TY) % NX, c = i / (TY * NX);
        S[(((size_t)c * NX + kx) * nz + z) * ny + y0 + ly] = lds[(c * TY + ly) * LS + pad16(kx)];
    }
}

// Inverse x transform from S back to psi lines; optionally normalize_psi
// (src/parallel.jl:162-165,184-185).  The 1/N of ifft is folded into the mask.  With
// oa != null the components go to two separate complex arrays oa, ob instead of psi.
// Without NORMALIZE, psin != null receives the normalized psi beside the raw one in psi.
template <int NX, int TY, bool NORMALIZE>
__global__ void __launch_bounds__(2 * TY * XCfg<NX>::T)
k_flow_x_inv(const float2* __restrict__ S, float4* __restrict__ psi, float2* __restrict__ oa, float2* __restrict__ ob,
             float4* __restrict__ psin, int ny, int nz, const float2* __restrict__ tw, Batch B) {
    constexpr int T = XCfg<NX>::T, E = XCfg<NX>::E, LS = XCfg<NX>::LS;
    const StTile tl = stencil_tile();  // x: y tile, y: z, z: member
    if (B.inst) {
        const Inst& m = B.inst[tl.bz];
        psi = m.psi;
        S = m.S;
        tw = m.tb.tw;
    }
    constexpr int NT = 2 * TY * T;
    __shared__ float2 lds[2 * TY * LS];
    __shared__ float2 twl[FftTw<NX, E>::SIZE];
    const int tid = threadIdx.x;
    const int y0 = tl.bx * TY, z = tl.by;
    constexpr int K = 2 * NX * TY / NT;
    float2 v[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int i = tid + k * NT, ly = i % TY, kx = (i / TY) % NX, c = i / (TY * NX);
        v[k] = S[(((size_t)c * NX + kx) * nz + z) * ny + y0 + ly];
    }
    FftTw<NX, E>::template fill<NT>(twl, tw, tid);
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int i = tid + k * NT, ly = i % TY, kx = (i / TY) % NX, c = i / (TY * NX);
        lds[(c * TY + ly) * LS + pad16(kx)] = v[k];
    }
    __syncthreads();
    const int line = tid / T;
    fft_line<NX, E, 1>(LayContig{lds + line * LS}, tid % T, true, twl);
    const size_t line0 = ((size_t)z * ny + y0) * NX;
    if (oa) {
        for (int i = tid; i < TY * NX; i += NT) {
            const int l = i / NX, x = i % NX;
            const float2 a = lds[l * LS + pad16(x)], b = lds[(TY + l) * LS + pad16(x)];
            float4 p = make_float4(a.x, a.y, b.x, b.y);
            if (NORMALIZE) p = normalized(p);
            oa[line0 + i] = make_float2(p.x, p.y);
            ob[line0 + i] = make_float2(p.z, p.w);
        }
    } else {
        float4* dst = psi + line0;
        for (int i = tid; i < TY * NX; i += NT) {
            const int l = i / NX, x = i % NX;
            const float2 a = lds[l * LS + pad16(x)], b = lds[(TY + l) * LS + pad16(x)];
            float4 p = make_float4(a.x, a.y, b.x, b.y);
            if (NORMALIZE) p = normalized(p);
            dst[i] = p;
            if (!NORMALIZE && psin) psin[line0 + i] = normalized(p);
        }
    }
}

// ======================================================================== plane pass

// Contiguous line in global memory.
struct LayGlobal {
    float2* p;
    __device__ __forceinline__ float2 ld(int e) const { return p[e]; }
    __device__ __forceinline__ void st(int e, float2 v) const { p[e] = v; }
};

// Every thread owns E elements of one y line during the y transforms and of one z line
// during the z transforms, so NT = NY * NZ / E threads: E = 16, or 8 when both sides are
// at most 64 (twice the wavefronts for the small planes, whose few workgroups are
// latency-bound); EO != 0 sets E (the batch autotuner may pick 16 for the small planes).  The first pass of each direction has radix R1 = E, the second
// N / R1.  With 1024 threads a thread has 128 VGPRs, too few for a second 16-element
// buffer, so the next-plane prefetch is used only when NT <= 512.
template <int NY, int NZ, int EO = 0> struct PlaneCfg {
    static constexpr int E = EO ? EO : (NY <= 64 && NZ <= 64) ? 8 : 16;
    static constexpr int R1 = E;
    static constexpr int TY = NY / E, TZ = NZ / E;   // threads per y line / z line
    static constexpr int NT = NZ * TY;
    static constexpr int RY2 = NY / R1, RZ2 = NZ / R1;  // second radix (1: single pass)
    static constexpr int LSY = padded_len(NY);
    static constexpr int LDS_BYTES = NZ * LSY * 8 + 16 * (NY + NZ);
    static constexpr bool PREFETCH = NT <= 512;
    static_assert(NY >= 16 && NZ >= 16 && NT <= 1024, "plane too large for one workgroup");
};

// PLANE_FLOW_ADJ: the adjoint of PLANE_FLOW, conj(mask) (the adjoint of F^-1 M F with an
// unnormalised forward transform F is F^-1 conj(M) F).
enum { PLANE_FLOW = 0, PLANE_POISSON = 1, PLANE_FLOW_ADJ = 2 };

// y and z transforms of one kx plane, spectral multiplier, inverse z and y; the block
// walks planes blockIdx.x, blockIdx.x + gridDim.x, ... and prefetches the next plane into
// registers while it transforms the current one.  Schedules: forward [R1, n/R1],
// inverse [n/R1, R1]; the multiplier and the first inverse z pass act on the registers of
// the last forward z pass (same elements), the first forward y pass reads global memory
// and the last inverse y pass writes it.  The threads of a y line are neighbouring lanes
// of one wavefront, so the exchanges between the two y passes of each direction stay
// inside the wavefront (wave_lds_sync).  The threads of a z column are NY lanes apart, so
// that a wavefront reads consecutive columns (no LDS bank conflicts); the z exchanges and
// the row / column switches take the workgroup barrier.  Giving each z column one
// wavefront (8 columns x 8 threads) would save those barriers too, but its strided column
// accesses conflict in the LDS banks: measured 59 us against 43 us for the flow plane.
//   PLANE_FLOW   : plane = c*nx + kx, multiplier mask_x[kx] mask_y[ky] mask_z[kz]
//                  (src/parallel.jl:58-62,182-183; 1/N folded into mask_x)
//   PLANE_POISSON: plane = kx in [0, nx/2], multiplier scale / (s2x+s2y+s2z), 0 at k = 0
//                  (src/parallel.jl:53-55,63,171; scale folds -0.25 and the 2/N of the
//                  half-length real inverse)
template <int NY, int NZ, int MODE, int EO>
__global__ void __launch_bounds__((PlaneCfg<NY, NZ, EO>::NT))
k_plane(float2* __restrict__ base, int nplanes, int nx, double scale, DevTables tb, Batch B) {
    using C = PlaneCfg<NY, NZ, EO>;
    if (B.inst) {
        const Inst& m = B.inst[blockIdx.y];
        base = MODE == PLANE_POISSON ? poisson_P(m) : m.S;
        scale = m.pscale;
        tb = m.tb;
    }
    constexpr int E = C::E, NT = C::NT, LSY = C::LSY;
    constexpr int RY2 = C::RY2, RZ2 = C::RZ2;
    constexpr int R1 = C::R1;
    using TY2 = PassTw<NY, RY2, R1>;   // forward y pass 2
    using TY2I = PassTw<NY, R1, RY2>;  // inverse y pass 2
    using TZ2 = PassTw<NZ, RZ2, R1>;
    using TZ2I = PassTw<NZ, R1, RZ2>;
    __shared__ float2 lds[NZ * LSY];
    __shared__ float2 twy[TY2::SIZE + TY2I::SIZE + 1];
    __shared__ float2 twz[TZ2::SIZE + TZ2I::SIZE + 1];
    __shared__ double mtab[NY + NZ];  // FLOW: float2 mask_y, mask_z ; POISSON: s2y, s2z
    const int tid = threadIdx.x;
    const int zl = tid / C::TY, ty = tid % C::TY;  // y transforms: line z = zl
    using Y1 = RPass<NY, R1, 1, E>;
    float2 v[E], pf[C::PREFETCH ? E : 1];
    // the first plane's load is issued before the table loads
    if (blockIdx.x < nplanes)
        Y1::load(C::PREFETCH ? pf : v, LayGlobal{base + (size_t)blockIdx.x * NY * NZ + zl * NY}, ty);
    TY2::template fill<NT>(twy, tb.tw + NY, tid);
    TY2I::template fill<NT>(twy + TY2::SIZE, tb.tw + NY, tid);
    TZ2::template fill<NT>(twz, tb.tw + NZ, tid);
    TZ2I::template fill<NT>(twz + TZ2::SIZE, tb.tw + NZ, tid);
    float2* mt2 = reinterpret_cast<float2*>(mtab);
    for (int i = tid; i < NY + NZ; i += NT) {
        if (MODE != PLANE_POISSON) mt2[i] = i < NY ? tb.mask_y[i] : tb.mask_z[i - NY];
        else mtab[i] = i < NY ? tb.s2y[i] : tb.s2z[i - NY];
    }
    const int yl = tid % NY, tz = tid / NY;        // z transforms: line y = yl
    __syncthreads();  // tables loaded
    const LayContig ly{lds + zl * LSY};
    const LayStrided lz{lds + pad16(yl), LSY};
    using Z1 = RPass<NZ, R1, 1, E>;

    int plane = blockIdx.x;
#pragma unroll 1
    for (; plane < nplanes; plane += gridDim.x) {
        float2* P = base + (size_t)plane * NY * NZ;
        if constexpr (C::PREFETCH) {
#pragma unroll
            for (int e = 0; e < E; e++) v[e] = pf[e];
            const int next = plane + gridDim.x;
            if (next < nplanes) Y1::load(pf, LayGlobal{base + (size_t)next * NY * NZ + zl * NY}, ty);
        } else if (plane != (int)blockIdx.x) {
            Y1::load(v, LayGlobal{P + zl * NY}, ty);
        }

        // ---- forward y (the first pass of each direction has no twiddles); the row's
        // previous readers (the last inverse y pass) are this wavefront
        Y1::template compute<-1>(v, ty, twy);
        wave_lds_sync();
        Y1::store(v, ly, ty);
        if constexpr (RY2 > 1) {
            using Y2 = RPass<NY, RY2, R1, E>;
            wave_lds_sync();
            Y2::load(v, ly, ty);
            Y2::template compute<-1>(v, ty, twy);  // TY2 table
            Y2::store(v, ly, ty);
        }
        __syncthreads();  // rows -> columns
        // ---- forward z, multiplier, inverse z
        Z1::load(v, lz, tz);
        Z1::template compute<-1>(v, tz, twz);
        using ZL = std::conditional_t<(RZ2 > 1), RPass<NZ, RZ2, R1, E>, Z1>;   // last forward pass
        using ZI = std::conditional_t<(RZ2 > 1), RPass<NZ, RZ2, 1, E>, Z1>;    // first inverse pass
        if constexpr (RZ2 > 1) {
            __syncthreads();
            Z1::store(v, lz, tz);
            __syncthreads();
            ZL::load(v, lz, tz);
            ZL::template compute<-1>(v, tz, twz);  // TZ2 table
        }
        constexpr int RL = E / ZL::NB;  // radix of the last forward z pass
        if (MODE != PLANE_POISSON) {
            const int kx = plane % nx;
            const float2 wxy = cmul(tb.mask_x[kx], mt2[yl]);
#pragma unroll
            for (int b = 0; b < ZL::NB; b++)
#pragma unroll
                for (int r = 0; r < RL; r++) {
                    const float2 m = cmul(wxy, mt2[NY + ZL::out(tz, b, r)]);
                    v[b * RL + r] = cmul(v[b * RL + r], MODE == PLANE_FLOW_ADJ ? cconj(m) : m);
                }
        } else {
            // eigenvalue sum in FP64 (the per-axis tables are FP64), rounded once to FP32;
            // scale = -0.5/N is a power of two, so scale * rcp(den) is within ~1.5 ulp of the
            // exact factor (v_rcp_f32 is accurate to 1 ulp)
            const int kx = plane;
            const double sxy = tb.s2x[kx] + mtab[yl];
            const float sc = (float)scale;
#pragma unroll
            for (int b = 0; b < ZL::NB; b++)
#pragma unroll
                for (int r = 0; r < RL; r++) {
                    const int kz = ZL::out(tz, b, r);
                    const float den = (float)(sxy + mtab[NY + kz]);
                    const float m = (kx == 0 && yl == 0 && kz == 0) ? 0.0f : sc * __builtin_amdgcn_rcpf(den);
                    v[b * RL + r] = cscale(v[b * RL + r], m);
                }
        }
        ZI::template compute<1>(v, tz, twz);
        __syncthreads();
        ZI::store(v, lz, tz);
        if constexpr (RZ2 > 1) {
            using Z2I = RPass<NZ, R1, RZ2, E>;
            __syncthreads();
            Z2I::load(v, lz, tz);
            Z2I::template compute<1>(v, tz, twz + TZ2::SIZE);
            __syncthreads();
            Z2I::store(v, lz, tz);
        }
        __syncthreads();  // columns -> rows
        // ---- inverse y; the last pass writes global memory
        if constexpr (RY2 > 1) {
            using Y1I = RPass<NY, RY2, 1, E>;
            using Y2I = RPass<NY, R1, RY2, E>;
            Y1I::load(v, ly, ty);
            Y1I::template compute<1>(v, ty, twy);
            Y1I::store(v, ly, ty);
            wave_lds_sync();
            Y2I::load(v, ly, ty);
            Y2I::template compute<1>(v, ty, twy + TY2::SIZE);
            Y2I::store(v, LayGlobal{P + zl * NY}, ty);
        } else {
            Y1::load(v, ly, ty);
            Y1::template compute<1>(v, ty, twy);
            Y1::store(v, LayGlobal{P + zl * NY}, ty);
        }
    }
}

// ======================================================================== Poisson x passes

// Real-to-complex x transform of TY lines of f (src/parallel.jl:169-170, x part):
// pack z[m] = f[2m] + i f[2m+1], length-nx/2 FFT, split into the half spectrum
// F[k] = E[k] + W^k O[k], k = 0..nx/2, written into the kx planes of P.
template <int NX, int TY>
__global__ void __launch_bounds__(TY * XCfg<NX / 2>::T)
k_r2c_x(const float* __restrict__ f, float2* __restrict__ P, int ny, int nz, const float2* __restrict__ tw, Batch B) {
    constexpr int M = NX / 2;
    const StTile tl = stencil_tile();  // x: y tile, y: z, z: member
    if (B.inst) {
        const Inst& m = B.inst[tl.bz];
        f = reinterpret_cast<const float*>(m.S);
        P = poisson_P(m);
        tw = m.tb.tw;
    }
    constexpr int T = XCfg<M>::T, E = XCfg<M>::E, LS = XCfg<M>::LS;
    constexpr int NT = TY * T;
    __shared__ float2 lds[TY * LS];
    __shared__ float2 twm[FftTw<M, E>::SIZE];
    __shared__ float2 twn[NX];
    const int tid = threadIdx.x;
    const int y0 = tl.bx * TY, z = tl.by;
    const float2* src = reinterpret_cast<const float2*>(f + ((size_t)z * ny + y0) * NX);
    constexpr int K = TY * M / NT;
    float2 v[K];
#pragma unroll
    for (int k = 0; k < K; k++) v[k] = src[tid + k * NT];  // issued before the table loads
    FftTw<M, E>::template fill<NT>(twm, tw, tid);
    load_twiddles<NX, NT>(twn, tw);
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int i = tid + k * NT;
        lds[(i / M) * LS + pad16(i % M)] = v[k];
    }
    __syncthreads();
    fft_line<M, E, -1>(LayContig{lds + (tid / T) * LS}, tid % T, true, twm);
    for (int i = tid; i < (M + 1) * TY; i += NT) {
        const int ly = i % TY, k = i / TY;
        const float2 zk = lds[ly * LS + pad16(k % M)];
        const float2 zm = cconj(lds[ly * LS + pad16((M - k) % M)]);
        const float2 e = cscale(cadd(zk, zm), 0.5f);
        const float2 dd = csub(zk, zm);
        const float2 o = make_float2(0.5f * dd.y, -0.5f * dd.x);  // (zk - zm) / (2i)
        P[((size_t)k * nz + z) * ny + y0 + ly] = cadd(e, cmul(twn[k], o));
    }
}

// ---- planes too large for one workgroup's LDS (ny or nz above 128)
// The plane pass is split into a y-line pass, a z-column pass that carries the spectral
// multiplier between its forward and inverse transforms, and an inverse y-line pass.
// Same multipliers and arithmetic as k_plane.

template <int NY> struct YLineCfg {
    static constexpr int E = DefaultE<NY>::value;
    static constexpr int T = NY / E;                   // threads per line
    static constexpr int R = T >= 256 ? 1 : 256 / T;   // lines (z rows) per block
    static constexpr int NT = R * T;
    static constexpr int LS = padded_len(NY);
};

// In-place y transforms (DIR -1 forward, +1 inverse) of R consecutive z rows of one plane.
template <int NY, int DIR>
__global__ void __launch_bounds__(YLineCfg<NY>::NT)
k_line_y(float2* __restrict__ base, int nz, const float2* __restrict__ tw) {
    using C = YLineCfg<NY>;
    __shared__ float2 lds[C::R * C::LS];
    __shared__ float2 twl[FftTw<NY, C::E>::SIZE];
    const int tid = threadIdx.x;
    float2* rows = base + ((size_t)blockIdx.y * nz + (size_t)blockIdx.x * C::R) * NY;
    stage<C::R * NY, C::NT>(
        tid, [&](int i) { return rows[i]; }, [&](int i, float2 v) { lds[(i / NY) * C::LS + pad16(i % NY)] = v; });
    FftTw<NY, C::E>::template fill<C::NT>(twl, tw, tid);
    __syncthreads();
    fft_line<NY, C::E, DIR>(LayContig{lds + (tid / C::T) * C::LS}, tid % C::T, true, twl);
    for (int i = tid; i < C::R * NY; i += C::NT) rows[i] = lds[(i / NY) * C::LS + pad16(i % NY)];
}

template <int NZ> struct ZColCfg {
    static constexpr int E = DefaultE<NZ>::value;
    static constexpr int T = NZ / E;                    // threads per column
    static constexpr int CW = T * 16 >= 256 ? 16 : 256 / T;  // columns per block (>= 16: 128 B rows)
    static constexpr int NT = CW * T;
    static constexpr int LS = padded_len(NZ);
};

// z transforms of CW consecutive y columns of one plane, the multiplier of k_plane, and the
// inverse z transforms.
template <int NZ, int MODE>
__global__ void __launch_bounds__(ZColCfg<NZ>::NT)
k_line_z(float2* __restrict__ base, int nx, int ny, double scale, DevTables tb) {
    using C = ZColCfg<NZ>;
    __shared__ float2 lds[C::CW * C::LS];
    __shared__ float2 twl[FftTw<NZ, C::E>::SIZE];
    const int tid = threadIdx.x;
    const int plane = blockIdx.y, y0 = blockIdx.x * C::CW;
    float2* P = base + (size_t)plane * ny * NZ;
    stage<C::CW * NZ, C::NT>(
        tid, [&](int i) { return P[(size_t)(i / C::CW) * ny + y0 + i % C::CW]; },
        [&](int i, float2 v) { lds[(i % C::CW) * C::LS + pad16(i / C::CW)] = v; });
    FftTw<NZ, C::E>::template fill<C::NT>(twl, tb.tw, tid);
    __syncthreads();
    const LayContig line{lds + (tid / C::T) * C::LS};
    fft_line<NZ, C::E, -1>(line, tid % C::T, true, twl);
    for (int i = tid; i < C::CW * NZ; i += C::NT) {
        const int c = i % C::CW, kz = i / C::CW, ky = y0 + c;
        float2& v = lds[c * C::LS + pad16(kz)];
        if (MODE != PLANE_POISSON) {
            const int kx = plane % nx;
            const float2 m = cmul(cmul(tb.mask_x[kx], tb.mask_y[ky]), tb.mask_z[kz]);
            v = cmul(v, MODE == PLANE_FLOW_ADJ ? cconj(m) : m);
        } else {
            const int kx = plane;
            const float den = (float)((tb.s2x[kx] + tb.s2y[ky]) + tb.s2z[kz]);
            const float m = (kx == 0 && ky == 0 && kz == 0) ? 0.0f : (float)scale * __builtin_amdgcn_rcpf(den);
            v = cscale(v, m);
        }
    }
    __syncthreads();
    fft_line<NZ, C::E, 1>(line, tid % C::T, true, twl);
    for (int i = tid; i < C::CW * NZ; i += C::NT) {
        const int c = i % C::CW, z = i / C::CW;
        P[(size_t)z * ny + y0 + c] = lds[c * C::LS + pad16(z)];
    }
}

// The same column transforms with the data in registers (columns of at most E^2 points,
// two radix passes per direction): the first forward pass reads global memory and the
// last inverse pass writes it, the multiplier and the first inverse pass act on the
// registers of the last forward pass, so each direction has one LDS exchange.  Same
// schedules and multipliers as the z passes of k_plane.
template <int NZ, int MODE>
__global__ void __launch_bounds__(ZColCfg<NZ>::NT)
k_line_z_reg(float2* __restrict__ base, int nx, int ny, double scale, DevTables tb) {
    using C = ZColCfg<NZ>;
    constexpr int E = C::E, CW = C::CW, LS = C::LS, NT = C::NT;
    constexpr int R1 = E, R2 = NZ / E;
    static_assert(R2 > 1 && R2 <= E, "two radix passes");
    using Z1 = RPass<NZ, R1, 1, E>;    // forward pass 1
    using ZL = RPass<NZ, R2, R1, E>;   // forward pass 2
    using ZI = RPass<NZ, R2, 1, E>;    // inverse pass 1 (reads ZL's outputs in place)
    using Z2I = RPass<NZ, R1, R2, E>;  // inverse pass 2
    using TF = PassTw<NZ, R2, R1>;
    using TI = PassTw<NZ, R1, R2>;
    __shared__ float2 lds[CW * LS];
    __shared__ float2 tw[TF::SIZE + TI::SIZE + 1];
    const int tid = threadIdx.x, c = tid % CW, t = tid / CW;
    const int plane = blockIdx.y, y0 = blockIdx.x * CW, ky = y0 + c;
    float2* P = base + (size_t)plane * ny * NZ;
    const LayStrided col{P + ky, ny};
    const LayContig lc{lds + c * LS};
    float2 v[E];
    Z1::load(v, col, t);  // issued before the table loads
    TF::template fill<NT>(tw, tb.tw + NZ, tid);
    TI::template fill<NT>(tw + TF::SIZE, tb.tw + NZ, tid);
    Z1::template compute<-1>(v, t, tw);
    Z1::store(v, lc, t);
    __syncthreads();  // column exchange, tables loaded
    ZL::load(v, lc, t);
    ZL::template compute<-1>(v, t, tw);
    constexpr int RL = E / ZL::NB;
#pragma unroll
    for (int b = 0; b < ZL::NB; b++)
#pragma unroll
        for (int r = 0; r < RL; r++) {
            const int kz = ZL::out(t, b, r);
            float2& x = v[b * RL + r];
            if (MODE != PLANE_POISSON) {
                const int kx = plane % nx;
                const float2 m = cmul(cmul(tb.mask_x[kx], tb.mask_y[ky]), tb.mask_z[kz]);
                x = cmul(x, MODE == PLANE_FLOW_ADJ ? cconj(m) : m);
            } else {
                const int kx = plane;
                const float den = (float)((tb.s2x[kx] + tb.s2y[ky]) + tb.s2z[kz]);
                const float m = (kx == 0 && ky == 0 && kz == 0) ? 0.0f : (float)scale * __builtin_amdgcn_rcpf(den);
                x = cscale(x, m);
            }
        }
    ZI::template compute<1>(v, t, tw);
    __syncthreads();  // the forward exchange's reads are done
    ZI::store(v, lc, t);
    __syncthreads();
    Z2I::load(v, lc, t);
    Z2I::template compute<1>(v, t, tw + TF::SIZE);
    Z2I::store(v, col, t);
}

enum { C2R_Q = 0, C2R_GAUGE = 1, C2R_GAUGE_RESTRICT = 2, C2R_GAUGE_AND_RESTRICTED = 3 };

// Complex-to-real inverse x transform of TY lines of P (src/parallel.jl:172) and, in the
// gauge modes, psi <- psi * exp(-i q) (src/parallel.jl:125-129,159) with the jet
// constraint applied afterwards in C2R_GAUGE_RESTRICT (src/particles.jl:12-27).
// Optional FP64 partial sums of |psi1|^2 + |psi2|^2 per block.
template <int NX, int TY, int MODE>
__global__ void __launch_bounds__(TY * C2rCfg<NX / 2>::T)
k_c2r_x(const float2* __restrict__ P, float* __restrict__ qout, float4* __restrict__ psi, GridInfo g, JetParams jet,
        double* __restrict__ partials, const float2* __restrict__ tw, Batch B, float4* __restrict__ rout) {
    constexpr int M = NX / 2;
    const StTile tl = stencil_tile();  // x: y tile, y: z, z: member
    if (B.inst) {
        const Inst& m = B.inst[tl.bz];
        P = poisson_P(m);
        psi = m.psi;
        g = m.g;
        jet = m.jet;
        jet.omgterm = m.omega * ((float)(*B.iter) * m.dt);  // test/runtests.jl:46,57
        partials = partials ? m.norm_parts : nullptr;
        tw = m.tb.tw;
    }
    constexpr int T = C2rCfg<M>::T, E = C2rCfg<M>::E;
    constexpr int LS = padded_len(M + 1);
    constexpr int NT = TY * T;
    __shared__ float2 lds[TY * LS];
    __shared__ float2 twm[FftTw<M, E>::SIZE];
    __shared__ float2 twn[NX];
    const int tid = threadIdx.x;
    const int ny = g.ny, nz = g.nz;
    const int y0 = tl.bx * TY, z = tl.by;
    {
        constexpr int CNT = (M + 1) * TY, K = (CNT + NT - 1) / NT;
        float2 v[K];
#pragma unroll
        for (int k = 0; k < K; k++) {  // issued before the table loads
            const int i = tid + k * NT, ly = i % TY, kk = i / TY;
            if (CNT % NT == 0 || i < CNT) v[k] = P[((size_t)kk * nz + z) * ny + y0 + ly];
        }
        FftTw<M, E>::template fill<NT>(twm, tw, tid);
        load_twiddles<NX, NT>(twn, tw);
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int i = tid + k * NT, ly = i % TY, kk = i / TY;
            if (CNT % NT == 0 || i < CNT) lds[ly * LS + pad16(kk)] = v[k];
        }
    }
    __syncthreads();
    for (int i = tid; i < (M / 2 + 1) * TY; i += NT) {
        const int ly = i % TY, k = i / TY;
        float2* L = lds + ly * LS;
        const float2 a = L[pad16(k)], b = L[pad16(M - k)];
        // Z[k] = (A + conj B)/2 + i (A - conj B) exp(+2 pi i k/nx) / 2
        {
            const float2 e = cscale(cadd(a, cconj(b)), 0.5f);
            const float2 o = cscale(cmul(csub(a, cconj(b)), cconj(twn[k])), 0.5f);
            L[pad16(k)] = make_float2(e.x - o.y, e.y + o.x);
        }
        if (k > 0 && k < M / 2) {
            const float2 e = cscale(cadd(b, cconj(a)), 0.5f);
            const float2 o = cscale(cmul(csub(b, cconj(a)), cconj(twn[M - k])), 0.5f);
            L[pad16(M - k)] = make_float2(e.x - o.y, e.y + o.x);
        }
    }
    __syncthreads();
    fft_line<M, E, 1>(LayContig{lds + (tid / T) * LS}, tid % T, true, twm);
    const size_t row0 = ((size_t)z * ny + y0) * NX;
    double acc = 0.0;
    // psi is read in groups of G cells per thread, the group's loads issued together
    constexpr int CNT = TY * NX, G = 8;
    static_assert(CNT % (NT * G) == 0, "c2r output tiling");
#pragma unroll 1
    for (int i0 = tid; i0 < CNT; i0 += NT * G) {
        float4 pg[G];
        if (MODE != C2R_Q) {
#pragma unroll
            for (int gi = 0; gi < G; gi++) pg[gi] = psi[row0 + i0 + gi * NT];
        }
#pragma unroll
        for (int gi = 0; gi < G; gi++) {
            const int i = i0 + gi * NT;
            const int l = i / NX, x = i % NX;
            const float2 zz = lds[l * LS + pad16(x >> 1)];
            const float q = (x & 1) ? zz.y : zz.x;
            if (MODE == C2R_Q) {
                qout[row0 + i] = q;
            } else {
                float4 p = pg[gi];
                float s, c;
                sincosf(q, &s, &c);  // exp(-i q) = (c, -s)
                if (qout) qout[row0 + i] = q;  // kept for the adjoint
                p = make_float4(p.x * c + p.y * s, p.y * c - p.x * s, p.z * c + p.w * s, p.w * c - p.z * s);
                if (MODE == C2R_GAUGE_RESTRICT && is_jet_cell(x, y0 + l, z, g, jet)) p = jet_phase(p, x, y0 + l, z, g, jet);
                if (partials) acc += (double)p.x * p.x + (double)p.y * p.y + (double)p.z * p.z + (double)p.w * p.w;
                psi[row0 + i] = p;
                if (MODE == C2R_GAUGE_AND_RESTRICTED)  // the gauge output in psi, its restriction in rout
                    rout[row0 + i] = is_jet_cell(x, y0 + l, z, g, jet) ? jet_phase(p, x, y0 + l, z, g, jet) : p;
            }
        }
    }
    if (MODE != C2R_Q && partials) {
        const double s = block_sum<NT>(acc);
        if (tid == 0) partials[tl.by * gridDim.x + tl.bx] = s;  // the tile within its member
    }
}

// ======================================================================== stencils


// ---- z-marching stencil tiles
// A block of BX x 4 threads covers BX consecutive x and 4 consecutive y and walks tz
// consecutive z (tz = 16, or fewer on small grids so that a launch has enough blocks).  psi(c + e_z) of one z step is psi(c) of the next, psi(c + e_x) comes
// from the neighbouring lane and psi(c + e_y) from the row above through LDS, so every
// psi row is fetched once per block (plus one halo row on each side in y and z) and
// every forward-edge angle
// v_a(c) = angle(conj psi1(c) psi1(c+e_a) + conj psi2(c) psi2(c+e_a)) (src/parallel.jl:100-113)
// is evaluated once per cell (plus a thin halo).  Blocks are mapped XCD-aware
// (stencil_tile), so the halo rows shared by neighbouring tiles stay in one L2.
constexpr int ST_BY = 4, ST_TZ = 16;  // rows per block, longest z walk

__device__ __forceinline__ float4 shfl4(float4 v, int src) {
    return make_float4(__shfl(v.x, src, 64), __shfl(v.y, src, 64), __shfl(v.z, src, 64), __shfl(v.w, src, 64));
}

// Fused velocity_one_form (hbar = 1) + div (src/parallel.jl:100-113,131-143,156-157):
// f(c) = sum_a (v_a(c) - v_a(c - e_a)) / d_a^2.
template <int BX>
__global__ void __launch_bounds__(BX * ST_BY)
k_div_psi(const float4* __restrict__ psi, float* __restrict__ f, GridInfo g, int tz, Batch B) {
    __shared__ float4 row_s[2][ST_BY][BX];
    __shared__ float vy_s[2][ST_BY][BX];
    const StTile tl = stencil_tile();
    int zslab = tl.bz;
    if (B.inst) {
        const int nzs = g.nz / tz;
        const Inst& m = B.inst[tl.bz / nzs];
        zslab = tl.bz % nzs;
        psi = m.psi;
        f = reinterpret_cast<float*>(m.S);
        g = m.g;
    }
    const int tx = threadIdx.x % BX, ty = threadIdx.x / BX;
    const int lane = threadIdx.x % 64;
    const int x = tl.bx * BX + tx, y = tl.by * ST_BY + ty, z0 = zslab * tz;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const size_t sy = nx, sz = (size_t)nx * ny;
    const int xp = x + 1 < nx ? x + 1 : 0, xm = x > 0 ? x - 1 : nx - 1;
    const int yp = y + 1 < ny ? y + 1 : 0, ym = y > 0 ? y - 1 : ny - 1;
    const bool top = ty == ST_BY - 1, bottom = ty == 0;
    const float ds0 = 1.0f / (g.dx * g.dx), ds1 = 1.0f / (g.dy * g.dy), ds2 = 1.0f / (g.dz * g.dz);
    const int zb = z0 > 0 ? z0 - 1 : nz - 1;
    const size_t cx = x + sy * y;
    // halo rows: y + 1 above the tile (top row), y - 1 below it (bottom row)
    const size_t chalo = x + sy * (top ? yp : ym);
    float4 pc = psi[cx + sz * z0];
    float vzm = edge_angle(psi[cx + sz * zb], pc);  // v_z(c - e_z) for the first z
    // loads of the next z step are issued one step ahead
    int zp = z0 + 1 < nz ? z0 + 1 : 0;
    float4 pz = psi[cx + sz * zp];
    float4 ph = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (top | bottom) ph = psi[chalo + sz * z0];
#pragma unroll 1
    for (int k = 0; k < tz; k++) {
        const int z = z0 + k;
        const size_t row = sy * y + sz * z;
        float4 pz_n = pz, ph_n = ph;
        if (k + 1 < tz) {
            const int zn = zp + 1 < nz ? zp + 1 : 0;
            pz_n = psi[cx + sz * zn];
            if (top | bottom) ph_n = psi[chalo + sz * zp];
            zp = zn;
        }
        row_s[k & 1][ty][tx] = pc;
        float4 px = shfl4(pc, tx + 1 < BX ? lane + 1 : lane);
        if (tx + 1 == BX) px = psi[xp + row];
        const float vx = edge_angle(pc, px), vz = edge_angle(pc, pz);
        float vxm = __shfl(vx, tx > 0 ? lane - 1 : lane, 64);
        if (tx == 0) vxm = edge_angle(psi[xm + row], pc);
        __syncthreads();
        float4 py = row_s[k & 1][top ? ty : ty + 1][tx];
        if (top) py = ph;
        const float vy = edge_angle(pc, py);
        vy_s[k & 1][ty][tx] = vy;
        __syncthreads();
        const float vym = bottom ? edge_angle(ph, pc) : vy_s[k & 1][ty - (bottom ? 0 : 1)][tx];
        f[x + row] = (vx - vxm) * ds0 + (vy - vym) * ds1 + (vz - vzm) * ds2;
        vzm = vz;
        pc = pz;
        pz = pz_n;
        ph = ph_n;
    }
}

// velocity_one_form! with hbar (src/parallel.jl:115-122) and, if sharp, the driver's
// in-place StaggeredSharp v_a *= 1/d_a (test/runtests.jl:61-62).
template <int BX>
__global__ void __launch_bounds__(BX * ST_BY)
k_velocity(const float4* __restrict__ psi, float* __restrict__ vel, GridInfo g, float hbar, int sharp, int tz,
           Batch B) {
    __shared__ float4 row_s[2][ST_BY][BX];
    const StTile tl = stencil_tile();
    int zslab = tl.bz;
    if (B.inst) {
        const int nzs = g.nz / tz;
        const Inst& m = B.inst[tl.bz / nzs];
        zslab = tl.bz % nzs;
        psi = m.psi;
        vel = m.vel;
        g = m.g;
        hbar = m.hbar;
    }
    const int tx = threadIdx.x % BX, ty = threadIdx.x / BX;
    const int lane = threadIdx.x % 64;
    const int x = tl.bx * BX + tx, y = tl.by * ST_BY + ty, z0 = zslab * tz;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const size_t sy = nx, sz = (size_t)nx * ny;
    const int xp = x + 1 < nx ? x + 1 : 0, yp = y + 1 < ny ? y + 1 : 0;
    const bool top = ty == ST_BY - 1;
    const float s0 = sharp ? 1.0f / g.dx : 1.0f, s1 = sharp ? 1.0f / g.dy : 1.0f, s2 = sharp ? 1.0f / g.dz : 1.0f;
    const size_t cx = x + sy * y, cyp = x + sy * yp;
    float4 pc = psi[cx + sz * z0];
    int zp = z0 + 1 < nz ? z0 + 1 : 0;
    float4 pz = psi[cx + sz * zp];
    float4 ph = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (top) ph = psi[cyp + sz * z0];
#pragma unroll 1
    for (int k = 0; k < tz; k++) {
        const int z = z0 + k;
        const size_t c = cx + sz * z;
        float4 pz_n = pz, ph_n = ph;
        if (k + 1 < tz) {
            const int zn = zp + 1 < nz ? zp + 1 : 0;
            pz_n = psi[cx + sz * zn];
            if (top) ph_n = psi[cyp + sz * zp];
            zp = zn;
        }
        row_s[k & 1][ty][tx] = pc;
        float4 px = shfl4(pc, tx + 1 < BX ? lane + 1 : lane);
        if (tx + 1 == BX) px = psi[xp + sy * y + sz * z];
        __syncthreads();
        float4 py = row_s[k & 1][top ? ty : ty + 1][tx];
        if (top) py = ph;
        float v0 = edge_angle(pc, px) * hbar, v1 = edge_angle(pc, py) * hbar, v2 = edge_angle(pc, pz) * hbar;
        if (sharp) {
            v0 = v0 * s0;
            v1 = v1 * s1;
            v2 = v2 * s2;
        }
        vel[3 * c + 0] = v0;
        vel[3 * c + 1] = v1;
        vel[3 * c + 2] = v2;
        pc = pz;
        pz = pz_n;
        ph = ph_n;
    }
}

// ---- fused divergence + R2C x pass

// Threads of the fused pass: 8 cells of the TY x NX tile each, at least one wavefront and
// at least the TY * T threads of the R2C transform.
template <int NX, int TY> struct DivR2cCfg {
    static constexpr int M = NX / 2;
    static constexpr int TF = TY * XCfg<M>::T;
    static constexpr int NC = TY * NX / 8;
    static constexpr int NT0 = NC > TF ? NC : TF;
    static constexpr int NT = NT0 > 64 ? NT0 : 64;
};

// pressure_project!'s f = div(velocity_one_form(psi, 1)) (src/parallel.jl:100-113,131-143,
// 156-157) fused with the real-to-complex x transform of k_r2c_x, so f never goes to
// memory.  A block computes f for TY x lines: TY/2 rows y0 .. y0+TY/2-1 at z0 and the same
// rows at z0 + 1 (z0 even).  A thread takes both cells of a (y, x) column, so the z edge
// between them, v_z at z0, is evaluated once for both (3.5 instead of 4 edge angles per
// cell), and the block reads one z halo row per column on each side.  psi(c + e_x) comes
// from the neighbouring lane, psi(c + e_y), the z halo rows and the row below each z plane
// of the tile through the caches.  The edge angles and their differences are those of
// k_div_psi (same operands).  Tiles are mapped XCD-aware like the stencils, so the halo
// rows a block reads are its neighbours' own rows, in the same L2.
template <int NX, int TY>
__global__ void __launch_bounds__((DivR2cCfg<NX, TY>::NT))
k_div_r2c(const float4* __restrict__ psi, float2* __restrict__ P, GridInfo g, const float2* __restrict__ tw,
          Batch B) {
    using C = DivR2cCfg<NX, TY>;
    constexpr int M = NX / 2, NT = C::NT, CELLS = TY * NX;
    constexpr int HT = TY / 2, COLS = HT * NX;  // rows per z plane, (row, x) columns
    constexpr int T = XCfg<M>::T, E = XCfg<M>::E, LS = XCfg<M>::LS;
    static_assert(CELLS % NT == 0 && COLS % NT == 0 && NT % 64 == 0, "div r2c tiling");
    const StTile tl = stencil_tile();  // x: y tile, y: z pair, z: member
    if (B.inst) {
        const Inst& m = B.inst[tl.bz];
        psi = m.psi;
        P = poisson_P(m);
        g = m.g;
        tw = m.tb.tw;
    }
    // STREAM (rows of 64 and more, so that every wavefront starts a row segment): f of a
    // column's two cells right after its edge angles (one block barrier per column makes
    // the row below visible), so neither v_x nor the z differences are staged and the
    // block needs 38 instead of 54 KB of LDS at 256 x 16 lines (4 blocks per CU instead
    // of 3); the 48-jet ensemble step 0.585 -> 0.579 ms with it at 64 and 128 wide
    constexpr bool STREAM = NX >= 64;
    constexpr int NSEG = NX / 64;  // 64-cell row segments (STREAM)
    __shared__ float vx_s[STREAM ? 1 : TY][NX];
    // STREAM: v_x of each row segment's last cell, the x - 1 neighbour of the next
    // segment's first cell
    __shared__ float vxb[STREAM ? TY * NSEG : 1];
    // z plane h, row ly: v_y at y0 + ly in row h (HT + 1) + ly + 1; at y0 - 1 in row h (HT + 1)
    __shared__ float vy_s[TY + 2][NX];
    __shared__ float2 lds[TY * LS];     // v_z differences, then f as packed complex lines
    __shared__ float2 twm[FftTw<M, E>::SIZE];
    float* dz_s = reinterpret_cast<float*>(lds);
    const int tid = threadIdx.x, lane = tid & 63;
    const int ny = g.ny, nz = g.nz;
    const int y0 = tl.bx * HT, z0 = 2 * tl.by;
    const size_t sy = NX, sz = (size_t)NX * ny;
    const size_t zo0 = sz * z0, zo1 = sz * (z0 + 1);
    const size_t zom = sz * (z0 > 0 ? z0 - 1 : nz - 1), zop = sz * (z0 + 2 < nz ? z0 + 2 : 0);
    FftTw<M, E>::template fill<NT>(twm, tw, tid);
    // v_y of the row below the tile, in both z planes
    {
        const int ym = y0 > 0 ? y0 - 1 : ny - 1;
        for (int i = tid; i < 2 * NX; i += NT) {
            const int h = i / NX, x = i % NX;
            const size_t zo = h ? zo1 : zo0;
            vy_s[h * (HT + 1)][x] = edge_angle(psi[sy * ym + zo + x], psi[sy * y0 + zo + x]);
        }
    }
    // edge angles of the tile's columns; a wavefront holds consecutive x of one row.
    // STREAM takes one column per thread at a time (57 VGPRs, 8 wavefronts per SIMD at
    // 256 x 16 lines; two columns' loads in flight needed 86 VGPRs, 4 wavefronts per SIMD:
    // 0.2626 -> 0.2598 ms per step at 256x128x128 for one column with v_x and the z
    // differences staged, 0.2738 -> 0.2688 on a slower box for STREAM).  Rows of 16 and 32
    // take G = 2 columns with their loads issued together.
    constexpr int K = COLS / NT, G = K < 2 ? K : 2;
    const float ds0 = 1.0f / (g.dx * g.dx), ds1 = 1.0f / (g.dy * g.dy), ds2 = 1.0f / (g.dz * g.dz);
    if constexpr (STREAM) {
#pragma unroll 1
        for (int k = 0; k < K; k++) {
            const int j = tid + k * NT, ly = j / NX, x = j % NX, y = y0 + ly;
            const size_t row = sy * y, rowy = sy * (y + 1 < ny ? y + 1 : 0);
            const float4 p0 = psi[row + zo0 + x], p1 = psi[row + zo1 + x];
            const float4 pm = psi[row + zom + x], pp = psi[row + zop + x];
            const float4 py0 = psi[rowy + zo0 + x], py1 = psi[rowy + zo1 + x];
            float4 ph0 = p0, ph1 = p1;
            if (lane == 63) {  // x + 1 is the next segment's first cell (or the row's first)
                const int xn = x + 1 < NX ? x + 1 : 0;
                ph0 = psi[row + zo0 + xn];
                ph1 = psi[row + zo1 + xn];
            }
            const int src = lane < 63 ? lane + 1 : lane;
            float4 px0 = shfl4(p0, src), px1 = shfl4(p1, src);
            if (lane == 63) {
                px0 = ph0;
                px1 = ph1;
            }
            const float vx0 = edge_angle(p0, px0), vx1 = edge_angle(p1, px1);
            vy_s[ly + 1][x] = edge_angle(p0, py0);
            vy_s[HT + 2 + ly][x] = edge_angle(p1, py1);
            const float ea = edge_angle(pm, p0), eb = edge_angle(p0, p1), ec = edge_angle(p1, pp);
            if (lane == 63) {
                vxb[ly * NSEG + x / 64] = vx0;
                vxb[(HT + ly) * NSEG + x / 64] = vx1;
            }
            // v_x(c - e_x): the lane to the left, or for a segment's first cell the previous
            // segment's last (the row's last cell for x = 0)
            float vxm0 = __shfl(vx0, lane > 0 ? lane - 1 : 0, 64), vxm1 = __shfl(vx1, lane > 0 ? lane - 1 : 0, 64);
            __syncthreads();  // the rows below and the segment ends of this column are written
            if (lane == 0) {
                const int sp = (x / 64 + NSEG - 1) % NSEG;
                vxm0 = vxb[ly * NSEG + sp];
                vxm1 = vxb[(HT + ly) * NSEG + sp];
            }
            const float f0 = (vx0 - vxm0) * ds0 + (vy_s[ly + 1][x] - vy_s[ly][x]) * ds1 + (eb - ea) * ds2;
            const float f1 =
                (vx1 - vxm1) * ds0 + (vy_s[HT + 2 + ly][x] - vy_s[HT + 1 + ly][x]) * ds1 + (ec - eb) * ds2;
            reinterpret_cast<float*>(lds + ly * LS + pad16(x >> 1))[x & 1] = f0;
            reinterpret_cast<float*>(lds + (HT + ly) * LS + pad16(x >> 1))[x & 1] = f1;
        }
    } else {
#pragma unroll 1
        for (int k0 = 0; k0 < K; k0 += G) {
            float4 p0[G], p1[G], pm[G], pp[G], py0[G], py1[G], ph0[G], ph1[G];
#pragma unroll
            for (int q = 0; q < G; q++) {
                const int j = tid + (k0 + q) * NT, ly = j / NX, x = j % NX, y = y0 + ly;
                const size_t row = sy * y, rowy = sy * (y + 1 < ny ? y + 1 : 0);
                p0[q] = psi[row + zo0 + x];
                p1[q] = psi[row + zo1 + x];
                pm[q] = psi[row + zom + x];
                pp[q] = psi[row + zop + x];
                py0[q] = psi[rowy + zo0 + x];
                py1[q] = psi[rowy + zo1 + x];
                ph0[q] = p0[q];
                ph1[q] = p1[q];
                if (lane == 63 || x + 1 == NX) {
                    const int xn = x + 1 < NX ? x + 1 : 0;
                    ph0[q] = psi[row + zo0 + xn];
                    ph1[q] = psi[row + zo1 + xn];
                }
            }
#pragma unroll
            for (int q = 0; q < G; q++) {
                const int j = tid + (k0 + q) * NT, ly = j / NX, x = j % NX;
                const int src = lane < 63 ? lane + 1 : lane;
                float4 px0 = shfl4(p0[q], src), px1 = shfl4(p1[q], src);
                if (lane == 63 || x + 1 == NX) {
                    px0 = ph0[q];
                    px1 = ph1[q];
                }
                vx_s[ly][x] = edge_angle(p0[q], px0);
                vx_s[HT + ly][x] = edge_angle(p1[q], px1);
                vy_s[ly + 1][x] = edge_angle(p0[q], py0[q]);
                vy_s[HT + 2 + ly][x] = edge_angle(p1[q], py1[q]);
                const float ea = edge_angle(pm[q], p0[q]), eb = edge_angle(p0[q], p1[q]), ec = edge_angle(p1[q], pp[q]);
                dz_s[j] = eb - ea;
                dz_s[COLS + j] = ec - eb;
            }
        }
        __syncthreads();
        // f(c) = sum_a (v_a(c) - v_a(c - e_a)) / d_a^2
        constexpr int KC = CELLS / NT;
        float fv[KC];
#pragma unroll
        for (int k = 0; k < KC; k++) {
            const int i = tid + k * NT, l = i / NX, x = i % NX;
            const float vx = vx_s[l][x], vxm = vx_s[l][x > 0 ? x - 1 : NX - 1];
            const int ry = l + l / HT + 1;
            const float vy = vy_s[ry][x], vym = vy_s[ry - 1][x];
            fv[k] = (vx - vxm) * ds0 + (vy - vym) * ds1 + dz_s[i] * ds2;
        }
        __syncthreads();  // dz_s is read; lds now receives f
#pragma unroll
        for (int k = 0; k < KC; k++) {
            const int i = tid + k * NT, l = i / NX, x = i % NX;
            reinterpret_cast<float*>(lds + l * LS + pad16(x >> 1))[x & 1] = fv[k];
        }
    }
    __syncthreads();
    // the R2C transform of k_r2c_x; line l is row y0 + l % HT of z plane z0 + l / HT
    fft_line<M, E, -1>(LayContig{lds + (tid / T) * LS}, tid % T, tid < TY * T, twm);
    for (int i = tid; i < (M + 1) * TY; i += NT) {
        const int lh = i % HT, k = (i / HT) % (M + 1), h = i / (HT * (M + 1));
        const int ly = h * HT + lh;
        const float2 zk = lds[ly * LS + pad16(k % M)];
        const float2 zm = cconj(lds[ly * LS + pad16((M - k) % M)]);
        const float2 e = cscale(cadd(zk, zm), 0.5f);
        const float2 dd = csub(zk, zm);
        const float2 o = make_float2(0.5f * dd.y, -0.5f * dd.x);  // (zk - zm) / (2i)
        P[((size_t)k * nz + z0 + h) * ny + y0 + lh] = cadd(e, cmul(tw[NX + k], o));
    }
}

// div! from a velocity field (src/parallel.jl:131-152).
__global__ void __launch_bounds__(256) k_div_vel(const float* __restrict__ vel, float* __restrict__ f, GridInfo g) {
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t N = (size_t)g.nx * g.ny * g.nz;
    if (c >= N) return;
    int x, y, z;
    cell_xyz(c, g, x, y, z);
    const size_t sy = g.nx, sz = (size_t)g.nx * g.ny;
    const size_t xm = x > 0 ? c - 1 : c + (size_t)(g.nx - 1);
    const size_t ym = y > 0 ? c - sy : c + (size_t)(g.ny - 1) * sy;
    const size_t zm = z > 0 ? c - sz : c + (size_t)(g.nz - 1) * sz;
    const float ds0 = 1.0f / (g.dx * g.dx), ds1 = 1.0f / (g.dy * g.dy), ds2 = 1.0f / (g.dz * g.dz);
    f[c] = (vel[3 * c] - vel[3 * xm]) * ds0 + (vel[3 * c + 1] - vel[3 * ym + 1]) * ds1 +
           (vel[3 * c + 2] - vel[3 * zm + 2]) * ds2;
}

// ======================================================================== pointwise

// StaggeredSharp as the driver applies it in place, velocity .* inv.(isf.d)
// (test/runtests.jl:61-62; src/MatlabPort.jl:99-104)
__global__ void __launch_bounds__(256) k_sharp(float* __restrict__ vel, float s0, float s1, float s2, size_t N) {
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    vel[3 * c + 0] *= s0;
    vel[3 * c + 1] *= s1;
    vel[3 * c + 2] *= s2;
}

__global__ void __launch_bounds__(256) k_normalize(float4* __restrict__ psi, size_t N) {
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    psi[c] = normalized(psi[c]);
}

__global__ void __launch_bounds__(256) k_gauge(float4* __restrict__ psi, const float* __restrict__ q, size_t N) {
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    float4 p = psi[c];
    float s, co;
    sincosf(q[c], &s, &co);
    psi[c] = make_float4(p.x * co + p.y * s, p.y * co - p.x * s, p.z * co + p.w * s, p.w * co - p.z * s);
}

__global__ void __launch_bounds__(256) k_isjet(unsigned char* __restrict__ out, GridInfo g, JetParams j) {
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t N = (size_t)g.nx * g.ny * g.nz;
    if (c >= N) return;
    int x, y, z;
    cell_xyz(c, g, x, y, z);
    out[c] = is_jet_cell(x, y, z, g, j) ? 1 : 0;
}

// restrict_velocity! with an explicit isjet array (src/particles.jl:25-27), or the
// analytic predicate when isjet is null.
__global__ void __launch_bounds__(256)
k_restrict(float4* __restrict__ psi, const unsigned char* __restrict__ isjet, GridInfo g, JetParams j) {
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t N = (size_t)g.nx * g.ny * g.nz;
    if (c >= N) return;
    int x, y, z;
    cell_xyz(c, g, x, y, z);
    const bool in = isjet ? (isjet[c] != 0) : is_jet_cell(x, y, z, g, j);
    if (in) psi[c] = jet_phase(psi[c], x, y, z, g, j);
}

template <int NT>
__global__ void __launch_bounds__(NT) k_norm2(const float4* __restrict__ psi, size_t N, double* __restrict__ partials) {
    double acc = 0.0;
    for (size_t c = (size_t)blockIdx.x * NT + threadIdx.x; c < N; c += (size_t)gridDim.x * NT) {
        const float4 p = psi[c];
        acc += (double)p.x * p.x + (double)p.y * p.y + (double)p.z * p.z + (double)p.w * p.w;
    }
    const double s = block_sum<NT>(acc);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

__global__ void __launch_bounds__(256) k_reduce_sum(const double* __restrict__ partials, size_t n, double* __restrict__ out) {
    double acc = 0.0;
    for (size_t i = threadIdx.x; i < n; i += 256) acc += partials[i];
    const double s = block_sum<256>(acc);
    if (threadIdx.x == 0) out[0] = s;
}

// ======================================================================== particles

// staggered_advect! (src/parallel.jl:190-216,248-261): RK4 for slots strictly inside
// the box, the dead marker (-1,-1,-1) otherwise.  Blocks walk ranges of 2 blockDim.x
// slots (grid-stride); each thread advances two slots of a range, k and k + blockDim.x
// (order == nullptr: slot k; else slot order[k]), one RK stage of both at a time: the two
// slots' gathers are independent, so each stage waits for memory once for both.  Every slot is independent, so neither
// the order nor the pairing changes a result.  A slot outside the box is stepped at a
// stand-in position and only its marker is written.
constexpr int ADVECT_SLOTS_PER_THREAD = 2;
__global__ void __launch_bounds__(256)
k_advect(float* __restrict__ parts, long long nslots, const int* __restrict__ order, const float* __restrict__ vel,
         GridInfo g, float dt, Batch B, const float* src) {
    if (B.inst) {
        const Inst& m = B.inst[blockIdx.y];
        parts = m.parts;
        nslots = m.nslots;
        order = nullptr;
        vel = (B.vel_by_step && ((*B.iter - 1) & 1)) ? m.vel_alt : m.vel;
        g = m.g;
        dt = m.dt;
        src = nullptr;
    }
    constexpr int Q = ADVECT_SLOTS_PER_THREAD;
    for (long long base = (long long)blockIdx.x * (Q * blockDim.x); base < nslots;
         base += (long long)gridDim.x * (Q * blockDim.x)) {
        const long long k0 = base + threadIdx.x;
        long long s[Q];
        bool act[Q], in[Q];
        float p[Q][3];
#pragma unroll
        for (int j = 0; j < Q; j++) {
            const long long k = k0 + (long long)j * blockDim.x;
            act[j] = k < nslots;
            s[j] = act[j] ? (order ? (long long)order[k] : k) : 0;
        }
#pragma unroll
        for (int j = 0; j < Q; j++) {
            p[j][0] = p[j][1] = p[j][2] = -1.0f;
            if (act[j]) {
                const F3 p3 = ld3((src ? src : parts) + 3 * s[j]);  // src: out of place, from a separate array
                p[j][0] = p3.x;
                p[j][1] = p3.y;
                p[j][2] = p3.z;
            }
            in[j] = act[j] & (p[j][0] > 0.0f) & (p[j][0] < g.Lx) & (p[j][1] > 0.0f) & (p[j][1] < g.Ly) &
                    (p[j][2] > 0.0f) & (p[j][2] < g.Lz);
            if (!in[j]) {
                p[j][0] = 0.5f * g.Lx;
                p[j][1] = 0.5f * g.Ly;
                p[j][2] = 0.5f * g.Lz;
            }
        }
        float k1[Q][3], k2[Q][3], k3[Q][3], k4[Q][3], t[Q][3];
#pragma unroll
        for (int j = 0; j < Q; j++) stag_vel(vel, p[j], g, k1[j]);
#pragma unroll
        for (int j = 0; j < Q; j++)
#pragma unroll
            for (int a = 0; a < 3; a++) t[j][a] = p[j][a] + k1[j][a] * dt * 0.5f;
#pragma unroll
        for (int j = 0; j < Q; j++) stag_vel(vel, t[j], g, k2[j]);
#pragma unroll
        for (int j = 0; j < Q; j++)
#pragma unroll
            for (int a = 0; a < 3; a++) t[j][a] = p[j][a] + k2[j][a] * dt * 0.5f;
#pragma unroll
        for (int j = 0; j < Q; j++) stag_vel(vel, t[j], g, k3[j]);
#pragma unroll
        for (int j = 0; j < Q; j++)
#pragma unroll
            for (int a = 0; a < 3; a++) t[j][a] = p[j][a] + k3[j][a] * dt;
#pragma unroll
        for (int j = 0; j < Q; j++) stag_vel(vel, t[j], g, k4[j]);
#pragma unroll
        for (int j = 0; j < Q; j++) {
            if (!act[j]) continue;
#pragma unroll
            for (int a = 0; a < 3; a++)
                parts[3 * s[j] + a] =
                    in[j] ? p[j][a] + dt / 6.0f * (((k1[j][a] + 2.0f * k2[j][a]) + 2.0f * k3[j][a]) + k4[j][a]) : -1.0f;
        }

    }
}

// add_particles! (src/particles.jl:34-54,75-94): n particles at
// (c_x - 0.1, c_y + 0.9 rad cos(rt), c_z + 0.9 rad sin(rt)) + u*0.1 - 0.05, rt = 2 pi u0.
// Draw j of the batch uses Philox4x32-10 with counter (j, step, 0, 0).  Each chunk of
// 1024 consecutive draws is stored in the chunk's slots first + chunk*1024 + k (mod
// nslots) in increasing order of key = (top 22 bits of raw word 0, i.e. the ring angle)
// << 10 | (j mod 1024), so neighbouring slots hold neighbouring points of the ring.  A
// block sorts blockDim.x keys (a power of two from 64 to 1024, at least the draw count
// when that is below 1024), padded with keys that sort last: with fewer than 1024 draws
// the order is the same as with 1024-key blocks, and at most 64 draws sort within one
// wavefront, without LDS stages.
constexpr int SPAWN_CHUNK = 1024;

__device__ __forceinline__ void spawn_draw(uint32_t j, unsigned step, unsigned long long seed, uint32_t raw[4]) {
    raw[0] = j;
    raw[1] = step;
    raw[2] = 0u;
    raw[3] = 0u;
    philox4x32_10(raw, (uint32_t)seed, (uint32_t)(seed >> 32));
}

__global__ void __launch_bounds__(SPAWN_CHUNK)
k_spawn(float* __restrict__ parts, long long nslots, long long first, long long n, JetParams jt,
        unsigned long long seed, unsigned step, Batch B) {
    __shared__ uint32_t key[SPAWN_CHUNK];
    if (B.inst) {
        // the ring-buffer slots of test/runtests.jl:52-53 for step *iter
        const Inst& m = B.inst[blockIdx.y];
        const int it = *B.iter;
        parts = m.parts;
        nslots = m.nslots;
        n = nslots > 0 ? m.n_spawn : 0;
        jt = m.jet;
        seed = m.seed;
        step = (unsigned)it;
        const long long start1 = nslots > 0 ? ((long long)(it - 1) * m.n_spawn + 1) % nslots : 0;
        first = start1 == 0 ? nslots - 1 : start1 - 1;
    }
    const int t = threadIdx.x, C = blockDim.x;
    const long long base = (long long)blockIdx.x * C;
    if (base >= n) return;  // uniform per block (batch members with fewer draws)
    const int m = (int)min((long long)C, n - base);
    uint32_t raw[4];
    uint32_t kv = 0xFFFFFFFFu;  // unused entries sort last
    if (t < m) {
        spawn_draw((uint32_t)(base + t), step, seed, raw);
        kv = ((raw[0] >> 10) << 10) | (uint32_t)t;
    }
    // bitonic sort of the 1024 keys, one key per thread: compare-exchange partners closer
    // than a wavefront trade through cross-lane shuffles, the rest through LDS
    for (int size = 2; size <= C; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            uint32_t other;
            if (stride >= 64) {
                key[t] = kv;
                __syncthreads();
                other = key[t ^ stride];
                __syncthreads();
            } else {
                other = (uint32_t)__shfl_xor((int)kv, stride, 64);
            }
            const bool up = (t & size) == 0, lower = (t & stride) == 0;
            kv = (lower == up) ? min(kv, other) : max(kv, other);
        }
    }
    if (t >= m) return;
    const uint32_t j = (uint32_t)base + (kv & (SPAWN_CHUNK - 1));  // t < C <= 1024
    spawn_draw(j, step, seed, raw);
    float u[4];
#pragma unroll
    for (int i = 0; i < 4; i++) u[i] = (float)(raw[i] >> 8) * (1.0f / 16777216.0f);
    const float rt = u[0] * 2.0f * 3.14159265358979323846f;
    float sn, cs;
    sincosf(rt, &sn, &cs);
    const long long s = (first + base + t) % nslots;
    parts[3 * s + 0] = (jt.cx - 0.1f) + (u[1] * 0.1f - 0.05f);
    parts[3 * s + 1] = (jt.cy + 0.9f * jt.rad * cs) + (u[2] * 0.1f - 0.05f);
    parts[3 * s + 2] = (jt.cz + 0.9f * jt.rad * sn) + (u[3] * 0.1f - 0.05f);
}

// ======================================================================== launchers

bool x_len_supported(int n) { return n == 16 || n == 32 || n == 64 || n == 128 || n == 256 || n == 512; }

static bool yz_len_ok(int n) { return n == 16 || n == 32 || n == 64 || n == 128 || n == 256 || n == 512; }
bool plane_supported(int ny, int nz) { return yz_len_ok(ny) && yz_len_ok(nz); }
// planes whose y-z transform runs in one workgroup (k_plane); larger ones are split
bool plane_in_lds(int ny, int nz) { return ny <= 128 && nz <= 128; }

// y lines per block of the x passes: g.ty when set (8, 16 or 32, not above ny; 32 only
// for nx <= 256), else 16
int x_tile_lines(int nx, int ny) {
    const int ty = 16;
    return ty < ny ? ty : ny;
}
static int x_tile(const GridInfo& g) {
    int ty = g.ty > 0 ? g.ty : x_tile_lines(g.nx, g.ny);
    if (ty > g.ny) ty = g.ny;
    if (ty == 32 && g.nx > 256) ty = 16;
    return ty;
}

#define ISF_LAUNCH(kernel, grid, block, stream, ...)                      \
    do {                                                                  \
        hipLaunchKernelGGL(kernel, grid, block, 0, stream, __VA_ARGS__);  \
        return (int)hipGetLastError();                                    \
    } while (0)

static inline unsigned members_of(const Batch& b) { return b.inst ? (unsigned)b.members : 1u; }

// done != null: the launch itself completes that event when the kernel ends
// (hipExtLaunchKernel's stop event), with no marker packet behind the kernel on the stream
template <int NX, int TY>
static int flow_x_fwd_t(const GridInfo& g, const float4* psi, const float2* sa, const float2* sb, float2* S,
                        const DevTables& t, hipStream_t s, Batch b, const VelOut* vo, hipEvent_t done = nullptr) {
    dim3 grid(g.ny / TY, g.nz, members_of(b));
    if (vo) {
        if (sa) return (int)hipErrorInvalidValue;
        if (done) {
            hipExtLaunchKernelGGL((k_flow_x_fwd<NX, TY, true>), grid, dim3(2 * TY * XCfg<NX>::T), 0, s, nullptr, done,
                                  0, psi, sa, sb, S, g.ny, g.nz, t.tw, b, *vo);
            return (int)hipGetLastError();
        }
        ISF_LAUNCH((k_flow_x_fwd<NX, TY, true>), grid, dim3(2 * TY * XCfg<NX>::T), s, psi, sa, sb, S, g.ny, g.nz,
                   t.tw, b, *vo);
    }
    if (done) return (int)hipErrorInvalidValue;
    ISF_LAUNCH((k_flow_x_fwd<NX, TY, false>), grid, dim3(2 * TY * XCfg<NX>::T), s, psi, sa, sb, S, g.ny, g.nz,
               t.tw, b, VelOut{});
}
template <int NX, int TY>
static int flow_x_inv_t(const GridInfo& g, const float2* S, float4* psi, float2* oa, float2* ob, float4* psin,
                        bool nrm, const DevTables& t, hipStream_t s, Batch b) {
    dim3 grid(g.ny / TY, g.nz, members_of(b));
    if (nrm)
        ISF_LAUNCH((k_flow_x_inv<NX, TY, true>), grid, dim3(2 * TY * XCfg<NX>::T), s, S, psi, oa, ob, nullptr, g.ny,
                   g.nz, t.tw, b);
    ISF_LAUNCH((k_flow_x_inv<NX, TY, false>), grid, dim3(2 * TY * XCfg<NX>::T), s, S, psi, oa, ob, psin, g.ny, g.nz,
               t.tw, b);
}
template <int NX, int TY>
static int r2c_x_t(const GridInfo& g, const float* f, float2* P, const DevTables& t, hipStream_t s, Batch b) {
    dim3 grid(g.ny / TY, g.nz, members_of(b));
    ISF_LAUNCH((k_r2c_x<NX, TY>), grid, dim3(TY * XCfg<NX / 2>::T), s, f, P, g.ny, g.nz, t.tw, b);
}
template <int NX, int TY>
static int c2r_x_t(const GridInfo& g, const float2* P, int mode, float* q, float4* psi, const JetParams* jet,
                   double* parts, const DevTables& t, hipStream_t s, Batch b, float4* rout) {
    dim3 grid(g.ny / TY, g.nz, members_of(b));
    dim3 block(TY * C2rCfg<NX / 2>::T);
    JetParams j{};
    if (jet) j = *jet;
    if (mode == C2R_Q) ISF_LAUNCH((k_c2r_x<NX, TY, C2R_Q>), grid, block, s, P, q, psi, g, j, parts, t.tw, b, nullptr);
    if (mode == C2R_GAUGE)
        ISF_LAUNCH((k_c2r_x<NX, TY, C2R_GAUGE>), grid, block, s, P, q, psi, g, j, parts, t.tw, b, nullptr);
    if (mode == C2R_GAUGE_AND_RESTRICTED) {
        if (!jet || !rout || b.inst) return (int)hipErrorInvalidValue;
        ISF_LAUNCH((k_c2r_x<NX, TY, C2R_GAUGE_AND_RESTRICTED>), grid, block, s, P, q, psi, g, j, parts, t.tw, b, rout);
    }
    ISF_LAUNCH((k_c2r_x<NX, TY, C2R_GAUGE_RESTRICT>), grid, block, s, P, q, psi, g, j, parts, t.tw, b, nullptr);
}

// dispatch on (nx, tile) for the x passes
#define ISF_X_CASE(NX, FN, ...)                                            \
    case NX:                                                               \
        if (ty == 32 && NX <= 256) return FN<NX, (NX <= 256 ? 32 : 8)>(__VA_ARGS__); \
        if (ty == 16 && NX <= 256) return FN<NX, (NX <= 256 ? 16 : 8)>(__VA_ARGS__); \
        return FN<NX, 8>(__VA_ARGS__);
#define ISF_X_DISPATCH(FN, g, ...)                                                              \
    do {                                                                                        \
        const int ty = x_tile(g);                                                               \
        switch (g.nx) {                                                                         \
            ISF_X_CASE(16, FN, g, __VA_ARGS__)                                                  \
            ISF_X_CASE(32, FN, g, __VA_ARGS__)                                                  \
            ISF_X_CASE(64, FN, g, __VA_ARGS__)                                                  \
            ISF_X_CASE(128, FN, g, __VA_ARGS__)                                                 \
            ISF_X_CASE(256, FN, g, __VA_ARGS__)                                                 \
            case 512:                                                                           \
                if (ty == 16) return FN<512, 16>(g, __VA_ARGS__);                               \
                return FN<512, 8>(g, __VA_ARGS__);                                              \
        }                                                                                       \
        return (int)hipErrorInvalidValue;                                                       \
    } while (0)

int launch_flow_x_fwd(const GridInfo& g, const float4* psi, float2* S, const DevTables& t, hipStream_t s, Batch b,
                      const VelOut* vo, hipEvent_t done) {
    ISF_X_DISPATCH(flow_x_fwd_t, g, psi, nullptr, nullptr, S, t, s, b, vo, done);
}
int launch_flow_x_inv(const GridInfo& g, const float2* S, float4* psi, bool normalize, const DevTables& t,
                      hipStream_t s, Batch b) {
    ISF_X_DISPATCH(flow_x_inv_t, g, S, psi, nullptr, nullptr, nullptr, normalize, t, s, b);
}
int launch_flow_x_inv_raw_normalized(const GridInfo& g, const float2* S, float4* psi, float4* psin,
                                     const DevTables& t, hipStream_t s) {
    ISF_X_DISPATCH(flow_x_inv_t, g, S, psi, nullptr, nullptr, psin, false, t, s, Batch{});
}
int launch_flow_x_fwd_split(const GridInfo& g, const float2* sa, const float2* sb, float2* S, const DevTables& t,
                            hipStream_t s) {
    ISF_X_DISPATCH(flow_x_fwd_t, g, nullptr, sa, sb, S, t, s, Batch{}, nullptr);
}
int launch_flow_x_inv_split(const GridInfo& g, const float2* S, float2* oa, float2* ob, const DevTables& t,
                            hipStream_t s) {
    ISF_X_DISPATCH(flow_x_inv_t, g, S, nullptr, oa, ob, nullptr, false, t, s, Batch{});
}
int launch_r2c_x(const GridInfo& g, const float* f, float2* P, const DevTables& t, hipStream_t s, Batch b) {
    ISF_X_DISPATCH(r2c_x_t, g, f, P, t, s, b);
}
template <int NX, int TY>
static int div_r2c_t(const GridInfo& g, const float4* psi, float2* P, const DevTables& t, hipStream_t s, Batch b) {
    dim3 grid(2 * g.ny / TY, g.nz / 2, members_of(b));  // TY/2 rows in each of two z planes
    ISF_LAUNCH((k_div_r2c<NX, TY>), grid, dim3(DivR2cCfg<NX, TY>::NT), s, psi, P, g, t.tw, b);
}
// Fused while a z plane of psi (nx ny 16 bytes) is at most 1 MB: the z +- 1 rows a block
// re-reads then stay in its XCD's L2.  Larger grids run k_div_psi (z-marching, each row
// read once) into f and the R2C pass.
static bool div_r2c_fused(const GridInfo& g) { return (size_t)g.nx * g.ny * 16 <= (size_t)1 << 20; }
int launch_div_r2c(const GridInfo& g, const float4* psi, float* f, float2* P, const DevTables& t, hipStream_t s,
                   Batch b) {
    if (!div_r2c_fused(g)) {
        const int e = launch_div_psi(g, psi, f, s, b);
        if (e) return e;
        return launch_r2c_x(g, f, P, t, s, b);
    }
    ISF_X_DISPATCH(div_r2c_t, g, psi, P, t, s, b);
}
int launch_c2r_x(const GridInfo& g, const float2* P, int mode, float* q, float4* psi, const JetParams* jet,
                 double* norm_partials, const DevTables& t, hipStream_t s, Batch b, float4* rout) {
    ISF_X_DISPATCH(c2r_x_t, g, P, mode, q, psi, jet, norm_partials, t, s, b, rout);
}
size_t c2r_x_blocks(const GridInfo& g) { return (size_t)(g.ny / x_tile(g)) * g.nz; }

static int cu_count() {
    static int cached[64] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    if (!cached[dev]) {
        int n = 0;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
        cached[dev] = n;
    }
    return cached[dev];
}

template <int NY, int NZ, int EO>
static int plane_t(int mode, float2* base, int nplanes, int nx, double scale, const DevTables& t, hipStream_t s,
                   Batch b) {
    using C = PlaneCfg<NY, NZ, EO>;
    const int by_lds = (160 * 1024) / C::LDS_BYTES, by_threads = 2048 / C::NT;
    const int per_cu = std::max(1, std::min(by_lds, by_threads));
    const int mem = (int)members_of(b);
    // blocks per member: enough to fill the GPU once across all members
    const int blocks = std::max(1, std::min(nplanes, (cu_count() * per_cu + mem - 1) / mem));
    const dim3 grid(blocks, mem);
    if (mode == PLANE_FLOW)
        ISF_LAUNCH((k_plane<NY, NZ, PLANE_FLOW, EO>), grid, dim3(C::NT), s, base, nplanes, nx, scale, t, b);
    if (mode == PLANE_FLOW_ADJ)
        ISF_LAUNCH((k_plane<NY, NZ, PLANE_FLOW_ADJ, EO>), grid, dim3(C::NT), s, base, nplanes, nx, scale, t, b);
    ISF_LAUNCH((k_plane<NY, NZ, PLANE_POISSON, EO>), grid, dim3(C::NT), s, base, nplanes, nx, scale, t, b);
}
// pe: elements per thread (0: the default; 16 for planes up to 64 x 64 selects the radix
// 16 schedule instead of radix 8)
template <int NY, int NZ>
static int plane_pe(int pe, int mode, float2* base, int nplanes, int nx, double scale, const DevTables& t,
                    hipStream_t s, Batch b) {
    if constexpr (NY <= 64 && NZ <= 64) {
        if (pe == 16) return plane_t<NY, NZ, 16>(mode, base, nplanes, nx, scale, t, s, b);
    }
    return plane_t<NY, NZ, 0>(mode, base, nplanes, nx, scale, t, s, b);
}

#define ISF_PLANE_Z(NY)                                                                     \
    switch (nz) {                                                                           \
        case 16: return plane_pe<NY, 16>(pe, mode, base, nplanes, nx, scale, t, s, b);     \
        case 32: return plane_pe<NY, 32>(pe, mode, base, nplanes, nx, scale, t, s, b);     \
        case 64: return plane_pe<NY, 64>(pe, mode, base, nplanes, nx, scale, t, s, b);     \
        case 128: return plane_pe<NY, 128>(pe, mode, base, nplanes, nx, scale, t, s, b);   \
    }                                                                                       \
    return (int)hipErrorInvalidValue;

template <int NY, int DIR>
static int line_y_t(float2* base, int nplanes, int nz, const DevTables& t, hipStream_t s) {
    using C = YLineCfg<NY>;
    if (nz % C::R) return (int)hipErrorInvalidValue;
    ISF_LAUNCH((k_line_y<NY, DIR>), dim3(nz / C::R, nplanes), dim3(C::NT), s, base, nz, t.tw);
}
template <int DIR>
static int line_y(int ny, float2* base, int nplanes, int nz, const DevTables& t, hipStream_t s) {
    switch (ny) {
        case 16: return line_y_t<16, DIR>(base, nplanes, nz, t, s);
        case 32: return line_y_t<32, DIR>(base, nplanes, nz, t, s);
        case 64: return line_y_t<64, DIR>(base, nplanes, nz, t, s);
        case 128: return line_y_t<128, DIR>(base, nplanes, nz, t, s);
        case 256: return line_y_t<256, DIR>(base, nplanes, nz, t, s);
        case 512: return line_y_t<512, DIR>(base, nplanes, nz, t, s);
    }
    return (int)hipErrorInvalidValue;
}
template <int NZ>
static int line_z_t(int mode, float2* base, int nplanes, int nx, int ny, double scale, const DevTables& t,
                    hipStream_t s) {
    using C = ZColCfg<NZ>;
    if (ny % C::CW) return (int)hipErrorInvalidValue;
    const dim3 grid(ny / C::CW, nplanes);
    if constexpr (NZ / C::E > 1 && NZ / C::E <= C::E) {
        if (mode == PLANE_FLOW) ISF_LAUNCH((k_line_z_reg<NZ, PLANE_FLOW>), grid, dim3(C::NT), s, base, nx, ny, scale, t);
        if (mode == PLANE_FLOW_ADJ)
            ISF_LAUNCH((k_line_z_reg<NZ, PLANE_FLOW_ADJ>), grid, dim3(C::NT), s, base, nx, ny, scale, t);
        ISF_LAUNCH((k_line_z_reg<NZ, PLANE_POISSON>), grid, dim3(C::NT), s, base, nx, ny, scale, t);
    }
    if (mode == PLANE_FLOW) ISF_LAUNCH((k_line_z<NZ, PLANE_FLOW>), grid, dim3(C::NT), s, base, nx, ny, scale, t);
    if (mode == PLANE_FLOW_ADJ) ISF_LAUNCH((k_line_z<NZ, PLANE_FLOW_ADJ>), grid, dim3(C::NT), s, base, nx, ny, scale, t);
    ISF_LAUNCH((k_line_z<NZ, PLANE_POISSON>), grid, dim3(C::NT), s, base, nx, ny, scale, t);
}
// y pass, z pass with the multiplier, inverse y pass
static int plane_split(int mode, int ny, int nz, float2* base, int nplanes, int nx, double scale, const DevTables& t,
                       hipStream_t s) {
    int e = line_y<-1>(ny, base, nplanes, nz, t, s);
    if (e) return e;
    switch (nz) {
        case 16: e = line_z_t<16>(mode, base, nplanes, nx, ny, scale, t, s); break;
        case 32: e = line_z_t<32>(mode, base, nplanes, nx, ny, scale, t, s); break;
        case 64: e = line_z_t<64>(mode, base, nplanes, nx, ny, scale, t, s); break;
        case 128: e = line_z_t<128>(mode, base, nplanes, nx, ny, scale, t, s); break;
        case 256: e = line_z_t<256>(mode, base, nplanes, nx, ny, scale, t, s); break;
        case 512: e = line_z_t<512>(mode, base, nplanes, nx, ny, scale, t, s); break;
        default: return (int)hipErrorInvalidValue;
    }
    if (e) return e;
    return line_y<1>(ny, base, nplanes, nz, t, s);
}

static int plane_dispatch(int mode, int ny, int nz, float2* base, int nplanes, int nx, double scale,
                          const DevTables& t, hipStream_t s, Batch b, int pe) {
    if (!plane_supported(ny, nz)) return (int)hipErrorInvalidValue;
    if (!plane_in_lds(ny, nz)) {
        if (b.inst) return (int)hipErrorInvalidValue;  // batches hold LDS-sized planes only
        return plane_split(mode, ny, nz, base, nplanes, nx, scale, t, s);
    }
    switch (ny) {
        case 16: { ISF_PLANE_Z(16) }
        case 32: { ISF_PLANE_Z(32) }
        case 64: { ISF_PLANE_Z(64) }
        case 128: { ISF_PLANE_Z(128) }
    }
    return (int)hipErrorInvalidValue;
}

int launch_flow_plane(const GridInfo& g, float2* S, const DevTables& t, hipStream_t s, Batch b) {
    return plane_dispatch(PLANE_FLOW, g.ny, g.nz, S, 2 * g.nx, g.nx, 0.0, t, s, b, g.pe);
}
int launch_flow_plane_adj(const GridInfo& g, float2* S, const DevTables& t, hipStream_t s) {
    return plane_dispatch(PLANE_FLOW_ADJ, g.ny, g.nz, S, 2 * g.nx, g.nx, 0.0, t, s, Batch{}, g.pe);
}
int launch_poisson_plane(const GridInfo& g, float2* P, double scale, const DevTables& t, hipStream_t s, Batch b) {
    return plane_dispatch(PLANE_POISSON, g.ny, g.nz, P, g.nx / 2 + 1, g.nx, scale, t, s, b, g.pe);
}

static inline dim3 grid1d(size_t n, int nt) { return dim3((unsigned)((n + nt - 1) / nt)); }
static inline size_t ncells(const GridInfo& g) { return (size_t)g.nx * g.ny * g.nz; }

// z planes walked per block: 16, halved (down to 1) while the launch would have fewer
// than kStencilBlocks blocks; nz >= 16 is a power of two, so tz divides it
constexpr long kStencilBlocks = 2048;
static int stencil_tz(const GridInfo& g, int bx, const Batch& b) {
    const long base = (long)(g.nx / bx) * (g.ny / ST_BY) * members_of(b);
    int tz = ST_TZ;
    while (tz > 1 && base * (g.nz / tz) < kStencilBlocks) tz >>= 1;
    return tz;
}
static dim3 stencil_grid(const GridInfo& g, int bx, int tz, const Batch& b) {
    return dim3(g.nx / bx, g.ny / ST_BY, (g.nz / tz) * members_of(b));
}

int launch_div_psi(const GridInfo& g, const float4* psi, float* f, hipStream_t s, Batch b) {
    const int bx = g.nx >= 64 ? 64 : g.nx, tz = stencil_tz(g, bx, b);
    const dim3 grid = stencil_grid(g, bx, tz, b);
    if (bx == 64) ISF_LAUNCH(k_div_psi<64>, grid, dim3(64 * ST_BY), s, psi, f, g, tz, b);
    if (bx == 32) ISF_LAUNCH(k_div_psi<32>, grid, dim3(32 * ST_BY), s, psi, f, g, tz, b);
    ISF_LAUNCH(k_div_psi<16>, grid, dim3(16 * ST_BY), s, psi, f, g, tz, b);
}
int launch_velocity(const GridInfo& g, const float4* psi, float hbar, int sharp, float* vel, hipStream_t s, Batch b) {
    const int bx = g.nx >= 64 ? 64 : g.nx, tz = stencil_tz(g, bx, b);
    const dim3 grid = stencil_grid(g, bx, tz, b);
    if (bx == 64) ISF_LAUNCH(k_velocity<64>, grid, dim3(64 * ST_BY), s, psi, vel, g, hbar, sharp, tz, b);
    if (bx == 32) ISF_LAUNCH(k_velocity<32>, grid, dim3(32 * ST_BY), s, psi, vel, g, hbar, sharp, tz, b);
    ISF_LAUNCH(k_velocity<16>, grid, dim3(16 * ST_BY), s, psi, vel, g, hbar, sharp, tz, b);
}
int launch_div_vel(const GridInfo& g, const float* vel, float* f, hipStream_t s) {
    ISF_LAUNCH(k_div_vel, grid1d(ncells(g), 256), dim3(256), s, vel, f, g);
}
int launch_sharp(const GridInfo& g, float* vel, hipStream_t s) {
    ISF_LAUNCH(k_sharp, grid1d(ncells(g), 256), dim3(256), s, vel, 1.0f / g.dx, 1.0f / g.dy, 1.0f / g.dz, ncells(g));
}
int launch_normalize(const GridInfo& g, float4* psi, hipStream_t s) {
    ISF_LAUNCH(k_normalize, grid1d(ncells(g), 256), dim3(256), s, psi, ncells(g));
}
int launch_gauge(const GridInfo& g, float4* psi, const float* q, hipStream_t s) {
    ISF_LAUNCH(k_gauge, grid1d(ncells(g), 256), dim3(256), s, psi, q, ncells(g));
}
int launch_isjet(const GridInfo& g, const JetParams& jet, unsigned char* out, hipStream_t s) {
    ISF_LAUNCH(k_isjet, grid1d(ncells(g), 256), dim3(256), s, out, g, jet);
}
int launch_restrict(const GridInfo& g, float4* psi, const unsigned char* isjet, const JetParams& jet, hipStream_t s) {
    ISF_LAUNCH(k_restrict, grid1d(ncells(g), 256), dim3(256), s, psi, isjet, g, jet);
}
int launch_norm2(const GridInfo& g, const float4* psi, double* partials, size_t* nparts, hipStream_t s) {
    const size_t nb = kNorm2Blocks;
    *nparts = nb;
    ISF_LAUNCH(k_norm2<256>, dim3((unsigned)nb), dim3(256), s, psi, ncells(g), partials);
}
int launch_reduce_sum(const double* partials, size_t n, double* out, hipStream_t s) {
    ISF_LAUNCH(k_reduce_sum, dim3(1), dim3(256), s, partials, n, out);
}
// batch mode: max_n = the largest n_spawn of the members (grid size)
int launch_spawn(float* parts, long long nslots, long long first, long long n, const JetParams& jet,
                 unsigned long long seed, unsigned step, hipStream_t s, Batch b, long long max_n) {
    const long long nn = b.inst ? max_n : n;
    if (nn <= 0) return 0;
    int chunk = 64;  // the sort size: the draw count rounded up to a power of two, 64..1024
    while (chunk < SPAWN_CHUNK && chunk < nn) chunk <<= 1;
    const dim3 grid((unsigned)((nn + chunk - 1) / chunk), members_of(b));
    // 16 KB of LDS reserved (unused): with its 4 KB the block cannot share a CU with a
    // 128 x 128 plane-pass workgroup (147 KB of the 160 KB).  Beside one it slows that
    // workgroup, and the plane pass ends with its slowest workgroup (flow plane 46 -> 68 us
    // in a trace where the spawn of the side stream landed there).
    constexpr unsigned kSpawnLdsReserve = 16384;
    hipLaunchKernelGGL(k_spawn, grid, dim3(chunk), kSpawnLdsReserve, s, parts, nslots, first, n, jet, seed, step, b);
    return (int)hipGetLastError();
}
// batch mode: max_slots = the largest nslots of the members (grid size)
int launch_advect(const GridInfo& g, float* parts, long long nslots, const int* order, const float* vel, float dt,
                  hipStream_t s, Batch b, long long max_slots, const float* src, bool beside) {
    const long long nn = b.inst ? max_slots : nslots;
    if (nn <= 0) return 0;
    if (src == parts) src = nullptr;
    constexpr long long per_block = 256LL * ADVECT_SLOTS_PER_THREAD;
    long long nb = (nn + per_block - 1) / per_block;
    // beside the grid stream's kernels: one block per CU.  Each plane-pass workgroup then
    // shares its CU with exactly one advection block, and the plane pass, which ends with
    // its slowest workgroup, is slowed evenly (step 0.282 -> 0.273 ms at 256x128x128 with
    // 1M particles; 192 blocks 0.278, 320 blocks 0.288, uncapped 1954 blocks 0.282)
    if (beside && !b.inst && nb > cu_count()) nb = cu_count();
    const dim3 grid((unsigned)nb, members_of(b));
    ISF_LAUNCH(k_advect, grid, dim3(256), s, parts, nslots, order, vel, g, dt, b, src);
}

__global__ void k_set_counter(int* c, int v) { c[0] = v; }
__global__ void k_bump_counter(int* c) { c[0] += 1; }
int launch_set_counter(int* counter, int value, hipStream_t s) {
    ISF_LAUNCH(k_set_counter, dim3(1), dim3(1), s, counter, value);
}
int launch_bump_counter(int* counter, hipStream_t s) { ISF_LAUNCH(k_bump_counter, dim3(1), dim3(1), s, counter); }

}  // namespace isf
