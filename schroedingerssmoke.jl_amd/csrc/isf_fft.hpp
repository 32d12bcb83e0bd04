// isf_fft.hpp -- device building blocks of the power-of-two complex FFT used by the ISF
// kernels (gfx950, 64-wide wavefronts).
//
// A line of N complex values lives in LDS; every thread of the line holds E = N/T
// elements in registers and performs radix-R butterflies (R = min(E, remaining length))
// of the Stockham autosort formulation (Govindaraju et al., "High performance discrete
// Fourier transforms on graphics processors", SC'08): pass p reads element
// j + r*N/R, multiplies by exp(-2 pi i r (j mod Ns) / (Ns R)), runs an R-point DFT and
// writes element (j/Ns)*Ns*R + (j mod Ns) + r*Ns.  No bit reversal is needed.
//
// Twiddles come from a table of N-th roots of unity, tw_N[m] = exp(-2 pi i m / N),
// computed in double on the host and rounded once; the table for length N starts at
// element N of the combined table in global memory (lengths 2..512 fit in 1024
// entries) and each kernel copies the tables it needs into LDS first.
#pragma once
#include <hip/hip_runtime.h>

namespace isf {

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }

// a * (DIR * i): DIR = -1 is the forward transform's -i, DIR = +1 the inverse's +i.
template <int DIR>
__device__ __forceinline__ float2 mul_i(float2 a) {
    return DIR < 0 ? make_float2(a.y, -a.x) : make_float2(-a.y, a.x);
}
// forward multiplies by w = exp(-i theta), inverse by conj(w)
template <int DIR>
__device__ __forceinline__ float2 twmul(float2 a, float2 w) {
    return DIR < 0 ? make_float2(a.x * w.x - a.y * w.y, a.x * w.y + a.y * w.x)
                   : make_float2(a.x * w.x + a.y * w.y, a.y * w.x - a.x * w.y);
}
// multiply by the constant (c, DIR*s)
template <int DIR>
__device__ __forceinline__ float2 rot(float2 a, float c, float s) {
    const float sd = DIR < 0 ? -s : s;
    return make_float2(a.x * c - a.y * sd, a.x * sd + a.y * c);
}

// ------------------------------------------------------------------ R-point DFTs
// X_k = sum_n x_n exp(DIR * 2 pi i n k / R), computed in place on v[0..R-1].
template <int R, int DIR> struct Dft;

template <int DIR> struct Dft<1, DIR> {
    __device__ __forceinline__ static void run(float2*) {}
};

template <int DIR> struct Dft<2, DIR> {
    __device__ __forceinline__ static void run(float2* v) {
        float2 a = v[0];
        v[0] = cadd(a, v[1]);
        v[1] = csub(a, v[1]);
    }
};

template <int DIR> struct Dft<4, DIR> {
    __device__ __forceinline__ static void run(float2* v) {
        float2 a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
        float2 a2 = cadd(v[1], v[3]), a3 = mul_i<DIR>(csub(v[1], v[3]));
        v[0] = cadd(a0, a2);
        v[2] = csub(a0, a2);
        v[1] = cadd(a1, a3);
        v[3] = csub(a1, a3);
    }
};

template <int DIR> struct Dft<8, DIR> {
    __device__ __forceinline__ static void run(float2* v) {
        constexpr float C = 0.70710678118654752440f;
        float2 e[4] = {v[0], v[2], v[4], v[6]};
        float2 o[4] = {v[1], v[3], v[5], v[7]};
        Dft<4, DIR>::run(e);
        Dft<4, DIR>::run(o);
        o[1] = rot<DIR>(o[1], C, C);
        o[2] = mul_i<DIR>(o[2]);
        o[3] = rot<DIR>(o[3], -C, C);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            v[k] = cadd(e[k], o[k]);
            v[k + 4] = csub(e[k], o[k]);
        }
    }
};

template <int DIR> struct Dft<16, DIR> {
    __device__ __forceinline__ static void run(float2* v) {
        constexpr float C1 = 0.92387953251128675613f;  // cos(pi/8)
        constexpr float S1 = 0.38268343236508977173f;  // sin(pi/8)
        constexpr float C = 0.70710678118654752440f;
        float2 e[8], o[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            e[k] = v[2 * k];
            o[k] = v[2 * k + 1];
        }
        Dft<8, DIR>::run(e);
        Dft<8, DIR>::run(o);
        o[1] = rot<DIR>(o[1], C1, S1);
        o[2] = rot<DIR>(o[2], C, C);
        o[3] = rot<DIR>(o[3], S1, C1);
        o[4] = mul_i<DIR>(o[4]);
        o[5] = rot<DIR>(o[5], -S1, C1);
        o[6] = rot<DIR>(o[6], -C, C);
        o[7] = rot<DIR>(o[7], -C1, S1);
#pragma unroll
        for (int k = 0; k < 8; k++) {
            v[k] = cadd(e[k], o[k]);
            v[k + 8] = csub(e[k], o[k]);
        }
    }
};

// ------------------------------------------------------------------ LDS line layouts
// Contiguous line with one pad slot every 16 elements (breaks the stride-R bank
// pattern of the Stockham stores).
__device__ __forceinline__ int pad16(int e) { return e + (e >> 4); }
// Padded length of a contiguous line of n elements, made odd so that lines placed back
// to back start on different banks.
__host__ __device__ constexpr int padded_len(int n) { return n + n / 16 + 1; }

struct LayContig {
    float2* p;
    __device__ __forceinline__ float2 ld(int e) const { return p[pad16(e)]; }
    __device__ __forceinline__ void st(int e, float2 v) const { p[pad16(e)] = v; }
};
struct LayStrided {
    float2* p;
    int s;
    __device__ __forceinline__ float2 ld(int e) const { return p[e * s]; }
    __device__ __forceinline__ void st(int e, float2 v) const { p[e * s] = v; }
};

// Global-to-LDS staging by NT threads (thread tid < NT): element i = tid + k NT,
// k < ceil(COUNT / NT), is read with ld(i) and written with st(i, value).  All of a
// thread's loads are issued before its first store, so their latencies overlap (a loop
// that stores each value as it arrives waits for every load in turn).
template <int COUNT, int NT, class Ld, class St>
__device__ __forceinline__ void stage(int tid, const Ld& ld, const St& st) {
    constexpr int K = (COUNT + NT - 1) / NT;
    decltype(ld(0)) v[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int i = tid + k * NT;
        if (COUNT % NT == 0 || i < COUNT) v[k] = ld(i);
    }
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int i = tid + k * NT;
        if (COUNT % NT == 0 || i < COUNT) st(i, v[k]);
    }
}

// ------------------------------------------------------------------ twiddle tables
// Pass (N, R, NS) needs w(kk, r) = exp(-2 pi i r kk / (NS R)) for kk < NS, r < R.  It is
// kept in LDS as a [kk][r] table with row pitch R + 1: lanes of one wavefront read
// different kk with the same r, and the odd pitch puts them in different banks (reading
// tw_N[r kk N/(NS R)] directly strides by a multiple of the bank count).
template <int N, int R, int NS>
struct PassTw {
    static constexpr int SIZE = NS > 1 ? NS * (R + 1) : 0;
    // fill from the global table of N-th roots (tw_n = tw_all + N) by NT threads; caller
    // synchronises
    template <int NT>
    __device__ __forceinline__ static void fill(float2* t, const float2* __restrict__ tw_n, int tid) {
        if constexpr (SIZE > 0)
            stage<SIZE, NT>(
                tid,
                [&](int i) {
                    const int kk = i / (R + 1), r = i % (R + 1);
                    return r < R ? tw_n[(r * kk * (N / (NS * R))) % N] : make_float2(1.0f, 0.0f);
                },
                [&](int i, float2 v) { t[i] = v; });
    }
    __device__ __forceinline__ static float2 get(const float2* t, int kk, int r) { return t[kk * (R + 1) + r]; }
};

// ------------------------------------------------------------------ Stockham passes
// Orders a wavefront's LDS accesses before the call before those after it, for exchanges
// among the lanes of one wavefront (a wavefront's LDS instructions execute in order; the
// fences keep the compiler from moving them across).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One radix-R pass over a line of N held in LDS by T = N/E threads (thread t of the
// line).  The T threads of a line are lanes of one wavefront (T divides 64 and the caller
// gives line l to threads l*T .. l*T + T - 1), so the pass orders its reads and writes
// with wavefront fences only; threads with active == false do nothing.  tw is this
// pass's PassTw table.
template <class Lay, int N, int R, int NS, int E, int DIR>
__device__ __forceinline__ void stockham_pass(const Lay& lay, int t, bool active, const float2* tw) {
    constexpr int T = N / E;
    constexpr int NB = E / R;
    static_assert(E <= N && NB >= 1 && R * NB == E, "bad FFT schedule");
    static_assert(64 % T == 0, "a line's threads must share a wavefront");
    float2 v[NB][R];
    if (active) {
#pragma unroll
        for (int b = 0; b < NB; b++) {
            const int j = t + T * b;
#pragma unroll
            for (int r = 0; r < R; r++) v[b][r] = lay.ld(j + r * (N / R));
            if (NS > 1) {
                const int kk = j % NS;
#pragma unroll
                for (int r = 1; r < R; r++) v[b][r] = twmul<DIR>(v[b][r], PassTw<N, R, NS>::get(tw, kk, r));
            }
            Dft<R, DIR>::run(v[b]);
        }
    }
    wave_lds_sync();
    if (active) {
#pragma unroll
        for (int b = 0; b < NB; b++) {
            const int j = t + T * b;
            const int kk = j % NS;
            const int idxd = (j / NS) * NS * R + kk;
#pragma unroll
            for (int r = 0; r < R; r++) lay.st(idxd + r * NS, v[b][r]);
        }
    }
    wave_lds_sync();
}

// The passes of a length-N transform with E elements per thread: radix
// R = min(E, N/NS) at every step.  Their twiddle tables lie back to back.
template <int N, int E, int NS>
struct Sched {
    static constexpr int REM = N / NS;
    static constexpr int R = E < REM ? E : REM;
    static constexpr int TW = PassTw<N, R, NS>::SIZE + Sched<N, E, NS * R>::TW;
    template <int NT>
    __device__ __forceinline__ static void fill(float2* t, const float2* __restrict__ tw_n, int tid) {
        PassTw<N, R, NS>::template fill<NT>(t, tw_n, tid);
        Sched<N, E, NS * R>::template fill<NT>(t + PassTw<N, R, NS>::SIZE, tw_n, tid);
    }
    template <class Lay, int DIR>
    __device__ __forceinline__ static void run(const Lay& lay, int t, bool active, const float2* tw) {
        stockham_pass<Lay, N, R, NS, E, DIR>(lay, t, active, tw);
        Sched<N, E, NS * R>::template run<Lay, DIR>(lay, t, active, tw + PassTw<N, R, NS>::SIZE);
    }
};
template <int N, int E>
struct Sched<N, E, N> {
    static constexpr int TW = 0;
    template <int NT>
    __device__ __forceinline__ static void fill(float2*, const float2* __restrict__, int) {}
    template <class Lay, int DIR>
    __device__ __forceinline__ static void run(const Lay&, int, bool, const float2*) {}
};

// Twiddle tables of the whole length-N schedule (SIZE entries, at least 1).
template <int N, int E>
struct FftTw {
    static constexpr int SIZE = Sched<N, E, 1>::TW > 0 ? Sched<N, E, 1>::TW : 1;
    // tw_all: combined table of roots (length-N roots at tw_all + N), NT threads; caller
    // synchronises
    template <int NT>
    __device__ __forceinline__ static void fill(float2* t, const float2* __restrict__ tw_all, int tid) {
        Sched<N, E, 1>::template fill<NT>(t, tw_all + N, tid);
    }
};

// Full length-N transform of one LDS line (DIR = -1 forward, +1 inverse, unnormalised).
// tw holds the schedule's tables (FftTw<N, E>).  All threads of the block call it: the
// line's data must be in LDS (block barrier) before, and it ends with a block barrier, so
// the caller may read other lines afterwards.
template <int N, int E, int DIR, class Lay>
__device__ __forceinline__ void fft_line(const Lay& lay, int t, bool active, const float2* tw) {
    Sched<N, E, 1>::template run<Lay, DIR>(lay, t, active, tw);
    __syncthreads();
}

// Copy the length-N twiddle table (at tw_all + N) into LDS; the caller synchronises.
template <int N, int NT>
__device__ __forceinline__ void load_twiddles(float2* dst, const float2* __restrict__ tw_all) {
    stage<N, NT>(
        (int)threadIdx.x, [&](int i) { return tw_all[N + i]; }, [&](int i, float2 v) { dst[i] = v; });
}

// ------------------------------------------------------------------ register passes
// The same Stockham pass split into load / compute / store, for kernels that keep a
// thread's E elements in registers between passes.  Thread t handles butterflies
// j = t*NB + b (b < NB, blocked so that a thread's elements are adjacent in the first
// and last passes); butterfly b uses v[b*R .. b*R+R-1], reads elements in(t,b,r) and
// writes out(t,b,r).  When pass B reads exactly the elements pass A wrote (same t, b, r
// -> same index), the LDS round trip between them is skipped.  compute() takes the pass's
// PassTw<N, R, NS> table.
template <int N, int R, int NS, int E>
struct RPass {
    static constexpr int T = N / E;
    static constexpr int NB = E / R;
    static_assert(E <= N && NB >= 1 && R * NB == E, "bad FFT schedule");
    __device__ __forceinline__ static int in(int t, int b, int r) { return t * NB + b + r * (N / R); }
    __device__ __forceinline__ static int out(int t, int b, int r) {
        const int j = t * NB + b;
        return (j / NS) * NS * R + (j % NS) + r * NS;
    }
    template <int DIR>
    __device__ __forceinline__ static void compute(float2* v, int t, const float2* tw) {
#pragma unroll
        for (int b = 0; b < NB; b++) {
            if (NS > 1) {
                const int kk = (t * NB + b) % NS;
#pragma unroll
                for (int r = 1; r < R; r++) v[b * R + r] = twmul<DIR>(v[b * R + r], PassTw<N, R, NS>::get(tw, kk, r));
            }
            Dft<R, DIR>::run(v + b * R);
        }
    }
    template <class Lay>
    __device__ __forceinline__ static void load(float2* v, const Lay& lay, int t) {
#pragma unroll
        for (int r = 0; r < R; r++)
#pragma unroll
            for (int b = 0; b < NB; b++) v[b * R + r] = lay.ld(in(t, b, r));
    }
    template <class Lay>
    __device__ __forceinline__ static void store(const float2* v, const Lay& lay, int t) {
#pragma unroll
        for (int r = 0; r < R; r++)
#pragma unroll
            for (int b = 0; b < NB; b++) lay.st(out(t, b, r), v[b * R + r]);
    }
};

// Elements per thread used for a line of length N.
template <int N> struct DefaultE { static constexpr int value = (N == 16 || N == 128 || N == 256) ? 16 : (N < 8 ? N : 8); };

}  // namespace isf
