// isf_rollout.hip -- C ABI of the differentiable rollout (isf_rollout_*): the driver's
// set-up and simloop (test/runtests.jl:16-66) as a function of the nozzle velocity, an
// initial phase field, the initial wavefunction and the initial particle positions, a
// particle-density loss, and the reverse sweep through the hand-written adjoints of
// isf_adjoint.hip.
//
// Forward: the step is run operator by operator (the same kernels and arithmetic as
// isf_step, with the fused normalize and restrict done as separate passes so that their
// inputs can be kept) and psi / particles are checkpointed every K steps; each step's
// particle part (spawn, velocity, advection) runs on a second stream beside the next
// step's psi passes.  Backward: segments are processed last to first; each is recomputed
// from its checkpoint keeping the per-step tape (step input psi0, flow output psi1, first
// gauge output psi3, the two gauge phases q1 q2, particles after the spawn, and with
// particles the step's velocity field, which its advection and advection adjoint read) and
// then reversed step by step.  The
// recompute of segment s-1 runs on a second stream, with its own scratch and the other of
// two tape sets, while segment s is reversed on the handle's stream.  The particle part
// of the reverse (velocity, advection adjoint, spawn adjoint) only feeds each step's
// velocity cotangent into the psi chain, so it runs ahead on a third stream, into one
// velocity-cotangent buffer per step of the tape set.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "isf_handle.hpp"
#include "isf_internal.hpp"

using namespace isf;

static constexpr uint32_t kRolloutMagic = 0x2011A0D7u;

// Per-step tape of one segment: step outputs inside the segment (mid), flow output psi1,
// first gauge output psi3, gauge phases q1 q2, velocity field of the step output vel
// (with particles only).
struct Tape {
    std::vector<float4*> mid, psi1, psi3;
    std::vector<float*> q1, q2, vel;
};
// Scratch and stream of one chain of forward steps.
struct Work {
    float2* S;     // spectrum (16N bytes); f at byte 0 and P at byte 4N in projections
    float* f;      // div scratch of the wide-plane projection path
    float* parts;  // particles being stepped
    hipStream_t st;
};

struct isf_rollout {
    uint32_t magic;
    isf_t h;
    isf_rollout_desc d;
    GridInfo g;
    size_t N;
    long long P;
    int T, K, C, nseg, M;
    SplatCfg sc;
    std::vector<void*> allocs;
    // set-up tape: cpsi[j] = input of constraint iteration j (cpsi[C] = step-1 input)
    std::vector<float4*> cpsi;
    std::vector<float*> cq;
    float* theta;
    bool use_theta;
    float* parts_init;
    // checkpoints at steps 0, K, 2K, ... (ck_psi[0] = cpsi[C])
    std::vector<float4*> ck_psi;
    // particle trajectory of the forward (with particles): traj[i] = the state of step
    // i + 1 after its spawn (its advection's input), traj[T] = the final state (parts)
    std::vector<float*> traj;
    // segment tapes: set s % 2 holds segment s during the backward
    Tape tape[2];
    // recompute stream of the backward with its own scratch (spectrum / Poisson, f,
    // particles); rec_done[i] / adj_done[i]: tape set i recomputed / reversed
    Work rw;
    hipEvent_t rec_done[2], adj_done[2], start;
    // particle stream of the forward and the backward (the handle's side stream, idle
    // otherwise during a rollout: a stream of its own would be the handle's fourth and
    // share a hardware queue with the recompute stream under the runtime's default of four
    // queues per process, which serialises the two); vbar[i][k] = velocity cotangent of step k of the segment in tape
    // set i (ready at pev[i][k]); pdone[i]: that segment's particle part done
    hipStream_t pst;
    std::vector<float*> vbar[2];
    std::vector<hipEvent_t> pev[2];
    hipEvent_t pdone[2];
    // forward: the particle part of each step runs on pst after that step's psi (fstep);
    // fseg[i]: the particle part of the last forward segment in tape set i done (its
    // velocities have read the set's step outputs)
    hipEvent_t fstep, fseg[2];
    // work
    float4* psibar;
    float *f, *fbar, *parts, *parts_final, *pbar, *phibar, *target, *resid;
    // spawn batches of steps 1..T (n_spawn x 3 each, in slot order): they depend only on
    // the nozzle geometry, the seed and the step, so a segment forward copies them in
    // instead of drawing and sorting them again; built for spawn_key
    float* spawn_tab;
    long long spawn_n;
    bool spawn_ready;
    float spawn_key[4];
    unsigned long long spawn_seed;
    double *rho, *partials, *loss_dev, *vacc;
    double* splat_part;  // per-block density grids (null: global atomics)
    int* adj_order;      // cell-sorted slot order of the advection adjoint
    char* adj_order_scratch;
    size_t adj_order_bytes;
    // parameters of the last forward
    float v[3];
    float omega;
    JetParams jet;  // kvec from v, omgterm set per use
    bool forward_done;
    bool last_tape;  // tape set (nseg - 1) % 2 still holds the forward's last segment
};

static int check_rollout(isf_rollout_t r) {
    if (!r || r->magic != kRolloutMagic) return fail(ISF_EHANDLE, "null or destroyed isf rollout");
    return ISF_OK;
}
#define CHECK_R(r)                      \
    do {                                \
        int r_ = check_rollout(r);      \
        if (r_) return r_;              \
    } while (0)

static void rollout_free(isf_rollout_t r) {
    if (r->h && r->h->stream) (void)hipStreamSynchronize(r->h->stream);
    if (r->rw.st) (void)hipStreamSynchronize(r->rw.st);
    if (r->pst) (void)hipStreamSynchronize(r->pst);
    for (int i = 0; i < 2; i++) {
        if (r->rec_done[i]) (void)hipEventDestroy(r->rec_done[i]);
        if (r->adj_done[i]) (void)hipEventDestroy(r->adj_done[i]);
    }
    if (r->start) (void)hipEventDestroy(r->start);
    for (int i = 0; i < 2; i++) {
        for (hipEvent_t e : r->pev[i])
            if (e) (void)hipEventDestroy(e);
        if (r->pdone[i]) (void)hipEventDestroy(r->pdone[i]);
        if (r->fseg[i]) (void)hipEventDestroy(r->fseg[i]);
    }
    if (r->fstep) (void)hipEventDestroy(r->fstep);
    if (r->rw.st) (void)hipStreamDestroy(r->rw.st);
    for (void* p : r->allocs) (void)hipFree(p);
    r->magic = 0;
    delete r;
}

template <class T>
static int ralloc(isf_rollout_t r, T** p, size_t count) {
    *p = nullptr;
    void* q = nullptr;
    const hipError_t e = hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T));
    if (e != hipSuccess) return fail(ISF_ENOMEM, "rollout buffers (%zu bytes): %s", count * sizeof(T), hipGetErrorString(e));
    r->allocs.push_back(q);
    *p = static_cast<T*>(q);
    return ISF_OK;
}

static bool pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }

// omega = sum(v.^2) / (2 hbar) with Float32 roundings and no contraction, as the driver
// computes it (test/runtests.jl:24)
static float driver_omega(const float v[3], float hbar) {
#pragma clang fp contract(off)
    const float s = (v[0] * v[0] + v[1] * v[1]) + v[2] * v[2];
    return s / (2.0f * hbar);
}

// ------------------------------------------------------------------ forward pieces

static float2* P_in(isf_rollout_t r, float2* S) {
    return reinterpret_cast<float2*>(reinterpret_cast<char*>(S) + 4 * r->N);
}
static float2* P_of(isf_rollout_t r) { return P_in(r, r->h->S); }
// the handle's stream with the handle's scratch and the rollout's work buffers
static Work main_work(isf_rollout_t r) { return Work{r->h->S, r->f, r->parts, r->h->stream}; }
static double poisson_scale(isf_rollout_t r) { return -0.25 * 2.0 / (double)r->N; }

// fout = Poisson(fin) (src/parallel.jl:168-174)
static int poisson(isf_rollout_t r, const float* fin, float* fout) {
    isf_t h = r->h;
    LAUNCH_CHECK(launch_r2c_x(r->g, fin, P_of(r), h->tb, h->stream), "poisson x forward");
    LAUNCH_CHECK(launch_poisson_plane(r->g, P_of(r), poisson_scale(r), h->tb, h->stream), "poisson plane");
    LAUNCH_CHECK(launch_c2r_x(r->g, P_of(r), 0, fout, nullptr, nullptr, nullptr, h->tb, h->stream), "poisson x inverse");
    return ISF_OK;
}
// pressure_project! in place keeping the gauge phase q (src/parallel.jl:155-160); with
// restricted != null the gauge pass also writes restrict_velocity! of its output there
// (src/particles.jl:12-27, jet parameters *j).
static int project_q(isf_rollout_t r, float4* psi, float* q, const Work& w, float4* restricted = nullptr,
                     const JetParams* j = nullptr) {
    isf_t h = r->h;
    float2* P = P_in(r, w.S);
    LAUNCH_CHECK(launch_div_r2c(r->g, psi, w.f, P, h->tb, w.st), "div(one_form(psi)) + poisson x forward");
    LAUNCH_CHECK(launch_poisson_plane(r->g, P, poisson_scale(r), h->tb, w.st), "poisson plane");
    if (restricted)
        LAUNCH_CHECK(launch_c2r_x(r->g, P, 3, q, psi, j, nullptr, h->tb, w.st, Batch{}, restricted),
                     "poisson x inverse + gauge + restrict_velocity");
    else
        LAUNCH_CHECK(launch_c2r_x(r->g, P, 1, q, psi, nullptr, nullptr, h->tb, w.st), "poisson x inverse + gauge");
    return ISF_OK;
}

static int seg_steps(isf_rollout_t r, int s) { return std::min(r->K, r->T - s * r->K); }
static float4* psi0_of(isf_rollout_t r, const Tape& tp, int s, int k) {
    if (k == 0) return r->ck_psi[s];
    if (k == seg_steps(r, s)) return r->ck_psi[s + 1];
    return tp.mid[k - 1];
}
static JetParams jet_at(isf_rollout_t r, int it) {
    JetParams j = r->jet;
    j.omgterm = it > 0 ? r->omega * ((float)it * r->h->dt) : 0.0f;  // test/runtests.jl:46,57
    return j;
}
static void spawn_slots(isf_rollout_t r, int it, long long* first0, long long* n) {
    *first0 = 0;
    *n = 0;
    if (r->P <= 0 || r->h->n_spawn <= 0) return;
    const long long start1 = ((long long)(it - 1) * r->h->n_spawn + 1) % r->P;  // test/runtests.jl:52
    *first0 = start1 == 0 ? r->P - 1 : start1 - 1;
    *n = r->h->n_spawn;
}

// Build the spawn table when missing or when the handle's nozzle or seed changed.
static int ensure_spawn_table(isf_rollout_t r, hipStream_t st) {
    isf_t h = r->h;
    if (!r->spawn_tab || h->n_spawn != r->spawn_n) return ISF_OK;
    const float key[4] = {h->jet.cx, h->jet.cy, h->jet.cz, h->jet.rad};
    if (r->spawn_ready && r->spawn_seed == h->seed && std::memcmp(key, r->spawn_key, sizeof key) == 0) return ISF_OK;
    for (int it = 1; it <= r->T; it++)
        LAUNCH_CHECK(launch_spawn(r->spawn_tab + 3 * (size_t)r->spawn_n * (it - 1), r->spawn_n, 0, r->spawn_n, h->jet,
                                  h->seed, (unsigned)it, st),
                     "add_particles table");
    std::memcpy(r->spawn_key, key, sizeof key);
    r->spawn_seed = h->seed;
    r->spawn_ready = true;
    return ISF_OK;
}
// add_particles! of step it into the ring slots first0 .. first0 + nsp - 1 (mod P), from
// the spawn table (the same values k_spawn writes there), else with k_spawn.
static int spawn_step(isf_rollout_t r, int it, long long first0, long long nsp, float* parts, hipStream_t st) {
    isf_t h = r->h;
    if (!r->spawn_ready || nsp != r->spawn_n || nsp > r->P) {
        LAUNCH_CHECK(launch_spawn(parts, r->P, first0, nsp, h->jet, h->seed, (unsigned)it, st), "add_particles");
        return ISF_OK;
    }
    const float* src = r->spawn_tab + 3 * (size_t)nsp * (it - 1);
    const long long n1 = std::min(nsp, r->P - first0);
    HIP_CHECK(hipMemcpyAsync(parts + 3 * first0, src, sizeof(float) * 3 * n1, hipMemcpyDeviceToDevice, st),
              "add_particles");
    if (n1 < nsp)
        HIP_CHECK(hipMemcpyAsync(parts, src + 3 * n1, sizeof(float) * 3 * (nsp - n1), hipMemcpyDeviceToDevice, st),
                  "add_particles");
    return ISF_OK;
}

// Steps of segment s from its checkpoint into tape tp, on w's stream and scratch.  The
// forward (checkpoint = true) also writes the next checkpoint; a recompute writes only
// the tape (its last step output is the next psi checkpoint again, rewritten with the
// same values).  As in isf_simloop, the velocity field of step k - 1 (into its tape entry)
// is formed by step k's forward x pass, which reads that step's output anyway, and step
// k - 1's advection follows it, then step k's spawn; the segment's last velocity pass and
// advection follow its last step.  With side, the particle part runs on that stream,
// ordered after each forward x pass by the event its launch completes (fstep), beside the
// step's remaining psi passes; the caller orders side after the set-up and waits for fseg
// at the end.  Steps [k0, k1) only (k1 < 0: to the segment's end), so that a caller can
// interleave a segment's enqueues with another stream's: the set-up comes with k0 = 0 and
// the segment's tail with its last step.
static int seg_forward(isf_rollout_t r, int s, const Tape& tp, const Work& w, bool checkpoint,
                       hipStream_t side = nullptr, int k0 = 0, int k1 = -1) {
    isf_t h = r->h;
    hipStream_t st = w.st;
    hipStream_t pt = side ? side : st;
    const size_t pb = sizeof(float) * 3 * (size_t)r->P;
    const int ks = seg_steps(r, s);
    if (k1 < 0 || k1 > ks) k1 = ks;
    // with particles every pass forms the velocity fields (the particle adjoint reads
    // them); the positions are stepped by the forward alone, into the trajectory, which
    // the backward reads instead of recomputing it
    const bool parts = r->P > 0, ptraj = parts && checkpoint;
    const VelOut vo0{nullptr, h->hbar, 1.0f / r->g.dx, 1.0f / r->g.dy, 1.0f / r->g.dz, 1};
    // the state of step s K + k + 1 after its spawn; the advection of step k reads it and
    // writes the next entry (out of place), so no step copies particles
    auto pos = [&](int k) { return r->traj[(size_t)s * r->K + k]; };
    if (k0 == 0) {
        // the step outputs of tape set s % 2 are read by the velocities of segment s - 2
        if (side && s >= 2) HIP_CHECK(hipStreamWaitEvent(st, r->fseg[s % 2], 0), "hipStreamWaitEvent");
        if (ptraj && s == 0)
            HIP_CHECK(hipMemcpyAsync(pos(0), r->parts_init, pb, hipMemcpyDeviceToDevice, pt), "initial particles");
    }
    // :63 of step k (velocity in its tape entry) on the particle stream
    auto advect = [&](int k) -> int {
        LAUNCH_CHECK(launch_advect(r->g, pos(k + 1), r->P, nullptr, tp.vel[k], h->dt, pt, Batch{}, 0, pos(k),
                                   side != nullptr),
                     "staggered_advect");
        return ISF_OK;
    };
    for (int k = k0; k < k1; k++) {
        const int it = s * r->K + k + 1;
        const float4* in = psi0_of(r, tp, s, k);
        float4* out = psi0_of(r, tp, s, k + 1);
        if (parts && k > 0) {
            // :59-62 of step k - 1 inside this step's forward x pass (in = its output)
            VelOut vo = vo0;
            vo.vel = tp.vel[k - 1];
            LAUNCH_CHECK(launch_flow_x_fwd(r->g, in, w.S, h->tb, st, Batch{}, &vo, side ? r->fstep : nullptr),
                         "flow x forward + velocity_one_form");
            if (side) HIP_CHECK(hipStreamWaitEvent(side, r->fstep, 0), "hipStreamWaitEvent");
            if (ptraj) {
                int e = advect(k - 1);
                if (e) return e;
            }
        } else {
            LAUNCH_CHECK(launch_flow_x_fwd(r->g, in, w.S, h->tb, st), "flow x forward");
        }
        long long first0, nsp;
        spawn_slots(r, it, &first0, &nsp);
        if (ptraj && nsp > 0) {
            int e2 = spawn_step(r, it, first0, nsp, pos(k), pt);
            if (e2) return e2;
        }
        LAUNCH_CHECK(launch_flow_plane(r->g, w.S, h->tb, st), "flow plane");
        // psi1 for the tape, normalize_psi of it into psi3 (projected in place next)
        LAUNCH_CHECK(launch_flow_x_inv_raw_normalized(r->g, w.S, tp.psi1[k], tp.psi3[k], h->tb, st),
                     "flow x inverse + normalize");
        const JetParams j = jet_at(r, it);
        int e = project_q(r, tp.psi3[k], tp.q1[k], w, out, &j);  // out = restrict_velocity!(psi3)
        if (e) return e;
        if ((e = project_q(r, out, tp.q2[k], w))) return e;
    }
    if (k1 < ks) return ISF_OK;
    if (parts) {
        // the last step's velocity and advection
        if (side) {
            HIP_CHECK(hipEventRecord(r->fstep, st), "hipEventRecord");
            HIP_CHECK(hipStreamWaitEvent(side, r->fstep, 0), "hipStreamWaitEvent");
        }
        LAUNCH_CHECK(launch_velocity(r->g, psi0_of(r, tp, s, ks), h->hbar, 1, tp.vel[ks - 1], pt), "velocity_one_form");
        if (ptraj) {
            int e = advect(ks - 1);
            if (e) return e;
        }
    }
    if (side) HIP_CHECK(hipEventRecord(r->fseg[s % 2], side), "hipEventRecord");
    return ISF_OK;
}

// Particle part of the reverse of step k of segment s (tape tp), on the particle stream:
// the advection adjoint at the step's taped velocity field (pbar in place, the velocity
// cotangent into vb) and the spawn adjoint.  sort: order the slots by the cells of this
// step's positions; otherwise the last order is reused.  The order only groups slots
// whose scatters share cells (any order gives the same sums up to the order of the float
// additions); a segment's particles move a few cells at most, so one order per segment
// can serve its steps.  Measured: with 1M slots at 256x128x128 one order per segment took
// the backward from 22.7 to 22.2 ms; with 100k slots at 128x64x64 one run gave 5.3 against
// 5.1-5.2 ms, inside that size's run-to-run spread of about 0.3 ms.  The rollout sorts once
// per segment from kSortPerSegment slots on and at every step below.
constexpr long long kSortPerSegment = 1 << 19;
static int step_adjoint_particles(isf_rollout_t r, const Tape& tp, int s, int k, float* vb, bool sort) {
    isf_t h = r->h;
    hipStream_t st = r->pst;
    const GridInfo& g = r->g;
    const int it = s * r->K + k + 1;
    // :59-63 velocity and advection
    LAUNCH_CHECK(launch_fill_f32(vb, 0.0f, 3 * r->N, st), "fill");
    if (r->P > 0) {
        // cell-sorted slot order: long same-cell runs for the scatter's wavefront reduction
        const float* pk = r->traj[(size_t)s * r->K + k];  // the step's advection input
        if (sort)
            LAUNCH_CHECK(launch_cell_order(g, pk, r->P, r->adj_order, r->adj_order_scratch, r->adj_order_bytes, st),
                         "particle order");
        LAUNCH_CHECK(launch_advect_adj(g, pk, r->P, r->adj_order, tp.vel[k], h->dt, r->pbar, vb, st),
                     "advect adjoint");
        long long first0, nsp;
        spawn_slots(r, it, &first0, &nsp);
        LAUNCH_CHECK(launch_zero_slots(r->pbar, r->P, first0, nsp, st), "spawn adjoint");
    }
    return ISF_OK;
}

// The psi part of the reverse of step k of segment s from its tape tp, on the handle's
// stream, with the step's velocity cotangent vb.
static int step_adjoint(isf_rollout_t r, const Tape& tp, int s, int k, const float* vb) {
    isf_t h = r->h;
    hipStream_t st = h->stream;
    const GridInfo& g = r->g;
    const int it = s * r->K + k + 1;
    const float4* psi5 = psi0_of(r, tp, s, k + 1);
    const JetParams j = jet_at(r, it);
    // the velocity adjoint also forms the second projection's gauge cotangent (output psi5)
    LAUNCH_CHECK(launch_oneform_adj_vel(g, r->psibar, psi5, vb, h->hbar, 1, st, r->f), "velocity adjoint");
    // :58 second projection (gauge output psi5, one-form at psi4 = restrict(psi3))
    int e = poisson(r, r->f, r->fbar);
    if (e) return e;
    // and :57 jet constraint, with the first projection's gauge cotangent (output psi3)
    LAUNCH_CHECK(launch_oneform_adj_div_restrict(g, r->psibar, tp.psi3[k], tp.q2[k], r->fbar, j, r->partials,
                                                 r->vacc + 4 * (size_t)(it - 1), r->f, st),
                 "projection and restrict adjoint");
    // :50 first projection (gauge output psi3, one-form at psi2 = normalize(psi1)) and
    // :49 normalize
    if ((e = poisson(r, r->f, r->fbar))) return e;
    LAUNCH_CHECK(launch_oneform_adj_div(g, r->psibar, tp.psi1[k], tp.q1[k], r->fbar, st, PRE_NORMALIZE),
                 "projection and normalize adjoint");
    // :48 schroedinger_flow
    LAUNCH_CHECK(launch_flow_x_fwd(g, r->psibar, h->S, h->tb, st), "flow adjoint x forward");
    LAUNCH_CHECK(launch_flow_plane_adj(g, h->S, h->tb, st), "flow adjoint plane");
    LAUNCH_CHECK(launch_flow_x_inv(g, h->S, r->psibar, false, h->tb, st), "flow adjoint x inverse");
    return ISF_OK;
}

extern "C" {

int isf_rollout_create(isf_rollout_t* out, isf_t h, const isf_rollout_desc* desc) {
    CHECK_PTR(out, "out");
    CHECK_PTR(desc, "desc");
    *out = nullptr;
    {
        int e = check_handle(h);
        if (e) return e;
    }
    if (!h->jet_set) return fail(ISF_EINVAL, "isf_jet_configure must be called first");
    const isf_rollout_desc& d = *desc;
    if (d.steps < 1) return fail(ISF_EINVAL, "steps must be >= 1");
    if (d.checkpoint_every < 1) return fail(ISF_EINVAL, "checkpoint_every must be >= 1");
    if (d.constraint_iters < 0) return fail(ISF_EINVAL, "constraint_iters must be >= 0");
    const int n3[3] = {h->g.nx, h->g.ny, h->g.nz};
    const bool no_phase = d.phase_res[0] == 0 && d.phase_res[1] == 0 && d.phase_res[2] == 0;
    for (int a = 0; a < 3 && !no_phase; a++)
        if (!pow2(d.phase_res[a]) || d.phase_res[a] > n3[a])
            return fail(ISF_EINVAL, "phase_res[%d] = %d must be a power of two <= %d", a, d.phase_res[a], n3[a]);
    for (int a = 0; a < 3; a++)
        if (d.splat_res[a] < 1 || d.splat_res[a] > 1024) return fail(ISF_EINVAL, "splat_res[%d] out of range", a);
    if (!(d.splat_sigma > 0.0f) || !(d.splat_cutoff > 0.0f)) return fail(ISF_EINVAL, "splat sigma and cutoff must be > 0");
    if (d.nslots < 0 || d.n_init < 0 || d.n_init > d.nslots) return fail(ISF_EINVAL, "bad particle counts");

    isf_rollout* r = new isf_rollout();
    r->magic = kRolloutMagic;
    r->h = h;
    r->d = d;
    r->g = h->g;
    r->N = h->N;
    r->P = d.nslots;
    r->T = d.steps;
    r->K = std::min(d.checkpoint_every, d.steps);
    r->C = d.constraint_iters;
    r->nseg = (r->T + r->K - 1) / r->K;
    r->M = d.splat_res[0] * d.splat_res[1] * d.splat_res[2];
    for (int a = 0; a < 3; a++) {
        r->sc.res[a] = d.splat_res[a];
        r->sc.L[a] = a == 0 ? h->g.Lx : a == 1 ? h->g.Ly : h->g.Lz;
        r->sc.h[a] = r->sc.L[a] / (float)d.splat_res[a];
    }
    r->sc.sigma = d.splat_sigma;
    r->sc.R = d.splat_cutoff * d.splat_sigma;
    r->sc.inv_n = d.nslots > 0 ? 1.0f / (float)d.nslots : 0.0f;
    r->use_theta = false;
    auto bail = [&](int code) {
        rollout_free(r);
        return code;
    };
    const size_t N = r->N, P3 = 3 * (size_t)r->P;
    int e = 0;
#define RA(ptr, count)                      \
    do {                                    \
        if ((e = ralloc(r, ptr, count))) return bail(e); \
    } while (0)
    r->cpsi.resize(r->C + 1);
    r->cq.resize(r->C);
    for (auto& p : r->cpsi) RA(&p, N);
    for (auto& p : r->cq) RA(&p, N);
    RA(&r->theta, no_phase ? 1 : (size_t)d.phase_res[0] * d.phase_res[1] * d.phase_res[2]);
    RA(&r->parts_init, P3);
    r->ck_psi.resize(r->nseg + 1);
    r->ck_psi[0] = r->cpsi[r->C];
    for (int s = 1; s <= r->nseg; s++) RA(&r->ck_psi[s], N);
    if (r->P > 0) {
        r->traj.resize((size_t)r->T + 1);
        for (auto& p : r->traj) RA(&p, P3);
    }
    for (Tape& tp : r->tape) {
        tp.mid.resize(r->K - 1);
        for (auto& p : tp.mid) RA(&p, N);
        tp.psi1.resize(r->K);
        tp.psi3.resize(r->K);
        tp.q1.resize(r->K);
        tp.q2.resize(r->K);
        tp.vel.resize(r->P > 0 ? r->K : 0);
        for (int k = 0; k < r->K; k++) {
            RA(&tp.psi1[k], N);
            RA(&tp.psi3[k], N);
            RA(&tp.q1[k], N);
            RA(&tp.q2[k], N);
            if (r->P > 0) RA(&tp.vel[k], 3 * N);
        }
    }
    RA(&r->rw.S, 2 * N);
    RA(&r->rw.f, N);
    r->rw.parts = nullptr;  // the recompute steps no particles
    RA(&r->psibar, N);
    RA(&r->f, N);
    RA(&r->fbar, N);
    for (int i = 0; i < 2; i++) {
        r->vbar[i].resize(r->K);
        for (auto& p : r->vbar[i]) RA(&p, 3 * N);
    }
    if (r->P > 0)
        r->parts = r->traj[r->T];  // the forward's final state
    else
        RA(&r->parts, P3);
    RA(&r->parts_final, P3);
    RA(&r->pbar, P3);
    RA(&r->phibar, N);
    RA(&r->target, (size_t)r->M);
    RA(&r->resid, (size_t)r->M);
    RA(&r->rho, (size_t)r->M);
    RA(&r->partials, std::max(restrict_adj_partials(r->g), splat_loss_blocks(r->M)));
    RA(&r->loss_dev, 1);
    RA(&r->vacc, 4 * (size_t)(r->T + r->C));
    r->spawn_n = h->n_spawn;
    if (r->P > 0 && r->spawn_n > 0 && r->spawn_n <= r->P) RA(&r->spawn_tab, 3 * (size_t)r->spawn_n * r->T);
    if (const size_t sp = splat_partial_doubles(r->sc)) RA(&r->splat_part, sp);
    if (r->P > 0) {
        r->adj_order_bytes = cell_order_scratch_bytes(r->g, r->P);
        RA(&r->adj_order, (size_t)r->P);
        RA(&r->adj_order_scratch, r->adj_order_bytes);
    }
#undef RA
    {
        // every rollout event orders one of the rollout's streams after another on this
        // device; none is waited for by the host, so none needs a system-scope release
        constexpr unsigned kStreamEv = hipEventDisableTiming | hipEventDisableSystemFence;
        hipError_t he = hipStreamCreate(&r->rw.st);
        for (int i = 0; i < 2 && he == hipSuccess; i++) {
            he = hipEventCreateWithFlags(&r->rec_done[i], kStreamEv);
            if (he == hipSuccess) he = hipEventCreateWithFlags(&r->adj_done[i], kStreamEv);
        }
        if (he == hipSuccess) he = hipEventCreateWithFlags(&r->start, kStreamEv);
        r->pst = h->pstream;
        for (int i = 0; i < 2 && he == hipSuccess; i++) {
            r->pev[i].assign(r->K, nullptr);
            for (int k = 0; k < r->K && he == hipSuccess; k++)
                he = hipEventCreateWithFlags(&r->pev[i][k], kStreamEv);
            if (he == hipSuccess) he = hipEventCreateWithFlags(&r->pdone[i], kStreamEv);
            if (he == hipSuccess) he = hipEventCreateWithFlags(&r->fseg[i], kStreamEv);
        }
        if (he == hipSuccess) he = hipEventCreateWithFlags(&r->fstep, kStreamEv);
        if (he != hipSuccess) return bail(hip_fail(he, "rollout streams"));
    }
    hipStream_t st = h->stream;
    if ((e = join(h))) return bail(e);
    if (launch_fill_f32(r->target, 0.0f, (size_t)r->M, st)) return bail(fail(ISF_EHIP, "fill"));
    if (r->P > 0) {
        if (launch_fill_f32(r->parts_init, -1.0f, P3, st)) return bail(fail(ISF_EHIP, "fill"));
        if (d.n_init > 0 && launch_fill_cylinder(r->parts_init, r->P, 0, d.n_init, h->jet, d.fill_seed, st))
            return bail(fail(ISF_EHIP, "fill_nozzle"));
    }
    const hipError_t he = hipStreamSynchronize(st);
    if (he != hipSuccess) return bail(hip_fail(he, "rollout set-up"));
    *out = r;
    return ISF_OK;
}

int isf_rollout_destroy(isf_rollout_t r) {
    CHECK_R(r);
    rollout_free(r);
    return ISF_OK;
}

int isf_rollout_set_target(isf_rollout_t r, const float* target) {
    CHECK_R(r);
    CHECK_PTR(target, "target");
    int e = join(r->h);
    if (e) return e;
    HIP_CHECK(hipMemcpyAsync(r->target, target, sizeof(float) * r->M, hipMemcpyDeviceToDevice, r->h->stream),
              "target copy");
    r->forward_done = false;
    return ISF_OK;
}

int isf_rollout_forward(isf_rollout_t r, const float velocity[3], const float* theta, const float* psi_init,
                        double* loss) {
    CHECK_R(r);
    CHECK_PTR(velocity, "velocity");
    CHECK_PTR(loss, "loss");
    const bool no_phase = r->d.phase_res[0] == 0;
    if (theta && no_phase) return fail(ISF_EINVAL, "this rollout has no phase field (phase_res = 0)");
    isf_t h = r->h;
    int e = join(h);
    if (e) return e;
    hipStream_t st = h->stream;
    r->forward_done = false;
    if ((e = ensure_spawn_table(r, st))) return e;
    // kvec = v / hbar, omega = sum(v.^2) / (2 hbar) in Float32 (test/runtests.jl:23-24)
    for (int a = 0; a < 3; a++) r->v[a] = velocity[a];
    r->jet = h->jet;
    r->jet.kx = velocity[0] / h->hbar;
    r->jet.ky = velocity[1] / h->hbar;
    r->jet.kz = velocity[2] / h->hbar;
    r->omega = driver_omega(velocity, h->hbar);
    // set-up: psi_init (default (1, 0.01) normalized, test/runtests.jl:19-21), phase field
    if (psi_init) {
        HIP_CHECK(hipMemcpyAsync(r->cpsi[0], psi_init, sizeof(float4) * r->N, hipMemcpyDeviceToDevice, st), "psi_init copy");
    } else {
        LAUNCH_CHECK(launch_fill_psi(r->cpsi[0], make_float4(1.0f, 0.0f, 0.01f, 0.0f), r->N, st), "fill");
        LAUNCH_CHECK(launch_normalize(r->g, r->cpsi[0], st), "normalize");
    }
    r->use_theta = theta != nullptr;
    if (theta) {
        const size_t nt = (size_t)r->d.phase_res[0] * r->d.phase_res[1] * r->d.phase_res[2];
        HIP_CHECK(hipMemcpyAsync(r->theta, theta, sizeof(float) * nt, hipMemcpyDeviceToDevice, st), "theta copy");
        LAUNCH_CHECK(launch_apply_phase(r->g, r->cpsi[0], r->theta, r->d.phase_res, st), "apply phase");
    }
    // test/runtests.jl:27-30
    const Work mw = main_work(r);
    // cpsi[j + 1] = project(restrict(cpsi[j])); each projection's gauge pass also writes
    // the next iteration's restricted input
    const JetParams j0 = jet_at(r, 0);
    if (r->C > 0)
        LAUNCH_CHECK(launch_restrict_to(r->g, r->cpsi[0], r->cpsi[1], j0, st), "restrict_velocity");
    for (int j = 0; j < r->C; j++)
        if ((e = project_q(r, r->cpsi[j + 1], r->cq[j], mw, j + 1 < r->C ? r->cpsi[j + 2] : nullptr, &j0))) return e;
    // the particle part on the particle stream, after the set-up above
    HIP_CHECK(hipEventRecord(r->start, st), "hipEventRecord");
    HIP_CHECK(hipStreamWaitEvent(r->pst, r->start, 0), "hipStreamWaitEvent");
    for (int s = 0; s < r->nseg; s++)
        if ((e = seg_forward(r, s, r->tape[s % 2], mw, true, r->pst))) return e;
    HIP_CHECK(hipStreamWaitEvent(st, r->fseg[(r->nseg - 1) % 2], 0), "hipStreamWaitEvent");
    if (r->P > 0)
        HIP_CHECK(hipMemcpyAsync(r->parts_final, r->parts, sizeof(float) * 3 * r->P, hipMemcpyDeviceToDevice, st),
                  "copy");
    LAUNCH_CHECK(launch_fill_f64(r->rho, 0.0, (size_t)r->M, st), "fill");
    LAUNCH_CHECK(launch_splat(r->parts_final, r->P, r->sc, r->rho, r->splat_part, st), "splat");
    LAUNCH_CHECK(launch_splat_loss(r->rho, r->target, r->M, r->resid, r->partials, r->loss_dev, st), "loss");
    double sum = 0.0;
    HIP_CHECK(hipMemcpyAsync(&sum, r->loss_dev, sizeof(double), hipMemcpyDeviceToHost, st), "loss copy");
    HIP_CHECK(hipStreamSynchronize(st), "hipStreamSynchronize");
    *loss = sum / (double)r->M;  // mean over the splat cells
    r->forward_done = true;
    r->last_tape = true;  // segment s went to tape set s % 2: the backward's first set
    return ISF_OK;
}

int isf_rollout_backward(isf_rollout_t r, double grad_velocity[3], float* grad_theta, float* grad_psi_init,
                         float* grad_parts_init) {
    CHECK_R(r);
    if (!r->forward_done) return fail(ISF_EINVAL, "isf_rollout_forward must run first");
    if (grad_theta && !r->use_theta) return fail(ISF_EINVAL, "the last forward had no phase field");
    isf_t h = r->h;
    int e = join(h);
    if (e) return e;
    hipStream_t st = h->stream;
    const GridInfo& g = r->g;
    LAUNCH_CHECK(launch_fill_psi(r->psibar, make_float4(0.0f, 0.0f, 0.0f, 0.0f), r->N, st), "fill");
    LAUNCH_CHECK(launch_fill_f64(r->vacc, 0.0, 4 * (size_t)(r->T + r->C), st), "fill");
    LAUNCH_CHECK(launch_splat_adj(r->parts_final, r->P, r->sc, r->resid, r->pbar, st), "loss adjoint");
    // segments last to first: segment s is recomputed into tape set s % 2 on the recompute
    // stream (after the adjoint of segment s + 2 has released that set) and reversed on
    // the handle's stream, so the recompute of s - 1 overlaps the reverse of s
    // The particle part of segment s runs on the particle stream after that recompute (and
    // after the reverse of segment s + 2 has read the cotangent buffers of the set); each
    // step's psi part waits for its cotangent.
    HIP_CHECK(hipEventRecord(r->start, st), "hipEventRecord");
    HIP_CHECK(hipStreamWaitEvent(r->rw.st, r->start, 0), "hipStreamWaitEvent");
    HIP_CHECK(hipStreamWaitEvent(r->pst, r->start, 0), "hipStreamWaitEvent");
    // steps [k0, k1) of segment s's recompute (k1 < 0: to the end)
    auto recompute = [&](int s, int k0 = 0, int k1 = -1) -> int {
        const int b = s % 2;
        if (k0 == 0 && s + 2 < r->nseg) {
            HIP_CHECK(hipStreamWaitEvent(r->rw.st, r->adj_done[b], 0), "hipStreamWaitEvent");
            HIP_CHECK(hipStreamWaitEvent(r->rw.st, r->pdone[b], 0), "hipStreamWaitEvent");
        }
        int e2 = seg_forward(r, s, r->tape[b], r->rw, false, nullptr, k0, k1);
        if (e2) return e2;
        if (k1 < 0 || k1 >= seg_steps(r, s)) HIP_CHECK(hipEventRecord(r->rec_done[b], r->rw.st), "hipEventRecord");
        return ISF_OK;
    };
    // the last segment's tape is still there after a forward (not after a backward, whose
    // recomputes reuse the set)
    if (r->last_tape) {
        HIP_CHECK(hipEventRecord(r->rec_done[(r->nseg - 1) % 2], st), "hipEventRecord");
        r->last_tape = false;
    } else if ((e = recompute(r->nseg - 1))) {
        return e;
    }
    // The host enqueues a segment's particle steps, its psi steps and the next recompute's
    // steps interleaved, one step of each at a time: the streams then start their work
    // together (enqueued one after another, the recompute's ~60 launches came before the
    // segment's first psi step and the psi chain idled for them at every segment).
    for (int s = r->nseg - 1; s >= 0; s--) {
        const int b = s % 2, ks = seg_steps(r, s), kr = s > 0 ? seg_steps(r, s - 1) : 0;
        HIP_CHECK(hipStreamWaitEvent(r->pst, r->rec_done[b], 0), "hipStreamWaitEvent");
        if (s + 2 < r->nseg) HIP_CHECK(hipStreamWaitEvent(r->pst, r->adj_done[b], 0), "hipStreamWaitEvent");
        HIP_CHECK(hipStreamWaitEvent(st, r->rec_done[b], 0), "hipStreamWaitEvent");
        for (int i = 0; i < ks || i < kr; i++) {
            const int k = ks - 1 - i;
            if (k >= 0) {
                const bool sort = r->P < kSortPerSegment || k == ks - 1;
                if ((e = step_adjoint_particles(r, r->tape[b], s, k, r->vbar[b][k], sort))) return e;
                HIP_CHECK(hipEventRecord(r->pev[b][k], r->pst), "hipEventRecord");
                if (k == 0) HIP_CHECK(hipEventRecord(r->pdone[b], r->pst), "hipEventRecord");
            }
            if (i < kr && (e = recompute(s - 1, i, i + 1))) return e;
            if (k >= 0) {
                HIP_CHECK(hipStreamWaitEvent(st, r->pev[b][k], 0), "hipStreamWaitEvent");
                if ((e = step_adjoint(r, r->tape[b], s, k, r->vbar[b][k]))) return e;
            }
        }
        HIP_CHECK(hipEventRecord(r->adj_done[b], st), "hipEventRecord");
    }
    HIP_CHECK(hipStreamWaitEvent(st, r->pdone[0], 0), "hipStreamWaitEvent");  // pbar complete
    // set-up constraint iterations in reverse
    // (each restrict adjoint also forms the next iteration's gauge cotangent: the gauge
    // output of iteration j - 1 is cpsi[j])
    if (r->C > 0) LAUNCH_CHECK(launch_gauge_qbar(g, r->psibar, r->cpsi[r->C], r->f, st), "gauge adjoint");
    for (int j = r->C - 1; j >= 0; j--) {
        const JetParams j0 = jet_at(r, 0);
        if ((e = poisson(r, r->f, r->fbar))) return e;
        LAUNCH_CHECK(launch_oneform_adj_div_restrict(g, r->psibar, r->cpsi[j], r->cq[j], r->fbar, j0, r->partials,
                                                     r->vacc + 4 * (size_t)(r->T + j), j > 0 ? r->f : nullptr, st),
                     "projection and restrict adjoint");
    }
    if (r->use_theta) {
        LAUNCH_CHECK(launch_apply_phase_adj(g, r->psibar, r->cpsi[0], r->theta, r->phibar, r->d.phase_res, st),
                     "phase adjoint");
        if (grad_theta) LAUNCH_CHECK(launch_phase_gather(g, r->phibar, grad_theta, r->d.phase_res, st), "phase gather");
    }
    if (grad_psi_init)
        HIP_CHECK(hipMemcpyAsync(grad_psi_init, r->psibar, sizeof(float4) * r->N, hipMemcpyDeviceToDevice, st), "copy");
    if (grad_parts_init && r->P > 0)
        HIP_CHECK(hipMemcpyAsync(grad_parts_init, r->pbar, sizeof(float) * 3 * r->P, hipMemcpyDeviceToDevice, st), "copy");
    std::vector<double> acc(4 * (size_t)(r->T + r->C));
    HIP_CHECK(hipMemcpyAsync(acc.data(), r->vacc, sizeof(double) * acc.size(), hipMemcpyDeviceToHost, st), "copy");
    HIP_CHECK(hipStreamSynchronize(st), "hipStreamSynchronize");
    if (grad_velocity) {
        // phi = k . pos - omega t (mode 0) or sum(k .+ pos) - omega t (mode 1), k = v / hbar,
        // omega = |v|^2 / (2 hbar):  d phi / d v_a = (pos_a or 1 - t v_a) / hbar
        const double hb = (double)h->hbar;
        double gv[3] = {0.0, 0.0, 0.0};
        for (int slot = 0; slot < r->T + r->C; slot++) {
            const double* S = acc.data() + 4 * slot;
            const double t = slot < r->T ? (double)((float)(slot + 1) * h->dt) : 0.0;
            for (int a = 0; a < 3; a++) {
                const double lin = r->jet.phase_mode == 1 ? S[0] : S[1 + a];
                gv[a] += (lin - t * (double)r->v[a] * S[0]) / hb;
            }
        }
        for (int a = 0; a < 3; a++) grad_velocity[a] = gv[a];
    }
    return ISF_OK;
}

int isf_rollout_state(isf_rollout_t r, const float** psi, const float** parts, const double** rho) {
    CHECK_R(r);
    if (!r->forward_done) return fail(ISF_EINVAL, "isf_rollout_forward must run first");
    if (psi) *psi = reinterpret_cast<const float*>(r->ck_psi[r->nseg]);
    if (parts) *parts = r->parts_final;
    if (rho) *rho = r->rho;
    return ISF_OK;
}


// ------------------------------------------------------------------ operator VJPs

}  // extern "C"

// Scratch of the operator VJPs (psi copy, q, f, fbar), allocated on first use.
static int vjp_scratch(isf_t h) {
    if (h->vjp_mem) return ISF_OK;
    const hipError_t e = hipMalloc(&h->vjp_mem, 16 * h->N + 12 * h->N + (4 + restrict_adj_partials(h->g)) * 8);
    if (e != hipSuccess) return fail(ISF_ENOMEM, "vjp scratch: %s", hipGetErrorString(e));
    return ISF_OK;
}
static float4* vjp_psi(isf_t h) { return static_cast<float4*>(h->vjp_mem); }
static float* vjp_q(isf_t h) { return reinterpret_cast<float*>(static_cast<char*>(h->vjp_mem) + 16 * h->N); }
static float* vjp_f(isf_t h) { return vjp_q(h) + h->N; }
static float* vjp_fbar(isf_t h) { return vjp_q(h) + 2 * h->N; }
static double* vjp_partials(isf_t h) {
    return reinterpret_cast<double*>(static_cast<char*>(h->vjp_mem) + 28 * h->N);
}
static float2* hP(isf_t h) { return reinterpret_cast<float2*>(reinterpret_cast<char*>(h->S) + 4 * h->N); }

extern "C" {

int isf_vjp_schroedinger_flow(isf_t h, float* psibar) {
    int e = check_handle(h);
    if (!e) e = join(h);
    if (e) return e;
    CHECK_PTR(psibar, "psibar");
    float4* b = reinterpret_cast<float4*>(psibar);
    LAUNCH_CHECK(launch_flow_x_fwd(h->g, b, h->S, h->tb, h->stream), "flow adjoint x forward");
    LAUNCH_CHECK(launch_flow_plane_adj(h->g, h->S, h->tb, h->stream), "flow adjoint plane");
    LAUNCH_CHECK(launch_flow_x_inv(h->g, h->S, b, false, h->tb, h->stream), "flow adjoint x inverse");
    return ISF_OK;
}

int isf_vjp_normalize(isf_t h, const float* psi_in, float* psibar) {
    int e = check_handle(h);
    if (!e) e = join(h);
    if (e) return e;
    CHECK_PTR(psi_in, "psi_in");
    CHECK_PTR(psibar, "psibar");
    LAUNCH_CHECK(launch_normalize_adj(h->g, reinterpret_cast<float4*>(psibar), reinterpret_cast<const float4*>(psi_in),
                                      h->stream),
                 "normalize adjoint");
    return ISF_OK;
}

int isf_vjp_pressure_project(isf_t h, const float* psi_in, float* psibar) {
    int e = check_handle(h);
    if (!e) e = join(h);
    if (!e) e = vjp_scratch(h);
    if (e) return e;
    CHECK_PTR(psi_in, "psi_in");
    CHECK_PTR(psibar, "psibar");
    hipStream_t st = h->stream;
    const double sc = -0.25 * 2.0 / (double)h->N;
    const float4* in = reinterpret_cast<const float4*>(psi_in);
    float4* b = reinterpret_cast<float4*>(psibar);
    float4* out = vjp_psi(h);
    // forward again, keeping q and the output
    HIP_CHECK(hipMemcpyAsync(out, in, 16 * h->N, hipMemcpyDeviceToDevice, st), "copy");
    LAUNCH_CHECK(launch_div_r2c(h->g, out, vjp_f(h), hP(h), h->tb, st), "div(one_form(psi)) + poisson x forward");
    LAUNCH_CHECK(launch_poisson_plane(h->g, hP(h), sc, h->tb, st), "poisson plane");
    LAUNCH_CHECK(launch_c2r_x(h->g, hP(h), 1, vjp_q(h), out, nullptr, nullptr, h->tb, st), "poisson x inverse + gauge");
    // reverse
    LAUNCH_CHECK(launch_gauge_qbar(h->g, b, out, vjp_f(h), st), "gauge adjoint");
    LAUNCH_CHECK(launch_r2c_x(h->g, vjp_f(h), hP(h), h->tb, st), "poisson x forward");
    LAUNCH_CHECK(launch_poisson_plane(h->g, hP(h), sc, h->tb, st), "poisson plane");
    LAUNCH_CHECK(launch_c2r_x(h->g, hP(h), 0, vjp_fbar(h), nullptr, nullptr, nullptr, h->tb, st), "poisson x inverse");
    LAUNCH_CHECK(launch_oneform_adj_div(h->g, b, in, vjp_q(h), vjp_fbar(h), st), "projection adjoint");
    return ISF_OK;
}

int isf_vjp_restrict_velocity(isf_t h, const float* psi_in, float* psibar, const float kvec[3], float omgterm,
                              int phase_mode, double dphase[4]) {
    int e = check_handle(h);
    if (!e) e = join(h);
    if (!e) e = vjp_scratch(h);
    if (e) return e;
    CHECK_PTR(psi_in, "psi_in");
    CHECK_PTR(psibar, "psibar");
    CHECK_PTR(kvec, "kvec");
    if (!h->jet_set) return fail(ISF_EINVAL, "isf_jet_configure must be called first (nozzle)");
    if (phase_mode != ISF_PHASE_DOT && phase_mode != ISF_PHASE_LEGACY) return fail(ISF_EINVAL, "bad phase_mode %d", phase_mode);
    JetParams j = h->jet;
    j.kx = kvec[0];
    j.ky = kvec[1];
    j.kz = kvec[2];
    j.omgterm = omgterm;
    j.phase_mode = phase_mode;
    double* acc = vjp_partials(h);
    double* parts = acc + 4;
    LAUNCH_CHECK(launch_fill_f64(acc, 0.0, 4, h->stream), "fill");
    LAUNCH_CHECK(launch_restrict_adj(h->g, reinterpret_cast<float4*>(psibar), reinterpret_cast<const float4*>(psi_in), j,
                                     parts, acc, h->stream),
                 "restrict adjoint");
    if (dphase) {
        HIP_CHECK(hipMemcpyAsync(dphase, acc, 4 * sizeof(double), hipMemcpyDeviceToHost, h->stream), "copy");
        HIP_CHECK(hipStreamSynchronize(h->stream), "hipStreamSynchronize");
    }
    return ISF_OK;
}

int isf_vjp_velocity_one_form(isf_t h, const float* psi, float hbar, int sharp, const float* velbar, float* psibar) {
    int e = check_handle(h);
    if (!e) e = join(h);
    if (e) return e;
    CHECK_PTR(psi, "psi");
    CHECK_PTR(velbar, "velbar");
    CHECK_PTR(psibar, "psibar");
    LAUNCH_CHECK(launch_oneform_adj_vel(h->g, reinterpret_cast<float4*>(psibar), reinterpret_cast<const float4*>(psi),
                                        velbar, hbar, sharp, h->stream),
                 "velocity adjoint");
    return ISF_OK;
}

int isf_vjp_staggered_advect(isf_t h, const float* parts0, long long nslots, const float* vel, float* pbar,
                             float* velbar) {
    int e = check_handle(h);
    if (!e) e = join(h);
    if (e) return e;
    CHECK_PTR(parts0, "parts0");
    CHECK_PTR(pbar, "pbar");
    CHECK_PTR(velbar, "velbar");
    if (nslots < 0) return fail(ISF_EINVAL, "bad nslots");
    if (nslots == 0) return ISF_OK;
    if (nslots > h->vjp_order_cap) {
        if (h->vjp_order) (void)hipFree(h->vjp_order);
        if (h->vjp_order_scratch) (void)hipFree(h->vjp_order_scratch);
        h->vjp_order = nullptr;
        h->vjp_order_scratch = nullptr;
        h->vjp_order_cap = 0;
        const size_t sb = cell_order_scratch_bytes(h->g, nslots);
        hipError_t he = hipMalloc(&h->vjp_order, sizeof(int) * (size_t)nslots);
        if (he == hipSuccess) he = hipMalloc(&h->vjp_order_scratch, sb);
        if (he != hipSuccess) return fail(ISF_ENOMEM, "advection VJP order: %s", hipGetErrorString(he));
        h->vjp_order_bytes = sb;
        h->vjp_order_cap = nslots;
    }
    // cell-sorted slot order: long same-cell runs for the scatter's wavefront reduction
    LAUNCH_CHECK(launch_cell_order(h->g, parts0, nslots, h->vjp_order, h->vjp_order_scratch,
                                             h->vjp_order_bytes, h->stream),
                 "particle order");
    LAUNCH_CHECK(launch_advect_adj(h->g, parts0, nslots, h->vjp_order, vel ? vel : h->vel, h->dt, pbar, velbar,
                                   h->stream),
                 "advect adjoint");
    return ISF_OK;
}

}  // extern "C"
