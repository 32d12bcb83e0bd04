/*
 * isf.h -- C ABI of libisf, the MI355X (gfx950) implementation of the incompressible
 * Schroedinger flow (ISF) step of SchroedingersSmoke.jl.
 *
 * Every entry point replaces one function of the reference's generic/GPU path; the
 * reference file:line is given beside it (paths relative to the SchroedingersSmoke.jl
 * repository root).  Conventions:
 *   - Return value: 0 on success, a negative ISF_E* code on failure; the message is
 *     available from isf_last_error() (thread-local).  Nothing aborts.
 *   - Array arguments are DEVICE pointers (from isf_malloc or any HIP allocation on the
 *     current device) in the reference's layouts, x fastest (column-major):
 *       psi       4 floats per cell  {re1, im1, re2, im2}   (ComplexF32, ComplexF32)
 *       velocity  3 floats per cell  {vx, vy, vz}           NTuple{3,Float32}
 *       f, q      1 float per cell
 *       isjet     1 byte per cell
 *       particles 3 floats per slot  {x, y, z}              NTuple{3,Float32}
 *   - All work is enqueued asynchronously on the handle's HIP stream (isf_step also uses a
 *     second internal stream, ordered by events); isf_synchronize waits for both
 *     (GPUArrays.synchronize, test/runtests.jl:65).
 *   - One host thread per handle.
 */
#ifndef ISF_H
#define ISF_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ISF_OK 0
#define ISF_EINVAL (-1)      /* bad argument (null pointer, unsupported size, ...) */
#define ISF_EHIP (-2)        /* HIP runtime error */
#define ISF_ENOMEM (-3)      /* device allocation failed */
#define ISF_EHANDLE (-4)     /* null or destroyed handle */

/* phase_mode values of the jet constraint */
#define ISF_PHASE_DOT 0      /* phase = k.x - omega t (src/MatlabPort driver, examples/matlabport/simple.jl:47,107) */
#define ISF_PHASE_LEGACY 1   /* phase = sum(k .+ x) - omega t (reference tuple_dot, src/particles.jl:1,15) */

typedef struct isf_handle* isf_t;

/* ---- library / memory -------------------------------------------------------- */
const char* isf_last_error(void);
const char* isf_version(void);
int isf_device_count(int* out);
int isf_malloc(void** ptr, size_t bytes);
int isf_free(void* ptr);
/* The copy/set helpers are synchronous and first wait for all work queued on the device. */
int isf_memcpy_h2d(void* dst, const void* src, size_t bytes);  /* ArrayType(hostarray), test/runtests.jl:19 */
int isf_memcpy_d2h(void* dst, const void* src, size_t bytes);  /* Array(devarray), examples/jet.jl:83 */
int isf_memcpy_d2d(void* dst, const void* src, size_t bytes);
int isf_memset(void* dst, int value, size_t bytes);

/* ---- ISF instance ------------------------------------------------------------- */
/* ISF{ArrayType,UInt32,Float32}(physical_size, dims, hbar, dt), src/parallel.jl:41-98.
 * res: grid_res (powers of two; x, y and z in 16..512; batch members y and z up to 128);
 * box: physical_size (integers, src/parallel.jl:12,45).  Builds the Schroedinger mask
 * and Poisson factor (:47-63) as per-axis FP64-derived tables, the FFT twiddles and
 * the scratch arrays. */
int isf_create(isf_t* out, const int res[3], const int box[3], float hbar, float dt);
int isf_destroy(isf_t h);
int isf_synchronize(isf_t h);                                         /* test/runtests.jl:65 */
/* Fields the drivers read (test/runtests.jl:25,44,59-63): d, dt, hbar, grid_res. */
int isf_get_grid(isf_t h, int res[3], float d[3], float box[3], float* hbar, float* dt);
/* isf.velocity (src/parallel.jl:14): device pointer of the handle's velocity field
 * (3 floats per cell), written by isf_step and by isf_velocity_one_form(..., NULL). */
int isf_velocity_ptr(isf_t h, float** out);

/* ---- reference operators -------------------------------------------------------- */
/* schroedinger_flow!(isf, psi), src/parallel.jl:176-187 (in place). */
int isf_schroedinger_flow(isf_t h, float* psi);
/* psi .= normalize_psi.(psi), src/parallel.jl:162-165, test/runtests.jl:49. */
int isf_normalize(isf_t h, float* psi);
/* velocity_one_form!(isf, psi, hbar), src/parallel.jl:100-122; sharp != 0 also applies
 * the driver's in-place StaggeredSharp v .*= 1 ./ d (test/runtests.jl:61-62).
 * vel == NULL writes the handle's own velocity field. */
int isf_velocity_one_form(isf_t h, const float* psi, float hbar, int sharp, float* vel);
/* StaggeredSharp in place, velocity .* inv.(isf.d) (test/runtests.jl:61-62,
 * src/MatlabPort.jl:99-104); vel NULL = the handle's velocity field. */
int isf_staggered_sharp(isf_t h, float* vel);
/* div!(isf), src/parallel.jl:131-152: f from a velocity field (vel == NULL: handle's). */
int isf_div(isf_t h, const float* vel, float* f);
/* poisson_solve(isf), src/parallel.jl:168-174: q = ifft(fac .* fft(f)).  q is returned
 * real (1 float per cell); the reference's imaginary part is FFT rounding noise. */
int isf_poisson_solve(isf_t h, const float* f, float* q);
/* psi .= gauge_transform.(psi, q), src/parallel.jl:125-129,159 with real q. */
int isf_gauge_transform(isf_t h, float* psi, const float* q);
/* pressure_project!(isf, psi), src/parallel.jl:155-160. */
int isf_pressure_project(isf_t h, float* psi);

/* isjet.(isf.positions, cen, len, rad), src/particles.jl:6-10, test/runtests.jl:25. */
int isf_isjet(isf_t h, const float cen[3], float len, float rad, unsigned char* out);
/* restrict_velocity!(isf, psi, kvec, isjet, omgterm), src/particles.jl:12-27.
 * isjet == NULL evaluates the nozzle predicate of isf_jet_configure analytically. */
int isf_restrict_velocity(isf_t h, float* psi, const unsigned char* isjet, const float kvec[3], float omgterm,
                          int phase_mode);
/* add_particles!(particles, first+1:first+n, cen, rad), src/particles.jl:34-94.
 * 0-based slots first..first+n-1 (wrapped modulo nslots).  Draw j comes from
 * Philox4x32-10 with counter (j, step, 0, 0) and key = seed; each chunk of 1024 draws
 * fills its slots in ring-angle order (neighbouring slots, neighbouring particles). */
int isf_add_particles(isf_t h, float* parts, long long nslots, long long first, long long n, const float cen[3],
                      float rad, unsigned long long seed, unsigned step);
/* staggered_advect!(particles, isf), src/parallel.jl:190-261 (vel == NULL: handle's). */
int isf_staggered_advect(isf_t h, float* parts, long long nslots, const float* vel);

/* ---- fused step (the simloop body, test/runtests.jl:45-64) -------------------------- */
/* Jet parameters of the driver (test/runtests.jl:7-12,23-25): kvec = v / hbar,
 * omega = |v|^2 / (2 hbar), nozzle centre/length/radius, particles spawned per step,
 * spawn seed and phase mode. */
int isf_jet_configure(isf_t h, const float kvec[3], float omega, const float cen[3], float len, float rad,
                      long long n_spawn, unsigned long long seed, int phase_mode);
/* Fill slots first..first+n-1 uniformly inside the configured nozzle cylinder (initial
 * particle cloud of the at-scale workloads; no reference counterpart).  Draw i comes from
 * Philox4x32-10 with counter (i, 0xFFFFFFFF, 1, 0); draws are stored in increasing order
 * of a 24-bit key of their raw words (x slab, theta wedge, r) so neighbouring slots hold
 * neighbouring particles.  Synchronous. */
int isf_fill_nozzle(isf_t h, float* parts, long long nslots, long long first, long long n, unsigned long long seed);
/* The 10 constraint iterations of the setup, test/runtests.jl:27-30. */
int isf_constrain(isf_t h, float* psi, int iterations);
/* One simulation step number iter (1-based, t = iter*dt):
 * flow, normalize, project, spawn, restrict, project, one-form(hbar) * 1/d, RK4 advect. */
int isf_step(isf_t h, float* psi, float* parts, long long nslots, int iter);
/* simloop(N, isf, psi, ...) (test/runtests.jl:40-66, examples/benchmark.jl:41-67): steps
 * iter0 .. iter0 + count - 1, the same results as count calls of isf_step (bit for bit),
 * pipelined across steps (each step's velocity pass is fused into the next step's forward
 * flow pass).  Asynchronous like isf_step; it may block the host until the GPU is at most
 * about two steps behind. */
int isf_simloop(isf_t h, float* psi, float* parts, long long nslots, int iter0, int count);
/* isf_step measures how scattered the slot order of the particle buffer is (when the
 * buffer changes and every 256 steps); if it is scattered, slots are advected in a
 * spatial order (slots grouped by cell) recomputed every `steps` steps (default 32;
 * 0 = always slot order).  Each slot is advanced independently, so the order never
 * changes a result.  No reference counterpart. */
int isf_set_sort_interval(isf_t h, int steps);
/* FP64 sum over cells of |psi1|^2 + |psi2|^2 (no reference counterpart; health check). */
int isf_norm2(isf_t h, const float* psi, double* out);
/* The same sum accumulated by the last isf_step (read after isf_synchronize). */
int isf_step_norm2(isf_t h, double* out);
/* The sum of isf_norm2 without allocating: no counterpart in the reference; for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops).
 * psi: ncells cells (1 .. ISF_NOALLOC_MAX_CELLS) of the interleaved layout; out: one
 * double in device memory; stream: a hipStream_t (NULL: the null stream).  Asynchronous,
 * needs no handle.  Its partial sums use one fixed 8 KiB device array per GPU, shared by
 * every call on that GPU: calls that overlap on one GPU race. */
#define ISF_NOALLOC_MAX_CELLS (1LL << 30)
int isf_norm2_noalloc(const float* psi, long long ncells, double* out, void* stream);

/* ---- MatlabPort particle store (src/MatlabPort.jl:178-230) ----------------------------- */
/* Particles{x, y, z, length, active}: Float32 structure-of-arrays positions with
 * `capacity` slots, the number of slots filled so far (length) and the active slot list.
 * Device-resident; every call is ordered on the handle's stream. */
typedef struct isf_particles* isf_particles_t;
int isf_particles_create(isf_particles_t* out, long long capacity);   /* Particles(zeros(...), 0, Int[]), simple.jl:74-78 */
int isf_particles_destroy(isf_particles_t p);
/* append!(p, x, y, z) (src/MatlabPort.jl:186-195): n device positions into slots length..;
 * like the reference, refuses unless length + n < capacity */
int isf_particles_append(isf_t h, isf_particles_t p, const float* x, const float* y, const float* z, long long n);
/* the driver's births (examples/matlabport/simple.jl:113-118): n particles on the ring of
 * radius 0.9 rad around the configured nozzle at x = c_x; angle from Philox4x32-10
 * counter (j, step, 2, 0) and key seed; appended */
int isf_particles_spawn_ring(isf_t h, isf_particles_t p, long long n, unsigned long long seed, unsigned step);
/* removethedead!(p, vol_size) (simple.jl:63-70): keeps the active slots strictly inside
 * the box, in order */
int isf_particles_remove_dead(isf_t h, isf_particles_t p);
/* StaggeredAdvect(p, isf, vx, vy, vz, dt) (src/MatlabPort.jl:201-222) over the active
 * slots; vel NULL = the handle's velocity field */
int isf_particles_advect(isf_t h, isf_particles_t p, const float* vel);
/* Keep(p, ind) (src/MatlabPort.jl:226-230): x, y, z become x[ind], y[ind], z[ind] (ind:
 * device, 0-based, n entries): slots 0..n-1 hold the gathered positions, the slots beyond
 * the end of the reference's n-long arrays hold -1, and append! then checks against n.
 * length and the active list are unchanged, as in the reference. */
int isf_particles_keep(isf_t h, isf_particles_t p, const int* ind, long long n);
/* filter!(p.active) of the driver's removethedead! (examples/matlabport/simple.jl:63-70),
 * or any other host-side edit of the active list: the list becomes the n slot indices in
 * `active` (host memory, 0-based, each below the length of the position arrays);
 * synchronous. */
int isf_particles_set_active(isf_t h, isf_particles_t p, const int* active, long long n);
/* length, active count (synchronises), device views of x, y, z and the active list */
int isf_particles_state(isf_t h, isf_particles_t p, long long* length, long long* active_count, const float** x,
                        const float** y, const float** z, const int** active);

/* ---- MatlabPort operators (src/MatlabPort.jl:87-273) on that module's layout ------------ */
/* psi1, psi2: separate complex arrays (2 floats per cell, re/im); vx, vy, vz, f: separate
 * real arrays; q: complex; all x fastest, N = nx*ny*nz cells.  Float32 arithmetic with the
 * cell formulas of the operators above.  Outputs may be the inputs (in place) except
 * where noted; results are returned in new arrays by the reference, in out* here. */
/* SchroedingerFlow(obj, psi1, psi2), src/MatlabPort.jl:87-93 */
int isf_mp_schroedinger_flow(isf_t h, const float* psi1, const float* psi2, float* out1, float* out2);
/* Normalize(psi1, psi2), src/MatlabPort.jl:155-160 */
int isf_mp_normalize(isf_t h, const float* psi1, const float* psi2, float* out1, float* out2);
/* VelocityOneForm(obj, psi1, psi2, hbar), src/MatlabPort.jl:106-126 (hbar 1 in the
 * reference's default); vx, vy, vz must not overlap psi1, psi2 */
int isf_mp_velocity_one_form(isf_t h, const float* psi1, const float* psi2, float hbar, float* vx, float* vy,
                             float* vz);
/* StaggeredSharp(obj, vx, vy, vz) = (vx/dx, vy/dy, vz/dz), src/MatlabPort.jl:99-104 */
int isf_mp_staggered_sharp(isf_t h, const float* vx, const float* vy, const float* vz, float* ux, float* uy,
                           float* uz);
/* Div(obj, vx, vy, vz), src/MatlabPort.jl:162-170; f must not overlap vx, vy, vz */
int isf_mp_div(isf_t h, const float* vx, const float* vy, const float* vz, float* f);
/* PoissonSolve(obj, f), src/MatlabPort.jl:128-138: q complex with imaginary part 0 (the
 * reference's is FFT rounding noise of a real solve) */
int isf_mp_poisson_solve(isf_t h, const float* f, float* q);
/* GaugeTransform(psi1, psi2, q) = psi * exp(i q) for complex q, src/MatlabPort.jl:140-145 */
int isf_mp_gauge_transform(isf_t h, const float* psi1, const float* psi2, const float* q, float* out1, float* out2);
/* PressureProject(obj, psi1, psi2), src/MatlabPort.jl:147-153 (q = -PoissonSolve(Div(
 * VelocityOneForm(psi))), psi exp(i q)); stages psi in a 16 N byte buffer of the handle,
 * allocated on first use */
int isf_mp_pressure_project(isf_t h, const float* psi1, const float* psi2, float* out1, float* out2);
/* The MatlabPort driver's jet constraint (examples/matlabport/simple.jl:104-110), in place:
 * psi_a[isJet] = |psi_a[isJet]| exp(i phase); phase = k.x - omgterm (phase_mode 0, the
 * driver's form) or sum(k + x) - omgterm (1, src/particles.jl:1); isjet NULL = the nozzle
 * of isf_jet_configure */
int isf_mp_restrict_velocity(isf_t h, float* psi1, float* psi2, const unsigned char* isjet, const float kvec[3],
                              float omgterm, int phase_mode);
/* StaggeredVelocity(px, py, pz, torus, vx, vy, vz) at n points, src/MatlabPort.jl:237-273 */
int isf_mp_staggered_velocity(isf_t h, const float* px, const float* py, const float* pz, long long n, const float* vx,
                              const float* vy, const float* vz, float* ux, float* uy, float* uz);
/* StaggeredAdvect(p, torus, vx, vy, vz, dt) over the active slots of a particle store,
 * src/MatlabPort.jl:201-222 */
int isf_mp_staggered_advect(isf_t h, isf_particles_t p, const float* vx, const float* vy, const float* vz, float dt);

/* ---- asynchronous particle export (examples/jet.jl:83, Array(particles) per frame) ------ */
/* Double-buffered copies of the particle positions (nslots x 3 floats) into pinned host
 * memory.  push stages the positions after all work queued on the handle, on the stream
 * where isf_step advects particles, and copies them to the host on a separate stream, so
 * later steps wait neither for the staging copy nor for the PCIe transfer.  previous returns push k-1 after push k (waiting for its copy if needed);
 * latest waits for the last push.  A returned buffer stays valid until the next push. */
typedef struct isf_export* isf_export_t;
int isf_export_create(isf_export_t* out, isf_t h, long long nslots);
int isf_export_destroy(isf_export_t e);
int isf_export_push(isf_export_t e, const float* parts);
int isf_export_previous(isf_export_t e, const float** host, long long* frame);
int isf_export_latest(isf_export_t e, const float** host, long long* frame);

/* ---- batched uneven ensembles (no reference counterpart) ------------------------------- */
/* One member of an ensemble: the arguments of isf_create and isf_jet_configure. */
typedef struct {
    int res[3];
    int box[3];
    float hbar, dt;
    float kvec[3];
    float omega;
    float cen[3];
    float len, rad;
    long long n_spawn;
    unsigned long long seed;
    int phase_mode;
} isf_member_desc;
typedef struct isf_batch* isf_batch_t;
/* Create n independent ISF instances (grids may differ). */
int isf_batch_create(isf_batch_t* out, int n, const isf_member_desc* members);
int isf_batch_destroy(isf_batch_t b);
/* The handle of member i (owned by the batch), for set-up and single operators. */
int isf_batch_member(isf_batch_t b, int i, isf_t* out);
/* Device buffers member i steps: psi, particles (nslots slots). */
int isf_batch_bind(isf_batch_t b, int i, float* psi, float* parts, long long nslots);
/* One step `iter` (1-based) of every member, the simloop body of each.  Members of the
 * same grid shape share one launch per kernel stage; shape groups run as parallel
 * branches of one hipGraph captured on first use and replayed each step (set_graph 0:
 * launched directly).  Spawned particles are written in slot order. */
int isf_batch_step(isf_batch_t b, int iter);
/* Steps iter0 .. iter0 + count - 1 of every member, the driver's simloop over the
 * ensemble (test/runtests.jl:40-66, one member per instance of its loop): equal to count
 * isf_batch_step calls, bit for bit.  Each later step's forward flow pass also writes the
 * previous step's velocity (two buffers per member, chosen by the step number), so the
 * velocity pass leaves the step; one graph for the first step and one for every later
 * step, captured on first use. */
int isf_batch_simloop(isf_batch_t b, int iter0, int count);
int isf_batch_set_graph(isf_batch_t b, int use_graph);
/* The same step as n separate isf_step calls, one member after another (baseline). */
int isf_batch_step_sequential(isf_batch_t b, int iter);
int isf_batch_synchronize(isf_batch_t b);
/* Pick the FFT schedule of every grid shape of the batch by timing it on scratch data:
 * the x-pass tile (y lines per block: 8, 16, 32) and, for planes of at most 64 x 64, the
 * plane pass's elements per thread (8: radix-8 passes, 16: radix-16 passes).  Results are
 * read from / appended to the text file cache_path when it is not NULL, one line per
 * shape: "nx ny nz members tile [plane_elements]". */
int isf_batch_tune(isf_batch_t b, const char* cache_path);
/* The tile chosen for a grid shape (0 when none was tuned). */
int isf_batch_tile(isf_batch_t b, const int res[3], int* ty);
/* The plane pass's elements per thread in use for a grid shape (8 or 16). */
int isf_batch_plane_elems(isf_batch_t b, const int res[3], int* pe);

/* ---- vector-Jacobian products of the operators (no reference counterpart) ------------- */
/* Each maps the cotangent of an operator's output to the cotangent of its input, under
 * the real inner product <a, b> = sum Re(conj(a) b) over cells and components; the
 * primal input is passed where the Jacobian depends on it.  In place on the cotangent
 * unless stated. */
int isf_vjp_schroedinger_flow(isf_t h, float* psibar);                        /* src/parallel.jl:176-187 */
int isf_vjp_normalize(isf_t h, const float* psi_in, float* psibar);          /* src/parallel.jl:162-165 */
int isf_vjp_pressure_project(isf_t h, const float* psi_in, float* psibar);   /* src/parallel.jl:155-160 */
/* restrict_velocity with the configured nozzle (src/particles.jl:12-27); dphase (host,
 * may be NULL) receives sum(phibar) and sum(phibar * pos_a), phibar the cotangent of the
 * jet phase of each jet cell: dL/dk_a = sum(phibar pos_a) (phase_mode 0) or sum(phibar)
 * (1), dL/d omgterm = -sum(phibar). */
int isf_vjp_restrict_velocity(isf_t h, const float* psi_in, float* psibar, const float kvec[3], float omgterm,
                              int phase_mode, double dphase[4]);
/* velocity_one_form (src/parallel.jl:115-122, with 1/d if sharp): psibar += J^T velbar */
int isf_vjp_velocity_one_form(isf_t h, const float* psi, float hbar, int sharp, const float* velbar, float* psibar);
/* staggered_advect (src/parallel.jl:190-261) at positions parts0: pbar (cotangent of the
 * advected positions) becomes that of parts0; velbar += the velocity cotangent. */
int isf_vjp_staggered_advect(isf_t h, const float* parts0, long long nslots, const float* vel, float* pbar,
                             float* velbar);

/* ---- differentiable rollout (no reference counterpart) -------------------------------- */
/* A rollout is the driver's set-up and simloop (test/runtests.jl:16-66) on a handle whose
 * jet is configured (isf_jet_configure: nozzle, spawn count, seed, phase mode), made a
 * function of differentiable inputs:
 *   velocity  the nozzle velocity v (kvec = v / hbar, omega = |v|^2 / (2 hbar),
 *             test/runtests.jl:23-24),
 *   theta     a coarse initial phase field: psi0 = psi_init * exp(i Phi), Phi its periodic
 *             trilinear upsampling to the grid (cell x at theta coordinate x pr / n),
 *   psi_init  the initial wavefunction (default: (1, 0.01) normalized, :19-21),
 *   particles the initial slot positions (n_init slots filled uniformly in the nozzle).
 * Set-up: constraint_iters x (restrict with omega t = 0, pressure_project), :27-30; then
 * `steps` simloop bodies; loss = mean over the splat grid of (rho - target)^2, rho(y) =
 * (1 / nslots) sum over live slots of exp(-|y - p|^2 / (2 sigma^2)) for |y - p| <=
 * cutoff sigma, y the centres of splat_res cells over the box.
 * The backward pass runs the hand-written adjoint of every operator in reverse, with psi
 * checkpointed every checkpoint_every steps and each segment's psi chain recomputed; the
 * forward keeps the whole particle trajectory ((steps + 1) x 12 nslots bytes), which the
 * backward reads. */
typedef struct {
    int steps;                 /* simloop bodies per rollout, >= 1 */
    int checkpoint_every;      /* K >= 1; K >= steps keeps the whole tape of one segment */
    int constraint_iters;      /* >= 0 */
    int phase_res[3];          /* theta grid (powers of two, each <= the grid); 0,0,0: none */
    int splat_res[3];          /* density grid, each 1..1024 */
    float splat_sigma;         /* physical units, > 0 */
    float splat_cutoff;        /* in units of sigma, > 0 */
    long long nslots;          /* particle slots, >= 0 */
    long long n_init;          /* slots filled in the nozzle cylinder at set-up, <= nslots */
    unsigned long long fill_seed;
} isf_rollout_desc;
typedef struct isf_rollout* isf_rollout_t;
int isf_rollout_create(isf_rollout_t* out, isf_t h, const isf_rollout_desc* desc);
int isf_rollout_destroy(isf_rollout_t r);
/* target density (device, splat_res cells, x fastest); copied */
int isf_rollout_set_target(isf_rollout_t r, const float* target);
/* velocity: host float[3]; theta: device (phase_res cells) or NULL (zero); psi_init:
 * device psi or NULL (the default).  Writes the loss (FP64) after synchronising.  The
 * spawn batches of all steps are drawn once, at the first forward and again after the
 * handle's nozzle or seed has changed (isf_jet_configure). */
int isf_rollout_forward(isf_rollout_t r, const float velocity[3], const float* theta, const float* psi_init,
                        double* loss);
/* Gradients of the last forward's loss; any output may be NULL.  grad_velocity: host
 * double[3]; grad_theta: device (phase_res cells); grad_psi_init: device psi layout;
 * grad_parts_init: device (nslots x 3).  Synchronous.  Besides the handle's stream it uses
 * two streams of the rollout (segment recompute, particle adjoint), ordered by events. */
int isf_rollout_backward(isf_rollout_t r, double grad_velocity[3], float* grad_theta, float* grad_psi_init,
                         float* grad_parts_init);
/* Device views of the last forward's final psi, particles and density (rho, FP64). */
int isf_rollout_state(isf_rollout_t r, const float** psi, const float** parts, const double** rho);

/* ---- timing helpers (HIP events on the handle's stream) ----------------------------- */
/* Number and names of the kernel stages of isf_step, in launch order. */
int isf_step_stage_count(void);
const char* isf_step_stage_name(int i);
/* isf_step with a HIP event between consecutive stages; waits for the step and writes
 * the duration of every stage (ms) to stage_ms[0 .. isf_step_stage_count()-1]. */
int isf_step_profile(isf_t h, float* psi, float* parts, long long nslots, int iter, float* stage_ms);
int isf_timer_start(isf_t h);
int isf_timer_stop(isf_t h, float* ms);

#ifdef __cplusplus
}
#endif
#endif /* ISF_H */
