/*
 * isf_oracle_impl.h -- body of the CPU oracle for the ISF step, included twice by
 * isf_oracle.c: once with R=float (the reference's ComplexF32 / Float32 path, used as
 * the CPU baseline) and once with R=double (the FP64 accuracy authority).
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or calls
 * this code; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
 * it, and only as the checker / the timed CPU baseline.
 *
 * Every function restates one function of the reference (paths relative to the
 * SchroedingersSmoke.jl repository root) and cites the file:line it follows.
 *
 * Layout (the reference's Julia column-major arrays, x fastest):
 *   cell index c = x + nx*(y + ny*z)
 *   psi      : 4 R per cell {re1, im1, re2, im2}    (test/runtests.jl:19, (ComplexF32,ComplexF32))
 *   velocity : 3 R per cell {vx, vy, vz}            (src/parallel.jl:14,66, NTuple{3,Float32})
 *   f        : 1 R per cell                         (src/parallel.jl:29,65)
 *   q, fc    : 2 R per cell {re, im}                 (src/parallel.jl:30,169)
 *   particles: 3 R per slot                         (test/runtests.jl:33)
 */

#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(name, SUF)

/* ---------------------------------------------------------------- math helpers */
/* Julia mod(x, y) for floating point: result has the sign of y
 * (Base.mod; used by src/parallel.jl:220 mod.(point, ph_size)). */
static inline R FN(jl_mod)(R x, R y) {
    R r = FN(r_fmod)(x, y);
    if (r != (R)0 && ((r < 0) != (y < 0))) r += y;
    return r;
}

/* ---------------------------------------------------------------- 1-D FFT */
/* Iterative radix-2 decimation-in-time FFT on a contiguous line of n complex values
 * (n a power of two).  sign = -1 forward (unnormalised), +1 backward (unnormalised).
 * Twiddles are formed in double and rounded to R.  This restates the transform that
 * FFTW computes for plan_fft!/plan_ifft! (src/parallel.jl:79); the 1/N of ifft is
 * applied by the 3-D driver below. */
static void FN(fft1d)(R* re, R* im, int n, int sign, const R* twr, const R* twi) {
    /* bit reversal permutation */
    for (int i = 1, j = 0; i < n; i++) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            R t = re[i]; re[i] = re[j]; re[j] = t;
            t = im[i]; im[i] = im[j]; im[j] = t;
        }
    }
    for (int len = 2; len <= n; len <<= 1) {
        int half = len >> 1, step = n / len;
        for (int i = 0; i < n; i += len) {
            for (int k = 0; k < half; k++) {
                R wr = twr[k * step], wi = -(R)sign * twi[k * step];  /* twi holds sin(-2 pi k/n) */
                R ar = re[i + k + half], ai = im[i + k + half];
                R br = ar * wr - ai * wi, bi = ar * wi + ai * wr;
                re[i + k + half] = re[i + k] - br; im[i + k + half] = im[i + k] - bi;
                re[i + k] += br; im[i + k] += bi;
            }
        }
    }
}

/* 3-D complex FFT in place on an interleaved array z[2*c + {0,1}] (c = cell index).
 * inverse != 0 applies the backward transform scaled by 1/N, the AbstractFFTs
 * convention of plan_ifft! (src/parallel.jl:79,172,184). */
void FN(isfo_fft3)(R* z, const int* n, int inverse) {
    if (isfo_fft_hook) {
        isfo_fft_hook(z, n, inverse, sizeof(R) == sizeof(double));
        return;
    }
    const int nx = n[0], ny = n[1], nz = n[2];
    const long N = (long)nx * ny * nz;
    const int sign = inverse ? 1 : -1;
    for (int ax = 0; ax < 3; ax++) {
        const int len = n[ax];
        const long stride = ax == 0 ? 1 : (ax == 1 ? nx : (long)nx * ny);
        const long nlines = N / len;
        R* twr = (R*)malloc(sizeof(R) * len);
        R* twi = (R*)malloc(sizeof(R) * len);
        for (int k = 0; k < len; k++) {
            double a = -2.0 * M_PI * (double)k / (double)len;
            twr[k] = (R)cos(a); twi[k] = (R)sin(a);
        }
#pragma omp parallel
        {
            R* lr = (R*)malloc(sizeof(R) * len);
            R* li = (R*)malloc(sizeof(R) * len);
#pragma omp for schedule(static)
            for (long l = 0; l < nlines; l++) {
                long base;
                if (ax == 0) base = l * nx;
                else if (ax == 1) base = (l % nx) + (l / nx) * (long)nx * ny;
                else base = l;
                for (int e = 0; e < len; e++) {
                    lr[e] = z[2 * (base + e * stride)];
                    li[e] = z[2 * (base + e * stride) + 1];
                }
                FN(fft1d)(lr, li, len, sign, twr, twi);
                for (int e = 0; e < len; e++) {
                    z[2 * (base + e * stride)] = lr[e];
                    z[2 * (base + e * stride) + 1] = li[e];
                }
            }
            free(lr); free(li);
        }
        free(twr); free(twi);
    }
    if (inverse) {
        const R s = (R)(1.0 / (double)N);
#pragma omp parallel for schedule(static)
        for (long c = 0; c < 2 * N; c++) z[c] *= s;
    }
}

/* ---------------------------------------------------------------- ISF tables */
/* Poisson factor and Schroedinger mask, src/parallel.jl:47-63.
 * The loop arithmetic is Float64 (pi, Int and /2 promote) and is stored as
 * Complex{FloatType}; mask is fftshift-ed (:62) so it is in FFT order, fac[1,1,1]=0 (:63).
 * hbar and dt are the FloatType values promoted to double, as in the reference where
 * they arrive as Float32 arguments.  Output arrays: fac (2 R/cell), mask (2 R/cell). */
void FN(isfo_tables)(const int* n, const int* L, double hbar, double dt, R* fac, R* mask) {
    const int nx = n[0], ny = n[1], nz = n[2];
    double d[3];
    for (int a = 0; a < 3; a++) d[a] = (double)(R)((double)L[a] / (double)n[a]);  /* :46 */
    const double f = -4.0 * M_PI * M_PI * hbar;                                    /* :49 */
#pragma omp parallel for schedule(static)
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) {
                const int xyz[3] = {x, y, z};  /* 0-based = Julia (xyz .- 1) */
                double denom = 0.0, ksq = 0.0;
                for (int a = 0; a < 3; a++) {
                    double s = sin(M_PI * xyz[a] / n[a]) / d[a];                   /* :53 */
                    denom += s * s;                                                /* :54 */
                    double k = (xyz[a] - n[a] / 2.0) / (double)L[a];               /* :58 */
                    ksq += k * k;
                }
                const long c = x + (long)nx * (y + (long)ny * z);
                fac[2 * c] = (R)(-0.25 / denom);                                   /* :55 */
                fac[2 * c + 1] = 0;
                /* fftshift (:62): the unshifted value at index i lands at (i + n/2) mod n */
                const long xs = (x + nx / 2) % nx, ys = (y + ny / 2) % ny, zs = (z + nz / 2) % nz;
                const long cs = xs + (long)nx * (ys + (long)ny * zs);
                const double ph = f * ksq * dt / 2.0;                              /* :59-60 */
                mask[2 * cs] = (R)cos(ph);
                mask[2 * cs + 1] = (R)sin(ph);
            }
    fac[0] = 0; fac[1] = 0;                                                        /* :63 */
}

/* ---------------------------------------------------------------- pointwise ops */
/* normalize_psi, src/parallel.jl:162-165 (applied as psi .= normalize_psi.(psi),
 * test/runtests.jl:49): psi_a <- psi_a / hypot(|psi1|, |psi2|). */
void FN(isfo_normalize)(R* psi, long N) {
#pragma omp parallel for schedule(static)
    for (long c = 0; c < N; c++) {
        R* p = psi + 4 * c;
        R a1 = FN(r_hypot)(p[0], p[1]), a2 = FN(r_hypot)(p[2], p[3]);
        R nrm = FN(r_hypot)(a1, a2);
        p[0] /= nrm; p[1] /= nrm; p[2] /= nrm; p[3] /= nrm;
    }
}

/* schroedinger_flow!, src/parallel.jl:176-187: psi_a <- ifft(mask .* fft(psi_a)),
 * a = 1, 2, through the SoA scratch arrays psi1/psi2 (:178-185). */
void FN(isfo_schroedinger_flow)(R* psi, const R* mask, const int* n) {
    const long N = (long)n[0] * n[1] * n[2];
    R* s = (R*)malloc(sizeof(R) * 4 * N);
    for (int comp = 0; comp < 2; comp++) {
        R* z = s + 2 * N * comp;
#pragma omp parallel for schedule(static)
        for (long c = 0; c < N; c++) {                                   /* :179-180 */
            z[2 * c] = psi[4 * c + 2 * comp];
            z[2 * c + 1] = psi[4 * c + 2 * comp + 1];
        }
        FN(isfo_fft3)(z, n, 0);                                          /* :181 */
#pragma omp parallel for schedule(static)
        for (long c = 0; c < N; c++) {                                   /* :182-183 */
            R ar = z[2 * c], ai = z[2 * c + 1], br = mask[2 * c], bi = mask[2 * c + 1];
            z[2 * c] = ar * br - ai * bi;
            z[2 * c + 1] = ar * bi + ai * br;
        }
        FN(isfo_fft3)(z, n, 1);                                          /* :184 */
#pragma omp parallel for schedule(static)
        for (long c = 0; c < N; c++) {                                   /* :185 */
            psi[4 * c + 2 * comp] = z[2 * c];
            psi[4 * c + 2 * comp + 1] = z[2 * c + 1];
        }
    }
    free(s);
}

/* velocity_one_form / velocity_one_form!, src/parallel.jl:100-122:
 * v_a(c) = hbar * angle(conj(psi1(c)) psi1(c+e_a) + conj(psi2(c)) psi2(c+e_a)),
 * periodic forward neighbour (index table i_shifted, :69-71).  If sharp != 0 the
 * driver's in-place StaggeredSharp (test/runtests.jl:61-62) v_a *= 1/d_a is applied. */
void FN(isfo_velocity_one_form)(const R* psi, const int* n, const R* d, R hbar, int sharp, R* vel) {
    const int nx = n[0], ny = n[1], nz = n[2];
    R dinv[3];
    for (int a = 0; a < 3; a++) dinv[a] = (R)1 / d[a];   /* inv.(isf.d) in Float32 */
#pragma omp parallel for schedule(static)
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) {
                const long c = x + (long)nx * (y + (long)ny * z);
                const long nb[3] = {
                    ((x + 1) % nx) + (long)nx * (y + (long)ny * z),
                    x + (long)nx * (((y + 1) % ny) + (long)ny * z),
                    x + (long)nx * (y + (long)ny * ((z + 1) % nz))};
                const R* p = psi + 4 * c;
                for (int a = 0; a < 3; a++) {
                    const R* q = psi + 4 * nb[a];
                    /* conj(p1)*q1 + conj(p2)*q2 */
                    R re = (p[0] * q[0] + p[1] * q[1]) + (p[2] * q[2] + p[3] * q[3]);
                    R im = (p[0] * q[1] - p[1] * q[0]) + (p[2] * q[3] - p[3] * q[2]);
                    R v = FN(r_atan2)(im, re) * hbar;
                    if (sharp) v = v * dinv[a];
                    vel[3 * c + a] = v;
                }
            }
}

/* div / div!, src/parallel.jl:131-152: f(c) = sum_a (v_a(c) - v_a(c - e_a)) / d_a^2,
 * periodic backward neighbour (index table i_shifted2, :72-74), ds = inv.(d.^2). */
void FN(isfo_div)(const R* vel, const int* n, const R* d, R* f) {
    const int nx = n[0], ny = n[1], nz = n[2];
    R ds[3];
    for (int a = 0; a < 3; a++) ds[a] = (R)1 / (d[a] * d[a]);
#pragma omp parallel for schedule(static)
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) {
                const long c = x + (long)nx * (y + (long)ny * z);
                const long mx = ((x + nx - 1) % nx) + (long)nx * (y + (long)ny * z);
                const long my = x + (long)nx * (((y + ny - 1) % ny) + (long)ny * z);
                const long mz = x + (long)nx * (y + (long)ny * ((z + nz - 1) % nz));
                R t0 = (vel[3 * c + 0] - vel[3 * mx + 0]) * ds[0];
                R t1 = (vel[3 * c + 1] - vel[3 * my + 1]) * ds[1];
                R t2 = (vel[3 * c + 2] - vel[3 * mz + 2]) * ds[2];
                f[c] = t0 + t1 + t2;
            }
}

/* poisson_solve, src/parallel.jl:168-174: q = ifft(fac .* fft(complex(f))).
 * Output q is complex (2 R/cell), as the reference returns isf.fc. */
void FN(isfo_poisson_solve)(const R* f, const R* fac, const int* n, R* q) {
    const long N = (long)n[0] * n[1] * n[2];
#pragma omp parallel for schedule(static)
    for (long c = 0; c < N; c++) { q[2 * c] = f[c]; q[2 * c + 1] = 0; }        /* :169 */
    FN(isfo_fft3)(q, n, 0);                                                     /* :170 */
#pragma omp parallel for schedule(static)
    for (long c = 0; c < N; c++) {                                              /* :171 */
        R ar = q[2 * c], ai = q[2 * c + 1], br = fac[2 * c], bi = fac[2 * c + 1];
        q[2 * c] = ar * br - ai * bi;
        q[2 * c + 1] = ar * bi + ai * br;
    }
    FN(isfo_fft3)(q, n, 1);                                                     /* :172 */
}

/* gauge_transform, src/parallel.jl:125-129: psi_a <- psi_a * exp(i*(-q)) with complex q,
 * i.e. exp(Im q) * (cos(Re q) - i sin(Re q)). */
void FN(isfo_gauge_transform)(R* psi, const R* q, long N) {
#pragma omp parallel for schedule(static)
    for (long c = 0; c < N; c++) {
        R xr = q[2 * c + 1], xi = -q[2 * c];       /* x = i*(-q) = Im(q) - i Re(q) */
        R e = FN(r_exp)(xr);
        R er = e * FN(r_cos)(xi), ei = e * FN(r_sin2)(xi);
        R* p = psi + 4 * c;
        R a = p[0], b = p[1];
        p[0] = a * er - b * ei; p[1] = a * ei + b * er;
        a = p[2]; b = p[3];
        p[2] = a * er - b * ei; p[3] = a * ei + b * er;
    }
}

/* pressure_project!, src/parallel.jl:155-160 (hbar = 1 inside, :115,156).
 * Scratch: vel (3N), f (N), q (2N) supplied by the caller. */
void FN(isfo_pressure_project)(R* psi, const R* fac, const int* n, const R* d,
                               R* vel, R* f, R* q) {
    const long N = (long)n[0] * n[1] * n[2];
    FN(isfo_velocity_one_form)(psi, n, d, (R)1, 0, vel);
    FN(isfo_div)(vel, n, d, f);
    FN(isfo_poisson_solve)(f, fac, n, q);
    FN(isfo_gauge_transform)(psi, q, N);
}

/* isjet, src/particles.jl:6-10, evaluated at isf.positions = (i-1)*d (src/parallel.jl:75-77),
 * as test/runtests.jl:25 does. out[c] = 1 inside the cylindrical nozzle. */
void FN(isfo_isjet)(const int* n, const R* d, const R* cen, R len, R rad, unsigned char* out) {
    const int nx = n[0], ny = n[1], nz = n[2];
#pragma omp parallel for schedule(static)
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) {
                const long c = x + (long)nx * (y + (long)ny * z);
                R px = (R)x * d[0], py = (R)y * d[1], pz = (R)z * d[2];
                R dy = py - cen[1], dz = pz - cen[2];
                out[c] = (FN(r_fabs)(px - cen[0]) <= len / 2) & (dy * dy + dz * dz <= rad * rad);
            }
}

/* restrict_velocity! + restrict_kernel, src/particles.jl:12-27.
 * Inside the jet: psi_a <- |psi_a| * exp(i*phase); outside unchanged.
 * phase_mode 0 ("dot", MatlabPort form, examples/matlabport/simple.jl:47,107):
 *     phase = k . x - omgterm
 * phase_mode 1 ("legacy_sum", the reference's tuple_dot(a,b) = sum(a .+ b),
 *     src/particles.jl:1,15): phase = sum_a (k_a + x_a) - omgterm */
void FN(isfo_restrict_velocity)(R* psi, const unsigned char* isjet, const int* n, const R* d,
                                const R* kvec, R omgterm, int phase_mode) {
    const int nx = n[0], ny = n[1], nz = n[2];
#pragma omp parallel for schedule(static)
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) {
                const long c = x + (long)nx * (y + (long)ny * z);
                if (!isjet[c]) continue;
                R pos[3] = {(R)x * d[0], (R)y * d[1], (R)z * d[2]};
                R ph;
                if (phase_mode == 1) ph = ((kvec[0] + pos[0]) + (kvec[1] + pos[1])) + (kvec[2] + pos[2]);
                else ph = (kvec[0] * pos[0] + kvec[1] * pos[1]) + kvec[2] * pos[2];
                ph = ph - omgterm;
                R cr = FN(r_cos)(ph), ci = FN(r_sin2)(ph);
                R* p = psi + 4 * c;
                R a1 = FN(r_hypot)(p[0], p[1]), a2 = FN(r_hypot)(p[2], p[3]);
                p[0] = a1 * cr; p[1] = a1 * ci; p[2] = a2 * cr; p[3] = a2 * ci;
            }
}

/* ---------------------------------------------------------------- particles */
/* staggered_velocity, src/parallel.jl:219-246 (same maths as StaggeredVelocity,
 * src/MatlabPort.jl:237-273).  Weights are NOT divided by d (:238), as in the reference.
 * The reference indexes out of bounds when mod(p, L) rounds to L (floor(p/d)+1 = res+1);
 * here that index wraps to 1 so the restatement is defined everywhere. */
static void FN(stag_vel)(const R* vel, const R* pt, const R* d, const R* L, const int* res, R* u) {
    R p[3]; long i[3], ip[3];
    for (int a = 0; a < 3; a++) {
        p[a] = FN(jl_mod)(pt[a], L[a]);                                  /* :220 */
        long ia = (long)FN(r_floor)(p[a] / d[a]);                        /* :221 (0-based) */
        if (ia >= res[a]) ia -= res[a];
        i[a] = ia;
        ip[a] = (ia + 1) % res[a];                                       /* :222 */
    }
    const long nx = res[0], nxy = (long)res[0] * res[1];
#define V(ix, iy, iz, comp) vel[3 * ((ix) + nx * (iy) + nxy * (iz)) + (comp)]
    R v0[3] = {V(i[0], i[1], i[2], 0), V(i[0], i[1], i[2], 1), V(i[0], i[1], i[2], 2)};
    R vn[3] = {V(i[0], ip[1], ip[2], 0), V(ip[0], i[1], ip[2], 1), V(ip[0], ip[1], i[2], 2)};
    R pp[3] = {V(i[0], ip[1], i[2], 0), V(ip[0], i[1], i[2], 1), V(ip[0], i[1], i[2], 2)};
    R pp2[3] = {V(i[0], i[1], ip[2], 0), V(i[0], i[1], ip[2], 1), V(i[0], ip[1], i[2], 2)};
#undef V
    R w[3];
    for (int a = 0; a < 3; a++) w[a] = p[a] - (R)i[a] * d[a];              /* :238 */
    const R w1[3] = {w[2], w[2], w[1]}, w2[3] = {w[1], w[0], w[0]};      /* :239-240 */
    for (int a = 0; a < 3; a++)                                           /* :242-245 */
        u[a] = ((R)1 - w1[a]) * (((R)1 - w2[a]) * v0[a] + w2[a] * pp[a]) +
               w1[a] * (((R)1 - w2[a]) * pp2[a] + w2[a] * vn[a]);
}

/* staggered_advect / staggered_advect! + in_grid, src/parallel.jl:190-216,248-261:
 * classic RK4 for every slot with 0 < p < L (strict), otherwise the slot becomes the
 * dead marker (-1,-1,-1) (:214). */
void FN(isfo_staggered_advect)(R* parts, long P, const R* vel, R dt, const R* d, const R* L,
                               const int* res) {
#pragma omp parallel for schedule(static)
    for (long s = 0; s < P; s++) {
        R* p = parts + 3 * s;
        int in = 1;
        for (int a = 0; a < 3; a++) in &= (p[a] > 0) & (p[a] < L[a]);   /* :190-195 */
        if (!in) { p[0] = p[1] = p[2] = (R)-1; continue; }
        R k1[3], k2[3], k3[3], k4[3], t[3];
        FN(stag_vel)(vel, p, d, L, res, k1);
        for (int a = 0; a < 3; a++) t[a] = p[a] + k1[a] * dt * (R)0.5;   /* :203 */
        FN(stag_vel)(vel, t, d, L, res, k2);
        for (int a = 0; a < 3; a++) t[a] = p[a] + k2[a] * dt * (R)0.5;   /* :206 */
        FN(stag_vel)(vel, t, d, L, res, k3);
        for (int a = 0; a < 3; a++) t[a] = p[a] + k3[a] * dt;            /* :209 */
        FN(stag_vel)(vel, t, d, L, res, k4);
        for (int a = 0; a < 3; a++)                                      /* :212 */
            p[a] = p[a] + dt / (R)6 * (((k1[a] + (R)2 * k2[a]) + (R)2 * k3[a]) + k4[a]);
    }
}

/* add_particles!, src/particles.jl:34-54 (CPU form: slot i for i in range).
 * n particles at (c_x - 0.1, c_y + 0.9 rad cos(rt), c_z + 0.9 rad sin(rt))
 * + (u1,u2,u3)*0.1 - 0.05 with rt = u0 * 2 pi, written to slots first .. first+n-1
 * (0-based, wrapped modulo P).  The reference draws u0..u3 from the unseeded
 * rand(Float32); here draw j of the batch comes from Philox4x32-10 (isf_philox.h) with
 * counter (j, step, 0, 0) and key (seed_lo, seed_hi), and each chunk of 1024 consecutive
 * draws fills the chunk's slots in increasing order of (top 22 bits of raw word 0) << 10
 * | (j mod 1024) (isf_spawn_order in isf_oracle.c). */
void FN(isfo_add_particles)(R* parts, long P, long first, long n, const R* cen, R rad,
                            unsigned long long seed, unsigned int step) {
    for (long c0 = 0; c0 < n; c0 += 1024) {
        const int m = (int)(n - c0 < 1024 ? n - c0 : 1024);
        int rank_to_draw[1024];
        isf_spawn_order((unsigned)c0, m, step, seed, rank_to_draw);
        for (int k = 0; k < m; k++) {
            const long s = (first + c0 + k) % P;
            float u[4];
            isf_philox_uniform4((unsigned)(c0 + rank_to_draw[k]), step, 0u, seed, u);
            R rt = (R)u[0] * (R)2 * (R)M_PI;
            R off[3];
            for (int a = 0; a < 3; a++) off[a] = (R)u[a + 1] * (R)0.1 - (R)0.05;
            R* p = parts + 3 * s;
            p[0] = (cen[0] - (R)0.1) + off[0];
            p[1] = (cen[1] + (R)0.9 * rad * FN(r_cos)(rt)) + off[1];
            p[2] = (cen[2] + (R)0.9 * rad * FN(r_sin2)(rt)) + off[2];
        }
    }
}

/* Initial particle cloud for the at-scale workloads: n particles uniform inside the
 * nozzle cylinder (axis along x, centre cen, length len, radius rad).  No reference
 * counterpart (the reference only spawns on a ring).  Draw i uses Philox4x32-10 with
 * counter (i, 0xFFFFFFFF, 1, 0); x = c_x + len (u0 - 1/2), r = rad sqrt(u1),
 * theta = 2 pi u2.  Draws go to slots first, first+1, ... (mod P) in increasing order of
 * the 24-bit key (x slab, theta wedge, r) built from the top bits of the raw words
 * (isf_fill_order in isf_oracle.c), ties by draw index. */
void FN(isfo_fill_cylinder)(R* parts, long long P, long long first, long long n, const R* cen, R len, R rad,
                            unsigned long long seed) {
    long long* order = (long long*)malloc(sizeof(long long) * (size_t)(n > 0 ? n : 1));
    isf_fill_order(n, seed, order);
    for (long long k = 0; k < n; k++) {
        const long long s = (first + k) % P;
        float u[4];
        isf_philox_uniform4((unsigned)order[k], 0xFFFFFFFFu, 1u, seed, u);
        R r = rad * (R)sqrt((double)u[1]);
        R th = (R)u[2] * (R)2 * (R)M_PI;
        R* p = parts + 3 * s;
        p[0] = cen[0] + len * ((R)u[0] - (R)0.5);
        p[1] = cen[1] + r * FN(r_cos)(th);
        p[2] = cen[2] + r * FN(r_sin2)(th);
    }
    free(order);
}

/* ---------------------------------------------------------------- one step */
/* The simloop body, test/runtests.jl:45-64 (identical in examples/benchmark.jl:46-65),
 * for iteration `iter` (1-based; t = iter*dt, :46).  Spawns n_spawn particles into the
 * ring-buffer slots start..stop of :52-54 (1-based start = mod((iter-1)*n+1, P)).
 * Scratch: vel (3N), f (N), q (2N). */
void FN(isfo_step)(R* psi, R* parts, long P, const R* fac, const R* mask, const unsigned char* isjet,
                   const int* n, const R* d, const R* L, R hbar, R dt, const R* kvec, R omega,
                   const R* cen, R rad, long n_spawn, int iter, unsigned long long seed,
                   int phase_mode, R* vel, R* f, R* q) {
    const long N = (long)n[0] * n[1] * n[2];
    const R t = (R)iter * dt;                                                   /* :46 */
    FN(isfo_schroedinger_flow)(psi, mask, n);                                   /* :48 */
    FN(isfo_normalize)(psi, N);                                                 /* :49 */
    FN(isfo_pressure_project)(psi, fac, n, d, vel, f, q);                       /* :50 */
    if (n_spawn > 0 && P > 0) {
        long start1 = ((long)(iter - 1) * n_spawn + 1) % P;                     /* :52 */
        long first0 = start1 == 0 ? P - 1 : start1 - 1;  /* Julia 1-based -> 0-based */
        FN(isfo_add_particles)(parts, P, first0, n_spawn, cen, rad, seed, (unsigned)iter);
    }
    FN(isfo_restrict_velocity)(psi, isjet, n, d, kvec, omega * t, phase_mode);  /* :57 */
    FN(isfo_pressure_project)(psi, fac, n, d, vel, f, q);                       /* :58 */
    FN(isfo_velocity_one_form)(psi, n, d, hbar, 1, vel);                        /* :59-62 */
    FN(isfo_staggered_advect)(parts, P, vel, dt, d, L, n);                      /* :63 */
}

/* FP64 norm sum_c (|psi1|^2 + |psi2|^2); no reference counterpart (health check). */
double FN(isfo_norm2)(const R* psi, long N) {
    double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (long c = 0; c < N; c++) {
        const R* p = psi + 4 * c;
        s += (double)p[0] * p[0] + (double)p[1] * p[1] + (double)p[2] * p[2] + (double)p[3] * p[3];
    }
    return s;
}

#undef FN
#undef CAT
#undef CAT2
