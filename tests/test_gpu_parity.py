"""Parity of the HIP path (libisf through the C ABI) with the CPU oracle.

Tolerances (the bar for ComplexF32 work is 1e-4 relative L2 on psi against the
reference CPU path):
  * single operators on psi / q / f: rel L2 <= 2e-6 vs the FP64 oracle
  * full steps: rel L2 of psi <= 1e-4 vs the FP64 oracle, with the FP32 oracle's own
    error printed beside it
  * particles: 99th percentile of |dp| <= 1e-4 over slots alive in both, alive-status
    mismatches <= 0.1 % of slots (a particle within ~1e-6 of a cell face can sample a
    different cell on GPU and CPU because interpolation weights are not divided by d,
    src/parallel.jl:238)
  * integer / byte / RNG work (isjet, spawn counters): bit-exact or float-identical.
"""
import numpy as np
import pytest

from oracle.oracle import JetConfig, Oracle
from _harness import GpuRun, oracle_run, particle_report, rel_l2, wrap_diff

pytestmark = pytest.mark.gpu

O64 = Oracle(np.float64)
O32 = Oracle(np.float32)


@pytest.fixture(scope="module")
def isf_mod():
    import isf_amd
    return isf_amd


def rnd(shape, seed):
    return np.random.default_rng(seed).standard_normal(shape).astype(np.float32)


def smooth_psi(n, L, seed=0):
    rng = np.random.default_rng(seed)
    z, y, x = np.meshgrid(*[np.arange(n[a]) * (L[a] / n[a]) for a in (2, 1, 0)], indexing="ij")
    ph1 = 0.3 * np.sin(2 * np.pi * x / L[0]) + 0.2 * np.cos(2 * np.pi * (y / L[1] + z / L[2]))
    ph2 = 0.25 * np.cos(4 * np.pi * x / L[0]) + 0.1 * np.sin(2 * np.pi * z / L[2])
    a1 = 1.0 + 0.1 * rng.random(x.shape)
    a2 = 0.1 + 0.05 * rng.random(x.shape)
    return np.stack([a1 * np.cos(ph1), a1 * np.sin(ph1), a2 * np.cos(ph2), a2 * np.sin(ph2)], -1).astype(np.float32)


def laplacian(q, d):
    """7-point periodic Laplacian with 1/d_a^2 (the operator div(grad) of the projection)."""
    out = np.zeros_like(q)
    for ax, a in ((2, 0), (1, 1), (0, 2)):
        out += (np.roll(q, -1, ax) - 2 * q + np.roll(q, 1, ax)) / d[a] ** 2
    return out


SHAPES = [((16, 16, 16), (2, 2, 2)), ((64, 32, 32), (4, 2, 2)), ((128, 64, 64), (4, 2, 2)),
          ((256, 128, 128), (8, 4, 4)), ((512, 64, 32), (8, 2, 1)), ((32, 128, 16), (2, 4, 1))]


@pytest.mark.parametrize("n,L", SHAPES)
def test_schroedinger_flow(isf_mod, n, L):
    print(f"[flow] n={n}", flush=True)
    isf = isf_mod.ISF(L, n, 0.1, 1 / 48)
    psi = rnd((n[2], n[1], n[0], 4), 1)
    _, mask = O64.tables(n, L, float(np.float32(0.1)), float(np.float32(1 / 48)))
    d = isf.psi_array(psi)
    isf.schroedinger_flow_(d)
    got = d.to_host()
    ref = O64.schroedinger_flow(psi.astype(np.float64), mask)
    err = rel_l2(got, ref)
    print(f"  rel L2 vs FP64 oracle = {err:.3e}")
    assert err < 2e-6


@pytest.mark.parametrize("n,L", SHAPES[:4])
def test_poisson_solve(isf_mod, n, L):
    print(f"[poisson] n={n}", flush=True)
    isf = isf_mod.ISF(L, n, 0.1, 1 / 48)
    f = rnd((n[2], n[1], n[0]), 2)
    fac, _ = O64.tables(n, L, 0.1, 1 / 48)
    q = isf.poisson_solve(isf_mod.DeviceArray.from_host(f)).to_host()
    ref = O64.poisson_solve(f.astype(np.float64), fac)
    err = rel_l2(q, ref[..., 0])
    # q is dominated by low wavenumbers (fac ~ 1/k^2); the discrete Laplacian of q weighs
    # every wavenumber of the result equally, so compare it with f - mean(f) as well
    d = [L[a] / n[a] for a in range(3)]
    lap_gpu = rel_l2(laplacian(q.astype(np.float64), d), f - f.mean())
    q32 = O32.poisson_solve(f, O32.tables(n, L, 0.1, 1 / 48)[0])[..., 0]
    lap_32 = rel_l2(laplacian(q32.astype(np.float64), d), f - f.mean())
    print(f"  rel L2 vs FP64 oracle = {err:.3e}; Laplacian residual {lap_gpu:.3e} (FP32 oracle {lap_32:.3e})")
    assert err < 2e-6
    assert lap_gpu < 3 * lap_32 + 1e-6


@pytest.mark.parametrize("n,L", SHAPES[:3])
def test_pointwise_and_stencils(isf_mod, n, L):
    print(f"[pointwise] n={n}", flush=True)
    isf = isf_mod.ISF(L, n, 0.1, 1 / 48)
    d = np.array(isf.d, np.float64)
    psi = rnd((n[2], n[1], n[0], 4), 3)
    # normalize
    dp = isf.psi_array(psi)
    isf.normalize_(dp)
    assert rel_l2(dp.to_host(), O64.normalize(psi.astype(np.float64))) < 1e-6
    # velocity one-form, hbar and sharp, angles compared modulo 2 pi
    dp = isf.psi_array(psi)
    for hbar, sharp in ((1.0, False), (0.1, True)):
        v = isf.velocity_one_form_(dp, hbar=hbar, sharp=sharp).to_host()
        ref = O64.velocity_one_form(psi.astype(np.float64), d, hbar=np.float32(hbar), sharp=sharp)
        scale = hbar / d if sharp else np.full(3, hbar)
        dv = wrap_diff(v / scale, ref / scale)
        # angle(z) is ill-conditioned where |z| is tiny (random psi); bound the bulk tightly
        assert np.percentile(dv, 99.9) < 2e-6 and np.max(dv) < 1e-3
    # standalone StaggeredSharp: velocity_one_form(hbar) then .* inv.(d) (test/runtests.jl:59-62)
    vs = isf.velocity_one_form_(dp, hbar=0.1, sharp=False).to_host()
    vb = isf_mod.DeviceArray.from_host(vs)
    isf.staggered_sharp_(vb)
    inv_d = (np.float32(1) / np.asarray(d, np.float32)).astype(np.float32)
    assert np.array_equal(vb.to_host(), (vs * inv_d).astype(np.float32))
    # div from a given velocity field
    vel = rnd((n[2], n[1], n[0], 3), 4)
    f = isf_mod.DeviceArray(isf.shape3, np.float32)
    isf.div_(f, isf_mod.DeviceArray.from_host(vel))
    assert rel_l2(f.to_host(), O64.div(vel.astype(np.float64), d)) < 1e-6
    # gauge with a real q
    q = rnd(isf.shape3, 5)
    dp = isf.psi_array(psi)
    isf.gauge_transform_(dp, isf_mod.DeviceArray.from_host(q))
    qq = np.stack([q, np.zeros_like(q)], -1).astype(np.float64)
    assert rel_l2(dp.to_host(), O64.gauge_transform(psi.astype(np.float64), qq)) < 1e-6
    # pressure projection of a smooth field (no one-form wraps)
    s = smooth_psi(n, L, 6)
    dp = isf.psi_array(s)
    isf.pressure_project_(dp)
    fac, _ = O64.tables(n, L, 0.1, 1 / 48)
    ref = O64.pressure_project(s.astype(np.float64), fac, d)
    err = rel_l2(dp.to_host(), ref)
    print(f"  pressure_project rel L2 = {err:.3e}")
    assert err < 2e-6
    # pressure_project! runs div(one_form) fused into the R2C x pass; the separate operators
    # (one-form with hbar 1, div!, poisson_solve, gauge_transform) give the same psi
    dc = isf.psi_array(s)
    v1 = isf.velocity_one_form_(dc, hbar=1.0, sharp=False, out=isf_mod.DeviceArray(isf.shape3 + (3,), np.float32))
    f1 = isf_mod.DeviceArray(isf.shape3, np.float32)
    isf.div_(f1, v1)
    isf.gauge_transform_(dc, isf.poisson_solve(f1))
    assert rel_l2(dc.to_host(), dp.to_host()) < 1e-6


# z planes of psi above 1 MB (nx ny > 65536): div(one_form) runs as its own z-marching
# stencil into f before the R2C x pass instead of fused into it; (512, 64, 32): the fused
# pass on the longest rows it takes (eight 64-cell segments per row)
@pytest.mark.parametrize("n,L", [((512, 256, 16), (8, 4, 1)), ((256, 512, 16), (4, 8, 1)), ((512, 64, 32), (16, 2, 1))])
def test_pressure_project_wide_planes(isf_mod, n, L):
    print(f"[project wide] n={n}", flush=True)
    isf = isf_mod.ISF(L, n, 0.1, 1 / 48)
    d = np.array(isf.d, np.float64)
    s = smooth_psi(n, L, 6)
    dp = isf.psi_array(s)
    isf.pressure_project_(dp)
    fac, _ = O64.tables(n, L, 0.1, 1 / 48)
    err = rel_l2(dp.to_host(), O64.pressure_project(s.astype(np.float64), fac, d))
    print(f"  pressure_project rel L2 = {err:.3e}")
    assert err < 2e-6


@pytest.mark.parametrize("mode", [0, 1])
def test_isjet_and_restrict(isf_mod, mode):
    n, L = (64, 32, 32), (4, 2, 2)
    isf = isf_mod.ISF(L, n, 0.1, 1 / 48)
    cen = np.float32([0.3, 0.966, 1.066])
    jet = isf.isjet(cen, 0.5, 0.5).to_host()
    ref = O32.isjet(n, np.array(isf.d, np.float32), cen, np.float32(0.5), np.float32(0.5))
    assert np.array_equal(jet, ref) and jet.sum() > 0
    psi = rnd((n[2], n[1], n[0], 4), 7)
    kvec = np.float32([10.0, 0.5, -0.25])
    dp = isf.psi_array(psi)
    isf.restrict_velocity_(dp, kvec, isf_mod.DeviceArray.from_host(jet), np.float32(0.3), mode)
    got = dp.to_host()
    want = O64.restrict_velocity(psi.astype(np.float64), jet, np.array(isf.d, np.float64), kvec.astype(np.float64),
                                 np.float64(np.float32(0.3)), mode)
    m = jet.astype(bool)
    assert np.array_equal(got[~m], psi[~m])
    assert np.max(np.abs(got[m] - want[m])) < 2e-5


def test_add_particles_matches_oracle(isf_mod):
    isf = isf_mod.ISF((4, 2, 2), (64, 32, 32), 0.1, 1 / 48)
    cen = np.float32([0.3, 0.966, 1.066])
    parts = np.full((1000, 3), -1.0, np.float32)
    dp = isf_mod.DeviceArray.from_host(parts)
    isf.add_particles_(dp, 950, 100, cen, np.float32(0.5), seed=99, step=7)
    got = dp.to_host()
    want = O32.add_particles(parts, 950, 100, cen, np.float32(0.5), 99, 7)
    assert np.array_equal(got == -1.0, want == -1.0)
    assert np.max(np.abs(got - want)) < 2e-6


def test_fill_nozzle_matches_oracle(isf_mod):
    isf = isf_mod.ISF((8, 4, 4), (256, 128, 128), 0.1, 1 / 48)
    cen = np.float32([0.3, 1.966, 2.066])
    isf.jet_configure(np.float32([10, 0, 0]), np.float32(5), cen, np.float32(0.5), np.float32(0.5), 0, 7)
    parts = np.full((300_000, 3), -1.0, np.float32)
    dp = isf_mod.DeviceArray.from_host(parts)
    isf.fill_nozzle_(dp, 1000, 250_000, 11)
    got = dp.to_host()
    want = O32.fill_cylinder(parts, 1000, 250_000, cen, np.float32(0.5), np.float32(0.5), 11)
    assert np.array_equal(got[:1000], parts[:1000]) and np.array_equal(got[251_000:], parts[251_000:])
    assert np.max(np.abs(got - want)) < 2e-6


@pytest.mark.parametrize("n,L", [((64, 32, 32), (4, 2, 2)), ((256, 128, 128), (8, 4, 4))])
def test_staggered_advect(isf_mod, n, L):
    isf = isf_mod.ISF(L, n, 0.1, 1 / 48)
    vel = rnd((n[2], n[1], n[0], 3), 8) * 2.0
    rng = np.random.default_rng(9)
    P = 200_000
    parts = (rng.random((P, 3)) * np.array(L) * 1.02 - 0.01 * np.array(L)).astype(np.float32)
    parts[:3] = [[-1, -1, -1], [0.0, 1.0, 1.0], [1.0, 1.0, L[2]]]
    dp = isf_mod.DeviceArray.from_host(parts)
    isf.staggered_advect_(dp, isf_mod.DeviceArray.from_host(vel))
    got = dp.to_host()
    want = O32.staggered_advect(parts, vel, isf.dt, np.array(isf.d), np.array(L, np.float32))
    rep = particle_report(got, want, L)
    print(f"  advect {rep}")
    assert np.all(got[:3] == -1.0)
    assert rep["status_mismatch"] <= P // 1000
    assert rep["p99"] < 1e-5


CFG_C1 = JetConfig(vol_size=(4, 2, 2), dims=(64, 32, 32), n_spawn=50, n_slots=48 * 50, constraint_iters=10)


def test_constraint_init_matches_oracle(isf_mod):
    g = GpuRun(CFG_C1)
    o = oracle_run(CFG_C1, np.float64)
    err = rel_l2(g.host_psi(), o.psi)
    print(f"  init rel L2 = {err:.3e}")
    assert err < 1e-5


def test_step_side_stream_ordering(isf_mod):
    """isf_step computes the velocity field and advects on a second stream.  A gauge
    transform queued right after the step (no synchronisation) rewrites psi with a phase
    that varies in x, which changes every edge angle: the step's velocity field and
    particles must still be those of the stepped psi, as in a run synchronised first."""
    cfg = JetConfig(n_spawn=200, n_slots=20_000, n_init=10_000, constraint_iters=4, seed=5)
    runs = []
    for sync_first in (True, False):
        g = GpuRun(cfg)
        g.step(3)
        if sync_first:
            g.isf.synchronize()
        nx = g.isf.grid_res[0]
        q = np.broadcast_to((0.4 * np.sin(2 * np.pi * np.arange(nx) / nx)).astype(np.float32), g.isf.shape3)
        g.isf.gauge_transform_(g.psi, isf_mod.DeviceArray.from_host(np.ascontiguousarray(q)))
        g.isf.synchronize()
        runs.append((g.isf.velocity.to_host(), g.host_parts(), g.host_psi()))
    (va, pa, sa), (vb, pb, sb) = runs
    assert np.array_equal(va, vb)
    assert np.array_equal(pa, pb)
    assert np.array_equal(sa, sb)
    # and the field is velocity_one_form of the psi the step produced (before the gauge)
    g = GpuRun(cfg)
    g.step(3)
    ref = g.isf.velocity_one_form_(g.psi, hbar=cfg.hbar, sharp=True,
                                   out=isf_mod.DeviceArray(g.isf.shape3 + (3,), np.float32)).to_host()
    assert np.array_equal(va, ref)


@pytest.mark.parametrize("phase_mode", [0, 1])
def test_full_steps_c1(isf_mod, phase_mode):
    """The CPU-runnable jet (64x32x32, 48 steps, 50 particles per step).

    In phase mode 0 the FP32 and FP64 oracles agree to ~1e-7 for the first ~24 steps;
    later a one-form edge angle crosses the +-pi branch cut (src/parallel.jl:108) on one
    side in FP32 and the other in FP64, and the Poisson solve spreads that 2 pi jump over
    the grid, so by step 48 every FP32 implementation (the FP32 CPU restatement
    included) is O(0.1) away from FP64.  The test therefore checks (a) free-running
    parity over the first 20 steps, and (b) step-local parity through the whole run: at
    steps 8, 16, 24, 32, 40 and 48 the GPU takes one step from the FP64 state of the step
    before, against the FP64 step (psi rel L2 <= 1e-5, particles p99 <= 1e-5)."""
    cfg = JetConfig(**{**CFG_C1.__dict__, "phase_mode": phase_mode})
    g = GpuRun(cfg)
    o = oracle_run(cfg, np.float64)
    o32 = oracle_run(cfg, np.float32)
    g.step(20)
    o.step(20)
    o32.step(20)
    err20 = rel_l2(g.host_psi(), o.psi)
    rep20 = particle_report(g.host_parts(), o.parts, cfg.vol_size)
    print(f"  C1 step 20: psi rel L2 vs FP64 = {err20:.3e} (FP32 oracle {rel_l2(o32.psi, o.psi):.3e}); {rep20}")
    assert err20 < 1e-4
    assert rep20["status_mismatch"] <= 2 and rep20["p99"] < 1e-4
    # step-local parity from the FP64 state (the oracle keeps running 21..48 on its own)
    checks = (8, 16, 24, 32, 40, 48)
    o = oracle_run(cfg, np.float64)
    for k in range(1, 49):
        if k in checks:
            g.psi.copy_from(o.psi.astype(np.float32))
            g.parts.copy_from(o.parts.astype(np.float32))
            g.iter = k - 1
            g.step(1)
        o.step(1)
        if k in checks:
            err1 = rel_l2(g.host_psi(), o.psi)
            rep1 = particle_report(g.host_parts(), o.parts, cfg.vol_size)
            print(f"  C1 step {k} from the FP64 state at {k - 1}: psi rel L2 = {err1:.3e}; particles p99 "
                  f"{rep1['p99']:.2e}, mismatches {rep1['status_mismatch']}", flush=True)
            assert err1 < 1e-5, k
            assert rep1["status_mismatch"] <= 2 and rep1["p99"] < 1e-5, k
            nrm = g.isf.step_norm2()
            assert abs(nrm - np.prod(cfg.dims)) < 1e-5 * np.prod(cfg.dims)
    # the free-running FP32 paths after the branch-cut event, reported (not a bar)
    print(f"  C1 step 48 free-running: FP32 oracle vs FP64 {rel_l2(o32.step(28).psi, o.psi):.3e}")


@pytest.mark.timeout(600)
def test_full_steps_c2(isf_mod):
    """The at-scale jet (256x128x128, box 8x4x4, 1M slots seeded in the nozzle, 1000
    spawned per step): 20 free-running steps against the FP64 oracle, psi within 1e-4
    relative L2 at steps 5, 10 and 20, particles p99 within 1e-4, plus the FP64 norm."""
    from oracle.oracle import cpu_engine
    cfg = JetConfig(vol_size=(8, 4, 4), dims=(256, 128, 128), nozzle_cen=(0.3, 2 - 0.034, 2 + 0.066),
                    n_spawn=1000, n_slots=1_000_000, n_init=1_000_000, constraint_iters=10)
    g = GpuRun(cfg)
    with cpu_engine("mkl"):
        o = oracle_run(cfg, np.float64)
        for k in (5, 10, 20):
            g.step(k - g.iter)
            o.step(k - o.iter)
            err = rel_l2(g.host_psi(), o.psi)
            rep = particle_report(g.host_parts(), o.parts, cfg.vol_size)
            print(f"  C2 step {k}: psi rel L2 vs FP64 = {err:.3e}; particles {rep}", flush=True)
            assert err < 1e-4
            assert rep["status_mismatch"] <= cfg.n_slots // 1000
            assert rep["p99"] < 1e-4
    nrm = g.isf.step_norm2()
    assert abs(nrm - np.prod(cfg.dims)) < 1e-5 * np.prod(cfg.dims)


@pytest.mark.timeout(900)
def test_full_steps_c5_large(isf_mod):
    """The large jet (512x256x256, box 16x8x8, 4M slots seeded in the nozzle, 1000
    spawned per step; split plane passes, unfused divergence) for 3 steps against the
    FP64 oracle, plus two size-independent invariants of the projected state: the FP64
    norm equals N, and the divergence of the recomputed one-form is at the level of
    rounding away from the cells whose edge angles wrap."""
    from oracle.oracle import cpu_engine
    cfg = JetConfig(vol_size=(16, 8, 8), dims=(512, 256, 256), nozzle_cen=(0.3, 4 - 0.034, 4 + 0.066),
                    n_spawn=1000, n_slots=4_000_000, n_init=4_000_000, constraint_iters=4)
    g = GpuRun(cfg)
    g.step(3)
    psi = g.host_psi()
    parts = g.host_parts()
    nrm = g.isf.step_norm2()
    N = int(np.prod(cfg.dims))
    assert abs(nrm - N) < 1e-5 * N
    # divergence of the one-form of the projected psi (hbar = 1), against the same
    # quantity of the FP64 oracle's state: rounding noise, plus 2 pi / d^2 at edges whose
    # angles wrap (none in this state)
    f = isf_mod.DeviceArray(g.isf.shape3, np.float32)
    g.isf.velocity_one_form_(g.psi, 1.0)
    g.isf.div_(f)
    fg = np.abs(f.to_host())
    del f
    g.psi.free()
    g.parts.free()
    g.isf.close()
    with cpu_engine("mkl"):
        o = oracle_run(cfg, np.float64)
        o.step(3)
    fo = np.abs(O64.div(O64.velocity_one_form(o.psi, o.d, 1.0), o.d))
    p99g, p99o = np.percentile(fg, 99), np.percentile(fo, 99)
    print(f"  C5 |div| after 3 steps: GPU median {np.median(fg):.3e} p99 {p99g:.3e}; FP64 median "
          f"{np.median(fo):.3e} p99 {p99o:.3e}", flush=True)
    # FP32 floor: angle rounding ~1e-7 times 1/d^2 = 1024 over six edge terms (measured
    # p99 4e-5); an unprojected state is O(1) and above in the jet
    assert p99g <= 2.0 * p99o + 1e-3
    err = rel_l2(psi, o.psi)
    rep = particle_report(parts, o.parts, cfg.vol_size)
    print(f"  C5 step 3: psi rel L2 vs FP64 = {err:.3e}; particles {rep}", flush=True)
    assert err < 1e-4
    assert rep["status_mismatch"] <= cfg.n_slots // 1000
    assert rep["p99"] < 1e-4


# planes larger than one workgroup's LDS (y or z above 128): split y / z passes
LARGE_SHAPES = [((32, 256, 128), (2, 16, 8)), ((16, 512, 64), (1, 32, 4)), ((64, 128, 256), (4, 8, 16)),
                ((16, 64, 512), (1, 4, 32))]


@pytest.mark.parametrize("n,L", LARGE_SHAPES)
def test_large_plane_flow_and_poisson(isf_mod, n, L):
    print(f"[large plane] n={n}", flush=True)
    isf = isf_mod.ISF(L, n, 0.1, 1 / 48)
    psi = rnd((n[2], n[1], n[0], 4), 3)
    fac, mask = O64.tables(n, L, float(np.float32(0.1)), float(np.float32(1 / 48)))
    d = isf.psi_array(psi)
    isf.schroedinger_flow_(d)
    err = rel_l2(d.to_host(), O64.schroedinger_flow(psi.astype(np.float64), mask))
    f = rnd((n[2], n[1], n[0]), 4)
    q = isf.poisson_solve(isf_mod.DeviceArray.from_host(f)).to_host()
    errq = rel_l2(q, O64.poisson_solve(f.astype(np.float64), fac)[..., 0])
    dd = [L[a] / n[a] for a in range(3)]
    lap = rel_l2(laplacian(q.astype(np.float64), dd), f - f.mean())
    print(f"  flow rel L2 {err:.3e}; poisson rel L2 {errq:.3e}, Laplacian residual {lap:.3e}")
    assert err < 2e-6
    assert errq < 2e-6
    assert lap < 1e-4
    # adjoint of the flow through the split passes: <A x, y> = <x, A^T y>
    y = rnd((n[2], n[1], n[0], 4), 5)
    b = isf.psi_array(y)
    isf.vjp_schroedinger_flow_(b)
    ax, aty = d.to_host().astype(np.float64), b.to_host().astype(np.float64)
    lhs = float(np.sum(ax * y))
    rhs = float(np.sum(psi.astype(np.float64) * aty))
    scale = float(np.linalg.norm(ax) * np.linalg.norm(y))
    print(f"  <Ax, y> = {lhs:.8e}, <x, A^T y> = {rhs:.8e} (scale {scale:.3e})")
    assert abs(lhs - rhs) < 1e-6 * scale


def test_full_steps_large_plane(isf_mod):
    """A jet on a grid with 256 x 128 planes (split plane passes) against the FP64 oracle."""
    cfg = JetConfig(vol_size=(2, 16, 8), dims=(32, 256, 128), nozzle_cen=(0.3, 8 - 0.034, 4 + 0.066),
                    n_spawn=200, n_slots=20_000, n_init=10_000, constraint_iters=4)
    g = GpuRun(cfg)
    g.step(3)
    o = oracle_run(cfg, np.float64)
    o.step(3)
    err = rel_l2(g.host_psi(), o.psi)
    rep = particle_report(g.host_parts(), o.parts, cfg.vol_size)
    print(f"  32x256x128 psi rel L2 vs FP64 = {err:.3e}; particles {rep}")
    assert err < 1e-4
    assert rep["status_mismatch"] <= 2 and rep["p99"] < 1e-4


@pytest.mark.parametrize("dims,box,n_slots,n_spawn", [((64, 32, 32), (4, 2, 2), 2400, 50),
                                                      ((256, 128, 128), (8, 4, 4), 1_000_000, 1000),
                                                      ((64, 32, 32), (4, 2, 2), 0, 0),
                                                      ((512, 32, 32), (16, 2, 2), 4000, 50),
                                                      ((64, 256, 32), (2, 8, 2), 4000, 50)])
def test_simloop_equals_isf_step(isf_mod, dims, box, n_slots, n_spawn):
    """isf_simloop (the driver's simloop, test/runtests.jl:40-66: the velocity pass of each
    step fused into the next step's forward x pass, two alternating velocity buffers)
    gives bit for bit the psi, particles, velocity field and FP64 norm of the same steps
    taken one isf_step at a time, over calls of 1, 2, 3 and 5 steps in a row."""
    cen = (0.3, box[1] / 2 - 0.034, box[2] / 2 + 0.066)
    cfg = JetConfig(vol_size=box, dims=dims, nozzle_cen=cen, n_spawn=n_spawn, n_slots=n_slots,
                    n_init=n_slots // 2, constraint_iters=4)
    a, b = GpuRun(cfg), GpuRun(cfg)
    it = 1
    for count in (1, 2, 3, 5):
        for k in range(count):
            a.isf.step_(a.psi, a.parts if n_slots else None, it + k)
        b.isf.simloop_(b.psi, b.parts if n_slots else None, it, count)
        it += count
        pa, pb = a.host_psi(), b.host_psi()
        print(f"  {dims} after {it - 1} steps (call of {count}): psi rel {rel_l2(pb, pa):.2e}", flush=True)
        assert np.array_equal(pa, pb)
        if n_slots:
            assert np.array_equal(a.host_parts(), b.host_parts())
        assert np.array_equal(a.isf.velocity.to_host(), b.isf.velocity.to_host())
        assert a.isf.step_norm2() == b.isf.step_norm2()
