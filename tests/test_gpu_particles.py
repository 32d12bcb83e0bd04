"""The MatlabPort particle store on the GPU (isf_particles_*) against the oracle's
restatement (oracle/matlabport.py Particles): the MatlabPort driver loop
(examples/matlabport/simple.jl:95-122) with the grid operators of libisf on one side
and the FP64 C oracle on the other.  Slot bookkeeping (length, the active list after
removethedead!) must be identical; positions of active slots within the particle
tolerances of the other parity tests."""
import numpy as np
import pytest

from oracle import matlabport as mp
from oracle.oracle import JetConfig, Oracle

pytestmark = pytest.mark.gpu


def run_driver(steps, births, capacity, cfg, keep=None):
    from isf_amd import ISF, Particles
    f32 = np.float32
    o = Oracle(np.float64)
    n, L = cfg.dims, cfg.vol_size
    # GPU side
    isf = ISF(L, n, cfg.hbar, cfg.dt)
    v = np.array(cfg.jet_velocity, f32)
    hb = f32(cfg.hbar)
    kvec = v / hb
    omega = f32(np.sum(v ** f32(2)) / (f32(2) * hb))
    cen = np.array(cfg.nozzle_cen, np.float64).astype(f32)
    isf.jet_configure(kvec, omega, cen, f32(cfg.nozzle_len), f32(cfg.nozzle_rad), 0, cfg.seed, 0)
    psi = np.zeros((n[2], n[1], n[0], 4), f32)
    psi[..., 0] = 1.0
    psi[..., 2] = 0.01
    dpsi = isf.psi_array(psi)
    isf.normalize_(dpsi)
    isf.constrain_(dpsi, cfg.constraint_iters)
    P = Particles(isf, capacity)
    # oracle side
    d = np.array([f32(L[a] / n[a]) for a in range(3)], np.float64)
    fac, mask = o.tables(n, L, float(hb), float(f32(cfg.dt)))
    isjet = o.isjet(n, d, cen.astype(np.float64), float(f32(cfg.nozzle_len)), float(f32(cfg.nozzle_rad)))
    opsi = np.zeros((n[2], n[1], n[0], 4))
    opsi[..., 0] = 1.0
    opsi[..., 2] = float(f32(0.01))
    opsi = o.normalize(opsi)
    k64, om64 = kvec.astype(np.float64), float(omega)
    for _ in range(cfg.constraint_iters):
        opsi = o.restrict_velocity(opsi, isjet, d, k64, 0.0, 0)
        opsi = o.pressure_project(opsi, fac, d)
    OP = mp.Particles(capacity)
    torus = mp.Torus(n, L, tuple(d))
    dt64 = float(f32(cfg.dt))
    for it in range(1, steps + 1):
        t32 = f32(it) * f32(cfg.dt)
        isf.schroedinger_flow_(dpsi)
        isf.normalize_(dpsi)
        isf.pressure_project_(dpsi)
        isf.restrict_velocity_(dpsi, kvec, None, f32(omega * t32), 0)
        isf.pressure_project_(dpsi)
        P.spawn_ring_(births, cfg.seed, it)
        P.remove_dead_()
        isf.velocity_one_form_(dpsi, float(hb), sharp=True)
        P.advect_()
        opsi = o.schroedinger_flow(opsi, mask)
        opsi = o.normalize(opsi)
        opsi = o.pressure_project(opsi, fac, d)
        opsi = o.restrict_velocity(opsi, isjet, d, k64, om64 * (it * dt64), 0)
        opsi = o.pressure_project(opsi, fac, d)
        OP.spawn_ring(births, cen.astype(np.float64), float(f32(cfg.nozzle_rad)), cfg.seed, it)
        OP.remove_dead(L)
        vel = o.velocity_one_form(opsi, d, float(hb), sharp=True)
        OP.advect(torus, mp.to_xyz(vel[..., 0]), mp.to_xyz(vel[..., 1]), mp.to_xyz(vel[..., 2]), dt64)
    if keep is not None:
        P.keep_(keep)
        OP.keep(keep)
    return P.state(), OP


def compare(state, OP, tol=1e-4):
    length, active, x, y, z = state
    g = np.stack([x, y, z], 1)
    r = np.stack([OP.x, OP.y, OP.z], 1)
    a = np.asarray(OP.active, np.int64)
    err = np.abs(g[a] - r[a]).max(1) if a.size else np.zeros(1)
    print(f"  length {length} / {OP.length}; active {len(active)} / {a.size}; |dp| p99 {np.percentile(err, 99):.2e} "
          f"max {err.max():.2e}")
    return length, active, a, err


def test_matlabport_driver_loop():
    """20 steps: before the FP32 / FP64 branch-cut event of this jet (see the C1 parity test)."""
    cfg = JetConfig(constraint_iters=10, seed=20261203)
    state, OP = run_driver(20, 10, 20_000, cfg)
    length, active, a, err = compare(state, OP)
    assert length == OP.length == 200
    assert np.array_equal(active, a)  # same survivors in the same order
    assert np.percentile(err, 99) < 1e-4


def test_remove_dead_on_the_device():
    from isf_amd import ISF, Particles
    isf = ISF((4, 2, 2), (64, 32, 32), 0.1, 1 / 48)
    rng = np.random.default_rng(9)
    n = 5000
    pts = rng.uniform(-0.5, 4.5, (n, 3)) * np.array([1.0, 0.5, 0.5])
    pts[:4] = [[0.0, 1.0, 1.0], [4.0, 1.0, 1.0], [1.0, 2.0, 1.0], [1.0, 1.0, 1e-7]]  # faces: out, out, out, in
    P = Particles(isf, 3 * n)
    OP = mp.Particles(3 * n)
    for k in range(2):  # two appends, removal after each
        P.append_(pts[:, 0], pts[:, 1], pts[:, 2])
        P.remove_dead_()
        OP.append(pts[:, 0], pts[:, 1], pts[:, 2])
        OP.remove_dead((4, 2, 2))
    length, active, x, y, z = P.state()
    print(f"  kept {len(active)} of {length}")
    assert length == OP.length == 2 * n
    assert np.array_equal(active, np.asarray(OP.active))
    assert 3 in active and 0 not in active and 1 not in active and 2 not in active


def test_host_filter_of_the_active_list():
    """The MatlabPort driver's own removethedead! (examples/matlabport/simple.jl:63-70)
    filters p.active on the host with isinside; written back with isf_particles_set_active
    it leaves the same active list as the device removethedead!, and the next advection
    moves the same slots."""
    from isf_amd import ISF, Particles
    isf = ISF((4, 2, 2), (64, 32, 32), 0.1, 1 / 48)
    rng = np.random.default_rng(5)
    n = 4000
    pts = rng.uniform(-0.3, 4.3, (n, 3)) * np.array([1.0, 0.5, 0.5])
    A, B = Particles(isf, 2 * n), Particles(isf, 2 * n)
    for P in (A, B):
        P.append_(pts[:, 0], pts[:, 1], pts[:, 2])
    A.remove_dead_()
    ln, act, x, y, z = B.state()
    keep = [i for i in act if 0 < x[i] < 4 and 0 < y[i] < 2 and 0 < z[i] < 2]  # isinside, strict
    B.set_active_(keep)
    la, acta, _, _, _ = A.state()
    lb, actb, _, _, _ = B.state()
    assert la == lb == n and np.array_equal(acta, actb) and 0 < len(actb) < n
    vel = isf.velocity
    vel.copy_from(np.full(isf.shape3 + (3,), 0.25, np.float32))
    A.advect_()
    B.advect_()
    xa, xb = A.state()[2], B.state()[2]
    assert np.array_equal(xa, xb)
    with pytest.raises(Exception):
        B.set_active_([2 * n])  # beyond the position arrays
    A.close()
    B.close()
    isf.close()


def test_keep_matches_the_reference():
    """Keep(p, ind) (src/MatlabPort.jl:226-230) replaces x, y, z by x[ind], y[ind], z[ind]
    and leaves length and the active list alone; the arrays are then len(ind) long, so the
    slots beyond hold the dead marker here and append! checks against that length."""
    cfg = JetConfig(constraint_iters=4, seed=3)
    keep = np.array([7, 3, 30, 0], np.int32)
    state, OP = run_driver(4, 10, 1_000, cfg)
    length0, active0 = state[0], state[1]
    from isf_amd import ISF, Particles
    state2, OP2 = run_driver(4, 10, 1_000, cfg, keep=keep)
    length, active, x, y, z = state2
    assert length == length0 == OP2.length == 40
    assert np.array_equal(active, active0) and np.array_equal(active, np.asarray(OP2.active))
    assert OP2.x.shape == (4,)
    assert np.allclose(x[:4], OP2.x, atol=1e-4) and np.allclose(y[:4], OP2.y, atol=1e-4)
    assert np.allclose(z[:4], OP2.z, atol=1e-4)
    assert np.all(x[4:] == -1.0) and np.all(y[4:] == -1.0) and np.all(z[4:] == -1.0)
    # append! after Keep: length(p.x) is now 4 and length 40, so any append is refused
    isf = ISF(cfg.vol_size, cfg.dims, cfg.hbar, cfg.dt)
    P = Particles(isf, 100)
    P.append_([1.0, 2.0], [0.5, 0.5], [0.5, 0.5])
    P.keep_(np.array([1, 0, 1], np.int32))
    with pytest.raises(Exception):
        P.append_([1.0], [0.5], [0.5])  # 2 + 1 < 3 fails, as the reference's assert does
    ln, act, x, _, _ = P.state()
    assert ln == 2 and list(act) == [0, 1] and list(x[:3]) == [2.0, 1.0, 2.0]
    P.close()
    isf.close()


def test_async_particle_export_frames():
    """Exported frames (the per-frame Array(particles) of examples/jet.jl:83) equal
    synchronous copies of the same deterministic run, and each frame is the FP64
    oracle's particle state after that step within the particle tolerance (p99 of |dp|
    <= 1e-4 over slots alive in both, at most 0.1 % alive-status mismatches)."""
    from _harness import GpuRun, oracle_run, particle_report
    from isf_amd import ParticleExport
    cfg = JetConfig(n_spawn=200, n_slots=50_000, n_init=30_000, constraint_iters=4, seed=11)
    ref = GpuRun(cfg)
    frames = []
    for _ in range(6):
        ref.step(1)
        frames.append(ref.host_parts())
    run = GpuRun(cfg)
    ex = ParticleExport(run.isf, cfg.n_slots)
    got = {}
    for k in range(6):
        run.step(1)  # no synchronisation between steps and exports
        ex.push(run.parts)
        if k >= 1:
            f, pos = ex.previous()
            got[f] = pos
    f, pos = ex.latest()
    got[f] = pos
    assert sorted(got) == list(range(6))
    for k in range(6):
        assert np.array_equal(got[k], frames[k]), k
    ex.close()
    o = oracle_run(cfg, np.float64)
    for k in range(6):
        o.step(1)
        rep = particle_report(got[k], o.parts, cfg.vol_size)
        print(f"  frame {k}: {rep}")
        assert rep["status_mismatch"] <= cfg.n_slots // 1000 and rep["p99"] < 1e-4, k
        assert rep["alive_gpu"] >= cfg.n_init, k
