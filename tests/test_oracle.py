"""Pins of the CPU oracle (oracle/isf_oracle.c) against closed forms, invariants, numpy's
FFT, brute force on tiny grids and the independent MatlabPort restatement
(oracle/matlabport.py, src/MatlabPort.jl).  CPU only."""
import numpy as np
import pytest

from oracle.oracle import JetConfig, Oracle, OracleRun, philox4x32_10, philox_uniform4
from oracle import matlabport as mp

O64 = Oracle(np.float64)
O32 = Oracle(np.float32)


def rnd_psi(n, seed=0, dtype=np.float64):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((n[2], n[1], n[0], 4)).astype(dtype)


def smooth_psi(n, L, seed=0):
    """Unit-amplitude-ish psi with a smooth phase, so no one-form edge wraps past pi."""
    rng = np.random.default_rng(seed)
    z, y, x = np.meshgrid(*[np.arange(n[a]) * (L[a] / n[a]) for a in (2, 1, 0)], indexing="ij")
    ph1 = 0.3 * np.sin(2 * np.pi * x / L[0]) + 0.2 * np.cos(2 * np.pi * (y / L[1] + z / L[2]))
    ph2 = 0.25 * np.cos(2 * np.pi * 2 * x / L[0]) + 0.1 * np.sin(2 * np.pi * z / L[2])
    a1 = 1.0 + 0.1 * rng.random(x.shape)
    a2 = 0.1 + 0.05 * rng.random(x.shape)
    psi = np.stack([a1 * np.cos(ph1), a1 * np.sin(ph1), a2 * np.cos(ph2), a2 * np.sin(ph2)], -1)
    return psi


def to_c(psi):
    return psi[..., 0] + 1j * psi[..., 1], psi[..., 2] + 1j * psi[..., 3]


def from_c(p1, p2):
    return np.stack([p1.real, p1.imag, p2.real, p2.imag], -1)


def rel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(np.asarray(b).ravel()))


# ------------------------------------------------------------------ FFT
@pytest.mark.parametrize("n", [(8, 4, 2), (16, 8, 4), (32, 16, 16), (64, 32, 32)])
def test_fft3_matches_numpy(n):
    rng = np.random.default_rng(1)
    z = rng.standard_normal((n[2], n[1], n[0], 2))
    zc = z[..., 0] + 1j * z[..., 1]
    F = O64.fft3(z)
    ref = np.fft.fftn(zc)
    assert rel(F[..., 0] + 1j * F[..., 1], ref) < 1e-13
    G = O64.fft3(z, inverse=True)
    assert rel(G[..., 0] + 1j * G[..., 1], np.fft.ifftn(zc)) < 1e-13
    F32 = O32.fft3(z.astype(np.float32))
    assert rel(F32[..., 0] + 1j * F32[..., 1], ref) < 2e-6


def test_fft3_impulse_planewave_parseval():
    n = (16, 8, 4)
    z = np.zeros((n[2], n[1], n[0], 2))
    z[0, 0, 0, 0] = 1.0
    F = O64.fft3(z)
    assert np.allclose(F[..., 0], 1.0) and np.allclose(F[..., 1], 0.0)
    m = (3, 5, 1)
    zz, yy, xx = np.meshgrid(np.arange(n[2]), np.arange(n[1]), np.arange(n[0]), indexing="ij")
    ph = 2 * np.pi * (m[0] * xx / n[0] + m[1] * yy / n[1] + m[2] * zz / n[2])
    w = np.stack([np.cos(ph), np.sin(ph)], -1)
    F = O64.fft3(w)
    mag = np.hypot(F[..., 0], F[..., 1])
    assert abs(mag[m[2], m[1], m[0]] - np.prod(n)) < 1e-9
    mag[m[2], m[1], m[0]] = 0
    assert mag.max() < 1e-9
    r = np.random.default_rng(3).standard_normal(z.shape)
    F = O64.fft3(r)
    assert abs(np.sum(F ** 2) / np.prod(n) - np.sum(r ** 2)) < 1e-9 * np.sum(r ** 2)


# ------------------------------------------------------------------ tables
@pytest.mark.parametrize("n,L", [((16, 8, 8), (4, 2, 2)), ((64, 32, 32), (4, 2, 2)), ((32, 16, 16), (8, 4, 4))])
def test_mask_is_fft_ordered_free_particle_phase(n, L):
    hbar, dt = float(np.float32(0.1)), float(np.float32(1 / 48))
    fac, mask = O64.tables(n, L, hbar, dt)
    kx = np.fft.fftfreq(n[0], d=L[0] / n[0])
    ky = np.fft.fftfreq(n[1], d=L[1] / n[1])
    kz = np.fft.fftfreq(n[2], d=L[2] / n[2])
    KZ, KY, KX = np.meshgrid(kz, ky, kx, indexing="ij")
    ph = -4 * np.pi ** 2 * hbar * (KX ** 2 + KY ** 2 + KZ ** 2) * dt / 2
    assert np.max(np.abs(mask[..., 0] - np.cos(ph))) < 1e-12
    assert np.max(np.abs(mask[..., 1] - np.sin(ph))) < 1e-12
    # MatlabPort builds the unshifted mask (src/MatlabPort.jl:75-83); fftshift of it is ours
    m = mp.MatlabISF(L, n, hbar, dt)
    ours = mp.to_xyz(mask[..., 0] + 1j * mask[..., 1])
    assert np.max(np.abs(np.fft.fftshift(m.mask) - ours)) < 1e-12


def laplacian(q, n, d):
    out = np.zeros_like(q)
    for ax, a in ((2, 0), (1, 1), (0, 2)):
        out += (np.roll(q, -1, ax) - 2 * q + np.roll(q, 1, ax)) / d[a] ** 2
    return out


@pytest.mark.parametrize("n,L", [((16, 8, 8), (4, 2, 2)), ((8, 8, 16), (2, 2, 4))])
def test_poisson_inverts_discrete_laplacian(n, L):
    fac, _ = O64.tables(n, L, 0.1, 0.02)
    assert fac[0, 0, 0, 0] == 0.0 and np.all(fac[..., 1] == 0)
    d = [L[a] / n[a] for a in range(3)]
    f = np.random.default_rng(5).standard_normal((n[2], n[1], n[0]))
    q = O64.poisson_solve(f, fac)
    assert np.max(np.abs(q[..., 1])) < 1e-12 * np.max(np.abs(q[..., 0]))
    lap = laplacian(q[..., 0], n, d)
    assert rel(lap, f - f.mean()) < 1e-11
    # brute force dense solve on a 4x4x4 grid (mean-zero solution of the same system)
    n2 = (4, 4, 4)
    fac2, _ = O64.tables(n2, (1, 1, 1), 0.1, 0.02)
    f2 = np.random.default_rng(6).standard_normal((4, 4, 4))
    f2 -= f2.mean()
    A = np.zeros((64, 64))
    for i in range(64):
        e = np.zeros(64)
        e[i] = 1
        A[:, i] = laplacian(e.reshape(4, 4, 4), n2, [0.25] * 3).ravel()
    A = np.vstack([A, np.ones((1, 64))])
    sol = np.linalg.lstsq(A, np.concatenate([f2.ravel(), [0.0]]), rcond=None)[0]
    q2 = O64.poisson_solve(f2, fac2)[..., 0]
    assert rel(q2.ravel(), sol) < 1e-10


# ------------------------------------------------------------------ operators
def test_schroedinger_flow_plane_wave_eigenfunction_and_unitarity():
    n, L = (16, 8, 8), (4, 2, 2)
    hbar, dt = 0.1, 1 / 48
    _, mask = O64.tables(n, L, hbar, dt)
    m = (2, -1, 3)
    zz, yy, xx = np.meshgrid(*[np.arange(n[a]) * L[a] / n[a] for a in (2, 1, 0)], indexing="ij")
    ph = 2 * np.pi * (m[0] * xx / L[0] + m[1] * yy / L[1] + m[2] * zz / L[2])
    psi = np.stack([np.cos(ph), np.sin(ph), 0.5 * np.cos(ph), 0.5 * np.sin(ph)], -1)
    out = O64.schroedinger_flow(psi, mask)
    k2 = sum((m[a] / L[a]) ** 2 for a in range(3))
    g = np.exp(1j * (-4 * np.pi ** 2 * hbar * k2 * dt / 2))
    p1, p2 = to_c(psi)
    assert rel(from_c(p1 * g, p2 * g), out) < 1e-12
    r = rnd_psi(n, 2)
    o2 = O64.schroedinger_flow(r, mask)
    assert abs(np.sum(o2 ** 2) - np.sum(r ** 2)) < 1e-10 * np.sum(r ** 2)


def test_normalize_unit_and_idempotent():
    psi = rnd_psi((8, 4, 4), 3)
    a = O64.normalize(psi)
    assert np.max(np.abs(np.sum(a ** 2, -1) - 1)) < 1e-14
    assert np.max(np.abs(O64.normalize(a) - a)) < 1e-15


def test_one_form_plane_wave_and_matlabport():
    n, L = (16, 8, 8), (4, 2, 2)
    d = [L[a] / n[a] for a in range(3)]
    m = (3, 1, -2)
    zz, yy, xx = np.meshgrid(*[np.arange(n[a]) * d[a] for a in (2, 1, 0)], indexing="ij")
    ph = 2 * np.pi * (m[0] * xx / L[0] + m[1] * yy / L[1] + m[2] * zz / L[2])
    psi = np.stack([np.cos(ph), np.sin(ph), 0.1 * np.cos(ph), 0.1 * np.sin(ph)], -1)
    v = O64.velocity_one_form(psi, d, hbar=0.5)
    for a in range(3):
        want = 0.5 * np.angle(np.exp(2j * np.pi * m[a] / n[a]))
        assert np.max(np.abs(v[..., a] - want)) < 1e-12
    vs = O64.velocity_one_form(psi, d, hbar=0.5, sharp=True)
    assert np.max(np.abs(vs - v / np.array(d))) < 1e-12
    # random psi against MatlabPort VelocityOneForm (src/MatlabPort.jl:106-126)
    r = rnd_psi(n, 4)
    obj = mp.MatlabISF(L, n, 0.1, 0.02)
    p1, p2 = to_c(r)
    vx, vy, vz = mp.velocity_one_form(obj, mp.to_xyz(p1), mp.to_xyz(p2), 0.1)
    vo = O64.velocity_one_form(r, d, hbar=0.1)
    for a, w in enumerate((vx, vy, vz)):
        assert np.max(np.abs(mp.to_xyz(vo[..., a]) - w)) < 1e-12


def test_div_sum_zero_constant_and_matlabport():
    n, L = (16, 8, 4), (4, 2, 1)
    d = [L[a] / n[a] for a in range(3)]
    rng = np.random.default_rng(7)
    v = rng.standard_normal((n[2], n[1], n[0], 3))
    f = O64.div(v, d)
    assert abs(f.sum()) < 1e-9 * np.abs(f).sum()
    assert np.max(np.abs(O64.div(np.ones_like(v) * [1.0, -2.0, 3.0], d))) == 0.0
    obj = mp.MatlabISF(L, n, 0.1, 0.02)
    want = mp.div(obj, mp.to_xyz(v[..., 0]), mp.to_xyz(v[..., 1]), mp.to_xyz(v[..., 2]))
    assert np.max(np.abs(mp.to_xyz(f) - want)) < 1e-9


def test_pressure_project_matches_matlabport_and_is_divergence_free():
    n, L = (16, 8, 8), (4, 2, 2)
    d = [L[a] / n[a] for a in range(3)]
    fac, _ = O64.tables(n, L, 0.1, 0.02)
    obj = mp.MatlabISF(L, n, 0.1, 0.02)
    r = rnd_psi(n, 8)
    out = O64.pressure_project(r, fac, d)
    p1, p2 = to_c(r)
    q1, q2 = mp.pressure_project(obj, mp.to_xyz(p1), mp.to_xyz(p2))
    o1, o2 = to_c(out)
    assert rel(mp.to_xyz(o1), q1) < 1e-12 and rel(mp.to_xyz(o2), q2) < 1e-12
    # amplitude preserved; projected one-form divergence-free (smooth psi, no wraps)
    s = smooth_psi(n, L, 9)
    ps = O64.pressure_project(s, fac, d)
    assert np.max(np.abs(np.sum(ps ** 2, -1) - np.sum(s ** 2, -1))) < 1e-12
    f_before = O64.div(O64.velocity_one_form(s, d), d)
    f_after = O64.div(O64.velocity_one_form(ps, d), d)
    assert np.max(np.abs(f_after)) < 1e-9 * np.max(np.abs(f_before))


def test_gauge_transform_complex_q():
    psi = rnd_psi((4, 4, 4), 10)
    q = np.random.default_rng(11).standard_normal((4, 4, 4, 2)) * 0.3
    out = O64.gauge_transform(psi, q)
    p1, p2 = to_c(psi)
    qc = q[..., 0] + 1j * q[..., 1]
    g = np.exp(1j * (-qc))
    assert rel(from_c(p1 * g, p2 * g), out) < 1e-14


# ------------------------------------------------------------------ jet
def test_isjet_brute_force():
    n, L = (32, 16, 16), (4, 2, 2)
    d = [np.float32(L[a] / n[a]) for a in range(3)]
    cen = np.float32([0.3, 0.966, 1.066])
    got = O32.isjet(n, d, cen, 0.5, 0.5)
    for z in range(n[2]):
        for y in range(n[1]):
            for x in range(n[0]):
                p = (np.float32(x) * d[0], np.float32(y) * d[1], np.float32(z) * d[2])
                want = (abs(p[0] - cen[0]) <= np.float32(0.25)) and \
                       ((p[1] - cen[1]) ** 2 + (p[2] - cen[2]) ** 2 <= np.float32(0.25))
                assert got[z, y, x] == want
    assert got.sum() > 0


@pytest.mark.parametrize("mode", [0, 1])
def test_restrict_velocity_inside_phase_outside_untouched(mode):
    n, L = (32, 16, 16), (4, 2, 2)
    d = np.array([L[a] / n[a] for a in range(3)])
    cen = np.array([0.3, 0.966, 1.066])
    jet = O64.isjet(n, d, cen, 0.5, 0.5)
    psi = rnd_psi(n, 12)
    kvec = np.array([10.0, 2.0, -1.0])
    out = O64.restrict_velocity(psi, jet, d, kvec, 0.7, mode)
    m = jet.astype(bool)
    assert np.array_equal(out[~m], psi[~m])
    zz, yy, xx = np.meshgrid(*[np.arange(n[a]) * d[a] for a in (2, 1, 0)], indexing="ij")
    if mode == 0:
        ph = kvec[0] * xx + kvec[1] * yy + kvec[2] * zz - 0.7
    else:
        ph = (kvec[0] + xx) + (kvec[1] + yy) + (kvec[2] + zz) - 0.7
    p1, p2 = to_c(psi)
    w1 = np.abs(p1) * np.exp(1j * ph)
    w2 = np.abs(p2) * np.exp(1j * ph)
    assert np.max(np.abs(from_c(w1, w2)[m] - out[m])) < 1e-12


# ------------------------------------------------------------------ particles
def test_philox_known_answers():
    # Random123 known-answer vectors for philox4x32_10
    assert philox4x32_10([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert philox4x32_10([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert philox4x32_10([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]
    u = philox_uniform4(0, 0, 0, 0)
    assert np.all((u >= 0) & (u < 1))
    assert u[0] == np.float32((0x6627E8D5 >> 8) / 2 ** 24)


def test_add_particles_ring_bounds_and_wrap():
    cen = np.array([0.3, 0.966, 1.066])
    parts = np.full((100, 3), -1.0)
    out = O64.add_particles(parts, 90, 20, cen, 0.5, seed=7, step=3)
    new = np.r_[out[90:], out[:10]]
    assert np.all(out[10:90] == -1.0)
    assert np.all(np.abs(new[:, 0] - (cen[0] - 0.1)) <= 0.05 + 1e-12)
    r = np.hypot(new[:, 1] - cen[1], new[:, 2] - cen[2])
    assert np.all(r <= 0.45 + 0.05 * np.sqrt(2) + 1e-12) and np.all(r >= 0.45 - 0.05 * np.sqrt(2) - 1e-12)
    out32 = O32.add_particles(parts.astype(np.float32), 90, 20, cen, 0.5, seed=7, step=3)
    assert np.max(np.abs(out32 - out)) < 1e-6
    cyl = O64.fill_cylinder(np.zeros((5000, 3)), 0, 5000, cen, 0.5, 0.5, seed=1)
    assert np.all(np.abs(cyl[:, 0] - cen[0]) <= 0.25)
    rr = np.hypot(cyl[:, 1] - cen[1], cyl[:, 2] - cen[2])
    assert np.all(rr <= 0.5 + 1e-12)
    # uniform in the disk: P(r < 0.25) = 1/4, P(r < 0.5 / sqrt 2) = 1/2
    assert abs(np.mean(rr < 0.25) - 0.25) < 0.03 and abs(np.mean(rr < 0.5 / np.sqrt(2)) - 0.5) < 0.03


def test_spawn_ring_orientation_from_the_draws():
    """src/particles.jl:37-49: angle rt = 2 pi u0 measured from +y towards +z, x = c_x - 0.1,
    jitter (u1,u2,u3)*0.1 - 0.05.  The draws come from the Philox KAT-pinned generator."""
    cen = np.array([0.3, 0.966, 1.066])
    n, step, seed = 1000, 5, 9
    out = O64.add_particles(np.full((n, 3), -1.0), 0, n, cen, 0.5, seed, step)
    u0 = np.array([philox_uniform4(j, step, 0, seed)[0] for j in range(n)], np.float64)
    ang = np.sort(np.mod(np.arctan2(out[:, 2] - cen[2], out[:, 1] - cen[1]), 2 * np.pi))
    # the set of ring angles is the set of 2 pi u0 up to the jitter (|jitter| <= 0.05*sqrt2 / 0.4)
    want = np.sort(2 * np.pi * u0)
    assert np.mean(np.abs(ang - want)) < 0.05
    assert np.all(np.abs(out[:, 0] - (cen[0] - 0.1)) <= 0.05 + 1e-12)
    assert abs(np.mean(out[:, 0]) - (cen[0] - 0.1)) < 0.005
    # slots are in increasing angle order within the batch (ring order)
    a_slot = np.mod(np.arctan2(out[:, 2] - cen[2], out[:, 1] - cen[1]), 2 * np.pi)
    assert np.mean(np.diff(a_slot) >= -0.3) > 0.95


def test_fill_cylinder_is_an_ordered_permutation_of_the_draws():
    cen = np.array([0.3, 0.966, 1.066])
    n = 4000
    parts = O64.fill_cylinder(np.zeros((n, 3)), 0, n, cen, 0.5, 0.5, seed=3)
    u = np.array([philox_uniform4(i, 0xFFFFFFFF, 1, 3) for i in range(n)], np.float64)
    r = 0.5 * np.sqrt(u[:, 1])
    th = u[:, 2].astype(np.float32).astype(np.float64) * 2 * np.pi
    draws = np.stack([cen[0] + 0.5 * (u[:, 0] - 0.5), cen[1] + r * np.cos(th), cen[2] + r * np.sin(th)], -1)
    a = parts[np.lexsort(parts.T)]
    b = draws[np.lexsort(draws.T)]
    assert np.max(np.abs(a - b)) < 1e-6
    # slot order = increasing (x slab, theta wedge, r) key of the raw Philox words
    raw = np.array([philox4x32_10([i, 0xFFFFFFFF, 1, 0], [3, 0]) for i in range(n)], np.uint64)
    key = ((raw[:, 0] >> 27) << 19) | ((raw[:, 2] >> 25) << 12) | (raw[:, 1] >> 20)
    order = np.argsort(key, kind="stable")
    assert np.max(np.abs(parts - draws[order])) < 1e-6
    assert np.all(np.abs(parts[:, 0] - cen[0]) <= 0.25)
    assert np.all(np.hypot(parts[:, 1] - cen[1], parts[:, 2] - cen[2]) <= 0.5 + 1e-12)


def test_advect_constant_field_out_of_box_and_matlabport():
    n, L = (16, 8, 8), (4, 2, 2)
    d = np.array([L[a] / n[a] for a in range(3)])
    vel = np.zeros((n[2], n[1], n[0], 3)) + np.array([0.7, -0.2, 0.1])
    rng = np.random.default_rng(13)
    parts = rng.random((200, 3)) * np.array(L) * 0.8 + 0.1 * np.array(L)
    parts[:5] = [[-0.1, 1, 1], [4.0, 1, 1], [1, 0.0, 1], [1, 1, 2.5], [1, 1, 2.0]]
    out = O64.staggered_advect(parts, vel, 0.05, d, L)
    assert np.all(out[:5] == -1.0)
    assert np.max(np.abs(out[5:] - (parts[5:] + 0.05 * np.array([0.7, -0.2, 0.1])))) < 1e-13
    # random staggered field against MatlabPort StaggeredAdvect (src/MatlabPort.jl:201-273)
    vel = rng.standard_normal((n[2], n[1], n[0], 3))
    out = O64.staggered_advect(parts[5:], vel, 0.05, d, L)
    obj = mp.MatlabISF(L, n, 0.1, 0.05)
    vx, vy, vz = (mp.to_xyz(vel[..., a]) for a in range(3))
    px, py, pz = mp.staggered_advect(parts[5:, 0], parts[5:, 1], parts[5:, 2], obj, vx, vy, vz, 0.05)
    assert np.max(np.abs(out - np.stack([px, py, pz], -1))) < 1e-12


# ------------------------------------------------------------------ full step
def test_step_fp32_vs_fp64_small_jet():
    cfg = JetConfig(dims=(32, 16, 16), n_spawn=50, n_slots=2000, constraint_iters=3)
    r64 = OracleRun(cfg, np.float64)
    r32 = OracleRun(cfg, np.float32)
    assert rel(r32.psi, r64.psi) < 1e-5
    r64.step(4)
    r32.step(4)
    assert rel(r32.psi, r64.psi) < 1e-4
    alive = r64.parts[:, 0] > 0
    assert alive.sum() == 200
    assert np.percentile(np.abs(r32.parts[alive] - r64.parts[alive]), 99) < 1e-4
    assert abs(O64.norm2(r64.psi) - np.prod(cfg.dims)) < 1e-9 * np.prod(cfg.dims)


def test_step_matches_matlabport_sequence():
    """One simloop step (test/runtests.jl:45-64) against the same sequence in the
    MatlabPort restatement (examples/matlabport/simple.jl:102-124 order, reference
    generic-path restrict), FP64."""
    cfg = JetConfig(dims=(16, 8, 8), n_spawn=20, n_slots=100, constraint_iters=2)
    run = OracleRun(cfg, np.float64)
    psi0 = run.psi.copy()
    parts0 = run.parts.copy()
    run.step(1)
    obj = mp.MatlabISF(run.Lbox, run.n, run.hbar, run.dt)
    p1, p2 = (mp.to_xyz(a) for a in to_c(psi0))
    p1, p2 = mp.schroedinger_flow(obj, p1, p2)
    p1, p2 = mp.normalize(p1, p2)
    p1, p2 = mp.pressure_project(obj, p1, p2)
    jet = mp.to_xyz(run.isjet).astype(bool)
    ph = run.kvec[0] * obj.px + run.kvec[1] * obj.py + run.kvec[2] * obj.pz - run.omega * run.dt
    p1 = np.where(jet, np.abs(p1) * np.exp(1j * ph), p1)
    p2 = np.where(jet, np.abs(p2) * np.exp(1j * ph), p2)
    p1, p2 = mp.pressure_project(obj, p1, p2)
    assert rel(mp.to_xyz(to_c(run.psi)[0]), p1) < 1e-12
    assert rel(mp.to_xyz(to_c(run.psi)[1]), p2) < 1e-12
    vx, vy, vz = mp.staggered_sharp(obj, *mp.velocity_one_form(obj, p1, p2, run.hbar))
    spawned = Oracle(np.float64).add_particles(parts0, 0, 20, run.cen, 0.5, cfg.seed, 1)[:20]
    px, py, pz = mp.staggered_advect(spawned[:, 0], spawned[:, 1], spawned[:, 2], obj, vx, vy, vz, run.dt)
    assert np.max(np.abs(run.parts[:20] - np.stack([px, py, pz], -1))) < 1e-12
    assert np.all(run.parts[20:] == -1.0)


def test_library_fft_engine_matches_the_radix2_restatement():
    """The CPU baseline's FFT engine (MKL through torch.fft, installed as the oracle's 3-D
    transform) against the oracle's own radix-2 transform and numpy, both precisions, and
    one full step of the jet in FP64 with each engine."""
    from oracle.oracle import cpu_engine
    rng = np.random.default_rng(21)
    z = rng.standard_normal((8, 16, 32, 2))
    for dt, tol in ((np.float64, 1e-13), (np.float32, 1e-6)):
        o = Oracle(dt)
        zz = z.astype(dt)
        for inv in (False, True):
            a = o.fft3(zz.copy(), inverse=inv)
            with cpu_engine("mkl", 2):
                b = o.fft3(zz.copy(), inverse=inv)
            c = zz[..., 0] + 1j * zz[..., 1]
            ref = np.fft.ifftn(c) if inv else np.fft.fftn(c)
            assert rel(a, b) < tol, (dt, inv)
            assert rel(b[..., 0] + 1j * b[..., 1], ref) < tol, (dt, inv)
    cfg = JetConfig(dims=(32, 16, 16), vol_size=(2, 1, 1), nozzle_cen=(0.3, 0.466, 0.566), n_spawn=20,
                    n_slots=200, constraint_iters=2)
    r1 = OracleRun(cfg, np.float64).step(2)
    with cpu_engine("mkl", 2):
        r2 = OracleRun(cfg, np.float64).step(2)
    assert rel(r1.psi, r2.psi) < 1e-12
    assert np.max(np.abs(r1.parts - r2.parts)) < 1e-12
