"""The hand-written adjoints on the GPU (isf_vjp_*, isf_rollout_*) against the FP64
torch restatement of the same operators (oracle/torch_ref.py), whose autograd gives the
reference vector-Jacobian products.

Tolerances: operator VJPs 2e-5 relative L2 (FP32 against FP64 on well-conditioned
inputs); rollout gradients 2e-3 relative L2 against FP64 autograd at the same inputs
(the FP32 trajectory differs from FP64 by ~1e-6 and the gradient inherits it);
checkpointed == full-tape within 1e-5 (only the order of float atomics differs);
directional derivatives within 1e-3 of FP64 central finite differences taken where the
loss is smooth (two step sizes agree).
"""
import numpy as np
import pytest
import torch

from oracle.torch_ref import TorchRollout

pytestmark = pytest.mark.gpu

DIMS, BOX = (32, 16, 16), (4, 2, 2)
CEN, LEN, RAD = (0.3, 1 - 0.034, 1 + 0.066), 0.5, 0.5
SRES = (16, 8, 8)


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def to_c(x):
    return torch.stack([torch.complex(x[..., 0], x[..., 1]), torch.complex(x[..., 2], x[..., 3])])


def to_r(p):
    return torch.stack([p[0].real, p[0].imag, p[1].real, p[1].imag], -1)


def smooth_psi(dims, seed, amp2=0.4):
    rng = np.random.default_rng(seed)
    nx, ny, nz = dims
    z, y, x = np.meshgrid(np.arange(nz) / nz, np.arange(ny) / ny, np.arange(nx) / nx, indexing="ij")
    a = amp2 * np.sin(2 * np.pi * (x + y) + rng.uniform(0, 6)) + 0.3
    b = 0.3 * np.sin(2 * np.pi * (y - z) + rng.uniform(0, 6)) + 0.2 * np.cos(2 * np.pi * x)
    c = 0.25 * np.cos(2 * np.pi * (x - z) + rng.uniform(0, 6))
    pi = np.zeros((nz, ny, nx, 4))
    pi[..., 0] = np.cos(a) * np.cos(b)
    pi[..., 1] = np.cos(a) * np.sin(b)
    pi[..., 2] = np.sin(a) * np.cos(c)
    pi[..., 3] = np.sin(a) * np.sin(c)
    return pi


def torch_vjp(fn, inputs, ybar):
    xs = [torch.tensor(np.asarray(x, np.float64), requires_grad=True) for x in inputs]
    y = fn(*xs)
    (y * torch.as_tensor(np.asarray(ybar, np.float64))).sum().backward()
    return [x.grad.numpy() for x in xs]


@pytest.fixture(scope="module")
def env():
    from isf_amd import ISF, DeviceArray
    tr = TorchRollout(DIMS, BOX, 0.1, 1 / 48, CEN, LEN, RAD, 0, 7, 0, 0, 0, 0, 1, 0, (8, 8, 8), SRES, 0.25, 4.0,
                      np.zeros(SRES[::-1]))
    isf = ISF(BOX, DIMS, 0.1, 1 / 48)
    isf.jet_configure((10.0, 0.0, 0.0), 5.0, np.array(CEN, np.float32), LEN, RAD, 0, 7, 0)
    return tr, isf, DeviceArray


def dev(DeviceArray, a):
    return DeviceArray.from_host(np.ascontiguousarray(a, np.float32))


def test_vjp_schroedinger_flow(env):
    tr, isf, DA = env
    rng = np.random.default_rng(0)
    ybar = rng.standard_normal(isf.shape3 + (4,))
    (ref,) = torch_vjp(lambda x: to_r(tr.flow(to_c(x))), [smooth_psi(DIMS, 1)], ybar)
    b = dev(DA, ybar)
    isf.vjp_schroedinger_flow_(b)
    e = rel(b.to_host(), ref)
    print(f"  flow vjp rel {e:.2e}")
    assert e < 2e-5


def test_vjp_normalize(env):
    tr, isf, DA = env
    rng = np.random.default_rng(1)
    x = rng.standard_normal(isf.shape3 + (4,)) + 0.5
    ybar = rng.standard_normal(x.shape)
    (ref,) = torch_vjp(lambda t: to_r(tr.normalize(to_c(t))), [x], ybar)
    b = dev(DA, ybar)
    isf.vjp_normalize_(dev(DA, x), b)
    e = rel(b.to_host(), ref)
    print(f"  normalize vjp rel {e:.2e}")
    assert e < 2e-5


def test_vjp_pressure_project(env):
    tr, isf, DA = env
    rng = np.random.default_rng(2)
    x = smooth_psi(DIMS, 3)
    ybar = rng.standard_normal(x.shape)
    (ref,) = torch_vjp(lambda t: to_r(tr.project(to_c(t))), [x], ybar)
    b = dev(DA, ybar)
    isf.vjp_pressure_project_(dev(DA, x), b)
    e = rel(b.to_host(), ref)
    print(f"  pressure_project vjp rel {e:.2e}")
    assert e < 2e-5


@pytest.mark.parametrize("mode", [0, 1])
def test_vjp_restrict_velocity(env, mode):
    tr, isf, DA = env
    rng = np.random.default_rng(4 + mode)
    x = smooth_psi(DIMS, 5)
    ybar = rng.standard_normal(x.shape)
    kvec = np.array([8.0, 2.0, -1.0])
    om = 0.3
    tr.phase_mode = mode
    try:
        gx, gk, go = torch_vjp(lambda t, k, w: to_r(tr.restrict(to_c(t), k, w)), [x, kvec, np.array(om)], ybar)
    finally:
        tr.phase_mode = 0
    b = dev(DA, ybar)
    S = isf.vjp_restrict_velocity_(dev(DA, x), b, kvec, om, mode)
    e = rel(b.to_host(), gx)
    gk_gpu = np.array([S[0]] * 3) if mode == 1 else np.array(S[1:])
    print(f"  restrict vjp rel {e:.2e}; dk gpu {gk_gpu} ref {gk}; domega gpu {-S[0]:.6e} ref {float(go):.6e}")
    assert e < 2e-5
    assert rel(gk_gpu, gk) < 1e-4
    assert abs(-S[0] - float(go)) < 1e-4 * abs(float(go))


def test_vjp_velocity_one_form(env):
    tr, isf, DA = env
    rng = np.random.default_rng(6)
    x = smooth_psi(DIMS, 7)
    vbar = rng.standard_normal(isf.shape3 + (3,))
    hb = 0.1

    def fn(t):
        v = tr.one_form(to_c(t), hb)
        return torch.stack([v[a] / tr.d[a] for a in range(3)], -1)

    (ref,) = torch_vjp(fn, [x], vbar)
    b = dev(DA, np.zeros_like(x))
    isf.vjp_velocity_one_form_(dev(DA, x), dev(DA, vbar), b, hbar=hb, sharp=True)
    e = rel(b.to_host(), ref)
    print(f"  velocity one-form vjp rel {e:.2e}")
    assert e < 2e-5


def test_vjp_one_forms_tiled():
    """The pressure projection's and the velocity one-form's VJPs on rows of 64 cells, the
    tiled edge-wise one-form adjoint, against FP64 autograd (32-cell rows take the
    per-cell kernel, tested above)."""
    from isf_amd import ISF, DeviceArray as DA
    dims, box = (64, 16, 16), (8, 2, 2)
    tr = TorchRollout(dims, box, 0.1, 1 / 48, CEN, LEN, RAD, 0, 7, 0, 0, 0, 0, 1, 0, (8, 8, 8), SRES, 0.25, 4.0,
                      np.zeros(SRES[::-1]))
    isf = ISF(box, dims, 0.1, 1 / 48)
    rng = np.random.default_rng(12)
    x = smooth_psi(dims, 13)
    ybar = rng.standard_normal(x.shape)
    (ref,) = torch_vjp(lambda t: to_r(tr.project(to_c(t))), [x], ybar)
    b = dev(DA, ybar)
    isf.vjp_pressure_project_(dev(DA, x), b)
    e = rel(b.to_host(), ref)
    vbar = rng.standard_normal(isf.shape3 + (3,))

    def fn(t):
        v = tr.one_form(to_c(t), 0.1)
        return torch.stack([v[a] / tr.d[a] for a in range(3)], -1)

    (ref_v,) = torch_vjp(fn, [x], vbar)
    bv = dev(DA, np.zeros_like(x))
    isf.vjp_velocity_one_form_(dev(DA, x), dev(DA, vbar), bv, hbar=0.1, sharp=True)
    ev = rel(bv.to_host(), ref_v)
    print(f"  tiled: pressure_project vjp rel {e:.2e}, velocity one-form vjp rel {ev:.2e}")
    assert e < 2e-5
    assert ev < 2e-5


def test_vjp_staggered_advect(env):
    tr, isf, DA = env
    rng = np.random.default_rng(8)
    P = 4000
    parts = rng.uniform(0.05, 0.95, (P, 3)) * np.array(BOX)
    parts[:10] = -1.0  # dead slots
    nx, ny, nz = DIMS
    z, y, x = np.meshgrid(np.arange(nz) / nz, np.arange(ny) / ny, np.arange(nx) / nx, indexing="ij")
    vel = np.stack([1 + 0.5 * np.sin(2 * np.pi * (y + z)), 0.4 * np.cos(2 * np.pi * x), 0.3 * np.sin(2 * np.pi * (x - y))], -1)
    pbar = rng.standard_normal((P, 3))
    gp, gv = torch_vjp(lambda p, v: tr.advect(p, [v[..., 0], v[..., 1], v[..., 2]]), [parts, vel], pbar)
    b = dev(DA, pbar)
    vb = dev(DA, np.zeros_like(vel))
    isf.vjp_staggered_advect_(dev(DA, parts), dev(DA, vel), b, vb)
    g = b.to_host()
    per = np.linalg.norm(g - gp, axis=1) / np.maximum(np.linalg.norm(gp, axis=1), 1e-12)
    bad = int((per > 1e-4).sum())
    print(f"  advect vjp: particles off {bad}/{P}, median rel {np.median(per):.2e}; velbar rel {rel(vb.to_host(), gv):.2e}")
    assert np.all(g[:10] == 0)
    assert bad <= P // 500  # a stage position within FP32 rounding of a cell face may pick the other cell
    assert rel(vb.to_host(), gv) < 1e-3


# ---------------------------------------------------------------- rollout

def rollout_case(steps=4, K=4, nslots=1500, n_init=1000, n_spawn=50, dims=DIMS, box=BOX):
    from isf_amd import ISF
    from isf_amd.rollout import Rollout, gaussian_target
    hb = 0.1
    v = np.array([1.0, 0.05, -0.03], np.float32)
    kvec = v / np.float32(hb)
    omega = np.float32(np.sum(v ** np.float32(2)) / (np.float32(2) * np.float32(hb)))
    isf = ISF(box, dims, hb, 1 / 48)
    isf.jet_configure(kvec, omega, np.array(CEN, np.float32), LEN, RAD, n_spawn, 7, 0)
    tgt = gaussian_target(box, SRES, 3, centre_lo=(0.2, 0.6, 0.6), centre_hi=(0.8, 1.4, 1.4))
    r = Rollout(isf, steps, checkpoint_every=K, constraint_iters=4, phase_res=(8, 8, 8), splat_res=SRES,
                splat_sigma=0.25, splat_cutoff=4.0, nslots=nslots, n_init=n_init, fill_seed=9)
    r.set_target(tgt)
    tr = TorchRollout(dims, box, hb, 1 / 48, CEN, LEN, RAD, n_spawn, 7, 0, nslots, n_init, 9, steps, 4, (8, 8, 8),
                      SRES, 0.25, 4.0, tgt)
    return isf, r, tr, v


def test_rollout_forward_equals_isf_step():
    from isf_amd import DeviceArray
    isf, r, tr, v = rollout_case(steps=5, K=2)
    loss = r.forward(v)
    psi_r, parts_r, rho = r.state()
    # the same driver through isf_step
    nx, ny, nz = DIMS
    psi = np.zeros((nz, ny, nx, 4), np.float32)
    psi[..., 0] = 1.0
    psi[..., 2] = 0.01
    dpsi = isf.psi_array(psi)
    isf.normalize_(dpsi)
    isf.constrain_(dpsi, 4)
    parts = DeviceArray.from_host(np.full((1500, 3), -1.0, np.float32))
    isf.fill_nozzle_(parts, 0, 1000, 9)
    for it in range(1, 6):
        isf.step_(dpsi, parts, it)
    # the tape runs normalize and restrict as separate passes; the compiler may contract
    # their neighbouring float expressions differently from the fused ones
    ps, pp = dpsi.to_host(), parts.to_host()
    print(f"  rollout vs isf_step: psi rel L2 {rel(psi_r, ps):.2e} (bit-identical: {np.array_equal(psi_r, ps)}), "
          f"particles max |d| {np.abs(parts_r - pp).max():.2e}")
    assert rel(psi_r, ps) < 1e-6
    assert np.array_equal((parts_r > 0).all(1), (pp > 0).all(1))
    assert np.abs(parts_r - pp).max() < 1e-5
    with torch.no_grad():
        L, _, _, rho_t = tr.run(torch.tensor(v, dtype=torch.float64), steps=5)
    print(f"  rollout loss gpu {loss:.8e} FP64 {float(L):.8e}")
    assert abs(loss - float(L)) < 1e-4 * float(L)


def test_rollout_spawn_table_follows_the_nozzle():
    """The rollout draws its spawn batches once per nozzle and seed and copies them in at
    each step; moving the handle's nozzle between two forwards must redraw them, so the
    second forward equals that of a rollout whose first forward sees the moved nozzle."""
    from isf_amd.rollout import Rollout, gaussian_target
    isf, r, tr, v = rollout_case(steps=3, K=2)
    r2 = Rollout(isf, 3, checkpoint_every=2, constraint_iters=4, phase_res=(8, 8, 8), splat_res=SRES,
                 splat_sigma=0.25, splat_cutoff=4.0, nslots=1500, n_init=1000, fill_seed=9)
    r2.set_target(gaussian_target(BOX, SRES, 3, centre_lo=(0.2, 0.6, 0.6), centre_hi=(0.8, 1.4, 1.4)))
    r.forward(v)
    _, parts_old, _ = r.state()
    hb = np.float32(0.1)
    omega = np.float32(np.sum(v ** np.float32(2)) / (np.float32(2) * hb))
    isf.jet_configure(v / hb, omega, np.array(CEN, np.float32) + np.array([0.0, 0.05, -0.05], np.float32), LEN, RAD,
                      50, 7, 0)
    l1 = r.forward(v)
    _, p1, _ = r.state()
    l2 = r2.forward(v)
    _, p2, _ = r2.state()
    assert not np.array_equal(p1, parts_old)
    assert np.array_equal(p1, p2)
    assert l1 == l2


def test_backward_twice_after_one_forward():
    """The first backward after a forward reverses the forward's own tape of the last
    segment; a second one recomputes that segment.  The particle gradient, which reads
    the tape's psi through each step's velocity field and involves no atomic sums, is
    bit-identical.  (In this 5-step case the psi_init and velocity gradients are about
    1e-13 and 1e-10, at the level of the order noise of the advection adjoint's atomic
    scatter, and differ between any two backward calls.)"""
    isf, r, tr, v = rollout_case(steps=5, K=2)
    r.forward(v)
    g1 = r.backward()
    g2 = r.backward()
    assert np.linalg.norm(g1["parts_init"]) > 0
    assert np.array_equal(g1["parts_init"], g2["parts_init"])


# (64, 16, 16): rows of 64 cells take the tiled one-form adjoints (every variant the
# reverse step uses), (32, 16, 16) the per-cell ones
@pytest.mark.parametrize("dims,box", [(DIMS, BOX), ((64, 16, 16), (8, 2, 2))])
def test_rollout_gradients_match_autograd(dims, box):
    isf, r, tr, v = rollout_case(dims=dims, box=box)
    rng = np.random.default_rng(1)
    th = (0.3 * rng.standard_normal((8, 8, 8))).astype(np.float32)
    pi = smooth_psi(dims, 2).astype(np.float32)
    loss = r.forward(v, theta=th, psi_init=pi)
    g = r.backward()
    L, gt = tr.grad(v.astype(np.float64), th.astype(np.float64), psi_init=pi.astype(np.float64),
                    parts_init=tr.parts0.numpy())
    scale = np.linalg.norm(gt["psi_init"])
    e_psi = rel(g["psi_init"], gt["psi_init"])
    e_parts = rel(g["parts_init"], gt["parts_init"])
    print(f"  loss gpu {loss:.8e} FP64 {L:.8e}; grad psi_init rel {e_psi:.2e}; grad parts_init rel {e_parts:.2e}")
    print(f"  grad v gpu {g['velocity']} FP64 {gt['velocity']}; |grad theta| gpu {np.linalg.norm(g['theta']):.2e} "
          f"FP64 {np.linalg.norm(gt['theta']):.2e}; |grad psi_init| {scale:.2e}")
    assert abs(loss - L) < 1e-4 * L
    assert e_psi < 2e-3
    assert e_parts < 2e-3
    # velocity and phase field enter only as a global phase after the projection: zero gradient
    assert np.abs(g["velocity"]).max() < 1e-3 * scale
    assert np.linalg.norm(g["theta"]) < 1e-3 * scale


@pytest.mark.parametrize("sres,nslots", [((32, 16, 16), 1500), ((48, 24, 24), 600)])
def test_splat_paths_match_autograd(sres, nslots):
    """The splat loss and its particle gradient on finer splat grids than the other cases:
    (32, 16, 16) (h = 0.125: a support 17 cells wide, the kernels' wide register path) and
    (48, 24, 24) (a grid too large for the per-block LDS grids, so global atomics, and a
    support 25 cells wide, the adjoint's per-cell path), against the FP64 restatement."""
    from isf_amd import ISF
    from isf_amd.rollout import Rollout, gaussian_target
    hb, steps, n_init = 0.1, 2, (2 * nslots) // 3
    v = np.array([1.0, 0.05, -0.03], np.float32)
    isf = ISF(BOX, DIMS, hb, 1 / 48)
    isf.jet_configure(v / np.float32(hb), np.float32(np.sum(v ** np.float32(2)) / (np.float32(2) * np.float32(hb))),
                      np.array(CEN, np.float32), LEN, RAD, 50, 7, 0)
    tgt = gaussian_target(BOX, sres, 3, centre_lo=(0.2, 0.6, 0.6), centre_hi=(0.8, 1.4, 1.4))
    r = Rollout(isf, steps, checkpoint_every=2, constraint_iters=4, phase_res=(8, 8, 8), splat_res=sres,
                splat_sigma=0.25, splat_cutoff=4.0, nslots=nslots, n_init=n_init, fill_seed=9)
    r.set_target(tgt)
    tr = TorchRollout(DIMS, BOX, hb, 1 / 48, CEN, LEN, RAD, 50, 7, 0, nslots, n_init, 9, steps, 4, (8, 8, 8),
                      sres, 0.25, 4.0, tgt)
    pi = smooth_psi(DIMS, 2).astype(np.float32)
    loss = r.forward(v, psi_init=pi)
    g = r.backward()
    L, gt = tr.grad(v.astype(np.float64), None, psi_init=pi.astype(np.float64), parts_init=tr.parts0.numpy())
    e_parts = rel(g["parts_init"], gt["parts_init"])
    print(f"  splat {sres}: loss gpu {loss:.8e} FP64 {L:.8e}; grad parts_init rel {e_parts:.2e}")
    assert abs(loss - L) < 1e-4 * L
    assert e_parts < 2e-3
    r.close()
    isf.close()


def test_checkpointed_equals_full_tape():
    isf, r4, tr, v = rollout_case(steps=6, K=6)
    _, r2, _, _ = rollout_case(steps=6, K=2)
    pi = smooth_psi(DIMS, 5).astype(np.float32)
    l4 = r4.forward(v, psi_init=pi)
    l2 = r2.forward(v, psi_init=pi)
    assert l4 == l2
    g4, g2 = r4.backward(), r2.backward()
    print(f"  K=2 vs K=6: psi rel {rel(g2['psi_init'], g4['psi_init']):.2e}, parts rel "
          f"{rel(g2['parts_init'], g4['parts_init']):.2e}")
    assert rel(g2["psi_init"], g4["psi_init"]) < 1e-5
    assert rel(g2["parts_init"], g4["parts_init"]) < 1e-5


def test_rollout_directional_derivative_matches_fd():
    isf, r, tr, v = rollout_case()
    rng = np.random.default_rng(11)
    pi = smooth_psi(DIMS, 12).astype(np.float32)
    r.forward(v, psi_init=pi)
    g = r.backward()
    d = rng.standard_normal(pi.shape)
    p64 = pi.astype(np.float64)
    v64 = v.astype(np.float64)
    fd = [(tr.loss(v64, psi_init=p64 + e * d) - tr.loss(v64, psi_init=p64 - e * d)) / (2 * e) for e in (1e-5, 1e-6)]
    ad = float((g["psi_init"].astype(np.float64) * d).sum())
    print(f"  directional derivative along psi_init: FD {fd}, adjoint {ad:.8e}")
    assert abs(fd[0] - fd[1]) < 1e-4 * abs(fd[1])
    assert abs(ad - fd[1]) < 1e-3 * abs(fd[1])
    # single-particle coordinates (perturbing one slot keeps the other trajectories fixed)
    p0 = tr.parts0.numpy()
    live = np.where((p0 > 0).all(1))[0]
    checked = 0
    for s in live[::97][:8]:
        for a in range(3):
            fds = []
            for e in (1e-6, 1e-7):
                dp = np.zeros_like(p0)
                dp[s, a] = e
                fds.append((tr.loss(v64, psi_init=p64, parts_init=p0 + dp) -
                            tr.loss(v64, psi_init=p64, parts_init=p0 - dp)) / (2 * e))
            if abs(fds[0] - fds[1]) > 1e-4 * max(abs(fds[1]), 1e-12) or abs(fds[1]) < 1e-9:
                continue  # not smooth here (a cell face within reach) or no influence
            assert abs(g["parts_init"][s, a] - fds[1]) < 1e-3 * abs(fds[1]), (s, a, fds, g["parts_init"][s, a])
            checked += 1
    print(f"  single-particle FD checks: {checked}")
    assert checked >= 6


def test_torch_autograd_function():
    from isf_amd.rollout import rollout_loss_function
    isf, r, tr, v = rollout_case(steps=3, K=2)
    RolloutLoss = rollout_loss_function()
    vt = torch.tensor(v.astype(np.float64), requires_grad=True)
    th = torch.zeros((8, 8, 8), dtype=torch.float64, requires_grad=True)
    loss = RolloutLoss.apply(r, vt, th)
    loss.backward()
    g = r.backward(want_psi=False, want_parts=False)
    assert np.allclose(vt.grad.numpy(), g["velocity"])
    assert np.allclose(th.grad.numpy(), g["theta"])


# ---------------------------------------------------------------- the 128x64x64 rollout

def c4_case(steps, K, nslots, n_spawn=1000):
    from isf_amd import ISF
    from isf_amd.rollout import Rollout, gaussian_target
    dims, box = (128, 64, 64), (8, 4, 4)
    cen = (0.3, 2 - 0.034, 2 + 0.066)
    sres = (32, 16, 16)
    hb = 0.1
    v = np.array([1.0, 0.0, 0.0], np.float32)
    isf = ISF(box, dims, hb, 1 / 48)
    isf.jet_configure(v / np.float32(hb), np.float32(np.sum(v ** np.float32(2)) / (np.float32(2) * np.float32(hb))),
                      np.array(cen, np.float32), 0.5, 0.5, n_spawn, 7, 0)
    tgt = gaussian_target(box, sres, 3, centre_lo=(0.3, 1.5, 1.5), centre_hi=(2.0, 2.5, 2.5))
    r = Rollout(isf, steps, checkpoint_every=K, constraint_iters=10, phase_res=(16, 16, 16), splat_res=sres,
                splat_sigma=0.25, splat_cutoff=4.0, nslots=nslots, n_init=nslots, fill_seed=9)
    r.set_target(tgt)
    nx, ny, nz = dims
    z, y, x = np.meshgrid(np.arange(nz) / nz, np.arange(ny) / ny, np.arange(nx) / nx, indexing="ij")
    a = 0.3 * np.sin(2 * np.pi * (x + y) + 1.0) + 0.3
    b = 0.3 * np.sin(2 * np.pi * (y - z) + 2.0) + 0.2 * np.cos(2 * np.pi * x)
    pi = np.zeros((nz, ny, nx, 4))
    pi[..., 0] = np.cos(a) * np.cos(b)
    pi[..., 1] = np.cos(a) * np.sin(b)
    pi[..., 2] = np.sin(a)
    tr = TorchRollout(dims, box, hb, 1 / 48, cen, 0.5, 0.5, n_spawn, 7, 0, nslots, nslots, 9, steps, 10, (16, 16, 16),
                      sres, 0.25, 4.0, tgt)
    return isf, r, tr, v, pi


def test_c4_grid_directional_derivative_matches_fd():
    """128x64x64, 8 steps, 20k particles: the adjoint's directional derivative along a
    random psi_init direction against FP64 central differences (smooth here: the step
    sizes 1e-6 and 1e-7 agree; at 1e-5 the second-order term is ~1e-3)."""
    torch.set_num_threads(16)
    isf, r, tr, v, pi = c4_case(8, 4, 20_000)
    r.forward(v, psi_init=pi.astype(np.float32))
    g = r.backward(want_parts=False)
    d = np.random.default_rng(3).standard_normal(pi.shape)
    v64 = v.astype(np.float64)
    fd = [(tr.loss(v64, psi_init=pi + e * d) - tr.loss(v64, psi_init=pi - e * d)) / (2 * e) for e in (1e-6, 1e-7)]
    ad = float((g["psi_init"].astype(np.float64) * d).sum())
    print(f"  128x64x64 / 8 steps: FD {fd}, adjoint {ad:.8e}")
    assert abs(fd[0] - fd[1]) < 1e-4 * abs(fd[1])
    assert abs(ad - fd[1]) < 1e-3 * abs(fd[1])


def test_c4_rollout_checkpointing_at_scale():
    """The full differentiable-rollout configuration (128x64x64, 32 steps, 100k
    particles, 16^3 phase field): checkpoints every 4 steps against the whole tape."""
    import time
    isf, r4, _, v, pi = c4_case(32, 4, 100_000)
    _, r32, _, _, _ = c4_case(32, 32, 100_000)
    th = (0.2 * np.random.default_rng(5).standard_normal((16, 16, 16))).astype(np.float32)
    t0 = time.perf_counter()
    l4 = r4.forward(v, theta=th)
    t1 = time.perf_counter()
    g4 = r4.backward()
    t2 = time.perf_counter()
    l32 = r32.forward(v, theta=th)
    g32 = r32.backward()
    print(f"  loss {l4:.8e}; forward {t1 - t0:.3f} s, backward {t2 - t1:.3f} s; "
          f"K=4 vs K=32 psi rel {rel(g4['psi_init'], g32['psi_init']):.2e} parts rel "
          f"{rel(g4['parts_init'], g32['parts_init']):.2e}; grad v {g4['velocity']}")
    assert l4 == l32
    for k in ("psi_init", "parts_init", "theta"):
        assert np.all(np.isfinite(g4[k]))
    assert np.linalg.norm(g4["psi_init"]) > 0 and np.linalg.norm(g4["parts_init"]) > 0
    assert rel(g4["psi_init"], g32["psi_init"]) < 1e-5
    assert rel(g4["parts_init"], g32["parts_init"]) < 1e-5


# ---------------------------------------------------------------- the 128^3 adjoint batch

def c5_case(steps, K, nslots, i=0, n_spawn=1000, torch_ref=False, constraint_iters=10):
    """Rollout i of the batch of adjoint rollouts at 128^3 (box 4x4x4): its own seeded
    nozzle velocity (|v| in [0.5, 1.5] within 20 degrees of +x), spawn seed, fill seed
    and target density."""
    from isf_amd import ISF
    from isf_amd.rollout import Rollout, gaussian_target
    dims, box = (128, 128, 128), (4, 4, 4)
    cen = (0.3, 2 - 0.034, 2 + 0.066)
    sres = (16, 16, 16)
    hb = np.float32(0.1)
    rng = np.random.default_rng(20261203 + i)
    ct, ph, sp = rng.uniform(np.cos(np.radians(20)), 1.0), rng.uniform(0, 2 * np.pi), rng.uniform(0.5, 1.5)
    st = np.sqrt(1 - ct * ct)
    v = np.array([sp * ct, sp * st * np.cos(ph), sp * st * np.sin(ph)], np.float32)
    isf = ISF(box, dims, hb, 1 / 48)
    isf.jet_configure(v / hb, np.float32(np.sum(v ** np.float32(2)) / (np.float32(2) * hb)),
                      np.array(cen, np.float32), 0.5, 0.5, n_spawn, 1000 + i, 0)
    tgt = gaussian_target(box, sres, 3000 + i, centre_lo=(0.3, 1.5, 1.5), centre_hi=(2.0, 2.5, 2.5))
    r = Rollout(isf, steps, checkpoint_every=K, constraint_iters=constraint_iters, phase_res=(16, 16, 16),
                splat_res=sres, splat_sigma=0.25, splat_cutoff=4.0, nslots=nslots, n_init=nslots,
                fill_seed=2000 + i)
    r.set_target(tgt)
    tr = None
    if torch_ref:
        tr = TorchRollout(dims, box, float(hb), 1 / 48, cen, 0.5, 0.5, n_spawn, 1000 + i, 0, nslots, nslots,
                          2000 + i, steps, constraint_iters, (16, 16, 16), sres, 0.25, 4.0, tgt)
    return isf, r, tr, v


@pytest.mark.timeout(900)
def test_c5_adjoint_rollouts_at_128_cubed():
    """The adjoint batch of the large configuration: 128^3 rollouts of 32 steps with
    100k particles.  (a) Checkpoints every 4 steps give the whole-tape gradients; (b) eight
    rollouts with distinct nozzles and targets run concurrently (one host thread and one
    handle each) give the same losses and gradients as each run alone; (c) the adjoint's
    directional derivative along psi_init matches FP64 central differences on a 4-step,
    20k-particle rollout of the same grid (2 constraint iterations, to bound the CPU time
    of the FP64 loss evaluations)."""
    from concurrent.futures import ThreadPoolExecutor
    # (a)
    isf4, r4, _, v = c5_case(32, 4, 100_000)
    isf32, r32, _, _ = c5_case(32, 32, 100_000)
    th = (0.2 * np.random.default_rng(5).standard_normal((16, 16, 16))).astype(np.float32)
    l4, l32 = r4.forward(v, theta=th), r32.forward(v, theta=th)
    g4, g32 = r4.backward(), r32.backward()
    print(f"  128^3 / 32 steps: loss {l4:.8e}; K=4 vs K=32 psi rel {rel(g4['psi_init'], g32['psi_init']):.2e}, "
          f"parts rel {rel(g4['parts_init'], g32['parts_init']):.2e}", flush=True)
    assert l4 == l32
    assert np.linalg.norm(g4["psi_init"]) > 0 and np.linalg.norm(g4["parts_init"]) > 0
    # the particle gradient has no atomic sums: the recomputed segments reproduce the tape
    # exactly; the psi gradient (|g| ~ 1e-12 here) carries the summation-order noise of
    # the float atomics of the velocity scatter, about 4e-5 of its norm at this size
    assert np.array_equal(g4["parts_init"], g32["parts_init"])
    assert rel(g4["psi_init"], g32["psi_init"]) < 2e-4
    for o in (r4, r32, isf4, isf32):
        o.close()
    # (b)
    jobs = [c5_case(32, 4, 100_000, i=i) for i in range(8)]
    solo = []
    for isf, r, _, v in jobs:
        loss = r.forward(v)
        solo.append((loss, r.backward(want_psi=False)))

    def fwd_bwd(j):
        loss = j[1].forward(j[3])
        return loss, j[1].backward(want_psi=False)

    with ThreadPoolExecutor(8) as ex:
        conc = list(ex.map(fwd_bwd, jobs))
    for i, ((la, ga), (lb, gb)) in enumerate(zip(solo, conc)):
        assert la == lb, i
        assert np.array_equal(ga["parts_init"], gb["parts_init"]), i  # no atomics on this path
    print(f"  8 concurrent rollouts: losses {[round(x[0], 9) for x in conc]}", flush=True)
    for isf, r, _, _ in jobs:
        r.close()
        isf.close()
    # (c)
    torch.set_num_threads(16)
    isf, r, tr, v = c5_case(4, 4, 20_000, n_spawn=200, torch_ref=True, constraint_iters=2)
    pi = smooth_psi((128, 128, 128), 17)
    r.forward(v, psi_init=pi.astype(np.float32))
    g = r.backward(want_parts=False)
    d = np.random.default_rng(4).standard_normal(pi.shape)
    v64 = v.astype(np.float64)
    # central differences at h and h/2 and their Richardson extrapolation (the h^2 term
    # removed): along this direction the curvature term is ~4e-4 of the slope at h = 1e-6
    # and FP64 roundoff ~4e-5 at h = 1e-7, so h = 4e-7, 2e-7 balance the two
    fd = [(tr.loss(v64, psi_init=pi + e * d) - tr.loss(v64, psi_init=pi - e * d)) / (2 * e) for e in (4e-7, 2e-7)]
    rich = (4.0 * fd[1] - fd[0]) / 3.0
    ad = float((g["psi_init"].astype(np.float64) * d).sum())
    print(f"  128^3 / 4 steps: FD {fd}, Richardson {rich:.8e}, adjoint {ad:.8e}", flush=True)
    assert abs(fd[0] - fd[1]) < 1e-3 * abs(rich)
    assert abs(ad - rich) < 1e-3 * abs(rich)
    r.close()
    isf.close()
