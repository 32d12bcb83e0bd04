"""CPU pins of the FP64 torch restatement of the rollout (oracle/torch_ref.py), the
reference gradient of the adjoint tests:
  * its forward equals the C oracle (set-up and steps) to rounding,
  * the phase-field upsampling and the splat density match closed forms,
  * its autograd gradient matches central finite differences where the loss is smooth,
  * the gradient with respect to the nozzle velocity and the initial phase field vanishes:
    both only add a single-valued scalar to the phase of psi, which the exact discrete
    pressure projection maps to a global phase (a property of the discretisation the
    reference uses, checked here by autograd and by finite differences)."""
import math

import numpy as np
import pytest
import torch

from oracle.oracle import JetConfig, OracleRun
from oracle.torch_ref import TorchRollout, psi_to_numpy


def small_rollout(steps=3, target=None, phase_res=(8, 8, 8), hbar=0.125, v=(1.0, 0.25, -0.125), n_spawn=40,
                  nslots=1500, n_init=600, constraint_iters=4):
    cfg = JetConfig(vol_size=(4, 2, 2), dims=(32, 16, 16), hbar=hbar, dt=1 / 48, jet_velocity=v,
                    nozzle_cen=(0.3, 1 - 0.034, 1 + 0.066), n_spawn=n_spawn, n_slots=nslots, n_init=n_init,
                    constraint_iters=constraint_iters, seed=77)
    sres = (16, 8, 8)
    tr = TorchRollout(cfg.dims, cfg.vol_size, cfg.hbar, cfg.dt, cfg.nozzle_cen, cfg.nozzle_len, cfg.nozzle_rad,
                      cfg.n_spawn, cfg.seed, cfg.phase_mode, cfg.n_slots, cfg.n_init, cfg.seed, steps,
                      cfg.constraint_iters, phase_res, sres, 0.25, 4.0,
                      np.zeros(sres[::-1]) if target is None else target)
    return cfg, tr


def test_forward_matches_c_oracle():
    # hbar = 1/8 and v with power-of-two parts make kvec and omega exact in Float32 too
    cfg, tr = small_rollout()
    o = OracleRun(cfg, np.float64)
    v = torch.tensor(cfg.jet_velocity, dtype=torch.float64)
    with torch.no_grad():
        psi0 = tr.initial_psi(v / tr.hbar, None)
        _, psi, parts, _ = tr.run(v, None)
    assert np.abs(psi_to_numpy(psi0) - o.psi).max() < 1e-13
    o.step(3)
    assert np.abs(psi_to_numpy(psi) - o.psi).max() < 1e-13
    assert np.abs(parts.numpy() - o.parts).max() < 1e-13


def test_phase_field_upsampling_closed_form():
    _, tr = small_rollout()
    rng = np.random.default_rng(0)
    th = torch.tensor(rng.standard_normal((8, 8, 8)))
    phi = tr.phase_field(th).numpy()  # grid 32x16x16, theta 8^3: x step 4 cells, y/z 2 cells
    assert phi.shape == (16, 16, 32)
    t = th.numpy()
    # nodes coincide with cells (0 mod 4 in x, 0 mod 2 in y, z)
    assert np.allclose(phi[::2, ::2, ::4], t, atol=1e-14)
    # midway in x between nodes: the mean of the two nodes (periodic)
    assert np.allclose(phi[::2, ::2, 2::4], 0.5 * (t + np.roll(t, -1, 2)), atol=1e-14)
    # a linear ramp in theta is reproduced linearly between nodes
    ramp = torch.tensor(np.broadcast_to(np.arange(8.0)[None, None, :], (8, 8, 8)).copy())
    pr = tr.phase_field(ramp).numpy()
    assert np.allclose(pr[0, 0, :28], np.arange(28) / 4.0, atol=1e-14)


def test_splat_density_closed_form():
    """One particle: the density sums to the Gaussian's integral over the truncation
    sphere divided by the cell volume (grid spacing = sigma, so the midpoint sum is exact
    to ~1e-8), times 1/nslots."""
    _, tr = small_rollout()
    p = torch.tensor([[2.01, 1.03, 0.97]], dtype=torch.float64)
    tr.nslots = 7
    rho = tr.splat(p).numpy()
    h = [tr.L[a] / tr.splat_res[a] for a in range(3)]
    s, c = tr.sigma, tr.cutoff
    # P(chi_3 <= c) for a standard 3-D normal
    frac = math.erf(c / math.sqrt(2)) - math.sqrt(2 / math.pi) * c * math.exp(-c * c / 2)
    expect = (2 * math.pi) ** 1.5 * s ** 3 * frac / (h[0] * h[1] * h[2]) / 7
    assert abs(rho.sum() - expect) / expect < 2e-3
    # dead particles contribute nothing
    assert tr.splat(torch.tensor([[-1.0, -1.0, -1.0]], dtype=torch.float64)).abs().max() == 0


def smooth_psi_init(dims, seed):
    rng = np.random.default_rng(seed)
    nx, ny, nz = dims
    z, y, x = np.meshgrid(np.arange(nz) / nz, np.arange(ny) / ny, np.arange(nx) / nx, indexing="ij")
    a = 0.3 * np.sin(2 * np.pi * (x + y) + rng.uniform(0, 6)) + 0.2 * np.cos(2 * np.pi * z)
    b = 0.2 * np.sin(2 * np.pi * (y - z) + rng.uniform(0, 6))
    pi = np.zeros((nz, ny, nx, 4))
    pi[..., 0] = np.cos(a) * np.cos(b)
    pi[..., 1] = np.cos(a) * np.sin(b)
    pi[..., 2] = np.sin(a)
    return pi


def test_autograd_matches_finite_differences_on_psi():
    from isf_amd.rollout import gaussian_target
    tgt = gaussian_target((4, 2, 2), (16, 8, 8), 3, centre_lo=(0.2, 0.6, 0.6), centre_hi=(0.8, 1.4, 1.4))
    _, tr = small_rollout(steps=4, target=tgt, hbar=0.1, v=(1.0, 0.05, -0.03), n_spawn=50, nslots=1500, n_init=1000)
    rng = np.random.default_rng(1)
    v = np.array([1.0, 0.05, -0.03])
    th = 0.3 * rng.standard_normal((8, 8, 8))
    pi = smooth_psi_init(tr.n, 2)
    L, g = tr.grad(v, th, psi_init=pi)
    d = rng.standard_normal(pi.shape)
    ad = float((g["psi_init"] * d).sum())
    fd = [(tr.loss(v, th, psi_init=pi + e * d) - tr.loss(v, th, psi_init=pi - e * d)) / (2 * e) for e in (1e-5, 1e-6)]
    assert abs(fd[0] - fd[1]) < 1e-4 * abs(fd[1])  # smooth around the base point
    assert abs(fd[1] - ad) < 1e-5 * abs(ad)


def test_velocity_and_phase_gradients_vanish():
    from isf_amd.rollout import gaussian_target
    tgt = gaussian_target((4, 2, 2), (16, 8, 8), 3, centre_lo=(0.2, 0.6, 0.6), centre_hi=(0.8, 1.4, 1.4))
    _, tr = small_rollout(steps=4, target=tgt, hbar=0.1, v=(1.0, 0.05, -0.03), n_spawn=50, nslots=1500, n_init=1000)
    rng = np.random.default_rng(3)
    v = np.array([1.0, 0.05, -0.03])
    th = 0.3 * rng.standard_normal((8, 8, 8))
    pi = smooth_psi_init(tr.n, 4)
    L, g = tr.grad(v, th, psi_init=pi)
    scale = np.linalg.norm(g["psi_init"])  # gradient size of a generic input of unit scale
    assert np.abs(g["velocity"]).max() < 1e-9 * scale
    assert np.linalg.norm(g["theta"]) < 1e-9 * scale
    e = 1e-6
    for a in range(3):
        dv = np.zeros(3)
        dv[a] = e
        fd = (tr.loss(v + dv, th, psi_init=pi) - tr.loss(v - dv, th, psi_init=pi)) / (2 * e)
        assert abs(fd) < 1e-6 * scale
