"""FP64 differentiable restatement of the ISF rollout (torch on the CPU).

TEST INFRASTRUCTURE ONLY: the reference gradient for the adjoint tests.  It restates
the same step as the C oracle (oracle/isf_oracle_impl.h, which cites the reference
file:line of every operation) with torch operations, so that torch.autograd gives the
gradient of the rollout loss with respect to the nozzle velocity and the initial phase
field.  Its forward values are pinned against the C oracle in tests/test_oracle.py.

Rollout (the reference driver, test/runtests.jl:16-66, plus the differentiable set-up):
  psi = normalize((1, 0.01)) * exp(i Phi),  Phi = trilinear periodic upsampling of a
        coarse phase field theta to the grid cells (cell x sits at theta coordinate
        x * pr / n per axis)
  constraint_iters x (restrict_velocity! with omega t = 0, pressure_project!)
  particles: n_init slots filled uniformly in the nozzle cylinder (constants)
  steps x simloop body (schroedinger_flow!, normalize, pressure_project!, add_particles!,
        restrict_velocity!(omega * it * dt), pressure_project!, velocity_one_form!(hbar)
        with 1/d, staggered_advect!)
  loss = mean over splat cells of (rho - target)^2, rho(y) = (1/nslots) * sum over live
        slots of exp(-|y - p|^2 / (2 sigma^2)) for |y - p| <= cutoff * sigma, y the
        centres of the splat grid cells covering the box.
Parameters: nozzle velocity v (kvec = v / hbar, omega = |v|^2 / (2 hbar),
test/runtests.jl:23-24) and theta.
"""
from __future__ import annotations

import math

import numpy as np
import torch

from .oracle import Oracle

AX = {0: -1, 1: -2, 2: -3}  # grid axis a -> tensor dim of a (nz, ny, nx) array


def psi_to_torch(psi_np):
    """(nz, ny, nx, 4) float -> complex128 (2, nz, ny, nx)."""
    p = torch.as_tensor(np.asarray(psi_np, np.float64))
    return torch.stack([torch.complex(p[..., 0], p[..., 1]), torch.complex(p[..., 2], p[..., 3])])


def psi_to_numpy(psi):
    p = psi.detach()
    return torch.stack([p[0].real, p[0].imag, p[1].real, p[1].imag], -1).numpy()


class TorchRollout:
    def __init__(self, dims, box, hbar, dt, cen, length, rad, n_spawn, seed, phase_mode, nslots, n_init,
                 fill_seed, steps, constraint_iters, phase_res, splat_res, sigma, cutoff, target):
        f32 = np.float32
        self.n = tuple(int(v) for v in dims)
        self.L = tuple(int(v) for v in box)
        nx, ny, nz = self.n
        self.N = nx * ny * nz
        self.hbar = float(f32(hbar))
        self.dt = float(f32(dt))
        self.d = [float(f32(self.L[a] / self.n[a])) for a in range(3)]
        self.cen = [float(f32(c)) for c in cen]
        self.length = float(f32(length))
        self.rad = float(f32(rad))
        self.n_spawn = int(n_spawn)
        self.seed = int(seed)
        self.phase_mode = int(phase_mode)
        self.nslots = int(nslots)
        self.steps = int(steps)
        self.constraint_iters = int(constraint_iters)
        self.phase_res = tuple(int(v) for v in phase_res)
        self.splat_res = tuple(int(v) for v in splat_res)
        self.sigma = float(f32(sigma))
        self.cutoff = float(f32(cutoff))
        self.o = Oracle(np.float64)
        fac, mask = self.o.tables(self.n, self.L, self.hbar, self.dt)
        self.fac = torch.as_tensor(fac[..., 0])
        self.mask = torch.complex(torch.as_tensor(mask[..., 0]), torch.as_tensor(mask[..., 1]))
        z, y, x = torch.meshgrid(torch.arange(nz, dtype=torch.float64), torch.arange(ny, dtype=torch.float64),
                                 torch.arange(nx, dtype=torch.float64), indexing="ij")
        self.pos = (x * self.d[0], y * self.d[1], z * self.d[2])
        jet = self.o.isjet(self.n, self.d, self.cen, self.length, self.rad)
        self.jet = torch.as_tensor(jet.astype(bool))
        parts = np.full((self.nslots, 3), -1.0)
        if n_init > 0:
            parts = self.o.fill_cylinder(parts, 0, int(n_init), self.cen, self.length, self.rad, int(fill_seed))
        self.parts0 = torch.as_tensor(parts)
        self.target = torch.as_tensor(np.asarray(target, np.float64).reshape(self.splat_res[::-1]))
        a = 1.0 / math.hypot(1.0, float(f32(0.01)))
        self.psi_c = (a, float(f32(0.01)) * a)

    # ---- operators (oracle/isf_oracle_impl.h restated)
    def flow(self, psi):
        return torch.fft.ifftn(self.mask * torch.fft.fftn(psi, dim=(-3, -2, -1)), dim=(-3, -2, -1))

    @staticmethod
    def normalize(psi):
        return psi / torch.sqrt((psi.real ** 2 + psi.imag ** 2).sum(0))

    def one_form(self, psi, hbar):
        out = []
        for a in range(3):
            z = (psi.conj() * torch.roll(psi, -1, AX[a])).sum(0)
            out.append(hbar * torch.atan2(z.imag, z.real))
        return out

    def div(self, v):
        return sum((v[a] - torch.roll(v[a], 1, AX[a])) / (self.d[a] * self.d[a]) for a in range(3))

    def poisson(self, f):
        return torch.fft.ifftn(self.fac * torch.fft.fftn(f.to(torch.complex128))).real

    def project(self, psi):
        q = self.poisson(self.div(self.one_form(psi, 1.0)))
        return psi * torch.exp(-1j * q)

    def restrict(self, psi, kvec, omgterm):
        if self.phase_mode == 1:
            ph = (kvec[0] + self.pos[0]) + (kvec[1] + self.pos[1]) + (kvec[2] + self.pos[2])
        else:
            ph = kvec[0] * self.pos[0] + kvec[1] * self.pos[1] + kvec[2] * self.pos[2]
        ph = ph - omgterm
        amp = torch.sqrt(psi.real ** 2 + psi.imag ** 2)
        new = amp * torch.exp(1j * ph)
        return torch.where(self.jet, new, psi)

    def spawn(self, parts, it):
        if self.n_spawn <= 0 or self.nslots <= 0:
            return parts
        P = self.nslots
        start1 = ((it - 1) * self.n_spawn + 1) % P
        first0 = P - 1 if start1 == 0 else start1 - 1
        fresh = self.o.add_particles(np.full((P, 3), -1.0), first0, self.n_spawn, self.cen, self.rad, self.seed, it)
        slots = (first0 + np.arange(self.n_spawn)) % P
        sel = torch.zeros(P, dtype=torch.bool)
        sel[torch.as_tensor(slots)] = True
        return torch.where(sel[:, None], torch.as_tensor(fresh), parts)

    def stag_vel(self, vel, x):
        """staggered_velocity with weights not divided by d; vel = list of 3 (nz, ny, nx)."""
        nx, ny, nz = self.n
        Lt = torch.tensor(self.L, dtype=torch.float64)
        dt_ = torch.tensor(self.d, dtype=torch.float64)
        nt = torch.tensor(self.n)
        xm = torch.remainder(x, Lt)
        i = torch.floor(xm.detach() / dt_).long()
        i = torch.where(i >= nt, i - nt, i)
        ip = torch.where(i + 1 == nt, torch.zeros_like(i), i + 1)
        w = xm - i.to(torch.float64) * dt_

        def V(comp, ix, iy, iz):
            return vel[comp].reshape(-1)[ix + nx * (iy + ny * iz)]

        I0, I1, I2 = i[:, 0], i[:, 1], i[:, 2]
        J0, J1, J2 = ip[:, 0], ip[:, 1], ip[:, 2]
        v0 = [V(0, I0, I1, I2), V(1, I0, I1, I2), V(2, I0, I1, I2)]
        vn = [V(0, I0, J1, J2), V(1, J0, I1, J2), V(2, J0, J1, I2)]
        pp = [V(0, I0, J1, I2), V(1, J0, I1, I2), V(2, J0, I1, I2)]
        pp2 = [V(0, I0, I1, J2), V(1, I0, I1, J2), V(2, I0, J1, I2)]
        w1 = [w[:, 2], w[:, 2], w[:, 1]]
        w2 = [w[:, 1], w[:, 0], w[:, 0]]
        u = [(1 - w1[a]) * ((1 - w2[a]) * v0[a] + w2[a] * pp[a]) + w1[a] * ((1 - w2[a]) * pp2[a] + w2[a] * vn[a])
             for a in range(3)]
        return torch.stack(u, 1)

    def advect(self, parts, vel):
        Lt = torch.tensor(self.L, dtype=torch.float64)
        live = ((parts > 0) & (parts < Lt)).all(1)
        dt = self.dt
        k1 = self.stag_vel(vel, parts)
        k2 = self.stag_vel(vel, parts + k1 * dt * 0.5)
        k3 = self.stag_vel(vel, parts + k2 * dt * 0.5)
        k4 = self.stag_vel(vel, parts + k3 * dt)
        new = parts + dt / 6.0 * (((k1 + 2 * k2) + 2 * k3) + k4)
        return torch.where(live[:, None], new, torch.full_like(parts, -1.0))

    def phase_field(self, theta):
        """Trilinear periodic upsampling of theta (pr_z, pr_y, pr_x) to the grid."""
        nx, ny, nz = self.n
        out = None
        coords = []
        for a, n in enumerate(self.n):
            pr = self.phase_res[a]
            s = torch.arange(n, dtype=torch.float64) * (pr / n)
            i0 = torch.floor(s).long()
            coords.append((i0 % pr, (i0 + 1) % pr, s - i0))
        (x0, x1, fx), (y0, y1, fy), (z0, z1, fz) = coords
        t = theta
        # interpolate along z, y, x
        tz = t[z0] * (1 - fz)[:, None, None] + t[z1] * fz[:, None, None]
        ty = tz[:, y0] * (1 - fy)[None, :, None] + tz[:, y1] * fy[None, :, None]
        out = ty[:, :, x0] * (1 - fx)[None, None, :] + ty[:, :, x1] * fx[None, None, :]
        return out

    def splat(self, parts, chunk=4096):
        h = [self.L[a] / self.splat_res[a] for a in range(3)]
        Lt = torch.tensor(self.L, dtype=torch.float64)
        ht = torch.tensor(h, dtype=torch.float64)
        live = ((parts > 0) & (parts < Lt)).all(1)
        P = parts[live]
        R = self.cutoff * self.sigma
        # cells j with |(j + 1/2) h - p| <= R lie within base - K .. base + K + 1,
        # base = floor(p / h - 1/2), K = ceil(R / h)
        K = [int(math.ceil(R / h[a])) for a in range(3)]
        offs = torch.stack(torch.meshgrid(*[torch.arange(-K[a], K[a] + 2) for a in range(3)], indexing="ij"), -1)
        offs = offs.reshape(-1, 3)
        res = torch.tensor(self.splat_res)
        M = self.splat_res[0] * self.splat_res[1] * self.splat_res[2]
        rho = torch.zeros(M, dtype=torch.float64)
        for c0 in range(0, P.shape[0], chunk):
            p = P[c0:c0 + chunk]
            base = torch.floor(p.detach() / ht - 0.5).long()
            idx = base[:, None, :] + offs[None, :, :]                 # (P, M, 3)
            inside = ((idx >= 0) & (idx < res)).all(-1)
            y = (idx.to(torch.float64) + 0.5) * ht
            r2 = ((y - p[:, None, :]) ** 2).sum(-1)
            ok = inside & (r2 <= R * R)
            g = torch.exp(-r2 / (2 * self.sigma ** 2)) * ok
            flat = idx[..., 0] + self.splat_res[0] * (idx[..., 1] + self.splat_res[1] * idx[..., 2])
            flat = torch.where(ok, flat, torch.zeros_like(flat))
            rho = rho.index_add(0, flat.reshape(-1), g.reshape(-1))
        return (rho / self.nslots).reshape(self.splat_res[::-1])

    # ---- rollout
    def initial_psi(self, kvec, theta, psi_init=None):
        nx, ny, nz = self.n
        if psi_init is None:
            ones = torch.ones((nz, ny, nx), dtype=torch.complex128)
            psi = torch.stack([ones * self.psi_c[0], ones * self.psi_c[1]])
        else:  # real (nz, ny, nx, 4) layout
            psi = torch.stack([torch.complex(psi_init[..., 0], psi_init[..., 1]),
                               torch.complex(psi_init[..., 2], psi_init[..., 3])])
        if theta is not None:
            psi = psi * torch.exp(1j * self.phase_field(theta))
        for _ in range(self.constraint_iters):
            psi = self.restrict(psi, kvec, 0.0)
            psi = self.project(psi)
        return psi

    def run(self, v, theta=None, steps=None, psi_init=None, parts_init=None):
        """(loss, psi, parts, rho) for nozzle velocity v (3,), phase field theta, initial
        psi (real (nz, ny, nx, 4) layout; default (1, 0.01) normalized) and initial
        particles (default: the nozzle fill)."""
        kvec = v / self.hbar
        omega = (v * v).sum() / (2 * self.hbar)
        psi = self.initial_psi(kvec, theta, psi_init)
        parts = self.parts0.clone() if parts_init is None else parts_init
        for it in range(1, (self.steps if steps is None else steps) + 1):
            psi = self.flow(psi)
            psi = self.normalize(psi)
            psi = self.project(psi)
            parts = self.spawn(parts, it)
            psi = self.restrict(psi, kvec, omega * (it * self.dt))
            psi = self.project(psi)
            vel = [u / self.d[a] for a, u in enumerate(self.one_form(psi, self.hbar))]
            parts = self.advect(parts, vel)
        rho = self.splat(parts)
        loss = ((rho - self.target) ** 2).mean()
        return loss, psi, parts, rho

    def grad(self, v, theta=None, psi_init=None, parts_init=None):
        """loss and gradients {velocity, theta, psi_init, parts_init} (None where the
        input was not given, except velocity)."""
        def leaf(a):
            return None if a is None else torch.tensor(np.asarray(a, np.float64), requires_grad=True)
        vt, th, pi, pa = leaf(v), leaf(theta), leaf(psi_init), leaf(parts_init)
        loss, _, _, _ = self.run(vt, th, psi_init=pi, parts_init=pa)
        loss.backward()
        g = {k: (None if t is None else t.grad.numpy().copy())
             for k, t in (("velocity", vt), ("theta", th), ("psi_init", pi), ("parts_init", pa))}
        return float(loss.detach()), g

    def loss(self, v, theta=None, psi_init=None, parts_init=None):
        def t(a):
            return None if a is None else torch.as_tensor(np.asarray(a, np.float64))
        with torch.no_grad():
            return float(self.run(t(v), t(theta), psi_init=t(psi_init), parts_init=t(parts_init))[0])
