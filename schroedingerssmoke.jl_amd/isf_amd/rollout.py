"""Differentiable rollout over libisf (isf_rollout_*): the driver's set-up and simloop as
a function of the nozzle velocity, a coarse initial phase field, the initial
wavefunction and the initial particle positions, with a Gaussian-splat particle-density
loss and gradients from the hand-written adjoint kernels (with step checkpointing)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import RolloutDesc, c_float3, check, load
from .api import DeviceArray, _ptr


def gaussian_target(box, res, seed, count=3, centre_lo=None, centre_hi=None, width=(0.15, 0.35),
                    amplitude=(0.002, 0.01)):
    """Target density on the splat grid (cell centres over the box): a sum of `count`
    isotropic Gaussians with seeded centres, widths and amplitudes.  float32, shape
    (res_z, res_y, res_x)."""
    rng = np.random.default_rng(seed)
    box = np.asarray(box, np.float64)
    lo = np.zeros(3) if centre_lo is None else np.asarray(centre_lo, np.float64)
    hi = box if centre_hi is None else np.asarray(centre_hi, np.float64)
    h = box / np.asarray(res)
    zz, yy, xx = np.meshgrid(*[(np.arange(res[a]) + 0.5) * h[a] for a in (2, 1, 0)], indexing="ij")
    out = np.zeros((res[2], res[1], res[0]))
    for _ in range(count):
        mu = rng.uniform(lo, hi)
        s = rng.uniform(*width)
        amp = rng.uniform(*amplitude)
        out += amp * np.exp(-((xx - mu[0]) ** 2 + (yy - mu[1]) ** 2 + (zz - mu[2]) ** 2) / (2 * s * s))
    return out.astype(np.float32)


class Rollout:
    """isf_rollout_t on an ISF handle whose jet is configured (ISF.jet_configure)."""

    def __init__(self, isf, steps, checkpoint_every=4, constraint_iters=10, phase_res=(16, 16, 16),
                 splat_res=(32, 16, 16), splat_sigma=0.25, splat_cutoff=4.0, nslots=100_000, n_init=None,
                 fill_seed=0):
        self.lib = load()
        self.isf = isf
        d = RolloutDesc()
        d.steps = int(steps)
        d.checkpoint_every = int(checkpoint_every)
        d.constraint_iters = int(constraint_iters)
        d.phase_res = (C.c_int * 3)(*[int(v) for v in (phase_res or (0, 0, 0))])
        d.splat_res = (C.c_int * 3)(*[int(v) for v in splat_res])
        d.splat_sigma = float(splat_sigma)
        d.splat_cutoff = float(splat_cutoff)
        d.nslots = int(nslots)
        d.n_init = int(nslots if n_init is None else n_init)
        d.fill_seed = int(fill_seed)
        self.desc = d
        self.phase_res = tuple(int(v) for v in (phase_res or (0, 0, 0)))
        self.splat_res = tuple(int(v) for v in splat_res)
        self.nslots = int(nslots)
        h = C.c_void_p()
        check(self.lib.isf_rollout_create(C.byref(h), isf.h, C.byref(d)), "isf_rollout_create")
        self.h = h.value
        self._theta = None
        self._psi_init = None

    @property
    def phase_shape(self):
        return self.phase_res[::-1]

    def set_target(self, target):
        t = DeviceArray.from_host(np.asarray(target, np.float32).reshape(self.splat_res[::-1]))
        check(self.lib.isf_rollout_set_target(self.h, t.ptr), "isf_rollout_set_target")

    def forward(self, velocity, theta=None, psi_init=None):
        """Loss (float) for nozzle velocity (3,), theta (host array of phase_shape or
        None) and psi_init (DeviceArray / host array in psi layout, or None)."""
        self._theta = None if theta is None else DeviceArray.from_host(
            np.asarray(theta, np.float32).reshape(self.phase_shape))
        if psi_init is not None and not isinstance(psi_init, DeviceArray):
            psi_init = self.isf.psi_array(np.asarray(psi_init, np.float32))
        self._psi_init = psi_init
        loss = C.c_double()
        check(self.lib.isf_rollout_forward(self.h, c_float3(*[float(np.float32(v)) for v in velocity]),
                                           _ptr(self._theta), _ptr(psi_init), C.byref(loss)), "isf_rollout_forward")
        return loss.value

    def backward(self, want_psi=True, want_parts=True):
        """Gradients of the last forward's loss: dict with velocity (3,), theta (or None),
        psi_init (psi layout) and parts_init (nslots, 3)."""
        gv = (C.c_double * 3)()
        gth = DeviceArray(self.phase_shape, np.float32) if self._theta is not None else None
        gpsi = DeviceArray(self.isf.shape3 + (4,), np.float32) if want_psi else None
        gparts = DeviceArray((self.nslots, 3), np.float32) if want_parts and self.nslots > 0 else None
        check(self.lib.isf_rollout_backward(self.h, gv, _ptr(gth), _ptr(gpsi), _ptr(gparts)), "isf_rollout_backward")
        return {"velocity": np.array(list(gv)), "theta": gth.to_host() if gth is not None else None,
                "psi_init": gpsi.to_host() if gpsi is not None else None,
                "parts_init": gparts.to_host() if gparts is not None else None}

    def state(self):
        """Final psi, particles and density of the last forward (host copies)."""
        ps, pa, rh = C.c_void_p(), C.c_void_p(), C.c_void_p()
        check(self.lib.isf_rollout_state(self.h, C.byref(ps), C.byref(pa), C.byref(rh)), "isf_rollout_state")
        psi = DeviceArray.borrowed(ps.value, self.isf.shape3 + (4,), np.float32).to_host()
        parts = DeviceArray.borrowed(pa.value, (self.nslots, 3), np.float32).to_host() if self.nslots else None
        rho = DeviceArray.borrowed(rh.value, self.splat_res[::-1], np.float64).to_host()
        return psi, parts, rho

    def close(self):
        if getattr(self, "h", None):
            self.lib.isf_rollout_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def rollout_loss_function():
    """A torch.autograd.Function over a Rollout: apply(rollout, velocity, theta) with CPU
    tensors returns the loss; backward returns the hand-written adjoint's gradients
    (the counterpart of the Julia binding's ChainRulesCore.rrule)."""
    import torch

    class RolloutLoss(torch.autograd.Function):
        @staticmethod
        def forward(ctx, rollout, velocity, theta):
            ctx.rollout = rollout
            ctx.has_theta = theta is not None
            loss = rollout.forward(velocity.detach().cpu().numpy(),
                                   None if theta is None else theta.detach().cpu().numpy())
            return torch.tensor(loss, dtype=torch.float64)

        @staticmethod
        def backward(ctx, gl):
            g = ctx.rollout.backward(want_psi=False, want_parts=False)
            gv = torch.as_tensor(g["velocity"], dtype=torch.float64) * gl
            gth = torch.as_tensor(g["theta"], dtype=torch.float64) * gl if ctx.has_theta else None
            return None, gv, gth

    return RolloutLoss
