"""Ensembles of independent jets for the batched path (isf_batch_*).

``draw_jets`` draws the parameters of an uneven ensemble: each member is the driver's
jet (test/runtests.jl:3-13,23-25) on its own grid, with the nozzle velocity, hbar and
seed varied per member.  ``init_member`` runs the driver's set-up on one member
(test/runtests.jl:19-30: psi = (1, 0.01) normalized, ten constraint iterations) and
fills its particle buffer inside the nozzle cylinder.
"""
from __future__ import annotations

import math
import os

import numpy as np

from .api import Batch, DeviceArray

# grid shapes of the uneven ensemble, each with 16 cells per unit length
ENSEMBLE_GRIDS = ((32, 32, 32), (64, 32, 32), (64, 64, 64), (128, 64, 64))

# Schedule cache (isf_batch_tune format: nx ny nz members x_tile plane_elements) written by
# the autotuner on an MI355X for the default 48-jet draw (draw_jets()); Batch.tune(path)
# reads it instead of timing the candidates again.
ENSEMBLE48_SCHEDULES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "schedules",
                                    "mi355x_ensemble48.txt")


def draw_jets(count=48, seed=20261203, grids=ENSEMBLE_GRIDS, n_slots=10_000, n_spawn=100, dt=1.0 / 48.0,
              hbar_range=(0.05, 0.2), speed_range=(0.5, 1.5), max_angle_deg=20.0, phase_mode=0):
    """Member descriptors (dicts with the fields of isf_member_desc, plus n_slots).

    grid uniform over ``grids`` with box = grid / 16; hbar uniform in ``hbar_range``; nozzle
    velocity of magnitude uniform in ``speed_range`` and direction uniform on the spherical
    cap within ``max_angle_deg`` of +x; kvec = v / hbar and omega = |v|^2 / (2 hbar) in
    Float32 (test/runtests.jl:23-24); nozzle centre 0.3 from the x = 0 face, near the
    middle of the y-z section, length and radius 0.5 (test/runtests.jl:9-12).
    """
    rng = np.random.default_rng(seed)
    f32 = np.float32
    out = []
    cmin = math.cos(math.radians(max_angle_deg))
    for i in range(count):
        res = tuple(int(v) for v in grids[int(rng.integers(len(grids)))])
        box = tuple(v // 16 for v in res)
        hbar = f32(rng.uniform(*hbar_range))
        speed = rng.uniform(*speed_range)
        ct = rng.uniform(cmin, 1.0)
        st = math.sqrt(max(0.0, 1.0 - ct * ct))
        ph = rng.uniform(0.0, 2.0 * math.pi)
        v = np.array([speed * ct, speed * st * math.cos(ph), speed * st * math.sin(ph)], f32)
        kvec = v / hbar
        omega = f32(np.sum(v ** f32(2)) / (f32(2) * hbar))
        cen = (f32(0.3), f32(box[1] / 2 - 0.034), f32(box[2] / 2 + 0.066))
        out.append(dict(res=res, box=box, hbar=hbar, dt=f32(dt), kvec=tuple(kvec), omega=omega, cen=cen,
                        len=f32(0.5), rad=f32(0.5), n_spawn=int(n_spawn), seed=int(seed) + 7919 * (i + 1),
                        phase_mode=int(phase_mode), n_slots=int(n_slots), velocity=tuple(v)))
    return out


def init_member(isf, member, n_init=None, constraint_iters=10):
    """Device psi and particles of one member after the driver's set-up."""
    nx, ny, nz = isf.grid_res
    psi = np.zeros((nz, ny, nx, 4), np.float32)
    psi[..., 0] = 1.0
    psi[..., 2] = 0.01
    dpsi = isf.psi_array(psi)
    isf.normalize_(dpsi)
    isf.constrain_(dpsi, constraint_iters)
    n_slots = int(member.get("n_slots", 0))
    parts = None
    if n_slots > 0:
        parts = DeviceArray.from_host(np.full((n_slots, 3), -1.0, np.float32))
        n_init = n_slots if n_init is None else int(n_init)
        if n_init > 0:
            isf.fill_nozzle_(parts, 0, n_init, member["seed"])
    return dpsi, parts


def build_batch(members, n_init=None, constraint_iters=10):
    """A Batch over ``members`` with every member set up and bound."""
    b = Batch(members)
    for i, m in enumerate(members):
        psi, parts = init_member(b.members[i], m, n_init, constraint_iters)
        b.bind(i, psi, parts)
    return b
