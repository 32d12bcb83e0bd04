"""isf_amd -- host-side mirror of SchroedingersSmoke.jl's ISF API over libisf (gfx950).

The compute path is libisf.so (hand-written HIP kernels behind the C ABI of
include/isf.h); this package only marshals arguments.  It raises ISFError when the
library is missing or a call fails; there is no CPU fallback.
"""
from ._lib import ISFError, LIB_PATH, SIGNATURES, MemberDesc, RolloutDesc, build, load
from .api import (ISF, Batch, DeviceArray, add_particles_, isjet, normalize_psi_, pressure_project_, restrict_velocity_,
                  schroedinger_flow_, staggered_advect_, velocity_one_form_)
from .particles import ParticleExport, Particles
from .rollout import Rollout, gaussian_target, rollout_loss_function

__all__ = ["ISF", "Batch", "Particles", "ParticleExport", "Rollout", "gaussian_target", "rollout_loss_function", "MemberDesc", "RolloutDesc", "DeviceArray", "ISFError", "LIB_PATH", "SIGNATURES", "build", "load", "schroedinger_flow_",
           "normalize_psi_", "pressure_project_", "velocity_one_form_", "isjet", "restrict_velocity_",
           "add_particles_", "staggered_advect_"]
