"""Loader of libisf.so (the HIP/gfx950 library behind include/isf.h).

There is no fallback: if the library is missing or a call fails, an exception says so.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_ROOT, "lib", "libisf.so")

_lib = None

c_float_p = C.POINTER(C.c_float)
c_int3 = C.c_int * 3
c_float3 = C.c_float * 3


class MemberDesc(C.Structure):
    """isf_member_desc of include/isf.h: one member of a batched ensemble."""
    _fields_ = [("res", C.c_int * 3), ("box", C.c_int * 3), ("hbar", C.c_float), ("dt", C.c_float),
                ("kvec", C.c_float * 3), ("omega", C.c_float), ("cen", C.c_float * 3), ("len", C.c_float),
                ("rad", C.c_float), ("n_spawn", C.c_longlong), ("seed", C.c_ulonglong), ("phase_mode", C.c_int)]

class RolloutDesc(C.Structure):
    """isf_rollout_desc of include/isf.h."""
    _fields_ = [("steps", C.c_int), ("checkpoint_every", C.c_int), ("constraint_iters", C.c_int),
                ("phase_res", C.c_int * 3), ("splat_res", C.c_int * 3), ("splat_sigma", C.c_float),
                ("splat_cutoff", C.c_float), ("nslots", C.c_longlong), ("n_init", C.c_longlong),
                ("fill_seed", C.c_ulonglong)]


# (name, argtypes) of every exported entry point of include/isf.h
SIGNATURES = {
    "isf_last_error": ([], C.c_char_p),
    "isf_version": ([], C.c_char_p),
    "isf_device_count": ([C.POINTER(C.c_int)], C.c_int),
    "isf_malloc": ([C.POINTER(C.c_void_p), C.c_size_t], C.c_int),
    "isf_free": ([C.c_void_p], C.c_int),
    "isf_memcpy_h2d": ([C.c_void_p, C.c_void_p, C.c_size_t], C.c_int),
    "isf_memcpy_d2h": ([C.c_void_p, C.c_void_p, C.c_size_t], C.c_int),
    "isf_memcpy_d2d": ([C.c_void_p, C.c_void_p, C.c_size_t], C.c_int),
    "isf_memset": ([C.c_void_p, C.c_int, C.c_size_t], C.c_int),
    "isf_create": ([C.POINTER(C.c_void_p), c_int3, c_int3, C.c_float, C.c_float], C.c_int),
    "isf_destroy": ([C.c_void_p], C.c_int),
    "isf_synchronize": ([C.c_void_p], C.c_int),
    "isf_get_grid": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "isf_velocity_ptr": ([C.c_void_p, C.POINTER(C.c_void_p)], C.c_int),
    "isf_schroedinger_flow": ([C.c_void_p, C.c_void_p], C.c_int),
    "isf_normalize": ([C.c_void_p, C.c_void_p], C.c_int),
    "isf_velocity_one_form": ([C.c_void_p, C.c_void_p, C.c_float, C.c_int, C.c_void_p], C.c_int),
    "isf_staggered_sharp": ([C.c_void_p, C.c_void_p], C.c_int),
    "isf_div": ([C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "isf_poisson_solve": ([C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "isf_gauge_transform": ([C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "isf_pressure_project": ([C.c_void_p, C.c_void_p], C.c_int),
    "isf_isjet": ([C.c_void_p, c_float3, C.c_float, C.c_float, C.c_void_p], C.c_int),
    "isf_restrict_velocity": ([C.c_void_p, C.c_void_p, C.c_void_p, c_float3, C.c_float, C.c_int], C.c_int),
    "isf_add_particles": ([C.c_void_p, C.c_void_p, C.c_longlong, C.c_longlong, C.c_longlong, c_float3, C.c_float,
                           C.c_ulonglong, C.c_uint], C.c_int),
    "isf_staggered_advect": ([C.c_void_p, C.c_void_p, C.c_longlong, C.c_void_p], C.c_int),
    "isf_jet_configure": ([C.c_void_p, c_float3, C.c_float, c_float3, C.c_float, C.c_float, C.c_longlong,
                           C.c_ulonglong, C.c_int], C.c_int),
    "isf_fill_nozzle": ([C.c_void_p, C.c_void_p, C.c_longlong, C.c_longlong, C.c_longlong, C.c_ulonglong], C.c_int),
    "isf_constrain": ([C.c_void_p, C.c_void_p, C.c_int], C.c_int),
    "isf_step": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_longlong, C.c_int], C.c_int),
    "isf_simloop": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_longlong, C.c_int, C.c_int], C.c_int),
    "isf_set_sort_interval": ([C.c_void_p, C.c_int], C.c_int),
    "isf_norm2": ([C.c_void_p, C.c_void_p, C.POINTER(C.c_double)], C.c_int),
    "isf_step_norm2": ([C.c_void_p, C.POINTER(C.c_double)], C.c_int),
    "isf_norm2_noalloc": ([C.c_void_p, C.c_longlong, C.c_void_p, C.c_void_p], C.c_int),
    "isf_step_stage_count": ([], C.c_int),
    "isf_step_stage_name": ([C.c_int], C.c_char_p),
    "isf_step_profile": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_longlong, C.c_int, C.POINTER(C.c_float)], C.c_int),
    "isf_particles_create": ([C.POINTER(C.c_void_p), C.c_longlong], C.c_int),
    "isf_particles_destroy": ([C.c_void_p], C.c_int),
    "isf_particles_append": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_longlong], C.c_int),
    "isf_particles_spawn_ring": ([C.c_void_p, C.c_void_p, C.c_longlong, C.c_ulonglong, C.c_uint], C.c_int),
    "isf_particles_remove_dead": ([C.c_void_p, C.c_void_p], C.c_int),
    "isf_particles_advect": ([C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "isf_particles_keep": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_longlong], C.c_int),
    "isf_particles_set_active": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_longlong], C.c_int),
    "isf_particles_state": ([C.c_void_p, C.c_void_p, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong),
                             C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                             C.POINTER(C.c_void_p)], C.c_int),
    "isf_mp_schroedinger_flow": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "isf_mp_normalize": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "isf_mp_velocity_one_form": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_float, C.c_void_p, C.c_void_p, C.c_void_p],
                                 C.c_int),
    "isf_mp_staggered_sharp": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "isf_mp_div": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "isf_mp_poisson_solve": ([C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "isf_mp_gauge_transform": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "isf_mp_pressure_project": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "isf_mp_restrict_velocity": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, c_float3, C.c_float, C.c_int],
                                 C.c_int),
    "isf_mp_staggered_velocity": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_longlong, C.c_void_p,
                                   C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "isf_mp_staggered_advect": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_float], C.c_int),
    "isf_export_create": ([C.POINTER(C.c_void_p), C.c_void_p, C.c_longlong], C.c_int),
    "isf_export_destroy": ([C.c_void_p], C.c_int),
    "isf_export_push": ([C.c_void_p, C.c_void_p], C.c_int),
    "isf_export_previous": ([C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_longlong)], C.c_int),
    "isf_export_latest": ([C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_longlong)], C.c_int),
    "isf_batch_create": ([C.POINTER(C.c_void_p), C.c_int, C.POINTER(MemberDesc)], C.c_int),
    "isf_batch_destroy": ([C.c_void_p], C.c_int),
    "isf_batch_member": ([C.c_void_p, C.c_int, C.POINTER(C.c_void_p)], C.c_int),
    "isf_batch_bind": ([C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_longlong], C.c_int),
    "isf_batch_step": ([C.c_void_p, C.c_int], C.c_int),
    "isf_batch_simloop": ([C.c_void_p, C.c_int, C.c_int], C.c_int),
    "isf_batch_set_graph": ([C.c_void_p, C.c_int], C.c_int),
    "isf_batch_step_sequential": ([C.c_void_p, C.c_int], C.c_int),
    "isf_batch_synchronize": ([C.c_void_p], C.c_int),
    "isf_batch_tune": ([C.c_void_p, C.c_char_p], C.c_int),
    "isf_batch_tile": ([C.c_void_p, c_int3, C.POINTER(C.c_int)], C.c_int),
    "isf_batch_plane_elems": ([C.c_void_p, c_int3, C.POINTER(C.c_int)], C.c_int),
    "isf_vjp_schroedinger_flow": ([C.c_void_p, C.c_void_p], C.c_int),
    "isf_vjp_normalize": ([C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "isf_vjp_pressure_project": ([C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "isf_vjp_restrict_velocity": ([C.c_void_p, C.c_void_p, C.c_void_p, c_float3, C.c_float, C.c_int,
                                   C.POINTER(C.c_double)], C.c_int),
    "isf_vjp_velocity_one_form": ([C.c_void_p, C.c_void_p, C.c_float, C.c_int, C.c_void_p, C.c_void_p], C.c_int),
    "isf_vjp_staggered_advect": ([C.c_void_p, C.c_void_p, C.c_longlong, C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "isf_rollout_create": ([C.POINTER(C.c_void_p), C.c_void_p, C.POINTER(RolloutDesc)], C.c_int),
    "isf_rollout_destroy": ([C.c_void_p], C.c_int),
    "isf_rollout_set_target": ([C.c_void_p, C.c_void_p], C.c_int),
    "isf_rollout_forward": ([C.c_void_p, c_float3, C.c_void_p, C.c_void_p, C.POINTER(C.c_double)], C.c_int),
    "isf_rollout_backward": ([C.c_void_p, C.POINTER(C.c_double), C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "isf_rollout_state": ([C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)], C.c_int),
    "isf_timer_start": ([C.c_void_p], C.c_int),
    "isf_timer_stop": ([C.c_void_p, C.POINTER(C.c_float)], C.c_int),
}


class ISFError(RuntimeError):
    pass


def build(verbose: bool = False) -> str:
    """Compile libisf.so for gfx950 with hipcc (the package Makefile)."""
    out = None if verbose else subprocess.DEVNULL
    subprocess.run(["make", "-s", "-C", PKG_ROOT, "-j4"], check=True, stdout=out)
    return LIB_PATH


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ISFError(f"{LIB_PATH} is not built; run __graft_entry__.build() (hipcc --offload-arch=gfx950)")
    lib = C.CDLL(LIB_PATH)
    for name, (args, res) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def check(rc: int, what: str = ""):
    if rc != 0:
        msg = load().isf_last_error().decode(errors="replace")
        raise ISFError(f"{what}: libisf error {rc}: {msg}")
