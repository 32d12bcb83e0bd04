"""Python mirror of the SchroedingersSmoke.jl API over libisf (the MI355X path).

Names follow the reference exports (src/parallel.jl:264-266, src/particles.jl:97); a
trailing underscore stands for Julia's ``!`` (in place).  Arrays are ``DeviceArray``s in
the reference's layouts (x fastest): psi ``(nz, ny, nx, 4)`` float32 =
(ComplexF32, ComplexF32) per cell, velocity ``(nz, ny, nx, 3)``, particles ``(P, 3)``.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import ISFError, MemberDesc, c_float3, c_int3, check, load


class DeviceArray:
    """A device allocation with a numpy-like shape/dtype (ArrayType(hostarray))."""

    def __init__(self, shape, dtype=np.float32):
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        check(load().isf_malloc(C.byref(p), self.nbytes), "isf_malloc")
        self.ptr = p.value or 0
        self.owned = True

    @classmethod
    def borrowed(cls, ptr, shape, dtype=np.float32):
        """A view of device memory owned elsewhere (never freed through this object)."""
        d = cls.__new__(cls)
        d.shape = tuple(int(s) for s in shape)
        d.dtype = np.dtype(dtype)
        d.nbytes = int(np.prod(d.shape)) * d.dtype.itemsize
        d.ptr = int(ptr)
        d.owned = False
        return d

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        if d.nbytes:
            check(load().isf_memcpy_h2d(d.ptr, a.ctypes.data, d.nbytes), "isf_memcpy_h2d")
        return d

    def to_host(self):
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            check(load().isf_memcpy_d2h(out.ctypes.data, self.ptr, self.nbytes), "isf_memcpy_d2h")
        return out

    def copy_from(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        check(load().isf_memcpy_h2d(self.ptr, a.ctypes.data, self.nbytes), "isf_memcpy_h2d")
        return self

    def clone(self):
        d = DeviceArray(self.shape, self.dtype)
        if self.nbytes:
            check(load().isf_memcpy_d2d(d.ptr, self.ptr, self.nbytes), "isf_memcpy_d2d")
        return d

    def free(self):
        if getattr(self, "ptr", 0) and getattr(self, "owned", False):
            load().isf_free(self.ptr)
        self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _ptr(a):
    if a is None:
        return None
    if isinstance(a, DeviceArray):
        return a.ptr
    raise TypeError("expected a DeviceArray")


class ISF:
    """ISF{ArrayType,UInt32,Float32}(physical_size, dims, hbar, dt), src/parallel.jl:41-98."""

    def __init__(self, physical_size, dims, hbar, dt):
        self.lib = load()
        h = C.c_void_p()
        check(self.lib.isf_create(C.byref(h), c_int3(*[int(v) for v in dims]),
                                  c_int3(*[int(v) for v in physical_size]),
                                  float(np.float32(hbar)), float(np.float32(dt))), "isf_create")
        self.h = h.value
        self.owned = True
        self._read_grid()

    @classmethod
    def _borrowed(cls, h):
        """A view of a handle owned elsewhere (a batch member); never destroyed here."""
        isf = cls.__new__(cls)
        isf.lib = load()
        isf.h = h
        isf.owned = False
        isf._read_grid()
        return isf

    def _read_grid(self):
        res = c_int3()
        box = (C.c_float * 3)()
        d = (C.c_float * 3)()
        hb, dt_ = C.c_float(), C.c_float()
        check(self.lib.isf_get_grid(self.h, res, d, box, C.byref(hb), C.byref(dt_)), "isf_get_grid")
        self.grid_res = tuple(int(v) for v in res)
        self.physical_size = tuple(int(round(v)) for v in box)
        self.d = tuple(np.float32(v) for v in d)
        self.hbar = np.float32(hb.value)
        self.dt = np.float32(dt_.value)
        self.N = int(np.prod(self.grid_res))

    @property
    def shape3(self):
        nx, ny, nz = self.grid_res
        return (nz, ny, nx)

    @property
    def velocity(self):
        """isf.velocity (src/parallel.jl:14): a view of the handle's own field."""
        p = C.c_void_p()
        check(self.lib.isf_velocity_ptr(self.h, C.byref(p)), "isf_velocity_ptr")
        return DeviceArray.borrowed(p.value, self.shape3 + (3,), np.float32)

    def psi_array(self, host=None):
        if host is None:
            return DeviceArray(self.shape3 + (4,), np.float32)
        return DeviceArray.from_host(np.asarray(host, np.float32).reshape(self.shape3 + (4,)))

    def synchronize(self):
        check(self.lib.isf_synchronize(self.h), "isf_synchronize")

    def close(self):
        if getattr(self, "h", None):
            if getattr(self, "owned", True):
                self.lib.isf_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- reference operators
    def schroedinger_flow_(self, psi):
        check(self.lib.isf_schroedinger_flow(self.h, _ptr(psi)), "schroedinger_flow!")
        return psi

    def normalize_(self, psi):
        check(self.lib.isf_normalize(self.h, _ptr(psi)), "normalize_psi")
        return psi

    def velocity_one_form_(self, psi, hbar=1.0, sharp=False, out=None):
        check(self.lib.isf_velocity_one_form(self.h, _ptr(psi), float(np.float32(hbar)), int(bool(sharp)),
                                             _ptr(out)), "velocity_one_form!")
        return out if out is not None else self.velocity

    def staggered_sharp_(self, velocity=None):
        """StaggeredSharp in place (test/runtests.jl:61-62); default: isf.velocity."""
        check(self.lib.isf_staggered_sharp(self.h, _ptr(velocity)), "isf_staggered_sharp")
        return velocity

    def div_(self, f, velocity=None):
        check(self.lib.isf_div(self.h, _ptr(velocity), _ptr(f)), "div!")
        return f

    def poisson_solve(self, f, q=None):
        if q is None:
            q = DeviceArray(self.shape3, np.float32)
        check(self.lib.isf_poisson_solve(self.h, _ptr(f), _ptr(q)), "poisson_solve")
        return q

    def gauge_transform_(self, psi, q):
        check(self.lib.isf_gauge_transform(self.h, _ptr(psi), _ptr(q)), "gauge_transform")
        return psi

    def pressure_project_(self, psi):
        check(self.lib.isf_pressure_project(self.h, _ptr(psi)), "pressure_project!")
        return psi

    def isjet(self, nozzle_cen, nozzle_len, nozzle_rad):
        out = DeviceArray(self.shape3, np.uint8)
        check(self.lib.isf_isjet(self.h, c_float3(*nozzle_cen), float(nozzle_len), float(nozzle_rad), out.ptr),
              "isjet")
        return out

    def restrict_velocity_(self, psi, kvec, isjet=None, omgterm=0.0, phase_mode=0):
        check(self.lib.isf_restrict_velocity(self.h, _ptr(psi), _ptr(isjet), c_float3(*kvec), float(omgterm),
                                             int(phase_mode)), "restrict_velocity!")
        return psi

    def add_particles_(self, particles, first, n, nozzle_cen, nozzle_rad, seed, step):
        check(self.lib.isf_add_particles(self.h, _ptr(particles), particles.shape[0], int(first), int(n),
                                         c_float3(*nozzle_cen), float(nozzle_rad), int(seed), int(step)),
              "add_particles!")
        return particles

    def staggered_advect_(self, particles, velocity=None):
        check(self.lib.isf_staggered_advect(self.h, _ptr(particles), particles.shape[0], _ptr(velocity)),
              "staggered_advect!")
        return particles

    # ---- fused driver
    def jet_configure(self, kvec, omega, nozzle_cen, nozzle_len, nozzle_rad, n_spawn, seed, phase_mode=0):
        check(self.lib.isf_jet_configure(self.h, c_float3(*kvec), float(omega), c_float3(*nozzle_cen),
                                         float(nozzle_len), float(nozzle_rad), int(n_spawn), int(seed),
                                         int(phase_mode)), "isf_jet_configure")

    def fill_nozzle_(self, particles, first, n, seed):
        check(self.lib.isf_fill_nozzle(self.h, _ptr(particles), particles.shape[0], int(first), int(n), int(seed)),
              "isf_fill_nozzle")
        return particles

    def constrain_(self, psi, iterations=10):
        check(self.lib.isf_constrain(self.h, _ptr(psi), int(iterations)), "isf_constrain")
        return psi

    def step_(self, psi, particles, iteration):
        nslots = particles.shape[0] if particles is not None else 0
        check(self.lib.isf_step(self.h, _ptr(psi), _ptr(particles), nslots, int(iteration)), "isf_step")
        return psi

    def simloop_(self, psi, particles, first_iteration, count):
        """Steps first_iteration .. first_iteration + count - 1 (the driver's simloop)."""
        nslots = particles.shape[0] if particles is not None else 0
        check(self.lib.isf_simloop(self.h, _ptr(psi), _ptr(particles), nslots, int(first_iteration), int(count)),
              "isf_simloop")
        return psi

    def step_profile(self, psi, particles, iteration):
        """One step with a HIP event between kernel stages: {stage name: ms}."""
        n = self.lib.isf_step_stage_count()
        ms = (C.c_float * n)()
        nslots = particles.shape[0] if particles is not None else 0
        check(self.lib.isf_step_profile(self.h, _ptr(psi), _ptr(particles), nslots, int(iteration), ms),
              "isf_step_profile")
        return {self.lib.isf_step_stage_name(i).decode(): float(ms[i]) for i in range(n)}

    def set_sort_interval(self, steps):
        check(self.lib.isf_set_sort_interval(self.h, int(steps)), "isf_set_sort_interval")

    def norm2(self, psi):
        out = C.c_double()
        check(self.lib.isf_norm2(self.h, _ptr(psi), C.byref(out)), "isf_norm2")
        return out.value

    def step_norm2(self):
        out = C.c_double()
        check(self.lib.isf_step_norm2(self.h, C.byref(out)), "isf_step_norm2")
        return out.value

    # ---- vector-Jacobian products (cotangents in the layouts of the primal arrays)
    def vjp_schroedinger_flow_(self, psibar):
        check(self.lib.isf_vjp_schroedinger_flow(self.h, _ptr(psibar)), "isf_vjp_schroedinger_flow")
        return psibar

    def vjp_normalize_(self, psi_in, psibar):
        check(self.lib.isf_vjp_normalize(self.h, _ptr(psi_in), _ptr(psibar)), "isf_vjp_normalize")
        return psibar

    def vjp_pressure_project_(self, psi_in, psibar):
        check(self.lib.isf_vjp_pressure_project(self.h, _ptr(psi_in), _ptr(psibar)), "isf_vjp_pressure_project")
        return psibar

    def vjp_restrict_velocity_(self, psi_in, psibar, kvec, omgterm=0.0, phase_mode=0):
        """psibar in place; returns (sum phibar, sum phibar x, sum phibar y, sum phibar z)."""
        out = (C.c_double * 4)()
        check(self.lib.isf_vjp_restrict_velocity(self.h, _ptr(psi_in), _ptr(psibar), c_float3(*[float(v) for v in kvec]),
                                                 float(omgterm), int(phase_mode), out), "isf_vjp_restrict_velocity")
        return tuple(out)

    def vjp_velocity_one_form_(self, psi, velbar, psibar, hbar=1.0, sharp=False):
        check(self.lib.isf_vjp_velocity_one_form(self.h, _ptr(psi), float(hbar), int(bool(sharp)), _ptr(velbar),
                                                 _ptr(psibar)), "isf_vjp_velocity_one_form")
        return psibar

    def vjp_staggered_advect_(self, parts0, vel, pbar, velbar):
        check(self.lib.isf_vjp_staggered_advect(self.h, _ptr(parts0), parts0.shape[0], _ptr(vel), _ptr(pbar),
                                                _ptr(velbar)), "isf_vjp_staggered_advect")
        return pbar, velbar

    def timer_start(self):
        check(self.lib.isf_timer_start(self.h), "isf_timer_start")

    def timer_stop(self):
        ms = C.c_float()
        check(self.lib.isf_timer_stop(self.h, C.byref(ms)), "isf_timer_stop")
        return ms.value


class Batch:
    """A batched uneven ensemble (isf_batch_*): independent jets of mixed grid sizes stepped
    together, one hipGraph replay per step.  ``members`` is a list of dicts with the keys
    of ``isf_member_desc`` (res, box, hbar, dt, kvec, omega, cen, len, rad, n_spawn, seed,
    phase_mode)."""

    def __init__(self, members):
        self.lib = load()
        n = len(members)
        arr = (MemberDesc * n)()
        for i, m in enumerate(members):
            d = arr[i]
            d.res = c_int3(*[int(v) for v in m["res"]])
            d.box = c_int3(*[int(v) for v in m["box"]])
            d.hbar = float(np.float32(m["hbar"]))
            d.dt = float(np.float32(m["dt"]))
            d.kvec = c_float3(*[float(np.float32(v)) for v in m["kvec"]])
            d.omega = float(np.float32(m["omega"]))
            d.cen = c_float3(*[float(np.float32(v)) for v in m["cen"]])
            d.len = float(np.float32(m["len"]))
            d.rad = float(np.float32(m["rad"]))
            d.n_spawn = int(m.get("n_spawn", 0))
            d.seed = int(m.get("seed", 0))
            d.phase_mode = int(m.get("phase_mode", 0))
        h = C.c_void_p()
        check(self.lib.isf_batch_create(C.byref(h), n, arr), "isf_batch_create")
        self.h = h.value
        self.n = n
        self.members = []
        for i in range(n):
            mh = C.c_void_p()
            check(self.lib.isf_batch_member(self.h, i, C.byref(mh)), "isf_batch_member")
            self.members.append(ISF._borrowed(mh.value))
        self._bound = [None] * n

    def bind(self, i, psi, particles=None):
        nslots = particles.shape[0] if particles is not None else 0
        check(self.lib.isf_batch_bind(self.h, int(i), _ptr(psi), _ptr(particles), nslots), "isf_batch_bind")
        self._bound[i] = (psi, particles)  # keep the buffers alive while bound

    def step_(self, iteration):
        check(self.lib.isf_batch_step(self.h, int(iteration)), "isf_batch_step")

    def simloop_(self, first_iteration, count):
        """Steps first_iteration .. first_iteration + count - 1 of every member (the
        driver's simloop over the ensemble); equal to count step_ calls."""
        check(self.lib.isf_batch_simloop(self.h, int(first_iteration), int(count)), "isf_batch_simloop")

    def step_sequential_(self, iteration):
        check(self.lib.isf_batch_step_sequential(self.h, int(iteration)), "isf_batch_step_sequential")

    def set_graph(self, use_graph):
        check(self.lib.isf_batch_set_graph(self.h, int(bool(use_graph))), "isf_batch_set_graph")

    def synchronize(self):
        check(self.lib.isf_batch_synchronize(self.h), "isf_batch_synchronize")

    def tune(self, cache_path=None):
        check(self.lib.isf_batch_tune(self.h, cache_path.encode() if cache_path else None), "isf_batch_tune")

    def tile(self, res):
        ty = C.c_int()
        check(self.lib.isf_batch_tile(self.h, c_int3(*[int(v) for v in res]), C.byref(ty)), "isf_batch_tile")
        return ty.value

    def plane_elems(self, res):
        """Elements per thread of the plane pass for grid shape res (8 or 16)."""
        pe = C.c_int()
        check(self.lib.isf_batch_plane_elems(self.h, c_int3(*[int(v) for v in res]), C.byref(pe)),
              "isf_batch_plane_elems")
        return pe.value

    def close(self):
        if getattr(self, "h", None):
            for m in self.members:
                m.h = None
            self.lib.isf_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- module-level functions with the reference's export names
def schroedinger_flow_(isf: ISF, psi):
    return isf.schroedinger_flow_(psi)


def normalize_psi_(isf: ISF, psi):
    return isf.normalize_(psi)


def pressure_project_(isf: ISF, psi):
    return isf.pressure_project_(psi)


def velocity_one_form_(isf: ISF, psi, hbar=1.0):
    return isf.velocity_one_form_(psi, hbar)


def isjet(isf: ISF, nozzle_cen, nozzle_len, nozzle_rad):
    return isf.isjet(nozzle_cen, nozzle_len, nozzle_rad)


def restrict_velocity_(isf: ISF, psi, kvec, restrict_positions, omgterm=0.0, phase_mode=0):
    return isf.restrict_velocity_(psi, kvec, restrict_positions, omgterm, phase_mode)


def add_particles_(isf: ISF, particles, rng, nozzle_cen, nozzle_rad, seed=0, step=0):
    """add_particles!(particles, range, cen, rad) with a Julia 1-based range."""
    first = rng.start - 1
    return isf.add_particles_(particles, first, len(rng), nozzle_cen, nozzle_rad, seed, step)


def staggered_advect_(particles, isf: ISF):
    return isf.staggered_advect_(particles)


__all__ = ["ISF", "Batch", "DeviceArray", "ISFError", "schroedinger_flow_", "normalize_psi_", "pressure_project_",
           "velocity_one_form_", "isjet", "restrict_velocity_", "add_particles_", "staggered_advect_"]
