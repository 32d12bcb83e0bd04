"""The MatlabPort particle store over libisf (isf_particles_*): Particles{x, y, z,
length, active} (src/MatlabPort.jl:178-230) held on the device, with append!, the
driver's ring births, removethedead! (examples/matlabport/simple.jl:63-70),
StaggeredAdvect over the active slots and Keep."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, load
from .api import DeviceArray, _ptr


class Particles:
    def __init__(self, isf, capacity):
        self.lib = load()
        self.isf = isf
        self.capacity = int(capacity)
        h = C.c_void_p()
        check(self.lib.isf_particles_create(C.byref(h), self.capacity), "isf_particles_create")
        self.h = h.value

    def append_(self, x, y, z):
        """append!(p, x, y, z) with host or device arrays."""
        arrs = [a if isinstance(a, DeviceArray) else DeviceArray.from_host(np.asarray(a, np.float32)) for a in (x, y, z)]
        n = arrs[0].shape[0]
        check(self.lib.isf_particles_append(self.isf.h, self.h, arrs[0].ptr, arrs[1].ptr, arrs[2].ptr, n),
              "isf_particles_append")
        self.isf.synchronize()  # the host arrays are freed on return

    def spawn_ring_(self, n, seed, step):
        check(self.lib.isf_particles_spawn_ring(self.isf.h, self.h, int(n), int(seed), int(step)),
              "isf_particles_spawn_ring")

    def remove_dead_(self):
        check(self.lib.isf_particles_remove_dead(self.isf.h, self.h), "isf_particles_remove_dead")

    def advect_(self, velocity=None):
        check(self.lib.isf_particles_advect(self.isf.h, self.h, _ptr(velocity)), "isf_particles_advect")

    def keep_(self, ind):
        d = DeviceArray.from_host(np.asarray(ind, np.int32))
        check(self.lib.isf_particles_keep(self.isf.h, self.h, d.ptr, d.shape[0]), "isf_particles_keep")

    def set_active_(self, active):
        """Replace the active list (0-based slot indices), e.g. after a host-side filter!."""
        a = np.ascontiguousarray(np.asarray(active, np.int32))
        check(self.lib.isf_particles_set_active(self.isf.h, self.h, a.ctypes.data if a.size else None, a.size),
              "isf_particles_set_active")

    def state(self):
        """(length, active indices, x, y, z) as host arrays (x, y, z over all capacity slots)."""
        ln, cnt = C.c_longlong(), C.c_longlong()
        px, py, pz, pa = C.c_void_p(), C.c_void_p(), C.c_void_p(), C.c_void_p()
        check(self.lib.isf_particles_state(self.isf.h, self.h, C.byref(ln), C.byref(cnt), C.byref(px), C.byref(py),
                                           C.byref(pz), C.byref(pa)), "isf_particles_state")
        cap = self.capacity
        x = DeviceArray.borrowed(px.value, (cap,), np.float32).to_host()
        y = DeviceArray.borrowed(py.value, (cap,), np.float32).to_host()
        z = DeviceArray.borrowed(pz.value, (cap,), np.float32).to_host()
        act = DeviceArray.borrowed(pa.value, (max(cnt.value, 1),), np.int32).to_host()[:cnt.value]
        return ln.value, act, x, y, z

    def close(self):
        if getattr(self, "h", None):
            self.lib.isf_particles_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ParticleExport:
    """isf_export_*: double-buffered asynchronous copies of a particle buffer (nslots x 3
    floats on the device) into pinned host memory, the per-frame `Array(particles)` of
    examples/jet.jl:83 without stalling the simulation stream."""

    def __init__(self, isf, nslots):
        self.lib = load()
        self.nslots = int(nslots)
        h = C.c_void_p()
        check(self.lib.isf_export_create(C.byref(h), isf.h, self.nslots), "isf_export_create")
        self.h = h.value

    def push(self, particles):
        check(self.lib.isf_export_push(self.h, _ptr(particles)), "isf_export_push")

    def _get(self, fn, name):
        p, f = C.c_void_p(), C.c_longlong()
        check(fn(self.h, C.byref(p), C.byref(f)), name)
        arr = np.ctypeslib.as_array(C.cast(p.value, C.POINTER(C.c_float)), shape=(self.nslots, 3))
        return f.value, arr.copy()

    def previous(self):
        """(frame index, positions) of the push before the last one."""
        return self._get(self.lib.isf_export_previous, "isf_export_previous")

    def latest(self):
        return self._get(self.lib.isf_export_latest, "isf_export_latest")

    def close(self):
        if getattr(self, "h", None):
            self.lib.isf_export_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
