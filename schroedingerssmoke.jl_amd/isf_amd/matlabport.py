"""MatlabPort names over libisf (isf_mp_*): the CamelCase API of src/MatlabPort.jl with
its array layout -- psi1 and psi2 as separate complex arrays, the velocity as separate
vx, vy, vz arrays -- computed by the gfx950 kernels in Float32.

Arrays are DeviceArrays of shape (nz, ny, nx) (x fastest, the reference's column-major
order): complex64 for psi and q, float32 for velocities and f.  Every operator also
accepts numpy arrays of that shape and then returns numpy arrays (uploaded, computed on
the GPU, downloaded).  Like the reference the operators return new arrays; pass `out=`
to write into existing ones (in place is allowed).

    isf = ISF((4, 2, 2), (64, 32, 32), 0.1, 1 / 48)       # src/MatlabPort.jl:38-73
    psi1, psi2 = Normalize(psi1, psi2)                     # :155-160
    psi1, psi2 = PressureProject(isf, psi1, psi2)          # :147-153
    psi1, psi2 = SchroedingerFlow(isf, psi1, psi2)         # :87-93
    vx, vy, vz = VelocityOneForm(isf, psi1, psi2, isf.hbar)
    vx, vy, vz = StaggeredSharp(isf, vx, vy, vz)           # :99-104
    StaggeredAdvect(particle, isf, vx, vy, vz, isf.dt)     # :201-222

Normalize and GaugeTransform take no ISF argument in the reference; they run on the
stream of the ISF constructed last (or the one given as `isf=`).
"""
from __future__ import annotations

import numpy as np

from ._lib import check, load
from .api import ISF as _ISF
from .api import DeviceArray, c_float3
from .particles import Particles as _Store

__all__ = ["ISF", "land", "SchroedingerFlow", "Normalize", "VelocityOneForm", "StaggeredSharp", "Div",
           "PoissonSolve", "GaugeTransform", "PressureProject", "RestrictVelocity", "StaggeredVelocity",
           "StaggeredAdvect", "Particles", "Keep"]

_current = [None]


class ISF(_ISF):
    """ISF(vol_size, vol_res, hbar, dt), src/MatlabPort.jl:38-73 (+ BuildSchroedinger
    :75-83, whose mask lives in the handle's tables).  The grid fields of the reference
    (sizex.., resx.., dx.., px, py, pz) are host values; px, py, pz are (nz, ny, nx)
    Float64 arrays of the cell positions (i - 1) d (ndgrid, :63-66)."""

    def __init__(self, vol_size, vol_res, hbar, dt):
        super().__init__(vol_size, vol_res, hbar, dt)
        self.sizex, self.sizey, self.sizez = self.physical_size
        self.resx, self.resy, self.resz = self.grid_res
        self.dx, self.dy, self.dz = (float(v) for v in self.d)
        _current[0] = self
        self._grid = None

    def _positions(self):
        if self._grid is None:
            nx, ny, nz = self.grid_res
            z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
            self._grid = (x * (self.sizex / nx), y * (self.sizey / ny), z * (self.sizez / nz))
        return self._grid

    px = property(lambda self: self._positions()[0])
    py = property(lambda self: self._positions()[1])
    pz = property(lambda self: self._positions()[2])


def land(a, b):
    """land(a, b), src/MatlabPort.jl:5-8: elementwise logical and."""
    return np.logical_and(a, b)


def _handle(isf):
    h = isf if isf is not None else _current[0]
    if h is None or not getattr(h, "h", None):
        raise ValueError("no ISF: construct one first or pass isf=")
    return h


def _args(isf, arrays, dtypes):
    """Device views of the inputs (numpy arrays are uploaded); True when any was numpy."""
    host = False
    dev = []
    for a, dt in zip(arrays, dtypes):
        if isinstance(a, DeviceArray):
            if a.dtype != np.dtype(dt) or a.shape != isf.shape3:
                raise ValueError(f"expected a {np.dtype(dt)} array of shape {isf.shape3}, got {a.dtype} {a.shape}")
            dev.append(a)
        else:
            host = True
            arr = np.asarray(a)
            if arr.shape != isf.shape3:
                raise ValueError(f"expected shape {isf.shape3}, got {arr.shape}")
            dev.append(DeviceArray.from_host(arr.astype(dt)))
    return dev, host


def _outs(isf, out, dtypes):
    if out is None:
        return [DeviceArray(isf.shape3, dt) for dt in dtypes]
    out = list(out) if isinstance(out, (tuple, list)) else [out]
    for o, dt in zip(out, dtypes):
        if not isinstance(o, DeviceArray) or o.dtype != np.dtype(dt) or o.shape != isf.shape3:
            raise ValueError(f"out must be {np.dtype(dt)} DeviceArrays of shape {isf.shape3}")
    return out


def _ret(res, host):
    if host:
        res = [r.to_host() for r in res]
    return res[0] if len(res) == 1 else tuple(res)


C64, F32 = np.complex64, np.float32


def SchroedingerFlow(obj, psi1, psi2, out=None):
    """src/MatlabPort.jl:87-93."""
    (a, b), host = _args(obj, (psi1, psi2), (C64, C64))
    o = _outs(obj, out, (C64, C64))
    check(obj.lib.isf_mp_schroedinger_flow(obj.h, a.ptr, b.ptr, o[0].ptr, o[1].ptr), "SchroedingerFlow")
    return _ret(o, host)


def Normalize(psi1, psi2, out=None, isf=None):
    """src/MatlabPort.jl:155-160."""
    h = _handle(isf)
    (a, b), host = _args(h, (psi1, psi2), (C64, C64))
    o = _outs(h, out, (C64, C64))
    check(h.lib.isf_mp_normalize(h.h, a.ptr, b.ptr, o[0].ptr, o[1].ptr), "Normalize")
    return _ret(o, host)


def VelocityOneForm(obj, psi1, psi2, hbar=1.0, out=None):
    """src/MatlabPort.jl:106-126."""
    (a, b), host = _args(obj, (psi1, psi2), (C64, C64))
    o = _outs(obj, out, (F32, F32, F32))
    check(obj.lib.isf_mp_velocity_one_form(obj.h, a.ptr, b.ptr, float(np.float32(hbar)), o[0].ptr, o[1].ptr,
                                           o[2].ptr), "VelocityOneForm")
    return _ret(o, host)


def StaggeredSharp(obj, vx, vy, vz, out=None):
    """src/MatlabPort.jl:99-104."""
    v, host = _args(obj, (vx, vy, vz), (F32, F32, F32))
    o = _outs(obj, out, (F32, F32, F32))
    check(obj.lib.isf_mp_staggered_sharp(obj.h, v[0].ptr, v[1].ptr, v[2].ptr, o[0].ptr, o[1].ptr, o[2].ptr),
          "StaggeredSharp")
    return _ret(o, host)


def Div(obj, vx, vy, vz, out=None):
    """src/MatlabPort.jl:162-170."""
    v, host = _args(obj, (vx, vy, vz), (F32, F32, F32))
    o = _outs(obj, out, (F32,))
    check(obj.lib.isf_mp_div(obj.h, v[0].ptr, v[1].ptr, v[2].ptr, o[0].ptr), "Div")
    return _ret(o, host)


def PoissonSolve(obj, f, out=None):
    """src/MatlabPort.jl:128-138; complex result with imaginary part 0."""
    (fd,), host = _args(obj, (f,), (F32,))
    o = _outs(obj, out, (C64,))
    check(obj.lib.isf_mp_poisson_solve(obj.h, fd.ptr, o[0].ptr), "PoissonSolve")
    return _ret(o, host)


def GaugeTransform(psi1, psi2, q, out=None, isf=None):
    """src/MatlabPort.jl:140-145: psi * exp(i q), q complex."""
    h = _handle(isf)
    (a, b, qd), host = _args(h, (psi1, psi2, q), (C64, C64, C64))
    o = _outs(h, out, (C64, C64))
    check(h.lib.isf_mp_gauge_transform(h.h, a.ptr, b.ptr, qd.ptr, o[0].ptr, o[1].ptr), "GaugeTransform")
    return _ret(o, host)


def PressureProject(obj, psi1, psi2, out=None):
    """src/MatlabPort.jl:147-153."""
    (a, b), host = _args(obj, (psi1, psi2), (C64, C64))
    o = _outs(obj, out, (C64, C64))
    check(obj.lib.isf_mp_pressure_project(obj.h, a.ptr, b.ptr, o[0].ptr, o[1].ptr), "PressureProject")
    return _ret(o, host)


def RestrictVelocity(obj, psi1, psi2, kvec, omgterm, isjet=None, phase_mode=0):
    """The driver's jet constraint in place (examples/matlabport/simple.jl:104-110):
    psi_a[isJet] = |psi_a[isJet]| exp(i (k.x - omgterm)).  isjet: boolean mask of shape
    (nz, ny, nx) or None for the nozzle of jet_configure.  Device arrays only."""
    for p in (psi1, psi2):
        if not isinstance(p, DeviceArray) or p.dtype != np.dtype(C64) or p.shape != obj.shape3:
            raise ValueError(f"psi must be complex64 DeviceArrays of shape {obj.shape3}")
    mask = None
    if isjet is not None:
        m = np.asarray(isjet)
        if m.shape != obj.shape3:
            raise ValueError(f"isjet must have shape {obj.shape3}")
        mask = isjet if isinstance(isjet, DeviceArray) else DeviceArray.from_host(m.astype(np.uint8))
    check(obj.lib.isf_mp_restrict_velocity(obj.h, psi1.ptr, psi2.ptr, mask.ptr if mask else None, c_float3(*kvec),
                                           float(np.float32(omgterm)), int(phase_mode)), "jet constraint")
    if mask is not None and not isinstance(isjet, DeviceArray):
        obj.synchronize()  # the uploaded mask is freed on return
    return psi1, psi2


def StaggeredVelocity(px, py, pz, torus, vx, vy, vz):
    """src/MatlabPort.jl:237-273 at n points (1-D position arrays)."""
    pts = [p if isinstance(p, DeviceArray) else DeviceArray.from_host(np.asarray(p, np.float32)) for p in (px, py, pz)]
    host = not isinstance(px, DeviceArray)
    n = pts[0].shape[0]
    if any(p.shape != (n,) or p.dtype != np.dtype(F32) for p in pts):
        raise ValueError("positions must be three float32 vectors of equal length")
    v, vhost = _args(torus, (vx, vy, vz), (F32, F32, F32))
    o = [DeviceArray((n,), F32) for _ in range(3)]
    check(torus.lib.isf_mp_staggered_velocity(torus.h, pts[0].ptr, pts[1].ptr, pts[2].ptr, n, v[0].ptr, v[1].ptr,
                                              v[2].ptr, o[0].ptr, o[1].ptr, o[2].ptr), "StaggeredVelocity")
    return _ret(o, host)


class Particles(_Store):
    """Particles(x, y, z, length, active), src/MatlabPort.jl:178-184, on the device:
    capacity len(x); the driver's empty start (length 0, no active slots,
    examples/matlabport/simple.jl:74-78), or slots 0..length-1 filled from x, y, z and
    all active."""

    def __init__(self, x, y, z, length=0, active=(), isf=None):
        h = _handle(isf)
        super().__init__(h, len(x))
        length = int(length)
        if length:
            if list(np.asarray(active).ravel()) != list(range(length)):
                raise ValueError("active must list slots 0..length-1")
            self.append_(np.asarray(x, np.float32)[:length], np.asarray(y, np.float32)[:length],
                         np.asarray(z, np.float32)[:length])
        self.length = length

    def append_(self, x, y, z):
        super().append_(x, y, z)
        self.length = getattr(self, "length", 0) + int(np.asarray(x).shape[0] if not isinstance(x, DeviceArray)
                                                        else x.shape[0])


def StaggeredAdvect(p, torus, vx, vy, vz, dt):
    """src/MatlabPort.jl:201-222 over the active slots of p, in place."""
    v, _ = _args(torus, (vx, vy, vz), (F32, F32, F32))
    check(torus.lib.isf_mp_staggered_advect(torus.h, p.h, v[0].ptr, v[1].ptr, v[2].ptr, float(np.float32(dt))),
          "StaggeredAdvect")
    if not all(isinstance(a, DeviceArray) for a in (vx, vy, vz)):
        torus.synchronize()  # uploaded velocities are freed on return


def Keep(particle, ind):
    """Keep(particle, ind), src/MatlabPort.jl:226-230 (0-based ind)."""
    particle.keep_(ind)
    particle.length = len(np.asarray(ind))
