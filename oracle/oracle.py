"""ctypes front end of the CPU oracle (oracle/isf_oracle.c).

TEST INFRASTRUCTURE ONLY: tests/, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg use this module as the checker / the timed CPU baseline. The
product path (``schroedingerssmoke.jl_amd``) never imports it.

``Oracle(np.float32)`` restates the reference's Float32 ``Array`` path
(test/runtests.jl:14-16); ``Oracle(np.float64)`` is the FP64 authority. Array layouts
follow the reference (x fastest, column-major): psi ``(nz, ny, nx, 4)``, velocity
``(nz, ny, nx, 3)``, f ``(nz, ny, nx)``, q ``(nz, ny, nx, 2)``, particles ``(P, 3)``.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libisf_oracle.so")
_lib = None


def build() -> str:
    """Compile the oracle with its Makefile (gcc + OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _i3(v):
    return (C.c_int * 3)(*[int(x) for x in v])


class Oracle:
    """One precision instance of the oracle (float32 or float64)."""

    def __init__(self, dtype=np.float64):
        self.dt = np.dtype(dtype)
        assert self.dt in (np.float32, np.float64)
        self.suf = "_f32" if self.dt == np.float32 else "_f64"
        self.R = C.c_float if self.dt == np.float32 else C.c_double
        self.L = lib()

    def _fn(self, name):
        return getattr(self.L, name + self.suf)

    def _arr(self, v):
        return np.ascontiguousarray(np.asarray(v, dtype=self.dt))

    def _r3(self, v):
        return self._arr(np.asarray(v, dtype=np.float64).astype(self.dt))

    # --- tables: src/parallel.jl:47-63
    def tables(self, n, L, hbar, dt):
        nx, ny, nz = n
        fac = np.zeros((nz, ny, nx, 2), self.dt)
        mask = np.zeros((nz, ny, nx, 2), self.dt)
        self._fn("isfo_tables")(_i3(n), _i3(L), C.c_double(hbar), C.c_double(dt), _p(fac), _p(mask))
        return fac, mask

    def fft3(self, z, inverse=False):
        z = self._arr(z).copy()
        n = (z.shape[2], z.shape[1], z.shape[0])
        self._fn("isfo_fft3")(_p(z), _i3(n), C.c_int(int(inverse)))
        return z

    def normalize(self, psi):
        psi = self._arr(psi).copy()
        self._fn("isfo_normalize")(_p(psi), C.c_long(psi.size // 4))
        return psi

    def schroedinger_flow(self, psi, mask):
        psi = self._arr(psi).copy()
        n = (psi.shape[2], psi.shape[1], psi.shape[0])
        self._fn("isfo_schroedinger_flow")(_p(psi), _p(self._arr(mask)), _i3(n))
        return psi

    def velocity_one_form(self, psi, d, hbar=1.0, sharp=False):
        psi = self._arr(psi)
        n = (psi.shape[2], psi.shape[1], psi.shape[0])
        vel = np.zeros(psi.shape[:3] + (3,), self.dt)
        self._fn("isfo_velocity_one_form")(_p(psi), _i3(n), _p(self._r3(d)), self.R(hbar),
                                          C.c_int(int(sharp)), _p(vel))
        return vel

    def div(self, vel, d):
        vel = self._arr(vel)
        n = (vel.shape[2], vel.shape[1], vel.shape[0])
        f = np.zeros(vel.shape[:3], self.dt)
        self._fn("isfo_div")(_p(vel), _i3(n), _p(self._r3(d)), _p(f))
        return f

    def poisson_solve(self, f, fac):
        f = self._arr(f)
        n = (f.shape[2], f.shape[1], f.shape[0])
        q = np.zeros(f.shape + (2,), self.dt)
        self._fn("isfo_poisson_solve")(_p(f), _p(self._arr(fac)), _i3(n), _p(q))
        return q

    def gauge_transform(self, psi, q):
        psi = self._arr(psi).copy()
        self._fn("isfo_gauge_transform")(_p(psi), _p(self._arr(q)), C.c_long(psi.size // 4))
        return psi

    def pressure_project(self, psi, fac, d):
        psi = self._arr(psi).copy()
        n = (psi.shape[2], psi.shape[1], psi.shape[0])
        N = psi.size // 4
        vel = np.zeros(3 * N, self.dt)
        f = np.zeros(N, self.dt)
        q = np.zeros(2 * N, self.dt)
        self._fn("isfo_pressure_project")(_p(psi), _p(self._arr(fac)), _i3(n), _p(self._r3(d)),
                                         _p(vel), _p(f), _p(q))
        return psi

    def isjet(self, n, d, cen, length, rad):
        out = np.zeros((n[2], n[1], n[0]), np.uint8)
        self._fn("isfo_isjet")(_i3(n), _p(self._r3(d)), _p(self._r3(cen)), self.R(length),
                               self.R(rad), _p(out))
        return out

    def restrict_velocity(self, psi, isjet, d, kvec, omgterm, phase_mode=0):
        psi = self._arr(psi).copy()
        n = (psi.shape[2], psi.shape[1], psi.shape[0])
        self._fn("isfo_restrict_velocity")(_p(psi), _p(np.ascontiguousarray(isjet, np.uint8)), _i3(n),
                                           _p(self._r3(d)), _p(self._r3(kvec)), self.R(omgterm),
                                           C.c_int(phase_mode))
        return psi

    def add_particles(self, parts, first, n, cen, rad, seed, step):
        parts = self._arr(parts).copy()
        self._fn("isfo_add_particles")(_p(parts), C.c_long(parts.shape[0]), C.c_long(first), C.c_long(n),
                                       _p(self._r3(cen)), self.R(rad), C.c_ulonglong(seed),
                                       C.c_uint(step))
        return parts

    def fill_cylinder(self, parts, first, n, cen, length, rad, seed):
        parts = self._arr(parts).copy()
        self._fn("isfo_fill_cylinder")(_p(parts), C.c_longlong(parts.shape[0]), C.c_longlong(first), C.c_longlong(n),
                                       _p(self._r3(cen)), self.R(length), self.R(rad), C.c_ulonglong(seed))
        return parts

    def staggered_advect(self, parts, vel, dt, d, L):
        parts = self._arr(parts).copy()
        vel = self._arr(vel)
        n = (vel.shape[2], vel.shape[1], vel.shape[0])
        self._fn("isfo_staggered_advect")(_p(parts), C.c_long(parts.shape[0]), _p(vel), self.R(dt),
                                          _p(self._r3(d)), _p(self._r3(L)), _i3(n))
        return parts

    def norm2(self, psi):
        f = self._fn("isfo_norm2")
        f.restype = C.c_double
        psi = self._arr(psi)
        return f(_p(psi), C.c_long(psi.size // 4))


def philox4x32_10(ctr, key):
    out = (C.c_uint * 4)()
    lib().isfo_philox4x32_10((C.c_uint * 4)(*ctr), (C.c_uint * 2)(*key), out)
    return list(out)


def philox_uniform4(a, b, c, seed):
    u = np.zeros(4, np.float32)
    lib().isfo_philox_uniform4(C.c_uint(a), C.c_uint(b), C.c_uint(c), C.c_ulonglong(seed), _p(u))
    return u


@dataclass
class JetConfig:
    """Parameters of a jet run, named after the reference driver (test/runtests.jl:3-13)."""
    vol_size: tuple = (4, 2, 2)
    dims: tuple = (64, 32, 32)
    hbar: float = 0.1
    dt: float = 1.0 / 48.0
    jet_velocity: tuple = (1.0, 0.0, 0.0)
    nozzle_cen: tuple = (2 - 1.7, 1 - 0.034, 1 + 0.066)
    nozzle_len: float = 0.5
    nozzle_rad: float = 0.5
    n_spawn: int = 50
    n_slots: int = 10_000
    n_init: int = 0            # particles drawn uniformly in the nozzle cylinder at setup
    constraint_iters: int = 10
    seed: int = 20261203
    phase_mode: int = 0        # 0 = k.x (MatlabPort), 1 = reference tuple_dot sum
    psi_init: tuple = (1.0, 0.01)


class OracleRun:
    """The reference driver's setup + simloop (test/runtests.jl:16-66) on the oracle."""

    def __init__(self, cfg: JetConfig, dtype=np.float64):
        self.cfg = cfg
        self.o = Oracle(dtype)
        dt_ = self.o.dt
        f32 = np.float32
        n = tuple(int(v) for v in cfg.dims)
        self.n = n
        self.N = n[0] * n[1] * n[2]
        self.Lbox = tuple(int(v) for v in cfg.vol_size)
        self.hbar = dt_.type(f32(cfg.hbar))
        self.dt = dt_.type(f32(cfg.dt))
        self.d = np.array([f32(self.Lbox[a] / n[a]) for a in range(3)], dt_)
        self.fac, self.mask = self.o.tables(n, self.Lbox, float(f32(cfg.hbar)), float(f32(cfg.dt)))
        # psi init + normalize, test/runtests.jl:19-21
        psi = np.zeros((n[2], n[1], n[0], 4), dt_)
        psi[..., 0] = f32(cfg.psi_init[0])
        psi[..., 2] = f32(cfg.psi_init[1])
        self.psi = self.o.normalize(psi)
        # kvec, omega, isjet: test/runtests.jl:23-25 (Float32 arithmetic)
        v = np.array(cfg.jet_velocity, f32)
        hb = f32(cfg.hbar)
        self.kvec = (v / hb).astype(dt_)
        self.omega = dt_.type(f32(np.sum(v ** f32(2)) / (f32(2) * hb)))
        self.cen = np.array(cfg.nozzle_cen, np.float64).astype(f32).astype(dt_)
        self.isjet = self.o.isjet(n, self.d, self.cen, f32(cfg.nozzle_len), f32(cfg.nozzle_rad))
        # constraint iterations, test/runtests.jl:27-30
        for _ in range(cfg.constraint_iters):
            self.psi = self.o.restrict_velocity(self.psi, self.isjet, self.d, self.kvec, 0.0, cfg.phase_mode)
            self.psi = self.o.pressure_project(self.psi, self.fac, self.d)
        # particle buffer of dead markers, test/runtests.jl:33
        self.parts = np.full((cfg.n_slots, 3), -1.0, dt_)
        if cfg.n_init > 0:
            self.parts = self.o.fill_cylinder(self.parts, 0, cfg.n_init, self.cen, f32(cfg.nozzle_len),
                                              f32(cfg.nozzle_rad), cfg.seed)
        self.iter = 0
        N = self.N
        self._vel = np.zeros(3 * N, dt_)
        self._f = np.zeros(N, dt_)
        self._q = np.zeros(2 * N, dt_)

    def step(self, k=1):
        o = self.o
        cfg = self.cfg
        for _ in range(k):
            self.iter += 1
            o._fn("isfo_step")(
                _p(self.psi), _p(self.parts), C.c_long(self.parts.shape[0]), _p(self.fac), _p(self.mask),
                _p(self.isjet), _i3(self.n), _p(self.d), _p(o._r3(self.Lbox)), o.R(self.hbar),
                o.R(self.dt), _p(self.kvec), o.R(self.omega), _p(self.cen), o.R(np.float32(cfg.nozzle_rad)),
                C.c_long(cfg.n_spawn), C.c_int(self.iter), C.c_ulonglong(cfg.seed), C.c_int(cfg.phase_mode),
                _p(self._vel), _p(self._f), _p(self._q))
        return self

    @property
    def velocity(self):
        return self._vel.reshape(self.n[2], self.n[1], self.n[0], 3)


# ---------------------------------------------------------------- library FFT for the CPU baseline
_FFT_HOOK_T = C.CFUNCTYPE(None, C.c_void_p, C.POINTER(C.c_int), C.c_int, C.c_int)


def _torch_fft3(ptr, n, inverse, dbl):
    """The 3-D transform of isfo_fft3 by MKL through torch.fft on the CPU: in place on the
    interleaved array at ptr (x fastest), forward unnormalised, inverse scaled by 1/N
    (plan_fft! / plan_ifft!, src/parallel.jl:79)."""
    import torch
    nx, ny, nz = n[0], n[1], n[2]
    ct = C.c_double if dbl else C.c_float
    a = np.ctypeslib.as_array(C.cast(ptr, C.POINTER(ct)), shape=(nz, ny, nx, 2))
    z = torch.view_as_complex(torch.from_numpy(a))
    z.copy_(torch.fft.ifftn(z, dim=(0, 1, 2)) if inverse else torch.fft.fftn(z, dim=(0, 1, 2)))


_fft_hook_ref = _FFT_HOOK_T(_torch_fft3)


class cpu_engine:
    """Context: the oracle's FFTs through MKL (torch.fft, fft="mkl") or the radix-2
    restatement (fft="radix2"), with `threads` OpenMP threads for the per-statement loops
    and the same number of torch intra-op threads."""

    def __init__(self, fft="mkl", threads=None):
        self.fft = fft
        self.threads = int(threads) if threads else os.cpu_count() or 1

    def __enter__(self):
        import torch
        L = lib()
        self._prev = (L.isfo_get_threads(), torch.get_num_threads())
        L.isfo_set_threads(C.c_int(self.threads))
        torch.set_num_threads(self.threads)
        L.isfo_set_fft_hook(_fft_hook_ref if self.fft == "mkl" else _FFT_HOOK_T())
        return self

    def __exit__(self, *exc):
        import torch
        L = lib()
        L.isfo_set_fft_hook(_FFT_HOOK_T())
        L.isfo_set_threads(C.c_int(self._prev[0]))
        torch.set_num_threads(self._prev[1])
        return False
