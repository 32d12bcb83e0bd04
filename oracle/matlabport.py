"""numpy restatement of the reference's FP64 "Baseline version" (src/MatlabPort.jl).

TEST INFRASTRUCTURE ONLY (see oracle/oracle.py). It is a second, independent CPU
restatement of the same mathematics as the generic path (src/parallel.jl) and is used
to cross-check the C oracle: it uses numpy's pocketfft instead of the oracle's own FFT,
whole-array ops instead of per-cell loops, fftshift around the mask (src/MatlabPort.jl:88,91)
instead of the pre-shifted mask (src/parallel.jl:62), separate psi1/psi2 arrays, and the
MATLAB ndgrid layout: arrays here are indexed [x, y, z] (shape (nx, ny, nz)).
"""
from __future__ import annotations

import numpy as np


class MatlabISF:
    """ISF, src/MatlabPort.jl:38-73 (+ BuildSchroedinger :75-83)."""

    def __init__(self, vol_size, vol_res, hbar, dt):
        self.sizex, self.sizey, self.sizez = [float(v) for v in vol_size]
        self.resx, self.resy, self.resz = [int(round(v)) for v in vol_res]
        self.dx = self.sizex / self.resx
        self.dy = self.sizey / self.resy
        self.dz = self.sizez / self.resz
        ix = np.arange(1, self.resx + 1)
        iy = np.arange(1, self.resy + 1)
        iz = np.arange(1, self.resz + 1)
        self.iix, self.iiy, self.iiz = np.meshgrid(ix, iy, iz, indexing="ij")  # ndgrid, :63
        self.px = (self.iix - 1) * self.dx
        self.py = (self.iiy - 1) * self.dy
        self.pz = (self.iiz - 1) * self.dz
        self.hbar = float(hbar)
        self.dt = float(dt)
        self._build_schroedinger()

    def _build_schroedinger(self):  # :75-83
        nx, ny, nz = self.resx, self.resy, self.resz
        fac = -4 * np.pi ** 2 * self.hbar
        kx = (self.iix - 1 - nx / 2) / self.sizex
        ky = (self.iiy - 1 - ny / 2) / self.sizey
        kz = (self.iiz - 1 - nz / 2) / self.sizez
        lam = fac * (kx ** 2 + ky ** 2 + kz ** 2)
        self.mask = np.exp(1j * lam * self.dt / 2.0)


def schroedinger_flow(obj, psi1, psi2):  # :87-93
    psi1 = np.fft.fftshift(np.fft.fftn(psi1))
    psi2 = np.fft.fftshift(np.fft.fftn(psi2))
    psi1 = psi1 * obj.mask
    psi2 = psi2 * obj.mask
    psi1 = np.fft.ifftn(np.fft.fftshift(psi1))
    psi2 = np.fft.ifftn(np.fft.fftshift(psi2))
    return psi1, psi2


def staggered_sharp(obj, vx, vy, vz):  # :99-104
    return vx / obj.dx, vy / obj.dy, vz / obj.dz


def velocity_one_form(obj, psi1, psi2, hbar=1.0):  # :106-126
    def shifted(a, axis):
        return np.roll(a, -1, axis=axis)  # view(psi, mod.(i, res)+1, ...)

    vx = np.angle(np.conj(psi1) * shifted(psi1, 0) + np.conj(psi2) * shifted(psi2, 0))
    vy = np.angle(np.conj(psi1) * shifted(psi1, 1) + np.conj(psi2) * shifted(psi2, 1))
    vz = np.angle(np.conj(psi1) * shifted(psi1, 2) + np.conj(psi2) * shifted(psi2, 2))
    return vx * hbar, vy * hbar, vz * hbar


def poisson_solve(obj, f):  # :128-138
    f = np.fft.fftn(f)
    sx = np.sin(np.pi * (obj.iix - 1) / obj.resx) / obj.dx
    sy = np.sin(np.pi * (obj.iiy - 1) / obj.resy) / obj.dy
    sz = np.sin(np.pi * (obj.iiz - 1) / obj.resz) / obj.dz
    denom = sx ** 2 + sy ** 2 + sz ** 2
    with np.errstate(divide="ignore"):
        fac = -0.25 / denom
    fac[0, 0, 0] = 0.0
    return np.fft.ifftn(f * fac)


def gauge_transform(psi1, psi2, q):  # :140-145
    eiq = np.exp(1j * q)
    return psi1 * eiq, psi2 * eiq


def div(obj, vx, vy, vz):  # :162-170
    f = (vx - np.roll(vx, 1, axis=0)) / obj.dx ** 2
    f = f + (vy - np.roll(vy, 1, axis=1)) / obj.dy ** 2
    f = f + (vz - np.roll(vz, 1, axis=2)) / obj.dz ** 2
    return f


def pressure_project(obj, psi1, psi2):  # :147-153
    vx, vy, vz = velocity_one_form(obj, psi1, psi2)
    d = div(obj, vx, vy, vz)
    q = -poisson_solve(obj, d)
    return gauge_transform(psi1, psi2, q)


def normalize(psi1, psi2):  # :155-160
    n = np.sqrt(np.abs(psi1) ** 2 + np.abs(psi2) ** 2)
    return psi1 / n, psi2 / n


def staggered_velocity(px, py, pz, torus, vx, vy, vz):  # :237-273
    px = np.mod(px, torus.sizex)
    py = np.mod(py, torus.sizey)
    pz = np.mod(pz, torus.sizez)
    ix = np.floor(px / torus.dx).astype(np.int64)  # 0-based here
    iy = np.floor(py / torus.dy).astype(np.int64)
    iz = np.floor(pz / torus.dz).astype(np.int64)
    ix %= torus.resx
    iy %= torus.resy
    iz %= torus.resz
    ixp = (ix + 1) % torus.resx
    iyp = (iy + 1) % torus.resy
    izp = (iz + 1) % torus.resz
    wx = px - ix * torus.dx
    wy = py - iy * torus.dy
    wz = pz - iz * torus.dz
    ux = ((1 - wz) * ((1 - wy) * vx[ix, iy, iz] + wy * vx[ix, iyp, iz])
          + wz * ((1 - wy) * vx[ix, iy, izp] + wy * vx[ix, iyp, izp]))
    uy = ((1 - wz) * ((1 - wx) * vy[ix, iy, iz] + wx * vy[ixp, iy, iz])
          + wz * ((1 - wx) * vy[ix, iy, izp] + wx * vy[ixp, iy, izp]))
    uz = ((1 - wy) * ((1 - wx) * vz[ix, iy, iz] + wx * vz[ixp, iy, iz])
          + wy * ((1 - wx) * vz[ix, iyp, iz] + wx * vz[ixp, iyp, iz]))
    return ux, uy, uz


def staggered_advect(px, py, pz, torus, vx, vy, vz, dt):  # :201-222
    k1x, k1y, k1z = staggered_velocity(px, py, pz, torus, vx, vy, vz)
    k2x, k2y, k2z = staggered_velocity(px + k1x * dt / 2, py + k1y * dt / 2, pz + k1z * dt / 2, torus, vx, vy, vz)
    k3x, k3y, k3z = staggered_velocity(px + k2x * dt / 2, py + k2y * dt / 2, pz + k2z * dt / 2, torus, vx, vy, vz)
    k4x, k4y, k4z = staggered_velocity(px + k3x * dt, py + k3y * dt, pz + k3z * dt, torus, vx, vy, vz)
    return (px + dt / 6 * (k1x + 2 * k2x + 2 * k3x + k4x),
            py + dt / 6 * (k1y + 2 * k2y + 2 * k3y + k4y),
            pz + dt / 6 * (k1z + 2 * k2z + 2 * k3z + k4z))


def to_xyz(a):
    """Oracle layout (nz, ny, nx, ...) -> MatlabPort layout (nx, ny, nz, ...)."""
    return np.swapaxes(a, 0, 2)


class Torus:
    """The grid fields StaggeredVelocity reads (sizex.., dx.., resx..)."""

    def __init__(self, n, L, d):
        self.resx, self.resy, self.resz = n
        self.sizex, self.sizey, self.sizez = L
        self.dx, self.dy, self.dz = d


class Particles:
    """Particles, src/MatlabPort.jl:178-196: Float32 positions, length, active list."""

    def __init__(self, capacity):
        self.x = np.zeros(capacity, np.float32)  # zeros(Float32, max_particles), simple.jl:74-78
        self.y = np.zeros(capacity, np.float32)
        self.z = np.zeros(capacity, np.float32)
        self.length = 0
        self.active = []

    def append(self, x, y, z):  # :186-195
        n = len(x)
        assert n + self.length < len(self.x)
        r = np.arange(self.length, self.length + n)
        self.x[r], self.y[r], self.z[r] = x, y, z
        self.length += n
        self.active.extend(int(i) for i in r)

    def spawn_ring(self, n, cen, rad, seed, step):
        """The driver's births (examples/matlabport/simple.jl:113-117) with u from
        Philox4x32-10 counter (k, step, 2, 0) (the reference draws rand(Float32))."""
        from .oracle import philox4x32_10
        u = np.array([np.float32((philox4x32_10((k, step, 2, 0), (seed & 0xFFFFFFFF, seed >> 32))[0] >> 8))
                      * np.float32(1.0 / 16777216.0) for k in range(n)], np.float32)
        rt = u.astype(np.float64) * 2 * np.pi
        self.append(np.full(n, cen[0]), cen[1] + 0.9 * rad * np.cos(rt), cen[2] + 0.9 * rad * np.sin(rt))

    def remove_dead(self, vol_size):  # removethedead!, simple.jl:63-70 (isinside: strict, :9-14)
        L = vol_size
        self.active = [i for i in self.active
                       if 0 < self.x[i] < L[0] and 0 < self.y[i] < L[1] and 0 < self.z[i] < L[2]]

    def advect(self, torus, vx, vy, vz, dt):  # StaggeredAdvect, src/MatlabPort.jl:201-222
        a = np.asarray(self.active, np.int64)
        if a.size == 0:
            return
        px, py, pz = (self.x[a].astype(np.float64), self.y[a].astype(np.float64), self.z[a].astype(np.float64))
        nx_, ny_, nz_ = staggered_advect(px, py, pz, torus, vx, vy, vz, dt)
        self.x[a], self.y[a], self.z[a] = nx_, ny_, nz_

    def keep(self, ind):  # Keep, :226-230
        ind = np.asarray(ind, np.int64)
        self.x, self.y, self.z = self.x[ind], self.y[ind], self.z[ind]
