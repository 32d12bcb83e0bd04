"""Shared drivers for the parity tests: the reference's setup + simloop
(test/runtests.jl:16-66) run on the oracle and on libisf with identical inputs."""
from __future__ import annotations

import numpy as np

from oracle.oracle import JetConfig, OracleRun


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def wrap_diff(a, b):
    """|a - b| for angles, modulo 2 pi."""
    d = np.abs(np.asarray(a, np.float64) - np.asarray(b, np.float64))
    return np.minimum(d, 2 * np.pi - d)


def particle_report(gpu, ref, L):
    """Compare particle buffers: alive-status mismatches, error percentiles of live slots."""
    gpu = np.asarray(gpu, np.float64)
    ref = np.asarray(ref, np.float64)
    ga = np.all(gpu > 0, axis=1) & np.all(gpu < np.asarray(L), axis=1)
    ra = np.all(ref > 0, axis=1) & np.all(ref < np.asarray(L), axis=1)
    both = ga & ra
    err = np.max(np.abs(gpu[both] - ref[both]), axis=1) if both.any() else np.zeros(1)
    return {
        "alive_gpu": int(ga.sum()),
        "alive_ref": int(ra.sum()),
        "status_mismatch": int((ga != ra).sum()),
        "p50": float(np.percentile(err, 50)),
        "p99": float(np.percentile(err, 99)),
        "max": float(err.max()),
        "outliers_1e-3": int((err > 1e-3).sum()),
    }


class GpuRun:
    """The same driver on libisf (isf_amd)."""

    def __init__(self, cfg: JetConfig):
        from isf_amd import ISF, DeviceArray
        self.cfg = cfg
        f32 = np.float32
        self.isf = ISF(cfg.vol_size, cfg.dims, cfg.hbar, cfg.dt)
        n = self.isf.grid_res
        psi = np.zeros((n[2], n[1], n[0], 4), f32)
        psi[..., 0] = f32(cfg.psi_init[0])
        psi[..., 2] = f32(cfg.psi_init[1])
        self.psi = self.isf.psi_array(psi)
        self.isf.normalize_(self.psi)
        v = np.array(cfg.jet_velocity, f32)
        hb = f32(cfg.hbar)
        self.kvec = v / hb
        self.omega = f32(np.sum(v ** f32(2)) / (f32(2) * hb))
        self.cen = np.array(cfg.nozzle_cen, np.float64).astype(f32)
        self.isf.jet_configure(self.kvec, self.omega, self.cen, f32(cfg.nozzle_len), f32(cfg.nozzle_rad),
                               cfg.n_spawn, cfg.seed, cfg.phase_mode)
        self.isf.constrain_(self.psi, cfg.constraint_iters)
        self.parts = DeviceArray.from_host(np.full((cfg.n_slots, 3), -1.0, f32))
        if cfg.n_init > 0:
            self.isf.fill_nozzle_(self.parts, 0, cfg.n_init, cfg.seed)
        self.iter = 0

    def step(self, k=1):
        for _ in range(k):
            self.iter += 1
            self.isf.step_(self.psi, self.parts, self.iter)
        return self

    def host_psi(self):
        self.isf.synchronize()
        return self.psi.to_host()

    def host_parts(self):
        self.isf.synchronize()
        return self.parts.to_host()


def oracle_run(cfg, dtype=np.float64):
    return OracleRun(cfg, dtype)
