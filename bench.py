#!/usr/bin/env python3
"""ISF step throughput on one MI355X (the 256x128x128 jet) + roofline + CPU baseline.

One step = the reference simloop body (test/runtests.jl:45-64): Schroedinger flow,
normalize, pressure projection, particle spawn, jet constraint, second projection,
hbar-scaled velocity one-form with 1/d, RK4 staggered advection of every particle slot.

Workload: grid 256x128x128, box 8x4x4, hbar 0.1, dt 1/48, psi = (1, 0.01) normalized,
jet velocity (1,0,0), nozzle radius 0.5 / length 0.5 at (0.3, 1.966, 2.066),
10 constraint iterations, 1,000,000 particle slots filled uniformly in the nozzle
cylinder (Philox4x32-10, seed 20261203), 1000 ring-spawned particles per step.
Inputs are resident in device memory before the timed region starts.
"""
import argparse
import csv
import glob
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "schroedingerssmoke.jl_amd"))
sys.path.insert(0, ROOT)

METRIC = "ISF steps/s at 256×128×128 (fwd, fwd+adjoint, batched) + % HBM peak"
HBM_PEAK_GBS = 8000.0

WORKLOAD = dict(vol_size=(8, 4, 4), dims=(256, 128, 128), hbar=0.1, dt=1.0 / 48.0, jet_velocity=(1.0, 0.0, 0.0),
                nozzle_cen=(0.3, 2 - 0.034, 2 + 0.066), nozzle_len=0.5, nozzle_rad=0.5, n_spawn=1000,
                n_slots=1_000_000, n_init=1_000_000, constraint_iters=10, seed=20261203, phase_mode=0)


def stage_bytes(nx, ny, nz, nslots, n_spawn):
    """Algorithmic (compulsory) HBM bytes of one launch of every stage of a step: each
    array the stage must read or write, counted once (tables and index math are free)."""
    N = nx * ny * nz
    half = 8 * (nx // 2 + 1) * ny * nz  # Poisson half spectrum, complex64
    b = {
        "flow_x_fwd": 16 * N + 16 * N,
        "flow_plane": 16 * N + 16 * N,
        "flow_x_inv_normalize": 16 * N + 16 * N,
        "spawn": 12 * n_spawn,
        "velocity_sharp": 16 * N + 12 * N,
        "particle_order": 0,
        "advect": 24 * nslots + 12 * N,
    }
    for p in ("p1", "p2"):
        b[f"{p}_div_r2c_x"] = 16 * N + half  # psi in, half spectrum out (f stays on chip)
        b[f"{p}_plane"] = 2 * half
    b["p1_c2r_x_gauge_restrict"] = half + 32 * N
    b["p2_c2r_x_gauge"] = half + 32 * N
    return b


def b_alg_step(N, nslots):
    """Compulsory bytes of one step: 248 B/cell + 24 B/particle slot."""
    return 248 * N + 24 * nslots


def setup_gpu(w):
    from isf_amd import ISF, DeviceArray
    f32 = np.float32
    isf = ISF(w["vol_size"], w["dims"], w["hbar"], w["dt"])
    nx, ny, nz = w["dims"]
    psi_h = np.zeros((nz, ny, nx, 4), f32)
    psi_h[..., 0] = 1.0
    psi_h[..., 2] = f32(0.01)
    psi = isf.psi_array(psi_h)
    isf.normalize_(psi)
    v = np.array(w["jet_velocity"], f32)
    hb = f32(w["hbar"])
    kvec = v / hb
    omega = f32(np.sum(v ** f32(2)) / (f32(2) * hb))
    cen = np.array(w["nozzle_cen"], np.float64).astype(f32)
    isf.jet_configure(kvec, omega, cen, f32(w["nozzle_len"]), f32(w["nozzle_rad"]), w["n_spawn"], w["seed"],
                      w["phase_mode"])
    isf.constrain_(psi, w["constraint_iters"])
    parts = DeviceArray.from_host(np.full((w["n_slots"], 3), -1.0, f32))
    isf.fill_nozzle_(parts, 0, w["n_init"], w["seed"])
    isf.synchronize()
    return isf, psi, parts


def cpu_baseline(w, budget_s=8.0):
    """The reference's Array-backend step restated on this host's CPU: one OpenMP loop per
    reference broadcast statement (the FP32 oracle, oracle/isf_oracle_impl.h, in the
    reference's operator order), with the eight 3-D FFTs of a step done either by MKL
    through torch.fft (FFTW-class; the reference plans FFTW, src/parallel.jl:79) or by the
    oracle's own radix-2 transform.  Both engines are timed with all host threads and with
    one thread (Julia 0.6 broadcasts are single-threaded); the reported value is the faster
    all-threads engine.  Each leg steps until its budget has elapsed (at least 2 steps)
    after one untimed step."""
    from oracle.oracle import JetConfig, OracleRun, cpu_engine
    cfg = JetConfig(**{k: w[k] for k in JetConfig.__dataclass_fields__ if k in w})
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    run = OracleRun(cfg, np.float32)

    def leg(fft, nthreads, budget):
        with cpu_engine(fft, nthreads):
            run.step(1)
            n, t0 = 0, time.perf_counter()
            while True:
                run.step(1)
                n += 1
                el = time.perf_counter() - t0
                if (el >= budget and n >= 2) or n >= 200:
                    break
        return {"value": round(n / el, 4), "cores": nthreads, "fft": fft, "sample": f"{n} steps in {el:.1f} s"}

    legs = [leg("mkl", threads, budget_s), leg("radix2", threads, budget_s),
            leg("mkl", 1, budget_s / 2), leg("radix2", 1, budget_s / 2)]
    best = max(legs[:2], key=lambda r: r["value"])
    one = max(legs[2:], key=lambda r: r["value"])
    print("[bench] CPU step: " + "; ".join(f"{r['fft']} {r['cores']} threads {r['value']:.3f} steps/s" for r in legs),
          file=sys.stderr, flush=True)
    name = {"mkl": "MKL via torch.fft", "radix2": "the oracle's radix-2 transform"}
    return {"value": best["value"], "unit": "steps/s", "cores": threads, "kind": "port",
            "one_thread": one, "legs": legs,
            "sample": f"FP32 restatement of the reference's Array step (oracle/isf_oracle_impl.h: one OpenMP loop "
                      f"per reference broadcast statement) with 3-D FFTs by {name[best['fft']]} (the faster engine "
                      f"on this host), same 256x128x128 / 1M-particle workload after 10 constraint iterations + 1 "
                      f"untimed step: {best['sample']} on {threads} threads"}


def bench_ensemble(steps, warmup, count=48, seed=20261203):
    """Batched uneven ensemble: `count` jets on grids drawn from {32^3, 64x32x32, 64^3,
    128x64x64} (box = grid/16), hbar in [0.05, 0.2], |v| in [0.5, 1.5] within 20 degrees of
    +x, 10k particle slots each filled in the nozzle, 100 spawned per step.  The same
    ensemble is stepped (a) batched through the driver's simloop, `steps` steps in one
    isf_batch_simloop call (the headline), (b) batched with one isf_batch_step (one graph
    replay) per step, and (c) sequentially: one isf_step per member, all with the
    autotuner's schedules cached for this draw on an MI355X (isf_amd/schedules); wall-clock
    time over `steps` ensemble steps after `warmup`, synchronised on both sides."""
    from isf_amd.ensemble import ENSEMBLE48_SCHEDULES, build_batch, draw_jets
    ms = draw_jets(count=count, seed=seed)
    cells = sum(int(np.prod(m["res"])) for m in ms)
    slots = sum(m["n_slots"] for m in ms)
    sched = os.path.join(tempfile.mkdtemp(prefix="isf_sched_", dir="/tmp"), "sched.txt")
    shutil.copy(ENSEMBLE48_SCHEDULES, sched)
    t0 = time.perf_counter()
    b = build_batch(ms)
    b.tune(sched)
    b.synchronize()
    setup_s = time.perf_counter() - t0
    shapes = sorted({m["res"] for m in ms})
    tiles = {"x".join(map(str, r)): b.tile(r) for r in shapes}
    plane_elems = {"x".join(map(str, r)): b.plane_elems(r) for r in shapes}
    counts = {"x".join(map(str, r)): sum(1 for m in ms if m["res"] == r) for r in shapes}

    def run(fn, it0):
        it = it0
        for _ in range(warmup):
            it += 1
            fn(it)
        b.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            it += 1
            fn(it)
        b.synchronize()
        return (time.perf_counter() - t) / steps, it

    t_seq, it = run(b.step_sequential_, 0)
    t_call, it = run(b.step_, it)
    # the driver's simloop over the ensemble: the same steps in one isf_batch_simloop call
    b.simloop_(it + 1, warmup)
    it += warmup
    b.synchronize()
    t0 = time.perf_counter()
    b.simloop_(it + 1, steps)
    b.synchronize()
    t_bat = (time.perf_counter() - t0) / steps
    it += steps
    norms = [b.members[i].step_norm2() / np.prod(ms[i]["res"]) for i in range(count)]
    b.close()
    balg = 248 * cells + 24 * slots
    print(f"[bench] ensemble of {count}: batched simloop {t_bat * 1e3:.3f} ms, one isf_batch_step per step "
          f"{t_call * 1e3:.3f} ms, sequential {t_seq * 1e3:.3f} ms per ensemble step (setup {setup_s:.1f} s)",
          file=sys.stderr, flush=True)
    return {
        "workload": f"{count} independent jets, grids {counts}, 10k particle slots each, 100 spawned per step",
        "ensemble_steps_s": round(1.0 / t_bat, 2),
        "ensemble_steps_s_sequential": round(1.0 / t_seq, 2),
        "speedup_vs_sequential": round(t_seq / t_bat, 2),
        "member_steps_s": round(count / t_bat, 1),
        "ms_per_step": round(t_bat * 1e3, 4),
        "ms_per_step_sequential": round(t_seq * 1e3, 4),
        "isf_batch_step_per_call": {"ensemble_steps_s": round(1.0 / t_call, 2), "ms_per_step": round(t_call * 1e3, 4),
                                    "speedup_vs_sequential": round(t_seq / t_call, 2)},
        "cells": cells,
        "B_alg_per_step": balg,
        "achieved_GB_s": round(balg / t_bat / 1e9, 1),
        "frac": round(balg / t_bat / 1e9 / HBM_PEAK_GBS, 4),
        "x_tiles": tiles,
        "plane_elements_per_thread": plane_elems,
        "norm2_over_N_range": [round(float(min(norms)), 9), round(float(max(norms)), 9)],
        "steps": steps,
        "warmup": warmup,
        "timing": "wall clock around the steps, device synchronised on both sides",
    }


def b_alg_fwd_adjoint(N, P, T, K):
    """Compulsory bytes per step of a forward + backward pass of the rollout (the adjoint
    DAG of one simloop step; each array read or written once per fused kernel, the
    Poisson half spectrum counted as 4 B/cell each way; the set-up constraint iterations
    are not counted, their time is):
      forward step   248 N + 24 P (the step, whose velocity field is written to the tape;
                     its advection writes the particle trajectory the backward reads)
                     + 24 N tape (raw flow output, q1, q2)
      recompute      the same without the particles (248 + 24) N, for the T - K steps
                     outside the last segment
      reverse step   psi chain 324 N: velocity adjoint 64 (psibar, psi, vbar -> psibar,
                     qbar), two Poisson solves 2 x 24, two one-form/div adjoints 56 + 60 (the
                     second with the restrict adjoint and its gauge cotangent fused), flow
                     adjoint 3 x 32; particle chain 24 N + 36 P:
                     vbar written 12, taped velocity read 12, positions 12 P, pbar read and
                     written 24 P."""
    fwd = 272 * N + 24 * P
    rec = 272 * N
    rev = 348 * N + 36 * P
    return fwd + rec * (T - K) / T + rev


def bench_fwd_adjoint(reps, dims=(128, 64, 64), box=(8, 4, 4), nslots=100_000, splat_res=(32, 16, 16)):
    """Differentiable rollout: a jet on `dims` (box, hbar 0.1, dt 1/48), 32 steps with
    checkpoints every 4, `nslots` particle slots filled in the nozzle and 1000 spawned per
    step, 16^3 seeded initial phase field, loss = MSE of the Gaussian-splatted particle
    density (splat_res grid, sigma 0.25, cutoff 4 sigma) against a sum of 3 seeded
    Gaussians.  Timed: `reps` forward + backward passes (set-up constraint iterations
    included), wall clock, synchronised; steps/s counts the 32 simloop steps of each pass."""
    from isf_amd import ISF
    from isf_amd.rollout import Rollout, gaussian_target
    steps, K = 32, 4
    hb = np.float32(0.1)
    v = np.array([1.0, 0.0, 0.0], np.float32)
    isf = ISF(box, dims, hb, 1.0 / 48.0)
    cen = np.array([0.3, box[1] / 2 - 0.034, box[2] / 2 + 0.066], np.float32)
    isf.jet_configure(v / hb, np.float32(np.sum(v ** np.float32(2)) / (np.float32(2) * hb)), cen, 0.5, 0.5, 1000,
                      20261203, 0)
    r = Rollout(isf, steps, checkpoint_every=K, constraint_iters=10, phase_res=(16, 16, 16), splat_res=splat_res,
                splat_sigma=0.25, splat_cutoff=4.0, nslots=nslots, n_init=nslots, fill_seed=20261203)
    lo = (0.3, box[1] / 2 - 0.5, box[2] / 2 - 0.5)
    hi = (2.0, box[1] / 2 + 0.5, box[2] / 2 + 0.5)
    r.set_target(gaussian_target(box, splat_res, 20261203, centre_lo=lo, centre_hi=hi))
    th = (0.2 * np.random.default_rng(20261203).standard_normal((16, 16, 16))).astype(np.float32)
    r.forward(v, theta=th)
    r.backward(want_psi=False, want_parts=False)  # warm-up
    tf = tb = 0.0
    for _ in range(reps):
        t0 = time.perf_counter()
        loss = r.forward(v, theta=th)
        t1 = time.perf_counter()
        g = r.backward(want_psi=False, want_parts=False)
        t2 = time.perf_counter()
        tf += t1 - t0
        tb += t2 - t1
    r.close()
    isf.close()
    tf /= reps
    tb /= reps
    N = int(np.prod(dims))
    balg = b_alg_fwd_adjoint(N, nslots, steps, K)
    gbs = balg * steps / (tf + tb) / 1e9
    name = "x".join(map(str, dims))
    print(f"[bench] rollout {name} x 32 steps: forward {tf * 1e3:.2f} ms, backward {tb * 1e3:.2f} ms",
          file=sys.stderr, flush=True)
    return {
        "workload": f"{name} jet, 32 steps, checkpoint every 4, {nslots // 1000}k particles, 16^3 phase field, "
                    "splat MSE loss",
        "fwd_adjoint_steps_s": round(steps / (tf + tb), 2),
        "fwd_steps_s": round(steps / tf, 2),
        "ms_forward": round(tf * 1e3, 3),
        "ms_backward": round(tb * 1e3, 3),
        "B_alg_fwd_adjoint_per_step": int(balg),
        "achieved_GB_s": round(gbs, 1),
        "frac": round(gbs / HBM_PEAK_GBS, 4),
        "loss": loss,
        "grad_velocity_noise_floor": [float(x) for x in g["velocity"]],
        "grad_theta_norm_noise_floor": float(np.linalg.norm(g["theta"])),
        "gradient_note": "dL/dv and dL/dtheta are zero almost everywhere in this discretisation (a single-valued "
                         "phase added by the nozzle or by theta is removed exactly by the projection); the figures "
                         "are FP32 cancellation noise of the atomic velocity scatter, not a signal",
        "reps": reps,
        "timing": "wall clock per pass (set-up and constraint iterations included), synchronised",
    }


# ---------------------------------------------------------------- HBM traffic per kernel
# Two rocprofv3 counter passes (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch) over a short
# child run of this script (--traffic-probe), started before this process touches the GPU.
# Each counter is scaled by what it reads for k_normalize, a streaming kernel that reads
# and writes exactly 16 N bytes (64 MiB each way at 256x128x128).

def traffic_probe(w):
    """Child of the counter passes: psi normalize (the calibration kernel) and two steps of
    the bench workload (1M particle slots filled in the nozzle)."""
    from isf_amd import ISF, DeviceArray
    f32 = np.float32
    isf = ISF(w["vol_size"], w["dims"], w["hbar"], w["dt"])
    nx, ny, nz = w["dims"]
    psi_h = np.zeros((nz, ny, nx, 4), f32)
    psi_h[..., 0] = 1.0
    psi_h[..., 2] = f32(0.01)
    psi = isf.psi_array(psi_h)
    isf.normalize_(psi)
    v = np.array(w["jet_velocity"], f32)
    hb = f32(w["hbar"])
    isf.jet_configure(v / hb, f32(np.sum(v ** f32(2)) / (f32(2) * hb)),
                      np.array(w["nozzle_cen"], np.float64).astype(f32), w["nozzle_len"], w["nozzle_rad"],
                      w["n_spawn"], w["seed"], w["phase_mode"])
    parts = DeviceArray.from_host(np.full((w["n_slots"], 3), -1.0, f32))
    isf.fill_nozzle_(parts, 0, w["n_init"], w["seed"])
    for it in (1, 2):
        isf.step_(psi, parts, it)
    isf.synchronize()


def _counter_pass(counter, timeout_s=150):
    rocprof = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    out = tempfile.mkdtemp(prefix="isf_pmc_", dir="/tmp")
    cmd = [rocprof, "--pmc", counter, "--output-format", "csv", "-d", out, "-o", "probe", "--",
           sys.executable, os.path.abspath(__file__), "--traffic-probe"]
    env = dict(os.environ, TMPDIR="/tmp")
    try:
        subprocess.run(cmd, cwd="/tmp", env=env, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL,
                       timeout=timeout_s)
        files = glob.glob(os.path.join(out, "**", "*counter_collection.csv"), recursive=True)
        if not files:
            return None
        acc = {}
        with open(files[0]) as fh:
            for r in csv.DictReader(fh):
                acc.setdefault(r["Kernel_Name"], []).append(float(r["Counter_Value"]))
        return {k: sum(v) / len(v) for k, v in acc.items()}
    except (OSError, subprocess.SubprocessError):
        return None
    finally:
        shutil.rmtree(out, ignore_errors=True)


PARTICLE_STAGES = ("spawn", "particle_order", "advect")

# stage name -> kernel name pattern (flow planes are k_plane<NY, NZ, 0, E>, Poisson planes <.., 1, E>)
STAGE_KERNELS = {
    "flow_x_fwd": r"k_flow_x_fwd<", "flow_plane": r"k_plane<\d+, \d+, 0[,>]", "flow_x_inv_normalize": r"k_flow_x_inv<",
    "p1_div_r2c_x": r"k_div_r2c<", "p1_plane": r"k_plane<\d+, \d+, 1[,>]",
    "p1_c2r_x_gauge_restrict": r"k_c2r_x<\d+, \d+, 2>", "spawn": r"k_spawn", "p2_div_r2c_x": r"k_div_r2c<",
    "p2_plane": r"k_plane<\d+, \d+, 1[,>]", "p2_c2r_x_gauge": r"k_c2r_x<\d+, \d+, 1>",
    "velocity_sharp": r"k_velocity<", "advect": r"k_advect\(",
}


def measure_traffic(N):
    """{stage: HBM bytes per launch} from the two counter passes, and the calibration."""
    fetch = _counter_pass("FETCH_SIZE")
    write = _counter_pass("WRITE_SIZE") if fetch else None
    if not fetch or not write:
        return None, "rocprofv3 counter passes unavailable"

    def pick(table, pat):
        vals = [v for k, v in table.items() if re.search(pat, k)]
        return sum(vals) / len(vals) if vals else None

    ref_kib = 16 * N / 1024.0
    fn, wn = pick(fetch, r"k_normalize\("), pick(write, r"k_normalize\(")
    if not fn or not wn:
        return None, "calibration kernel missing from the counter passes"
    ff, fw = ref_kib / fn, ref_kib / wn
    out = {}
    for stage, pat in STAGE_KERNELS.items():
        f, wv = pick(fetch, pat), pick(write, pat)
        if f is not None and wv is not None:
            out[stage] = 1024.0 * (f * ff + wv * fw)
    note = (f"FETCH_SIZE x {ff:.3f}, WRITE_SIZE x {fw:.3f} (k_normalize, 16N bytes each way); "
            f"counts L2-to-fabric bytes, Infinity-Cache hits included")
    return out, note


def bench_export(isf, psi, parts, it, steps, every):
    """Step rate with the particle positions exported to pinned host memory every `every`
    steps (isf_export_push; the host takes the previous frame after each push)."""
    from isf_amd import ParticleExport
    ex = ParticleExport(isf, parts.shape[0])
    isf.synchronize()
    t0 = time.perf_counter()
    pushes = 0
    for k in range(steps):
        it += 1
        isf.step_(psi, parts, it)
        if k % every == 0:
            ex.push(parts)
            pushes += 1
            if pushes >= 2:
                ex.previous()
    ex.latest()
    isf.synchronize()
    dt = time.perf_counter() - t0
    ex.close()
    return it, {"every": every, "steps_s": round(steps / dt, 2), "pushes": pushes,
                "MB_per_push": round(parts.shape[0] * 12 / 1e6, 2)}


def bench_large(steps, warmup):
    """Large single jet: 512x256x256 (box 16x8x8, d = 1/32), 4M particle slots filled in the
    nozzle, 1000 spawned per step; forward steps through the driver's simloop (one
    isf_simloop call, as the headline), and one isf_step per step beside it; HIP events on
    the handle's stream."""
    w = dict(WORKLOAD, vol_size=(16, 8, 8), dims=(512, 256, 256), nozzle_cen=(0.3, 4 - 0.034, 4 + 0.066),
             n_slots=4_000_000, n_init=4_000_000)
    isf, psi, parts = setup_gpu(w)
    it = 0
    for _ in range(warmup):
        it += 1
        isf.step_(psi, parts, it)
    isf.synchronize()
    isf.timer_start()
    for _ in range(steps):
        it += 1
        isf.step_(psi, parts, it)
    ms_call = isf.timer_stop() / steps
    isf.simloop_(psi, parts, it + 1, warmup)
    it += warmup
    isf.timer_start()
    isf.simloop_(psi, parts, it + 1, steps)
    it += steps
    ms = isf.timer_stop() / steps
    isf.synchronize()
    N = 512 * 256 * 256
    norm = isf.step_norm2() / N
    balg = b_alg_step(N, w["n_slots"])
    psi.free()
    parts.free()
    isf.close()
    print(f"[bench] 512x256x256 / 4M particles: simloop {ms:.4f} ms/step, one isf_step per step {ms_call:.4f} ms",
          file=sys.stderr, flush=True)
    return {"workload": "single jet 512x256x256 (box 16x8x8), 4M particles seeded in the nozzle, 1000 spawned per step, "
                        "the driver's simloop (isf_simloop)",
            "steps_s": round(1000.0 / ms, 2), "ms_per_step": round(ms, 4),
            "isf_step_per_call": {"ms_per_step": round(ms_call, 4), "steps_s": round(1000.0 / ms_call, 2)},
            "B_alg_per_step": balg,
            "achieved_GB_s": round(balg / (ms * 1e-3) / 1e9, 1), "frac": round(balg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "norm2_over_N": round(norm, 9), "steps": steps, "warmup": warmup}


def bench_adjoint_batch(count=8, steps=32):
    """A batch of `count` differentiable rollouts at 128^3 (box 4x4x4), each with its own
    seeded nozzle velocity and target density, 32 steps with checkpoints every 4, 100k
    particles, run concurrently (one handle and one host thread per rollout)."""
    from concurrent.futures import ThreadPoolExecutor
    from isf_amd import ISF
    from isf_amd.rollout import Rollout, gaussian_target
    dims, box = (128, 128, 128), (4, 4, 4)
    hb = np.float32(0.1)
    rng = np.random.default_rng(20261203)
    jobs = []
    for i in range(count):
        ct = rng.uniform(np.cos(np.radians(20)), 1.0)
        ph = rng.uniform(0, 2 * np.pi)
        sp = rng.uniform(0.5, 1.5)
        st = np.sqrt(1 - ct * ct)
        v = np.array([sp * ct, sp * st * np.cos(ph), sp * st * np.sin(ph)], np.float32)
        isf = ISF(box, dims, hb, 1.0 / 48.0)
        isf.jet_configure(v / hb, np.float32(np.sum(v ** np.float32(2)) / (np.float32(2) * hb)),
                          np.array([0.3, 2 - 0.034, 2 + 0.066], np.float32), 0.5, 0.5, 1000, 1000 + i, 0)
        r = Rollout(isf, steps, checkpoint_every=4, constraint_iters=10, phase_res=(16, 16, 16),
                    splat_res=(16, 16, 16), splat_sigma=0.25, splat_cutoff=4.0, nslots=100_000, n_init=100_000,
                    fill_seed=2000 + i)
        r.set_target(gaussian_target(box, (16, 16, 16), 3000 + i, centre_lo=(0.3, 1.5, 1.5), centre_hi=(2.0, 2.5, 2.5)))
        th = (0.2 * rng.standard_normal((16, 16, 16))).astype(np.float32)
        jobs.append((isf, r, v, th))

    def fwd(j):
        return j[1].forward(j[2], theta=j[3])

    def fwd_bwd(j):
        j[1].forward(j[2], theta=j[3])
        return j[1].backward(want_psi=False, want_parts=False)

    with ThreadPoolExecutor(count) as ex:
        list(ex.map(fwd_bwd, jobs))  # warm-up
        t0 = time.perf_counter()
        losses = list(ex.map(fwd, jobs))
        t1 = time.perf_counter()
        list(ex.map(fwd_bwd, jobs))
        t2 = time.perf_counter()
    for isf, r, _, _ in jobs:
        r.close()
        isf.close()
    print(f"[bench] {count} rollouts 128^3 x {steps} steps: forward {t1 - t0:.3f} s, forward+adjoint {t2 - t1:.3f} s",
          file=sys.stderr, flush=True)
    return {"workload": f"{count} concurrent rollouts, 128^3 (box 4x4x4), {steps} steps, checkpoint every 4, "
                        "100k particles each, distinct seeded nozzle velocities and targets",
            "fwd_steps_s": round(count * steps / (t1 - t0), 2), "fwd_adjoint_steps_s": round(count * steps / (t2 - t1), 2),
            "s_forward": round(t1 - t0, 4), "s_forward_adjoint": round(t2 - t1, 4),
            "losses": [round(float(x), 9) for x in losses],
            "timing": "wall clock over all rollouts, set-up and constraint iterations included"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=90)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--profile-steps", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-ensemble leg")
    ap.add_argument("--batch-steps", type=int, default=60)
    ap.add_argument("--no-adjoint", action="store_true", help="skip the forward+adjoint rollout leg")
    ap.add_argument("--adjoint-reps", type=int, default=5)
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 counter passes")
    ap.add_argument("--no-large", action="store_true", help="skip the 512x256x256 and 128^3 adjoint-batch legs")
    ap.add_argument("--no-export", action="store_true", help="skip the particle export timing")
    ap.add_argument("--traffic-probe", action="store_true", help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.traffic_probe:
        traffic_probe(WORKLOAD)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only; --gpus {args.gpus} is not supported "
              f"(the multi-GPU slab decomposition is a design, not an implementation)", file=sys.stderr)
        sys.exit(2)

    w = WORKLOAD
    nx, ny, nz = w["dims"]
    N = nx * ny * nz
    traffic, traffic_note = None, "not measured (--no-traffic)"
    if not args.no_traffic:  # child processes, before this one touches the GPU
        print("[bench] rocprofv3 counter passes (FETCH_SIZE, WRITE_SIZE)", file=sys.stderr, flush=True)
        traffic, traffic_note = measure_traffic(N)
    print(f"[bench] setup {nx}x{ny}x{nz}, {w['n_slots']} particle slots", file=sys.stderr, flush=True)
    isf, psi, parts = setup_gpu(w)
    it = 0
    for _ in range(args.warmup):
        it += 1
        isf.step_(psi, parts, it)
    isf.synchronize()
    print(f"[bench] warmup {args.warmup} steps done", file=sys.stderr, flush=True)
    # one isf_step call per step (the step as a single C-ABI call)
    isf.timer_start()
    for _ in range(args.steps):
        it += 1
        isf.step_(psi, parts, it)
    ms_call = isf.timer_stop() / args.steps
    isf.synchronize()
    # the driver's simloop (test/runtests.jl:40-66, timed by examples/benchmark.jl:70): the
    # same steps in one isf_simloop call -- the headline figure
    isf.simloop_(psi, parts, it + 1, args.warmup)
    it += args.warmup
    isf.synchronize()
    t0 = time.perf_counter()
    isf.timer_start()
    isf.simloop_(psi, parts, it + 1, args.steps)
    it += args.steps
    ms = isf.timer_stop()
    isf.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = ms / args.steps
    steps_s = 1000.0 / ms_per_step
    norm = isf.step_norm2()
    print(f"[bench] timed {args.steps} steps: simloop {ms_per_step:.4f} ms/step (events), wall {wall:.3f} s, "
          f"norm2/N = {norm / N:.9f}; one isf_step per step {ms_call:.4f} ms", file=sys.stderr, flush=True)

    # per-stage HIP-event timing of the same step on the same stream
    acc = {}
    for _ in range(args.profile_steps):
        it += 1
        for k, v in isf.step_profile(psi, parts, it).items():
            acc.setdefault(k, []).append(v)
    stage_ms = {k: float(np.mean(v)) for k, v in acc.items()}
    sb = stage_bytes(nx, ny, nz, w["n_slots"], w["n_spawn"])
    # the dominant kernel: the largest share of the step's critical-path time, with the
    # launches of one kernel per step added up (both projections run the divergence + R2C
    # pass and the Poisson plane pass); the particle stages (spawn, order, advect) run on
    # the side stream, overlapped with the grid stages
    kernel_of = {"p1_div_r2c_x": "div_r2c_x", "p2_div_r2c_x": "div_r2c_x", "p1_plane": "poisson_plane",
                 "p2_plane": "poisson_plane"}
    per_kernel = {}
    for k, v in stage_ms.items():
        if k not in PARTICLE_STAGES and k != "velocity_sharp":
            per_kernel.setdefault(kernel_of.get(k, k), []).append(k)
    dom_kernel = max(per_kernel, key=lambda kk: sum(stage_ms[x] for x in per_kernel[kk])) if per_kernel else "flow_plane"
    dom_stages = per_kernel.get(dom_kernel, ["flow_plane"])
    dom = dom_stages[0]
    dom_ms = float(np.mean([stage_ms[x] for x in dom_stages])) if stage_ms else 1.0
    achieved = sb[dom] / (dom_ms * 1e-3) / 1e9 if stage_ms else 0.0
    stages = {k: {"ms": round(stage_ms[k], 5), "GB": round(sb[k] / 1e9, 4),
                  "GB_s": round(sb[k] / max(stage_ms[k] * 1e-3, 1e-9) / 1e9, 1)} for k in stage_ms}
    if traffic:
        for k in stages:
            if k in traffic:
                stages[k]["traffic_GB"] = round(traffic[k] / 1e9, 4)
    balg = b_alg_step(N, w["n_slots"])
    exports = []
    if not args.no_export:
        for every in (1, 4):
            it, rec = bench_export(isf, psi, parts, it, 60, every)
            exports.append(rec)
        print(f"[bench] export: {exports}", file=sys.stderr, flush=True)

    out = {
        "metric": METRIC,
        "value": round(steps_s, 2),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "plain single jet 256x128x128 (box 8x4x4, hbar 0.1, dt 1/48), 1M particles "
                               "seeded in the nozzle, 1000 spawned per step, forward steps through the driver's "
                               "simloop (isf_simloop)",
                   "grid": [nx, ny, nz], "particles": w["n_slots"], "parallelism": "single GPU"},
        "isf_step_per_call": {"ms_per_step": round(ms_call, 5), "steps_s": round(1000.0 / ms_call, 2),
                              "frac": round(b_alg_step(N, w["n_slots"]) * 1000.0 / ms_call / 1e9 / HBM_PEAK_GBS, 4)},
        "roofline": {"bound": "hbm", "kernel": dom_kernel, "launches_per_step": len(dom_stages),
                     "avg_launch_us": round(dom_ms * 1e3, 2), "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": round(traffic[dom]) if traffic and dom in traffic else None,
                     "algorithmic_bytes": sb.get(dom), "traffic_note": traffic_note},
        "step_roofline": {"B_alg_per_step": balg, "achieved_GB_s": round(balg * steps_s / 1e9, 1),
                          "frac": round(balg * steps_s / 1e9 / HBM_PEAK_GBS, 4)},
        "stages": stages,
        "export": exports or None,
    }
    # release the step leg's handle (its streams hold hardware queues, which the runtime
    # shares out among the live streams of the process) before the other legs
    psi.free()
    parts.free()
    isf.close()
    if not args.no_batch:
        print("[bench] batched ensemble leg", file=sys.stderr, flush=True)
        out["batched"] = bench_ensemble(args.batch_steps, 5)
    if not args.no_adjoint:
        print("[bench] forward+adjoint rollout legs", file=sys.stderr, flush=True)
        out["fwd_adjoint"] = bench_fwd_adjoint(args.adjoint_reps, dims=(256, 128, 128), box=(8, 4, 4),
                                               nslots=1_000_000, splat_res=(32, 16, 16))
        out["fwd_adjoint_128x64x64"] = bench_fwd_adjoint(args.adjoint_reps)
    if not args.no_large:
        print("[bench] large jet leg", file=sys.stderr, flush=True)
        out["large"] = bench_large(20, 3)
        print("[bench] adjoint batch leg", file=sys.stderr, flush=True)
        out["adjoint_batch"] = bench_adjoint_batch()
    if not args.no_cpu:
        print("[bench] CPU baseline leg", file=sys.stderr, flush=True)
        cb = cpu_baseline(w)
        out["cpu_baseline"] = cb
        out["gpu_over_cpu"] = round(steps_s / cb["value"], 1)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
