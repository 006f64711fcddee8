#!/usr/bin/env python3
"""Throughput of the MLS-MPM substep on one MI355X.

Workload (the scene the headline metric is quoted on): Drucker-Prager sand
(E = 3.5e5, nu = 0.3, friction angle 30 deg) on a 512 x 512 grid with
SlipBoundary(2), 1,048,576 particles: a 256 x 256-cell column at spacing 0.25
(16 particles per cell; 1M particles at 4 per cell do not fit the 508^2
interior cells), jittered by +/-0.25 cell from seed 42, at rest, F = I,
gravity (0, -9.81), step dt 1/60 with the reference's adaptive CFL substeps.

A "step" is one Simulation::step (solver/step.rs:23-60): every adaptive
substep of config.dt.  `value` is particle-substeps per second over the K
timed steps; particle-steps per second is reported beside it.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "emerge-engine_amd"))

import numpy as np  # noqa: E402

import emerge_hip as eh  # noqa: E402

METRIC = "particle-steps/sec (1M sand, 512^2)"
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s

# Algorithmic bytes per particle-substep (Drucker-Prager, f32 SoA, 4-byte
# material id): compulsory reads x 8, v 8, C 16, F 16, mass 4, V0 4,
# material 4, density 4, q 4, log-volume-strain 4 (72 B) and writes x, v, C,
# F, q, log-volume-strain, volume, density (64 B).  k_substep carries the
# whole substep's particle traffic (G2P + update + the next P2G), so one
# launch moves 136 B per particle algorithmically.
SUBSTEP_BYTES = 136
ALGO_BYTES = {"k_substep": SUBSTEP_BYTES}
# HBM bytes per k_substep launch on this workload, measured with rocprofv3
# FETCH_SIZE and WRITE_SIZE (one run each, mean over launches) and scaled by
# the same counters' reading for a 256 MiB device-to-device copy
# (profiles/k_substep_sand_1m_512.md): 26861 KiB x 2.000 + 78440 KiB x 1.000.
MEASURED_TRAFFIC = {"k_substep": (26861.0 * 2.000 + 78440.0 * 1.000) * 1024.0}


def build_workload(backend: str = "gpu"):
    """The 1M-particle sand column on either backend (the oracle for the CPU baseline)."""
    cfg = eh.SimConfig.standard(512, 1.0 / 60.0, (0.0, -9.81))
    spawn = eh.SpawnRegion(spacing=0.25, box_size=(256, 256), box_center=(256.0, 130.0), position_jitter=1.0,
                           rng_seed=42)
    if backend == "gpu":
        lam, mu = eh.lame_from_young(3.5e5, 0.3)
        mat = eh.DruckerPragerMaterial.new(lam, mu)
        mat.params.dp_h0 = eh.to_radians(30.0)
        sim = eh.Simulation.empty(cfg).with_default_material(mat).with_boundary(eh.SlipBoundary.new(2))
    else:
        sys.path.insert(0, os.path.join(REPO, "oracle"))
        import oracle_py
        lam, mu = oracle_py.lame_from_young(3.5e5, 0.3)
        mat = oracle_py.material("drucker_prager", lam, mu)
        mat.params.dp_h0 = oracle_py.to_radians(30.0)
        sim = oracle_py.OracleSim.empty(cfg, 32).with_default_material(mat).with_boundary(eh.SlipBoundary.new(2))
    sim.add_body(spawn)
    return sim


def cpu_baseline(target_s: float, first_dt: float) -> dict:
    """The oracle (f32 restatement of the reference's CPU Simulation::step,
    single thread) on a bounded sample: the same 1M-particle scene, a few
    substeps at the GPU run's first CFL dt."""
    sim = build_workload("oracle")
    n = sim.num_particles()
    t1 = sim.time_substeps(1, first_dt)
    k = max(1, min(400, int(target_s / max(t1, 1e-6))))
    t = sim.time_substeps(k, first_dt)
    return {"value": n * k / t, "unit": "particle-substeps/s", "cores": 1, "kind": "port",
            "sample": f"{n} particles x {k} substeps (dt {first_dt:.3e}) of the same workload, after 1 untimed "
                      f"substep; oracle f32, 1 thread, {t:.1f} s"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--profile-steps", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="after the bench line, run three device-to-device copies of this size (a known-size "
                         "kernel for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the substep has no multi-GPU "
              "path in this build; see the sharding plan in DESIGN.md)", file=sys.stderr)
        return 2

    t_build = time.perf_counter()
    sim = build_workload("gpu")
    n = sim.num_particles()
    sim.step_n(max(args.warmup, 1))
    sim.synchronize()
    first_dt = sim.effective_dt()
    build_s = time.perf_counter() - t_build

    s0 = sim.stats().total_substeps
    t0 = time.perf_counter()
    sim.step_n(args.steps)
    sim.synchronize()
    t1 = time.perf_counter()
    s1 = sim.stats().total_substeps
    elapsed = t1 - t0
    substeps = s1 - s0
    rate = n * substeps / elapsed

    # live per-kernel timing (HIP events on the library's stream)
    sim.set_profiling(True)
    sim.step_n(args.profile_steps)
    sim.synchronize()
    times = sim.kernel_times()
    sim.set_profiling(False)
    active = {k: v for k, v in times.items() if v[0] > 0}
    dominant = max(active, key=lambda k: active[k][1])
    launches, total_ms = active[dominant]
    avg_s = total_ms / launches / 1e3
    algo = ALGO_BYTES.get(dominant, SUBSTEP_BYTES) * n
    achieved = algo / avg_s / 1e9
    substep_gbs = rate * SUBSTEP_BYTES / 1e9

    p = sim.particles()
    F = p["deformation_gradient"]
    finite = bool(np.all(np.isfinite(p["x"])) and np.all(np.isfinite(F)))

    out = {
        "metric": METRIC,
        "value": rate,
        "unit": "particle-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "drucker_prager_sand_column_1m_512", "grid_res": 512, "particles": n,
                   "particles_per_cell": 16, "material": "DruckerPrager E=3.5e5 nu=0.3 phi=30deg",
                   "step_dt": 1.0 / 60.0, "gravity": [0.0, -9.81], "boundary": "SlipBoundary(2)",
                   "parallelism": "single GPU"},
        "particle_steps_per_s": n * args.steps / elapsed,
        "substeps_timed": substeps,
        "substeps_per_step": substeps / args.steps,
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": MEASURED_TRAFFIC.get(dominant),
                     "algorithmic_bytes_per_launch": algo, "avg_launch_us": avg_s * 1e6},
        "substep_roofline": {"achieved": substep_gbs, "frac": substep_gbs / HBM_PEAK_GBS,
                             "bytes_per_particle_substep": SUBSTEP_BYTES},
        "kernel_avg_us": {k: v[1] / v[0] * 1e3 for k, v in active.items()},
        "state_finite": finite,
        "setup_s": build_s,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.cpu_seconds, first_dt)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)
    if args.copy_calibration_mib > 0:
        copy_calibration(args.copy_calibration_mib << 20)
    return 0


def copy_calibration(nbytes: int) -> None:
    """Three device-to-device copies of `nbytes` through the HIP runtime."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) or hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)):
        raise RuntimeError("hipMalloc failed")
    try:
        if hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)):
            raise RuntimeError("hipMemset failed")
        for _ in range(3):
            if hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3):  # hipMemcpyDeviceToDevice
                raise RuntimeError("hipMemcpy failed")
        if hip.hipDeviceSynchronize():
            raise RuntimeError("hipDeviceSynchronize failed")
    finally:
        hip.hipFree(a)
        hip.hipFree(b)


if __name__ == "__main__":
    sys.exit(main())
