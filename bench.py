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

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline] [--no-surrogate] [--no-batch]
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
# FETCH_SIZE and WRITE_SIZE (one run each, median over the launches) and
# scaled by the same counters' reading for a 256 MiB device-to-device copy
# (profiles/k_substep_sand_1m_512.md): 29039 KiB x 2.000 + 80910 KiB x 1.000.
MEASURED_TRAFFIC = {"k_substep": (29039.0 * 2.000 + 80910.0 * 1.000) * 1024.0}


def build_workload(backend: str = "gpu", geometry: str = "column"):
    """The 1M-particle sand column on the GPU ("gpu") or the CPU oracle ("oracle" = f32,
    "oracle64" = the FP64 oracle).  geometry "column": 1,048,576 particles at 16 per
    cell in a 256 x 256-cell column on the floor (the headline); "interior": the other
    way to hold ~1M particles at 512^2, 4 per cell filling the whole 508 x 508-cell
    interior inside the slip walls (1,032,256 particles, jitter +-0.25 cell)."""
    cfg = eh.SimConfig.standard(512, 1.0 / 60.0, (0.0, -9.81))
    if geometry == "column":
        spawn = eh.SpawnRegion(spacing=0.25, box_size=(256, 256), box_center=(256.0, 130.0), position_jitter=1.0,
                               rng_seed=42)
    elif geometry == "interior":
        spawn = eh.SpawnRegion(spacing=0.5, box_size=(508, 508), box_center=(256.0, 256.0), position_jitter=0.5,
                               rng_seed=42)
    else:
        raise ValueError(f"unknown geometry {geometry!r}")
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
        sim = oracle_py.OracleSim.empty(cfg, 64 if backend == "oracle64" else 32).with_default_material(
            mat).with_boundary(eh.SlipBoundary.new(2))
    sim.add_body(spawn)
    return sim


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(target_s: float, first_dt: float) -> dict:
    """The oracle (f32 restatement of the reference's CPU Simulation::step) on a
    bounded sample: the same 1M-particle scene, a few substeps at the GPU run's
    first CFL dt, on one thread and then on every core this process may use
    (G2P phase 1 in parallel, as the reference's rayon loop runs it; P2G and
    the plastic update stay sequential, as in the reference)."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle_py
    cores = len(os.sched_getaffinity(0))
    threads = max(1, min(cores, int(os.environ.get("OMP_NUM_THREADS", cores))))
    rates = {}
    for t_count, share in ((1, 0.6), (threads, 0.4)):
        oracle_py.set_threads(t_count)
        try:
            sim = build_workload("oracle")
            n = sim.num_particles()
            t1 = sim.time_substeps(1, first_dt)
            k = max(1, min(400, int(target_s * share / max(t1, 1e-6))))
            t = sim.time_substeps(k, first_dt)
            sim.close()
        finally:
            oracle_py.set_threads(1)
        rates[t_count] = (n * k / t, k, t)
    v1, k1, t1 = rates[1]
    vn, kn, tn = rates[threads]
    return {"value": vn, "unit": "particle-substeps/s", "cores": threads, "kind": "port",
            "sample": f"{n} particles x {kn} substeps (dt {first_dt:.3e}) of the same workload, after 1 untimed "
                      f"substep; oracle f32, G2P phase 1 on {threads} threads, {tn:.1f} s",
            "single_thread": {"value": v1, "cores": 1, "sample": f"{n} particles x {k1} substeps, {t1:.1f} s"},
            "nproc": os.cpu_count(), "cores_available": cores, "cpu_model": _cpu_model()}


def full_run_section(args, tuning) -> dict:
    """The headline scene's whole run as its config states it (500 steps from
    rest, the column collapsing), timed as one run after the same autotuned
    launch settings, with the substep kernel's live average over all of it."""
    sim = build_workload("gpu")
    n = sim.num_particles()
    if tuning is not None:
        sim.set_tuning(tuning["chosen"]["reorder_interval"], bool(tuning["chosen"]["graphs"]))
    sim.set_profiling(True)
    t0 = time.perf_counter()
    sim.step_n(args.full_run_steps)
    sim.synchronize()
    elapsed = time.perf_counter() - t0
    substeps = sim.stats().total_substeps
    times = {k: v for k, v in sim.kernel_times().items() if v[0] > 0}
    sim.set_profiling(False)
    sim.close()
    launches, total_ms = times["k_substep"]
    avg_s = total_ms / launches / 1e3
    achieved = SUBSTEP_BYTES * n / avg_s / 1e9
    return {"steps": args.full_run_steps, "substeps": substeps,
            "particle_substeps_per_s": n * substeps / elapsed, "particle_steps_per_s": n * args.full_run_steps / elapsed,
            "seconds": elapsed, "k_substep_avg_us": avg_s * 1e6,
            "k_substep_roofline_frac": achieved / HBM_PEAK_GBS,
            "note": "includes the first step's upload, the physical reorders and the whole collapse"}


class _Hip:
    """The HIP runtime the library itself loaded (libamdhip64), through ctypes:
    device buffers, a stream and events for the stand-alone sort."""

    def __init__(self):
        import ctypes
        self.ct = ctypes
        self.lib = ctypes.CDLL("libamdhip64.so")

    def check(self, rc, what):
        if rc:
            raise RuntimeError(f"{what} failed ({rc})")

    def malloc(self, nbytes: int) -> int:
        p = self.ct.c_void_p()
        self.check(self.lib.hipMalloc(self.ct.byref(p), self.ct.c_size_t(nbytes)), "hipMalloc")
        return p.value

    def free(self, p: int) -> None:
        self.lib.hipFree(self.ct.c_void_p(p))

    def h2d(self, dst: int, a: np.ndarray) -> None:
        self.check(self.lib.hipMemcpy(self.ct.c_void_p(dst), self.ct.c_void_p(a.ctypes.data),
                                      self.ct.c_size_t(a.nbytes), 1), "hipMemcpy H2D")

    def d2h(self, a: np.ndarray, src: int) -> None:
        self.check(self.lib.hipMemcpy(self.ct.c_void_p(a.ctypes.data), self.ct.c_void_p(src),
                                      self.ct.c_size_t(a.nbytes), 2), "hipMemcpy D2H")

    def stream(self) -> int:
        s = self.ct.c_void_p()
        self.check(self.lib.hipStreamCreateWithFlags(self.ct.byref(s), 1), "hipStreamCreateWithFlags")
        return s.value

    def destroy_stream(self, s: int) -> None:
        self.lib.hipStreamDestroy(self.ct.c_void_p(s))

    def sync(self, s: int) -> None:
        self.check(self.lib.hipStreamSynchronize(self.ct.c_void_p(s)), "hipStreamSynchronize")

    def time_ms(self, fn, reps: int, stream: int) -> float:
        """Milliseconds between two events recorded on `stream` around `reps` calls of fn."""
        ct, hip = self.ct, self.lib
        ev = [ct.c_void_p(), ct.c_void_p()]
        for e in ev:
            self.check(hip.hipEventCreate(ct.byref(e)), "hipEventCreate")
        try:
            self.check(hip.hipEventRecord(ev[0], ct.c_void_p(stream)), "hipEventRecord")
            for _ in range(reps):
                fn()
            self.check(hip.hipEventRecord(ev[1], ct.c_void_p(stream)), "hipEventRecord")
            self.check(hip.hipEventSynchronize(ev[1]), "hipEventSynchronize")
            ms = ct.c_float()
            self.check(hip.hipEventElapsedTime(ct.byref(ms), ev[0], ev[1]), "hipEventElapsedTime")
            return float(ms.value)
        finally:
            for e in ev:
                hip.hipEventDestroy(e)


def interior_section(args, tuning) -> dict:
    """The same material and grid at 4 particles per cell over the whole interior
    (1,032,256 particles in 508 x 508 cells): `--interior-steps` steps after 20,
    with the headline's tuned launch settings and the substep kernel's live mean."""
    sim = build_workload("gpu", "interior")
    n = sim.num_particles()
    if tuning is not None:
        sim.set_tuning(tuning["chosen"]["reorder_interval"], bool(tuning["chosen"]["graphs"]))
    sim.step_n(20)
    sim.synchronize()
    sim.set_profiling(True)
    s0 = sim.stats().total_substeps
    t0 = time.perf_counter()
    sim.step_n(args.interior_steps)
    sim.synchronize()
    elapsed = time.perf_counter() - t0
    substeps = sim.stats().total_substeps - s0
    times = {k: v for k, v in sim.kernel_times().items() if v[0] > 0}
    sim.set_profiling(False)
    p = sim.particles()
    sim.close()
    launches, total_ms = times["k_substep"]
    avg_s = total_ms / launches / 1e3
    achieved = SUBSTEP_BYTES * n / avg_s / 1e9
    return {"workload": "drucker_prager_interior_4ppc_512", "particles": n, "steps": args.interior_steps,
            "substeps": substeps, "particle_substeps_per_s": n * substeps / elapsed,
            "k_substep_avg_us": avg_s * 1e6, "roofline_frac": achieved / HBM_PEAK_GBS,
            "state_finite": bool(np.all(np.isfinite(p["x"])))}


SORT_BYTES_PER_KEY = 16  # compulsory: the position read (8 B), the key and the permutation entry written


def sort_section(args) -> dict:
    """The cell-key sort on its own (emh_sort_cell_keys_noalloc on device buffers):
    the headline scene's 1,048,576 positions on its 512^2 grid (18-bit keys, two
    9-bit passes), in the order the scene generator writes them and in a random
    order, each sorted `--sort-reps` times back to back and timed with HIP events
    on the stream the sorts are enqueued on.  Every result is checked on the host
    (keys non-decreasing along the permutation, equal keys in increasing slot
    order, the permutation a bijection)."""
    sim = build_workload("gpu")
    x = np.ascontiguousarray(sim.particles()["x"], dtype=np.float32)
    sim.close()
    n, R = len(x), 512
    hip = _Hip()
    stream = hip.stream()
    d_xy, d_keys, d_perm = hip.malloc(8 * n), hip.malloc(4 * n), hip.malloc(4 * n)
    out = {"workload": "cell_key_sort_1m_512", "keys": n, "grid_res": R,
           "bytes_per_key": SORT_BYTES_PER_KEY, "reps": args.sort_reps}
    try:
        orders = {"generated_order": np.arange(n), "random_order": np.random.default_rng(11).permutation(n)}
        for name, order in orders.items():
            hip.h2d(d_xy, np.ascontiguousarray(x[order]))

            def run():
                eh.sort_cell_keys_noalloc(d_xy, n, R, d_keys, d_perm, stream)

            run()
            hip.sync(stream)
            k, p = np.empty(n, np.uint32), np.empty(n, np.uint32)
            hip.d2h(k, d_keys)
            hip.d2h(p, d_perm)
            ks, pi = k.astype(np.int64)[p], p.astype(np.int64)
            dk, dp = np.diff(ks), np.diff(pi)
            ok = bool(np.all(dk >= 0) and np.all((dk > 0) | (dp > 0)) and np.array_equal(np.sort(pi), np.arange(n)))
            us = hip.time_ms(run, args.sort_reps, stream) * 1e3 / args.sort_reps
            gbs = SORT_BYTES_PER_KEY * n / us / 1e3
            out[name] = {"us_per_sort": us, "keys_per_s": n / us * 1e6, "achieved_GBs": gbs,
                         "roofline_frac": gbs / HBM_PEAK_GBS, "stable_permutation": ok}
    finally:
        for d in (d_xy, d_keys, d_perm):
            hip.free(d)
        hip.destroy_stream(stream)
    return out


def build_multi_material(segments: bool = True):
    """Four-material scene at its stated size: 1024^2 grid, 370 seeded boxes with
    sides U[8, 96] at 4 particles per cell (3.98M particles; the stated 40 boxes
    hold only ~0.54M), materials in turn NeoHookean (400, 200), water (1000, 1e4),
    Stomakhin snow (E = 1.4e5, nu = 0.2) and corotated (400, 200), initial
    velocity scale 0.5, gravity (0, -9.81), step dt 0.02.  The water makes every
    substep run the density re-estimate."""
    cfg = eh.SimConfig.standard(1024, 0.02, (0.0, -9.81))
    lam, mu = eh.lame_from_young(1.4e5, 0.2)
    sim = (eh.Simulation.empty(cfg)
           .with_default_material(eh.NeoHookeanMaterial.new(400.0, 200.0))
           .with_material(1, eh.NewtonianFluidMaterial.low_viscosity(1000.0, 1e4))
           .with_material(2, eh.StomakhinMaterial.new(lam, mu, 10.0, 0.025, 0.0075, 0.6, 20.0))
           .with_material(3, eh.CorotatedMaterial.new(400.0, 200.0))
           .with_boundary(eh.SlipBoundary.new(2)))
    sim.set_tuning(sim.tuning()["reorder_interval"], False, segments)
    rng = np.random.default_rng(3)
    for k in range(370):
        w, h = (int(v) for v in rng.integers(8, 97, size=2))
        cx = float(rng.uniform(4 + w / 2 + 1, 1020 - w / 2 - 1))
        cy = float(rng.uniform(4 + h / 2 + 1, 1020 - h / 2 - 1))
        sim.add_body(eh.SpawnRegion(spacing=0.5, box_size=(w, h), box_center=(cx, cy), material_id=k % 4,
                                    rng_seed=1000 + k, initial_velocity_scale=0.5, position_jitter=0.1))
    return sim


# Per particle-substep for the NeoHookean / corotated / water / snow mix: the
# sand figure's fields minus friction_hardening and log_volume_strain (128 B).
MIXED_BYTES = 128


def multi_material_section(args) -> dict:
    """The four-material scene (density pass every substep) from its upload:
    the autotuner times reorder interval x hipGraph replay x per-material
    segments (the reorder grouping particles by model, so each wave runs one
    model's update) from the state after two steps and puts that state back;
    then `--mixed-steps` steps are timed with its choice.  Reports the share of
    waves that hold more than one model (where the update's per-model switch
    diverges) in the chosen order, and every candidate's cost."""
    out = {"workload": "multi_material_4m_1024", "bytes_per_particle_substep": MIXED_BYTES}
    sim = build_multi_material(False)
    n = sim.num_particles()
    sim.step_n(2)
    sim.synchronize()
    chosen, timed = sim.autotune(args.tune_steps)
    sim.step_n(1)
    sim.synchronize()
    # waves (64 consecutive device slots) holding more than one model
    mid = sim.particles()["material_id"][sim.slot_order()]
    nw = len(mid) // 64
    w = mid[:nw * 64].reshape(nw, 64)
    mixed_waves = float(np.mean(np.any(w != w[:, :1], axis=1)))
    sim.set_profiling(True)
    s0 = sim.stats().total_substeps
    t0 = time.perf_counter()
    sim.step_n(args.mixed_steps)
    sim.synchronize()
    elapsed = time.perf_counter() - t0
    subs = sim.stats().total_substeps - s0
    times = {k: v for k, v in sim.kernel_times().items() if v[0] > 0}
    sim.set_profiling(False)
    key = max((k for k in times if k != "step_setup"), key=lambda k: times[k][1])
    launches, total_ms = times[key]
    avg_s = total_ms / launches / 1e3
    p = sim.particles()
    sim.close()
    best = {seg: min(t["ns_per_particle_substep"] for t in timed if t["model_segments"] == seg) for seg in (0, 1)}
    out.update({"particles": n, "steps": args.mixed_steps, "particle_substeps_per_s": n * subs / elapsed,
                "substeps": subs, "seconds": elapsed, "kernels": key, "avg_substep_us": avg_s * 1e6,
                "roofline_frac": MIXED_BYTES * n / avg_s / 1e9 / HBM_PEAK_GBS, "mixed_model_waves": mixed_waves,
                "state_finite": bool(np.all(np.isfinite(p["x"]))),
                "tuning": {"chosen": chosen, "candidates": timed, "trial_steps": args.tune_steps},
                "segments_speedup": best[0] / best[1]})
    return out


# ------------------------------------------------------------ NACC surrogate
# Algorithmic MFMA work of one network evaluation: 2 * (6*64 + 64*64 + 64*64 + 64*3).
SURROGATE_FLOPS = 2 * (6 * 64 + 64 * 64 + 64 * 64 + 64 * 3)
MFMA_BF16_PEAK_TFLOPS = 2500.0  # MI355X dense bf16


def build_nacc_slope(seed: int):
    """NACC soil (E = 2e5, nu = 0.3, M = 1.85, beta = 0.5, xi = 0.8, F advanced
    before the projection) in 24 seeded random blocks (spacing 0.25, 16 per
    cell, about 2.0M particles) above a heightmap slope (terrain 120 -> 20
    cells across the 512^2 grid, terrain friction 0.3); gravity (0, -9.81),
    step dt 1/60."""
    cfg = eh.SimConfig.standard(512, 1.0 / 60.0, (0.0, -9.81))
    mat = eh.NaccMaterial.from_young_modulus(2e5, 0.3, 1.85, 0.5, 0.8, advance_f=True)
    sim = eh.Simulation.empty(cfg).with_default_material(mat)
    sim.clear_boundaries()
    sim.with_boundary(eh.HeightmapBoundary.new([20.0 + 100.0 * (1.0 - x / 511.0) for x in range(512)], 0.3, 2))
    rng = np.random.default_rng(seed)
    k = 0
    for row in range(4):
        for col in range(6):
            w, h = int(rng.integers(62, 79)), int(rng.integers(66, 85))
            cx = 16.0 + 80.0 * col + 40.0 + rng.uniform(-0.5, 0.5) * (80 - w - 2)
            cy = 150.0 + 87.5 * row + 43.75 + rng.uniform(-0.5, 0.5) * (87.5 - h - 2)
            sim.add_body(eh.SpawnRegion(spacing=0.25, box_size=(w, h), box_center=(float(cx), float(cy)),
                                        position_jitter=1.0, rng_seed=seed * 100 + k))
            k += 1
    return sim, mat


def surrogate_section(args) -> dict:
    """Learned NACC return mapping on the slope scene: train on trial states
    harvested from a plain run (seed 101), evaluate on every trial state of
    one substep of another run (seed 202) after it has landed."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import surrogate_ref  # the checker: Kirchhoff stress of a returned F

    t_all = time.perf_counter()
    sg = eh.NaccSurrogate(7)
    train_sim, mat = build_nacc_slope(101)
    n_train = train_sim.num_particles()
    train_sim.step_n(args.sg_warm_steps)
    cap = args.sg_records
    stride = max(1, int(n_train * 10 * args.sg_harvest_steps // cap))  # about 10 substeps per step
    train_sim.set_surrogate_mode(sg, "harvest", cap, stride)
    train_sim.step_n(args.sg_harvest_steps)
    train_sim.synchronize()
    train_sim.set_surrogate_mode(None, "plain")
    offered, stored = sg.harvest_count()
    train_sim.close()
    n_set = sg.build_training_set()
    t0 = time.perf_counter()
    loss = sg.train(epochs=args.sg_epochs, batch=args.sg_batch, lr=args.sg_lr, lr_final=args.sg_lr * 1e-2, seed=3)
    train_s = time.perf_counter() - t0

    ev, mat = build_nacc_slope(202)
    n = ev.num_particles()
    ev.step_n(args.sg_eval_steps)
    ev.set_surrogate_mode(sg, "harvest", 2 * n, 1)
    ev.step_n(1)
    ev.synchronize()
    ev.set_surrogate_mode(None, "plain")
    rec = sg.records(0, n)  # every particle's trial state in the step's first substep
    reps = args.sg_reps
    ms_plain, fp, ap = sg.eval_records(mat, 0, n, False, 1)
    sg.counts(reset=True)
    ms_sg, fs, as_ = sg.eval_records(mat, 0, n, True, 1)
    counts = sg.counts(reset=True)
    ms_plain, _, _ = sg.eval_records(mat, 0, n, False, reps, outputs=False)
    ms_sg, _, _ = sg.eval_records(mat, 0, n, True, reps, outputs=False)
    sg.counts(reset=True)

    f_trial = rec[:, :4]
    yl = ~np.all(fp == f_trial, axis=1)
    mu, kappa = float(rec[0, 5]), float(rec[0, 6])
    tp = surrogate_ref.nacc_kirchhoff(fp[yl], mu, kappa)
    ts = surrogate_ref.nacc_kirchhoff(fs[yl], mu, kappa)
    tt = surrogate_ref.nacc_kirchhoff(f_trial[yl], mu, kappa)
    agg = float(np.linalg.norm(ts - tp) / np.linalg.norm(tp))
    agg_trial = float(np.linalg.norm(tt - tp) / np.linalg.norm(tp))
    # per-particle error in units of the RMS stress of the yielding states
    per = np.linalg.norm(ts - tp, axis=1) / np.sqrt(np.mean(np.sum(tp * tp, axis=1)))

    # FP64 oracle on a sample of the yielding states
    import oracle_py
    oracle_py.load()
    om = oracle_py.material("nacc", mu, kappa, mat.params.stretch_limit, mat.params.compression_limit,
                            mat.params.hardening_exponent)
    om.ext.nacc_advance_f = 1
    idx = np.flatnonzero(yl)
    idx = idx[np.random.default_rng(0).permutation(len(idx))[:args.sg_fp64_sample]]
    f64, _ = oracle_py.nacc_project_batch(om, f_trial[idx], rec[idx, 4], 64)
    t64 = surrogate_ref.nacc_kirchhoff(f64, mu, kappa)
    e64 = lambda t: float(np.linalg.norm(t - t64) / np.linalg.norm(t64))
    e_plain64 = e64(surrogate_ref.nacc_kirchhoff(fp[idx], mu, kappa))
    e_sg64 = e64(surrogate_ref.nacc_kirchhoff(fs[idx], mu, kappa))

    # the whole substep with either return mapping (the scene continues)
    def substep_us(mode):
        ev.set_surrogate_mode(sg if mode == "surrogate" else None, mode)
        ev.set_profiling(True)
        ev.step_n(args.sg_substep_steps)
        ev.synchronize()
        t = ev.kernel_times().get("k_substep", (0, 0.0))
        ev.set_profiling(False)
        return t[1] / max(t[0], 1) * 1e3
    sub_plain = substep_us("plain")
    sub_sg = substep_us("surrogate")
    ev.set_surrogate_mode(None, "plain")
    ev.close()

    # CPU: the oracle's f32 return mapping, one thread, on a sample of the same states
    m_cpu = min(n, args.sg_cpu_sample)
    t0 = time.perf_counter()
    oracle_py.nacc_project_batch(om, f_trial[:m_cpu], rec[:m_cpu, 4], 32)
    cpu_s = time.perf_counter() - t0

    inferred = counts["surrogate"]
    tflops = SURROGATE_FLOPS * inferred / (ms_sg * 1e-3) / 1e12
    sg.close()
    return {
        "workload": "nacc_slope_2m_512",
        "particles": n,
        "speedup": ms_plain / ms_sg,
        "plain_ms": ms_plain,
        "surrogate_ms": ms_sg,
        "yielding": int(yl.sum()),
        "counts": counts,
        "stress_rel_err_vs_plain": agg,
        "stress_err_per_particle_p50": float(np.percentile(per, 50)),
        "stress_err_per_particle_p99": float(np.percentile(per, 99)),
        "stress_rel_err_trial_vs_plain": agg_trial,
        "stress_rel_err_vs_fp64": {"surrogate": e_sg64, "plain_f32": e_plain64, "sample": int(len(idx))},
        "target_err": 1e-3,
        "meets_target": bool(agg <= 1e-3 and ms_plain / ms_sg >= 1.5),
        "substep_us": {"plain": sub_plain, "surrogate": sub_sg, "speedup": sub_plain / sub_sg},
        "roofline": {"bound": "mfma", "kernel": "k_nacc_return_surrogate", "achieved": tflops,
                     "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": tflops / MFMA_BF16_PEAK_TFLOPS,
                     "flops_per_inference": SURROGATE_FLOPS},
        "training": {"records_offered": offered, "records_stored": stored, "harvest_stride": stride,
                     "samples": n_set, "epochs": args.sg_epochs, "batch": args.sg_batch, "seconds": train_s,
                     "loss_first_epoch": float(loss[0]), "loss_last_epoch": float(loss[-1]),
                     "scene": "nacc_slope seed 101", "particles": n_train},
        "cpu_baseline": {"value": m_cpu / cpu_s, "unit": "return-mappings/s", "cores": 1, "kind": "port",
                         "sample": f"{m_cpu} trial states of the evaluation substep, oracle f32 NaccMaterial::project"},
        "seconds": time.perf_counter() - t_all,
    }


# ------------------------------------------------------------ batched scenes
def batch_scene_specs(seed: int, n_scenes: int) -> list:
    """Independent scenes: grid side uniform over {64, 128, 192, 256}; particle
    count log-uniform from 2k to 64k (capped at 16 per cell over 35 % of the
    interior, so the smallest grids hold at most ~17.5k); 1-4 boxes, one per
    interior quadrant, each of Drucker-Prager sand (E = 3.5e5, nu = 0.3,
    30 deg), NeoHookean (400, 200) or Newtonian water (1000, 1e4) at random;
    gravity (0, -1), step dt 0.05 with each scene's own CFL substeps.  Plain
    data: (grid side, [(w, h, cx, cy, seed, material slot), ...]) per scene."""
    rng = np.random.default_rng(seed)
    specs = []
    for k in range(n_scenes):
        r = int(rng.choice([64, 128, 192, 256]))
        cap = min(64000, int(0.35 * (r - 8) ** 2 * 16))
        n_target = float(np.exp(rng.uniform(np.log(2000.0), np.log(cap))))
        boxes = int(rng.integers(1, 5))
        q = (r - 8) / 2.0
        bl = []
        for b in range(boxes):
            area = n_target / boxes / 16.0
            aspect = rng.uniform(0.6, 1.6)
            w = int(max(2, min(q - 2, round(np.sqrt(area * aspect)))))
            h = int(max(2, min(q - 2, round(area / w))))
            cx = 4.0 + q * (b % 2) + q / 2 + rng.uniform(-0.5, 0.5) * (q - w - 1)
            cy = 4.0 + q * (b // 2) + q / 2 + rng.uniform(-0.5, 0.5) * (q - h - 1)
            bl.append((w, h, float(cx), float(cy), seed * 1000 + k * 4 + b, int(rng.integers(0, 3))))
        specs.append((r, bl))
    return specs


def build_batch_scene(spec, backend: str = "gpu"):
    """One scene of batch_scene_specs on the GPU ("gpu") or the CPU oracle ("oracle32" / "oracle64")."""
    r, boxes = spec
    cfg = eh.SimConfig.standard(r, 0.05, (0.0, -1.0))
    if backend == "gpu":
        lam, mu = eh.lame_from_young(3.5e5, 0.3)
        sand = eh.DruckerPragerMaterial.new(lam, mu)
        sand.params.dp_h0 = eh.to_radians(30.0)
        solid = eh.NeoHookeanMaterial.new(400.0, 200.0)
        water = eh.NewtonianFluidMaterial.low_viscosity(1000.0, 1e4)
        sim = eh.Simulation.empty(cfg)
    else:
        sys.path.insert(0, os.path.join(REPO, "oracle"))
        import oracle_py
        lam, mu = oracle_py.lame_from_young(3.5e5, 0.3)
        sand = oracle_py.material("drucker_prager", lam, mu)
        sand.params.dp_h0 = oracle_py.to_radians(30.0)
        solid = oracle_py.material("neohookean", 400.0, 200.0)
        water = oracle_py.material("fluid", 1000.0, 1.0e-3, 1e4, 7.0)  # low_viscosity  fluid.rs:73-75
        sim = oracle_py.OracleSim.empty(cfg, 64 if backend == "oracle64" else 32)
    sim = sim.with_material(0, sand).with_material(1, solid).with_material(2, water).with_boundary(
        eh.SlipBoundary.new(2))
    for w, h, cx, cy, rs, m in boxes:
        sim.add_body(eh.SpawnRegion(spacing=0.25, box_size=(w, h), box_center=(cx, cy), position_jitter=0.5,
                                    rng_seed=rs, material_id=m))
    return sim


def build_batch_scenes(seed: int, n_scenes: int):
    return [build_batch_scene(sp) for sp in batch_scene_specs(seed, n_scenes)]


def batch_section(args) -> dict:
    """The scenes stepped as one batch against the same scenes stepped one
    after another (separate runs of the single-scene path)."""
    t0 = time.perf_counter()
    alone = build_batch_scenes(31, args.batch_scenes)
    together = build_batch_scenes(31, args.batch_scenes)
    n = np.array([s.num_particles() for s in alone], dtype=np.float64)
    for s in alone:
        s.set_reorder_interval(8)  # the batch path's reorder interval
        s.step_n(1)
    batch = eh.SceneBatch(together)
    batch.step_n(1)
    tuning = None
    if not args.no_autotune:
        # trials as long as the other sections': over fewer steps the decay of
        # the particle order between reorders does not show, and every
        # candidate times the same
        chosen, timed = batch.autotune(args.tune_steps)
        tuning = {"chosen": chosen, "candidates": timed, "trial_steps": args.tune_steps}
        batch.step_n(1)  # the re-upload after the trials, outside the timed region
        for s in alone:
            s.set_tuning(chosen["reorder_interval"], bool(chosen["graphs"]))
            s.step_n(1)
    build_s = time.perf_counter() - t0
    sa0 = np.array([s.stats().total_substeps for s in alone])
    sb0 = np.array([s.stats().total_substeps for s in together])
    t0 = time.perf_counter()
    for s in alone:
        s.step_n(args.batch_steps)
    t_alone = time.perf_counter() - t0
    batch.set_profiling(True)  # per-stage HIP-event spans, recorded while the steps run
    t0 = time.perf_counter()
    batch.step_n(args.batch_steps)
    t_batch = time.perf_counter() - t0
    stages = {k: {"launches": v[0], "ms": v[1]} for k, v in batch.kernel_times().items()}
    batch.set_profiling(False)
    sa = np.array([s.stats().total_substeps for s in alone]) - sa0
    sb = np.array([s.stats().total_substeps for s in together]) - sb0
    dx = 0.0
    for a, b in zip(alone, together):
        dx = max(dx, float(np.abs(a.particles()["x"].astype(np.float64) - b.particles()["x"]).max()))
    batch.close()
    for s in alone + together:
        s.close()
    work_a, work_b = float((n * sa).sum()), float((n * sb).sum())
    return {
        "workload": "batched_scenes",
        "scenes": len(n),
        "particles": int(n.sum()),
        "particles_min_max": [int(n.min()), int(n.max())],
        "steps": args.batch_steps,
        "substeps_per_scene_min_max": [int(sb.min()), int(sb.max())],
        "substeps_equal": bool(np.array_equal(sa, sb)),
        "max_abs_dx_cells": dx,
        "batch": {"particle_substeps_per_s": work_b / t_batch, "seconds": t_batch, "stages": stages},
        "separate": {"particle_substeps_per_s": work_a / t_alone, "seconds": t_alone},
        "speedup": (work_b / t_batch) / (work_a / t_alone),
        "build_s": build_s,
        "tuning": tuning,
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--profile-steps", type=int, default=0,
                    help="extra steps after the timed ones, added to the live kernel timing only")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-surrogate", action="store_true", help="skip the NACC surrogate section")
    ap.add_argument("--sg-warm-steps", type=int, default=200)
    ap.add_argument("--sg-harvest-steps", type=int, default=200)
    ap.add_argument("--sg-records", type=int, default=24_000_000)
    ap.add_argument("--sg-epochs", type=int, default=400)
    ap.add_argument("--sg-batch", type=int, default=16384)
    ap.add_argument("--sg-lr", type=float, default=1e-2)
    ap.add_argument("--sg-eval-steps", type=int, default=400)
    ap.add_argument("--sg-reps", type=int, default=20)
    ap.add_argument("--sg-substep-steps", type=int, default=10)
    ap.add_argument("--sg-fp64-sample", type=int, default=50_000)
    ap.add_argument("--sg-cpu-sample", type=int, default=200_000)
    ap.add_argument("--no-autotune", action="store_true", help="time the default launch settings")
    ap.add_argument("--tune-steps", type=int, default=10, help="timed steps per autotuner candidate")
    ap.add_argument("--headline-tune-steps", type=int, default=40,
                    help="timed steps per autotuner candidate on the headline scene: 40 steps are 520 substeps, "
                         "long enough to hold the first reorder of the 512-substep candidate, which shorter "
                         "trials cannot tell from 4096")
    ap.add_argument("--full-run-steps", type=int, default=500,
                    help="the whole collapse from rest, timed as one run (0: skip)")
    ap.add_argument("--mixed-steps", type=int, default=20,
                    help="timed steps of the four-material 1024^2 scene, per ordering (0: skip)")
    ap.add_argument("--interior-steps", type=int, default=50,
                    help="timed steps of the 4-per-cell whole-interior geometry (0: skip)")
    ap.add_argument("--sort-reps", type=int, default=50,
                    help="timed sorts of the headline positions, per input order (0: skip)")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-scenes section")
    ap.add_argument("--batch-scenes", type=int, default=256)
    ap.add_argument("--batch-steps", type=int, default=20)
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="after the bench line, run three device-to-device copies of this size (a known-size "
                         "kernel for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported: the substep path runs on a "
              "single GPU", file=sys.stderr)
        return 2

    t_build = time.perf_counter()
    sim = build_workload("gpu")
    n = sim.num_particles()
    sim.step_n(max(args.warmup, 1))
    sim.synchronize()
    first_dt = sim.effective_dt()
    # launch settings (reorder interval, hipGraph replay) timed on this scene
    # from this state, which the autotuner then puts back; one untimed step
    # takes the re-upload before the timed region
    tuning = None
    if not args.no_autotune:
        chosen, timed = sim.autotune(args.headline_tune_steps)
        tuning = {"chosen": chosen, "candidates": timed, "trial_steps": args.headline_tune_steps}
        sim.step_n(1)
        sim.synchronize()
    build_s = time.perf_counter() - t_build

    # live per-kernel timing over the timed steps: HIP events on the
    # library's stream, two per step, recorded without stopping the queue
    sim.set_profiling(True)
    s0 = sim.stats().total_substeps
    t0 = time.perf_counter()
    sim.step_n(args.steps)
    sim.synchronize()
    t1 = time.perf_counter()
    s1 = sim.stats().total_substeps
    elapsed = t1 - t0
    substeps = s1 - s0
    rate = n * substeps / elapsed
    if args.profile_steps > 0:
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
        "tuning": tuning,
    }
    _progress(f"headline: {rate:.4g} particle-substeps/s, {dominant} {avg_s * 1e6:.1f} us")
    sim.close()
    sections = [("cpu_baseline", not args.no_cpu_baseline, lambda: cpu_baseline(args.cpu_seconds, first_dt)),
                ("cell_sort", args.sort_reps > 0, lambda: sort_section(args)),
                ("full_run", args.full_run_steps > 0, lambda: full_run_section(args, tuning)),
                ("interior_4ppc", args.interior_steps > 0, lambda: interior_section(args, tuning)),
                ("multi_material", args.mixed_steps > 0, lambda: multi_material_section(args)),
                ("nacc_surrogate", not args.no_surrogate, lambda: surrogate_section(args)),
                ("batched_scenes", not args.no_batch, lambda: batch_section(args))]
    for name, on, run in sections:
        if on:
            _progress(f"{name} ...")
            t_sec = time.perf_counter()
        out[name] = run() if on else None
        if on:
            _progress(f"{name} done in {time.perf_counter() - t_sec:.1f} s")
    print(json.dumps(out), flush=True)
    if args.copy_calibration_mib > 0:
        copy_calibration(args.copy_calibration_mib << 20)
    return 0


def _progress(msg: str) -> None:
    """One line per finished part of the run, on stderr (stdout carries the JSON line)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


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
