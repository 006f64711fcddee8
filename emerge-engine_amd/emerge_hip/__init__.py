"""Python mirror of the reference crate's Simulation API over libemerge_hip.so.

The shared library (built from ``csrc/``) is the product: gfx950 HIP kernels
behind the C-ABI declared in ``include/emerge_hip.h``.  This module is a thin
ctypes layer with the reference's names, so that scenes read as they do in
Rust::

    config = SimConfig.standard(64, 0.05, (0.0, -1.0))
    sim = (Simulation.empty(config)
           .with_default_material(NeoHookeanMaterial.new(400.0, 200.0))
           .with_material(1, NewtonianFluidMaterial.low_viscosity(1000.0, 1e4))
           .with_boundary(SlipBoundary.new(2)))
    sim.add_body(SpawnRegion.for_sim(config).at(24.0, 40.0).box_of(12, 12))
    sim.step_n(60)
    particles = sim.particles()        # numpy structured array, 128 B records

(README.md:20-51 of the reference is the same scene.)  There is no CPU
fallback: if the library or a GPU is missing, calls raise ``EmergeError``.
"""
from __future__ import annotations

import ctypes
import dataclasses
import os
from typing import Optional, Sequence, Tuple

import numpy as np

__all__ = [
    "EmergeError", "SimConfig", "SpawnRegion", "SlipBoundary", "Material", "NeoHookeanMaterial",
    "CorotatedMaterial", "NewtonianFluidMaterial", "DruckerPragerMaterial", "StomakhinMaterial",
    "NaccMaterial", "FallbackMaterial", "Simulation", "PARTICLE_DTYPE", "StepStats", "lame_from_young",
    "to_radians", "sort_cell_keys", "cell_key", "library_path", "load_library", "device_count",
]

_HERE = os.path.dirname(os.path.abspath(__file__))
_PKG_ROOT = os.path.dirname(_HERE)

EMH_OK = 0
MODEL_FALLBACK, MODEL_FLUID, MODEL_NEOHOOKEAN, MODEL_COROTATED, MODEL_SNOW, MODEL_DRUCKER_PRAGER, MODEL_NACC = (
    0, 1, 2, 3, 4, 5, 10)


class EmergeError(RuntimeError):
    """A C-ABI call failed (the message is emh_last_error())."""

    def __init__(self, code: int, message: str):
        super().__init__(f"emerge_hip error {code}: {message}")
        self.code = code


# --------------------------------------------------------------- C structs
class CConfig(ctypes.Structure):
    """emh_config: SimConfig mirror (solver/config.rs:26-121)."""
    _fields_ = [
        ("grid_res", ctypes.c_uint32), ("grid_cell_size", ctypes.c_float), ("dt", ctypes.c_float),
        ("adaptive_timestep", ctypes.c_uint32), ("cfl_include_affine_speed", ctypes.c_uint32),
        ("cfl_coefficient", ctypes.c_float), ("material_cfl_coefficient", ctypes.c_float),
        ("viscous_timestep_coefficient", ctypes.c_float), ("min_dt", ctypes.c_float),
        ("project_invalid_state", ctypes.c_uint32), ("projection_min_density", ctypes.c_float),
        ("projection_min_volume", ctypes.c_float), ("projection_min_deformation_j", ctypes.c_float),
        ("gravity", ctypes.c_float * 2), ("boundary_thickness", ctypes.c_uint32),
        ("default_initial_volume", ctypes.c_float), ("recompute_density_each_step", ctypes.c_uint32),
        ("particle_mass", ctypes.c_float), ("max_substeps_per_step", ctypes.c_uint32),
        ("apic_blend", ctypes.c_float), ("j_max", ctypes.c_float), ("sleep_threshold", ctypes.c_float),
        ("contact_friction", ctypes.c_float), ("asflip_blend", ctypes.c_float),
        ("mixture_drag_coefficient", ctypes.c_float), ("mixture_pressure_iterations", ctypes.c_uint32),
        ("dx_meters", ctypes.c_float), ("dt_seconds", ctypes.c_float),
    ]


class CMaterialParams(ctypes.Structure):
    """emh_material_params: MaterialParams mirror, 96 bytes (materials/params.rs:8-101)."""
    _fields_ = [("model", ctypes.c_uint32)] + [(n, ctypes.c_float) for n in (
        "lambda_", "mu", "hardening_exponent", "compression_limit", "stretch_limit", "rest_density",
        "eos_stiffness", "eos_power", "dynamic_viscosity", "volume_ratio_min", "volume_ratio_max", "dp_h0",
        "dp_h1", "dp_h2", "dp_h3", "active_stress_coeff", "hardening_modulus", "thermal_viscosity_coeff",
        "thermal_expansion", "pressure_floor", "bulk_viscosity", "surface_tension_coeff", "cohesion_coeff")]


class CMaterialExt(ctypes.Structure):
    """emh_material_ext: model fields that MaterialParams drops."""
    _fields_ = [("min_density", ctypes.c_float), ("min_volume", ctypes.c_float),
                ("dp_volume_correction", ctypes.c_float), ("nacc_kappa", ctypes.c_float),
                ("nacc_hardening_enabled", ctypes.c_uint32), ("nacc_advance_f", ctypes.c_uint32),
                ("bingham_critical_shear_rate", ctypes.c_float), ("fluid_bingham", ctypes.c_uint32)]


class CSpawnRegion(ctypes.Structure):
    """emh_spawn_region: SpawnRegion mirror (solver/config.rs:317-364)."""
    _fields_ = [("spacing", ctypes.c_float), ("box_size", ctypes.c_int32 * 2), ("box_center", ctypes.c_float * 2),
                ("shape", ctypes.c_uint32), ("disk_radius", ctypes.c_float),
                ("initial_deformation_gradient", ctypes.c_float * 4),
                ("precompute_initial_volumes", ctypes.c_uint32), ("initial_velocity_scale", ctypes.c_float),
                ("position_jitter", ctypes.c_float), ("rng_seed", ctypes.c_uint32),
                ("material_id", ctypes.c_uint32), ("has_mass_override", ctypes.c_uint32),
                ("mass_override", ctypes.c_float)]


class CTrainOptions(ctypes.Structure):
    _fields_ = [("epochs", ctypes.c_uint32), ("batch", ctypes.c_uint32), ("lr", ctypes.c_float),
                ("lr_final", ctypes.c_float), ("beta1", ctypes.c_float), ("beta2", ctypes.c_float),
                ("eps", ctypes.c_float), ("pad", ctypes.c_uint32), ("seed", ctypes.c_uint64),
                ("max_steps", ctypes.c_uint64)]


class CSurrogateCounts(ctypes.Structure):
    _fields_ = [("yielding", ctypes.c_uint64), ("surrogate", ctypes.c_uint64), ("fallback", ctypes.c_uint64)]


class CKernelTime(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 32), ("launches", ctypes.c_uint64), ("total_ms", ctypes.c_double)]


class CTuning(ctypes.Structure):
    _fields_ = [("reorder_interval", ctypes.c_uint32), ("graphs", ctypes.c_uint32),
                ("ns_per_particle_substep", ctypes.c_float), ("model_segments", ctypes.c_uint32)]


class CStepStats(ctypes.Structure):
    _fields_ = [("last_substeps", ctypes.c_uint32), ("last_step_dt", ctypes.c_float),
                ("last_sim_time_dropped", ctypes.c_float), ("last_vel_clamp_count", ctypes.c_uint32),
                ("last_j_projection_count", ctypes.c_uint32), ("total_substeps", ctypes.c_uint32),
                ("frame_index", ctypes.c_uint64)]


class CSnapshot(ctypes.Structure):
    """SimSnapshot mirror (systems/diagnostics/snapshot.rs:8-80, without timing)."""
    _fields_ = [("frame_index", ctypes.c_uint64), ("configured_dt", ctypes.c_float), ("effective_dt", ctypes.c_float),
                ("substeps_last_step", ctypes.c_uint32), ("particle_count", ctypes.c_uint32),
                ("valid_particle_count", ctypes.c_uint32), ("active_grid_cells", ctypes.c_uint32),
                ("particles_per_active_cell", ctypes.c_float), ("mixed_material_cell_ratio", ctypes.c_float),
                ("mixed_material_particle_ratio", ctypes.c_float), ("total_particle_mass", ctypes.c_float),
                ("total_grid_mass", ctypes.c_float), ("relative_mass_error", ctypes.c_float),
                ("total_particle_momentum", ctypes.c_float * 2), ("total_grid_momentum", ctypes.c_float * 2),
                ("relative_momentum_error", ctypes.c_float), ("max_particle_speed", ctypes.c_float),
                ("max_grid_speed", ctypes.c_float), ("cfl_number", ctypes.c_float),
                ("min_deformation_j", ctypes.c_float), ("max_deformation_j", ctypes.c_float),
                ("out_of_bounds_particles", ctypes.c_uint32), ("invalid_physical_particle_values", ctypes.c_uint32),
                ("non_finite_particle_values", ctypes.c_uint32), ("non_finite_grid_values", ctypes.c_uint32),
                ("recommended_max_dt_from_velocity_cfl", ctypes.c_float), ("avg_plastic_jacobian", ctypes.c_float),
                ("min_plastic_jacobian", ctypes.c_float), ("avg_elastic_hardening", ctypes.c_float),
                ("active_count", ctypes.c_uint32), ("sleeping_count", ctypes.c_uint32),
                ("vel_clamp_count", ctypes.c_uint32), ("j_projection_count", ctypes.c_uint32),
                ("sim_time_dropped", ctypes.c_float), ("total_kinetic_energy", ctypes.c_float),
                ("max_pinned_particle_speed", ctypes.c_float), ("_pad", ctypes.c_uint32)]


assert ctypes.sizeof(CConfig) == 116
assert ctypes.sizeof(CMaterialParams) == 96
assert ctypes.sizeof(CMaterialExt) == 32

# Particle: 128 B repr(C) record (matter/particle.rs:11-144).
PARTICLE_DTYPE = np.dtype([
    ("x", "<f4", (2,)), ("v", "<f4", (2,)), ("velocity_gradient", "<f4", (4,)),
    ("deformation_gradient", "<f4", (4,)), ("mass", "<f4"), ("initial_volume", "<f4"), ("volume", "<f4"),
    ("density", "<f4"), ("material_id", "<u4"), ("plastic_volume_ratio", "<f4"), ("hardening_scale", "<f4"),
    ("friction_hardening", "<f4"), ("log_volume_strain", "<f4"), ("temperature", "<f4"), ("user_tag", "<u4"),
    ("activation", "<f4"), ("activation_dir", "<f4", (2,)), ("muscle_group_id", "<u4"), ("contact_group", "<u4"),
    ("sleeping", "<u4"), ("pinned", "<u4"), ("scalar_field", "<f4"), ("_pad", "<u4"),
])
assert PARTICLE_DTYPE.itemsize == 128


def zeroed_particles(n: int) -> np.ndarray:
    """Particle::zeroed() x n (particle.rs:148-174): F = I, Jp = h = 1, all else 0."""
    p = np.zeros(n, dtype=PARTICLE_DTYPE)
    p["deformation_gradient"][:] = (1.0, 0.0, 0.0, 1.0)
    p["plastic_volume_ratio"] = 1.0
    p["hardening_scale"] = 1.0
    return p


# ---------------------------------------------------------------- library
_lib = None


def library_path() -> str:
    return os.environ.get("EMERGE_HIP_LIB", os.path.join(_PKG_ROOT, "libemerge_hip.so"))


def load_library() -> ctypes.CDLL:
    """Load libemerge_hip.so (raises if it is not built: there is no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    path = library_path()
    if not os.path.exists(path):
        raise EmergeError(-3, f"{path} is not built; run `make -C emerge-engine_amd` or __graft_entry__.build()")
    lib = ctypes.CDLL(path)
    P = ctypes.POINTER
    sig = {
        "emh_abi_version": (ctypes.c_uint32, []),
        "emh_last_error": (ctypes.c_char_p, []),
        "emh_device_count": (ctypes.c_int, [P(ctypes.c_int)]),
        "emh_config_default": (ctypes.c_int, [P(CConfig)]),
        "emh_config_standard": (ctypes.c_int, [ctypes.c_uint32, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                               P(CConfig)]),
        "emh_config_unsafe_defaults": (ctypes.c_int, [P(CConfig)]),
        "emh_config_validate": (ctypes.c_int, [P(CConfig)]),
        "emh_lame_from_young": (ctypes.c_int, [ctypes.c_float, ctypes.c_float, P(ctypes.c_float),
                                               P(ctypes.c_float)]),
        "emh_to_radians": (ctypes.c_float, [ctypes.c_float]),
        "emh_material_neohookean": (ctypes.c_int, [ctypes.c_float, ctypes.c_float, P(CMaterialParams),
                                                   P(CMaterialExt)]),
        "emh_material_corotated": (ctypes.c_int, [ctypes.c_float, ctypes.c_float, P(CMaterialParams),
                                                  P(CMaterialExt)]),
        "emh_material_fluid": (ctypes.c_int, [ctypes.c_float] * 4 + [P(CMaterialParams), P(CMaterialExt)]),
        "emh_material_fluid_low_viscosity": (ctypes.c_int, [ctypes.c_float] * 2 + [P(CMaterialParams),
                                                                                   P(CMaterialExt)]),
        "emh_material_drucker_prager": (ctypes.c_int, [ctypes.c_float] * 2 + [P(CMaterialParams),
                                                                              P(CMaterialExt)]),
        "emh_material_snow": (ctypes.c_int, [ctypes.c_float] * 7 + [P(CMaterialParams), P(CMaterialExt)]),
        "emh_material_von_mises": (ctypes.c_int, [ctypes.c_float] * 4 + [P(CMaterialParams), P(CMaterialExt)]),
        "emh_material_viscoelastic": (ctypes.c_int, [ctypes.c_float] * 3 + [P(CMaterialParams), P(CMaterialExt)]),
        "emh_material_rankine": (ctypes.c_int, [ctypes.c_float] * 4 + [P(CMaterialParams), P(CMaterialExt)]),
        "emh_material_mu_i": (ctypes.c_int, [ctypes.c_float] * 5 + [P(CMaterialParams), P(CMaterialExt)]),
        "emh_material_mu_i_default": (ctypes.c_int, [ctypes.c_float] * 2 + [P(CMaterialParams), P(CMaterialExt)]),
        "emh_material_granular_fluid": (ctypes.c_int, [ctypes.c_float] * 6 + [P(CMaterialParams), P(CMaterialExt)]),
        "emh_material_bingham": (ctypes.c_int, [ctypes.c_float] * 5 + [P(CMaterialParams), P(CMaterialExt)]),
        "emh_material_snow_from_young": (ctypes.c_int, [ctypes.c_float] * 2 + [P(CMaterialParams),
                                                                               P(CMaterialExt)]),
        "emh_material_nacc": (ctypes.c_int, [ctypes.c_float] * 5 + [P(CMaterialParams), P(CMaterialExt)]),
        "emh_material_nacc_from_young": (ctypes.c_int, [ctypes.c_float] * 5 + [P(CMaterialParams),
                                                                               P(CMaterialExt)]),
        "emh_sim_create": (ctypes.c_int, [P(CConfig), P(ctypes.c_void_p)]),
        "emh_sim_create_from_spawn": (ctypes.c_int, [P(CConfig), P(CSpawnRegion), P(ctypes.c_void_p)]),
        "emh_sim_destroy": (None, [ctypes.c_void_p]),
        "emh_sim_set_material": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, P(CMaterialParams),
                                                P(CMaterialExt)]),
        "emh_sim_reinit_particle_state": (ctypes.c_int, [ctypes.c_void_p]),
        "emh_sim_add_slip_boundary": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32]),
        "emh_sim_export_particles_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]),
        "emh_sim_export_grid_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]),
        "emh_sim_set_gravity": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_float, ctypes.c_float]),
        "emh_sim_apply_impulse": (ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_float] * 5),
        "emh_sim_apply_radial_impulse": (ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_float] * 4),
        "emh_sim_add_friction_boundary": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_float]),
        "emh_sim_add_heightmap_boundary": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32,
                                                          ctypes.c_float, ctypes.c_uint32]),
        "emh_sim_add_predictive_boundary": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_float]),
        "emh_sim_clear_boundaries": (ctypes.c_int, [ctypes.c_void_p]),
        "emh_sim_add_body": (ctypes.c_int, [ctypes.c_void_p, P(CSpawnRegion), P(ctypes.c_uint32)]),
        "emh_sim_set_particles": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]),
        "emh_sim_num_particles": (ctypes.c_int, [ctypes.c_void_p, P(ctypes.c_size_t)]),
        "emh_sim_get_particles": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]),
        "emh_sim_get_config": (ctypes.c_int, [ctypes.c_void_p, P(CConfig)]),
        "emh_sim_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32]),
        "emh_sim_stats": (ctypes.c_int, [ctypes.c_void_p, P(CStepStats)]),
        "emh_sim_get_grid": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]),
        "emh_sim_snapshot": (ctypes.c_int, [ctypes.c_void_p, P(CSnapshot)]),
        "emh_sim_synchronize": (ctypes.c_int, [ctypes.c_void_p]),
        "emh_sim_set_profiling": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
        "emh_sim_set_reorder_interval": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32]),
        "emh_sim_get_slot_order": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]),
        "emh_sort_cell_keys_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint32,
                                                      ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
        "emh_sim_kernel_times": (ctypes.c_int, [ctypes.c_void_p, P(CKernelTime), ctypes.c_uint32,
                                                P(ctypes.c_uint32)]),
        "emh_surrogate_create": (ctypes.c_int, [ctypes.c_uint64, P(ctypes.c_void_p)]),
        "emh_surrogate_destroy": (ctypes.c_int, [ctypes.c_void_p]),
        "emh_surrogate_get_params": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
        "emh_surrogate_set_params": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
        "emh_sim_set_surrogate_mode": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32,
                                                      ctypes.c_uint64, ctypes.c_uint32]),
        "emh_surrogate_eval_records": (ctypes.c_int, [ctypes.c_void_p, P(CMaterialParams), P(CMaterialExt),
                                                      ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int, ctypes.c_uint32,
                                                      P(ctypes.c_float), ctypes.c_void_p, ctypes.c_void_p]),
        "emh_surrogate_harvest_count": (ctypes.c_int, [ctypes.c_void_p, P(ctypes.c_uint64), P(ctypes.c_uint64)]),
        "emh_surrogate_build_training_set": (ctypes.c_int, [ctypes.c_void_p, P(ctypes.c_uint64)]),
        "emh_surrogate_get_records": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64,
                                                     ctypes.c_void_p]),
        "emh_surrogate_set_training_set": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                                          ctypes.c_uint64]),
        "emh_surrogate_get_training_set": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64,
                                                          ctypes.c_void_p, ctypes.c_void_p]),
        "emh_surrogate_train": (ctypes.c_int, [ctypes.c_void_p, P(CTrainOptions), ctypes.c_void_p, ctypes.c_uint32]),
        "emh_surrogate_gradient": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32,
                                                  ctypes.c_void_p, P(ctypes.c_float)]),
        "emh_surrogate_infer": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p]),
        "emh_surrogate_get_counts": (ctypes.c_int, [ctypes.c_void_p, P(CSurrogateCounts), ctypes.c_int]),
        "emh_nacc_return_batch": (ctypes.c_int, [ctypes.c_void_p, P(CMaterialParams), P(CMaterialExt),
                                                 ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p,
                                                 ctypes.c_void_p, ctypes.c_uint32, P(ctypes.c_float)]),
        "emh_sim_sleep_tag": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32]),
        "emh_sim_wake_tag": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32]),
        "emh_batch_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, P(ctypes.c_void_p)]),
        "emh_batch_destroy": (ctypes.c_int, [ctypes.c_void_p]),
        "emh_batch_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32]),
        "emh_batch_get_tuning": (ctypes.c_int, [ctypes.c_void_p, P(CTuning)]),
        "emh_batch_set_tuning": (ctypes.c_int, [ctypes.c_void_p, P(CTuning)]),
        "emh_batch_autotune": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, P(CTuning), P(CTuning),
                                              ctypes.c_uint32, P(ctypes.c_uint32)]),
        "emh_batch_set_profiling": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
        "emh_batch_kernel_times": (ctypes.c_int, [ctypes.c_void_p, P(CKernelTime), ctypes.c_uint32,
                                                  P(ctypes.c_uint32)]),
        "emh_sim_get_tuning": (ctypes.c_int, [ctypes.c_void_p, P(CTuning)]),
        "emh_sim_set_tuning": (ctypes.c_int, [ctypes.c_void_p, P(CTuning)]),
        "emh_sim_autotune": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, P(CTuning), P(CTuning),
                                            ctypes.c_uint32, P(ctypes.c_uint32)]),
        "emh_cell_key": (ctypes.c_uint32, [ctypes.c_int32, ctypes.c_int32, ctypes.c_uint32]),
        "emh_sort_cell_keys": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint32, ctypes.c_void_p,
                                              ctypes.c_void_p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _check(rc: int) -> None:
    if rc != EMH_OK:
        raise EmergeError(rc, load_library().emh_last_error().decode("utf-8", "replace"))


def device_count() -> int:
    n = ctypes.c_int(0)
    _check(load_library().emh_device_count(ctypes.byref(n)))
    return n.value


def lame_from_young(young_modulus: float, poisson_ratio: float) -> Tuple[float, float]:
    """materials/utils.rs:183-193 (f32)."""
    lam, mu = ctypes.c_float(), ctypes.c_float()
    _check(load_library().emh_lame_from_young(young_modulus, poisson_ratio, ctypes.byref(lam), ctypes.byref(mu)))
    return lam.value, mu.value


def to_radians(deg: float) -> float:
    """f32::to_radians as Rust evaluates it."""
    return load_library().emh_to_radians(deg)


def cell_key(cx: int, cy: int, grid_res: int) -> int:
    return int(load_library().emh_cell_key(cx, cy, grid_res))


def sort_cell_keys(xy: np.ndarray, grid_res: int) -> Tuple[np.ndarray, np.ndarray]:
    """The device cell-key sort on its own: (keys per input, stable permutation)."""
    xy = np.ascontiguousarray(xy, dtype=np.float32).reshape(-1, 2)
    n = xy.shape[0]
    keys = np.zeros(n, dtype=np.uint32)
    perm = np.zeros(n, dtype=np.uint32)
    _check(load_library().emh_sort_cell_keys(xy.ctypes.data, n, grid_res, keys.ctypes.data, perm.ctypes.data))
    return keys, perm


SORT_NOALLOC_MAX_KEYS = 1048576  # EMH_SORT_NOALLOC_MAX_KEYS


def sort_cell_keys_noalloc(xy_dev: int, n: int, grid_res: int, keys_dev: int, perm_dev: int, stream: int = 0) -> None:
    """The allocation-free sort on device pointers (xy n x 2 f32, keys and perm n u32),
    enqueued on `stream` (a hipStream_t handle, 0 = null stream) and not waited for."""
    _check(load_library().emh_sort_cell_keys_noalloc(ctypes.c_void_p(xy_dev), n, grid_res, ctypes.c_void_p(keys_dev),
                                                     ctypes.c_void_p(perm_dev), ctypes.c_void_p(stream)))


# ----------------------------------------------------------------- config
class SimConfig:
    """SimConfig (solver/config.rs:26-121) backed by the C struct."""

    def __init__(self, c: Optional[CConfig] = None):
        if c is None:
            c = CConfig()
            _check(load_library().emh_config_default(ctypes.byref(c)))
        object.__setattr__(self, "_c", c)

    @classmethod
    def default(cls) -> "SimConfig":
        return cls()

    @classmethod
    def standard(cls, grid_res: int, dt: float, gravity: Sequence[float]) -> "SimConfig":
        c = CConfig()
        _check(load_library().emh_config_standard(grid_res, dt, gravity[0], gravity[1], ctypes.byref(c)))
        return cls(c)

    @classmethod
    def unsafe_defaults(cls) -> "SimConfig":
        c = CConfig()
        _check(load_library().emh_config_unsafe_defaults(ctypes.byref(c)))
        return cls(c)

    def validate(self) -> None:
        _check(load_library().emh_config_validate(ctypes.byref(self._c)))

    def copy(self) -> "SimConfig":
        c = CConfig()
        ctypes.pointer(c)[0] = self._c
        return SimConfig(c)

    def replace(self, **kw) -> "SimConfig":
        out = self.copy()
        for k, v in kw.items():
            setattr(out, k, v)
        return out

    def __getattr__(self, name):
        c = object.__getattribute__(self, "_c")
        v = getattr(c, name)
        if name == "gravity":
            return (v[0], v[1])
        return v

    def __setattr__(self, name, value):
        if name == "gravity":
            self._c.gravity[0], self._c.gravity[1] = value
        elif isinstance(value, bool):
            setattr(self._c, name, int(value))
        else:
            setattr(self._c, name, value)

    @property
    def c(self) -> CConfig:
        return self._c


# ------------------------------------------------------------------ spawn
@dataclasses.dataclass
class SpawnRegion:
    """SpawnRegion (solver/config.rs:317-375) with its fluent builders."""
    spacing: float = 1.0
    box_size: Tuple[int, int] = (16, 16)
    box_center: Tuple[float, float] = (32.0, 32.0)
    disk_radius: Optional[float] = None
    initial_deformation_gradient: Tuple[float, float, float, float] = (1.0, 0.0, 0.0, 1.0)
    precompute_initial_volumes: bool = False
    initial_velocity_scale: float = 0.0
    position_jitter: float = 0.0
    rng_seed: int = 1
    material_id: int = 0
    mass_override: Optional[float] = None

    @classmethod
    def for_sim(cls, config: SimConfig) -> "SpawnRegion":
        h = config.grid_res * 0.5
        return cls(box_center=(float(np.float32(h)), float(np.float32(h))))

    def at(self, x: float, y: float) -> "SpawnRegion":
        return dataclasses.replace(self, box_center=(x, y))

    def box_of(self, w: int, h: int) -> "SpawnRegion":
        return dataclasses.replace(self, box_size=(w, h))

    def disk(self, r: float) -> "SpawnRegion":
        side = int(np.ceil(np.float32(2.0) * np.float32(r))) + 1
        return dataclasses.replace(self, disk_radius=r, box_size=(side, side))

    def with_spacing(self, s: float) -> "SpawnRegion":
        return dataclasses.replace(self, spacing=s)

    def material(self, mid: int) -> "SpawnRegion":
        return dataclasses.replace(self, material_id=mid)

    def mass(self, m: float) -> "SpawnRegion":
        return dataclasses.replace(self, mass_override=m)

    def precompute_volumes(self) -> "SpawnRegion":
        return dataclasses.replace(self, precompute_initial_volumes=True)

    def velocity_scale(self, s: float) -> "SpawnRegion":
        return dataclasses.replace(self, initial_velocity_scale=s)

    def jitter(self, s: float) -> "SpawnRegion":
        return dataclasses.replace(self, position_jitter=s)

    def seed(self, seed: int) -> "SpawnRegion":
        return dataclasses.replace(self, rng_seed=seed)

    def to_c(self) -> CSpawnRegion:
        c = CSpawnRegion()
        c.spacing = self.spacing
        c.box_size[0], c.box_size[1] = int(self.box_size[0]), int(self.box_size[1])
        c.box_center[0], c.box_center[1] = self.box_center
        c.shape = 1 if self.disk_radius is not None else 0
        c.disk_radius = self.disk_radius or 0.0
        for k in range(4):
            c.initial_deformation_gradient[k] = self.initial_deformation_gradient[k]
        c.precompute_initial_volumes = int(bool(self.precompute_initial_volumes))
        c.initial_velocity_scale = self.initial_velocity_scale
        c.position_jitter = self.position_jitter
        c.rng_seed = int(self.rng_seed) & 0xFFFFFFFF
        c.material_id = int(self.material_id)
        c.has_mass_override = 0 if self.mass_override is None else 1
        c.mass_override = 0.0 if self.mass_override is None else self.mass_override
        return c


@dataclasses.dataclass(frozen=True)
class SlipBoundary:
    """SlipBoundary (forces/boundary/slip.rs:5-24)."""
    thickness: int

    @classmethod
    def new(cls, thickness: int) -> "SlipBoundary":
        return cls(thickness)


@dataclasses.dataclass(frozen=True)
class FrictionBoundary:
    """FrictionBoundary (forces/boundary/friction.rs): no-penetration walls with
    Coulomb friction on the tangential grid velocity."""
    thickness: int
    friction_coefficient: float

    @classmethod
    def new(cls, thickness: int, friction_coefficient: float) -> "FrictionBoundary":
        return cls(thickness, friction_coefficient)


@dataclasses.dataclass(frozen=True)
class HeightmapBoundary:
    """HeightmapBoundary (forces/boundary/heightmap.rs): terrain heights per x
    column (Coulomb friction on it) inside slip side and top walls."""
    heights: Tuple[float, ...]
    friction: float
    wall_thickness: int

    @classmethod
    def new(cls, heights, friction: float, wall_thickness: int) -> "HeightmapBoundary":
        return cls(tuple(float(h) for h in heights), friction, wall_thickness)

    @classmethod
    def flat_floor(cls, grid_res: int, floor_height: float, friction: float) -> "HeightmapBoundary":
        return cls.new([floor_height] * grid_res, friction, 2)


@dataclasses.dataclass(frozen=True)
class PredictiveBoundary:
    """PredictiveBoundary (forces/boundary/predictive.rs): slip walls with particles
    kept inside [wall_min, R - 1 - wall_min]."""
    thickness: int
    predictive_wall_min: float

    @classmethod
    def new(cls, thickness: int, predictive_wall_min: float) -> "PredictiveBoundary":
        return cls(thickness, predictive_wall_min)


# --------------------------------------------------------------- materials
class Material:
    """A constitutive model as the 96 B MaterialParams + the extension record."""

    def __init__(self, params: CMaterialParams, ext: CMaterialExt):
        self.params = params
        self.ext = ext

    @property
    def model(self) -> int:
        return self.params.model

    def __repr__(self) -> str:
        fields = {n: getattr(self.params, n) for n, _ in CMaterialParams._fields_ if getattr(self.params, n)}
        return f"{type(self).__name__}({fields})"


def _build(fn_name: str, *args) -> Tuple[CMaterialParams, CMaterialExt]:
    p, e = CMaterialParams(), CMaterialExt()
    _check(getattr(load_library(), fn_name)(*args, ctypes.byref(p), ctypes.byref(e)))
    return p, e


class FallbackMaterial(Material):
    """FallbackMaterial (materials/mod.rs:362-364): zero stress, no update."""

    def __init__(self):
        p, e = CMaterialParams(), CMaterialExt()
        e.min_density, e.min_volume, e.dp_volume_correction, e.nacc_hardening_enabled = 1e-6, 1e-6, 1.0, 2
        super().__init__(p, e)


class NeoHookeanMaterial(Material):
    """NeoHookeanMaterial (matter/materials/elastic.rs)."""

    @classmethod
    def new(cls, lam: float, mu: float) -> "NeoHookeanMaterial":
        return cls(*_build("emh_material_neohookean", lam, mu))

    @classmethod
    def from_young_modulus(cls, e: float, nu: float) -> "NeoHookeanMaterial":
        return cls.new(*lame_from_young(e, nu))


class CorotatedMaterial(Material):
    """CorotatedMaterial (matter/materials/corotated.rs)."""

    @classmethod
    def new(cls, lam: float, mu: float) -> "CorotatedMaterial":
        return cls(*_build("emh_material_corotated", lam, mu))

    @classmethod
    def from_young_modulus(cls, e: float, nu: float) -> "CorotatedMaterial":
        return cls.new(*lame_from_young(e, nu))


class NewtonianFluidMaterial(Material):
    """NewtonianFluidMaterial (matter/materials/fluid.rs)."""

    @classmethod
    def new(cls, rest_density: float, dynamic_viscosity: float, eos_stiffness: float,
            eos_power: float) -> "NewtonianFluidMaterial":
        return cls(*_build("emh_material_fluid", rest_density, dynamic_viscosity, eos_stiffness, eos_power))

    @classmethod
    def low_viscosity(cls, rest_density: float, eos_stiffness: float) -> "NewtonianFluidMaterial":
        return cls(*_build("emh_material_fluid_low_viscosity", rest_density, eos_stiffness))


class DruckerPragerMaterial(Material):
    """DruckerPragerMaterial (matter/materials/sand.rs)."""

    @classmethod
    def new(cls, lam: float, mu: float) -> "DruckerPragerMaterial":
        return cls(*_build("emh_material_drucker_prager", lam, mu))

    @classmethod
    def from_young_modulus(cls, e: float, nu: float) -> "DruckerPragerMaterial":
        return cls.new(*lame_from_young(e, nu))

    def with_friction_angle_deg(self, deg: float) -> "DruckerPragerMaterial":
        self.params.dp_h0 = to_radians(deg)
        return self


class VonMisesMaterial(Material):
    """VonMisesMaterial (matter/materials/von_mises.rs): J2 plasticity on the
    Hencky strain, yield stress + H * friction_hardening."""

    @classmethod
    def new(cls, lam: float, mu: float, yield_stress: float) -> "VonMisesMaterial":
        return cls.with_hardening(lam, mu, yield_stress, 0.0)

    @classmethod
    def with_hardening(cls, lam: float, mu: float, yield_stress: float, hardening_modulus: float) -> "VonMisesMaterial":
        return cls(*_build("emh_material_von_mises", lam, mu, yield_stress, hardening_modulus))

    @classmethod
    def from_young_modulus(cls, e: float, nu: float, yield_stress: float) -> "VonMisesMaterial":
        return cls.new(*lame_from_young(e, nu), yield_stress)

    @classmethod
    def soft_ductile(cls, e: float, nu: float) -> "VonMisesMaterial":
        """von_mises.rs:71-74: yield stress E / 200."""
        return cls.from_young_modulus(e, nu, e * 0.005)


class ViscoelasticMaterial(Material):
    """ViscoelasticMaterial (matter/materials/viscoelastic.rs): Kelvin-Voigt,
    neo-Hookean spring in parallel with a deviatoric dashpot."""

    @classmethod
    def new(cls, lam: float, mu: float, viscosity: float) -> "ViscoelasticMaterial":
        return cls(*_build("emh_material_viscoelastic", lam, mu, viscosity))

    @classmethod
    def from_young_modulus(cls, e: float, nu: float, viscosity: float) -> "ViscoelasticMaterial":
        return cls.new(*lame_from_young(e, nu), viscosity)

    @classmethod
    def near_incompressible(cls, e: float, viscosity: float) -> "ViscoelasticMaterial":
        return cls.from_young_modulus(e, 0.45, viscosity)


class RankineMaterial(Material):
    """RankineMaterial (matter/materials/rankine.rs): corotated elasticity with a
    softening tensile cutoff on the principal Kirchhoff stresses."""

    @classmethod
    def new(cls, lam: float, mu: float, tensile_strength: float, softening_rate: float) -> "RankineMaterial":
        return cls(*_build("emh_material_rankine", lam, mu, tensile_strength, softening_rate))

    @classmethod
    def from_young_modulus(cls, e: float, nu: float, tensile_strength: float, softening_rate: float):
        return cls.new(*lame_from_young(e, nu), tensile_strength, softening_rate)

    @classmethod
    def stiff_brittle(cls, e: float, nu: float) -> "RankineMaterial":
        """rankine.rs:78-86: tensile strength 2.5e-4 E, softening 2."""
        return cls.from_young_modulus(e, nu, e * np.float32(2.5e-4), 2.0)


class MuIRheologyMaterial(Material):
    """MuIRheologyMaterial (matter/materials/sand_mui.rs): rate-dependent
    Drucker-Prager, mu(I) stored in friction_hardening."""

    @classmethod
    def new(cls, lam: float, mu: float) -> "MuIRheologyMaterial":
        return cls(*_build("emh_material_mu_i_default", lam, mu))

    @classmethod
    def with_params(cls, lam, mu, mu_static, mu_dynamic, inertial_q) -> "MuIRheologyMaterial":
        return cls(*_build("emh_material_mu_i", lam, mu, mu_static, mu_dynamic, inertial_q))

    @classmethod
    def from_young_modulus(cls, e: float, nu: float) -> "MuIRheologyMaterial":
        return cls.new(*lame_from_young(e, nu))


class GranularFluidMaterial(Material):
    """GranularFluidMaterial (matter/materials/granular_fluid.rs): Tait EOS on
    rho0 / J plus corotated deviatoric stress, Stomakhin-style clamp."""

    @classmethod
    def new(cls, lam, mu, rest_density, eos_stiffness, hardening_exponent, compression_limit):
        return cls(*_build("emh_material_granular_fluid", lam, mu, rest_density, eos_stiffness,
                           hardening_exponent, compression_limit))

    @classmethod
    def saturated_loam(cls, e: float, nu: float) -> "GranularFluidMaterial":
        """granular_fluid.rs:101-117."""
        return cls.new(*lame_from_young(e, nu), 1.0, 200.0, 5.0, 0.4)


class BinghamFluidMaterial(Material):
    """BinghamFluidMaterial (matter/materials/bingham.rs): Tait pressure plus
    a viscoplastic deviatoric law; uploaded as a Fluid record."""

    @classmethod
    def new(cls, rest_density, dynamic_viscosity, eos_stiffness, eos_power, yield_stress):
        return cls(*_build("emh_material_bingham", rest_density, dynamic_viscosity, eos_stiffness,
                           eos_power, yield_stress))

    @classmethod
    def high_yield(cls, rest_density: float, eos_stiffness: float) -> "BinghamFluidMaterial":
        """bingham.rs:74-77: yield 100, viscosity 0.5."""
        return cls.new(rest_density, 0.5, eos_stiffness, 7.0, 100.0)


class StomakhinMaterial(Material):
    """StomakhinMaterial (matter/materials/snow.rs)."""

    @classmethod
    def new(cls, lam, mu, xi, theta_c, theta_s, jp_min, jp_max) -> "StomakhinMaterial":
        return cls(*_build("emh_material_snow", lam, mu, xi, theta_c, theta_s, jp_min, jp_max))

    @classmethod
    def from_young_modulus(cls, e: float, nu: float) -> "StomakhinMaterial":
        return cls(*_build("emh_material_snow_from_young", e, nu))


class NaccMaterial(Material):
    """NaccMaterial (matter/materials/nacc.rs).  advance_f: apply F <- (I + dt C) F
    before the projection (what the reference GPU path does); off by default,
    as in the CPU Simulation."""

    @classmethod
    def new(cls, mu, kappa, friction, cohesion, hardening_factor, advance_f=False) -> "NaccMaterial":
        m = cls(*_build("emh_material_nacc", mu, kappa, friction, cohesion, hardening_factor))
        m.ext.nacc_advance_f = int(bool(advance_f))
        return m

    @classmethod
    def from_young_modulus(cls, e, nu, friction, cohesion, hardening_factor, advance_f=False) -> "NaccMaterial":
        m = cls(*_build("emh_material_nacc_from_young", e, nu, friction, cohesion, hardening_factor))
        m.ext.nacc_advance_f = int(bool(advance_f))
        return m


# -------------------------------------------------------------- simulation
@dataclasses.dataclass
class StepStats:
    last_substeps: int
    last_step_dt: float
    last_sim_time_dropped: float
    last_vel_clamp_count: int
    last_j_projection_count: int
    total_substeps: int
    frame_index: int


def _tuning_dict(t: CTuning) -> dict:
    return {"reorder_interval": int(t.reorder_interval), "graphs": int(t.graphs),
            "model_segments": int(t.model_segments), "ns_per_particle_substep": float(t.ns_per_particle_substep)}


class _Tunable:
    """Launch tuning (reorder interval, hipGraph replay) and the autotuner
    (emh_sim_* / emh_batch_* tuning entry points)."""
    _tune_prefix = ""

    def _fn(self, name):
        return getattr(load_library(), f"emh_{self._tune_prefix}_{name}")

    def tuning(self) -> dict:
        t = CTuning()
        _check(self._fn("get_tuning")(self._h, ctypes.byref(t)))
        return _tuning_dict(t)

    def set_tuning(self, reorder_interval: int, graphs: bool, model_segments: Optional[bool] = None) -> None:
        if model_segments is None:
            model_segments = bool(self.tuning()["model_segments"])
        t = CTuning(int(reorder_interval), 1 if graphs else 0, 0.0, 1 if model_segments else 0)
        _check(self._fn("set_tuning")(self._h, ctypes.byref(t)))

    def autotune(self, trial_steps: int = 10) -> Tuple[dict, list]:
        """Time every candidate setting on this scene (state put back afterwards), apply
        the fastest; returns (chosen, [every candidate with its cost])."""
        best = CTuning()
        timed = (CTuning * 64)()
        n = ctypes.c_uint32()
        _check(self._fn("autotune")(self._h, int(trial_steps), ctypes.byref(best), timed, 64, ctypes.byref(n)))
        return _tuning_dict(best), [_tuning_dict(timed[i]) for i in range(min(n.value, 64))]

    def autotune_cached(self, key: str, cache_path: str, trial_steps: int = 10) -> dict:
        """autotune() once per `key` (a workload signature), the choice kept in the JSON
        file `cache_path`; later calls with the same key apply the cached choice."""
        import json
        cache = {}
        if os.path.exists(cache_path):
            with open(cache_path) as f:
                cache = json.load(f)
        if key in cache:
            c = cache[key]
            self.set_tuning(c["reorder_interval"], bool(c["graphs"]), bool(c.get("model_segments", 1)))
            return dict(c, cached=True)
        best, timed = self.autotune(trial_steps)
        cache[key] = dict(best, candidates=timed)
        tmp = cache_path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(cache, f, indent=1)
        os.replace(tmp, cache_path)
        return dict(best, candidates=timed, cached=False)


class Simulation(_Tunable):
    """Simulation (spacetime/solver/mod.rs:37-75) on one GPU."""
    _tune_prefix = "sim"

    def __init__(self, handle: ctypes.c_void_p, config: SimConfig):
        self._h = handle
        self.config = config

    @classmethod
    def empty(cls, config: SimConfig) -> "Simulation":
        h = ctypes.c_void_p()
        _check(load_library().emh_sim_create(ctypes.byref(config.c), ctypes.byref(h)))
        return cls(h, config.copy())

    @classmethod
    def new(cls, config: SimConfig, spawn: SpawnRegion) -> "Simulation":
        h = ctypes.c_void_p()
        sc = spawn.to_c()
        _check(load_library().emh_sim_create_from_spawn(ctypes.byref(config.c), ctypes.byref(sc), ctypes.byref(h)))
        return cls(h, config.copy())

    def close(self) -> None:
        if self._h:
            load_library().emh_sim_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # builders (lifecycle.rs)
    def set_material(self, material_id: int, material: Material) -> None:
        _check(load_library().emh_sim_set_material(self._h, material_id, ctypes.byref(material.params),
                                                   ctypes.byref(material.ext)))

    def with_default_material(self, material: Material) -> "Simulation":
        self.set_material(0, material)
        return self

    def with_material(self, material_id: int, material: Material) -> "Simulation":
        self.set_material(material_id, material)
        return self

    def with_boundary(self, boundary) -> "Simulation":
        self.add_boundary_condition(boundary)
        return self

    def add_boundary_condition(self, boundary) -> None:
        lib = load_library()
        if isinstance(boundary, SlipBoundary):
            _check(lib.emh_sim_add_slip_boundary(self._h, boundary.thickness))
        elif isinstance(boundary, FrictionBoundary):
            _check(lib.emh_sim_add_friction_boundary(self._h, boundary.thickness, boundary.friction_coefficient))
        elif isinstance(boundary, HeightmapBoundary):
            h = np.ascontiguousarray(boundary.heights, np.float32)
            _check(lib.emh_sim_add_heightmap_boundary(self._h, h.ctypes.data, len(h), boundary.friction,
                                                      boundary.wall_thickness))
        elif isinstance(boundary, PredictiveBoundary):
            _check(lib.emh_sim_add_predictive_boundary(self._h, boundary.thickness, boundary.predictive_wall_min))
        else:
            raise EmergeError(-2, f"{type(boundary).__name__} is not supported on the GPU path")

    def clear_boundaries(self) -> None:
        _check(load_library().emh_sim_clear_boundaries(self._h))

    def reinit_all_particle_state(self) -> None:
        _check(load_library().emh_sim_reinit_particle_state(self._h))

    def add_body(self, spawn: SpawnRegion) -> int:
        tag = ctypes.c_uint32()
        sc = spawn.to_c()
        _check(load_library().emh_sim_add_body(self._h, ctypes.byref(sc), ctypes.byref(tag)))
        return tag.value

    def set_particles(self, particles: np.ndarray) -> None:
        arr = np.ascontiguousarray(particles, dtype=PARTICLE_DTYPE)
        _check(load_library().emh_sim_set_particles(self._h, arr.ctypes.data, arr.shape[0]))

    def num_particles(self) -> int:
        n = ctypes.c_size_t()
        _check(load_library().emh_sim_num_particles(self._h, ctypes.byref(n)))
        return n.value

    def particles(self) -> np.ndarray:
        n = self.num_particles()
        out = np.zeros(n, dtype=PARTICLE_DTYPE)
        _check(load_library().emh_sim_get_particles(self._h, out.ctypes.data, n))
        return out

    # the hot path
    def step(self) -> None:
        _check(load_library().emh_sim_step(self._h, 1))

    def step_n(self, steps: int) -> None:
        _check(load_library().emh_sim_step(self._h, steps))

    def synchronize(self) -> None:
        _check(load_library().emh_sim_synchronize(self._h))

    def stats(self) -> StepStats:
        s = CStepStats()
        _check(load_library().emh_sim_stats(self._h, ctypes.byref(s)))
        return StepStats(s.last_substeps, s.last_step_dt, s.last_sim_time_dropped, s.last_vel_clamp_count,
                         s.last_j_projection_count, s.total_substeps, s.frame_index)

    def effective_dt(self) -> float:
        return self.stats().last_step_dt

    def last_substeps(self) -> int:
        return self.stats().last_substeps

    def set_profiling(self, enable: bool) -> None:
        """Step-level HIP-event timing on the library stream, without stopping its queue (resets the totals)."""
        _check(load_library().emh_sim_set_profiling(self._h, int(bool(enable))))

    def set_reorder_interval(self, substeps: int) -> None:
        """Physical cell-key reorder every `substeps` substeps (and at the first step after an upload)."""
        _check(load_library().emh_sim_set_reorder_interval(self._h, int(substeps)))

    def slot_order(self) -> np.ndarray:
        """Caller index of the particle in each device slot (the stored order)."""
        n = self.num_particles()
        out = np.zeros(n, np.uint32)
        _check(load_library().emh_sim_get_slot_order(self._h, out.ctypes.data, n))
        return out

    def kernel_times(self) -> dict:
        """{"step_setup": (steps, ms), "k_substep": (substeps, ms)} since profiling was enabled."""
        arr = (CKernelTime * 16)()
        n = ctypes.c_uint32()
        _check(load_library().emh_sim_kernel_times(self._h, arr, 16, ctypes.byref(n)))
        return {arr[i].name.decode(): (int(arr[i].launches), float(arr[i].total_ms)) for i in range(min(n.value, 16))}

    def diagnostics_snapshot(self) -> dict:
        """Simulation::diagnostics_snapshot (solver/queries.rs:15-30) as a dict."""
        out = CSnapshot()
        _check(load_library().emh_sim_snapshot(self._h, ctypes.byref(out)))
        d = {}
        for name, _ in CSnapshot._fields_:
            if name == "_pad":
                continue
            v = getattr(out, name)
            d[name] = tuple(v) if isinstance(v, ctypes.Array) else v
        return d

    def set_surrogate_mode(self, surrogate: Optional["NaccSurrogate"], mode: str,
                           harvest_capacity: int = 0, harvest_stride: int = 1) -> None:
        """NACC return mapping of the following substeps: "plain", "harvest" (plain,
        recording every stride-th trial state into the surrogate's buffer) or "surrogate"."""
        code = {"plain": 0, "harvest": 1, "surrogate": 2}[mode]
        _check(load_library().emh_sim_set_surrogate_mode(self._h, surrogate._h if surrogate else None, code,
                                                          int(harvest_capacity), int(harvest_stride)))
        self._surrogate = surrogate  # keeps it alive while the sim uses it

    def sleep_tag(self, tag: int) -> None:
        """Simulation::sleep_tag (solver/particles.rs:293-312)."""
        _check(load_library().emh_sim_sleep_tag(self._h, int(tag)))

    def wake_tag(self, tag: int) -> None:
        """Simulation::wake_tag (solver/particles.rs:270-291)."""
        _check(load_library().emh_sim_wake_tag(self._h, int(tag)))

    def set_gravity(self, gravity) -> None:
        """Simulation::set_gravity (lifecycle.rs:403-405)."""
        _check(load_library().emh_sim_set_gravity(self._h, float(gravity[0]), float(gravity[1])))
        self.config = self.config.replace(gravity=(float(np.float32(gravity[0])), float(np.float32(gravity[1]))))

    def apply_impulse(self, center, radius: float, force) -> None:
        """Simulation::apply_impulse (solver/particles.rs:115-139), on the device."""
        _check(load_library().emh_sim_apply_impulse(self._h, float(center[0]), float(center[1]), float(radius),
                                                    float(force[0]), float(force[1])))

    def apply_radial_impulse(self, center, radius: float, strength: float) -> None:
        """Simulation::apply_radial_impulse (solver/particles.rs:141-165), on the device."""
        _check(load_library().emh_sim_apply_radial_impulse(self._h, float(center[0]), float(center[1]),
                                                           float(radius), float(strength)))

    def export_particles_device(self, dst_ptr: int, capacity: int) -> None:
        """Writes the particles (128 B records, caller order) into device memory at dst_ptr."""
        _check(load_library().emh_sim_export_particles_device(self._h, ctypes.c_void_p(dst_ptr), capacity))

    def export_grid_device(self, dst_ptr: int, capacity_cells: int) -> None:
        """Writes grid_res^2 cells {momentum.x, momentum.y, mass, 0} into device memory at dst_ptr."""
        _check(load_library().emh_sim_export_grid_device(self._h, ctypes.c_void_p(dst_ptr), capacity_cells))

    def grid(self) -> Tuple[np.ndarray, np.ndarray]:
        """(mass[R*R], momentum[R*R, 2]) of the last substep's P2G, flat index x*R+y."""
        r = self.config.grid_res
        m = np.zeros(r * r, dtype=np.float32)
        mom = np.zeros((r * r, 2), dtype=np.float32)
        _check(load_library().emh_sim_get_grid(self._h, m.ctypes.data, mom.ctypes.data, r * r))
        return m, mom


# -------------------------------------------------------------- NACC surrogate
SURROGATE_PARAMS = 8963
SURROGATE_NORM = 20


class NaccSurrogate:
    """Learned NACC return mapping (a 6-64-64-64-3 SiLU MLP on BF16 MFMA) on one GPU."""

    def __init__(self, seed: int = 0):
        h = ctypes.c_void_p()
        _check(load_library().emh_surrogate_create(int(seed), ctypes.byref(h)))
        self._h = h

    def close(self) -> None:
        if self._h:
            load_library().emh_surrogate_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self) -> Tuple[np.ndarray, np.ndarray]:
        w = np.zeros(SURROGATE_PARAMS, dtype=np.float32)
        nm = np.zeros(SURROGATE_NORM, dtype=np.float32)
        _check(load_library().emh_surrogate_get_params(self._h, w.ctypes.data, nm.ctypes.data))
        return w, nm

    def set_params(self, params: Optional[np.ndarray] = None, norm: Optional[np.ndarray] = None) -> None:
        w = None if params is None else np.ascontiguousarray(params, dtype=np.float32)
        nm = None if norm is None else np.ascontiguousarray(norm, dtype=np.float32)
        if w is not None and w.size != SURROGATE_PARAMS:
            raise ValueError("params has the wrong size")
        if nm is not None and nm.size != SURROGATE_NORM:
            raise ValueError("norm has the wrong size")
        _check(load_library().emh_surrogate_set_params(self._h, None if w is None else w.ctypes.data,
                                                       None if nm is None else nm.ctypes.data))

    def harvest_count(self) -> Tuple[int, int]:
        a, b = ctypes.c_uint64(0), ctypes.c_uint64(0)
        _check(load_library().emh_surrogate_harvest_count(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def records(self, first: int = 0, n: Optional[int] = None) -> np.ndarray:
        """Harvested trial states [n, 12]: F_trial (a, b, c, d), alpha, mu, kappa, M, beta, xi, flags, 0."""
        if n is None:
            n = self.harvest_count()[1] - first
        out = np.zeros((n, 12), dtype=np.float32)
        _check(load_library().emh_surrogate_get_records(self._h, first, n, out.ctypes.data))
        return out

    def eval_records(self, material: "NaccMaterial", first: int, n: int, use_surrogate: bool, reps: int = 1,
                     outputs: bool = True):
        """Batch return mapping of harvested records on the device; (ms per launch, F [n, 4], alpha [n])."""
        ms = ctypes.c_float(0.0)
        f = np.zeros((n, 4), dtype=np.float32) if outputs else None
        a = np.zeros(n, dtype=np.float32) if outputs else None
        _check(load_library().emh_surrogate_eval_records(
            self._h, ctypes.byref(material.params), ctypes.byref(material.ext), first, n, int(use_surrogate), reps,
            ctypes.byref(ms), None if f is None else f.ctypes.data, None if a is None else a.ctypes.data))
        return ms.value, f, a

    def build_training_set(self) -> int:
        n = ctypes.c_uint64(0)
        _check(load_library().emh_surrogate_build_training_set(self._h, ctypes.byref(n)))
        return n.value

    def set_training_set(self, x: np.ndarray, y: np.ndarray) -> None:
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 6)
        y = np.ascontiguousarray(y, dtype=np.float32).reshape(-1, 3)
        if len(x) != len(y):
            raise ValueError("x and y differ in length")
        _check(load_library().emh_surrogate_set_training_set(self._h, x.ctypes.data, y.ctypes.data, len(x)))

    def training_set(self, first: int, n: int) -> Tuple[np.ndarray, np.ndarray]:
        x = np.zeros((n, 6), dtype=np.float32)
        y = np.zeros((n, 3), dtype=np.float32)
        _check(load_library().emh_surrogate_get_training_set(self._h, first, n, x.ctypes.data, y.ctypes.data))
        return x, y

    def train(self, epochs: int, batch: int = 4096, lr: float = 2e-3, lr_final: float = 1e-5, beta1: float = 0.9,
              beta2: float = 0.999, eps: float = 1e-8, seed: int = 1, max_steps: int = 0) -> np.ndarray:
        o = CTrainOptions(epochs, batch, lr, lr_final, beta1, beta2, eps, 0, seed, max_steps)
        loss = np.zeros(epochs, dtype=np.float32)
        _check(load_library().emh_surrogate_train(self._h, ctypes.byref(o), loss.ctypes.data, epochs))
        return loss

    def gradient(self, x: np.ndarray, y: np.ndarray) -> Tuple[np.ndarray, float]:
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 6)
        y = np.ascontiguousarray(y, dtype=np.float32).reshape(-1, 3)
        g = np.zeros(SURROGATE_PARAMS, dtype=np.float32)
        loss = ctypes.c_float(0.0)
        _check(load_library().emh_surrogate_gradient(self._h, x.ctypes.data, y.ctypes.data, len(x), g.ctypes.data,
                                                     ctypes.byref(loss)))
        return g, loss.value

    def infer(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 6)
        y = np.zeros((len(x), 3), dtype=np.float32)
        _check(load_library().emh_surrogate_infer(self._h, x.ctypes.data, len(x), y.ctypes.data))
        return y

    def counts(self, reset: bool = False) -> dict:
        c = CSurrogateCounts()
        _check(load_library().emh_surrogate_get_counts(self._h, ctypes.byref(c), int(reset)))
        return {"yielding": c.yielding, "surrogate": c.surrogate, "fallback": c.fallback}


def nacc_return_batch(material: "NaccMaterial", f_ptr: int, alpha_ptr: int, n: int, f_out_ptr: int,
                      alpha_out_ptr: int, surrogate: Optional[NaccSurrogate] = None, reps: int = 1) -> float:
    """NACC return mapping of n device-resident trial states (plain, or through
    the surrogate); returns the milliseconds per launch."""
    ms = ctypes.c_float(0.0)
    _check(load_library().emh_nacc_return_batch(surrogate._h if surrogate else None, ctypes.byref(material.params),
                                                ctypes.byref(material.ext), ctypes.c_void_p(f_ptr),
                                                ctypes.c_void_p(alpha_ptr), n, ctypes.c_void_p(f_out_ptr),
                                                ctypes.c_void_p(alpha_out_ptr), reps, ctypes.byref(ms)))
    return ms.value


# -------------------------------------------------------------- batched scenes
class SceneBatch(_Tunable):
    """Many independent simulations stepped together (one launch per stage for
    all of them); each keeps its own adaptive substeps."""
    _tune_prefix = "batch"

    def __init__(self, sims: Sequence["Simulation"]):
        self._sims = list(sims)  # keeps them alive while batched
        arr = (ctypes.c_void_p * len(self._sims))(*[sim._h for sim in self._sims])
        h = ctypes.c_void_p()
        _check(load_library().emh_batch_create(arr, len(self._sims), ctypes.byref(h)))
        self._h = h

    def step_n(self, steps: int) -> None:
        _check(load_library().emh_batch_step(self._h, int(steps)))

    def set_profiling(self, enable: bool) -> None:
        """Live per-stage timing of the batch steps (resets the totals)."""
        _check(load_library().emh_batch_set_profiling(self._h, int(bool(enable))))

    def kernel_times(self) -> dict:
        """{"kb_begin_step" | "kb_sort" | "kb_substep" | "kb_gather_ctrl": (launches, ms)}
        since profiling was enabled."""
        arr = (CKernelTime * 8)()
        n = ctypes.c_uint32()
        _check(load_library().emh_batch_kernel_times(self._h, arr, 8, ctypes.byref(n)))
        return {arr[i].name.decode(): (int(arr[i].launches), float(arr[i].total_ms)) for i in range(min(n.value, 8))}

    def close(self) -> None:
        if self._h:
            load_library().emh_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
