//! `hip` backend of `Simulation::step` (src/spacetime/solver/step.rs:23-60) for
//! MI355X, bound over the C-ABI of `include/emerge_hip.h`.
//!
//! Drop-in module for the emerge-engine crate (e.g. `src/systems/hip.rs` behind
//! a `hip` cargo feature).  Not compiled in this repository: the image that
//! builds it has no Rust toolchain.  Link with `libemerge_hip.so`:
//!
//! ```text
//! // build.rs
//! println!("cargo:rustc-link-search=native=/path/to/emerge-engine_amd");
//! println!("cargo:rustc-link-lib=dylib=emerge_hip");
//! ```
//!
//! `HipSimulation` mirrors the `Simulation` methods the substep path needs
//! (lifecycle.rs:26-365, particles.rs:355, step.rs:23-331).  Particles and
//! `MaterialParams` cross the boundary as their own `repr(C)` bytes
//! (particle.rs:11-144, params.rs:8-101); `SimConfig` is copied field by field
//! into the fixed-width `EmhConfig` (config.rs:26-121: `usize` -> u32,
//! `bool` -> u32).

use core::ffi::{c_char, c_int, c_void, CStr};

use glam::Vec2;

use crate::forces::boundary::friction::FrictionBoundary;
use crate::forces::boundary::heightmap::HeightmapBoundary;
use crate::forces::boundary::predictive::PredictiveBoundary;
use crate::forces::boundary::slip::SlipBoundary;
use crate::matter::materials::bingham::BinghamFluidMaterial;
use crate::matter::materials::elastic::NeoHookeanMaterial;
use crate::matter::materials::fluid::NewtonianFluidMaterial;
use crate::matter::materials::nacc::NaccMaterial;
use crate::matter::materials::sand::DruckerPragerMaterial;
use crate::matter::materials::{ConstitutiveModel, MaterialModel, MaterialParams};
use crate::matter::particle::Particle;
use crate::spacetime::solver::config::{SimConfig, SpawnRegion, SpawnShape};

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct EmhConfig {
    pub grid_res: u32,
    pub grid_cell_size: f32,
    pub dt: f32,
    pub adaptive_timestep: u32,
    pub cfl_include_affine_speed: u32,
    pub cfl_coefficient: f32,
    pub material_cfl_coefficient: f32,
    pub viscous_timestep_coefficient: f32,
    pub min_dt: f32,
    pub project_invalid_state: u32,
    pub projection_min_density: f32,
    pub projection_min_volume: f32,
    pub projection_min_deformation_j: f32,
    pub gravity: [f32; 2],
    pub boundary_thickness: u32,
    pub default_initial_volume: f32,
    pub recompute_density_each_step: u32,
    pub particle_mass: f32,
    pub max_substeps_per_step: u32,
    pub apic_blend: f32,
    pub j_max: f32,
    pub sleep_threshold: f32,
    pub contact_friction: f32,
    pub asflip_blend: f32,
    pub mixture_drag_coefficient: f32,
    pub mixture_pressure_iterations: u32,
    pub dx_meters: f32,
    pub dt_seconds: f32,
}
const _: () = assert!(core::mem::size_of::<EmhConfig>() == 116);

impl From<&SimConfig> for EmhConfig {
    fn from(c: &SimConfig) -> Self {
        Self {
            grid_res: c.grid_res as u32,
            grid_cell_size: c.grid_cell_size,
            dt: c.dt,
            adaptive_timestep: c.adaptive_timestep as u32,
            cfl_include_affine_speed: c.cfl_include_affine_speed as u32,
            cfl_coefficient: c.cfl_coefficient,
            material_cfl_coefficient: c.material_cfl_coefficient,
            viscous_timestep_coefficient: c.viscous_timestep_coefficient,
            min_dt: c.min_dt,
            project_invalid_state: c.project_invalid_state as u32,
            projection_min_density: c.projection_min_density,
            projection_min_volume: c.projection_min_volume,
            projection_min_deformation_j: c.projection_min_deformation_j,
            gravity: [c.gravity.x, c.gravity.y],
            boundary_thickness: c.boundary_thickness as u32,
            default_initial_volume: c.default_initial_volume,
            recompute_density_each_step: c.recompute_density_each_step as u32,
            particle_mass: c.particle_mass,
            max_substeps_per_step: c.max_substeps_per_step as u32,
            apic_blend: c.apic_blend,
            j_max: c.j_max,
            sleep_threshold: c.sleep_threshold,
            contact_friction: c.contact_friction,
            asflip_blend: c.asflip_blend,
            mixture_drag_coefficient: c.mixture_drag_coefficient,
            mixture_pressure_iterations: c.mixture_pressure_iterations as u32,
            dx_meters: c.dx_meters,
            dt_seconds: c.dt_seconds,
        }
    }
}

/// Fields `MaterialModel::params()` drops (emh_material_ext).
#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct EmhMaterialExt {
    pub min_density: f32,
    pub min_volume: f32,
    pub dp_volume_correction: f32,
    pub nacc_kappa: f32,
    pub nacc_hardening_enabled: u32,
    pub nacc_advance_f: u32,
    pub bingham_critical_shear_rate: f32,
    pub fluid_bingham: u32,
}
impl Default for EmhMaterialExt {
    fn default() -> Self {
        Self { min_density: 1e-6, min_volume: 1e-6, dp_volume_correction: 1.0, nacc_kappa: 0.0,
               nacc_hardening_enabled: 2, nacc_advance_f: 0,
               bingham_critical_shear_rate: 1e-4, fluid_bingham: 0 }
    }
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct EmhSpawnRegion {
    pub spacing: f32,
    pub box_size: [i32; 2],
    pub box_center: [f32; 2],
    pub shape: u32,
    pub disk_radius: f32,
    pub initial_deformation_gradient: [f32; 4],
    pub precompute_initial_volumes: u32,
    pub initial_velocity_scale: f32,
    pub position_jitter: f32,
    pub rng_seed: u32,
    pub material_id: u32,
    pub has_mass_override: u32,
    pub mass_override: f32,
}

impl From<&SpawnRegion> for EmhSpawnRegion {
    fn from(s: &SpawnRegion) -> Self {
        let (shape, disk_radius) = match s.shape {
            SpawnShape::Box => (0, 0.0),
            SpawnShape::Disk { radius } => (1, radius),
        };
        Self {
            spacing: s.spacing,
            box_size: [s.box_size.x, s.box_size.y],
            box_center: [s.box_center.x, s.box_center.y],
            shape,
            disk_radius,
            initial_deformation_gradient: s.initial_deformation_gradient.to_cols_array(),
            precompute_initial_volumes: s.precompute_initial_volumes as u32,
            initial_velocity_scale: s.initial_velocity_scale,
            position_jitter: s.position_jitter,
            rng_seed: s.rng_seed,
            material_id: s.material_id,
            has_mass_override: s.mass_override.is_some() as u32,
            mass_override: s.mass_override.unwrap_or(0.0),
        }
    }
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct EmhStepStats {
    pub last_substeps: u32,
    pub last_step_dt: f32,
    pub last_sim_time_dropped: f32,
    pub last_vel_clamp_count: u32,
    pub last_j_projection_count: u32,
    pub total_substeps: u32,
    pub frame_index: u64,
}

#[repr(C)]
pub struct EmhSim {
    _opaque: [u8; 0],
}
#[repr(C)]
pub struct EmhBatch {
    _opaque: [u8; 0],
}
#[repr(C)]
pub struct EmhSurrogate {
    _opaque: [u8; 0],
}

/// `emh_snapshot`: `SimSnapshot` (systems/diagnostics/snapshot.rs:8-80) over the device state.
#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct EmhSnapshot {
    pub frame_index: u64,
    pub configured_dt: f32,
    pub effective_dt: f32,
    pub substeps_last_step: u32,
    pub particle_count: u32,
    pub valid_particle_count: u32,
    pub active_grid_cells: u32,
    pub particles_per_active_cell: f32,
    pub mixed_material_cell_ratio: f32,
    pub mixed_material_particle_ratio: f32,
    pub total_particle_mass: f32,
    pub total_grid_mass: f32,
    pub relative_mass_error: f32,
    pub total_particle_momentum: [f32; 2],
    pub total_grid_momentum: [f32; 2],
    pub relative_momentum_error: f32,
    pub max_particle_speed: f32,
    pub max_grid_speed: f32,
    pub cfl_number: f32,
    pub min_deformation_j: f32,
    pub max_deformation_j: f32,
    pub out_of_bounds_particles: u32,
    pub invalid_physical_particle_values: u32,
    pub non_finite_particle_values: u32,
    pub non_finite_grid_values: u32,
    pub recommended_max_dt_from_velocity_cfl: f32,
    pub avg_plastic_jacobian: f32,
    pub min_plastic_jacobian: f32,
    pub avg_elastic_hardening: f32,
    pub active_count: u32,
    pub sleeping_count: u32,
    pub vel_clamp_count: u32,
    pub j_projection_count: u32,
    pub sim_time_dropped: f32,
    pub total_kinetic_energy: f32,
    pub max_pinned_particle_speed: f32,
    pub _pad: u32,
}

/// `emh_kernel_time`.
#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct EmhKernelTime {
    pub name: [c_char; 32],
    pub launches: u64,
    pub total_ms: f64,
}
impl Default for EmhKernelTime {
    fn default() -> Self {
        Self { name: [0; 32], launches: 0, total_ms: 0.0 }
    }
}

/// `emh_tuning`: launch settings and the autotuner's measurement.
#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct EmhTuning {
    pub reorder_interval: u32,
    pub graphs: u32,
    pub ns_per_particle_substep: f32,
    pub model_segments: u32,
}

/// The C-ABI of `include/emerge_hip.h`, every entry point in declaration order.
extern "C" {
    fn emh_abi_version() -> u32;
    fn emh_last_error() -> *const c_char;
    fn emh_device_count(count: *mut c_int) -> c_int;
    fn emh_config_default(out: *mut EmhConfig) -> c_int;
    fn emh_config_standard(grid_res: u32, dt: f32, gx: f32, gy: f32, out: *mut EmhConfig) -> c_int;
    fn emh_config_unsafe_defaults(out: *mut EmhConfig) -> c_int;
    fn emh_config_validate(cfg: *const EmhConfig) -> c_int;
    fn emh_lame_from_young(young_modulus: f32, poisson_ratio: f32, lambda: *mut f32, mu: *mut f32) -> c_int;
    fn emh_material_neohookean(lambda: f32, mu: f32, p: *mut MaterialParams, e: *mut EmhMaterialExt) -> c_int;
    fn emh_material_corotated(lambda: f32, mu: f32, p: *mut MaterialParams, e: *mut EmhMaterialExt) -> c_int;
    fn emh_material_fluid(rest_density: f32, dynamic_viscosity: f32, eos_stiffness: f32, eos_power: f32,
                          p: *mut MaterialParams, e: *mut EmhMaterialExt) -> c_int;
    fn emh_material_fluid_low_viscosity(rest_density: f32, eos_stiffness: f32, p: *mut MaterialParams,
                                        e: *mut EmhMaterialExt) -> c_int;
    fn emh_material_drucker_prager(lambda: f32, mu: f32, p: *mut MaterialParams, e: *mut EmhMaterialExt) -> c_int;
    fn emh_material_snow(lambda: f32, mu: f32, hardening_exponent: f32, compression_limit: f32, stretch_limit: f32,
                         min_plastic_jacobian: f32, max_plastic_jacobian: f32, p: *mut MaterialParams,
                         e: *mut EmhMaterialExt) -> c_int;
    fn emh_material_snow_from_young(young_modulus: f32, poisson_ratio: f32, p: *mut MaterialParams,
                                    e: *mut EmhMaterialExt) -> c_int;
    fn emh_material_von_mises(lambda: f32, mu: f32, yield_stress: f32, hardening_modulus: f32,
                              p: *mut MaterialParams, e: *mut EmhMaterialExt) -> c_int;
    fn emh_material_rankine(lambda: f32, mu: f32, tensile_strength: f32, softening_rate: f32,
                            p: *mut MaterialParams, e: *mut EmhMaterialExt) -> c_int;
    fn emh_material_mu_i(lambda: f32, mu: f32, mu_static: f32, mu_dynamic: f32, inertial_q: f32,
                         p: *mut MaterialParams, e: *mut EmhMaterialExt) -> c_int;
    fn emh_material_mu_i_default(lambda: f32, mu: f32, p: *mut MaterialParams, e: *mut EmhMaterialExt) -> c_int;
    fn emh_material_granular_fluid(lambda: f32, mu: f32, rest_density: f32, eos_stiffness: f32,
                                   hardening_exponent: f32, compression_limit: f32, p: *mut MaterialParams,
                                   e: *mut EmhMaterialExt) -> c_int;
    fn emh_material_bingham(rest_density: f32, dynamic_viscosity: f32, eos_stiffness: f32, eos_power: f32,
                            yield_stress: f32, p: *mut MaterialParams, e: *mut EmhMaterialExt) -> c_int;
    fn emh_material_viscoelastic(lambda: f32, mu: f32, viscosity: f32, p: *mut MaterialParams,
                                 e: *mut EmhMaterialExt) -> c_int;
    fn emh_material_nacc(mu: f32, kappa: f32, friction: f32, cohesion: f32, hardening_factor: f32,
                         p: *mut MaterialParams, e: *mut EmhMaterialExt) -> c_int;
    fn emh_material_nacc_from_young(young_modulus: f32, poisson_ratio: f32, friction: f32, cohesion: f32,
                                    hardening_factor: f32, p: *mut MaterialParams, e: *mut EmhMaterialExt)
        -> c_int;
    fn emh_to_radians(degrees: f32) -> f32;
    fn emh_sim_create(cfg: *const EmhConfig, out: *mut *mut EmhSim) -> c_int;
    fn emh_sim_create_from_spawn(cfg: *const EmhConfig, spawn: *const EmhSpawnRegion, out: *mut *mut EmhSim)
        -> c_int;
    fn emh_sim_destroy(sim: *mut EmhSim);
    fn emh_sim_set_material(sim: *mut EmhSim, material_id: u32, params: *const MaterialParams,
                            ext: *const EmhMaterialExt) -> c_int;
    fn emh_sim_reinit_particle_state(sim: *mut EmhSim) -> c_int;
    fn emh_sim_add_slip_boundary(sim: *mut EmhSim, thickness: u32) -> c_int;
    fn emh_sim_add_friction_boundary(sim: *mut EmhSim, thickness: u32, friction_coefficient: f32) -> c_int;
    fn emh_sim_add_heightmap_boundary(sim: *mut EmhSim, heights: *const f32, n_heights: u32, friction: f32,
                                      wall_thickness: u32) -> c_int;
    fn emh_sim_add_predictive_boundary(sim: *mut EmhSim, thickness: u32, predictive_wall_min: f32) -> c_int;
    fn emh_sim_clear_boundaries(sim: *mut EmhSim) -> c_int;
    fn emh_sim_add_body(sim: *mut EmhSim, spawn: *const EmhSpawnRegion, tag_out: *mut u32) -> c_int;
    fn emh_sim_set_particles(sim: *mut EmhSim, particles: *const Particle, n: usize) -> c_int;
    fn emh_sim_num_particles(sim: *const EmhSim, n: *mut usize) -> c_int;
    fn emh_sim_get_particles(sim: *mut EmhSim, out: *mut Particle, n: usize) -> c_int;
    fn emh_sim_get_config(sim: *const EmhSim, out: *mut EmhConfig) -> c_int;
    fn emh_sim_step(sim: *mut EmhSim, n_steps: u32) -> c_int;
    fn emh_sim_stats(sim: *const EmhSim, out: *mut EmhStepStats) -> c_int;
    fn emh_sim_get_grid(sim: *mut EmhSim, mass: *mut f32, momentum: *mut f32, n_cells: usize) -> c_int;
    fn emh_sim_snapshot(sim: *mut EmhSim, out: *mut EmhSnapshot) -> c_int;
    fn emh_sim_synchronize(sim: *mut EmhSim) -> c_int;
    fn emh_sim_export_particles_device(sim: *mut EmhSim, dst: *mut c_void, capacity: usize) -> c_int;
    fn emh_sim_export_grid_device(sim: *mut EmhSim, dst: *mut c_void, capacity_cells: usize) -> c_int;
    fn emh_sim_sleep_tag(sim: *mut EmhSim, tag: u32) -> c_int;
    fn emh_sim_wake_tag(sim: *mut EmhSim, tag: u32) -> c_int;
    fn emh_sim_set_gravity(sim: *mut EmhSim, gx: f32, gy: f32) -> c_int;
    fn emh_sim_apply_impulse(sim: *mut EmhSim, cx: f32, cy: f32, radius: f32, fx: f32, fy: f32) -> c_int;
    fn emh_sim_apply_radial_impulse(sim: *mut EmhSim, cx: f32, cy: f32, radius: f32, strength: f32) -> c_int;
    fn emh_sim_set_profiling(sim: *mut EmhSim, enable: c_int) -> c_int;
    fn emh_sim_set_reorder_interval(sim: *mut EmhSim, substeps: u32) -> c_int;
    fn emh_sim_get_slot_order(sim: *mut EmhSim, ids_out: *mut u32, n: usize) -> c_int;
    fn emh_sim_kernel_times(sim: *mut EmhSim, out: *mut EmhKernelTime, cap: u32, n_out: *mut u32) -> c_int;
    fn emh_sim_get_tuning(sim: *mut EmhSim, out: *mut EmhTuning) -> c_int;
    fn emh_sim_set_tuning(sim: *mut EmhSim, t: *const EmhTuning) -> c_int;
    fn emh_sim_autotune(sim: *mut EmhSim, trial_steps: u32, chosen: *mut EmhTuning, timed: *mut EmhTuning,
                        timed_cap: u32, n_timed: *mut u32) -> c_int;
    fn emh_cell_key(cx: i32, cy: i32, grid_res: u32) -> u32;
    fn emh_sort_cell_keys(xy: *const f32, n: usize, grid_res: u32, keys_out: *mut u32, perm_out: *mut u32) -> c_int;
    fn emh_sort_cell_keys_noalloc(xy_dev: *const f32, n: usize, grid_res: u32, keys_out_dev: *mut u32,
                                  perm_out_dev: *mut u32, stream: *mut c_void) -> c_int;
    fn emh_batch_create(sims: *const *mut EmhSim, n: u32, out: *mut *mut EmhBatch) -> c_int;
    fn emh_batch_destroy(b: *mut EmhBatch) -> c_int;
    fn emh_batch_step(b: *mut EmhBatch, n_steps: u32) -> c_int;
    fn emh_batch_get_tuning(b: *mut EmhBatch, out: *mut EmhTuning) -> c_int;
    fn emh_batch_set_tuning(b: *mut EmhBatch, t: *const EmhTuning) -> c_int;
    fn emh_batch_autotune(b: *mut EmhBatch, trial_steps: u32, chosen: *mut EmhTuning, timed: *mut EmhTuning,
                          timed_cap: u32, n_timed: *mut u32) -> c_int;
    fn emh_batch_set_profiling(b: *mut EmhBatch, enable: c_int) -> c_int;
    fn emh_batch_kernel_times(b: *mut EmhBatch, out: *mut EmhKernelTime, cap: u32, n_out: *mut u32) -> c_int;
    fn emh_surrogate_create(seed: u64, out: *mut *mut EmhSurrogate) -> c_int;
    fn emh_surrogate_destroy(sg: *mut EmhSurrogate) -> c_int;
    fn emh_surrogate_get_params(sg: *mut EmhSurrogate, params: *mut f32, norm: *mut f32) -> c_int;
    fn emh_surrogate_set_params(sg: *mut EmhSurrogate, params: *const f32, norm: *const f32) -> c_int;
    fn emh_sim_set_surrogate_mode(sim: *mut EmhSim, sg: *mut EmhSurrogate, mode: u32, harvest_capacity: u64,
                                  harvest_stride: u32) -> c_int;
    fn emh_surrogate_harvest_count(sg: *mut EmhSurrogate, offered: *mut u64, stored: *mut u64) -> c_int;
    fn emh_surrogate_get_records(sg: *mut EmhSurrogate, first: u64, n: u64, out: *mut f32) -> c_int;
    fn emh_surrogate_eval_records(sg: *mut EmhSurrogate, p: *const MaterialParams, ext: *const EmhMaterialExt,
                                  first: u64, n: u64, use_surrogate: c_int, reps: u32, ms_per_rep: *mut f32,
                                  f_out: *mut f32, alpha_out: *mut f32) -> c_int;
    fn emh_surrogate_build_training_set(sg: *mut EmhSurrogate, n_samples: *mut u64) -> c_int;
    fn emh_surrogate_set_training_set(sg: *mut EmhSurrogate, x: *const f32, y: *const f32, n: u64) -> c_int;
    fn emh_surrogate_get_training_set(sg: *mut EmhSurrogate, first: u64, n: u64, x: *mut f32, y: *mut f32)
        -> c_int;
    fn emh_surrogate_train(sg: *mut EmhSurrogate, o: *const EmhTrainOptions, epoch_loss: *mut f32,
                           n_epoch_loss: u32) -> c_int;
    fn emh_surrogate_gradient(sg: *mut EmhSurrogate, x: *const f32, y: *const f32, n: u32, grad: *mut f32,
                              loss: *mut f32) -> c_int;
    fn emh_surrogate_infer(sg: *mut EmhSurrogate, x: *const f32, n: u64, y: *mut f32) -> c_int;
    fn emh_surrogate_get_counts(sg: *mut EmhSurrogate, out: *mut EmhSurrogateCounts, reset: c_int) -> c_int;
    fn emh_nacc_return_batch(sg: *mut EmhSurrogate, p: *const MaterialParams, ext: *const EmhMaterialExt,
                             f_trial: *const f32, alpha: *const f32, n: u64, f_out: *mut f32,
                             alpha_out: *mut f32, reps: u32, ms_per_rep: *mut f32) -> c_int;
}

/// `emh_surrogate_train_options`.
#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct EmhTrainOptions {
    pub epochs: u32,
    pub batch: u32,
    pub lr: f32,
    pub lr_final: f32,
    pub beta1: f32,
    pub beta2: f32,
    pub eps: f32,
    pub pad: u32,
    pub seed: u64,
    pub max_steps: u64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct EmhSurrogateCounts {
    pub yielding: u64,
    pub surrogate: u64,
    pub fallback: u64,
}

/// A failed C-ABI call: the status code and the library's message.
#[derive(Debug, Clone)]
pub struct HipError {
    pub code: i32,
    pub message: String,
}

fn check(rc: c_int) -> Result<(), HipError> {
    if rc == 0 {
        return Ok(());
    }
    let message = unsafe { CStr::from_ptr(emh_last_error()) }.to_string_lossy().into_owned();
    Err(HipError { code: rc, message })
}

/// The parameters this backend needs for one material: `params()` as is, plus
/// the fields it drops at their constructors' defaults.  A material whose
/// dropped fields were changed after construction -- NeoHookean or Newtonian
/// `min_density` / `min_volume`, Drucker-Prager `volume_correction`, Bingham
/// `critical_shear_rate` -- goes through its typed record below instead, and
/// NACC always does (its `params()` uploads a NeoHookean stand-in,
/// nacc.rs:300-314).
pub fn material_record(m: &dyn MaterialModel) -> (MaterialParams, EmhMaterialExt) {
    let mut p = m.params();
    p.model = m.constitutive_model() as u32;
    (p, EmhMaterialExt::default())
}
/// NeoHookeanMaterial (elastic.rs:20-60): `min_density` rides in the extension.
pub fn neohookean_record(m: &NeoHookeanMaterial) -> (MaterialParams, EmhMaterialExt) {
    let (p, mut e) = material_record(m);
    e.min_density = m.min_density;
    (p, e)
}
/// NewtonianFluidMaterial (fluid.rs:10-62): `min_density` and `min_volume`.
pub fn fluid_record(m: &NewtonianFluidMaterial) -> (MaterialParams, EmhMaterialExt) {
    let (p, mut e) = material_record(m);
    e.min_density = m.min_density;
    e.min_volume = m.min_volume;
    (p, e)
}
/// DruckerPragerMaterial (sand.rs:11-90): `volume_correction`.
pub fn drucker_prager_record(m: &DruckerPragerMaterial) -> (MaterialParams, EmhMaterialExt) {
    let (p, mut e) = material_record(m);
    e.dp_volume_correction = m.volume_correction;
    (p, e)
}
/// BinghamFluidMaterial (bingham.rs:26-70): a Fluid record marked as Bingham,
/// with its `critical_shear_rate`, `min_density` and `min_volume`.
pub fn bingham_record(m: &BinghamFluidMaterial) -> (MaterialParams, EmhMaterialExt) {
    let (p, mut e) = material_record(m);
    e.fluid_bingham = 1;
    e.bingham_critical_shear_rate = m.critical_shear_rate;
    e.min_density = m.min_density;
    e.min_volume = m.min_volume;
    (p, e)
}
pub fn nacc_record(m: &NaccMaterial, advance_f: bool) -> (MaterialParams, EmhMaterialExt) {
    let mut p = m.params();
    p.model = ConstitutiveModel::Nacc as u32;
    p.lambda = m.kappa - m.mu;
    let ext = EmhMaterialExt {
        nacc_kappa: m.kappa,
        nacc_advance_f: advance_f as u32,
        nacc_hardening_enabled: m.hardening_enabled as u32,
        min_density: m.min_density,
        ..Default::default()
    };
    (p, ext)
}

/// `Simulation` with its substep on the GPU.  One host thread per handle.
pub struct HipSimulation {
    raw: *mut EmhSim,
    grid_res: usize,
}

impl Drop for HipSimulation {
    fn drop(&mut self) {
        unsafe { emh_sim_destroy(self.raw) }
    }
}

impl HipSimulation {
    /// `Simulation::empty` (lifecycle.rs:26-55).
    pub fn empty(config: SimConfig) -> Result<Self, HipError> {
        let c = EmhConfig::from(&config);
        let mut raw = core::ptr::null_mut();
        check(unsafe { emh_sim_create(&c, &mut raw) })?;
        Ok(Self { raw, grid_res: config.grid_res })
    }
    /// `Simulation::new` (lifecycle.rs:57-105).
    pub fn new(config: SimConfig, spawn: SpawnRegion) -> Result<Self, HipError> {
        let c = EmhConfig::from(&config);
        let s = EmhSpawnRegion::from(&spawn);
        let mut raw = core::ptr::null_mut();
        check(unsafe { emh_sim_create_from_spawn(&c, &s, &mut raw) })?;
        Ok(Self { raw, grid_res: config.grid_res })
    }
    /// `set_material` / `with_material` (lifecycle.rs:197, 348); id 0 is
    /// `with_default_material` (lifecycle.rs:107).
    pub fn set_material(&mut self, id: u32, record: (MaterialParams, EmhMaterialExt)) -> Result<(), HipError> {
        check(unsafe { emh_sim_set_material(self.raw, id, &record.0, &record.1) })
    }
    pub fn with_material(mut self, id: u32, material: &dyn MaterialModel) -> Result<Self, HipError> {
        self.set_material(id, material_record(material))?;
        Ok(self)
    }
    /// `add_boundary_condition(Box::new(SlipBoundary))` (lifecycle.rs:359).
    pub fn with_boundary(self, b: SlipBoundary) -> Result<Self, HipError> {
        check(unsafe { emh_sim_add_slip_boundary(self.raw, b.thickness as u32) })?;
        Ok(self)
    }
    pub fn clear_boundaries(&mut self) -> Result<(), HipError> {
        check(unsafe { emh_sim_clear_boundaries(self.raw) })
    }
    pub fn reinit_all_particle_state(&mut self) -> Result<(), HipError> {
        check(unsafe { emh_sim_reinit_particle_state(self.raw) })
    }
    /// `add_body` (particles.rs:355-424); returns the body's tag.
    pub fn add_body(&mut self, spawn: SpawnRegion) -> Result<u32, HipError> {
        let s = EmhSpawnRegion::from(&spawn);
        let mut tag = 0u32;
        check(unsafe { emh_sim_add_body(self.raw, &s, &mut tag) })?;
        Ok(tag)
    }
    /// `step` (step.rs:23-60).
    pub fn step(&mut self) -> Result<(), HipError> {
        check(unsafe { emh_sim_step(self.raw, 1) })
    }
    /// `step_n` (step.rs:327-331), queued on the device and synchronised once.
    pub fn step_n(&mut self, steps: usize) -> Result<(), HipError> {
        check(unsafe { emh_sim_step(self.raw, steps as u32) })
    }

    /// Physical cell-key reorder interval in substeps (no counterpart in the
    /// reference; the upload is always reordered at the first step).
    pub fn set_reorder_interval(&mut self, substeps: u32) -> Result<(), HipError> {
        check(unsafe { emh_sim_set_reorder_interval(self.raw, substeps) })
    }
    pub fn set_particles(&mut self, particles: &[Particle]) -> Result<(), HipError> {
        check(unsafe { emh_sim_set_particles(self.raw, particles.as_ptr(), particles.len()) })
    }
    /// `collect_particles` (particles.rs:331): the particles in their original order.
    pub fn collect_particles(&mut self) -> Result<Vec<Particle>, HipError> {
        let mut n = 0usize;
        check(unsafe { emh_sim_num_particles(self.raw, &mut n) })?;
        let mut out = vec![Particle::default(); n];
        check(unsafe { emh_sim_get_particles(self.raw, out.as_mut_ptr(), n) })?;
        Ok(out)
    }
    fn stats(&self) -> EmhStepStats {
        let mut s = EmhStepStats::default();
        let _ = check(unsafe { emh_sim_stats(self.raw, &mut s) });
        s
    }
    /// `effective_dt` (step.rs:319).
    pub fn effective_dt(&self) -> f32 {
        self.stats().last_step_dt
    }
    /// `last_substeps` (step.rs:323).
    pub fn last_substeps(&self) -> usize {
        self.stats().last_substeps as usize
    }
    /// Mass and momentum of the last substep's P2G, flat index x * R + y.
    pub fn grid_mass_momentum(&mut self) -> Result<(Vec<f32>, Vec<[f32; 2]>), HipError> {
        let n = self.grid_res * self.grid_res;
        let mut m = vec![0f32; n];
        let mut p = vec![[0f32; 2]; n];
        check(unsafe { emh_sim_get_grid(self.raw, m.as_mut_ptr(), p.as_mut_ptr() as *mut f32, n) })?;
        Ok((m, p))
    }

    /// `add_boundary_condition(Box::new(FrictionBoundary))` (friction.rs:20-59).
    pub fn with_friction_boundary(self, b: FrictionBoundary) -> Result<Self, HipError> {
        check(unsafe { emh_sim_add_friction_boundary(self.raw, b.thickness as u32, b.friction_coefficient) })?;
        Ok(self)
    }
    /// `add_boundary_condition(Box::new(HeightmapBoundary))` (heightmap.rs:29-114).
    pub fn with_heightmap_boundary(self, b: HeightmapBoundary) -> Result<Self, HipError> {
        check(unsafe {
            emh_sim_add_heightmap_boundary(self.raw, b.heights.as_ptr(), b.heights.len() as u32, b.friction,
                                           b.wall_thickness as u32)
        })?;
        Ok(self)
    }
    /// `add_boundary_condition(Box::new(PredictiveBoundary))` (predictive.rs:12-35).
    pub fn with_predictive_boundary(self, b: PredictiveBoundary) -> Result<Self, HipError> {
        check(unsafe { emh_sim_add_predictive_boundary(self.raw, b.thickness as u32, b.predictive_wall_min) })?;
        Ok(self)
    }
    /// `set_gravity` (lifecycle.rs:403-405), from the next step.
    pub fn set_gravity(&mut self, gravity: Vec2) -> Result<(), HipError> {
        check(unsafe { emh_sim_set_gravity(self.raw, gravity.x, gravity.y) })
    }
    /// `apply_impulse` (particles.rs:115-139), on the device.
    pub fn apply_impulse(&mut self, center: Vec2, radius: f32, force: Vec2) -> Result<(), HipError> {
        check(unsafe { emh_sim_apply_impulse(self.raw, center.x, center.y, radius, force.x, force.y) })
    }
    /// `apply_radial_impulse` (particles.rs:141-165), on the device.
    pub fn apply_radial_impulse(&mut self, center: Vec2, radius: f32, strength: f32) -> Result<(), HipError> {
        check(unsafe { emh_sim_apply_radial_impulse(self.raw, center.x, center.y, radius, strength) })
    }
    /// `sleep_tag` / `wake_tag` (particles.rs:270-312).
    pub fn sleep_tag(&mut self, tag: u32) -> Result<(), HipError> {
        check(unsafe { emh_sim_sleep_tag(self.raw, tag) })
    }
    pub fn wake_tag(&mut self, tag: u32) -> Result<(), HipError> {
        check(unsafe { emh_sim_wake_tag(self.raw, tag) })
    }
    /// `diagnostics_snapshot` (queries.rs:15-30) computed on the device.
    pub fn diagnostics_snapshot(&mut self) -> Result<EmhSnapshot, HipError> {
        let mut out = EmhSnapshot::default();
        check(unsafe { emh_sim_snapshot(self.raw, &mut out) })?;
        Ok(out)
    }
    /// `GpuSimulation::particle_buffer` (gpu/solver/mod.rs:466-471): every particle as
    /// its 128 B record, in caller order, into device memory the caller owns.
    ///
    /// # Safety
    /// `dst` must be device memory of this GPU holding `capacity` records.
    pub unsafe fn export_particles_device(&mut self, dst: *mut c_void, capacity: usize) -> Result<(), HipError> {
        check(emh_sim_export_particles_device(self.raw, dst, capacity))
    }
    /// `GpuSimulation::grid_buffer` (gpu/solver/mod.rs:510-515): R^2 cells of 16 B.
    ///
    /// # Safety
    /// `dst` must be device memory of this GPU holding `capacity_cells` cells.
    pub unsafe fn export_grid_device(&mut self, dst: *mut c_void, capacity_cells: usize) -> Result<(), HipError> {
        check(emh_sim_export_grid_device(self.raw, dst, capacity_cells))
    }
    /// Waits for the device work this simulation queued.
    pub fn synchronize(&mut self) -> Result<(), HipError> {
        check(unsafe { emh_sim_synchronize(self.raw) })
    }
    /// `GpuSimulation::enable_profiling` / `last_pass_timings_ns` (gpu/solver/profiling.rs:9-75).
    pub fn set_profiling(&mut self, enable: bool) -> Result<(), HipError> {
        check(unsafe { emh_sim_set_profiling(self.raw, enable as c_int) })
    }
    pub fn kernel_times(&mut self) -> Result<Vec<EmhKernelTime>, HipError> {
        let mut out = [EmhKernelTime::default(); 8];
        let mut n = 0u32;
        check(unsafe { emh_sim_kernel_times(self.raw, out.as_mut_ptr(), out.len() as u32, &mut n) })?;
        Ok(out[..(n as usize).min(out.len())].to_vec())
    }
    /// The caller index of the particle in each device slot (no counterpart in the reference).
    pub fn slot_order(&mut self) -> Result<Vec<u32>, HipError> {
        let mut n = 0usize;
        check(unsafe { emh_sim_num_particles(self.raw, &mut n) })?;
        let mut ids = vec![0u32; n];
        check(unsafe { emh_sim_get_slot_order(self.raw, ids.as_mut_ptr(), n) })?;
        Ok(ids)
    }
    /// Launch settings (no counterpart in the reference).
    pub fn tuning(&mut self) -> Result<EmhTuning, HipError> {
        let mut t = EmhTuning::default();
        check(unsafe { emh_sim_get_tuning(self.raw, &mut t) })?;
        Ok(t)
    }
    pub fn set_tuning(&mut self, t: EmhTuning) -> Result<(), HipError> {
        check(unsafe { emh_sim_set_tuning(self.raw, &t) })
    }
    /// Times every candidate setting on this scene (state put back), applies the fastest.
    pub fn autotune(&mut self, trial_steps: u32) -> Result<(EmhTuning, Vec<EmhTuning>), HipError> {
        let mut best = EmhTuning::default();
        let mut timed = [EmhTuning::default(); 64];
        let mut n = 0u32;
        check(unsafe { emh_sim_autotune(self.raw, trial_steps, &mut best, timed.as_mut_ptr(), 64, &mut n) })?;
        Ok((best, timed[..(n as usize).min(64)].to_vec()))
    }
}

/// Many independent simulations stepped together, one launch per stage (no
/// counterpart in the reference).  The batch borrows its scenes mutably for its
/// lifetime: they are stepped through it.
pub struct HipBatch<'a> {
    raw: *mut EmhBatch,
    _scenes: core::marker::PhantomData<&'a mut HipSimulation>,
}

impl Drop for HipBatch<'_> {
    fn drop(&mut self) {
        unsafe {
            emh_batch_destroy(self.raw);
        }
    }
}

impl<'a> HipBatch<'a> {
    pub fn new(scenes: &'a mut [HipSimulation]) -> Result<Self, HipError> {
        let raws: Vec<*mut EmhSim> = scenes.iter().map(|s| s.raw).collect();
        let mut raw = core::ptr::null_mut();
        check(unsafe { emh_batch_create(raws.as_ptr(), raws.len() as u32, &mut raw) })?;
        Ok(Self { raw, _scenes: core::marker::PhantomData })
    }
    /// `step_n` of every scene, each with its own adaptive substeps.
    pub fn step_n(&mut self, steps: u32) -> Result<(), HipError> {
        check(unsafe { emh_batch_step(self.raw, steps) })
    }
    pub fn tuning(&mut self) -> Result<EmhTuning, HipError> {
        let mut t = EmhTuning::default();
        check(unsafe { emh_batch_get_tuning(self.raw, &mut t) })?;
        Ok(t)
    }
    pub fn set_tuning(&mut self, t: EmhTuning) -> Result<(), HipError> {
        check(unsafe { emh_batch_set_tuning(self.raw, &t) })
    }
    pub fn autotune(&mut self, trial_steps: u32) -> Result<(EmhTuning, Vec<EmhTuning>), HipError> {
        let mut best = EmhTuning::default();
        let mut timed = [EmhTuning::default(); 64];
        let mut n = 0u32;
        check(unsafe { emh_batch_autotune(self.raw, trial_steps, &mut best, timed.as_mut_ptr(), 64, &mut n) })?;
        Ok((best, timed[..(n as usize).min(64)].to_vec()))
    }
    /// Live per-stage timing of the batch steps (resets the totals).
    pub fn set_profiling(&mut self, enable: bool) -> Result<(), HipError> {
        check(unsafe { emh_batch_set_profiling(self.raw, enable as c_int) })
    }
    pub fn kernel_times(&mut self) -> Result<Vec<EmhKernelTime>, HipError> {
        let mut out = [EmhKernelTime::default(); 8];
        let mut n = 0u32;
        check(unsafe { emh_batch_kernel_times(self.raw, out.as_mut_ptr(), out.len() as u32, &mut n) })?;
        Ok(out[..(n as usize).min(out.len())].to_vec())
    }
}

/// The cell-key sort on its own: (key of each input, stable permutation).
pub fn sort_cell_keys(xy: &[[f32; 2]], grid_res: u32) -> Result<(Vec<u32>, Vec<u32>), HipError> {
    let mut keys = vec![0u32; xy.len()];
    let mut perm = vec![0u32; xy.len()];
    check(unsafe {
        emh_sort_cell_keys(xy.as_ptr() as *const f32, xy.len(), grid_res, keys.as_mut_ptr(), perm.as_mut_ptr())
    })?;
    Ok((keys, perm))
}

/// Learned NACC return mapping on the GPU (no counterpart in the crate): harvest
/// trial states from a plain run, train, then let a NACC-only simulation use it.
pub struct NaccSurrogate {
    raw: *mut EmhSurrogate,
}

impl Drop for NaccSurrogate {
    fn drop(&mut self) {
        unsafe {
            emh_surrogate_destroy(self.raw);
        }
    }
}

impl NaccSurrogate {
    pub fn new(seed: u64) -> Result<Self, HipError> {
        let mut raw = std::ptr::null_mut();
        check(unsafe { emh_surrogate_create(seed, &mut raw) })?;
        Ok(NaccSurrogate { raw })
    }
    /// Plain return mapping while recording every `stride`-th trial state (at most `capacity`).
    pub fn harvest_from(&mut self, sim: &mut HipSimulation, capacity: u64, stride: u32) -> Result<(), HipError> {
        check(unsafe { emh_sim_set_surrogate_mode(sim.raw, self.raw, 1, capacity, stride) })
    }
    /// Returns the number of training samples (Case C states).
    pub fn build_training_set(&mut self) -> Result<u64, HipError> {
        let mut n = 0u64;
        check(unsafe { emh_surrogate_build_training_set(self.raw, &mut n) })?;
        Ok(n)
    }
    pub fn train(&mut self, o: &EmhTrainOptions) -> Result<Vec<f32>, HipError> {
        let mut loss = vec![0f32; o.epochs as usize];
        check(unsafe { emh_surrogate_train(self.raw, o, loss.as_mut_ptr(), o.epochs) })?;
        Ok(loss)
    }
    /// The simulation's NACC return mapping goes through this network from the next step.
    /// The simulation keeps a pointer: call `detach` (or drop the simulation) before dropping `self`.
    pub fn attach(&self, sim: &mut HipSimulation) -> Result<(), HipError> {
        check(unsafe { emh_sim_set_surrogate_mode(sim.raw, self.raw, 2, 0, 1) })
    }
    pub fn detach(sim: &mut HipSimulation) -> Result<(), HipError> {
        check(unsafe { emh_sim_set_surrogate_mode(sim.raw, std::ptr::null_mut(), 0, 0, 1) })
    }
    pub fn counts(&self, reset: bool) -> Result<EmhSurrogateCounts, HipError> {
        let mut c = EmhSurrogateCounts::default();
        check(unsafe { emh_surrogate_get_counts(self.raw, &mut c, reset as c_int) })?;
        Ok(c)
    }
}
