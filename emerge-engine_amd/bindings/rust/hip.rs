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

use core::ffi::{c_char, c_int, CStr};

use crate::forces::boundary::slip::SlipBoundary;
use crate::matter::materials::nacc::NaccMaterial;
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

extern "C" {
    fn emh_last_error() -> *const c_char;
    fn emh_sim_create(cfg: *const EmhConfig, out: *mut *mut EmhSim) -> c_int;
    fn emh_sim_create_from_spawn(cfg: *const EmhConfig, spawn: *const EmhSpawnRegion, out: *mut *mut EmhSim)
        -> c_int;
    fn emh_sim_destroy(sim: *mut EmhSim);
    fn emh_sim_set_material(sim: *mut EmhSim, id: u32, p: *const MaterialParams, e: *const EmhMaterialExt)
        -> c_int;
    fn emh_sim_reinit_particle_state(sim: *mut EmhSim) -> c_int;
    fn emh_sim_add_slip_boundary(sim: *mut EmhSim, thickness: u32) -> c_int;
    fn emh_sim_clear_boundaries(sim: *mut EmhSim) -> c_int;
    fn emh_sim_add_body(sim: *mut EmhSim, spawn: *const EmhSpawnRegion, tag_out: *mut u32) -> c_int;
    fn emh_sim_set_particles(sim: *mut EmhSim, p: *const Particle, n: usize) -> c_int;
    fn emh_sim_num_particles(sim: *const EmhSim, n: *mut usize) -> c_int;
    fn emh_sim_get_particles(sim: *mut EmhSim, out: *mut Particle, n: usize) -> c_int;
    fn emh_sim_step(sim: *mut EmhSim, n_steps: u32) -> c_int;
    fn emh_sim_set_reorder_interval(sim: *mut EmhSim, substeps: u32) -> c_int;
    fn emh_sim_stats(sim: *const EmhSim, out: *mut EmhStepStats) -> c_int;
    fn emh_sim_get_grid(sim: *mut EmhSim, mass: *mut f32, momentum: *mut f32, n_cells: usize) -> c_int;
    // learned NACC return mapping
    fn emh_surrogate_create(seed: u64, out: *mut *mut EmhSurrogate) -> c_int;
    fn emh_surrogate_destroy(sg: *mut EmhSurrogate) -> c_int;
    fn emh_sim_set_surrogate_mode(sim: *mut EmhSim, sg: *mut EmhSurrogate, mode: u32, harvest_capacity: u64,
                                  harvest_stride: u32) -> c_int;
    fn emh_surrogate_build_training_set(sg: *mut EmhSurrogate, n_samples: *mut u64) -> c_int;
    fn emh_surrogate_train(sg: *mut EmhSurrogate, o: *const EmhTrainOptions, epoch_loss: *mut f32, n: u32) -> c_int;
    fn emh_surrogate_get_counts(sg: *mut EmhSurrogate, out: *mut EmhSurrogateCounts, reset: c_int) -> c_int;
}

#[repr(C)]
pub struct EmhSurrogate {
    _opaque: [u8; 0],
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
/// the fields it drops.  NACC needs its true parameters (its `params()` uploads
/// a NeoHookean stand-in, nacc.rs:300-314).
pub fn material_record(m: &dyn MaterialModel) -> (MaterialParams, EmhMaterialExt) {
    let mut p = m.params();
    p.model = m.constitutive_model() as u32;
    (p, EmhMaterialExt::default())
}
pub fn nacc_record(m: &NaccMaterial, advance_f: bool) -> (MaterialParams, EmhMaterialExt) {
    let mut p = m.params();
    p.model = ConstitutiveModel::Nacc as u32;
    p.lambda = m.kappa - m.mu;
    let ext = EmhMaterialExt { nacc_kappa: m.kappa, nacc_advance_f: advance_f as u32, ..Default::default() };
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
