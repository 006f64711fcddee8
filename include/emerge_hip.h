/*
 * emerge_hip.h -- C-ABI of the MI355X (gfx950) MLS-MPM substep library.
 *
 * The library is a drop-in backend for the reference crate's CPU
 * `Simulation::step` (src/spacetime/solver/step.rs:23-60): a caller builds a
 * simulation from the same SimConfig / MaterialParams / SlipBoundary / Particle
 * data the reference uses, and `emh_sim_step(sim, n)` advances it exactly as
 * `n` calls of `Simulation::step` would (adaptive CFL substeps, state
 * projection, density re-estimate, P2G, grid update with gravity + slip walls,
 * G2P and the constitutive/plasticity update), on one GPU.
 *
 * Conventions
 *   - Every function returns EMH_OK (0) or a negative EMH_ERR_* code; the
 *     message of the last failure on the calling thread is emh_last_error().
 *     Nothing aborts the process (the reference panics/asserts instead,
 *     src/spacetime/solver/config.rs:247-299, materials/registry.rs:29-49).
 *   - The caller owns every host buffer; the library copies in and out and
 *     keeps no caller pointer.  A handle is not thread-safe: drive one handle
 *     from one host thread.
 *   - Struct layouts below are byte-compatible with the reference's
 *     repr(C) types where the reference has one (Particle 128 B,
 *     MaterialParams 96 B); `emh_config` is a fixed-width mirror of SimConfig
 *     (usize and bool fields become uint32_t).
 */
#ifndef EMERGE_HIP_H
#define EMERGE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EMH_ABI_VERSION 1u

/* ---- status codes ------------------------------------------------------- */
#define EMH_OK                 0
#define EMH_ERR_INVALID_ARG   -1  /* null pointer, bad size, failed validate() */
#define EMH_ERR_UNSUPPORTED   -2  /* a material/boundary/feature the GPU path does not implement */
#define EMH_ERR_DEVICE        -3  /* HIP runtime error or device loss */
#define EMH_ERR_OUT_OF_MEMORY -4
#define EMH_ERR_STATE         -5  /* call made in a state that does not allow it */

/* ---- constitutive model ids: ConstitutiveModel discriminants -------------
 * src/matter/materials/mod.rs:53-66 (ABI-stable values asserted at :71-85). */
#define EMH_MODEL_FALLBACK        0u
#define EMH_MODEL_FLUID           1u
#define EMH_MODEL_NEOHOOKEAN      2u
#define EMH_MODEL_COROTATED       3u
#define EMH_MODEL_SNOW            4u
#define EMH_MODEL_DRUCKER_PRAGER  5u
#define EMH_MODEL_VON_MISES       6u
#define EMH_MODEL_RANKINE         7u
#define EMH_MODEL_MU_I            8u  /* DruckerPragerMuI */
#define EMH_MODEL_VISCOELASTIC    9u
#define EMH_MODEL_NACC           10u
#define EMH_MODEL_GRANULAR_FLUID 11u

#define EMH_MAX_MATERIALS  64u   /* MAX_MATERIAL_SLOTS, materials/registry.rs:6 */
#define EMH_MAX_BOUNDARIES  8u

/* ---- SimConfig mirror: src/spacetime/solver/config.rs:26-121 ------------ */
typedef struct emh_config {
    uint32_t grid_res;                    /* usize */
    float    grid_cell_size;
    float    dt;
    uint32_t adaptive_timestep;           /* bool */
    uint32_t cfl_include_affine_speed;    /* bool */
    float    cfl_coefficient;
    float    material_cfl_coefficient;
    float    viscous_timestep_coefficient;
    float    min_dt;
    uint32_t project_invalid_state;       /* bool */
    float    projection_min_density;
    float    projection_min_volume;
    float    projection_min_deformation_j;
    float    gravity[2];
    uint32_t boundary_thickness;          /* usize */
    float    default_initial_volume;
    uint32_t recompute_density_each_step; /* bool */
    float    particle_mass;
    uint32_t max_substeps_per_step;       /* usize */
    float    apic_blend;
    float    j_max;
    float    sleep_threshold;             /* > 0: the sleep partition (see emh_sim_sleep_tag) */
    float    contact_friction;
    float    asflip_blend;                /* must be 0: ASFLIP is not supported */
    float    mixture_drag_coefficient;
    uint32_t mixture_pressure_iterations;
    float    dx_meters;
    float    dt_seconds;
} emh_config;                             /* 116 bytes */

/* ---- MaterialParams mirror (96 B): src/matter/materials/params.rs:8-101 --
 * Field meaning is per model, exactly as each model's `params()` fills it:
 *   NeoHookean  elastic.rs:275-288   Corotated corotated.rs:97-106
 *   Fluid       fluid.rs:209-226     DruckerPrager sand.rs:321-343
 *   Snow        snow.rs:178-191      VonMises von_mises.rs:143-152
 *   Viscoelastic viscoelastic.rs:149-159 (active_stress_coeff must be 0)
 *   Rankine     rankine.rs:206-216   DruckerPragerMuI sand_mui.rs:219-229
 *   GranularFluid granular_fluid.rs:232-247
 *   Bingham     bingham.rs:205-218 (model Fluid, yield stress in compression_limit)
 *   Nacc        nacc.rs:300-314 uploads model=NeoHookean; this library needs
 *               model=EMH_MODEL_NACC (10) instead, with the same field use
 *               (lambda = kappa - mu, hardening_exponent = xi,
 *                compression_limit = beta, stretch_limit = M). */
typedef struct emh_material_params {
    uint32_t model;
    float lambda;
    float mu;
    float hardening_exponent;
    float compression_limit;
    float stretch_limit;
    float rest_density;
    float eos_stiffness;
    float eos_power;
    float dynamic_viscosity;
    float volume_ratio_min;
    float volume_ratio_max;
    float dp_h0;
    float dp_h1;
    float dp_h2;
    float dp_h3;
    float active_stress_coeff;
    float hardening_modulus;
    float thermal_viscosity_coeff;
    float thermal_expansion;
    float pressure_floor;
    float bulk_viscosity;
    float surface_tension_coeff;
    float cohesion_coeff;
} emh_material_params;

/* Model fields that MaterialParams drops (each model struct's constructor
 * default in brackets).  Pass NULL to emh_sim_set_material for the defaults. */
typedef struct emh_material_ext {
    float    min_density;            /* NH elastic.rs:55, Fluid fluid.rs:60, NACC nacc.rs:71 [1e-6] */
    float    min_volume;             /* Fluid fluid.rs:61 [1e-6] */
    float    dp_volume_correction;   /* DP sand.rs:89 [1.0] */
    float    nacc_kappa;             /* NACC bulk modulus kappa (nacc.rs:40); 0 = lambda + mu */
    uint32_t nacc_hardening_enabled; /* nacc.rs:70 [hardening_factor > 0]; 2 = derive */
    uint32_t nacc_advance_f;         /* 0: F <- project(F_old) as CPU Simulation does (nacc.rs:255-269);
                                        1: F <- project((I + dt C) F_old) as the reference GPU path does */
    float    bingham_critical_shear_rate; /* Bingham bingham.rs:67 [1e-4]; 0 = the default */
    uint32_t fluid_bingham;          /* a Fluid record is a BinghamFluidMaterial even with yield stress 0
                                        (a positive compression_limit selects Bingham by itself,
                                        as BinghamFluidMaterial::params encodes it, bingham.rs:205-218) */
} emh_material_ext;                  /* 32 bytes */

/* ---- Particle mirror (128 B): src/matter/particle.rs:11-144 ------------- */
typedef struct emh_particle {
    float    x[2];
    float    v[2];
    float    velocity_gradient[4];    /* glam Mat2, column-major: c00, c10, c01, c11 */
    float    deformation_gradient[4]; /* column-major */
    float    mass;
    float    initial_volume;
    float    volume;
    float    density;
    uint32_t material_id;
    float    plastic_volume_ratio;
    float    hardening_scale;
    float    friction_hardening;
    float    log_volume_strain;
    float    temperature;
    uint32_t user_tag;
    float    activation;
    float    activation_dir[2];
    uint32_t muscle_group_id;
    uint32_t contact_group;
    uint32_t sleeping;
    uint32_t pinned;
    float    scalar_field;
    uint32_t _pad;
} emh_particle;

/* ---- SpawnRegion mirror: src/spacetime/solver/config.rs:317-364 --------- */
#define EMH_SHAPE_BOX  0u
#define EMH_SHAPE_DISK 1u
typedef struct emh_spawn_region {
    float    spacing;
    int32_t  box_size[2];
    float    box_center[2];
    uint32_t shape;                    /* EMH_SHAPE_* */
    float    disk_radius;
    float    initial_deformation_gradient[4];
    uint32_t precompute_initial_volumes;
    float    initial_velocity_scale;
    float    position_jitter;
    uint32_t rng_seed;
    uint32_t material_id;
    uint32_t has_mass_override;
    float    mass_override;
} emh_spawn_region;

/* ---- per-step diagnostics (Simulation fields, solver/mod.rs:60-66) ------ */
typedef struct emh_step_stats {
    uint32_t last_substeps;           /* last_substeps()  step.rs:323 */
    float    last_step_dt;            /* effective_dt()   step.rs:319 */
    float    last_sim_time_dropped;   /* step.rs:51 */
    uint32_t last_vel_clamp_count;    /* step.rs:196 */
    uint32_t last_j_projection_count; /* step.rs:69-71 */
    uint32_t total_substeps;          /* since creation */
    uint64_t frame_index;             /* step.rs:59 */
} emh_step_stats;

/* ---- diagnostics: SimSnapshot (src/systems/diagnostics/snapshot.rs:8-80),
 * the fields Simulation::diagnostics_snapshot fills
 * (src/spacetime/solver/queries.rs:15-30), without the wall-clock timing.
 * Particle sums are accumulated in f64 (the reference sums in f32 in
 * particle order); the grid fields are those of the last substep's updated
 * grid.  active_count / sleeping_count split the particles by the sleep
 * partition. */
typedef struct emh_snapshot {
    uint64_t frame_index;
    float    configured_dt;
    float    effective_dt;
    uint32_t substeps_last_step;
    uint32_t particle_count;
    uint32_t valid_particle_count;
    uint32_t active_grid_cells;
    float    particles_per_active_cell;
    float    mixed_material_cell_ratio;
    float    mixed_material_particle_ratio;
    float    total_particle_mass;
    float    total_grid_mass;
    float    relative_mass_error;
    float    total_particle_momentum[2];
    float    total_grid_momentum[2];
    float    relative_momentum_error;
    float    max_particle_speed;
    float    max_grid_speed;
    float    cfl_number;
    float    min_deformation_j;
    float    max_deformation_j;
    uint32_t out_of_bounds_particles;
    uint32_t invalid_physical_particle_values;
    uint32_t non_finite_particle_values;
    uint32_t non_finite_grid_values;
    float    recommended_max_dt_from_velocity_cfl;
    float    avg_plastic_jacobian;
    float    min_plastic_jacobian;
    float    avg_elastic_hardening;
    uint32_t active_count;
    uint32_t sleeping_count;
    uint32_t vel_clamp_count;
    uint32_t j_projection_count;
    float    sim_time_dropped;
    float    total_kinetic_energy;
    float    max_pinned_particle_speed;
    uint32_t _pad;
} emh_snapshot;

typedef struct emh_sim emh_sim;

/* ---- library ------------------------------------------------------------ */
uint32_t    emh_abi_version(void);
const char* emh_last_error(void);
/* Number of HIP devices visible; 0 without a GPU (does not fail). */
int         emh_device_count(int* count);

/* ---- config helpers ----------------------------------------------------- */
/* SimConfig::default()  config.rs:123-159 */
int emh_config_default(emh_config* out);
/* SimConfig::standard(grid_res, dt, gravity)  config.rs:165-172 */
int emh_config_standard(uint32_t grid_res, float dt, float gx, float gy, emh_config* out);
/* SimConfig::unsafe_defaults()  config.rs:179-185 */
int emh_config_unsafe_defaults(emh_config* out);
/* SimConfig::validate()  config.rs:247-299 (error code instead of a panic) */
int emh_config_validate(const emh_config* cfg);

/* ---- material constructors (each fills the 96 B params + the extension) -- */
/* lame_from_young  materials/utils.rs:183-193 */
int emh_lame_from_young(float young_modulus, float poisson_ratio, float* lambda, float* mu);
/* NeoHookeanMaterial::new  elastic.rs:51-61 */
int emh_material_neohookean(float lambda, float mu, emh_material_params* p, emh_material_ext* e);
/* CorotatedMaterial::new  corotated.rs:29-36 */
int emh_material_corotated(float lambda, float mu, emh_material_params* p, emh_material_ext* e);
/* NewtonianFluidMaterial::new  fluid.rs:48-67 */
int emh_material_fluid(float rest_density, float dynamic_viscosity, float eos_stiffness,
                       float eos_power, emh_material_params* p, emh_material_ext* e);
/* NewtonianFluidMaterial::low_viscosity  fluid.rs:73-75 */
int emh_material_fluid_low_viscosity(float rest_density, float eos_stiffness,
                                     emh_material_params* p, emh_material_ext* e);
/* DruckerPragerMaterial::new  sand.rs:81-94 */
int emh_material_drucker_prager(float lambda, float mu, emh_material_params* p, emh_material_ext* e);
/* StomakhinMaterial::new  snow.rs:40-59 */
int emh_material_snow(float lambda, float mu, float hardening_exponent, float compression_limit,
                      float stretch_limit, float min_plastic_jacobian, float max_plastic_jacobian,
                      emh_material_params* p, emh_material_ext* e);
/* StomakhinMaterial::from_young_modulus  snow.rs:68-71 */
int emh_material_snow_from_young(float young_modulus, float poisson_ratio,
                                 emh_material_params* p, emh_material_ext* e);
/* VonMisesMaterial::with_hardening(lambda, mu, yield_stress, H)  von_mises.rs:60-69;
 * ::new is H = 0 (von_mises.rs:49-56).  H < 0 is EMH_ERR_INVALID_ARG (the
 * reference asserts). */
int emh_material_von_mises(float lambda, float mu, float yield_stress, float hardening_modulus,
                           emh_material_params* p, emh_material_ext* e);
/* RankineMaterial::new(lambda, mu, tensile_strength, softening_rate)  rankine.rs:53-60 */
int emh_material_rankine(float lambda, float mu, float tensile_strength, float softening_rate,
                         emh_material_params* p, emh_material_ext* e);
/* MuIRheologyMaterial with explicit (mu_static, mu_dynamic, Q): the fields of sand_mui.rs:44-58 */
int emh_material_mu_i(float lambda, float mu, float mu_static, float mu_dynamic, float inertial_q,
                      emh_material_params* p, emh_material_ext* e);
/* MuIRheologyMaterial::new(lambda, mu)  sand_mui.rs:62-70 (tan 20.9 deg, tan 32.8 deg, Q 5.58) */
int emh_material_mu_i_default(float lambda, float mu, emh_material_params* p, emh_material_ext* e);
/* GranularFluidMaterial::new(lambda, mu, rho0, k, xi, theta_c)  granular_fluid.rs:61-83 */
int emh_material_granular_fluid(float lambda, float mu, float rest_density, float eos_stiffness,
                                float hardening_exponent, float compression_limit,
                                emh_material_params* p, emh_material_ext* e);
/* BinghamFluidMaterial::new(rho0, eta, k, gamma, yield_stress)  bingham.rs:53-71: a model-Fluid
 * record (bingham.rs:205-218) with ext.fluid_bingham = 1 */
int emh_material_bingham(float rest_density, float dynamic_viscosity, float eos_stiffness,
                         float eos_power, float yield_stress,
                         emh_material_params* p, emh_material_ext* e);
/* ViscoelasticMaterial::new(lambda, mu, viscosity)  viscoelastic.rs:53-62 (j_min 0.01) */
int emh_material_viscoelastic(float lambda, float mu, float viscosity, emh_material_params* p,
                              emh_material_ext* e);
/* NaccMaterial::new(mu, kappa, friction M, cohesion beta, hardening xi)  nacc.rs:63-73 */
int emh_material_nacc(float mu, float kappa, float friction, float cohesion, float hardening_factor,
                      emh_material_params* p, emh_material_ext* e);
/* NaccMaterial::from_young_modulus  nacc.rs:77-90 */
int emh_material_nacc_from_young(float young_modulus, float poisson_ratio, float friction,
                                 float cohesion, float hardening_factor,
                                 emh_material_params* p, emh_material_ext* e);
/* f32::to_radians as Rust computes it (for friction angles) */
float emh_to_radians(float degrees);

/* ---- simulation lifecycle (solver/lifecycle.rs) -------------------------- */
/* Simulation::empty(config)  lifecycle.rs:26-55: Fallback material 0 and one
 * SlipBoundary(config.boundary_thickness). */
int emh_sim_create(const emh_config* cfg, emh_sim** out);
/* Simulation::new(config, spawn)  lifecycle.rs:57-105 */
int emh_sim_create_from_spawn(const emh_config* cfg, const emh_spawn_region* spawn, emh_sim** out);
void emh_sim_destroy(emh_sim* sim);
/* with_material / set_material (id == 0: with_default_material, which also
 * re-runs init_particle on every particle, lifecycle.rs:107-111,328-334).
 * ext may be NULL.  Ids must be contiguous (registry.rs:29-49). */
int emh_sim_set_material(emh_sim* sim, uint32_t material_id, const emh_material_params* params,
                         const emh_material_ext* ext);
/* Simulation::reinit_all_particle_state  lifecycle.rs:328-334 */
int emh_sim_reinit_particle_state(emh_sim* sim);
/* with_boundary(Box::new(SlipBoundary::new(t)))  lifecycle.rs:114-117, slip.rs:5-24 */
int emh_sim_add_slip_boundary(emh_sim* sim, uint32_t thickness);
/* with_boundary(Box::new(FrictionBoundary::new(t, mu)))  friction.rs:28-63: no-penetration
 * walls with Coulomb friction (boundary/mod.rs:57-75); mu outside [0, 1] is EMH_ERR_INVALID_ARG. */
int emh_sim_add_friction_boundary(emh_sim* sim, uint32_t thickness, float friction_coefficient);
/* with_boundary(Box::new(HeightmapBoundary::new(heights, friction, wall_thickness)))
 * heightmap.rs:44-114: terrain height per x column (the last height repeats past
 * the end), Coulomb friction on it, slip side and top walls.  The heights are copied. */
int emh_sim_add_heightmap_boundary(emh_sim* sim, const float* heights, uint32_t n_heights,
                                   float friction, uint32_t wall_thickness);
/* with_boundary(Box::new(PredictiveBoundary::new(t, wall_min)))  predictive.rs:17-35 */
int emh_sim_add_predictive_boundary(emh_sim* sim, uint32_t thickness, float predictive_wall_min);
/* clear_boundaries  lifecycle.rs:364-366 */
int emh_sim_clear_boundaries(emh_sim* sim);
/* add_body(spawn) -> tag  solver/particles.rs:355-424 */
int emh_sim_add_body(emh_sim* sim, const emh_spawn_region* spawn, uint32_t* tag_out);
/* Replace all particles (all active).  Particle::material_id must be registered. */
int emh_sim_set_particles(emh_sim* sim, const emh_particle* particles, size_t n);
int emh_sim_num_particles(const emh_sim* sim, size_t* n);
/* particles() in index order (copies device state to the host). */
int emh_sim_get_particles(emh_sim* sim, emh_particle* out, size_t n);
int emh_sim_get_config(const emh_sim* sim, emh_config* out);

/* ---- the hot path ------------------------------------------------------- */
/* n_steps x Simulation::step  step.rs:23-60 / step_n step.rs:327-331 */
int emh_sim_step(emh_sim* sim, uint32_t n_steps);
int emh_sim_stats(const emh_sim* sim, emh_step_stats* out);
/* Grid after the last substep's P2G (before normalisation), dense, flat index
 * x * grid_res + y as Grid uses (grid/mod.rs:89-99).  mass: grid_res^2 floats,
 * momentum: 2 * grid_res^2 floats (x, y interleaved).  Either may be NULL. */
int emh_sim_get_grid(emh_sim* sim, float* mass, float* momentum, size_t n_cells);
/* Simulation::diagnostics_snapshot (src/spacetime/solver/queries.rs:15-30),
 * computed on the device. */
int emh_sim_snapshot(emh_sim* sim, emh_snapshot* out);
/* Wait for all queued device work of this handle. */
int emh_sim_synchronize(emh_sim* sim);

/* ---- device-resident access (GpuSimulation's buffer API) -----------------
 * For consumers on the same GPU, e.g. a renderer: the state is written into
 * caller-owned device memory (hipMalloc'd or interop-mapped) with no host
 * round trip.  Each call returns once the copy is complete. */
/* GpuSimulation::particle_buffer (gpu/solver/mod.rs:466-471): every particle as
 * its 128 B Particle record, in caller order, into `dst` (capacity in records). */
int emh_sim_export_particles_device(emh_sim* sim, void* dst, size_t capacity);
/* GpuSimulation::grid_buffer (gpu/solver/mod.rs:510-515): grid_res^2 cells of
 * 16 B {momentum.x, momentum.y, mass, 0} (x-major), the same grid as
 * emh_sim_get_grid. */
int emh_sim_export_grid_device(emh_sim* sim, void* dst, size_t capacity_cells);
/* Simulation::set_gravity  lifecycle.rs:403-405 (applies from the next step) */
/* Sleep partition (solver/mod.rs:40-42, step.rs:107-131,299-315): with
 * config.sleep_threshold > 0, a particle without activation whose speed falls
 * below the threshold sleeps from the end of that substep (skipped by P2G,
 * G2P, the CFL bound and projection, its state frozen) until this substep's
 * P2G reaches one of its stencil cells, or an impulse covers it.  Particles
 * uploaded with Particle::sleeping = 1 start asleep.  The reference also
 * moves sleeping particles to the end of its arrays; here particles keep the
 * caller's order and carry the state in Particle::sleeping.
 * sleep_tag / wake_tag = Simulation::sleep_tag / wake_tag
 * (solver/particles.rs:270-312): every particle with that user_tag. */
int emh_sim_sleep_tag(emh_sim* sim, uint32_t tag);
int emh_sim_wake_tag(emh_sim* sim, uint32_t tag);
int emh_sim_set_gravity(emh_sim* sim, float gx, float gy);
/* Simulation::apply_impulse(center, radius, force)  solver/particles.rs:115-139:
 * v += force (1 - |d| / radius) within radius, speed clamped to
 * grid_cell_size / min_dt.  Applied on the device. */
int emh_sim_apply_impulse(emh_sim* sim, float cx, float cy, float radius, float fx, float fy);
/* Simulation::apply_radial_impulse(center, radius, strength)  solver/particles.rs:141-165 */
int emh_sim_apply_radial_impulse(emh_sim* sim, float cx, float cy, float radius, float strength);

/* ---- per-kernel timing ---------------------------------------------------
 * HIP events on the library's stream, two per step and one where a step loop
 * stops to read its control block, recorded without stopping the queue (the
 * spans are read at those stops).  Two entries, summed over the steps taken
 * since enabling -- "step_setup" (launches = steps): the step's start kernel
 * and, on reorder steps, the cell-key and scan kernels; "k_substep"
 * ("k_density_p2g+k_substep" when a density pass runs; launches = substeps
 * taken): the back-to-back substep launches of each step.  The counterpart of
 * GpuSimulation::enable_profiling / last_pass_timings_ns
 * (src/systems/gpu/solver/profiling.rs:9-75).  Enabling resets the totals. */
typedef struct emh_kernel_time {
    char     name[32];
    uint64_t launches;
    double   total_ms;
} emh_kernel_time;
int emh_sim_set_profiling(emh_sim* sim, int enable);

/* ---- particle order -------------------------------------------------------
 * No counterpart in the reference (its CPU path has no particle sort; the WGSL
 * path re-sorts every frame, gpu/solver/step.rs:339-385).  The particles are
 * physically reordered into cell-key order at the first step after an upload
 * and then at a step start once `substeps` substeps have run since the last
 * reorder (default 4096; >= 1).  Results do not depend on it beyond f32
 * summation order; it trades the reorder's cost against chunk locality. */
int emh_sim_set_reorder_interval(emh_sim* sim, uint32_t substeps);
/* The stored order itself: ids_out[j] is the caller index of the particle in
 * device slot j (n = the particle count).  Right after a reorder this is the
 * stable sort of the keys of the positions the reorder saw, by stored slot:
 * the cell key (emh_sort_cell_keys' permutation of them, for an upload) or,
 * with per-material segments in a scene of several models, model << bits |
 * cell key (emh_tuning).  EMH_ERR_STATE before the first step after an upload
 * (the particles are not on the device yet). */
int emh_sim_get_slot_order(emh_sim* sim, uint32_t* ids_out, size_t n);
int emh_sim_kernel_times(emh_sim* sim, emh_kernel_time* out, uint32_t cap, uint32_t* n_out);

/* ---- launch tuning and the autotuner --------------------------------------
 * No counterpart in the reference (its WGSL backend submits up to 64
 * substeps per frame in one command buffer, gpu/solver/step.rs:404-441).
 * Three run-time settings change how a step is launched, never its results
 * beyond f32 summation order:
 *   reorder_interval  substeps between physical cell-key reorders (>= 1);
 *   graphs            1: each step's run of substep launches is captured
 *                     once into a hipGraph (per launch-number residue mod 6,
 *                     run length and kernel variant) and replayed; 0: the
 *                     launches are queued one by one;
 *   model_segments    1: a scene whose particles use several constitutive
 *                     models is reordered by (model, cell key), so that a
 *                     chunk of the substep -- and each wave of it -- runs one
 *                     model's update; 0 (default): by cell key alone, which
 *                     keeps each chunk spatially compact.  It takes effect at
 *                     the next reorder.
 * emh_*_autotune times every candidate (reorder intervals x graphs, x model
 * segments on / off for scenes that mix models) on the
 * scene itself: from the current state, one untimed step, then trial_steps
 * steps timed on the host clock to their completion; it then puts the state
 * back (the particles are re-uploaded at the next step, the step counters
 * restored), applies the fastest candidate and reports it in *chosen (and
 * every candidate's cost in timed[0 .. min(cap, *n_timed))).  A scene in a
 * surrogate mode (harvest or inference) is refused with EMH_ERR_STATE: its
 * trial steps would add records or counts to the surrogate that the restore
 * cannot take back. */
typedef struct emh_tuning {
    uint32_t reorder_interval;
    uint32_t graphs;
    float    ns_per_particle_substep; /* set by the autotuners: wall time / (particles x substeps) */
    uint32_t model_segments;
} emh_tuning;
int emh_sim_get_tuning(emh_sim* sim, emh_tuning* out);
int emh_sim_set_tuning(emh_sim* sim, const emh_tuning* t);
int emh_sim_autotune(emh_sim* sim, uint32_t trial_steps, emh_tuning* chosen, emh_tuning* timed,
                     uint32_t timed_cap, uint32_t* n_timed);

/* ---- the cell-key sort on its own ---------------------------------------
 * No CPU counterpart in the reference (its WGSL sort is a coarse 16x16-block
 * counting sort with a non-deterministic order inside a block,
 * src/systems/gpu/shaders/particle_sort.wgsl:94-245).  This is the device
 * cell key the substep reorders particles by: key = emh_cell_key(floor(x),
 * floor(y)) -- the cell's index along the Hilbert curve over the smallest
 * 2^k x 2^k square (k >= 1) that holds the grid, coordinates clamped into the
 * grid -- stable by input index.  keys_out[i] is the key of input i; perm_out[j] is the input
 * index at sorted position j. */
uint32_t emh_cell_key(int32_t cx, int32_t cy, uint32_t grid_res);
int emh_sort_cell_keys(const float* xy, size_t n, uint32_t grid_res,
                       uint32_t* keys_out, uint32_t* perm_out);

/* ---- the same sort without allocating ------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory -- code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc: emh_sort_cell_keys on caller-owned device buffers (xy:
 * n x 2 floats, 8-byte aligned; keys_out, perm_out: n each, not aliasing xy or
 * each other), with the same results, calling no allocator.  Its scratch is
 * a static device array of EMH_SORT_NOALLOC_SCRATCH_BYTES on each GPU it runs
 * on, held until the process ends; n above EMH_SORT_NOALLOC_MAX_KEYS fails
 * with EMH_ERR_INVALID_ARG before any work.  The work is enqueued on `stream`
 * (a hipStream_t; NULL: the null stream) and not waited for.  All calls on
 * one GPU share its scratch: calls whose work can overlap (different streams
 * or threads) race, so the caller serialises them.  Nothing in the scratch
 * needs clearing between calls. */
#define EMH_SORT_NOALLOC_MAX_KEYS 1048576u
#define EMH_SORT_NOALLOC_SCRATCH_BYTES 18878464u /* (4 x 1048576 keys + 1024 x 513 counts, totals) x 4 B */
int emh_sort_cell_keys_noalloc(const float* xy_dev, size_t n, uint32_t grid_res, uint32_t* keys_out_dev,
                               uint32_t* perm_out_dev, void* stream);

/* ---------------------------------------------------------------------
 * Batched scenes.
 *
 * No counterpart in the reference (it steps one Simulation at a time,
 * step.rs:23-60): many independent simulations -- each with its own grid
 * size, particle count, materials and boundaries -- stepped together, one
 * kernel launch per stage for all of them.  Each scene keeps its own
 * adaptive substep schedule: emh_batch_step(b, n) leaves every scene where
 * n calls of emh_sim_step on it would (up to f32 roundoff of the grid sums).
 * The batch does not own the scenes.  While batched, a scene is stepped only
 * through the batch; reading its particles or stats, or editing it, between
 * batch steps is fine (an edit makes the next batch step restart every
 * scene's loop from its stored state, as an upload does).  Every scene needs
 * particles; the surrogate modes are not batched.
 * ------------------------------------------------------------------- */
typedef struct emh_batch emh_batch;
int emh_batch_create(emh_sim* const* sims, uint32_t n, emh_batch** out);
int emh_batch_destroy(emh_batch* b);
int emh_batch_step(emh_batch* b, uint32_t n_steps);
/* Launch tuning of the batch (see emh_sim_autotune; the batch's default
 * reorder interval is 8 substeps) and its autotuner. */
int emh_batch_get_tuning(emh_batch* b, emh_tuning* out);
int emh_batch_set_tuning(emh_batch* b, const emh_tuning* t);
int emh_batch_autotune(emh_batch* b, uint32_t trial_steps, emh_tuning* chosen, emh_tuning* timed,
                       uint32_t timed_cap, uint32_t* n_timed);
/* Live timing of the batch's stages, the counterpart of emh_sim_set_profiling
 * for a batch: HIP events on the batch stream before each stage's launches,
 * read at the drains.  Four entries, summed over the batch steps since
 * enabling: "kb_begin_step" (launches = batch steps), "kb_sort" (the radix
 * sort of every scene, launches = reorder steps), "kb_substep" (the density
 * pass where a scene has one and the substep kernels; launches = substep
 * launch slots, each covering every scene) and "kb_gather_ctrl" (the control
 * gather of each drain).  Enabling resets the totals; the autotuner refuses
 * to run while profiling is on. */
int emh_batch_set_profiling(emh_batch* b, int enable);
int emh_batch_kernel_times(emh_batch* b, emh_kernel_time* out, uint32_t cap, uint32_t* n_out);

/* ---------------------------------------------------------------------
 * Learned NACC return-mapping surrogate.
 *
 * No counterpart in the reference: the crate's return mapping is
 * NaccMaterial::project (materials/nacc.rs:129-223), applied by
 * update_particle (nacc.rs:255-269); these entry points add the learned
 * replacement for it.  A 6-64-64-64-3 SiLU MLP maps the trial state
 * (ln s1, ln s2, alpha, M, beta, xi) -- s the singular values of F_trial,
 * alpha the hardening state (Particle::log_volume_strain slot), M / beta /
 * xi the material's stretch_limit / compression_limit / hardening_exponent
 * fields -- to the projected (ln s1', ln s2', alpha').  Weights are BF16 on
 * MFMA with FP32 accumulation; training (forward, backward, Adam) runs on
 * the device.  Each particle is gated by the exact yield-function check:
 * a state inside the yield surface keeps F; the tension and compression
 * caps (the closed-form Cases A and B of nacc.rs) keep the plain return; a
 * state projected onto the yield surface (Case C) goes through the network
 * unless its inputs leave the training range (more than z_max standard
 * deviations) or s2 <= 0, in which case the plain return mapping runs
 * instead.
 * ------------------------------------------------------------------- */
typedef struct emh_surrogate emh_surrogate;

#define EMH_SURROGATE_PARAMS 8963u /* W1[64][6] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[3][64] b4[3] */
#define EMH_SURROGATE_NORM   20u   /* in_mu[6] in_sd[6] out_mu[3] out_sd[3] z_max pad */

typedef struct emh_surrogate_train_options {
    uint32_t epochs;
    uint32_t batch;      /* rounded up to a multiple of 256 */
    float    lr;         /* cosine schedule from lr to lr_final over the run */
    float    lr_final;
    float    beta1, beta2, eps;
    uint32_t pad;
    uint64_t seed;       /* per-epoch sample order */
    uint64_t max_steps;  /* 0: epochs * (n_set / batch) */
} emh_surrogate_train_options;

typedef struct emh_surrogate_counts {
    uint64_t yielding;   /* trial states that failed the yield check */
    uint64_t surrogate;  /* ... returned through the network */
    uint64_t fallback;   /* ... returned by the plain mapping (caps A / B, out of range, s2 <= 0) */
} emh_surrogate_counts;

/* Glorot-uniform weights from `seed`, zero biases, identity normalisation. */
int emh_surrogate_create(uint64_t seed, emh_surrogate** out);
int emh_surrogate_destroy(emh_surrogate* sg);
/* FP32 master weights (EMH_SURROGATE_PARAMS floats) and the normalisation
 * block (EMH_SURROGATE_NORM floats); either pointer may be NULL.  set_params
 * resets the Adam moments. */
int emh_surrogate_get_params(emh_surrogate* sg, float* params, float* norm);
int emh_surrogate_set_params(emh_surrogate* sg, const float* params, const float* norm);
/* mode 0: plain return mapping; 1: plain, and record every NACC particle's
 * trial state (F_trial as update_particle forms it, alpha, material) into
 * sg's buffer of harvest_capacity records (every harvest_stride-th trial
 * state offered; 0 or 1: every one); 2: return mapping through sg.
 * Modes 1 and 2 need a scene whose particles are all NACC (the next step
 * fails with EMH_ERR_UNSUPPORTED otherwise).  sg must outlive its use by the
 * sim (or be detached with mode 0). */
int emh_sim_set_surrogate_mode(emh_sim* sim, emh_surrogate* sg, uint32_t mode, uint64_t harvest_capacity,
                               uint32_t harvest_stride);
int emh_surrogate_harvest_count(emh_surrogate* sg, uint64_t* offered, uint64_t* stored);
/* Harvested records [first, first + n), 12 floats each: F_trial (column-major
 * a, b, c, d), alpha, mu, kappa, M, beta, xi, flags (uint32 bits), 0. */
int emh_surrogate_get_records(emh_surrogate* sg, uint64_t first, uint64_t n, float* out);
/* The batch return mapping (emh_nacc_return_batch) of harvested records
 * [first, first + n), plain or through sg; outputs copied to the host
 * arrays when non-NULL (f_out [n][4], alpha_out [n]). */
int emh_surrogate_eval_records(emh_surrogate* sg, const emh_material_params* p, const emh_material_ext* ext,
                               uint64_t first, uint64_t n, int use_surrogate, uint32_t reps, float* ms_per_rep,
                               float* f_out, float* alpha_out);
/* Training set from the harvested records: those projected onto the yield
 * surface (Case C) with s2 > 0, labelled by the plain return mapping; sets
 * the normalisation. */
int emh_surrogate_build_training_set(emh_surrogate* sg, uint64_t* n_samples);
/* Host samples x[n][6] (raw inputs), y[n][3] (raw targets); sets the normalisation. */
int emh_surrogate_set_training_set(emh_surrogate* sg, const float* x, const float* y, uint64_t n);
int emh_surrogate_get_training_set(emh_surrogate* sg, uint64_t first, uint64_t n, float* x, float* y);
/* Adam on the training set; epoch_loss[e] = mean over samples of the summed
 * squared error of the normalised residual outputs. */
int emh_surrogate_train(emh_surrogate* sg, const emh_surrogate_train_options* o, float* epoch_loss,
                        uint32_t n_epoch_loss);
/* Gradient of that loss over the given samples (no update): grad has
 * EMH_SURROGATE_PARAMS floats; loss is the mean. */
int emh_surrogate_gradient(emh_surrogate* sg, const float* x, const float* y, uint32_t n, float* grad, float* loss);
/* Network outputs y[n][3] = x[:, 0:3] + denormalised residual for raw x[n][6]. */
int emh_surrogate_infer(emh_surrogate* sg, const float* x, uint64_t n, float* y);
int emh_surrogate_get_counts(emh_surrogate* sg, emh_surrogate_counts* out, int reset);
/* NACC return mapping of n trial states on the device (device pointers:
 * f_trial / f_out [n][4] column-major 2x2, alpha / alpha_out [n]), through
 * sg (gated as above) or, with sg NULL, the plain mapping; `reps` back-to-back
 * launches timed with HIP events on the launching stream. */
int emh_nacc_return_batch(emh_surrogate* sg, const emh_material_params* p, const emh_material_ext* ext,
                          const float* f_trial, const float* alpha, uint64_t n, float* f_out, float* alpha_out,
                          uint32_t reps, float* ms_per_rep);

#ifdef __cplusplus
}
#endif

#endif /* EMERGE_HIP_H */
