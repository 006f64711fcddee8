// capi.cpp -- the C-ABI of include/emerge_hip.h: a C++ mirror of the
// reference's Simulation lifecycle (solver/lifecycle.rs, solver/particles.rs)
// that owns device memory and drives the substep kernels.
//
// There is no CPU compute path here: every substep, and the spawn-time volume
// estimate, runs on the GPU; a call that would need the GPU fails with
// EMH_ERR_DEVICE when no device is present.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "../../include/emerge_hip.h"
#include "interop_kernels.hpp"
#include "kernels.hpp"
#include "sort_static.hpp"
#include "spawn_kernels.hpp"
#include "surrogate_kernels.hpp"

using namespace emh;

static_assert(sizeof(emh_particle) == 128, "Particle is 128 bytes (particle.rs:144)");
static_assert(sizeof(emh_material_params) == 96, "MaterialParams is 96 bytes (params.rs:101)");
static_assert(sizeof(emh_config) == 116, "emh_config layout");
static_assert(sizeof(emh_material_ext) == 32, "emh_material_ext layout");

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIP_TRY(expr)                                                                             \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess) return fail(EMH_ERR_DEVICE, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

int device_present() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

void ext_defaults(emh_material_ext* e) {
    std::memset(e, 0, sizeof(*e));
    e->min_density = 1.0e-6f;
    e->min_volume = 1.0e-6f;
    e->dp_volume_correction = 1.0f;
    e->nacc_hardening_enabled = 2u;
    e->bingham_critical_shear_rate = 1.0e-4f;
}

// f32::to_radians: self * (PI / 180.0) evaluated in f32.
float to_radians(float deg) { return deg * (3.14159265358979323846f / 180.0f); }

int validate_config(const emh_config& c) {  // SimConfig::validate  config.rs:247-299
    if (c.grid_res < 4) return fail(EMH_ERR_INVALID_ARG, "grid_res must be >= 4");
    if (!(c.grid_cell_size > 0.0f)) return fail(EMH_ERR_INVALID_ARG, "grid_cell_size must be positive");
    if (!(c.dt > 0.0f)) return fail(EMH_ERR_INVALID_ARG, "dt must be positive");
    if (!(c.cfl_coefficient > 0.0f)) return fail(EMH_ERR_INVALID_ARG, "cfl_coefficient must be positive");
    if (!(c.material_cfl_coefficient > 0.0f))
        return fail(EMH_ERR_INVALID_ARG, "material_cfl_coefficient must be positive");
    if (!(c.viscous_timestep_coefficient > 0.0f))
        return fail(EMH_ERR_INVALID_ARG, "viscous_timestep_coefficient must be positive");
    if (!(c.min_dt > 0.0f)) return fail(EMH_ERR_INVALID_ARG, "min_dt must be positive");
    if (!(c.min_dt <= c.dt)) return fail(EMH_ERR_INVALID_ARG, "min_dt must be <= dt");
    if (!(c.projection_min_density > 0.0f))
        return fail(EMH_ERR_INVALID_ARG, "projection_min_density must be positive");
    if (!(c.projection_min_volume > 0.0f))
        return fail(EMH_ERR_INVALID_ARG, "projection_min_volume must be positive");
    if (!(c.projection_min_deformation_j > 0.0f))
        return fail(EMH_ERR_INVALID_ARG, "projection_min_deformation_j must be positive");
    if (!(c.particle_mass > 0.0f)) return fail(EMH_ERR_INVALID_ARG, "particle_mass must be positive");
    if (!(c.contact_friction >= 0.0f)) return fail(EMH_ERR_INVALID_ARG, "contact_friction must be non-negative");
    if (c.max_substeps_per_step == 0) return fail(EMH_ERR_INVALID_ARG, "max_substeps_per_step must be > 0");
    if (!(c.default_initial_volume > 0.0f))
        return fail(EMH_ERR_INVALID_ARG, "default_initial_volume must be positive");
    if (!(c.j_max > 1.0f)) return fail(EMH_ERR_INVALID_ARG, "j_max must be > 1.0");
    if (!(c.apic_blend >= 0.0f && c.apic_blend <= 1.0f))
        return fail(EMH_ERR_INVALID_ARG, "apic_blend must be in [0, 1]");
    if (!(c.boundary_thickness > 0 && c.boundary_thickness < c.grid_res - 1))
        return fail(EMH_ERR_INVALID_ARG, "boundary_thickness must be in [1, grid_res-2]");
    return EMH_OK;
}

int check_supported_config(const emh_config& c) {
    if (!(c.sleep_threshold >= 0.0f) || !std::isfinite(c.sleep_threshold))
        return fail(EMH_ERR_INVALID_ARG, "sleep_threshold must be finite and >= 0");
    if (c.asflip_blend != 0.0f) return fail(EMH_ERR_UNSUPPORTED, "asflip_blend > 0 is not supported");
    if (c.grid_res > 16384) return fail(EMH_ERR_UNSUPPORTED, "grid_res above 16384 is not supported");
    return EMH_OK;
}

int check_supported_material(const emh_material_params& p) {
    switch (p.model) {
        case EMH_MODEL_FALLBACK:
        case EMH_MODEL_FLUID:
        case EMH_MODEL_NEOHOOKEAN:
        case EMH_MODEL_COROTATED:
        case EMH_MODEL_SNOW:
        case EMH_MODEL_DRUCKER_PRAGER:
        case EMH_MODEL_VON_MISES:
        case EMH_MODEL_RANKINE:
        case EMH_MODEL_MU_I:
        case EMH_MODEL_VISCOELASTIC:
        case EMH_MODEL_NACC:
        case EMH_MODEL_GRANULAR_FLUID:
            break;
        default:
            return fail(EMH_ERR_UNSUPPORTED, "constitutive model %u is not implemented on the GPU path", p.model);
    }
    if (p.model == EMH_MODEL_VISCOELASTIC && !(p.volume_ratio_min > 0.0f))
        return fail(EMH_ERR_INVALID_ARG, "viscoelastic j_min (volume_ratio_min) must be positive");
    if (p.active_stress_coeff != 0.0f &&
        (p.model == EMH_MODEL_NEOHOOKEAN || p.model == EMH_MODEL_COROTATED || p.model == EMH_MODEL_VISCOELASTIC))
        return fail(EMH_ERR_UNSUPPORTED, "active stress (active_stress_coeff != 0) is not supported");
    if (p.thermal_expansion != 0.0f && (p.model == EMH_MODEL_NEOHOOKEAN || p.model == EMH_MODEL_COROTATED))
        return fail(EMH_ERR_UNSUPPORTED, "thermal_expansion != 0 is not supported");
    if (p.thermal_viscosity_coeff != 0.0f && p.model == EMH_MODEL_FLUID)
        return fail(EMH_ERR_UNSUPPORTED, "thermal_viscosity_coeff != 0 is not supported");
    return EMH_OK;
}

// Device record of one material: the model's own fields recovered from the
// 96 B MaterialParams (each model's params(), see emerge_hip.h) plus the
// extension, and host-evaluated constants.
DevMat make_dev_mat(const emh_material_params& p, const emh_material_ext& e) {
    DevMat d;
    std::memset(&d, 0, sizeof(d));
    d.model = p.model;
    d.lambda = p.lambda;
    d.mu = p.mu;
    d.min_density = e.min_density;
    d.min_volume = e.min_volume;
    d.wave_lambda = p.lambda;
    switch (p.model) {
        case EMH_MODEL_NEOHOOKEAN:
            d.damage_rate = p.cohesion_coeff;
            d.nh_viscosity = p.dynamic_viscosity;
            break;
        case EMH_MODEL_FLUID:
            d.rest_density = p.rest_density;
            d.dyn_visc = p.dynamic_viscosity;
            d.eos_k = p.eos_stiffness;
            d.eos_p = p.eos_power;
            d.p_floor = p.pressure_floor;
            d.bulk_visc = p.bulk_viscosity;
            d.surf_tension = p.surface_tension_coeff;
            d.settling = p.dp_h0;
            d.fl_c2_denom = std::fmax(p.rest_density, e.min_density);
            // BinghamFluidMaterial::params reports model Fluid with the yield
            // stress in compression_limit (bingham.rs:205-218)
            if (p.compression_limit > 0.0f || e.fluid_bingham != 0u) d.flags |= kMatBingham;
            d.bing_yield = p.compression_limit;
            d.bing_crit = e.bingham_critical_shear_rate > 0.0f ? e.bingham_critical_shear_rate : 1.0e-4f;
            break;
        case EMH_MODEL_GRANULAR_FLUID:  // granular_fluid.rs:232-247
            d.rest_density = p.rest_density;
            d.eos_k = p.eos_stiffness;
            d.eos_p = p.eos_power;
            d.p_floor = p.pressure_floor;
            d.gf_rest_floor = std::fmax(p.rest_density, 1.0e-6f);
            d.sn_xi = p.hardening_exponent;
            d.sn_lo = 1.0f - p.compression_limit;
            d.sn_hi = 1.0f + p.stretch_limit;
            d.sn_jmin = p.volume_ratio_min;
            d.sn_jmax = p.volume_ratio_max;
            if (p.compression_limit > 0.0f || p.stretch_limit > 0.0f) d.flags |= kMatGfPlastic;
            break;
        case EMH_MODEL_RANKINE:  // rankine.rs:206-216; utils.rs:28,43-49
            d.rk_ts = p.hardening_exponent;
            d.rk_sr = p.hardening_modulus;
            d.rk_floor = p.hardening_exponent * 0.05f;
            d.rk_sat = p.hardening_modulus <= 0.0f ? std::numeric_limits<float>::infinity()
                                                   : -logf(0.05f) / p.hardening_modulus;
            break;
        case EMH_MODEL_MU_I:  // sand_mui.rs:219-229
            d.mui_s = p.dp_h0;
            d.mui_d = p.dp_h1;
            d.mui_q = p.dp_h2;
            break;
        case EMH_MODEL_DRUCKER_PRAGER:
            d.dp_phi0 = p.dp_h0;
            d.dp_h1 = p.dp_h1;
            d.dp_h2 = p.dp_h2;
            d.dp_h3 = p.dp_h3;
            d.dp_vol_corr = e.dp_volume_correction;
            d.dp_dilat = p.compression_limit;
            d.dp_sin_dilat = sinf(p.compression_limit);
            d.dp_cohesion_term = p.stretch_limit / (2.0f * p.mu);
            d.dp_ratio = (p.lambda + p.mu) / p.mu;
            d.dp_min_vj = p.volume_ratio_min;
            d.dp_qmax = 5.0f / std::fmax(p.dp_h2, 1e-6f);
            d.dp_sqrt23 = sqrtf(2.0f / 3.0f);
            break;
        case EMH_MODEL_SNOW:
            d.sn_xi = p.hardening_exponent;
            d.sn_lo = 1.0f - p.compression_limit;
            d.sn_hi = 1.0f + p.stretch_limit;
            d.sn_jmin = p.volume_ratio_min;
            d.sn_jmax = p.volume_ratio_max;
            d.sn_coh = p.cohesion_coeff;
            break;
        case EMH_MODEL_VON_MISES:  // von_mises.rs:143-152: yield stress in hardening_exponent
            d.vm_yield = p.hardening_exponent;
            d.vm_hmod = p.hardening_modulus;
            break;
        case EMH_MODEL_VISCOELASTIC:  // viscoelastic.rs:149-159
            d.ve_visc = p.dynamic_viscosity;
            d.ve_jmin = p.volume_ratio_min;
            break;
        case EMH_MODEL_NACC: {
            float kappa = e.nacc_kappa != 0.0f ? e.nacc_kappa : p.lambda + p.mu;
            d.na_kappa = kappa;
            d.na_m = p.stretch_limit;
            d.na_beta = p.compression_limit;
            d.na_xi = p.hardening_exponent;
            bool hard = e.nacc_hardening_enabled == 2u ? (p.hardening_exponent > 0.0f) : (e.nacc_hardening_enabled != 0u);
            d.flags = (hard ? kMatNaccHardening : 0u) | (e.nacc_advance_f ? kMatNaccAdvance : 0u);
            d.wave_lambda = kappa - p.mu;
            break;
        }
        default:
            break;
    }
    return d;
}

// init_particle: corotated.rs:57-60, sand.rs:237-246, snow.rs:104-107, nacc.rs:271-273,
// sand_mui.rs:146-149, granular_fluid.rs:166-169
void init_particle(const emh_material_params& p, emh_particle& q) {
    switch (p.model) {
        case EMH_MODEL_COROTATED:
            q.hardening_scale = 1.0f;
            q.plastic_volume_ratio = 1.0f;
            break;
        case EMH_MODEL_DRUCKER_PRAGER:
            q.friction_hardening = p.dp_h1 > 0.0f ? p.dp_h3 / p.dp_h1 : 0.0f;
            break;
        case EMH_MODEL_SNOW:
            q.plastic_volume_ratio = 1.0f;
            q.hardening_scale = 1.0f;
            break;
        case EMH_MODEL_NACC:
            q.log_volume_strain = 0.0f;
            break;
        case EMH_MODEL_MU_I:
            q.friction_hardening = p.dp_h0;
            break;
        case EMH_MODEL_GRANULAR_FLUID:
            q.plastic_volume_ratio = 1.0f;
            q.hardening_scale = 1.0f;
            break;
        default:
            break;
    }
}

struct LcgRng {  // solver/mod.rs:154-175
    uint32_t state;
    uint32_t next_u32() {
        state = state * 1664525u + 1013904223u;
        return state;
    }
    float next_f32() { return (float)next_u32() / ((float)4294967295u + 1.0f); }
};

// initialize_particles  solver/mod.rs:88-152
std::vector<emh_particle> lattice(const emh_config& cfg, const emh_spawn_region& sp) {
    std::vector<emh_particle> out;
    LcgRng rng{sp.rng_seed};
    const float mass = sp.has_mass_override ? sp.mass_override : cfg.particle_mass;
    const float hx = (float)sp.box_size[0] * 0.5f, hy = (float)sp.box_size[1] * 0.5f;
    const float x0 = sp.box_center[0] - hx, y0 = sp.box_center[1] - hy;
    const float x1 = sp.box_center[0] + hx, y1 = sp.box_center[1] + hy;
    for (float i = x0; i < x1; i += sp.spacing) {
        for (float j = y0; j < y1; j += sp.spacing) {
            if (sp.shape == EMH_SHAPE_DISK) {
                float dx = i - sp.box_center[0], dy = j - sp.box_center[1];
                if (!(sqrtf(dx * dx + dy * dy) <= sp.disk_radius)) continue;
            }
            const float jm = sp.position_jitter * sp.spacing;
            const float jx = (rng.next_f32() - 0.5f) * 2.0f * jm;
            const float jy = (rng.next_f32() - 0.5f) * 2.0f * jm;
            const float rx = rng.next_f32(), ry = rng.next_f32();
            emh_particle p;
            std::memset(&p, 0, sizeof(p));
            p.x[0] = i + jx;
            p.x[1] = j + jy;
            p.v[0] = (rx - 0.5f) * sp.initial_velocity_scale;
            p.v[1] = (ry - 0.5f) * sp.initial_velocity_scale;
            std::memcpy(p.deformation_gradient, sp.initial_deformation_gradient, sizeof(float) * 4);
            p.mass = mass;
            p.initial_volume = cfg.default_initial_volume;
            p.volume = cfg.default_initial_volume;
            p.density = mass / cfg.default_initial_volume;
            p.material_id = sp.material_id;
            p.plastic_volume_ratio = 1.0f;
            p.hardening_scale = 1.0f;
            out.push_back(p);
        }
    }
    return out;
}

bool spawn_fits(const emh_spawn_region& sp, const emh_config& cfg) {  // config.rs:472-482
    if (!(sp.spacing > 0.0f) || sp.box_size[0] <= 0 || sp.box_size[1] <= 0) return false;
    const float hx = (float)sp.box_size[0] * 0.5f, hy = (float)sp.box_size[1] * 0.5f;
    const float dmin = (float)cfg.boundary_thickness;
    const float dmax = (float)(cfg.grid_res > cfg.boundary_thickness ? cfg.grid_res - cfg.boundary_thickness : 0u);
    return sp.box_center[0] - hx >= dmin && sp.box_center[1] - hy >= dmin && sp.box_center[0] + hx <= dmax &&
           sp.box_center[1] + hy <= dmax;
}

template <class T> T* dalloc(size_t n) {
    void* p = nullptr;
    if (n == 0) n = 1;
    if (hipMalloc(&p, n * sizeof(T)) != hipSuccess) return nullptr;
    return (T*)p;
}

}  // namespace

// ------------------------------------------------------------------ emh_sim
// The settings the autotuners time: reorder intervals (substeps) x graph
// replay, x per-material segments on and off when some scene mixes models.
static std::vector<emh_tuning> tuning_candidates(const uint32_t* intervals, size_t n, bool mixed) {
    std::vector<emh_tuning> v;
    for (uint32_t seg = 0; seg < (mixed ? 2u : 1u); ++seg)
        for (uint32_t g = 0; g < 2; ++g)
            for (size_t i = 0; i < n; ++i) {
                emh_tuning t{};
                t.reorder_interval = intervals[i];
                t.graphs = g;
                t.model_segments = seg;
                v.push_back(t);
            }
    return v;
}
static std::vector<emh_tuning> tuning_candidates(bool mixed) {
    static const uint32_t kIntervals[] = {8u, 64u, 512u, kReorderSubsteps};
    return tuning_candidates(kIntervals, sizeof(kIntervals) / sizeof(kIntervals[0]), mixed);
}
static std::vector<emh_tuning> batch_tuning_candidates(bool mixed) {
    static const uint32_t kIntervals[] = {4u, kBatchReorderSubsteps, 32u, 256u};
    return tuning_candidates(kIntervals, sizeof(kIntervals) / sizeof(kIntervals[0]), mixed);
}

// Instantiated graphs of runs of substep launches, keyed by (launch number
// mod 6, launches, kernel variant); valid for one kernel argument (`pipe`).
struct GraphCache {
    struct Entry {
        uint32_t r, slots, variant;
        hipGraphExec_t exec;
    };
    static constexpr size_t kMax = 24;
    std::vector<Entry> entries;
    std::vector<unsigned char> arg;  // the argument bytes the entries captured

    ~GraphCache() { clear(); }
    void clear() {
        for (auto& e : entries) (void)hipGraphExecDestroy(e.exec);
        entries.clear();
    }
    // Drops every entry when the argument differs from the captured one.
    void check_arg(const void* a, size_t n) {
        if (arg.size() == n && std::memcmp(arg.data(), a, n) == 0) return;
        clear();
        arg.assign((const unsigned char*)a, (const unsigned char*)a + n);
    }
    hipGraphExec_t find(uint32_t r, uint32_t slots, uint32_t variant) const {
        for (const auto& e : entries)
            if (e.r == r && e.slots == slots && e.variant == variant) return e.exec;
        return nullptr;
    }
    // Captures `launch(k)` for k < slots on `stream` into a new entry.
    template <class F>
    hipError_t capture(hipStream_t stream, uint32_t r, uint32_t slots, uint32_t variant, F&& launch,
                       hipGraphExec_t* out) {
        *out = nullptr;
        if (entries.size() >= kMax) {
            (void)hipGraphExecDestroy(entries.front().exec);
            entries.erase(entries.begin());
        }
        hipError_t e = hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) return e;
        for (uint32_t k = 0; k < slots && e == hipSuccess; ++k) e = launch(k);
        hipGraph_t g = nullptr;
        const hipError_t e2 = hipStreamEndCapture(stream, &g);
        if (e == hipSuccess) e = e2;
        if (e == hipSuccess) e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
        if (g) (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
            *out = nullptr;
            return e;
        }
        entries.push_back(Entry{r, slots, variant, *out});
        return hipSuccess;
    }
};

struct emh_sim {
    emh_config cfg{};
    std::vector<emh_material_params> mats;
    std::vector<emh_material_ext> exts;
    struct HostBoundary {  // one installed BoundaryCondition (boundary/*.rs)
        uint32_t kind, thickness;
        float mu, wall_min;
        std::vector<float> heights;
    };
    std::vector<HostBoundary> bnds;
    float* d_heights = nullptr;   // terrain heights of the heightmap boundaries
    size_t heights_cap = 0;
    emh_particle* d_cold = nullptr;  // the last upload's 128 B records, caller order (export)
    size_t cold_cap = 0;
    bool cold_valid = false;
    std::vector<emh_particle> host;  // authoritative when !device_ahead
    uint32_t next_tag = 1;
    bool host_dirty = true;     // host changed since the last upload
    bool device_ahead = false;  // device stepped since the last download
    emh_step_stats stats{};
    uint32_t predicted_substeps = 1;
    uint32_t substeps_since_reorder = 0;
    uint32_t reorder_substeps = kReorderSubsteps;  // emh_sim_set_reorder_interval
    bool reorder_pending = false;                  // an upload: the next step reorders
    // One-shot upload order (autotune's restore): slot s of the next upload
    // takes host[upload_order[s]], and the reorder is pending only if it was
    // when the order was saved, so a restored scene keeps its stored order.
    std::vector<uint32_t> upload_order;
    bool upload_order_pending = false;
    uint64_t seq = 0;  // sequence number of the next control launch (see Ctrl)
    // learned NACC return mapping: 0 plain, 1 harvest trial states, 2 infer
    uint32_t sg_mode = 0;
    emh_surrogate* sg_attached = nullptr;  // the surrogate of modes 1 and 2
    uint64_t cfg_epoch = 0;                // bumped by in-place edits of pipe.cfg (set_gravity)
    bool model_segments = false;  // emh_tuning::model_segments
    // hipGraph replay of a step's substep launches (emh_sim_set_graphs): the
    // kernels read their launch number only modulo 6 (control-state parity
    // and reduction slot), so a run of `slots` launches captured at residue r
    // replays for every later run at that residue while the pipeline (the
    // launches' argument) is unchanged.
    bool use_graphs = false;
    GraphCache graphs;
    // sleep partition: on when sleep_threshold > 0 or an uploaded particle sleeps
    bool sleep = false;
    float sleep_sq() const {
        return cfg.sleep_threshold > 0.0f ? cfg.sleep_threshold * cfg.sleep_threshold : 0.0f;
    }
    SgDev sg_host{};  // the descriptor Ctrl::sg holds

    // Live timing: event marks on the library stream, recorded without
    // stopping the queue -- before each step's set-up launches (k_begin_step
    // and, on reorder steps, the key / scan kernels), before its substep
    // launches, and where a drain stops to read the control block (and around
    // the extra substep launches a drain makes).  A mark's kind says what the
    // span up to the next mark holds; a drain folds the spans once its stream
    // synchronisation has completed every mark.  The substep launches run back
    // to back, so their spans divided by the launches is the substep kernel's
    // average duration (inter-kernel gaps included, no host wait).
    enum : uint8_t { kMarkSetup = 0, kMarkSubsteps = 1, kMarkIdle = 2 };
    bool profiling = false;
    std::vector<hipEvent_t> prof_pool;
    std::vector<uint8_t> prof_kind;
    size_t prof_used = 0;
    double prof_ms[2] = {0.0, 0.0};       // set-up, substep kernels
    uint64_t prof_count[2] = {0, 0};      // steps, substeps taken
    uint32_t prof_sub_base = 0;           // total_substeps at the last fold

    // device
    int device = 0;
    hipStream_t stream = nullptr;
    Pipeline pipe{};
    uint32_t cap = 0;  // particle capacity of the device buffers
    DevMat* d_mats = nullptr;
    void* grid_block = nullptr;
    Ctrl* d_ctrl = nullptr;
    DevSnapshot* d_snap = nullptr;     // diagnostics scratch, allocated on first use
    uint32_t* d_snap_cells = nullptr;  // 3 R^2 per-cell material ranges and counts
    std::vector<void*> particle_allocs;

    ~emh_sim();

    void release_particles() {
        for (void* p : particle_allocs) (void)hipFree(p);
        particle_allocs.clear();
        cap = 0;
    }
    void release_events() {
        for (auto& e : prof_pool) (void)hipEventDestroy(e);
        prof_pool.clear();
        prof_kind.clear();
        prof_used = 0;
    }
    // Record a mark of `kind` on the stream (no-op unless profiling).
    int prof_mark(uint8_t kind) {
        if (!profiling) return EMH_OK;
        if (prof_used == prof_pool.size()) {
            // timing only (read after a stream synchronisation): no system-scope
            // fence, whose cache writeback would slow the kernels that follow
            hipEvent_t e = nullptr;
            HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
            prof_pool.push_back(e);
            prof_kind.push_back(kMarkIdle);
        }
        HIP_TRY(hipEventRecord(prof_pool[prof_used], stream));
        prof_kind[prof_used++] = kind;
        return EMH_OK;
    }
    // Fold the recorded spans into the totals; every mark must have completed
    // (the caller has synchronised the stream after the last one).
    int prof_fold() {
        for (size_t i = 0; i + 1 < prof_used; ++i) {
            if (prof_kind[i] == kMarkIdle) continue;
            float ms = 0.0f;
            HIP_TRY(hipEventElapsedTime(&ms, prof_pool[i], prof_pool[i + 1]));
            prof_ms[prof_kind[i] == kMarkSetup ? 0 : 1] += ms;
        }
        prof_used = 0;
        return EMH_OK;
    }

    void release() {
        if (stream) (void)hipStreamSynchronize(stream);
        graphs.clear();
        release_events();
        release_particles();
        if (d_mats) (void)hipFree(d_mats);
        if (d_heights) (void)hipFree(d_heights);
        if (d_cold) (void)hipFree(d_cold);
        if (grid_block) (void)hipFree(grid_block);
        if (d_ctrl) (void)hipFree(d_ctrl);
        if (d_snap) (void)hipFree(d_snap);
        if (d_snap_cells) (void)hipFree(d_snap_cells);
        d_snap = nullptr;
        d_snap_cells = nullptr;
        if (stream) (void)hipStreamDestroy(stream);
        d_mats = nullptr;
        d_heights = nullptr;
        heights_cap = 0;
        d_cold = nullptr;
        cold_cap = 0;
        cold_valid = false;
        grid_block = nullptr;
        d_ctrl = nullptr;
        stream = nullptr;
    }

    uint32_t res() const { return cfg.grid_res; }
    uint32_t key_order() const { return cell_key_order(cfg.grid_res); }
    uint32_t n_bins() const { return 1u << (2u * key_order()); }

    int init_device() {
        if (device_present() == 0) return fail(EMH_ERR_DEVICE, "no HIP device is available");
        HIP_TRY(hipGetDevice(&device));
        HIP_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        const size_t R2 = (size_t)res() * res();
        // grid: three sets of 5 P2G accumulators + three density-pass mass grids
        HIP_TRY(hipMalloc(&grid_block, sizeof(float) * R2 * 3 * (kAccChannels + 1)));
        float* g = (float*)grid_block;
        for (int set = 0; set < 3; ++set) {
            for (int ch = 0; ch < kAccChannels; ++ch) pipe.grid.acc[set][ch] = g + (size_t)(set * kAccChannels + ch) * R2;
            pipe.grid.dm[set] = g + (size_t)(3 * kAccChannels + set) * R2;
        }
        d_ctrl = dalloc<Ctrl>(1);
        d_mats = dalloc<DevMat>(EMH_MAX_MATERIALS);
        if (!d_ctrl || !d_mats) return fail(EMH_ERR_OUT_OF_MEMORY, "device allocation failed");
        pipe.ctrl = d_ctrl;
        pipe.mats = d_mats;
        return EMH_OK;
    }

    int ensure_capacity(uint32_t n) {
        if (n <= cap && cap > 0) return EMH_OK;
        HIP_TRY(hipStreamSynchronize(stream));
        release_particles();
        uint32_t c = std::max<uint32_t>(n, 1024u);
        auto take = [&](size_t bytes) -> void* {
            void* p = nullptr;
            if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
            particle_allocs.push_back(p);
            return p;
        };
        for (int b = 0; b < 2; ++b) {
            PBuf& pb = pipe.buf[b];
            pb.x = (float2*)take(sizeof(float2) * c);
            pb.v = (float2*)take(sizeof(float2) * c);
            pb.c = (float4*)take(sizeof(float4) * c);
            pb.f = (float4*)take(sizeof(float4) * c);
            pb.mass = (float*)take(sizeof(float) * c);
            pb.v0 = (float*)take(sizeof(float) * c);
            pb.vol = (float*)take(sizeof(float) * c);
            pb.dens = (float*)take(sizeof(float) * c);
            pb.jp = (float*)take(sizeof(float) * c);
            pb.h = (float*)take(sizeof(float) * c);
            pb.q = (float*)take(sizeof(float) * c);
            pb.lvs = (float*)take(sizeof(float) * c);
            pb.mat = (uint32_t*)take(sizeof(uint32_t) * c);
            pb.id = (uint32_t*)take(sizeof(uint32_t) * c);
        }
        // the reorder's radix sort; nothing reads its input-order keys, so
        // pass 0 does not write them (key = null)
        pipe.sort.key = nullptr;
        pipe.sort.kbuf[0] = (uint32_t*)take(sizeof(uint32_t) * c);
        pipe.sort.kbuf[1] = (uint32_t*)take(sizeof(uint32_t) * c);
        pipe.sort.vbuf[0] = (uint32_t*)take(sizeof(uint32_t) * c);
        pipe.sort.vbuf[1] = (uint32_t*)take(sizeof(uint32_t) * c);
        pipe.sort.out = (uint32_t*)take(sizeof(uint32_t) * c);
        pipe.sort.hist = (uint32_t*)take(sizeof(uint32_t) * sort_hist_len_max(c));
        pipe.sort.perm = 0u;
        for (void* p : particle_allocs)
            if (!p) {
                release_particles();
                return fail(EMH_ERR_OUT_OF_MEMORY, "device allocation of %u particles failed", c);
            }
        cap = c;
        return EMH_OK;
    }

    void fill_dev_config() {
        DevConfig& d = pipe.cfg;
        std::memset(&d, 0, sizeof(d));
        d.grid_res = cfg.grid_res;
        d.cell = cfg.grid_cell_size;
        d.dt = cfg.dt;
        d.adaptive = cfg.adaptive_timestep;
        d.cfl_affine = cfg.cfl_include_affine_speed;
        d.cfl_coefficient = cfg.cfl_coefficient;
        d.material_cfl = cfg.material_cfl_coefficient;
        d.viscous_cfl = cfg.viscous_timestep_coefficient;
        d.min_dt = cfg.min_dt;
        d.project = cfg.project_invalid_state;
        d.proj_min_density = cfg.projection_min_density;
        d.proj_min_volume = cfg.projection_min_volume;
        d.proj_min_j = cfg.projection_min_deformation_j;
        d.gx = cfg.gravity[0];
        d.gy = cfg.gravity[1];
        d.bt = cfg.boundary_thickness;
        d.default_initial_volume = cfg.default_initial_volume;
        d.particle_mass = cfg.particle_mass;
        d.max_substeps = cfg.max_substeps_per_step;
        d.apic_blend = cfg.apic_blend;
        d.j_max = cfg.j_max;
        d.n_bnd = (uint32_t)bnds.size();
        uint32_t hoff = 0;
        for (size_t i = 0; i < bnds.size(); ++i) {
            const HostBoundary& b = bnds[i];
            d.bnd_kind[i] = b.kind;
            d.bnd_t[i] = b.thickness;
            d.bnd_mu[i] = b.mu;
            if (b.kind == kBndPredictive) {  // predictive.rs:31-35
                d.bnd_lo[i] = b.wall_min;
                d.bnd_hi[i] = ((float)cfg.grid_res - 1.0f) - b.wall_min;
            } else {  // clamp_position_inside_grid  boundary/mod.rs:103-111
                d.bnd_lo[i] = (float)(b.thickness > 0 ? b.thickness - 1 : 0u);
                d.bnd_hi[i] = (float)(cfg.grid_res > b.thickness ? cfg.grid_res - b.thickness : 0u);
            }
            d.bnd_hoff[i] = hoff;
            d.bnd_hlen[i] = (uint32_t)b.heights.size();
            hoff += (uint32_t)b.heights.size();
        }
        d.heights = d_heights;
        const uint32_t bt = cfg.boundary_thickness;
        d.proj_lo = (float)(bt > 0 ? bt - 1 : 0u);
        d.proj_hi = (float)(cfg.grid_res > bt ? cfg.grid_res - bt : 0u);
        d.key_order = key_order();
        d.n_bins = n_bins();
        bool dens = cfg.recompute_density_each_step != 0;
        for (auto& m : mats)
            if (m.model == EMH_MODEL_FLUID) dens = true;  // needs_density_recompute  fluid.rs:261-263
        d.density_pass = dens ? 1u : 0u;
        pipe.n_mats = (uint32_t)mats.size();
        // optional fields: what the materials the particles resolve to use
        std::vector<bool> used(mats.size(), false);
        for (const auto& p : host) {
            const uint32_t m = p.material_id > kMatMask ? kMatMask : p.material_id;
            used[m < mats.size() ? m : 0u] = true;
        }
        uint32_t fields = 0, only = kModelMixed;
        bool first = true;
        for (size_t i = 0; i < mats.size(); ++i) {
            if (!used[i]) continue;
            if (first)
                only = mats[i].model;
            else if (only != mats[i].model)
                only = kModelMixed;
            first = false;
            fields |= model_fields(mats[i].model);  // fluid, fallback: volume/density carried over
        }
        d.fields = fields;
        d.only_model = only;
        // per-material segments (a tuning option): a scene of several models is
        // reordered by (model, cell key), so a chunk -- and each wave -- runs
        // one model
        pipe.sort.model_bits = (only == kModelMixed && model_segments) ? kModelKeyBits : 0u;
    }

    // Terrain heights of every heightmap boundary, concatenated in boundary order.
    int upload_heights() {
        std::vector<float> all;
        for (const HostBoundary& b : bnds) all.insert(all.end(), b.heights.begin(), b.heights.end());
        if (all.size() > heights_cap) {
            if (d_heights) (void)hipFree(d_heights);
            d_heights = nullptr;
            heights_cap = 0;
            if (hipMalloc((void**)&d_heights, all.size() * sizeof(float)) != hipSuccess) {
                d_heights = nullptr;
                return fail(EMH_ERR_OUT_OF_MEMORY, "heightmap of %zu heights", all.size());
            }
            heights_cap = all.size();
        }
        if (!all.empty()) {
            HIP_TRY(hipMemcpyAsync(d_heights, all.data(), all.size() * sizeof(float), hipMemcpyHostToDevice, stream));
            HIP_TRY(hipStreamSynchronize(stream));
        }
        return EMH_OK;
    }

    int upload_materials() {
        std::vector<DevMat> dm(mats.size());
        for (size_t i = 0; i < mats.size(); ++i) dm[i] = make_dev_mat(mats[i], exts[i]);
        HIP_TRY(hipMemcpyAsync(d_mats, dm.data(), sizeof(DevMat) * dm.size(), hipMemcpyHostToDevice, stream));
        HIP_TRY(hipStreamSynchronize(stream));
        return EMH_OK;
    }

    // Host AoS -> device SoA (buffer 0, caller order), then keys/ranks/CFL.
    int upload() {
        for (auto& p : host) {
            if (p.contact_group != 0)
                return fail(EMH_ERR_UNSUPPORTED, "particles with contact_group != 0 (multi-field contact) are not supported");
        }
        const uint32_t n = (uint32_t)host.size();
        int rc = ensure_capacity(n);
        if (rc) return rc;
        rc = upload_heights();
        if (rc) return rc;
        fill_dev_config();
        rc = upload_materials();
        if (rc) return rc;
        pipe.n = n;
        std::vector<float2> x(n), v(n);
        std::vector<float4> c(n), f(n);
        std::vector<float> mass(n), v0(n), vol(n), dens(n), jp(n), h(n), q(n), lvs(n);
        std::vector<uint32_t> mat(n), id(n);
        bool any_sleeping = false;
        const bool keep_order = upload_order.size() == n;
        for (uint32_t i = 0; i < n; ++i) {
            const uint32_t src = keep_order ? upload_order[i] : i;
            const emh_particle& p = host[src];
            x[i] = make_float2(p.x[0], p.x[1]);
            v[i] = make_float2(p.v[0], p.v[1]);
            c[i] = make_float4(p.velocity_gradient[0], p.velocity_gradient[1], p.velocity_gradient[2],
                               p.velocity_gradient[3]);
            f[i] = make_float4(p.deformation_gradient[0], p.deformation_gradient[1], p.deformation_gradient[2],
                               p.deformation_gradient[3]);
            mass[i] = p.mass;
            v0[i] = p.initial_volume;
            vol[i] = p.volume;
            dens[i] = p.density;
            jp[i] = p.plastic_volume_ratio;
            h[i] = p.hardening_scale;
            q[i] = p.friction_hardening;
            lvs[i] = p.log_volume_strain;
            mat[i] = (p.material_id & kMatMask) | (p.pinned ? kPinnedBit : 0u);
            if (p.material_id > kMatMask) mat[i] = kMatMask;  // unregistered: falls back to slot 0
            if (p.sleeping) mat[i] |= kSleepBit;
            if (p.activation != 0.0f) mat[i] |= kNoSleepBit;
            any_sleeping = any_sleeping || p.sleeping != 0;
            id[i] = src;
        }
        const PBuf& b = pipe.buf[0];
        auto put = [&](void* dst, const void* src, size_t bytes) -> hipError_t {
            return bytes ? hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream) : hipSuccess;
        };
        HIP_TRY(put(b.x, x.data(), sizeof(float2) * n));
        HIP_TRY(put(b.v, v.data(), sizeof(float2) * n));
        HIP_TRY(put(b.c, c.data(), sizeof(float4) * n));
        HIP_TRY(put(b.f, f.data(), sizeof(float4) * n));
        HIP_TRY(put(b.mass, mass.data(), sizeof(float) * n));
        HIP_TRY(put(b.v0, v0.data(), sizeof(float) * n));
        HIP_TRY(put(b.vol, vol.data(), sizeof(float) * n));
        HIP_TRY(put(b.dens, dens.data(), sizeof(float) * n));
        HIP_TRY(put(b.jp, jp.data(), sizeof(float) * n));
        HIP_TRY(put(b.h, h.data(), sizeof(float) * n));
        HIP_TRY(put(b.q, q.data(), sizeof(float) * n));
        HIP_TRY(put(b.lvs, lvs.data(), sizeof(float) * n));
        HIP_TRY(put(b.mat, mat.data(), sizeof(uint32_t) * n));
        sleep = cfg.sleep_threshold > 0.0f || any_sleeping;
        HIP_TRY(put(b.id, id.data(), sizeof(uint32_t) * n));
        Ctrl ctrl;
        std::memset(&ctrl, 0, sizeof(ctrl));
        ctrl.st[0].total_substeps = stats.total_substeps;
        ctrl.st[0].last_dt = stats.last_step_dt;
        ctrl.st[0].dt = stats.last_step_dt;
        ctrl.st[0].free_dt = cfg.dt;
        for (auto& r : ctrl.red) {
            r.max_speed = 0u;
            r.min_mdt = kInfBits;
            r.dirty = 0u;
        }
        ctrl.sg = sg_host;
        ctrl.sleep_sq = sleep_sq();
        seq = 0;
        HIP_TRY(hipMemcpyAsync(d_ctrl, &ctrl, sizeof(Ctrl), hipMemcpyHostToDevice, stream));
        HIP_TRY(launch_zero_all(pipe, stream));
        HIP_TRY(launch_prologue(pipe, seq++, stream, sleep));
        HIP_TRY(hipStreamSynchronize(stream));
        // the first step puts the upload in cell-key order (unless it restores
        // a saved stored order)
        reorder_pending = keep_order ? upload_order_pending : true;
        upload_order.clear();
        host_dirty = false;
        device_ahead = false;
        cold_valid = false;
        return EMH_OK;
    }

    // The control state the last launch left (st[seq & 1]).
    int read_ctrl(CtrlState& st) {
        Ctrl ctrl;
        HIP_TRY(hipMemcpyAsync(&ctrl, d_ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipStreamSynchronize(stream));
        st = ctrl.st[seq & 1u];
        return EMH_OK;
    }

    // After a device-side edit of the stored particles between steps: the
    // loop's pending CFL reduction and the next substep's P2G were formed
    // from the old state, so restart the loop from the stored particles in
    // buffer `cur_buf` as an upload does (no reorder is forced).
    int restart_loop(uint32_t cur_buf) {
        Ctrl ctrl;
        std::memset(&ctrl, 0, sizeof(ctrl));
        ctrl.st[0].total_substeps = stats.total_substeps;
        ctrl.st[0].last_dt = stats.last_step_dt;
        ctrl.st[0].dt = stats.last_step_dt;
        ctrl.st[0].free_dt = cfg.dt;
        ctrl.st[0].cur_buf = cur_buf & 1u;
        for (auto& r : ctrl.red) {
            r.max_speed = 0u;
            r.min_mdt = kInfBits;
            r.dirty = 0u;
        }
        ctrl.sg = sg_host;
        ctrl.sleep_sq = sleep_sq();
        seq = 0;
        HIP_TRY(hipMemcpyAsync(d_ctrl, &ctrl, sizeof(Ctrl), hipMemcpyHostToDevice, stream));
        HIP_TRY(launch_zero_all(pipe, stream));
        HIP_TRY(launch_prologue(pipe, seq++, stream, sleep));
        HIP_TRY(hipStreamSynchronize(stream));
        device_ahead = true;
        return EMH_OK;
    }

    // Simulation::apply_impulse / apply_radial_impulse  solver/particles.rs:115-165
    int impulse(float cx, float cy, float radius, float fx, float fy, float strength, uint32_t radial) {
        if (host_dirty) {
            int rc = upload();
            if (rc) return rc;
        }
        if (pipe.n == 0) return EMH_OK;
        CtrlState st;
        int rc = read_ctrl(st);
        if (rc) return rc;
        const float vel_limit = cfg.grid_cell_size / cfg.min_dt;
        HIP_TRY(launch_impulse(pipe.buf[st.cur_buf & 1u], pipe.n, cx, cy, radius, fx, fy, strength, radial, vel_limit,
                               stream));
        return restart_loop(st.cur_buf);
    }

    // Time each tuning candidate on this scene: from the same saved state,
    // one untimed step (upload, graph capture), then `trial` steps timed on
    // the host clock up to their completion.  The state is put back afterwards
    // (re-uploaded from the saved copy at the next step in the stored order it
    // had, stats restored) and the fastest candidate applied.  Keeping the
    // stored order matters: a freshly reordered state of the 1M-particle sand
    // column runs later substeps up to ~8 us slower than the order it replaced,
    // so a restore that reordered would both bias the trials and slow the run.
    int autotune(uint32_t trial, emh_tuning* best, emh_tuning* timed, uint32_t cap, uint32_t* n_timed) {
        if (trial == 0) return fail(EMH_ERR_INVALID_ARG, "trial_steps must be at least 1");
        if (profiling) return fail(EMH_ERR_STATE, "turn profiling off before autotuning");
        if (pipe.n == 0 && host.empty()) return fail(EMH_ERR_STATE, "no particles to tune on");
        if (sg_mode != 0)
            return fail(EMH_ERR_STATE, "switch the surrogate mode to plain before autotuning (the trial steps "
                                       "would add harvest records or counts to the surrogate)");
        int rc = host_current();
        if (rc) return rc;
        const std::vector<emh_particle> saved = host;
        std::vector<uint32_t> saved_order;  // caller index of each slot, when on the device
        const bool saved_pending = reorder_pending;
        if (!host_dirty && pipe.n == (uint32_t)host.size() && pipe.n > 0) {
            saved_order.resize(pipe.n);
            rc = slot_order(saved_order.data());
            if (rc) return rc;
        }
        const emh_step_stats saved_stats = stats;
        const uint32_t saved_since = substeps_since_reorder, saved_pred = predicted_substeps;
        const emh_tuning saved_tuning = tuning();
        if (host_dirty) {  // the material mix (mixed_models) is known after an upload
            rc = upload();
            if (rc) return rc;
        }
        const std::vector<emh_tuning> cands = tuning_candidates(mixed_models());
        std::vector<float> us(cands.size(), 0.0f);
        // a candidate whose reorder key differs from the one the stored order
        // was sorted by (per-material segments) starts with a reorder
        auto restore = [&](bool rekey) {
            host = saved;
            upload_order = saved_order;
            upload_order_pending = saved_pending || rekey;
            host_dirty = true;
            device_ahead = false;  // the host copy, not the last trial's device state, is current
            stats = saved_stats;
            substeps_since_reorder = saved_since;
            predicted_substeps = saved_pred;
        };
        for (size_t c = 0; c < cands.size() && rc == EMH_OK; ++c) {
            restore(cands[c].model_segments != saved_tuning.model_segments);
            apply_tuning(cands[c]);
            rc = step(1);
            if (rc) break;
            const uint64_t s0 = stats.total_substeps;
            const auto t0 = std::chrono::steady_clock::now();
            rc = step(trial);
            const auto t1 = std::chrono::steady_clock::now();
            const double work = (double)std::max<uint64_t>(stats.total_substeps - s0, 1) * (double)pipe.n;
            us[c] = (float)(std::chrono::duration<double, std::nano>(t1 - t0).count() / work);
        }
        size_t bi = 0;
        for (size_t c = 1; c < cands.size(); ++c)
            if (us[c] < us[bi]) bi = c;
        restore(rc == EMH_OK && cands[bi].model_segments != saved_tuning.model_segments);
        apply_tuning(saved_tuning);
        if (rc) return rc;
        for (size_t c = 0; c < cands.size() && c < cap && timed; ++c) {
            timed[c] = cands[c];
            timed[c].ns_per_particle_substep = us[c];
        }
        if (n_timed) *n_timed = (uint32_t)cands.size();
        apply_tuning(cands[bi]);
        if (best) {
            *best = cands[bi];
            best->ns_per_particle_substep = us[bi];
        }
        return EMH_OK;
    }

    // Launch settings (emh_tuning); the model grouping of the reorder key
    // takes effect at the next reorder.
    void apply_tuning(const emh_tuning& t) {
        reorder_substeps = t.reorder_interval;
        use_graphs = t.graphs != 0;
        model_segments = t.model_segments != 0;
        pipe.sort.model_bits = (pipe.cfg.only_model == kModelMixed && model_segments) ? kModelKeyBits : 0u;
        ++cfg_epoch;
    }
    emh_tuning tuning() const {
        emh_tuning t{};
        t.reorder_interval = reorder_substeps;
        t.graphs = use_graphs ? 1u : 0u;
        t.model_segments = model_segments ? 1u : 0u;
        return t;
    }
    bool mixed_models() const { return pipe.cfg.only_model == kModelMixed; }

    // The caller index of the particle in each device slot (the stored order).
    int slot_order(uint32_t* out) {
        if (host_dirty) return fail(EMH_ERR_STATE, "the particles are not on the device yet: step first");
        if (pipe.n == 0) return EMH_OK;
        CtrlState st;
        int rc = read_ctrl(st);
        if (rc) return rc;
        HIP_TRY(hipMemcpyAsync(out, pipe.buf[st.cur_buf & 1u].id, sizeof(uint32_t) * pipe.n, hipMemcpyDeviceToHost,
                               stream));
        HIP_TRY(hipStreamSynchronize(stream));
        return EMH_OK;
    }

    // Particle records in caller order into a device buffer of `cap` records.
    int export_particles(void* dst, size_t cap) {
        if (host_dirty) {
            int rc = upload();
            if (rc) return rc;
        }
        const uint32_t n = pipe.n;
        if (cap < n) return fail(EMH_ERR_INVALID_ARG, "buffer holds %zu particles, %u needed", cap, n);
        if (n == 0) return EMH_OK;
        if (!cold_valid) {
            if (cold_cap < n) {
                if (d_cold) (void)hipFree(d_cold);
                d_cold = nullptr;
                cold_cap = 0;
                if (hipMalloc((void**)&d_cold, sizeof(emh_particle) * n) != hipSuccess) {
                    d_cold = nullptr;
                    return fail(EMH_ERR_OUT_OF_MEMORY, "export staging of %u particles", n);
                }
                cold_cap = n;
            }
            HIP_TRY(hipMemcpyAsync(d_cold, host.data(), sizeof(emh_particle) * n, hipMemcpyHostToDevice, stream));
            cold_valid = true;
        }
        CtrlState st;
        int rc = read_ctrl(st);
        if (rc) return rc;
        HIP_TRY(launch_export_particles(pipe.buf[st.cur_buf & 1u], n, d_cold, (emh_particle*)dst, stream));
        HIP_TRY(hipStreamSynchronize(stream));
        return EMH_OK;
    }

    // Device SoA (current buffer, sorted order) -> host AoS in caller order.
    int download() {
        if (!device_ahead) return EMH_OK;
        const uint32_t n = pipe.n;
        Ctrl ctrl;
        HIP_TRY(hipMemcpyAsync(&ctrl, d_ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipStreamSynchronize(stream));
        const PBuf& b = pipe.buf[ctrl.st[seq & 1u].cur_buf & 1u];
        std::vector<float2> x(n), v(n);
        std::vector<float4> c(n), f(n);
        std::vector<float> mass(n), v0(n), vol(n), dens(n), jp(n), h(n), q(n), lvs(n);
        std::vector<uint32_t> id(n);
        auto get = [&](void* dst, const void* src, size_t bytes) -> hipError_t {
            return bytes ? hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream) : hipSuccess;
        };
        HIP_TRY(get(x.data(), b.x, sizeof(float2) * n));
        HIP_TRY(get(v.data(), b.v, sizeof(float2) * n));
        HIP_TRY(get(c.data(), b.c, sizeof(float4) * n));
        HIP_TRY(get(f.data(), b.f, sizeof(float4) * n));
        HIP_TRY(get(mass.data(), b.mass, sizeof(float) * n));
        HIP_TRY(get(v0.data(), b.v0, sizeof(float) * n));
        HIP_TRY(get(vol.data(), b.vol, sizeof(float) * n));
        HIP_TRY(get(dens.data(), b.dens, sizeof(float) * n));
        HIP_TRY(get(jp.data(), b.jp, sizeof(float) * n));
        HIP_TRY(get(h.data(), b.h, sizeof(float) * n));
        HIP_TRY(get(q.data(), b.q, sizeof(float) * n));
        HIP_TRY(get(lvs.data(), b.lvs, sizeof(float) * n));
        HIP_TRY(get(id.data(), b.id, sizeof(uint32_t) * n));
        std::vector<uint32_t> mat;
        if (sleep) {
            mat.resize(n);
            HIP_TRY(get(mat.data(), b.mat, sizeof(uint32_t) * n));
        }
        HIP_TRY(hipStreamSynchronize(stream));
        for (uint32_t k = 0; k < n; ++k) {
            const uint32_t i = id[k];
            if (i >= n) return fail(EMH_ERR_DEVICE, "corrupt particle id %u at slot %u", i, k);
            emh_particle& p = host[i];
            p.x[0] = x[k].x;
            p.x[1] = x[k].y;
            p.v[0] = v[k].x;
            p.v[1] = v[k].y;
            p.velocity_gradient[0] = c[k].x;
            p.velocity_gradient[1] = c[k].y;
            p.velocity_gradient[2] = c[k].z;
            p.velocity_gradient[3] = c[k].w;
            p.deformation_gradient[0] = f[k].x;
            p.deformation_gradient[1] = f[k].y;
            p.deformation_gradient[2] = f[k].z;
            p.deformation_gradient[3] = f[k].w;
            p.mass = mass[k];
            p.initial_volume = v0[k];
            p.volume = vol[k];
            p.density = dens[k];
            p.plastic_volume_ratio = jp[k];
            p.hardening_scale = h[k];
            p.friction_hardening = q[k];
            p.log_volume_strain = lvs[k];
            if (sleep) p.sleeping = (mat[k] & kSleepBit) ? 1u : 0u;
        }
        device_ahead = false;
        return EMH_OK;
    }

    int host_current() { return download(); }

    // `slots` back-to-back substep launches from launch number seq, replayed
    // from a captured graph in graph mode.
    int launch_substep_run(uint32_t slots) {
        if (!use_graphs) {
            for (uint32_t k = 0; k < slots; ++k) HIP_TRY(launch_substep(pipe, seq++, stream, sg_mode, sleep));
            return EMH_OK;
        }
        graphs.check_arg(&pipe, sizeof(Pipeline));
        const uint32_t r = (uint32_t)(seq % 6u), variant = sg_mode | (sleep ? 4u : 0u);
        hipGraphExec_t ex = graphs.find(r, slots, variant);
        if (!ex) {
            const uint64_t s0 = seq;
            HIP_TRY(graphs.capture(
                stream, r, slots, variant,
                [&](uint32_t k) { return launch_substep(pipe, s0 + k, stream, sg_mode, sleep); }, &ex));
        }
        HIP_TRY(hipGraphLaunch(ex, stream));
        seq += slots;
        return EMH_OK;
    }

    // Running per the loop condition of step.rs:34.
    bool running(const CtrlState& st) const {
        return st.in_step && st.remaining > kF32Eps && st.substeps < cfg.max_substeps_per_step;
    }
    // Wait until every queued step has finished, launching further substep
    // slots while a step is still running or queued; then publish the last
    // step's records (step.rs:51-59).
    int drain() {
        Ctrl c;
        CtrlState st;
        uint32_t extra = 0;
        const uint32_t limit = (cfg.max_substeps_per_step + 1) * (kStepRing + 1);
        while (true) {
            {
                int rc = prof_mark(kMarkIdle);
                if (rc) return rc;
            }
            HIP_TRY(hipMemcpyAsync(&c, d_ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, stream));
            HIP_TRY(hipStreamSynchronize(stream));
            st = c.st[seq & 1u];  // the state the next launch would decide from
            if (!running(st) && st.pending == 0) break;
            uint32_t more = (running(st) ? 1u : 0u) + st.pending * std::max<uint32_t>(predicted_substeps, 1u);
            more = std::min<uint32_t>(std::max<uint32_t>(more, 1u), 64u);
            if (extra + more > limit) return fail(EMH_ERR_DEVICE, "substep loop did not terminate");
            {
                int rc = prof_mark(kMarkSubsteps);
                if (rc) return rc;
            }
            for (uint32_t k = 0; k < more; ++k) HIP_TRY(launch_substep(pipe, seq++, stream, sg_mode, sleep));
            extra += more;
        }
        if (profiling) {
            int rc = prof_fold();
            if (rc) return rc;
            prof_count[1] += (uint32_t)(st.total_substeps - prof_sub_base);
            prof_sub_base = st.total_substeps;
        }
        return publish(c, st);
    }

    // The records of the last finished step (step.rs:51-59) from the control
    // block the loop left; st = c.st[seq & 1].
    int publish(const Ctrl& c, const CtrlState& st) {
        // the last step: finished but not yet recorded if no launch followed it
        uint32_t subs, step;
        float last_dt, dropped;
        if (st.in_step) {
            subs = st.substeps;
            step = st.step;
            last_dt = st.last_dt;
            dropped = std::fmax(st.remaining, 0.0f);
        } else {
            subs = st.done_substeps;
            step = st.done_step;
            last_dt = st.done_dt;
            dropped = st.done_dropped;
        }
        predicted_substeps = subs;
        stats.last_substeps = subs;
        stats.last_step_dt = last_dt;
        stats.last_sim_time_dropped = dropped;
        stats.last_vel_clamp_count = c.clamp_count[step % kStepRing];
        stats.last_j_projection_count = c.proj_count[step % kStepRing];
        stats.total_substeps = st.total_substeps;
        return EMH_OK;
    }

    // Simulation::step x n  (step.rs:23-60): queued on the device, drained
    // at the end.
    int step(uint32_t n_steps) {
        if (host_dirty) {
            int rc = upload();
            if (rc) return rc;
        }
        if (sg_mode != 0 && pipe.n > 0 && (pipe.cfg.only_model != kModelNacc || pipe.cfg.density_pass))
            return fail(EMH_ERR_UNSUPPORTED, "the surrogate harvest / inference modes need a scene whose particles "
                                             "are all NACC");
        if (sg_mode != 0 && sleep)
            return fail(EMH_ERR_UNSUPPORTED, "the surrogate harvest / inference modes have no sleep partition");
        if (pipe.n == 0) {  // no particles: the loop still advances time
            for (uint32_t s = 0; s < n_steps; ++s) {
                float remaining = cfg.dt;
                uint32_t k = 0;
                while (remaining > kF32Eps && k < cfg.max_substeps_per_step) {
                    float dt = cfg.adaptive_timestep ? remaining : std::fmin(remaining, cfg.dt);
                    dt = std::fmin(dt, remaining);
                    remaining -= dt;
                    stats.last_step_dt = dt;
                    ++k;
                }
                stats.last_substeps = k;
                stats.total_substeps += k;
                stats.last_sim_time_dropped = std::fmax(remaining, 0.0f);
                stats.last_vel_clamp_count = 0;
                stats.last_j_projection_count = 0;
                stats.frame_index++;
            }
            return EMH_OK;
        }
        // Queue every step without waiting: the device loop starts each queued
        // step once the previous one has finished.  Drain (wait and, if some
        // step needed more substep launches than queued, launch more) every
        // kDrainSteps steps and at the end.
        constexpr uint32_t kDrainSteps = kStepRing / 2;
        for (uint32_t s = 0; s < n_steps; ++s) {
            const bool reorder = reorder_pending || substeps_since_reorder >= reorder_substeps;
            {
                int rc = prof_mark(kMarkSetup);
                if (rc) return rc;
            }
            HIP_TRY(launch_begin_step(pipe, seq++, reorder, stream));
            {
                int rc = prof_mark(kMarkSubsteps);
                if (rc) return rc;
            }
            if (reorder) {
                substeps_since_reorder = 0;
                reorder_pending = false;
            }
            uint32_t slots = std::max<uint32_t>(predicted_substeps, 1u);
            slots = std::min<uint32_t>(slots, cfg.max_substeps_per_step);
            {
                int rc = launch_substep_run(slots);
                if (rc) return rc;
            }
            if (profiling) prof_count[0] += 1;
            substeps_since_reorder += slots;
            if ((s + 1) % kDrainSteps == 0 || s + 1 == n_steps) {
                int rc = drain();
                if (rc) return rc;
            }
        }
        stats.frame_index += n_steps;
        device_ahead = true;
        return EMH_OK;
    }

    // estimate_particle_volumes(_local) on the device for particles [from, n).
    int estimate_volumes(uint32_t from, bool use_box, const float box[4], bool write_initial) {
        const uint32_t n = (uint32_t)host.size();
        if (from >= n) return EMH_OK;
        float2* dx = dalloc<float2>(n);
        float* dm = dalloc<float>(n);
        float* dd = dalloc<float>(n);
        float* dv = dalloc<float>(n);
        uint32_t* dw = dalloc<uint32_t>(n);
        int rc = EMH_OK;
        if (!dx || !dm || !dd || !dv || !dw) rc = fail(EMH_ERR_OUT_OF_MEMORY, "device allocation failed");
        std::vector<float2> x(n);
        std::vector<float> m(n), d(n), v(n);
        std::vector<uint32_t> w(n);
        if (rc == EMH_OK) {
            for (uint32_t i = 0; i < n; ++i) {
                x[i] = make_float2(host[i].x[0], host[i].x[1]);
                m[i] = host[i].mass;
            }
            hipError_t e = hipMemcpyAsync(dx, x.data(), sizeof(float2) * n, hipMemcpyHostToDevice, stream);
            if (e == hipSuccess) e = hipMemcpyAsync(dm, m.data(), sizeof(float) * n, hipMemcpyHostToDevice, stream);
            if (e == hipSuccess)
                e = estimate_volumes_device(dx, dm, n, from, res(), use_box, box, pipe.grid.dm[0], dd, dv, dw, stream);
            if (e == hipSuccess) e = hipMemcpyAsync(d.data(), dd, sizeof(float) * n, hipMemcpyDeviceToHost, stream);
            if (e == hipSuccess) e = hipMemcpyAsync(v.data(), dv, sizeof(float) * n, hipMemcpyDeviceToHost, stream);
            if (e == hipSuccess) e = hipMemcpyAsync(w.data(), dw, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, stream);
            if (e == hipSuccess) e = hipStreamSynchronize(stream);
            if (e != hipSuccess) rc = fail(EMH_ERR_DEVICE, "volume estimate: %s", hipGetErrorString(e));
            // the dense mass grid doubles as the substep's density-pass buffer: clear it again
            if (rc == EMH_OK) {
                e = hipMemsetAsync(pipe.grid.dm[0], 0, sizeof(float) * (size_t)res() * res(), stream);
                if (e == hipSuccess) e = hipStreamSynchronize(stream);
                if (e != hipSuccess) rc = fail(EMH_ERR_DEVICE, "volume estimate clear: %s", hipGetErrorString(e));
            }
        }
        if (rc == EMH_OK) {
            for (uint32_t i = from; i < n; ++i) {
                if (!w[i]) continue;
                host[i].density = d[i];
                host[i].volume = v[i];
                if (write_initial) host[i].initial_volume = v[i];
            }
        }
        (void)hipFree(dx);
        (void)hipFree(dm);
        (void)hipFree(dd);
        (void)hipFree(dv);
        (void)hipFree(dw);
        return rc;
    }

    int add_body(const emh_spawn_region& sp, uint32_t* tag_out) {  // solver/particles.rs:355-424
        if (!spawn_fits(sp, cfg)) return fail(EMH_ERR_INVALID_ARG, "spawn region must stay inside the simulation domain");
        int rc = host_current();
        if (rc) return rc;
        const uint32_t tag = next_tag++;
        const uint32_t old_len = (uint32_t)host.size();
        std::vector<emh_particle> np = lattice(cfg, sp);
        const bool registered = sp.material_id < mats.size();
        for (auto& p : np) {
            p.user_tag = tag;
            if (registered) init_particle(mats[sp.material_id], p);
            host.push_back(p);
        }
        host_dirty = true;
        // estimate_particle_volumes_local  density.rs:256-319
        if (old_len < host.size()) {
            float lo[2] = {std::numeric_limits<float>::max(), std::numeric_limits<float>::max()};
            float hi[2] = {std::numeric_limits<float>::lowest(), std::numeric_limits<float>::lowest()};
            for (size_t i = old_len; i < host.size(); ++i)
                for (int k = 0; k < 2; ++k) {
                    lo[k] = std::fmin(lo[k], host[i].x[k]);
                    hi[k] = std::fmax(hi[k], host[i].x[k]);
                }
            const float box[4] = {lo[0] - 3.0f, lo[1] - 3.0f, hi[0] + 3.0f, hi[1] + 3.0f};
            rc = estimate_volumes(old_len, true, box, true);
            if (rc) return rc;
        }
        if (tag_out) *tag_out = tag;
        return EMH_OK;
    }
};

// -------------------------------------------------------------------- C-ABI
extern "C" {

uint32_t emh_abi_version(void) { return EMH_ABI_VERSION; }
const char* emh_last_error(void) { return g_err.c_str(); }

int emh_device_count(int* count) {
    if (!count) return fail(EMH_ERR_INVALID_ARG, "null count");
    *count = device_present();
    return EMH_OK;
}

int emh_config_default(emh_config* o) {  // config.rs:123-159
    if (!o) return fail(EMH_ERR_INVALID_ARG, "null config");
    std::memset(o, 0, sizeof(*o));
    o->grid_res = 64;
    o->grid_cell_size = 1.0f;
    o->dt = 1.0f;
    o->adaptive_timestep = 1;
    o->cfl_include_affine_speed = 1;
    o->cfl_coefficient = 0.9f;
    o->material_cfl_coefficient = 0.5f;
    o->viscous_timestep_coefficient = 0.5f;
    o->min_dt = 1.0e-3f;
    o->project_invalid_state = 1;
    o->projection_min_density = 1.0e-6f;
    o->projection_min_volume = 1.0e-6f;
    o->projection_min_deformation_j = 1.0e-6f;
    o->gravity[0] = 0.0f;
    o->gravity[1] = -0.05f;
    o->boundary_thickness = 2;
    o->default_initial_volume = 1.0f;
    o->recompute_density_each_step = 0;
    o->particle_mass = 1.0f;
    o->max_substeps_per_step = 64;
    o->apic_blend = 1.0f;
    o->j_max = 50.0f;
    o->sleep_threshold = 0.0f;
    o->contact_friction = 0.5f;
    o->asflip_blend = 0.0f;
    o->mixture_drag_coefficient = 0.0f;
    o->mixture_pressure_iterations = 0;
    o->dx_meters = 1.0f;
    o->dt_seconds = 1.0f;
    return EMH_OK;
}

int emh_config_standard(uint32_t grid_res, float dt, float gx, float gy, emh_config* o) {  // config.rs:165-172
    int rc = emh_config_default(o);
    if (rc) return rc;
    o->grid_res = grid_res;
    o->dt = dt;
    o->gravity[0] = gx;
    o->gravity[1] = gy;
    return EMH_OK;
}

int emh_config_unsafe_defaults(emh_config* o) {  // config.rs:179-185
    int rc = emh_config_default(o);
    if (rc) return rc;
    o->adaptive_timestep = 0;
    o->project_invalid_state = 0;
    return EMH_OK;
}

int emh_config_validate(const emh_config* c) {
    if (!c) return fail(EMH_ERR_INVALID_ARG, "null config");
    return validate_config(*c);
}

int emh_lame_from_young(float e, float nu, float* lambda, float* mu) {  // utils.rs:183-193
    if (!lambda || !mu) return fail(EMH_ERR_INVALID_ARG, "null output");
    *lambda = e * nu / ((1.0f + nu) * (1.0f - 2.0f * nu));
    *mu = e / (2.0f * (1.0f + nu));
    return EMH_OK;
}

float emh_to_radians(float degrees) { return to_radians(degrees); }

int emh_material_neohookean(float lambda, float mu, emh_material_params* p, emh_material_ext* e) {
    if (!p || !e) return fail(EMH_ERR_INVALID_ARG, "null output");
    std::memset(p, 0, sizeof(*p));
    ext_defaults(e);
    p->model = EMH_MODEL_NEOHOOKEAN;
    p->lambda = lambda;
    p->mu = mu;
    return EMH_OK;
}

int emh_material_corotated(float lambda, float mu, emh_material_params* p, emh_material_ext* e) {
    if (!p || !e) return fail(EMH_ERR_INVALID_ARG, "null output");
    std::memset(p, 0, sizeof(*p));
    ext_defaults(e);
    p->model = EMH_MODEL_COROTATED;
    p->lambda = lambda;
    p->mu = mu;
    return EMH_OK;
}

int emh_material_fluid(float rest_density, float visc, float eos_k, float eos_p, emh_material_params* p,
                       emh_material_ext* e) {
    if (!p || !e) return fail(EMH_ERR_INVALID_ARG, "null output");
    std::memset(p, 0, sizeof(*p));
    ext_defaults(e);
    p->model = EMH_MODEL_FLUID;
    p->rest_density = rest_density;
    p->eos_stiffness = eos_k;
    p->eos_power = eos_p;
    p->dynamic_viscosity = visc;
    p->volume_ratio_max = 2.0f;
    p->pressure_floor = -0.1f;
    return EMH_OK;
}

int emh_material_fluid_low_viscosity(float rest_density, float eos_k, emh_material_params* p, emh_material_ext* e) {
    return emh_material_fluid(rest_density, 1.0e-3f, eos_k, 7.0f, p, e);
}

int emh_material_drucker_prager(float lambda, float mu, emh_material_params* p, emh_material_ext* e) {
    if (!p || !e) return fail(EMH_ERR_INVALID_ARG, "null output");
    std::memset(p, 0, sizeof(*p));
    ext_defaults(e);
    p->model = EMH_MODEL_DRUCKER_PRAGER;
    p->lambda = lambda;
    p->mu = mu;
    p->dp_h0 = to_radians(35.0f);
    p->dp_h1 = to_radians(9.0f);
    p->dp_h2 = 0.2f;
    p->dp_h3 = to_radians(10.0f);
    p->volume_ratio_min = 0.6f;
    return EMH_OK;
}

int emh_material_snow(float lambda, float mu, float xi, float tc, float ts, float jmin, float jmax,
                      emh_material_params* p, emh_material_ext* e) {
    if (!p || !e) return fail(EMH_ERR_INVALID_ARG, "null output");
    std::memset(p, 0, sizeof(*p));
    ext_defaults(e);
    p->model = EMH_MODEL_SNOW;
    p->lambda = lambda;
    p->mu = mu;
    p->hardening_exponent = xi;
    p->compression_limit = tc;
    p->stretch_limit = ts;
    p->volume_ratio_min = jmin;
    p->volume_ratio_max = jmax;
    return EMH_OK;
}

int emh_material_snow_from_young(float young, float nu, emh_material_params* p, emh_material_ext* e) {
    float l, m;
    emh_lame_from_young(young, nu, &l, &m);
    return emh_material_snow(l, m, 10.0f, 0.025f, 0.0075f, 0.6f, 20.0f, p, e);
}

int emh_material_von_mises(float lambda, float mu, float yield_stress, float hardening_modulus,
                           emh_material_params* p, emh_material_ext* e) {
    if (!p || !e) return fail(EMH_ERR_INVALID_ARG, "null output");
    if (!(hardening_modulus >= 0.0f)) return fail(EMH_ERR_INVALID_ARG, "hardening_modulus must be >= 0");
    std::memset(p, 0, sizeof(*p));
    ext_defaults(e);
    p->model = EMH_MODEL_VON_MISES;
    p->lambda = lambda;
    p->mu = mu;
    p->hardening_exponent = yield_stress;
    p->hardening_modulus = hardening_modulus;
    return EMH_OK;
}

int emh_material_rankine(float lambda, float mu, float tensile_strength, float softening_rate,
                         emh_material_params* p, emh_material_ext* e) {
    if (!p || !e) return fail(EMH_ERR_INVALID_ARG, "null output");
    std::memset(p, 0, sizeof(*p));
    ext_defaults(e);
    p->model = EMH_MODEL_RANKINE;
    p->lambda = lambda;
    p->mu = mu;
    p->hardening_exponent = tensile_strength;
    p->hardening_modulus = softening_rate;
    return EMH_OK;
}

int emh_material_mu_i(float lambda, float mu, float mu_static, float mu_dynamic, float inertial_q,
                      emh_material_params* p, emh_material_ext* e) {
    if (!p || !e) return fail(EMH_ERR_INVALID_ARG, "null output");
    std::memset(p, 0, sizeof(*p));
    ext_defaults(e);
    p->model = EMH_MODEL_MU_I;
    p->lambda = lambda;
    p->mu = mu;
    p->dp_h0 = mu_static;
    p->dp_h1 = mu_dynamic;
    p->dp_h2 = inertial_q;
    return EMH_OK;
}

int emh_material_mu_i_default(float lambda, float mu, emh_material_params* p, emh_material_ext* e) {
    return emh_material_mu_i(lambda, mu, tanf(to_radians(20.9f)), tanf(to_radians(32.8f)), 5.58f, p, e);
}

int emh_material_granular_fluid(float lambda, float mu, float rest_density, float eos_stiffness,
                                float hardening_exponent, float compression_limit, emh_material_params* p,
                                emh_material_ext* e) {
    if (!p || !e) return fail(EMH_ERR_INVALID_ARG, "null output");
    std::memset(p, 0, sizeof(*p));
    ext_defaults(e);
    p->model = EMH_MODEL_GRANULAR_FLUID;
    p->lambda = lambda;
    p->mu = mu;
    p->rest_density = rest_density;
    p->eos_stiffness = eos_stiffness;
    p->eos_power = 7.0f;
    p->hardening_exponent = hardening_exponent;
    p->compression_limit = compression_limit;
    p->stretch_limit = 0.01f;
    p->volume_ratio_min = 0.2f;
    p->volume_ratio_max = 3.0f;
    p->pressure_floor = 0.0f;
    return EMH_OK;
}

int emh_material_bingham(float rest_density, float dynamic_viscosity, float eos_stiffness, float eos_power,
                         float yield_stress, emh_material_params* p, emh_material_ext* e) {
    if (!p || !e) return fail(EMH_ERR_INVALID_ARG, "null output");
    std::memset(p, 0, sizeof(*p));
    ext_defaults(e);
    p->model = EMH_MODEL_FLUID;
    p->rest_density = rest_density;
    p->eos_stiffness = eos_stiffness;
    p->eos_power = eos_power;
    p->dynamic_viscosity = dynamic_viscosity;
    p->compression_limit = yield_stress;
    p->volume_ratio_max = 2.0f;
    p->pressure_floor = 0.0f;
    e->fluid_bingham = 1u;
    return EMH_OK;
}

int emh_material_viscoelastic(float lambda, float mu, float viscosity, emh_material_params* p,
                              emh_material_ext* e) {
    if (!p || !e) return fail(EMH_ERR_INVALID_ARG, "null output");
    std::memset(p, 0, sizeof(*p));
    ext_defaults(e);
    p->model = EMH_MODEL_VISCOELASTIC;
    p->lambda = lambda;
    p->mu = mu;
    p->dynamic_viscosity = viscosity;
    p->volume_ratio_min = 0.01f;
    return EMH_OK;
}

int emh_material_nacc(float mu, float kappa, float friction, float cohesion, float xi, emh_material_params* p,
                      emh_material_ext* e) {
    if (!p || !e) return fail(EMH_ERR_INVALID_ARG, "null output");
    std::memset(p, 0, sizeof(*p));
    ext_defaults(e);
    p->model = EMH_MODEL_NACC;
    p->lambda = kappa - mu;
    p->mu = mu;
    p->hardening_exponent = xi;
    p->compression_limit = cohesion;
    p->stretch_limit = friction;
    e->nacc_kappa = kappa;
    e->nacc_hardening_enabled = xi > 0.0f ? 1u : 0u;
    return EMH_OK;
}

int emh_material_nacc_from_young(float young, float nu, float friction, float cohesion, float xi,
                                 emh_material_params* p, emh_material_ext* e) {
    float l, m;
    emh_lame_from_young(young, nu, &l, &m);
    return emh_material_nacc(m, l + m, friction, cohesion, xi, p, e);
}

int emh_sim_create(const emh_config* cfg, emh_sim** out) {  // Simulation::empty  lifecycle.rs:26-55
    if (!cfg || !out) return fail(EMH_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    int rc = validate_config(*cfg);
    if (rc) return rc;
    rc = check_supported_config(*cfg);
    if (rc) return rc;
    emh_sim* s = new emh_sim();
    s->cfg = *cfg;
    emh_material_params fb;
    std::memset(&fb, 0, sizeof(fb));
    emh_material_ext fe;
    ext_defaults(&fe);
    s->mats.push_back(fb);  // FallbackMaterial
    s->exts.push_back(fe);
    s->bnds.push_back(emh_sim::HostBoundary{kBndSlip, cfg->boundary_thickness, 0.0f, 0.0f, {}});
    s->stats.last_step_dt = cfg->dt;
    rc = s->init_device();
    if (rc) {
        delete s;
        return rc;
    }
    *out = s;
    return EMH_OK;
}

int emh_sim_create_from_spawn(const emh_config* cfg, const emh_spawn_region* sp, emh_sim** out) {
    if (!sp) return fail(EMH_ERR_INVALID_ARG, "null spawn");
    int rc = emh_sim_create(cfg, out);
    if (rc) return rc;
    emh_sim* s = *out;
    if (!spawn_fits(*sp, s->cfg)) {
        emh_sim_destroy(s);
        *out = nullptr;
        return fail(EMH_ERR_INVALID_ARG, "spawn region must stay inside the simulation domain");
    }
    s->host = lattice(s->cfg, *sp);  // Simulation::new  lifecycle.rs:57-105
    s->host_dirty = true;
    if (sp->precompute_initial_volumes) {
        const float box[4] = {0, 0, 0, 0};
        rc = s->estimate_volumes(0, false, box, true);
        if (rc) {
            emh_sim_destroy(s);
            *out = nullptr;
            return rc;
        }
    }
    return EMH_OK;
}

void emh_sim_destroy(emh_sim* sim) { delete sim; }

int emh_sim_set_material(emh_sim* s, uint32_t id, const emh_material_params* p, const emh_material_ext* e) {
    if (!s || !p) return fail(EMH_ERR_INVALID_ARG, "null argument");
    if (id >= EMH_MAX_MATERIALS) return fail(EMH_ERR_INVALID_ARG, "material_id %u exceeds the limit of %u", id, EMH_MAX_MATERIALS);
    if (id > s->mats.size())
        return fail(EMH_ERR_INVALID_ARG, "material IDs must be registered contiguously starting at 0; expected id %zu, got %u",
                    s->mats.size(), id);
    int rc = check_supported_material(*p);
    if (rc) return rc;
    rc = s->host_current();
    if (rc) return rc;
    emh_material_ext ext;
    if (e) ext = *e;
    else ext_defaults(&ext);
    if (id < s->mats.size()) {
        s->mats[id] = *p;
        s->exts[id] = ext;
    } else {
        s->mats.push_back(*p);
        s->exts.push_back(ext);
    }
    if (id == 0) {  // with_default_material re-runs init_particle on every particle
        for (auto& q : s->host) {
            const uint32_t m = q.material_id < s->mats.size() ? q.material_id : 0u;
            init_particle(s->mats[m], q);
        }
    }
    s->host_dirty = true;
    return EMH_OK;
}

int emh_sim_reinit_particle_state(emh_sim* s) {
    if (!s) return fail(EMH_ERR_INVALID_ARG, "null sim");
    int rc = s->host_current();
    if (rc) return rc;
    for (auto& q : s->host) {
        const uint32_t m = q.material_id < s->mats.size() ? q.material_id : 0u;
        init_particle(s->mats[m], q);
    }
    s->host_dirty = true;
    return EMH_OK;
}

static int add_boundary(emh_sim* s, uint32_t kind, uint32_t t, float mu, float wall_min, const float* heights,
                        uint32_t n_heights) {
    if (!s) return fail(EMH_ERR_INVALID_ARG, "null sim");
    if (s->bnds.size() >= EMH_MAX_BOUNDARIES) return fail(EMH_ERR_UNSUPPORTED, "at most %u boundaries", EMH_MAX_BOUNDARIES);
    if (t + 1 > s->cfg.grid_res) return fail(EMH_ERR_INVALID_ARG, "boundary thickness %u too large for grid_res %u", t, s->cfg.grid_res);
    int rc = s->host_current();
    if (rc) return rc;
    emh_sim::HostBoundary b{kind, t, mu, wall_min, {}};
    if (heights) b.heights.assign(heights, heights + n_heights);
    s->bnds.push_back(std::move(b));
    s->host_dirty = true;
    return EMH_OK;
}

int emh_sim_add_slip_boundary(emh_sim* s, uint32_t t) { return add_boundary(s, kBndSlip, t, 0.0f, 0.0f, nullptr, 0); }

int emh_sim_add_friction_boundary(emh_sim* s, uint32_t t, float mu) {
    if (!(mu >= 0.0f && mu <= 1.0f))  // friction.rs:29-32
        return fail(EMH_ERR_INVALID_ARG, "friction_coefficient must be in [0.0, 1.0], got %g", (double)mu);
    return add_boundary(s, kBndFriction, t, mu, 0.0f, nullptr, 0);
}

int emh_sim_add_heightmap_boundary(emh_sim* s, const float* heights, uint32_t n_heights, float friction,
                                   uint32_t wall_thickness) {
    if (n_heights && !heights) return fail(EMH_ERR_INVALID_ARG, "null heights");
    if (n_heights > (1u << 24)) return fail(EMH_ERR_INVALID_ARG, "too many heights");
    return add_boundary(s, kBndHeightmap, wall_thickness, friction, 0.0f, heights, n_heights);
}

int emh_sim_add_predictive_boundary(emh_sim* s, uint32_t t, float predictive_wall_min) {
    return add_boundary(s, kBndPredictive, t, 0.0f, predictive_wall_min, nullptr, 0);
}

int emh_sim_clear_boundaries(emh_sim* s) {
    if (!s) return fail(EMH_ERR_INVALID_ARG, "null sim");
    int rc = s->host_current();
    if (rc) return rc;
    s->bnds.clear();
    s->host_dirty = true;
    return EMH_OK;
}

int emh_sim_add_body(emh_sim* s, const emh_spawn_region* sp, uint32_t* tag) {
    if (!s || !sp) return fail(EMH_ERR_INVALID_ARG, "null argument");
    return s->add_body(*sp, tag);
}

int emh_sim_set_particles(emh_sim* s, const emh_particle* p, size_t n) {
    if (!s || (!p && n)) return fail(EMH_ERR_INVALID_ARG, "null argument");
    if (n > 0x7fffffffu) return fail(EMH_ERR_INVALID_ARG, "too many particles");
    s->host.assign(p, p + n);
    s->device_ahead = false;
    s->host_dirty = true;
    return EMH_OK;
}

int emh_sim_num_particles(const emh_sim* s, size_t* n) {
    if (!s || !n) return fail(EMH_ERR_INVALID_ARG, "null argument");
    *n = s->host.size();
    return EMH_OK;
}

int emh_sim_get_particles(emh_sim* s, emh_particle* out, size_t n) {
    if (!s || (!out && n)) return fail(EMH_ERR_INVALID_ARG, "null argument");
    if (n != s->host.size()) return fail(EMH_ERR_INVALID_ARG, "expected %zu particles, got %zu", s->host.size(), n);
    int rc = s->host_current();
    if (rc) return rc;
    if (n) std::memcpy(out, s->host.data(), sizeof(emh_particle) * n);
    return EMH_OK;
}

int emh_sim_get_config(const emh_sim* s, emh_config* out) {
    if (!s || !out) return fail(EMH_ERR_INVALID_ARG, "null argument");
    *out = s->cfg;
    return EMH_OK;
}

int emh_sim_step(emh_sim* s, uint32_t n) {
    if (!s) return fail(EMH_ERR_INVALID_ARG, "null sim");
    return s->step(n);
}

int emh_sim_stats(const emh_sim* s, emh_step_stats* out) {
    if (!s || !out) return fail(EMH_ERR_INVALID_ARG, "null argument");
    *out = s->stats;
    return EMH_OK;
}

// Simulation::sleep_tag / wake_tag (solver/particles.rs:270-312) on the
// host copy; the next step uploads it.
static int set_tag_sleep(emh_sim* s, uint32_t tag, uint32_t sleeping) {
    if (!s) return fail(EMH_ERR_INVALID_ARG, "null sim");
    int rc = s->host_current();
    if (rc) return rc;
    bool any = false;
    for (emh_particle& p : s->host)
        if (p.user_tag == tag && p.sleeping != sleeping) {
            p.sleeping = sleeping;
            any = true;
        }
    if (any) s->host_dirty = true;
    return EMH_OK;
}
int emh_sim_sleep_tag(emh_sim* s, uint32_t tag) { return set_tag_sleep(s, tag, 1u); }
int emh_sim_wake_tag(emh_sim* s, uint32_t tag) { return set_tag_sleep(s, tag, 0u); }

int emh_sim_set_gravity(emh_sim* s, float gx, float gy) {  // lifecycle.rs:403-405
    if (!s) return fail(EMH_ERR_INVALID_ARG, "null sim");
    s->cfg.gravity[0] = gx;
    s->cfg.gravity[1] = gy;
    s->pipe.cfg.gx = gx;  // read by the next grid update; nothing queued holds the old value
    s->pipe.cfg.gy = gy;
    ++s->cfg_epoch;  // a batch holding this scene refreshes its device copy of the pipeline
    return EMH_OK;
}

int emh_sim_apply_impulse(emh_sim* s, float cx, float cy, float radius, float fx, float fy) {
    if (!s) return fail(EMH_ERR_INVALID_ARG, "null sim");
    return s->impulse(cx, cy, radius, fx, fy, 0.0f, 0u);
}

int emh_sim_apply_radial_impulse(emh_sim* s, float cx, float cy, float radius, float strength) {
    if (!s) return fail(EMH_ERR_INVALID_ARG, "null sim");
    return s->impulse(cx, cy, radius, 0.0f, 0.0f, strength, 1u);
}

int emh_sim_get_slot_order(emh_sim* s, uint32_t* ids_out, size_t n) {
    if (!s || (!ids_out && n)) return fail(EMH_ERR_INVALID_ARG, "null argument");
    if (n != s->host.size()) return fail(EMH_ERR_INVALID_ARG, "expected %zu particles, got %zu", s->host.size(), n);
    return s->slot_order(ids_out);
}

int emh_sim_export_particles_device(emh_sim* s, void* dst, size_t capacity) {
    if (!s || !dst) return fail(EMH_ERR_INVALID_ARG, "null argument");
    return s->export_particles(dst, capacity);
}

int emh_sim_export_grid_device(emh_sim* s, void* dst, size_t capacity_cells) {
    if (!s || !dst) return fail(EMH_ERR_INVALID_ARG, "null argument");
    const size_t R2 = (size_t)s->res() * s->res();
    if (capacity_cells < R2) return fail(EMH_ERR_INVALID_ARG, "buffer holds %zu cells, %zu needed", capacity_cells, R2);
    HIP_TRY(launch_export_grid(s->pipe, s->seq, (float4*)dst, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return EMH_OK;
}

int emh_sim_get_grid(emh_sim* s, float* mass, float* momentum, size_t n_cells) {
    if (!s) return fail(EMH_ERR_INVALID_ARG, "null sim");
    const size_t R2 = (size_t)s->res() * s->res();
    if (n_cells != R2) return fail(EMH_ERR_INVALID_ARG, "expected %zu cells", R2);
    Ctrl ctrl;
    HIP_TRY(hipMemcpyAsync(&ctrl, s->d_ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    // the accumulators the last substep's grid update read; momentum as that
    // update formed it: A + (dt - dt') S (see GridBuf)
    const CtrlState st = ctrl.st[s->seq & 1u];
    const uint32_t g = st.cur_grid % 3u;
    if (mass) HIP_TRY(hipMemcpy(mass, s->pipe.grid.acc[g][kAccM], sizeof(float) * R2, hipMemcpyDeviceToHost));
    if (momentum) {
        std::vector<float> ch[4];
        for (int k = 0; k < 4; ++k) {
            ch[k].resize(R2);
            HIP_TRY(hipMemcpy(ch[k].data(), s->pipe.grid.acc[g][kAccPx + k], sizeof(float) * R2, hipMemcpyDeviceToHost));
        }
        const float ddt = st.last_ddt;
        for (size_t i = 0; i < R2; ++i) {
            momentum[2 * i] = ch[0][i] + ddt * ch[2][i];
            momentum[2 * i + 1] = ch[1][i] + ddt * ch[3][i];
        }
    }
    return EMH_OK;
}

int emh_sim_snapshot(emh_sim* s, emh_snapshot* out) {  // snapshot.rs:227-372, queries.rs:15-30
    if (!s || !out) return fail(EMH_ERR_INVALID_ARG, "null argument");
    if (s->host_dirty) {
        int rc = s->upload();
        if (rc) return rc;
    }
    const size_t R2 = (size_t)s->res() * s->res();
    if (!s->d_snap) {
        s->d_snap = dalloc<DevSnapshot>(1);
        s->d_snap_cells = dalloc<uint32_t>(3 * R2);
        if (!s->d_snap || !s->d_snap_cells) return fail(EMH_ERR_OUT_OF_MEMORY, "device allocation failed");
        HIP_TRY(hipMemsetAsync(s->d_snap_cells, 0xff, sizeof(uint32_t) * R2, s->stream));
        HIP_TRY(hipMemsetAsync(s->d_snap_cells + R2, 0, sizeof(uint32_t) * 2 * R2, s->stream));
    }
    Ctrl ctrl;
    HIP_TRY(hipMemcpyAsync(&ctrl, s->d_ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    const CtrlState st = ctrl.st[s->seq & 1u];
    DevSnapshot z;
    std::memset(&z, 0, sizeof(z));
    z.min_j = z.min_jp = 0xffffffffu;
    HIP_TRY(hipMemcpyAsync(s->d_snap, &z, sizeof(z), hipMemcpyHostToDevice, s->stream));
    HIP_TRY(launch_snapshot(s->pipe, st.cur_buf & 1u, st.cur_grid % 3u, st.last_dt, st.last_ddt, s->d_snap,
                            s->d_snap_cells, s->d_snap_cells + R2, s->d_snap_cells + 2 * R2, s->stream));
    DevSnapshot d;
    HIP_TRY(hipMemcpyAsync(&d, s->d_snap, sizeof(d), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    const emh_config& cfg = s->cfg;
    const uint32_t n = s->pipe.n;
    emh_snapshot o;
    std::memset(&o, 0, sizeof(o));
    o.frame_index = s->stats.frame_index;
    o.configured_dt = cfg.dt;
    o.effective_dt = s->stats.last_step_dt;
    o.substeps_last_step = s->stats.last_substeps;
    o.particle_count = n;
    o.valid_particle_count = d.valid;
    o.active_grid_cells = d.active_cells;
    if (d.active_cells > 0) o.particles_per_active_cell = (float)n / (float)d.active_cells;
    if (d.touched_cells > 0) o.mixed_material_cell_ratio = (float)d.mixed_cells / (float)d.touched_cells;
    if (n > 0) o.mixed_material_particle_ratio = (float)d.particles_in_mixed / (float)n;
    o.total_particle_mass = (float)d.mass;
    o.total_grid_mass = (float)d.grid_mass;
    if (o.total_particle_mass > kF32Eps)
        o.relative_mass_error = std::fabs(o.total_grid_mass - o.total_particle_mass) / o.total_particle_mass;
    o.total_particle_momentum[0] = (float)d.mom_x;
    o.total_particle_momentum[1] = (float)d.mom_y;
    o.total_grid_momentum[0] = (float)d.grid_mom_x;
    o.total_grid_momentum[1] = (float)d.grid_mom_y;
    const float pm = std::hypot(o.total_particle_momentum[0], o.total_particle_momentum[1]);
    const float gm = std::hypot(o.total_grid_momentum[0], o.total_grid_momentum[1]);
    const float scale = std::fmax(std::fmax(pm, gm), o.total_particle_mass * 1.0e-3f);
    if (scale > 0.0f)
        o.relative_momentum_error = std::hypot(o.total_grid_momentum[0] - o.total_particle_momentum[0],
                                               o.total_grid_momentum[1] - o.total_particle_momentum[1]) / scale;
    float bits2f;
    std::memcpy(&bits2f, &d.max_speed, 4);
    o.max_particle_speed = bits2f;
    std::memcpy(&bits2f, &d.max_grid_speed, 4);
    o.max_grid_speed = bits2f;
    std::memcpy(&bits2f, &d.max_pinned_speed, 4);
    o.max_pinned_particle_speed = bits2f;
    o.min_deformation_j = d.min_j == 0xffffffffu ? std::numeric_limits<float>::infinity() : order_key_inverse(d.min_j);
    o.max_deformation_j = d.max_j == 0u ? -std::numeric_limits<float>::infinity() : order_key_inverse(d.max_j);
    o.out_of_bounds_particles = d.out_of_bounds;
    o.invalid_physical_particle_values = d.invalid;
    o.non_finite_particle_values = d.non_finite;
    o.non_finite_grid_values = d.non_finite_grid;
    o.avg_plastic_jacobian = 1.0f;
    o.avg_elastic_hardening = 1.0f;
    o.min_plastic_jacobian = std::numeric_limits<float>::infinity();
    if (n > 0) {
        o.avg_plastic_jacobian = (float)(d.jp_sum / (double)n);
        o.avg_elastic_hardening = (float)(d.h_sum / (double)n);
        o.min_plastic_jacobian = d.jp_finite ? order_key_inverse(d.min_jp) : 1.0f;
    }
    const float cell = std::fmax(cfg.grid_cell_size, kF32Eps);
    o.cfl_number = o.max_particle_speed * o.effective_dt / cell;
    o.recommended_max_dt_from_velocity_cfl =
        o.max_particle_speed > kF32Eps ? cell / o.max_particle_speed : std::numeric_limits<float>::infinity();
    o.active_count = n - d.sleeping;
    o.sleeping_count = d.sleeping;
    o.vel_clamp_count = s->stats.last_vel_clamp_count;
    o.j_projection_count = s->stats.last_j_projection_count;
    o.sim_time_dropped = s->stats.last_sim_time_dropped;
    o.total_kinetic_energy = (float)d.kinetic;
    *out = o;
    return EMH_OK;
}

int emh_sim_synchronize(emh_sim* s) {
    if (!s) return fail(EMH_ERR_INVALID_ARG, "null sim");
    HIP_TRY(hipStreamSynchronize(s->stream));
    return EMH_OK;
}

int emh_sim_set_reorder_interval(emh_sim* s, uint32_t substeps) {  // the tuning's reorder_interval alone
    if (!s) return fail(EMH_ERR_INVALID_ARG, "null sim");
    if (substeps == 0) return fail(EMH_ERR_INVALID_ARG, "reorder interval must be at least one substep");
    s->reorder_substeps = substeps;
    return EMH_OK;
}

static int check_tuning(const emh_tuning* t) {
    if (t->reorder_interval == 0) return fail(EMH_ERR_INVALID_ARG, "reorder interval must be at least one substep");
    if (t->graphs > 1) return fail(EMH_ERR_INVALID_ARG, "graphs must be 0 or 1");
    if (t->model_segments > 1) return fail(EMH_ERR_INVALID_ARG, "model_segments must be 0 or 1");
    return EMH_OK;
}

int emh_sim_get_tuning(emh_sim* s, emh_tuning* out) {
    if (!s || !out) return fail(EMH_ERR_INVALID_ARG, "null argument");
    *out = s->tuning();
    return EMH_OK;
}

int emh_sim_set_tuning(emh_sim* s, const emh_tuning* t) {
    if (!s || !t) return fail(EMH_ERR_INVALID_ARG, "null argument");
    int rc = check_tuning(t);
    if (rc) return rc;
    s->apply_tuning(*t);
    return EMH_OK;
}

int emh_sim_autotune(emh_sim* s, uint32_t trial_steps, emh_tuning* chosen, emh_tuning* timed, uint32_t timed_cap,
                     uint32_t* n_timed) {
    if (!s) return fail(EMH_ERR_INVALID_ARG, "null sim");
    return s->autotune(trial_steps, chosen, timed, timed_cap, n_timed);
}

int emh_sim_set_profiling(emh_sim* s, int enable) {
    if (!s) return fail(EMH_ERR_INVALID_ARG, "null sim");
    s->profiling = enable != 0;
    s->prof_used = 0;
    s->prof_sub_base = s->stats.total_substeps;
    for (int k = 0; k < 2; ++k) {
        s->prof_ms[k] = 0.0;
        s->prof_count[k] = 0;
    }
    return EMH_OK;
}

int emh_sim_kernel_times(emh_sim* s, emh_kernel_time* out, uint32_t cap, uint32_t* n_out) {
    if (!s || (!out && cap)) return fail(EMH_ERR_INVALID_ARG, "null argument");
    const char* names[2] = {"step_setup", s->pipe.cfg.density_pass ? "k_density_p2g+k_substep" : "k_substep"};
    uint32_t n = 0;
    for (int k = 0; k < 2; ++k) {
        if (n < cap) {
            std::memset(&out[n], 0, sizeof(out[n]));
            std::snprintf(out[n].name, sizeof(out[n].name), "%s", names[k]);
            out[n].launches = s->prof_count[k];
            out[n].total_ms = s->prof_ms[k];
        }
        ++n;
    }
    if (n_out) *n_out = n;
    return EMH_OK;
}

uint32_t emh_cell_key(int32_t cx, int32_t cy, uint32_t R) {
    int32_t hi = (int32_t)R - 1;
    cx = cx < 0 ? 0 : (cx > hi ? hi : cx);
    cy = cy < 0 ? 0 : (cy > hi ? hi : cy);
    return emh::cell_key(cx, cy, R, emh::cell_key_order(R));
}

int emh_sort_cell_keys(const float* xy, size_t n, uint32_t R, uint32_t* keys_out, uint32_t* perm_out) {
    if ((!xy || !keys_out || !perm_out) && n) return fail(EMH_ERR_INVALID_ARG, "null argument");
    if (R < 1 || R > 16384) return fail(EMH_ERR_INVALID_ARG, "grid_res out of range");
    if (n > 0x7fffffffu) return fail(EMH_ERR_INVALID_ARG, "too many keys");
    if (device_present() == 0) return fail(EMH_ERR_DEVICE, "no HIP device is available");
    const uint32_t n_bins = 1u << (2u * cell_key_order(R));
    const uint32_t nn = (uint32_t)n;
    if (nn == 0) return EMH_OK;
    hipStream_t st = nullptr;
    HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    float2* d_xy = dalloc<float2>(nn);
    uint32_t* d_keys = dalloc<uint32_t>(nn);
    uint32_t* d_perm = dalloc<uint32_t>(nn);
    uint32_t* d_tmp = dalloc<uint32_t>(4 * (size_t)nn);
    uint32_t* d_hist = dalloc<uint32_t>(sort_hist_len(nn, n_bins));
    int rc = EMH_OK;
    if (!d_xy || !d_keys || !d_perm || !d_tmp || !d_hist) {
        rc = fail(EMH_ERR_OUT_OF_MEMORY, "device allocation failed");
    } else {
        SortBuf sb{};
        for (int k = 0; k < 2; ++k) {
            sb.kbuf[k] = d_tmp + (size_t)(2 * k) * nn;
            sb.vbuf[k] = d_tmp + (size_t)(2 * k + 1) * nn;
        }
        sb.hist = d_hist;
        hipError_t e = hipMemcpyAsync(d_xy, xy, sizeof(float2) * nn, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = sort_cell_keys_device(d_xy, nn, R, d_keys, d_perm, sb, st);
        if (e == hipSuccess) e = hipMemcpyAsync(keys_out, d_keys, sizeof(uint32_t) * nn, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipMemcpyAsync(perm_out, d_perm, sizeof(uint32_t) * nn, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) rc = fail(EMH_ERR_DEVICE, "sort: %s", hipGetErrorString(e));
    }
    (void)hipFree(d_xy);
    (void)hipFree(d_keys);
    (void)hipFree(d_perm);
    (void)hipFree(d_tmp);
    (void)hipFree(d_hist);
    (void)hipStreamDestroy(st);
    return rc;
}

int emh_sort_cell_keys_noalloc(const float* xy, size_t n, uint32_t R, uint32_t* keys_out, uint32_t* perm_out,
                               void* stream) {
    if ((!xy || !keys_out || !perm_out) && n) return fail(EMH_ERR_INVALID_ARG, "null argument");
    if (R < 1 || R > 16384) return fail(EMH_ERR_INVALID_ARG, "grid_res out of range");
    if (n > EMH_SORT_NOALLOC_MAX_KEYS)
        return fail(EMH_ERR_INVALID_ARG, "%zu keys exceed the static scratch maximum of %u", n,
                    EMH_SORT_NOALLOC_MAX_KEYS);
    if (n == 0) return EMH_OK;
    SortBuf sb;
    HIP_TRY(noalloc_sort_scratch(sb));
    HIP_TRY(sort_cell_keys_device(reinterpret_cast<const float2*>(xy), (uint32_t)n, R, keys_out, perm_out, sb,
                                  static_cast<hipStream_t>(stream)));
    return EMH_OK;
}

}  // extern "C"

// ------------------------------------------------------ learned NACC surrogate
// Device state of one surrogate network: FP32 master weights with Adam
// moments, the packed BF16 operand tiles, the normalisation block, harvested
// trial-state records and the training set.
struct emh_surrogate {
    int device = 0;
    hipStream_t stream = nullptr;
    float* d_w = nullptr;     // kSgParams master weights
    float* d_m = nullptr;     // Adam first moments
    float* d_v = nullptr;     // Adam second moments
    float* d_grad = nullptr;  // gradient accumulator (cleared by each Adam step)
    float* d_norm = nullptr;  // kSgNorm
    uint4* d_pk = nullptr;    // kPkTiles * 64 packed operand lanes
    float* d_bias = nullptr;  // compact bias block (kSgBias)
    unsigned long long* d_counts = nullptr;  // 3 x kSgCountSlots return-mapping counts, then the harvest offers
    float* d_rec = nullptr;   // harvested records, 12 floats each
    uint64_t rec_cap = 0;     // allocated records
    uint64_t rec_used = 0;    // capacity of the current harvest
    uint64_t rec_stride = 1;  // its stride
    float* d_xs = nullptr;    // training inputs [n][8]
    float* d_ts = nullptr;    // training targets [n][4]
    uint64_t n_set = 0, set_cap = 0;
    float* d_scratch = nullptr;  // per-slot GEMM operands of one batch
    uint32_t scratch_nb = 0;
    SgTrainBufs bufs{};
    double* d_sums = nullptr;  // 18 normalisation sums
    float* d_loss = nullptr;
    uint64_t adam_t = 0;
    float norm[kSgNorm] = {};
    std::vector<emh_sim*> users;  // sims in harvest or surrogate mode with this network

    ~emh_surrogate() {
        for (emh_sim* u : users) {  // (the caller was to detach them first)
            u->sg_attached = nullptr;
            u->sg_mode = 0;
            u->sg_host = SgDev{};
        }
        release();
    }
    void detach(emh_sim* s) { users.erase(std::remove(users.begin(), users.end(), s), users.end()); }
    bool other_harvester(const emh_sim* s) const {
        for (const emh_sim* u : users)
            if (u != s && u->sg_mode == 1) return true;
        return false;
    }
    void release() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (void* q : {(void*)d_w, (void*)d_m, (void*)d_v, (void*)d_grad, (void*)d_norm, (void*)d_pk, (void*)d_bias,
                        (void*)d_counts, (void*)d_rec, (void*)d_xs, (void*)d_ts, (void*)d_scratch, (void*)d_sums,
                        (void*)d_loss})
            if (q) (void)hipFree(q);
        if (stream) (void)hipStreamDestroy(stream);
        stream = nullptr;
    }
    SgDev dev(bool harvest) const {
        SgDev d{};
        d.pk = d_pk;
        d.bias = d_bias;
        d.norm = d_norm;
        d.counts = d_counts;
        if (harvest) {
            d.harvest = d_rec;
            d.harvest_count = d_counts + 3 * kSgCountSlots;
            d.harvest_cap = rec_cap;
        }
        return d;
    }
    int upload_norm() {
        HIP_TRY(hipMemcpyAsync(d_norm, norm, sizeof(norm), hipMemcpyHostToDevice, stream));
        return EMH_OK;
    }
    int pack() {
        HIP_TRY(sg_launch_pack(d_w, d_pk, d_bias, stream));
        return EMH_OK;
    }
    // Normalisation from the sums of x (6), residual y - x[0:3] (3), and their squares.
    void set_norm(const double* sums, uint64_t n) {
        auto sd_of = [](double mean, double sq, double cnt) {
            const double var = std::max(sq / cnt - mean * mean, 0.0);
            const double floor_ = std::max(1.0e-3 * std::fabs(mean), 1.0e-6);
            return (float)std::max(std::sqrt(var), floor_);
        };
        const double cnt = (double)std::max<uint64_t>(n, 1);
        for (int k = 0; k < kSgIn; ++k) {
            const double mean = sums[k] / cnt;
            norm[kSgInMu + k] = (float)mean;
            norm[kSgInSd + k] = sd_of(mean, sums[9 + k], cnt);
        }
        for (int k = 0; k < kSgOut; ++k) {
            const double mean = sums[6 + k] / cnt;
            norm[kSgOutMu + k] = (float)mean;
            norm[kSgOutSd + k] = sd_of(mean, sums[15 + k], cnt);
        }
        norm[kSgZMax] = 8.0f;
    }
    int ensure_set(uint64_t n) {
        if (n <= set_cap) return EMH_OK;
        if (d_xs) (void)hipFree(d_xs);
        if (d_ts) (void)hipFree(d_ts);
        d_xs = dalloc<float>(n * 8);
        d_ts = dalloc<float>(n * 4);
        set_cap = (d_xs && d_ts) ? n : 0;
        if (!set_cap) return fail(EMH_ERR_OUT_OF_MEMORY, "training-set allocation failed");
        return EMH_OK;
    }
    int ensure_scratch(uint32_t nb) {
        if (nb <= scratch_nb) return EMH_OK;
        if (d_scratch) (void)hipFree(d_scratch);
        const size_t per = 8 + 3 * kSgHid + 3 * kSgHid + 4;
        d_scratch = dalloc<float>(per * nb);
        if (!d_scratch) {
            scratch_nb = 0;
            return fail(EMH_ERR_OUT_OF_MEMORY, "training scratch allocation failed");
        }
        scratch_nb = nb;
        float* q = d_scratch;
        bufs.hin = q;
        q += 8 * (size_t)nb;
        float** hs[6] = {&bufs.h1, &bufs.h2, &bufs.h3, &bufs.dz1, &bufs.dz2, &bufs.dz3};
        for (float** h : hs) {
            *h = q;
            q += (size_t)kSgHid * nb;
        }
        bufs.dout = q;
        bufs.loss = d_loss;
        return EMH_OK;
    }
    SgBatch batch(const float* xs, const float* ts, uint64_t n_set_, uint32_t nb, uint32_t n_valid) const {
        SgBatch b{};
        b.xs = xs;
        b.ts = ts;
        b.n_set = n_set_;
        b.nb = nb;
        b.n_valid = n_valid;
        b.inv_n = 1.0f / (float)std::max<uint32_t>(n_valid, 1u);
        b.pk = d_pk;
        b.bias = d_bias;
        b.norm = d_norm;
        return b;
    }
};

emh_sim::~emh_sim() {
    if (sg_attached) sg_attached->detach(this);
    release();
}

namespace {
uint64_t gcd_u64(uint64_t a, uint64_t b) {
    while (b) {
        const uint64_t t = a % b;
        a = b;
        b = t;
    }
    return a;
}
// splitmix64
uint64_t mix64(uint64_t& s) {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
uint32_t round_up_256(uint64_t n) { return (uint32_t)((n + 255u) / 256u * 256u); }
}  // namespace

extern "C" {

int emh_surrogate_create(uint64_t seed, emh_surrogate** out) {
    if (!out) return fail(EMH_ERR_INVALID_ARG, "null out");
    *out = nullptr;
    if (device_present() == 0) return fail(EMH_ERR_DEVICE, "no HIP device is available");
    auto* sg = new emh_surrogate();
    auto bail = [&](int rc) {
        delete sg;
        return rc;
    };
    if (hipGetDevice(&sg->device) != hipSuccess || hipStreamCreateWithFlags(&sg->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(EMH_ERR_DEVICE, "stream creation failed"));
    sg->d_w = dalloc<float>(kSgParams);
    sg->d_m = dalloc<float>(kSgParams);
    sg->d_v = dalloc<float>(kSgParams);
    sg->d_grad = dalloc<float>(kSgParams);
    sg->d_norm = dalloc<float>(kSgNorm);
    sg->d_pk = dalloc<uint4>((size_t)kPkTiles * 64);
    sg->d_bias = dalloc<float>(kSgBias);
    sg->d_counts = dalloc<unsigned long long>(3 * kSgCountSlots + 1);
    sg->d_sums = dalloc<double>(18);
    sg->d_loss = dalloc<float>(1);
    if (!sg->d_w || !sg->d_m || !sg->d_v || !sg->d_grad || !sg->d_norm || !sg->d_pk || !sg->d_bias || !sg->d_counts || !sg->d_sums ||
        !sg->d_loss)
        return bail(fail(EMH_ERR_OUT_OF_MEMORY, "device allocation failed"));
    // uniform(-1, 1) * sqrt(6 / (fan_in + fan_out)) weights, zero biases
    std::vector<float> w(kSgParams, 0.0f);
    uint64_t st = seed;
    auto fill = [&](int off, int n_out, int n_in) {
        const float lim = std::sqrt(6.0f / (float)(n_in + n_out));
        for (int i = 0; i < n_out * n_in; ++i) {
            const double u = (double)(mix64(st) >> 11) * (1.0 / 9007199254740992.0);
            w[off + i] = (float)((2.0 * u - 1.0) * lim);
        }
    };
    fill(kSgW1, kSgHid, kSgIn);
    fill(kSgW2, kSgHid, kSgHid);
    fill(kSgW3, kSgHid, kSgHid);
    fill(kSgW4, kSgOut, kSgHid);
    for (int k = 0; k < kSgIn; ++k) sg->norm[kSgInSd + k] = 1.0f;
    for (int k = 0; k < kSgOut; ++k) sg->norm[kSgOutSd + k] = 1.0f;
    sg->norm[kSgZMax] = 8.0f;
    const hipStream_t s = sg->stream;
    if (hipMemcpyAsync(sg->d_w, w.data(), sizeof(float) * kSgParams, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemsetAsync(sg->d_m, 0, sizeof(float) * kSgParams, s) != hipSuccess ||
        hipMemsetAsync(sg->d_v, 0, sizeof(float) * kSgParams, s) != hipSuccess ||
        hipMemsetAsync(sg->d_grad, 0, sizeof(float) * kSgParams, s) != hipSuccess ||
        hipMemsetAsync(sg->d_counts, 0, sizeof(unsigned long long) * (3 * kSgCountSlots + 1), s) != hipSuccess)
        return bail(fail(EMH_ERR_DEVICE, "surrogate initialisation failed"));
    int rc = sg->upload_norm();
    if (!rc) rc = sg->pack();
    if (!rc && hipStreamSynchronize(s) != hipSuccess) rc = fail(EMH_ERR_DEVICE, "surrogate initialisation failed");
    if (rc) return bail(rc);
    *out = sg;
    return EMH_OK;
}

int emh_surrogate_destroy(emh_surrogate* sg) {
    delete sg;
    return EMH_OK;
}

int emh_surrogate_get_params(emh_surrogate* sg, float* params, float* norm) {
    if (!sg) return fail(EMH_ERR_INVALID_ARG, "null surrogate");
    if (params) HIP_TRY(hipMemcpyAsync(params, sg->d_w, sizeof(float) * kSgParams, hipMemcpyDeviceToHost, sg->stream));
    HIP_TRY(hipStreamSynchronize(sg->stream));
    if (norm) std::memcpy(norm, sg->norm, sizeof(sg->norm));
    return EMH_OK;
}

int emh_surrogate_set_params(emh_surrogate* sg, const float* params, const float* norm) {
    if (!sg) return fail(EMH_ERR_INVALID_ARG, "null surrogate");
    if (norm) {
        for (int k = 0; k < kSgIn; ++k)
            if (!(norm[kSgInSd + k] > 0.0f)) return fail(EMH_ERR_INVALID_ARG, "input scales must be positive");
        for (int k = 0; k < kSgOut; ++k)
            if (!(norm[kSgOutSd + k] > 0.0f)) return fail(EMH_ERR_INVALID_ARG, "output scales must be positive");
        std::memcpy(sg->norm, norm, sizeof(sg->norm));
        int rc = sg->upload_norm();
        if (rc) return rc;
    }
    if (params) {
        HIP_TRY(hipMemcpyAsync(sg->d_w, params, sizeof(float) * kSgParams, hipMemcpyHostToDevice, sg->stream));
        HIP_TRY(hipMemsetAsync(sg->d_m, 0, sizeof(float) * kSgParams, sg->stream));
        HIP_TRY(hipMemsetAsync(sg->d_v, 0, sizeof(float) * kSgParams, sg->stream));
        sg->adam_t = 0;
        int rc = sg->pack();
        if (rc) return rc;
    }
    HIP_TRY(hipStreamSynchronize(sg->stream));
    return EMH_OK;
}

int emh_sim_set_surrogate_mode(emh_sim* s, emh_surrogate* sg, uint32_t mode, uint64_t harvest_capacity,
                               uint32_t harvest_stride) {
    if (!s) return fail(EMH_ERR_INVALID_ARG, "null sim");
    if (mode > 2) return fail(EMH_ERR_INVALID_ARG, "mode must be 0 (plain), 1 (harvest) or 2 (surrogate)");
    if (mode != 0 && !sg) return fail(EMH_ERR_INVALID_ARG, "null surrogate");
    if (mode == 1 && harvest_capacity == 0) return fail(EMH_ERR_INVALID_ARG, "harvest capacity must be > 0");
    if (mode != 0 && sg->device != s->device) return fail(EMH_ERR_INVALID_ARG, "surrogate and sim are on different devices");
    if (mode == 1) {
        if (harvest_capacity > sg->rec_cap) {
            // another scene's device descriptor points at the current buffer
            if (sg->other_harvester(s))
                return fail(EMH_ERR_STATE,
                            "another simulation harvests into this surrogate's buffer: its capacity (%llu records) "
                            "cannot grow until that simulation leaves harvest mode",
                            (unsigned long long)sg->rec_cap);
            if (sg->d_rec) (void)hipFree(sg->d_rec);
            sg->d_rec = dalloc<float>(harvest_capacity * 12);
            sg->rec_cap = sg->d_rec ? harvest_capacity : 0;
            if (!sg->d_rec) return fail(EMH_ERR_OUT_OF_MEMORY, "harvest buffer allocation failed");
        }
        HIP_TRY(hipMemsetAsync(sg->d_counts + 3 * kSgCountSlots, 0, sizeof(unsigned long long), sg->stream));
        HIP_TRY(hipStreamSynchronize(sg->stream));
    }
    // drain queued work before the descriptor changes under it
    if (s->stream) HIP_TRY(hipStreamSynchronize(s->stream));
    s->sg_host = mode == 0 ? SgDev{} : sg->dev(mode == 1);
    if (mode == 1) {
        s->sg_host.harvest_cap = harvest_capacity;
        s->sg_host.harvest_stride = std::max<uint32_t>(harvest_stride, 1u);
        sg->rec_stride = s->sg_host.harvest_stride;
        sg->rec_used = harvest_capacity;
    }
    s->sg_mode = mode;
    if (s->sg_attached) s->sg_attached->detach(s);
    s->sg_attached = mode == 0 ? nullptr : sg;
    if (s->sg_attached) s->sg_attached->users.push_back(s);
    if (s->d_ctrl) {
        HIP_TRY(hipMemcpyAsync(reinterpret_cast<char*>(s->d_ctrl) + offsetof(Ctrl, sg), &s->sg_host, sizeof(SgDev),
                               hipMemcpyHostToDevice, s->stream));
        HIP_TRY(hipStreamSynchronize(s->stream));
    }
    return EMH_OK;
}

int emh_surrogate_harvest_count(emh_surrogate* sg, uint64_t* offered, uint64_t* stored) {
    if (!sg) return fail(EMH_ERR_INVALID_ARG, "null surrogate");
    unsigned long long c = 0;
    HIP_TRY(hipMemcpyAsync(&c, sg->d_counts + 3 * kSgCountSlots, sizeof(c), hipMemcpyDeviceToHost, sg->stream));
    HIP_TRY(hipStreamSynchronize(sg->stream));
    if (offered) *offered = c;
    if (stored) *stored = std::min<uint64_t>((c + sg->rec_stride - 1) / sg->rec_stride, sg->rec_used);
    return EMH_OK;
}

int emh_surrogate_get_records(emh_surrogate* sg, uint64_t first, uint64_t n, float* out) {
    if (!sg || (n && !out)) return fail(EMH_ERR_INVALID_ARG, "null argument");
    uint64_t stored = 0;
    int rc = emh_surrogate_harvest_count(sg, nullptr, &stored);
    if (rc) return rc;
    if (first + n > stored) return fail(EMH_ERR_INVALID_ARG, "range beyond the harvested records");
    if (n == 0) return EMH_OK;
    HIP_TRY(hipMemcpyAsync(out, sg->d_rec + first * 12, sizeof(float) * 12 * n, hipMemcpyDeviceToHost, sg->stream));
    HIP_TRY(hipStreamSynchronize(sg->stream));
    return EMH_OK;
}

int emh_surrogate_eval_records(emh_surrogate* sg, const emh_material_params* p, const emh_material_ext* ext,
                               uint64_t first, uint64_t n, int use_surrogate, uint32_t reps, float* ms_per_rep,
                               float* f_out, float* alpha_out) {
    if (!sg || !p) return fail(EMH_ERR_INVALID_ARG, "null argument");
    uint64_t stored = 0;
    int rc = emh_surrogate_harvest_count(sg, nullptr, &stored);
    if (rc) return rc;
    if (n == 0 || first + n > stored) return fail(EMH_ERR_INVALID_ARG, "range beyond the harvested records");
    float* d = dalloc<float>(n * 10);  // trial F, alpha, then the outputs
    if (!d) return fail(EMH_ERR_OUT_OF_MEMORY, "device allocation failed");
    float *df = d, *da = d + 4 * n, *dfo = d + 5 * n, *dao = d + 9 * n;
    const float* src = sg->d_rec + first * 12;
    hipError_t e = hipMemcpy2DAsync(df, 16, src, 48, 16, n, hipMemcpyDeviceToDevice, sg->stream);
    if (e == hipSuccess) e = hipMemcpy2DAsync(da, 4, src + 4, 48, 4, n, hipMemcpyDeviceToDevice, sg->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(sg->stream);
    if (e != hipSuccess) rc = fail(EMH_ERR_DEVICE, "staging: %s", hipGetErrorString(e));
    if (!rc) rc = emh_nacc_return_batch(use_surrogate ? sg : nullptr, p, ext, df, da, n, dfo, dao, reps, ms_per_rep);
    if (!rc && f_out) e = hipMemcpy(f_out, dfo, sizeof(float) * 4 * n, hipMemcpyDeviceToHost);
    if (!rc && e == hipSuccess && alpha_out) e = hipMemcpy(alpha_out, dao, sizeof(float) * n, hipMemcpyDeviceToHost);
    if (!rc && e != hipSuccess) rc = fail(EMH_ERR_DEVICE, "copy-out: %s", hipGetErrorString(e));
    (void)hipFree(d);
    return rc;
}

int emh_surrogate_build_training_set(emh_surrogate* sg, uint64_t* n_samples) {
    if (!sg) return fail(EMH_ERR_INVALID_ARG, "null surrogate");
    uint64_t offered = 0, n_rec = 0;
    int rc = emh_surrogate_harvest_count(sg, &offered, &n_rec);
    if (rc) return rc;
    if (n_rec == 0) return fail(EMH_ERR_STATE, "no harvested trial states");
    rc = sg->ensure_set(n_rec);
    if (rc) return rc;
    unsigned long long* d_n = sg->d_counts;  // reuse count slot 0 as the output counter, reset below
    HIP_TRY(hipMemsetAsync(d_n, 0, sizeof(unsigned long long), sg->stream));
    HIP_TRY(hipMemsetAsync(sg->d_sums, 0, sizeof(double) * 18, sg->stream));
    HIP_TRY(sg_launch_prepare(sg->d_rec, n_rec, sg->d_xs, sg->d_ts, d_n, sg->d_sums, sg->stream));
    unsigned long long n = 0;
    double sums[18];
    HIP_TRY(hipMemcpyAsync(&n, d_n, sizeof(n), hipMemcpyDeviceToHost, sg->stream));
    HIP_TRY(hipMemcpyAsync(sums, sg->d_sums, sizeof(sums), hipMemcpyDeviceToHost, sg->stream));
    HIP_TRY(hipMemsetAsync(d_n, 0, sizeof(unsigned long long), sg->stream));
    HIP_TRY(hipStreamSynchronize(sg->stream));
    sg->n_set = n;
    if (n == 0) return fail(EMH_ERR_STATE, "no harvested trial state fails the yield check");
    sg->set_norm(sums, n);
    rc = sg->upload_norm();
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(sg->stream));
    if (n_samples) *n_samples = n;
    return EMH_OK;
}

int emh_surrogate_set_training_set(emh_surrogate* sg, const float* x, const float* y, uint64_t n) {
    if (!sg || (n && (!x || !y))) return fail(EMH_ERR_INVALID_ARG, "null argument");
    if (n == 0) return fail(EMH_ERR_INVALID_ARG, "empty training set");
    int rc = sg->ensure_set(n);
    if (rc) return rc;
    std::vector<float> xs(n * 8, 0.0f), ts(n * 4, 0.0f);
    double sums[18] = {};
    for (uint64_t i = 0; i < n; ++i) {
        for (int k = 0; k < kSgIn; ++k) {
            const double v = x[i * kSgIn + k];
            xs[i * 8 + k] = x[i * kSgIn + k];
            sums[k] += v;
            sums[9 + k] += v * v;
        }
        for (int k = 0; k < kSgOut; ++k) {
            ts[i * 4 + k] = y[i * kSgOut + k];
            const double d = (double)y[i * kSgOut + k] - (double)x[i * kSgIn + k];
            sums[6 + k] += d;
            sums[15 + k] += d * d;
        }
    }
    HIP_TRY(hipMemcpyAsync(sg->d_xs, xs.data(), sizeof(float) * xs.size(), hipMemcpyHostToDevice, sg->stream));
    HIP_TRY(hipMemcpyAsync(sg->d_ts, ts.data(), sizeof(float) * ts.size(), hipMemcpyHostToDevice, sg->stream));
    sg->n_set = n;
    sg->set_norm(sums, n);
    rc = sg->upload_norm();
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(sg->stream));
    return EMH_OK;
}

int emh_surrogate_get_training_set(emh_surrogate* sg, uint64_t first, uint64_t n, float* x, float* y) {
    if (!sg || (n && (!x || !y))) return fail(EMH_ERR_INVALID_ARG, "null argument");
    if (first + n > sg->n_set) return fail(EMH_ERR_INVALID_ARG, "range beyond the training set");
    if (n == 0) return EMH_OK;
    std::vector<float> xs(n * 8), ts(n * 4);
    HIP_TRY(hipMemcpyAsync(xs.data(), sg->d_xs + first * 8, sizeof(float) * xs.size(), hipMemcpyDeviceToHost, sg->stream));
    HIP_TRY(hipMemcpyAsync(ts.data(), sg->d_ts + first * 4, sizeof(float) * ts.size(), hipMemcpyDeviceToHost, sg->stream));
    HIP_TRY(hipStreamSynchronize(sg->stream));
    for (uint64_t i = 0; i < n; ++i) {
        for (int k = 0; k < kSgIn; ++k) x[i * kSgIn + k] = xs[i * 8 + k];
        for (int k = 0; k < kSgOut; ++k) y[i * kSgOut + k] = ts[i * 4 + k];
    }
    return EMH_OK;
}

int emh_surrogate_train(emh_surrogate* sg, const emh_surrogate_train_options* o, float* epoch_loss,
                        uint32_t n_epoch_loss) {
    if (!sg || !o) return fail(EMH_ERR_INVALID_ARG, "null argument");
    if (sg->n_set == 0) return fail(EMH_ERR_STATE, "no training set");
    if (o->epochs == 0 || o->batch == 0) return fail(EMH_ERR_INVALID_ARG, "epochs and batch must be > 0");
    if (!(o->lr > 0.0f) || !(o->lr_final >= 0.0f) || !(o->beta1 >= 0.0f && o->beta1 < 1.0f) ||
        !(o->beta2 >= 0.0f && o->beta2 < 1.0f) || !(o->eps > 0.0f))
        return fail(EMH_ERR_INVALID_ARG, "invalid Adam hyper-parameters");
    const uint32_t nb = round_up_256(o->batch);
    if (nb > sg->n_set) return fail(EMH_ERR_INVALID_ARG, "batch larger than the training set");
    int rc = sg->ensure_scratch(nb);
    if (rc) return rc;
    const uint64_t per_epoch = sg->n_set / nb;
    uint64_t total = (uint64_t)o->epochs * per_epoch;
    if (o->max_steps && o->max_steps < total) total = o->max_steps;
    uint64_t rng = o->seed;
    SgBatch b = sg->batch(sg->d_xs, sg->d_ts, sg->n_set, nb, nb);
    uint64_t step = 0;
    for (uint32_t e = 0; e < o->epochs && step < total; ++e) {
        // this epoch's visiting order: q -> (a q + c) mod n, a coprime to n
        uint64_t a = 1 + mix64(rng) % (sg->n_set - 1 > 0 ? sg->n_set - 1 : 1);
        while (gcd_u64(a, sg->n_set) != 1) a = a % sg->n_set + 1;
        b.perm_a = a;
        b.perm_b = mix64(rng) % sg->n_set;
        HIP_TRY(hipMemsetAsync(sg->d_loss, 0, sizeof(float), sg->stream));
        uint64_t done = 0;
        for (uint64_t k = 0; k < per_epoch && step < total; ++k, ++step, ++done) {
            b.first = k * nb;
            const double frac = total > 1 ? (double)step / (double)(total - 1) : 1.0;
            SgAdam h{};
            h.lr = (float)(o->lr_final + 0.5 * (o->lr - o->lr_final) * (1.0 + std::cos(3.141592653589793 * frac)));
            h.beta1 = o->beta1;
            h.beta2 = o->beta2;
            h.eps = o->eps;
            ++sg->adam_t;
            h.c1 = (float)(1.0 / (1.0 - std::pow((double)o->beta1, (double)sg->adam_t)));
            h.c2 = (float)(1.0 / (1.0 - std::pow((double)o->beta2, (double)sg->adam_t)));
            HIP_TRY(sg_launch_train_step(b, sg->bufs, sg->d_w, sg->d_m, sg->d_v, sg->d_grad, &h, sg->d_bias, sg->stream));
        }
        float l = 0.0f;
        HIP_TRY(hipMemcpyAsync(&l, sg->d_loss, sizeof(float), hipMemcpyDeviceToHost, sg->stream));
        HIP_TRY(hipStreamSynchronize(sg->stream));
        if (epoch_loss && e < n_epoch_loss) epoch_loss[e] = done ? l / (float)(done * nb) : 0.0f;
    }
    return EMH_OK;
}

int emh_surrogate_gradient(emh_surrogate* sg, const float* x, const float* y, uint32_t n, float* grad, float* loss) {
    if (!sg || !x || !y || n == 0) return fail(EMH_ERR_INVALID_ARG, "null or empty argument");
    const uint32_t nb = round_up_256(n);
    int rc = sg->ensure_scratch(nb);
    if (rc) return rc;
    std::vector<float> xs((size_t)n * 8, 0.0f), ts((size_t)n * 4, 0.0f);
    for (uint32_t i = 0; i < n; ++i) {
        for (int k = 0; k < kSgIn; ++k) xs[(size_t)i * 8 + k] = x[(size_t)i * kSgIn + k];
        for (int k = 0; k < kSgOut; ++k) ts[(size_t)i * 4 + k] = y[(size_t)i * kSgOut + k];
    }
    float* d_x = dalloc<float>(xs.size());
    float* d_t = dalloc<float>(ts.size());
    rc = EMH_OK;
    if (!d_x || !d_t) rc = fail(EMH_ERR_OUT_OF_MEMORY, "device allocation failed");
    hipError_t e = hipSuccess;
    float l = 0.0f;
    if (!rc) {
        const SgBatch b = sg->batch(d_x, d_t, n, nb, n);
        e = hipMemcpyAsync(d_x, xs.data(), sizeof(float) * xs.size(), hipMemcpyHostToDevice, sg->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(d_t, ts.data(), sizeof(float) * ts.size(), hipMemcpyHostToDevice, sg->stream);
        if (e == hipSuccess) e = hipMemsetAsync(sg->d_grad, 0, sizeof(float) * kSgParams, sg->stream);
        if (e == hipSuccess) e = hipMemsetAsync(sg->d_loss, 0, sizeof(float), sg->stream);
        if (e == hipSuccess) e = sg_launch_train_step(b, sg->bufs, sg->d_w, sg->d_m, sg->d_v, sg->d_grad, nullptr, sg->d_bias, sg->stream);
        if (e == hipSuccess && grad)
            e = hipMemcpyAsync(grad, sg->d_grad, sizeof(float) * kSgParams, hipMemcpyDeviceToHost, sg->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(&l, sg->d_loss, sizeof(float), hipMemcpyDeviceToHost, sg->stream);
        if (e == hipSuccess) e = hipMemsetAsync(sg->d_grad, 0, sizeof(float) * kSgParams, sg->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(sg->stream);
        if (e != hipSuccess) rc = fail(EMH_ERR_DEVICE, "gradient: %s", hipGetErrorString(e));
    }
    if (d_x) (void)hipFree(d_x);
    if (d_t) (void)hipFree(d_t);
    if (!rc && loss) *loss = l / (float)n;
    return rc;
}

int emh_surrogate_infer(emh_surrogate* sg, const float* x, uint64_t n, float* y) {
    if (!sg || (n && (!x || !y))) return fail(EMH_ERR_INVALID_ARG, "null argument");
    if (n == 0) return EMH_OK;
    float* d_x = dalloc<float>(n * kSgIn);
    float* d_y = dalloc<float>(n * kSgOut);
    int rc = EMH_OK;
    if (!d_x || !d_y) rc = fail(EMH_ERR_OUT_OF_MEMORY, "device allocation failed");
    if (!rc) {
        hipError_t e = hipMemcpyAsync(d_x, x, sizeof(float) * n * kSgIn, hipMemcpyHostToDevice, sg->stream);
        if (e == hipSuccess) e = sg_launch_infer(d_x, n, d_y, sg->dev(false), sg->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(y, d_y, sizeof(float) * n * kSgOut, hipMemcpyDeviceToHost, sg->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(sg->stream);
        if (e != hipSuccess) rc = fail(EMH_ERR_DEVICE, "infer: %s", hipGetErrorString(e));
    }
    if (d_x) (void)hipFree(d_x);
    if (d_y) (void)hipFree(d_y);
    return rc;
}

int emh_surrogate_get_counts(emh_surrogate* sg, emh_surrogate_counts* out, int reset) {
    if (!sg || !out) return fail(EMH_ERR_INVALID_ARG, "null argument");
    std::vector<unsigned long long> c(3 * kSgCountSlots);
    HIP_TRY(hipMemcpyAsync(c.data(), sg->d_counts, sizeof(unsigned long long) * c.size(), hipMemcpyDeviceToHost,
                           sg->stream));
    if (reset) HIP_TRY(hipMemsetAsync(sg->d_counts, 0, sizeof(unsigned long long) * c.size(), sg->stream));
    HIP_TRY(hipStreamSynchronize(sg->stream));
    uint64_t sum[3] = {0, 0, 0};
    for (int k = 0; k < 3; ++k)
        for (int i = 0; i < kSgCountSlots; ++i) sum[k] += c[(size_t)k * kSgCountSlots + i];
    out->yielding = sum[0];
    out->surrogate = sum[1];
    out->fallback = sum[2];
    return EMH_OK;
}

int emh_nacc_return_batch(emh_surrogate* sg, const emh_material_params* p, const emh_material_ext* ext,
                          const float* f_trial, const float* alpha, uint64_t n, float* f_out, float* alpha_out,
                          uint32_t reps, float* ms_per_rep) {
    if (!p || (n && (!f_trial || !alpha || !f_out || !alpha_out))) return fail(EMH_ERR_INVALID_ARG, "null argument");
    if (p->model != EMH_MODEL_NACC) return fail(EMH_ERR_INVALID_ARG, "material is not NACC");
    emh_material_ext e{};
    if (ext)
        e = *ext;
    else
        ext_defaults(&e);
    const DevMat m = make_dev_mat(*p, e);
    hipStream_t st = nullptr;
    bool own = false;
    if (sg) {
        st = sg->stream;
    } else {
        if (device_present() == 0) return fail(EMH_ERR_DEVICE, "no HIP device is available");
        HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        own = true;
    }
    hipEvent_t ev[2] = {nullptr, nullptr};
    int rc = EMH_OK;
    hipError_t err = hipEventCreate(&ev[0]);
    if (err == hipSuccess) err = hipEventCreate(&ev[1]);
    const uint32_t r = std::max<uint32_t>(reps, 1u);
    if (err == hipSuccess) err = hipEventRecord(ev[0], st);
    for (uint32_t k = 0; k < r && err == hipSuccess; ++k) {
        if (sg)
            err = sg_launch_return_surrogate(m, reinterpret_cast<const float4*>(f_trial), alpha, n,
                                             reinterpret_cast<float4*>(f_out), alpha_out, sg->dev(false), st);
        else
            err = sg_launch_return_plain(m, reinterpret_cast<const float4*>(f_trial), alpha, n,
                                         reinterpret_cast<float4*>(f_out), alpha_out, st);
    }
    if (err == hipSuccess) err = hipEventRecord(ev[1], st);
    if (err == hipSuccess) err = hipEventSynchronize(ev[1]);
    float ms = 0.0f;
    if (err == hipSuccess) err = hipEventElapsedTime(&ms, ev[0], ev[1]);
    if (err != hipSuccess) rc = fail(EMH_ERR_DEVICE, "return batch: %s", hipGetErrorString(err));
    if (ms_per_rep) *ms_per_rep = ms / (float)r;
    for (auto& x : ev)
        if (x) (void)hipEventDestroy(x);
    if (own) (void)hipStreamDestroy(st);
    return rc;
}

}  // extern "C"

// ------------------------------------------------------------ batched scenes
// Many independent simulations stepped together: one launch per stage over
// all of them (kernels.hip, "batched scenes").  Each keeps its own
// device-side control, so each takes its own adaptive substeps.
struct emh_batch {
    std::vector<emh_sim*> sims;
    int device = 0;
    hipStream_t stream = nullptr;
    Pipeline* d_pipes = nullptr;
    uint32_t* d_starts = nullptr;  // five work lists of n + 1 entries
    Ctrl* d_ctrl = nullptr;        // gathered control blocks
    std::vector<Ctrl> h_ctrl;
    BatchLists L{};
    uint64_t seq = 0;
    uint32_t reorder_substeps = kBatchReorderSubsteps;  // emh_batch_set_tuning
    uint32_t substeps_since_reorder = kBatchReorderSubsteps;
    uint32_t predicted = 1;
    uint32_t max_substeps = 1;
    std::vector<uint64_t> epochs;  // each scene's cfg_epoch in d_pipes
    // graph replay of runs of batch substep launches (the work lists are the
    // launches' arguments; the scenes' pipelines are read from d_pipes)
    bool use_graphs = false;
    GraphCache graphs;

    // Live timing of the batch's stages (emh_batch_set_profiling): event marks
    // on the batch stream before each stage's launches, recorded without
    // stopping the queue and read at the drains, whose synchronisation has
    // completed them.  A span runs from a mark to the next one.
    enum : uint8_t { kStBegin = 0, kStSort = 1, kStSubsteps = 2, kStGather = 3, kStages = 4, kStIdle = 4 };
    bool profiling = false;
    std::vector<hipEvent_t> prof_pool;
    std::vector<uint8_t> prof_kind;
    size_t prof_used = 0;
    double prof_ms[kStages] = {};
    uint64_t prof_count[kStages] = {};
    int prof_mark(uint8_t kind) {
        if (!profiling) return EMH_OK;
        if (prof_used == prof_pool.size()) {
            hipEvent_t e = nullptr;  // timing only: no system-scope fence (see emh_sim::prof_mark)
            HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
            prof_pool.push_back(e);
            prof_kind.push_back(kStIdle);
        }
        HIP_TRY(hipEventRecord(prof_pool[prof_used], stream));
        prof_kind[prof_used++] = kind;
        return EMH_OK;
    }
    int prof_fold() {
        for (size_t i = 0; i + 1 < prof_used; ++i) {
            if (prof_kind[i] >= kStages) continue;
            float ms = 0.0f;
            HIP_TRY(hipEventElapsedTime(&ms, prof_pool[i], prof_pool[i + 1]));
            prof_ms[prof_kind[i]] += ms;
        }
        prof_used = 0;
        return EMH_OK;
    }
    void prof_count_add(uint8_t kind, uint64_t n) {
        if (profiling) prof_count[kind] += n;
    }

    ~emh_batch() {
        if (stream) (void)hipStreamSynchronize(stream);
        graphs.clear();
        for (auto& e : prof_pool) (void)hipEventDestroy(e);
        if (d_pipes) (void)hipFree(d_pipes);
        if (d_starts) (void)hipFree(d_starts);
        if (d_ctrl) (void)hipFree(d_ctrl);
        if (stream) (void)hipStreamDestroy(stream);
    }

    // Bring every scene's loop to the same launch sequence (each restarts
    // from its stored particles, as an upload does) and rebuild the work lists.
    int sync() {
        const uint32_t n = (uint32_t)sims.size();
        std::vector<Pipeline> pipes(n);
        epochs.assign(n, 0);
        std::vector<uint32_t> st(5 * (n + 1), 0u);
        uint32_t* one = st.data();
        uint32_t* sorts = one + (n + 1);
        uint32_t* fused = sorts + (n + 1);
        uint32_t* dens = fused + (n + 1);
        uint32_t* scans = dens + (n + 1);
        max_substeps = 1;
        predicted = 1;
        uint32_t passes = 0;
        for (uint32_t i = 0; i < n; ++i) {
            emh_sim* s = sims[i];
            int rc;
            if (s->host_dirty) {
                rc = s->upload();
            } else {
                CtrlState cs;
                rc = s->read_ctrl(cs);
                if (!rc) rc = s->restart_loop(cs.cur_buf);
            }
            if (rc) return rc;
            if (s->pipe.n == 0) return fail(EMH_ERR_INVALID_ARG, "scene %u of the batch has no particles", i);
            if (s->sg_mode != 0)
                return fail(EMH_ERR_UNSUPPORTED, "scene %u uses the surrogate harvest / inference mode", i);
            if (s->sleep) return fail(EMH_ERR_UNSUPPORTED, "scene %u uses the sleep partition", i);
            if (s->seq != 1) return fail(EMH_ERR_STATE, "scene %u did not restart its loop", i);
            pipes[i] = s->pipe;
            epochs[i] = s->cfg_epoch;
            const uint32_t chunks = (s->pipe.n + kChunk - 1) / kChunk;
            const bool d = s->pipe.cfg.density_pass != 0;
            one[i + 1] = one[i] + 1;
            sorts[i + 1] = sorts[i] + sort_tiles(s->pipe.n);
            const uint32_t key_bits = cell_key_bits(s->n_bins()) + s->pipe.sort.model_bits;
            scans[i + 1] = scans[i] + sort_scan_groups(key_bits);
            passes = std::max<uint32_t>(passes, sort_plan_bits(key_bits).passes);
            fused[i + 1] = fused[i] + (d ? 0u : chunks);
            dens[i + 1] = dens[i] + (d ? chunks : 0u);
            max_substeps = std::max<uint32_t>(max_substeps, s->cfg.max_substeps_per_step);
            predicted = std::max<uint32_t>(predicted, s->predicted_substeps);
        }
        HIP_TRY(hipMemcpyAsync(d_pipes, pipes.data(), sizeof(Pipeline) * n, hipMemcpyHostToDevice, stream));
        HIP_TRY(hipMemcpyAsync(d_starts, st.data(), sizeof(uint32_t) * st.size(), hipMemcpyHostToDevice, stream));
        HIP_TRY(hipStreamSynchronize(stream));
        auto view = [&](int k) { return BatchView{d_pipes, d_starts + (size_t)k * (n + 1), n}; };
        L.one = view(0);
        L.sort = view(1);
        L.fused = view(2);
        L.density = view(3);
        L.scan = view(4);
        L.sort_blocks = sorts[n];
        L.scan_blocks = scans[n];
        L.sort_passes = passes;
        L.fused_blocks = fused[n];
        L.density_blocks = dens[n];
        seq = 1;
        substeps_since_reorder = reorder_substeps;
        return EMH_OK;
    }

    bool in_sync() const {
        for (emh_sim* s : sims)
            if (s->host_dirty || s->seq != seq || s->sg_mode != 0) return false;
        return seq != 0;
    }

    // Wait for every scene's queued steps, launching further substeps while
    // any scene still runs one; then publish each scene's last step.
    int drain() {
        const uint32_t n = (uint32_t)sims.size();
        uint32_t extra = 0;
        const uint32_t limit = (max_substeps + 1) * (kStepRing + 1);
        while (true) {
            int rc = prof_mark(kStGather);
            if (rc) return rc;
            HIP_TRY(launch_batch_gather_ctrl(L, d_ctrl, stream));
            rc = prof_mark(kStIdle);
            if (rc) return rc;
            prof_count_add(kStGather, 1);
            HIP_TRY(hipMemcpyAsync(h_ctrl.data(), d_ctrl, sizeof(Ctrl) * n, hipMemcpyDeviceToHost, stream));
            HIP_TRY(hipStreamSynchronize(stream));
            uint32_t more = 0;
            for (uint32_t i = 0; i < n; ++i) {
                const CtrlState& st = h_ctrl[i].st[seq & 1u];
                if (!sims[i]->running(st) && st.pending == 0) continue;
                more = std::max<uint32_t>(more, (sims[i]->running(st) ? 1u : 0u) +
                                                    st.pending * std::max<uint32_t>(sims[i]->predicted_substeps, 1u));
            }
            if (more == 0) break;
            more = std::min<uint32_t>(more, 64u);
            if (extra + more > limit) return fail(EMH_ERR_DEVICE, "batched substep loop did not terminate");
            rc = prof_mark(kStSubsteps);
            if (rc) return rc;
            prof_count_add(kStSubsteps, more);
            for (uint32_t k = 0; k < more; ++k) HIP_TRY(launch_batch_substep(L, seq++, stream));
            extra += more;
        }
        {
            int rc = prof_fold();  // the last synchronisation completed every mark
            if (rc) return rc;
        }
        predicted = 1;
        for (uint32_t i = 0; i < n; ++i) {
            int rc = sims[i]->publish(h_ctrl[i], h_ctrl[i].st[seq & 1u]);
            if (rc) return rc;
            predicted = std::max<uint32_t>(predicted, sims[i]->predicted_substeps);
        }
        return EMH_OK;
    }

    // Scenes whose configuration was edited in place since the batch copied
    // their pipelines (set_gravity): refresh those copies, stream-ordered
    // after the launches already queued.
    int refresh_configs() {
        for (size_t i = 0; i < sims.size(); ++i) {
            if (sims[i]->cfg_epoch == epochs[i]) continue;
            HIP_TRY(hipMemcpyAsync(d_pipes + i, &sims[i]->pipe, sizeof(Pipeline), hipMemcpyHostToDevice, stream));
            epochs[i] = sims[i]->cfg_epoch;
        }
        return EMH_OK;
    }

    // Launch settings of the batch; a change of the scenes' model grouping
    // rebuilds the work lists (the sort's pass count) at the next step.
    bool lists_stale = false;
    void apply_tuning(const emh_tuning& t) {
        reorder_substeps = t.reorder_interval;
        use_graphs = t.graphs != 0;
        for (emh_sim* sim : sims) {
            if (sim->model_segments == (t.model_segments != 0)) continue;
            emh_tuning st = sim->tuning();
            st.model_segments = t.model_segments;
            sim->apply_tuning(st);
            lists_stale = true;
        }
    }
    emh_tuning tuning() const {
        emh_tuning t{};
        t.reorder_interval = reorder_substeps;
        t.graphs = use_graphs ? 1u : 0u;
        t.model_segments = !sims.empty() && sims[0]->model_segments ? 1u : 0u;
        return t;
    }

    // The batch counterpart of emh_sim::autotune: every scene is put back to
    // its saved state between candidates and after the last.
    int autotune(uint32_t trial, emh_tuning* best, emh_tuning* timed, uint32_t cap, uint32_t* n_timed) {
        if (trial == 0) return fail(EMH_ERR_INVALID_ARG, "trial_steps must be at least 1");
        if (profiling) return fail(EMH_ERR_STATE, "turn profiling off before autotuning");
        const size_t n = sims.size();
        std::vector<std::vector<emh_particle>> saved(n);
        std::vector<emh_step_stats> saved_stats(n);
        std::vector<uint32_t> saved_pred(n);
        for (size_t i = 0; i < n; ++i) {
            int rc = sims[i]->host_current();
            if (rc) return rc;
            saved[i] = sims[i]->host;
            saved_stats[i] = sims[i]->stats;
            saved_pred[i] = sims[i]->predicted_substeps;
        }
        const emh_tuning saved_tuning = tuning();
        auto restore = [&]() {
            for (size_t i = 0; i < n; ++i) {
                sims[i]->host = saved[i];
                sims[i]->host_dirty = true;
                sims[i]->device_ahead = false;
                sims[i]->stats = saved_stats[i];
                sims[i]->predicted_substeps = saved_pred[i];
            }
        };
        bool mixed = false;
        for (emh_sim* sim : sims) mixed = mixed || sim->mixed_models();
        const std::vector<emh_tuning> cands = batch_tuning_candidates(mixed);
        std::vector<float> cost(cands.size(), 0.0f);
        int rc = EMH_OK;
        for (size_t c = 0; c < cands.size() && rc == EMH_OK; ++c) {
            restore();
            apply_tuning(cands[c]);
            rc = step(1);
            if (rc) break;
            std::vector<uint64_t> s0(n);
            for (size_t i = 0; i < n; ++i) s0[i] = sims[i]->stats.total_substeps;
            const auto t0 = std::chrono::steady_clock::now();
            rc = step(trial);
            const auto t1 = std::chrono::steady_clock::now();
            double work = 0.0;
            for (size_t i = 0; i < n; ++i)
                work += (double)(sims[i]->stats.total_substeps - s0[i]) * (double)sims[i]->pipe.n;
            cost[c] = (float)(std::chrono::duration<double, std::nano>(t1 - t0).count() / std::max(work, 1.0));
        }
        restore();
        apply_tuning(saved_tuning);
        if (rc) return rc;
        size_t bi = 0;
        for (size_t c = 1; c < cands.size(); ++c)
            if (cost[c] < cost[bi]) bi = c;
        for (size_t c = 0; c < cands.size() && c < cap && timed; ++c) {
            timed[c] = cands[c];
            timed[c].ns_per_particle_substep = cost[c];
        }
        if (n_timed) *n_timed = (uint32_t)cands.size();
        apply_tuning(cands[bi]);
        if (best) {
            *best = cands[bi];
            best->ns_per_particle_substep = cost[bi];
        }
        return EMH_OK;
    }

    int step(uint32_t n_steps) {
        if (!in_sync() || lists_stale) {
            int rc = sync();
            if (rc) return rc;
            lists_stale = false;
        }
        int rc = refresh_configs();
        if (rc) return rc;
        constexpr uint32_t kDrainSteps = kStepRing / 2;
        for (uint32_t s = 0; s < n_steps; ++s) {
            const bool reorder = substeps_since_reorder >= reorder_substeps;
            rc = prof_mark(kStBegin);
            if (rc) return rc;
            prof_count_add(kStBegin, 1);
            HIP_TRY(launch_batch_begin_step(L, seq, reorder, stream));
            if (reorder) {
                rc = prof_mark(kStSort);
                if (rc) return rc;
                prof_count_add(kStSort, 1);
                HIP_TRY(launch_batch_sort(L, seq, stream));
                substeps_since_reorder = 0;
            }
            ++seq;
            const uint32_t slots = std::min<uint32_t>(std::max<uint32_t>(predicted, 1u), max_substeps);
            rc = prof_mark(kStSubsteps);
            if (rc) return rc;
            prof_count_add(kStSubsteps, slots);
            if (use_graphs) {
                graphs.check_arg(&L, sizeof(L));
                const uint32_t r = (uint32_t)(seq % 6u);
                hipGraphExec_t ex = graphs.find(r, slots, 0u);
                if (!ex) {
                    const uint64_t s0 = seq;
                    HIP_TRY(graphs.capture(stream, r, slots, 0u,
                                           [&](uint32_t k) { return launch_batch_substep(L, s0 + k, stream); }, &ex));
                }
                HIP_TRY(hipGraphLaunch(ex, stream));
                seq += slots;
            } else {
                for (uint32_t k = 0; k < slots; ++k) HIP_TRY(launch_batch_substep(L, seq++, stream));
            }
            substeps_since_reorder += slots;
            if ((s + 1) % kDrainSteps == 0 || s + 1 == n_steps) {
                int rc = drain();
                if (rc) return rc;
            }
        }
        for (emh_sim* sim : sims) {
            sim->seq = seq;
            sim->device_ahead = true;
            sim->stats.frame_index += n_steps;
        }
        return EMH_OK;
    }
};

extern "C" {

int emh_batch_create(emh_sim* const* sims, uint32_t n, emh_batch** out) {
    if (!out || !sims || n == 0) return fail(EMH_ERR_INVALID_ARG, "null or empty argument");
    *out = nullptr;
    auto* b = new emh_batch();
    auto bail = [&](int rc) {
        delete b;
        return rc;
    };
    for (uint32_t i = 0; i < n; ++i) {
        if (!sims[i]) return bail(fail(EMH_ERR_INVALID_ARG, "scene %u is null", i));
        for (uint32_t k = 0; k < i; ++k)
            if (sims[k] == sims[i]) return bail(fail(EMH_ERR_INVALID_ARG, "scene %u appears twice", i));
        if (sims[i]->sg_mode != 0)
            return bail(fail(EMH_ERR_UNSUPPORTED, "scene %u uses the surrogate harvest / inference mode", i));
        if (sims[i]->cfg.sleep_threshold > 0.0f || sims[i]->sleep)
            return bail(fail(EMH_ERR_UNSUPPORTED, "scene %u uses the sleep partition", i));
        if (i > 0 && sims[i]->device != sims[0]->device)
            return bail(fail(EMH_ERR_INVALID_ARG, "scenes are on different devices"));
        b->sims.push_back(sims[i]);
    }
    b->device = sims[0]->device;
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(EMH_ERR_DEVICE, "stream creation failed"));
    b->d_pipes = dalloc<Pipeline>(n);
    b->d_starts = dalloc<uint32_t>(5 * ((size_t)n + 1));
    b->d_ctrl = dalloc<Ctrl>(n);
    if (!b->d_pipes || !b->d_starts || !b->d_ctrl) return bail(fail(EMH_ERR_OUT_OF_MEMORY, "device allocation failed"));
    b->h_ctrl.resize(n);
    int rc = b->sync();
    if (rc) return bail(rc);
    *out = b;
    return EMH_OK;
}

int emh_batch_destroy(emh_batch* b) {
    delete b;
    return EMH_OK;
}

int emh_batch_step(emh_batch* b, uint32_t n_steps) {
    if (!b) return fail(EMH_ERR_INVALID_ARG, "null batch");
    return b->step(n_steps);
}

int emh_batch_get_tuning(emh_batch* b, emh_tuning* out) {
    if (!b || !out) return fail(EMH_ERR_INVALID_ARG, "null argument");
    *out = b->tuning();
    return EMH_OK;
}

int emh_batch_set_tuning(emh_batch* b, const emh_tuning* t) {
    if (!b || !t) return fail(EMH_ERR_INVALID_ARG, "null argument");
    int rc = check_tuning(t);
    if (rc) return rc;
    b->apply_tuning(*t);
    return EMH_OK;
}

int emh_batch_autotune(emh_batch* b, uint32_t trial_steps, emh_tuning* chosen, emh_tuning* timed, uint32_t timed_cap,
                       uint32_t* n_timed) {
    if (!b) return fail(EMH_ERR_INVALID_ARG, "null batch");
    return b->autotune(trial_steps, chosen, timed, timed_cap, n_timed);
}

int emh_batch_set_profiling(emh_batch* b, int enable) {
    if (!b) return fail(EMH_ERR_INVALID_ARG, "null batch");
    b->profiling = enable != 0;
    b->prof_used = 0;
    for (int k = 0; k < emh_batch::kStages; ++k) {
        b->prof_ms[k] = 0.0;
        b->prof_count[k] = 0;
    }
    return EMH_OK;
}

int emh_batch_kernel_times(emh_batch* b, emh_kernel_time* out, uint32_t cap, uint32_t* n_out) {
    if (!b || (!out && cap)) return fail(EMH_ERR_INVALID_ARG, "null argument");
    static const char* const names[emh_batch::kStages] = {"kb_begin_step", "kb_sort", "kb_substep",
                                                          "kb_gather_ctrl"};
    uint32_t n = 0;
    for (int k = 0; k < emh_batch::kStages; ++k) {
        if (n < cap) {
            std::memset(&out[n], 0, sizeof(out[n]));
            std::snprintf(out[n].name, sizeof(out[n].name), "%s", names[k]);
            out[n].launches = b->prof_count[k];
            out[n].total_ms = b->prof_ms[k];
        }
        ++n;
    }
    if (n_out) *n_out = n;
    return EMH_OK;
}

}  // extern "C"
