// oracle_capi.cpp -- extern "C" surface of the CPU oracle (mpm_oracle.hpp).
//
// TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py through ctypes, never by the product library.
// `precision` 32 runs the f32 restatement of the reference's Simulation::step;
// 64 runs the same algorithm in f64 (the FP64 oracle).
#include <algorithm>
#include <chrono>
#include <cstring>
#include <numeric>

#include "mpm_oracle.hpp"

using namespace orc;

namespace {

struct Handle {
    int precision = 32;
    Simulation<float> s32;
    Simulation<double> s64;
};

template <class F> auto with_sim(Handle* h, F&& f) {
    if (h->precision == 64) return f(h->s64);
    return f(h->s32);
}

bool valid_config(const emh_config& c) {  // SimConfig::validate  config.rs:247-299
    if (c.grid_res < 4) return false;
    if (!(c.grid_cell_size > 0.0f) || !(c.dt > 0.0f)) return false;
    if (!(c.cfl_coefficient > 0.0f) || !(c.material_cfl_coefficient > 0.0f)) return false;
    if (!(c.viscous_timestep_coefficient > 0.0f) || !(c.min_dt > 0.0f) || !(c.min_dt <= c.dt)) return false;
    if (!(c.projection_min_density > 0.0f) || !(c.projection_min_volume > 0.0f)) return false;
    if (!(c.projection_min_deformation_j > 0.0f) || !(c.particle_mass > 0.0f)) return false;
    if (!(c.contact_friction >= 0.0f) || c.max_substeps_per_step == 0) return false;
    if (!(c.default_initial_volume > 0.0f) || !(c.j_max > 1.0f)) return false;
    if (!(c.apic_blend >= 0.0f && c.apic_blend <= 1.0f)) return false;
    if (!(c.boundary_thickness > 0 && c.boundary_thickness < c.grid_res - 1)) return false;
    return true;
}

template <class T> void export_f64(const Particles<T>& P, size_t i, double* o) {
    o[0] = (double)P.x[i].x; o[1] = (double)P.x[i].y;
    o[2] = (double)P.v[i].x; o[3] = (double)P.v[i].y;
    o[4] = (double)P.c[i].m00; o[5] = (double)P.c[i].m10; o[6] = (double)P.c[i].m01; o[7] = (double)P.c[i].m11;
    o[8] = (double)P.f[i].m00; o[9] = (double)P.f[i].m10; o[10] = (double)P.f[i].m01; o[11] = (double)P.f[i].m11;
    o[12] = (double)P.mass[i];
    o[13] = (double)P.initial_volume[i];
    o[14] = (double)P.volume[i];
    o[15] = (double)P.density[i];
    o[16] = (double)P.plastic_volume_ratio[i];
    o[17] = (double)P.hardening_scale[i];
    o[18] = (double)P.friction_hardening[i];
    o[19] = (double)P.log_volume_strain[i];
}

}  // namespace

extern "C" {

// ------------------------------------------------------------- simulation
void* orc_create(int precision, const emh_config* cfg) {
    if (!cfg || !valid_config(*cfg)) return nullptr;
    Handle* h = new Handle();
    h->precision = precision == 64 ? 64 : 32;
    with_sim(h, [&](auto& s) { s.init_empty(*cfg); return 0; });
    return h;
}

void* orc_create_from_spawn(int precision, const emh_config* cfg, const emh_spawn_region* sp) {
    if (!cfg || !sp || !valid_config(*cfg)) return nullptr;
    Handle* h = new Handle();
    h->precision = precision == 64 ? 64 : 32;
    int rc = with_sim(h, [&](auto& s) { return s.init_from_spawn(*cfg, *sp); });
    if (rc != EMH_OK) {
        delete h;
        return nullptr;
    }
    return h;
}

void orc_destroy(void* h) { delete (Handle*)h; }

int orc_set_material(void* hv, uint32_t id, const emh_material_params* p, const emh_material_ext* e) {
    Handle* h = (Handle*)hv;
    if (!h || !p || id >= EMH_MAX_MATERIALS) return EMH_ERR_INVALID_ARG;
    return with_sim(h, [&](auto& s) {
        using T = typename std::decay_t<decltype(s.P.mass)>::value_type;
        if (id > s.materials.size()) return EMH_ERR_INVALID_ARG;  // registry.rs:39-46
        Material<T> m = Material<T>::from_params(*p, e);
        if (id < s.materials.size()) s.materials[id] = m;
        else s.materials.push_back(m);
        if (id == 0) s.reinit_all();  // with_default_material  lifecycle.rs:107-111
        return EMH_OK;
    });
}

int orc_reinit_particle_state(void* hv) {
    Handle* h = (Handle*)hv;
    if (!h) return EMH_ERR_INVALID_ARG;
    return with_sim(h, [&](auto& s) { s.reinit_all(); return EMH_OK; });
}

static int add_boundary(void* hv, const Boundary& b) {
    Handle* h = (Handle*)hv;
    if (!h) return EMH_ERR_INVALID_ARG;
    return with_sim(h, [&](auto& s) { s.boundaries.push_back(b); return EMH_OK; });
}

int orc_add_slip_boundary(void* hv, uint32_t t) {  // SlipBoundary::new  slip.rs:9-13
    Boundary b;
    b.kind = BND_SLIP;
    b.thickness = t;
    return add_boundary(hv, b);
}

int orc_add_friction_boundary(void* hv, uint32_t t, float mu) {  // FrictionBoundary::new  friction.rs:28-37
    if (!(mu >= 0.0f && mu <= 1.0f)) return EMH_ERR_INVALID_ARG;
    Boundary b;
    b.kind = BND_FRICTION;
    b.thickness = t;
    b.mu = mu;
    return add_boundary(hv, b);
}

int orc_add_heightmap_boundary(void* hv, const float* heights, uint32_t n, float friction, uint32_t wall_t) {
    Boundary b;  // HeightmapBoundary::new  heightmap.rs:44-50
    b.kind = BND_HEIGHTMAP;
    b.thickness = wall_t;
    b.mu = friction;
    b.heights.assign(heights, heights + n);
    return add_boundary(hv, b);
}

int orc_add_predictive_boundary(void* hv, uint32_t t, float wall_min) {  // predictive.rs:17-24
    Boundary b;
    b.kind = BND_PREDICTIVE;
    b.thickness = t;
    b.wall_min = wall_min;
    return add_boundary(hv, b);
}

int orc_apply_impulse(void* hv, float cx, float cy, float radius, float fx, float fy, float strength,
                      uint32_t radial) {
    Handle* h = (Handle*)hv;
    if (!h) return EMH_ERR_INVALID_ARG;
    return with_sim(h, [&](auto& s) {
        s.apply_impulse_f(cx, cy, radius, fx, fy, strength, radial != 0u);
        return EMH_OK;
    });
}

int orc_set_gravity(void* hv, float gx, float gy) {  // lifecycle.rs:403-405
    Handle* h = (Handle*)hv;
    if (!h) return EMH_ERR_INVALID_ARG;
    return with_sim(h, [&](auto& s) {
        s.cfg.gravity[0] = gx;
        s.cfg.gravity[1] = gy;
        return EMH_OK;
    });
}

int orc_clear_boundaries(void* hv) {
    Handle* h = (Handle*)hv;
    if (!h) return EMH_ERR_INVALID_ARG;
    return with_sim(h, [&](auto& s) { s.boundaries.clear(); return EMH_OK; });
}

// One boundary's grid-velocity law / particle clamp on one cell or position (f32).
void orc_boundary_velocity(uint32_t kind, uint32_t t, float mu, float wall_min, const float* heights,
                           uint32_t n_heights, uint32_t cell_index, uint32_t grid_res, float v[2]) {
    Boundary b;
    b.kind = kind;
    b.thickness = t;
    b.mu = mu;
    b.wall_min = wall_min;
    if (heights) b.heights.assign(heights, heights + n_heights);
    V2<float> vv{v[0], v[1]};
    boundary_velocity(b, cell_index, grid_res, vv);
    v[0] = vv.x;
    v[1] = vv.y;
}

void orc_boundary_clamp(uint32_t kind, uint32_t t, float wall_min, const float* heights, uint32_t n_heights,
                        uint32_t grid_res, float pos[2]) {
    Boundary b;
    b.kind = kind;
    b.thickness = t;
    b.wall_min = wall_min;
    if (heights) b.heights.assign(heights, heights + n_heights);
    V2<float> p = boundary_clamp(b, V2<float>{pos[0], pos[1]}, grid_res);
    pos[0] = p.x;
    pos[1] = p.y;
}

int orc_add_body(void* hv, const emh_spawn_region* sp, uint32_t* tag) {
    Handle* h = (Handle*)hv;
    if (!h || !sp) return EMH_ERR_INVALID_ARG;
    return with_sim(h, [&](auto& s) { return s.add_body(*sp, tag); });
}

int orc_set_particles(void* hv, const emh_particle* p, size_t n) {
    Handle* h = (Handle*)hv;
    if (!h || (!p && n)) return EMH_ERR_INVALID_ARG;
    return with_sim(h, [&](auto& s) {
        using T = typename std::decay_t<decltype(s.P.mass)>::value_type;
        s.P = Particles<T>();
        for (size_t i = 0; i < n; ++i) s.P.push(p[i]);
        s.reset_partition();
        return EMH_OK;
    });
}

int orc_sleep_tag(void* hv, uint32_t tag, uint32_t sleep) {
    Handle* h = (Handle*)hv;
    if (!h) return EMH_ERR_INVALID_ARG;
    return with_sim(h, [&](auto& s) {
        if (sleep)
            s.sleep_tag(tag);
        else
            s.wake_tag(tag);
        return EMH_OK;
    });
}

size_t orc_active_count(void* hv) {
    Handle* h = (Handle*)hv;
    if (!h) return 0;
    return with_sim(h, [&](auto& s) { return s.active_count; });
}

size_t orc_num_particles(void* hv) {
    Handle* h = (Handle*)hv;
    if (!h) return 0;
    return with_sim(h, [&](auto& s) { return s.P.size(); });
}

int orc_get_particles(void* hv, emh_particle* out, size_t n) {
    Handle* h = (Handle*)hv;
    if (!h || !out) return EMH_ERR_INVALID_ARG;
    return with_sim(h, [&](auto& s) {
        if (n != s.P.size()) return EMH_ERR_INVALID_ARG;
        for (size_t i = 0; i < n; ++i) out[i] = s.P.get(i);
        return EMH_OK;
    });
}

// 20 doubles per particle: x(2) v(2) C(4) F(4) mass V0 volume density Jp h q lvs
int orc_get_particles_f64(void* hv, double* out, size_t n) {
    Handle* h = (Handle*)hv;
    if (!h || !out) return EMH_ERR_INVALID_ARG;
    return with_sim(h, [&](auto& s) {
        if (n != s.P.size()) return EMH_ERR_INVALID_ARG;
        for (size_t i = 0; i < n; ++i) export_f64(s.P, i, out + 20 * i);
        return EMH_OK;
    });
}

int orc_step(void* hv, uint32_t n) {
    Handle* h = (Handle*)hv;
    if (!h) return EMH_ERR_INVALID_ARG;
    return with_sim(h, [&](auto& s) {
        for (uint32_t k = 0; k < n; ++k) s.step();
        return EMH_OK;
    });
}

// Runs exactly `n` substeps of size `dt` (do_substep, step.rs:62-317) with no
// CFL selection: used to time a bounded sample of the workload.
int orc_substeps(void* hv, uint32_t n, float dt) {
    Handle* h = (Handle*)hv;
    if (!h) return EMH_ERR_INVALID_ARG;
    return with_sim(h, [&](auto& s) {
        using T = typename std::decay_t<decltype(s.P.mass)>::value_type;
        for (uint32_t k = 0; k < n; ++k) s.do_substep(T(dt));
        return EMH_OK;
    });
}

int orc_stats(void* hv, emh_step_stats* out) {
    Handle* h = (Handle*)hv;
    if (!h || !out) return EMH_ERR_INVALID_ARG;
    return with_sim(h, [&](auto& s) {
        out->last_substeps = s.last_substeps;
        out->last_step_dt = (float)s.last_step_dt;
        out->last_sim_time_dropped = (float)s.last_sim_time_dropped;
        out->last_vel_clamp_count = s.last_vel_clamp_count;
        out->last_j_projection_count = s.last_j_projection_count;
        out->total_substeps = s.total_substeps;
        out->frame_index = s.frame_index;
        return EMH_OK;
    });
}

// Every substep dt taken so far, as f64.
size_t orc_substep_dts(void* hv, double* out, size_t cap) {
    Handle* h = (Handle*)hv;
    if (!h) return 0;
    return with_sim(h, [&](auto& s) {
        size_t n = s.substep_dts.size();
        if (out)
            for (size_t i = 0; i < n && i < cap; ++i) out[i] = (double)s.substep_dts[i];
        return n;
    });
}

// Dense export of the sparse grid as it stands (after the last substep: mass
// and the post-update velocity held in `momentum`, grid/mod.rs:101-110).
int orc_get_grid(void* hv, double* mass, double* momentum, size_t n_cells) {
    Handle* h = (Handle*)hv;
    if (!h) return EMH_ERR_INVALID_ARG;
    return with_sim(h, [&](auto& s) {
        size_t R = s.cfg.grid_res;
        if (n_cells != R * R) return EMH_ERR_INVALID_ARG;
        if (mass) std::fill(mass, mass + n_cells, 0.0);
        if (momentum) std::fill(momentum, momentum + 2 * n_cells, 0.0);
        for (uint32_t slot : s.grid.dirty_slots) {
            uint32_t idx = s.grid.keys[slot];
            if (mass) mass[idx] = (double)s.grid.cells[slot].mass;
            if (momentum) {
                momentum[2 * idx] = (double)s.grid.cells[slot].momentum.x;
                momentum[2 * idx + 1] = (double)s.grid.cells[slot].momentum.y;
            }
        }
        return EMH_OK;
    });
}

// P2G alone on the current state (clear + scatter_particles_to_grid,
// transfer/p2g.rs:24-72) -- the grid before normalisation, for conservation.
int orc_p2g_only(void* hv, float dt, double* mass, double* momentum, size_t n_cells) {
    Handle* h = (Handle*)hv;
    if (!h) return EMH_ERR_INVALID_ARG;
    return with_sim(h, [&](auto& s) {
        using T = typename std::decay_t<decltype(s.P.mass)>::value_type;
        s.grid.clear();
        s.p2g(T(dt));
        return 0;
    }) == 0 ? orc_get_grid(hv, mass, momentum, n_cells) : EMH_ERR_STATE;
}

// choose_substep_dt (step.rs:434-479) over the current state, for max_dt.
double orc_choose_substep_dt(void* hv, float max_dt) {
    Handle* h = (Handle*)hv;
    if (!h) return -1.0;
    return with_sim(h, [&](auto& s) {
        using T = typename std::decay_t<decltype(s.P.mass)>::value_type;
        return (double)s.choose_substep_dt(T(max_dt));
    });
}

// project_particle_state_to_admissible (step.rs:347-429) of particle i, in
// place; returns 1 if a field was corrected, 0 if not, -1 on a bad index.
int orc_project_particle(void* hv, size_t i) {
    Handle* h = (Handle*)hv;
    if (!h) return -1;
    return with_sim(h, [&](auto& s) {
        if (i >= s.P.mass.size()) return -1;
        return s.project(i) ? 1 : 0;
    });
}

// Planted mistakes (Mutation in mpm_oracle.hpp) for the oracle's own tests.
void orc_set_mutation(int m) { g_mutation = m; }

// Threads of G2P phase 1 (the reference's rayon loop); 1 = serial.
void orc_set_threads(int n) { g_threads = n < 1 ? 1 : n; }

// Seconds of wall time for `n` full steps (the CPU baseline timer).
double orc_time_steps(void* hv, uint32_t n) {
    auto t0 = std::chrono::steady_clock::now();
    orc_step(hv, n);
    auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count();
}

double orc_time_substeps(void* hv, uint32_t n, float dt) {
    auto t0 = std::chrono::steady_clock::now();
    orc_substeps(hv, n, dt);
    auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count();
}

// ------------------------------------------------------- unit-level access
void orc_quadratic_weights(float x, float y, int32_t base[2], float wx[3], float wy[3]) {
    Weights<float> w = quadratic_weights(V2<float>{x, y});
    base[0] = w.bx;
    base[1] = w.by;
    for (int k = 0; k < 3; ++k) {
        wx[k] = w.wx[k];
        wy[k] = w.wy[k];
    }
}

void orc_svd2(const float f[4], float u[4], float sigma[2], float vt[4]) {
    M2<float> F{f[0], f[1], f[2], f[3]}, U, VT;
    V2<float> S;
    svd2(F, U, S, VT);
    u[0] = U.m00; u[1] = U.m10; u[2] = U.m01; u[3] = U.m11;
    vt[0] = VT.m00; vt[1] = VT.m10; vt[2] = VT.m01; vt[3] = VT.m11;
    sigma[0] = S.x;
    sigma[1] = S.y;
}

void orc_svd2_f64(const double f[4], double u[4], double sigma[2], double vt[4]) {
    M2<double> F{f[0], f[1], f[2], f[3]}, U, VT;
    V2<double> S;
    svd2(F, U, S, VT);
    u[0] = U.m00; u[1] = U.m10; u[2] = U.m01; u[3] = U.m11;
    vt[0] = VT.m00; vt[1] = VT.m10; vt[2] = VT.m01; vt[3] = VT.m11;
    sigma[0] = S.x;
    sigma[1] = S.y;
}

void orc_polar(const float f[4], float r[4]) {
    M2<float> R = polar_decomposition_2d(M2<float>{f[0], f[1], f[2], f[3]});
    r[0] = R.m00; r[1] = R.m10; r[2] = R.m01; r[3] = R.m11;
}

// Kirchhoff stress of one particle under one material (f32).
void orc_kirchhoff(const emh_material_params* p, const emh_material_ext* e, const emh_particle* part,
                   float tau[4]) {
    Material<float> m = Material<float>::from_params(*p, e);
    Particles<float> P;
    P.push(*part);
    M2<float> t = m.kirchhoff_stress(P, 0);
    tau[0] = t.m00; tau[1] = t.m10; tau[2] = t.m01; tau[3] = t.m11;
}

// update_particle on one particle (f32 or f64 state rounded back to f32).
void orc_update_particle(const emh_material_params* p, const emh_material_ext* e, emh_particle* part, float dt,
                         int precision) {
    if (precision == 64) {
        Material<double> m = Material<double>::from_params(*p, e);
        Particles<double> P;
        P.push(*part);
        m.update_particle(P, 0, (double)dt);
        *part = P.get(0);
    } else {
        Material<float> m = Material<float>::from_params(*p, e);
        Particles<float> P;
        P.push(*part);
        m.update_particle(P, 0, dt);
        *part = P.get(0);
    }
}

float orc_timestep_bound(const emh_material_params* p, const emh_material_ext* e, float density, float h,
                         float cell, float mcfl, float vcfl) {
    Material<float> m = Material<float>::from_params(*p, e);
    return m.timestep_bound(density, h, cell, mcfl, vcfl);
}

float orc_dp_alpha(const emh_material_params* p, float q) {
    return Material<float>::from_params(*p, nullptr).dp_alpha(q);
}

// NaccMaterial::project on a deformation gradient (f32).
void orc_nacc_project(const emh_material_params* p, const emh_material_ext* e, const float f[4], float alpha,
                      float f_out[4], float* alpha_out) {
    Material<float> m = Material<float>::from_params(*p, e);
    M2<float> fo;
    float ao;
    m.nacc_project(M2<float>{f[0], f[1], f[2], f[3]}, alpha, fo, ao);
    f_out[0] = fo.m00; f_out[1] = fo.m10; f_out[2] = fo.m01; f_out[3] = fo.m11;
    *alpha_out = ao;
}

// NaccMaterial::project (nacc.rs:129-223) over n trial states (f [n][4]
// column-major, alpha [n]) in precision 32 or 64; results widened to double.
void orc_nacc_project_batch(const emh_material_params* p, const emh_material_ext* e, int precision, size_t n,
                            const float* f, const float* alpha, double* f_out, double* alpha_out) {
    auto run = [&](auto zero) {
        using T = decltype(zero);
        Material<T> m = Material<T>::from_params(*p, e);
        for (size_t i = 0; i < n; ++i) {
            const float* q = f + 4 * i;
            M2<T> fo;
            T ao;
            m.nacc_project(M2<T>{(T)q[0], (T)q[1], (T)q[2], (T)q[3]}, (T)alpha[i], fo, ao);
            f_out[4 * i + 0] = (double)fo.m00;
            f_out[4 * i + 1] = (double)fo.m10;
            f_out[4 * i + 2] = (double)fo.m01;
            f_out[4 * i + 3] = (double)fo.m11;
            alpha_out[i] = (double)ao;
        }
    };
    if (precision == 64)
        run(0.0);
    else
        run(0.0f);
}

void orc_slip_velocity(uint32_t thickness, uint32_t cell_index, uint32_t grid_res, float v[2]) {
    uint32_t hi = grid_res - (thickness + 1);
    uint32_t x = cell_index / grid_res, y = cell_index % grid_res;
    if (x < thickness) v[0] = rmax(v[0], 0.0f);
    if (x > hi) v[0] = rmin(v[0], 0.0f);
    if (y < thickness) v[1] = rmax(v[1], 0.0f);
    if (y > hi) v[1] = rmin(v[1], 0.0f);
}

void orc_lame_from_young(float e, float nu, float* lambda, float* mu) {  // utils.rs:183-193
    *lambda = e * nu / ((1.0f + nu) * (1.0f - 2.0f * nu));
    *mu = e / (2.0f * (1.0f + nu));
}

float orc_to_radians(float deg) { return deg * (3.14159265358979323846f / 180.0f); }

// Material constructors restated from the model files (f32), filling the
// 96 B params (each model's params()) and the extension.
static void ext_defaults(emh_material_ext* e) {
    std::memset(e, 0, sizeof(*e));
    e->min_density = 1.0e-6f;
    e->min_volume = 1.0e-6f;
    e->dp_volume_correction = 1.0f;
    e->nacc_hardening_enabled = 2u;
    e->bingham_critical_shear_rate = 1.0e-4f;
}

void orc_material_neohookean(float lambda, float mu, emh_material_params* p, emh_material_ext* e) {
    std::memset(p, 0, sizeof(*p));
    ext_defaults(e);
    p->model = EMH_MODEL_NEOHOOKEAN;  // elastic.rs:275-288
    p->lambda = lambda;
    p->mu = mu;
}

void orc_material_corotated(float lambda, float mu, emh_material_params* p, emh_material_ext* e) {
    std::memset(p, 0, sizeof(*p));
    ext_defaults(e);
    p->model = EMH_MODEL_COROTATED;  // corotated.rs:97-106
    p->lambda = lambda;
    p->mu = mu;
}

void orc_material_von_mises(float lambda, float mu, float yield_stress, float hardening_modulus,
                            emh_material_params* p, emh_material_ext* e) {
    std::memset(p, 0, sizeof(*p));
    ext_defaults(e);
    p->model = EMH_MODEL_VON_MISES;  // von_mises.rs:47-70, 150-159
    p->lambda = lambda;
    p->mu = mu;
    p->hardening_exponent = yield_stress;
    p->hardening_modulus = hardening_modulus;
}

void orc_material_rankine(float lambda, float mu, float tensile_strength, float softening_rate,
                          emh_material_params* p, emh_material_ext* e) {
    std::memset(p, 0, sizeof(*p));
    ext_defaults(e);
    p->model = EMH_MODEL_RANKINE;  // rankine.rs:53-60, 206-216
    p->lambda = lambda;
    p->mu = mu;
    p->hardening_exponent = tensile_strength;
    p->hardening_modulus = softening_rate;
}

// f32::to_radians then f32::tan (sand_mui.rs:62-70)
static float tan_deg(float deg) { return tanf(deg * (3.14159265358979323846f / 180.0f)); }

void orc_material_mu_i(float lambda, float mu, float mu_static, float mu_dynamic, float inertial_q,
                       emh_material_params* p, emh_material_ext* e) {
    std::memset(p, 0, sizeof(*p));
    ext_defaults(e);
    p->model = EMH_MODEL_MU_I;  // sand_mui.rs:62-70, 219-229
    p->lambda = lambda;
    p->mu = mu;
    p->dp_h0 = mu_static;
    p->dp_h1 = mu_dynamic;
    p->dp_h2 = inertial_q;
}

void orc_material_mu_i_default(float lambda, float mu, emh_material_params* p, emh_material_ext* e) {
    orc_material_mu_i(lambda, mu, tan_deg(20.9f), tan_deg(32.8f), 5.58f, p, e);
}

void orc_material_granular_fluid(float lambda, float mu, float rest_density, float eos_stiffness,
                                 float hardening_exponent, float compression_limit, emh_material_params* p,
                                 emh_material_ext* e) {
    std::memset(p, 0, sizeof(*p));
    ext_defaults(e);
    p->model = EMH_MODEL_GRANULAR_FLUID;  // granular_fluid.rs:61-83, 232-247
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
}

void orc_material_bingham(float rest_density, float dynamic_viscosity, float eos_stiffness, float eos_power,
                          float yield_stress, emh_material_params* p, emh_material_ext* e) {
    std::memset(p, 0, sizeof(*p));
    ext_defaults(e);
    p->model = EMH_MODEL_FLUID;  // bingham.rs:53-71, 205-218
    p->rest_density = rest_density;
    p->eos_stiffness = eos_stiffness;
    p->eos_power = eos_power;
    p->dynamic_viscosity = dynamic_viscosity;
    p->compression_limit = yield_stress;
    p->volume_ratio_max = 2.0f;
    p->pressure_floor = 0.0f;
    e->fluid_bingham = 1u;
}

void orc_material_viscoelastic(float lambda, float mu, float viscosity, emh_material_params* p,
                               emh_material_ext* e) {
    std::memset(p, 0, sizeof(*p));
    ext_defaults(e);
    p->model = EMH_MODEL_VISCOELASTIC;  // viscoelastic.rs:47-56, 149-159
    p->lambda = lambda;
    p->mu = mu;
    p->dynamic_viscosity = viscosity;
    p->volume_ratio_min = 0.01f;
}

void orc_material_fluid(float rest_density, float visc, float eos_k, float eos_p, emh_material_params* p,
                        emh_material_ext* e) {
    std::memset(p, 0, sizeof(*p));
    ext_defaults(e);
    p->model = EMH_MODEL_FLUID;  // fluid.rs:48-67, 209-226
    p->rest_density = rest_density;
    p->eos_stiffness = eos_k;
    p->eos_power = eos_p;
    p->dynamic_viscosity = visc;
    p->volume_ratio_max = 2.0f;
    p->pressure_floor = -0.1f;
}

void orc_material_drucker_prager(float lambda, float mu, emh_material_params* p, emh_material_ext* e) {
    std::memset(p, 0, sizeof(*p));
    ext_defaults(e);
    p->model = EMH_MODEL_DRUCKER_PRAGER;  // sand.rs:81-94, 321-343
    p->lambda = lambda;
    p->mu = mu;
    p->dp_h0 = orc_to_radians(35.0f);
    p->dp_h1 = orc_to_radians(9.0f);
    p->dp_h2 = 0.2f;
    p->dp_h3 = orc_to_radians(10.0f);
    p->compression_limit = 0.0f;
    p->stretch_limit = 0.0f;
    p->volume_ratio_min = 0.6f;
}

void orc_material_snow(float lambda, float mu, float xi, float tc, float ts, float jmin, float jmax,
                       emh_material_params* p, emh_material_ext* e) {
    std::memset(p, 0, sizeof(*p));
    ext_defaults(e);
    p->model = EMH_MODEL_SNOW;  // snow.rs:40-59, 178-191
    p->lambda = lambda;
    p->mu = mu;
    p->hardening_exponent = xi;
    p->compression_limit = tc;
    p->stretch_limit = ts;
    p->volume_ratio_min = jmin;
    p->volume_ratio_max = jmax;
}

void orc_material_nacc(float mu, float kappa, float friction, float cohesion, float xi, emh_material_params* p,
                       emh_material_ext* e) {
    std::memset(p, 0, sizeof(*p));
    ext_defaults(e);
    p->model = EMH_MODEL_NACC;  // nacc.rs:63-73, 300-314 (model id 10 instead of 2)
    p->lambda = kappa - mu;
    p->mu = mu;
    p->hardening_exponent = xi;
    p->compression_limit = cohesion;
    p->stretch_limit = friction;
    e->nacc_kappa = kappa;
    e->nacc_hardening_enabled = xi > 0.0f ? 1u : 0u;
}

// ---------------------------------------------------------------- sort
// Cell key (no reference counterpart: the reference's CPU path has no sort):
// the index of the clamped cell along the Hilbert curve over the smallest
// 2^k x 2^k square (k >= 1) holding the R x R grid, built from the top
// quadrant down: quadrant digit (3 rx) ^ ry, then the lower bits reflected
// (ry = 0, rx = 1) and transposed (ry = 0).
uint32_t orc_cell_key(int32_t cx, int32_t cy, uint32_t R) {
    int32_t hi = (int32_t)R - 1;
    cx = std::min(std::max(cx, 0), hi);
    cy = std::min(std::max(cy, 0), hi);
    uint32_t k = 1;
    while (k < 16u && (1u << k) < R) ++k;
    const uint32_t side = 1u << k;
    uint32_t x = (uint32_t)cx, y = (uint32_t)cy, key = 0;
    for (uint32_t q = side / 2u; q >= 1u; q /= 2u) {
        const uint32_t rx = (x & q) != 0u, ry = (y & q) != 0u;
        key += q * q * ((3u * rx) ^ ry);
        if (ry == 0u) {
            if (rx == 1u) {
                x = side - 1u - x;
                y = side - 1u - y;
            }
            std::swap(x, y);
        }
    }
    return key;
}

// std::stable_sort of input indices by the cell key of floor(x).
void orc_stable_sort_cell_keys(const float* xy, size_t n, uint32_t R, uint32_t* keys, uint32_t* perm) {
    for (size_t i = 0; i < n; ++i)
        keys[i] = orc_cell_key(sat_i32(std::floor(xy[2 * i])), sat_i32(std::floor(xy[2 * i + 1])), R);
    std::vector<uint32_t> idx(n);
    std::iota(idx.begin(), idx.end(), 0u);
    std::stable_sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) { return keys[a] < keys[b]; });
    for (size_t j = 0; j < n; ++j) perm[j] = idx[j];
}

// LcgRng sequence (solver/mod.rs:154-175) for the known-answer test.
void orc_lcg(uint32_t seed, uint32_t n, uint32_t* u32_out, float* f32_out) {
    LcgRng r(seed), r2(seed);
    for (uint32_t i = 0; i < n; ++i) {
        u32_out[i] = r.next_u32();
        f32_out[i] = r2.next_f32();
    }
}

}  // extern "C"
