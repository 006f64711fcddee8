// mpm_oracle.hpp -- CPU restatement of the reference crate's Simulation::step.
//
// TEST INFRASTRUCTURE ONLY.  This is the correctness oracle that the GPU
// library is checked against; it is never linked into, loaded by or called
// from the product library.  Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py may load it.
//
// Each function names the reference file:line it restates (paths relative to
// the reference crate root).  The arithmetic follows the Rust source operation
// by operation so that the float instantiation reproduces the reference's f32
// results: glam Vec2 is scalar f32, glam Mat2 products are plain a*b + c*d
// (no fused multiply-add), and f32::ln/exp/sin/sinh/powf lower to glibc's
// logf/expf/sinf/sinhf/powf on linux-gnu.  Build with -ffp-contract=off.
// The double instantiation runs the same algorithm in f64 and is the "FP64
// oracle" every reported error is stated against.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "../include/emerge_hip.h"

namespace orc {

// ---------------------------------------------------------------- scalar math
inline float o_ln(float x) { return logf(x); }
inline double o_ln(double x) { return std::log(x); }
inline float o_exp(float x) { return expf(x); }
inline double o_exp(double x) { return std::exp(x); }
inline float o_sin(float x) { return sinf(x); }
inline double o_sin(double x) { return std::sin(x); }
inline float o_sinh(float x) { return sinhf(x); }
inline double o_sinh(double x) { return std::sinh(x); }
inline float o_pow(float x, float y) { return powf(x, y); }
inline double o_pow(double x, double y) { return std::pow(x, y); }
inline float o_sqrt(float x) { return sqrtf(x); }
inline double o_sqrt(double x) { return std::sqrt(x); }

// Rust f32::max / f32::min: IEEE maxNum/minNum (a NaN operand is ignored).
template <class T> inline T rmax(T a, T b) { return std::fmax(a, b); }
template <class T> inline T rmin(T a, T b) { return std::fmin(a, b); }
// Rust f32::clamp: NaN passes through.
template <class T> inline T rclamp(T x, T lo, T hi) {
    if (x < lo) x = lo;
    if (x > hi) x = hi;
    return x;
}
template <class T> inline bool finite(T x) { return std::isfinite(x); }

constexpr float F32_EPSILON = 1.1920929e-07f;  // f32::EPSILON

// Planted mistakes for the oracle's own tests: a test switches one on to show
// that its closed-form check fails on a plausible error, then switches it off.
// 0 (the default) is the faithful restatement.
enum Mutation : int {
    kMutNone = 0,
    kMutCflNoAffine = 1,       // choose_substep_dt without the affine speed term
    kMutCflNoMinClamp = 2,     // cfl_bound without the lower clamp to min(min_dt, max_dt)
    kMutProjJScaleLinear = 3,  // project: F scaled by j_max/J instead of sqrt(j_max/J)
    kMutProjNoDensityFloor = 4,  // project: a valid density is not floored
    kMutNaccOtherRoot = 5,     // NACC hardening: the intersection on the far side of p_c
    kMutNaccCaseANoAlpha = 6,  // NACC Case A: alpha not updated
    kMutNaccHardenSign = 7,    // NACC hardening: alpha -= ln(J_tr / J_x)
};
inline int g_mutation = kMutNone;
inline bool mutated(int m) { return g_mutation == m; }

// Threads of G2P phase 1, the one loop the reference runs in parallel (rayon's
// par_iter over particles, transfer/g2p.rs:189-196).  Each particle reads the
// grid and writes only its own state, so the result does not depend on it.
inline int g_threads = 1;

// ---------------------------------------------------------------- glam mirror
template <class T> struct V2 {
    T x, y;
};
template <class T> inline V2<T> v2(T x, T y) { return V2<T>{x, y}; }
template <class T> inline V2<T> operator+(V2<T> a, V2<T> b) { return {a.x + b.x, a.y + b.y}; }
template <class T> inline V2<T> operator-(V2<T> a, V2<T> b) { return {a.x - b.x, a.y - b.y}; }
template <class T> inline V2<T> operator*(V2<T> a, T s) { return {a.x * s, a.y * s}; }
template <class T> inline V2<T> operator*(T s, V2<T> a) { return {s * a.x, s * a.y}; }
template <class T> inline V2<T> operator/(V2<T> a, T s) { return {a.x / s, a.y / s}; }
template <class T> inline T dot(V2<T> a, V2<T> b) { return a.x * b.x + a.y * b.y; }
template <class T> inline T length(V2<T> a) { return o_sqrt(dot(a, a)); }
template <class T> inline bool finite(V2<T> a) { return finite(a.x) && finite(a.y); }
// glam Vec2::normalize_or_zero: multiply by the reciprocal length.
template <class T> inline V2<T> normalize_or_zero(V2<T> a) {
    T rcp = T(1) / length(a);
    if (finite(rcp) && rcp > T(0)) return a * rcp;
    return {T(0), T(0)};
}

// glam Mat2 stores columns: x_axis = (m00, m10), y_axis = (m01, m11).
template <class T> struct M2 {
    T m00, m10, m01, m11;
};
template <class T> inline M2<T> m2_zero() { return {T(0), T(0), T(0), T(0)}; }
template <class T> inline M2<T> m2_identity() { return {T(1), T(0), T(0), T(1)}; }
template <class T> inline M2<T> m2_diag(T a, T b) { return {a, T(0), T(0), b}; }
template <class T> inline M2<T> m2_cols(V2<T> c0, V2<T> c1) { return {c0.x, c0.y, c1.x, c1.y}; }
template <class T> inline M2<T> operator+(M2<T> a, M2<T> b) {
    return {a.m00 + b.m00, a.m10 + b.m10, a.m01 + b.m01, a.m11 + b.m11};
}
template <class T> inline M2<T> operator-(M2<T> a, M2<T> b) {
    return {a.m00 - b.m00, a.m10 - b.m10, a.m01 - b.m01, a.m11 - b.m11};
}
template <class T> inline M2<T> operator*(M2<T> a, T s) { return {a.m00 * s, a.m10 * s, a.m01 * s, a.m11 * s}; }
template <class T> inline M2<T> operator*(T s, M2<T> a) { return {s * a.m00, s * a.m10, s * a.m01, s * a.m11}; }
template <class T> inline V2<T> operator*(M2<T> a, V2<T> v) {
    return {a.m00 * v.x + a.m01 * v.y, a.m10 * v.x + a.m11 * v.y};
}
template <class T> inline M2<T> operator*(M2<T> a, M2<T> b) {
    return {a.m00 * b.m00 + a.m01 * b.m10, a.m10 * b.m00 + a.m11 * b.m10,
            a.m00 * b.m01 + a.m01 * b.m11, a.m10 * b.m01 + a.m11 * b.m11};
}
template <class T> inline M2<T> transpose(M2<T> a) { return {a.m00, a.m01, a.m10, a.m11}; }
template <class T> inline T det(M2<T> a) { return a.m00 * a.m11 - a.m10 * a.m01; }
template <class T> inline bool finite_cols(M2<T> a) {
    return finite(a.m00) && finite(a.m10) && finite(a.m01) && finite(a.m11);
}

// ------------------------------------------------------------- kernel weights
// grid/kernel.rs:10-25
template <class T> struct Weights {
    int32_t bx, by;
    T wx[3], wy[3];
};
template <class T> inline void axis_weights(T d, T w[3]) {
    T a = T(0.5) - d;
    T b = T(0.5) + d;
    w[0] = T(0.5) * (a * a);
    w[1] = T(0.75) - d * d;
    w[2] = T(0.5) * (b * b);
}
// glam `as_ivec2` is Rust's saturating `as i32`.
template <class T> inline int32_t sat_i32(T f) {
    if (std::isnan(f)) return 0;
    if (f >= T(2147483647.0)) return 2147483647;
    if (f <= T(-2147483648.0)) return (-2147483647 - 1);
    return (int32_t)f;
}
template <class T> inline Weights<T> quadratic_weights(V2<T> p) {
    Weights<T> w;
    w.bx = sat_i32(std::floor(p.x));
    w.by = sat_i32(std::floor(p.y));
    T dx = (p.x - T(w.bx)) - T(0.5);
    T dy = (p.y - T(w.by)) - T(0.5);
    axis_weights(dx, w.wx);
    axis_weights(dy, w.wy);
    return w;
}

// --------------------------------------------------------------------- svd2
// materials/svd.rs:14-85
template <class T> inline V2<T> orthogonal_complement(V2<T> v) {
    V2<T> perp{-v.y, v.x};
    T len = length(perp);
    if (len > T(1e-10f)) return perp / len;
    return {T(1), T(0)};
}
template <class T> inline void svd2(M2<T> f, M2<T>& u, V2<T>& sigma, M2<T>& vt) {
    T f00 = f.m00, f10 = f.m10, f01 = f.m01, f11 = f.m11;
    T c00 = f00 * f00 + f10 * f10;
    T c01 = f00 * f01 + f10 * f11;
    T c11 = f01 * f01 + f11 * f11;
    T mean = (c00 + c11) * T(0.5);
    T half_diff = (c00 - c11) * T(0.5);
    T disc = o_sqrt(half_diff * half_diff + c01 * c01);
    T lambda1 = mean + disc;
    T lambda2 = mean - disc;
    V2<T> v1;
    if (std::fabs(c01) > T(1e-10f) * rmax(c00 + c11, T(1e-30f))) {
        V2<T> raw{c01, lambda1 - c00};
        v1 = raw / length(raw);
    } else {
        v1 = (c00 >= c11) ? V2<T>{T(1), T(0)} : V2<T>{T(0), T(1)};
    }
    V2<T> v2_{-v1.y, v1.x};
    M2<T> v_sorted = m2_cols(v1, v2_);
    vt = transpose(v_sorted);
    sigma = {o_sqrt(rmax(lambda1, T(0))), o_sqrt(rmax(lambda2, T(0)))};
    V2<T> fv1 = f * v1;
    V2<T> fv2 = f * v2_;
    V2<T> u1 = sigma.x > T(1e-10f) ? fv1 / sigma.x : orthogonal_complement(fv2);
    V2<T> u2 = sigma.y > T(1e-10f) ? fv2 / sigma.y : orthogonal_complement(u1);
    u = m2_cols(u1, u2);
    if (det(u) < T(0)) {
        u.m01 = -u.m01;
        u.m11 = -u.m11;
        sigma.y = -sigma.y;
    }
}

// materials/utils.rs:129-138
template <class T> inline M2<T> polar_decomposition_2d(M2<T> f) {
    T x = f.m00 + f.m11;
    T y = f.m10 - f.m01;
    T norm = o_sqrt(x * x + y * y);
    if (norm > T(F32_EPSILON)) return m2_cols(V2<T>{x, y} / norm, V2<T>{-y, x} / norm);
    return m2_identity<T>();
}

// materials/utils.rs:145-164
template <class T>
inline T elastic_wave_dt(T lambda, T mu, T hardening, T density, T min_density, T cell_width, T material_cfl) {
    T rho = rmax(density, min_density);
    T modulus = rmax((lambda + T(2) * mu) * hardening, T(0));
    if (modulus <= T(F32_EPSILON)) return std::numeric_limits<T>::infinity();
    T c = o_sqrt(modulus / rho);
    if (c <= T(F32_EPSILON)) return std::numeric_limits<T>::infinity();
    return material_cfl * cell_width / c;
}

constexpr float LOG_CLAMP = 1e-10f;  // utils.rs:6
constexpr float MIN_J = 1e-6f;       // utils.rs:11
constexpr float KERNEL_D_INVERSE = 4.0f;  // solver/config.rs:22

// ------------------------------------------------------------------ particles
// matter/particle.rs:206-243 (the fields the substep touches + pass-through)
template <class T> struct Particles {
    std::vector<V2<T>> x, v;
    std::vector<M2<T>> c, f;
    std::vector<T> mass, initial_volume, volume, density;
    std::vector<uint32_t> material_id;
    std::vector<T> plastic_volume_ratio, hardening_scale, friction_hardening, log_volume_strain;
    std::vector<T> temperature, activation;
    std::vector<V2<T>> activation_dir;
    std::vector<uint32_t> user_tag, muscle_group_id, contact_group, sleeping, pinned;
    std::vector<T> scalar_field;
    size_t size() const { return x.size(); }
    void push(const emh_particle& p) {
        x.push_back({T(p.x[0]), T(p.x[1])});
        v.push_back({T(p.v[0]), T(p.v[1])});
        c.push_back({T(p.velocity_gradient[0]), T(p.velocity_gradient[1]), T(p.velocity_gradient[2]),
                     T(p.velocity_gradient[3])});
        f.push_back({T(p.deformation_gradient[0]), T(p.deformation_gradient[1]), T(p.deformation_gradient[2]),
                     T(p.deformation_gradient[3])});
        mass.push_back(T(p.mass));
        initial_volume.push_back(T(p.initial_volume));
        volume.push_back(T(p.volume));
        density.push_back(T(p.density));
        material_id.push_back(p.material_id);
        plastic_volume_ratio.push_back(T(p.plastic_volume_ratio));
        hardening_scale.push_back(T(p.hardening_scale));
        friction_hardening.push_back(T(p.friction_hardening));
        log_volume_strain.push_back(T(p.log_volume_strain));
        temperature.push_back(T(p.temperature));
        activation.push_back(T(p.activation));
        activation_dir.push_back({T(p.activation_dir[0]), T(p.activation_dir[1])});
        user_tag.push_back(p.user_tag);
        muscle_group_id.push_back(p.muscle_group_id);
        contact_group.push_back(p.contact_group);
        sleeping.push_back(p.sleeping);
        pinned.push_back(p.pinned);
        scalar_field.push_back(T(p.scalar_field));
    }
    emh_particle get(size_t i) const {
        emh_particle p;
        std::memset(&p, 0, sizeof(p));
        p.x[0] = (float)x[i].x; p.x[1] = (float)x[i].y;
        p.v[0] = (float)v[i].x; p.v[1] = (float)v[i].y;
        p.velocity_gradient[0] = (float)c[i].m00; p.velocity_gradient[1] = (float)c[i].m10;
        p.velocity_gradient[2] = (float)c[i].m01; p.velocity_gradient[3] = (float)c[i].m11;
        p.deformation_gradient[0] = (float)f[i].m00; p.deformation_gradient[1] = (float)f[i].m10;
        p.deformation_gradient[2] = (float)f[i].m01; p.deformation_gradient[3] = (float)f[i].m11;
        p.mass = (float)mass[i];
        p.initial_volume = (float)initial_volume[i];
        p.volume = (float)volume[i];
        p.density = (float)density[i];
        p.material_id = material_id[i];
        p.plastic_volume_ratio = (float)plastic_volume_ratio[i];
        p.hardening_scale = (float)hardening_scale[i];
        p.friction_hardening = (float)friction_hardening[i];
        p.log_volume_strain = (float)log_volume_strain[i];
        p.temperature = (float)temperature[i];
        p.user_tag = user_tag[i];
        p.activation = (float)activation[i];
        p.activation_dir[0] = (float)activation_dir[i].x; p.activation_dir[1] = (float)activation_dir[i].y;
        p.muscle_group_id = muscle_group_id[i];
        p.contact_group = contact_group[i];
        p.sleeping = sleeping[i];
        p.pinned = pinned[i];
        p.scalar_field = (float)scalar_field[i];
        return p;
    }
    // Particles::swap / rotate_range (every field), the moves of the sleep partition
    template <class F> void each_field(F&& fn) {
        fn(x); fn(v); fn(c); fn(f); fn(mass); fn(initial_volume); fn(volume); fn(density); fn(material_id);
        fn(plastic_volume_ratio); fn(hardening_scale); fn(friction_hardening); fn(log_volume_strain);
        fn(temperature); fn(activation); fn(activation_dir); fn(user_tag); fn(muscle_group_id); fn(contact_group);
        fn(sleeping); fn(pinned); fn(scalar_field);
    }
    void swap(size_t i, size_t j) {
        each_field([&](auto& vec) { std::swap(vec[i], vec[j]); });
    }
    // [a, mid) and [mid, end) trade places: [mid, end) moves to a
    void rotate_range(size_t a, size_t mid, size_t end) {
        each_field([&](auto& vec) { std::rotate(vec.begin() + a, vec.begin() + mid, vec.begin() + end); });
    }
    void set(size_t i, const emh_particle& p) {
        Particles<T> one;
        one.push(p);
        x[i] = one.x[0]; v[i] = one.v[0]; c[i] = one.c[0]; f[i] = one.f[0];
        mass[i] = one.mass[0]; initial_volume[i] = one.initial_volume[0];
        volume[i] = one.volume[0]; density[i] = one.density[0];
        material_id[i] = one.material_id[0];
        plastic_volume_ratio[i] = one.plastic_volume_ratio[0];
        hardening_scale[i] = one.hardening_scale[0];
        friction_hardening[i] = one.friction_hardening[0];
        log_volume_strain[i] = one.log_volume_strain[0];
        temperature[i] = one.temperature[0];
        activation[i] = one.activation[0];
        activation_dir[i] = one.activation_dir[0];
        user_tag[i] = one.user_tag[0];
        muscle_group_id[i] = one.muscle_group_id[0];
        contact_group[i] = one.contact_group[0];
        sleeping[i] = one.sleeping[0];
        pinned[i] = one.pinned[0];
        scalar_field[i] = one.scalar_field[0];
    }
};

// ------------------------------------------------------------------ materials
// One struct covers the six core constitutive models plus Fallback; `model` picks the
// branch.  Fields are the model structs' own (not the 96 B upload), rebuilt
// from MaterialParams + the extension exactly as each `params()` encodes them.
template <class T> struct Material {
    uint32_t model = EMH_MODEL_FALLBACK;
    T lambda = 0, mu = 0;
    // NeoHookean (elastic.rs:20-48)
    T min_density = T(1e-6f), thermal_expansion = 0, active_stress_coeff = 0, damage_softening_rate = 0,
      viscosity = 0;
    // Fluid (fluid.rs:10-45)
    T rest_density = 0, dynamic_viscosity = 0, eos_stiffness = 0, eos_power = 0, pressure_floor = 0,
      min_volume = T(1e-6f), thermal_viscosity_coeff = 0, bulk_viscosity = 0, surface_tension_coeff = 0,
      settling_damping = 0;
    // DruckerPrager (sand.rs:11-75)
    T friction_angle = 0, hardening_peak = 0, hardening_decay = 0, friction_residual = 0, volume_correction = 1,
      dilatancy_angle = 0, cohesion = 0, min_volume_jacobian = 0;
    // Snow (snow.rs:19-37)
    T hardening_exponent = 0, compression_limit = 0, stretch_limit = 0, min_plastic_jacobian = 0,
      max_plastic_jacobian = 0, cohesion_coeff = 0;
    // NACC (nacc.rs:34-60)
    T kappa = 0, friction_m = 0, cohesion_beta = 0, hardening_factor = 0;
    // VonMises (von_mises.rs:29-39), Viscoelastic (viscoelastic.rs:9-16)
    T yield_stress = 0, hardening_modulus = 0, ve_viscosity = 0, ve_j_min = 0;
    // Rankine (rankine.rs:37-49)
    T tensile_strength = 0, softening_rate = 0;
    // DruckerPragerMuI (sand_mui.rs:44-58)
    T mu_static = 0, mu_dynamic = 0, inertial_q = 0;
    // Bingham (bingham.rs:26-48): a Fluid record whose params() carries the
    // yield stress in compression_limit (bingham.rs:205-218)
    bool bingham = false;
    T bingham_yield = 0, critical_shear_rate = T(1.0e-4f);
    bool hardening_enabled = false;
    bool nacc_advance_f = false;

    static Material from_params(const emh_material_params& p, const emh_material_ext* e) {
        Material m;
        m.model = p.model;
        m.lambda = T(p.lambda);
        m.mu = T(p.mu);
        float min_density = e ? e->min_density : 1e-6f;
        float min_volume = e ? e->min_volume : 1e-6f;
        switch (p.model) {
            case EMH_MODEL_NEOHOOKEAN:
                m.min_density = T(min_density);
                m.thermal_expansion = T(p.thermal_expansion);
                m.active_stress_coeff = T(p.active_stress_coeff);
                m.damage_softening_rate = T(p.cohesion_coeff);
                m.viscosity = T(p.dynamic_viscosity);
                break;
            case EMH_MODEL_COROTATED:
                m.thermal_expansion = T(p.thermal_expansion);
                m.active_stress_coeff = T(p.active_stress_coeff);
                break;
            case EMH_MODEL_FLUID:
                m.rest_density = T(p.rest_density);
                m.eos_stiffness = T(p.eos_stiffness);
                m.eos_power = T(p.eos_power);
                m.dynamic_viscosity = T(p.dynamic_viscosity);
                m.thermal_viscosity_coeff = T(p.thermal_viscosity_coeff);
                m.pressure_floor = T(p.pressure_floor);
                m.bulk_viscosity = T(p.bulk_viscosity);
                m.surface_tension_coeff = T(p.surface_tension_coeff);
                m.settling_damping = T(p.dp_h0);
                m.min_density = T(min_density);
                m.min_volume = T(min_volume);
                m.bingham = p.compression_limit > 0.0f || (e && e->fluid_bingham != 0u);
                m.bingham_yield = T(p.compression_limit);
                if (e && e->bingham_critical_shear_rate > 0.0f) m.critical_shear_rate = T(e->bingham_critical_shear_rate);
                break;
            case EMH_MODEL_RANKINE:  // rankine.rs:206-216
                m.tensile_strength = T(p.hardening_exponent);
                m.softening_rate = T(p.hardening_modulus);
                break;
            case EMH_MODEL_MU_I:  // sand_mui.rs:219-229
                m.mu_static = T(p.dp_h0);
                m.mu_dynamic = T(p.dp_h1);
                m.inertial_q = T(p.dp_h2);
                break;
            case EMH_MODEL_GRANULAR_FLUID:  // granular_fluid.rs:232-247
                m.rest_density = T(p.rest_density);
                m.eos_stiffness = T(p.eos_stiffness);
                m.eos_power = T(p.eos_power);
                m.hardening_exponent = T(p.hardening_exponent);
                m.compression_limit = T(p.compression_limit);
                m.stretch_limit = T(p.stretch_limit);
                m.min_plastic_jacobian = T(p.volume_ratio_min);
                m.max_plastic_jacobian = T(p.volume_ratio_max);
                m.pressure_floor = T(p.pressure_floor);
                break;
            case EMH_MODEL_DRUCKER_PRAGER:
                m.friction_angle = T(p.dp_h0);
                m.hardening_peak = T(p.dp_h1);
                m.hardening_decay = T(p.dp_h2);
                m.friction_residual = T(p.dp_h3);
                m.dilatancy_angle = T(p.compression_limit);
                m.cohesion = T(p.stretch_limit);
                m.min_volume_jacobian = T(p.volume_ratio_min);
                m.volume_correction = T(e ? e->dp_volume_correction : 1.0f);
                break;
            case EMH_MODEL_SNOW:
                m.hardening_exponent = T(p.hardening_exponent);
                m.compression_limit = T(p.compression_limit);
                m.stretch_limit = T(p.stretch_limit);
                m.min_plastic_jacobian = T(p.volume_ratio_min);
                m.max_plastic_jacobian = T(p.volume_ratio_max);
                m.cohesion_coeff = T(p.cohesion_coeff);
                break;
            case EMH_MODEL_VON_MISES:  // von_mises.rs:150-159
                m.yield_stress = T(p.hardening_exponent);
                m.hardening_modulus = T(p.hardening_modulus);
                break;
            case EMH_MODEL_VISCOELASTIC:  // viscoelastic.rs:149-159
                m.ve_viscosity = T(p.dynamic_viscosity);
                m.active_stress_coeff = T(p.active_stress_coeff);
                m.ve_j_min = T(p.volume_ratio_min);
                break;
            case EMH_MODEL_NACC: {
                m.min_density = T(min_density);
                float kappa = (e && e->nacc_kappa != 0.0f) ? e->nacc_kappa : p.lambda + p.mu;
                m.kappa = T(kappa);
                m.hardening_factor = T(p.hardening_exponent);
                m.cohesion_beta = T(p.compression_limit);
                m.friction_m = T(p.stretch_limit);
                uint32_t he = e ? e->nacc_hardening_enabled : 2u;
                m.hardening_enabled = (he == 2u) ? (p.hardening_exponent > 0.0f) : (he != 0u);
                m.nacc_advance_f = e ? (e->nacc_advance_f != 0u) : false;
                break;
            }
            default:
                break;
        }
        return m;
    }

    bool needs_density_recompute() const { return model == EMH_MODEL_FLUID; }  // fluid.rs:261-263

    // init_particle: corotated.rs:57-60, sand.rs:237-246, snow.rs:104-107, nacc.rs:271-273
    void init_particle(emh_particle& p) const {
        switch (model) {
            case EMH_MODEL_COROTATED:
                p.hardening_scale = 1.0f;
                p.plastic_volume_ratio = 1.0f;
                break;
            case EMH_MODEL_DRUCKER_PRAGER:
                p.friction_hardening = hardening_peak > T(0) ? (float)(friction_residual / hardening_peak) : 0.0f;
                break;
            case EMH_MODEL_SNOW:
                p.plastic_volume_ratio = 1.0f;
                p.hardening_scale = 1.0f;
                break;
            case EMH_MODEL_NACC:
                p.log_volume_strain = 0.0f;
                break;
            case EMH_MODEL_MU_I:  // sand_mui.rs:146-149
                p.friction_hardening = (float)mu_static;
                break;
            case EMH_MODEL_GRANULAR_FLUID:  // granular_fluid.rs:166-169
                p.plastic_volume_ratio = 1.0f;
                p.hardening_scale = 1.0f;
                break;
            default:
                break;
        }
    }

    // stress_volume: materials/mod.rs:118-120 (default), fluid.rs:193-195
    T stress_volume(const Particles<T>& P, size_t i) const {
        if (model == EMH_MODEL_FLUID) return rmax(P.volume[i], min_volume);
        if (model == EMH_MODEL_GRANULAR_FLUID) return rmax(P.volume[i], T(1.0e-6f));  // granular_fluid.rs:196-198
        return P.initial_volume[i];
    }

    M2<T> kirchhoff_stress(const Particles<T>& P, size_t i) const {
        switch (model) {
            case EMH_MODEL_NEOHOOKEAN: {  // elastic.rs:180-255
                M2<T> f = P.f[i];
                T j = det(f);
                if (j <= T(MIN_J)) return m2_zero<T>();
                T t_scale = T(1) + thermal_expansion * P.temperature[i];
                T damage_scale = o_exp(-damage_softening_rate * P.friction_hardening[i]);
                T mu_ = mu * t_scale * damage_scale;
                T lambda_ = lambda * t_scale * damage_scale;
                M2<T> b = f * transpose(f);
                T tr_b = b.m00 + b.m11;
                M2<T> dev_b = b - m2_diag(tr_b * T(0.5), tr_b * T(0.5));
                T k = lambda_ + mu_;
                M2<T> dev_stress = (mu_ / j) * dev_b;
                M2<T> vol_stress = (k * o_ln(j)) * m2_identity<T>();
                M2<T> viscous = m2_zero<T>();
                if (viscosity > T(0)) {
                    M2<T> c = P.c[i];
                    M2<T> sym = c + transpose(c);
                    M2<T> d = sym * T(0.5);
                    T d_trace = d.m00 + d.m11;
                    M2<T> d_dev = d - m2_diag(d_trace * T(0.5), d_trace * T(0.5));
                    viscous = viscosity * d_dev;
                }
                return dev_stress + vol_stress + viscous;
            }
            case EMH_MODEL_COROTATED: {  // corotated.rs:62-78
                M2<T> f = P.f[i];
                T j = det(f);
                if (j <= T(MIN_J)) return m2_zero<T>();
                M2<T> r = polar_decomposition_2d(f);
                T h = P.hardening_scale[i];
                T t_scale = T(1) + thermal_expansion * P.temperature[i];
                T mu_eff = mu * h * t_scale;
                T lambda_eff = lambda * h * t_scale;
                M2<T> f_t = transpose(f);
                return ((T(2) * mu_eff) * (f - r)) * f_t + ((lambda_eff * (j - T(1))) * j) * m2_identity<T>();
            }
            case EMH_MODEL_DRUCKER_PRAGER:  // sand.rs:220-231
            case EMH_MODEL_RANKINE:         // rankine.rs:157-166 (the same corotated form)
            case EMH_MODEL_MU_I:            // sand_mui.rs:132-140
            case EMH_MODEL_VON_MISES: {     // von_mises.rs:96-104
                M2<T> f = P.f[i];
                T j = det(f);
                if (j <= T(MIN_J)) return m2_zero<T>();
                M2<T> r = polar_decomposition_2d(f);
                M2<T> f_t = transpose(f);
                return ((T(2) * mu) * (f - r)) * f_t + ((lambda * (j - T(1))) * j) * m2_identity<T>();
            }
            case EMH_MODEL_VISCOELASTIC: {  // viscoelastic.rs:103-130 (temperature scale 1: no thermal term)
                M2<T> f = P.f[i];
                T j = rmin(rmax(f.m00 * f.m11 - f.m10 * f.m01, ve_j_min), T(1) / ve_j_min);
                const T t_scale = T(1);
                T mu_t = mu * t_scale, lambda_t = lambda * t_scale;
                M2<T> b = f * transpose(f);
                M2<T> elastic = mu_t * (b - m2_identity<T>()) + m2_diag(lambda_t * o_ln(j), lambda_t * o_ln(j));
                M2<T> c = P.c[i];
                M2<T> d = (c + transpose(c)) * T(0.5);
                T trace = d.m00 + d.m11;
                M2<T> d_dev = d - m2_diag(trace * T(0.5), trace * T(0.5));
                M2<T> viscous = (ve_viscosity * t_scale) * d_dev;
                return elastic + viscous;
            }
            case EMH_MODEL_SNOW: {  // snow.rs:109-136
                M2<T> f = P.f[i];
                T j = det(f);
                if (j <= T(MIN_J)) return m2_zero<T>();
                M2<T> r = polar_decomposition_2d(f);
                T h = P.hardening_scale[i];
                T mu_eff = mu * h;
                T lambda_eff = lambda * h;
                M2<T> f_t = transpose(f);
                M2<T> tau = ((T(2) * mu_eff) * (f - r)) * f_t + ((lambda_eff * (j - T(1))) * j) * m2_identity<T>();
                T jp = P.plastic_volume_ratio[i];
                if (cohesion_coeff > T(0) && jp < T(1) && j > T(1)) {
                    tau = tau + (((cohesion_coeff * jp) * (j - T(1))) * j) * m2_identity<T>();
                }
                return tau;
            }
            case EMH_MODEL_GRANULAR_FLUID: {  // granular_fluid.rs:171-194
                M2<T> f = P.f[i];
                T j = rmax(det(f), T(MIN_J));
                T density = rmax(rest_density / j, T(1.0e-6f));
                T ratio = rmax(density / rmax(rest_density, T(1.0e-6f)), T(1.0e-6f));
                T pressure = rmax(eos_stiffness * (o_pow(ratio, eos_power) - T(1)), pressure_floor);
                T h = P.hardening_scale[i];
                M2<T> r = polar_decomposition_2d(f);
                T mu_eff = mu * h;
                M2<T> coro = ((T(2) * mu_eff) * (f - r)) * transpose(f);
                T tr = coro.m00 + coro.m11;
                M2<T> dev_coro = coro - m2_diag(tr * T(0.5), tr * T(0.5));
                M2<T> lam_vol = (((lambda * h) * (j - T(1))) * j) * m2_identity<T>();
                return (m2_diag(-pressure, -pressure) + dev_coro) + lam_vol;
            }
            case EMH_MODEL_FLUID: {  // fluid.rs:135-191
                if (bingham) return bingham_stress(P, i);
                T density = rmin(rmax(P.density[i], min_density), rest_density * T(2));
                T pressure = rmax(eos_stiffness * (o_pow(density / rest_density, eos_power) - T(1)), pressure_floor);
                M2<T> stress = m2_diag(-pressure, -pressure);
                T eff_visc = dynamic_viscosity;
                if (thermal_viscosity_coeff > T(0))
                    eff_visc = dynamic_viscosity * o_exp(-thermal_viscosity_coeff * P.temperature[i]);
                M2<T> c = P.c[i];
                M2<T> sym = c + transpose(c);
                T div_v = sym.m00 + sym.m11;
                M2<T> strain_dev = sym - m2_diag(div_v * T(0.5), div_v * T(0.5));
                stress = stress + eff_visc * strain_dev;
                if (bulk_viscosity > T(0)) {
                    T s = bulk_viscosity * div_v * T(0.5);
                    stress = stress + m2_diag(s, s);
                }
                if (surface_tension_coeff != T(0)) {
                    M2<T> f = P.f[i];
                    T j = f.m00 * f.m11 - f.m10 * f.m01;
                    T s = surface_tension_coeff * j;
                    stress = stress + m2_diag(s, s);
                }
                return stress;
            }
            case EMH_MODEL_NACC: {  // nacc.rs:236-249
                M2<T> f = P.f[i];
                T j = rmax(det(f), T(MIN_J));
                M2<T> b = f * transpose(f);
                T tr_b = b.m00 + b.m11;
                M2<T> dev_b = b - m2_diag(tr_b * T(0.5), tr_b * T(0.5));
                M2<T> dev_stress = (mu / j) * dev_b;
                M2<T> vol_stress = (kappa * T(0.5) * (j * j - T(1))) * m2_identity<T>();
                return dev_stress + vol_stress;
            }
            default:
                return m2_zero<T>();  // Fallback: materials/mod.rs:113-115
        }
    }

    // BinghamFluidMaterial::deviatoric_stress + kirchhoff_stress  bingham.rs:96-126,147-178
    M2<T> bingham_deviatoric(const M2<T>& c) const {
        M2<T> sym = c + transpose(c);
        M2<T> d = sym * T(0.5);
        T trace = d.m00 + d.m11;
        M2<T> d_dev = d - m2_diag(trace * T(0.5), trace * T(0.5));
        T dxx = d_dev.m00, dyy = d_dev.m11, dxy = d_dev.m10;
        T d_sq = dxx * dxx + dyy * dyy + T(2) * dxy * dxy;
        T shear_rate = o_sqrt(T(2) * d_sq);
        if (shear_rate < critical_shear_rate) return m2_zero<T>();
        T eta_app = bingham_yield / shear_rate + dynamic_viscosity;
        return d_dev * eta_app;
    }
    M2<T> bingham_stress(const Particles<T>& P, size_t i) const {
        T density = rmin(rmax(P.density[i], min_density), rest_density * T(2));
        T pressure = rmax(eos_stiffness * (o_pow(density / rest_density, eos_power) - T(1)), pressure_floor);
        M2<T> hydrostatic = m2_diag(-pressure, -pressure);
        M2<T> deviatoric = bingham_deviatoric(P.c[i]);
        M2<T> surface = m2_zero<T>();
        if (surface_tension_coeff != T(0)) {
            M2<T> f = P.f[i];
            T j = f.m00 * f.m11 - f.m10 * f.m01;
            surface = m2_diag(surface_tension_coeff * j, surface_tension_coeff * j);
        }
        return (hydrostatic + deviatoric) + surface;
    }

    T activation_scale() const {
        // elastic.rs, corotated.rs, viscoelastic.rs:145-147 report their coefficient
        if (model == EMH_MODEL_NEOHOOKEAN || model == EMH_MODEL_COROTATED || model == EMH_MODEL_VISCOELASTIC)
            return active_stress_coeff;
        return T(0);
    }

    // timestep_bound: elastic.rs:290-327, corotated.rs:108-125, fluid.rs:228-259,
    // sand.rs:345-362, snow.rs:193-211, nacc.rs:279-298
    T timestep_bound(T density, T hardening_scale, T cell_width, T material_cfl, T viscous_cfl) const {
        const T inf = std::numeric_limits<T>::infinity();
        switch (model) {
            case EMH_MODEL_NEOHOOKEAN: {
                T elastic_dt = elastic_wave_dt(lambda, mu, T(1), density, min_density, cell_width, material_cfl);
                T viscous_dt = inf;
                if (viscosity > T(0)) {
                    T d = rmax(density, T(1e-6f));
                    T kinematic = viscosity / d;
                    if (kinematic > T(F32_EPSILON)) viscous_dt = viscous_cfl * cell_width * cell_width / kinematic;
                }
                return rmin(elastic_dt, viscous_dt);
            }
            case EMH_MODEL_COROTATED:
            case EMH_MODEL_SNOW:
                return elastic_wave_dt(lambda, mu, hardening_scale, density, T(MIN_J), cell_width, material_cfl);
            case EMH_MODEL_DRUCKER_PRAGER:
            case EMH_MODEL_RANKINE:    // rankine.rs:218-229
            case EMH_MODEL_MU_I:       // sand_mui.rs:232-249
            case EMH_MODEL_VON_MISES:  // von_mises.rs:161-178
                return elastic_wave_dt(lambda, mu, T(1), density, T(MIN_J), cell_width, material_cfl);
            case EMH_MODEL_GRANULAR_FLUID:  // granular_fluid.rs:250-268
                return elastic_wave_dt(lambda, mu, hardening_scale, density, T(MIN_J), cell_width, material_cfl);
            case EMH_MODEL_VISCOELASTIC: {  // viscoelastic.rs:161-190
                T elastic_dt = elastic_wave_dt(lambda, mu, T(1), density, T(1.0e-6f), cell_width, material_cfl);
                T viscous_dt = inf;
                if (ve_viscosity > T(0)) {
                    T d = rmax(density, T(1.0e-6f));
                    T kinematic = ve_viscosity / d;
                    if (kinematic > T(F32_EPSILON)) viscous_dt = viscous_cfl * cell_width * cell_width / kinematic;
                }
                return rmin(elastic_dt, viscous_dt);
            }
            case EMH_MODEL_NACC:
                return elastic_wave_dt(kappa - mu, mu, hardening_scale, density, min_density, cell_width,
                                       material_cfl);
            case EMH_MODEL_FLUID: {
                const T MIN_DENSITY_RATIO = T(1e-6f);
                T d = rmax(density, min_density);
                T ratio = rmax(d / rmax(rest_density, min_density), MIN_DENSITY_RATIO);
                T dt_bound = inf;
                T c2 = eos_stiffness * eos_power * o_pow(ratio, eos_power - T(1)) / rmax(rest_density, min_density);
                if (finite(c2) && c2 > T(F32_EPSILON)) dt_bound = rmin(dt_bound, material_cfl * cell_width / o_sqrt(c2));
                if (dynamic_viscosity > T(0)) {
                    T kinematic = dynamic_viscosity / d;
                    if (kinematic > T(F32_EPSILON))
                        dt_bound = rmin(dt_bound, viscous_cfl * cell_width * cell_width / kinematic);
                }
                return dt_bound;
            }
            default:
                return inf;  // materials/mod.rs:127-136
        }
    }

    // DruckerPragerMaterial::alpha  sand.rs:135-141
    T dp_alpha(T q) const {
        T phi = friction_angle + (hardening_peak * q - friction_residual) * o_exp(-hardening_decay * q);
        T s = o_sin(phi);
        return o_sqrt(T(2.0f / 3.0f)) * (T(2) * s) / (T(3) - s);
    }
    // DruckerPragerMaterial::project  sand.rs:156-194; returns false for an elastic step.
    bool dp_project(V2<T> sigma, T log_volume_strain, T q, V2<T>& out_sigma, T& dq) const {
        V2<T> s{rmax(std::fabs(sigma.x), T(LOG_CLAMP)), rmax(std::fabs(sigma.y), T(LOG_CLAMP))};
        V2<T> eps{o_ln(s.x) + log_volume_strain * T(0.5), o_ln(s.y) + log_volume_strain * T(0.5)};
        T trace = eps.x + eps.y;
        V2<T> dev{eps.x - trace * T(0.5), eps.y - trace * T(0.5)};
        T dev_norm = length(dev);
        if (dev_norm == T(0) || trace > T(0)) {
            out_sigma = {T(1), T(1)};
            dq = dev_norm;
            return true;
        }
        T ratio = (lambda + mu) / mu;
        T alpha = dp_alpha(q);
        T cohesion_term = cohesion / (T(2) * mu);
        T gamma = dev_norm + ratio * trace * alpha - cohesion_term;
        if (gamma <= T(0)) return false;
        V2<T> h = eps - gamma * (dev / dev_norm);
        out_sigma = {o_exp(h.x), o_exp(h.y)};
        dq = gamma;
        return true;
    }

    // NaccMaterial::project  nacc.rs:129-223
    void nacc_project(M2<T> f, T alpha, M2<T>& f_out, T& alpha_out) const {
        T xi = hardening_factor, beta = cohesion_beta, m = friction_m;
        M2<T> u, vt;
        V2<T> sigma;
        svd2(f, u, sigma, vt);
        V2<T> sv = sigma;
        V2<T> sv_sq{sv.x * sv.x, sv.y * sv.y};
        T sv_sq_trace = sv_sq.x + sv_sq.y;
        T p0 = kappa * (T(1e-5f) + o_sinh(xi * rmax(-alpha, T(0))));
        T j_e_tr = rmax(sv.x * sv.y, T(1.0e-6f));
        T inv_j = T(1) / j_e_tr;
        V2<T> s_tr = (mu * inv_j) * V2<T>{sv_sq.x - sv_sq_trace * T(0.5), sv_sq.y - sv_sq_trace * T(0.5)};
        T psi_kappa = kappa * T(0.5) * (j_e_tr - T(1) / j_e_tr);
        T p_tr = -psi_kappa * j_e_tr;
        auto reconstruct = [&](V2<T> s) { return (u * m2_diag(s.x, s.y)) * vt; };
        if (p_tr > p0) {  // case A
            T j_n1 = o_sqrt(rmax(T(-2) * p0 / kappa + T(1), T(1.0e-8f)));
            T sv_new = o_pow(j_n1, T(0.5));
            if (hardening_enabled && !mutated(kMutNaccCaseANoAlpha)) alpha = alpha + o_ln(j_e_tr / j_n1);
            f_out = reconstruct({sv_new, sv_new});
            alpha_out = alpha;
            return;
        }
        if (p_tr < -beta * p0) {  // case B
            T j_n1 = o_sqrt(rmax(T(2) * beta * p0 / kappa + T(1), T(1.0e-8f)));
            T sv_new = o_pow(j_n1, T(0.5));
            if (hardening_enabled) alpha = alpha + o_ln(j_e_tr / j_n1);
            f_out = reconstruct({sv_new, sv_new});
            alpha_out = alpha;
            return;
        }
        T y0 = (T(1) + T(2) * beta) * T(2);
        T y1 = m * m * (p_tr + beta * p0) * (p_tr - p0);
        T s_norm_sq = s_tr.x * s_tr.x + s_tr.y * s_tr.y;
        T y = y0 * s_norm_sq + y1;
        if (y < T(1.0e-4f)) {
            f_out = f;
            alpha_out = alpha;
            return;
        }
        if (hardening_enabled && p0 > T(1.0e-4f) && p_tr < p0 - T(1.0e-4f) && p_tr > -beta * p0 + T(1.0e-4f)) {
            T p_c = (T(1) - beta) * p0 * T(0.5);
            T q_tr = o_sqrt(T(2)) * length(s_tr);
            V2<T> dir = normalize_or_zero(V2<T>{p_c - p_tr, -q_tr});
            T c = m * m * (p_c + beta * p0) * (p_c - p0);
            T b = m * m * dir.x * (T(2) * p_c - p0 + beta * p0);
            T a = m * m * dir.x * dir.x + (T(1) + T(2) * beta) * dir.y * dir.y;
            T discr = o_sqrt(rmax(b * b - T(4) * a * c, T(0)));
            T l1 = (-b + discr) / (T(2) * a);
            T l2 = (-b - discr) / (T(2) * a);
            T p1 = p_c + l1 * dir.x;
            T p2 = p_c + l2 * dir.x;
            T p_x = ((p_tr - p_c) * (p1 - p_c) > T(0)) ? p1 : p2;
            if (mutated(kMutNaccOtherRoot)) p_x = (p_x == p1) ? p2 : p1;
            T j_e_x = o_sqrt(rmax(std::fabs(T(-2) * p_x / kappa + T(1)), T(1.0e-8f)));
            if (j_e_x > T(1.0e-4f))
                alpha = mutated(kMutNaccHardenSign) ? alpha - o_ln(j_e_tr / j_e_x) : alpha + o_ln(j_e_tr / j_e_x);
        }
        T scale = o_sqrt(rmax(-y1 / rmax(y0, T(1.0e-10f)), T(0))) * (j_e_tr / mu);
        V2<T> b_n1 = scale * normalize_or_zero(s_tr) + V2<T>{sv_sq_trace / T(2), sv_sq_trace / T(2)};
        V2<T> sv_new{o_sqrt(rmax(b_n1.x, T(1.0e-8f))), o_sqrt(rmax(b_n1.y, T(1.0e-8f)))};
        f_out = reconstruct(sv_new);
        alpha_out = alpha;
    }

    void update_particle(Particles<T>& P, size_t i, T dt) const {
        switch (model) {
            case EMH_MODEL_NEOHOOKEAN:
            case EMH_MODEL_COROTATED:      // elastic.rs:262-269, corotated.rs:84-91
            case EMH_MODEL_VISCOELASTIC: {  // viscoelastic.rs:132-139
                M2<T> fp_new = m2_identity<T>() + dt * P.c[i];
                P.f[i] = fp_new * P.f[i];
                T j = rmax(det(P.f[i]), T(MIN_J));
                T v = rmax(P.initial_volume[i] * j, T(1.0e-6f));
                P.volume[i] = v;
                P.density[i] = P.mass[i] / v;
                break;
            }
            case EMH_MODEL_DRUCKER_PRAGER: {  // sand.rs:248-319
                M2<T> f_trial = (m2_identity<T>() + dt * P.c[i]) * P.f[i];
                M2<T> u, vt;
                V2<T> sigma;
                svd2(f_trial, u, sigma, vt);
                V2<T> new_sigma = sigma;
                V2<T> proj;
                T dq;
                if (dp_project(sigma, P.log_volume_strain[i], P.friction_hardening[i], proj, dq)) {
                    V2<T> sa{rmax(std::fabs(sigma.x), T(LOG_CLAMP)), rmax(std::fabs(sigma.y), T(LOG_CLAMP))};
                    T prev_det = sa.x * sa.y;
                    T new_det = proj.x * proj.y;
                    T diff = new_det - prev_det;
                    T corrected = diff > T(0) ? new_det : prev_det + diff * volume_correction;
                    P.log_volume_strain[i] = P.log_volume_strain[i] + (o_ln(prev_det) - o_ln(corrected));
                    T q_max = T(5) / rmax(hardening_decay, T(1e-6f));
                    P.friction_hardening[i] = rmin(P.friction_hardening[i] + dq, q_max);
                    if (dilatancy_angle > T(0))
                        P.log_volume_strain[i] = P.log_volume_strain[i] + o_sin(dilatancy_angle) * dq;
                    new_sigma = proj;
                }
                new_sigma = {std::fabs(new_sigma.x), std::fabs(new_sigma.y)};
                T j_new = new_sigma.x * new_sigma.y;
                if (j_new < min_volume_jacobian) {
                    T s = o_sqrt(min_volume_jacobian / rmax(j_new, T(1e-6f)));
                    new_sigma = new_sigma * s;
                }
                P.f[i] = (u * m2_diag(new_sigma.x, new_sigma.y)) * vt;
                T j = rmax(det(P.f[i]), T(MIN_J));
                T v = rmax(P.initial_volume[i] * j, T(1.0e-6f));
                P.volume[i] = v;
                P.density[i] = P.mass[i] / v;
                break;
            }
            case EMH_MODEL_VON_MISES: {  // von_mises.rs:113-147
                M2<T> f_trial = (m2_identity<T>() + dt * P.c[i]) * P.f[i];
                M2<T> u, vt;
                V2<T> sigma;
                svd2(f_trial, u, sigma, vt);
                // hencky_strains  utils.rs:58-61
                V2<T> eps{o_ln(rmax(std::fabs(sigma.x), T(LOG_CLAMP))), o_ln(rmax(std::fabs(sigma.y), T(LOG_CLAMP)))};
                T tr = eps.x + eps.y;
                V2<T> dev{eps.x - tr * T(0.5), eps.y - tr * T(0.5)};
                T dev_norm = length(dev);
                T kappa = P.friction_hardening[i];
                T effective_yield = yield_stress + hardening_modulus * kappa;
                T elastic_dev = (T(2) * mu) * dev_norm;
                V2<T> sigma_new = sigma;
                if (elastic_dev > effective_yield && dev_norm > T(LOG_CLAMP)) {
                    T denom = T(2) * mu + hardening_modulus;
                    T gamma = denom > T(F32_EPSILON) ? (elastic_dev - effective_yield) / denom : T(0);
                    P.friction_hardening[i] = kappa + gamma;
                    T k = effective_yield / elastic_dev;
                    V2<T> eps_proj{dev.x * k + tr * T(0.5), dev.y * k + tr * T(0.5)};
                    sigma_new = {o_exp(eps_proj.x), o_exp(eps_proj.y)};
                }
                P.f[i] = (u * m2_diag(sigma_new.x, sigma_new.y)) * vt;
                T j = rmax(det(P.f[i]), T(MIN_J));
                T v = rmax(P.initial_volume[i] * j, T(1.0e-6f));
                P.volume[i] = v;
                P.density[i] = P.mass[i] / v;
                break;
            }
            case EMH_MODEL_RANKINE: {  // rankine.rs:168-203
                M2<T> f_trial = (m2_identity<T>() + dt * P.c[i]) * P.f[i];
                M2<T> u, vt;
                V2<T> sigma;
                svd2(f_trial, u, sigma, vt);
                V2<T> eps{o_ln(rmax(std::fabs(sigma.x), T(LOG_CLAMP))), o_ln(rmax(std::fabs(sigma.y), T(LOG_CLAMP)))};
                T a = T(2) * mu + lambda;
                V2<T> tau{a * eps.x + lambda * eps.y, lambda * eps.x + a * eps.y};
                T damage = P.friction_hardening[i];
                // tensile_strength_eff  rankine.rs:113-117 (residual 5 %, utils.rs:28)
                T t_eff = rmax(tensile_strength * o_exp(-softening_rate * damage), tensile_strength * T(0.05f));
                // project_stress  rankine.rs:137-146
                bool t1 = tau.x > t_eff, t2 = tau.y > t_eff;
                V2<T> sigma_new = sigma;
                if (t1 || t2) {
                    V2<T> tp{t1 ? t_eff : tau.x, t2 ? t_eff : tau.y};
                    // stress_to_hencky  utils.rs:76-83
                    auto to_hencky = [&](V2<T> t) {
                        T dd = T(4) * mu * (mu + lambda);
                        T aa = T(2) * mu + lambda;
                        return V2<T>{(aa * t.x - lambda * t.y) / dd, (aa * t.y - lambda * t.x) / dd};
                    };
                    V2<T> eps_proj = to_hencky(tp);
                    V2<T> eps_trial = to_hencky(tau);
                    // rankine_damage_saturation_point  utils.rs:43-49
                    T sat = softening_rate <= T(0) ? std::numeric_limits<T>::infinity()
                                                   : -o_ln(T(0.05f)) / softening_rate;
                    V2<T> de{eps_trial.x - eps_proj.x, eps_trial.y - eps_proj.y};
                    P.friction_hardening[i] = rmin(damage + length(de), sat);
                    sigma_new = {o_exp(eps_proj.x), o_exp(eps_proj.y)};
                }
                P.f[i] = (u * m2_diag(sigma_new.x, sigma_new.y)) * vt;
                T j = rmax(det(P.f[i]), T(MIN_J));
                T v = rmax(P.initial_volume[i] * j, T(1.0e-6f));
                P.volume[i] = v;
                P.density[i] = P.mass[i] / v;
                break;
            }
            case EMH_MODEL_MU_I: {  // sand_mui.rs:151-217
                const T SQRT_2 = T(1.41421356237309504880);
                M2<T> f_trial = (m2_identity<T>() + dt * P.c[i]) * P.f[i];
                M2<T> u, vt;
                V2<T> sigma;
                svd2(f_trial, u, sigma, vt);
                V2<T> eps{o_ln(rmax(std::fabs(sigma.x), T(LOG_CLAMP))), o_ln(rmax(std::fabs(sigma.y), T(LOG_CLAMP)))};
                T tr = eps.x + eps.y;
                T k_2d = lambda + mu;
                T p_trial = -k_2d * tr;
                V2<T> sigma_new = sigma;
                T mu_out = mu_static;
                if (p_trial <= T(0)) {
                    sigma_new = {T(1), T(1)};
                } else {
                    V2<T> dev{eps.x - tr * T(0.5), eps.y - tr * T(0.5)};
                    T dev_norm = length(dev);
                    T q_trial = SQRT_2 * mu * dev_norm;
                    T q_yield = mu_static * p_trial;
                    if (!(q_trial <= q_yield || dev_norm < T(F32_EPSILON))) {
                        T delta_q = q_trial - q_yield;
                        T sqrt_p = o_sqrt(p_trial);
                        T a = mu * dt;
                        T b = p_trial * (mu_dynamic - mu_static) + a * inertial_q * sqrt_p - delta_q;
                        T c = -delta_q * inertial_q * sqrt_p;
                        T gamma_dot = (-b + o_sqrt(b * b - T(4) * a * c)) / (T(2) * a);
                        gamma_dot = rmax(gamma_dot, T(0));
                        if (gamma_dot > T(F32_EPSILON))
                            mu_out = mu_static + (mu_dynamic - mu_static) / (inertial_q * sqrt_p / gamma_dot + T(1));
                        T delta_gamma = gamma_dot * dt;
                        V2<T> n_hat = dev / dev_norm;
                        T s = delta_gamma / SQRT_2;
                        V2<T> eps_new{eps.x - n_hat.x * s, eps.y - n_hat.y * s};
                        sigma_new = {o_exp(eps_new.x), o_exp(eps_new.y)};
                    }
                }
                P.f[i] = (u * m2_diag(sigma_new.x, sigma_new.y)) * vt;
                P.friction_hardening[i] = mu_out;
                T j = rmax(det(P.f[i]), T(MIN_J));
                T v = rmax(P.initial_volume[i] * j, T(1.0e-6f));
                P.volume[i] = v;
                P.density[i] = P.mass[i] / v;
                break;
            }
            case EMH_MODEL_GRANULAR_FLUID: {  // granular_fluid.rs:200-230
                M2<T> f_trial = (m2_identity<T>() + dt * P.c[i]) * P.f[i];
                if (compression_limit > T(0) || stretch_limit > T(0)) {
                    M2<T> u, vt;
                    V2<T> sigma;
                    svd2(f_trial, u, sigma, vt);
                    T lo = T(1) - compression_limit, hi = T(1) + stretch_limit;
                    V2<T> sc{rclamp(sigma.x, lo, hi), rclamp(sigma.y, lo, hi)};
                    T jp_new = P.plastic_volume_ratio[i] * (sigma.x * sigma.y) / rmax(sc.x * sc.y, T(1.0e-10f));
                    P.plastic_volume_ratio[i] = rclamp(jp_new, min_plastic_jacobian, max_plastic_jacobian);
                    P.hardening_scale[i] =
                        rclamp(o_exp(hardening_exponent * (T(1) - P.plastic_volume_ratio[i])), T(0.1f), T(7.0f));
                    P.f[i] = (u * m2_diag(sc.x, sc.y)) * vt;
                } else {
                    P.f[i] = f_trial;
                }
                T j = rmax(det(P.f[i]), T(MIN_J));
                T v = rmax(P.initial_volume[i] * j, T(1.0e-6f));
                P.volume[i] = v;
                P.density[i] = P.mass[i] / v;
                break;
            }
            case EMH_MODEL_SNOW: {  // snow.rs:142-176
                M2<T> f_trial = (m2_identity<T>() + dt * P.c[i]) * P.f[i];
                M2<T> u, vt;
                V2<T> sigma;
                svd2(f_trial, u, sigma, vt);
                T lo = T(1) - compression_limit, hi = T(1) + stretch_limit;
                V2<T> sc{rclamp(sigma.x, lo, hi), rclamp(sigma.y, lo, hi)};
                T jp_new = P.plastic_volume_ratio[i] * (sigma.x * sigma.y) / (sc.x * sc.y);
                P.plastic_volume_ratio[i] = rclamp(jp_new, min_plastic_jacobian, max_plastic_jacobian);
                P.hardening_scale[i] =
                    rclamp(o_exp(hardening_exponent * (T(1) - P.plastic_volume_ratio[i])), T(0.1f), T(7.0f));
                P.f[i] = (u * m2_diag(sc.x, sc.y)) * vt;
                T j = rmax(det(P.f[i]), T(MIN_J));
                T v = rmax(P.initial_volume[i] * j, T(1.0e-6f));
                P.volume[i] = v;
                P.density[i] = P.mass[i] / v;
                break;
            }
            case EMH_MODEL_FLUID: {  // fluid.rs:197-207
                T j = rclamp(det(P.f[i]), T(0.5f), T(2.0f));
                T s = o_sqrt(j);
                P.f[i] = m2_diag(s, s);
                if (settling_damping > T(0)) {
                    T k = T(1) - rmin(settling_damping * dt, T(0.5f));
                    P.v[i] = P.v[i] * k;
                }
                break;
            }
            case EMH_MODEL_NACC: {  // nacc.rs:255-269 (+ optional advance, see emh_material_ext)
                M2<T> f_in = P.f[i];
                if (nacc_advance_f) f_in = (m2_identity<T>() + dt * P.c[i]) * f_in;
                M2<T> new_f;
                T new_alpha;
                nacc_project(f_in, P.log_volume_strain[i], new_f, new_alpha);
                P.f[i] = new_f;
                P.log_volume_strain[i] = new_alpha;
                T j = rmax(det(new_f), T(MIN_J));
                T vol = P.initial_volume[i] * j;
                P.volume[i] = vol;
                if (vol > T(MIN_J)) P.density[i] = P.mass[i] / vol;
                break;
            }
            default:
                break;  // Fallback: no-op
        }
    }
};

// ------------------------------------------------------------------- grid
// Sparse grid, grid/mod.rs:89-381: FxHash-keyed map from the flat index
// x * res + y to Cell{momentum, mass}, plus the insertion-ordered dirty list.
template <class T> struct Cell {
    V2<T> momentum;
    T mass;
};
template <class T> struct Grid {
    static constexpr uint64_t FXHASH_K = 0x517cc1b727220a95ULL;  // grid/mod.rs:41
    static constexpr uint32_t EMPTY = 0xFFFFFFFFu;
    uint32_t res = 0;
    std::vector<uint32_t> keys;
    std::vector<Cell<T>> cells;
    std::vector<uint32_t> dirty_slots;  // slot per touched cell, insertion order
    uint32_t mask = 0;
    uint32_t shift = 0;

    void init(uint32_t resolution) {
        res = resolution;
        rehash(1024);
    }
    void rehash(uint32_t cap) {
        std::vector<uint32_t> old_keys;
        std::vector<Cell<T>> old_cells;
        std::vector<uint32_t> old_dirty;
        old_keys.swap(keys);
        old_cells.swap(cells);
        old_dirty.swap(dirty_slots);
        keys.assign(cap, EMPTY);
        cells.assign(cap, Cell<T>{});
        mask = cap - 1;
        shift = 64u - (uint32_t)__builtin_ctz(cap);
        for (uint32_t s : old_dirty) insert_new(old_keys[s], old_cells[s]);
    }
    inline uint32_t home(uint32_t key) const { return (uint32_t)(((uint64_t)key * FXHASH_K) >> shift); }
    inline int64_t find(uint32_t key) const {
        uint32_t s = home(key);
        while (true) {
            uint32_t k = keys[s];
            if (k == key) return s;
            if (k == EMPTY) return -1;
            s = (s + 1) & mask;
        }
    }
    void insert_new(uint32_t key, const Cell<T>& c) {
        uint32_t s = home(key);
        while (keys[s] != EMPTY) s = (s + 1) & mask;
        keys[s] = key;
        cells[s] = c;
        dirty_slots.push_back(s);
    }
    // Grid::clear  grid/mod.rs:167-180
    void clear() {
        for (uint32_t s : dirty_slots) keys[s] = EMPTY;
        dirty_slots.clear();
    }
    // flat_index  grid/mod.rs:89-99
    inline bool flat_index(int32_t cx, int32_t cy, uint32_t& idx) const {
        if (cx < 0 || cy < 0) return false;
        if ((uint32_t)cx >= res || (uint32_t)cy >= res) return false;
        idx = (uint32_t)cx * res + (uint32_t)cy;
        return true;
    }
    // add_mass_momentum / accumulate  grid/mod.rs:190-212
    void add_mass_momentum(int32_t cx, int32_t cy, T mass, V2<T> momentum) {
        uint32_t idx;
        if (!flat_index(cx, cy, idx)) return;
        int64_t s = find(idx);
        if (s >= 0) {
            Cell<T>& c = cells[(size_t)s];
            c.mass = c.mass + mass;
            c.momentum = c.momentum + momentum;
        } else {
            if ((dirty_slots.size() + 1) * 2 > keys.size()) rehash((uint32_t)keys.size() * 2);
            insert_new(idx, Cell<T>{momentum, mass});
        }
    }
    // velocity_at / mass_at  grid/mod.rs:215-255
    V2<T> velocity_at(int32_t cx, int32_t cy) const {
        uint32_t idx;
        if (!flat_index(cx, cy, idx)) return {T(0), T(0)};
        int64_t s = find(idx);
        return s >= 0 ? cells[(size_t)s].momentum : V2<T>{T(0), T(0)};
    }
    T mass_at(int32_t cx, int32_t cy) const {
        uint32_t idx;
        if (!flat_index(cx, cy, idx)) return T(0);
        int64_t s = find(idx);
        return s >= 0 ? cells[(size_t)s].mass : T(0);
    }
    // cell_is_active  grid/mod.rs:231-241: touched by this substep's P2G
    bool cell_is_active(int32_t cx, int32_t cy) const {
        uint32_t idx;
        if (!flat_index(cx, cy, idx)) return false;
        return find(idx) >= 0;
    }
};

// ----------------------------------------------------------------- spawning
// LcgRng  solver/mod.rs:154-175
struct LcgRng {
    uint32_t state;
    explicit LcgRng(uint32_t seed) : state(seed) {}
    uint32_t next_u32() {
        state = state * 1664525u + 1013904223u;
        return state;
    }
    float next_f32() { return (float)next_u32() / ((float)4294967295u + 1.0f); }
};

// initialize_particles  solver/mod.rs:88-152 (f32 throughout, as in Rust)
inline std::vector<emh_particle> initialize_particles(const emh_config& cfg, const emh_spawn_region& sp,
                                                      LcgRng& rng) {
    std::vector<emh_particle> out;
    float mass = sp.has_mass_override ? sp.mass_override : cfg.particle_mass;
    float hx = (float)sp.box_size[0] * 0.5f, hy = (float)sp.box_size[1] * 0.5f;
    float minx = sp.box_center[0] - hx, miny = sp.box_center[1] - hy;
    float maxx = sp.box_center[0] + hx, maxy = sp.box_center[1] + hy;
    for (float i = minx; i < maxx; i += sp.spacing) {
        for (float j = miny; j < maxy; j += sp.spacing) {
            bool inside = true;
            if (sp.shape == EMH_SHAPE_DISK) {
                float dx = i - sp.box_center[0], dy = j - sp.box_center[1];
                inside = sqrtf(dx * dx + dy * dy) <= sp.disk_radius;
            }
            if (inside) {
                float jitter_mag = sp.position_jitter * sp.spacing;
                float jx = (rng.next_f32() - 0.5f) * 2.0f * jitter_mag;
                float jy = (rng.next_f32() - 0.5f) * 2.0f * jitter_mag;
                float rx = rng.next_f32();
                float ry = rng.next_f32();
                emh_particle p;
                std::memset(&p, 0, sizeof(p));
                p.x[0] = i + jx;
                p.x[1] = j + jy;
                p.v[0] = (rx - 0.5f) * sp.initial_velocity_scale;
                p.v[1] = (ry - 0.5f) * sp.initial_velocity_scale;
                for (int k = 0; k < 4; ++k) p.deformation_gradient[k] = sp.initial_deformation_gradient[k];
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
    }
    return out;
}

// SpawnRegion::fits_in_sim / validate_for_sim  config.rs:472-505
inline bool spawn_fits(const emh_spawn_region& sp, const emh_config& cfg) {
    if (!(sp.spacing > 0.0f) || sp.box_size[0] <= 0 || sp.box_size[1] <= 0) return false;
    float hx = (float)sp.box_size[0] * 0.5f, hy = (float)sp.box_size[1] * 0.5f;
    float minx = sp.box_center[0] - hx, miny = sp.box_center[1] - hy;
    float maxx = sp.box_center[0] + hx, maxy = sp.box_center[1] + hy;
    float dmin = (float)cfg.boundary_thickness;
    float dmax = (float)(cfg.grid_res > cfg.boundary_thickness ? cfg.grid_res - cfg.boundary_thickness : 0u);
    return minx >= dmin && miny >= dmin && maxx <= dmax && maxy <= dmax;
}

// ---------------------------------------------------------------- boundaries
// One record per installed BoundaryCondition, applied in installation order
// (step.rs:334-344 per cell, g2p.rs:316-320 per particle).
enum : uint32_t { BND_SLIP = 0, BND_FRICTION = 1, BND_HEIGHTMAP = 2, BND_PREDICTIVE = 3 };
struct Boundary {
    uint32_t kind = BND_SLIP;
    uint32_t thickness = 0;         // Slip / Friction thickness, Heightmap wall_thickness
    float mu = 0.0f;                // Friction coefficient, Heightmap terrain friction
    float wall_min = 0.0f;          // Predictive predictive_wall_min
    std::vector<float> heights;     // Heightmap terrain height per x column
};

// Coulomb wall  boundary/mod.rs:57-75
template <class T> inline void coulomb_wall(V2<T>& v, V2<T> n, T mu) {
    T vn = dot(v, n);
    if (vn >= T(0)) return;
    T normal_speed = std::fabs(vn);
    V2<T> vt = v - n * vn;
    T vt_len = length(vt);
    T fi = mu * normal_speed;
    v = vt_len > fi ? vt * ((vt_len - fi) / vt_len) : V2<T>{T(0), T(0)};
}

// BoundaryCondition::apply_to_grid_velocity: slip.rs:16-18 -> mod.rs:78-101,
// friction.rs:40-59, heightmap.rs:66-98, predictive.rs:27-29
template <class T> inline void boundary_velocity(const Boundary& b, uint32_t idx, uint32_t R, V2<T>& v) {
    const uint32_t x = idx / R, y = idx % R, t = b.thickness;
    switch (b.kind) {
        case BND_SLIP:
        case BND_PREDICTIVE: {
            uint32_t hi = R - (t + 1);
            if (x < t) v.x = rmax(v.x, T(0));
            if (x > hi) v.x = rmin(v.x, T(0));
            if (y < t) v.y = rmax(v.y, T(0));
            if (y > hi) v.y = rmin(v.y, T(0));
            break;
        }
        case BND_FRICTION: {
            uint32_t hi = R > t + 1 ? R - (t + 1) : 0u;
            T mu = T(b.mu);
            if (x < t) coulomb_wall(v, V2<T>{T(1), T(0)}, mu);
            if (x > hi) coulomb_wall(v, V2<T>{T(-1), T(0)}, mu);  // Vec2::NEG_X
            if (y < t) coulomb_wall(v, V2<T>{T(0), T(1)}, mu);
            if (y > hi) coulomb_wall(v, V2<T>{T(0), T(-1)}, mu);  // Vec2::NEG_Y
            break;
        }
        case BND_HEIGHTMAP: {
            uint32_t hi = R > t + 1 ? R - (t + 1) : 0u;
            if (x < t) v.x = rmax(v.x, T(0));
            if (x > hi) v.x = rmin(v.x, T(0));
            if (y > hi) v.y = rmin(v.y, T(0));
            T h = b.heights.empty() ? T(0) : T(b.heights[std::min<size_t>(x, b.heights.size() - 1)]);
            if (T((float)y) <= h && v.y < T(0)) {
                T vn = std::fabs(v.y);
                v.y = T(0);
                if (T(b.mu) > T(0)) {
                    T fi = T(b.mu) * vn;
                    T vt = std::fabs(v.x);
                    v.x = vt > fi ? v.x * (T(1) - fi / vt) : T(0);
                }
            }
            break;
        }
    }
}

// BoundaryCondition::clamp_particle_position: mod.rs:103-111 (Slip, Friction),
// heightmap.rs:100-114, predictive.rs:31-35
template <class T> inline V2<T> boundary_clamp(const Boundary& b, V2<T> pos, uint32_t R) {
    auto clamp2 = [](V2<T> p, T lo, T hi) { return V2<T>{rmin(rmax(p.x, lo), hi), rmin(rmax(p.y, lo), hi)}; };
    const uint32_t t = b.thickness;
    if (b.kind == BND_PREDICTIVE) {
        T lo = T(b.wall_min);
        T hi = T((float)R - 1.0f) - T(b.wall_min);
        return clamp2(pos, lo, hi);
    }
    pos = clamp2(pos, T((float)(t > 0 ? t - 1 : 0u)), T((float)(R > t ? R - t : 0u)));
    if (b.kind == BND_HEIGHTMAP) {
        // `as usize` saturates: negative and NaN give 0
        uint32_t xc = (pos.x >= T(1)) ? (uint32_t)std::min<T>(pos.x, T(4294967295.0)) : 0u;
        xc = std::min<uint32_t>(xc, R > 0 ? R - 1 : 0u);
        T h = b.heights.empty() ? T(0) : T(b.heights[std::min<size_t>(xc, b.heights.size() - 1)]);
        if (pos.y < h + T(1)) pos.y = h + T(1);
    }
    return pos;
}

// ---------------------------------------------------------------- simulation
template <class T> struct Simulation {
    emh_config cfg{};
    std::vector<Material<T>> materials;
    std::vector<Boundary> boundaries;
    Particles<T> P;
    Grid<T> grid;
    // sleep partition (solver/mod.rs:40-42): P[0, active_count) active,
    // P[active_count, N) sleeping; P2G / G2P / CFL / projection visit the active part
    size_t active_count = 0;
    uint32_t next_tag = 1;
    T last_step_dt = 0;
    uint32_t last_substeps = 0, last_vel_clamp_count = 0, last_j_projection_count = 0, total_substeps = 0;
    T last_sim_time_dropped = 0;
    uint64_t frame_index = 0;
    std::vector<T> substep_dts;  // record of every substep dt taken (diagnostics)

    // sleep_particle / wake_particle  solver/particles.rs:227-268 (tag_index,
    // a lookup structure, is not kept: nothing here queries by tag)
    void sleep_particle(size_t i) {
        if (P.sleeping[i] || active_count == 0) return;
        P.sleeping[i] = 1;
        const size_t last_active = active_count - 1;
        if (i != last_active) P.swap(i, last_active);
        active_count -= 1;
    }
    void wake_particle(size_t i) {
        if (!P.sleeping[i]) return;
        P.sleeping[i] = 0;
        const size_t first_sleeping = active_count;
        if (i != first_sleeping) P.swap(i, first_sleeping);
        active_count += 1;
    }
    // wake_tag / sleep_tag  solver/particles.rs:270-312: ascending wakes, descending sleeps
    void wake_tag(uint32_t tag) {
        std::vector<size_t> idx;
        for (size_t i = 0; i < P.size(); ++i)
            if (P.user_tag[i] == tag && P.sleeping[i]) idx.push_back(i);
        for (size_t i : idx) wake_particle(i);
    }
    void sleep_tag(uint32_t tag) {
        std::vector<size_t> idx;
        for (size_t i = 0; i < P.size(); ++i)
            if (P.user_tag[i] == tag && !P.sleeping[i]) idx.push_back(i);
        for (size_t k = idx.size(); k-- > 0;) sleep_particle(idx[k]);
    }
    // A particle set handed over whole: awake ones first (stable), then the sleeping ones.
    void reset_partition() {
        std::vector<size_t> order;
        for (size_t i = 0; i < P.size(); ++i)
            if (!P.sleeping[i]) order.push_back(i);
        active_count = order.size();
        for (size_t i = 0; i < P.size(); ++i)
            if (P.sleeping[i]) order.push_back(i);
        Particles<T> q;
        for (size_t i : order) q.push(P.get(i));
        P = q;
    }

    // Simulation::apply_impulse / apply_radial_impulse  solver/particles.rs:115-165:
    // every particle in range, sleeping ones woken afterwards in index order
    void apply_impulse(V2<T> center, T radius, V2<T> force, T strength, bool radial) {
        T vel_limit = T(cfg.grid_cell_size) / T(cfg.min_dt);
        T r2 = radius * radius;
        std::vector<size_t> to_wake;
        for (size_t i = 0; i < P.size(); ++i) {
            V2<T> d = P.x[i] - center;
            T dist2 = dot(d, d);
            if (!(dist2 <= r2 && dist2 > T(1e-8f))) continue;
            if (P.sleeping[i]) to_wake.push_back(i);
            if (radial) {
                T dist = o_sqrt(dist2);
                T falloff = T(1) - dist / radius;
                P.v[i] = P.v[i] + ((d / dist) * strength) * falloff;
            } else {
                T falloff = T(1) - o_sqrt(dist2 / r2);
                P.v[i] = P.v[i] + force * falloff;
            }
            T spd = length(P.v[i]);
            if (spd > vel_limit) P.v[i] = P.v[i] * (vel_limit / spd);
        }
        for (size_t i : to_wake) wake_particle(i);
    }

    void apply_impulse_f(float cx, float cy, float radius, float fx, float fy, float strength, bool radial) {
        apply_impulse(V2<T>{T(cx), T(cy)}, T(radius), V2<T>{T(fx), T(fy)}, T(strength), radial);
    }

    // Simulation::empty  lifecycle.rs:26-55
    void init_empty(const emh_config& c) {
        cfg = c;
        materials.assign(1, Material<T>{});  // FallbackMaterial
        Boundary b;
        b.kind = BND_SLIP;
        b.thickness = c.boundary_thickness;
        boundaries.assign(1, b);
        grid.init(c.grid_res);
        last_step_dt = T(c.dt);
    }
    const Material<T>& mat(uint32_t id) const {  // registry.rs:66-76
        return id < materials.size() ? materials[id] : materials[0];
    }
    bool any_needs_density_recompute() const {
        for (auto& m : materials)
            if (m.needs_density_recompute()) return true;
        return false;
    }
    // reinit_all_particle_state  lifecycle.rs:328-334
    void reinit_all() {
        for (size_t i = 0; i < P.size(); ++i) {
            emh_particle p = P.get(i);
            mat(p.material_id).init_particle(p);
            P.set(i, p);
        }
    }

    // estimate_particle_volumes  density.rs:216-247
    void estimate_particle_volumes(size_t count, bool write_initial) {
        grid.clear();
        for (size_t i = 0; i < count; ++i) {  // scatter_particle_mass  transfer/p2g.rs:230-243
            Weights<T> w = quadratic_weights(P.x[i]);
            for (int gx = 0; gx < 3; ++gx)
                for (int gy = 0; gy < 3; ++gy) {
                    T weight = w.wx[gx] * w.wy[gy];
                    grid.add_mass_momentum(w.bx + gx - 1, w.by + gy - 1, weight * P.mass[i], V2<T>{T(0), T(0)});
                }
        }
        gather_density(0, count, write_initial);
    }
    void gather_density(size_t from, size_t to, bool write_initial) {
        for (size_t i = from; i < to; ++i) {
            Weights<T> w = quadratic_weights(P.x[i]);
            T density = T(0);
            for (int gx = 0; gx < 3; ++gx)
                for (int gy = 0; gy < 3; ++gy) {
                    T weight = w.wx[gx] * w.wy[gy];
                    density = density + grid.mass_at(w.bx + gx - 1, w.by + gy - 1) * weight;
                }
            if (density > T(F32_EPSILON)) {
                P.density[i] = density;
                P.volume[i] = P.mass[i] / density;
                if (write_initial) P.initial_volume[i] = P.volume[i];
            }
        }
    }
    // estimate_particle_volumes_local  density.rs:256-319
    void estimate_particle_volumes_local(size_t active_count, size_t new_start, bool write_initial) {
        if (new_start >= active_count) return;
        T lox = std::numeric_limits<float>::max(), loy = std::numeric_limits<float>::max();
        T hix = std::numeric_limits<float>::lowest(), hiy = std::numeric_limits<float>::lowest();
        for (size_t i = new_start; i < active_count; ++i) {
            lox = rmin(lox, P.x[i].x);
            loy = rmin(loy, P.x[i].y);
            hix = rmax(hix, P.x[i].x);
            hiy = rmax(hiy, P.x[i].y);
        }
        const T MARGIN = T(3);
        lox = lox - MARGIN; loy = loy - MARGIN;
        hix = hix + MARGIN; hiy = hiy + MARGIN;
        grid.clear();
        for (size_t i = 0; i < active_count; ++i) {
            V2<T> x = P.x[i];
            if (x.x < lox || x.y < loy || x.x > hix || x.y > hiy) continue;
            Weights<T> w = quadratic_weights(x);
            for (int gx = 0; gx < 3; ++gx)
                for (int gy = 0; gy < 3; ++gy) {
                    T weight = w.wx[gx] * w.wy[gy];
                    grid.add_mass_momentum(w.bx + gx - 1, w.by + gy - 1, weight * P.mass[i], V2<T>{T(0), T(0)});
                }
        }
        gather_density(new_start, active_count, write_initial);
    }

    // add_body  solver/particles.rs:355-424: the new particles join the active
    // part, in front of any sleeping ones
    int add_body(const emh_spawn_region& sp, uint32_t* tag_out) {
        if (!spawn_fits(sp, cfg)) return EMH_ERR_INVALID_ARG;
        uint32_t tag = next_tag++;
        const size_t old_active = active_count;
        size_t old_len = P.size();
        LcgRng rng(sp.rng_seed);
        std::vector<emh_particle> np = initialize_particles(cfg, sp, rng);
        bool registered = sp.material_id < materials.size();
        for (auto& p : np) {
            p.user_tag = tag;
            if (registered) materials[sp.material_id].init_particle(p);
            P.push(p);
        }
        const size_t new_count = P.size() - old_len;
        if (old_len > old_active) P.rotate_range(old_active, old_len, P.size());
        active_count = old_active + new_count;
        estimate_particle_volumes_local(active_count, old_active, true);
        if (tag_out) *tag_out = tag;
        return EMH_OK;
    }

    // Simulation::new(config, spawn)  lifecycle.rs:57-105
    int init_from_spawn(const emh_config& c, const emh_spawn_region& sp) {
        init_empty(c);
        if (!spawn_fits(sp, cfg)) return EMH_ERR_INVALID_ARG;
        LcgRng rng(sp.rng_seed);
        for (auto& p : initialize_particles(cfg, sp, rng)) P.push(p);
        active_count = P.size();
        if (sp.precompute_initial_volumes) estimate_particle_volumes(active_count, true);
        return EMH_OK;
    }

    // affine_cfl_speed_contribution  step.rs:481-488
    T affine_cfl(const M2<T>& c) const {
        const T STENCIL_CORNER_DISTANCE = T(1.5f * 1.41421356237309504880f);
        T ls0 = c.m00 * c.m00 + c.m10 * c.m10;
        T ls1 = c.m01 * c.m01 + c.m11 * c.m11;
        T grad_norm = o_sqrt(ls0 + ls1);
        return grad_norm * STENCIL_CORNER_DISTANCE * T(cfg.grid_cell_size);
    }
    // choose_substep_dt + cfl_bound  step.rs:434-479
    T choose_substep_dt(T max_dt) const {
        if (!cfg.adaptive_timestep) return rmin(max_dt, T(cfg.dt));
        T max_speed = T(0);
        T min_mat_dt = max_dt;
        for (size_t i = 0; i < active_count; ++i) {
            T s = length(P.v[i]);
            if (cfg.cfl_include_affine_speed && !mutated(kMutCflNoAffine)) s = s + affine_cfl(P.c[i]);
            max_speed = rmax(max_speed, s);
            T mdt = mat(P.material_id[i])
                        .timestep_bound(P.density[i], P.hardening_scale[i], T(cfg.grid_cell_size),
                                        T(cfg.material_cfl_coefficient), T(cfg.viscous_timestep_coefficient));
            if (finite(mdt) && mdt > T(0)) min_mat_dt = rmin(min_mat_dt, mdt);
        }
        T dt = max_dt;
        if (max_speed > T(F32_EPSILON)) dt = rmin(dt, T(cfg.cfl_coefficient) * T(cfg.grid_cell_size) / max_speed);
        dt = rmin(dt, min_mat_dt);
        if (mutated(kMutCflNoMinClamp)) return rmin(dt, max_dt);
        return rclamp(dt, rmin(T(cfg.min_dt), max_dt), max_dt);
    }

    // project_particle_state_to_admissible  step.rs:347-429
    bool project(size_t i) {
        bool projected = false;
        uint32_t bt = cfg.boundary_thickness, R = cfg.grid_res;
        T lo = T((float)(bt > 0 ? bt - 1 : 0));
        T hi = T((float)(R > bt ? R - bt : 0));
        T center = (lo + hi) * T(0.5);
        if (!finite(P.x[i])) {
            P.x[i] = {center, center};
            projected = true;
        } else {
            P.x[i] = {rmin(rmax(P.x[i].x, lo), hi), rmin(rmax(P.x[i].y, lo), hi)};
        }
        if (!finite(P.v[i])) {
            P.v[i] = {T(0), T(0)};
            projected = true;
        }
        if (!finite_cols(P.c[i])) {
            P.c[i] = m2_zero<T>();
            projected = true;
        }
        M2<T> f = P.f[i];
        if (!finite_cols(f) || det(f) <= T(cfg.projection_min_deformation_j)) {
            P.f[i] = m2_identity<T>();
            projected = true;
        } else {
            T j = det(f);
            if (j > T(cfg.j_max)) {
                P.f[i] = P.f[i] * (mutated(kMutProjJScaleLinear) ? T(cfg.j_max) / j : o_sqrt(T(cfg.j_max) / j));
                projected = true;
            }
        }
        if (!finite(P.plastic_volume_ratio[i]) || P.plastic_volume_ratio[i] <= T(0)) {
            P.plastic_volume_ratio[i] = T(1);
            projected = true;
        }
        if (!finite(P.hardening_scale[i]) || P.hardening_scale[i] <= T(0)) {
            P.hardening_scale[i] = T(1);
            projected = true;
        }
        if (!finite(P.friction_hardening[i])) {
            P.friction_hardening[i] = T(0);
            projected = true;
        }
        if (!finite(P.log_volume_strain[i])) {
            P.log_volume_strain[i] = T(0);
            projected = true;
        }
        if (!finite(P.mass[i]) || P.mass[i] <= T(0)) {
            P.mass[i] = T(cfg.particle_mass);
            projected = true;
        }
        if (!finite(P.initial_volume[i]) || P.initial_volume[i] <= T(0)) {
            P.initial_volume[i] = rmax(T(cfg.default_initial_volume), T(cfg.projection_min_volume));
            projected = true;
        }
        if (!finite(P.volume[i]) || P.volume[i] <= T(0)) {
            P.volume[i] = rmax(P.initial_volume[i], T(cfg.projection_min_volume));
            projected = true;
        }
        if (!finite(P.density[i]) || P.density[i] <= T(0)) {
            P.density[i] = rmax(P.mass[i] / P.volume[i], T(cfg.projection_min_density));
            projected = true;
        } else if (!mutated(kMutProjNoDensityFloor)) {
            P.density[i] = rmax(P.density[i], T(cfg.projection_min_density));
        }
        return projected;
    }

    // combined_kirchhoff_stress  transfer.rs:77-104
    M2<T> combined_stress(const Material<T>& m, size_t i) const {
        M2<T> tau = m.kirchhoff_stress(P, i);
        T coeff = m.activation_scale();
        if (P.activation[i] <= T(0) || coeff <= T(0)) return tau;
        T a = P.activation[i] * coeff;
        V2<T> n = P.activation_dir[i];
        T len_sq = dot(n, n);
        M2<T> tau_active;
        if (m.model != EMH_MODEL_VISCOELASTIC && len_sq > T(F32_EPSILON)) {  // Viscoelastic: isotropic (transfer.rs:87-89)
            V2<T> n0 = n / o_sqrt(len_sq);
            M2<T> n_outer = m2_cols(n0 * n0.x, n0 * n0.y);
            M2<T> a_mat = n_outer * a;
            M2<T> f = P.f[i];
            tau_active = (f * a_mat) * transpose(f);
        } else {
            tau_active = m2_diag(a, a);
        }
        return tau + tau_active;
    }

    // scatter_particles_to_grid  transfer/p2g.rs:24-72
    void p2g(T dt) {
        for (size_t i = 0; i < active_count; ++i) {
            const Material<T>& m = mat(P.material_id[i]);
            V2<T> x = P.x[i];
            T mass_i = P.mass[i];
            V2<T> v_i = P.v[i];
            M2<T> c_i = P.c[i];
            M2<T> stress = combined_stress(m, i);
            T stress_coeff = ((-m.stress_volume(P, i)) * T(KERNEL_D_INVERSE)) * dt;
            Weights<T> w = quadratic_weights(x);
            for (int gx = 0; gx < 3; ++gx)
                for (int gy = 0; gy < 3; ++gy) {
                    T weight = w.wx[gx] * w.wy[gy];
                    int32_t cx = w.bx + gx - 1, cy = w.by + gy - 1;
                    V2<T> dist{(T(cx) - x.x) + T(0.5), (T(cy) - x.y) + T(0.5)};
                    V2<T> momentum = weight * (mass_i * (v_i + c_i * dist) + stress_coeff * (stress * dist));
                    grid.add_mass_momentum(cx, cy, weight * mass_i, momentum);
                }
        }
    }

    // Grid::update_velocities + boundaries + clamp  grid/mod.rs:264-291, step.rs:146-165,334-344
    void grid_update(T dt) {
        V2<T> g{T(cfg.gravity[0]), T(cfg.gravity[1])};
        V2<T> gdt = g * dt;
        T vel_limit = T(cfg.grid_cell_size) / dt;
        uint32_t R = cfg.grid_res;
        for (uint32_t s : grid.dirty_slots) {
            Cell<T>& c = grid.cells[s];
            if (c.mass > T(0)) c.momentum = c.momentum / c.mass;
        }
        for (uint32_t s : grid.dirty_slots) {
            Cell<T>& c = grid.cells[s];
            if (c.mass > T(0)) c.momentum = c.momentum + gdt;
        }
        for (const Boundary& b : boundaries) {  // apply_boundary_conditions_to_grid  step.rs:334-344
            for (uint32_t s : grid.dirty_slots) {
                Cell<T>& c = grid.cells[s];
                if (!(c.mass > T(0))) continue;
                boundary_velocity(b, grid.keys[s], R, c.momentum);
            }
        }
        for (uint32_t s : grid.dirty_slots) {
            Cell<T>& c = grid.cells[s];
            if (c.mass > T(0)) {
                T spd = length(c.momentum);
                if (spd > vel_limit) c.momentum = c.momentum * (vel_limit / spd);
            }
        }
    }

    // gather_grid_to_particles  transfer/g2p.rs:151-343
    uint32_t g2p(T dt) {
        T vel_limit = T(cfg.grid_cell_size) / dt;
        T apic = T(cfg.apic_blend);
        uint32_t R = cfg.grid_res;
        uint32_t clamped = 0;
        const long n_active = (long)active_count;
#pragma omp parallel for schedule(static) reduction(+ : clamped) num_threads(g_threads) if (g_threads > 1)
        for (long i = 0; i < n_active; ++i) {  // phase 1
            V2<T> x = P.x[i];
            Weights<T> w = quadratic_weights(x);
            V2<T> new_v{T(0), T(0)};
            M2<T> b = m2_zero<T>();
            for (int gx = 0; gx < 3; ++gx)
                for (int gy = 0; gy < 3; ++gy) {
                    T weight = w.wx[gx] * w.wy[gy];
                    int32_t cx = w.bx + gx - 1, cy = w.by + gy - 1;
                    V2<T> dist{(T(cx) - x.x) + T(0.5), (T(cy) - x.y) + T(0.5)};
                    V2<T> wv = grid.velocity_at(cx, cy) * weight;
                    b = b + m2_cols(wv * dist.x, wv * dist.y);
                    new_v = new_v + wv;
                }
            if (P.pinned[i] != 0) {
                P.v[i] = {T(0), T(0)};
                P.c[i] = m2_zero<T>();
                continue;
            }
            T spd = length(new_v);
            if (spd > vel_limit) {
                T scale = vel_limit / spd;
                new_v = new_v * scale;
                clamped++;
            }
            V2<T> np = x + new_v * dt;
            for (const Boundary& bd : boundaries) np = boundary_clamp(bd, np, R);  // g2p.rs:316-320
            P.v[i] = new_v;
            P.c[i] = (b * T(KERNEL_D_INVERSE)) * apic;
            P.x[i] = np;
        }
        for (size_t i = 0; i < active_count; ++i) mat(P.material_id[i]).update_particle(P, i, dt);  // phase 2
        return clamped;
    }

    // do_substep  step.rs:62-317 (the parts that run for slip-only scenes)
    void do_substep(T sub_dt) {
        if (cfg.project_invalid_state) {
            for (size_t i = 0; i < active_count; ++i)
                if (project(i)) last_j_projection_count++;
        }
        if (cfg.recompute_density_each_step || any_needs_density_recompute())
            estimate_particle_volumes(active_count, false);
        grid.clear();
        p2g(sub_dt);
        // wake every sleeping particle whose stencil touches a cell this P2G reached
        // (step.rs:107-131), in index order
        if (active_count < P.size()) {
            std::vector<size_t> wake;
            for (size_t i = active_count; i < P.size(); ++i) {
                Weights<T> w = quadratic_weights(P.x[i]);
                bool hit = false;
                for (int gx = 0; gx < 3 && !hit; ++gx)
                    for (int gy = 0; gy < 3 && !hit; ++gy) hit = grid.cell_is_active(w.bx + gx - 1, w.by + gy - 1);
                if (hit) wake.push_back(i);
            }
            for (size_t i : wake) wake_particle(i);
        }
        grid_update(sub_dt);
        last_vel_clamp_count += g2p(sub_dt);
        // sleep scoring (step.rs:299-315): slow particles without activation,
        // put to sleep in descending index order
        const T threshold = T(cfg.sleep_threshold);
        if (threshold > T(0)) {
            const T threshold_sq = threshold * threshold;
            std::vector<size_t> slow;
            for (size_t i = 0; i < active_count; ++i)
                if (P.activation[i] == T(0) && dot(P.v[i], P.v[i]) < threshold_sq) slow.push_back(i);
            for (size_t k = slow.size(); k-- > 0;) sleep_particle(slow[k]);
        }
    }

    // Simulation::step  step.rs:23-60
    void step() {
        T remaining = T(cfg.dt);
        uint32_t substeps = 0;
        last_vel_clamp_count = 0;
        last_j_projection_count = 0;
        while (remaining > T(F32_EPSILON) && substeps < cfg.max_substeps_per_step) {
            T sub_dt = choose_substep_dt(remaining);
            do_substep(sub_dt);
            remaining = remaining - sub_dt;
            last_step_dt = sub_dt;
            substeps++;
            substep_dts.push_back(sub_dt);
        }
        last_substeps = substeps;
        total_substeps += substeps;
        last_sim_time_dropped = rmax(remaining, T(0));
        frame_index++;
    }
};

}  // namespace orc
