// device_math.hpp -- per-particle MLS-MPM math for the gfx950 kernels.
//
// Every formula follows the reference crate's CPU path operation by operation
// (file:line cited per function) so that, for the same inputs, a kernel lane
// computes the same f32 value the reference computes: the library is built
// with -ffp-contract=off (no a*b+c fusion, as Rust does not fuse) and with
// correctly rounded f32 division and square root.  Transcendentals
// (logf/expf/sinf/sinhf/powf) are the device library's and may differ from
// glibc's by an ulp; that is the only source of per-particle difference.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace emh {

#define EMH_DEV __device__ __forceinline__

constexpr float kF32Eps = 1.1920929e-07f;  // f32::EPSILON
constexpr float kLogClamp = 1e-10f;        // materials/utils.rs:6
constexpr float kMinJ = 1e-6f;             // materials/utils.rs:11
constexpr float kDInv = 4.0f;              // KERNEL_D_INVERSE, solver/config.rs:22
constexpr float kInf = __builtin_huge_valf();

// Packed flags in the per-particle material word.
constexpr uint32_t kPinnedBit = 0x80000000u;
// sleep partition (solver/mod.rs:40-42): the particle sleeps (skipped by P2G,
// G2P, the CFL bound and projection); it has non-zero activation and never
// falls asleep (step.rs:299-306)
constexpr uint32_t kSleepBit = 0x40000000u;
constexpr uint32_t kNoSleepBit = 0x20000000u;
constexpr uint32_t kMatMask = 0x0000ffffu;

// ------------------------------------------------------------------ config
// Device copy of the SimConfig fields the substep reads, plus derived values.
struct DevConfig {
    uint32_t grid_res;
    float cell;  // grid_cell_size
    float dt;    // config.dt (the step length)
    uint32_t adaptive;
    uint32_t cfl_affine;
    float cfl_coefficient;
    float material_cfl;
    float viscous_cfl;
    float min_dt;
    uint32_t project;
    float proj_min_density;
    float proj_min_volume;
    float proj_min_j;
    float gx, gy;
    uint32_t bt;  // boundary_thickness
    float default_initial_volume;
    float particle_mass;
    uint32_t max_substeps;
    float apic_blend;
    float j_max;
    // installed boundaries, applied in order (kBnd* kinds)
    uint32_t n_bnd;
    uint32_t bnd_kind[8];
    uint32_t bnd_t[8];            // Slip / Friction / Predictive thickness, Heightmap wall thickness
    float bnd_mu[8];              // Friction coefficient, Heightmap terrain friction
    uint32_t bnd_hoff[8], bnd_hlen[8];  // Heightmap: its heights in `heights`
    const float* heights;         // device: terrain heights of every heightmap boundary
    // derived
    float proj_lo, proj_hi;  // particle clamp range of project_particle_state_to_admissible
    float bnd_lo[8], bnd_hi[8];   // each boundary's particle clamp box
    uint32_t key_order;       // cell keys: Hilbert curve over a 2^key_order square (cell_key)
    uint32_t n_bins;          // 4^key_order
    uint32_t density_pass;    // any material needs_density_recompute || config flag
    uint32_t fields;          // kF* bits: optional particle fields the particles' materials use
    uint32_t only_model;      // the model every particle's material has, or kModelMixed
};

constexpr uint32_t kBndSlip = 0, kBndFriction = 1, kBndHeightmap = 2, kBndPredictive = 3;

// Optional particle fields (DevConfig::fields).  A field no registered
// material reads or writes stays untouched in memory between reorders.
constexpr uint32_t kFQ = 1u;        // friction_hardening (NeoHookean damage, Drucker-Prager)
constexpr uint32_t kFLvs = 2u;      // log_volume_strain (Drucker-Prager, NACC alpha)
constexpr uint32_t kFJp = 4u;       // plastic_volume_ratio (snow)
constexpr uint32_t kFH = 8u;        // hardening_scale (corotated, snow, NACC)
constexpr uint32_t kFVolDens = 16u; // volume/density read back (fluid, NACC, fallback)
constexpr uint32_t kFAll = 31u;

// ---------------------------------------------------------------- materials
// Device material record, filled on the host from MaterialParams + extension.
// Constants that the reference recomputes from material fields only (for
// example sin(dilatancy_angle)) are evaluated once on the host with the same
// f32 arithmetic / libm the reference uses.
struct DevMat {
    uint32_t model;
    uint32_t flags;  // kMatFlag* bits
    float lambda, mu;
    float min_density, min_volume;
    // NeoHookean
    float damage_rate, nh_viscosity;
    // Fluid
    float rest_density, dyn_visc, eos_k, eos_p, p_floor, bulk_visc, surf_tension, settling;
    float fl_c2_denom;  // rest_density.max(min_density)
    // Drucker-Prager
    float dp_phi0, dp_h1, dp_h2, dp_h3, dp_vol_corr, dp_dilat, dp_sin_dilat, dp_cohesion_term, dp_ratio,
        dp_min_vj, dp_qmax, dp_sqrt23;
    // Snow
    float sn_xi, sn_lo, sn_hi, sn_jmin, sn_jmax, sn_coh;
    // NACC
    float na_kappa, na_m, na_beta, na_xi;
    // Von Mises
    float vm_yield, vm_hmod;
    // Viscoelastic
    float ve_visc, ve_jmin;
    // Rankine: tensile strength, softening rate, 5 % residual strength, damage cap
    float rk_ts, rk_sr, rk_floor, rk_sat;
    // DruckerPragerMuI: mu_static, mu_dynamic, inertial Q
    float mui_s, mui_d, mui_q;
    // Bingham (a Fluid record): yield stress, critical shear rate
    float bing_yield, bing_crit;
    // GranularFluid: rest_density.max(1e-6) (EOS on the Fluid fields, clamp on the Snow fields)
    float gf_rest_floor;
    // CFL modulus helpers
    float wave_lambda;  // lambda used by elastic_wave_dt (NACC: kappa - mu)
    float pad[3];
};

constexpr uint32_t kModelFallback = 0, kModelFluid = 1, kModelNH = 2, kModelCorot = 3, kModelSnow = 4,
                   kModelDP = 5, kModelVonMises = 6, kModelRankine = 7, kModelMuI = 8, kModelViscoelastic = 9,
                   kModelNacc = 10, kModelGranular = 11, kModelMixed = 0xffffffffu;
// The optional particle fields (kF*) a particle of `model` reads or writes;
// DevConfig::fields is their union over the scene's models.
__host__ __device__ constexpr uint32_t model_fields(uint32_t model) {
    return (model == kModelNH || model == kModelVonMises || model == kModelRankine || model == kModelMuI) ? kFQ
           : model == kModelDP                                                                      ? (kFQ | kFLvs)
           : model == kModelGranular                                                     ? (kFJp | kFH | kFVolDens)
           : model == kModelViscoelastic                                                                   ? 0u
           : model == kModelCorot                                                                          ? kFH
           : model == kModelSnow                                                                  ? (kFJp | kFH)
           : model == kModelNacc                                                       ? (kFLvs | kFH | kFVolDens)
                                                                                                    : kFVolDens;
}
constexpr uint32_t kMatNaccHardening = 1u;  // NACC: hardening enabled
constexpr uint32_t kMatNaccAdvance = 2u;    // NACC: project (I + dt C) F
constexpr uint32_t kMatBingham = 4u;        // Fluid: Bingham viscoplastic deviatoric law
constexpr uint32_t kMatGfPlastic = 8u;      // GranularFluid: compression or stretch limit > 0

// ------------------------------------------------------------ small algebra
struct M2 {
    float a, b, c, d;  // column-major: (a, b) first column, (c, d) second
};
EMH_DEV M2 m2(float a, float b, float c, float d) { return M2{a, b, c, d}; }
EMH_DEV M2 mmul(const M2& p, const M2& q) {
    return M2{p.a * q.a + p.c * q.b, p.b * q.a + p.d * q.b, p.a * q.c + p.c * q.d, p.b * q.c + p.d * q.d};
}
EMH_DEV float2 mvec(const M2& p, float x, float y) { return make_float2(p.a * x + p.c * y, p.b * x + p.d * y); }
EMH_DEV float mdet(const M2& p) { return p.a * p.d - p.b * p.c; }
EMH_DEV M2 mtr(const M2& p) { return M2{p.a, p.c, p.b, p.d}; }
EMH_DEV M2 mscale(const M2& p, float s) { return M2{p.a * s, p.b * s, p.c * s, p.d * s}; }
EMH_DEV M2 madd(const M2& p, const M2& q) { return M2{p.a + q.a, p.b + q.b, p.c + q.c, p.d + q.d}; }
EMH_DEV M2 msub(const M2& p, const M2& q) { return M2{p.a - q.a, p.b - q.b, p.c - q.c, p.d - q.d}; }
EMH_DEV bool mfinite(const M2& p) {
    return __builtin_isfinite(p.a) && __builtin_isfinite(p.b) && __builtin_isfinite(p.c) && __builtin_isfinite(p.d);
}
EMH_DEV M2 from4(float4 v) { return M2{v.x, v.y, v.z, v.w}; }
EMH_DEV float4 to4(const M2& p) { return make_float4(p.a, p.b, p.c, p.d); }
// Rust f32::max/min ignore a NaN operand, as fmaxf/fminf do.
EMH_DEV float rmax(float a, float b) { return fmaxf(a, b); }
EMH_DEV float rmin(float a, float b) { return fminf(a, b); }
EMH_DEV float rclamp(float x, float lo, float hi) {  // Rust f32::clamp (NaN passes through)
    if (x < lo) x = lo;
    if (x > hi) x = hi;
    return x;
}
EMH_DEV float len2(float x, float y) { return sqrtf(x * x + y * y); }

// ------------------------------------------------------------ B-spline weights
// grid/kernel.rs:10-25.  Base cell = floor(x); d = (x - base) - 0.5.
struct Wts {
    int bx, by;
    float wx[3], wy[3];
};
EMH_DEV int sat_i32(float f) {  // Rust `as i32`: saturating, NaN -> 0
    if (__builtin_isnan(f)) return 0;
    if (f >= 2147483647.0f) return 2147483647;
    if (f <= -2147483648.0f) return (-2147483647 - 1);
    return (int)f;
}
EMH_DEV void axis_w(float d, float w[3]) {
    float a = 0.5f - d, b = 0.5f + d;
    w[0] = 0.5f * (a * a);
    w[1] = 0.75f - d * d;
    w[2] = 0.5f * (b * b);
}
EMH_DEV Wts weights(float x, float y) {
    Wts w;
    w.bx = sat_i32(floorf(x));
    w.by = sat_i32(floorf(y));
    axis_w((x - (float)w.bx) - 0.5f, w.wx);
    axis_w((y - (float)w.by) - 0.5f, w.wy);
    return w;
}

// ---------------------------------------------------------------- svd2
// materials/svd.rs:14-85 (signed second singular value, det U = det V = +1).
EMH_DEV float2 orth_complement(float x, float y) {
    float px = -y, py = x;
    float l = len2(px, py);
    if (l > 1e-10f) return make_float2(px / l, py / l);
    return make_float2(1.0f, 0.0f);
}
EMH_DEV void svd2(const M2& f, M2& u, float2& sig, M2& vt) {
    float c00 = f.a * f.a + f.b * f.b;
    float c01 = f.a * f.c + f.b * f.d;
    float c11 = f.c * f.c + f.d * f.d;
    float mean = (c00 + c11) * 0.5f;
    float hd = (c00 - c11) * 0.5f;
    float disc = sqrtf(hd * hd + c01 * c01);
    float l1 = mean + disc, l2 = mean - disc;
    float v1x, v1y;
    if (fabsf(c01) > 1e-10f * rmax(c00 + c11, 1e-30f)) {
        float rx = c01, ry = l1 - c00;
        float rl = len2(rx, ry);
        v1x = rx / rl;
        v1y = ry / rl;
    } else if (c00 >= c11) {
        v1x = 1.0f;
        v1y = 0.0f;
    } else {
        v1x = 0.0f;
        v1y = 1.0f;
    }
    float v2x = -v1y, v2y = v1x;
    vt = M2{v1x, v2x, v1y, v2y};  // transpose of [v1 | v2]
    float s1 = sqrtf(rmax(l1, 0.0f)), s2 = sqrtf(rmax(l2, 0.0f));
    float2 fv1 = mvec(f, v1x, v1y), fv2 = mvec(f, v2x, v2y);
    float2 u1 = s1 > 1e-10f ? make_float2(fv1.x / s1, fv1.y / s1) : orth_complement(fv2.x, fv2.y);
    float2 u2 = s2 > 1e-10f ? make_float2(fv2.x / s2, fv2.y / s2) : orth_complement(u1.x, u1.y);
    u = M2{u1.x, u1.y, u2.x, u2.y};
    if (mdet(u) < 0.0f) {
        u.c = -u.c;
        u.d = -u.d;
        s2 = -s2;
    }
    sig = make_float2(s1, s2);
}
// U diag(s) V^T  (utils.rs:67-69, nacc.rs:227-229)
EMH_DEV M2 usv(const M2& u, float s1, float s2, const M2& vt) { return mmul(mmul(u, M2{s1, 0.0f, 0.0f, s2}), vt); }

// materials/utils.rs:129-138
EMH_DEV M2 polar_r(const M2& f) {
    float x = f.a + f.d;
    float y = f.b - f.c;
    float n = sqrtf(x * x + y * y);
    if (n > kF32Eps) return M2{x / n, y / n, -y / n, x / n};
    return M2{1.0f, 0.0f, 0.0f, 1.0f};
}

// materials/utils.rs:145-164
EMH_DEV float elastic_wave_dt(float lambda, float mu, float h, float density, float min_density, float cell,
                              float mcfl) {
    float rho = rmax(density, min_density);
    float modulus = rmax((lambda + 2.0f * mu) * h, 0.0f);
    if (modulus <= kF32Eps) return kInf;
    float c = sqrtf(modulus / rho);
    if (c <= kF32Eps) return kInf;
    return mcfl * cell / c;
}

// --------------------------------------------------------- particle state
struct PState {
    float x, y, vx, vy;
    M2 c, f;
    float mass, v0, vol, dens, jp, h, q, lvs;
    uint32_t mat;  // material id | pinned bit
};

// project_particle_state_to_admissible  solver/step.rs:347-429
EMH_DEV bool project_state(PState& s, const DevConfig& cfg) {
    bool p = false;
    if (!(__builtin_isfinite(s.x) && __builtin_isfinite(s.y))) {
        float c = (cfg.proj_lo + cfg.proj_hi) * 0.5f;
        s.x = c;
        s.y = c;
        p = true;
    } else {
        s.x = rmin(rmax(s.x, cfg.proj_lo), cfg.proj_hi);
        s.y = rmin(rmax(s.y, cfg.proj_lo), cfg.proj_hi);
    }
    if (!(__builtin_isfinite(s.vx) && __builtin_isfinite(s.vy))) {
        s.vx = 0.0f;
        s.vy = 0.0f;
        p = true;
    }
    if (!mfinite(s.c)) {
        s.c = M2{0.0f, 0.0f, 0.0f, 0.0f};
        p = true;
    }
    if (!mfinite(s.f) || mdet(s.f) <= cfg.proj_min_j) {
        s.f = M2{1.0f, 0.0f, 0.0f, 1.0f};
        p = true;
    } else {
        float j = mdet(s.f);
        if (j > cfg.j_max) {
            s.f = mscale(s.f, sqrtf(cfg.j_max / j));
            p = true;
        }
    }
    if (!__builtin_isfinite(s.jp) || s.jp <= 0.0f) {
        s.jp = 1.0f;
        p = true;
    }
    if (!__builtin_isfinite(s.h) || s.h <= 0.0f) {
        s.h = 1.0f;
        p = true;
    }
    if (!__builtin_isfinite(s.q)) {
        s.q = 0.0f;
        p = true;
    }
    if (!__builtin_isfinite(s.lvs)) {
        s.lvs = 0.0f;
        p = true;
    }
    if (!__builtin_isfinite(s.mass) || s.mass <= 0.0f) {
        s.mass = cfg.particle_mass;
        p = true;
    }
    if (!__builtin_isfinite(s.v0) || s.v0 <= 0.0f) {
        s.v0 = rmax(cfg.default_initial_volume, cfg.proj_min_volume);
        p = true;
    }
    if (!__builtin_isfinite(s.vol) || s.vol <= 0.0f) {
        s.vol = rmax(s.v0, cfg.proj_min_volume);
        p = true;
    }
    if (!__builtin_isfinite(s.dens) || s.dens <= 0.0f) {
        s.dens = rmax(s.mass / s.vol, cfg.proj_min_density);
        p = true;
    } else {
        s.dens = rmax(s.dens, cfg.proj_min_density);
    }
    return p;
}

// ------------------------------------------------------- constitutive models
// Template argument ONLY of the per-model functions here and below: -1 dispatches on
// the material's model id at run time; a model id compiles that model alone
// (used when every particle resolves to materials of one model).

// stress_volume: materials/mod.rs:118-120, fluid.rs:193-195, bingham.rs:180-182,
// granular_fluid.rs:196-198 (the other models restate the default)
template <int ONLY = -1>
EMH_DEV float stress_volume(const DevMat& m, const PState& s) {
    const uint32_t model = ONLY >= 0 ? (uint32_t)ONLY : m.model;
    if (model == kModelFluid) return rmax(s.vol, m.min_volume);
    if (model == kModelGranular) return rmax(s.vol, 1.0e-6f);  // granular_fluid.rs:196-198
    return s.v0;
}

EMH_DEV M2 corot_tau(const M2& f, float j, float mu, float lambda) {
    M2 r = polar_r(f);
    M2 t1 = mmul(mscale(msub(f, r), 2.0f * mu), mtr(f));
    float iso = (lambda * (j - 1.0f)) * j;
    return madd(t1, M2{iso, 0.0f, 0.0f, iso});
}

// Kirchhoff stress: elastic.rs:180-255, corotated.rs:62-78, sand.rs:220-231,
// von_mises.rs:94-102, rankine.rs:157-166, sand_mui.rs:132-140, viscoelastic.rs:103-130,
// snow.rs:109-136, fluid.rs:135-191, bingham.rs:96-178, granular_fluid.rs:171-194,
// nacc.rs:236-249, Fallback mod.rs:113-115.
template <int ONLY = -1>
EMH_DEV M2 kirchhoff(const DevMat& m, const PState& s) {
    const M2 zero{0.0f, 0.0f, 0.0f, 0.0f};
    const uint32_t model = ONLY >= 0 ? (uint32_t)ONLY : m.model;
    switch (model) {
        case kModelNH: {
            float j = mdet(s.f);
            if (j <= kMinJ) return zero;
            float dscale = expf(-m.damage_rate * s.q);
            float mu = m.mu * 1.0f * dscale;
            float lambda = m.lambda * 1.0f * dscale;
            M2 b = mmul(s.f, mtr(s.f));
            float trb = b.a + b.d;
            float hb = trb * 0.5f;
            M2 devb = M2{b.a - hb, b.b - 0.0f, b.c - 0.0f, b.d - hb};
            float k = lambda + mu;
            M2 dev = mscale(devb, mu / j);
            float vol = k * logf(j);
            M2 t = madd(dev, M2{vol, 0.0f * vol, 0.0f * vol, vol});
            if (m.nh_viscosity > 0.0f) {
                M2 sym = madd(s.c, mtr(s.c));
                M2 d = mscale(sym, 0.5f);
                float hdt = (d.a + d.d) * 0.5f;
                M2 ddev = M2{d.a - hdt, d.b - 0.0f, d.c - 0.0f, d.d - hdt};
                t = madd(t, mscale(ddev, m.nh_viscosity));
            } else {
                t = madd(t, zero);
            }
            return t;
        }
        case kModelCorot: {
            float j = mdet(s.f);
            if (j <= kMinJ) return zero;
            return corot_tau(s.f, j, m.mu * s.h * 1.0f, m.lambda * s.h * 1.0f);
        }
        case kModelDP:
        case kModelVonMises:
        case kModelRankine:
        case kModelMuI: {
            float j = mdet(s.f);
            if (j <= kMinJ) return zero;
            return corot_tau(s.f, j, m.mu, m.lambda);
        }
        case kModelGranular: {
            // Tait pressure on rho0 / J, floored; deviatoric part of the
            // corotated term scaled by h, plus its lambda h (J - 1) J I
            float j = rmax(mdet(s.f), kMinJ);
            float dens = rmax(m.rest_density / j, 1.0e-6f);
            float ratio = rmax(dens / m.gf_rest_floor, 1.0e-6f);
            float p = rmax(m.eos_k * (powf(ratio, m.eos_p) - 1.0f), m.p_floor);
            float h = s.h;
            M2 r = polar_r(s.f);
            float mu_eff = m.mu * h;
            M2 coro = mmul(mscale(msub(s.f, r), 2.0f * mu_eff), mtr(s.f));
            float ht = (coro.a + coro.d) * 0.5f;
            M2 t{-p + (coro.a - ht), 0.0f + (coro.b - 0.0f), 0.0f + (coro.c - 0.0f), -p + (coro.d - ht)};
            float lv = ((m.lambda * h) * (j - 1.0f)) * j;
            return madd(t, M2{lv, lv * 0.0f, lv * 0.0f, lv});
        }
        case kModelViscoelastic: {
            // neo-Hookean spring mu (F F^T - I) + lambda ln(J) I with J clamped to
            // [j_min, 1/j_min], in parallel with the dashpot visc dev(D); the
            // thermal scale is 1 (thermal_expansion is not part of MaterialParams)
            float j = rmin(rmax(s.f.a * s.f.d - s.f.c * s.f.b, m.ve_jmin), 1.0f / m.ve_jmin);
            M2 b = mmul(s.f, mtr(s.f));
            M2 el = mscale(M2{b.a - 1.0f, b.b - 0.0f, b.c - 0.0f, b.d - 1.0f}, m.mu);
            float lj = m.lambda * logf(j);
            el = madd(el, M2{lj, 0.0f, 0.0f, lj});
            M2 d = mscale(madd(s.c, mtr(s.c)), 0.5f);
            float ht = (d.a + d.d) * 0.5f;
            M2 ddev = M2{d.a - ht, d.b - 0.0f, d.c - 0.0f, d.d - ht};
            return madd(el, mscale(ddev, m.ve_visc));
        }
        case kModelSnow: {
            float j = mdet(s.f);
            if (j <= kMinJ) return zero;
            M2 t = corot_tau(s.f, j, m.mu * s.h, m.lambda * s.h);
            if (m.sn_coh > 0.0f && s.jp < 1.0f && j > 1.0f) {
                float iso = ((m.sn_coh * s.jp) * (j - 1.0f)) * j;
                t = madd(t, M2{iso, 0.0f, 0.0f, iso});
            }
            return t;
        }
        case kModelFluid: {
            float dens = rmin(rmax(s.dens, m.min_density), m.rest_density * 2.0f);
            float p = rmax(m.eos_k * (powf(dens / m.rest_density, m.eos_p) - 1.0f), m.p_floor);
            M2 t{-p, 0.0f, 0.0f, -p};
            if (m.flags & kMatBingham) {
                // apparent viscosity yield / rate + eta on D_dev above the
                // critical rate of sqrt(2 D_dev : D_dev), rigid plug below
                M2 d = mscale(madd(s.c, mtr(s.c)), 0.5f);
                float ht = (d.a + d.d) * 0.5f;
                M2 dd = M2{d.a - ht, d.b - 0.0f, d.c - 0.0f, d.d - ht};
                float dsq = (dd.a * dd.a + dd.d * dd.d) + 2.0f * dd.b * dd.b;
                float rate = sqrtf(2.0f * dsq);
                M2 dev = zero;
                if (!(rate < m.bing_crit)) dev = mscale(dd, m.bing_yield / rate + m.dyn_visc);
                t = madd(t, dev);
                if (m.surf_tension != 0.0f) {
                    float j = s.f.a * s.f.d - s.f.b * s.f.c;
                    float st = m.surf_tension * j;
                    t = madd(t, M2{st, 0.0f, 0.0f, st});
                } else {
                    t = madd(t, zero);
                }
                return t;
            }
            M2 sym = madd(s.c, mtr(s.c));
            float div = sym.a + sym.d;
            float hd = div * 0.5f;
            M2 sdev = M2{sym.a - hd, sym.b - 0.0f, sym.c - 0.0f, sym.d - hd};
            t = madd(t, mscale(sdev, m.dyn_visc));
            if (m.bulk_visc > 0.0f) {
                float bv = m.bulk_visc * div * 0.5f;
                t = madd(t, M2{bv, 0.0f, 0.0f, bv});
            }
            if (m.surf_tension != 0.0f) {
                float j = s.f.a * s.f.d - s.f.b * s.f.c;
                float st = m.surf_tension * j;
                t = madd(t, M2{st, 0.0f, 0.0f, st});
            }
            return t;
        }
        case kModelNacc: {
            float j = rmax(mdet(s.f), kMinJ);
            M2 b = mmul(s.f, mtr(s.f));
            float hb = (b.a + b.d) * 0.5f;
            M2 devb = M2{b.a - hb, b.b - 0.0f, b.c - 0.0f, b.d - hb};
            M2 dev = mscale(devb, m.mu / j);
            float vol = (m.na_kappa * 0.5f) * (j * j - 1.0f);
            return madd(dev, M2{vol, 0.0f, 0.0f, vol});
        }
        default:
            return zero;
    }
}

// timestep_bound: elastic.rs:290-327, corotated.rs:108-125, fluid.rs:228-259,
// sand.rs:345-362, von_mises.rs:154-178, rankine.rs:218-229, sand_mui.rs:232-249,
// viscoelastic.rs:161-190, granular_fluid.rs:250-268, snow.rs:193-211, bingham.rs:220-244
// (the Fluid form), nacc.rs:279-298 (Fallback: +inf).
template <int ONLY = -1>
EMH_DEV float timestep_bound(const DevMat& m, float density, float h, const DevConfig& cfg) {
    const uint32_t model = ONLY >= 0 ? (uint32_t)ONLY : m.model;
    switch (model) {
        case kModelNH: {
            float e = elastic_wave_dt(m.lambda, m.mu, 1.0f, density, m.min_density, cfg.cell, cfg.material_cfl);
            float v = kInf;
            if (m.nh_viscosity > 0.0f) {
                float d = rmax(density, 1.0e-6f);
                float kin = m.nh_viscosity / d;
                if (kin > kF32Eps) v = cfg.viscous_cfl * cfg.cell * cfg.cell / kin;
            }
            return rmin(e, v);
        }
        case kModelCorot:
        case kModelSnow:
            return elastic_wave_dt(m.lambda, m.mu, h, density, kMinJ, cfg.cell, cfg.material_cfl);
        case kModelDP:
        case kModelVonMises:
        case kModelRankine:
        case kModelMuI:
            return elastic_wave_dt(m.lambda, m.mu, 1.0f, density, kMinJ, cfg.cell, cfg.material_cfl);
        case kModelGranular:
            return elastic_wave_dt(m.lambda, m.mu, h, density, kMinJ, cfg.cell, cfg.material_cfl);
        case kModelViscoelastic: {
            float e = elastic_wave_dt(m.lambda, m.mu, 1.0f, density, 1.0e-6f, cfg.cell, cfg.material_cfl);
            float v = kInf;
            if (m.ve_visc > 0.0f) {
                float d = rmax(density, 1.0e-6f);
                float kin = m.ve_visc / d;
                if (kin > kF32Eps) v = cfg.viscous_cfl * cfg.cell * cfg.cell / kin;
            }
            return rmin(e, v);
        }
        case kModelNacc:
            return elastic_wave_dt(m.wave_lambda, m.mu, h, density, m.min_density, cfg.cell, cfg.material_cfl);
        case kModelFluid: {
            float d = rmax(density, m.min_density);
            float ratio = rmax(d / m.fl_c2_denom, 1.0e-6f);
            float out = kInf;
            float c2 = m.eos_k * m.eos_p * powf(ratio, m.eos_p - 1.0f) / m.fl_c2_denom;
            if (__builtin_isfinite(c2) && c2 > kF32Eps) out = rmin(out, cfg.material_cfl * cfg.cell / sqrtf(c2));
            if (m.dyn_visc > 0.0f) {
                float kin = m.dyn_visc / d;
                if (kin > kF32Eps) out = rmin(out, cfg.viscous_cfl * cfg.cell * cfg.cell / kin);
            }
            return out;
        }
        default:
            return kInf;
    }
}

// DruckerPragerMaterial::alpha + project  sand.rs:135-194
EMH_DEV bool dp_project(const DevMat& m, float s1, float s2, float lvs, float q, float& o1, float& o2,
                        float& dq) {
    float a1 = rmax(fabsf(s1), kLogClamp), a2 = rmax(fabsf(s2), kLogClamp);
    float e1 = logf(a1) + lvs * 0.5f;
    float e2 = logf(a2) + lvs * 0.5f;
    float tr = e1 + e2;
    float ht = tr * 0.5f;
    float d1 = e1 - ht, d2 = e2 - ht;
    float dn = len2(d1, d2);
    if (dn == 0.0f || tr > 0.0f) {
        o1 = 1.0f;
        o2 = 1.0f;
        dq = dn;
        return true;
    }
    float phi = m.dp_phi0 + (m.dp_h1 * q - m.dp_h3) * expf(-m.dp_h2 * q);
    float sp = sinf(phi);
    float alpha = m.dp_sqrt23 * (2.0f * sp) / (3.0f - sp);
    float gamma = dn + m.dp_ratio * tr * alpha - m.dp_cohesion_term;
    if (gamma <= 0.0f) return false;
    float h1 = e1 - gamma * (d1 / dn);
    float h2 = e2 - gamma * (d2 / dn);
    o1 = expf(h1);
    o2 = expf(h2);
    dq = gamma;
    return true;
}

// NaccMaterial::project  nacc.rs:129-223
// NaccMaterial::project (nacc.rs:129-223) in singular-value space, split at
// the yield-function check: nacc_trial evaluates the trial state and the
// check, nacc_return the projection of a state that fails it.  nacc_project
// chains them; the learned surrogate replaces nacc_return.
struct NaccTrial {
    float q1, q2, qt, p0, jtr, st1, st2, ptr, y0, y1;
    int yield;  // 0 inside the yield surface (F unchanged), 1 Case A, 2 Case B, 3 Case C
};
EMH_DEV NaccTrial nacc_trial(const DevMat& m, float2 sg, float alpha) {
    const float xi = m.na_xi, beta = m.na_beta, mm = m.na_m, kappa = m.na_kappa;
    NaccTrial t;
    t.q1 = sg.x * sg.x;
    t.q2 = sg.y * sg.y;
    t.qt = t.q1 + t.q2;
    t.p0 = kappa * (1e-5f + sinhf(xi * rmax(-alpha, 0.0f)));
    t.jtr = rmax(sg.x * sg.y, 1.0e-6f);
    float mu_inv = m.mu * (1.0f / t.jtr);
    t.st1 = mu_inv * (t.q1 - t.qt * 0.5f);
    t.st2 = mu_inv * (t.q2 - t.qt * 0.5f);
    float psi = (kappa * 0.5f) * (t.jtr - 1.0f / t.jtr);
    t.ptr = -psi * t.jtr;
    t.y0 = 0.0f;
    t.y1 = 0.0f;
    if (t.ptr > t.p0) {
        t.yield = 1;
        return t;
    }
    if (t.ptr < -beta * t.p0) {
        t.yield = 2;
        return t;
    }
    t.y0 = (1.0f + 2.0f * beta) * 2.0f;
    t.y1 = mm * mm * (t.ptr + beta * t.p0) * (t.ptr - t.p0);
    float sn2 = t.st1 * t.st1 + t.st2 * t.st2;
    float y = t.y0 * sn2 + t.y1;
    t.yield = y < 1.0e-4f ? 0 : 3;
    return t;
}
// Projected singular values (s1, s2) and hardening state of a trial state
// with t.yield != 0.
EMH_DEV void nacc_return(const DevMat& m, const NaccTrial& t, float alpha, float& s1, float& s2, float& ao) {
    const float beta = m.na_beta, mm = m.na_m, kappa = m.na_kappa;
    const bool hard = (m.flags & kMatNaccHardening) != 0u;
    const float p0 = t.p0, jtr = t.jtr, ptr = t.ptr;
    if (t.yield == 1 || t.yield == 2) {
        float jn = t.yield == 1 ? sqrtf(rmax(-2.0f * p0 / kappa + 1.0f, 1.0e-8f))
                                : sqrtf(rmax(2.0f * beta * p0 / kappa + 1.0f, 1.0e-8f));
        float sv = powf(jn, 0.5f);
        if (hard) alpha = alpha + logf(jtr / jn);
        s1 = sv;
        s2 = sv;
        ao = alpha;
        return;
    }
    if (hard && p0 > 1.0e-4f && ptr < p0 - 1.0e-4f && ptr > -beta * p0 + 1.0e-4f) {
        float pc = (1.0f - beta) * p0 * 0.5f;
        float qtr = sqrtf(2.0f) * len2(t.st1, t.st2);
        float dx = pc - ptr, dy = -qtr;
        float rcp = 1.0f / len2(dx, dy);
        if (__builtin_isfinite(rcp) && rcp > 0.0f) {
            dx = dx * rcp;
            dy = dy * rcp;
        } else {
            dx = 0.0f;
            dy = 0.0f;
        }
        float c = mm * mm * (pc + beta * p0) * (pc - p0);
        float b = mm * mm * dx * (2.0f * pc - p0 + beta * p0);
        float a = mm * mm * dx * dx + (1.0f + 2.0f * beta) * dy * dy;
        float disc = sqrtf(rmax(b * b - 4.0f * a * c, 0.0f));
        float l1 = (-b + disc) / (2.0f * a);
        float l2 = (-b - disc) / (2.0f * a);
        float p1 = pc + l1 * dx;
        float p2 = pc + l2 * dx;
        float px = ((ptr - pc) * (p1 - pc) > 0.0f) ? p1 : p2;
        float jex = sqrtf(rmax(fabsf(-2.0f * px / kappa + 1.0f), 1.0e-8f));
        if (jex > 1.0e-4f) alpha = alpha + logf(jtr / jex);
    }
    float sc = sqrtf(rmax(-t.y1 / rmax(t.y0, 1.0e-10f), 0.0f)) * (jtr / m.mu);
    float nx = t.st1, ny = t.st2;
    float rcp = 1.0f / len2(nx, ny);
    if (__builtin_isfinite(rcp) && rcp > 0.0f) {
        nx = nx * rcp;
        ny = ny * rcp;
    } else {
        nx = 0.0f;
        ny = 0.0f;
    }
    float b1 = sc * nx + t.qt / 2.0f;
    float b2 = sc * ny + t.qt / 2.0f;
    s1 = sqrtf(rmax(b1, 1.0e-8f));
    s2 = sqrtf(rmax(b2, 1.0e-8f));
    ao = alpha;
}
EMH_DEV void nacc_project(const DevMat& m, const M2& f, float alpha, M2& fo, float& ao) {
    M2 u, vt;
    float2 sg;
    svd2(f, u, sg, vt);
    const NaccTrial t = nacc_trial(m, sg, alpha);
    if (t.yield == 0) {
        fo = f;
        ao = alpha;
        return;
    }
    float s1, s2;
    nacc_return(m, t, alpha, s1, s2, ao);
    fo = usv(u, s1, s2, vt);
}

EMH_DEV void sync_volume(PState& s) {
    float j = rmax(mdet(s.f), kMinJ);
    float v = rmax(s.v0 * j, 1.0e-6f);
    s.vol = v;
    s.dens = s.mass / v;
}

// update_particle: elastic.rs:262-269, corotated.rs:84-91, viscoelastic.rs:132-139,
// sand.rs:248-319, von_mises.rs:108-141, rankine.rs:168-203, sand_mui.rs:151-217,
// granular_fluid.rs:200-230, snow.rs:142-176, fluid.rs:197-207 (bingham.rs:184-193 is
// the same), nacc.rs:255-269; Fallback: no-op.
template <int ONLY = -1>
EMH_DEV void update_particle(const DevMat& m, PState& s, float dt) {
    const uint32_t model = ONLY >= 0 ? (uint32_t)ONLY : m.model;
    switch (model) {
        case kModelNH:
        case kModelCorot:
        case kModelViscoelastic: {
            M2 fp = madd(M2{1.0f, 0.0f, 0.0f, 1.0f}, mscale(s.c, dt));
            s.f = mmul(fp, s.f);
            sync_volume(s);
            break;
        }
        case kModelDP: {
            M2 ft = mmul(madd(M2{1.0f, 0.0f, 0.0f, 1.0f}, mscale(s.c, dt)), s.f);
            M2 u, vt;
            float2 sg;
            svd2(ft, u, sg, vt);
            float n1 = sg.x, n2 = sg.y;
            float p1, p2, dq;
            if (dp_project(m, sg.x, sg.y, s.lvs, s.q, p1, p2, dq)) {
                float a1 = rmax(fabsf(sg.x), kLogClamp), a2 = rmax(fabsf(sg.y), kLogClamp);
                float prev = a1 * a2;
                float nd = p1 * p2;
                float diff = nd - prev;
                float corr = diff > 0.0f ? nd : prev + diff * m.dp_vol_corr;
                s.lvs = s.lvs + (logf(prev) - logf(corr));
                s.q = rmin(s.q + dq, m.dp_qmax);
                if (m.dp_dilat > 0.0f) s.lvs = s.lvs + m.dp_sin_dilat * dq;
                n1 = p1;
                n2 = p2;
            }
            n1 = fabsf(n1);
            n2 = fabsf(n2);
            float jn = n1 * n2;
            if (jn < m.dp_min_vj) {
                float k = sqrtf(m.dp_min_vj / rmax(jn, 1e-6f));
                n1 = n1 * k;
                n2 = n2 * k;
            }
            s.f = usv(u, n1, n2, vt);
            sync_volume(s);
            break;
        }
        case kModelVonMises: {
            // J2 return map on the Hencky strain: the deviatoric part is scaled
            // back onto 2 mu |dev| = yield + H kappa, the trace is kept
            M2 ft = mmul(madd(M2{1.0f, 0.0f, 0.0f, 1.0f}, mscale(s.c, dt)), s.f);
            M2 u, vt;
            float2 sg;
            svd2(ft, u, sg, vt);
            float e1 = logf(rmax(fabsf(sg.x), kLogClamp)), e2 = logf(rmax(fabsf(sg.y), kLogClamp));
            float tr = e1 + e2;
            float ht = tr * 0.5f;
            float d1 = e1 - ht, d2 = e2 - ht;
            float dn = len2(d1, d2);
            float kappa = s.q;
            float ey = m.vm_yield + m.vm_hmod * kappa;
            float el = (2.0f * m.mu) * dn;
            float n1 = sg.x, n2 = sg.y;
            if (el > ey && dn > kLogClamp) {
                float den = 2.0f * m.mu + m.vm_hmod;
                float gamma = den > kF32Eps ? (el - ey) / den : 0.0f;
                s.q = kappa + gamma;
                float k = ey / el;
                n1 = expf(d1 * k + ht);
                n2 = expf(d2 * k + ht);
            }
            s.f = usv(u, n1, n2, vt);
            sync_volume(s);
            break;
        }
        case kModelRankine: {
            // principal Kirchhoff stresses of the Hencky strain; components
            // above the softened cutoff return to it, the strain distance
            // moved accumulates as damage (capped where the cutoff floors)
            M2 ft = mmul(madd(M2{1.0f, 0.0f, 0.0f, 1.0f}, mscale(s.c, dt)), s.f);
            M2 u, vt;
            float2 sg;
            svd2(ft, u, sg, vt);
            float e1 = logf(rmax(fabsf(sg.x), kLogClamp)), e2 = logf(rmax(fabsf(sg.y), kLogClamp));
            float a = 2.0f * m.mu + m.lambda;
            float t1 = a * e1 + m.lambda * e2, t2 = m.lambda * e1 + a * e2;
            float damage = s.q;
            float teff = rmax(m.rk_ts * expf(-m.rk_sr * damage), m.rk_floor);
            bool y1 = t1 > teff, y2 = t2 > teff;
            float n1 = sg.x, n2 = sg.y;
            if (y1 || y2) {
                float p1 = y1 ? teff : t1, p2 = y2 ? teff : t2;
                float den = 4.0f * m.mu * (m.mu + m.lambda);
                float ep1 = (a * p1 - m.lambda * p2) / den, ep2 = (a * p2 - m.lambda * p1) / den;
                float et1 = (a * t1 - m.lambda * t2) / den, et2 = (a * t2 - m.lambda * t1) / den;
                s.q = rmin(damage + len2(et1 - ep1, et2 - ep2), m.rk_sat);
                n1 = expf(ep1);
                n2 = expf(ep2);
            }
            s.f = usv(u, n1, n2, vt);
            sync_volume(s);
            break;
        }
        case kModelMuI: {
            // rate-dependent Drucker-Prager: tension resets to sigma = 1;
            // beyond mu_static p the plastic rate solves the quadratic in
            // closed form and mu(I) is stored in friction_hardening
            const float kSqrt2 = 1.41421356237309504880f;
            M2 ft = mmul(madd(M2{1.0f, 0.0f, 0.0f, 1.0f}, mscale(s.c, dt)), s.f);
            M2 u, vt;
            float2 sg;
            svd2(ft, u, sg, vt);
            float e1 = logf(rmax(fabsf(sg.x), kLogClamp)), e2 = logf(rmax(fabsf(sg.y), kLogClamp));
            float tr = e1 + e2;
            float ptr = -(m.lambda + m.mu) * tr;
            float n1 = sg.x, n2 = sg.y;
            float mu_out = m.mui_s;
            if (ptr <= 0.0f) {
                n1 = 1.0f;
                n2 = 1.0f;
            } else {
                float ht = tr * 0.5f;
                float d1 = e1 - ht, d2 = e2 - ht;
                float dn = len2(d1, d2);
                float qtr = kSqrt2 * m.mu * dn;
                float qy = m.mui_s * ptr;
                if (!(qtr <= qy || dn < kF32Eps)) {
                    float dq = qtr - qy;
                    float sp = sqrtf(ptr);
                    float a = m.mu * dt;
                    float b = ptr * (m.mui_d - m.mui_s) + a * m.mui_q * sp - dq;
                    float c = -dq * m.mui_q * sp;
                    float g = rmax((-b + sqrtf(b * b - 4.0f * a * c)) / (2.0f * a), 0.0f);
                    if (g > kF32Eps) mu_out = m.mui_s + (m.mui_d - m.mui_s) / (m.mui_q * sp / g + 1.0f);
                    float sh = g * dt / kSqrt2;
                    n1 = expf(e1 - (d1 / dn) * sh);
                    n2 = expf(e2 - (d2 / dn) * sh);
                }
            }
            s.f = usv(u, n1, n2, vt);
            s.q = mu_out;
            sync_volume(s);
            break;
        }
        case kModelGranular: {
            M2 ft = mmul(madd(M2{1.0f, 0.0f, 0.0f, 1.0f}, mscale(s.c, dt)), s.f);
            if (m.flags & kMatGfPlastic) {
                M2 u, vt;
                float2 sg;
                svd2(ft, u, sg, vt);
                float c1 = rclamp(sg.x, m.sn_lo, m.sn_hi), c2 = rclamp(sg.y, m.sn_lo, m.sn_hi);
                float jpn = s.jp * (sg.x * sg.y) / rmax(c1 * c2, 1.0e-10f);
                s.jp = rclamp(jpn, m.sn_jmin, m.sn_jmax);
                s.h = rclamp(expf(m.sn_xi * (1.0f - s.jp)), 0.1f, 7.0f);
                s.f = usv(u, c1, c2, vt);
            } else {
                s.f = ft;
            }
            sync_volume(s);
            break;
        }
        case kModelSnow: {
            M2 ft = mmul(madd(M2{1.0f, 0.0f, 0.0f, 1.0f}, mscale(s.c, dt)), s.f);
            M2 u, vt;
            float2 sg;
            svd2(ft, u, sg, vt);
            float c1 = rclamp(sg.x, m.sn_lo, m.sn_hi), c2 = rclamp(sg.y, m.sn_lo, m.sn_hi);
            float jpn = s.jp * (sg.x * sg.y) / (c1 * c2);
            s.jp = rclamp(jpn, m.sn_jmin, m.sn_jmax);
            s.h = rclamp(expf(m.sn_xi * (1.0f - s.jp)), 0.1f, 7.0f);
            s.f = usv(u, c1, c2, vt);
            sync_volume(s);
            break;
        }
        case kModelFluid: {
            float j = rclamp(mdet(s.f), 0.5f, 2.0f);
            float r = sqrtf(j);
            s.f = M2{r, 0.0f, 0.0f, r};
            if (m.settling > 0.0f) {
                float k = 1.0f - rmin(m.settling * dt, 0.5f);
                s.vx = s.vx * k;
                s.vy = s.vy * k;
            }
            break;
        }
        case kModelNacc: {
            M2 fin = s.f;
            if (m.flags & kMatNaccAdvance) fin = mmul(madd(M2{1.0f, 0.0f, 0.0f, 1.0f}, mscale(s.c, dt)), fin);
            M2 nf;
            float na;
            nacc_project(m, fin, s.lvs, nf, na);
            s.f = nf;
            s.lvs = na;
            float j = rmax(mdet(nf), kMinJ);
            float vol = s.v0 * j;
            s.vol = vol;
            if (vol > kMinJ) s.dens = s.mass / vol;
            break;
        }
        default:
            break;
    }
}

// affine_cfl_speed_contribution  solver/step.rs:481-488
EMH_DEV float affine_cfl(const M2& c, float cell) {
    const float kCorner = 1.5f * 1.41421356237309504880f;
    float g = sqrtf((c.a * c.a + c.b * c.b) + (c.c * c.c + c.d * c.d));
    return g * kCorner * cell;
}

// Cell key: the cell's index along the Hilbert curve over the 2^order x
// 2^order square that holds the grid (order = cell_key_order(R)).  Cells
// whose keys are consecutive are edge neighbours, and every aligned run of
// 4^j keys fills a 2^j x 2^j square, so any run of keys -- a workgroup's
// chunk of particles, or a wave's -- covers a compact patch also where the
// cells it takes particles from are sparse (a thin layer at a body's edge).
// Cell coordinates are clamped into the grid.  The curve is the classic one
// (quadrant digits (3 rx) ^ ry from the top bit down, the lower bits rotated
// or reflected when ry = 0).
__host__ __device__ inline uint32_t cell_key_order(uint32_t R) {
    uint32_t k = 1;
    while (k < 16u && (1u << k) < R) ++k;
    return k;
}
__host__ __device__ inline uint32_t cell_key(int cx, int cy, uint32_t R, uint32_t order) {
    int hi = (int)R - 1;
    cx = cx < 0 ? 0 : (cx > hi ? hi : cx);
    cy = cy < 0 ? 0 : (cy > hi ? hi : cy);
    const uint32_t n = 1u << order;
    uint32_t x = (uint32_t)cx, y = (uint32_t)cy, d = 0;
    for (uint32_t s = n >> 1; s > 0u; s >>= 1) {
        const uint32_t rx = (x & s) ? 1u : 0u, ry = (y & s) ? 1u : 0u;
        d += s * s * ((3u * rx) ^ ry);
        if (ry == 0u) {
            if (rx == 1u) {
                x = n - 1u - x;
                y = n - 1u - y;
            }
            const uint32_t t = x;
            x = y;
            y = t;
        }
    }
    return d;
}

}  // namespace emh
