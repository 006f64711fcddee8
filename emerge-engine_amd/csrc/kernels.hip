// kernels.hip -- the MLS-MPM substep as gfx950 HIP kernels.
//
// Simulation::step (solver/step.rs:23-60) runs on the device as:
//   k_begin_step     1 lane: queues the start of a step (step.rs:24-33).
//  [k_sort_*]        on reorder steps: a stable radix sort of the particles'
//                    cell keys (count, scan and scatter per pass) giving
//                    each slot's position in cell-key order.
// then, per substep:
//  [k_density_p2g]   only when a material re-estimates density
//                    (solver/density.rs:216-247): gathers the density of the
//                    mass grid the previous substep scattered, then P2G.
//   k_substep        one workgroup per 1024 particles, in cell-key order:
//                    0. the substep's dt (choose_substep_dt/cfl_bound,
//                       step.rs:434-479) and the loop condition (step.rs:34),
//                       derived by every workgroup from the control state and
//                       the CFL reduction the previous launch left;
//                    1. grid update (grid/mod.rs:264-291 + SlipBoundary
//                       boundary/mod.rs:78-101 + speed clamp step.rs:150-165)
//                       of the chunk's bounding box of cells, staged in LDS;
//                    2. G2P + update_particle (transfer/g2p.rs:151-343);
//                    3. the CFL contribution of the new state (step.rs:445-466);
//                    4. the NEXT substep's P2G (transfer/p2g.rs:24-72) from the
//                       new state held in registers, accumulated in LDS with
//                       wave-level segmented sums and flushed with one global
//                       f32 atomic per touched cell and channel.  The next dt
//                       is not known until every particle is done, so the
//                       stress coefficient -4 V0 dt (p2g.rs:42-44) uses a
//                       predicted dt' and the stress moment is kept beside the
//                       momentum for the grid update's (dt - dt') correction;
//                    5. the CFL contribution (max speed, min material dt) of
//                       the new state, reduced with integer atomics on the
//                       f32 bit patterns (both are non-negative).
// The particle state is read once and written once per substep.
//
// Projection (project_particle_state_to_admissible, step.rs:347-429) is a
// pure function of the stored state.  A kernel that writes a state checks
// whether projecting it would change anything; only then does the next reader
// project (and count) on the fly and persist the projected fields.
#include <hip/hip_runtime.h>

#include <climits>
#include <cstring>
#include <type_traits>

#include "kernels.hpp"
#include "surrogate_kernels.hpp"

namespace emh {

namespace {

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63u; }

__device__ __forceinline__ PBuf sel_buf(const Pipeline& p, uint32_t i) { return i ? p.buf[1] : p.buf[0]; }
// The buffer a reorder substep stores into, read where it stores: the index
// passes through an empty asm definition, so the compiler cannot hoist the
// buffer's fourteen pointers out of the particle loop and keep them live in
// SGPRs through every substep (those spilled into VGPR lanes, and their
// reloads were VALU instructions on the loop's path).
// A uniform index made opaque to the optimizer, for loads that only a rare
// path needs inside the particle loop (see reorder_buf).
__device__ __forceinline__ uint32_t opaque_index(uint32_t i) {
    i = __builtin_amdgcn_readfirstlane(i);
    asm volatile("" : "+s"(i));
    return i;
}
__device__ __forceinline__ PBuf reorder_buf(const Pipeline& p, uint32_t i) {
    i = __builtin_amdgcn_readfirstlane(i);  // uniform: the pipeline's control state
    asm volatile("" : "+s"(i));
    return p.buf[i & 1u];
}

// Global float add (every caller passes grid memory): the address-space cast
// keeps it a global atomic where the pointer was loaded from memory (the
// batch kernels' per-scene pipelines) and would otherwise be a flat one.
__device__ __forceinline__ void gatomic_add(float* addr, float v) {
    __hip_atomic_fetch_add((__attribute__((address_space(1))) float*)addr, v, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}
// LDS add into a box accumulator, workgroup scope.  The box accumulates in
// f64: gfx950's LDS executes an f64 add (ds_add_f64) at about 0.14 CU cycles
// per lane, its f32 add (ds_add_f32) one lane at a time at about 3 cycles per
// lane; each box cell is rounded to f32 once, when it is flushed.  Keeping it
// distinct from gatomic_add also stops the compiler from merging an LDS add
// and a global add on two sides of a branch into one flat atomic.
__device__ __forceinline__ void latomic_add(double* addr, double v) {
    __hip_atomic_fetch_add(addr, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    const uint32_t lane = lane_id();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t n = __shfl_up(v, o, 64);
        if (lane >= (uint32_t)o) v += n;
    }
    return v;
}
__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ int wave_imin(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ int wave_imax(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}

// Block-wide exclusive scan (blockDim.x a multiple of 64, at most 1024).
__device__ uint32_t block_excl_scan(uint32_t v, uint32_t* s_wave, uint32_t& total) {
    const uint32_t lane = lane_id(), wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    uint32_t incl = wave_incl_scan(v);
    if (lane == 63) s_wave[wave] = incl;
    __syncthreads();
    if (wave == 0) {
        uint32_t w = lane < nw ? s_wave[lane] : 0u;
        uint32_t wi = wave_incl_scan(w);
        if (lane < nw) s_wave[lane] = wi - w;  // exclusive prefix of wave totals
        if (lane == nw - 1) s_wave[16] = wi;
    }
    __syncthreads();
    uint32_t out = s_wave[wave] + incl - v;
    total = s_wave[16];
    __syncthreads();
    return out;
}

// Block-wide (max, min) of a float pair; every thread gets the result.
__device__ void block_maxmin(float& a, float& d, float* s_red) {
    a = wave_max(a);
    d = wave_min(d);
    const uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
    if (lane == 0) {
        s_red[2 * wave] = a;
        s_red[2 * wave + 1] = d;
    }
    __syncthreads();
    a = s_red[0];
    d = s_red[1];
    for (uint32_t w = 1; w < (blockDim.x >> 6); ++w) {
        a = fmaxf(a, s_red[2 * w]);
        d = fminf(d, s_red[2 * w + 1]);
    }
    __syncthreads();
}

// ------------------------------------------------------------ particle I/O
// Fields a kernel does not load keep values projection leaves unchanged.
__device__ __forceinline__ PState neutral_state() {
    PState s;
    s.x = s.y = s.vx = s.vy = 0.0f;
    s.c = M2{0.0f, 0.0f, 0.0f, 0.0f};
    s.f = M2{1.0f, 0.0f, 0.0f, 1.0f};
    s.mass = s.v0 = s.vol = s.dens = 1.0f;
    s.jp = s.h = 1.0f;
    s.q = s.lvs = 0.0f;
    s.mat = 0;
    return s;
}
// x is loaded by the caller; v and C only when `vc`.
__device__ __forceinline__ void load_fields(PState& s, const PBuf& b, uint32_t i, uint32_t mask, bool vc) {
    if (vc) {
        const float2 v = b.v[i];
        s.vx = v.x;
        s.vy = v.y;
        s.c = from4(b.c[i]);
    }
    s.f = from4(b.f[i]);
    s.mass = b.mass[i];
    s.v0 = b.v0[i];
    s.mat = b.mat[i];
    if (mask & kFVolDens) {
        s.vol = b.vol[i];
        s.dens = b.dens[i];
    }
    if (mask & kFJp) s.jp = b.jp[i];
    if (mask & kFH) s.h = b.h[i];
    if (mask & kFQ) s.q = b.q[i];
    if (mask & kFLvs) s.lvs = b.lvs[i];
}
// x, v, C, F, volume and density always; the optional fields in `mask`;
// mass, initial volume and material word when `meta`.
__device__ __forceinline__ void store_fields(const PBuf& b, uint32_t i, const PState& s, uint32_t mask, bool meta) {
    b.x[i] = make_float2(s.x, s.y);
    b.v[i] = make_float2(s.vx, s.vy);
    b.c[i] = to4(s.c);
    b.f[i] = to4(s.f);
    b.vol[i] = s.vol;
    b.dens[i] = s.dens;
    if (mask & kFJp) b.jp[i] = s.jp;
    if (mask & kFH) b.h[i] = s.h;
    if (mask & kFQ) b.q[i] = s.q;
    if (mask & kFLvs) b.lvs[i] = s.lvs;
    if (meta) {
        b.mass[i] = s.mass;
        b.v0[i] = s.v0;
        b.mat[i] = s.mat;
    }
}

// Position part of project_particle_state_to_admissible (step.rs:357-362).
__device__ __forceinline__ void project_x(float& x, float& y, float lo, float hi) {
    if (!(__builtin_isfinite(x) && __builtin_isfinite(y))) {
        float c = (lo + hi) * 0.5f;
        x = c;
        y = c;
    } else {
        x = rmin(rmax(x, lo), hi);
        y = rmin(rmax(y, lo), hi);
    }
}
__device__ __forceinline__ void project_x(float& x, float& y, const DevConfig& cfg) {
    project_x(x, y, cfg.proj_lo, cfg.proj_hi);
}

// Would projecting `s` change it?  (`sp` receives the projected state.)
// Every lane of the wave calls this.  A state that project_state leaves
// untouched is recognised directly (finite fields, positive scalars, x inside
// the clamp range, J within (proj_min_j, j_max], density at or above its
// floor); only a wave with some other lane runs project_state itself.
__device__ __forceinline__ bool needs_projection(const PState& s, PState& sp, const DevConfig& cfg) {
    sp = s;
    if (!cfg.project) return false;
    const float j = mdet(s.f);
    const bool ok = s.x >= cfg.proj_lo && s.x <= cfg.proj_hi && s.y >= cfg.proj_lo && s.y <= cfg.proj_hi &&
                    __builtin_isfinite(s.vx) && __builtin_isfinite(s.vy) && mfinite(s.c) && mfinite(s.f) &&
                    j > cfg.proj_min_j && j <= cfg.j_max && s.jp > 0.0f && __builtin_isfinite(s.jp) && s.h > 0.0f &&
                    __builtin_isfinite(s.h) && __builtin_isfinite(s.q) && __builtin_isfinite(s.lvs) &&
                    s.mass > 0.0f && __builtin_isfinite(s.mass) && s.v0 > 0.0f && __builtin_isfinite(s.v0) &&
                    s.vol > 0.0f && __builtin_isfinite(s.vol) && s.dens >= cfg.proj_min_density &&
                    __builtin_isfinite(s.dens);
    if (__ballot(!ok) == 0ull) return false;
    const bool reset = project_state(sp, cfg);
    return reset || __float_as_uint(sp.x) != __float_as_uint(s.x) || __float_as_uint(sp.y) != __float_as_uint(s.y) ||
           __float_as_uint(sp.dens) != __float_as_uint(s.dens);
}

__device__ __forceinline__ const DevMat& mat_of(const Pipeline& p, uint32_t word) {
    uint32_t m = word & kMatMask;  // MaterialRegistry::get falls back to slot 0 (registry.rs:66-76)
    return p.mats[m < p.n_mats ? m : 0u];
}

// CFL contribution of one particle as choose_substep_dt reads it, before
// projection (step.rs:445-466): speed incl. affine term, material dt bound.
// For the models whose bound is elastic_wave_dt(lambda, mu, 1, density, ...)
// the bound depends on the particle only through max(density, kMinJ) and
// never decreases with it (a correctly rounded quotient and square root are
// monotone, and c <= eps gives +inf), so in a scene of one material the
// minimum over particles is the bound at the smallest clamped density:
// density_bound_models() scenes reduce that density (mdt holds it) and
// cfl_bound() evaluates the bound once per workgroup.
template <int ONLY>
constexpr bool density_bound_model() {
    return ONLY == (int)kModelDP || ONLY == (int)kModelVonMises || ONLY == (int)kModelRankine ||
           ONLY == (int)kModelMuI;
}
template <int ONLY>
__device__ __forceinline__ void cfl_of(const Pipeline& p, const PState& s, float& speed, float& mdt) {
    float sp = len2(s.vx, s.vy);
    if (p.cfg.cfl_affine) sp = sp + affine_cfl(s.c, p.cfg.cell);
    speed = fmaxf(speed, sp);
    if (density_bound_model<ONLY>() && p.n_mats == 1u) {
        mdt = fminf(mdt, rmax(s.dens, kMinJ));
        return;
    }
    const float b = timestep_bound<ONLY>(mat_of(p, s.mat), s.dens, s.h, p.cfg);
    if (__builtin_isfinite(b) && b > 0.0f) mdt = fminf(mdt, b);
}
// The workgroup's material dt bound from what cfl_of accumulated.
template <int ONLY>
__device__ __forceinline__ float cfl_bound(const Pipeline& p, float mdt) {
    if (!(density_bound_model<ONLY>() && p.n_mats == 1u) || !(mdt < kInf)) return mdt;
    const float b = timestep_bound<ONLY>(p.mats[0], mdt, 1.0f, p.cfg);
    return (__builtin_isfinite(b) && b > 0.0f) ? b : kInf;
}

// The decision of control launch s from the state launch s - 1 left and its
// CFL reduction: the loop condition of Simulation::step (step.rs:34) --
// finishing the step and starting the next queued one when it fails -- then
// choose_substep_dt / cfl_bound (step.rs:434-479).
__device__ __forceinline__ CtrlState decide_substep(const CtrlState& P, const CtrlRed& Q, const DevConfig& cfg) {
    CtrlState N = P;
    N.active = 0;
    bool running = P.in_step && P.remaining > kF32Eps && P.substeps < cfg.max_substeps;
    if (!running) {
        if (P.in_step) {  // step.rs:51-59: the step's records
            N.in_step = 0;
            N.done_steps = P.done_steps + 1;
            N.done_step = P.step;
            N.done_substeps = P.substeps;
            N.done_dt = P.last_dt;
            N.done_dropped = rmax(P.remaining, 0.0f);
        }
        if (N.pending == 0) return N;
        N.pending -= 1;  // step.rs:24-33
        N.in_step = 1;
        N.step = P.step + 1;
        N.remaining = cfg.dt;
        N.substeps = 0;
        if (!(N.remaining > kF32Eps)) return N;  // a step too short for any substep ends next launch
    }
    const float a = __uint_as_float(Q.max_speed), d = __uint_as_float(Q.min_mdt);
    const float remaining = N.remaining;
    float dt, free = cfg.dt;
    if (!cfg.adaptive) {
        dt = rmin(remaining, cfg.dt);
    } else {
        float min_mat_dt = rmin(remaining, d);
        dt = remaining;
        if (a > kF32Eps) dt = rmin(dt, cfg.cfl_coefficient * cfg.cell / a);
        dt = rmin(dt, min_mat_dt);
        dt = rclamp(dt, rmin(cfg.min_dt, remaining), remaining);
        free = d;
        if (a > kF32Eps) free = rmin(free, cfg.cfl_coefficient * cfg.cell / a);
        free = rmax(free, cfg.min_dt);
    }
    N.dt = dt;
    N.free_dt = free;
    N.last_dt = dt;
    N.active = 1;
    N.substeps = N.substeps + 1;
    N.remaining = remaining - dt;
    N.total_substeps = P.total_substeps + 1;
    N.cur_grid = P.cur_grid == 2u ? 0u : P.cur_grid + 1u;
    N.cur_buf = P.cur_buf ^ (P.reorder ? 1u : 0u);
    N.reorder = 0;
    return N;
}
// Predicted dt of the substep after the one state N describes: the same
// clamps as decide_substep with the CFL bound of N's own decision (the CFL
// reduction of the new state is not known yet).  Exact for fixed dt and for
// a step's last, time-limited substep.
__device__ __forceinline__ float predict_dt(const CtrlState& N, const DevConfig& cfg) {
    const float rem = (N.remaining > kF32Eps && N.substeps < cfg.max_substeps) ? N.remaining : cfg.dt;
    if (!cfg.adaptive) return rmin(rem, cfg.dt);
    float dt = rmin(rem, N.free_dt);
    return rclamp(dt, rmin(cfg.min_dt, rem), rem);
}
__device__ __forceinline__ void reset_red(CtrlRed& r) {
    r.max_speed = 0u;
    r.min_mdt = kInfBits;
    r.dirty = 0u;
}
__device__ __forceinline__ uint32_t seq_mod3(uint64_t s) { return (uint32_t)(s % 3ull); }

// Grid velocity of one cell after Grid::update_velocities, gravity, the slip
// walls and the speed clamp (grid/mod.rs:264-291, step.rs:140-165), from the
// P2G accumulators of the current substep; (0, 0) for untouched cells
// (velocity_at of a missing cell, grid/mod.rs:215-233).
// Coulomb wall  boundary/mod.rs:57-75 (outward normal (nx, ny))
__device__ __forceinline__ void coulomb_wall(float& vx, float& vy, float nx, float ny, float mu) {
    const float vn = vx * nx + vy * ny;
    if (vn >= 0.0f) return;
    const float ns = fabsf(vn);
    const float tx = vx - nx * vn, ty = vy - ny * vn;
    const float tl = len2(tx, ty);
    const float fi = mu * ns;
    if (tl > fi) {
        const float k = (tl - fi) / tl;
        vx = tx * k;
        vy = ty * k;
    } else {
        vx = 0.0f;
        vy = 0.0f;
    }
}
__device__ __forceinline__ float terrain_height(const DevConfig& cfg, uint32_t b, uint32_t x) {
    const uint32_t n = cfg.bnd_hlen[b];
    return n == 0u ? 0.0f : cfg.heights[cfg.bnd_hoff[b] + (x < n - 1u ? x : n - 1u)];
}
// BoundaryCondition::apply_to_grid_velocity of boundary b on cell (x, y):
// slip.rs:16-18 (mod.rs:78-101), friction.rs:40-59, heightmap.rs:66-98, predictive.rs:27-29
__device__ __forceinline__ void boundary_velocity(const DevConfig& cfg, uint32_t b, uint32_t x, uint32_t y,
                                                  float& vx, float& vy) {
    const uint32_t R = cfg.grid_res, t = cfg.bnd_t[b], kind = cfg.bnd_kind[b];
    const uint32_t hi = R > t + 1u ? R - (t + 1u) : 0u;
    if (kind == kBndSlip || kind == kBndPredictive) {
        if (x < t) vx = rmax(vx, 0.0f);
        if (x > hi) vx = rmin(vx, 0.0f);
        if (y < t) vy = rmax(vy, 0.0f);
        if (y > hi) vy = rmin(vy, 0.0f);
    } else if (kind == kBndFriction) {
        const float mu = cfg.bnd_mu[b];
        if (x < t) coulomb_wall(vx, vy, 1.0f, 0.0f, mu);
        if (x > hi) coulomb_wall(vx, vy, -1.0f, 0.0f, mu);
        if (y < t) coulomb_wall(vx, vy, 0.0f, 1.0f, mu);
        if (y > hi) coulomb_wall(vx, vy, 0.0f, -1.0f, mu);
    } else {  // heightmap: slip side and top walls, Coulomb floor on the terrain
        if (x < t) vx = rmax(vx, 0.0f);
        if (x > hi) vx = rmin(vx, 0.0f);
        if (y > hi) vy = rmin(vy, 0.0f);
        if ((float)y <= terrain_height(cfg, b, x) && vy < 0.0f) {
            const float vn = fabsf(vy);
            vy = 0.0f;
            const float mu = cfg.bnd_mu[b];
            if (mu > 0.0f) {
                const float fi = mu * vn;
                const float vt = fabsf(vx);
                vx = vt > fi ? vx * (1.0f - fi / vt) : 0.0f;
            }
        }
    }
}
// BoundaryCondition::clamp_particle_position of every boundary in order:
// mod.rs:103-111, predictive.rs:31-35, heightmap.rs:100-114
__device__ __forceinline__ void boundary_clamp(const DevConfig& cfg, float& nx, float& ny) {
    for (uint32_t q = 0; q < cfg.n_bnd; ++q) {
        nx = rmin(rmax(nx, cfg.bnd_lo[q]), cfg.bnd_hi[q]);
        ny = rmin(rmax(ny, cfg.bnd_lo[q]), cfg.bnd_hi[q]);
        if (cfg.bnd_kind[q] == kBndHeightmap) {
            uint32_t xc = nx >= 1.0f ? (uint32_t)nx : 0u;  // `as usize` saturates at 0
            xc = xc < cfg.grid_res - 1u ? xc : cfg.grid_res - 1u;
            const float h = terrain_height(cfg, q, xc);
            if (ny < h + 1.0f) ny = h + 1.0f;
        }
    }
}

__device__ __forceinline__ float2 cell_velocity(const Pipeline& p, uint32_t set, float dt, float ddt, int cx, int cy) {
    const DevConfig& cfg = p.cfg;
    const int R = (int)cfg.grid_res;
    if (cx < 0 || cy < 0 || cx >= R || cy >= R) return make_float2(0.0f, 0.0f);
    const uint32_t gi = (uint32_t)cx * (uint32_t)R + (uint32_t)cy;
    // every channel the cell needs is read before the mass is tested: one
    // round trip to memory per cell instead of two (the stencil's nine cells
    // of a particle outside its LDS box then go out together)
    const float m = p.grid.acc[set][kAccM][gi];
    float vx = p.grid.acc[set][kAccPx][gi];
    float vy = p.grid.acc[set][kAccPy][gi];
    float sx = 0.0f, sy = 0.0f;
    if (ddt != 0.0f) {
        sx = p.grid.acc[set][kAccSx][gi];
        sy = p.grid.acc[set][kAccSy][gi];
    }
    if (!(m > 0.0f)) return make_float2(0.0f, 0.0f);
    // momentum = A + (dt - dt') S (ddt = dt - dt' is 0 when the prediction held,
    // and always for scenes with a density pass, which never write S)
    if (ddt != 0.0f) {
        vx = vx + ddt * sx;
        vy = vy + ddt * sy;
    }
    vx = vx / m;
    vy = vy / m;
    vx = vx + cfg.gx * dt;
    vy = vy + cfg.gy * dt;
    const uint32_t x = (uint32_t)cx, y = (uint32_t)cy;
    for (uint32_t b = 0; b < cfg.n_bnd; ++b) boundary_velocity(cfg, b, x, y, vx, vy);
    const float vel_limit = cfg.cell / dt;
    const float spd = len2(vx, vy);
    if (spd > vel_limit) {
        const float k = vel_limit / spd;
        vx = vx * k;
        vy = vy * k;
    }
    return make_float2(vx, vy);
}

// ---------------------------------------------------- wave segmented sums
// Lanes of a wave hold consecutive particles of a cell-sorted chunk, so runs
// of lanes share a base cell and with it all nine stencil cells.  A run's
// contributions are summed across its lanes with a segmented inclusive scan
// of six DPP steps (row shifts 1/2/4/8 inside 16-lane rows, then row
// broadcasts 15 and 31; VALU work, no LDS traffic), leaving each run's total
// in its last lane, which alone issues the atomics.  A step adds the shifted
// partial sum only if the source lane is in the same run: as fma(t, p, v)
// with p = 1 or 0, which is exactly v + t or v when t is finite (one
// v_fmac_f32_dpp per step), or as a select (3 instructions) for waves whose
// terms are not all small enough for every partial sum to stay finite.
// kScanSteps = 4 splits runs at the 16-lane row boundaries (every row start
// is a run head), so that the scan needs only the four row shifts; 6 keeps
// runs across rows with the two row-broadcast steps.
constexpr int kScanSteps = 4;
constexpr uint64_t kRowHeads = kScanSteps == 2   ? 0x1111111111111111ull
                               : kScanSteps == 3 ? 0x0101010101010101ull
                               : kScanSteps == 4 ? 0x0001000100010001ull
                                                 : 1ull;
struct SegMask {
    uint64_t heads;  // bit L: lane L starts a run
    uint32_t lane;
};
template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp_take(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROWS, 0xf, true));
}
struct ScanPred {
    bool p[6];   // per step: the source lane is in this lane's run
    float f[6];  // the same as 1.0f / 0.0f
};
// Source lane src < lane is in this lane's run exactly when src >= the run's
// first lane (the highest run head at or below this lane); the row-broadcast
// steps read lane 15 (into rows 1 and 3) and lane 31 (into rows 2 and 3).
__device__ __forceinline__ ScanPred scan_pred(const SegMask& m) {
    ScanPred q;
    const int l = (int)m.lane, r = l & 15;
    const uint64_t upto = (m.lane == 63) ? ~0ull : ((2ull << m.lane) - 1ull);
    const int start = 63 - __clzll((long long)(m.heads & upto));  // lane 0 is always a head
    q.p[0] = r >= 1 && l - 1 >= start;
    q.p[1] = r >= 2 && l - 2 >= start;
    q.p[2] = r >= 4 && l - 4 >= start;
    q.p[3] = r >= 8 && l - 8 >= start;
    q.p[4] = ((l >> 4) == 1 || (l >> 4) == 3) && (l & ~15) - 1 >= start;
    q.p[5] = l >= 32 && 31 >= start;
#pragma unroll
    for (int k = 0; k < 6; ++k) q.f[k] = q.p[k] ? 1.0f : 0.0f;
    return q;
}
template <bool FMA, int CTRL, int ROWS>
__device__ __forceinline__ float scan_step(float v, bool p, float f) {
    const float t = dpp_take<CTRL, ROWS>(v);
    return FMA ? __builtin_fmaf(t, f, v) : v + (p ? t : 0.0f);
}
// The fma form of the nine-value scan with the DPP read folded into the fma:
// v_fmac_f32_dpp v, v, f computes fma(dpp(v), f, v), the same rounding as the
// two-instruction form, in one VALU instruction per step, and a row_bcast
// step leaves the rows it does not write unchanged instead of adding a zero
// (which only differs in the sign of a zero partial sum, and the sums only
// ever reach memory through additions to +0-initialised cells).  Step-major
// order puts eight instructions between a value's write and its next DPP
// read, more than the two wait states the read needs; the leading s_nop 1
// covers whatever VALU write precedes the block.
#define EMH_SCAN9(F, DPP)                                                                                   \
    "v_fmac_f32_dpp %0, %0, " F " " DPP "\n\tv_fmac_f32_dpp %1, %1, " F " " DPP "\n\t"                     \
    "v_fmac_f32_dpp %2, %2, " F " " DPP "\n\tv_fmac_f32_dpp %3, %3, " F " " DPP "\n\t"                     \
    "v_fmac_f32_dpp %4, %4, " F " " DPP "\n\tv_fmac_f32_dpp %5, %5, " F " " DPP "\n\t"                     \
    "v_fmac_f32_dpp %6, %6, " F " " DPP "\n\tv_fmac_f32_dpp %7, %7, " F " " DPP "\n\t"                     \
    "v_fmac_f32_dpp %8, %8, " F " " DPP "\n\t"
__device__ __forceinline__ void seg_scan_fma9(float (&v)[9], const ScanPred& q) {
    if constexpr (kScanSteps == 2) {
        asm("s_nop 1\n\t"
            EMH_SCAN9("%9", "row_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1")
            EMH_SCAN9("%10", "row_shr:2 row_mask:0xf bank_mask:0xf bound_ctrl:1")
            : "+&v"(v[0]), "+&v"(v[1]), "+&v"(v[2]), "+&v"(v[3]), "+&v"(v[4]), "+&v"(v[5]), "+&v"(v[6]), "+&v"(v[7]),
              "+&v"(v[8])
            : "v"(q.f[0]), "v"(q.f[1]));
    } else if constexpr (kScanSteps == 3) {
        asm("s_nop 1\n\t"
            EMH_SCAN9("%9", "row_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1")
            EMH_SCAN9("%10", "row_shr:2 row_mask:0xf bank_mask:0xf bound_ctrl:1")
            EMH_SCAN9("%11", "row_shr:4 row_mask:0xf bank_mask:0xf bound_ctrl:1")
            : "+&v"(v[0]), "+&v"(v[1]), "+&v"(v[2]), "+&v"(v[3]), "+&v"(v[4]), "+&v"(v[5]), "+&v"(v[6]), "+&v"(v[7]),
              "+&v"(v[8])
            : "v"(q.f[0]), "v"(q.f[1]), "v"(q.f[2]));
    } else if constexpr (kScanSteps == 4) {
        asm("s_nop 1\n\t"
            EMH_SCAN9("%9", "row_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1")
            EMH_SCAN9("%10", "row_shr:2 row_mask:0xf bank_mask:0xf bound_ctrl:1")
            EMH_SCAN9("%11", "row_shr:4 row_mask:0xf bank_mask:0xf bound_ctrl:1")
            EMH_SCAN9("%12", "row_shr:8 row_mask:0xf bank_mask:0xf bound_ctrl:1")
            : "+&v"(v[0]), "+&v"(v[1]), "+&v"(v[2]), "+&v"(v[3]), "+&v"(v[4]), "+&v"(v[5]), "+&v"(v[6]), "+&v"(v[7]),
              "+&v"(v[8])
            : "v"(q.f[0]), "v"(q.f[1]), "v"(q.f[2]), "v"(q.f[3]));
    } else {
        asm("s_nop 1\n\t"
            EMH_SCAN9("%9", "row_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1")
            EMH_SCAN9("%10", "row_shr:2 row_mask:0xf bank_mask:0xf bound_ctrl:1")
            EMH_SCAN9("%11", "row_shr:4 row_mask:0xf bank_mask:0xf bound_ctrl:1")
            EMH_SCAN9("%12", "row_shr:8 row_mask:0xf bank_mask:0xf bound_ctrl:1")
            EMH_SCAN9("%13", "row_bcast:15 row_mask:0xa bank_mask:0xf")
            EMH_SCAN9("%14", "row_bcast:31 row_mask:0xc bank_mask:0xf")
            : "+&v"(v[0]), "+&v"(v[1]), "+&v"(v[2]), "+&v"(v[3]), "+&v"(v[4]), "+&v"(v[5]), "+&v"(v[6]), "+&v"(v[7]),
              "+&v"(v[8])
            : "v"(q.f[0]), "v"(q.f[1]), "v"(q.f[2]), "v"(q.f[3]), "v"(q.f[4]), "v"(q.f[5]));
    }
}
#undef EMH_SCAN9
template <bool FMA, int N>
__device__ __forceinline__ void seg_scan(float (&v)[N], const ScanPred& q) {
    if constexpr (FMA && N == 9) {
        seg_scan_fma9(v, q);
        return;
    }
#pragma unroll
    for (int i = 0; i < N; ++i) {
        float x = v[i];
        x = scan_step<FMA, 0x111, 0xf>(x, q.p[0], q.f[0]);  // row_shr:1
        x = scan_step<FMA, 0x112, 0xf>(x, q.p[1], q.f[1]);  // row_shr:2
        if constexpr (kScanSteps >= 3) x = scan_step<FMA, 0x114, 0xf>(x, q.p[2], q.f[2]);  // row_shr:4
        if constexpr (kScanSteps >= 4) x = scan_step<FMA, 0x118, 0xf>(x, q.p[3], q.f[3]);  // row_shr:8
        if constexpr (kScanSteps == 6) {
            x = scan_step<FMA, 0x142, 0xa>(x, q.p[4], q.f[4]);  // row_bcast:15 into rows 1 and 3
            x = scan_step<FMA, 0x143, 0xc>(x, q.p[5], q.f[5]);  // row_bcast:31 into rows 2 and 3
        }
        v[i] = x;
    }
}
// Run structure of the base cells (bx, by) across the wave; every lane must
// call.  Invalid lanes pass a cell no valid lane can have (bx = INT_MIN + lane).
__device__ __forceinline__ SegMask run_mask(int bx, int by) {
    SegMask m;
    m.lane = lane_id();
    const int pbx = __shfl_up(bx, 1, 64), pby = __shfl_up(by, 1, 64);
    const bool head = (m.lane == 0) || pbx != bx || pby != by;
    m.heads = __ballot(head) | kRowHeads;
    return m;
}
__device__ __forceinline__ bool run_tail(const SegMask& m) {
    return m.lane == 63 || ((m.heads >> (m.lane + 1)) & 1ull);
}

// ------------------------------------------------------------- chunk box
// LDS box of grid cells: [ox, ox + w) x [oy, oy + h), cell (cx, cy) at
// lo + (cx - ox) * h + (cy - oy) of the LDS arrays (lo: 0, or the offset of
// one slice's box among the boxes of a chunk's slices).  `lds` is false when
// the box exceeds the budget (sparse particles); those read and scatter the
// global grid directly.
struct Box {
    int ox, oy, w, h;
    int lo;
    bool lds;
    __device__ __forceinline__ int index(int cx, int cy) const { return lo + (cx - ox) * h + (cy - oy); }
    // LDS index of cell (bx - 1, by - 1) when the 3x3 stencil at base cell
    // (bx, by) (cells bx-1..bx+1, by-1..by+1) lies in the box, else -1;
    // stencil cell (gx, gy) is then at + gx * h + gy.  The test 0 <= bx - 1 -
    // ox <= w - 3 runs on the wrapped 32-bit difference: the true difference
    // of any int bx spans less than 2^32 values, so wrapping maps it one to
    // one, with the negative ones above every w - 3 (an LDS box has
    // w, h >= 3: margin >= 1).
    __device__ __forceinline__ int stencil_at(int bx, int by) const {
        const uint32_t lx = (uint32_t)bx - 1u - (uint32_t)ox, ly = (uint32_t)by - 1u - (uint32_t)oy;
        return (lds && lx <= (uint32_t)(w - 3) && ly <= (uint32_t)(h - 3)) ? lo + (int)(lx * (uint32_t)h + ly) : -1;
    }
};
__device__ Box block_box(int mnx, int mny, int mxx, int mxy, int margin, int* s_box) {
    mnx = wave_imin(mnx);
    mny = wave_imin(mny);
    mxx = wave_imax(mxx);
    mxy = wave_imax(mxy);
    const uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
    if (lane == 0) {
        s_box[4 * wave + 0] = mnx;
        s_box[4 * wave + 1] = mny;
        s_box[4 * wave + 2] = mxx;
        s_box[4 * wave + 3] = mxy;
    }
    __syncthreads();
    for (int w = 0; w < kWaves; ++w) {
        mnx = min(mnx, s_box[4 * w + 0]);
        mny = min(mny, s_box[4 * w + 1]);
        mxx = max(mxx, s_box[4 * w + 2]);
        mxy = max(mxy, s_box[4 * w + 3]);
    }
    __syncthreads();
    Box b;
    const long long w = (long long)mxx - (long long)mnx + 1 + 2 * margin;
    const long long h = (long long)mxy - (long long)mny + 1 + 2 * margin;
    b.lds = (mnx <= mxx) && (long long)mnx - margin > INT_MIN / 2 && (long long)mxx + margin < INT_MAX / 2 &&
            (long long)mny - margin > INT_MIN / 2 && (long long)mxy + margin < INT_MAX / 2 && w * h <= kLdsCells;
    b.ox = b.lds ? mnx - margin : 0;
    b.oy = b.lds ? mny - margin : 0;
    b.w = b.lds ? (int)w : 0;
    b.h = b.lds ? (int)h : 0;
    b.lo = 0;
    return b;
}


// P2G contribution of one particle to the 3x3 stencil, one channel at a time
// (p2g.rs:45-68): mass, the momentum m (v + C d) + (-4 V0 dt') (tau d) per
// axis, and the stress moment -4 V0 (tau d) per axis that the grid update
// multiplies by dt - dt'.  The momentum is linear in the cell offset d, so it
// is formed as u + K d with u = m v and K = m C + (-4 V0 dt') tau folded once
// per particle: per cell and axis two products and two sums before the
// weight, instead of the reference's three matrix-vector terms.  The offsets
// d = (cell - x) + 1/2 are exact in f32 (Sterbenz), three per axis.
struct P2GTerms {
    Wts w;
    float x, y, mass;
    float ux, uy;  // m v
    M2 k;          // m C + coeff dt' tau   (coeff = -4 V0)
    M2 st;         // coeff tau
};
template <int CH>
__device__ __forceinline__ void p2g_channel(const P2GTerms& t, float (&v)[9]) {
    float dx[3], dy[3];
#pragma unroll
    for (int g = 0; g < 3; ++g) {
        dx[g] = ((float)(t.w.bx + g - 1) - t.x) + 0.5f;
        dy[g] = ((float)(t.w.by + g - 1) - t.y) + 0.5f;
    }
#pragma unroll
    for (int gx = 0; gx < 3; ++gx)
#pragma unroll
        for (int gy = 0; gy < 3; ++gy) {
            const float weight = t.w.wx[gx] * t.w.wy[gy];
            float r;
            if (CH == kAccM) {
                r = weight * t.mass;
            } else if (CH == kAccPx) {
                r = weight * (t.ux + (t.k.a * dx[gx] + t.k.c * dy[gy]));
            } else if (CH == kAccPy) {
                r = weight * (t.uy + (t.k.b * dx[gx] + t.k.d * dy[gy]));
            } else if (CH == kAccSx) {
                r = weight * (t.st.a * dx[gx] + t.st.c * dy[gy]);
            } else {
                r = weight * (t.st.b * dx[gx] + t.st.d * dy[gy]);
            }
            v[gx * 3 + gy] = r;
        }
}
// Segmented wave sum of one channel, then the run tails add it to the LDS box
// (or to the global grid when the stencil leaves the box).  Lanes that do not
// scatter carry zero terms (scatter_terms), so their values are zero.
// at: the stencil's LDS index (Box::stencil_at), -1 for the global grid.
// The accumulator set a P2G scatters into, for its global-memory fallback:
// channel ch of acc[set], or the density pass's mass grid dm[set].  The
// pointer is read where a stencil outside the LDS box needs it, through an
// opaque index, instead of being held in SGPRs through the particle loop.
struct GridDst {
    uint32_t set;
    uint32_t mass;
};
__device__ __forceinline__ float* grid_dst(const Pipeline& p, GridDst d, int ch) {
    const uint32_t s = opaque_index(d.set);
    return d.mass ? p.grid.dm[s] : p.grid.acc[s][ch];
}

template <bool FMA, int CH>
__device__ __forceinline__ void scatter_channel(const P2GTerms& t, bool tail, const ScanPred& sq, const Box& box, int at,
                                                double* s_ch, const Pipeline& p, GridDst gd, int R) {
    float v[9];
    p2g_channel<CH>(t, v);
    seg_scan<FMA>(v, sq);
    if (tail) {
        const int bx = t.w.bx, by = t.w.by;
        if (at >= 0) {
            double* c = s_ch + at;
#pragma unroll
            for (int gx = 0; gx < 3; ++gx)
#pragma unroll
                for (int gy = 0; gy < 3; ++gy) latomic_add(&c[gx * box.h + gy], (double)v[gx * 3 + gy]);
        } else {
            float* g_ch = grid_dst(p, gd, CH);
#pragma unroll
            for (int gx = 0; gx < 3; ++gx)
#pragma unroll
                for (int gy = 0; gy < 3; ++gy) {
                    const int cx = bx + gx - 1, cy = by + gy - 1;
                    if (cx >= 0 && cy >= 0 && cx < R && cy < R)
                        gatomic_add(&g_ch[(uint32_t)cx * (uint32_t)R + (uint32_t)cy], v[gx * 3 + gy]);
                }
        }
    }
}
// Terms as scattered: a lane that does not scatter (no particle, or a
// sleeping one) gets zero weights, position and terms, so that every stencil
// value it forms is exactly zero, whatever its state held.
__device__ __forceinline__ P2GTerms scatter_terms(const P2GTerms& t, bool valid) {
    P2GTerms z = t;
    if (!valid) {
#pragma unroll
        for (int g = 0; g < 3; ++g) z.w.wx[g] = z.w.wy[g] = 0.0f;
        z.x = z.y = z.mass = z.ux = z.uy = 0.0f;
        z.k = M2{0.0f, 0.0f, 0.0f, 0.0f};
        z.st = M2{0.0f, 0.0f, 0.0f, 0.0f};
    }
    return z;
}

// Workgroups are dealt to the 8 XCDs round-robin.  Chunks go to them in
// groups of kXcdGroup consecutive chunks (a compact patch of cells along the
// key curve, whose grid cells its workgroups then share in one XCD's L2), the
// groups dealt round-robin, so that every XCD gets patches from all over the
// scene: a launch ends with its slowest XCD, and one contiguous 1/8 of the
// scene each would give one XCD all of a flowing region and another a resting
// one.  Measured on the 1M-particle sand column (500 steps): 84.8 us per
// substep in groups of 4, 84.4 in groups of 16, 91.1 with one contiguous run
// per XCD; one chunk at a time (no groups) loses the L2 sharing (the
// 4-per-cell interior: 78.0 against 70.3 us).  Workgroups beyond the last
// whole round of groups keep their own chunk.
constexpr uint32_t kXcdGroup = 4;
__device__ __forceinline__ uint32_t xcd_chunk(uint32_t b, uint32_t nb) {
    const uint32_t m = nb / (8u * kXcdGroup) * (8u * kXcdGroup);
    if (b >= m) return b;
    const uint32_t x = b & 7u, i = b >> 3;
    return ((i / kXcdGroup) * 8u + x) * kXcdGroup + i % kXcdGroup;
}

// Block reduction of the CFL contribution, one atomic pair per workgroup.
// f32::max / f32::min ignore NaN (as cfl_bound's fold does, step.rs:445-466);
// a NaN block result means every lane was NaN and contributes nothing.
__device__ void reduce_cfl(CtrlRed* red, float speed, float mdt, float* s_red) {
    block_maxmin(speed, mdt, s_red);
    if (threadIdx.x == 0) {
        if (!__builtin_isnan(speed) && speed > 0.0f) atomicMax(&red->max_speed, __float_as_uint(speed));
        if (!__builtin_isnan(mdt) && mdt < kInf) atomicMin(&red->min_mdt, __float_as_uint(mdt));
    }
}

// ------------------------------------------------------------------ kernels

__global__ void k_zero_all(const Pipeline p) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t R2 = p.cfg.grid_res * p.cfg.grid_res;
    if (i < R2) {
        for (int s = 0; s < 3; ++s) {
            for (int ch = 0; ch < kAccChannels; ++ch) p.grid.acc[s][ch][i] = 0.0f;
            p.grid.dm[s][i] = 0.0f;
        }
    }
}

// Control launch `seq` that queues a step start (remaining = config.dt, no
// substeps yet: step.rs:24-33, applied by the substep launch that begins it)
// and clears that step's counters; the CFL reduction carries over.
__device__ __forceinline__ void begin_step_body(const Pipeline& p, uint64_t seq, uint32_t reorder) {
    Ctrl* c = p.ctrl;
    CtrlState N = c->st[seq & 1u];
    const uint32_t step = N.step + N.pending + 1u;  // the step this marker starts
    N.pending += 1u;
    N.reorder |= reorder;
    c->st[(seq + 1u) & 1u] = N;
    c->red[seq_mod3(seq)] = c->red[seq_mod3(seq + 2u)];
    reset_red(c->red[seq_mod3(seq + 1u)]);
    c->clamp_count[step % kStepRing] = 0;
    c->proj_count[step % kStepRing] = 0;
}
__global__ void k_begin_step(const Pipeline p, uint64_t seq, uint32_t reorder) { begin_step_body(p, seq, reorder); }

// ------------------------------------------------- stable cell-key sort
// LSD radix sort of the cell keys (SortPlan): every pass is stable, so the
// result is the order std::stable_sort gives -- by key, then by slot.
//   count   per (digit, workgroup) counts of one pass (pass 0 also computes
//           the keys, from the stored and, when flagged, projected positions);
//   scan    one wave per digit: exclusive scan of the digit's counts over
//           the workgroups in place, and the digit's total;
//   scatter each key's position = the totals of the lower digits + the
//           digit's start in its workgroup + its rank among the keys of the
//           workgroup with the same digit and a lower slot.
// A workgroup's kSortTile keys are 64 x kSortItems rows; wave w takes the
// kSortItems consecutive rows from w * kSortItems, so wave order, row order
// and lane order together are slot order.  A key's rank among the equal
// digits of its row comes from `bits` ballots (the lanes agreeing on every
// bit), its rank among the wave's earlier rows from a wave-private LDS count
// that the row's lowest such lane advances: no atomics, no arrival order.
// Plain values only: a pointer into a kernel argument (say &p.cfg) would make
// the compiler copy the whole argument to every lane's scratch.
struct SortJob {
    const float2* x;         // positions whose cell keys pass 0 computes
    uint32_t project;        // clamp those positions to [proj_lo, proj_hi] first
    float proj_lo, proj_hi;
    const uint32_t* mat;    // material words (model segments only)
    const DevMat* mats;     // material table (model segments only)
    uint32_t n_mats;
    uint32_t n, grid_res, key_order, n_bins;
    uint32_t cell_bits, key_bits;  // key = model << cell_bits | cell key when s.model_bits
    SortBuf s;
};

constexpr uint32_t kSortBuckets = 1u << kSortMaxDigit;

__device__ __forceinline__ uint64_t lanes_below() {
    const uint32_t lane = lane_id();
    return lane == 0u ? 0ull : (~0ull >> (64u - lane));
}

// Digit width and shift of pass `pass` of a key of plan `pl`.
__device__ __forceinline__ void pass_digit(const SortPlan& pl, uint32_t pass, uint32_t& shift, uint32_t& bits) {
    shift = pass * pl.digit;
    bits = min(pl.digit, pl.key_bits - shift);
}

// Lanes of the wave with this lane's digit (`bits` low bits), among the valid ones.
__device__ __forceinline__ uint64_t digit_peers(uint32_t digit, uint32_t bits, bool valid) {
    uint64_t peers = __ballot(valid);
    for (uint32_t b = 0; b < bits; ++b) {
        const bool bit = ((digit >> b) & 1u) != 0u;
        const uint64_t m = __ballot(bit);
        peers &= bit ? m : ~m;
    }
    return peers;
}

// Buffer k of a pair, by select: a runtime index into a SortJob built in
// registers would send the whole job to scratch.
__device__ __forceinline__ uint32_t* pick(uint32_t* const (&buf)[2], uint32_t k) { return k ? buf[1] : buf[0]; }

// Key of slot i as pass `pass` reads it.
__device__ __forceinline__ uint32_t sort_key_in(const SortJob& J, uint32_t pass, uint32_t i) {
    if (pass > 0) return pick(J.s.kbuf, (pass - 1u) & 1u)[i];
    float2 x = J.x[i];
    if (J.project) project_x(x.x, x.y, J.proj_lo, J.proj_hi);
    const uint32_t key = cell_key(sat_i32(floorf(x.x)), sat_i32(floorf(x.y)), J.grid_res, J.key_order);
    if (!J.s.model_bits) return key;
    const uint32_t slot = J.mat[i] & kMatMask;  // the slot mat_of reads (slot 0 for an unregistered id)
    const uint32_t model = min(J.mats[slot < J.n_mats ? slot : 0u].model, (1u << kModelKeyBits) - 1u);
    return (model << J.cell_bits) | key;
}

// Ranks of workgroup blk's keys on the digit of `pass`.  On return s_cnt[w][d]
// holds the number of digit-d keys in waves before w plus `add(d)` (the
// scatter's start slot, 0 for the count), loc[k] the key's rank among the
// equal digits before it in its wave, key[k] the key (0 beyond n); the
// return value of each thread's buckets d = tid, tid + 256, ... is their
// total in the workgroup, written by `total(d, t)`.
template <class Add, class Total>
__device__ __forceinline__ void sort_rank_tile(const SortJob& J, uint32_t pass, uint32_t blk, uint32_t shift,
                                               uint32_t bits, uint32_t (*s_cnt)[kSortBuckets],
                                               uint32_t (&key)[kSortItems], uint32_t (&loc)[kSortItems], Add add,
                                               Total total) {
    const uint32_t lane = lane_id(), w = threadIdx.x >> 6;
    const uint32_t buckets = 1u << bits, mask = buckets - 1u;
    for (uint32_t d = lane; d < buckets; d += 64u) s_cnt[w][d] = 0u;
    __syncthreads();
    const uint64_t below = lanes_below();
    const uint32_t row0 = blk * kSortTile + w * (uint32_t)kSortItems * 64u;
#pragma unroll
    for (int k = 0; k < kSortItems; ++k) {
        const uint32_t i = row0 + (uint32_t)k * 64u + lane;
        const bool valid = i < J.n;
        key[k] = valid ? sort_key_in(J, pass, i) : 0u;
        if (pass == 0 && valid && J.s.key) J.s.key[i] = key[k];
        const uint32_t d = (key[k] >> shift) & mask;
        const uint64_t peers = digit_peers(d, bits, valid);
        const uint32_t before = (uint32_t)__popcll(peers & below);
        uint32_t b = 0u;
        if (valid) b = s_cnt[w][d];
        // the read above is issued before the write below; a wave's LDS
        // operations complete in order, so every lane reads the old count
        if (valid && before == 0u) s_cnt[w][d] = b + (uint32_t)__popcll(peers);
        loc[k] = b + before;
    }
    __syncthreads();
    for (uint32_t d = threadIdx.x; d < buckets; d += blockDim.x) {
        const uint32_t a = add(d);
        uint32_t run = a;
#pragma unroll
        for (int v = 0; v < kWaves; ++v) {
            const uint32_t c = s_cnt[v][d];
            s_cnt[v][d] = run;
            run += c;
        }
        total(d, run - a);
    }
    __syncthreads();
}

__device__ __forceinline__ void sort_count_body(const SortJob& J, uint32_t pass, uint32_t blk, uint32_t nblk) {
    __shared__ uint32_t s_cnt[kWaves][kSortBuckets];
    const SortPlan pl = sort_plan_bits(J.key_bits);
    if (pass >= pl.passes) return;
    uint32_t shift, bits;
    pass_digit(pl, pass, shift, bits);
    uint32_t key[kSortItems], loc[kSortItems];
    sort_rank_tile(J, pass, blk, shift, bits, s_cnt, key, loc, [](uint32_t) { return 0u; },
                   [&](uint32_t d, uint32_t t) { J.s.hist[(size_t)d * nblk + blk] = t; });
}

// The digit totals of a pass follow its (1 << digit) x nblk counts.
__device__ __forceinline__ uint32_t* sort_totals(const SortJob& J, const SortPlan& pl, uint32_t nblk) {
    return J.s.hist + ((size_t)1 << pl.digit) * nblk;
}

// Scan group g of a pass: wave w scans the counts of digit g * kSortRowsPerGroup
// + w over the workgroups (exclusive, in place) and writes the digit's total.
// Waves past the pass's digits return at once (the body has no barrier).
__device__ __forceinline__ void sort_scan_body(const SortJob& J, uint32_t pass, uint32_t g) {
    const SortPlan pl = sort_plan_bits(J.key_bits);
    if (pass >= pl.passes) return;
    uint32_t shift, bits;
    pass_digit(pl, pass, shift, bits);
    const uint32_t d = g * kSortRowsPerGroup + (threadIdx.x >> 6);
    if (d >= (1u << bits)) return;
    const uint32_t nblk = sort_tiles(J.n), lane = lane_id();
    uint32_t* row = J.s.hist + (size_t)d * nblk;
    constexpr uint32_t kPerLane = 8;  // 512 workgroups per round: all of them up to 1M keys
    uint32_t carry = 0;
    for (uint32_t base = 0; base < nblk; base += 64u * kPerLane) {
        const uint32_t t0 = base + lane * kPerLane;
        uint32_t v[kPerLane], sum = 0;
#pragma unroll
        for (uint32_t j = 0; j < kPerLane; ++j) {
            v[j] = t0 + j < nblk ? row[t0 + j] : 0u;
            sum += v[j];
        }
        const uint32_t incl = wave_incl_scan(sum);
        uint32_t run = carry + incl - sum;
#pragma unroll
        for (uint32_t j = 0; j < kPerLane; ++j) {
            if (t0 + j < nblk) row[t0 + j] = run;
            run += v[j];
        }
        carry += __shfl(incl, 63, 64);
    }
    if (lane == 0) sort_totals(J, pl, nblk)[d] = carry;
}

__device__ __forceinline__ void sort_scatter_body(const SortJob& J, uint32_t pass, uint32_t blk, uint32_t nblk) {
    __shared__ uint32_t s_cnt[kWaves][kSortBuckets];
    const SortPlan pl = sort_plan_bits(J.key_bits);
    if (pass >= pl.passes) return;
    uint32_t shift, bits;
    pass_digit(pl, pass, shift, bits);
    const bool last = pass + 1u == pl.passes;
    // the first slot of each digit: exclusive scan of the digit totals
    __shared__ uint32_t s_base[kSortBuckets];
    __shared__ uint32_t s_scan[17];
    static_assert(kBlock * 4 == kSortBuckets, "four digit totals per thread");
    {
        const uint32_t* tot = sort_totals(J, pl, nblk);
        const uint32_t d0 = threadIdx.x * 4u, buckets = 1u << bits;
        uint32_t v[4], sum = 0, total;
#pragma unroll
        for (uint32_t j = 0; j < 4; ++j) {
            v[j] = d0 + j < buckets ? tot[d0 + j] : 0u;
            sum += v[j];
        }
        uint32_t run = block_excl_scan(sum, s_scan, total);
#pragma unroll
        for (uint32_t j = 0; j < 4; ++j) {
            if (d0 + j < buckets) s_base[d0 + j] = run;
            run += v[j];
        }
        // block_excl_scan ends on a barrier and sort_rank_tile barriers
        // before its first read of s_base
    }
    uint32_t key[kSortItems], loc[kSortItems];
    const uint32_t* hist = J.s.hist;
    sort_rank_tile(J, pass, blk, shift, bits, s_cnt, key, loc,
                   [&](uint32_t d) { return s_base[d] + hist[(size_t)d * nblk + blk]; }, [](uint32_t, uint32_t) {});
    const uint32_t lane = lane_id(), w = threadIdx.x >> 6, mask = (1u << bits) - 1u;
    const uint32_t row0 = blk * kSortTile + w * (uint32_t)kSortItems * 64u;
    const uint32_t* vin = pass > 0 ? pick(J.s.vbuf, (pass - 1u) & 1u) : nullptr;
    uint32_t* kout = pick(J.s.kbuf, pass & 1u);
    uint32_t* vout = pick(J.s.vbuf, pass & 1u);
#pragma unroll
    for (int k = 0; k < kSortItems; ++k) {
        const uint32_t i = row0 + (uint32_t)k * 64u + lane;
        if (i >= J.n) continue;
        const uint32_t pos = s_cnt[w][(key[k] >> shift) & mask] + loc[k];
        const uint32_t val = vin ? vin[i] : i;
        if (!last) {
            kout[pos] = key[k];
            vout[pos] = val;
        } else if (J.s.perm) {
            J.s.out[pos] = val;
        } else {
            J.s.out[val] = pos;
        }
    }
}

// The pipeline's reorder: keys of the stored state of control launch `seq`
// (the launch that will move the particles), projected when flagged;
// s.out = each slot's position in cell-key order.
__device__ __forceinline__ SortJob pipeline_sort_job(const Pipeline& p, uint64_t seq) {
    const Ctrl* c = p.ctrl;
    const bool project = p.cfg.project && c->red[seq_mod3(seq + 2u)].dirty;
    SortJob J;
    const PBuf b = sel_buf(p, c->st[seq & 1u].cur_buf);
    J.x = b.x;
    J.project = project ? 1u : 0u;
    J.proj_lo = p.cfg.proj_lo;
    J.proj_hi = p.cfg.proj_hi;
    J.mat = b.mat;
    J.mats = p.mats;
    J.n_mats = p.n_mats;
    J.n = p.n;
    J.grid_res = p.cfg.grid_res;
    J.key_order = p.cfg.key_order;
    J.n_bins = p.cfg.n_bins;
    J.cell_bits = cell_key_bits(p.cfg.n_bins);
    J.key_bits = J.cell_bits + p.sort.model_bits;
    J.s = p.sort;
    return J;
}
__global__ __launch_bounds__(kBlock) void k_sort_count(const Pipeline p, uint64_t seq, uint32_t pass) {
    sort_count_body(pipeline_sort_job(p, seq), pass, blockIdx.x, gridDim.x);
}
__global__ __launch_bounds__(kBlock) void k_sort_scan(const Pipeline p, uint64_t seq, uint32_t pass) {
    sort_scan_body(pipeline_sort_job(p, seq), pass, blockIdx.x);
}
__global__ __launch_bounds__(kBlock) void k_sort_scatter(const Pipeline p, uint64_t seq, uint32_t pass) {
    sort_scatter_body(pipeline_sort_job(p, seq), pass, blockIdx.x, gridDim.x);
}

constexpr int kModeG2P = 1;   // gather + update the stored state (a substep)
constexpr int kModeP2G = 2;   // scatter mass, momentum and stress moment
constexpr int kModeMass = 4;  // scatter mass only (density re-estimate)
constexpr int kModeHarvest = 8;     // record NACC trial states (training data of the surrogate)
constexpr int kModeSurrogate = 16;  // NACC return mapping through the learned surrogate
constexpr int kModeSleep = 32;      // sleep partition on (sleeping particles are skipped)

// wake_particle's test (step.rs:107-131): a stencil cell of the sleeping
// particle holds mass from this substep's P2G (the reference asks whether P2G
// touched the cell; the two differ only for a stencil weight of exactly 0).
__device__ __forceinline__ bool stencil_touched(const Pipeline& p, uint32_t set, float x, float y) {
    const Wts w = weights(x, y);
    const int R = (int)p.cfg.grid_res;
    const float* m = p.grid.acc[set][kAccM];
    bool hit = false;
#pragma unroll
    for (int gx = 0; gx < 3; ++gx)
#pragma unroll
        for (int gy = 0; gy < 3; ++gy) {
            const int cx = w.bx + gx - 1, cy = w.by + gy - 1;
            if (cx >= 0 && cy >= 0 && cx < R && cy < R) hit = hit || m[(uint32_t)cx * (uint32_t)R + (uint32_t)cy] > 0.0f;
        }
    return hit;
}
// Sleep scoring (step.rs:299-315): a particle without activation whose
// speed^2 is below threshold^2 falls asleep at the end of the substep.
__device__ __forceinline__ void sleep_test(float sleep_sq, PState& s, bool& asleep, bool& changed) {
    if (sleep_sq > 0.0f && !(s.mat & kNoSleepBit) && s.vx * s.vx + s.vy * s.vy < sleep_sq) {
        s.mat |= kSleepBit;
        asleep = true;
        changed = true;
    }
}

// Bounding box of the chunk's base cells (positions projected when the state
// is flagged), widened by `margin` cells.
__device__ __forceinline__ Box chunk_box(const Pipeline& p, const PBuf& a, uint32_t base, bool project, int margin,
                                         int* s_box) {
    int mnx = INT_MAX, mny = INT_MAX, mxx = INT_MIN, mxy = INT_MIN;
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const uint32_t j = base + k * kBlock + threadIdx.x;
        if (j < p.n) {
            float2 x = a.x[j];
            if (project) project_x(x.x, x.y, p.cfg);
            const int bx = sat_i32(floorf(x.x)), by = sat_i32(floorf(x.y));
            mnx = min(mnx, bx);
            mny = min(mny, by);
            mxx = max(mxx, bx);
            mxy = max(mxy, by);
        }
    }
    return block_box(mnx, mny, mxx, mxy, margin, s_box);
}

// Sort bin of a base cell among nb bins for particles that scatter through
// global memory: a hash of the cell, so that the particles of one cell end up
// side by side (one run, one set of atomics per run) whatever the cells'
// spread; cells that share a bin only interleave.
__device__ __forceinline__ uint32_t hash_bin(int cx, int cy, uint32_t nb) {
    const uint32_t h = ((uint32_t)cx * 0x9E3779B1u) ^ ((uint32_t)cy * 0x85EBCA77u);
    return (h ^ (h >> 15)) % nb;
}
// Bins of a slice that runs through global memory inside a chunk's slice
// boxes (the other slices' boxes permitting).
constexpr int kGlobalSliceBins = 12;

// Boxes of a chunk's kPer slices (slice k: slots k * kBlock .. + kBlock, the
// particles of round k in stored order), one after another in the LDS
// budget: a slice whose own box fits takes its w * h cells at offset lo; one
// whose box does not fit runs through global memory and keeps
// kGlobalSliceBins cells as sort bins (h holds their count; one for a slice
// without particles).  Returns the cells used;
// when that is within the budget the chunk stages, sorts and flushes all its
// slices' boxes at once, like a chunk with one box, instead of running slice
// by slice.  s_box: 4 * kPer * kWaves ints; s_mbox receives the boxes.
__device__ __forceinline__ int slice_boxes(const Pipeline& p, const PBuf& a, uint32_t base, bool project, int margin,
                                           int* s_box, Box* s_mbox) {
    const uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const uint32_t j = base + k * kBlock + threadIdx.x;
        int bx = INT_MAX, by = INT_MAX, ex = INT_MIN, ey = INT_MIN;
        if (j < p.n) {
            float2 x = a.x[j];
            if (project) project_x(x.x, x.y, p.cfg);
            bx = ex = sat_i32(floorf(x.x));
            by = ey = sat_i32(floorf(x.y));
        }
        bx = wave_imin(bx);
        by = wave_imin(by);
        ex = wave_imax(ex);
        ey = wave_imax(ey);
        if (lane == 0) {
            int* o = s_box + 4 * (k * kWaves + (int)wave);
            o[0] = bx;
            o[1] = by;
            o[2] = ex;
            o[3] = ey;
        }
    }
    __syncthreads();
    int lo = 0;
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        int mnx = INT_MAX, mny = INT_MAX, mxx = INT_MIN, mxy = INT_MIN;
        for (int w = 0; w < kWaves; ++w) {
            const int* o = s_box + 4 * (k * kWaves + w);
            mnx = min(mnx, o[0]);
            mny = min(mny, o[1]);
            mxx = max(mxx, o[2]);
            mxy = max(mxy, o[3]);
        }
        const long long w = (long long)mxx - (long long)mnx + 1 + 2 * margin;
        const long long h = (long long)mxy - (long long)mny + 1 + 2 * margin;
        Box b;
        b.lds = (mnx <= mxx) && (long long)mnx - margin > INT_MIN / 2 && (long long)mxx + margin < INT_MAX / 2 &&
                (long long)mny - margin > INT_MIN / 2 && (long long)mxy + margin < INT_MAX / 2 && w * h <= kLdsCells;
        b.ox = b.lds ? mnx - margin : 0;
        b.oy = b.lds ? mny - margin : 0;
        b.w = b.lds ? (int)w : 0;
        b.h = b.lds ? (int)h : (mnx <= mxx ? kGlobalSliceBins : 1);
        b.lo = lo;
        lo += b.lds ? b.w * b.h : b.h;
        if (threadIdx.x == 0) s_mbox[k] = b;
    }
    __syncthreads();
    return lo;
}

// Processing order of a chunk: its nvalid particles sorted by base cell (a
// counting sort over the box cells in LDS, arrival order inside a cell), so
// that the lanes of a wave hold long runs of one cell however far the stored
// order has drifted since the last physical reorder.  Lane t's k-th particle
// is slot s_perm[k * kBlock + t] of the chunk.  s_cnt: kLdsCells scratch.
// A chunk keeps its stored order when that order's runs (cell changes between
// neighbouring lanes, counted per wave) are within 5/4 of what the sort would
// leave, one run per distinct cell and wave: right after a physical reorder,
// where particles barely move, and in sparse scenes whose cells hold few
// particles, the sort costs time and shortens nothing (1M-particle sand: 75.1
// us per substep never sorting, 68.7 sorting; a four-material scene of 4
// particles per cell 9 % faster never sorting than above 8 runs per wave).
// The distinct cells come from the sort's own counting pass: integer LDS
// atomics run at full rate on gfx950.
// With slice boxes (`mbox`, nbins cells in all), a particle's bin is its cell
// in its slice's box, or its slice's single bin when that slice runs through
// global memory; every slice holds kBlock particles but the last, so each
// slice's particles stay in the slots of its round.
__device__ __forceinline__ void chunk_order(const Pipeline& p, const PBuf& a, uint32_t base, bool project, const Box& box,
                                            const Box* mbox, int nbins, uint32_t nvalid, uint32_t* s_cnt,
                                            uint16_t* s_perm, uint32_t* s_scan) {
    if (!box.lds && !mbox) {  // the chunk runs slice by slice; each slice is sorted on its own (slice_order)
#pragma unroll
        for (int k = 0; k < kPer; ++k) s_perm[k * kBlock + threadIdx.x] = (uint16_t)(k * kBlock + threadIdx.x);
        __syncthreads();
        return;
    }
    __shared__ uint32_t s_runs, s_distinct;
    const int cells = mbox ? nbins : box.w * box.h;
    for (int i = threadIdx.x; i < cells; i += kBlock) s_cnt[i] = 0u;
    if (threadIdx.x == 0) s_runs = s_distinct = 0u;
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t cell[kPer], rank[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const uint32_t q = k * kBlock + threadIdx.x;
        const bool valid = q < nvalid;
        cell[k] = 0u;
        if (valid) {  // the base cell chunk_box saw (same projection)
            float2 x = a.x[base + q];
            if (project) project_x(x.x, x.y, p.cfg);
            const int cx = sat_i32(floorf(x.x)), cy = sat_i32(floorf(x.y));
            if (mbox) {
                const Box& b = mbox[k];
                cell[k] = (uint32_t)(b.lds ? b.index(cx, cy) : b.lo + (int)hash_bin(cx, cy, (uint32_t)b.h));
            } else {
                cell[k] = (uint32_t)box.index(cx, cy);
            }
        }
        // runs of the stored order: lanes whose cell differs from the lane before
        const uint32_t prev = (uint32_t)__shfl_up((int)cell[k], 1);
        const uint64_t starts = __ballot(valid && (lane == 0u || prev != cell[k]));
        if (lane == 0u && starts != 0ull) atomicAdd(&s_runs, (uint32_t)__popcll(starts));
        // one returning LDS atomic per particle: same-cell lanes serialise in
        // the LDS unit, with no per-cell loop of dependent round trips
        rank[k] = valid ? atomicAdd(&s_cnt[cell[k]], 1u) : 0u;
    }
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const uint64_t firsts = __ballot(k * kBlock + threadIdx.x < nvalid && rank[k] == 0u);
        if (lane == 0u && firsts != 0ull) atomicAdd(&s_distinct, (uint32_t)__popcll(firsts));
    }
    __syncthreads();
    if (4u * s_runs <= 5u * (s_distinct + (uint32_t)(kChunk / 64))) {
#pragma unroll
        for (int k = 0; k < kPer; ++k) s_perm[k * kBlock + threadIdx.x] = (uint16_t)(k * kBlock + threadIdx.x);
        __syncthreads();
        return;
    }
    constexpr int kCellsPerThread = kLdsCells / kBlock;
    uint32_t cnt[kCellsPerThread], sum = 0;
#pragma unroll
    for (int i = 0; i < kCellsPerThread; ++i) {
        const int idx = threadIdx.x * kCellsPerThread + i;
        cnt[i] = idx < cells ? s_cnt[idx] : 0u;
        sum += cnt[i];
    }
    uint32_t total;
    uint32_t off = block_excl_scan(sum, s_scan, total);
#pragma unroll
    for (int i = 0; i < kCellsPerThread; ++i) {
        const int idx = threadIdx.x * kCellsPerThread + i;
        if (idx < cells) s_cnt[idx] = off;
        off += cnt[i];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const uint32_t q = k * kBlock + threadIdx.x;
        if (q < nvalid) s_perm[s_cnt[cell[k]] + rank[k]] = (uint16_t)q;
    }
    __syncthreads();
}

// P2G terms of a (projected) particle state (p2g.rs:24-72, tau from
// transfer.rs:77-104); the stress coefficient as p2g.rs:42-44 forms it,
// ((-V) * 4) * dt', and without dt' for the correction moment.
template <int ONLY>
__device__ __forceinline__ P2GTerms p2g_terms(const Pipeline& p, const PState& s, float pred_dt) {
    P2GTerms t;
    t.w = weights(s.x, s.y);
    t.x = s.x;
    t.y = s.y;
    t.mass = s.mass;
    const DevMat& mt = mat_of(p, s.mat);
    const M2 tau = kirchhoff<ONLY>(mt, s);
    const float coeff = (-stress_volume<ONLY>(mt, s)) * kDInv;
    const float coeff_dt = coeff * pred_dt;
    t.ux = s.mass * s.vx;
    t.uy = s.mass * s.vy;
    t.k = madd(mscale(s.c, s.mass), mscale(tau, coeff_dt));
    t.st = mscale(tau, coeff);
    return t;
}

template <bool FMA, int NCH>
__device__ __forceinline__ void scatter_channels(const P2GTerms& z, bool tail, const ScanPred& sq, const Box& box, int at,
                                                 double (*s_acc)[kLdsCells], const Pipeline& p, GridDst gd, int R) {
    scatter_channel<FMA, kAccM>(z, tail, sq, box, at, s_acc[kAccM], p, gd, R);
    if constexpr (NCH > 1) {
        scatter_channel<FMA, kAccPx>(z, tail, sq, box, at, s_acc[kAccPx], p, gd, R);
        scatter_channel<FMA, kAccPy>(z, tail, sq, box, at, s_acc[kAccPy], p, gd, R);
    }
    if constexpr (NCH == kAccChannels) {
        scatter_channel<FMA, kAccSx>(z, tail, sq, box, at, s_acc[kAccSx], p, gd, R);
        scatter_channel<FMA, kAccSy>(z, tail, sq, box, at, s_acc[kAccSy], p, gd, R);
    }
}

// The first NCH channels, one at a time (keeps 9 partial sums live, not 45):
// all five, or mass and momentum alone when the dt the terms were formed with
// is the substep's own (the density pass).  Terms below 1e18 in magnitude
// keep every stencil value below 1e37 and every partial sum of 64 of them
// finite, so the fma form of the scan is exact.
template <int NCH = kAccChannels>
__device__ __forceinline__ void scatter_p2g(const P2GTerms& t, bool valid, const Box& box,
                                            double (*s_acc)[kLdsCells], const Pipeline& p, GridDst gd, int R) {
    static_assert(NCH == kAccChannels || NCH == kAccPy + 1 || NCH == 1, "all channels, mass and momentum, or mass");
    const SegMask sm = run_mask(valid ? t.w.bx : INT_MIN + (int)lane_id(), valid ? t.w.by : 0);
    const ScanPred sq = scan_pred(sm);
    float big = fmaxf(fabsf(t.mass), fmaxf(fabsf(t.ux), fabsf(t.uy)));
    big = fmaxf(big, fmaxf(fmaxf(fabsf(t.k.a), fabsf(t.k.b)), fmaxf(fabsf(t.k.c), fabsf(t.k.d))));
    big = fmaxf(big, fmaxf(fmaxf(fabsf(t.st.a), fabsf(t.st.b)), fmaxf(fabsf(t.st.c), fabsf(t.st.d))));
    const bool small = !valid || (big < 1e18f && !__builtin_isnan(t.mass + t.ux + t.uy + t.k.a + t.k.b + t.k.c + t.k.d +
                                                                  t.st.a + t.st.b + t.st.c + t.st.d));
    const bool fma_ok = __ballot(!small) == 0ull;
    const int at = box.stencil_at(t.w.bx, t.w.by);
    const bool tail = valid && run_tail(sm);
    const P2GTerms z = scatter_terms(t, valid);
    // one branch per particle round: the scans of the fast path feed the
    // atomics directly
    if (fma_ok)
        scatter_channels<true, NCH>(z, tail, sq, box, at, s_acc, p, gd, R);
    else
        scatter_channels<false, NCH>(z, tail, sq, box, at, s_acc, p, gd, R);
}

// Processing order of slice k of a sliced chunk (slots k*kBlock .. +kBlock),
// sorted by base cell inside the slice's box: s_perm[k*kBlock + r].
__device__ __forceinline__ void slice_order(const Pipeline& p, const PBuf& a, uint32_t base, bool project, const Box& box,
                                            uint32_t k, uint32_t nvalid, uint32_t* s_cnt, uint16_t* s_perm,
                                            uint32_t* s_scan) {
    const uint32_t q = k * kBlock + threadIdx.x;
    const bool valid = q < nvalid;
    // a slice without an LDS box is sorted by hashed cell (its P2G atomics go
    // to global memory once per run)
    const int cells = box.lds ? box.w * box.h : kLdsCells;
    for (int i = threadIdx.x; i < cells; i += kBlock) s_cnt[i] = 0u;
    __syncthreads();
    uint32_t cell = 0u, rank = 0u;
    if (valid) {
        float2 x = a.x[base + q];
        if (project) project_x(x.x, x.y, p.cfg);
        const int cx = sat_i32(floorf(x.x)), cy = sat_i32(floorf(x.y));
        cell = box.lds ? (uint32_t)box.index(cx, cy) : hash_bin(cx, cy, (uint32_t)kLdsCells);
        rank = atomicAdd(&s_cnt[cell], 1u);
    }
    __syncthreads();
    constexpr int kCellsPerThread = kLdsCells / kBlock;
    uint32_t cnt[kCellsPerThread], sum = 0;
#pragma unroll
    for (int i = 0; i < kCellsPerThread; ++i) {
        const int idx = threadIdx.x * kCellsPerThread + i;
        cnt[i] = idx < cells ? s_cnt[idx] : 0u;
        sum += cnt[i];
    }
    uint32_t total;
    uint32_t off = block_excl_scan(sum, s_scan, total);
#pragma unroll
    for (int i = 0; i < kCellsPerThread; ++i) {
        const int idx = threadIdx.x * kCellsPerThread + i;
        if (idx < cells) s_cnt[idx] = off;
        off += cnt[i];
    }
    __syncthreads();
    if (valid) s_perm[k * kBlock + s_cnt[cell] + rank] = (uint16_t)q;
    __syncthreads();
}

// Box of one kBlock-particle slice (lane's particle j when valid).
__device__ __forceinline__ Box slice_box(const Pipeline& p, const PBuf& a, uint32_t j, bool valid, bool project, int margin,
                                         int* s_box) {
    int mnx = INT_MAX, mny = INT_MAX, mxx = INT_MIN, mxy = INT_MIN;
    if (valid) {
        float2 x = a.x[j];
        if (project) project_x(x.x, x.y, p.cfg);
        mnx = mxx = sat_i32(floorf(x.x));
        mny = mxy = sat_i32(floorf(x.y));
    }
    return block_box(mnx, mny, mxx, mxy, margin, s_box);
}

// Clear the box's LDS accumulators and stage its updated grid velocity.
template <bool G2P, int NCH>
__device__ __forceinline__ void stage_box(const Pipeline& p, const Box& box, uint32_t cur, float dt, float ddt,
                                          double (*s_acc)[kLdsCells], float2* s_vel) {
    if (!box.lds) return;
    for (int i = threadIdx.x; i < box.w * box.h; i += kBlock) {
#pragma unroll
        for (int ch = 0; ch < NCH; ++ch) s_acc[ch][i] = 0.0;
        if (G2P) s_vel[i] = cell_velocity(p, cur, dt, ddt, box.ox + i / box.h, box.oy + i % box.h);
    }
    __syncthreads();
}

// Stage and clear every LDS slice box of a chunk (one barrier for all).
template <bool G2P, int NCH>
__device__ __forceinline__ void stage_boxes(const Pipeline& p, const Box* mbox, uint32_t cur, float dt, float ddt,
                                            double (*s_acc)[kLdsCells], float2* s_vel) {
#pragma unroll 1
    for (int k = 0; k < kPer; ++k) {
        const Box b = mbox[k];
        if (!b.lds) continue;
        for (int i = threadIdx.x; i < b.w * b.h; i += kBlock) {
#pragma unroll
            for (int ch = 0; ch < NCH; ++ch) s_acc[ch][b.lo + i] = 0.0;
            if (G2P) s_vel[b.lo + i] = cell_velocity(p, cur, dt, ddt, b.ox + i / b.h, b.oy + i % b.h);
        }
    }
    __syncthreads();
}

// Add the LDS box into the global grid: one atomic per touched cell and channel.
// (`sync`: barrier first, for the LDS adds of the whole workgroup)
template <int NCH>
__device__ __forceinline__ void flush_box(const Box& box, double (*s_acc)[kLdsCells], const Pipeline& p, GridDst gd,
                                          int R, bool sync = true) {
    if (!box.lds) return;
    if (sync) __syncthreads();
    float* g[NCH];
#pragma unroll
    for (int ch = 0; ch < NCH; ++ch) g[ch] = grid_dst(p, gd, ch);
    for (int i = threadIdx.x; i < box.w * box.h; i += kBlock) {
        const int cx = box.ox + i / box.h, cy = box.oy + i % box.h;
        if (cx < 0 || cy < 0 || cx >= R || cy >= R) continue;
        const uint32_t gi = (uint32_t)cx * (uint32_t)R + (uint32_t)cy;
#pragma unroll
        for (int ch = 0; ch < NCH; ++ch) {
            const float v = (float)s_acc[ch][box.lo + i];
            if (v != 0.0f) gatomic_add(&g[ch][gi], v);
        }
    }
}

// The substep kernel (MODE = G2P | P2G or G2P | Mass) and the prologue of a
// freshly uploaded state (MODE = P2G or Mass: no gather, no store).  ONLY:
// the one constitutive model every particle uses, or -1 for any mix.
// NACC trial state (F_trial as update_particle forms it, alpha and the
// material's parameters) of every stride-th offer into the harvest buffer;
// offers beyond its capacity are counted and dropped.
__device__ __forceinline__ void harvest_nacc(const SgDev& sg, const DevMat& m, const PState& s, float dt) {
    const bool want = sg.harvest != nullptr && m.model == kModelNacc;
    const uint64_t act = __ballot(want);
    if (act == 0ull) return;
    M2 fin = s.f;
    if (m.flags & kMatNaccAdvance) fin = mmul(madd(M2{1.0f, 0.0f, 0.0f, 1.0f}, mscale(s.c, dt)), fin);
    const uint32_t lane = lane_id();
    const uint32_t leader = (uint32_t)__ffsll((unsigned long long)act) - 1u;
    unsigned long long at = 0;
    if (lane == leader) at = atomicAdd(sg.harvest_count, (unsigned long long)__popcll(act));
    at = __shfl(at, (int)leader);
    if (!want) return;
    at += (unsigned long long)__popcll(act & ((1ull << lane) - 1ull));
    if (at % sg.harvest_stride != 0ull) return;
    at /= sg.harvest_stride;
    if (at >= sg.harvest_cap) return;
    float4* r = reinterpret_cast<float4*>(sg.harvest + at * 12ull);
    r[0] = make_float4(fin.a, fin.b, fin.c, fin.d);
    r[1] = make_float4(s.lvs, m.mu, m.na_kappa, m.na_m);
    r[2] = make_float4(m.na_beta, m.na_xi, __uint_as_float(m.flags), 0.0f);
}

// update_particle with the NACC return mapping through the learned
// surrogate (nacc.rs:255-269 otherwise unchanged).  Every lane of the
// workgroup calls this.
template <int ONLY>
__device__ __forceinline__ void surrogate_update(const SgDev& sg, const SgLds& net, const DevMat& m, bool valid,
                                                 PState& s, float dt, float* s_io) {
    const bool nacc = valid && m.model == kModelNacc;
    M2 fin = s.f;
    if (nacc && (m.flags & kMatNaccAdvance)) fin = mmul(madd(M2{1.0f, 0.0f, 0.0f, 1.0f}, mscale(s.c, dt)), fin);
    M2 nf;
    float na;
    sg_return<1>(m, sg, net, nacc, fin, s.lvs, nf, na, s_io);
    if (nacc) {
        s.f = nf;
        s.lvs = na;
        float j = rmax(mdet(nf), kMinJ);
        float vol = s.v0 * j;
        s.vol = vol;
        if (vol > kMinJ) s.dens = s.mass / vol;
    } else if (valid) {
        update_particle<ONLY>(m, s, dt);
    }
}

// One substep launch for one scene's chunks: blk is this workgroup's index
// among the nblk workgroups the scene's particles get (the whole grid of a
// single-scene launch; the scene's share of a batch launch, whose chunks
// are dealt without the XCD remap).
template <int MODE, int ONLY, bool BATCH>
__device__ __forceinline__ void substep_body(const Pipeline& p, uint64_t seq, uint32_t blk, uint32_t nblk) {
    constexpr bool G2P = (MODE & kModeG2P) != 0;
    constexpr bool P2G = (MODE & kModeP2G) != 0;
    constexpr bool HARVEST = (MODE & kModeHarvest) != 0;
    constexpr bool SURR = (MODE & kModeSurrogate) != 0;
    constexpr bool SLEEP = (MODE & kModeSleep) != 0;
    static_assert(!(SURR && SLEEP), "the surrogate substep has no sleep partition");
    constexpr int NCH = P2G ? kAccChannels : 1;
    __shared__ float s_sg[SURR ? kWaves : 1][SURR ? kSgQueueLds<1> : 1];
    __shared__ std::conditional_t<SURR, SgLds, char> s_net;
    __shared__ double s_acc[NCH][kLdsCells];
    __shared__ float2 s_vel[G2P ? kLdsCells : 1];
    __shared__ uint16_t s_perm[kChunk];
    __shared__ uint32_t s_scan[17];
    __shared__ int s_box[4 * kPer * kWaves];
    __shared__ Box s_mbox[kPer];
    __shared__ float s_red[2 * kWaves];

    // 0. this launch's substep, as every workgroup derives it
    Ctrl* c = p.ctrl;
    const DevConfig& cfg = p.cfg;
    const CtrlState P = c->st[seq & 1u];
    const CtrlRed Q = c->red[seq_mod3(seq + 2u)];
    const CtrlState N = G2P ? decide_substep(P, Q, cfg) : P;
    // dt' of this launch's P2G, and dt - dt' of the accumulators it gathers
    // (a density pass scatters with the exact dt)
    const float pred = predict_dt(N, cfg);
    const float ddt = (G2P && !cfg.density_pass) ? N.dt - P.pred_dt : 0.0f;
    if (blk == 0 && threadIdx.x == 0) {
        CtrlState W = N;
        if (!G2P) W.active = 0;
        if (P2G) W.pred_dt = pred;
        if (G2P && N.active) W.last_ddt = ddt;
        c->st[(seq + 1u) & 1u] = W;
        if (G2P && !N.active) c->red[seq_mod3(seq)] = Q;  // nothing ran: the reduction carries over
        reset_red(c->red[seq_mod3(seq + 1u)]);
    }
    if (G2P && !N.active) return;
    if constexpr (SURR) sg_stage(c->sg.pk, c->sg.bias, s_net);
    CtrlRed* red = &c->red[seq_mod3(seq)];
    const int R = (int)cfg.grid_res;
    const uint32_t cur = N.cur_grid;
    const uint32_t nxt = cur == 2u ? 0u : cur + 1u;
    const uint32_t old = nxt == 2u ? 0u : nxt + 1u;
    const float dt = G2P ? N.dt : 0.0f;
    // the prologue projects the uploaded state in registers; a substep
    // projects its input only if the writer flagged it (and no density pass
    // ran in between, which would have projected and persisted it)
    const bool project_in = cfg.project && (G2P ? (Q.dirty && !cfg.density_pass) : true);
    const bool reorder = G2P && P.reorder;
    const bool full = !G2P || project_in || reorder;
    // a one-model instance's fields are a constant (the host's fields of that model alone)
    const uint32_t mask = full ? kFAll : (ONLY >= 0 ? model_fields((uint32_t)ONLY) : cfg.fields);
    const PBuf a = sel_buf(p, P.cur_buf);
    const uint32_t other_buf = P.cur_buf ^ 1u;
    const uint32_t chunk = BATCH ? blk : xcd_chunk(blk, nblk);
    const uint32_t base = chunk * kChunk;

    // 1. the chunk's box; a substep's output base cell moves at most one cell
    //    (the G2P speed clamp bounds |v dt| by one cell), so the box is
    //    widened by two for G2P and one for a scatter-only pass; then the
    //    chunk's processing order, sorted by base cell
    //    A chunk whose box exceeds the LDS budget (sparse, scattered
    //    particles) runs slice by slice instead: each kBlock-particle slice
    //    with a box of its own, so that one such chunk does not fall back to
    //    global-memory G2P and P2G and set the launch's duration.
    //    When the chunk's box does not fit but its slices' boxes fit side by
    //    side, each round uses its slice's box, all staged, sorted and
    //    flushed once (a chunk at a body's edge, or one whose run of keys
    //    spans two patches, without the four passes of slicing).
    const int margin = G2P ? 2 : 1;
    Box box = chunk_box(p, a, base, project_in, margin, s_box);
    int nbins = 0;
    if (!box.lds) nbins = slice_boxes(p, a, base, project_in, margin, s_box, s_mbox);
    const bool multi = !box.lds && nbins <= kLdsCells;
    const bool sliced = !box.lds && !multi;
    const uint32_t nvalid = base < p.n ? min((uint32_t)kChunk, p.n - base) : 0u;
    chunk_order(p, a, base, project_in, box, multi ? s_mbox : nullptr, nbins, nvalid,
                reinterpret_cast<uint32_t*>(s_acc[0]), s_perm, s_scan);

    // 2. clear the LDS accumulators; stage the updated grid velocity
    if (multi)
        stage_boxes<G2P, NCH>(p, s_mbox, cur, dt, ddt, s_acc, s_vel);
    else if (!sliced)
        stage_box<G2P, NCH>(p, box, cur, dt, ddt, s_acc, s_vel);

    const GridDst gd{nxt, P2G ? 0u : 1u};

    float speed = 0.0f, mdt = kInf;
    uint32_t n_clamp = 0, n_proj = 0;
    bool any_dirty = false;
#pragma unroll 1
    for (int k = 0; k < kPer; ++k) {
        const uint32_t q = k * kBlock + threadIdx.x;
        const bool valid = q < nvalid;
        if (multi) box = s_mbox[k];
        if (sliced) {
            box = slice_box(p, a, base + q, valid, project_in, margin, s_box);
            slice_order(p, a, base, project_in, box, k, nvalid, reinterpret_cast<uint32_t*>(s_acc[0]), s_perm, s_scan);
            stage_box<G2P, NCH>(p, box, cur, dt, ddt, s_acc, s_vel);
        }
        const uint32_t j = base + (valid ? s_perm[q] : q);
        PState sp = neutral_state();
        PState s = neutral_state();
        bool asleep = false, mat_changed = false;
        if (valid) {
            const float2 x = a.x[j];
            s.x = x.x;
            s.y = x.y;
            load_fields(s, a, j, mask, !G2P || project_in);
            if constexpr (SLEEP) {
                asleep = (s.mat & kSleepBit) != 0u;
                if (G2P && asleep && stencil_touched(p, cur, s.x, s.y)) {
                    s.mat &= ~kSleepBit;
                    asleep = false;
                    mat_changed = true;
                }
            }
            if (G2P && !asleep) {
                if (project_in && project_state(s, cfg)) ++n_proj;
                // 3. gather_grid_to_particles phase 1 (g2p.rs:203-325)
                const Wts w = weights(s.x, s.y);
                const int at = box.stencil_at(w.bx, w.by);
                const bool lds = at >= 0;
                float nvx = 0.0f, nvy = 0.0f;
                M2 b{0.0f, 0.0f, 0.0f, 0.0f};
                auto gather = [&](int gx, int gy, float2 nv) {
                    const float weight = w.wx[gx] * w.wy[gy];
                    const float dx = ((float)(w.bx + gx - 1) - s.x) + 0.5f;
                    const float dy = ((float)(w.by + gy - 1) - s.y) + 0.5f;
                    const float wvx = nv.x * weight, wvy = nv.y * weight;
                    b.a = b.a + wvx * dx;
                    b.b = b.b + wvy * dx;
                    b.c = b.c + wvx * dy;
                    b.d = b.d + wvy * dy;
                    nvx = nvx + wvx;
                    nvy = nvy + wvy;
                };
                // the nine LDS reads go out together; a lane whose stencil
                // leaves the box reads from base 0 and replaces what it read
                // (in bounds: 2 h + 2 < w h <= kLdsCells for an LDS box,
                // h = 0 otherwise)
                const int lb = lds ? at : 0;
                float2 nv[9];
#pragma unroll
                for (int gx = 0; gx < 3; ++gx)
#pragma unroll
                    for (int gy = 0; gy < 3; ++gy) nv[gx * 3 + gy] = s_vel[lb + gx * box.h + gy];
#pragma unroll
                for (int gx = 0; gx < 3; ++gx)
#pragma unroll
                    for (int gy = 0; gy < 3; ++gy) {
                        if (!lds)
                            nv[gx * 3 + gy] =
                                cell_velocity(p, opaque_index(cur), dt, ddt, w.bx + gx - 1, w.by + gy - 1);
                        gather(gx, gy, nv[gx * 3 + gy]);
                    }
                if (s.mat & kPinnedBit) {
                    s.vx = 0.0f;
                    s.vy = 0.0f;
                    s.c = M2{0.0f, 0.0f, 0.0f, 0.0f};
                } else {
                    const float vel_limit = cfg.cell / dt;
                    const float spd = len2(nvx, nvy);
                    if (spd > vel_limit) {
                        const float kk = vel_limit / spd;
                        nvx = nvx * kk;
                        nvy = nvy * kk;
                        ++n_clamp;
                    }
                    float nx = s.x + nvx * dt, ny = s.y + nvy * dt;
                    boundary_clamp(cfg, nx, ny);
                    s.vx = nvx;
                    s.vy = nvy;
                    s.c = M2{(b.a * kDInv) * cfg.apic_blend, (b.b * kDInv) * cfg.apic_blend,
                             (b.c * kDInv) * cfg.apic_blend, (b.d * kDInv) * cfg.apic_blend};
                    s.x = nx;
                    s.y = ny;
                }
                // phase 2: update_particle (g2p.rs:333-340)
                if constexpr (HARVEST) harvest_nacc(p.ctrl->sg, mat_of(p, s.mat), s, dt);
                if constexpr (!SURR) update_particle<ONLY>(mat_of(p, s.mat), s, dt);
                if constexpr (SLEEP) sleep_test(c->sleep_sq, s, asleep, mat_changed);
            }
            if constexpr (!SURR) {
                // 4. the state the next substep starts from: CFL, projection check
                if (!asleep) {
                    cfl_of<ONLY>(p, s, speed, mdt);
                    any_dirty |= needs_projection(s, sp, cfg);
                }
                if (G2P && (!SLEEP || !asleep || mat_changed || reorder)) {
                    if (reorder) {
                        const PBuf o = reorder_buf(p, other_buf);
                        const uint32_t dest = p.sort.out[j];
                        store_fields(o, dest, s, mask, full);
                        o.id[dest] = a.id[j];
                        if (SLEEP && mat_changed && !full) o.mat[dest] = s.mat;
                    } else {
                        store_fields(a, j, s, mask, full);
                        if (SLEEP && mat_changed && !full) a.mat[j] = s.mat;
                    }
                }
            }
        }
        if constexpr (SURR) {
            // phase 2 with the learned return mapping (every lane of the wave)
            surrogate_update<ONLY>(p.ctrl->sg, s_net, mat_of(p, s.mat), valid, s, dt, s_sg[threadIdx.x >> 6]);
            if (valid) {
                cfl_of<ONLY>(p, s, speed, mdt);
                any_dirty |= needs_projection(s, sp, cfg);
                const uint32_t dest = reorder ? p.sort.out[j] : j;
                if (reorder) {
                    const PBuf o = reorder_buf(p, other_buf);
                    store_fields(o, dest, s, mask, full);
                    o.id[dest] = a.id[j];
                } else {
                    store_fields(a, dest, s, mask, full);
                }
            }
        }
        // 5. P2G of the projected state into the next accumulator set (the
        //    awake particles only)
        const bool scat = valid && !asleep;
        if (P2G) {
            const P2GTerms t = p2g_terms<ONLY>(p, sp, pred);
            scatter_p2g(t, scat, box, s_acc, p, gd, R);
        } else {
            P2GTerms t;
            t.w = weights(sp.x, sp.y);
            t.x = sp.x;
            t.y = sp.y;
            t.mass = sp.mass;
            t.ux = t.uy = 0.0f;
            t.k = t.st = M2{0.0f, 0.0f, 0.0f, 0.0f};
            scatter_p2g<1>(t, scat, box, s_acc, p, gd, R);
        }
        if (sliced) {
            flush_box<NCH>(box, s_acc, p, gd, R);
            __syncthreads();  // the next slice stages into the same LDS
        }
    }

    // 6. flush the LDS box (or boxes) into the next accumulator set
    if (multi) {
        __syncthreads();
#pragma unroll 1
        for (int k = 0; k < kPer; ++k) flush_box<NCH>(s_mbox[k], s_acc, p, gd, R, false);
    } else if (!sliced) {
        flush_box<NCH>(box, s_acc, p, gd, R);
    }
    // 7. clear this workgroup's slice of the set the substep after next uses
    if (G2P) {
        const uint32_t R2 = (uint32_t)R * (uint32_t)R;
        const uint32_t per = (R2 + nblk - 1) / nblk;
        const uint32_t lo = blk * per, hi = min(R2, lo + per);
        // (a density-pass scene leaves the S channels zero from the start)
        const int nch = cfg.density_pass ? kAccPy + 1 : kAccChannels;
        for (uint32_t i = lo + threadIdx.x; i < hi; i += kBlock) {
            for (int ch = 0; ch < nch; ++ch) p.grid.acc[old][ch][i] = 0.0f;
            if (cfg.density_pass) p.grid.dm[old][i] = 0.0f;
        }
    }
    // 8. counters and the CFL reduction the next launch decides from
    const uint32_t wc = wave_sum(n_clamp), wp = wave_sum(n_proj);
    const bool wd = __ballot(any_dirty) != 0ull;
    if (lane_id() == 0) {
        if (wc) atomicAdd(&c->clamp_count[N.step % kStepRing], wc);
        if (wp) atomicAdd(&c->proj_count[N.step % kStepRing], wp);
        if (wd) atomicOr(&red->dirty, 1u);
    }
    reduce_cfl(red, speed, cfl_bound<ONLY>(p, mdt), s_red);
}

template <int MODE, int ONLY>
__global__ __launch_bounds__(kBlock, (MODE & kModeSurrogate) ? 2 : 4) void k_substep(const Pipeline p, uint64_t seq) {
    substep_body<MODE, ONLY, false>(p, seq, blockIdx.x, gridDim.x);
}

// Density re-estimate (density.rs:216-247) from the mass grid the previous
// substep scattered, then this substep's P2G.  This kernel is the first to
// read the substep's state, so it also projects (and persists) it if flagged.
template <bool BATCH, bool SLEEP>
__device__ __forceinline__ void density_p2g_body(const Pipeline& p, uint64_t seq, uint32_t blk, uint32_t nblk) {
    // dt is decided here, before the terms are formed: the S channels (the
    // correction for a mispredicted dt) stay empty, and the grid update of a
    // density-pass scene never reads them
    constexpr int kDensCh = kAccPy + 1;
    __shared__ double s_acc[kDensCh][kLdsCells];
    __shared__ float s_dm[kLdsCells];
    __shared__ int s_box[4 * kPer * kWaves];
    __shared__ Box s_mbox[kPer];
    Ctrl* c = p.ctrl;
    const DevConfig& cfg = p.cfg;
    // same launch number as the substep kernel that follows: same decision
    const CtrlState P = c->st[seq & 1u];
    const CtrlRed Q = c->red[seq_mod3(seq + 2u)];
    const CtrlState N = decide_substep(P, Q, cfg);
    if (!N.active) return;
    const int R = (int)cfg.grid_res;
    const uint32_t cur = N.cur_grid;
    const bool project_in = cfg.project && Q.dirty;
    const uint32_t mask = project_in ? kFAll : (cfg.fields | kFVolDens);
    const PBuf a = sel_buf(p, P.cur_buf);
    const uint32_t chunk = BATCH ? blk : xcd_chunk(blk, nblk);
    const uint32_t base = chunk * kChunk;
    const float* dm = p.grid.dm[cur];
    const GridDst gd{cur, 0u};

    // a chunk whose box exceeds the budget takes its slices' boxes side by
    // side when they fit together (slice_boxes), the particles of round k in
    // slice k's box
    Box box = chunk_box(p, a, base, project_in, 1, s_box);
    bool multi = false;
    if (!box.lds) multi = slice_boxes(p, a, base, project_in, 1, s_box, s_mbox) <= kLdsCells;
    if (box.lds || multi) {
#pragma unroll 1
        for (int b = 0; b < (multi ? kPer : 1); ++b) {
            const Box x = multi ? s_mbox[b] : box;
            if (!x.lds) continue;
            for (int i = threadIdx.x; i < x.w * x.h; i += kBlock) {
#pragma unroll
                for (int ch = 0; ch < kDensCh; ++ch) s_acc[ch][x.lo + i] = 0.0;
                const int cx = x.ox + i / x.h, cy = x.oy + i % x.h;
                s_dm[x.lo + i] =
                    (cx >= 0 && cy >= 0 && cx < R && cy < R) ? dm[(uint32_t)cx * (uint32_t)R + (uint32_t)cy] : 0.0f;
            }
        }
        __syncthreads();
    }
    uint32_t n_proj = 0;
#pragma unroll 1
    for (int k = 0; k < kPer; ++k) {
        const uint32_t j = base + k * kBlock + threadIdx.x;
        const bool valid = j < p.n;
        if (multi) box = s_mbox[k];
        PState s = neutral_state();
        bool asleep = false;
        if (valid) {
            const float2 x = a.x[j];
            s.x = x.x;
            s.y = x.y;
            load_fields(s, a, j, mask, true);
            if (SLEEP) asleep = (s.mat & kSleepBit) != 0u;  // estimate_particle_volumes visits the awake ones
        }
        if (valid && !asleep) {
            if (project_in && project_state(s, cfg)) ++n_proj;
            const Wts w = weights(s.x, s.y);
            const int at = box.stencil_at(w.bx, w.by);
            float density = 0.0f;
            for (int gx = 0; gx < 3; ++gx)
                for (int gy = 0; gy < 3; ++gy) {
                    const float weight = w.wx[gx] * w.wy[gy];
                    const int cx = w.bx + gx - 1, cy = w.by + gy - 1;
                    float m = 0.0f;
                    if (at >= 0)
                        m = s_dm[at + gx * box.h + gy];
                    else if (cx >= 0 && cy >= 0 && cx < R && cy < R)
                        m = dm[(uint32_t)cx * (uint32_t)R + (uint32_t)cy];
                    density = density + m * weight;
                }
            if (density > kF32Eps) {
                s.dens = density;
                s.vol = s.mass / density;
            }
            if (project_in) {
                store_fields(a, j, s, kFAll, true);  // persist the projection
            } else {
                a.dens[j] = s.dens;
                a.vol[j] = s.vol;
            }
        }
        const P2GTerms t = p2g_terms<-1>(p, s, N.dt);
        scatter_p2g<kDensCh>(t, valid && !asleep, box, s_acc, p, gd, R);
    }
    if (multi) {
        __syncthreads();
#pragma unroll 1
        for (int k = 0; k < kPer; ++k) flush_box<kDensCh>(s_mbox[k], s_acc, p, gd, R, false);
    } else {
        flush_box<kDensCh>(box, s_acc, p, gd, R);
    }
    const uint32_t wp = wave_sum(n_proj);
    if (lane_id() == 0 && wp) atomicAdd(&c->proj_count[N.step % kStepRing], wp);
}
template <bool SLEEP>
__global__ __launch_bounds__(kBlock) void k_density_p2g(const Pipeline p, uint64_t seq) {
    density_p2g_body<false, SLEEP>(p, seq, blockIdx.x, gridDim.x);
}

// ------------------------------------------------------ batched scenes
// One launch per stage for many independent scenes: workgroup b of a batch
// launch belongs to the scene s with start[s] <= b < start[s + 1] (a
// per-stage work list of each scene's workgroup count), and runs the
// single-scene body on that scene's pipeline.  Every scene keeps its own
// device-side control (its own substep dt from its own CFL bound); scenes
// whose step has finished idle through the launches the others still need.
__device__ __forceinline__ uint32_t batch_scene(const BatchView& b, uint32_t block, uint32_t& blk, uint32_t& nblk) {
    uint32_t lo = 0, hi = b.n;  // the last scene whose start is <= block
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (b.start[mid] <= block)
            lo = mid;
        else
            hi = mid;
    }
    blk = block - b.start[lo];
    nblk = b.start[lo + 1] - b.start[lo];
    return lo;
}
__global__ void kb_begin_step(BatchView b, uint64_t seq, uint32_t reorder) {
    uint32_t blk, nblk;
    const uint32_t s = batch_scene(b, blockIdx.x, blk, nblk);
    if (threadIdx.x == 0) begin_step_body(b.pipes[s], seq, reorder);
}
__global__ __launch_bounds__(kBlock) void kb_sort_count(BatchView b, uint64_t seq, uint32_t pass) {
    uint32_t blk, nblk;
    const uint32_t s = batch_scene(b, blockIdx.x, blk, nblk);
    sort_count_body(pipeline_sort_job(b.pipes[s], seq), pass, blk, nblk);
}
__global__ __launch_bounds__(kBlock) void kb_sort_scan(BatchView b, uint64_t seq, uint32_t pass) {
    uint32_t blk, nblk;
    const uint32_t s = batch_scene(b, blockIdx.x, blk, nblk);
    sort_scan_body(pipeline_sort_job(b.pipes[s], seq), pass, blk);
}
__global__ __launch_bounds__(kBlock) void kb_sort_scatter(BatchView b, uint64_t seq, uint32_t pass) {
    uint32_t blk, nblk;
    const uint32_t s = batch_scene(b, blockIdx.x, blk, nblk);
    sort_scatter_body(pipeline_sort_job(b.pipes[s], seq), pass, blk, nblk);
}
// 3 waves per SIMD (at most 168 VGPRs): the scene's pipeline is read through
// a pointer, so the uniform fields the single-scene kernel keeps in SGPRs
// compete for VGPRs; measured on 256 batched scenes, 3 waves (24-72 B/lane of
// spills) beat 2 (no spills) by 11 % and 4 (190-200 B/lane) by 10 %.
template <int MODE>
__global__ __launch_bounds__(kBlock, 3) void kb_substep(BatchView b, uint64_t seq) {
    uint32_t blk, nblk;
    const uint32_t s = batch_scene(b, blockIdx.x, blk, nblk);
    substep_body<MODE, -1, true>(b.pipes[s], seq, blk, nblk);
}
// 4 waves per SIMD: fits 128 VGPRs without spills (3.5 % faster on the batch).
__global__ __launch_bounds__(kBlock, 4) void kb_density_p2g(BatchView b, uint64_t seq) {
    uint32_t blk, nblk;
    const uint32_t s = batch_scene(b, blockIdx.x, blk, nblk);
    density_p2g_body<true, false>(b.pipes[s], seq, blk, nblk);
}
// Every scene's control block, for the host's drain.
__global__ void kb_gather_ctrl(BatchView b, Ctrl* out) {
    uint32_t blk, nblk;
    const uint32_t s = batch_scene(b, blockIdx.x, blk, nblk);
    const uint32_t* src = reinterpret_cast<const uint32_t*>(b.pipes[s].ctrl);
    uint32_t* dst = reinterpret_cast<uint32_t*>(out + s);
    for (uint32_t i = threadIdx.x; i < sizeof(Ctrl) / 4; i += blockDim.x) dst[i] = src[i];
}

// ------------------------------------------------------ diagnostics
// SimSnapshot (systems/diagnostics/snapshot.rs:227-372) over the device state.
__device__ __forceinline__ uint32_t dev_order_key(float f) {
    const uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__global__ __launch_bounds__(kBlock) void k_snap_particles(const Pipeline p, uint32_t cur_buf, DevSnapshot* out,
                                                           uint32_t* cell_lo, uint32_t* cell_hi, uint32_t* cell_count) {
    const PBuf b = sel_buf(p, cur_buf);
    const DevConfig& cfg = p.cfg;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < p.n;
    double mass = 0.0, mx = 0.0, my = 0.0, ke = 0.0, jp_sum = 0.0, h_sum = 0.0;
    float speed = 0.0f, pinned_speed = 0.0f;
    uint32_t min_j = 0xffffffffu, max_j = 0u, min_jp = 0xffffffffu;
    uint32_t n_valid = 0, n_nonfinite = 0, n_invalid = 0, n_oob = 0, n_jp = 0, n_sleep = 0;
    if (valid) {
        const float2 x = b.x[i], v = b.v[i];
        const float4 c = b.c[i], f = b.f[i];
        const float m = b.mass[i], v0 = b.v0[i], vol = b.vol[i], dens = b.dens[i], jp = b.jp[i], h = b.h[i];
        const uint32_t word = b.mat[i];
        n_sleep = (word & kSleepBit) ? 1u : 0u;
        const float j = f.x * f.w - f.z * f.y;  // Mat2::determinant, column-major
        mass = m;
        mx = (double)(m * v.x);
        my = (double)(m * v.y);
        const float sp = len2(v.x, v.y);
        speed = sp;
        ke = (double)((0.5f * m) * (v.x * v.x + v.y * v.y));
        if (word & kPinnedBit) pinned_speed = sp;
        if (__builtin_isfinite(j)) {
            min_j = max_j = dev_order_key(j);
        }
        const float lo = (float)(cfg.bt > 0 ? cfg.bt - 1 : 0), hi = (float)(cfg.grid_res > cfg.bt ? cfg.grid_res - cfg.bt : 0);
        if (x.x < lo || x.x > hi || x.y < lo || x.y > hi) n_oob = 1;
        if (__builtin_isfinite(x.x) && __builtin_isfinite(x.y)) {  // particle_cell_index
            const int ix = sat_i32(floorf(x.x)), iy = sat_i32(floorf(x.y));
            if (ix >= 0 && iy >= 0 && ix < (int)cfg.grid_res && iy < (int)cfg.grid_res) {
                const uint32_t ci = (uint32_t)ix * cfg.grid_res + (uint32_t)iy;
                const uint32_t id = word & kMatMask;
                atomicMin(&cell_lo[ci], id);
                atomicMax(&cell_hi[ci], id);
                atomicAdd(&cell_count[ci], 1u);
            }
        }
        if (__builtin_isfinite(jp)) {
            jp_sum = jp;
            min_jp = dev_order_key(jp);
            n_jp = 1;
        }
        if (__builtin_isfinite(h)) h_sum = h;
        const float vals[17] = {x.x, x.y, v.x, v.y, c.x, c.y, c.z, c.w, f.x, f.y, f.z, f.w, m, v0, vol, dens, jp};
        for (int k = 0; k < 17; ++k) n_nonfinite += __builtin_isfinite(vals[k]) ? 0u : 1u;
        n_valid = n_nonfinite == 0 ? 1u : 0u;
        n_invalid = (m <= 0.0f) + (vol <= 0.0f) + (v0 <= 0.0f) + (dens <= 0.0f) + (j <= 0.0f) + (jp <= 0.0f);
    }
    mass = wave_sum_d(mass);
    mx = wave_sum_d(mx);
    my = wave_sum_d(my);
    ke = wave_sum_d(ke);
    jp_sum = wave_sum_d(jp_sum);
    h_sum = wave_sum_d(h_sum);
    speed = wave_max(speed);
    pinned_speed = wave_max(pinned_speed);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        min_j = min(min_j, (uint32_t)__shfl_xor((int)min_j, o, 64));
        max_j = max(max_j, (uint32_t)__shfl_xor((int)max_j, o, 64));
        min_jp = min(min_jp, (uint32_t)__shfl_xor((int)min_jp, o, 64));
    }
    n_valid = wave_sum(n_valid);
    n_nonfinite = wave_sum(n_nonfinite);
    n_invalid = wave_sum(n_invalid);
    n_oob = wave_sum(n_oob);
    n_jp = wave_sum(n_jp);
    n_sleep = wave_sum(n_sleep);
    if (lane_id() == 0) {
        atomicAdd(&out->mass, mass);
        atomicAdd(&out->mom_x, mx);
        atomicAdd(&out->mom_y, my);
        atomicAdd(&out->kinetic, ke);
        atomicAdd(&out->jp_sum, jp_sum);
        atomicAdd(&out->h_sum, h_sum);
        if (speed > 0.0f) atomicMax(&out->max_speed, __float_as_uint(speed));
        if (pinned_speed > 0.0f) atomicMax(&out->max_pinned_speed, __float_as_uint(pinned_speed));
        atomicMin(&out->min_j, min_j);
        atomicMax(&out->max_j, max_j);
        atomicMin(&out->min_jp, min_jp);
        if (n_valid) atomicAdd(&out->valid, n_valid);
        if (n_nonfinite) atomicAdd(&out->non_finite, n_nonfinite);
        if (n_invalid) atomicAdd(&out->invalid, n_invalid);
        if (n_oob) atomicAdd(&out->out_of_bounds, n_oob);
        if (n_sleep) atomicAdd(&out->sleeping, n_sleep);
        if (n_jp) atomicAdd(&out->jp_finite, n_jp);
    }
}
__global__ __launch_bounds__(kBlock) void k_snap_grid(const Pipeline p, uint32_t set, float dt, float ddt,
                                                      DevSnapshot* out, uint32_t* cell_lo, uint32_t* cell_hi,
                                                      uint32_t* cell_count) {
    const uint32_t R = p.cfg.grid_res;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    double gm = 0.0, gx = 0.0, gy = 0.0;
    float gspeed = 0.0f;
    uint32_t active = 0, nonfinite = 0, touched = 0, mixed = 0, in_mixed = 0;
    if (i < R * R) {
        const float m = p.grid.acc[set][kAccM][i];
        if (m > 0.0f) {
            const float2 v = cell_velocity(p, set, dt, ddt, (int)(i / R), (int)(i % R));
            active = 1;
            gm = m;
            gx = (double)(m * v.x);
            gy = (double)(m * v.y);
            gspeed = len2(v.x, v.y);
            nonfinite = (__builtin_isfinite(m) ? 0u : 1u) + (__builtin_isfinite(v.x) ? 0u : 1u) +
                        (__builtin_isfinite(v.y) ? 0u : 1u);
        }
        const uint32_t cnt = cell_count[i];
        if (cnt) {
            touched = 1;
            if (cell_lo[i] != cell_hi[i]) {
                mixed = 1;
                in_mixed = cnt;
            }
            cell_lo[i] = 0xffffffffu;
            cell_hi[i] = 0u;
            cell_count[i] = 0u;
        }
    }
    gm = wave_sum_d(gm);
    gx = wave_sum_d(gx);
    gy = wave_sum_d(gy);
    gspeed = wave_max(gspeed);
    active = wave_sum(active);
    nonfinite = wave_sum(nonfinite);
    touched = wave_sum(touched);
    mixed = wave_sum(mixed);
    in_mixed = wave_sum(in_mixed);
    if (lane_id() == 0) {
        atomicAdd(&out->grid_mass, gm);
        atomicAdd(&out->grid_mom_x, gx);
        atomicAdd(&out->grid_mom_y, gy);
        if (gspeed > 0.0f) atomicMax(&out->max_grid_speed, __float_as_uint(gspeed));
        if (active) atomicAdd(&out->active_cells, active);
        if (nonfinite) atomicAdd(&out->non_finite_grid, nonfinite);
        if (touched) atomicAdd(&out->touched_cells, touched);
        if (mixed) atomicAdd(&out->mixed_cells, mixed);
        if (in_mixed) atomicAdd(&out->particles_in_mixed, in_mixed);
    }
}

// ------------------------------------------------- stand-alone cell sort
__global__ __launch_bounds__(kBlock) void k_sortx_count(const SortJob J, uint32_t pass) {
    sort_count_body(J, pass, blockIdx.x, gridDim.x);
}
__global__ __launch_bounds__(kBlock) void k_sortx_scan(const SortJob J, uint32_t pass) {
    sort_scan_body(J, pass, blockIdx.x);
}
__global__ __launch_bounds__(kBlock) void k_sortx_scatter(const SortJob J, uint32_t pass) {
    sort_scatter_body(J, pass, blockIdx.x, gridDim.x);
}

inline uint32_t div_up(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

}  // namespace

// ------------------------------------------------------------ launchers

hipError_t launch_zero_all(const Pipeline& p, hipStream_t stream) {
    const uint32_t R2 = p.cfg.grid_res * p.cfg.grid_res;
    k_zero_all<<<div_up(R2, kBlock), kBlock, 0, stream>>>(p);
    return hipGetLastError();
}

hipError_t launch_prologue(const Pipeline& p, uint64_t seq, hipStream_t stream, bool sleep) {
    const uint32_t nb = p.n == 0 ? 1u : div_up(p.n, kChunk);
    if (p.cfg.density_pass) {
        if (sleep)
            k_substep<kModeMass | kModeSleep, -1><<<nb, kBlock, 0, stream>>>(p, seq);
        else
            k_substep<kModeMass, -1><<<nb, kBlock, 0, stream>>>(p, seq);
    } else {
        if (sleep)
            k_substep<kModeP2G | kModeSleep, -1><<<nb, kBlock, 0, stream>>>(p, seq);
        else
            k_substep<kModeP2G, -1><<<nb, kBlock, 0, stream>>>(p, seq);
    }
    return hipGetLastError();
}

hipError_t launch_batch_begin_step(const BatchLists& L, uint64_t seq, bool reorder, hipStream_t stream) {
    kb_begin_step<<<L.one.n, 64, 0, stream>>>(L.one, seq, reorder ? 1u : 0u);
    return hipGetLastError();
}

hipError_t launch_batch_sort(const BatchLists& L, uint64_t seq, hipStream_t stream) {
    if (!L.sort_blocks) return hipSuccess;
    // the keys are those of the state launch seq + 1 will move
    for (uint32_t pass = 0; pass < L.sort_passes; ++pass) {
        kb_sort_count<<<L.sort_blocks, kBlock, 0, stream>>>(L.sort, seq + 1u, pass);
        kb_sort_scan<<<L.scan_blocks, kBlock, 0, stream>>>(L.scan, seq + 1u, pass);
        kb_sort_scatter<<<L.sort_blocks, kBlock, 0, stream>>>(L.sort, seq + 1u, pass);
    }
    return hipGetLastError();
}

hipError_t launch_batch_substep(const BatchLists& L, uint64_t seq, hipStream_t stream) {
    if (L.density_blocks) {
        kb_density_p2g<<<L.density_blocks, kBlock, 0, stream>>>(L.density, seq);
        kb_substep<kModeG2P | kModeMass><<<L.density_blocks, kBlock, 0, stream>>>(L.density, seq);
    }
    if (L.fused_blocks) kb_substep<kModeG2P | kModeP2G><<<L.fused_blocks, kBlock, 0, stream>>>(L.fused, seq);
    return hipGetLastError();
}

hipError_t launch_batch_gather_ctrl(const BatchLists& L, Ctrl* out, hipStream_t stream) {
    kb_gather_ctrl<<<L.one.n, 256, 0, stream>>>(L.one, out);
    return hipGetLastError();
}

hipError_t launch_begin_step(const Pipeline& p, uint64_t seq, bool reorder, hipStream_t stream) {
    reorder = reorder && p.n > 0;
    k_begin_step<<<1, 1, 0, stream>>>(p, seq, reorder ? 1u : 0u);
    if (reorder) {
        // the keys are those of the state launch seq + 1 will move
        const uint32_t nb = sort_tiles(p.n);
        const uint32_t key_bits = cell_key_bits(p.cfg.n_bins) + p.sort.model_bits;
        const uint32_t passes = sort_plan_bits(key_bits).passes, ng = sort_scan_groups(key_bits);
        for (uint32_t pass = 0; pass < passes; ++pass) {
            k_sort_count<<<nb, kBlock, 0, stream>>>(p, seq + 1u, pass);
            k_sort_scan<<<ng, kBlock, 0, stream>>>(p, seq + 1u, pass);
            k_sort_scatter<<<nb, kBlock, 0, stream>>>(p, seq + 1u, pass);
        }
    }
    return hipGetLastError();
}

hipError_t launch_substep(const Pipeline& p, uint64_t seq, hipStream_t stream, uint32_t sg_mode, bool sleep) {
    const uint32_t nb = p.n == 0 ? 1u : div_up(p.n, kChunk);
    if (sleep) {  // the sleep partition: mixed-model instances only
        if (p.cfg.density_pass) {
            k_density_p2g<true><<<nb, kBlock, 0, stream>>>(p, seq);
            k_substep<kModeG2P | kModeMass | kModeSleep, -1><<<nb, kBlock, 0, stream>>>(p, seq);
        } else {
            k_substep<kModeG2P | kModeP2G | kModeSleep, -1><<<nb, kBlock, 0, stream>>>(p, seq);
        }
        return hipGetLastError();
    }
    if (p.cfg.density_pass) {
        k_density_p2g<false><<<nb, kBlock, 0, stream>>>(p, seq);
        k_substep<kModeG2P | kModeMass, -1><<<nb, kBlock, 0, stream>>>(p, seq);
    } else {
        constexpr int kFused = kModeG2P | kModeP2G;
        switch (p.cfg.only_model) {
            case kModelNH: k_substep<kFused, kModelNH><<<nb, kBlock, 0, stream>>>(p, seq); break;
            case kModelCorot: k_substep<kFused, kModelCorot><<<nb, kBlock, 0, stream>>>(p, seq); break;
            case kModelSnow: k_substep<kFused, kModelSnow><<<nb, kBlock, 0, stream>>>(p, seq); break;
            case kModelDP: k_substep<kFused, kModelDP><<<nb, kBlock, 0, stream>>>(p, seq); break;
            case kModelNacc:
                if (sg_mode == 1u)
                    k_substep<kFused | kModeHarvest, kModelNacc><<<nb, kBlock, 0, stream>>>(p, seq);
                else if (sg_mode == 2u)
                    k_substep<kFused | kModeSurrogate, kModelNacc><<<nb, kBlock, 0, stream>>>(p, seq);
                else
                    k_substep<kFused, kModelNacc><<<nb, kBlock, 0, stream>>>(p, seq);
                break;
            default: k_substep<kFused, -1><<<nb, kBlock, 0, stream>>>(p, seq); break;
        }
    }
    return hipGetLastError();
}

uint32_t order_key(float f) {
    uint32_t b;
    std::memcpy(&b, &f, 4);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
float order_key_inverse(uint32_t k) {
    const uint32_t b = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
    float f;
    std::memcpy(&f, &b, 4);
    return f;
}

hipError_t launch_snapshot(const Pipeline& p, uint32_t cur_buf, uint32_t grid_set, float dt, float ddt,
                           DevSnapshot* out, uint32_t* cell_lo, uint32_t* cell_hi, uint32_t* cell_count,
                           hipStream_t stream) {
    if (p.n > 0)
        k_snap_particles<<<div_up(p.n, kBlock), kBlock, 0, stream>>>(p, cur_buf, out, cell_lo, cell_hi, cell_count);
    const uint32_t R2 = p.cfg.grid_res * p.cfg.grid_res;
    k_snap_grid<<<div_up(R2, kBlock), kBlock, 0, stream>>>(p, grid_set, dt, ddt, out, cell_lo, cell_hi, cell_count);
    return hipGetLastError();
}

hipError_t sort_cell_keys_device(const float2* xy, uint32_t n, uint32_t grid_res, uint32_t* keys_out,
                                 uint32_t* perm_out, const SortBuf& s, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    SortJob J;
    J.x = xy;
    J.project = 0u;
    J.proj_lo = J.proj_hi = 0.0f;
    J.n = n;
    J.grid_res = grid_res;
    J.key_order = cell_key_order(grid_res);
    J.n_bins = 1u << (2u * J.key_order);
    J.mat = nullptr;
    J.mats = nullptr;
    J.n_mats = 0;
    J.cell_bits = J.key_bits = cell_key_bits(J.n_bins);
    J.s = s;
    J.s.key = keys_out;
    J.s.out = perm_out;
    J.s.perm = 1u;
    J.s.model_bits = 0u;
    const uint32_t nb = sort_tiles(n), passes = sort_plan(J.n_bins).passes, ng = sort_scan_groups(J.key_bits);
    for (uint32_t pass = 0; pass < passes; ++pass) {
        k_sortx_count<<<nb, kBlock, 0, stream>>>(J, pass);
        k_sortx_scan<<<ng, kBlock, 0, stream>>>(J, pass);
        k_sortx_scatter<<<nb, kBlock, 0, stream>>>(J, pass);
    }
    return hipGetLastError();
}

}  // namespace emh
