// kernels.hip -- the MLS-MPM substep as gfx950 HIP kernels.
//
// One substep (solver/step.rs:62-317) is, on the device:
//   k_control      1 workgroup: the CFL dt of choose_substep_dt/cfl_bound
//                  (step.rs:434-479) from per-block partials, the device-side
//                  substep loop of Simulation::step (step.rs:34-49), and the
//                  exclusive scan of the per-tile particle counts.
//   k_cell_scan    per 16x16 tile: exclusive scan of per-cell counts -> the
//                  start slot of every cell-key bin.
//  [k_density_*]   estimate_particle_volumes (solver/density.rs:216-247) when
//                  a fluid material needs it.
//   k_p2g          scatter_particles_to_grid (transfer/p2g.rs:24-72): chunks
//                  of 1024 particles (already in cell-key order from the
//                  previous substep) accumulate their 3x3 stencils in an LDS
//                  box and flush it with one global f32 atomic per touched
//                  cell; the chunk also scatters the permutation for this
//                  substep's sort.
//   k_grid_update  Grid::update_velocities + SlipBoundary + speed clamp
//                  (grid/mod.rs:264-291, boundary/mod.rs:78-101,
//                  step.rs:150-165), and clears the buffers the next substep
//                  accumulates into.
//   k_g2p          gather_grid_to_particles + update_particle
//                  (transfer/g2p.rs:151-343) for the particle at each sorted
//                  slot, written out in cell-key order; its epilogue computes
//                  the next substep's CFL partials, cell key and bin rank.
// Particle state projection (step.rs:347-429) is applied on the fly by every
// kernel that reads the state of a substep (it is a pure function of that
// state), and persisted by the kernel that writes the substep's output.
#include <hip/hip_runtime.h>

#include <climits>

#include "kernels.hpp"

namespace emh {

namespace {

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63u; }

__device__ __forceinline__ PBuf sel_buf(const Pipeline& p, uint32_t i) { return i ? p.buf[1] : p.buf[0]; }

__device__ __forceinline__ void gatomic_add(float* addr, float v) {
    __hip_atomic_fetch_add(addr, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Wave-level exclusive/inclusive scans and reductions (64 lanes).
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    const uint32_t lane = lane_id();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t n = __shfl_up(v, o, 64);
        if (lane >= (uint32_t)o) v += n;
    }
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

// Arrival rank of `key` among the lanes and waves that add to counts[key];
// same-key lanes of a wave share one atomic.  Every lane of the wave must call.
__device__ uint32_t agg_rank(uint32_t* counts, uint32_t key, bool valid) {
    const uint32_t lane = lane_id();
    const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
    uint32_t rank = 0;
    bool pending = valid;
    while (true) {
        uint64_t pm = __ballot(pending);
        if (pm == 0) break;
        int leader = __ffsll((unsigned long long)pm) - 1;
        uint32_t lk = __shfl(key, leader, 64);
        bool mine = pending && key == lk;
        uint64_t same = __ballot(mine);
        uint32_t base = 0;
        if (lane == (uint32_t)leader) base = atomicAdd(&counts[lk], (uint32_t)__popcll(same));
        base = __shfl(base, leader, 64);
        if (mine) {
            rank = base + (uint32_t)__popcll(same & lt);
            pending = false;
        }
    }
    return rank;
}
__device__ void agg_count(uint32_t* counts, uint32_t key, bool valid) {
    const uint32_t lane = lane_id();
    bool pending = valid;
    while (true) {
        uint64_t pm = __ballot(pending);
        if (pm == 0) break;
        int leader = __ffsll((unsigned long long)pm) - 1;
        uint32_t lk = __shfl(key, leader, 64);
        bool mine = pending && key == lk;
        uint64_t same = __ballot(mine);
        if (lane == (uint32_t)leader) atomicAdd(&counts[lk], (uint32_t)__popcll(same));
        if (mine) pending = false;
    }
}

__device__ __forceinline__ PState load_state(const PBuf& b, uint32_t i) {
    PState s;
    float2 x = b.x[i], v = b.v[i];
    s.x = x.x;
    s.y = x.y;
    s.vx = v.x;
    s.vy = v.y;
    s.c = from4(b.c[i]);
    s.f = from4(b.f[i]);
    s.mass = b.mass[i];
    s.v0 = b.v0[i];
    s.vol = b.vol[i];
    s.dens = b.dens[i];
    s.jp = b.jp[i];
    s.h = b.h[i];
    s.q = b.q[i];
    s.lvs = b.lvs[i];
    s.mat = b.mat[i];
    return s;
}
__device__ __forceinline__ void store_state(const PBuf& b, uint32_t i, const PState& s) {
    b.x[i] = make_float2(s.x, s.y);
    b.v[i] = make_float2(s.vx, s.vy);
    b.c[i] = to4(s.c);
    b.f[i] = to4(s.f);
    b.mass[i] = s.mass;
    b.v0[i] = s.v0;
    b.vol[i] = s.vol;
    b.dens[i] = s.dens;
    b.jp[i] = s.jp;
    b.h[i] = s.h;
    b.q[i] = s.q;
    b.lvs[i] = s.lvs;
    b.mat[i] = s.mat;
}

// Position part of project_particle_state_to_admissible (step.rs:357-362).
__device__ __forceinline__ void project_x(float& x, float& y, const DevConfig& cfg) {
    if (!(__builtin_isfinite(x) && __builtin_isfinite(y))) {
        float c = (cfg.proj_lo + cfg.proj_hi) * 0.5f;
        x = c;
        y = c;
    } else {
        x = rmin(rmax(x, cfg.proj_lo), cfg.proj_hi);
        y = rmin(rmax(y, cfg.proj_lo), cfg.proj_hi);
    }
}

__device__ __forceinline__ const DevMat& mat_of(const Pipeline& p, uint32_t word) {
    uint32_t m = word & kMatMask;  // MaterialRegistry::get falls back to slot 0 (registry.rs:66-76)
    return p.mats[m < p.n_mats ? m : 0u];
}

// Epilogue shared by the prologue and G2P: CFL contribution of the state as
// it stands (choose_substep_dt reads it before projection, step.rs:445-466),
// then the cell key of the projected position and its arrival rank.
__device__ void finish_particle(const Pipeline& p, const PState& s, bool valid, const PBuf& out, uint32_t dest,
                                float* s_red) {
    const DevConfig& cfg = p.cfg;
    float speed = 0.0f, mdt = kInf;
    uint32_t key = 0;
    if (valid) {
        float sp = len2(s.vx, s.vy);
        if (cfg.cfl_affine) sp = sp + affine_cfl(s.c, cfg.cell);
        speed = sp;
        float b = timestep_bound(mat_of(p, s.mat), s.dens, s.h, cfg);
        mdt = (__builtin_isfinite(b) && b > 0.0f) ? b : kInf;
        float px = s.x, py = s.y;
        if (cfg.project) project_x(px, py, cfg);
        key = cell_key(sat_i32(floorf(px)), sat_i32(floorf(py)), cfg.grid_res, cfg.tiles_per_dim);
    }
    uint32_t rank = agg_rank(p.sort.cell_count, key, valid);
    agg_count(p.sort.tile_count, key >> 8, valid);
    if (valid) {
        out.key[dest] = key;
        out.rank[dest] = rank;
    }
    // block reduction of (max speed, min dt) -> one partial per block
    float ws = wave_max(speed), wd = wave_min(mdt);
    const uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
    if (lane == 0) {
        s_red[2 * wave] = ws;
        s_red[2 * wave + 1] = wd;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        float a = s_red[0], d = s_red[1];
        for (uint32_t w = 1; w < (blockDim.x >> 6); ++w) {
            a = fmaxf(a, s_red[2 * w]);
            d = fminf(d, s_red[2 * w + 1]);
        }
        p.sort.cfl_partial[blockIdx.x] = make_float2(a, d);
    }
}

// ------------------------------------------------------------------ kernels

__global__ void k_zero_counters(const Pipeline p) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t R2 = p.cfg.grid_res * p.cfg.grid_res;
    if (i < p.cfg.n_bins) p.sort.cell_count[i] = 0;
    if (i < p.cfg.tiles_per_dim * p.cfg.tiles_per_dim) p.sort.tile_count[i] = 0;
    if (i < R2) {
        for (int g = 0; g < 2; ++g) {
            p.grid.m[g][i] = 0.0f;
            p.grid.mx[g][i] = 0.0f;
            p.grid.my[g][i] = 0.0f;
        }
        p.grid.vel[i] = make_float2(0.0f, 0.0f);
        p.grid.dmass[i] = 0.0f;
    }
}

// Keys, ranks and CFL partials of a freshly uploaded state.
__global__ __launch_bounds__(kBlock) void k_prologue(const Pipeline p) {
    __shared__ float s_red[2 * (kBlock / 64)];
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const PBuf b = sel_buf(p, p.ctrl->cur_buf);
    bool valid = i < p.n;
    PState s{};
    if (valid) s = load_state(b, i);
    finish_particle(p, s, valid, b, i, s_red);
}

__global__ void k_begin_step(const Pipeline p) {
    Ctrl* c = p.ctrl;
    c->remaining = p.cfg.dt;
    c->substeps = 0;
    c->active = 0;
    c->clamp_count = 0;
    c->proj_count = 0;
}

// One workgroup: substep loop control, CFL dt and the tile scan.
__global__ __launch_bounds__(1024) void k_control(const Pipeline p) {
    __shared__ float s_a[16], s_d[16];
    __shared__ uint32_t s_scan[17];
    Ctrl* c = p.ctrl;
    const DevConfig& cfg = p.cfg;
    const float remaining = c->remaining;
    const uint32_t subs = c->substeps;
    // while remaining > f32::EPSILON && substeps < max_substeps_per_step  (step.rs:34)
    if (!(remaining > kF32Eps && subs < cfg.max_substeps)) {
        if (threadIdx.x == 0) c->active = 0;
        return;
    }
    float ms = 0.0f, md = kInf;
    for (uint32_t i = threadIdx.x; i < c->n_cfl_blocks; i += blockDim.x) {
        float2 v = p.sort.cfl_partial[i];
        ms = fmaxf(ms, v.x);
        md = fminf(md, v.y);
    }
    ms = wave_max(ms);
    md = wave_min(md);
    const uint32_t lane = lane_id(), wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    if (lane == 0) {
        s_a[wave] = ms;
        s_d[wave] = md;
    }
    // tile scan: each thread sums a contiguous run of tiles
    const uint32_t n_tiles = cfg.tiles_per_dim * cfg.tiles_per_dim;
    const uint32_t per = (n_tiles + blockDim.x - 1) / blockDim.x;
    const uint32_t beg = threadIdx.x * per;
    uint32_t sum = 0;
    for (uint32_t k = 0; k < per; ++k)
        if (beg + k < n_tiles) sum += p.sort.tile_count[beg + k];
    uint32_t total;
    uint32_t run = block_excl_scan(sum, s_scan, total);  // contains __syncthreads
    for (uint32_t k = 0; k < per; ++k) {
        if (beg + k < n_tiles) {
            p.sort.tile_start[beg + k] = run;
            run += p.sort.tile_count[beg + k];
        }
    }
    if (threadIdx.x == 0) {
        p.sort.tile_start[n_tiles] = total;
        float a = s_a[0], d = s_d[0];
        for (uint32_t w = 1; w < nw; ++w) {
            a = fmaxf(a, s_a[w]);
            d = fminf(d, s_d[w]);
        }
        // choose_substep_dt / cfl_bound  step.rs:434-479
        float dt;
        if (!cfg.adaptive) {
            dt = rmin(remaining, cfg.dt);
        } else {
            float min_mat_dt = rmin(remaining, d);
            dt = remaining;
            if (a > kF32Eps) dt = rmin(dt, cfg.cfl_coefficient * cfg.cell / a);
            dt = rmin(dt, min_mat_dt);
            dt = rclamp(dt, rmin(cfg.min_dt, remaining), remaining);
        }
        c->dt = dt;
        c->active = 1;
        c->substeps = subs + 1;
        c->remaining = remaining - dt;
        c->last_dt = dt;
        c->total_substeps = c->total_substeps + 1;
        c->cur_grid ^= 1u;
        c->cur_buf ^= 1u;  // the state after this substep lives in the other buffer
    }
}

// Cell-level exclusive scan inside each 16x16 tile.
__global__ __launch_bounds__(kTileCells) void k_cell_scan(const Pipeline p) {
    __shared__ uint32_t s_scan[17];
    if (!p.ctrl->active) return;
    const uint32_t tile = blockIdx.x;
    const uint32_t idx = tile * kTileCells + threadIdx.x;
    uint32_t cnt = p.sort.cell_count[idx];
    uint32_t total;
    uint32_t ex = block_excl_scan(cnt, s_scan, total);
    p.sort.cell_start[idx] = p.sort.tile_start[tile] + ex;
    if (tile == gridDim.x - 1 && threadIdx.x == kTileCells - 1)
        p.sort.cell_start[p.cfg.n_bins] = p.sort.tile_start[gridDim.x];
}

// ---------------------------------------------------- wave segmented sums
// Lanes of a wave hold consecutive particles of a cell-sorted chunk, so runs
// of lanes share a base cell and with it all nine stencil cells.  The run's
// 27 contributions are summed across its lanes with DPP row shifts and row
// broadcasts (VALU work, no LDS traffic), leaving each run's total in its
// last lane, which alone issues the atomics.
struct SegMask {
    uint64_t heads;  // bit L: lane L starts a run
    uint32_t lane;
};
// Lanes in (src, lane] contain no run head <=> lane src is in this lane's run.
__device__ __forceinline__ bool same_run(const SegMask& m, int src) {
    if (src < 0) return false;
    const uint64_t upto_lane = (m.lane == 63) ? ~0ull : ((2ull << m.lane) - 1ull);
    const uint64_t upto_src = (2ull << (uint32_t)src) - 1ull;
    return (m.heads & (upto_lane & ~upto_src)) == 0ull;
}
template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp_take(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROWS, 0xf, true));
}
// Six-step segmented inclusive scan over the wave (row_shr 1/2/4/8 inside
// 16-lane rows, then row_bcast 15 and 31 across rows).
template <int N>
__device__ __forceinline__ void seg_scan(float (&v)[N], const SegMask& m) {
    const int l = (int)m.lane, r = l & 15;
    const bool p1 = r >= 1 && same_run(m, l - 1);
    const bool p2 = r >= 2 && same_run(m, l - 2);
    const bool p4 = r >= 4 && same_run(m, l - 4);
    const bool p8 = r >= 8 && same_run(m, l - 8);
    const bool pb15 = (l >= 16) && same_run(m, (l & ~15) - 1) && ((l >> 4) == 1 || (l >> 4) == 3);
    const bool pb31 = (l >= 32) && same_run(m, 31);
#pragma unroll
    for (int i = 0; i < N; ++i) {
        float t = dpp_take<0x111, 0xf>(v[i]);  // row_shr:1
        v[i] = v[i] + (p1 ? t : 0.0f);
    }
#pragma unroll
    for (int i = 0; i < N; ++i) {
        float t = dpp_take<0x112, 0xf>(v[i]);  // row_shr:2
        v[i] = v[i] + (p2 ? t : 0.0f);
    }
#pragma unroll
    for (int i = 0; i < N; ++i) {
        float t = dpp_take<0x114, 0xf>(v[i]);  // row_shr:4
        v[i] = v[i] + (p4 ? t : 0.0f);
    }
#pragma unroll
    for (int i = 0; i < N; ++i) {
        float t = dpp_take<0x118, 0xf>(v[i]);  // row_shr:8
        v[i] = v[i] + (p8 ? t : 0.0f);
    }
#pragma unroll
    for (int i = 0; i < N; ++i) {
        float t = dpp_take<0x142, 0xa>(v[i]);  // row_bcast:15 into rows 1 and 3
        v[i] = v[i] + (pb15 ? t : 0.0f);
    }
#pragma unroll
    for (int i = 0; i < N; ++i) {
        float t = dpp_take<0x143, 0xc>(v[i]);  // row_bcast:31 into rows 2 and 3
        v[i] = v[i] + (pb31 ? t : 0.0f);
    }
}
// Run structure of the base cells (bx, by) across the wave; every lane must
// call.  Invalid lanes pass a cell no valid lane can have (bx = INT_MIN + lane).
__device__ __forceinline__ SegMask run_mask(int bx, int by) {
    SegMask m;
    m.lane = lane_id();
    const int pbx = __shfl_up(bx, 1, 64), pby = __shfl_up(by, 1, 64);
    const bool head = (m.lane == 0) || pbx != bx || pby != by;
    m.heads = __ballot(head);
    return m;
}
__device__ __forceinline__ bool run_tail(const SegMask& m) {
    return m.lane == 63 || ((m.heads >> (m.lane + 1)) & 1ull);
}

// Chunk bounding box of base cells (positions projected) -> LDS box origin/size.
struct Box {
    int ox, oy, w, h;
    bool lds;
};
__device__ Box chunk_box(const Pipeline& p, const PBuf& a, uint32_t base, bool project, int* s_box) {
    int mnx = INT_MAX, mny = INT_MAX, mxx = INT_MIN, mxy = INT_MIN;
    for (int k = 0; k < kChunk / kBlock; ++k) {
        uint32_t j = base + k * kBlock + threadIdx.x;
        if (j < p.n) {
            float2 x = a.x[j];
            float px = x.x, py = x.y;
            if (project) project_x(px, py, p.cfg);
            int bx = sat_i32(floorf(px)), by = sat_i32(floorf(py));
            mnx = min(mnx, bx);
            mny = min(mny, by);
            mxx = max(mxx, bx);
            mxy = max(mxy, by);
        }
    }
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
    for (int w = 0; w < kBlock / 64; ++w) {
        mnx = min(mnx, s_box[4 * w + 0]);
        mny = min(mny, s_box[4 * w + 1]);
        mxx = max(mxx, s_box[4 * w + 2]);
        mxy = max(mxy, s_box[4 * w + 3]);
    }
    __syncthreads();
    Box b;
    b.ox = mnx - 1;
    b.oy = mny - 1;
    // widths in 64-bit to stay safe for any position range
    long long w = (long long)mxx - (long long)mnx + 3, h = (long long)mxy - (long long)mny + 3;
    b.lds = (w > 0 && h > 0 && w * h <= (long long)kLdsCells);
    b.w = b.lds ? (int)w : 0;
    b.h = b.lds ? (int)h : 0;
    return b;
}

// Density re-estimate, scatter half: Σ w·m per cell (p2g.rs:230-243).
__global__ __launch_bounds__(kBlock) void k_density_scatter(const Pipeline p) {
    __shared__ float s_m[kLdsCells];
    __shared__ int s_box[4 * (kBlock / 64)];
    const Ctrl* c = p.ctrl;
    if (!c->active) return;
    const PBuf a = sel_buf(p, c->cur_buf ^ 1u);  // state at the start of this substep
    const uint32_t base = blockIdx.x * kChunk;
    const uint32_t R = p.cfg.grid_res;
    Box bx = chunk_box(p, a, base, p.cfg.project != 0, s_box);
    if (bx.lds) {
        for (int i = threadIdx.x; i < bx.w * bx.h; i += kBlock) s_m[i] = 0.0f;
        __syncthreads();
    }
    for (int k = 0; k < kChunk / kBlock; ++k) {
        const uint32_t j = base + k * kBlock + threadIdx.x;
        const bool valid = j < p.n;
        float vals[9];
        int bxc = INT_MIN + (int)lane_id(), byc = 0;
        if (valid) {
            PState s = load_state(a, j);
            if (p.cfg.project) project_state(s, p.cfg);
            Wts w = weights(s.x, s.y);
            bxc = w.bx;
            byc = w.by;
#pragma unroll
            for (int gx = 0; gx < 3; ++gx)
#pragma unroll
                for (int gy = 0; gy < 3; ++gy) vals[gx * 3 + gy] = (w.wx[gx] * w.wy[gy]) * s.mass;
        } else {
#pragma unroll
            for (int q = 0; q < 9; ++q) vals[q] = 0.0f;
        }
        const SegMask sm = run_mask(bxc, byc);
        seg_scan(vals, sm);
        if (valid && run_tail(sm)) {
            for (int gx = 0; gx < 3; ++gx)
                for (int gy = 0; gy < 3; ++gy) {
                    const float wm = vals[gx * 3 + gy];
                    const int cx = bxc + gx - 1, cy = byc + gy - 1;
                    if (bx.lds) {
                        atomicAdd(&s_m[(cx - bx.ox) * bx.h + (cy - bx.oy)], wm);
                    } else if (cx >= 0 && cy >= 0 && cx < (int)R && cy < (int)R) {
                        gatomic_add(&p.grid.dmass[(uint32_t)cx * R + (uint32_t)cy], wm);
                    }
                }
        }
    }
    if (bx.lds) {
        __syncthreads();
        for (int i = threadIdx.x; i < bx.w * bx.h; i += kBlock) {
            float m = s_m[i];
            if (m == 0.0f) continue;
            int cx = bx.ox + i / bx.h, cy = bx.oy + i % bx.h;
            if (cx >= 0 && cy >= 0 && cx < (int)R && cy < (int)R) gatomic_add(&p.grid.dmass[(uint32_t)cx * R + cy], m);
        }
    }
}

// Density re-estimate, gather half (density.rs:225-246), in place; this pass
// also persists the substep's projected state, so later kernels of the
// substep do not project again.
__global__ __launch_bounds__(kBlock) void k_density_gather(const Pipeline p) {
    const Ctrl* c = p.ctrl;
    if (!c->active) return;
    const PBuf a = sel_buf(p, c->cur_buf ^ 1u);
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= p.n) return;
    const uint32_t R = p.cfg.grid_res;
    PState s = load_state(a, j);
    bool projected = false;
    if (p.cfg.project) {
        projected = project_state(s, p.cfg);
        if (projected) atomicAdd(&p.ctrl->proj_count, 1u);
    }
    Wts w = weights(s.x, s.y);
    float density = 0.0f;
    for (int gx = 0; gx < 3; ++gx)
        for (int gy = 0; gy < 3; ++gy) {
            float weight = w.wx[gx] * w.wy[gy];
            int cx = w.bx + gx - 1, cy = w.by + gy - 1;
            float m = 0.0f;
            if (cx >= 0 && cy >= 0 && cx < (int)R && cy < (int)R) m = p.grid.dmass[(uint32_t)cx * R + cy];
            density = density + m * weight;
        }
    if (density > kF32Eps) {
        s.dens = density;
        s.vol = s.mass / density;
    }
    if (projected) {
        store_state(a, j, s);
    } else {
        a.x[j] = make_float2(s.x, s.y);
        a.dens[j] = s.dens;
        a.vol[j] = s.vol;
    }
}

// P2G.  PROJECT: apply the state projection on the fly (no density pass ran).
template <bool PROJECT>
__global__ __launch_bounds__(kBlock) void k_p2g(const Pipeline p) {
    __shared__ float s_m[kLdsCells], s_mx[kLdsCells], s_my[kLdsCells];
    __shared__ int s_box[4 * (kBlock / 64)];
    const Ctrl* c = p.ctrl;
    if (!c->active) return;
    const PBuf a = sel_buf(p, c->cur_buf ^ 1u);
    const float dt = c->dt;
    const uint32_t g = c->cur_grid;
    float* gm = g ? p.grid.m[1] : p.grid.m[0];
    float* gmx = g ? p.grid.mx[1] : p.grid.mx[0];
    float* gmy = g ? p.grid.my[1] : p.grid.my[0];
    const uint32_t base = blockIdx.x * kChunk;
    const uint32_t R = p.cfg.grid_res;

    // this substep's counting-sort scatter: slot = bin start + arrival rank
    for (int k = 0; k < kChunk / kBlock; ++k) {
        uint32_t j = base + k * kBlock + threadIdx.x;
        if (j < p.n) p.sort.perm[p.sort.cell_start[a.key[j]] + a.rank[j]] = j;
    }

    Box bx = chunk_box(p, a, base, PROJECT && p.cfg.project, s_box);
    if (bx.lds) {
        for (int i = threadIdx.x; i < bx.w * bx.h; i += kBlock) {
            s_m[i] = 0.0f;
            s_mx[i] = 0.0f;
            s_my[i] = 0.0f;
        }
        __syncthreads();
    }
    for (int k = 0; k < kChunk / kBlock; ++k) {
        const uint32_t j = base + k * kBlock + threadIdx.x;
        const bool valid = j < p.n;
        // contributions to the 3x3 stencil: mass, momentum x, momentum y
        float vals[27];
        int bxc = INT_MIN + (int)lane_id(), byc = 0;
        if (valid) {
            PState s = load_state(a, j);
            if (PROJECT && p.cfg.project) project_state(s, p.cfg);
            const DevMat& mt = mat_of(p, s.mat);
            const M2 tau = kirchhoff(mt, s);
            const float coeff = ((-stress_volume(mt, s)) * kDInv) * dt;
            const Wts w = weights(s.x, s.y);
            bxc = w.bx;
            byc = w.by;
#pragma unroll
            for (int gx = 0; gx < 3; ++gx)
#pragma unroll
                for (int gy = 0; gy < 3; ++gy) {
                    const float weight = w.wx[gx] * w.wy[gy];
                    const float dx = ((float)(w.bx + gx - 1) - s.x) + 0.5f;
                    const float dy = ((float)(w.by + gy - 1) - s.y) + 0.5f;
                    const float vcx = s.vx + (s.c.a * dx + s.c.c * dy);
                    const float vcy = s.vy + (s.c.b * dx + s.c.d * dy);
                    const float sdx = tau.a * dx + tau.c * dy;
                    const float sdy = tau.b * dx + tau.d * dy;
                    vals[gx * 3 + gy] = weight * s.mass;
                    vals[9 + gx * 3 + gy] = weight * (s.mass * vcx + coeff * sdx);
                    vals[18 + gx * 3 + gy] = weight * (s.mass * vcy + coeff * sdy);
                }
        } else {
#pragma unroll
            for (int q = 0; q < 27; ++q) vals[q] = 0.0f;
        }
        const SegMask sm = run_mask(bxc, byc);
        seg_scan(vals, sm);
        if (valid && run_tail(sm)) {
            for (int gx = 0; gx < 3; ++gx)
                for (int gy = 0; gy < 3; ++gy) {
                    const int cx = bxc + gx - 1, cy = byc + gy - 1;
                    const float mm = vals[gx * 3 + gy], mx = vals[9 + gx * 3 + gy], my = vals[18 + gx * 3 + gy];
                    if (bx.lds) {
                        const int li = (cx - bx.ox) * bx.h + (cy - bx.oy);
                        atomicAdd(&s_m[li], mm);
                        atomicAdd(&s_mx[li], mx);
                        atomicAdd(&s_my[li], my);
                    } else if (cx >= 0 && cy >= 0 && cx < (int)R && cy < (int)R) {
                        const uint32_t gi = (uint32_t)cx * R + (uint32_t)cy;
                        gatomic_add(&gm[gi], mm);
                        gatomic_add(&gmx[gi], mx);
                        gatomic_add(&gmy[gi], my);
                    }
                }
        }
    }
    if (bx.lds) {
        __syncthreads();
        for (int i = threadIdx.x; i < bx.w * bx.h; i += kBlock) {
            float m = s_m[i], mx = s_mx[i], my = s_my[i];
            if (m == 0.0f && mx == 0.0f && my == 0.0f) continue;
            int cx = bx.ox + i / bx.h, cy = bx.oy + i % bx.h;
            if (cx < 0 || cy < 0 || cx >= (int)R || cy >= (int)R) continue;
            uint32_t gi = (uint32_t)cx * R + (uint32_t)cy;
            if (m != 0.0f) gatomic_add(&gm[gi], m);
            if (mx != 0.0f) gatomic_add(&gmx[gi], mx);
            if (my != 0.0f) gatomic_add(&gmy[gi], my);
        }
    }
}

// Grid update (normalise, gravity, slip walls, speed clamp) and clears.
__global__ __launch_bounds__(kBlock) void k_grid_update(const Pipeline p) {
    const Ctrl* c = p.ctrl;
    if (!c->active) return;
    const DevConfig& cfg = p.cfg;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t R = cfg.grid_res, R2 = R * R;
    if (i < cfg.n_bins) p.sort.cell_count[i] = 0;
    if (i < cfg.tiles_per_dim * cfg.tiles_per_dim) p.sort.tile_count[i] = 0;
    if (i >= R2) return;
    const uint32_t g = c->cur_grid;
    const float dt = c->dt;
    float* gm = g ? p.grid.m[1] : p.grid.m[0];
    float* gmx = g ? p.grid.mx[1] : p.grid.mx[0];
    float* gmy = g ? p.grid.my[1] : p.grid.my[0];
    float m = gm[i], vx = gmx[i], vy = gmy[i];
    if (m > 0.0f) {
        vx = vx / m;
        vy = vy / m;
        vx = vx + cfg.gx * dt;
        vy = vy + cfg.gy * dt;
        const uint32_t x = i / R, y = i % R;
        for (uint32_t b = 0; b < cfg.n_slip; ++b) {
            const uint32_t t = cfg.slip[b];
            const uint32_t hi = R - (t + 1);
            if (x < t) vx = rmax(vx, 0.0f);
            if (x > hi) vx = rmin(vx, 0.0f);
            if (y < t) vy = rmax(vy, 0.0f);
            if (y > hi) vy = rmin(vy, 0.0f);
        }
        const float vel_limit = cfg.cell / dt;
        float spd = len2(vx, vy);
        if (spd > vel_limit) {
            float k = vel_limit / spd;
            vx = vx * k;
            vy = vy * k;
        }
    }
    p.grid.vel[i] = make_float2(vx, vy);
    // clear the accumulators the next substep's P2G (and density pass) will use
    float* om = g ? p.grid.m[0] : p.grid.m[1];
    float* omx = g ? p.grid.mx[0] : p.grid.mx[1];
    float* omy = g ? p.grid.my[0] : p.grid.my[1];
    om[i] = 0.0f;
    omx[i] = 0.0f;
    omy[i] = 0.0f;
    if (cfg.density_pass) p.grid.dmass[i] = 0.0f;
}

// G2P + constitutive update for the particle at each sorted slot.
template <bool PROJECT>
__global__ __launch_bounds__(kBlock) void k_g2p(const Pipeline p) {
    __shared__ float s_red[2 * (kBlock / 64)];
    Ctrl* c = p.ctrl;
    if (!c->active) return;
    const DevConfig& cfg = p.cfg;
    const uint32_t wb = c->cur_buf;  // output buffer (state after this substep)
    const PBuf a = sel_buf(p, wb ^ 1u);
    const PBuf o = sel_buf(p, wb);
    const float dt = c->dt;
    const uint32_t R = cfg.grid_res;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = t < p.n;
    PState s{};
    uint32_t dest = 0;
    if (valid) {
        // stable position of this slot's particle inside its bin: the bin is
        // ordered by the particles' positions in the input buffer
        const uint32_t e = p.sort.perm[t];
        const uint32_t key = a.key[e];
        const uint32_t start = p.sort.cell_start[key];
        const uint32_t cnt = p.sort.cell_start[key + 1] - start;
        uint32_t rank = 0;
        for (uint32_t u = 0; u < cnt; ++u) rank += (p.sort.perm[start + u] < e) ? 1u : 0u;
        dest = start + rank;

        s = load_state(a, e);
        if (PROJECT && cfg.project) {
            if (project_state(s, cfg)) atomicAdd(&c->proj_count, 1u);
        }
        // gather_grid_to_particles phase 1 (g2p.rs:203-325)
        Wts w = weights(s.x, s.y);
        float nvx = 0.0f, nvy = 0.0f;
        M2 b{0.0f, 0.0f, 0.0f, 0.0f};
        for (int gx = 0; gx < 3; ++gx)
            for (int gy = 0; gy < 3; ++gy) {
                float weight = w.wx[gx] * w.wy[gy];
                int cx = w.bx + gx - 1, cy = w.by + gy - 1;
                float dx = ((float)cx - s.x) + 0.5f;
                float dy = ((float)cy - s.y) + 0.5f;
                float2 nv = make_float2(0.0f, 0.0f);
                if (cx >= 0 && cy >= 0 && cx < (int)R && cy < (int)R) nv = p.grid.vel[(uint32_t)cx * R + cy];
                float wvx = nv.x * weight, wvy = nv.y * weight;
                b.a = b.a + wvx * dx;
                b.b = b.b + wvy * dx;
                b.c = b.c + wvx * dy;
                b.d = b.d + wvy * dy;
                nvx = nvx + wvx;
                nvy = nvy + wvy;
            }
        if (s.mat & kPinnedBit) {
            s.vx = 0.0f;
            s.vy = 0.0f;
            s.c = M2{0.0f, 0.0f, 0.0f, 0.0f};
        } else {
            const float vel_limit = cfg.cell / dt;
            float spd = len2(nvx, nvy);
            if (spd > vel_limit) {
                float k = vel_limit / spd;
                nvx = nvx * k;
                nvy = nvy * k;
                atomicAdd(&c->clamp_count, 1u);
            }
            float nx = s.x + nvx * dt, ny = s.y + nvy * dt;
            for (uint32_t q = 0; q < cfg.n_slip; ++q) {
                nx = rmin(rmax(nx, cfg.slip_lo[q]), cfg.slip_hi[q]);
                ny = rmin(rmax(ny, cfg.slip_lo[q]), cfg.slip_hi[q]);
            }
            s.vx = nvx;
            s.vy = nvy;
            s.c = M2{(b.a * kDInv) * cfg.apic_blend, (b.b * kDInv) * cfg.apic_blend, (b.c * kDInv) * cfg.apic_blend,
                     (b.d * kDInv) * cfg.apic_blend};
            s.x = nx;
            s.y = ny;
        }
        // phase 2: update_particle (g2p.rs:333-340)
        update_particle(mat_of(p, s.mat), s, dt);
        store_state(o, dest, s);
        o.id[dest] = a.id[e];
    }
    finish_particle(p, s, valid, o, dest, s_red);
}

// ------------------------------------------------- stand-alone cell sort
__global__ void k_sort_keys(const float2* xy, uint32_t n, uint32_t R, uint32_t tpd, uint32_t* keys, uint32_t* rank,
                            uint32_t* cell_count, uint32_t* tile_count) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    bool valid = i < n;
    uint32_t key = 0;
    if (valid) {
        float2 x = xy[i];
        key = cell_key(sat_i32(floorf(x.x)), sat_i32(floorf(x.y)), R, tpd);
    }
    uint32_t r = agg_rank(cell_count, key, valid);
    agg_count(tile_count, key >> 8, valid);
    if (valid) {
        keys[i] = key;
        rank[i] = r;
    }
}
__global__ __launch_bounds__(1024) void k_sort_tile_scan(uint32_t n_tiles, const uint32_t* tile_count,
                                                         uint32_t* tile_start) {
    __shared__ uint32_t s_scan[17];
    const uint32_t per = (n_tiles + blockDim.x - 1) / blockDim.x;
    const uint32_t beg = threadIdx.x * per;
    uint32_t sum = 0;
    for (uint32_t k = 0; k < per; ++k)
        if (beg + k < n_tiles) sum += tile_count[beg + k];
    uint32_t total;
    uint32_t run = block_excl_scan(sum, s_scan, total);
    for (uint32_t k = 0; k < per; ++k)
        if (beg + k < n_tiles) {
            tile_start[beg + k] = run;
            run += tile_count[beg + k];
        }
    if (threadIdx.x == 0) tile_start[n_tiles] = total;
}
__global__ __launch_bounds__(kTileCells) void k_sort_cell_scan(uint32_t n_bins, const uint32_t* cell_count,
                                                               const uint32_t* tile_start, uint32_t* cell_start) {
    __shared__ uint32_t s_scan[17];
    const uint32_t tile = blockIdx.x;
    const uint32_t idx = tile * kTileCells + threadIdx.x;
    uint32_t total;
    uint32_t ex = block_excl_scan(cell_count[idx], s_scan, total);
    cell_start[idx] = tile_start[tile] + ex;
    if (tile == gridDim.x - 1 && threadIdx.x == kTileCells - 1) cell_start[n_bins] = tile_start[gridDim.x];
}
__global__ void k_sort_scatter(uint32_t n, const uint32_t* keys, const uint32_t* rank, const uint32_t* cell_start,
                               uint32_t* perm_unordered) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) perm_unordered[cell_start[keys[i]] + rank[i]] = i;
}
__global__ void k_sort_finalize(uint32_t n, const uint32_t* keys, const uint32_t* cell_start,
                                const uint32_t* perm_unordered, uint32_t* perm_sorted) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    uint32_t e = perm_unordered[t];
    uint32_t key = keys[e];
    uint32_t start = cell_start[key], cnt = cell_start[key + 1] - start;
    uint32_t r = 0;
    for (uint32_t u = 0; u < cnt; ++u) r += (perm_unordered[start + u] < e) ? 1u : 0u;
    perm_sorted[start + r] = e;
}

inline uint32_t div_up(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

}  // namespace

// ------------------------------------------------------------ launchers

hipError_t launch_zero_counters(const Pipeline& p, hipStream_t stream) {
    uint32_t R2 = p.cfg.grid_res * p.cfg.grid_res;
    uint32_t n = R2 > p.cfg.n_bins ? R2 : p.cfg.n_bins;
    k_zero_counters<<<div_up(n, kBlock), kBlock, 0, stream>>>(p);
    return hipGetLastError();
}

hipError_t launch_prologue(const Pipeline& p, hipStream_t stream) {
    if (p.n == 0) return hipSuccess;
    k_prologue<<<div_up(p.n, kBlock), kBlock, 0, stream>>>(p);
    return hipGetLastError();
}

hipError_t launch_begin_step(const Pipeline& p, hipStream_t stream) {
    k_begin_step<<<1, 1, 0, stream>>>(p);
    return hipGetLastError();
}

const char* const kKernelNames[kKernelCount] = {"k_control",        "k_cell_scan", "k_density_scatter",
                                                  "k_density_gather", "k_p2g",       "k_grid_update",
                                                  "k_g2p"};

hipError_t launch_substep(const Pipeline& p, hipStream_t stream, SlotEvents* events) {
    if (p.n == 0) return hipSuccess;
    const uint32_t n_tiles = p.cfg.tiles_per_dim * p.cfg.tiles_per_dim;
    const uint32_t R2 = p.cfg.grid_res * p.cfg.grid_res;
    const uint32_t cells = R2 > p.cfg.n_bins ? R2 : p.cfg.n_bins;
    const bool dens = p.cfg.density_pass != 0;
    auto mark = [&](int id, int which) {
        if (!events) return;
        (void)hipEventRecord(events->ev[id][which], stream);
        events->used[id] = true;
    };
    if (events)
        for (int k = 0; k < kKernelCount; ++k) events->used[k] = false;
    mark(kKControl, 0);
    k_control<<<1, 1024, 0, stream>>>(p);
    mark(kKControl, 1);
    mark(kKCellScan, 0);
    k_cell_scan<<<n_tiles, kTileCells, 0, stream>>>(p);
    mark(kKCellScan, 1);
    if (dens) {
        mark(kKDensityScatter, 0);
        k_density_scatter<<<div_up(p.n, kChunk), kBlock, 0, stream>>>(p);
        mark(kKDensityScatter, 1);
        mark(kKDensityGather, 0);
        k_density_gather<<<div_up(p.n, kBlock), kBlock, 0, stream>>>(p);
        mark(kKDensityGather, 1);
        mark(kKP2G, 0);
        k_p2g<false><<<div_up(p.n, kChunk), kBlock, 0, stream>>>(p);
        mark(kKP2G, 1);
    } else {
        mark(kKP2G, 0);
        k_p2g<true><<<div_up(p.n, kChunk), kBlock, 0, stream>>>(p);
        mark(kKP2G, 1);
    }
    mark(kKGridUpdate, 0);
    k_grid_update<<<div_up(cells, kBlock), kBlock, 0, stream>>>(p);
    mark(kKGridUpdate, 1);
    mark(kKG2P, 0);
    if (dens) {
        k_g2p<false><<<div_up(p.n, kBlock), kBlock, 0, stream>>>(p);
    } else {
        k_g2p<true><<<div_up(p.n, kBlock), kBlock, 0, stream>>>(p);
    }
    mark(kKG2P, 1);
    return hipGetLastError();
}

hipError_t sort_cell_keys_device(const float2* xy, uint32_t n, uint32_t grid_res, uint32_t* keys,
                                 uint32_t* perm_sorted, uint32_t* scratch_rank, uint32_t* cell_count,
                                 uint32_t* cell_start, uint32_t* tile_count, uint32_t* tile_start,
                                 uint32_t* perm_unordered, hipStream_t stream) {
    const uint32_t tpd = (grid_res + 15u) / 16u;
    const uint32_t n_tiles = tpd * tpd, n_bins = n_tiles * kTileCells;
    hipError_t e = hipMemsetAsync(cell_count, 0, sizeof(uint32_t) * n_bins, stream);
    if (e != hipSuccess) return e;
    e = hipMemsetAsync(tile_count, 0, sizeof(uint32_t) * n_tiles, stream);
    if (e != hipSuccess) return e;
    if (n == 0) return hipSuccess;
    k_sort_keys<<<div_up(n, kBlock), kBlock, 0, stream>>>(xy, n, grid_res, tpd, keys, scratch_rank, cell_count,
                                                          tile_count);
    k_sort_tile_scan<<<1, 1024, 0, stream>>>(n_tiles, tile_count, tile_start);
    k_sort_cell_scan<<<n_tiles, kTileCells, 0, stream>>>(n_bins, cell_count, tile_start, cell_start);
    k_sort_scatter<<<div_up(n, kBlock), kBlock, 0, stream>>>(n, keys, scratch_rank, cell_start, perm_unordered);
    k_sort_finalize<<<div_up(n, kBlock), kBlock, 0, stream>>>(n, keys, cell_start, perm_unordered, perm_sorted);
    return hipGetLastError();
}

}  // namespace emh
