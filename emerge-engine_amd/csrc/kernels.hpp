// kernels.hpp -- device buffers and launch wrappers of the substep pipeline.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_math.hpp"
#include "surrogate.hpp"

namespace emh {

constexpr int kBlock = 256;        // threads per workgroup for particle kernels (4 waves of 64)
constexpr int kWaves = kBlock / 64;
constexpr int kChunk = 1024;       // particles per substep workgroup (4 per thread)
constexpr int kPer = kChunk / kBlock;
// LDS box budget in grid cells: five f64 accumulators and the staged f32
// velocity per cell (48 B: 36 KiB, with the u16 processing order under the
// 40 KiB that keeps four 256-lane workgroups per CU).  A multiple of kBlock.
constexpr int kLdsCells = 768;
// Default interval of the physical cell-key reorder, in substeps (the upload
// is always reordered at the first step).  The in-chunk sort keeps the runs
// short between reorders, and on the 1M-particle sand column no interval pays:
// 8-160 substeps run at 11.5e9-11.8e9 particle-substeps/s, no reorder after
// the upload at 12.8e9 (a reorder mid-collapse leaves the later substeps
// ~6 us slower).  Long runs still restore cell-key order now and then.
constexpr uint32_t kReorderSubsteps = 4096;
// The batched scenes' interval: their small grids and hundreds of substeps per
// step do gain from frequent reorders (256 scenes: 5.7e9 particle-substeps/s
// at 8 substeps, 4.5e9 at 4096).
constexpr uint32_t kBatchReorderSubsteps = 8;

// Particle state, structure of arrays.  Two of these ping-pong only when the
// particles are physically reordered; other substeps update in place.
struct PBuf {
    float2* x;
    float2* v;
    float4* c;  // velocity_gradient, column-major
    float4* f;  // deformation_gradient, column-major
    float* mass;
    float* v0;  // initial_volume
    float* vol;
    float* dens;
    float* jp;   // plastic_volume_ratio
    float* h;    // hardening_scale
    float* q;    // friction_hardening
    float* lvs;  // log_volume_strain
    uint32_t* mat;  // material_id | pinned bit
    uint32_t* id;   // index of the particle in the caller's order
};

// Dense grid, x-major (cell (x, y) at x * R + y).  Three rotating sets of P2G
// accumulators: the substep being gathered reads one, the particle kernel
// scatters the next substep's P2G into the second and clears the third.
// P2G runs before the CFL reduction has fixed the substep's dt, so it uses a
// predicted dt' and also keeps the stress moment S = sum w (-4 V0) (tau d):
// the grid update forms momentum = A + (dt - dt') S, which is the reference's
// sum w (m (v + C d) + (-4 V0 dt) (tau d)) exactly when dt' = dt.
constexpr int kAccM = 0, kAccPx = 1, kAccPy = 2, kAccSx = 3, kAccSy = 4, kAccChannels = 5;
struct GridBuf {
    float* acc[3][kAccChannels];  // mass, A (x, y), S (x, y)
    float* dm[3];  // mass grid of the density re-estimate pass
};

// Stable LSD radix sort by cell key (the pipeline's reorder and the
// stand-alone sort share it).  A pass ranks kSortTile keys per workgroup on
// one digit of at most kSortMaxDigit bits: a per-(digit, workgroup) count, an
// exclusive scan of each digit's counts over the workgroups with the digit's
// total beside it, and a scatter that adds the digits' scanned totals and
// ranks each key among the equal digits before it in index order.
// Keys of k bits take ceil(k / kSortMaxDigit) passes of equal width.
constexpr int kSortItems = 8;                                 // keys per lane
constexpr uint32_t kSortTile = (uint32_t)kBlock * kSortItems;  // keys per workgroup
constexpr uint32_t kSortMaxDigit = 10;                         // at most 1024 buckets
struct SortPlan {
    uint32_t key_bits, passes, digit;
};
__host__ __device__ inline uint32_t cell_key_bits(uint32_t n_bins) {
    uint32_t b = 1;
    while (b < 32u && (1u << b) < n_bins) ++b;
    return b;
}
__host__ __device__ inline SortPlan sort_plan_bits(uint32_t bits) {
    SortPlan s;
    s.key_bits = bits;
    s.passes = (bits + kSortMaxDigit - 1) / kSortMaxDigit;
    s.digit = (bits + s.passes - 1) / s.passes;
    return s;
}
__host__ __device__ inline SortPlan sort_plan(uint32_t n_bins) { return sort_plan_bits(cell_key_bits(n_bins)); }
// Bits of the constitutive model above the cell key when the reorder groups
// particles by model (per-material segments): model ids are below 16.
constexpr uint32_t kModelKeyBits = 4;
__host__ __device__ constexpr uint32_t sort_tiles(uint32_t n) { return (n + kSortTile - 1) / kSortTile; }
// Counts of one pass, (1 << digit) buckets x sort_tiles(n) workgroups, then
// the (1 << digit) digit totals.
__host__ __device__ inline size_t sort_hist_len(uint32_t n, uint32_t n_bins) {
    return ((size_t)1 << sort_plan(n_bins).digit) * ((size_t)sort_tiles(n) + 1u);
}
// Enough for any key width: at most 2^kSortMaxDigit buckets.
__host__ __device__ constexpr size_t sort_hist_len_max(uint32_t n) {
    return ((size_t)1 << kSortMaxDigit) * ((size_t)sort_tiles(n) + 1u);
}
// Workgroups of a pass's scan: one wave per digit row.
constexpr uint32_t kSortRowsPerGroup = (uint32_t)kBlock / 64u;
__host__ __device__ inline uint32_t sort_scan_groups(uint32_t key_bits) {
    return ((1u << sort_plan_bits(key_bits).digit) + kSortRowsPerGroup - 1u) / kSortRowsPerGroup;
}

struct SortBuf {
    uint32_t* key;      // N: cell key of the particle in each slot (written by pass 0 unless null)
    uint32_t* kbuf[2];  // N each: keys in the order after pass 0, 1
    uint32_t* vbuf[2];  // N each: the slot each of those keys came from
    uint32_t* out;      // N: the result (see perm)
    uint32_t* hist;     // per-pass counts, then start slots, and the digit totals (sort_hist_len*)
    uint32_t perm;      // 1: out[sorted position] = slot; 0: out[slot] = sorted position
    uint32_t model_bits;  // pipeline reorder: kModelKeyBits to sort by (model, cell key), 0 by cell key
};

// Device-resident control: the adaptive substep loop runs on the device so
// that neither a substep nor a step needs a host round trip, and no kernel
// synchronises its workgroups with each other.  The host queues step starts
// (k_begin_step adds one to `pending`) and substep launches; a substep launch
// that finds the current step finished (step.rs:34) starts the next pending
// step, so launches left over by a short step serve the next one and the
// host only checks at the end that every queued step has finished.  Every launch that takes part in the loop has a
// sequence number s (counted by the host).  Launch s reads st[s & 1] (written
// by launch s - 1) and the reduction red[(s + 2) % 3] of launch s - 1; every
// workgroup derives the same decision for its substep from those, workgroup 0
// alone writes st[(s + 1) & 1] and clears red[(s + 1) % 3], and all
// workgroups reduce into red[s % 3] with atomics.
struct CtrlState {
    float remaining;   // simulated time left in the current step
    float dt;          // dt of the last decided substep
    float last_dt;
    uint32_t substeps;  // substeps taken in the current step
    uint32_t active;    // the launch that wrote this state ran a substep
    uint32_t total_substeps;
    uint32_t cur_buf;   // which PBuf holds the state
    uint32_t cur_grid;  // accumulator set the last substep gathered from (0..2)
    uint32_t reorder;   // the next substep writes the particles in cell-key order
    float pred_dt;      // dt' the accumulators of the next substep were scattered with
    float free_dt;      // the last decision's dt before the clamp to the time left
    float last_ddt;     // dt - dt' the last substep's grid update applied
    uint32_t in_step;   // a step is running
    uint32_t pending;   // queued step starts
    uint32_t step;      // number of the running (or last) step, from 1
    uint32_t done_steps;     // finished steps
    uint32_t done_step;      // number of the last finished step
    uint32_t done_substeps;  // its substeps
    float done_dt;           // its last substep's dt
    float done_dropped;      // its unsimulated time, max(remaining, 0)
};
struct CtrlRed {
    uint32_t max_speed;  // f32 bits (non-negative: ordered as integers)
    uint32_t min_mdt;    // f32 bits (positive or +inf)
    uint32_t dirty;      // some stored particle is changed by projection
    uint32_t pad;
};
constexpr uint32_t kStepRing = 64;  // per-step counters, indexed by step % kStepRing
struct Ctrl {
    CtrlState st[2];
    CtrlRed red[3];
    uint32_t clamp_count[kStepRing];  // G2P velocity clamps of each step
    uint32_t proj_count[kStepRing];   // state projections of each step
    SgDev sg;  // learned NACC return mapping (read by the harvest and surrogate substeps only)
    float sleep_sq;  // sleep_threshold^2 (f32), read by the sleep-partition substeps only
    uint32_t sleep_pad;
};
constexpr uint32_t kInfBits = 0x7f800000u;

struct Pipeline {
    DevConfig cfg;
    const DevMat* mats;  // device table
    uint32_t n_mats;
    uint32_t n;          // particle count
    PBuf buf[2];
    GridBuf grid;
    SortBuf sort;
    Ctrl* ctrl;
};

// Launch wrappers (all asynchronous on `stream`); `seq` is the launch's
// sequence number in the device-side loop (see Ctrl).
hipError_t launch_zero_all(const Pipeline& p, hipStream_t stream);
hipError_t launch_prologue(const Pipeline& p, uint64_t seq, hipStream_t stream, bool sleep = false);
hipError_t launch_begin_step(const Pipeline& p, uint64_t seq, bool reorder, hipStream_t stream);
// sg_mode: 0 plain, 1 harvest NACC trial states, 2 NACC return mapping through
// the learned surrogate (Ctrl::sg); modes 1 and 2 need a NACC-only scene.
hipError_t launch_substep(const Pipeline& p, uint64_t seq, hipStream_t stream, uint32_t sg_mode = 0,
                          bool sleep = false);

// Work list of one batch launch stage: workgroup b belongs to scene s with
// start[s] <= b < start[s + 1] (device arrays of n + 1 entries).
struct BatchView {
    const Pipeline* __restrict__ pipes;  // device array of the scenes' pipelines (read-only in every kernel)
    const uint32_t* __restrict__ start;
    uint32_t n;
};
struct BatchLists {
    BatchView one;      // one workgroup per scene (step start, control gather)
    BatchView sort;     // sort_tiles(n_s) workgroups (cell-key sort counts and scatters)
    BatchView scan;     // sort_scan_groups(key bits of s) workgroups (the sort's count scans)
    BatchView fused;    // chunks of the scenes without a density pass (fused substep)
    BatchView density;  // chunks of the scenes with one (density + P2G, then G2P + mass)
    uint32_t sort_blocks, scan_blocks, fused_blocks, density_blocks;
    uint32_t sort_passes;  // the most radix passes any scene's key needs
};
// One launch per stage over every scene of the batch (the single-scene
// kernels' bodies, ONLY = mixed).
hipError_t launch_batch_begin_step(const BatchLists& L, uint64_t seq, bool reorder, hipStream_t stream);
// The radix sort of every scene's keys, after a reorder step's begin_step (same seq).
hipError_t launch_batch_sort(const BatchLists& L, uint64_t seq, hipStream_t stream);
hipError_t launch_batch_substep(const BatchLists& L, uint64_t seq, hipStream_t stream);
hipError_t launch_batch_gather_ctrl(const BatchLists& L, Ctrl* out, hipStream_t stream);

// Device accumulator of the diagnostics snapshot (see launch_snapshot).
struct DevSnapshot {
    double mass, mom_x, mom_y, kinetic, jp_sum, h_sum;
    double grid_mass, grid_mom_x, grid_mom_y;
    uint32_t max_speed, max_pinned_speed, max_grid_speed;  // f32 bits (non-negative)
    uint32_t min_j, max_j, min_jp;                        // order-preserving f32 keys
    uint32_t valid, non_finite, invalid, out_of_bounds, jp_finite;
    uint32_t active_cells, non_finite_grid, touched_cells, mixed_cells, particles_in_mixed;
    uint32_t sleeping;
};
// Diagnostics over the stored particles and the last substep's grid: set
// `grid_set` updated with `dt` / `ddt`; `cell_lo/hi/count` are R^2 scratch
// arrays (lo = UINT_MAX, hi = count = 0 between calls; left that way).
hipError_t launch_snapshot(const Pipeline& p, uint32_t cur_buf, uint32_t grid_set, float dt, float ddt,
                           DevSnapshot* out, uint32_t* cell_lo, uint32_t* cell_hi, uint32_t* cell_count,
                           hipStream_t stream);
uint32_t order_key(float f);       // the key launch_snapshot uses for signed min / max
float order_key_inverse(uint32_t k);

// Stand-alone cell-key sort (stable LSD radix sort): keys_out[i] is the key
// of xy[i], perm_out[j] the index at sorted position j.  `s` supplies the
// scratch: kbuf and vbuf of n entries each (none may alias keys_out or
// perm_out) and hist of sort_hist_len(n, n_bins) entries; s.key and s.out
// are ignored.  Clears nothing with hipMemset (capturable).
hipError_t sort_cell_keys_device(const float2* xy, uint32_t n, uint32_t grid_res, uint32_t* keys_out,
                                 uint32_t* perm_out, const SortBuf& s, hipStream_t stream);

}  // namespace emh
