// kernels.hpp -- device buffers and launch wrappers of the substep pipeline.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_math.hpp"

namespace emh {

constexpr int kBlock = 256;        // threads per workgroup for particle kernels (4 waves of 64)
constexpr int kChunk = 1024;       // particles per P2G workgroup (4 per thread)
constexpr int kLdsCells = 3072;    // P2G bounding-box budget in LDS cells (3 floats each: 36 KiB)
constexpr int kTileCells = 256;    // 16 x 16 cells per sort tile

// Particle state, structure of arrays.  Two of these ping-pong: a substep
// reads one (in the order the previous substep left it) and writes the other
// in cell-key order.
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
    uint32_t* mat;   // material_id | pinned bit
    uint32_t* id;    // index of the particle in the caller's order
    uint32_t* key;   // cell key of the (projected) position, for the next sort
    uint32_t* rank;  // arrival rank of the particle inside its key bin
};

struct GridBuf {
    float* m[2];  // P2G mass, ping-pong by substep parity
    float* mx[2];
    float* my[2];
    float2* vel;  // grid velocity after the update (what G2P gathers)
    float* dmass; // mass of the density re-estimate pass
};

struct SortBuf {
    uint32_t* cell_count;  // n_bins
    uint32_t* cell_start;  // n_bins + 1
    uint32_t* tile_count;  // n_tiles
    uint32_t* tile_start;  // n_tiles + 1
    uint32_t* perm;        // N: particle at each slot, unordered inside a bin
    float2* cfl_partial;   // per G2P block: (max speed, min material dt)
};

// Device-resident control block: the adaptive substep loop runs on the
// device so that no substep needs a host round trip.
struct Ctrl {
    float remaining;      // simulated time left in the current step
    float dt;             // current substep dt
    uint32_t substeps;    // substeps taken in the current step
    uint32_t active;      // 1 while the current substep slot is live
    uint32_t clamp_count; // G2P velocity clamps in the current step
    uint32_t proj_count;  // state projections in the current step
    uint32_t total_substeps;
    float last_dt;
    uint32_t cur_buf;     // which PBuf holds the current state
    uint32_t cur_grid;    // which grid accumulator the current substep uses
    uint32_t n_cfl_blocks;
    uint32_t pad[5];
};

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

// Kernels of one substep, in launch order (profiling identifiers).
enum KernelId : int {
    kKControl = 0,
    kKCellScan,
    kKDensityScatter,
    kKDensityGather,
    kKP2G,
    kKGridUpdate,
    kKG2P,
    kKernelCount
};
extern const char* const kKernelNames[kKernelCount];

// Optional per-kernel event pairs of one substep slot (null: no timing).
struct SlotEvents {
    hipEvent_t ev[kKernelCount][2];
    bool used[kKernelCount];
};

// Launch wrappers (all asynchronous on `stream`).
hipError_t launch_zero_counters(const Pipeline& p, hipStream_t stream);
hipError_t launch_prologue(const Pipeline& p, hipStream_t stream);
hipError_t launch_begin_step(const Pipeline& p, hipStream_t stream);
hipError_t launch_substep(const Pipeline& p, hipStream_t stream, SlotEvents* events = nullptr);

// Stand-alone cell-key sort (same kernels the substep uses).
hipError_t sort_cell_keys_device(const float2* xy, uint32_t n, uint32_t grid_res, uint32_t* keys,
                                 uint32_t* perm_sorted, uint32_t* scratch_rank, uint32_t* cell_count,
                                 uint32_t* cell_start, uint32_t* tile_count, uint32_t* tile_start,
                                 uint32_t* perm_unordered, hipStream_t stream);

}  // namespace emh
