// interop_kernels.hpp -- device-side access to the simulation state for code
// that shares the GPU with it (renderers, gameplay logic).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/emerge_hip.h"
#include "kernels.hpp"

namespace emh {

// Writes every particle as a 128 B Particle record at its caller index in
// `out`: the device-owned fields from buffer `b`, the rest from `cold` (the
// caller's last upload, same order).
hipError_t launch_export_particles(const PBuf& b, uint32_t n, const emh_particle* cold, emh_particle* out,
                                   hipStream_t stream);

// Writes the grid as R^2 cells {momentum.x, momentum.y, mass, 0} (x-major):
// the accumulator set the last substep gathered, momentum = A + last_ddt S.
// The set and last_ddt are read on the device from launch `seq`'s state.
hipError_t launch_export_grid(const Pipeline& p, uint64_t seq, float4* out, hipStream_t stream);

// Simulation::apply_impulse (radial = 0) / apply_radial_impulse (radial = 1)
// on buffer b: v += falloff * (force | strength * d / |d|) for particles within
// `radius` of the centre, then the speed is clamped to vel_limit.
hipError_t launch_impulse(const PBuf& b, uint32_t n, float cx, float cy, float radius, float fx, float fy,
                          float strength, uint32_t radial, float vel_limit, hipStream_t stream);

}  // namespace emh
