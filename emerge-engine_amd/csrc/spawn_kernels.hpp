// spawn_kernels.hpp -- device volume estimate used when bodies are spawned.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace emh {

// Mass scatter of particles [0, n) (only those inside box = {lox, loy, hix, hiy}
// when use_box), then density / volume gather for particles [from, n).
// wrote[i] = 1 where the density exceeded f32::EPSILON and dens_out/vol_out
// were written.
hipError_t estimate_volumes_device(const float2* x, const float* mass, uint32_t n, uint32_t from, uint32_t R,
                                   bool use_box, const float box[4], float* dmass, float* dens_out, float* vol_out,
                                   uint32_t* wrote, hipStream_t stream);

}  // namespace emh
