// spawn_kernels.hip -- device side of add_body's volume estimate.
//
// estimate_particle_volumes_local (solver/density.rs:256-319): scatter w·m of
// every particle inside the new group's bounding box (grown by 3 cells) onto
// a dense mass grid, then gather ρ = Σ w·m_cell for the new particles and set
// density, volume and initial_volume where ρ > f32::EPSILON.  Also used for
// estimate_particle_volumes over all particles (Simulation::new with
// precompute_initial_volumes, lifecycle.rs:64-67).
#include <hip/hip_runtime.h>

#include "device_math.hpp"
#include "spawn_kernels.hpp"

namespace emh {

namespace {

__global__ void k_local_mass_scatter(const float2* x, const float* mass, uint32_t n, uint32_t R, float lox, float loy,
                                     float hix, float hiy, uint32_t use_box, float* dmass) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float2 p = x[i];
    if (use_box && (p.x < lox || p.y < loy || p.x > hix || p.y > hiy)) return;
    Wts w = weights(p.x, p.y);
    float m = mass[i];
    for (int gx = 0; gx < 3; ++gx)
        for (int gy = 0; gy < 3; ++gy) {
            int cx = w.bx + gx - 1, cy = w.by + gy - 1;
            if (cx < 0 || cy < 0 || cx >= (int)R || cy >= (int)R) continue;
            __hip_atomic_fetch_add(&dmass[(uint32_t)cx * R + (uint32_t)cy], (w.wx[gx] * w.wy[gy]) * m,
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
}

__global__ void k_local_density_gather(const float2* x, const float* mass, uint32_t from, uint32_t to, uint32_t R,
                                       const float* dmass, float* dens_out, float* vol_out, uint32_t* wrote) {
    uint32_t i = from + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= to) return;
    float2 p = x[i];
    Wts w = weights(p.x, p.y);
    float d = 0.0f;
    for (int gx = 0; gx < 3; ++gx)
        for (int gy = 0; gy < 3; ++gy) {
            float weight = w.wx[gx] * w.wy[gy];
            int cx = w.bx + gx - 1, cy = w.by + gy - 1;
            float m = 0.0f;
            if (cx >= 0 && cy >= 0 && cx < (int)R && cy < (int)R) m = dmass[(uint32_t)cx * R + (uint32_t)cy];
            d = d + m * weight;
        }
    if (d > kF32Eps) {
        dens_out[i] = d;
        vol_out[i] = mass[i] / d;
        wrote[i] = 1u;
    } else {
        wrote[i] = 0u;
    }
}

}  // namespace

hipError_t estimate_volumes_device(const float2* x, const float* mass, uint32_t n, uint32_t from, uint32_t R,
                                   bool use_box, const float box[4], float* dmass, float* dens_out, float* vol_out,
                                   uint32_t* wrote, hipStream_t stream) {
    hipError_t e = hipMemsetAsync(dmass, 0, sizeof(float) * (size_t)R * R, stream);
    if (e != hipSuccess) return e;
    if (n == 0 || from >= n) return hipSuccess;
    const uint32_t B = 256;
    k_local_mass_scatter<<<(n + B - 1) / B, B, 0, stream>>>(x, mass, n, R, box[0], box[1], box[2], box[3],
                                                            use_box ? 1u : 0u, dmass);
    k_local_density_gather<<<(n - from + B - 1) / B, B, 0, stream>>>(x, mass, from, n, R, dmass, dens_out, vol_out,
                                                                     wrote);
    return hipGetLastError();
}

}  // namespace emh
