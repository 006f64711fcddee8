// interop_kernels.hip -- particle / grid export into caller-owned device
// buffers (the layouts GpuSimulation::particle_buffer / grid_buffer hand to a
// renderer, gpu/solver/mod.rs:466-532) and the velocity impulses of
// solver/particles.rs:115-165, all without a host round trip.
#include "interop_kernels.hpp"

#include <cstddef>

namespace emh {

namespace {

__host__ __device__ inline uint32_t div_up_u(uint32_t a, uint32_t b) { return (a + b - 1u) / b; }

// One thread per particle slot; the 128 B record is moved as 8 x 16 B.
__global__ __launch_bounds__(256) void k_export_particles(const PBuf b, uint32_t n, const emh_particle* cold,
                                                          emh_particle* out) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const uint32_t i = b.id[s];
    if (i >= n) return;
    const float4* src = reinterpret_cast<const float4*>(cold + i);
    float4 r[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) r[k] = src[k];
    // device-owned fields, at their Particle offsets (particle.rs:11-144)
    static_assert(offsetof(emh_particle, x) == 0 && offsetof(emh_particle, velocity_gradient) == 16 &&
                      offsetof(emh_particle, deformation_gradient) == 32 && offsetof(emh_particle, mass) == 48 &&
                      offsetof(emh_particle, material_id) == 64 &&
                      offsetof(emh_particle, plastic_volume_ratio) == 68 &&
                      offsetof(emh_particle, log_volume_strain) == 80,
                  "Particle layout");
    const float2 x = b.x[s], v = b.v[s];
    r[0] = make_float4(x.x, x.y, v.x, v.y);
    r[1] = b.c[s];
    r[2] = b.f[s];
    r[3] = make_float4(b.mass[s], b.v0[s], b.vol[s], b.dens[s]);
    r[4] = make_float4(r[4].x, b.jp[s], b.h[s], b.q[s]);  // material_id stays the caller's
    r[5] = make_float4(b.lvs[s], r[5].y, r[5].z, r[5].w);
    static_assert(offsetof(emh_particle, sleeping) == 112, "Particle layout");
    r[7].x = __uint_as_float((b.mat[s] & kSleepBit) ? 1u : 0u);
    float4* dst = reinterpret_cast<float4*>(out + i);
#pragma unroll
    for (int k = 0; k < 8; ++k) dst[k] = r[k];
}

__global__ __launch_bounds__(256) void k_export_grid(const Pipeline p, uint64_t seq, float4* out) {
    const uint32_t R = p.cfg.grid_res;
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= R * R) return;
    const CtrlState st = p.ctrl->st[seq & 1u];
    const uint32_t set = st.cur_grid % 3u;
    const float ddt = st.last_ddt;
    const float m = p.grid.acc[set][kAccM][g];
    const float mx = p.grid.acc[set][kAccPx][g] + ddt * p.grid.acc[set][kAccSx][g];
    const float my = p.grid.acc[set][kAccPy][g] + ddt * p.grid.acc[set][kAccSy][g];
    out[g] = make_float4(mx, my, m, 0.0f);
}

__global__ __launch_bounds__(256) void k_impulse(const PBuf b, uint32_t n, float cx, float cy, float radius,
                                                 float fx, float fy, float strength, uint32_t radial,
                                                 float vel_limit) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const float2 x = b.x[s];
    const float dx = x.x - cx, dy = x.y - cy;
    const float r2 = radius * radius;
    const float dist2 = dx * dx + dy * dy;
    if (!(dist2 <= r2 && dist2 > 1e-8f)) return;
    float2 v = b.v[s];
    if (radial) {  // particles.rs:141-165
        const float dist = sqrtf(dist2);
        const float falloff = 1.0f - dist / radius;
        v.x = v.x + ((dx / dist) * strength) * falloff;
        v.y = v.y + ((dy / dist) * strength) * falloff;
    } else {  // particles.rs:115-139
        const float falloff = 1.0f - sqrtf(dist2 / r2);
        v.x = v.x + fx * falloff;
        v.y = v.y + fy * falloff;
    }
    const float spd = sqrtf(v.x * v.x + v.y * v.y);
    if (spd > vel_limit) {
        const float k = vel_limit / spd;
        v.x = v.x * k;
        v.y = v.y * k;
    }
    b.v[s] = v;
    const uint32_t word = b.mat[s];
    if (word & kSleepBit) b.mat[s] = word & ~kSleepBit;  // particles in range wake (particles.rs:122-137)
}

}  // namespace

hipError_t launch_export_particles(const PBuf& b, uint32_t n, const emh_particle* cold, emh_particle* out,
                                   hipStream_t stream) {
    if (n == 0) return hipSuccess;
    k_export_particles<<<div_up_u(n, 256u), 256, 0, stream>>>(b, n, cold, out);
    return hipGetLastError();
}

hipError_t launch_export_grid(const Pipeline& p, uint64_t seq, float4* out, hipStream_t stream) {
    const uint32_t cells = p.cfg.grid_res * p.cfg.grid_res;
    k_export_grid<<<div_up_u(cells, 256u), 256, 0, stream>>>(p, seq, out);
    return hipGetLastError();
}

hipError_t launch_impulse(const PBuf& b, uint32_t n, float cx, float cy, float radius, float fx, float fy,
                          float strength, uint32_t radial, float vel_limit, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    k_impulse<<<div_up_u(n, 256u), 256, 0, stream>>>(b, n, cx, cy, radius, fx, fy, strength, radial, vel_limit);
    return hipGetLastError();
}

}  // namespace emh
