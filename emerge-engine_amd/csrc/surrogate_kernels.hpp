// Surrogate return mapping shared by the batch kernels and the substep
// kernel, and the launch wrappers of the training pipeline.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_math.hpp"
#include "surrogate.hpp"

namespace emh {

// Floats of LDS the network queue of sg_return<POOL> needs: 12 per lane of
// the POOL waves that share it.
constexpr int kSgBlockLanes = 256;  // the lanes of a workgroup (kernels using POOL = 4)
template <uint32_t POOL>
constexpr int kSgQueueLds = (int)POOL * 64 * 12;

// NaccMaterial::project (nacc.rs:129-223) with the exact yield-function check
// as the gate: a state inside the yield surface keeps F; the tension and
// compression caps (Cases A and B, closed forms of one sqrt and one pow) keep
// the plain return; a state outside the yield surface between them (Case C,
// the projection onto the surface with the hardening update) goes through the
// network when its singular values are positive and every normalised input
// lies within z_max, and through the plain return mapping (nacc_return)
// otherwise.
// The network's lanes of POOL waves are queued together in LDS and the queue
// is evaluated in groups of 16 (one MFMA column group), dealt to the POOL
// waves in turn, so the groups are full except the last.  POOL = 4 (the
// workgroup) for the batch return mapping: about 41 of 256 lanes reach the
// network in the NACC slope scene, 3 groups where four separate waves run 4
// part-filled ones (112-117 -> 87-89 us per 1.97M states).  POOL = 1 inside
// the substep, where waiting for the workgroup's other waves costs more than
// the fuller groups save (284-287 us per substep with one-wave queues against
// 295-321 us with workgroup queues).
// Every lane of the workgroup calls this (it synchronises the workgroup);
// `valid` marks the lanes with a particle.  net: the network staged in LDS
// (sg_stage); s_io: kSgQueueLds<POOL> floats of LDS, this wave's own (POOL = 1)
// or the workgroup's (POOL = 4).
template <uint32_t POOL>
__device__ __forceinline__ void sg_return(const DevMat& m, const SgDev& sg, const SgLds& net, bool valid,
                                          const M2& fin, float alpha, M2& fo, float& ao, float* s_io) {
    static_assert(POOL == 1 || POOL * 64 == kSgBlockLanes, "one wave or the whole workgroup");
    __shared__ uint32_t s_wcnt[POOL];
    const uint32_t lane = threadIdx.x & 63u, w = POOL == 1 ? 0u : threadIdx.x >> 6, c = lane & 15u, g = lane >> 4;
    M2 u{1.0f, 0.0f, 0.0f, 1.0f}, vt{1.0f, 0.0f, 0.0f, 1.0f};
    NaccTrial tr{};
    tr.yield = 0;
    float x[kSgIn] = {};
    bool use = false;
    if (valid) {
        float2 sv;
        svd2(fin, u, sv, vt);
        tr = nacc_trial(m, sv, alpha);
        if (tr.yield == 3 && sv.y > 0.0f) {  // svd2 orders s1 >= |s2|
            x[0] = logf(sv.x);
            x[1] = logf(sv.y);
            x[2] = alpha;
            x[3] = m.na_m;
            x[4] = m.na_beta;
            x[5] = m.na_xi;
            use = true;
#pragma unroll
            for (int k = 0; k < kSgIn; ++k) {
                const float z = (x[k] - sg.norm[kSgInMu + k]) / sg.norm[kSgInSd + k];
                use = use && fabsf(z) <= sg.norm[kSgZMax];
            }
        }
    }
    // queue slot: the network lanes of the waves before this one, then this
    // wave's in lane order
    const uint64_t mask = __ballot(use);
    const uint32_t cnt = (uint32_t)__popcll(mask);
    uint32_t off = 0, total = cnt;
    if (POOL > 1) {
        if (lane == 0) s_wcnt[w] = cnt;
        __syncthreads();
        total = 0;
#pragma unroll
        for (uint32_t k = 0; k < POOL; ++k) {
            const uint32_t v = s_wcnt[k];
            off += k < w ? v : 0u;
            total += v;
        }
    }
    const uint32_t rank = off + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
    if (use) {
#pragma unroll
        for (int k = 0; k < kSgIn; ++k) s_io[rank * 12 + k] = (x[k] - sg.norm[kSgInMu + k]) / sg.norm[kSgInSd + k];
    }
    __syncthreads();
    const uint32_t groups = (total + 15u) / 16u;
    for (uint32_t grp = w; grp < groups; grp += POOL) {  // wave-uniform
        const uint32_t slot = grp * 16u + c;
        const bool mine = g == 0 && slot < total;
        float z[kSgIn] = {};
        if (mine) {
#pragma unroll
            for (int k = 0; k < kSgIn; ++k) z[k] = s_io[slot * 12 + k];
        }
        const f32x4 o = sg_forward_infer(net.pk, net.bias, sg_pack8(z[0], z[1], z[2], z[3], z[4], z[5], 0.0f, 0.0f));
        if (mine) {
#pragma unroll
            for (int k = 0; k < kSgOut; ++k) s_io[slot * 12 + 8 + k] = o[k];
        }
    }
    __syncthreads();
    if (valid) {
        if (tr.yield == 0) {
            fo = fin;
            ao = alpha;
        } else if (use) {
            float y[kSgOut];
#pragma unroll
            for (int k = 0; k < kSgOut; ++k)
                y[k] = x[k] + (s_io[rank * 12 + 8 + k] * sg.norm[kSgOutSd + k] + sg.norm[kSgOutMu + k]);
            ao = (m.flags & kMatNaccHardening) ? y[2] : alpha;
            fo = usv(u, expf(y[0]), expf(y[1]), vt);
        } else {
            float s1, s2;
            nacc_return(m, tr, alpha, s1, s2, ao);
            fo = usv(u, s1, s2, vt);
        }
    }
    if (sg.counts) {  // spread over slots: one address per counter would serialise every wave's atomics
        const uint32_t yielding = (uint32_t)__popcll(__ballot(valid && tr.yield != 0));
        if (lane == 0 && yielding) {
            unsigned long long* c = sg.counts + (blockIdx.x % kSgCountSlots);
            atomicAdd(&c[0], (unsigned long long)yielding);
            atomicAdd(&c[kSgCountSlots], (unsigned long long)cnt);
            atomicAdd(&c[2 * kSgCountSlots], (unsigned long long)(yielding - cnt));
        }
    }
}

// One training batch: slots 0..nb-1 map to samples (perm_a * (first + slot)
// + perm_b) % n_set (perm_a = 0: first + slot); slots >= n_valid are padding.
struct SgBatch {
    const float* xs;  // [n_set][8] raw inputs
    const float* ts;  // [n_set][4] raw targets
    unsigned long long n_set, perm_a, perm_b, first;
    uint32_t nb, n_valid;  // nb: multiple of 256
    float inv_n;           // 1 / n_valid
    const uint4* pk;
    const float* bias;
    const float* norm;
};
// Per-slot operands of the weight-gradient GEMMs ([slot][feature]) and the
// loss accumulator.
struct SgTrainBufs {
    float *hin, *h1, *h2, *h3, *dz1, *dz2, *dz3, *dout;
    float* loss;
};
struct SgAdam {
    float lr, beta1, beta2, eps, c1, c2;  // c1 = 1 / (1 - beta1^t), c2 = 1 / (1 - beta2^t)
};

hipError_t sg_launch_pack(const float* w, uint4* pk, float* bias, hipStream_t st);
hipError_t sg_launch_prepare(const float* rec, unsigned long long n_rec, float* xs, float* ts,
                             unsigned long long* n_out, double* sums, hipStream_t st);
// fwd/bwd + weight gradients (accumulated into grad); with adam, the update
// and the re-packing of the operands as well.
hipError_t sg_launch_train_step(const SgBatch& b, const SgTrainBufs& o, float* w, float* m, float* v, float* grad,
                                const SgAdam* adam, float* bias, hipStream_t st);
hipError_t sg_launch_return_plain(const DevMat& m, const float4* f, const float* alpha, unsigned long long n,
                                  float4* fo, float* ao, hipStream_t st);
hipError_t sg_launch_return_surrogate(const DevMat& m, const float4* f, const float* alpha, unsigned long long n,
                                      float4* fo, float* ao, const SgDev& sg, hipStream_t st);
hipError_t sg_launch_infer(const float* xs, unsigned long long n, float* ys, const SgDev& sg, hipStream_t st);

}  // namespace emh
