// Learned surrogate of the NACC (Cam-Clay) return mapping: a 6-64-64-64-3
// SiLU MLP on BF16 MFMA with FP32 accumulation.
//
//   inputs  x = (ln s1, ln s2, alpha, M, beta, xi)   trial singular values of
//               F_trial, hardening state (log Jp) and the material's NACC
//               parameters (nacc.rs:129-223 reads exactly these)
//   outputs y = (ln s1', ln s2', alpha')             projected singular values
//               and hardening state
//
// The network predicts the normalised residual r = (y - x[0:3] - out_mu) /
// out_sd from the normalised input (x - in_mu) / in_sd; the return mapping
// rebuilds F' = U diag(exp y0, exp y1) V^T from the trial SVD.
//
// Device layout.  Activations live in registers in the 16x16 MFMA output
// layout (features on rows, 16 particles on columns): lane l holds rows
// 4*(l/16)+r (r = 0..3) of column l%16 in each 16-row tile.  The next layer's
// B operand (v_mfma_f32_16x16x32_bf16: lane l holds k-slots 8*(l/16)+j of
// column l%16) is built from those registers without lane exchange by
// permuting the K index of the weights instead:
//   k-slot (kb, g, j)  <->  feature (2*kb + j/4)*16 + 4*g + j%4
// Weights are stored pre-packed in exactly the per-lane order each MFMA A
// operand needs (one 1 KiB tile = 64 lanes x 8 bf16 per 16x32 block).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace emh {

constexpr int kSgIn = 6, kSgHid = 64, kSgOut = 3;
// FP32 master parameters, W_l row-major [out][in]
constexpr int kSgW1 = 0, kSgB1 = kSgW1 + kSgHid * kSgIn, kSgW2 = kSgB1 + kSgHid, kSgB2 = kSgW2 + kSgHid * kSgHid,
              kSgW3 = kSgB2 + kSgHid, kSgB3 = kSgW3 + kSgHid * kSgHid, kSgW4 = kSgB3 + kSgHid,
              kSgB4 = kSgW4 + kSgOut * kSgHid, kSgParams = kSgB4 + kSgOut;  // 8963
// normalisation block
constexpr int kSgInMu = 0, kSgInSd = 6, kSgOutMu = 12, kSgOutSd = 15, kSgZMax = 18, kSgNorm = 20;
// packed operand tiles: forward L1 (4 out tiles), L2 / L3 (4 out tiles x 2
// k-blocks), L4 (2 k-blocks); backward-data L4^T (4 in tiles), L3^T / L2^T
// (4 in tiles x 2 k-blocks)
constexpr int kPkF1 = 0, kPkF2 = 4, kPkF3 = 12, kPkF4 = 20, kPkB4 = 22, kPkB3 = 26, kPkB2 = 34, kPkTiles = 42;
constexpr int kPkFwdTiles = kPkB4;  // the forward pass reads tiles 0..21
// compact bias block the forward pass reads: b1 at 0, b2 at 64, b3 at 128, b4 at 192
constexpr int kSgBias = 196;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// f32 -> bf16 bits, round to nearest even (finite inputs)
__host__ __device__ inline uint32_t sg_bf16_bits(float f) {
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return u >> 16;
}
// Two f32 -> packed bf16 pair: one v_cvt_pk_bf16_f32 on gfx950 (round to
// nearest even, as sg_bf16_bits for finite inputs).
typedef __bf16 sg_bf16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t sg_pack2(float lo, float hi) {
    const sg_bf16x2 v = {(__bf16)lo, (__bf16)hi};
    return __builtin_bit_cast(uint32_t, v);
}
__device__ __forceinline__ bf16x8 sg_pack8(float a0, float a1, float a2, float a3, float a4, float a5, float a6,
                                           float a7) {
    const uint4 u = make_uint4(sg_pack2(a0, a1), sg_pack2(a2, a3), sg_pack2(a4, a5), sg_pack2(a6, a7));
    return __builtin_bit_cast(bf16x8, u);
}
__device__ __forceinline__ bf16x8 sg_tile(const uint4* __restrict__ pk, int tile) {
    return __builtin_bit_cast(bf16x8, pk[tile * 64 + (threadIdx.x & 63)]);
}
__device__ __forceinline__ f32x4 sg_mfma(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// SiLU z * sigmoid(z) with the hardware exp2 and reciprocal (about 1 ulp each)
__device__ __forceinline__ float sg_sigmoid(float z) {
    return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-z * 1.4426950408889634f));
}
__device__ __forceinline__ float sg_silu(float z) { return z * sg_sigmoid(z); }
__device__ __forceinline__ float sg_silu_grad(float z) {
    const float s = sg_sigmoid(z);
    return s * (1.0f + z * (1.0f - s));
}

// Four 16-row tiles in output layout -> the two chained B operands.
__device__ __forceinline__ void sg_chain(const f32x4 (&h)[4], bf16x8 (&b)[2]) {
#pragma unroll
    for (int kb = 0; kb < 2; ++kb)
        b[kb] = sg_pack8(h[2 * kb][0], h[2 * kb][1], h[2 * kb][2], h[2 * kb][3], h[2 * kb + 1][0], h[2 * kb + 1][1],
                         h[2 * kb + 1][2], h[2 * kb + 1][3]);
}

// Hidden layer: z = W h + b over the chained input, four output tiles.
__device__ __forceinline__ void sg_hidden(const uint4* __restrict__ pk, int t0, const float* __restrict__ bias,
                                          const bf16x8 (&in)[2], f32x4 (&z)[4]) {
    const int g = (threadIdx.x & 63) >> 4;
#pragma unroll
    for (int ot = 0; ot < 4; ++ot) {
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        acc = sg_mfma(sg_tile(pk, t0 + ot * 2), in[0], acc);
        acc = sg_mfma(sg_tile(pk, t0 + ot * 2 + 1), in[1], acc);
        const float4 bb = *reinterpret_cast<const float4*>(bias + ot * 16 + 4 * g);
        z[ot] = acc + f32x4{bb.x, bb.y, bb.z, bb.w};
    }
}

// Forward pass of one 16-particle group.  x: the layer-1 B operand (lanes
// 0..15 hold the six normalised inputs of particle l in slots 0..5, every
// other slot zero).  Returns the output tile: lanes 0..15 hold outputs 0..2
// of particle l in elements 0..2.  With zs, the pre-activations of the three
// hidden layers are kept for the backward pass.
__device__ __forceinline__ f32x4 sg_forward(const uint4* __restrict__ pk, const float* __restrict__ bias, bf16x8 x,
                                            f32x4 (*zs)[4]) {
    const int g = (threadIdx.x & 63) >> 4;
    f32x4 z[4], h[4];
    bf16x8 b[2];
#pragma unroll
    for (int ot = 0; ot < 4; ++ot) {
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        acc = sg_mfma(sg_tile(pk, kPkF1 + ot), x, acc);
        const float4 bb = *reinterpret_cast<const float4*>(bias + ot * 16 + 4 * g);
        z[ot] = acc + f32x4{bb.x, bb.y, bb.z, bb.w};
    }
#pragma unroll
    for (int layer = 0; layer < 3; ++layer) {
        if (zs) {
#pragma unroll
            for (int t = 0; t < 4; ++t) zs[layer][t] = z[t];
        }
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) h[t][r] = sg_silu(z[t][r]);
        sg_chain(h, b);
        if (layer == 0) sg_hidden(pk, kPkF2, bias + 64, b, z);
        if (layer == 1) sg_hidden(pk, kPkF3, bias + 128, b, z);
    }
    f32x4 o = {0.0f, 0.0f, 0.0f, 0.0f};
    o = sg_mfma(sg_tile(pk, kPkF4), b[0], o);
    o = sg_mfma(sg_tile(pk, kPkF4 + 1), b[1], o);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int row = 4 * g + r;
        o[r] = row < kSgOut ? o[r] + bias[192 + row] : 0.0f;
    }
    return o;
}

// The same forward pass without saved pre-activations, the two hidden
// layers in a rolled loop: each layer's operand loads then stay behind the
// previous layer's arithmetic instead of all being hoisted (22 tiles would
// hold 88 registers).
__device__ __forceinline__ f32x4 sg_forward_infer(const uint4* __restrict__ pk, const float* __restrict__ bias,
                                                  bf16x8 x) {
    const int g = (threadIdx.x & 63) >> 4;
    f32x4 z[4], h[4];
    bf16x8 b[2];
#pragma unroll
    for (int ot = 0; ot < 4; ++ot) {
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        acc = sg_mfma(sg_tile(pk, kPkF1 + ot), x, acc);
        const float4 bb = *reinterpret_cast<const float4*>(bias + ot * 16 + 4 * g);
        z[ot] = acc + f32x4{bb.x, bb.y, bb.z, bb.w};
    }
#pragma unroll 1
    for (int layer = 0; layer < 2; ++layer) {
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) h[t][r] = sg_silu(z[t][r]);
        sg_chain(h, b);
        sg_hidden(pk, kPkF2 + 8 * layer, bias + 64 * (layer + 1), b, z);
    }
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) h[t][r] = sg_silu(z[t][r]);
    sg_chain(h, b);
    f32x4 o = {0.0f, 0.0f, 0.0f, 0.0f};
    o = sg_mfma(sg_tile(pk, kPkF4), b[0], o);
    o = sg_mfma(sg_tile(pk, kPkF4 + 1), b[1], o);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int row = 4 * g + r;
        o[r] = row < kSgOut ? o[r] + bias[192 + row] : 0.0f;
    }
    return o;
}

// Backward-data pass of one group: from dO (output-tile layout, rows >= 3
// zero) and the saved pre-activations, the pre-activation gradients dz of
// the three hidden layers (same layout as zs).
__device__ __forceinline__ void sg_backward(const uint4* __restrict__ pk, const f32x4 (&zs)[3][4], f32x4 d_out,
                                            f32x4 (&dz)[3][4]) {
    const bf16x8 bo = sg_pack8(d_out[0], d_out[1], d_out[2], d_out[3], 0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
    for (int it = 0; it < 4; ++it) {
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        acc = sg_mfma(sg_tile(pk, kPkB4 + it), bo, acc);
#pragma unroll
        for (int r = 0; r < 4; ++r) dz[2][it][r] = acc[r] * sg_silu_grad(zs[2][it][r]);
    }
#pragma unroll
    for (int layer = 2; layer >= 1; --layer) {
        bf16x8 b[2];
        sg_chain(dz[layer], b);
        const int t0 = layer == 2 ? kPkB3 : kPkB2;
#pragma unroll
        for (int it = 0; it < 4; ++it) {
            f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
            acc = sg_mfma(sg_tile(pk, t0 + it * 2), b[0], acc);
            acc = sg_mfma(sg_tile(pk, t0 + it * 2 + 1), b[1], acc);
#pragma unroll
            for (int r = 0; r < 4; ++r) dz[layer - 1][it][r] = acc[r] * sg_silu_grad(zs[layer - 1][it][r]);
        }
    }
}

// The forward tiles and the bias block, staged in LDS by the whole workgroup
// (ends with a barrier): inference kernels read them from there.
struct SgLds {
    uint4 pk[kPkFwdTiles * 64];
    float bias[kSgBias];
};
__device__ __forceinline__ void sg_stage(const uint4* __restrict__ pk, const float* __restrict__ bias, SgLds& s) {
    for (int i = threadIdx.x; i < kPkFwdTiles * 64; i += blockDim.x) s.pk[i] = pk[i];
    for (int i = threadIdx.x; i < kSgBias; i += blockDim.x) s.bias[i] = bias[i];
    __syncthreads();
}

// Host/device view of a surrogate's device state used by the substep kernel.
struct SgDev {
    const uint4* pk;    // packed bf16 operand tiles (kPkTiles * 64 uint4)
    const float* bias;  // compact bias block (kSgBias floats)
    const float* norm;  // normalisation block
    float* harvest;     // NACC trial-state records (12 floats each), or null
    unsigned long long* harvest_count;  // trial states offered
    unsigned long long harvest_cap;     // records the buffer holds
    unsigned long long harvest_stride;  // offer k is recorded (as record k / stride) when k % stride == 0
    // [c * kSgCountSlots + slot]: c = 0 yielding trial states, 1 surrogate
    // returns, 2 plain fallbacks; a workgroup adds into slot blockIdx % kSgCountSlots
    unsigned long long* counts;
};
constexpr int kSgCountSlots = 256;

}  // namespace emh
