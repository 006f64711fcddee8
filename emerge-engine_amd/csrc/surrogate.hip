// Learned NACC return-mapping surrogate: training-set preparation, the
// training kernels (forward + backward-data on BF16 MFMA, split-K weight
// gradients on BF16 MFMA, Adam on the FP32 master weights), operand packing,
// and the batch return-mapping kernels (plain and gated surrogate) the
// speed and error comparison runs.
#include "surrogate_kernels.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>

namespace emh {
namespace {

constexpr int kSgBlock = 256;

inline uint32_t div_up_u(uint32_t a, uint32_t b) { return (a + b - 1u) / b; }

__device__ __forceinline__ uint32_t sg_lane() { return threadIdx.x & 63u; }

// ------------------------------------------------------------ packing
// One thread per (tile, lane, element pair): the bf16 A operands of every
// forward and backward-data MFMA from the FP32 master weights.
__device__ float pk_value(const float* __restrict__ w, int tile, int lane, int j) {
    const int c = lane & 15, g = lane >> 4;
    auto chained = [&](int kb) { return (2 * kb + j / 4) * 16 + 4 * g + (j & 3); };
    if (tile < kPkF2) {  // layer 1: out tile `tile`, inputs in slots 0..5 of lanes 0..15
        const int out = tile * 16 + c;
        return (g == 0 && j < kSgIn) ? w[kSgW1 + out * kSgIn + j] : 0.0f;
    }
    if (tile < kPkF4) {  // layers 2, 3: (out tile, k-block)
        const bool l2 = tile < kPkF3;
        const int t = tile - (l2 ? kPkF2 : kPkF3), ot = t >> 1, kb = t & 1;
        return w[(l2 ? kSgW2 : kSgW3) + (ot * 16 + c) * kSgHid + chained(kb)];
    }
    if (tile < kPkB4) {  // layer 4: k-block
        const int kb = tile - kPkF4;
        return c < kSgOut ? w[kSgW4 + c * kSgHid + chained(kb)] : 0.0f;
    }
    if (tile < kPkB3) {  // dH3 = W4^T dO: in tile; k-slot (g, j<4) is output 4g+j
        const int it = tile - kPkB4, out = 4 * g + j;
        return (j < 4 && out < kSgOut) ? w[kSgW4 + out * kSgHid + it * 16 + c] : 0.0f;
    }
    // dH2 = W3^T dZ3, dH1 = W2^T dZ2: (in tile, k-block over the outputs)
    const bool l3 = tile < kPkB2;
    const int t = tile - (l3 ? kPkB3 : kPkB2), it = t >> 1, kb = t & 1;
    return w[(l3 ? kSgW3 : kSgW2) + chained(kb) * kSgHid + it * 16 + c];
}

__global__ void k_sg_pack(const float* __restrict__ w, uint32_t* __restrict__ pk, float* __restrict__ bias) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;  // (tile, lane, pair)
    if (i < kSgBias) {
        float b = 0.0f;
        if (i < 64)
            b = w[kSgB1 + i];
        else if (i < 128)
            b = w[kSgB2 + i - 64];
        else if (i < 192)
            b = w[kSgB3 + i - 128];
        else if (i < 192 + kSgOut)
            b = w[kSgB4 + i - 192];
        bias[i] = b;
    }
    if (i >= kPkTiles * 64 * 4) return;
    const int pair = i & 3, lane = (i >> 2) & 63, tile = i >> 8;
    pk[i] = sg_bf16_bits(pk_value(w, tile, lane, 2 * pair)) | (sg_bf16_bits(pk_value(w, tile, lane, 2 * pair + 1)) << 16);
}

// ------------------------------------------------------------ training set
// A harvested record (12 floats): F_trial (a, b, c, d), alpha, mu, kappa, M,
// beta, xi, flags (bits), 0.
__device__ DevMat record_mat(const float* r) {
    DevMat m{};
    m.model = kModelNacc;
    m.mu = r[5];
    m.na_kappa = r[6];
    m.na_m = r[7];
    m.na_beta = r[8];
    m.na_xi = r[9];
    m.flags = __float_as_uint(r[10]);
    return m;
}

// Keeps the records whose return is the projection onto the yield surface
// (Case C, the states the network replaces the return mapping for) with
// positive singular values, labels them with the plain return mapping, and
// accumulates the normalisation sums.
__global__ void k_sg_prepare(const float* __restrict__ rec, unsigned long long n_rec, float* __restrict__ xs,
                             float* __restrict__ ts, unsigned long long* __restrict__ n_out, double* __restrict__ sums) {
    const unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
    bool keep = false;
    float x[kSgIn] = {}, t[kSgOut] = {};
    if (i < n_rec) {
        const float* r = rec + i * 12;
        const DevMat m = record_mat(r);
        const M2 f{r[0], r[1], r[2], r[3]};
        M2 u, vt;
        float2 sg;
        svd2(f, u, sg, vt);
        if (sg.y > 0.0f && sg.x > 0.0f) {
            const NaccTrial tr = nacc_trial(m, sg, r[4]);
            if (tr.yield == 3) {
                float s1, s2, ao;
                nacc_return(m, tr, r[4], s1, s2, ao);
                x[0] = logf(sg.x);
                x[1] = logf(sg.y);
                x[2] = r[4];
                x[3] = m.na_m;
                x[4] = m.na_beta;
                x[5] = m.na_xi;
                t[0] = logf(s1);
                t[1] = logf(s2);
                t[2] = ao;
                keep = __builtin_isfinite(x[0] + x[1] + x[2] + t[0] + t[1] + t[2]);
            }
        }
    }
    const uint64_t mask = __ballot(keep);
    if (mask == 0ull) return;
    unsigned long long base = 0;
    const uint32_t lane = sg_lane();
    const uint32_t leader = __ffsll((unsigned long long)mask) - 1;
    if (lane == leader) base = atomicAdd(n_out, (unsigned long long)__popcll(mask));
    base = __shfl(base, leader);
    if (keep) {
        const unsigned long long o = base + __popcll(mask & ((1ull << lane) - 1ull));
        float* xo = xs + o * 8;
        for (int k = 0; k < kSgIn; ++k) xo[k] = x[k];
        xo[6] = xo[7] = 0.0f;
        float* to = ts + o * 4;
        for (int k = 0; k < kSgOut; ++k) to[k] = t[k];
        to[3] = 0.0f;
        for (int k = 0; k < kSgIn; ++k) {
            atomicAdd(&sums[k], (double)x[k]);
            atomicAdd(&sums[9 + k], (double)x[k] * (double)x[k]);
        }
        for (int k = 0; k < kSgOut; ++k) {
            const double d = (double)t[k] - (double)x[k];
            atomicAdd(&sums[6 + k], d);
            atomicAdd(&sums[15 + k], d * d);
        }
    }
}

// ------------------------------------------------------------ training
__device__ __forceinline__ uint64_t sample_of(const SgBatch& b, uint32_t slot) {
    const uint64_t q = b.first + slot;
    return b.perm_a ? (b.perm_a * q + b.perm_b) % b.n_set : q;
}

// One wave per 16-sample group: forward with saved pre-activations, MSE loss
// on the normalised residual, backward-data; the per-sample operands of the
// weight-gradient GEMMs are written out in [sample][feature] order.
__global__ __launch_bounds__(kSgBlock) void k_sg_fwd_bwd(SgBatch b, SgTrainBufs o) {
    const uint32_t lane = sg_lane(), c = lane & 15u, g = lane >> 4;
    const uint32_t slot = (blockIdx.x * (kSgBlock / 64) + (threadIdx.x >> 6)) * 16u + c;
    if ((blockIdx.x * (kSgBlock / 64) + (threadIdx.x >> 6)) * 16u >= b.nb) return;  // whole wave
    const bool valid = slot < b.n_valid;
    const float* nm = b.norm;
    float x[8] = {};
    float t[kSgOut] = {};
    if (valid && g == 0) {
        const uint64_t s = sample_of(b, slot);
        const float* xr = b.xs + s * 8;
        for (int k = 0; k < kSgIn; ++k) x[k] = (xr[k] - nm[kSgInMu + k]) / nm[kSgInSd + k];
        const float* tr = b.ts + s * 4;
        for (int k = 0; k < kSgOut; ++k) t[k] = ((tr[k] - xr[k]) - nm[kSgOutMu + k]) / nm[kSgOutSd + k];
    }
    const bf16x8 xb = sg_pack8(x[0], x[1], x[2], x[3], x[4], x[5], 0.0f, 0.0f);
    f32x4 zs[3][4];
    const f32x4 out = sg_forward(b.pk, b.bias, xb, zs);
    f32x4 d_out = {0.0f, 0.0f, 0.0f, 0.0f};
    float se = 0.0f;
    if (valid && g == 0) {
#pragma unroll
        for (int r = 0; r < kSgOut; ++r) {
            const float e = out[r] - t[r];
            se += e * e;
            d_out[r] = 2.0f * e * b.inv_n;
        }
    }
    f32x4 dz[3][4];
    sg_backward(b.pk, zs, d_out, dz);
    // operands of the weight gradients
    if (g == 0) {
        float4* hi = reinterpret_cast<float4*>(o.hin + (size_t)slot * 8);
        hi[0] = make_float4(x[0], x[1], x[2], x[3]);
        hi[1] = make_float4(x[4], x[5], 0.0f, 0.0f);
        *reinterpret_cast<float4*>(o.dout + (size_t)slot * 4) = make_float4(d_out[0], d_out[1], d_out[2], d_out[3]);
    }
    float* hs[3] = {o.h1, o.h2, o.h3};
    float* ds[3] = {o.dz1, o.dz2, o.dz3};
#pragma unroll
    for (int layer = 0; layer < 3; ++layer)
#pragma unroll
        for (int tt = 0; tt < 4; ++tt) {
            const size_t at = (size_t)slot * kSgHid + tt * 16 + 4 * g;
            *reinterpret_cast<float4*>(hs[layer] + at) =
                make_float4(sg_silu(zs[layer][tt][0]), sg_silu(zs[layer][tt][1]), sg_silu(zs[layer][tt][2]),
                            sg_silu(zs[layer][tt][3]));
            *reinterpret_cast<float4*>(ds[layer] + at) =
                make_float4(dz[layer][tt][0], dz[layer][tt][1], dz[layer][tt][2], dz[layer][tt][3]);
        }
    // loss
    for (int off = 32; off > 0; off >>= 1) se += __shfl_xor(se, off);
    if (lane == 0 && se != 0.0f) atomicAdd(o.loss, se);
}

// Weight and bias gradients dW_l = sum_s dZ_l[s] (x) H_{l-1}[s] with the bias
// as an extra input column of ones.  One wave per (16x16 output tile, slice
// of kSgSlice samples); K runs over samples in steps of 32.
constexpr int kSgSlice = 256;
struct SgLayerDesc {
    int out_dim, in_dim, ot_tiles, it_tiles, w_off, b_off;
};
__constant__ SgLayerDesc c_layers[4] = {
    {kSgHid, kSgIn, 4, 1, kSgW1, kSgB1},
    {kSgHid, kSgHid, 4, 5, kSgW2, kSgB2},
    {kSgHid, kSgHid, 4, 5, kSgW3, kSgB3},
    {kSgOut, kSgHid, 1, 5, kSgW4, kSgB4},
};
constexpr int kSgTiles = 4 * 1 + 4 * 5 + 4 * 5 + 1 * 5;  // 49

__global__ __launch_bounds__(kSgBlock) void k_sg_wgrad(SgTrainBufs o, uint32_t nb, float* __restrict__ grad) {
    const uint32_t wave = blockIdx.x * (kSgBlock / 64) + (threadIdx.x >> 6);
    const uint32_t n_slices = nb / kSgSlice;
    if (wave >= (uint32_t)kSgTiles * n_slices) return;
    const uint32_t slice = wave % n_slices;
    int tile = (int)(wave / n_slices), layer = 0;
    while (tile >= c_layers[layer].ot_tiles * c_layers[layer].it_tiles) {
        tile -= c_layers[layer].ot_tiles * c_layers[layer].it_tiles;
        ++layer;
    }
    const SgLayerDesc L = c_layers[layer];
    const int ot = tile / L.it_tiles, it = tile % L.it_tiles;
    const float* hp = layer == 0 ? o.hin : (layer == 1 ? o.h1 : (layer == 2 ? o.h2 : o.h3));
    const int hs = layer == 0 ? 8 : kSgHid;
    const float* dp = layer == 0 ? o.dz1 : (layer == 1 ? o.dz2 : (layer == 2 ? o.dz3 : o.dout));
    const int ds = layer == 3 ? 4 : kSgHid;
    const uint32_t lane = sg_lane(), c = lane & 15u, g = lane >> 4;
    const int orow = ot * 16 + (int)c, icol = it * 16 + (int)c;
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int k0 = 0; k0 < kSgSlice; k0 += 32) {
        float a[8], bv[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const size_t s = (size_t)slice * kSgSlice + k0 + 8 * g + j;
            a[j] = orow < L.out_dim ? dp[s * ds + orow] : 0.0f;
            bv[j] = icol < L.in_dim ? hp[s * hs + icol] : (icol == L.in_dim ? 1.0f : 0.0f);
        }
        acc = sg_mfma(sg_pack8(a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7]),
                      sg_pack8(bv[0], bv[1], bv[2], bv[3], bv[4], bv[5], bv[6], bv[7]), acc);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int oo = ot * 16 + 4 * (int)g + r;
        if (oo >= L.out_dim) continue;
        if (icol < L.in_dim)
            atomicAdd(&grad[L.w_off + oo * L.in_dim + icol], acc[r]);
        else if (icol == L.in_dim)
            atomicAdd(&grad[L.b_off + oo], acc[r]);
    }
}

// Adam on the FP32 master weights; consumes and clears the gradient.
__global__ void k_sg_adam(float* __restrict__ w, float* __restrict__ m, float* __restrict__ v,
                          float* __restrict__ grad, SgAdam h) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kSgParams) return;
    const float gr = grad[i];
    grad[i] = 0.0f;
    const float mi = h.beta1 * m[i] + (1.0f - h.beta1) * gr;
    const float vi = h.beta2 * v[i] + (1.0f - h.beta2) * gr * gr;
    m[i] = mi;
    v[i] = vi;
    w[i] = w[i] - h.lr * (mi * h.c1) / (sqrtf(vi * h.c2) + h.eps);
}

// ------------------------------------------------------------ batch return mapping
__global__ void k_nacc_return_plain(DevMat m, const float4* __restrict__ f, const float* __restrict__ alpha,
                                    unsigned long long n, float4* __restrict__ fo, float* __restrict__ ao) {
    const unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 q = f[i];
    M2 nf;
    float na;
    nacc_project(m, M2{q.x, q.y, q.z, q.w}, alpha[i], nf, na);
    fo[i] = make_float4(nf.a, nf.b, nf.c, nf.d);
    ao[i] = na;
}

__global__ __launch_bounds__(kSgBlock, 4) void k_nacc_return_surrogate(DevMat m, const float4* __restrict__ f,
                                                                    const float* __restrict__ alpha,
                                                                    unsigned long long n, float4* __restrict__ fo,
                                                                    float* __restrict__ ao, SgDev sg) {
    static_assert(kSgBlock == kSgBlockLanes, "sg_return<4> queues one workgroup of kSgBlockLanes lanes");
    __shared__ float s_io[kSgQueueLds<4>];
    __shared__ SgLds s_net;
    sg_stage(sg.pk, sg.bias, s_net);  // once per workgroup: the grid strides over the particles
    const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
    for (unsigned long long b0 = (unsigned long long)blockIdx.x * blockDim.x; b0 < n; b0 += stride) {
        const unsigned long long i = b0 + threadIdx.x;
        const bool valid = i < n;
        M2 fin{1.0f, 0.0f, 0.0f, 1.0f};
        float a = 0.0f;
        if (valid) {
            const float4 q = f[i];
            fin = M2{q.x, q.y, q.z, q.w};
            a = alpha[i];
        }
        M2 nf;
        float na;
        sg_return<4>(m, sg, s_net, valid, fin, a, nf, na, s_io);
        if (valid) {
            fo[i] = make_float4(nf.a, nf.b, nf.c, nf.d);
            ao[i] = na;
        }
    }
}

// Network outputs for given raw inputs (testing and inspection).
__global__ __launch_bounds__(kSgBlock) void k_sg_infer(const float* __restrict__ xs, unsigned long long n,
                                                       float* __restrict__ ys, SgDev sg) {
    const uint32_t lane = sg_lane(), c = lane & 15u, g = lane >> 4;
    const unsigned long long grp = (unsigned long long)blockIdx.x * (kSgBlock / 64) + (threadIdx.x >> 6);
    const unsigned long long s = grp * 16 + c;
    if (grp * 16 >= n) return;
    float x[kSgIn] = {};
    if (s < n && g == 0)
        for (int k = 0; k < kSgIn; ++k) x[k] = (xs[s * kSgIn + k] - sg.norm[kSgInMu + k]) / sg.norm[kSgInSd + k];
    const f32x4 out = sg_forward(sg.pk, sg.bias, sg_pack8(x[0], x[1], x[2], x[3], x[4], x[5], 0.0f, 0.0f), nullptr);
    if (s < n && g == 0)
        for (int k = 0; k < kSgOut; ++k)
            ys[s * kSgOut + k] = xs[s * kSgIn + k] + (out[k] * sg.norm[kSgOutSd + k] + sg.norm[kSgOutMu + k]);
}

}  // namespace

hipError_t sg_launch_pack(const float* w, uint4* pk, float* bias, hipStream_t st) {
    k_sg_pack<<<div_up_u(kPkTiles * 64 * 4, kSgBlock), kSgBlock, 0, st>>>(w, reinterpret_cast<uint32_t*>(pk), bias);
    return hipGetLastError();
}
hipError_t sg_launch_prepare(const float* rec, unsigned long long n_rec, float* xs, float* ts,
                             unsigned long long* n_out, double* sums, hipStream_t st) {
    if (n_rec == 0) return hipSuccess;
    k_sg_prepare<<<(unsigned)((n_rec + kSgBlock - 1) / kSgBlock), kSgBlock, 0, st>>>(rec, n_rec, xs, ts, n_out, sums);
    return hipGetLastError();
}
hipError_t sg_launch_train_step(const SgBatch& b, const SgTrainBufs& o, float* w, float* m, float* v, float* grad,
                                const SgAdam* adam, float* bias, hipStream_t st) {
    const uint32_t groups = b.nb / 16u;
    k_sg_fwd_bwd<<<div_up_u(groups, kSgBlock / 64), kSgBlock, 0, st>>>(b, o);
    const uint32_t waves = (uint32_t)kSgTiles * (b.nb / kSgSlice);
    k_sg_wgrad<<<div_up_u(waves, kSgBlock / 64), kSgBlock, 0, st>>>(o, b.nb, grad);
    if (adam) {
        k_sg_adam<<<div_up_u(kSgParams, kSgBlock), kSgBlock, 0, st>>>(w, m, v, grad, *adam);
        k_sg_pack<<<div_up_u(kPkTiles * 64 * 4, kSgBlock), kSgBlock, 0, st>>>(
            w, reinterpret_cast<uint32_t*>(const_cast<uint4*>(b.pk)), bias);
    }
    return hipGetLastError();
}
hipError_t sg_launch_return_plain(const DevMat& m, const float4* f, const float* alpha, unsigned long long n,
                                  float4* fo, float* ao, hipStream_t st) {
    if (n == 0) return hipSuccess;
    k_nacc_return_plain<<<(unsigned)((n + kSgBlock - 1) / kSgBlock), kSgBlock, 0, st>>>(m, f, alpha, n, fo, ao);
    return hipGetLastError();
}
hipError_t sg_launch_return_surrogate(const DevMat& m, const float4* f, const float* alpha, unsigned long long n,
                                      float4* fo, float* ao, const SgDev& sg, hipStream_t st) {
    if (n == 0) return hipSuccess;
    // about four resident workgroups per CU (256 CUs), each staging the network once
    const unsigned long long blocks = std::min<unsigned long long>((n + kSgBlock - 1) / kSgBlock, 1024ull);
    k_nacc_return_surrogate<<<(unsigned)blocks, kSgBlock, 0, st>>>(m, f, alpha, n, fo, ao, sg);
    return hipGetLastError();
}
hipError_t sg_launch_infer(const float* xs, unsigned long long n, float* ys, const SgDev& sg, hipStream_t st) {
    if (n == 0) return hipSuccess;
    const unsigned long long groups = (n + 15) / 16;
    k_sg_infer<<<(unsigned)((groups + 3) / 4), kSgBlock, 0, st>>>(xs, n, ys, sg);
    return hipGetLastError();
}

}  // namespace emh
