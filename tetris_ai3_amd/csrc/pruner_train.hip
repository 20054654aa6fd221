// Training kernels of the placement-proposal network (the pruner): forward, backward and Adam in FP32 with
// the inference roundings (bfloat16 W1, W2 and first hidden layer, straight-through gradients).
//
// Network (build-defined): 265 inputs -> 64 ReLU -> 64 ReLU -> 1 logit. Inputs of a candidate placement:
// board slot 24c + r = bit (3 + r) of the parent board's column c (r < 20; r = 20..23 never set), then
// one-hots piece (240 + p), rotation (247 + o), column (251 + x + 2) and ply (262 + ply - 1).
// Parameters: W1[64][265], b1[64], W2[64][64], b2[64], W3[64], b3 (21,249 floats).
// Sample = trace record: parent board (10 words) + meta = piece | o << 3 | (x + 2) << 5 | ply << 9 | label << 11.
// Loss: batch-mean weighted binary cross-entropy on the logit (weight posWeight for label 1, else 1).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "pruner.hpp"

namespace ta3h {
namespace {

__device__ inline float bf16r(float v) {
    uint32_t u = __float_as_uint(v);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return __uint_as_float(u & 0xFFFF0000u);
}

struct SampleFields {
    int piece, o, x, ply;
    float y;
};

__device__ inline SampleFields fields(uint32_t meta) {
    return {static_cast<int>(meta & 7u), static_cast<int>((meta >> 3) & 3u), static_cast<int>((meta >> 5) & 15u) - 2,
            static_cast<int>((meta >> 9) & 3u), static_cast<float>((meta >> 11) & 1u)};
}

// input k of a sample: 1 or 0
__device__ inline float input_bit(uint32_t const* rec, SampleFields const& f, int k) {
    if(k < 240) {
        int const c = k / 24, r = k % 24;
        return (r < 20 && ((rec[c] >> (3 + r)) & 1u)) ? 1.0f : 0.0f;
    }
    bool on = k == 240 + f.piece || k == 247 + f.o || k == 251 + f.x + 2 || k == 262 + f.ply - 1;
    return on ? 1.0f : 0.0f;
}

// One wave per sample, lane = hidden unit. Stores what the gradient kernel needs:
// act[b] = {pre1[64], h1[64], pre2[64], h2[64]}, delta[b] = {dpre1[64], dpre2[64]}, dscore[b], loss[b].
__global__ __launch_bounds__(64) void pr_fwd_bwd(float const* __restrict__ w, uint32_t const* __restrict__ samples,
                                                 uint32_t const* __restrict__ order, uint32_t n, float invBatch,
                                                 float posWeight, float* __restrict__ act, float* __restrict__ delta,
                                                 float* __restrict__ dscore, float* __restrict__ lossOut) {
    __shared__ float sh1[PR_H], sdp2[PR_H];
    uint32_t const b = blockIdx.x;
    if(b >= n) return;
    uint32_t const u = threadIdx.x;
    uint32_t const* rec = samples + 11ull * (order ? order[b] : b);
    SampleFields const f = fields(rec[10]);
    // layer 1: bias then the active inputs in ascending order
    float pre1 = w[PR_B1 + u];
    float const* w1 = w + PR_W1 + u * PR_IN;
    for(int c = 0; c < 10; ++c) {
        uint32_t bits = (rec[c] >> 3) & 0xFFFFFu;
        while(bits) {
            int const r = __builtin_ctz(bits);
            bits &= bits - 1;
            pre1 += bf16r(w1[24 * c + r]);
        }
    }
    pre1 += bf16r(w1[240 + f.piece]);
    pre1 += bf16r(w1[247 + f.o]);
    pre1 += bf16r(w1[251 + f.x + 2]);
    pre1 += bf16r(w1[262 + f.ply - 1]);
    float const h1 = bf16r(pre1 > 0.0f ? pre1 : 0.0f);
    sh1[u] = h1;
    __syncthreads();
    float pre2 = w[PR_B2 + u];
    for(int i = 0; i < PR_H; ++i) pre2 = __fmaf_rn(sh1[i], bf16r(w[PR_W2 + u * PR_H + i]), pre2);
    float const h2 = pre2 > 0.0f ? pre2 : 0.0f;
    float part = h2 * w[PR_W3 + u];
    for(int m = 32; m > 0; m >>= 1) part += __shfl_xor(part, m);
    float const s = part + w[PR_B3];
    float const wt = f.y > 0.5f ? posWeight : 1.0f;
    float const sig = 1.0f / (1.0f + expf(-s));
    float const ds = wt * (sig - f.y) * invBatch;
    float const dp2 = pre2 > 0.0f ? ds * w[PR_W3 + u] : 0.0f;
    sdp2[u] = dp2;
    __syncthreads();
    float dh1 = 0.0f;
    for(int v = 0; v < PR_H; ++v) dh1 = __fmaf_rn(sdp2[v], bf16r(w[PR_W2 + v * PR_H + u]), dh1);
    float const dp1 = pre1 > 0.0f ? dh1 : 0.0f;
    float* a = act + 4ull * PR_H * b;
    a[u] = pre1;
    a[PR_H + u] = h1;
    a[2 * PR_H + u] = pre2;
    a[3 * PR_H + u] = h2;
    float* d = delta + 2ull * PR_H * b;
    d[u] = dp1;
    d[PR_H + u] = dp2;
    if(u == 0) {
        dscore[b] = ds;
        float const sp = s > 0.0f ? s + log1pf(expf(-s)) : log1pf(expf(s));
        lossOut[b] = wt * (sp - f.y * s) * invBatch;
    }
}

// The batch gradient in PR_SPLIT fixed sample ranges: thread (p, s) sums parameter p's per-sample terms over
// range s in sample order into partial[s][p]. Each term's loads depend on the previous load (sample index,
// record, delta), so splitting the batch shortens every thread's chain of L2 round trips and fills the GPU with
// PR_SPLIT times as many waves: a 512-sample batch's gradient takes 0.027 ms (both kernels) against 0.27 ms for
// one thread per parameter over the whole batch.
__global__ void pr_grad_partial(float* __restrict__ partial, uint32_t const* __restrict__ samples,
                                uint32_t const* __restrict__ order, uint32_t n, float const* __restrict__ act,
                                float const* __restrict__ delta, float const* __restrict__ dscore) {
    uint32_t const p = blockIdx.x * blockDim.x + threadIdx.x;
    if(p >= static_cast<uint32_t>(PR_PARAMS)) return;
    uint32_t const span = (n + PR_SPLIT - 1) / PR_SPLIT;
    uint32_t const b0 = blockIdx.y * span;
    uint32_t const b1 = b0 + span < n ? b0 + span : n;
    float g = 0.0f;
    if(p < static_cast<uint32_t>(PR_B1)) { // W1[u][k]
        uint32_t const u = p / PR_IN, k = p % PR_IN;
        for(uint32_t b = b0; b < b1; ++b) {
            uint32_t const* rec = samples + 11ull * (order ? order[b] : b);
            if(input_bit(rec, fields(rec[10]), static_cast<int>(k)) != 0.0f) g += delta[2ull * PR_H * b + u];
        }
    } else if(p < static_cast<uint32_t>(PR_W2)) { // b1[u]
        uint32_t const u = p - PR_B1;
        for(uint32_t b = b0; b < b1; ++b) g += delta[2ull * PR_H * b + u];
    } else if(p < static_cast<uint32_t>(PR_B2)) { // W2[v][u]
        uint32_t const q = p - PR_W2, vv = q / PR_H, u = q % PR_H;
        for(uint32_t b = b0; b < b1; ++b) g = __fmaf_rn(delta[2ull * PR_H * b + PR_H + vv], act[4ull * PR_H * b + PR_H + u], g);
    } else if(p < static_cast<uint32_t>(PR_W3)) { // b2[v]
        uint32_t const vv = p - PR_B2;
        for(uint32_t b = b0; b < b1; ++b) g += delta[2ull * PR_H * b + PR_H + vv];
    } else if(p < static_cast<uint32_t>(PR_B3)) { // W3[v]
        uint32_t const vv = p - PR_W3;
        for(uint32_t b = b0; b < b1; ++b) g = __fmaf_rn(dscore[b], act[4ull * PR_H * b + 3 * PR_H + vv], g);
    } else { // b3
        for(uint32_t b = b0; b < b1; ++b) g += dscore[b];
    }
    partial[static_cast<uint64_t>(blockIdx.y) * PR_PARAMS + p] = g;
}

// One thread per parameter: the partial sums added in range order, then (optionally) the Adam step.
__global__ void pr_grad_adam(float* __restrict__ w, float* __restrict__ m, float* __restrict__ v,
                             float* __restrict__ gradOut, float const* __restrict__ partial, int adam, float lr,
                             float beta1, float beta2, float eps, float corr1, float corr2) {
    uint32_t const p = blockIdx.x * blockDim.x + threadIdx.x;
    if(p >= static_cast<uint32_t>(PR_PARAMS)) return;
    float g = 0.0f;
#pragma unroll
    for(int s = 0; s < PR_SPLIT; ++s) g += partial[static_cast<uint64_t>(s) * PR_PARAMS + p];
    if(gradOut) gradOut[p] = g;
    if(!adam) return;
    float const mm = beta1 * m[p] + (1.0f - beta1) * g;
    float const vv2 = beta2 * v[p] + (1.0f - beta2) * g * g;
    m[p] = mm;
    v[p] = vv2;
    float const mhat = mm / corr1, vhat = vv2 / corr2;
    w[p] -= lr * mhat / (sqrtf(vhat) + eps);
}

// Plain-FP32 pruner logits of samples (same arithmetic as the training forward pass).
__global__ __launch_bounds__(64) void pr_score(float const* __restrict__ w, uint32_t const* __restrict__ samples,
                                               uint32_t n, float* __restrict__ out) {
    __shared__ float sh1[PR_H];
    uint32_t const b = blockIdx.x;
    if(b >= n) return;
    uint32_t const u = threadIdx.x;
    uint32_t const* rec = samples + 11ull * b;
    SampleFields const f = fields(rec[10]);
    float pre1 = w[PR_B1 + u];
    float const* w1 = w + PR_W1 + u * PR_IN;
    for(int c = 0; c < 10; ++c) {
        uint32_t bits = (rec[c] >> 3) & 0xFFFFFu;
        while(bits) {
            int const r = __builtin_ctz(bits);
            bits &= bits - 1;
            pre1 += bf16r(w1[24 * c + r]);
        }
    }
    pre1 += bf16r(w1[240 + f.piece]);
    pre1 += bf16r(w1[247 + f.o]);
    pre1 += bf16r(w1[251 + f.x + 2]);
    pre1 += bf16r(w1[262 + f.ply - 1]);
    sh1[u] = bf16r(pre1 > 0.0f ? pre1 : 0.0f);
    __syncthreads();
    float pre2 = w[PR_B2 + u];
    for(int i = 0; i < PR_H; ++i) pre2 = __fmaf_rn(sh1[i], bf16r(w[PR_W2 + u * PR_H + i]), pre2);
    float part = (pre2 > 0.0f ? pre2 : 0.0f) * w[PR_W3 + u];
    for(int mm = 32; mm > 0; mm >>= 1) part += __shfl_xor(part, mm);
    if(u == 0) out[b] = part + w[PR_B3];
}

// First hidden layer of samples as trained (bfloat16 ReLU outputs): out[b][u].
__global__ __launch_bounds__(64) void pr_hidden(float const* __restrict__ w, uint32_t const* __restrict__ samples,
                                                uint32_t n, float* __restrict__ out) {
    uint32_t const b = blockIdx.x;
    if(b >= n) return;
    uint32_t const u = threadIdx.x;
    uint32_t const* rec = samples + 11ull * b;
    SampleFields const f = fields(rec[10]);
    float pre1 = w[PR_B1 + u];
    float const* w1 = w + PR_W1 + u * PR_IN;
    for(int c = 0; c < 10; ++c) {
        uint32_t bits = (rec[c] >> 3) & 0xFFFFFu;
        while(bits) {
            int const r = __builtin_ctz(bits);
            bits &= bits - 1;
            pre1 += bf16r(w1[24 * c + r]);
        }
    }
    pre1 += bf16r(w1[240 + f.piece]);
    pre1 += bf16r(w1[247 + f.o]);
    pre1 += bf16r(w1[251 + f.x + 2]);
    pre1 += bf16r(w1[262 + f.ply - 1]);
    out[64ull * b + u] = bf16r(pre1 > 0.0f ? pre1 : 0.0f);
}

} // namespace

hipError_t pruner_hidden_samples(float const* w, uint32_t const* samples, uint32_t n, float* out, hipStream_t s) {
    if(n == 0) return hipSuccess;
    hipLaunchKernelGGL(pr_hidden, dim3(n), dim3(64), 0, s, w, samples, n, out);
    return hipGetLastError();
}

hipError_t pruner_fwd_bwd(float const* w, uint32_t const* samples, uint32_t const* order, uint32_t n, float invBatch,
                          float posWeight, float* act, float* delta, float* dscore, float* loss, hipStream_t s) {
    hipLaunchKernelGGL(pr_fwd_bwd, dim3(n), dim3(64), 0, s, w, samples, order, n, invBatch, posWeight, act, delta,
                       dscore, loss);
    return hipGetLastError();
}

hipError_t pruner_grad_adam(float* w, float* m, float* v, float* gradOut, float* partial, uint32_t const* samples,
                            uint32_t const* order, uint32_t n, float const* act, float const* delta,
                            float const* dscore, int adam, float lr, float beta1, float beta2, float eps, float corr1,
                            float corr2, hipStream_t s) {
    hipLaunchKernelGGL(pr_grad_partial, dim3((PR_PARAMS + 255) / 256, PR_SPLIT), dim3(256), 0, s, partial, samples,
                       order, n, act, delta, dscore);
    hipError_t const e = hipGetLastError();
    if(e != hipSuccess) return e;
    hipLaunchKernelGGL(pr_grad_adam, dim3((PR_PARAMS + 255) / 256), dim3(256), 0, s, w, m, v, gradOut, partial, adam,
                       lr, beta1, beta2, eps, corr1, corr2);
    return hipGetLastError();
}

hipError_t pruner_score_samples(float const* w, uint32_t const* samples, uint32_t n, float* out, hipStream_t s) {
    hipLaunchKernelGGL(pr_score, dim3(n), dim3(64), 0, s, w, samples, n, out);
    return hipGetLastError();
}

} // namespace ta3h
