// ta3h_eval_noalloc: the evaluation of ta3h_eval for callers that cannot allocate device memory (code under HIP
// graph capture, tight serving loops, harnesses that forbid hipMalloc). Inputs and outputs are device buffers of
// the caller, the work is enqueued on the caller's stream, and every scratch buffer the search needs lives in one
// static device array per GPU:
//   [0, 256)            work-queue counter, parked-game counter
//   weights             the models' weights repacked into the network's device layout (a kernel, not the host)
//   parked games        park_words(kind) words per item (two-phase schedule)
//   resume lists        parked list [items], resume order [items], per-model parked counts [models]
// Nothing here calls hipMalloc, hipFree or any other allocating routine. Calls that overlap on one GPU share the
// array and race; every counter the search reads is cleared by a kernel at the start of the call.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <mutex>

#include "kernels.hpp"
#include "ta3h.h"

namespace ta3h {
namespace {

constexpr size_t SCRATCH_BYTES = TA3H_NOALLOC_SCRATCH_BYTES;
static_assert(SCRATCH_BYTES % 256 == 0, "sections are 256-byte aligned");

// each GPU the library runs on holds its own copy until the process ends
static __device__ __align__(256) uint8_t scratch[SCRATCH_BYTES];

// device address of `scratch` per device, from hipGetSymbolAddress (the host-side name is not a device address)
std::mutex g_mu;
void* g_scratch[MAX_DEVICES] = {};
uint32_t g_numCUs[MAX_DEVICES] = {};

constexpr uint32_t WIDE_BYTES = f32w::BYTES;

__device__ inline uint32_t bf16_bits_rne(float v) { // the host repack's rounding (round to nearest even)
    uint32_t u = __float_as_uint(v);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return u >> 16;
}

// 32-bit word j of the wide network's device layout of one model (host layout src: W1[128][32], b1, W2[128][128],
// b2, W3[128], b3); the fragment halves are bf16 elements 2j and 2j + 1 of their region
__device__ inline uint32_t wide_word(float const* src, uint32_t j) {
    if(j < f32w::W2F / 4) { // w1f[t][q][r][e] = W1[16 t + r][8 q + e]
        uint32_t out = 0;
        for(uint32_t h2 = 0; h2 < 2; ++h2) {
            uint32_t const h = 2 * j + h2;
            uint32_t const e = h & 7u, r = (h >> 3) & 15u, q = (h >> 7) & 3u, t = h >> 9;
            out |= bf16_bits_rne(src[(16 * t + r) * 32 + 8 * q + e]) << (16 * h2);
        }
        return out;
    }
    if(j < f32w::B1 / 4) { // w2f[t][kb][q][r][e] = W2[16 t + r][unit(kb, q, e)]
        uint32_t out = 0;
        for(uint32_t h2 = 0; h2 < 2; ++h2) {
            uint32_t const h = 2 * (j - f32w::W2F / 4) + h2;
            uint32_t const e = h & 7u, r = (h >> 3) & 15u, q = (h >> 7) & 3u, kb = (h >> 9) & 3u, t = h >> 11;
            uint32_t const unit = e < 4 ? 32 * kb + 4 * q + e : 32 * kb + 16 + 4 * q + (e - 4);
            out |= bf16_bits_rne(src[4224 + (16 * t + r) * 128 + unit]) << (16 * h2);
        }
        return out;
    }
    if(j < f32w::B2 / 4) return __float_as_uint(src[4096 + (j - f32w::B1 / 4)]);
    if(j < f32w::W3 / 4) return __float_as_uint(src[4224 + 16384 + (j - f32w::B2 / 4)]);
    if(j < f32w::W3 / 4 + 129) return __float_as_uint(src[4224 + 16512 + (j - f32w::W3 / 4)]); // W3, b3
    return 0u;
}

// host layout (params floats per model) -> device layout (stride floats per model), one word per thread
__global__ void repack_kernel(int kind, float const* __restrict__ src, uint32_t params, uint32_t stride,
                              uint32_t models, uint32_t* __restrict__ dst) {
    uint64_t const i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if(i >= static_cast<uint64_t>(models) * stride) return;
    uint32_t const m = static_cast<uint32_t>(i / stride), j = static_cast<uint32_t>(i % stride);
    float const* const w = src + static_cast<uint64_t>(m) * params;
    uint32_t v = 0;
    if(kind == MODEL_F32W) {
        v = wide_word(w, j);
    } else if(kind == MODEL_F16 && j >= f16f::DENSE) { // the MFMA operands of the hidden layers (f16f)
        uint32_t const lane = j < f16f::C1 ? (j - f16f::DENSE) % 64 : ((j - f16f::C1) % 256) / 4;
        uint32_t const r = lane % 16, q = lane / 16; // row and quarter of the lane (q is k in the A fragments)
        if(j < f16f::A2) {
            uint32_t const g = (j - f16f::A1) / 64, rt = g / 4, s = g % 4;
            v = __float_as_uint(w[(16 * rt + 4 * (r % 4) + r / 4) * 16 + 4 * s + q]);
        } else if(j < f16f::C1) {
            uint32_t const g = (j - f16f::A2) / 64, rt2 = g / 8, rt1 = (g / 4) % 2, s = g % 4;
            v = __float_as_uint(w[544 + (16 * rt2 + r) * 32 + 16 * rt1 + 4 * s + q]);
        } else if(j < f16f::C2) {
            uint32_t const t = (j - f16f::C1) / 256, jj = (j - f16f::C1) % 4;
            v = __float_as_uint(w[512 + 16 * t + 4 * jj + q]);
        } else {
            uint32_t const t = (j - f16f::C2) / 256, i = (j - f16f::C2) % 4;
            v = __float_as_uint(w[1568 + 16 * t + 4 * q + i]);
        }
    } else if(kind == MODEL_F16) { // W2 [32 out][32 in] -> W2T [in][out]; the rest unchanged
        if(j >= 544 && j < 1568) {
            uint32_t const in = (j - 544) / 32, out = (j - 544) % 32;
            v = __float_as_uint(w[544 + out * 32 + in]);
        } else if(j < params) {
            v = __float_as_uint(w[j]);
        }
    } else if(j < params) {
        v = __float_as_uint(w[j]);
    }
    dst[i] = v;
}

uint32_t params_of(int kind) { return kind == MODEL_V11 ? 97u : kind == MODEL_F16 ? 1633u : 20865u; }
uint32_t stride_of(int kind) {
    uint32_t const n = kind == MODEL_F32W ? WIDE_BYTES / 4 : kind == MODEL_F16 ? f16f::FLOATS : params_of(kind);
    return (n + 15u) & ~15u;
}
size_t round256(size_t b) { return (b + 255) & ~static_cast<size_t>(255); }

// set-up once per device (with that device current): the device address of `scratch`, the CU count and the
// constant tables
int device_setup(int dev, void** base, uint32_t* cus) {
    std::lock_guard<std::mutex> const lock(g_mu);
    if(g_scratch[dev] == nullptr) {
        void* p = nullptr;
        if(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || p == nullptr) return TA3H_E_HIP;
        int n = 0;
        if(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0)
            return TA3H_E_HIP;
        if(upload_constants() != hipSuccess) return TA3H_E_HIP;
        g_numCUs[dev] = static_cast<uint32_t>(n);
        g_scratch[dev] = p;
    }
    *base = g_scratch[dev];
    *cus = g_numCUs[dev];
    return TA3H_OK;
}

} // namespace
} // namespace ta3h

extern "C" int ta3h_eval_noalloc(int kind, const float* d_weights, uint32_t num_models, const uint64_t* d_seeds,
                                 uint32_t games, uint32_t max_moves, float* d_fitness, uint32_t* d_placements,
                                 void* stream) {
    using namespace ta3h;
    if(kind != MODEL_V11 && kind != MODEL_F16 && kind != MODEL_F32W) return TA3H_E_ARG;
    uint64_t const items = static_cast<uint64_t>(num_models) * games;
    if(items == 0) return TA3H_OK;
    if(d_weights == nullptr || d_seeds == nullptr || d_fitness == nullptr) return TA3H_E_ARG;
    if(items > 0xFFFFFFFFull / 2) return TA3H_E_ARG;
    // the two-phase schedule with the default budget: max_moves / 32 clamped to [64, 512]
    uint32_t const budget = max_moves / 32 < 64 ? 64u : (max_moves / 32 > 512 ? 512u : max_moves / 32);
    uint64_t const ringSlots = budget < max_moves ? ring_slots(items, max_moves, budget) : 0;
    if(ringSlots > 0xFFFFFFFFull) return TA3H_E_SCRATCH;
    // scratch sections: counters, repacked weights, parked games, resume list, per-model counts and work ring
    size_t const wBytes = round256(static_cast<size_t>(num_models) * stride_of(kind) * 4);
    size_t const parkBytes = round256(static_cast<size_t>(items) * park_words(kind) * 4);
    size_t const resumeBytes = round256((static_cast<size_t>(items) + num_models + 2 + ringSlots) * 4);
    if(256 + wBytes + parkBytes + resumeBytes > SCRATCH_BYTES) return TA3H_E_SCRATCH; // before touching the device
    int dev = 0;
    if(hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES) return TA3H_E_HIP;
    void* base = nullptr;
    uint32_t cus = 0;
    int r = device_setup(dev, &base, &cus);
    if(r != TA3H_OK) return r;
    int const wpg = 4; // the networks' measured launch shape
    int perCU = 0;
    if(max_active_blocks(kind, wpg, 0, &perCU) != hipSuccess) return TA3H_E_HIP;
    uint64_t blocks = static_cast<uint64_t>(cus) * static_cast<uint32_t>(perCU > 0 ? perCU : 1);
    if(blocks > items) blocks = items;
    auto* const s = static_cast<hipStream_t>(stream);
    uint8_t* const p = static_cast<uint8_t*>(base);
    uint32_t* const counters = reinterpret_cast<uint32_t*>(p);
    float* const weights = reinterpret_cast<float*>(p + 256);
    uint32_t* const parked = reinterpret_cast<uint32_t*>(p + 256 + wBytes);
    uint32_t* const resume = reinterpret_cast<uint32_t*>(p + 256 + wBytes + parkBytes);

    uint64_t const words = static_cast<uint64_t>(num_models) * stride_of(kind);
    hipLaunchKernelGGL(repack_kernel, dim3(static_cast<uint32_t>((words + 255) / 256)), dim3(256), 0, s, kind,
                       d_weights, params_of(kind), stride_of(kind), num_models, reinterpret_cast<uint32_t*>(weights));
    if(hipGetLastError() != hipSuccess) return TA3H_E_HIP;

    EvalArgs a{};
    a.weights = weights;
    a.seeds = d_seeds;
    a.fitness = d_fitness;
    a.placements = d_placements;
    a.queue = counters;
    a.items = static_cast<uint32_t>(items);
    a.games = games;
    a.maxMoves = max_moves;
    a.weightStride = stride_of(kind);
    a.beamWidth[0] = 16; // the reference's widths (variation_sequences.hpp:22)
    a.beamWidth[1] = 32;
    a.beamWidth[2] = 32;
    a.persistent = 1;
    if(budget < max_moves) {
        a.phaseMoves = budget;
        a.parked = parked;
        a.resumeList = resume;
        a.resumeCount = counters + 1;
        a.modelParked = resume + items;
        a.ring = resume + items + num_models;
        a.ringCap = static_cast<uint32_t>(ringSlots);
    }
    if(launch_beam_eval(kind, a, wpg, static_cast<uint32_t>(blocks), s) != hipSuccess) return TA3H_E_HIP;
    return TA3H_OK;
}
