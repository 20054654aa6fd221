// Host driver and C-ABI of the evaluator (declared in include/ta3h.h).
//
// Mirrors the reference host driver's contract (ta3_gpu/src/cuda/model_eval.cu:46-153): one
// non-blocking stream per evaluator, grow-only device buffers reused across calls, one host
// synchronisation per evaluation, every runtime error reported as "hip: <what>: <message>".
// Differences by design: the games are seeded on the device (a 144 B engine per item is never built
// or copied by the host), weights are repacked into the device layout of the selected network, and
// the search kernel is persistent over a device work queue.
#include "ta3h.h"

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "kernels.hpp"
#include "tables.hpp"

namespace {

int params_of(int kind) {
    switch(kind) {
        case TA3H_MODEL_V11: return 97;
        case TA3H_MODEL_F16: return 1633;
        default: return -1;
    }
}

// floats per model in device memory (a multiple of 16 so every model starts 64-byte aligned)
uint32_t stride_of(int kind) { return (static_cast<uint32_t>(params_of(kind)) + 15u) & ~15u; }

// host layout (reference dense(), utility.hpp:330-347) -> device layout of the selected network
void repack(int kind, float const* src, float* dst) {
    int const P = params_of(kind);
    if(kind == TA3H_MODEL_F16) {
        // W1[32][16] (512), b1 (32) unchanged; W2[32 out][32 in] -> W2T[in][out]; b2, W3, b3 unchanged
        std::memcpy(dst, src, 544 * sizeof(float));
        for(int o = 0; o < 32; ++o)
            for(int i = 0; i < 32; ++i) dst[544 + i * 32 + o] = src[544 + o * 32 + i];
        std::memcpy(dst + 1568, src + 1568, 65 * sizeof(float));
    } else
        std::memcpy(dst, src, static_cast<size_t>(P) * sizeof(float));
    for(uint32_t i = static_cast<uint32_t>(P); i < stride_of(kind); ++i) dst[i] = 0.0f;
}

} // namespace

struct ta3h_evaluator {
    hipStream_t stream = nullptr;
    hipEvent_t evStart = nullptr, evStop = nullptr;
    bool ready = false;
    int device = 0;
    uint32_t numCUs = 0;

    float* dWeights = nullptr;
    size_t weightsCap = 0;
    uint64_t* dSeeds = nullptr;
    size_t seedsCap = 0;
    float* dFitness = nullptr;
    size_t fitnessCap = 0;
    uint32_t* dPlacements = nullptr;
    size_t placementsCap = 0;
    uint64_t* dCounters = nullptr;
    size_t countersCap = 0;
    uint64_t* dTimeline = nullptr;
    size_t timelineCap = 0;
    bool timelineOn = false;
    uint32_t* dQueue = nullptr;

    int kind = TA3H_MODEL_V11;
    int wavesPerBlock = 2; // waves cooperating on one game
    int persistent = 1;
    uint32_t gridBlocks = 0; // 0: occupancy-derived
    uint32_t lastGrid = 0;

    // staged batch
    bool staged = false;
    bool launched = false;
    uint32_t items = 0, games = 0, maxMoves = 0;

    std::vector<float> packed;
    std::string err;

    int fail(char const* what, hipError_t e) {
        err = std::string("hip: ") + what + ": " + hipGetErrorString(e);
        return TA3H_E_HIP;
    }
    int fail_arg(std::string msg) {
        err = std::move(msg);
        return TA3H_E_ARG;
    }

    int ensure_ready() {
        if(ready) return TA3H_OK;
        hipError_t e = hipGetDevice(&device);
        if(e != hipSuccess) return fail("get device", e);
        int cus = 0;
        e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
        if(e != hipSuccess) return fail("device attribute", e);
        numCUs = static_cast<uint32_t>(cus);
        e = hipStreamCreateWithFlags(&stream, hipStreamNonBlocking);
        if(e != hipSuccess) return fail("create stream", e);
        e = hipEventCreate(&evStart);
        if(e != hipSuccess) return fail("create event", e);
        e = hipEventCreate(&evStop);
        if(e != hipSuccess) return fail("create event", e);
        e = hipMalloc(&dQueue, 64);
        if(e != hipSuccess) return fail("malloc queue", e);
        e = ta3h::upload_constants(); // per process and device; idempotent
        if(e != hipSuccess) return fail("upload constants", e);
        ready = true;
        return TA3H_OK;
    }

    template<class T>
    int grow(T*& ptr, size_t& cap, size_t need, char const* what) {
        if(need <= cap) return TA3H_OK;
        if(ptr != nullptr) {
            hipError_t const e = hipFree(ptr);
            ptr = nullptr;
            cap = 0;
            if(e != hipSuccess) return fail(what, e);
        }
        hipError_t const e = hipMalloc(&ptr, need * sizeof(T));
        if(e != hipSuccess) return fail(what, e);
        cap = need;
        return TA3H_OK;
    }

    ~ta3h_evaluator() {
        // best effort: a dying evaluator never reports (model_eval.cu:72-83)
        if(dWeights) static_cast<void>(hipFree(dWeights));
        if(dSeeds) static_cast<void>(hipFree(dSeeds));
        if(dFitness) static_cast<void>(hipFree(dFitness));
        if(dPlacements) static_cast<void>(hipFree(dPlacements));
        if(dCounters) static_cast<void>(hipFree(dCounters));
        if(dTimeline) static_cast<void>(hipFree(dTimeline));
        if(dQueue) static_cast<void>(hipFree(dQueue));
        if(evStart) static_cast<void>(hipEventDestroy(evStart));
        if(evStop) static_cast<void>(hipEventDestroy(evStop));
        if(stream) static_cast<void>(hipStreamDestroy(stream));
    }
};

extern "C" {

int ta3h_create(ta3h_evaluator** out) {
    if(out == nullptr) return TA3H_E_ARG;
    *out = new(std::nothrow) ta3h_evaluator();
    return *out ? TA3H_OK : TA3H_E_ARG;
}

void ta3h_destroy(ta3h_evaluator* ev) { delete ev; }

const char* ta3h_last_error(const ta3h_evaluator* ev) { return ev ? ev->err.c_str() : "null evaluator"; }

int ta3h_num_params(int kind) { return params_of(kind); }

int ta3h_set_model(ta3h_evaluator* ev, int kind) {
    if(!ev) return TA3H_E_ARG;
    if(params_of(kind) < 0) return ev->fail_arg("set_model: unknown model kind " + std::to_string(kind));
    ev->kind = kind;
    ev->staged = false;
    return TA3H_OK;
}

int ta3h_set_launch(ta3h_evaluator* ev, int waves_per_game, int persistent, uint32_t grid_blocks) {
    if(!ev) return TA3H_E_ARG;
    if(waves_per_game != 1 && waves_per_game != 2 && waves_per_game != 4)
        return ev->fail_arg("set_launch: waves_per_game must be 1, 2 or 4");
    ev->wavesPerBlock = waves_per_game;
    ev->persistent = persistent ? 1 : 0;
    ev->gridBlocks = grid_blocks;
    return TA3H_OK;
}

int ta3h_stage(ta3h_evaluator* ev, const float* weights, uint32_t num_models, uint32_t num_params,
               const uint64_t* seeds, uint32_t games, uint32_t max_moves) {
    if(!ev) return TA3H_E_ARG;
    ev->staged = false;
    ev->launched = false;
    uint64_t const items = static_cast<uint64_t>(num_models) * games;
    if(items > 0xFFFFFFFFull / 2) return ev->fail_arg("eval: num_models * games too large");
    int const P = params_of(ev->kind);
    if(static_cast<int64_t>(num_params) != P)
        return ev->fail_arg("eval: weights per model must equal NUM_PARAMS (" + std::to_string(P) + ") of the model");
    if(items > 0 && (weights == nullptr || seeds == nullptr)) return ev->fail_arg("eval: null weights or seeds");
    ev->items = static_cast<uint32_t>(items);
    ev->games = games;
    ev->maxMoves = max_moves;
    if(items == 0) {
        ev->staged = true;
        return TA3H_OK;
    }
    int r = ev->ensure_ready();
    if(r) return r;
    uint32_t const stride = stride_of(ev->kind);
    size_t const wCount = static_cast<size_t>(num_models) * stride;
    if((r = ev->grow(ev->dWeights, ev->weightsCap, wCount, "malloc weights"))) return r;
    if((r = ev->grow(ev->dSeeds, ev->seedsCap, games, "malloc seeds"))) return r;
    if((r = ev->grow(ev->dFitness, ev->fitnessCap, items, "malloc fitness"))) return r;
    if((r = ev->grow(ev->dPlacements, ev->placementsCap, items, "malloc placements"))) return r;
    if((r = ev->grow(ev->dCounters, ev->countersCap, 2 * items, "malloc counters"))) return r;
    if(ev->timelineOn && (r = ev->grow(ev->dTimeline, ev->timelineCap, 2 * items, "malloc timeline"))) return r;
    ev->packed.resize(wCount);
    for(uint32_t m = 0; m < num_models; ++m)
        repack(ev->kind, weights + static_cast<size_t>(m) * num_params, ev->packed.data() + static_cast<size_t>(m) * stride);
    hipError_t e = hipMemcpyAsync(ev->dWeights, ev->packed.data(), wCount * sizeof(float), hipMemcpyHostToDevice, ev->stream);
    if(e != hipSuccess) return ev->fail("copy weights", e);
    e = hipMemcpyAsync(ev->dSeeds, seeds, games * sizeof(uint64_t), hipMemcpyHostToDevice, ev->stream);
    if(e != hipSuccess) return ev->fail("copy seeds", e);
    e = hipStreamSynchronize(ev->stream); // the host staging vector may be reused by the next stage
    if(e != hipSuccess) return ev->fail("stage", e);
    ev->staged = true;
    return TA3H_OK;
}

int ta3h_launch(ta3h_evaluator* ev) {
    if(!ev) return TA3H_E_ARG;
    if(!ev->staged) {
        ev->err = "launch: nothing staged";
        return TA3H_E_STATE;
    }
    ev->launched = true;
    if(ev->items == 0) return TA3H_OK;
    ta3h::EvalArgs a{};
    a.weights = ev->dWeights;
    a.seeds = ev->dSeeds;
    a.fitness = ev->dFitness;
    a.placements = ev->dPlacements;
    a.counters = ev->dCounters;
    a.timeline = ev->timelineOn ? ev->dTimeline : nullptr;
    a.queue = ev->dQueue;
    a.items = ev->items;
    a.games = ev->games;
    a.maxMoves = ev->maxMoves;
    a.weightStride = stride_of(ev->kind);
    a.persistent = static_cast<uint32_t>(ev->persistent);
    // one workgroup (of wavesPerBlock cooperating waves) per game in flight
    uint32_t const needed = ev->items;
    uint32_t blocks;
    hipError_t e0 = hipSuccess;
    if(ev->persistent) {
        blocks = ev->gridBlocks;
        if(blocks == 0) {
            // fill every CU up to the kernel's occupancy limit; workgroups beyond it would only queue
            int perCU = 0;
            e0 = ta3h::max_active_blocks(ev->kind, ev->wavesPerBlock, &perCU);
            if(e0 != hipSuccess) return ev->fail("occupancy", e0);
            blocks = ev->numCUs * static_cast<uint32_t>(perCU > 0 ? perCU : 1);
        }
        if(blocks > needed) blocks = needed;
    } else
        blocks = needed;
    ev->lastGrid = blocks;
    hipError_t e = hipEventRecord(ev->evStart, ev->stream);
    if(e != hipSuccess) return ev->fail("event record", e);
    e = ta3h::launch_beam_eval(ev->kind, a, ev->wavesPerBlock, blocks, ev->stream);
    if(e != hipSuccess) return ev->fail("kernel launch", e);
    e = hipEventRecord(ev->evStop, ev->stream);
    if(e != hipSuccess) return ev->fail("event record", e);
    return TA3H_OK;
}

int ta3h_sync(ta3h_evaluator* ev) {
    if(!ev) return TA3H_E_ARG;
    if(!ev->ready) return TA3H_OK;
    hipError_t const e = hipStreamSynchronize(ev->stream);
    if(e != hipSuccess) return ev->fail("kernel run", e);
    return TA3H_OK;
}

int ta3h_fetch(ta3h_evaluator* ev, float* out_fitness, uint32_t* out_placements) {
    if(!ev) return TA3H_E_ARG;
    if(!ev->launched) {
        ev->err = "fetch: nothing launched";
        return TA3H_E_STATE;
    }
    if(ev->items == 0) return TA3H_OK;
    hipError_t e;
    if(out_fitness) {
        e = hipMemcpyAsync(out_fitness, ev->dFitness, ev->items * sizeof(float), hipMemcpyDeviceToHost, ev->stream);
        if(e != hipSuccess) return ev->fail("copy fitness", e);
    }
    if(out_placements) {
        e = hipMemcpyAsync(out_placements, ev->dPlacements, ev->items * sizeof(uint32_t), hipMemcpyDeviceToHost,
                           ev->stream);
        if(e != hipSuccess) return ev->fail("copy placements", e);
    }
    e = hipStreamSynchronize(ev->stream);
    if(e != hipSuccess) return ev->fail("kernel run", e);
    return TA3H_OK;
}

int ta3h_fetch_counters(ta3h_evaluator* ev, uint64_t* out_counters) {
    if(!ev || !out_counters) return TA3H_E_ARG;
    if(!ev->launched) {
        ev->err = "fetch: nothing launched";
        return TA3H_E_STATE;
    }
    if(ev->items == 0) return TA3H_OK;
    hipError_t e = hipMemcpyAsync(out_counters, ev->dCounters, 2 * static_cast<size_t>(ev->items) * sizeof(uint64_t),
                                  hipMemcpyDeviceToHost, ev->stream);
    if(e != hipSuccess) return ev->fail("copy counters", e);
    e = hipStreamSynchronize(ev->stream);
    if(e != hipSuccess) return ev->fail("kernel run", e);
    return TA3H_OK;
}

int ta3h_set_timeline(ta3h_evaluator* ev, int enabled) {
    if(!ev) return TA3H_E_ARG;
    ev->timelineOn = enabled != 0;
    ev->staged = false;
    return TA3H_OK;
}

int ta3h_fetch_timeline(ta3h_evaluator* ev, uint64_t* out, double* ticks_per_second) {
    if(!ev || !out) return TA3H_E_ARG;
    if(!ev->launched || !ev->timelineOn) {
        ev->err = "fetch_timeline: no launch with the timeline enabled";
        return TA3H_E_STATE;
    }
    if(ticks_per_second) {
        int khz = 0;
        hipError_t const e = hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, ev->device);
        if(e != hipSuccess) return ev->fail("wall clock rate", e);
        *ticks_per_second = 1000.0 * khz;
    }
    if(ev->items == 0) return TA3H_OK;
    hipError_t e = hipMemcpyAsync(out, ev->dTimeline, 2 * static_cast<size_t>(ev->items) * sizeof(uint64_t),
                                  hipMemcpyDeviceToHost, ev->stream);
    if(e != hipSuccess) return ev->fail("copy timeline", e);
    e = hipStreamSynchronize(ev->stream);
    if(e != hipSuccess) return ev->fail("kernel run", e);
    return TA3H_OK;
}

int ta3h_eval(ta3h_evaluator* ev, const float* weights, uint32_t num_models, uint32_t num_params,
              const uint64_t* seeds, uint32_t games, uint32_t max_moves, float* out_fitness) {
    int r = ta3h_stage(ev, weights, num_models, num_params, seeds, games, max_moves);
    if(r) return r;
    if((r = ta3h_launch(ev))) return r;
    return ta3h_fetch(ev, out_fitness, nullptr);
}

int ta3h_last_kernel_ms(ta3h_evaluator* ev, float* ms) {
    if(!ev || !ms) return TA3H_E_ARG;
    if(!ev->ready || !ev->launched || ev->items == 0) {
        *ms = 0.0f;
        return TA3H_OK;
    }
    hipError_t const e = hipEventElapsedTime(ms, ev->evStart, ev->evStop);
    if(e != hipSuccess) return ev->fail("event elapsed", e);
    return TA3H_OK;
}

int ta3h_launch_info(ta3h_evaluator* ev, uint32_t* grid_blocks, uint32_t* waves_per_game, uint32_t* num_cus) {
    if(!ev) return TA3H_E_ARG;
    if(grid_blocks) *grid_blocks = ev->lastGrid;
    if(waves_per_game) *waves_per_game = static_cast<uint32_t>(ev->wavesPerBlock);
    if(num_cus) *num_cus = ev->numCUs;
    return TA3H_OK;
}

int ta3h_device_bits_canary(int* ok) {
    if(!ok) return TA3H_E_ARG;
    *ok = 0;
    uint32_t* d = nullptr;
    if(hipMalloc(&d, sizeof(uint32_t)) != hipSuccess) return TA3H_E_HIP;
    uint32_t h = 0;
    int r = TA3H_OK;
    if(ta3h::launch_bits_canary(d, nullptr) != hipSuccess || hipMemcpy(&h, d, 4, hipMemcpyDeviceToHost) != hipSuccess)
        r = TA3H_E_HIP;
    static_cast<void>(hipFree(d));
    *ok = h == 1u ? 1 : 0;
    return r;
}

int ta3h_feature_norms(float* out, uint32_t n) {
    if(!out || n == 0) return TA3H_E_ARG;
    float* d = nullptr;
    if(hipMalloc(&d, static_cast<size_t>(n) * 6 * sizeof(float)) != hipSuccess) return TA3H_E_HIP;
    int r = TA3H_OK;
    if(ta3h::launch_feature_norms(d, n, nullptr) != hipSuccess
       || hipMemcpy(out, d, static_cast<size_t>(n) * 6 * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
        r = TA3H_E_HIP;
    static_cast<void>(hipFree(d));
    return r;
}

void ta3h_tables(int8_t* lut_ox, uint8_t* lut_off, uint8_t* tree, uint8_t* tree_count) {
    constexpr ta3h::Tables const& T = ta3h::TABLES;
    for(int i = 0; i < ta3h::LUT_TOTAL; ++i) {
        lut_ox[2 * i] = static_cast<int8_t>(T.lut[i].ox & 3);
        lut_ox[2 * i + 1] = static_cast<int8_t>((T.lut[i].ox >> 2) - 2);
    }
    for(int p = 0; p <= ta3h::NUM_PIECES; ++p) lut_off[p] = T.lutOff[p];
    for(int h = 0; h < 2; ++h)
        for(int l = 0; l < ta3h::DEPTH; ++l) {
            tree_count[h * ta3h::DEPTH + l] = T.treeCount[h][l];
            for(int n = 0; n < ta3h::TREE_NODES; ++n) {
                ta3h::TreeNode const& v = T.tree[h][l][n];
                uint8_t* o = tree + ((h * ta3h::DEPTH + l) * ta3h::TREE_NODES + n) * 6;
                o[0] = v.slot;
                o[1] = v.parent;
                o[2] = v.childBegin;
                o[3] = v.childEnd;
                o[4] = v.rootHold;
                o[5] = v.heldSlot;
            }
        }
}

} // extern "C"
