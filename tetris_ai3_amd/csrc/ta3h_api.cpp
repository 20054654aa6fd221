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

#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "kernels.hpp"
#include "tables.hpp"

namespace {

int params_of(int kind) {
    switch(kind) {
        case TA3H_MODEL_V11: return 97;
        case TA3H_MODEL_F16: return 1633;
        case TA3H_MODEL_F32W: return 20865;
        default: return -1;
    }
}

namespace wide = ta3h::f32w; // device layout of the wide network (bytes)

// floats per model in device memory (a multiple of 16 so every model starts 64-byte aligned)
uint32_t stride_of(int kind) {
    uint32_t const n = kind == TA3H_MODEL_F32W ? wide::BYTES / 4
                       : kind == TA3H_MODEL_F16 ? ta3h::f16f::FLOATS
                                                : static_cast<uint32_t>(params_of(kind));
    return (n + 15u) & ~15u;
}

// 16-32-32-1 network, host layout -> the MFMA operands after its dense device layout (ta3h::f16f)
void repack_f16_fragments(float const* src, float* dst) {
    namespace F = ta3h::f16f;
    for(int rt = 0; rt < 2; ++rt)
        for(int s = 0; s < 4; ++s)
            for(int l = 0; l < 64; ++l) {
                int const r = l % 16, k = l / 16;
                dst[F::A1 + (rt * 4 + s) * 64 + l] = src[(16 * rt + 4 * (r % 4) + r / 4) * 16 + 4 * s + k];
            }
    for(int rt2 = 0; rt2 < 2; ++rt2)
        for(int rt1 = 0; rt1 < 2; ++rt1)
            for(int s = 0; s < 4; ++s)
                for(int l = 0; l < 64; ++l) {
                    int const r = l % 16, k = l / 16;
                    dst[F::A2 + ((rt2 * 2 + rt1) * 4 + s) * 64 + l] = src[544 + (16 * rt2 + r) * 32 + 16 * rt1 + 4 * s + k];
                }
    for(int t = 0; t < 2; ++t)
        for(int l = 0; l < 64; ++l)
            for(int j = 0; j < 4; ++j) {
                int const q = l / 16;
                dst[F::C1 + (t * 64 + l) * 4 + j] = src[512 + 16 * t + 4 * j + q];
                dst[F::C2 + (t * 64 + l) * 4 + j] = src[1568 + 16 * t + 4 * q + j];
            }
}

uint16_t bf16_bits(float v) { // round to nearest even
    uint32_t u;
    std::memcpy(&u, &v, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return static_cast<uint16_t>(u >> 16);
}

// wide network, host layout W1[128][32], b1, W2[128][128], b2, W3[128], b3 -> MFMA fragment layout, lane-major
// (lane 16 q + r of a wave reads the 16 bytes [.][q][r]):
// w1f[t][q][r][e] = W1[16 t + r][8 q + e]; w2f[t][kb][q][r][e] = W2[16 t + r][unit(kb, q, e)] with
// unit = 32 kb + 4 q + e (e < 4) or 32 kb + 16 + 4 q + e - 4 (the units lane quarter q holds in layer 1's
// output tiles 2 kb, 2 kb + 1); biases and W3 as float.
void repack_f32w(float const* src, uint8_t* dst) {
    uint16_t* w1f = reinterpret_cast<uint16_t*>(dst + wide::W1F);
    for(int t = 0; t < 8; ++t)
        for(int r = 0; r < 16; ++r)
            for(int q = 0; q < 4; ++q)
                for(int e = 0; e < 8; ++e) w1f[((t * 4 + q) * 16 + r) * 8 + e] = bf16_bits(src[(16 * t + r) * 32 + 8 * q + e]);
    float const* w2 = src + 4224;
    uint16_t* w2f = reinterpret_cast<uint16_t*>(dst + wide::W2F);
    for(int t = 0; t < 8; ++t)
        for(int kb = 0; kb < 4; ++kb)
            for(int r = 0; r < 16; ++r)
                for(int q = 0; q < 4; ++q)
                    for(int e = 0; e < 8; ++e) {
                        int const unit = e < 4 ? 32 * kb + 4 * q + e : 32 * kb + 16 + 4 * q + (e - 4);
                        w2f[(((t * 4 + kb) * 4 + q) * 16 + r) * 8 + e] = bf16_bits(w2[(16 * t + r) * 128 + unit]);
                    }
    std::memcpy(dst + wide::B1, src + 4096, 128 * sizeof(float));
    std::memcpy(dst + wide::B2, src + 4224 + 16384, 128 * sizeof(float));
    std::memcpy(dst + wide::W3, src + 4224 + 16512, 129 * sizeof(float)); // W3 and b3
}

// host layout (reference dense(), utility.hpp:29-46) -> device layout of the selected network
void repack(int kind, float const* src, float* dst) {
    int const P = params_of(kind);
    if(kind == TA3H_MODEL_F32W) {
        std::memset(dst, 0, stride_of(kind) * sizeof(float));
        repack_f32w(src, reinterpret_cast<uint8_t*>(dst));
        return;
    }
    if(kind == TA3H_MODEL_F16) {
        // W1[32][16] (512), b1 (32) unchanged; W2[32 out][32 in] -> W2T[in][out]; b2, W3, b3 unchanged
        std::memcpy(dst, src, 544 * sizeof(float));
        for(int o = 0; o < 32; ++o)
            for(int i = 0; i < 32; ++i) dst[544 + i * 32 + o] = src[544 + o * 32 + i];
        std::memcpy(dst + 1568, src + 1568, 65 * sizeof(float));
        for(uint32_t i = static_cast<uint32_t>(P); i < ta3h::f16f::DENSE; ++i) dst[i] = 0.0f;
        repack_f16_fragments(src, dst);
        return;
    }
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
    uint8_t* dMoves = nullptr; // move log [items][maxMoves]
    size_t movesCap = 0;
    bool movesOn = false;
    bool movesLogged = false; // the last launch wrote the move log
    uint32_t* dQueue = nullptr;
    // two-phase schedule
    uint32_t phaseMoves = TA3H_PHASES_AUTO; // first-phase move budget of the two-phase schedule; 0 = one phase
    uint32_t* dParked = nullptr;
    size_t parkedCap = 0;
    uint32_t* dResume = nullptr; // parked list [items], its counter, resume order [items], per-model counts
    size_t resumeCap = 0;

    // pruned beam
    ta3h::PrunerDev* dPruner = nullptr;
    bool prunerOn = false;
    uint32_t pruneK = 34, pruneMask = 0;
    // full-beam traces
    uint32_t* dTrace = nullptr;
    size_t traceWords = 0; // capacity of dTrace in 32-bit words
    size_t traceCap = 0;   // records the launches keep (set_trace's cap_records; traceCap * 11 <= traceWords)
    uint32_t* dTraceCount = nullptr;
    uint64_t traceRecords = 0; // records the trace buffer holds after a launch (0 = none requested)
    uint32_t traceEvery = 0;   // 0 = off
    // pruner training scratch (grow-only)
    float* dPw = nullptr;
    size_t pwCap = 0;
    float* dAdam = nullptr; // m then v
    size_t adamCap = 0;
    uint32_t* dSamples = nullptr;
    size_t samplesCap = 0;
    uint32_t* dOrder = nullptr;
    size_t orderCap = 0;
    float* dAct = nullptr;
    size_t actCap = 0;
    float* dDelta = nullptr;
    size_t deltaCap = 0;
    float* dScore = nullptr;
    size_t scoreCap = 0;
    float* dPartial = nullptr; // gradient partial sums, PR_SPLIT x PR_PARAMS
    size_t partialCap = 0;
    float* dLoss = nullptr;
    size_t lossCap = 0;
    float* dOut = nullptr;
    size_t outCap = 0;

    // one CMA-ES generation as a captured graph: pinned weights -> device, search, fitness -> pinned
    hipGraph_t graph = nullptr;
    hipGraphExec_t graphExec = nullptr;
    float* hWeights = nullptr; // pinned, repacked layout
    size_t hWeightsCap = 0;
    float* hFitness = nullptr; // pinned
    size_t hFitnessCap = 0;
    uint32_t* hPlacements = nullptr;
    size_t hPlacementsCap = 0;
    uint64_t* hSeeds = nullptr; // pinned staging of the seeds
    size_t hSeedsCap = 0;
    // copies from the pinned staging buffers may still be in flight (no stream synchronisation since they were
    // enqueued): the next stage waits before it rewrites the buffers
    bool stagingInFlight = false;
    uint32_t graphModels = 0;

    int kind = TA3H_MODEL_V11;
    int garbage = 0;
    uint32_t beamWidth[3] = {16, 32, 32}; // beam widths per level (variation_sequences.hpp:22)
    int wavesPerBlock = 4; // waves cooperating on one game (the measured best of every network at the reference batch)
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
        // each resource is created once: a call after a partial failure picks up where that one stopped
        if(!stream && (e = hipStreamCreateWithFlags(&stream, hipStreamNonBlocking)) != hipSuccess)
            return fail("create stream", e);
        if(!evStart && (e = hipEventCreate(&evStart)) != hipSuccess) return fail("create event", e);
        if(!evStop && (e = hipEventCreate(&evStop)) != hipSuccess) return fail("create event", e);
        if(!dQueue && (e = hipMalloc(&dQueue, 64)) != hipSuccess) return fail("malloc queue", e);
        e = ta3h::upload_constants(); // once per process and device
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

    void drop_graph() {
        if(graphExec) static_cast<void>(hipGraphExecDestroy(graphExec));
        if(graph) static_cast<void>(hipGraphDestroy(graph));
        graphExec = nullptr;
        graph = nullptr;
    }

    template<class T>
    int grow_pinned(T*& ptr, size_t& cap, size_t need, char const* what) {
        if(need <= cap) return TA3H_OK;
        if(ptr != nullptr) {
            hipError_t const e = hipHostFree(ptr);
            ptr = nullptr;
            cap = 0;
            if(e != hipSuccess) return fail(what, e);
        }
        hipError_t const e = hipHostMalloc(reinterpret_cast<void**>(&ptr), need * sizeof(T), hipHostMallocDefault);
        if(e != hipSuccess) return fail(what, e);
        cap = need;
        return TA3H_OK;
    }

    ~ta3h_evaluator() {
        // best effort: a dying evaluator never reports (model_eval.cu:72-83)
        drop_graph();
        if(hWeights) static_cast<void>(hipHostFree(hWeights));
        if(hFitness) static_cast<void>(hipHostFree(hFitness));
        if(hPlacements) static_cast<void>(hipHostFree(hPlacements));
        if(hSeeds) static_cast<void>(hipHostFree(hSeeds));
        if(dWeights) static_cast<void>(hipFree(dWeights));
        if(dSeeds) static_cast<void>(hipFree(dSeeds));
        if(dFitness) static_cast<void>(hipFree(dFitness));
        if(dPlacements) static_cast<void>(hipFree(dPlacements));
        if(dCounters) static_cast<void>(hipFree(dCounters));
        if(dTimeline) static_cast<void>(hipFree(dTimeline));
        if(dMoves) static_cast<void>(hipFree(dMoves));
        if(dQueue) static_cast<void>(hipFree(dQueue));
        for(void* p : {static_cast<void*>(dParked), static_cast<void*>(dResume), static_cast<void*>(dPruner),
                       static_cast<void*>(dTrace), static_cast<void*>(dTraceCount),
                       static_cast<void*>(dPw), static_cast<void*>(dAdam), static_cast<void*>(dSamples),
                       static_cast<void*>(dOrder), static_cast<void*>(dAct), static_cast<void*>(dDelta),
                       static_cast<void*>(dScore), static_cast<void*>(dLoss), static_cast<void*>(dOut),
                       static_cast<void*>(dPartial)})
            if(p) static_cast<void>(hipFree(p));
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
    ev->drop_graph(); // captured settings and buffers change
    if(params_of(kind) < 0) return ev->fail_arg("set_model: unknown model kind " + std::to_string(kind));
    ev->kind = kind;
    ev->staged = false;
    return TA3H_OK;
}

int ta3h_set_phases(ta3h_evaluator* ev, uint32_t first_phase_moves) {
    if(!ev) return TA3H_E_ARG;
    ev->drop_graph(); // captured settings and buffers change
    ev->phaseMoves = first_phase_moves;
    return TA3H_OK;
}

int ta3h_set_garbage(ta3h_evaluator* ev, int enabled) {
    if(!ev) return TA3H_E_ARG;
    ev->drop_graph(); // captured settings and buffers change
    ev->garbage = enabled ? 1 : 0;
    return TA3H_OK;
}

int ta3h_set_beam_widths(ta3h_evaluator* ev, uint32_t k0, uint32_t k1, uint32_t k2) {
    if(!ev) return TA3H_E_ARG;
    ev->drop_graph(); // captured settings change
    if(k0 < 1 || k0 > 16 || k1 < 1 || k1 > 32 || k2 < 1 || k2 > 32)
        return ev->fail_arg("set_beam_widths: widths must be in [1, 16], [1, 32], [1, 32]");
    ev->beamWidth[0] = k0;
    ev->beamWidth[1] = k1;
    ev->beamWidth[2] = k2;
    return TA3H_OK;
}

int ta3h_set_launch(ta3h_evaluator* ev, int waves_per_game, int persistent, uint32_t grid_blocks) {
    if(!ev) return TA3H_E_ARG;
    ev->drop_graph(); // captured settings and buffers change
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
    ev->drop_graph(); // buffers may be regrown
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
    hipError_t e;
    if(ev->stagingInFlight) { // a stage without a synchronisation since: its copies may still read the buffers
        if((e = hipStreamSynchronize(ev->stream)) != hipSuccess) return ev->fail("stage", e);
        ev->stagingInFlight = false;
    }
    uint32_t const stride = stride_of(ev->kind);
    size_t const wCount = static_cast<size_t>(num_models) * stride;
    if((r = ev->grow(ev->dWeights, ev->weightsCap, wCount, "malloc weights"))) return r;
    if((r = ev->grow(ev->dSeeds, ev->seedsCap, games, "malloc seeds"))) return r;
    if((r = ev->grow(ev->dFitness, ev->fitnessCap, items, "malloc fitness"))) return r;
    if((r = ev->grow(ev->dPlacements, ev->placementsCap, items, "malloc placements"))) return r;
    if((r = ev->grow(ev->dCounters, ev->countersCap, 3 * items, "malloc counters"))) return r;
    if(ev->timelineOn && (r = ev->grow(ev->dTimeline, ev->timelineCap, 2 * items, "malloc timeline"))) return r;
    // inputs go through pinned staging buffers, so the two copies are asynchronous (model_eval.cu:123-134)
    if((r = ev->grow_pinned(ev->hWeights, ev->hWeightsCap, wCount, "pinned weights"))) return r;
    if((r = ev->grow_pinned(ev->hSeeds, ev->hSeedsCap, games, "pinned seeds"))) return r;
    for(uint32_t m = 0; m < num_models; ++m)
        repack(ev->kind, weights + static_cast<size_t>(m) * num_params, ev->hWeights + static_cast<size_t>(m) * stride);
    std::memcpy(ev->hSeeds, seeds, games * sizeof(uint64_t));
    e = hipMemcpyAsync(ev->dWeights, ev->hWeights, wCount * sizeof(float), hipMemcpyHostToDevice, ev->stream);
    if(e != hipSuccess) return ev->fail("copy weights", e);
    ev->stagingInFlight = true;
    e = hipMemcpyAsync(ev->dSeeds, ev->hSeeds, games * sizeof(uint64_t), hipMemcpyHostToDevice, ev->stream);
    if(e != hipSuccess) return ev->fail("copy seeds", e);
    ev->staged = true;
    return TA3H_OK;
}

namespace {

// First-phase move budget of the two-phase schedule for the staged batch; 0 = one phase.
uint32_t phase_budget(ta3h_evaluator const* ev) {
    uint32_t const phase = ev->phaseMoves != TA3H_PHASES_AUTO
                               ? ev->phaseMoves
                               : (ev->maxMoves / 32 < 64 ? 64u : (ev->maxMoves / 32 > 512 ? 512u : ev->maxMoves / 32));
    return phase != 0 && ev->persistent && phase < ev->maxMoves ? phase : 0u;
}

// resume list [items], its counter, per-model parked counts, then the second phase's work ring (head, tail, slots)
size_t resume_words(ta3h_evaluator const* ev) {
    return static_cast<size_t>(ev->items) + 1 + ev->items / ev->games + 2
           + ta3h::ring_slots(ev->items, ev->maxMoves, phase_budget(ev));
}

// Kernel mode (0 full beam, 1 traces or move log, 2 pruned), waves per game and workgroups of the staged batch's
// launch: one workgroup (of wpb cooperating waves) per game in flight.
int launch_shape(ta3h_evaluator* ev, int* mode, int* wpb, uint32_t* blocks) {
    *mode = ev->prunerOn ? 2 : (ev->traceEvery != 0 || ev->movesOn ? 1 : 0);
    *wpb = *mode != 0 ? 4 : ev->wavesPerBlock; // the pruned and trace kernels run four waves per game
    uint32_t const needed = ev->items;
    uint32_t b = needed;
    if(ev->persistent) {
        b = ev->gridBlocks;
        if(b == 0) {
            // fill every CU up to the kernel's occupancy limit; workgroups beyond it would only queue
            int perCU = 0;
            hipError_t const e0 = ta3h::max_active_blocks(ev->kind, *wpb, *mode, &perCU);
            if(e0 != hipSuccess) return ev->fail("occupancy", e0);
            b = ev->numCUs * static_cast<uint32_t>(perCU > 0 ? perCU : 1);
        }
        if(b > needed) b = needed;
    }
    *blocks = b;
    return TA3H_OK;
}

// Device buffers the launch of the staged batch needs beyond the staged ones (the parked games of the two-phase
// schedule). Called before enqueue_eval and never while the stream is being captured.
int prepare_launch(ta3h_evaluator* ev) {
    int r;
    if(ev->movesOn
       && (r = ev->grow(ev->dMoves, ev->movesCap, static_cast<size_t>(ev->items) * ev->maxMoves, "malloc move log")))
        return r;
    if(phase_budget(ev) == 0) return TA3H_OK;
    if(ta3h::ring_slots(ev->items, ev->maxMoves, phase_budget(ev)) > 0xFFFFFFF0ull) {
        ev->err = "launch: too many second-phase turns (raise the first-phase budget)";
        return TA3H_E_ARG;
    }
    if((r = ev->grow(ev->dParked, ev->parkedCap, static_cast<size_t>(ev->items) * ta3h::park_words(ev->kind), "malloc parked")))
        return r;
    return ev->grow(ev->dResume, ev->resumeCap, resume_words(ev), "malloc resume list");
}

// Enqueues the search of the staged batch on the evaluator's stream (optionally bracketed by the timing events);
// allocates nothing, so it may run inside a stream capture.
int enqueue_eval(ta3h_evaluator* ev, bool events) {
    ta3h::EvalArgs a{};
    a.weights = ev->dWeights;
    a.seeds = ev->dSeeds;
    a.fitness = ev->dFitness;
    a.placements = ev->dPlacements;
    a.counters = ev->dCounters;
    a.timeline = ev->timelineOn ? ev->dTimeline : nullptr;
    if(ev->movesOn) {
        if(ev->movesCap < static_cast<size_t>(ev->items) * ev->maxMoves) {
            ev->err = "launch: move log not prepared";
            return TA3H_E_STATE;
        }
        a.moves = ev->dMoves;
    }
    ev->movesLogged = ev->movesOn;
    a.queue = ev->dQueue;
    a.items = ev->items;
    a.games = ev->games;
    a.maxMoves = ev->maxMoves;
    a.weightStride = stride_of(ev->kind);
    for(int l = 0; l < 3; ++l) a.beamWidth[l] = ev->beamWidth[l];
    a.persistent = static_cast<uint32_t>(ev->persistent);
    a.garbage = static_cast<uint32_t>(ev->garbage);
    uint32_t const phase = phase_budget(ev);
    if(phase != 0) {
        if(ev->parkedCap < static_cast<size_t>(ev->items) * ta3h::park_words(ev->kind) || ev->resumeCap < resume_words(ev)) {
            ev->err = "launch: two-phase buffers not prepared";
            return TA3H_E_STATE;
        }
        a.phaseMoves = phase;
        a.parked = ev->dParked;
        a.resumeList = ev->dResume;
        a.resumeCount = ev->dResume + ev->items;
        a.modelParked = ev->dResume + ev->items + 1;
        a.ring = a.modelParked + ev->items / ev->games;
        a.ringCap = static_cast<uint32_t>(ta3h::ring_slots(ev->items, ev->maxMoves, phase));
    }
    int mode = 0, wpb = 0;
    uint32_t blocks = 0;
    int const r = launch_shape(ev, &mode, &wpb, &blocks);
    if(r) return r;
    if(ev->prunerOn) {
        a.pruner = ev->dPruner;
        a.pruneK = ev->pruneK;
        a.pruneMask = ev->pruneMask;
    } else if(ev->traceEvery != 0) {
        a.trace = ev->dTrace;
        a.traceCount = ev->dTraceCount;
        a.traceCap = static_cast<uint32_t>(ev->traceCap);
        a.traceEvery = ev->traceEvery;
    } // (a move log alone: the trace kernel, recording no traces, writes the move log of a full beam)
    ev->lastGrid = blocks;
    hipError_t e;
    if(events && (e = hipEventRecord(ev->evStart, ev->stream)) != hipSuccess) return ev->fail("event record", e);
    e = ta3h::launch_beam_eval(ev->kind, a, wpb, blocks, ev->stream);
    if(e != hipSuccess) return ev->fail("kernel launch", e);
    if(events && (e = hipEventRecord(ev->evStop, ev->stream)) != hipSuccess) return ev->fail("event record", e);
    return TA3H_OK;
}

} // namespace

int ta3h_launch(ta3h_evaluator* ev) {
    if(!ev) return TA3H_E_ARG;
    if(!ev->staged) {
        ev->err = "launch: nothing staged";
        return TA3H_E_STATE;
    }
    if(ev->items == 0) {
        ev->launched = true;
        return TA3H_OK;
    }
    int r = prepare_launch(ev);
    if(r == TA3H_OK) r = enqueue_eval(ev, true);
    ev->launched = r == TA3H_OK; // a failed launch leaves nothing to fetch
    return r;
}

int ta3h_sync(ta3h_evaluator* ev) {
    if(!ev) return TA3H_E_ARG;
    if(!ev->ready) return TA3H_OK;
    hipError_t const e = hipStreamSynchronize(ev->stream);
    if(e != hipSuccess) return ev->fail("kernel run", e);
    ev->stagingInFlight = false;
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
    ev->stagingInFlight = false;
    return TA3H_OK;
}

int ta3h_fetch_counters(ta3h_evaluator* ev, uint64_t* out_counters) {
    if(!ev || !out_counters) return TA3H_E_ARG;
    if(!ev->launched) {
        ev->err = "fetch: nothing launched";
        return TA3H_E_STATE;
    }
    if(ev->items == 0) return TA3H_OK;
    hipError_t e = hipMemcpyAsync(out_counters, ev->dCounters, 3 * static_cast<size_t>(ev->items) * sizeof(uint64_t),
                                  hipMemcpyDeviceToHost, ev->stream);
    if(e != hipSuccess) return ev->fail("copy counters", e);
    e = hipStreamSynchronize(ev->stream);
    if(e != hipSuccess) return ev->fail("kernel run", e);
    return TA3H_OK;
}

int ta3h_set_timeline(ta3h_evaluator* ev, int enabled) {
    if(!ev) return TA3H_E_ARG;
    ev->drop_graph(); // captured settings and buffers change
    ev->timelineOn = enabled != 0;
    ev->staged = false;
    return TA3H_OK;
}

int ta3h_set_move_log(ta3h_evaluator* ev, int enabled) {
    if(!ev) return TA3H_E_ARG;
    ev->drop_graph(); // captured settings and buffers change
    ev->movesOn = enabled != 0;
    return TA3H_OK;
}

int ta3h_fetch_moves(ta3h_evaluator* ev, uint8_t* out) {
    if(!ev || !out) return TA3H_E_ARG;
    if(!ev->launched || !ev->movesLogged) {
        ev->err = "fetch_moves: no launch with the move log enabled";
        return TA3H_E_STATE;
    }
    size_t const bytes = static_cast<size_t>(ev->items) * ev->maxMoves;
    if(bytes == 0) return TA3H_OK;
    hipError_t e = hipMemcpyAsync(out, ev->dMoves, bytes, hipMemcpyDeviceToHost, ev->stream);
    if(e != hipSuccess) return ev->fail("copy move log", e);
    e = hipStreamSynchronize(ev->stream);
    if(e != hipSuccess) return ev->fail("kernel run", e);
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

// As the reference's eval (model_eval.cu:97-153): two asynchronous input copies, the search, one result copy,
// then exactly one host synchronisation.
int ta3h_eval(ta3h_evaluator* ev, const float* weights, uint32_t num_models, uint32_t num_params,
              const uint64_t* seeds, uint32_t games, uint32_t max_moves, float* out_fitness) {
    int r = ta3h_stage(ev, weights, num_models, num_params, seeds, games, max_moves);
    if(r) return r;
    if((r = ta3h_launch(ev))) return r;
    if(ev->items == 0) return TA3H_OK;
    if(!out_fitness) return ev->fail_arg("eval: null output");
    if((r = ev->grow_pinned(ev->hFitness, ev->hFitnessCap, ev->items, "pinned fitness"))) return r;
    hipError_t e = hipMemcpyAsync(ev->hFitness, ev->dFitness, ev->items * sizeof(float), hipMemcpyDeviceToHost,
                                  ev->stream);
    if(e != hipSuccess) return ev->fail("copy fitness", e);
    e = hipStreamSynchronize(ev->stream); // the one host synchronisation of the call
    if(e != hipSuccess) return ev->fail("kernel run", e);
    ev->stagingInFlight = false;
    std::memcpy(out_fitness, ev->hFitness, ev->items * sizeof(float));
    return TA3H_OK;
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
    if(waves_per_game)
        *waves_per_game = static_cast<uint32_t>(ev->prunerOn || ev->traceEvery != 0 || ev->movesOn ? 4 : ev->wavesPerBlock);
    if(num_cus) *num_cus = ev->numCUs;
    return TA3H_OK;
}

int ta3h_model_values(int kind, const float* weights, const uint32_t* boards, const uint32_t* clears,
                      const uint8_t* held, uint32_t n, float* out) {
    if(params_of(kind) < 0 || !weights || !boards || !clears || !held || !out) return TA3H_E_ARG;
    if(n == 0) return TA3H_OK;
    std::vector<float> packed(stride_of(kind));
    repack(kind, weights, packed.data());
    size_t const bytes = packed.size() * 4 + static_cast<size_t>(n) * (40 + 4 + 1 + 4) + 64;
    uint8_t* d = nullptr;
    if(hipMalloc(&d, bytes) != hipSuccess) return TA3H_E_HIP;
    float* dw = reinterpret_cast<float*>(d);
    uint32_t* db = reinterpret_cast<uint32_t*>(d + packed.size() * 4);
    uint32_t* dc = db + 10ull * n;
    float* dout = reinterpret_cast<float*>(dc + n);
    uint8_t* dh = reinterpret_cast<uint8_t*>(dout + n);
    int r = TA3H_OK;
    if(hipMemcpy(dw, packed.data(), packed.size() * 4, hipMemcpyHostToDevice) != hipSuccess
       || hipMemcpy(db, boards, 40ull * n, hipMemcpyHostToDevice) != hipSuccess
       || hipMemcpy(dc, clears, 4ull * n, hipMemcpyHostToDevice) != hipSuccess
       || hipMemcpy(dh, held, n, hipMemcpyHostToDevice) != hipSuccess
       || ta3h::upload_constants() != hipSuccess
       || ta3h::launch_model_values(kind, dw, db, dc, dh, n, dout, nullptr) != hipSuccess
       || hipMemcpy(out, dout, 4ull * n, hipMemcpyDeviceToHost) != hipSuccess)
        r = TA3H_E_HIP;
    static_cast<void>(hipFree(d));
    return r;
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

int ta3h_graph_prepare(ta3h_evaluator* ev, uint32_t num_models, uint32_t num_params, const uint64_t* seeds,
                       uint32_t games, uint32_t max_moves) {
    if(!ev) return TA3H_E_ARG;
    ev->drop_graph();
    ev->graphModels = 0;
    uint64_t const items = static_cast<uint64_t>(num_models) * games;
    if(items == 0) return ev->fail_arg("graph_prepare: no (model, game) items");
    if(items > 0xFFFFFFFFull / 2) return ev->fail_arg("graph_prepare: num_models * games too large");
    int const P = params_of(ev->kind);
    if(static_cast<int64_t>(num_params) != P)
        return ev->fail_arg("graph_prepare: weights per model must equal NUM_PARAMS (" + std::to_string(P) + ")");
    if(seeds == nullptr) return ev->fail_arg("graph_prepare: null seeds");
    int r = ev->ensure_ready();
    if(r) return r;
    uint32_t const stride = stride_of(ev->kind);
    size_t const wCount = static_cast<size_t>(num_models) * stride;
    if((r = ev->grow(ev->dWeights, ev->weightsCap, wCount, "malloc weights"))) return r;
    if((r = ev->grow(ev->dSeeds, ev->seedsCap, games, "malloc seeds"))) return r;
    if((r = ev->grow(ev->dFitness, ev->fitnessCap, items, "malloc fitness"))) return r;
    if((r = ev->grow(ev->dPlacements, ev->placementsCap, items, "malloc placements"))) return r;
    if((r = ev->grow(ev->dCounters, ev->countersCap, 3 * items, "malloc counters"))) return r;
    if(ev->timelineOn && (r = ev->grow(ev->dTimeline, ev->timelineCap, 2 * items, "malloc timeline"))) return r;
    hipError_t e;
    if(ev->stagingInFlight) { // copies of an earlier stage may still read the pinned weights
        if((e = hipStreamSynchronize(ev->stream)) != hipSuccess) return ev->fail("stage", e);
        ev->stagingInFlight = false;
    }
    if((r = ev->grow_pinned(ev->hWeights, ev->hWeightsCap, wCount, "pinned weights"))) return r;
    if((r = ev->grow_pinned(ev->hFitness, ev->hFitnessCap, items, "pinned fitness"))) return r;
    if((r = ev->grow_pinned(ev->hPlacements, ev->hPlacementsCap, items, "pinned placements"))) return r;
    std::memset(ev->hWeights, 0, wCount * sizeof(float)); // padding floats stay zero
    e = hipMemcpyAsync(ev->dSeeds, seeds, games * sizeof(uint64_t), hipMemcpyHostToDevice, ev->stream);
    if(e == hipSuccess) e = hipStreamSynchronize(ev->stream);
    if(e != hipSuccess) return ev->fail("copy seeds", e);
    ev->items = static_cast<uint32_t>(items);
    ev->games = games;
    ev->maxMoves = max_moves;
    ev->staged = true;
    ev->launched = false;
    if((r = prepare_launch(ev))) return r; // no allocation inside the capture
    // capture: weights in, search, fitness and placements out
    e = hipStreamBeginCapture(ev->stream, hipStreamCaptureModeThreadLocal);
    if(e != hipSuccess) return ev->fail("begin capture", e);
    hipError_t ec = hipMemcpyAsync(ev->dWeights, ev->hWeights, wCount * sizeof(float), hipMemcpyHostToDevice, ev->stream);
    int rc = TA3H_OK;
    if(ec == hipSuccess) rc = enqueue_eval(ev, false);
    if(ec == hipSuccess && rc == TA3H_OK)
        ec = hipMemcpyAsync(ev->hFitness, ev->dFitness, items * sizeof(float), hipMemcpyDeviceToHost, ev->stream);
    if(ec == hipSuccess && rc == TA3H_OK)
        ec = hipMemcpyAsync(ev->hPlacements, ev->dPlacements, items * sizeof(uint32_t), hipMemcpyDeviceToHost, ev->stream);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(ev->stream, &g);
    if(rc != TA3H_OK) {
        if(g) static_cast<void>(hipGraphDestroy(g));
        return rc;
    }
    if(ec != hipSuccess || e != hipSuccess) {
        if(g) static_cast<void>(hipGraphDestroy(g));
        return ev->fail("capture generation", ec != hipSuccess ? ec : e);
    }
    ev->graph = g;
    e = hipGraphInstantiate(&ev->graphExec, g, nullptr, nullptr, 0);
    if(e != hipSuccess) {
        ev->drop_graph();
        return ev->fail("instantiate generation graph", e);
    }
    ev->graphModels = num_models;
    return TA3H_OK;
}

int ta3h_graph_run(ta3h_evaluator* ev, const float* weights, float* out_fitness, uint32_t* out_placements) {
    if(!ev) return TA3H_E_ARG;
    if(ev->graphExec == nullptr) {
        ev->err = "graph_run: no generation graph prepared";
        return TA3H_E_STATE;
    }
    if(weights == nullptr) return ev->fail_arg("graph_run: null weights");
    int const P = params_of(ev->kind);
    uint32_t const stride = stride_of(ev->kind);
    for(uint32_t m = 0; m < ev->graphModels; ++m)
        repack(ev->kind, weights + static_cast<size_t>(m) * P, ev->hWeights + static_cast<size_t>(m) * stride);
    hipError_t e = hipEventRecord(ev->evStart, ev->stream);
    if(e == hipSuccess) e = hipGraphLaunch(ev->graphExec, ev->stream);
    if(e == hipSuccess) e = hipEventRecord(ev->evStop, ev->stream);
    if(e != hipSuccess) return ev->fail("graph launch", e);
    e = hipStreamSynchronize(ev->stream); // the one host synchronisation of the generation
    if(e != hipSuccess) return ev->fail("graph run", e);
    ev->launched = true;
    if(out_fitness) std::memcpy(out_fitness, ev->hFitness, ev->items * sizeof(float));
    if(out_placements) std::memcpy(out_placements, ev->hPlacements, ev->items * sizeof(uint32_t));
    return TA3H_OK;
}

int ta3h_pruner_num_params(void) { return ta3h::PR_PARAMS; }

int ta3h_pruner_quant_info(const float* w, float act_range, int* info) {
    if(!w || !info || !(act_range > 0.0f) || !std::isfinite(act_range)) return TA3H_E_ARG;
    auto t = std::make_unique<ta3h::PrunerDev>();
    ta3h::build_pruner_tables(w, act_range, *t, info);
    return TA3H_OK;
}

int ta3h_set_pruner(ta3h_evaluator* ev, const float* w, float act_range, uint32_t k, uint32_t ply_mask) {
    if(!ev) return TA3H_E_ARG;
    ev->drop_graph(); // captured settings and buffers change
    if(w == nullptr) {
        ev->prunerOn = false;
        return TA3H_OK;
    }
    if(!(act_range > 0.0f) || !std::isfinite(act_range)) return ev->fail_arg("set_pruner: act_range must be > 0");
    if(k < 1 || k > 34) return ev->fail_arg("set_pruner: k must be in [1, 34]");
    if((ply_mask & ~0xEu) != 0) return ev->fail_arg("set_pruner: ply_mask may hold bits 1, 2 and 3 only");
    for(int i = 0; i < ta3h::PR_PARAMS; ++i)
        if(!std::isfinite(w[i])) return ev->fail_arg("set_pruner: non-finite weight");
    int r = ev->ensure_ready();
    if(r) return r;
    if(ev->dPruner == nullptr) {
        hipError_t const e = hipMalloc(&ev->dPruner, sizeof(ta3h::PrunerDev));
        if(e != hipSuccess) return ev->fail("malloc pruner", e);
    }
    auto t = std::make_unique<ta3h::PrunerDev>();
    ta3h::build_pruner_tables(w, act_range, *t, nullptr);
    hipError_t e = hipMemcpyAsync(ev->dPruner, t.get(), sizeof(ta3h::PrunerDev), hipMemcpyHostToDevice, ev->stream);
    if(e != hipSuccess) return ev->fail("copy pruner", e);
    e = hipStreamSynchronize(ev->stream);
    if(e != hipSuccess) return ev->fail("copy pruner", e);
    ev->prunerOn = true;
    ev->pruneK = k;
    ev->pruneMask = ply_mask;
    ev->traceEvery = 0; // a pruned search records no traces
    return TA3H_OK;
}

int ta3h_set_trace(ta3h_evaluator* ev, uint64_t cap_records, uint32_t every) {
    if(!ev) return TA3H_E_ARG;
    ev->drop_graph(); // captured settings and buffers change
    if(cap_records == 0 || every == 0) {
        ev->traceEvery = 0;
        return TA3H_OK;
    }
    if(cap_records > 0xFFFFFFFFull) return ev->fail_arg("set_trace: at most 2^32 - 1 records");
    int r = ev->ensure_ready();
    if(r) return r;
    if((r = ev->grow(ev->dTrace, ev->traceWords, static_cast<size_t>(cap_records) * 11, "malloc trace"))) return r;
    ev->traceCap = static_cast<size_t>(cap_records); // records (the buffer may be larger)
    if(ev->dTraceCount == nullptr) {
        hipError_t const e = hipMalloc(&ev->dTraceCount, 64);
        if(e != hipSuccess) return ev->fail("malloc trace count", e);
    }
    ev->traceEvery = every;
    ev->prunerOn = false; // traces come from the full beam
    return TA3H_OK;
}

int ta3h_fetch_trace(ta3h_evaluator* ev, uint32_t* out, uint64_t cap_records, uint64_t* produced, uint64_t* copied) {
    if(!ev) return TA3H_E_ARG;
    if(!ev->launched || ev->traceEvery == 0) {
        ev->err = "fetch_trace: no launch with traces enabled";
        return TA3H_E_STATE;
    }
    uint32_t count = 0;
    if(ev->items != 0) {
        hipError_t e = hipMemcpyAsync(&count, ev->dTraceCount, 4, hipMemcpyDeviceToHost, ev->stream);
        if(e != hipSuccess) return ev->fail("copy trace count", e);
        e = hipStreamSynchronize(ev->stream);
        if(e != hipSuccess) return ev->fail("kernel run", e);
    }
    if(produced) *produced = count;
    uint64_t n = count;
    if(n > ev->traceCap) n = ev->traceCap;
    if(n > cap_records) n = cap_records;
    if(out == nullptr) n = 0;
    if(copied) *copied = n;
    if(n) {
        hipError_t e = hipMemcpyAsync(out, ev->dTrace, n * 11 * sizeof(uint32_t), hipMemcpyDeviceToHost, ev->stream);
        if(e != hipSuccess) return ev->fail("copy trace", e);
        e = hipStreamSynchronize(ev->stream);
        if(e != hipSuccess) return ev->fail("copy trace", e);
    }
    return TA3H_OK;
}

namespace {

// device copy of the master weights and of caller samples (null samples = the evaluator's trace buffer)
int stage_training(ta3h_evaluator* ev, const float* w, const uint32_t* samples, uint64_t n, uint32_t const** dSmp) {
    int r = ev->ensure_ready();
    if(r) return r;
    if((r = ev->grow(ev->dPw, ev->pwCap, ta3h::PR_PARAMS, "malloc pruner weights"))) return r;
    hipError_t e = hipMemcpyAsync(ev->dPw, w, ta3h::PR_PARAMS * sizeof(float), hipMemcpyHostToDevice, ev->stream);
    if(e != hipSuccess) return ev->fail("copy pruner weights", e);
    if(samples == nullptr) {
        if(ev->dTrace == nullptr || n > ev->traceCap) return ev->fail_arg("pruner: trace buffer holds fewer records");
        *dSmp = ev->dTrace;
    } else {
        if((r = ev->grow(ev->dSamples, ev->samplesCap, static_cast<size_t>(n) * 11, "malloc samples"))) return r;
        e = hipMemcpyAsync(ev->dSamples, samples, static_cast<size_t>(n) * 11 * sizeof(uint32_t), hipMemcpyHostToDevice,
                           ev->stream);
        if(e != hipSuccess) return ev->fail("copy samples", e);
        *dSmp = ev->dSamples;
    }
    return TA3H_OK;
}

int grow_batch(ta3h_evaluator* ev, size_t b) {
    int r;
    if((r = ev->grow(ev->dAct, ev->actCap, b * 4 * ta3h::PR_H, "malloc activations"))) return r;
    if((r = ev->grow(ev->dDelta, ev->deltaCap, b * 2 * ta3h::PR_H, "malloc deltas"))) return r;
    if((r = ev->grow(ev->dScore, ev->scoreCap, b, "malloc dscore"))) return r;
    if((r = ev->grow(ev->dPartial, ev->partialCap, static_cast<size_t>(ta3h::PR_SPLIT) * ta3h::PR_PARAMS,
                     "malloc gradient partials")))
        return r;
    return TA3H_OK;
}

} // namespace

int ta3h_pruner_grad(ta3h_evaluator* ev, const float* w, const uint32_t* samples, uint32_t n, float pos_weight,
                     float* grad, float* loss) {
    if(!ev || !w || !grad || n == 0) return TA3H_E_ARG;
    uint32_t const* dSmp = nullptr;
    int r = stage_training(ev, w, samples, n, &dSmp);
    if(r) return r;
    if((r = grow_batch(ev, n))) return r;
    if((r = ev->grow(ev->dLoss, ev->lossCap, n, "malloc loss"))) return r;
    if((r = ev->grow(ev->dOut, ev->outCap, ta3h::PR_PARAMS, "malloc gradient"))) return r;
    hipError_t e = ta3h::pruner_fwd_bwd(ev->dPw, dSmp, nullptr, n, 1.0f / static_cast<float>(n), pos_weight, ev->dAct,
                                        ev->dDelta, ev->dScore, ev->dLoss, ev->stream);
    if(e != hipSuccess) return ev->fail("pruner forward/backward", e);
    e = ta3h::pruner_grad_adam(ev->dPw, nullptr, nullptr, ev->dOut, ev->dPartial, dSmp, nullptr, n, ev->dAct,
                               ev->dDelta, ev->dScore, 0, 0.0f, 0.0f, 0.0f, 0.0f, 1.0f, 1.0f, ev->stream);
    if(e != hipSuccess) return ev->fail("pruner gradient", e);
    e = hipMemcpyAsync(grad, ev->dOut, ta3h::PR_PARAMS * sizeof(float), hipMemcpyDeviceToHost, ev->stream);
    if(e != hipSuccess) return ev->fail("copy gradient", e);
    std::vector<float> l(n);
    e = hipMemcpyAsync(l.data(), ev->dLoss, n * sizeof(float), hipMemcpyDeviceToHost, ev->stream);
    if(e != hipSuccess) return ev->fail("copy loss", e);
    e = hipStreamSynchronize(ev->stream);
    if(e != hipSuccess) return ev->fail("pruner gradient", e);
    if(loss) {
        double s = 0.0;
        for(float v : l) s += v;
        *loss = static_cast<float>(s);
    }
    return TA3H_OK;
}

int ta3h_pruner_train(ta3h_evaluator* ev, float* w, const uint32_t* samples, uint64_t n, uint32_t epochs,
                      uint32_t batch, float lr, float pos_weight, uint64_t seed, float* epoch_loss) {
    if(!ev || !w || n == 0 || batch == 0 || !(lr > 0.0f)) return TA3H_E_ARG;
    if(n > 0xFFFFFFFFull) return ev->fail_arg("pruner_train: at most 2^32 - 1 samples");
    uint32_t const* dSmp = nullptr;
    int r = stage_training(ev, w, samples, n, &dSmp);
    if(r) return r;
    size_t const b = batch < n ? batch : static_cast<size_t>(n);
    if((r = grow_batch(ev, b))) return r;
    if((r = ev->grow(ev->dLoss, ev->lossCap, static_cast<size_t>(n), "malloc loss"))) return r;
    if((r = ev->grow(ev->dOrder, ev->orderCap, static_cast<size_t>(n), "malloc order"))) return r;
    if((r = ev->grow(ev->dAdam, ev->adamCap, 2 * static_cast<size_t>(ta3h::PR_PARAMS), "malloc adam"))) return r;
    hipError_t e = hipMemsetAsync(ev->dAdam, 0, 2 * ta3h::PR_PARAMS * sizeof(float), ev->stream);
    if(e != hipSuccess) return ev->fail("clear adam", e);
    float* const m = ev->dAdam;
    float* const v = ev->dAdam + ta3h::PR_PARAMS;
    float const beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
    std::vector<uint32_t> order(static_cast<size_t>(n));
    std::vector<float> losses(static_cast<size_t>(n));
    uint64_t st = seed;
    auto next = [&st]() { // splitmix64
        uint64_t z = (st += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    };
    uint64_t step = 0;
    for(uint32_t ep = 0; ep < epochs; ++ep) {
        for(size_t i = 0; i < order.size(); ++i) order[i] = static_cast<uint32_t>(i);
        for(size_t i = order.size() - 1; i > 0; --i) std::swap(order[i], order[next() % (i + 1)]); // Fisher-Yates
        e = hipMemcpyAsync(ev->dOrder, order.data(), order.size() * sizeof(uint32_t), hipMemcpyHostToDevice, ev->stream);
        if(e != hipSuccess) return ev->fail("copy order", e);
        for(uint64_t b0 = 0; b0 < n; b0 += b) {
            uint32_t const nb = static_cast<uint32_t>(n - b0 < b ? n - b0 : b);
            ++step;
            e = ta3h::pruner_fwd_bwd(ev->dPw, dSmp, ev->dOrder + b0, nb, 1.0f / static_cast<float>(nb), pos_weight,
                                     ev->dAct, ev->dDelta, ev->dScore, ev->dLoss + b0, ev->stream);
            if(e != hipSuccess) return ev->fail("pruner forward/backward", e);
            float const c1 = 1.0f - static_cast<float>(std::pow(0.9, static_cast<double>(step)));
            float const c2 = 1.0f - static_cast<float>(std::pow(0.999, static_cast<double>(step)));
            e = ta3h::pruner_grad_adam(ev->dPw, m, v, nullptr, ev->dPartial, dSmp, ev->dOrder + b0, nb, ev->dAct,
                                       ev->dDelta, ev->dScore, 1, lr, beta1, beta2, eps, c1, c2, ev->stream);
            if(e != hipSuccess) return ev->fail("pruner adam", e);
        }
        // the next epoch's order upload must not overtake this epoch's kernels
        e = hipMemcpyAsync(losses.data(), ev->dLoss, losses.size() * sizeof(float), hipMemcpyDeviceToHost, ev->stream);
        if(e != hipSuccess) return ev->fail("copy loss", e);
        e = hipStreamSynchronize(ev->stream);
        if(e != hipSuccess) return ev->fail("pruner training", e);
        if(epoch_loss) { // mean over samples: each batch stored loss / batch size
            double s = 0.0;
            for(uint64_t b0 = 0; b0 < n; b0 += b) {
                uint64_t const nb = n - b0 < b ? n - b0 : b;
                double bs = 0.0;
                for(uint64_t i = b0; i < b0 + nb; ++i) bs += losses[i];
                s += bs * static_cast<double>(nb);
            }
            epoch_loss[ep] = static_cast<float>(s / static_cast<double>(n));
        }
    }
    e = hipMemcpyAsync(w, ev->dPw, ta3h::PR_PARAMS * sizeof(float), hipMemcpyDeviceToHost, ev->stream);
    if(e != hipSuccess) return ev->fail("copy pruner weights", e);
    e = hipStreamSynchronize(ev->stream);
    if(e != hipSuccess) return ev->fail("pruner training", e);
    return TA3H_OK;
}

int ta3h_pruner_hidden(ta3h_evaluator* ev, const float* w, const uint32_t* samples, uint32_t n, float* out) {
    if(!ev || !w || !out || n == 0) return TA3H_E_ARG;
    uint32_t const* dSmp = nullptr;
    int r = stage_training(ev, w, samples, n, &dSmp);
    if(r) return r;
    if((r = ev->grow(ev->dOut, ev->outCap, static_cast<size_t>(n) * ta3h::PR_H, "malloc hidden"))) return r;
    hipError_t e = ta3h::pruner_hidden_samples(ev->dPw, dSmp, n, ev->dOut, ev->stream);
    if(e != hipSuccess) return ev->fail("pruner hidden", e);
    e = hipMemcpyAsync(out, ev->dOut, static_cast<size_t>(n) * ta3h::PR_H * sizeof(float), hipMemcpyDeviceToHost,
                       ev->stream);
    if(e != hipSuccess) return ev->fail("copy hidden", e);
    e = hipStreamSynchronize(ev->stream);
    if(e != hipSuccess) return ev->fail("pruner hidden", e);
    return TA3H_OK;
}

int ta3h_pruner_logits(ta3h_evaluator* ev, const float* w, float act_range, const uint32_t* samples, uint32_t n,
                       float* out) {
    if(!ev || !w || !out || n == 0) return TA3H_E_ARG;
    if(!(act_range > 0.0f) || !std::isfinite(act_range)) return ev->fail_arg("pruner_logits: act_range must be > 0");
    uint32_t const* dSmp = nullptr;
    int r = stage_training(ev, w, samples, n, &dSmp);
    if(r) return r;
    if((r = ev->grow(ev->dOut, ev->outCap, n, "malloc logits"))) return r;
    // a private table copy: the evaluator's pruner (if any) stays as set
    ta3h::PrunerDev* dTab = nullptr;
    hipError_t e = hipMalloc(&dTab, sizeof(ta3h::PrunerDev));
    if(e != hipSuccess) return ev->fail("malloc pruner", e);
    auto t = std::make_unique<ta3h::PrunerDev>();
    ta3h::build_pruner_tables(w, act_range, *t, nullptr);
    e = hipMemcpyAsync(dTab, t.get(), sizeof(ta3h::PrunerDev), hipMemcpyHostToDevice, ev->stream);
    if(e == hipSuccess) e = ta3h::pruner_infer_samples(dTab, dSmp, n, ev->dOut, ev->stream);
    if(e == hipSuccess) e = hipMemcpyAsync(out, ev->dOut, n * sizeof(float), hipMemcpyDeviceToHost, ev->stream);
    hipError_t const es = hipStreamSynchronize(ev->stream);
    static_cast<void>(hipFree(dTab));
    if(e != hipSuccess) return ev->fail("pruner logits", e);
    if(es != hipSuccess) return ev->fail("pruner logits", es);
    return TA3H_OK;
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
