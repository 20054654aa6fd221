// Launch interface between the host driver and the evaluator kernels (no device code here).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "pruner.hpp"

namespace ta3h {

enum ModelKind : int {
    MODEL_V11 = 0, // reference default value network: 10 features -> 8 ReLU -> 1 (97 params)
    MODEL_F16 = 1, // 16 features -> 32 ReLU -> 32 ReLU -> 1 (1633 params)
    MODEL_F32W = 3, // 32 features -> 128 ReLU -> 128 ReLU -> 1 on BF16 MFMA (20865 params)
};

// Device weight layout of the wide network, per model (bytes): w1f [8 t][4 q][16 r] x 8 bf16 = MFMA A fragment
// (rows 16 t + r, features 8 q .. + 7) of layer 1; w2f [8 t][4 kb][4 q][16 r] x 8 bf16 = A fragment of layer 2
// for output rows 16 t + r, K block kb, whose element e is hidden unit 32 kb + 4 q + e (e < 4) or
// 32 kb + 16 + 4 q + (e - 4) -- the units lane quarter q holds in layer 1's output tiles 2 kb and 2 kb + 1, so
// layer 1's outputs are layer 2's B operands in place; then b1[128], b2[128], W3[128], b3 as float.
namespace f32w {
inline constexpr uint32_t W1F = 0, W2F = 8192, B1 = 40960, B2 = 41472, W3 = 41984, B3 = 42496, BYTES = 42512;
}

// Device layout of the 16-32-32-1 network, per model (floats): the dense part (W1[32][16], b1, W2T[32 in][32 out],
// b2, W3, b3 at 0 / 512 / 544 / 1568 / 1600 / 1632, padded to 1,648), then the operands of its two hidden layers on
// v_mfma_f32_16x16x4_f32, lane-major (lane 16 k + r of a wave reads word [lane]): with row r of a 16-unit layer-1
// tile standing for unit 4 (r % 4) + r / 4, the layer-1 outputs land in the layout layer 2 takes as its B operand.
//   a1[rt][s][lane r + 16 k] = W1[16 rt + 4 (r % 4) + r / 4][4 s + k]
//   a2[rt2][rt1][s][lane r + 16 k] = W2[16 rt2 + r][16 rt1 + 4 s + k]
//   c1[rt][lane n + 16 q][j] = b1[16 rt + 4 j + q]; c2[rt2][lane n + 16 q][i] = b2[16 rt2 + 4 q + i]
namespace f16f {
inline constexpr uint32_t DENSE = 1648, A1 = DENSE, A2 = A1 + 512, C1 = A2 + 1024, C2 = C1 + 512, FLOATS = C2 + 512;
}

inline constexpr uint32_t PARK_WORDS = 128; // parked game: engine state, counters, move index, root scores (512 B)
// the 16-32-32-1 network's plain kernels also carry the previous move's level-2 and leaf scores, list infos,
// legality and kept slots (the first words of DeepLds in the kernel source)
inline constexpr uint32_t DEEP_PARK_WORDS = 4520;
inline constexpr uint32_t park_words(int kind) { return kind == MODEL_F16 ? PARK_WORDS + DEEP_PARK_WORDS : PARK_WORDS; }

// Work-ring slots of a two-phase launch: every parked game enters once after the first phase and once after
// each further turn of `phase` moves it ends unfinished, at most (maxMoves - phase) / phase times.
inline constexpr uint64_t ring_slots(uint64_t items, uint32_t maxMoves, uint32_t phase) {
    return phase == 0 ? 0 : items * (1 + (maxMoves - (phase < maxMoves ? phase : maxMoves)) / phase);
}
inline constexpr int MAX_DEVICES = 64;      // per-device set-up tables of the library (fixed size)

struct EvalArgs {
    float const* weights;     // device layout per model, weightStride floats apart
    uint64_t const* seeds;    // one per game
    float* fitness;           // [items], model-major (item = m * games + g)
    uint32_t* placements;     // [items] engine.place calls per item (may be null)
    uint64_t* counters;       // [items][3]: candidates fitted, candidates legal, lanes issued (may be null)
    uint64_t* timeline;       // [items][2]: device wall clock at item start / end (may be null)
    uint8_t* moves;           // [items][maxMoves] committed moves, hold << 7 | rotation << 5 | (x + 2) (may be null;
                              // written by the trace and pruned kernels)
    uint32_t* queue;          // work-queue counter (persistent mode)
    uint32_t items;
    uint32_t games;
    uint32_t maxMoves;
    uint32_t weightStride;
    // beam widths per level, at most the compiled 16 / 32 / 32 (variation_sequences.hpp:22): a level keeps the
    // first beamWidth[l] keys of its exact top-K list, which is the top-beamWidth[l] selection; the order ids
    // keep their relative order, so moves equal those of a beam built with these widths
    uint32_t beamWidth[3];
    uint32_t persistent;
    uint32_t garbage; // start every game with 0-4 garbage rows (seeded from the game seed)
    // Two-phase schedule (persistent mode): the first launch plays every item for at most phaseMoves moves and
    // parks the unfinished ones (engine state, counters, move index) in parked[item], listing them in
    // resumeList[0 .. *resumeCount); the second launch (resumePhase = 1) plays only the parked games, taking them
    // from a work ring in turns of phaseMoves moves: a game not done after its turn is parked again and appended
    // to the ring, so the unfinished games advance together and the longest ones end together, instead of the
    // last ones to start running alone at the end. The ring starts with the parked games of the models with the
    // most parked games first. phaseMoves = 0: one phase.
    uint32_t phaseMoves;
    uint32_t resumePhase;
    uint32_t* parked;      // [items][park_words(kind)]
    uint32_t* resumeList;  // [items], in parking order
    uint32_t* resumeCount; // one counter
    uint32_t* modelParked; // [items / games] parked games per model
    uint32_t* ring;        // [0] head (next slot to take), [1] tail (slots taken by writers), [2 + s] item + 1 of
                           // slot s, 0 until its writer publishes it
    uint32_t ringCap;      // slots; ring_slots() bounds what a second launch can append
    // pruned beam: at plies in pruneMask (bit 1, 2 = middle levels, bit 3 = leaves) only the pruner's top pruneK
    // legal placements of every list are scored by the value network (null pruner = full beam)
    PrunerDev const* pruner;
    uint32_t pruneK;
    uint32_t pruneMask;
    // full-beam traces: every legal candidate of plies 1..3 of the moves whose index is congruent to the game
    // seed modulo traceEvery (null = off)
    uint32_t* trace;      // [traceCap][11] records: parent board (10 words) + meta
    uint32_t* traceCount; // records produced (may exceed traceCap; the excess is dropped)
    uint32_t traceCap;
    uint32_t traceEvery;
};

hipError_t launch_beam_eval(int kind, EvalArgs const& a, int wavesPerBlock, uint32_t blocks, hipStream_t stream);
// uploads the constant tables computed on the host before the first launch on the current device (once per
// process and device; later calls return at once)
hipError_t upload_constants();
// leaf values of n boards (10 words each) as the search scores candidate boards; weights in device layout
hipError_t launch_model_values(int kind, float const* weights, uint32_t const* boards, uint32_t const* clears,
                               uint8_t const* held, uint32_t n, float* out, hipStream_t s);
// mode: 0 full beam, 1 full beam with traces, 2 pruned beam
hipError_t max_active_blocks(int kind, int wavesPerBlock, int mode, int* blocksPerCU);
hipError_t launch_bits_canary(uint32_t* ok, hipStream_t stream);
hipError_t launch_feature_norms(float* out, uint32_t n, hipStream_t stream);

} // namespace ta3h
