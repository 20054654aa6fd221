// Beam-search evaluator kernels for CDNA4 (gfx950), wave64.
//
// A workgroup of G waves (G = 1, 2 or 4; one per SIMD) plays one (model, game) item at a time and pulls
// the next item from a device work queue when its game ends, so uneven, early-terminating games keep the
// device busy. The G waves split every stage's candidates and merge their exact partial top-K lists
// through LDS, which cuts the per-move latency that bounds the longest (2,000-piece) games. Per move it runs
// the depth-4 beam search of the reference (ta3_ai/incl/ta3/ai/search/beam.hpp:483-570):
//   roots (<= 2 x 34 candidates) -> top-16 -> level 1 (<= 16 x 2 x 34) -> top-32
//   -> level 2 (<= 32 x 2 x 34) -> top-32 -> leaves (<= 32 x 2 x 34) -> argmax -> commit.
// Each stage is a compacted candidate space: segment s = 2k + c is (kept node k, child c); its
// placements are laid out back to back, one candidate per lane, so pieces with 9, 17 or 34
// placements all fill the 64 lanes. The slot / order id of a candidate is s * 34 + i, which is the
// reference's slot numbering (beam.hpp:296, :360, :423).
//
// Exact top-K ("max score, lowest slot on ties", beam.hpp:244-265) is kept per wave as a lane-distributed
// sorted list of unique 64-bit keys (score << 32 | ~slot), merged chunk by chunk with in-wave bitonic
// networks; a chunk with no key above the current K-th is skipped.
//
// Floating point: every multiply and add of the value network is rounded separately (__fmul_rn /
// __fadd_rn) in the reference order, divisions and square roots are IEEE correctly rounded, so leaf
// values are bit-identical to the reference host build.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <mutex>
#include <type_traits>
#include <vector>

#include "kernels.hpp"
#include "tables.hpp"

namespace ta3h {
namespace {

__constant__ Tables c_tab = TABLES;

// sqrt(sv) / 100 for every surface-variance sum sv in [0, 9 * 23^2], from host IEEE arithmetic (the device
// square root is not correctly rounded)
constexpr int SV_MAX = 9 * HEIGHT * HEIGHT;
__constant__ float c_svNorm[SV_MAX + 1];

constexpr int K0 = 16, K1 = 32, K2 = 32;        // beam widths (variation_sequences.hpp:22)
constexpr uint32_t LEAF_ORDERS = K2 * 2 * 34;   // leaf_orders (beam.hpp:188)
constexpr uint32_t SPAN = LEAF_ORDERS + K0;     // order_span (beam.hpp:196-198)
constexpr int COMMITTED = 64;                   // board index of the committed board in wave LDS

// ---------------------------------------------------------------------------------------------------
// wave helpers
// ---------------------------------------------------------------------------------------------------
__device__ inline uint32_t lane_id() { return __lane_id(); }

// orders this wave's LDS writes before its later LDS reads (all lanes of a wave run in lockstep)
__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ inline uint32_t uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ inline uint64_t uniform64(uint64_t v) {
    return static_cast<uint64_t>(uniform(static_cast<uint32_t>(v))) | static_cast<uint64_t>(uniform(static_cast<uint32_t>(v >> 32))) << 32;
}

__device__ inline uint32_t shfl_xor_u32(uint32_t v, int m) { return static_cast<uint32_t>(__shfl_xor(static_cast<int>(v), m)); }
__device__ inline uint32_t shfl_u32(uint32_t v, int src) { return static_cast<uint32_t>(__shfl(static_cast<int>(v), src)); }
__device__ inline uint64_t shfl_xor_u64(uint64_t v, int m) {
    uint32_t const lo = shfl_xor_u32(static_cast<uint32_t>(v), m);
    uint32_t const hi = shfl_xor_u32(static_cast<uint32_t>(v >> 32), m);
    return (static_cast<uint64_t>(hi) << 32) | lo;
}
__device__ inline uint64_t shfl_u64(uint64_t v, int src) {
    uint32_t const lo = shfl_u32(static_cast<uint32_t>(v), src);
    uint32_t const hi = shfl_u32(static_cast<uint32_t>(v >> 32), src);
    return (static_cast<uint64_t>(hi) << 32) | lo;
}
__device__ inline uint64_t readlane_u64(uint64_t v, int src) {
    uint32_t const lo = __builtin_amdgcn_readlane(static_cast<uint32_t>(v), src);
    uint32_t const hi = __builtin_amdgcn_readlane(static_cast<uint32_t>(v >> 32), src);
    return (static_cast<uint64_t>(hi) << 32) | lo;
}

__device__ inline uint32_t wave_excl_scan(uint32_t x) {
    uint32_t const lane = lane_id();
    uint32_t v = x;
#pragma unroll
    for(int d = 1; d < 64; d <<= 1) {
        uint32_t const t = static_cast<uint32_t>(__shfl_up(static_cast<int>(v), d));
        if(lane >= static_cast<uint32_t>(d)) v += t;
    }
    return v - x;
}


__device__ inline uint64_t wave_or_u64(uint64_t x) {
#pragma unroll
    for(int m = 32; m > 0; m >>= 1) x |= shfl_xor_u64(x, m);
    return x;
}

__device__ inline uint64_t wave_max_u64(uint64_t x) {
#pragma unroll
    for(int m = 32; m > 0; m >>= 1) {
        uint64_t const p = shfl_xor_u64(x, m);
        x = p > x ? p : x;
    }
    return x;
}

// descending bitonic sort of one key per lane (64 keys)
__device__ inline uint64_t bitonic_sort_desc(uint64_t v) {
    uint32_t const lane = lane_id();
#pragma unroll
    for(int k = 2; k <= 64; k <<= 1) {
#pragma unroll
        for(int j = k >> 1; j > 0; j >>= 1) {
            uint64_t const p = shfl_xor_u64(v, j);
            bool const takeMax = ((lane & j) == 0) == ((lane & k) == 0);
            uint64_t const mx = v > p ? v : p;
            uint64_t const mn = v > p ? p : v;
            v = takeMax ? mx : mn;
        }
    }
    return v;
}

// chunks with at most this many keys above the running K-th best insert them one by one instead of sorting
constexpr int TOPK_INSERT_MAX = 16;

// Running exact top-K (K <= 32): lanes 0..K-1 hold the best keys seen so far in descending order,
// other lanes hold 0. Folds one chunk of 64 candidate keys (0 = no candidate) into the list. Keys are unique,
// so the list after a chunk is the same whether its qualifying keys are inserted one at a time (few of them:
// rank by one ballot, shift the tail down one lane; a key no longer above the K-th is dropped) or the chunk is
// sorted and merged (many).
template<int K>
__device__ inline uint64_t topk_merge(uint64_t running, uint64_t key) {
    uint64_t const thr = readlane_u64(running, K - 1);
    uint64_t better = __ballot(key > thr);
    if(better == 0) return running;
    uint32_t const lane = lane_id();
    if(__popcll(better) <= TOPK_INSERT_MAX) {
        uint32_t const up = lane == 0 ? 0u : lane - 1;
        do {
            int const src = __builtin_ctzll(better);
            better &= better - 1;
            uint64_t const k = readlane_u64(key, src);
            uint32_t const pos = static_cast<uint32_t>(__popcll(__ballot(running > k))); // keys above k
            if(pos < static_cast<uint32_t>(K)) {
                uint64_t const prev = shfl_u64(running, static_cast<int>(up));
                running = lane < pos ? running : (lane == pos ? k : (lane < static_cast<uint32_t>(K) ? prev : 0));
            }
        } while(better != 0);
        return running;
    }
    uint64_t const sorted = bitonic_sort_desc(key);
    uint64_t const mirrored = shfl_u64(sorted, static_cast<int>(63 - lane));
    uint64_t v = lane < 32 ? running : mirrored; // descending then ascending: a bitonic sequence
#pragma unroll
    for(int j = 32; j > 0; j >>= 1) {
        uint64_t const p = shfl_xor_u64(v, j);
        uint64_t const mx = v > p ? v : p;
        uint64_t const mn = v > p ? p : v;
        v = (lane & j) == 0 ? mx : mn;
    }
    return lane < static_cast<uint32_t>(K) ? v : 0;
}

// ---------------------------------------------------------------------------------------------------
// bit helpers (semantics of ta3_sim/incl/ta3/sim/utility/bits.hpp:27-58)
// ---------------------------------------------------------------------------------------------------
__device__ inline int ctz_nz(uint32_t v) { return __builtin_ctz(v); }          // v != 0
__device__ inline int bit_width(uint32_t v) { return 32 - __clz(static_cast<int>(v)); } // __clz(0) == 32
__device__ inline int popc(uint32_t v) { return __popc(v); }

__device__ inline uint32_t ordered_bits(float v) { // beam.hpp:204-207
    uint32_t const u = __float_as_uint(v);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// ---------------------------------------------------------------------------------------------------
// engine state (one per wave, in LDS); the serial engine steps run on lane 0
// ---------------------------------------------------------------------------------------------------
struct EngineState {
    uint32_t cols[WIDTH];
    uint32_t queue[6];
    uint32_t held;
    uint32_t gameOver;
    uint32_t bag[7];
    uint32_t bagPos;
    uint64_t rng[4];
    // stats_v4 fields read by score_v1 (ta3_ai/incl/ta3/ai/models/metrics/tetris_stats_v4.hpp:8-15)
    uint32_t pieces;
    uint32_t clears[5];
    uint64_t sumMaxHeight;
    uint64_t sumHoleDepths;
    uint32_t placements;
};

constexpr uint32_t ENGINE_WORDS = sizeof(EngineState) / 4;
// root scores carried to the next move (RootCacheMeta): valid, two list keys, two lists of 34 scores
constexpr uint32_t PARK_RC = ENGINE_WORDS + 6, RC_WORDS = 3 + 2 * 34;
static_assert(sizeof(EngineState) % 4 == 0 && PARK_RC + RC_WORDS <= PARK_WORDS && ENGINE_WORDS <= 64,
              "a parked game fits its record and one wave copies it");

constexpr int MAX_WAVES = 4; // waves cooperating on one game: 1, 2 or 4

struct alignas(16) GameLds {
    uint32_t boards[(COMMITTED + 1) * WIDTH]; // frontier ping-pong: [buf*32 + j], committed at 64
    uint32_t clears[COMMITTED + 1];           // clear histogram along the path (model_defs.hpp:37-58)
    uint32_t score[COMMITTED + 1];            // ordered_bits score of the kept node
    uint32_t info[COMMITTED + 1];             // tree node | root rank << 8
    uint32_t segOff[64];                      // exclusive prefix offsets of the stage's segments
    uint32_t segInfo[64];                     // parent board | piece << 8 | heldIsI << 11 | child << 12
    // a stage's candidates c = 64 k + l (chunk k, lane l) without a per-candidate table (the wide plain kernel): the
    // scored lists in order (listOf[j] = segment), the list holding chunk k's first candidate (chunkList[k]) and
    // the lanes of chunk k where another list starts (chunkStarts[k]); see cand_of
    uint64_t chunkStarts[36];
    uint8_t listOf[64];
    uint8_t chunkList[36];
    uint32_t rootSlot[32];
    uint32_t slotPiece[8];                    // pieces per abstract window slot (beam.hpp:226-237)
    uint32_t treeIdx;
    uint32_t stageTotal;
    uint32_t item;
    uint64_t partTop[MAX_WAVES][32];          // per-wave top-K lists of a stage
    uint64_t partBest[MAX_WAVES];             // per-wave best leaf keys
    uint64_t partMask[MAX_WAVES];             // per-wave kept-node / segment legality masks
    uint32_t partLegal[MAX_WAVES];            // per-wave legal-candidate counts of the item
    EngineState eng;
    uint32_t moveItem; // the item being played (its move-log row; trace and pruned kernels)
};

// A stage's flat candidate index -> segment << 6 | placement index (every kernel but the wide network's plain one,
// whose LDS has no room for it and which uses the chunk headers of GameLds instead; with the table the 16-32-32-1
// generation runs 1.5% faster, V11's alike).
struct alignas(16) CandLds {
    uint16_t seg[32 * 2 * 34];
};

// Extra LDS of the trace / pruned kernels: the pruner's board projection and kept-candidate list (the pruner
// logits share the lanes' candidate-board scratch, which is idle until the value network runs).
struct alignas(16) PruneLds {
    float bp[32][PR_H];
    uint16_t kept[32 * 2 * 34];
    uint64_t segBest[64];   // traces: best leaf key of every leaf list
    uint32_t survivors[68]; // traces: slots kept by a middle level (2,176 bits)
    uint32_t keptCount;     // pruned stages: candidates kept for the value network
};


// Root lists scored by the previous move. When the committed placement cleared no line, the next move's root lists
// (the committed board, an empty clear histogram, a piece, held-is-I) are level-1 lists of the chosen root node
// whenever piece and held-is-I agree, which the previous search scored: the same candidates with the same values.
// Level 1 keeps the scores of the lists of the best RC_ROOTS kept roots in LDS (`level1`; all 16 but for the wide
// network, whose LDS leaves room for the best 6: the chosen root is among them in ~86% of moves); the commit records
// where the chosen root's two lists are (a repeat's at its origin, an earlier list, so of a better root); the next
// root stage, which runs before the next level 1 rewrites them, takes its keys from there when every root list
// matches. A parked game carries `valid`, `nextKey` and the two lists' scores in its record, so the choice
// depends only on the game.
struct RootCacheMeta {
    uint32_t valid;      // the last committed placement cleared no line (nextKey / src apply)
    uint32_t nextKey[2]; // piece | heldIsI << 3 | 16 of the chosen root's level-1 lists (0: none)
    uint32_t src[2];     // their list indices in level1 (scores of list l at level1[34 l ..])
    uint32_t info1[32];  // segInfo of the level-1 lists
};
template<int KIND>
constexpr uint32_t RC_ROOTS = KIND == MODEL_F32W ? 6 : 16;
template<int KIND>
struct alignas(16) RootScoresLds {
    uint32_t s[RC_ROOTS<KIND> * 2 * 34]; // ordered-bits score of level-1 candidate slot s * 34 + i (0: illegal)
};
static_assert(RC_WORDS == 3 + 2 * 34, "a parked game's root scores: valid, nextKey, two lists");

// Lists of levels 1 and 2 scored by the previous move (the plain 16-32-32-1 kernels). Where the root lists come
// from the previous move (RootCacheMeta), a root of this move is a level-1 candidate of the chosen root; if that
// candidate was a kept level-1 node of the previous search, its level-2 lists with this list's piece and held-is-I
// are this move's level-1 lists of that root, with the same content (the board follows from the path; the clear
// histograms agree because the committed placement cleared no line). One level deeper likewise: a kept level-1
// node of this move in such a list was a level-2 candidate of the previous search, and if that was kept, its leaf
// lists are this move's level-2 lists. Such lists, when not repeats, are not scored: their candidates follow the
// scored ones in the stage's candidate table, and their keys come from the previous move's scores. Level 2 and
// the leaves keep their scores, list infos and legality, levels 1 and 2 their kept slots, for the next move; a
// parked game carries them (the record part) when its roots may be reused.
struct alignas(16) DeepLds {
    uint32_t s2[64 * 34]; // level-2 candidate scores by slot (0: illegal or 0), read by the next move's level 1
    uint32_t sL[64 * 34]; // leaf candidate scores by slot, read by the next move's level 2
    uint32_t info2[64];   // segInfo of the level-2 lists
    uint32_t infoL[64];   // segInfo of the leaf lists
    uint64_t legal2;      // level-2 lists with a legal placement
    uint64_t legalL;      // leaf lists with a legal placement
    uint16_t kept1[32];   // slots of the kept level-1 nodes, in rank order
    uint16_t kept2[32];   // slots of the kept level-2 nodes
    uint32_t n1, n2;      // their counts
    uint32_t rank;        // rank of the committed root: its level-1 lists were 2 rank and 2 rank + 1
    // the current move (not in the record)
    uint32_t rootChild[2]; // roots reused: child of the previous chosen root whose level-1 list is root list g
    uint32_t reusedTotal;  // candidates of the stage's lists taken from the previous move
    uint64_t reusedLegal;  // those lists with a legal placement
    int8_t id1[64];        // level-1 list -> the previous move's level-2 list with its content, or -1
    uint8_t src[64];       // a list taken from the previous move: the previous list holding its scores
};
constexpr uint32_t DEEP_RECORD_WORDS = offsetof(DeepLds, rootChild) / 4;
static_assert(DEEP_RECORD_WORDS <= DEEP_PARK_WORDS, "the record part of DeepLds fits a parked game's record");

__device__ inline uint64_t rng_next(uint64_t* s) { // xoshiro256ss.hpp:33-45
    uint64_t const r = ((s[1] * 5) << 7 | (s[1] * 5) >> 57) * 9;
    uint64_t const t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = (s[3] << 45) | (s[3] >> 19);
    return r;
}

__device__ inline uint32_t draw_piece(EngineState& e) { // tetris_engine.hpp:116-129
    if(e.bagPos > 0) return e.bag[--e.bagPos];
    for(uint32_t i = 0; i < 7; ++i) e.bag[i] = i;
    for(uint32_t i = 6; i > 0; --i) {
        uint32_t const j = static_cast<uint32_t>(rng_next(e.rng) % (i + 1));
        uint32_t const t = e.bag[i];
        e.bag[i] = e.bag[j];
        e.bag[j] = t;
    }
    e.bagPos = 6;
    return e.bag[6];
}

__device__ inline uint32_t column_of(uint32_t cols, int col0, int x) { // board2.hpp:165-168
    uint32_t const d = static_cast<uint32_t>(x - col0);
    return d < 4 ? (cols >> (4 * d)) & 0xFu : 0u;
}

__device__ inline void update_game_over(EngineState& e) { // tetris_engine.hpp:159-161, board2.hpp:236-256
    uint32_t const t = e.queue[0];
    uint32_t const cols = c_tab.shapeCols[t][0];
    int const col0 = 3 + c_tab.shapeLeft[t][0];
    int const w = (bit_width(cols) + 3) / 4;
    bool ok = col0 >= 0 && col0 + w <= WIDTH;
    uint32_t hit = 0;
    for(int x = 0; x < WIDTH; ++x) hit |= e.cols[x] & column_of(cols, col0, x);
    e.gameOver = (ok && hit == 0) ? 0u : 1u;
}

__device__ inline void advance_queue(EngineState& e) { // tetris_engine.hpp:131-135
    for(int i = 1; i < 6; ++i) e.queue[i - 1] = e.queue[i];
    e.queue[5] = draw_piece(e);
}

__device__ void engine_reset(EngineState& e, uint64_t seed) { // tetris_engine.hpp:103-114
    for(int x = 0; x < WIDTH; ++x) e.cols[x] = 0;
    e.held = NO_PIECE;
    e.gameOver = 0;
    e.bagPos = 0;
    for(int i = 0; i < 4; ++i) { // splitmix64 seeding (xoshiro256ss.hpp:20-28)
        seed += 0x9e3779b97f4a7c15ULL;
        uint64_t z = seed;
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
        e.rng[i] = z ^ (z >> 31);
    }
    for(int i = 0; i < 6; ++i) e.queue[i] = draw_piece(e);
    update_game_over(e);
    e.pieces = 0;
    for(int k = 0; k < 5; ++k) e.clears[k] = 0;
    e.sumMaxHeight = 0;
    e.sumHoleDepths = 0;
    e.placements = 0;
}

// Start-position extension (build-defined): 0-4 garbage rows at the bottom, each full but for one hole, from a
// second xoshiro256** seeded with seed ^ 0x6A09E667F3BCC908 (the piece sequence is unchanged):
// n = next() % 5, then for rows y = 22, 21, ... (n of them): hole column = next() % 10.
__device__ void add_garbage_rows(EngineState& e, uint64_t seed) {
    uint64_t r[4];
    seed ^= 0x6A09E667F3BCC908ULL;
    for(int i = 0; i < 4; ++i) {
        seed += 0x9e3779b97f4a7c15ULL;
        uint64_t z = seed;
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
        r[i] = z ^ (z >> 31);
    }
    uint32_t const n = static_cast<uint32_t>(rng_next(r) % 5u);
    for(uint32_t k = 0; k < n; ++k) {
        uint32_t const hole = static_cast<uint32_t>(rng_next(r) % WIDTH);
        uint32_t const row = 1u << (HEIGHT - 1 - k);
        for(uint32_t x = 0; x < WIDTH; ++x)
            if(x != hole) e.cols[x] |= row;
    }
    update_game_over(e);
}

__device__ inline int clear_full_rows(uint32_t* cols) { // board2.hpp:333-342 with :218-234
    uint32_t full = FIELD;
    for(int x = 0; x < WIDTH; ++x) full &= cols[x];
    if(full == 0) return 0;
    int const n = popc(full);
    for(int x = 0; x < WIDTH; ++x) {
        uint32_t col = cols[x], rows = full;
        for(int i = 0; i < n; ++i) {
            uint32_t const row = rows & (~rows + 1);
            col = ((col & (row - 1)) << 1) | (col & ~((row << 1) - 1));
            rows ^= row;
        }
        cols[x] = col;
    }
    return n;
}

// hard drop of the current piece; -1 = DIED (tetris_engine.hpp:137-146)
__device__ int engine_place(EngineState& e, int o, int x) {
    uint32_t const t = e.queue[0];
    uint32_t const cols = c_tab.shapeCols[t][o];
    int const col0 = x + c_tab.shapeLeft[t][o];
    int minDrop = HEIGHT; // dropDistance at y = 0 (board2.hpp:262-275)
    for(int c = 0; c < WIDTH; ++c) {
        uint32_t const nib = column_of(cols, col0, c);
        int const low = bit_width(nib) - 1;
        uint32_t const mask = ~((1u << (low + 1)) - 1);
        int const drop = ctz_nz((e.cols[c] & mask) | FLOOR) - low - 1;
        if(nib != 0) minDrop = drop < minDrop ? drop : minDrop;
    }
    for(int c = 0; c < WIDTH; ++c) e.cols[c] |= column_of(cols, col0, c) << minDrop;
    int const cleared = clear_full_rows(e.cols);
    advance_queue(e);
    update_game_over(e);
    return e.gameOver ? -1 : cleared;
}

__device__ void engine_hold(EngineState& e) { // tetris_engine.hpp:148-157
    if(e.held == NO_PIECE) {
        e.held = e.queue[0];
        advance_queue(e);
    } else {
        uint32_t const t = e.held;
        e.held = e.queue[0];
        e.queue[0] = t;
    }
    update_game_over(e);
}

// stats_v4::advance for the fields score_v1 reads (tetris_stats.hpp:19-23, metrics.hpp:94-127)
__device__ void stats_advance(EngineState& e, int cleared) {
    ++e.pieces;
    ++e.clears[cleared];
    int maxH = 0, hd = 0;
    for(int x = 0; x < WIDTH; ++x) {
        uint32_t const col = e.cols[x];
        int const h = HEIGHT - ctz_nz(col | FLOOR);
        maxH = h > maxH ? h : maxH;
        int const n = popc(col);
        int s = 0;
        for(int y = 0; y < HEIGHT; ++y) s += ((col >> y) & 1u) ? y : 0;
        hd += (HEIGHT - 1) * n + n * (n + 1) / 2 - (s + n * n);
    }
    e.sumMaxHeight += static_cast<uint64_t>(maxH);
    e.sumHoleDepths += static_cast<uint64_t>(hd);
}

__device__ float final_score(EngineState const& e) { // score_v1.hpp:20-34; model_eval_host.cpp:51
    if(e.pieces == 0) return -1.0e6f;
    double const payouts[5] = {0.0, 0.2, 0.8, 2.0, 10.0};
    double bonus = 0.0;
    for(int k = 1; k < 5; ++k) bonus = __dadd_rn(bonus, __dmul_rn(payouts[k], static_cast<double>(e.clears[k])));
    double const pieces = static_cast<double>(e.pieces);
    double const survival = __dmul_rn(0.02, pieces);
    double const avgMaxH = __ddiv_rn(static_cast<double>(e.sumMaxHeight), pieces);
    double const over = __dadd_rn(avgMaxH, -12.0);
    double const heightPenalty = __dmul_rn(over > 0.0 ? over : 0.0, 1.0);
    double const avgHd = __ddiv_rn(static_cast<double>(e.sumHoleDepths), pieces);
    double const holesPenalty = __dmul_rn(avgHd, 1.5);
    double s = __dadd_rn(bonus, survival);
    s = __dsub_rn(s, heightPenalty);
    s = __dsub_rn(s, holesPenalty);
    return static_cast<float>(s);
}

// ---------------------------------------------------------------------------------------------------
// candidate evaluation: apply (search.hpp:61-69) + features + value network
// ---------------------------------------------------------------------------------------------------
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef short i16x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// Weights are read through the constant address space: the per-model base address is wave-uniform, so
// every weight becomes a scalar load into an SGPR that VALU instructions take as an operand for free.
using wptr = float const __attribute__((address_space(4)))*;

// Hides the weight base address from loop-invariant code motion: weights are then re-loaded from the
// scalar cache where they are used instead of being hoisted into (and spilled from) scarce SGPRs.
__device__ inline wptr opaque(wptr w) {
    uint64_t v = reinterpret_cast<uint64_t>(w);
    asm volatile("" : "+s"(v));
    return reinterpret_cast<wptr>(v);
}

// the same for any wave-uniform pointer
template<typename T>
__device__ inline T* opaque_ptr(T* p) {
    uint64_t v = reinterpret_cast<uint64_t>(p);
    asm volatile("" : "+s"(v));
    return reinterpret_cast<T*>(v);
}

// the same for a pointer into device global memory, kept in the global address space so that its loads are
// global loads (counted by vmcnt only) rather than flat loads, whose waits also drain the wave's LDS accesses
template<typename T>
__device__ inline T __attribute__((address_space(1)))* opaque_global(T* p) {
    uint64_t v = reinterpret_cast<uint64_t>(p);
    asm volatile("" : "+s"(v));
    return reinterpret_cast<T __attribute__((address_space(1)))*>(v);
}

__device__ inline float relu(float v) { return (0.0f < v) ? v : 0.0f; } // std::max<float>(0, v)
__device__ inline float mad_rn(float acc, float a, float b) { return __fadd_rn(acc, __fmul_rn(a, b)); }
__device__ inline float fma_rn(float acc, float a, float b) { return __fmaf_rn(a, b, acc); }

// V11 value: 10 features -> 8 ReLU -> 1 (model_v11.hpp:166-180; utility.hpp:29-46). The zero skipping of the
// 16-feature network does not pay here: with three variants the V11 generation ran 4% slower, with one ballot
// and two copies (skipping the multi-line clear counts) 5% slower.
__device__ inline float value_v11(wptr w, float const (&f)[16]) {
    float out = w[96];
#pragma unroll
    for(int o = 0; o < 8; ++o) {
        float v = w[80 + o];
#pragma unroll
        for(int i = 0; i < 10; ++i) v = mad_rn(v, f[i], w[o * 10 + i]);
        out = mad_rn(out, relu(v), w[88 + o]);
    }
    return out;
}

// features of the 16-feature set that are zero across a whole chunk of candidates most of the time: the
// multi-line clear counts (1-3); with them the single-clear count, the held-I flag and the two line-clear
// totals (0, 8, 13, 14)
constexpr uint32_t RARE_FEATURES = 0xEu, F16_OFTEN_ZERO = 0x610Fu;

// 16 -> 32 ReLU -> 32 ReLU -> 1 on the device layout: W1[32][16], b1[32], W2 transposed [32 in][32 out],
// b2[32], W3[32], b3. Each step is one fused multiply-add (the network's defined arithmetic); layer 2
// accumulates in ascending input order, the order of a row-wise dense().
constexpr int F16_W1 = 0, F16_B1 = 512, F16_W2T = 544, F16_B2 = 1568, F16_W3 = 1600, F16_B3 = 1632;

// Hidden layers with exact skipping (SPARSE, finite weights only): a step fma(acc, x, w) whose x is zero in every
// active lane adds a zero product, which can change only the sign of a zero accumulator, and every layer-1 and
// layer-2 accumulator reaches the output only through relu, which maps both zeros to +0. SKIP lists the features
// known to be zero in every active lane; with SPARSE a first-layer unit that is zero in every active lane skips
// its 32 second-layer steps (a wave-uniform branch).
template<uint32_t SKIP, bool SPARSE>
__device__ inline void f16_hidden(wptr w, float const (&f)[16], float (&acc)[32]) {
#pragma unroll 2
    for(int o = 0; o < 32; ++o) {
        float v = w[F16_B1 + o];
#pragma unroll
        for(int i = 0; i < 16; ++i)
            if(((SKIP >> i) & 1u) == 0) v = fma_rn(v, f[i], w[F16_W1 + o * 16 + i]);
        float const h = relu(v);
        if(!SPARSE || __ballot(h != 0.0f) != 0) {
#pragma unroll
            for(int j = 0; j < 32; ++j) acc[j] = fma_rn(acc[j], h, w[F16_W2T + o * 32 + j]);
        }
    }
}

// sparse: the model's weights are all finite (checked per item), which makes the skipping exact
__device__ inline float value_f16(wptr w, float const (&f)[16], bool sparse) {
    float acc[32];
#pragma unroll
    for(int j = 0; j < 32; ++j) acc[j] = w[F16_B2 + j];
    if(!sparse) {
        f16_hidden<0u, false>(w, f, acc);
    } else {
        bool rare = false, often = false;
#pragma unroll
        for(int i = 0; i < 16; ++i) {
            if((RARE_FEATURES >> i) & 1u) rare |= f[i] != 0.0f;
            if((F16_OFTEN_ZERO >> i) & 1u) often |= f[i] != 0.0f;
        }
        if(__ballot(often) == 0) f16_hidden<F16_OFTEN_ZERO, true>(w, f, acc);
        else if(__ballot(rare) == 0) f16_hidden<RARE_FEATURES, true>(w, f, acc);
        else f16_hidden<0u, true>(w, f, acc);
    }
    float out = w[F16_B3];
#pragma unroll
    for(int j = 0; j < 32; ++j) out = fma_rn(out, relu(acc[j]), w[F16_W3 + j]);
    return out;
}

// Two chunks of candidates through the network at once (fa / fb: the features of this lane's candidates of the
// two chunks; la / lb: whether they are legal, the lanes whose values are used): every weight is loaded into an
// SGPR once for both, which halves the scalar loads and the scalar-cache misses per candidate. Each candidate's
// operations are value_f16's in the same order; the exact skipping considers the legal lanes of the chunk it
// skips for (a unit or feature skipped for one chunk may still be computed for the other).
template<uint32_t SKIP, bool SPARSE>
__device__ inline void f16_hidden2(wptr w, float const (&fa)[16], float const (&fb)[16], bool la, bool lb,
                                   float (&accA)[32], float (&accB)[32]) {
#pragma unroll 2
    for(int o = 0; o < 32; ++o) {
        float va = w[F16_B1 + o], vb = va;
#pragma unroll
        for(int i = 0; i < 16; ++i)
            if(((SKIP >> i) & 1u) == 0) {
                float const wi = w[F16_W1 + o * 16 + i];
                va = fma_rn(va, fa[i], wi);
                vb = fma_rn(vb, fb[i], wi);
            }
        float const ha = relu(va), hb = relu(vb);
        bool const doA = !SPARSE || __ballot(la && ha != 0.0f) != 0;
        bool const doB = !SPARSE || __ballot(lb && hb != 0.0f) != 0;
        if(doA) {
#pragma unroll
            for(int j = 0; j < 32; ++j) accA[j] = fma_rn(accA[j], ha, w[F16_W2T + o * 32 + j]);
        }
        if(doB) {
#pragma unroll
            for(int j = 0; j < 32; ++j) accB[j] = fma_rn(accB[j], hb, w[F16_W2T + o * 32 + j]);
        }
    }
}

__device__ inline void value_f16x2(wptr w, float const (&fa)[16], float const (&fb)[16], bool la, bool lb,
                                   bool sparse, float& outA, float& outB) {
    float accA[32], accB[32];
#pragma unroll
    for(int j = 0; j < 32; ++j) accA[j] = accB[j] = w[F16_B2 + j];
    if(!sparse) {
        f16_hidden2<0u, false>(w, fa, fb, la, lb, accA, accB);
    } else {
        bool rare = false, often = false;
#pragma unroll
        for(int i = 0; i < 16; ++i) {
            if((RARE_FEATURES >> i) & 1u) rare |= (la && fa[i] != 0.0f) || (lb && fb[i] != 0.0f);
            if((F16_OFTEN_ZERO >> i) & 1u) often |= (la && fa[i] != 0.0f) || (lb && fb[i] != 0.0f);
        }
        if(__ballot(often) == 0) f16_hidden2<F16_OFTEN_ZERO, true>(w, fa, fb, la, lb, accA, accB);
        else if(__ballot(rare) == 0) f16_hidden2<RARE_FEATURES, true>(w, fa, fb, la, lb, accA, accB);
        else f16_hidden2<0u, true>(w, fa, fb, la, lb, accA, accB);
    }
    outA = w[F16_B3];
    outB = outA;
#pragma unroll
    for(int j = 0; j < 32; ++j) {
        float const w3 = w[F16_W3 + j];
        outA = fma_rn(outA, relu(accA[j]), w3);
        outB = fma_rn(outB, relu(accB[j]), w3);
    }
}

// 4 x 4 transpose of (register, 16-lane row) blocks: t[b] in row a = r_a in row b. v_permlane32_swap exchanges the
// off-diagonal 2 x 2 blocks of rows, v_permlane16_swap the single rows inside them (the whole wave takes part).
__device__ inline void transpose_rows(uint32_t r0, uint32_t r1, uint32_t r2, uint32_t r3, uint32_t (&t)[4]) {
    auto const s02 = __builtin_amdgcn_permlane32_swap(r0, r2, false, false);
    auto const s13 = __builtin_amdgcn_permlane32_swap(r1, r3, false, false);
    auto const s01 = __builtin_amdgcn_permlane16_swap(s02[0], s13[0], false, false);
    auto const s23 = __builtin_amdgcn_permlane16_swap(s02[1], s13[1], false, false);
    t[0] = s01[0];
    t[1] = s01[1];
    t[2] = s23[0];
    t[3] = s23[1];
}

// The 16-32-32-1 network's value for the 64 candidates of a chunk (fin: this lane's features; the whole wave takes
// part), its two hidden layers on v_mfma_f32_16x16x4_f32 (rows = units, columns = 16 candidates, K = 4 inputs per
// instruction). That instruction adds its four products to the accumulator one after another, each rounded once
// (as fmaf, in k order; checked on the device over normal, subnormal, zero, infinite and tie-rounding operands), so
// chaining the K blocks in input order computes exactly value_f16's chains (v = fma(x_i, w_i, v) for i ascending
// from the bias), with no step skipped. frag: the model's operands (f16f); the output layer (W3, b3) runs on the
// VALU from SGPRs as in value_f16, after the outputs are moved back to one candidate per lane.
__device__ inline float value_f16_mfma(float const* __restrict__ frag, wptr w, float const (&fin)[16]) {
    uint32_t const lane = lane_id();
    auto const g = opaque_global(frag);
    // B operands of layer 1: for inputs 4 s .. 4 s + 3, lane 16 k + n holds input 4 s + k of candidate 16 ct + n
    uint32_t b1[4][4];
#pragma unroll
    for(int s = 0; s < 4; ++s)
        transpose_rows(__float_as_uint(fin[4 * s]), __float_as_uint(fin[4 * s + 1]), __float_as_uint(fin[4 * s + 2]),
                       __float_as_uint(fin[4 * s + 3]), b1[s]);
    // layer 1: acc1[rt][ct] register j of lane 16 q + n = unit 16 rt + 4 j + q of candidate 16 ct + n
    // (the wave's issue priority is raised while it feeds the MFMA unit, so the other waves' VALU work fills the
    // MFMA latency instead of delaying the chain)
    __builtin_amdgcn_s_setprio(1);
    f32x4 acc1[2][4];
#pragma unroll
    for(int rt = 0; rt < 2; ++rt) {
        f32x4 const c = *reinterpret_cast<f32x4 const __attribute__((address_space(1)))*>(&g[f16f::C1 + (rt * 64 + lane) * 4]);
#pragma unroll
        for(int ct = 0; ct < 4; ++ct) acc1[rt][ct] = c;
#pragma unroll
        for(int s = 0; s < 4; ++s) {
            float const a = g[f16f::A1 + (rt * 4 + s) * 64 + lane];
#pragma unroll
            for(int ct = 0; ct < 4; ++ct)
                acc1[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, __uint_as_float(b1[s][ct]), acc1[rt][ct], 0, 0, 0);
        }
    }
#pragma unroll
    for(int rt = 0; rt < 2; ++rt)
#pragma unroll
        for(int ct = 0; ct < 4; ++ct)
#pragma unroll
            for(int j = 0; j < 4; ++j) acc1[rt][ct][j] = relu(acc1[rt][ct][j]);
    // layer 2: register s of acc1[rt1][ct] is the B operand of inputs 16 rt1 + 4 s .. + 3 (lane 16 k + n: input
    // 16 rt1 + 4 s + k); acc2[rt2][ct] register i of lane 16 q + n = output 16 rt2 + 4 q + i of candidate 16 ct + n
    f32x4 acc2[2][4];
#pragma unroll
    for(int rt2 = 0; rt2 < 2; ++rt2) {
        f32x4 const c = *reinterpret_cast<f32x4 const __attribute__((address_space(1)))*>(&g[f16f::C2 + (rt2 * 64 + lane) * 4]);
#pragma unroll
        for(int ct = 0; ct < 4; ++ct) acc2[rt2][ct] = c;
#pragma unroll
        for(int rt1 = 0; rt1 < 2; ++rt1)
#pragma unroll
            for(int s = 0; s < 4; ++s) {
                float const a = g[f16f::A2 + ((rt2 * 2 + rt1) * 4 + s) * 64 + lane];
#pragma unroll
                for(int ct = 0; ct < 4; ++ct)
                    acc2[rt2][ct] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, acc1[rt1][ct][s], acc2[rt2][ct], 0, 0, 0);
            }
    }
    __builtin_amdgcn_s_setprio(0);
    // outputs back to one candidate per lane (output 16 rt2 + 4 b + i in o[rt2][i][b]), then the output chain
    float out = w[F16_B3];
#pragma unroll
    for(int rt2 = 0; rt2 < 2; ++rt2) {
        uint32_t o[4][4];
#pragma unroll
        for(int i = 0; i < 4; ++i)
            transpose_rows(__float_as_uint(acc2[rt2][0][i]), __float_as_uint(acc2[rt2][1][i]), __float_as_uint(acc2[rt2][2][i]),
                           __float_as_uint(acc2[rt2][3][i]), o[i]);
#pragma unroll
        for(int b = 0; b < 4; ++b)
#pragma unroll
            for(int i = 0; i < 4; ++i)
                out = fma_rn(out, relu(__uint_as_float(o[i][b])), w[F16_W3 + 16 * rt2 + 4 * b + i]);
    }
    return out;
}

// x / c for an integer-valued x: reciprocal multiply plus one FMA correction. Equals the IEEE quotient for
// every numerator the features can take (checked exhaustively on the device by the feature-norm kernel).
__device__ inline float div_const(float x, float c, float r) {
    float const q = __fmul_rn(x, r);
    float const e = __fmaf_rn(-q, c, x);
    return __fmaf_rn(e, r, q);
}
__device__ inline float div20(int x) { return div_const(static_cast<float>(x), 20.0f, 1.0f / 20.0f); }
__device__ inline float div200(int x) { return div_const(static_cast<float>(x), 200.0f, 1.0f / 200.0f); }
__device__ inline float div5(int x) { return div_const(static_cast<float>(x), 5.0f, 1.0f / 5.0f); }
__device__ inline float div40(int x) { return div_const(static_cast<float>(x), 40.0f, 1.0f / 40.0f); }
__device__ inline float div100(int x) { return div_const(static_cast<float>(x), 100.0f, 1.0f / 100.0f); }

// set-bit index planes for 23 rows: plane b has bit y set when bit b of y is set
constexpr uint32_t PLANE0 = 0x2AAAAAu, PLANE1 = 0x4CCCCCu, PLANE2 = 0x70F0F0u, PLANE3 = 0x00FF00u, PLANE4 = 0x7F0000u;

// Sum over a board's filled cells of their row index (bit index y of the column words), for a board in registers:
// a carry-save count of the ten columns gives every row's filled-cell count as four bit planes (count =
// p0 + 2 p1 + 4 p2 + 8 p3 per row), and each plane's row-index sum is its five masked population counts.
__device__ inline int row_index_sum(uint32_t const* b) {
    auto fa = [](uint32_t x, uint32_t y, uint32_t z, uint32_t& carry) { // full adder of three one-bit planes
        carry = (x & y) | (z & (x ^ y));
        return x ^ y ^ z;
    };
    uint32_t ka, kb, kc, kd, kf, kg;
    uint32_t const sa = fa(b[0], b[1], b[2], ka), sb = fa(b[3], b[4], b[5], kb), sc = fa(b[6], b[7], b[8], kc);
    uint32_t const s1 = fa(sa, sb, sc, kd);
    uint32_t const p0 = s1 ^ b[9], ke = s1 & b[9];                // weight 1
    uint32_t const s2 = fa(ka, kb, kc, kf);                        // ka..ke: weight 2
    uint32_t const p1 = fa(s2, kd, ke, kg);
    uint32_t const p2 = kf ^ kg, p3 = kf & kg;                     // kf, kg: weight 4
    uint32_t const planes[4] = {p0, p1, p2, p3};
    constexpr uint32_t IDX[5] = {PLANE0, PLANE1, PLANE2, PLANE3, PLANE4};
    int byWeight[8] = {}; // population counts of weight 2^(k + j)
#pragma unroll
    for(int k = 0; k < 4; ++k)
#pragma unroll
        for(int j = 0; j < 5; ++j) byWeight[k + j] += popc(planes[k] & IDX[j]);
    int sum = 0;
#pragma unroll
    for(int e = 7; e >= 0; --e) sum = 2 * sum + byWeight[e];
    return sum;
}

// Candidate boards in LDS (the wide network's kernels with a feature tile, and the leaf-value kernel), one per
// lane: column x of the lane's board is col[x * SCR_STRIDE], so the 64 lanes of a wave touch 64 consecutive words
// (no bank conflicts) and the column loops stay rolled, which keeps the register footprint small. The search's
// other kernels keep the board in ten registers (place_regs, leaf_features<KIND, 1>): their loops unroll fully.
constexpr int SCR_STRIDE = 64;

// Features of the 64 candidates of a chunk for the wide network, bfloat16, one 64-byte row per lane; the 16-byte
// chunk q of row r sits at position q ^ (r & 3), which spreads the 16 rows a fragment read touches over the banks.
__device__ inline uint32_t feat_index(uint32_t row, uint32_t q) { return row * 32 + 8 * (q ^ (row & 3u)); }

__device__ inline uint16_t bf16_bits(float v) { return __builtin_bit_cast(uint16_t, static_cast<__bf16>(v)); }

// Feature pass over the candidate board (one streaming pass over the columns, v_board_metrics.hpp:130-225; a
// column's well flag is settled once its right neighbour is known). Fills f[0..9] (V11) or f[0..15] (16-feature
// set); the 32-feature set also writes its features 16..31 (bfloat16) straight into the lane's row of featTile or,
// for a board held in registers, into g[0..15] (float; the caller rounds them to bfloat16).
// STRIDE: distance between the board's columns, SCR_STRIDE for a lane's LDS board, 1 for a board held in registers
// (the column loop is then unrolled completely so that the array stays in registers).
template<int KIND, int STRIDE = SCR_STRIDE>
__device__ inline void leaf_features(uint32_t const* col, uint32_t clears, bool heldIsI, float* f,
                                     uint16_t* featTile, float* g = nullptr) {
    constexpr bool WIDE = KIND != MODEL_V11;
    constexpr bool F32 = KIND == MODEL_F32W;
    // x / c: the IEEE quotient; the wide network rounds its features to bfloat16, and there one multiply by the
    // rounded reciprocal gives the same value as the quotient for every numerator 0..2,600 (the features' range;
    // checked exhaustively on the host)
    auto nd = [](int x, float c, float r) {
        if constexpr(F32) return __fmul_rn(static_cast<float>(x), r);
        else return div_const(static_cast<float>(x), c, r);
    };
    // column sums: heights, filled cells and their squares, row indices of the filled cells
    int agg = 0, cells = 0, sq = 0, rows = 0;
    int maxH = 0, minH = HEIGHT, sv = 0, bump = 0, wells = 0, ytr = 0;
    int holeCols = 0, rowTr = 0, deepest = 0;
    uint32_t holeRows = 0, cPrev = 0;
    int hPrev2 = 0, hPrev = 0;
    uint32_t const lane = lane_id();
#pragma unroll(STRIDE == 1 ? WIDTH : 2)
    for(int x = 0; x < WIDTH; ++x) {
        uint32_t const c = col[x * STRIDE];
        uint32_t const yc = c | FLOOR;
        int const top = ctz_nz(yc);
        int const hx = HEIGHT - top;
        int const n = popc(c);
        agg += hx;
        cells += n;
        sq += n * n;
        maxH = hx > maxH ? hx : maxH;
        if constexpr(STRIDE != 1) // (a board in registers gets the sum from row_index_sum after the loop)
            rows += popc(c & PLANE0) + 2 * popc(c & PLANE1) + 4 * popc(c & PLANE2) + 8 * popc(c & PLANE3)
                + 16 * popc(c & PLANE4);
        if(x > 0) {
            int const d = hx - hPrev;
            sv += d * d;
            if(WIDE) bump += d < 0 ? -d : d;
            // column x-1 against its lower neighbour (only the right one at the wall): a well of depth >= 3
            // counts (wells_t, v_board_metrics.hpp:204-225); the wide set also keeps the deepest
            int const lo = x == 1 ? hx : (hPrev2 < hx ? hPrev2 : hx);
            int const dp = lo - hPrev;
            wells += dp >= 3 ? 1 : 0;
            if(F32) deepest = dp > deepest ? dp : deepest;
        }
        if(WIDE) {
            minH = hx < minH ? hx : minH;
            ytr += popc((yc ^ (yc >> 1)) & FIELD);
        }
        if(F32) {
            if constexpr(STRIDE == 1) {
                g[x] = nd(hx, 20.0f, 1.0f / 20.0f);
            } else {
                uint32_t const k = 16 + static_cast<uint32_t>(x);
                featTile[feat_index(lane, k >> 3) + (k & 7u)] = bf16_bits(nd(hx, 20.0f, 1.0f / 20.0f));
            }
            // empty cells under the column's top: c - 1 fills the rows above its top filled cell (all rows of an
            // empty column)
            uint32_t const hole = ~(c | (c - 1u)) & FIELD;
            holeRows |= hole;
            holeCols += hole != 0 ? 1 : 0;
            rowTr += popc((x > 0 ? (cPrev ^ c) : ~c) & FIELD); // the left wall counts as filled
            cPrev = c;
        }
        hPrev2 = hPrev;
        hPrev = hx;
    }
    int const d9 = hPrev2 - hPrev; // column 9 against its single neighbour, column 8
    wells += d9 >= 3 ? 1 : 0;
    if(F32) deepest = d9 > deepest ? d9 : deepest;
    if constexpr(STRIDE == 1) rows = row_index_sum(col);
    int const holes = agg - cells; // empty cells under the tops
    // hole depths, sum over the columns of (HEIGHT - 1) n + n (n + 1) / 2 - (n^2 + the filled rows' indices)
    int const hd = (HEIGHT - 1) * cells + (cells - sq) / 2 - rows;
#pragma unroll
    for(int k = 0; k < 4; ++k) f[k] = __fmul_rn(static_cast<float>((clears >> (8 * k)) & 0xFFu), 0.25f);
    f[4] = c_svNorm[sv];
    f[5] = nd(holes, 20.0f, 1.0f / 20.0f);
    f[6] = nd(hd, 200.0f, 1.0f / 200.0f);
    f[7] = static_cast<float>(maxH);
    f[8] = heldIsI ? 1.0f : 0.0f;
    f[9] = nd(wells, 5.0f, 1.0f / 5.0f);
    if constexpr(WIDE) {
        uint32_t const c0 = clears & 0xFFu, c1 = (clears >> 8) & 0xFFu, c2 = (clears >> 16) & 0xFFu,
                       c3 = clears >> 24;
        int const lines = static_cast<int>(c0 + 2 * c1 + 3 * c2 + 4 * c3);
        int const cscore = -static_cast<int>(c0) - 2 * static_cast<int>(c1) - 3 * static_cast<int>(c2)
            + 10 * static_cast<int>(c3);
        f[10] = nd(agg, 200.0f, 1.0f / 200.0f);
        f[11] = nd(ytr, 40.0f, 1.0f / 40.0f);
        f[12] = nd(bump, 100.0f, 1.0f / 100.0f);
        f[13] = __fmul_rn(static_cast<float>(lines), 0.1f);
        f[14] = __fmul_rn(static_cast<float>(cscore), 0.1f);
        f[15] = static_cast<float>(maxH - minH);
    }
    if constexpr(F32) {
        rowTr += popc(~cPrev & FIELD); // the right wall
        float const t[6] = {nd(popc(holeRows), 20.0f, 1.0f / 20.0f), nd(rowTr, 100.0f, 1.0f / 100.0f),
                            nd(deepest, 20.0f, 1.0f / 20.0f),    nd(cells, 200.0f, 1.0f / 200.0f),
                            nd(holeCols, 5.0f, 1.0f / 5.0f),     nd(minH, 20.0f, 1.0f / 20.0f)}; // features 26..31
        if constexpr(STRIDE == 1) {
#pragma unroll
            for(int k = 0; k < 6; ++k) g[10 + k] = t[k];
        } else {
            uint16_t* const c3 = &featTile[feat_index(lane, 3)]; // features 24..31
#pragma unroll
            for(int k = 0; k < 6; ++k) c3[2 + k] = bf16_bits(t[k]);
        }
    }
}

template<int KIND>
__device__ inline float leaf_value(uint32_t const* col, uint32_t clears, bool heldIsI, wptr w, bool sparse) {
    float f[16];
    leaf_features<KIND>(col, clears, heldIsI, f, nullptr);
    if constexpr(KIND == MODEL_F16) return value_f16(w, f, sparse);
    else return value_v11(w, f);
}

// ---------------------------------------------------------------------------------------------------
// wide network 32 -> 128 -> 128 -> 1 on BF16 MFMA (v_mfma_f32_16x16x32_bf16, FP32 accumulate)
// ---------------------------------------------------------------------------------------------------
// (weight layout: f32w in kernels.hpp)

// features 0..15 of the lane's candidate into chunks 0 and 1 of its row
__device__ inline void store_features16(uint16_t* featTile, float const (&f)[16]) {
    uint32_t const lane = lane_id();
#pragma unroll
    for(uint32_t q = 0; q < 2; ++q) {
        bf16x8 v;
#pragma unroll
        for(int j = 0; j < 8; ++j) v[j] = static_cast<__bf16>(f[8 * q + j]);
        *reinterpret_cast<bf16x8*>(&featTile[feat_index(lane, q)]) = v;
    }
}

// Value of every lane's candidate (the 64 rows of featWave; rows of illegal lanes hold stale values whose
// results are discarded: MFMA columns are independent). The 64 candidates are four 16-candidate column tiles,
// taken CB at a time: each weight fragment is loaded once for the CB tiles (the model's 42 KB of fragments
// exceed the vector L1, so every reload is an L2 read; four tiles per block run the wide generation 1.8x
// faster than one). Per column tile: layer 1 = 8 MFMAs, ReLU, bfloat16; layers 2 and 3 = 32 MFMAs, ReLU, then
// the FP32 output FMA chain of lane quarter q over o = 16 t + 4 q + j; value = ((p0 + p1) + (p2 + p3)) + b3.
// Every tile's operations and their order do not depend on CB.
// relu of two values already rounded to bfloat16: the signed 16-bit maximum of their bits and 0 (one v_pk_max_i16)
// equals bfloat16(relu(x)) for every x that is not NaN (negative values and -0 have the sign bit set). A NaN
// (only from non-finite or enormous weights) with the sign bit clear stays NaN.
__device__ inline bf16x2 relu_bf16x2(float a, float b) {
    bf16x2 const v = __builtin_convertvector(f32x2{a, b}, bf16x2);
    i16x2 const m = __builtin_elementwise_max(__builtin_bit_cast(i16x2, v), i16x2{0, 0});
    return __builtin_bit_cast(bf16x2, m);
}

// relu of an FP32 value as the signed maximum of its bits and 0: (0 < x) ? x : +0 for every x that is not NaN
__device__ inline float relu_bits(float x) {
    int const b = __builtin_bit_cast(int, x);
    return __builtin_bit_cast(float, b > 0 ? b : 0);
}

// a load from LDS at byte address addr
template<typename T>
__device__ inline T lds_load(uint32_t addr) {
    return *reinterpret_cast<T const __attribute__((address_space(3)))*>(static_cast<uintptr_t>(addr));
}

// Lane l holds the 32 bfloat16 features of candidate l in F (word w = features 2 w, 2 w + 1). Returns in x[t] the
// B operand of column tile t (candidates 16 t .. 16 t + 15): lane 16 q + n gets features 8 q .. 8 q + 7 of
// candidate 16 t + n, the operand the feature tile in LDS gives. For each word d of an 8-feature chunk, moving word
// group q of row t to row q of operand t is a 4 x 4 transpose of (register, 16-lane row) blocks:
// v_permlane32_swap exchanges the off-diagonal 2 x 2 blocks of rows, v_permlane16_swap the single rows inside them.
__device__ inline void feature_tiles(uint32_t const (&F)[16], bf16x8 (&x)[4]) {
    u32x4 t[4];
#pragma unroll
    for(int d = 0; d < 4; ++d) {
        uint32_t r0 = F[d], r1 = F[4 + d], r2 = F[8 + d], r3 = F[12 + d];
        auto const s02 = __builtin_amdgcn_permlane32_swap(r0, r2, false, false); // rows 2, 3 of r0 <-> rows 0, 1 of r2
        auto const s13 = __builtin_amdgcn_permlane32_swap(r1, r3, false, false);
        auto const s01 = __builtin_amdgcn_permlane16_swap(s02[0], s13[0], false, false); // rows 1, 3 <-> rows 0, 2
        auto const s23 = __builtin_amdgcn_permlane16_swap(s02[1], s13[1], false, false);
        t[0][d] = s01[0];
        t[1][d] = s01[1];
        t[2][d] = s23[0];
        t[3][d] = s23[1];
    }
#pragma unroll
    for(int i = 0; i < 4; ++i) x[i] = __builtin_bit_cast(bf16x8, t[i]);
}

// LDSW: the model's weights were staged in LDS at byte address wlds (wb unused); else they are read from wb in
// device memory. Fragment reads are lane-major (lane l reads 16 bytes at (fragment group) * 1024 + 16 l), so a
// wave reads 1 KB in one piece from either. REGF (CB = 4): xin holds the four tiles' B operands, built in
// registers by feature_tiles (featWave unused); else the operands are read from the wave's feature tile featWave.
template<int CB, bool LDSW = false, bool REGF = false>
__device__ inline float value_f32w(uint8_t const* __restrict__ wb, uint16_t const* featWave, uint32_t wlds = 0,
                                   bf16x8 const* xin = nullptr) {
    static_assert(CB == 1 || CB == 2 || CB == 4);
    static_assert(!REGF || CB == 4, "the register operands hold all four tiles");
    uint32_t const lane = lane_id(), r16 = lane & 15u, q = lane >> 4;
    bf16x8 const* w1f = reinterpret_cast<bf16x8 const*>(wb + f32w::W1F);
    bf16x8 const* w2f = reinterpret_cast<bf16x8 const*>(wb + f32w::W2F);
    f32x4 const* b1 = reinterpret_cast<f32x4 const*>(wb + f32w::B1);
    f32x4 const* b2 = reinterpret_cast<f32x4 const*>(wb + f32w::B2);
    f32x4 const* w3 = reinterpret_cast<f32x4 const*>(wb + f32w::W3);
    if constexpr(LDSW) asm volatile("" : "+s"(wlds)); // re-read the staged weights for every chunk
    auto frag1 = [&](uint32_t t) -> bf16x8 {
        if constexpr(LDSW) return lds_load<bf16x8>(wlds + f32w::W1F + (t * 64 + lane) * 16);
        else {
            auto const w1l = opaque_global(w1f); // one tile's fragment in flight at a time
            return w1l[t * 64 + lane];
        }
    };
    auto frag2 = [&](uint32_t g) -> bf16x8 {
        if constexpr(LDSW) return lds_load<bf16x8>(wlds + f32w::W2F + (g * 64 + lane) * 16);
        else return w2f[g * 64 + lane];
    };
    auto vec4 = [&](f32x4 const* gp, uint32_t off, uint32_t i) -> f32x4 {
        if constexpr(LDSW) return lds_load<f32x4>(wlds + off + i * 16);
        else return gp[i];
    };
    float const b3 = LDSW ? lds_load<float>(wlds + f32w::B3) : *reinterpret_cast<float const*>(wb + f32w::B3);
    float mine = 0.0f;
#pragma unroll 1
    for(uint32_t cb = 0; cb < 4; cb += CB) {
        u32x4 h[CB][4]; // layer-2 B operands of each tile (bfloat16 pairs): K block kb = layer-1 tiles 2 kb, 2 kb + 1
        __builtin_amdgcn_s_setprio(1); // the MFMA layers at raised issue priority (the other waves fill their latency)
        {
            bf16x8 x[CB];
#pragma unroll
            for(int c = 0; c < CB; ++c)
                x[c] = REGF ? xin[c] : *reinterpret_cast<bf16x8 const*>(&featWave[feat_index(16 * (cb + c) + r16, q)]);
#pragma unroll
            for(uint32_t t = 0; t < 8; ++t) {
                bf16x8 const a = frag1(t);
                f32x4 const bias = vec4(b1, f32w::B1, 4 * t + q); // D rows 4 q + j of tile t start at b1[16 t + 4 q + j]
#pragma unroll
                for(int c = 0; c < CB; ++c) {
                    f32x4 const acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, x[c], bias, 0, 0, 0);
                    h[c][t >> 1][2 * (t & 1u)] = __builtin_bit_cast(uint32_t, relu_bf16x2(acc[0], acc[1]));
                    h[c][t >> 1][2 * (t & 1u) + 1] = __builtin_bit_cast(uint32_t, relu_bf16x2(acc[2], acc[3]));
                }
            }
        }
        float part[CB];
#pragma unroll
        for(int c = 0; c < CB; ++c) part[c] = 0.0f;
#pragma unroll 2 // lets the next tile's MFMAs overlap this tile's ReLU / output chain (0.4% on the wide generation)
        for(uint32_t t = 0; t < 8; ++t) {
            f32x4 acc[CB];
            f32x4 const bias = vec4(b2, f32w::B2, 4 * t + q);
#pragma unroll
            for(int c = 0; c < CB; ++c) acc[c] = bias;
#pragma unroll
            for(uint32_t kb = 0; kb < 4; ++kb) {
                bf16x8 const a = frag2(t * 4 + kb);
#pragma unroll
                for(int c = 0; c < CB; ++c)
                    acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, __builtin_bit_cast(bf16x8, h[c][kb]), acc[c], 0, 0, 0);
            }
            f32x4 const w = vec4(w3, f32w::W3, 4 * t + q);
#pragma unroll
            for(int c = 0; c < CB; ++c)
#pragma unroll
                for(int j = 0; j < 4; ++j) part[c] = __fmaf_rn(relu_bits(acc[c][j]), w[j], part[c]);
        }
        __builtin_amdgcn_s_setprio(0);
#pragma unroll
        for(int c = 0; c < CB; ++c) {
            float p = part[c];
            p += __shfl_xor(p, 16); // (p0 + p1), (p2 + p3)
            p += __shfl_xor(p, 32); // (p0 + p1) + (p2 + p3)
            float const v = p + b3;
            mine = q == cb + static_cast<uint32_t>(c) ? v : mine; // lane 16 ct + r16 owns candidate 16 ct + r16
        }
    }
    return mine;
}

// fit (board2.hpp:306-320) over the piece's occupied columns only: a column without a piece nibble can
// neither collide nor bound the drop. Board words past the piece's width are read but never used (their
// nibble is 0; the reads stay inside the LDS state).
__device__ inline bool fit_span(uint32_t const* parent, uint32_t cols, int col0, int& minDrop) {
    uint32_t hit = 0;
    int md = HEIGHT;
#pragma unroll
    for(int c = 0; c < 4; ++c) {
        uint32_t const nib = (cols >> (4 * c)) & 0xFu;
        uint32_t const bc = parent[col0 + c];
        hit |= bc & nib;
        int const low = bit_width(nib) - 1;
        uint32_t const mask = ~((1u << (low + 1)) - 1);
        int const drop = ctz_nz((bc & mask) | FLOOR) - low - 1;
        md = (nib != 0 && drop < md) ? drop : md;
    }
    minDrop = md;
    return hit == 0;
}

// placed (board2.hpp:322-327) then clearLines (:333-342, :218-234) into the lane's scratch board;
// returns the number of rows cleared
__device__ inline int place_into(uint32_t const* parent, uint32_t cols, int col0, int drop, uint32_t* col) {
    uint32_t full = FIELD;
#pragma unroll
    for(int x = 0; x < WIDTH; ++x) {
        uint32_t const v = parent[x] | (column_of(cols, col0, x) << drop);
        col[x * SCR_STRIDE] = v;
        full &= v;
    }
    if(full == 0) return 0;
    int const n = popc(full);
#pragma unroll 1
    for(int x = 0; x < WIDTH; ++x) {
        uint32_t c = col[x * SCR_STRIDE], rows = full;
        for(int i = 0; i < n; ++i) {
            uint32_t const row = rows & (~rows + 1); // top-most remaining full row
            c = ((c & (row - 1)) << 1) | (c & ~((row << 1) - 1));
            rows ^= row;
        }
        col[x * SCR_STRIDE] = c;
    }
    return n;
}

// placed + clearLines into a board held in registers (b[WIDTH], every loop over columns unrolled)
__device__ inline int place_regs(uint32_t const* parent, uint32_t cols, int col0, int drop, uint32_t (&b)[WIDTH]) {
    uint32_t full = FIELD;
#pragma unroll
    for(int x = 0; x < WIDTH; ++x) {
        b[x] = parent[x] | (column_of(cols, col0, x) << drop);
        full &= b[x];
    }
    if(full == 0) return 0;
    int const n = popc(full);
    // one pass per full row over all ten columns (the row's masks are computed once): the rows above the
    // top-most remaining full row move down one, the rows under it stay
    uint32_t rows = full;
    for(int i = 0; i < n; ++i) {
        uint32_t const row = rows & (~rows + 1);
        uint32_t const under = ~((row << 1) - 1);
#pragma unroll
        for(int x = 0; x < WIDTH; ++x) b[x] = (b[x] & under) | ((b[x] << 1) & ~under);
        rows ^= row;
    }
    return n;
}

__device__ inline uint32_t hist_added(uint32_t h, int cleared) { // clear_hist_t::added (model_defs.hpp:53-55)
    return cleared == 0 ? h : h + (1u << (8 * (cleared - 1)));
}

__device__ inline bool held_is_i(GameLds const& L, uint8_t heldSlot) {
    return heldSlot != HELD_NONE && L.slotPiece[heldSlot] == PIECE_I;
}

// ---------------------------------------------------------------------------------------------------
// one stage of the beam over a compacted candidate space
// ---------------------------------------------------------------------------------------------------
enum StageKind { STAGE_ROOT = 0, STAGE_MID = 1, STAGE_LEAF = 2 };

// segInfo fields: parent board (bits 0-6) | piece << 8 | heldIsI << 11 | tree child << 12 | origin << 20 |
// SEG_EXISTS | SEG_REPEAT
constexpr uint32_t SEG_EXISTS = 1u << 30; // the list has placements (every piece has at least 9)
constexpr uint32_t SEG_REPEAT = 1u << 31; // an earlier list of the stage has the same content: not scored
__device__ inline uint32_t seg_origin(uint32_t info) { return (info >> 20) & 63u; }

// Hash of kept node q's content (board, clear histogram), read from LDS, for finding equal nodes.
__device__ inline uint32_t node_hash(GameLds const& L, uint32_t q) {
    uint32_t h = L.clears[q] * 0x9E3779B1u;
#pragma unroll
    for(int x = 0; x < WIDTH; ++x) {
        h = (h ^ L.boards[q * WIDTH + x]) * 0x85EBCA6Bu;
        h ^= h >> 13;
    }
    return h ^ (h >> 16);
}

// Lanes d of `valid` whose BITS-bit key theirs(d) equals this lane's `mine`: one ballot per key bit.
template<int BITS>
__device__ inline uint64_t match_keys(uint32_t mine, uint32_t theirs, uint64_t valid) {
    uint64_t m = valid;
#pragma unroll
    for(int b = 0; b < BITS; ++b) {
        uint64_t const set = __ballot(((theirs >> b) & 1u) != 0);
        m &= ((mine >> b) & 1u) != 0 ? set : ~set;
    }
    return m;
}

// Kept nodes q and r (frontier indices) hold the same board and clear histogram.
__device__ inline bool same_nodes(GameLds const& L, uint32_t q, uint32_t r) {
    bool same = L.clears[q] == L.clears[r];
#pragma unroll
    for(int x = 0; x < WIDTH; ++x) same = same && L.boards[q * WIDTH + x] == L.boards[r * WIDTH + x];
    return same;
}

// Repeated lists. Two lists of a stage with equal parent board, parent clear histogram, piece and held-is-I have
// the same legal placements with the same values (a candidate's board, clear histogram and value are functions of
// these and the placement), which the reference computes twice (beam.hpp:337-372, :407-443): the hold plans reach
// equal positions by different orders, and equal candidates of one level are both kept and expanded again. With
// DEDUP a list is marked a repeat of the earliest list with its content (its origin, which is scored): kept nodes
// are matched by a 16-bit hash of board and clears and checked exactly (a collision falls back to a scan), then
// lists by (earliest equal node, piece, held-is-I).

// Segment << 6 | placement index of candidate f of the stage: from the candidate table when the kernel keeps one
// (seg), else from the chunk headers: lane l of chunk k is in scored list chunkList[k] + (lists starting at lanes
// 1..l of the chunk), at placement f - segOff of that list. The calling lane must be lane f % 64 (chunks are
// processed one candidate per lane).
__device__ inline uint32_t cand_of(GameLds const& L, uint16_t const* seg, uint32_t f) {
    if(seg != nullptr) return seg[f];
    uint32_t const k = f >> 6, l = f & 63u;
    uint64_t const m = L.chunkStarts[k];
    uint32_t const below = __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(m >> 32),
                                                     __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(m), 0u));
    uint32_t const s = L.listOf[L.chunkList[k] + below + static_cast<uint32_t>((m >> l) & 1ull)];
    return (s << 6) | (f - L.segOff[s]);
}

// Builds the stage's segment table in LDS. Segment s = 2k + c: child c of kept node k of buffer `cur`
// (roots: s = root group, parent = the committed board). With DEDUP (middle levels and leaves), repeated lists
// get no candidates. With D (DeepLds; middle levels and leaves), the middle levels' lists with the content of a
// previous-move list (deepUse: the roots were reused) get no candidates either; theirs follow in the candidate
// table, D->reusedTotal of them, and the level-2 and leaf lists' infos are recorded for the next move. Returns the
// count of candidates to score.
template<int STAGE, bool DEDUP = false>
__device__ inline uint32_t build_segments(GameLds& L, uint16_t* cand, uint32_t ti, int layer, uint32_t keptCount,
                                          uint32_t cur, DeepLds* D = nullptr, bool deepUse = false) {
    uint32_t const s = lane_id();
    uint32_t n = 0, info = 0;
    if(STAGE == STAGE_ROOT) {
        if(s < c_tab.treeCount[ti][0]) {
            TreeNode const node = c_tab.tree[ti][0][s];
            uint32_t const piece = L.slotPiece[node.slot];
            n = c_tab.lutOff[piece + 1] - c_tab.lutOff[piece];
            info = COMMITTED | (piece << 8) | (held_is_i(L, node.heldSlot) ? 1u << 11 : 0u) | (s << 12);
        }
    } else {
        uint32_t const k = s >> 1, c = s & 1u;
        if(k < keptCount) {
            uint32_t const pn = L.info[cur * 32 + k] & 0xFFu;
            TreeNode const pnode = c_tab.tree[ti][layer - 1][pn];
            if(c < static_cast<uint32_t>(pnode.childEnd - pnode.childBegin)) {
                uint32_t const child = pnode.childBegin + c;
                TreeNode const cn = c_tab.tree[ti][layer][child];
                uint32_t const piece = L.slotPiece[cn.slot];
                n = c_tab.lutOff[piece + 1] - c_tab.lutOff[piece];
                info = (cur * 32 + k) | (piece << 8) | (held_is_i(L, cn.heldSlot) ? 1u << 11 : 0u) | (child << 12);
            }
        }
    }
    if(n != 0) info |= SEG_EXISTS | (s << 20);
    if constexpr(DEDUP && STAGE != STAGE_ROOT) {
        // lane k < keptCount: the earliest kept node with the same board and clear histogram
        bool const node = s < keptCount;
        uint32_t const q = cur * 32 + (s & 31u);
        uint32_t const h = node ? node_hash(L, q) & 0xFFFFu : 0u;
        uint64_t const eq = match_keys<16>(h, h, __ballot(node));
        uint32_t canon = node ? static_cast<uint32_t>(__builtin_ctzll(eq)) : s; // eq holds the lane itself
        if(canon != s && !same_nodes(L, cur * 32 + canon, q)) {
            canon = s;
            for(uint32_t t = 0; t < s; ++t)
                if(same_nodes(L, cur * 32 + t, q)) {
                    canon = t;
                    break;
                }
        }
        // lane s: the earliest list with the same (node content, piece, held-is-I)
        uint32_t const key = static_cast<uint32_t>(__shfl(static_cast<int>(canon), static_cast<int>(s >> 1)))
                             | (((info >> 8) & 15u) << 5);
        uint64_t const same = match_keys<9>(key, key, __ballot(n != 0));
        uint32_t const origin = n != 0 ? static_cast<uint32_t>(__builtin_ctzll(same)) : s;
        if(origin != s) {
            info = (info & ~(63u << 20)) | (origin << 20) | SEG_REPEAT;
            n = 0;
        }
    }
    uint32_t nReused = 0; // candidates of this list taken from the previous move (DeepLds)
    if(STAGE == STAGE_MID && D != nullptr) {
        bool const lvl1 = layer == 1;
        // lane k < keptCount: kept node k as a candidate slot of the previous search one level down
        int32_t prevSlot = -1;
        if(deepUse && s < keptCount) {
            if(lvl1) { // a root: a level-1 candidate of the previous chosen root
                uint32_t const rs = L.rootSlot[s];
                prevSlot = static_cast<int32_t>((2 * D->rank + D->rootChild[rs / 34]) * 34 + rs % 34);
            } else { // a kept level-1 node: a level-2 candidate where its list has a previous-move match
                uint32_t const ks = D->kept1[s];
                int32_t const id = D->id1[ks / 34];
                prevSlot = id >= 0 ? id * 34 + static_cast<int32_t>(ks % 34) : -1;
            }
        }
        // ... and the previous search's kept node of that slot (slots are unique; 12 bits)
        uint32_t const prevCount = lvl1 ? D->n1 : D->n2;
        bool const theirValid = s < 32 && s < prevCount;
        uint32_t const theirs = theirValid ? static_cast<uint32_t>(lvl1 ? D->kept1[s] : D->kept2[s]) : 0u;
        uint64_t const hit = match_keys<12>(static_cast<uint32_t>(prevSlot) & 0xFFFu, theirs, __ballot(theirValid));
        int32_t const prevNode = prevSlot >= 0 && hit != 0 ? static_cast<int32_t>(__builtin_ctzll(hit)) : -1;
        int32_t const pn = __shfl(prevNode, static_cast<int>(s >> 1)); // this list's parent's
        // the previous node's list with this list's piece and held-is-I
        int32_t prevList = -1;
        uint32_t prevInfo = 0;
        if(pn >= 0 && (info & SEG_EXISTS) != 0) {
            uint32_t const* const pinfo = lvl1 ? D->info2 : D->infoL;
#pragma unroll
            for(uint32_t c2 = 0; c2 < 2; ++c2) {
                uint32_t const q = pinfo[2 * static_cast<uint32_t>(pn) + c2];
                if(prevList < 0 && (q & SEG_EXISTS) != 0 && ((q >> 8) & 15u) == ((info >> 8) & 15u)) {
                    prevList = 2 * pn + static_cast<int32_t>(c2);
                    prevInfo = q;
                }
            }
        }
        if(lvl1) D->id1[s] = static_cast<int8_t>(prevList);
        bool const reuse = prevList >= 0 && (info & SEG_REPEAT) == 0;
        uint64_t const prevLegal = lvl1 ? D->legal2 : D->legalL;
        uint64_t const reusedLegal = __ballot(reuse && ((prevLegal >> (static_cast<uint32_t>(prevList) & 63u)) & 1ull) != 0);
        if(reuse) { // its scores: the previous list's, or its origin's when that was a repeat
            D->src[s] = static_cast<uint8_t>((prevInfo & SEG_REPEAT) != 0 ? seg_origin(prevInfo)
                                                                          : static_cast<uint32_t>(prevList));
            nReused = n;
            n = 0;
        }
        if(s == 0) D->reusedLegal = reusedLegal;
    }
    uint32_t const off = wave_excl_scan(n);
    L.segOff[s] = off;
    L.segInfo[s] = info;
    uint32_t const total = __builtin_amdgcn_readlane(off + n, 63);
    if(STAGE != STAGE_ROOT && D != nullptr) {
        if(STAGE == STAGE_MID) { // the reused lists' candidates after the scored ones
            uint32_t const roff = wave_excl_scan(nReused);
            for(uint32_t i = 0; i < nReused; ++i) cand[total + roff + i] = static_cast<uint16_t>((s << 6) | i);
            uint32_t const rt = __builtin_amdgcn_readlane(roff + nReused, 63);
            if(s == 0) D->reusedTotal = rt;
        }
        if(layer == 2) D->info2[s] = info; // for the next move
        if(STAGE == STAGE_LEAF) D->infoL[s] = info;
    }
    if(cand != nullptr) {
        for(uint32_t i = 0; i < n; ++i) cand[off + i] = static_cast<uint16_t>((s << 6) | i);
    } else { // chunk headers (cand_of); every list has at most 34 < 64 placements
        uint32_t const j = wave_excl_scan(n != 0 ? 1u : 0u);
        if(n != 0) L.listOf[j] = static_cast<uint8_t>(s);
        if(s < (total + 63u) / 64u) L.chunkStarts[s] = 0;
        wave_sync();
        if(n != 0) {
            uint32_t const k = off >> 6;
            if((off & 63u) == 0) {
                L.chunkList[k] = static_cast<uint8_t>(j); // the list starts chunk k
            } else {
                atomicOr(reinterpret_cast<unsigned long long*>(&L.chunkStarts[k]), 1ull << (off & 63u));
                if(64u * (k + 1) < off + n) L.chunkList[k + 1] = static_cast<uint8_t>(j); // it runs into chunk k + 1
            }
        }
    }
    wave_sync();
    return total;
}


// Synchronises the G waves that cooperate on one game (one workgroup); a lone wave needs no barrier.
template<int G>
__device__ inline void group_sync() {
    if constexpr(G == 1) wave_sync();
    else __syncthreads();
}

// Scores this wave's share of a stage's candidates (chunks wv, wv + G, ...).
//  ROOT / MID: returns the exact top-K keys (score << 32 | ~slot) of the legal candidates with a
//    nonzero score (select_top, beam.hpp:244-265); MID also returns in *mask the segments with at least one
//    legal placement (a kept node has a legal continuation when one of its two does, score_mid_level,
//    beam.hpp:337-372).
//  LEAF: returns the best (score, order) key over all legal leaves and, in *mask, the segments with at
//    least one legal placement (expand_leaves, beam.hpp:407-443).
// per-game data of the wide network (its weights and the wave's feature tile) and, in trace launches, the
// leaf lists' best keys
struct WideArgs {
    uint8_t const* wb;
    uint16_t* feat;
    uint64_t* segBest;
    bool finite; // every weight of the game's model is finite (enables the exact zero skipping of value_f16)
    uint32_t wlds; // LDS byte address of the model's staged weights (wide network, plain kernel)
};

template<int KIND, int STAGE, int K, int G, bool TRACE = false, int WIDE_CB = 4, bool LDSW = false>
__device__ inline uint64_t run_stage(GameLds& L, uint16_t const* seg, uint32_t total, wptr w, uint32_t wv, uint32_t* col,
                                     uint64_t* mask, uint32_t& nLegal, WideArgs const& wide, uint32_t* keep = nullptr,
                                     uint32_t keepSlots = 0, uint64_t acc0 = 0) {
    uint32_t const lane = lane_id();
    uint64_t acc = acc0; // ROOT/MID: running top-K (acc0: keys merged before); LEAF: per-lane best key
    uint64_t ms = 0;
    constexpr bool REGF = KIND == MODEL_F32W && LDSW; // wide features stay in registers (no feature tile)
    // a scored candidate f (this lane's) of list s, placement i: counts, kept score, key into the wave's result
    auto consume = [&](uint32_t f, bool legal, uint32_t score, uint32_t s, uint32_t i) {
        nLegal += static_cast<uint32_t>(__popcll(__ballot(legal)));
        uint32_t const slot = s * 34 + i;
        if(keep != nullptr && f < total && slot < keepSlots) // scores kept for the next move (its roots, DeepLds)
            keep[slot] = legal ? score : 0u;
        if constexpr(STAGE == STAGE_LEAF) {
            uint64_t const key = legal ? ((static_cast<uint64_t>(score) << 32) | (SPAN - slot)) : 0;
            acc = key > acc ? key : acc;
            if constexpr(TRACE) {
                if(legal) atomicMax(reinterpret_cast<unsigned long long*>(&wide.segBest[s]), static_cast<unsigned long long>(key));
            }
            ms |= legal ? (1ull << s) : 0ull;
        } else {
            if(STAGE == STAGE_MID) ms |= legal ? (1ull << s) : 0ull;
            uint64_t const key = (legal && score != 0) ? ((static_cast<uint64_t>(score) << 32) | (0xFFFFFFFFu - slot)) : 0;
            acc = topk_merge<K>(acc, key);
        }
    };
    // the plain 4-wave 16-32-32-1 kernel: a model with finite weights runs its hidden layers on MFMA, one chunk at a
    // time; otherwise two of the wave's chunks (c and c + G) go through the VALU network at once
    constexpr bool PAIR = KIND == MODEL_F16 && LDSW;
    if constexpr(PAIR) {
        // this lane's candidate f: its list and placement, whether it fits, and its features when it does
        auto features_of = [&](uint32_t f, bool& legal, uint32_t& s, uint32_t& i, float (&ft)[16]) {
            if(f >= total) return;
            uint32_t const si = cand_of(L, seg, f);
            s = si >> 6;
            i = si & 63u;
            uint32_t const info = L.segInfo[s];
            uint32_t const piece = (info >> 8) & 7u;
            uint32_t const* parent = &L.boards[(info & 0x7Fu) * WIDTH];
            LutEntry const e = c_tab.lut[c_tab.lutOff[piece] + i];
            int drop = 0;
            legal = fit_span(parent, e.cols, e.col0, drop);
            if(legal) {
                uint32_t b[WIDTH];
                int const cleared = place_regs(parent, e.cols, e.col0, drop, b);
                uint32_t const clears = hist_added(STAGE == STAGE_ROOT ? 0u : L.clears[info & 0x7Fu], cleared);
                leaf_features<KIND, 1>(b, clears, ((info >> 11) & 1u) != 0, ft, nullptr);
            }
        };
        if(wide.finite) {
            for(uint32_t base = wv * 64; base < total; base += 64 * G) {
                wptr const wl = opaque(w);
                float ft[16] = {};
                bool legal = false;
                uint32_t s = 0, i = 0;
                features_of(base + lane, legal, s, i, ft);
                float const v = value_f16_mfma(reinterpret_cast<float const*>(wide.wb), wl, ft);
                consume(base + lane, legal, legal ? ordered_bits(v) : 0u, s, i);
            }
        }
        for(uint32_t base = wv * 64; !wide.finite && base < total; base += 128 * G) {
            wptr const wl = opaque(w); // wave-uniform point of the loop: reload weights per pair of chunks
            uint32_t const baseB = base + 64 * G;
            bool const hasB = baseB < total;
            float fa[16] = {}, fb[16] = {};
            bool la = false, lb = false;
            uint32_t sa = 0, ia = 0, sb = 0, ib = 0, scoreA = 0, scoreB = 0;
            features_of(base + lane, la, sa, ia, fa);
            if(hasB) {
                features_of(baseB + lane, lb, sb, ib, fb);
                float va, vb;
                value_f16x2(wl, fa, fb, la, lb, wide.finite, va, vb);
                scoreA = la ? ordered_bits(va) : 0u;
                scoreB = lb ? ordered_bits(vb) : 0u;
            } else if(la) {
                scoreA = ordered_bits(value_f16(wl, fa, wide.finite));
            }
            consume(base + lane, la, scoreA, sa, ia);
            if(hasB) consume(baseB + lane, lb, scoreB, sb, ib);
        }
    }
    for(uint32_t base = wv * 64; !PAIR && base < total; base += 64 * G) {
        wptr const wl = opaque(w); // wave-uniform point of the loop: reload weights per chunk
        uint32_t const f = base + lane;
        bool legal = false;
        uint32_t score = 0, s = 0, i = 0;
        uint32_t fpk[REGF ? 16 : 1] = {}; // REGF: the lane's 32 features as bfloat16 pairs (zero for an illegal lane)
        if(f < total) {
            uint32_t const si = cand_of(L, seg, f);
            s = si >> 6;
            i = si & 63u;
            uint32_t const info = L.segInfo[s];
            uint32_t const piece = (info >> 8) & 7u;
            uint32_t const* parent = &L.boards[(info & 0x7Fu) * WIDTH];
            LutEntry const e = c_tab.lut[c_tab.lutOff[piece] + i];
            int drop = 0;
            legal = fit_span(parent, e.cols, e.col0, drop);
            if(legal) {
                if constexpr(KIND != MODEL_F32W) { // the candidate board in registers through place, clear, features
                    uint32_t b[WIDTH];
                    int const cleared = place_regs(parent, e.cols, e.col0, drop, b);
                    uint32_t const clears = hist_added(STAGE == STAGE_ROOT ? 0u : L.clears[info & 0x7Fu], cleared);
                    float ft[16];
                    leaf_features<KIND, 1>(b, clears, ((info >> 11) & 1u) != 0, ft, nullptr);
                    if constexpr(KIND == MODEL_F16) score = ordered_bits(value_f16(wl, ft, wide.finite));
                    else score = ordered_bits(value_v11(wl, ft));
                } else if constexpr(REGF) {
                    // the 4-wave plain kernel keeps the candidate board and its features in registers
                    uint32_t b[WIDTH];
                    int const cleared = place_regs(parent, e.cols, e.col0, drop, b);
                    uint32_t const clears = hist_added(STAGE == STAGE_ROOT ? 0u : L.clears[info & 0x7Fu], cleared);
                    float ft[32];
                    leaf_features<KIND, 1>(b, clears, ((info >> 11) & 1u) != 0, ft, nullptr, ft + 16);
#pragma unroll
                    for(int k = 0; k < 16; ++k)
                        fpk[k] = __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2{ft[2 * k], ft[2 * k + 1]}), bf16x2));
                } else { // the wide network's kernels with a feature tile: the lane's LDS board
                    int const cleared = place_into(parent, e.cols, e.col0, drop, col);
                    uint32_t const clears = hist_added(STAGE == STAGE_ROOT ? 0u : L.clears[info & 0x7Fu], cleared);
                    float ft[16];
                    leaf_features<KIND>(col, clears, ((info >> 11) & 1u) != 0, ft, wide.feat);
                    store_features16(wide.feat, ft);
                }
            }
        }
        if constexpr(REGF) { // the whole wave runs the network over the chunk's 64 candidates
            bf16x8 x[4];
            feature_tiles(fpk, x);
            float const v = value_f32w<WIDE_CB, LDSW, true>(wide.wb, nullptr, wide.wlds, x);
            if(legal) score = ordered_bits(v);
        } else if constexpr(KIND == MODEL_F32W) { // the whole wave runs the network over the chunk's feature tile
            wave_sync();
            float const v = value_f32w<WIDE_CB, LDSW>(wide.wb, wide.feat, wide.wlds);
            if(legal) score = ordered_bits(v);
            wave_sync(); // the tile is rewritten by the next chunk
        }
        consume(f, legal, score, s, i);
    }
    if(STAGE != STAGE_ROOT) *mask = wave_or_u64(ms);
    return acc;
}

// merges two lane-distributed descending lists (lanes 0..31, zero padded) into the top K of their union
template<int K>
__device__ inline uint64_t merge_sorted(uint64_t a, uint64_t b) {
    uint32_t const lane = lane_id();
    uint64_t const mirrored = shfl_u64(b, static_cast<int>(63 - lane));
    uint64_t v = lane < 32 ? a : mirrored; // descending then ascending: a bitonic sequence
#pragma unroll
    for(int j = 32; j > 0; j >>= 1) {
        uint64_t const p = shfl_xor_u64(v, j);
        uint64_t const mx = v > p ? v : p;
        uint64_t const mn = v > p ? p : v;
        v = (lane & j) == 0 ? mx : mn;
    }
    return lane < static_cast<uint32_t>(K) ? v : 0;
}

// Combines the G waves' partial results of a stage; every wave returns the same values. The union of
// the per-wave exact top-K lists holds the global top-K, so the merged list is exact.
template<int G, int K, bool TOPK>
__device__ inline uint64_t combine_parts(GameLds& L, uint64_t part, uint64_t& mask, uint32_t wv) {
    static_assert(G == 1 || G == 2 || G == 4);
    if constexpr(G == 1) {
        return part;
    } else {
        uint32_t const lane = lane_id();
        if constexpr(TOPK) {
            if(lane < 32) L.partTop[wv][lane] = part;
        } else {
            part = wave_max_u64(part);
            if(lane == 0) L.partBest[wv] = part;
        }
        if(lane == 0) L.partMask[wv] = mask;
        group_sync<G>();
        uint64_t m = 0;
#pragma unroll
        for(int g = 0; g < G; ++g) m |= L.partMask[g];
        mask = m;
        if constexpr(TOPK) {
            auto part_of = [&](int g) { return lane < 32 ? L.partTop[g][lane] : uint64_t{0}; };
            uint64_t r = merge_sorted<K>(part_of(0), part_of(1));
#pragma unroll
            for(int g = 2; g + 1 < G; g += 2) r = merge_sorted<K>(r, merge_sorted<K>(part_of(g), part_of(g + 1)));
            if constexpr(G % 2 == 1) r = merge_sorted<K>(r, part_of(G - 1));
            return r;
        } else {
            uint64_t best = 0;
#pragma unroll
            for(int g = 0; g < G; ++g) best = L.partBest[g] > best ? L.partBest[g] : best;
            return best;
        }
    }
}


// Kept nodes (lanes 0..31) with a legal continuation, from the legality of their two segments.
__device__ inline uint64_t node_mask(uint64_t segMask) {
    uint32_t const lane = lane_id();
    return __ballot(lane < 32 && ((segMask >> (2 * (lane & 31u))) & 3ull) != 0);
}

// Legality of the stage's repeated lists, which were not scored: a repeat's placements fit where its origin's do.
__device__ inline uint64_t repeat_legality(GameLds const& L, uint64_t segMask) {
    uint32_t const info = L.segInfo[lane_id()];
    return segMask | __ballot((info & SEG_REPEAT) != 0 && ((segMask >> seg_origin(info)) & 1ull) != 0);
}

// Adds the candidates of a middle level's repeated lists to its exact top-K of the scored candidates. A repeat's
// candidate has the score of its origin's candidate at the same placement and a higher slot, so it ranks in the
// top K of all candidates only if that candidate ranks in the top K of the scored ones: merging, for every kept
// key, the keys of the same placement in the repeats of its list gives the top K of all candidates, as
// select_top over the whole level would (beam.hpp:244-265).
template<int K>
__device__ inline uint64_t add_repeats(GameLds const& L, uint64_t top) {
    uint32_t const info = L.segInfo[lane_id()];
    uint64_t const reps = __ballot((info & SEG_REPEAT) != 0);
    if(reps == 0) return top;
    // lane j < K holds kept key j: its list, placement and score; `mine` = the repeats of its list
    uint32_t const slot = 0xFFFFFFFFu - static_cast<uint32_t>(top);
    uint32_t const seg = slot / 34, i = slot - seg * 34;
    uint64_t mine = match_keys<6>(seg & 63u, seg_origin(info), reps);
    mine = top != 0 ? mine : 0ull;
    uint64_t const score = top & 0xFFFFFFFF00000000ull;
    while(__ballot(mine != 0) != 0) { // one repeat per kept key and round
        uint64_t key = 0;
        if(mine != 0) {
            uint32_t const d = static_cast<uint32_t>(__builtin_ctzll(mine));
            mine &= mine - 1;
            key = score | (0xFFFFFFFFu - (d * 34 + i));
        }
        top = topk_merge<K>(top, key);
    }
    return top;
}

// ---------------------------------------------------------------------------------------------------
// pruned stages: placement-proposal network on BF16 MFMA (v_mfma_f32_16x16x32_bf16, FP32 accumulate)
// ---------------------------------------------------------------------------------------------------

// 8 board bits -> 8 bfloat16 values 0.0 / 1.0
__device__ inline bf16x8 bits_bf16(uint32_t byte) {
    u32x4 v;
#pragma unroll
    for(int j = 0; j < 4; ++j)
        v[j] = (((byte >> (2 * j)) & 1u) ? 0x3F80u : 0u) | (((byte >> (2 * j + 1)) & 1u) ? 0x3F800000u : 0u);
    return __builtin_bit_cast(bf16x8, v);
}

__device__ inline bf16x8 load_bf16x8(uint16_t const* p) { return *reinterpret_cast<bf16x8 const*>(p); }

// One 16 x 16 tile of the board projection, bp[row][unit] = sum over the row board's set inputs k of N1[unit][k]:
// rows = 16 boards (this lane's A row is `board`; rowValid = false gives an empty row), columns = hidden units
// 16 nt .. + 15. v_mfma_f32_16x16x32_bf16 fragments: lane l holds A row / B column l % 16 and K elements
// 8 (l / 16) .. + 7; D element j of lane l is row 4 (l / 16) + j, column l % 16. Every partial sum is an integer
// below 2^24.
__device__ inline f32x4 project_tile(PrunerDev const* __restrict__ pr, uint32_t const* board, bool rowValid,
                                     uint32_t nt) {
    uint32_t const lane = lane_id(), q = lane >> 4;
    uint32_t const unit = 16 * nt + (lane & 15u);
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll 2
    for(uint32_t ks = 0; ks < 8; ++ks) {
        PrunerDev const* const pl = opaque_ptr(pr); // keep the weight loads inside the loop
        uint32_t const k0 = 32 * ks + 8 * q;         // board slots k0 .. k0 + 7 lie in one column
        uint32_t const c = k0 < 240 ? k0 / 24 : 0, r0 = k0 % 24;
        uint32_t const colw = board[c];
        uint32_t byte = (colw >> (3 + r0)) & (r0 == 16 ? 0xFu : 0xFFu);
        byte = rowValid && k0 < 240 ? byte : 0u;
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bits_bf16(byte), load_bf16x8(&pl->w1b[unit][k0]), acc, 0, 0, 0);
    }
    return acc;
}

// First hidden layer codes of one candidate column for this lane's MFMA B fragments (hidden units
// {8 (l / 16) .. + 7} and {32 + 8 (l / 16) .. + 7}) from its scaled board projection row and placement row (both
// always readable; valid = false gives an empty column). rne(P 2^-d) clamped to [0, 255] is exact
// (P 2^-d = projection + ptab exactly).
__device__ inline void pruner_codes(float const* bpRow, float const* ptRow, bool valid, bf16x8& h1a, bf16x8& h1b) {
    uint32_t const q = lane_id() >> 4;
    f32x4 const b0 = *reinterpret_cast<f32x4 const*>(bpRow + 8 * q);
    f32x4 const b1 = *reinterpret_cast<f32x4 const*>(bpRow + 8 * q + 4);
    f32x4 const b2 = *reinterpret_cast<f32x4 const*>(bpRow + 32 + 8 * q);
    f32x4 const b3 = *reinterpret_cast<f32x4 const*>(bpRow + 32 + 8 * q + 4);
    f32x4 const p0 = *reinterpret_cast<f32x4 const*>(ptRow + 8 * q);
    f32x4 const p1 = *reinterpret_cast<f32x4 const*>(ptRow + 8 * q + 4);
    f32x4 const p2 = *reinterpret_cast<f32x4 const*>(ptRow + 32 + 8 * q);
    f32x4 const p3 = *reinterpret_cast<f32x4 const*>(ptRow + 32 + 8 * q + 4);
    float const hi = valid ? 255.0f : 0.0f; // an empty column clamps to 0
#pragma unroll
    for(int j = 0; j < 4; ++j) {
        h1a[j] = static_cast<__bf16>(__builtin_amdgcn_fmed3f(__builtin_rintf(b0[j] + p0[j]), 0.0f, hi));
        h1a[4 + j] = static_cast<__bf16>(__builtin_amdgcn_fmed3f(__builtin_rintf(b1[j] + p1[j]), 0.0f, hi));
        h1b[j] = static_cast<__bf16>(__builtin_amdgcn_fmed3f(__builtin_rintf(b2[j] + p2[j]), 0.0f, hi));
        h1b[4 + j] = static_cast<__bf16>(__builtin_amdgcn_fmed3f(__builtin_rintf(b3[j] + p3[j]), 0.0f, hi));
    }
}

// Layers 2 and 3 for T candidate columns (MFMA B operands h1a/h1b[c]): the A fragments of output tile t
// (rows 16 t .. + 15) and its bias / W3 slice are loaded once for all T columns, and the next tile's are in flight
// while this tile's MFMAs run. Returns in every lane of column c its logit (p0 + p1) + (p2 + p3), p_q the FMA
// chain of lane quarter q over v = 16 t + 4 q + j.
template<int T>
__device__ inline void pruner_layers23(PrunerDev const* __restrict__ pr, bf16x8 const (&h1a)[T],
                                       bf16x8 const (&h1b)[T], float (&logit)[T]) {
    uint32_t const lane = lane_id(), r16 = lane & 15u, q = lane >> 4;
    float part[T];
#pragma unroll
    for(int c = 0; c < T; ++c) part[c] = 0.0f;
    bf16x8 const* fr = reinterpret_cast<bf16x8 const*>(&pr->w2f[r16][q][0]);
    f32x4 const* bw = reinterpret_cast<f32x4 const*>(&pr->bw[0][q][0]);
    constexpr uint32_t FR_STRIDE = 16 * 4 * 16 * 2 / 16; // bf16x8 per output tile (16 rows)
    constexpr uint32_t BW_STRIDE = 4 * 8 * 4 / 16;      // f32x4 per output tile
    bf16x8 a0 = fr[0], a1 = fr[1];
    f32x4 bias = bw[0], w3 = bw[1];
#pragma unroll 1
    for(uint32_t t = 0; t < 4; ++t) {
        uint32_t const nt = t < 3 ? t + 1 : t; // prefetch the next tile (the last re-reads its own)
        bf16x8 const na0 = fr[nt * FR_STRIDE], na1 = fr[nt * FR_STRIDE + 1];
        f32x4 const nbias = bw[nt * BW_STRIDE], nw3 = bw[nt * BW_STRIDE + 1];
#pragma unroll
        for(int c = 0; c < T; ++c) {
            f32x4 acc = bias; // D rows 4 q + j start at Nb2
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, h1a[c], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, h1b[c], acc, 0, 0, 0);
#pragma unroll
            for(int j = 0; j < 4; ++j) part[c] = __fmaf_rn(relu_bits(acc[j]), w3[j], part[c]); // acc: integers < 2^24
        }
        a0 = na0;
        a1 = na1;
        bias = nbias;
        w3 = nw3;
    }
#pragma unroll
    for(int c = 0; c < T; ++c) {
        part[c] += __shfl_xor(part[c], 16); // (p0 + p1), (p2 + p3)
        part[c] += __shfl_xor(part[c], 32); // (p0 + p1) + (p2 + p3)
        logit[c] = part[c];
    }
}

// Board projection of the frontier's kept nodes: bp[p][u] for p < count (32 x 64, eight tiles).
template<int G>
__device__ void pruner_board_projection(GameLds const& L, PrunerDev const* __restrict__ pr, uint32_t cur,
                                        uint32_t count, float (*bp)[PR_H], uint32_t wv) {
    uint32_t const lane = lane_id(), r16 = lane & 15u, q = lane >> 4;
#pragma unroll 1
    for(uint32_t t = wv; t < 8; t += G) {
        uint32_t const mt = t >> 2, nt = t & 3u;
        uint32_t const p = 16 * mt + r16;
        f32x4 const acc = project_tile(pr, &L.boards[(cur * 32 + (p < count ? p : 0u)) * WIDTH], p < count, nt);
        float const sc = pr->projScale; // a power of two: exact
#pragma unroll
        for(int j = 0; j < 4; ++j) bp[16 * mt + 4 * q + j][16 * nt + r16] = acc[j] * sc;
    }
}

// pruner logits are stored list-major: list s occupies prs[36 s .. 36 s + 35] (16-byte aligned rows)
constexpr uint32_t PRS_ROW = 36;

// Legality of a stage's candidates, one per lane: prs[36 s + i] = -inf for illegal placement i of list s; the
// legal ones are listed (any order, as list << 6 | placement index) in `legalList`. Returns their count.
template<int G>
__device__ uint32_t pruner_legal(GameLds& L, uint16_t const* cand, uint32_t& count, uint32_t total, float* prs,
                                 uint16_t* legalList, uint32_t wv) {
    uint32_t const lane = lane_id();
    if(wv == 0 && lane == 0) count = 0;
    group_sync<G>();
    for(uint32_t base = wv * 64; base < total; base += 64 * G) {
        uint32_t const f = base + lane;
        bool legal = false;
        if(f < total) {
            uint32_t const si = cand[f];
            uint32_t const info = L.segInfo[si >> 6];
            uint32_t const piece = (info >> 8) & 7u;
            LutEntry const e = c_tab.lut[c_tab.lutOff[piece] + (si & 63u)];
            int drop = 0;
            legal = fit_span(&L.boards[(info & 0x7Fu) * WIDTH], e.cols, e.col0, drop);
            if(!legal) prs[(si >> 6) * PRS_ROW + (si & 63u)] = -__builtin_huge_valf();
        }
        uint64_t const bal = __ballot(legal);
        uint32_t const cnt = static_cast<uint32_t>(__popcll(bal));
        uint32_t b = 0;
        if(lane == 0 && cnt != 0) b = atomicAdd(&count, cnt);
        b = uniform(b);
        if(legal)
            legalList[b + static_cast<uint32_t>(__popcll(bal & ((1ull << lane) - 1)))] =
                static_cast<uint16_t>(cand[f]);
    }
    group_sync<G>();
    return uniform(count);
}

// first placement-list entry of each piece (lutOff), one byte per piece
__device__ inline uint32_t lut_base(uint32_t piece) {
    constexpr uint64_t packed = [] {
        uint64_t v = 0;
        for(int p = 0; p < NUM_PIECES; ++p) v |= static_cast<uint64_t>(TABLES.lutOff[p]) << (8 * p);
        return v;
    }();
    return static_cast<uint32_t>(packed >> (8 * piece)) & 0xFFu;
}

// Pruner logits of the listed legal candidates, 32 per pass (two MFMA column tiles share the weight loads):
// prs[36 s + i] = logit.
template<int G>
__device__ void pruner_scores(GameLds const& L, PrunerDev const* __restrict__ pr, uint16_t const* legalList,
                              uint32_t nLegal, uint32_t cur, uint32_t ply, float const (*bp)[PR_H], float* prs,
                              uint32_t wv) {
    constexpr int T = 2;
    uint32_t const lane = lane_id(), r16 = lane & 15u;
    uint32_t const groups = (nLegal + 16 * T - 1) / (16 * T);
    for(uint32_t gi = wv; gi < groups; gi += G) {
        bf16x8 h1a[T], h1b[T];
        uint32_t f[T];
        bool valid[T];
#pragma unroll
        for(int c = 0; c < T; ++c) {
            uint32_t const t = 16 * (T * gi + c) + r16;
            valid[c] = t < nLegal;
            uint32_t const si = legalList[valid[c] ? t : 0u]; // nLegal > 0 here
            uint32_t const s = si >> 6, i = si & 63u;
            uint32_t const info = L.segInfo[s];
            f[c] = s * PRS_ROW + i;
            pruner_codes(bp[(info & 0x7Fu) - cur * 32], pr->ptab[ply - 1][lut_base((info >> 8) & 7u) + i], valid[c],
                         h1a[c], h1b[c]);
        }
        float logit[T];
        pruner_layers23<T>(pr, h1a, h1b, logit);
#pragma unroll
        for(int c = 0; c < T; ++c)
            if(lane < 16 && valid[c]) prs[f[c]] = logit[c];
    }
}

// Keeps the K best legal placements of every list by pruner logit (logit descending, placement index ascending);
// returns how many were kept and lists them in `kept` (list << 6 | placement index, any order). A lane reads its
// whole list's logits with nine 16-byte loads (lanes of one list read the same addresses).
template<int G>
__device__ uint32_t pruner_select(GameLds& L, uint16_t const* cand, uint32_t& count, uint32_t total, float const* prs,
                                  uint16_t* kept, uint32_t K, uint32_t wv) {
    uint32_t const lane = lane_id();
    if(wv == 0 && lane == 0) count = 0;
    group_sync<G>();
    for(uint32_t base = wv * 64; base < total; base += 64 * G) {
        uint32_t const f = base + lane;
        bool keep = false;
        uint32_t si = 0;
        if(f < total) {
            si = cand[f];
            uint32_t const s = si >> 6, i = si & 63u;
            f32x4 const* row = reinterpret_cast<f32x4 const*>(prs + s * PRS_ROW);
            float const pf = prs[s * PRS_ROW + i];
            if(pf != -__builtin_huge_valf()) {
                uint32_t const n = (s < 63 ? L.segOff[s + 1] : total) - L.segOff[s];
                uint32_t rank = 0;
#pragma unroll
                for(uint32_t v = 0; v < PRS_ROW / 4; ++v) {
                    f32x4 const x = row[v];
#pragma unroll
                    for(uint32_t e = 0; e < 4; ++e) {
                        uint32_t const j = 4 * v + e;
                        rank += (j < n && (x[e] > pf || (x[e] == pf && j < i))) ? 1u : 0u;
                    }
                }
                keep = rank < K;
            }
        }
        uint64_t const bal = __ballot(keep);
        uint32_t const cnt = static_cast<uint32_t>(__popcll(bal));
        uint32_t b = 0;
        if(lane == 0 && cnt != 0) b = atomicAdd(&count, cnt);
        b = uniform(b);
        if(keep) kept[b + static_cast<uint32_t>(__popcll(bal & ((1ull << lane) - 1)))] = static_cast<uint16_t>(si);
    }
    group_sync<G>();
    return uniform(count);
}

// ---------------------------------------------------------------------------------------------------
// full-beam traces for training the pruner
// ---------------------------------------------------------------------------------------------------
// every lane of the wave calls this; lanes with `want` append one record
__device__ inline void emit_record(EvalArgs const& a, bool want, uint32_t const* parent, uint32_t meta) {
    uint32_t const lane = lane_id();
    uint64_t const bal = __ballot(want);
    uint32_t const cnt = static_cast<uint32_t>(__popcll(bal));
    if(cnt == 0) return;
    uint32_t b = 0;
    if(lane == 0) b = atomicAdd(a.traceCount, cnt);
    b = uniform(b);
    uint32_t const idx = b + static_cast<uint32_t>(__popcll(bal & ((1ull << lane) - 1)));
    if(want && idx < a.traceCap) {
        uint32_t* r = a.trace + 11ull * idx;
#pragma unroll
        for(int x = 0; x < WIDTH; ++x) r[x] = parent[x];
        r[10] = meta;
    }
}

// one record per legal candidate of the stage; label = slot kept by the level (ply 1, 2) or the list's best
// leaf (ply 3)
template<int G>
__device__ void emit_stage(EvalArgs const& a, GameLds const& L, uint16_t const* cand, PruneLds const* P, uint32_t total,
                           uint32_t ply, uint32_t wv) {
    uint32_t const lane = lane_id();
    for(uint32_t base = wv * 64; base < total; base += 64 * G) {
        uint32_t const f = base + lane;
        bool legal = false;
        uint32_t meta = 0, pb = COMMITTED;
        if(f < total) {
            uint32_t const si = cand[f];
            uint32_t const s = si >> 6, i = si & 63u;
            uint32_t const info = L.segInfo[s];
            pb = info & 0x7Fu;
            uint32_t const piece = (info >> 8) & 7u;
            LutEntry const e = c_tab.lut[c_tab.lutOff[piece] + i];
            int drop = 0;
            legal = fit_span(&L.boards[pb * WIDTH], e.cols, e.col0, drop);
            uint32_t const slot = s * 34 + i;
            bool const label = ply < 3 ? ((P->survivors[slot >> 5] >> (slot & 31u)) & 1u) != 0
                                       : (SPAN - static_cast<uint32_t>(P->segBest[s])) == slot;
            meta = piece | static_cast<uint32_t>(e.ox & 3) << 3 | static_cast<uint32_t>(e.ox >> 2) << 5 | ply << 9
                 | (label ? 1u << 11 : 0u);
        }
        emit_record(a, legal, &L.boards[pb * WIDTH], meta);
    }
}

// Re-applies the selected placement of rank j (lane j < count) to its parent and writes kept node j of
// frontier buffer `dst` (seed_frontier / advance_frontier, beam.hpp:307-329, 379-403).
// REGS: the board is built in registers (col unused), for kernels without the per-lane LDS board.
template<bool REGS = false>
__device__ inline void write_frontier(GameLds& L, uint64_t key, uint32_t count, uint32_t ti, int layer, uint32_t src,
                                      uint32_t dst, bool root, uint32_t* col) {
    uint32_t const j = lane_id();
    if(j >= count) return;
    uint32_t const slot = 0xFFFFFFFFu - static_cast<uint32_t>(key);
    uint32_t const seg = slot / 34, i = slot % 34;
    uint32_t parentBoard, node, rank, parentClears;
    if(root) {
        parentBoard = COMMITTED;
        node = seg;
        rank = j;
        parentClears = 0;
    } else {
        uint32_t const k = seg >> 1, c = seg & 1u;
        parentBoard = src * 32 + k;
        uint32_t const pinfo = L.info[parentBoard];
        node = c_tab.tree[ti][layer - 1][pinfo & 0xFFu].childBegin + c;
        rank = pinfo >> 8;
        parentClears = L.clears[parentBoard];
    }
    uint32_t const piece = L.slotPiece[c_tab.tree[ti][layer][node].slot];
    LutEntry const e = c_tab.lut[c_tab.lutOff[piece] + i];
    uint32_t const* parent = &L.boards[parentBoard * WIDTH];
    int drop = 0;
    fit_span(parent, e.cols, e.col0, drop); // a selected placement always fits
    uint32_t const d = dst * 32 + j;
    int cleared;
    if constexpr(REGS) {
        uint32_t b[WIDTH];
        cleared = place_regs(parent, e.cols, e.col0, drop, b);
#pragma unroll
        for(int x = 0; x < WIDTH; ++x) L.boards[d * WIDTH + x] = b[x];
    } else {
        cleared = place_into(parent, e.cols, e.col0, drop, col);
#pragma unroll
        for(int x = 0; x < WIDTH; ++x) L.boards[d * WIDTH + x] = col[x * SCR_STRIDE];
    }
    L.clears[d] = hist_added(parentClears, cleared);
    L.score[d] = static_cast<uint32_t>(key >> 32);
    L.info[d] = node | (rank << 8);
}

// wave 0 builds the stage's segment table; every wave gets the candidate count
template<int STAGE, int G, bool DEDUP = false>
__device__ inline uint32_t stage_segments(GameLds& L, uint16_t* cand, uint32_t ti, int layer, uint32_t keptCount,
                                          uint32_t cur, uint32_t wv, DeepLds* D = nullptr, bool deepUse = false) {
    if(wv == 0) {
        uint32_t const total = build_segments<STAGE, DEDUP>(L, cand, ti, layer, keptCount, cur, D, deepUse);
        if(lane_id() == 0) L.stageTotal = total;
    }
    group_sync<G>();
    return uniform(L.stageTotal);
}

// feature tiles of the wide network, one per wave
template<int G>
struct alignas(16) FeatLds {
    uint16_t v[G][64 * 32];
};

// the wide network's weights of the game's model, in their device layout
struct alignas(16) WideWeightsLds {
    uint8_t b[f32w::BYTES];
};
static_assert(f32w::BYTES % 16 == 0, "copied as 16-byte words");


enum KernelMode { MODE_PLAIN = 0, MODE_TRACE = 1, MODE_PRUNED = 2 };
// the kernels that take lists of levels 1 and 2 from the previous move (DeepLds): the full-beam 16-32-32-1 search
// (V11's workgroups are LDS-bound at 2 waves per game, and the wide network's LDS is full)
template<int KIND, int MODE>
constexpr bool DEEP_REUSE = KIND == MODEL_F16 && MODE == MODE_PLAIN;

// Work ring of the second phase (EvalArgs::ring; thread 0 of a workgroup). Returns the next parked game, or
// RING_EMPTY when no slot is left to take. A workgroup that finds the ring empty ends: every unfinished game is then
// held by a running workgroup, which appends it and takes a slot again before it could end, so no game is left
// behind. A taken slot's writer has reserved it and publishes it at once; the wait for it is bounded all the same.
constexpr uint32_t RING_EMPTY = 0xFFFFFFFFu;
__device__ inline uint32_t ring_pop(uint32_t* ring) {
    for(uint32_t tries = 0; tries < (1u << 20); ++tries) {
        uint32_t h = __hip_atomic_load(&ring[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        uint32_t const t = __hip_atomic_load(&ring[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if(h >= t) return RING_EMPTY;
        if(__hip_atomic_compare_exchange_strong(&ring[0], &h, h + 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT)) {
            for(uint32_t w = 0; w < (1u << 26); ++w) {
                uint32_t const v = __hip_atomic_load(&ring[2 + h], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
                if(v != 0) return v - 1;
                __builtin_amdgcn_s_sleep(2);
            }
            return RING_EMPTY;
        }
    }
    return RING_EMPTY;
}

// Appends a parked game (thread 0, after its wave wrote the parked record): the release makes the record visible
// to the workgroup that takes the slot, on any XCD.
__device__ inline void ring_push(uint32_t* ring, uint32_t cap, uint32_t item) {
    uint32_t const s = atomicAdd(&ring[1], 1u);
    if(s < cap) __hip_atomic_store(&ring[2 + s], item + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// One move of the beamed search plus its commit (search_move_beam, beam.hpp:483-570; the eval_host loop
// body, model_eval_host.cpp:37-47), played by the G waves of the workgroup. Returns false when the game
// ends (no move, or a top-out). MODE_PRUNED scores only the pruner's top-k placements of every list at the
// pruned plies; MODE_TRACE records the candidates of plies 1..3 when traceMove is set.
// The wave's top-K of a middle level's lists taken from the previous move (DeepLds), which seeds the top-K of its
// scored chunks: their candidates follow the scored ones in the stage's table (chunks wv, wv + G, ...), their keys
// come from the previous move's scores (`prev`, at the list that holds them), and `keep` gets their scores for the
// next move as a scored candidate's. Taken first, these keys (the previous chosen root's best continuations) raise
// the threshold that most scored chunks then fail without a sort: the 16-32-32-1 generation ran 1.2% faster than
// with these keys merged after the scored chunks.
template<int K, int G>
__device__ inline uint64_t deep_topk(uint16_t const* cand, DeepLds const& D, uint32_t const* prev, uint32_t total,
                                     uint32_t* keep, uint32_t keepSlots, uint32_t wv) {
    uint32_t const rt = uniform(D.reusedTotal);
    uint32_t const lane = lane_id();
    uint64_t acc = 0;
    for(uint32_t base = wv * 64; base < rt; base += 64 * G) {
        uint32_t const f = base + lane;
        uint64_t key = 0;
        if(f < rt) {
            uint32_t const si = cand[total + f];
            uint32_t const s = si >> 6, i = si & 63u, slot = s * 34 + i;
            uint32_t const score = prev[D.src[s] * 34u + i];
            if(keep != nullptr && slot < keepSlots) keep[slot] = score;
            key = score != 0 ? ((static_cast<uint64_t>(score) << 32) | (0xFFFFFFFFu - slot)) : 0ull;
        }
        acc = topk_merge<K>(acc, key);
    }
    return acc;
}

// The root stage from the previous move's scores (RootCacheMeta): when every root list matches one of the chosen
// root's level-1 lists, each wave builds the exact top-K of the root keys from them (no scoring) and returns true.
template<int K>
__device__ inline bool root_reuse(GameLds const& L, uint16_t const* cand, RootCacheMeta const& RC,
                                  uint32_t const* level1, uint32_t total, uint64_t& top) {
    if(uniform(RC.valid) == 0) return false;
    uint32_t const lane = lane_id();
    uint32_t const info = L.segInfo[lane & 1u]; // lanes 0, 1: the root lists
    uint32_t const key = ((info >> 8) & 15u) | 16u;
    uint32_t const c = RC.nextKey[0] == key ? 0u : (RC.nextKey[1] == key ? 1u : 2u);
    if(__ballot(lane < 2 && (info & SEG_EXISTS) != 0 && c == 2u) != 0) return false;
    uint32_t const src0 = uniform(RC.src[__builtin_amdgcn_readlane(c, 0) & 1u]);
    uint32_t const src1 = uniform(RC.src[__builtin_amdgcn_readlane(c, 1) & 1u]);
    uint64_t acc = 0;
    for(uint32_t base = 0; base < total; base += 64) {
        uint32_t const f = base + lane;
        uint64_t k = 0;
        if(f < total) {
            uint32_t const si = cand_of(L, cand, f);
            uint32_t const s = si >> 6, i = si & 63u;
            uint32_t const score = level1[(s == 0 ? src0 : src1) * 34 + i];
            k = score != 0 ? ((static_cast<uint64_t>(score) << 32) | (0xFFFFFFFFu - (s * 34 + i))) : 0ull;
        }
        acc = topk_merge<K>(acc, k);
    }
    top = acc;
    return true;
}

template<int KIND, int G, int MODE>
__device__ bool play_move(EvalArgs const& a, GameLds& L, uint16_t* cand, PruneLds* P, RootCacheMeta* RC,
                          uint32_t* level1, DeepLds* D, wptr w, WideArgs const& wide, uint32_t wv, uint32_t* col,
                          bool traceMove, uint64_t& nApply, uint32_t& nLegal, uint64_t& nSlots) {
    uint32_t const lane = lane_id();
    constexpr int CB = MODE == MODE_PLAIN ? 4 : 2; // wide network: column tiles per weight load (register budget)
    // candidate boards in registers: the FP32 networks' kernels and the wide network's 4-wave plain kernel
    constexpr bool REG_BOARDS = KIND != MODEL_F32W || (MODE == MODE_PLAIN && G == 4);
    // repeated lists of the middle levels and the leaves are not scored (their results are their origins'; a
    // repeat's pruner logits are its origin's too); trace launches record every candidate of every list
    constexpr bool DEDUP = MODE != MODE_TRACE;
    // roots from the previous move's level-1 scores (RootCacheMeta; a pruned search whose level 1 is pruned keeps
    // only some of its scores and takes none)
    constexpr bool ROOT_CACHE = MODE != MODE_TRACE;
    // lists of levels 1 and 2 from the previous move, when its roots are reused (DeepLds)
    constexpr bool DEEP = DEEP_REUSE<KIND, MODE>;
    constexpr uint32_t LEVEL1_SLOTS = RC_ROOTS<KIND> * 2 * 34, DEEP_SLOTS = 64 * 34;
    // gather the committed state (beam.hpp:503-505)
    if(wv == 0) {
        if(lane < WIDTH) L.boards[COMMITTED * WIDTH + lane] = L.eng.cols[lane];
        if(lane == 0) {
            L.clears[COMMITTED] = 0;
            L.slotPiece[0] = L.eng.held;
            for(int i = 0; i < 5; ++i) L.slotPiece[1 + i] = L.eng.queue[i];
            L.treeIdx = L.eng.held == NO_PIECE ? 1u : 0u;
        }
    }
    group_sync<G>();
    uint32_t const ti = uniform(L.treeIdx);
    // the per-lane candidate-board scratch doubles as the pruner logits of a stage (2,176 floats)
    float* const prs = MODE == MODE_PRUNED ? reinterpret_cast<float*>(col - lane - wv * (WIDTH * SCR_STRIDE)) : nullptr;

    // stage 0: roots -> top K0
    uint32_t total = stage_segments<STAGE_ROOT, G>(L, cand, ti, 0, 0, 0, wv);
    uint64_t mask = 0;
    uint64_t top = 0;
    bool reused = false;
    if constexpr(ROOT_CACHE) reused = root_reuse<K0>(L, cand, *RC, level1, total, top);
    if(!reused) {
        nApply += total;
        top = run_stage<KIND, STAGE_ROOT, K0, G, false, CB, MODE == MODE_PLAIN && G == 4>(L, cand, total, w, wv, col, nullptr, nLegal, wide);
        nSlots += (total + 63u) & ~63u; // lanes the stage's chunks occupied
        top = combine_parts<G, K0, true>(L, top, mask, wv);
    }
    // a narrower beam keeps the first beamWidth[0] keys of the exact, descending top-K0 list
    top = lane < a.beamWidth[0] ? top : 0;
    uint32_t count = __popcll(__ballot(top != 0));
    if(count == 0) return false;
    if(wv == 0) {
        if(lane < K0) L.rootSlot[lane] = 0xFFFFFFFFu - static_cast<uint32_t>(top);
        if constexpr(DEEP) { // root list g is the previous chosen root's level-1 list of child rootChild[g]
            uint32_t const key = ((L.segInfo[lane & 1u] >> 8) & 15u) | 16u;
            if(lane < 2) D->rootChild[lane] = RC->nextKey[0] == key ? 0u : 1u;
        }
        write_frontier<REG_BOARDS>(L, top, count, ti, 0, 0, 0, true, col);
    }
    group_sync<G>();

    uint64_t bestKey = 0;
    uint32_t cur = 0;
    // middle levels 1 and 2 -> top K1, K2
#pragma unroll
    for(int level = 1; level <= 2; ++level) {
        total = stage_segments<STAGE_MID, G, DEDUP>(L, cand, ti, level, count, cur, wv, DEEP ? D : nullptr, reused);
        nApply += total;
        uint32_t* keep = nullptr;
        uint32_t keepSlots = 0;
        if constexpr(ROOT_CACHE) {
            if(level == 1) {
                if(wv == 0 && lane < 32) RC->info1[lane] = L.segInfo[lane];
                keep = level1;
                keepSlots = LEVEL1_SLOTS;
            }
        }
        if constexpr(DEEP) {
            if(level == 2) {
                keep = D->s2;
                keepSlots = DEEP_SLOTS;
            }
        }
        uint16_t const* seg = cand;
        uint32_t evalTotal = total;
        if constexpr(MODE == MODE_PRUNED) {
            if((a.pruneMask >> level) & 1u) {
                uint32_t const nl = pruner_legal<G>(L, cand, P->keptCount, total, prs, P->kept, wv);
                pruner_board_projection<G>(L, a.pruner, cur, count, P->bp, wv);
                group_sync<G>();
                pruner_scores<G>(L, a.pruner, P->kept, nl, cur, static_cast<uint32_t>(level), P->bp, prs, wv);
                group_sync<G>();
                evalTotal = pruner_select<G>(L, cand, P->keptCount, total, prs, P->kept, a.pruneK, wv);
                seg = P->kept;
            }
        }
        mask = 0;
        // the lists taken from the previous move (their scores: its level 2 / its leaves) seed the top-K
        uint64_t first = 0;
        if constexpr(DEEP)
            first = level == 1 ? deep_topk<K1, G>(cand, *D, D->s2, total, keep, keepSlots, wv)
                               : deep_topk<K2, G>(cand, *D, D->sL, total, keep, keepSlots, wv);
        top = level == 1 ? run_stage<KIND, STAGE_MID, K1, G, false, CB, MODE == MODE_PLAIN && G == 4>(L, seg, evalTotal, w, wv, col, &mask, nLegal, wide, keep, keepSlots, first)
                         : run_stage<KIND, STAGE_MID, K2, G, false, CB, MODE == MODE_PLAIN && G == 4>(L, seg, evalTotal, w, wv, col, &mask, nLegal, wide, keep, keepSlots, first);
        if constexpr(DEEP) // the taken lists' legality (read here: no register holds it across the scoring)
            if(wv == 0) mask |= D->reusedLegal;
        nSlots += (evalTotal + 63u) & ~63u;
        top = combine_parts<G, K2, true>(L, top, mask, wv);
        if constexpr(DEDUP) {
            mask = repeat_legality(L, mask);
            top = add_repeats<K2>(L, top);
        }
        if constexpr(DEEP) {
            if(level == 2 && wv == 0 && lane == 0) D->legal2 = mask; // for the next move
        }
        top = lane < a.beamWidth[level] ? top : 0; // narrower beam: the first beamWidth[level] keys
        mask = node_mask(mask);
        // kept nodes without a legal continuation become leaves at this depth, keyed by root rank
        uint64_t fb = 0;
        if(lane < count && ((mask >> lane) & 1ull) == 0) {
            uint32_t const rank = L.info[cur * 32 + lane] >> 8;
            fb = (static_cast<uint64_t>(L.score[cur * 32 + lane]) << 32) | (SPAN - (LEAF_ORDERS + rank));
        }
        fb = wave_max_u64(fb);
        bestKey = uniform64(fb > bestKey ? fb : bestKey); // wave-uniform: kept in SGPRs
        uint32_t const next = __popcll(__ballot(top != 0));
        if constexpr(MODE == MODE_TRACE) {
            if(traceMove) {
                if(wv == 0) {
                    for(uint32_t i = lane; i < 68; i += 64) P->survivors[i] = 0;
                    wave_sync();
                    if(lane < next) {
                        uint32_t const slot = 0xFFFFFFFFu - static_cast<uint32_t>(top);
                        atomicOr(&P->survivors[slot >> 5], 1u << (slot & 31u));
                    }
                }
                group_sync<G>();
                emit_stage<G>(a, L, cand, P, total, static_cast<uint32_t>(level), wv);
            }
        }
        if(wv == 0) {
            if constexpr(DEEP) { // the level's kept slots, for the next level and the next move
                if(lane < 32) (level == 1 ? D->kept1 : D->kept2)[lane] =
                    static_cast<uint16_t>(0xFFFFFFFFu - static_cast<uint32_t>(top));
                if(lane == 0) (level == 1 ? D->n1 : D->n2) = next;
            }
            write_frontier<REG_BOARDS>(L, top, next, ti, level, cur, cur ^ 1u, false, col);
        }
        group_sync<G>();
        cur ^= 1u;
        count = next;
    }

    // leaves
    total = stage_segments<STAGE_LEAF, G, DEDUP>(L, cand, ti, 3, count, cur, wv, DEEP ? D : nullptr);
    nApply += total;
    uint16_t const* seg = cand;
    uint32_t evalTotal = total;
    if constexpr(MODE == MODE_PRUNED) {
        if((a.pruneMask >> 3) & 1u) {
            uint32_t const nl = pruner_legal<G>(L, cand, P->keptCount, total, prs, P->kept, wv);
            pruner_board_projection<G>(L, a.pruner, cur, count, P->bp, wv);
            group_sync<G>();
            pruner_scores<G>(L, a.pruner, P->kept, nl, cur, 3u, P->bp, prs, wv);
            group_sync<G>();
            evalTotal = pruner_select<G>(L, cand, P->keptCount, total, prs, P->kept, a.pruneK, wv);
            seg = P->kept;
        }
    }
    mask = 0;
    uint64_t leaf;
    if constexpr(MODE == MODE_TRACE) {
        if(traceMove) {
            if(wv == 0) P->segBest[lane] = 0;
            group_sync<G>();
            leaf = run_stage<KIND, STAGE_LEAF, 1, G, true, CB>(L, seg, evalTotal, w, wv, col, &mask, nLegal, wide);
        } else {
            leaf = run_stage<KIND, STAGE_LEAF, 1, G, false, CB, MODE == MODE_PLAIN && G == 4>(L, seg, evalTotal, w, wv, col, &mask, nLegal, wide);
        }
    } else {
        leaf = run_stage<KIND, STAGE_LEAF, 1, G, false, CB, MODE == MODE_PLAIN && G == 4>(L, seg, evalTotal, w, wv, col, &mask, nLegal, wide,
                                                                                     DEEP ? D->sL : nullptr, DEEP_SLOTS);
    }
    nSlots += (evalTotal + 63u) & ~63u;
    leaf = combine_parts<G, 1, false>(L, leaf, mask, wv);
    if constexpr(DEDUP) mask = repeat_legality(L, mask); // a repeat's leaves never beat its origin's (higher orders)
    if constexpr(DEEP) {
        if(wv == 0 && lane == 0) D->legalL = mask; // for the next move
    }
    if constexpr(MODE == MODE_TRACE) {
        if(traceMove) emit_stage<G>(a, L, cand, P, total, 3u, wv);
    }
    {
        // a list where no placement fits scores the frontier board itself at order base (beam.hpp:439-440)
        uint32_t const s = lane;
        uint64_t fb = 0;
        if((L.segInfo[s] & SEG_EXISTS) != 0 && ((mask >> s) & 1ull) == 0) {
            uint32_t const k = s >> 1;
            fb = (static_cast<uint64_t>(L.score[cur * 32 + k]) << 32) | (SPAN - s * 34);
        }
        leaf = fb > leaf ? fb : leaf;
    }
    leaf = wave_max_u64(leaf);
    bestKey = leaf > bestKey ? leaf : bestKey;

    // decode_move (beam.hpp:450-470) and commit on one lane
    uint32_t const order = SPAN - static_cast<uint32_t>(bestKey);
    int cleared = -1;
    if(wv == 0 && lane == 0) {
        uint32_t const rank = order >= LEAF_ORDERS ? order - LEAF_ORDERS : (L.info[cur * 32 + order / 68] >> 8);
        uint32_t const rootSlot = L.rootSlot[rank];
        TreeNode const rn = c_tab.tree[ti][0][rootSlot / 34];
        uint32_t const piece = L.slotPiece[rn.slot];
        LutEntry const e = c_tab.lut[c_tab.lutOff[piece] + rootSlot % 34];
        int const o = e.ox & 3, x = (e.ox >> 2) - 2;
        EngineState& eng = L.eng;
        if constexpr(MODE != MODE_PLAIN) // the item's move log, one byte per placement (not kept by the plain kernels)
            if(a.moves != nullptr && eng.placements < a.maxMoves)
                a.moves[static_cast<uint64_t>(L.moveItem) * a.maxMoves + eng.placements] =
                    static_cast<uint8_t>((rn.rootHold ? 0x80u : 0u) | (e.ox & 3u) << 5 | (e.ox >> 2));
        if(rn.rootHold) engine_hold(eng);
        cleared = engine_place(eng, o, x);
        ++eng.placements;
        if(cleared >= 0) stats_advance(eng, cleared);
        else eng.gameOver = 1;
    }
    if constexpr(ROOT_CACHE) { // the chosen root's level-1 lists become the next move's root lists (RootCacheMeta)
        if(wv == 0 && lane == 0) {
            uint32_t const rank = order >= LEAF_ORDERS ? order - LEAF_ORDERS : (L.info[cur * 32 + order / 68] >> 8);
            bool const valid = rank < RC_ROOTS<KIND> && cleared == 0
                               && (MODE != MODE_PRUNED || ((a.pruneMask >> 1) & 1u) == 0);
#pragma unroll
            for(uint32_t c = 0; c < 2; ++c) {
                uint32_t const li = RC->info1[2 * rank + c];
                RC->nextKey[c] = (li & SEG_EXISTS) != 0 ? (((li >> 8) & 15u) | 16u) : 0u;
                // a repeat's origin is an earlier list, of a better root, so also kept when valid
                RC->src[c] = !valid ? c : ((li & SEG_REPEAT) != 0 ? seg_origin(li) : 2 * rank + c);
            }
            RC->valid = valid ? 1u : 0u;
            if constexpr(DEEP) D->rank = rank;
        }
    }
    group_sync<G>();
    return uniform(L.eng.gameOver) == 0;
}

// ---------------------------------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------------------------------
// Waves per SIMD the register allocation is held to (measured fastest): V11 and 16-32-32-1 at 4 (128 VGPRs, no
// spills; V11 at 5 waves per SIMD, 96 VGPRs, spilled 92 B per lane once the repeated lists and the root scores
// of the previous move were added, and was 15% slower; 16-32-32-1 with the candidate board in registers was 5%
// faster at 4 than at 5 with spills). The trace and pruned kernels' LDS caps them at 4 workgroups of 4 waves per
// CU, so they get the 128 registers of 4 waves per SIMD; one wave per game needs more registers (3 waves per
// SIMD).
template<int KIND, int MODE, int G>
constexpr int WAVES_PER_SIMD = G == 1 ? 3 : (KIND == MODEL_F32W ? 3 : 4);
// (the wide network at 3: four column tiles per weight load need ~168 VGPRs, and beat 4 waves/SIMD with two; its
// 4-wave plain kernel, with the model's weights staged in LDS, fits 3 workgroups per CU in LDS)

// One workgroup of G waves plays one (model, game) item at a time; in persistent mode it then pulls the
// next item from the device work queue.
template<int KIND, int G, int MODE = MODE_PLAIN>
__global__ __launch_bounds__(64 * G) __attribute__((amdgpu_waves_per_eu(WAVES_PER_SIMD<KIND, MODE, G>, 8)))
void beam_eval_kernel(EvalArgs a) {
    static_assert(MODE == MODE_PLAIN || G * WIDTH * SCR_STRIDE >= 64 * PRS_ROW, "pruner logits share the scratch");
    // the wide network's weights, staged once per item (plain kernel): every chunk reads its fragments from LDS,
    // and the candidate boards and features stay in registers (no per-lane board, no feature tile), which leaves
    // ~52 KB of LDS per workgroup: 3 workgroups per CU
    constexpr bool WIDE_LDS = KIND == MODEL_F32W && MODE == MODE_PLAIN && G == 4;
    __shared__ GameLds L;
    // one candidate board per lane (the wide network's kernels with a feature tile) or the pruner's logits
    constexpr bool SCRATCH = (KIND == MODEL_F32W && !WIDE_LDS) || MODE == MODE_PRUNED;
    __shared__ __attribute__((aligned(16))) uint32_t scratch[SCRATCH ? G * WIDTH * SCR_STRIDE : 1];
    __shared__ std::conditional_t<MODE == MODE_PLAIN, uint32_t, PruneLds> pruneStore;
    // the candidate table (the wide network's plain kernel uses the chunk headers of GameLds instead)
    constexpr bool CAND_TABLE = !(KIND == MODEL_F32W && MODE == MODE_PLAIN);
    __shared__ std::conditional_t<CAND_TABLE, CandLds, uint32_t> candStore;
    uint16_t* const cand = CAND_TABLE ? reinterpret_cast<CandLds*>(&candStore)->seg : nullptr;
    PruneLds* const pruneLds = MODE == MODE_PLAIN ? nullptr : reinterpret_cast<PruneLds*>(&pruneStore);
    // wide network's feature tiles (kernels that keep the candidate boards in LDS)
    __shared__ std::conditional_t<KIND == MODEL_F32W && !WIDE_LDS, FeatLds<G>, uint32_t> featStore;
    __shared__ std::conditional_t<WIDE_LDS, WideWeightsLds, uint32_t> wideWeights;
    // the previous move's root scores (RootCacheMeta, level-1 scores in LDS)
    constexpr bool ROOT_CACHE = MODE != MODE_TRACE;
    __shared__ std::conditional_t<ROOT_CACHE, RootCacheMeta, uint32_t> rcMeta;
    __shared__ std::conditional_t<ROOT_CACHE, RootScoresLds<KIND>, uint32_t> rcScores;
    RootCacheMeta* const RC = ROOT_CACHE ? reinterpret_cast<RootCacheMeta*>(&rcMeta) : nullptr;
    uint32_t* const level1 = ROOT_CACHE ? reinterpret_cast<RootScoresLds<KIND>*>(&rcScores)->s : nullptr;
    // the previous move's level-2 and leaf lists (16-32-32-1 full beam)
    constexpr bool DEEP = DEEP_REUSE<KIND, MODE>;
    __shared__ std::conditional_t<DEEP, DeepLds, uint32_t> deepStore;
    DeepLds* const D = DEEP ? reinterpret_cast<DeepLds*>(&deepStore) : nullptr;
    constexpr uint32_t PW = park_words(KIND);
    uint32_t const wv = uniform(threadIdx.x >> 6);
    uint32_t const lane = lane_id();
    uint32_t* const col = SCRATCH ? scratch + wv * (WIDTH * SCR_STRIDE) + lane : nullptr;
    uint32_t item = a.persistent ? 0u : blockIdx.x;
    for(;;) {
        if(a.persistent) {
            // second phase of a two-phase schedule: the parked games, in turns, from the work ring
            if(threadIdx.x == 0) L.item = a.resumePhase ? ring_pop(a.ring) : atomicAdd(a.queue, 1u);
            group_sync<G>();
            item = uniform(L.item);
        }
        if(item >= a.items) break;
        // integer division runs on the VALU; re-declare the results wave-uniform (SGPRs)
        uint32_t const m = uniform(item / a.games), g = uniform(item % a.games);
        wptr const w = (wptr)(a.weights + static_cast<uint64_t>(m) * a.weightStride);
        WideArgs wide{reinterpret_cast<uint8_t const*>(a.weights + static_cast<uint64_t>(m) * a.weightStride), nullptr,
                      MODE == MODE_TRACE ? pruneLds->segBest : nullptr, false};
        if constexpr(KIND == MODEL_F32W && !WIDE_LDS) wide.feat = reinterpret_cast<FeatLds<G>*>(&featStore)->v[wv];
        if constexpr(WIDE_LDS) { // all waves copy the model's 42,512 bytes; the group sync below publishes them
            uint4 const* const src = reinterpret_cast<uint4 const*>(a.weights + static_cast<uint64_t>(m) * a.weightStride);
            uint4* const dst = reinterpret_cast<uint4*>(&wideWeights);
            for(uint32_t k = threadIdx.x; k < f32w::BYTES / 16; k += 64 * G) dst[k] = src[k];
            wide.wlds = static_cast<uint32_t>(
                reinterpret_cast<uintptr_t>((uint8_t __attribute__((address_space(3)))*)(&wideWeights.b[0])));
        }
        if constexpr(KIND == MODEL_F16) { // each wave checks that the model's weights are all finite
            uint32_t const* const wbits =
                reinterpret_cast<uint32_t const*>(a.weights + static_cast<uint64_t>(m) * a.weightStride);
            bool inf = false;
            for(uint32_t k = lane; k <= static_cast<uint32_t>(F16_B3); k += 64)
                inf |= (wbits[k] & 0x7F800000u) == 0x7F800000u;
            wide.finite = __ballot(inf) == 0;
        }
        uint32_t const* const park = a.parked + static_cast<uint64_t>(item) * PW;
        uint64_t nApply = 0, nSlots = 0;
        uint32_t nLegal = 0, mv = 0;
        if constexpr(MODE != MODE_PLAIN)
            if(threadIdx.x == 0) L.moveItem = item;
        if(a.resumePhase) { // a parked game: its engine state, counters and move index
            if(wv == 0 && lane < ENGINE_WORDS) reinterpret_cast<uint32_t*>(&L.eng)[lane] = park[lane];
            nApply = uniform(static_cast<uint64_t>(park[ENGINE_WORDS]) | static_cast<uint64_t>(park[ENGINE_WORDS + 1]) << 32);
            nLegal = wv == 0 ? uniform(park[ENGINE_WORDS + 2]) : 0u;
            mv = uniform(park[ENGINE_WORDS + 3]);
            nSlots = uniform(static_cast<uint64_t>(park[ENGINE_WORDS + 4]) | static_cast<uint64_t>(park[ENGINE_WORDS + 5]) << 32);
            if constexpr(ROOT_CACHE) { // the root scores carried in the record: lists 0 and 1 of level1
                if(wv == 0) {
                    if(lane < 3) (&RC->valid)[lane] = park[PARK_RC + lane]; // valid, nextKey[2]
                    if(lane < 2) RC->src[lane] = lane;
                    for(uint32_t k = lane; k < 2 * 34; k += 64) level1[k] = park[PARK_RC + 3 + k];
                }
            }
            if constexpr(DEEP) { // the deeper lists, when the roots may be reused
                if(wv == 0 && uniform(park[PARK_RC]) != 0) {
                    uint32_t* const d = reinterpret_cast<uint32_t*>(D);
                    for(uint32_t k = lane; k < DEEP_RECORD_WORDS; k += 64) d[k] = park[PARK_WORDS + k];
                }
            }
        } else if(threadIdx.x == 0) {
            if(a.timeline) a.timeline[2 * static_cast<uint64_t>(item)] = wall_clock64();
            engine_reset(L.eng, a.seeds[g]);
            if(a.garbage) add_garbage_rows(L.eng, a.seeds[g]);
            if constexpr(ROOT_CACHE) RC->valid = 0; // the first move has no previous search
            if constexpr(DEEP) D->n1 = D->n2 = 0;
        }
        group_sync<G>();
        // traced moves: index congruent to the game seed modulo traceEvery (spreads them over the games)
        uint32_t const tracePhase =
            MODE == MODE_TRACE && a.traceEvery != 0 ? uniform(static_cast<uint32_t>(a.seeds[g] % a.traceEvery)) : 0u;
        // first phase: phaseMoves moves; second phase: a turn of phaseMoves more
        uint32_t const limit = a.phaseMoves == 0 || !a.persistent ? a.maxMoves
                               : !a.resumePhase                   ? (a.phaseMoves < a.maxMoves ? a.phaseMoves : a.maxMoves)
                               : (a.maxMoves - mv > a.phaseMoves ? mv + a.phaseMoves : a.maxMoves);
        bool ended = false;
        for(; mv < limit; ++mv) {
            if(uniform(L.eng.gameOver) != 0 || !play_move<KIND, G, MODE>(a, L, cand, pruneLds, RC, level1, D, w, wide, wv, col,
                                                                         MODE == MODE_TRACE && a.traceEvery != 0
                                                                             && mv % a.traceEvery == tracePhase,
                                                                         nApply, nLegal, nSlots)) {
                ended = true;
                break;
            }
        }
        if(lane == 0) L.partLegal[wv] = nLegal;
        group_sync<G>();
        if(!ended && limit < a.maxMoves) { // park the game for the second phase or its next turn
            // (wave 0 writes the whole record, so the release of ring_push, in the same wave, covers it)
            uint32_t* const out = a.parked + static_cast<uint64_t>(item) * PW;
            if(wv == 0 && lane < ENGINE_WORDS) out[lane] = reinterpret_cast<uint32_t const*>(&L.eng)[lane];
            if constexpr(ROOT_CACHE) {
                if(wv == 0) {
                    if(lane < 3) out[PARK_RC + lane] = (&RC->valid)[lane];
                    for(uint32_t k = lane; k < 2 * 34; k += 64)
                        out[PARK_RC + 3 + k] = level1[RC->src[k / 34] * 34 + k % 34];
                }
            }
            if constexpr(DEEP) {
                if(wv == 0 && uniform(RC->valid) != 0) {
                    uint32_t const* const d = reinterpret_cast<uint32_t const*>(D);
                    for(uint32_t k = lane; k < DEEP_RECORD_WORDS; k += 64) out[PARK_WORDS + k] = d[k];
                }
            }
            if(threadIdx.x == 0) {
                uint32_t legal = 0;
                for(int k = 0; k < G; ++k) legal += L.partLegal[k];
                out[ENGINE_WORDS] = static_cast<uint32_t>(nApply);
                out[ENGINE_WORDS + 1] = static_cast<uint32_t>(nApply >> 32);
                out[ENGINE_WORDS + 2] = legal;
                out[ENGINE_WORDS + 3] = mv;
                out[ENGINE_WORDS + 4] = static_cast<uint32_t>(nSlots);
                out[ENGINE_WORDS + 5] = static_cast<uint32_t>(nSlots >> 32);
                if(a.resumePhase) ring_push(a.ring, a.ringCap, item);
                else a.resumeList[atomicAdd(a.resumeCount, 1u)] = item;
            }
        } else if(threadIdx.x == 0) {
            a.fitness[item] = final_score(L.eng);
            if(a.placements) a.placements[item] = L.eng.placements;
            if(a.counters) {
                uint64_t legal = 0;
                for(int k = 0; k < G; ++k) legal += L.partLegal[k];
                a.counters[3 * static_cast<uint64_t>(item)] = nApply;
                a.counters[3 * static_cast<uint64_t>(item) + 1] = legal;
                a.counters[3 * static_cast<uint64_t>(item) + 2] = nSlots;
            }
            if(a.timeline) a.timeline[2 * static_cast<uint64_t>(item) + 1] = wall_clock64();
        }
        group_sync<G>();
        if(!a.persistent) break;
    }
}

// Leaf values of given boards, one board per lane, exactly as the search scores a candidate board (features of
// the board, then the network). Weights in device layout.
template<int KIND>
__global__ __launch_bounds__(64) void model_value_kernel(float const* __restrict__ weights,
                                                         uint32_t const* __restrict__ boards,
                                                         uint32_t const* __restrict__ clears,
                                                         uint8_t const* __restrict__ held, uint32_t n,
                                                         float* __restrict__ out) {
    __shared__ __attribute__((aligned(16))) uint32_t scratch[WIDTH * SCR_STRIDE];
    __shared__ std::conditional_t<KIND == MODEL_F32W, FeatLds<1>, uint32_t> featStore;
    uint32_t const lane = lane_id();
    uint32_t const b = blockIdx.x * 64 + lane;
    uint32_t const src = b < n ? b : 0u; // the last block's spare lanes recompute board 0
    uint32_t* const col = scratch + lane;
    for(int x = 0; x < WIDTH; ++x) col[x * SCR_STRIDE] = boards[10ull * src + x];
    bool const heldI = held[src] != 0;
    float v;
    if constexpr(KIND == MODEL_F32W) {
        uint16_t* const feat = reinterpret_cast<FeatLds<1>*>(&featStore)->v[0];
        float f[16];
        leaf_features<KIND>(col, clears[src], heldI, f, feat);
        store_features16(feat, f);
        wave_sync();
        v = value_f32w<4>(reinterpret_cast<uint8_t const*>(weights), feat);
    } else {
        v = leaf_value<KIND>(col, clears[src], heldI, (wptr)weights, false);
    }
    if(b < n) out[b] = v;
}

__global__ void reset_counter_kernel(uint32_t* q) {
    if(threadIdx.x == 0) q[0] = 0;
}

// Clears the work ring's slots past the parked games (before order_resume_kernel fills the first ones).
__global__ void ring_clear_kernel(uint32_t* ring, uint32_t const* count, uint32_t cap) {
    uint32_t const n = *count;
    for(uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < cap; s += gridDim.x * blockDim.x)
        if(s >= n) ring[2 + s] = 0;
}

// Between the two phases (one workgroup): counts the parked games of every model, then writes the parked list
// into the work ring in descending order of that count (a counting sort on min(count, ORDER_BINS - 1); the order
// inside a bin is whatever the atomics give, which changes no result, only the order of work).
constexpr uint32_t ORDER_BINS = 1024;
__global__ void __launch_bounds__(ORDER_BINS) order_resume_kernel(uint32_t const* list, uint32_t const* count,
                                                                   uint32_t* ring, uint32_t* modelParked,
                                                                   uint32_t models, uint32_t games) {
    __shared__ uint32_t bins[ORDER_BINS];
    uint32_t const t = threadIdx.x, n = *count;
    if(t == 0) {
        ring[0] = 0; // head
        ring[1] = n; // tail
    }
    for(uint32_t m = t; m < models; m += ORDER_BINS) modelParked[m] = 0;
    bins[t] = 0;
    __syncthreads();
    for(uint32_t i = t; i < n; i += ORDER_BINS) atomicAdd(&modelParked[list[i] / games], 1u);
    __threadfence_block();
    __syncthreads();
    auto bin_of = [&](uint32_t item) { // most parked games -> bin 0
        uint32_t const c = __hip_atomic_load(&modelParked[item / games], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return ORDER_BINS - 1 - (c < ORDER_BINS - 1 ? c : ORDER_BINS - 1);
    };
    for(uint32_t i = t; i < n; i += ORDER_BINS) atomicAdd(&bins[bin_of(list[i])], 1u);
    __syncthreads();
    if(t == 0) { // exclusive prefix sum over the bins (1024 adds)
        uint32_t run = 0;
        for(uint32_t b = 0; b < ORDER_BINS; ++b) {
            uint32_t const c = bins[b];
            bins[b] = run;
            run += c;
        }
    }
    __syncthreads();
    for(uint32_t i = t; i < n; i += ORDER_BINS) {
        uint32_t const item = list[i];
        ring[2 + atomicAdd(&bins[bin_of(item)], 1u)] = item + 1;
    }
}

// Device self-check of the bit intrinsics the board code relies on (device_canary.hpp:16); writes 1 on
// success. Mirrors bits.hpp:27-58: ctz(0) = 32, bit_width(0) = 0.
__global__ void bits_canary_kernel(uint32_t* ok) {
    if(threadIdx.x != 0) return;
    bool good = true;
    uint32_t const vals[6] = {1u, 2u, 0x80000000u, 0x00F000u, 0x7FFFFFu, 0xFFFFFFFFu};
    int const ctzs[6] = {0, 1, 31, 12, 0, 0};
    int const pops[6] = {1, 1, 1, 4, 23, 32};
    int const widths[6] = {1, 2, 32, 16, 23, 32};
    for(int i = 0; i < 6; ++i) {
        good = good && ctz_nz(vals[i]) == ctzs[i] && popc(vals[i]) == pops[i] && bit_width(vals[i]) == widths[i];
    }
    good = good && bit_width(0u) == 0 && ctz_nz(0u | FLOOR) == 23;
    ok[0] = good ? 1u : 0u;
}

// Evaluates the normalised features exactly as the search computes them, over integer numerators
// 0..n-1, so a test can compare them bit for bit with host IEEE arithmetic.
__global__ void feature_norm_kernel(float* out, uint32_t n) {
    uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= n) return;
    int const v = static_cast<int>(i);
    out[6 * i + 0] = i <= static_cast<uint32_t>(SV_MAX) ? c_svNorm[i] : 0.0f;
    out[6 * i + 1] = div20(v);
    out[6 * i + 2] = div200(v);
    out[6 * i + 3] = div5(v);
    out[6 * i + 4] = div40(v);
    out[6 * i + 5] = div100(v);
}

template<int KIND>
hipError_t launch_kind(EvalArgs const& a, int wpb, uint32_t blocks, hipStream_t stream) {
    // pruned and trace kernels (the latter also for a move log alone) run 4 waves per game
    if(a.pruner != nullptr || a.trace != nullptr || a.moves != nullptr) {
        if(wpb != 4) return hipErrorInvalidValue;
        if(a.pruner != nullptr)
            hipLaunchKernelGGL((beam_eval_kernel<KIND, 4, MODE_PRUNED>), dim3(blocks), dim3(256), 0, stream, a);
        else
            hipLaunchKernelGGL((beam_eval_kernel<KIND, 4, MODE_TRACE>), dim3(blocks), dim3(256), 0, stream, a);
        return hipGetLastError();
    }
    switch(wpb) {
        case 1: hipLaunchKernelGGL((beam_eval_kernel<KIND, 1>), dim3(blocks), dim3(64), 0, stream, a); break;
        case 2: hipLaunchKernelGGL((beam_eval_kernel<KIND, 2>), dim3(blocks), dim3(128), 0, stream, a); break;
        case 4: hipLaunchKernelGGL((beam_eval_kernel<KIND, 4>), dim3(blocks), dim3(256), 0, stream, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}


// Integer-grid pruner logits of trace records, 16 per wave (the pruned kernel's projection and layer-2 code on
// rows taken from memory); a record whose (rotation, column) is no placement of its piece gets NaN.
__global__ __launch_bounds__(64) void pruner_infer_kernel(PrunerDev const* __restrict__ pr,
                                                          uint32_t const* __restrict__ samples, uint32_t n,
                                                          float* __restrict__ out) {
    __shared__ float bp[16][PR_H];
    uint32_t const lane = lane_id(), r16 = lane & 15u, q = lane >> 4;
    uint32_t const base = blockIdx.x * 16;
    uint32_t const t = base + r16;
    uint32_t const* rec = samples + 11ull * (t < n ? t : 0u);
#pragma unroll 1
    for(uint32_t nt = 0; nt < 4; ++nt) {
        f32x4 const acc = project_tile(pr, rec, t < n, nt);
        float const sc = pr->projScale;
#pragma unroll
        for(int j = 0; j < 4; ++j) bp[4 * q + j][16 * nt + r16] = acc[j] * sc;
    }
    __syncthreads();
    float const* ptRow = pr->ptab[0][0];
    bool found = false;
    if(t < n) {
        uint32_t const meta = rec[10];
        uint32_t const piece = meta & 7u, ox = ((meta >> 3) & 3u) | (((meta >> 5) & 15u) << 2);
        uint32_t const ply = (meta >> 9) & 3u;
        if(piece < 7 && ply >= 1 && ply <= 3) {
            for(uint32_t li = c_tab.lutOff[piece]; li < c_tab.lutOff[piece + 1]; ++li) {
                if(c_tab.lut[li].ox == ox) {
                    found = true;
                    ptRow = pr->ptab[ply - 1][li];
                }
            }
        }
    }
    bf16x8 h1a[1], h1b[1];
    pruner_codes(bp[r16], ptRow, found, h1a[0], h1b[0]);
    float logit[1];
    pruner_layers23<1>(pr, h1a, h1b, logit);
    if(lane < 16 && t < n) out[t] = found ? logit[0] : __builtin_nanf("");
}
} // namespace

hipError_t launch_model(int kind, EvalArgs const& a, int wpb, uint32_t blocks, hipStream_t stream);

hipError_t launch_beam_eval(int kind, EvalArgs const& a, int wpb, uint32_t blocks, hipStream_t stream) {
    if(a.persistent) {
        hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(64), 0, stream, a.queue);
        hipError_t const e = hipGetLastError();
        if(e != hipSuccess) return e;
    }
    if(a.trace != nullptr) {
        hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(64), 0, stream, a.traceCount);
        hipError_t const e = hipGetLastError();
        if(e != hipSuccess) return e;
    }
    if(a.phaseMoves != 0 && a.persistent) { // two phases: all items to phaseMoves, then the parked games
        hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(64), 0, stream, a.resumeCount);
        hipError_t e = hipGetLastError();
        if(e != hipSuccess) return e;
        EvalArgs b = a;
        b.resumePhase = 0;
        e = launch_model(kind, b, wpb, blocks, stream);
        if(e != hipSuccess) return e;
        hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(64), 0, stream, a.queue);
        if((e = hipGetLastError()) != hipSuccess) return e;
        hipLaunchKernelGGL(ring_clear_kernel, dim3(256), dim3(256), 0, stream, a.ring, a.resumeCount, a.ringCap);
        if((e = hipGetLastError()) != hipSuccess) return e;
        hipLaunchKernelGGL(order_resume_kernel, dim3(1), dim3(ORDER_BINS), 0, stream, a.resumeList, a.resumeCount,
                           a.ring, a.modelParked, a.items / a.games, a.games);
        if((e = hipGetLastError()) != hipSuccess) return e;
        b.resumePhase = 1;
        return launch_model(kind, b, wpb, blocks, stream);
    }
    return launch_model(kind, a, wpb, blocks, stream);
}

hipError_t launch_model(int kind, EvalArgs const& a, int wpb, uint32_t blocks, hipStream_t stream) {
    switch(kind) {
        case MODEL_V11: return launch_kind<MODEL_V11>(a, wpb, blocks, stream);
        case MODEL_F16: return launch_kind<MODEL_F16>(a, wpb, blocks, stream);
        case MODEL_F32W: return launch_kind<MODEL_F32W>(a, wpb, blocks, stream);
        default: return hipErrorInvalidValue;
    }
}

namespace {
template<int KIND>
hipError_t occupancy_kind(int wpb, int* n, int mode) {
    if(mode == MODE_PRUNED) return hipOccupancyMaxActiveBlocksPerMultiprocessor(n, beam_eval_kernel<KIND, 4, MODE_PRUNED>, 256, 0);
    if(mode == MODE_TRACE) return hipOccupancyMaxActiveBlocksPerMultiprocessor(n, beam_eval_kernel<KIND, 4, MODE_TRACE>, 256, 0);
    switch(wpb) {
        case 1: return hipOccupancyMaxActiveBlocksPerMultiprocessor(n, beam_eval_kernel<KIND, 1>, 64, 0);
        case 2: return hipOccupancyMaxActiveBlocksPerMultiprocessor(n, beam_eval_kernel<KIND, 2>, 128, 0);
        case 4: return hipOccupancyMaxActiveBlocksPerMultiprocessor(n, beam_eval_kernel<KIND, 4>, 256, 0);
        default: return hipErrorInvalidValue;
    }
}
} // namespace

hipError_t launch_model_values(int kind, float const* weights, uint32_t const* boards, uint32_t const* clears,
                               uint8_t const* held, uint32_t n, float* out, hipStream_t s) {
    if(n == 0) return hipSuccess;
    dim3 const grid((n + 63) / 64);
    switch(kind) {
        case MODEL_V11: hipLaunchKernelGGL(model_value_kernel<MODEL_V11>, grid, dim3(64), 0, s, weights, boards, clears, held, n, out); break;
        case MODEL_F16: hipLaunchKernelGGL(model_value_kernel<MODEL_F16>, grid, dim3(64), 0, s, weights, boards, clears, held, n, out); break;
        case MODEL_F32W: hipLaunchKernelGGL(model_value_kernel<MODEL_F32W>, grid, dim3(64), 0, s, weights, boards, clears, held, n, out); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t pruner_infer_samples(PrunerDev const* pr, uint32_t const* samples, uint32_t n, float* out, hipStream_t s) {
    if(n == 0) return hipSuccess;
    hipLaunchKernelGGL(pruner_infer_kernel, dim3((n + 15) / 16), dim3(64), 0, s, pr, samples, n, out);
    return hipGetLastError();
}

hipError_t max_active_blocks(int kind, int wavesPerBlock, int mode, int* blocksPerCU) {
    switch(kind) {
        case MODEL_V11: return occupancy_kind<MODEL_V11>(wavesPerBlock, blocksPerCU, mode);
        case MODEL_F16: return occupancy_kind<MODEL_F16>(wavesPerBlock, blocksPerCU, mode);
        case MODEL_F32W: return occupancy_kind<MODEL_F32W>(wavesPerBlock, blocksPerCU, mode);
        default: return hipErrorInvalidValue;
    }
}

// Once per process and device (later calls return at once), so that an evaluator created while another one's
// kernels run on the same device never rewrites a table those kernels read.
hipError_t upload_constants() {
    static std::mutex mu;
    static bool done[MAX_DEVICES] = {}; // devices whose table is in place (a fixed table: no host allocation)
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if(e != hipSuccess) return e;
    if(dev < 0 || dev >= MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> const lock(mu);
    if(done[dev]) return hipSuccess;
    // fsqrt (v_board_metrics.hpp:23-37) then the /100 normaliser (:229), both IEEE on the host
    float lut[SV_MAX + 1];
    for(int sv = 0; sv <= SV_MAX; ++sv) {
        float const x = static_cast<float>(sv);
        float const r = x <= 0.0f ? 0.0f : std::sqrt(x);
        lut[sv] = r / 100.0f;
    }
    e = hipMemcpyToSymbol(HIP_SYMBOL(c_svNorm), lut, sizeof(lut), 0, hipMemcpyHostToDevice);
    if(e == hipSuccess) done[dev] = true;
    return e;
}

hipError_t launch_bits_canary(uint32_t* ok, hipStream_t stream) {
    hipLaunchKernelGGL(bits_canary_kernel, dim3(1), dim3(64), 0, stream, ok);
    return hipGetLastError();
}

hipError_t launch_feature_norms(float* out, uint32_t n, hipStream_t stream) {
    hipError_t const e = upload_constants();
    if(e != hipSuccess) return e;
    hipLaunchKernelGGL(feature_norm_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, out, n);
    return hipGetLastError();
}

} // namespace ta3h
