/*
 * ta3_oracle.cpp -- CPU oracle: a scalar C++ restatement of the reference's CPU evaluation path.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker the HIP evaluator is compared against; only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it. Nothing in the product
 * (tetris_ai3_amd/) includes, links or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to the reference root).
 * Deliberate generalisations over the reference, all reducing to it at the reference defaults:
 *  - the beam depth and widths are runtime values (reference: compile-time BEAM_WIDTHS {16,32,32},
 *    ta3_ai/incl/ta3/ai/search/variation_sequences.hpp:22) so the reference's own
 *    "full-width beam == exhaustive search" check (ta3_ai/tests/src/search/beam_test.cpp:113-141)
 *    can run at depths where it is feasible;
 *  - the value model is selected at runtime: the reference model_t = ModelV11
 *    (ta3_ai/incl/ta3/ai/model.hpp:44), the build-defined 16->32->32->1 network, and the
 *    height_model test double of ta3_ai/tests/src/search/beam_test.cpp:20-35.
 *
 * Floating point: compile with -ffp-contract=off; every multiply and add is rounded separately, as in
 * the reference's host build (ta3_ai/incl/ta3/ai/models/models/utility.hpp:39-44).
 */
#include "ta3_oracle.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <memory>
#include <unordered_set>
#include <vector>

namespace {

constexpr int WIDTH = 10, HEIGHT = 23;                      // tetris_defs.hpp:9-10
constexpr uint32_t FIELD = (1u << HEIGHT) - 1;              // board2.hpp:154
constexpr uint32_t FLOOR = 1u << HEIGHT;                    // board2.hpp:157
constexpr int NUM_PIECES = 7, NO_PIECE = 7, NUM_ORI = 4;    // piece_defs.hpp:19-31, 129-136
constexpr int PIECE_I = 0;
constexpr int MAX_PLACEMENTS = 34;                          // placements.hpp:223-231 (pinned by search_test.cpp:60-67)
constexpr int QUEUE = 6;                                    // tetris_defs.hpp:12
constexpr int MAX_DEPTH = 5;                                // variation_sequences.hpp:33: DEPTH + 1 <= PIECE_QUEUE_SIZE
constexpr uint8_t HELD_NONE = 7;                            // variation_sequences.hpp:137

// bits.hpp:27-58: ctz(0) = 32, bit_width(0) = 0
inline int ctz32(uint32_t v) { return v == 0 ? 32 : __builtin_ctz(v); }
inline int popc32(uint32_t v) { return __builtin_popcount(v); }
inline int bitw32(uint32_t v) { return v == 0 ? 0 : 32 - __builtin_clz(v); }

// ---------------------------------------------------------------------------------------------------
// pieces: block coordinates {x, y} per (type, orientation), given in the reference's
// set(type, BOTTOM, TOP, LEFT, RIGHT) argument order (ta3_sim/incl/ta3/sim/pieces/piece_offsets.hpp:19-85)
// ---------------------------------------------------------------------------------------------------
struct Cell { int x, y; };
using Blocks = Cell[4];

struct PieceTables {
    Cell blocks[NUM_PIECES][NUM_ORI][4];
    uint16_t shapeCols[NUM_PIECES][NUM_ORI]; // nibble-packed occupied columns (piece_offsets.hpp:186-201)
    int8_t shapeLeft[NUM_PIECES][NUM_ORI];
    uint32_t box[NUM_PIECES][NUM_ORI][4];    // full 4-wide column masks (piece_offsets.hpp:112-119)

    PieceTables() {
        auto set = [&](int t, std::initializer_list<Cell> bottom, std::initializer_list<Cell> top,
                       std::initializer_list<Cell> left, std::initializer_list<Cell> right) {
            auto put = [&](int o, std::initializer_list<Cell> c) {
                int i = 0;
                for(auto const& v : c) blocks[t][o][i++] = v;
            };
            put(0, top);    // TOP
            put(1, right);  // RIGHT
            put(2, bottom); // BOTTOM
            put(3, left);   // LEFT
        };
        // J, L, I, S, T, Z, O -- order of the reference table; enum values I=0 O=1 T=2 S=3 Z=4 J=5 L=6
        set(5, {{0, 1}, {1, 1}, {2, 1}, {2, 2}}, {{0, 0}, {0, 1}, {1, 1}, {2, 1}},
               {{1, 0}, {1, 1}, {0, 2}, {1, 2}}, {{1, 0}, {2, 0}, {1, 1}, {1, 2}});
        set(6, {{0, 1}, {1, 1}, {2, 1}, {0, 2}}, {{2, 0}, {0, 1}, {1, 1}, {2, 1}},
               {{0, 0}, {1, 0}, {1, 1}, {1, 2}}, {{1, 0}, {1, 1}, {1, 2}, {2, 2}});
        set(0, {{0, 2}, {1, 2}, {2, 2}, {3, 2}}, {{0, 1}, {1, 1}, {2, 1}, {3, 1}},
               {{1, 0}, {1, 1}, {1, 2}, {1, 3}}, {{2, 0}, {2, 1}, {2, 2}, {2, 3}});
        set(3, {{1, 1}, {2, 1}, {1, 2}, {0, 2}}, {{1, 0}, {2, 0}, {0, 1}, {1, 1}},
               {{1, 0}, {1, 1}, {2, 1}, {2, 2}}, {{0, 0}, {0, 1}, {1, 1}, {1, 2}});
        set(2, {{0, 1}, {1, 1}, {2, 1}, {1, 2}}, {{1, 0}, {0, 1}, {1, 1}, {2, 1}},
               {{1, 0}, {0, 1}, {1, 1}, {1, 2}}, {{1, 0}, {1, 1}, {2, 1}, {1, 2}});
        set(4, {{1, 1}, {0, 1}, {1, 2}, {2, 2}}, {{0, 0}, {1, 0}, {1, 1}, {2, 1}},
               {{1, 0}, {0, 1}, {1, 1}, {0, 2}}, {{2, 0}, {1, 1}, {2, 1}, {1, 2}});
        set(1, {{1, 0}, {2, 0}, {1, 1}, {2, 1}}, {{1, 0}, {2, 0}, {1, 1}, {2, 1}},
               {{1, 0}, {2, 0}, {1, 1}, {2, 1}}, {{1, 0}, {2, 0}, {1, 1}, {2, 1}});

        for(int t = 0; t < NUM_PIECES; ++t)
            for(int o = 0; o < NUM_ORI; ++o) {
                for(int c = 0; c < 4; ++c) box[t][o][c] = 0;
                for(auto const& b : blocks[t][o]) box[t][o][b.x] |= 1u << b.y;
                // piece_columns2 (piece_offsets.hpp:132-143): trim empty box columns
                int left = 0;
                while(left < 4 && box[t][o][left] == 0) ++left;
                int right = 4;
                while(right > left && box[t][o][right - 1] == 0) --right;
                uint16_t cols = 0;
                for(int c = left; c < right; ++c) cols = static_cast<uint16_t>(cols | (box[t][o][c] << (4 * (c - left))));
                shapeCols[t][o] = cols;
                shapeLeft[t][o] = static_cast<int8_t>(left);
            }
    }
};
const PieceTables PT;

// piece_offsets.hpp:157-160
struct Shape { uint16_t cols; int col0; };
inline Shape shape_at(int t, int o, int x) { return {PT.shapeCols[t][o], x + PT.shapeLeft[t][o]}; } // :206-209
inline int shape_width(uint16_t cols) { return (bitw32(cols) + 3) / 4; }                             // :163

// ---------------------------------------------------------------------------------------------------
// Board2 (ta3_sim/incl/ta3/sim/board2.hpp)
// ---------------------------------------------------------------------------------------------------
struct Board {
    uint32_t c[WIDTH];
};

inline uint32_t column_of(Shape s, int x) { // board2.hpp:165-168
    auto const d = static_cast<uint32_t>(x - s.col0);
    return d < 4 ? (static_cast<uint32_t>(s.cols) >> (4 * d)) & 0xFu : 0u;
}
inline int column_top(uint32_t col) { return ctz32(col | FLOOR); } // :174
inline int height(Board const& b, int x) { return HEIGHT - column_top(b.c[x]); } // :127-129

struct Fit { int drop; bool legal; };
inline Fit fit(Board const& b, Shape s) { // board2.hpp:306-320
    uint32_t hit = 0;
    int minDrop = HEIGHT;
    for(int x = 0; x < WIDTH; ++x) {
        uint32_t const nib = column_of(s, x);
        hit |= b.c[x] & nib;
        int const low = bitw32(nib) - 1;
        uint32_t const mask = ~((1u << (low + 1)) - 1);
        int const drop = column_top(b.c[x] & mask) - low - 1;
        if(nib != 0) minDrop = std::min(minDrop, drop);
    }
    return {minDrop, hit == 0};
}

inline Board placed(Board const& b, Shape s, int y) { // :322-327
    Board n;
    for(int x = 0; x < WIDTH; ++x) n.c[x] = b.c[x] | (column_of(s, x) << y);
    return n;
}

inline uint32_t full_rows(Board const& b) { // :211-216
    uint32_t m = FIELD;
    for(int x = 0; x < WIDTH; ++x) m &= b.c[x];
    return m;
}

inline uint32_t clear_line(uint32_t col, uint32_t row) { // :218-222
    uint32_t const above = col & (row - 1);
    uint32_t const below = col & ~((row << 1) - 1);
    return (above << 1) | below;
}

inline int clear_lines(Board& b) { // :333-342 with :224-234
    uint32_t const full = full_rows(b);
    if(full == 0) return 0;
    for(int x = 0; x < WIDTH; ++x) {
        uint32_t col = b.c[x], rows = full;
        int const n = popc32(full);
        for(int i = 0; i < n; ++i) {
            uint32_t const row = rows & (~rows + 1); // top-most (lowest set bit)
            col = clear_line(col, row);
            rows ^= row;
        }
        b.c[x] = col;
    }
    return popc32(full);
}

inline bool available_shape(Board const& b, Shape s, int y) { // :241-250
    uint32_t hit = 0;
    for(int x = 0; x < WIDTH; ++x) {
        uint32_t const shifted = column_of(s, x) << y;
        hit |= shifted & ~FIELD;
        hit |= b.c[x] & shifted;
    }
    return hit == 0;
}

inline bool available(Board const& b, int t, int o, int ox, int oy) { // :252-256 with :236-239
    Shape const s = shape_at(t, o, ox);
    int const w = shape_width(s.cols);
    bool const inBounds = 0 <= s.col0 && s.col0 + w <= WIDTH && 0 <= oy && oy < HEIGHT;
    return inBounds && available_shape(b, s, oy);
}

inline int drop_distance(Board const& b, Shape s, int y) { // :262-275
    int minDrop = HEIGHT;
    for(int x = 0; x < WIDTH; ++x) {
        uint32_t const nib = column_of(s, x);
        int const low = y + bitw32(nib) - 1;
        uint32_t const mask = ~((1u << (low + 1)) - 1);
        int const drop = column_top(b.c[x] & mask) - low - 1;
        if(nib != 0) minDrop = std::min(minDrop, drop);
    }
    return minDrop;
}

inline void place_shape(Board& b, Shape s, int y) { // :301-304
    for(int x = 0; x < WIDTH; ++x) b.c[x] |= column_of(s, x) << y;
}

inline int holes(Board const& b) { // :344-353
    int sum = 0;
    for(int x = 0; x < WIDTH; ++x) sum += HEIGHT - column_top(b.c[x]) - popc32(b.c[x]);
    return sum;
}

inline Board canonical(Board const& b) { // :132-143: the lexicographically smaller of board and mirror
    Board f;
    for(int x = 0; x < WIDTH; ++x) f.c[x] = b.c[WIDTH - 1 - x];
    return std::lexicographical_compare(f.c, f.c + WIDTH, b.c, b.c + WIDTH) ? f : b;
}

// ---------------------------------------------------------------------------------------------------
// xoshiro256** (ta3_sim/incl/ta3/sim/utility/xoshiro256ss.hpp:20-48)
// ---------------------------------------------------------------------------------------------------
inline uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

void rng_seed(uint64_t s, uint64_t st[4]) {
    for(int i = 0; i < 4; ++i) {
        s += 0x9e3779b97f4a7c15ULL;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
        st[i] = z ^ (z >> 31);
    }
}

uint64_t rng_next(uint64_t st[4]) {
    uint64_t const result = rotl64(st[1] * 5, 7) * 9;
    uint64_t const t = st[1] << 17;
    st[2] ^= st[0];
    st[3] ^= st[1];
    st[1] ^= st[2];
    st[0] ^= st[3];
    st[2] ^= t;
    st[3] = rotl64(st[3], 45);
    return result;
}

// ---------------------------------------------------------------------------------------------------
// TetrisEngine (ta3_sim/incl/ta3/sim/tetris_engine.hpp:103-161)
// ---------------------------------------------------------------------------------------------------
struct Engine {
    Board board;
    uint32_t queue[QUEUE];
    uint32_t held;
    bool gameOver;
    uint32_t bag[NUM_PIECES];
    uint32_t bagPos;
    uint64_t rng[4];
};

uint32_t draw_piece(Engine& e) { // :116-129
    if(e.bagPos > 0) return e.bag[--e.bagPos];
    for(uint32_t i = 0; i < NUM_PIECES; ++i) e.bag[i] = i;
    for(uint32_t i = NUM_PIECES - 1; i > 0; --i) {
        uint64_t const j = rng_next(e.rng) % (i + 1);
        std::swap(e.bag[i], e.bag[j]);
    }
    e.bagPos = NUM_PIECES - 1;
    return e.bag[NUM_PIECES - 1];
}

void update_game_over(Engine& e) { e.gameOver = !available(e.board, e.queue[0], 0, 3, 0); } // :159-161, SPAWN {3,0}

void engine_reset(Engine& e, uint64_t seed) { // :103-114
    std::memset(&e.board, 0, sizeof(e.board));
    e.held = NO_PIECE;
    e.gameOver = false;
    e.bagPos = 0;
    rng_seed(seed, e.rng);
    for(int i = 0; i < QUEUE; ++i) e.queue[i] = draw_piece(e);
    update_game_over(e);
}

// Build-defined start-position extension: 0-4 garbage rows at the bottom, each full but for one hole, drawn
// from a second xoshiro256** seeded with seed ^ 0x6A09E667F3BCC908 (so the piece sequence is unchanged):
// n = next() % 5; then for rows y = 22, 21, ... (n of them): hole column = next() % 10.
constexpr uint64_t GARBAGE_SALT = 0x6A09E667F3BCC908ull;
void add_garbage_rows(Engine& e, uint64_t seed) {
    uint64_t r[4];
    rng_seed(seed ^ GARBAGE_SALT, r);
    int const n = static_cast<int>(rng_next(r) % 5);
    for(int k = 0; k < n; ++k) {
        int const hole = static_cast<int>(rng_next(r) % WIDTH);
        int const y = HEIGHT - 1 - k;
        for(int x = 0; x < WIDTH; ++x)
            if(x != hole) e.board.c[x] |= 1u << y;
    }
    update_game_over(e);
}
int g_garbage = 0; // process-wide option of the oracle: start every game with garbage rows

void advance_queue(Engine& e) { // :131-135
    for(int i = 1; i < QUEUE; ++i) e.queue[i - 1] = e.queue[i];
    e.queue[QUEUE - 1] = draw_piece(e);
}

int engine_place(Engine& e, int o, int x) { // :137-146; -1 = DIED
    uint32_t const t = e.queue[0];
    Shape const s = shape_at(t, o, x);
    int const y = drop_distance(e.board, s, 0); // dropLocation(type, o, {x, 0})
    place_shape(e.board, s, y);
    int const cleared = clear_lines(e.board);
    advance_queue(e);
    update_game_over(e);
    return e.gameOver ? -1 : cleared;
}

void engine_hold(Engine& e) { // :148-157
    if(e.held == NO_PIECE) {
        e.held = e.queue[0];
        advance_queue(e);
    } else
        std::swap(e.held, e.queue[0]);
    update_game_over(e);
}

// ---------------------------------------------------------------------------------------------------
// placement LUT (ta3_ai/incl/ta3/ai/search/placements.hpp:70-205)
// ---------------------------------------------------------------------------------------------------
struct Placement { int o, x; Shape shape; };

struct Lut {
    uint8_t data[256];
    int offset[NUM_PIECES + 1];

    Lut() {
        int w = 0;
        for(int p = 0; p < NUM_PIECES; ++p) {
            offset[p] = w;
            // representatives (:142-164): lowest orientation per vertically-normalised column signature
            uint32_t seen[4][4];
            int seenN[4];
            int nSeen = 0;
            for(int o = 0; o < NUM_ORI; ++o) {
                int const left = PT.shapeLeft[p][o];
                int const n = shape_width(PT.shapeCols[p][o]);
                uint32_t cols[4] = {0, 0, 0, 0};
                uint32_t minbit = 32;
                for(int c = 0; c < n; ++c) {
                    cols[c] = PT.box[p][o][left + c];
                    if(cols[c] != 0) minbit = std::min<uint32_t>(minbit, static_cast<uint32_t>(ctz32(cols[c])));
                }
                if(minbit == 32) minbit = 0;
                for(int c = 0; c < n; ++c) cols[c] >>= minbit;
                bool dup = false;
                for(int j = 0; j < nSeen && !dup; ++j) {
                    if(seenN[j] != n) continue;
                    bool same = true;
                    for(int c = 0; c < n; ++c) same = same && seen[j][c] == cols[c];
                    dup = same;
                }
                if(dup) continue;
                for(int c = 0; c < 4; ++c) seen[nSeen][c] = cols[c];
                seenN[nSeen++] = n;
                // orientation_range (:70-75): x in [-left, WIDTH - left - ncols]
                for(int x = -left; x <= WIDTH - left - n; ++x)
                    data[w++] = static_cast<uint8_t>((o & 3) | (((x + 2) & 0xF) << 2)); // encode (:83-86)
            }
        }
        offset[NUM_PIECES] = w;
    }
    int count(int p) const { return offset[p + 1] - offset[p]; }
    Placement nth(int p, int i) const { // decode (:88-92)
        uint8_t const b = data[offset[p] + i];
        int const o = b & 3, x = ((b >> 2) & 0xF) - 2;
        return {o, x, shape_at(p, o, x)};
    }
};
const Lut LUT;

// ---------------------------------------------------------------------------------------------------
// hold plans and plan trees (variation_sequences.hpp:216-255; beam.hpp:69-141)
// ---------------------------------------------------------------------------------------------------
struct Plan {
    uint8_t slot[MAX_DEPTH];
    uint8_t held[MAX_DEPTH];
    bool rootHold;
};

std::vector<Plan> build_plans(int depth, bool heldEmpty) {
    std::vector<Plan> plans;
    for(uint32_t d = 0; d < (1u << depth); ++d) {
        uint8_t cur = 1, hld = 0; // win_slot(0) = 1, SEQ_HELD_SLOT = 0
        bool empty = heldEmpty;
        uint8_t k = 1;
        Plan p{};
        for(int i = 0; i < depth; ++i) {
            bool didHold = false;
            if(d & (1u << i)) {
                if(empty) {
                    hld = cur;
                    cur = static_cast<uint8_t>(1 + k++);
                    empty = false;
                } else
                    std::swap(cur, hld);
                didHold = true;
            }
            p.slot[i] = cur;
            if(i == 0) p.rootHold = didHold;
            p.held[i] = empty ? HELD_NONE : hld;
            if(i + 1 < depth) cur = static_cast<uint8_t>(1 + k++);
        }
        plans.push_back(p); // no dedup (:197-206 keeps it commented out)
    }
    return plans;
}

struct Node { uint8_t slot, parent, childBegin, childEnd; bool rootHold; uint8_t heldSlot; };

struct Tree {
    int depth = 0;
    std::vector<std::vector<Node>> layers;
};

Tree build_tree(int depth, bool heldEmpty) { // beam.hpp:69-126
    auto const plans = build_plans(depth, heldEmpty);
    Tree tree;
    tree.depth = depth;
    tree.layers.assign(depth, {});
    std::vector<int> planNode(plans.size());
    for(size_t pi = 0; pi < plans.size(); ++pi) {
        auto const& p = plans[pi];
        auto& L = tree.layers[0];
        size_t n = 0;
        while(n < L.size() && !(L[n].slot == p.slot[0] && L[n].rootHold == p.rootHold)) ++n;
        if(n == L.size()) L.push_back(Node{p.slot[0], 0, 0, 0, p.rootHold, p.held[0]});
        planNode[pi] = static_cast<int>(n);
    }
    for(int level = 1; level < depth; ++level) {
        std::vector<int> nextNode(plans.size());
        auto& P = tree.layers[level - 1];
        auto& L = tree.layers[level];
        for(size_t parent = 0; parent < P.size(); ++parent) {
            P[parent].childBegin = static_cast<uint8_t>(L.size());
            for(size_t pi = 0; pi < plans.size(); ++pi) {
                if(planNode[pi] != static_cast<int>(parent)) continue;
                auto const& p = plans[pi];
                size_t n = P[parent].childBegin;
                while(n < L.size() && L[n].slot != p.slot[level]) ++n;
                if(n == L.size()) L.push_back(Node{p.slot[level], static_cast<uint8_t>(parent), 0, 0, false, p.held[level]});
                nextNode[pi] = static_cast<int>(n);
            }
            P[parent].childEnd = static_cast<uint8_t>(L.size());
        }
        planNode = nextNode;
    }
    return tree;
}

struct Trees {
    Tree t[MAX_DEPTH + 1][2];
    int maxRootGroups[MAX_DEPTH + 1];
    int maxChildMid[MAX_DEPTH + 1];
    int maxLast[MAX_DEPTH + 1];
    Trees() {
        for(int d = 2; d <= MAX_DEPTH; ++d) {
            t[d][0] = build_tree(d, false);
            t[d][1] = build_tree(d, true);
            // beam.hpp:131-156
            auto span = [&](int lo, int hi) {
                int m = 1;
                for(int h = 0; h < 2; ++h)
                    for(int l = lo; l <= hi && l < d; ++l)
                        for(auto const& n : t[d][h].layers[l]) m = std::max(m, n.childEnd - n.childBegin);
                return m;
            };
            maxRootGroups[d] = static_cast<int>(std::max(t[d][0].layers[0].size(), t[d][1].layers[0].size()));
            maxChildMid[d] = d > 2 ? span(0, d - 3) : 1;
            maxLast[d] = span(d - 2, d - 2);
        }
    }
};
const Trees TREES;

// ---------------------------------------------------------------------------------------------------
// value models
// ---------------------------------------------------------------------------------------------------
// clear histogram (ta3_ai/incl/ta3/ai/model_defs.hpp:37-58): four 8-bit lanes, lane k = (k+1)-line clears
inline uint32_t hist_count(uint32_t h, int k) { return (h >> (8 * k)) & 0xFFu; }
inline uint32_t hist_added(uint32_t h, int cleared) { return cleared == 0 ? h : h + (1u << (8 * (cleared - 1))); }
inline uint32_t hist_lines(uint32_t h) {
    uint32_t n = 0;
    for(int k = 0; k < 4; ++k) n += static_cast<uint32_t>(k + 1) * hist_count(h, k);
    return n;
}

// column metrics of one board (ta3_ai/incl/ta3/ai/models/metrics/v_board_metrics.hpp:130-225)
struct Metrics {
    int heights[WIDTH];
    int holes = 0, aggHeight = 0, maxHeight = 0, minHeight = HEIGHT, surfaceVar = 0, yTransitions = 0;
    int holeDepths = 0, bumpiness = 0, wells = 0;
};

inline int sum_set_bit_indices(uint32_t c) { // v_board_metrics.hpp:41-56
    int s = 0;
    for(int y = 0; y < HEIGHT; ++y)
        if((c >> y) & 1u) s += y;
    return s;
}

Metrics metrics(Board const& b) {
    Metrics m;
    for(int x = 0; x < WIDTH; ++x) {
        uint32_t const col = b.c[x];
        int const h = height(b, x);
        int const n = popc32(col);
        m.heights[x] = h;
        m.holes += h - n;                                   // holes_t :130-136
        m.aggHeight += h;                                   // agg_height_t :140-144
        m.maxHeight = std::max(m.maxHeight, h);             // max_height_t :146-150
        m.minHeight = std::min(m.minHeight, h);
        if(x > 0) {
            int const diff = h - m.heights[x - 1];
            m.surfaceVar += diff * diff;                    // surface_variance_t :152-165
            m.bumpiness += diff < 0 ? -diff : diff;         // bumpiness_t :191-202
        }
        uint32_t const yc = col | FLOOR;                    // y_transitions_t :167-175
        m.yTransitions += popc32((yc ^ (yc >> 1)) & FIELD);
        int const s = sum_set_bit_indices(col);             // hole_depths_t :177-189
        m.holeDepths += (HEIGHT - 1) * n + n * (n + 1) / 2 - (s + n * n);
    }
    for(int x = 0; x < WIDTH; ++x) { // wells_t :204-225, WELL_DEPTH = 3
        int const h = m.heights[x];
        bool well;
        if(x == 0) well = m.heights[1] >= h + 3;
        else if(x == WIDTH - 1) well = m.heights[WIDTH - 2] >= h + 3;
        else well = m.heights[x - 1] >= h + 3 && m.heights[x + 1] >= h + 3;
        m.wells += well ? 1 : 0;
    }
    return m;
}

// fsqrt (v_board_metrics.hpp:23-37): 0 for x <= 0, else the correctly rounded sqrt
inline float fsqrt(float x) { return x <= 0 ? 0.0f : std::sqrt(x); }

// ModelV11::extractInputs (ta3_ai/incl/ta3/ai/models/models/model_v11.hpp:166-180)
void features_v11(uint32_t clears, Board const& b, bool heldIsI, float* out) {
    for(int k = 0; k < 4; ++k) out[k] = static_cast<float>(hist_count(clears, k)) * 0.25f; // CLEAR_COUNT_NORM
    Metrics const m = metrics(b);
    out[4] = fsqrt(static_cast<float>(m.surfaceVar)) / 100.0f; // norm_surface_var_t :229
    out[5] = static_cast<float>(m.holes) / 20.0f;              // norm_holes_t :227
    out[6] = static_cast<float>(m.holeDepths) / 200.0f;        // norm_hole_depths_t :228
    out[7] = static_cast<float>(m.maxHeight);                  // max_height_t, raw
    out[8] = heldIsI ? 1.0f : 0.0f;
    out[9] = static_cast<float>(m.wells) / 5.0f;               // norm_wells_t :232
}

// Build-defined 16-feature superset of V11 drawn from the same metric library.
// CLEAR_SCORE_TABLE (ta3_ai/incl/ta3/ai/models/metrics/metrics.hpp:22): {0, -1, -2, -3, 10}
void features_f16(uint32_t clears, Board const& b, bool heldIsI, float* out) {
    features_v11(clears, b, heldIsI, out);
    Metrics const m = metrics(b);
    out[10] = static_cast<float>(m.aggHeight) / 200.0f;    // norm_agg_height_t :230
    out[11] = static_cast<float>(m.yTransitions) / 40.0f;  // norm_y_transitions_t :231
    out[12] = static_cast<float>(m.bumpiness) / 100.0f;
    out[13] = static_cast<float>(hist_lines(clears)) * 0.1f; // CLEAR_SCORE_NORM (model_v11.hpp:170)
    int const clearScore = -1 * static_cast<int>(hist_count(clears, 0)) - 2 * static_cast<int>(hist_count(clears, 1))
        - 3 * static_cast<int>(hist_count(clears, 2)) + 10 * static_cast<int>(hist_count(clears, 3));
    out[14] = static_cast<float>(clearScore) * 0.1f;
    out[15] = static_cast<float>(m.maxHeight - m.minHeight);
}

// Build-defined 32-feature set of the wide network: the 16 above, then 16..25 column heights / 20, 26 rows
// holding a hole / 20, 27 row transitions / 100 (the side walls count as filled), 28 deepest well / 20 (depth =
// lower neighbour height minus the column's; one neighbour at the walls), 29 filled cells / 200, 30 columns
// holding a hole / 5, 31 minimum height / 20.
void features_f32w(uint32_t clears, Board const& b, bool heldIsI, float* out) {
    features_f16(clears, b, heldIsI, out);
    Metrics const m = metrics(b);
    uint32_t holeRows = 0;
    int holeCols = 0, cells = 0, rowTrans = 0, deepest = 0;
    for(int x = 0; x < WIDTH; ++x) {
        uint32_t const col = b.c[x];
        out[16 + x] = static_cast<float>(m.heights[x]) / 20.0f;
        int const top = HEIGHT - m.heights[x]; // first filled row (HEIGHT when empty)
        uint32_t const below = FIELD & ~((1u << top) - 1u);
        uint32_t const holes = below & ~col;
        holeRows |= holes;
        holeCols += holes != 0 ? 1 : 0;
        cells += popc32(col);
        if(x + 1 < WIDTH) rowTrans += popc32((col ^ b.c[x + 1]) & FIELD);
        int const left = x > 0 ? m.heights[x - 1] : m.heights[x + 1];
        int const right = x + 1 < WIDTH ? m.heights[x + 1] : m.heights[x - 1];
        deepest = std::max(deepest, std::min(left, right) - m.heights[x]);
    }
    rowTrans += popc32(~b.c[0] & FIELD) + popc32(~b.c[WIDTH - 1] & FIELD);
    out[26] = static_cast<float>(popc32(holeRows)) / 20.0f;
    out[27] = static_cast<float>(rowTrans) / 100.0f;
    out[28] = static_cast<float>(deepest) / 20.0f;
    out[29] = static_cast<float>(cells) / 200.0f;
    out[30] = static_cast<float>(holeCols) / 5.0f;
    out[31] = static_cast<float>(m.minHeight) / 20.0f;
}

// bfloat16 rounding (nearest, ties to even) of a float32
inline float bf16_rne(float v) {
    uint32_t u;
    std::memcpy(&u, &v, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    u &= 0xFFFF0000u;
    float r;
    std::memcpy(&r, &u, 4);
    return r;
}

// Wide network 32 -> 128 ReLU -> 128 ReLU -> 1, layout as dense(): W1[128][32], b1, W2[128][128], b2, W3[128], b3.
// Device arithmetic (BF16 MFMA, FP32 accumulate): features, W1 and W2 rounded to bfloat16; each hidden
// pre-activation = bias + exact sum of the bfloat16 products, rounded once to float32 (the device's FP32
// accumulation may differ in the last bits); layer 1 output rounded to bfloat16 after ReLU; the output layer in
// FP32: p_q = FMA chain over o = 16 t + 4 q + j (t = 0..7, then j = 0..3) of h2[o] W3[o], value =
// ((p0 + p1) + (p2 + p3)) + b3.
// The device's relu of the wide network: the signed maximum of the value's bits and 0, i.e. +0 for a value with
// the sign bit set and the value itself otherwise. Equal to (0 < v) ? v : +0 for every v that is not NaN (a NaN
// arises only from non-finite or enormous weights; the bfloat16 rounding of a NaN is not restated).
inline float relu_bits(float v) { return std::signbit(v) ? 0.0f : v; }

float value_f32w(float const* w, float const* in) {
    float x[32], h1[128], h2[128];
    for(int i = 0; i < 32; ++i) x[i] = bf16_rne(in[i]);
    for(int u = 0; u < 128; ++u) {
        double v = w[4096 + u];
        for(int i = 0; i < 32; ++i) v += static_cast<double>(x[i]) * bf16_rne(w[u * 32 + i]);
        float const f = static_cast<float>(v);
        h1[u] = relu_bits(bf16_rne(f)); // the device rounds to bfloat16, then takes relu
    }
    for(int o = 0; o < 128; ++o) {
        double v = w[4224 + 16384 + o];
        for(int u = 0; u < 128; ++u) v += static_cast<double>(h1[u]) * bf16_rne(w[4224 + o * 128 + u]);
        float const f = static_cast<float>(v);
        h2[o] = relu_bits(f);
    }
    float const* w3 = w + 4224 + 16512;
    float p[4];
    for(int q = 0; q < 4; ++q) {
        p[q] = 0.0f;
        for(int t = 0; t < 8; ++t)
            for(int j = 0; j < 4; ++j) p[q] = std::fmaf(h2[16 * t + 4 * q + j], w3[16 * t + 4 * q + j], p[q]);
    }
    return ((p[0] + p[1]) + (p[2] + p[3])) + w3[128];
}

// dense layer (ta3_ai/incl/ta3/ai/models/models/utility.hpp:29-46): W[o*In + i], biases after the
// matrix, v = bias then += in[i] * w in ascending i, ReLU = std::max<float>(0, v)
void dense(float const* in, int nIn, int nOut, bool relu, float const* w, float* out) {
    for(int o = 0; o < nOut; ++o) {
        float v = w[nIn * nOut + o];
        for(int i = 0; i < nIn; ++i) {
            float const p = in[i] * w[o * nIn + i];
            v = v + p;
        }
        out[o] = relu ? ((0.0f < v) ? v : 0.0f) : v;
    }
}

// Build-defined dense layer of the 16-feature network: the same layout and order as dense() above, but
// each step is one fused multiply-add, v = fma(in[i], w, v), rounded once (IEEE 754 fusedMultiplyAdd).
void dense_fma(float const* in, int nIn, int nOut, bool relu, float const* w, float* out) {
    for(int o = 0; o < nOut; ++o) {
        float v = w[nIn * nOut + o];
        for(int i = 0; i < nIn; ++i) v = std::fma(in[i], w[o * nIn + i], v);
        out[o] = relu ? ((0.0f < v) ? v : 0.0f) : v;
    }
}

struct ModelRef {
    int kind;
    float const* w;

    float value(uint32_t clears, Board const& b, bool heldIsI) const {
        switch(kind) {
            case TA3O_MODEL_V11: { // ModelV11::forward (model_v11.hpp:166-180): 10 -> 8 ReLU -> 1
                float in[10], h1[8], out[1];
                features_v11(clears, b, heldIsI, in);
                dense(in, 10, 8, true, w, h1);
                dense(h1, 8, 1, false, w + 88, out);
                return out[0];
            }
            case TA3O_MODEL_F16: { // 16 -> 32 ReLU -> 32 ReLU -> 1, fused multiply-adds
                float in[16], h1[32], h2[32], out[1];
                features_f16(clears, b, heldIsI, in);
                dense_fma(in, 16, 32, true, w, h1);
                dense_fma(h1, 32, 32, true, w + 544, h2);
                dense_fma(h2, 32, 1, false, w + 1600, out);
                return out[0];
            }
            case TA3O_MODEL_F32W: { // 32 -> 128 ReLU -> 128 ReLU -> 1, device (BF16 MFMA) arithmetic
                float in[32];
                features_f32w(clears, b, heldIsI, in);
                return value_f32w(w, in);
            }
            case TA3O_MODEL_HEIGHT: { // beam_test.cpp:20-35
                uint32_t agg = 0;
                for(int x = 0; x < WIDTH; ++x) agg += static_cast<uint32_t>(height(b, x));
                float const a = 2.0f * static_cast<float>(hist_lines(clears));
                float const c = 0.3f * static_cast<float>(agg);
                float const d = 1.0f * static_cast<float>(holes(b));
                return (a - c) - d;
            }
            default: return 0.0f;
        }
    }
};

// ---------------------------------------------------------------------------------------------------
// beamed search (ta3_ai/incl/ta3/ai/search/beam.hpp)
// ---------------------------------------------------------------------------------------------------
inline uint32_t ordered_bits(float v) { // :204-207
    uint32_t u;
    std::memcpy(&u, &v, 4);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
inline float ordered_bits_inv(uint32_t s) { // :209-211
    uint32_t const u = (s & 0x80000000u) ? (s ^ 0x80000000u) : ~s;
    float v;
    std::memcpy(&v, &u, 4);
    return v;
}
inline uint64_t pack_key(uint32_t score, uint32_t pos, uint32_t span) { // :217-219
    return (static_cast<uint64_t>(score) << 32) | static_cast<uint64_t>(span - pos);
}
inline uint32_t key_pos(uint64_t key, uint32_t span) { return span - static_cast<uint32_t>(key & 0xFFFFFFFFull); } // :221-223

// select_top (:244-265): repeatedly take the max nonzero score, lowest slot on ties, zero it
std::atomic<int> g_selectLinear{0}; // 1: the argmax loop for every width

uint32_t select_top(uint32_t* scores, uint32_t size, uint32_t k, uint32_t* outSlots, uint32_t* outScores) {
    if(k > 64 && g_selectLinear.load(std::memory_order_relaxed) == 0) {
        // very wide beams (only the full-width checks use them): the same total order -- score
        // descending, slot ascending, zeros excluded -- by one sort instead of k linear scans
        std::vector<uint32_t> idx;
        for(uint32_t i = 0; i < size; ++i)
            if(scores[i] != 0) idx.push_back(i);
        std::sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) {
            return scores[a] != scores[b] ? scores[a] > scores[b] : a < b;
        });
        uint32_t const n = std::min<uint32_t>(k, static_cast<uint32_t>(idx.size()));
        for(uint32_t j = 0; j < n; ++j) {
            outSlots[j] = idx[j];
            outScores[j] = scores[idx[j]];
            scores[idx[j]] = 0;
        }
        return n;
    }
    uint32_t count = 0;
    for(; count < k; ++count) {
        uint32_t best = 0, bestSlot = 0;
        for(uint32_t i = 0; i < size; ++i)
            if(scores[i] > best) {
                best = scores[i];
                bestSlot = i;
            }
        if(best == 0) break;
        outSlots[count] = bestSlot;
        outScores[count] = best;
        scores[bestSlot] = 0;
    }
    return count;
}
uint32_t select_top(std::vector<uint32_t>& scores, uint32_t k, uint32_t* outSlots, uint32_t* outScores) {
    return select_top(scores.data(), static_cast<uint32_t>(scores.size()), k, outSlots, outScores);
}

struct Step { Board board; uint32_t clears; bool legal; };
inline Step apply(Board const& b, uint32_t clears, Placement const& pl) { // search.hpp:61-69
    Fit const f = fit(b, pl.shape);
    if(!f.legal) return {b, clears, false};
    Board n = placed(b, pl.shape, f.drop);
    int const cleared = clear_lines(n);
    return {n, hist_added(clears, cleared), true};
}

struct Frontier {
    std::vector<Board> board;
    std::vector<uint32_t> clears, score;
    std::vector<uint8_t> node, rootRank;
    void resize(size_t n) {
        board.resize(n);
        clears.resize(n);
        score.resize(n);
        node.resize(n);
        rootRank.resize(n);
    }
};

struct SearchResult { bool some; bool hold; int o, x; float value; };

struct Counters {
    uint64_t moves = 0, applies = 0, legal = 0, reapplies = 0, placements = 0;
    uint64_t deadMid = 0, deadLeaf = 0, fallbackMidBest = 0, fallbackLeafBest = 0; // dead ends, and moves they decided
    uint64_t repeatApplies = 0, repeatLegal = 0; // of applies / legal: in repeated lists (ListSeen)
    uint64_t reuseApplies = 0, reuseLegal = 0;   // of applies / legal: roots scored by the previous move (RootReuse)
    void add(Counters const& c) {
        reuseApplies += c.reuseApplies;
        reuseLegal += c.reuseLegal;
        repeatApplies += c.repeatApplies;
        repeatLegal += c.repeatLegal;
        moves += c.moves;
        applies += c.applies;
        legal += c.legal;
        reapplies += c.reapplies;
        placements += c.placements;
        deadMid += c.deadMid;
        deadLeaf += c.deadLeaf;
        fallbackMidBest += c.fallbackMidBest;
        fallbackLeafBest += c.fallbackLeafBest;
    }
};


// ---------------------------------------------------------------------------------------------------
// placement-proposal network (build-defined; no reference counterpart)
// input 265: board slots 24c + r = bit (3 + r) of column c (r < 20; r in [20, 24) always 0), then one-hots:
// piece (240 + p), rotation (247 + o), column (251 + x + 2), ply (262 + ply - 1). 265 -> 64 ReLU -> 64 ReLU -> 1.
// layout: W1[64][265], b1[64], W2[64][64], b2[64], W3[64], b3 (21,249 floats). W1 and W2 are used rounded to
// bfloat16 (round to nearest even), the first hidden layer is rounded to bfloat16 before layer 2.
// ---------------------------------------------------------------------------------------------------
constexpr int PR_IN = 265, PR_H = 64;
constexpr int PR_W1 = 0, PR_B1 = PR_W1 + PR_H * PR_IN, PR_W2 = PR_B1 + PR_H, PR_B2 = PR_W2 + PR_H * PR_H,
              PR_W3 = PR_B2 + PR_H, PR_B3 = PR_W3 + PR_H, PR_PARAMS = PR_B3 + 1;

inline float bf16_round(float v) {
    uint32_t u;
    std::memcpy(&u, &v, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    u &= 0xFFFF0000u;
    float r;
    std::memcpy(&r, &u, 4);
    return r;
}

// active input indices of a candidate (board bits then the four one-hots); returns the count
int pruner_inputs(Board const& b, int piece, int o, int x, int ply, int* idx) {
    int n = 0;
    for(int c = 0; c < WIDTH; ++c)
        for(int r = 0; r < 20; ++r)
            if((b.c[c] >> (3 + r)) & 1u) idx[n++] = 24 * c + r;
    idx[n++] = 240 + piece;
    idx[n++] = 247 + o;
    idx[n++] = 251 + x + 2;
    idx[n++] = 262 + ply - 1;
    return n;
}

// Inference form: layers 1 and 2 are an integer network on power-of-two grids (each grid's exponent is the
// largest that keeps the layer's largest magnitude inside its integer range), so every sum is an exact integer
// below 2^24 whatever the summation order; layer 3 is float32 in a fixed order:
//   N1 = rne(W1 2^e1), Nb1 = rne(b1 2^e1) (|N1| <= 255, |Nb1| <= 2^20), e1 from max |W1| <= 255
//   P  = sum over active inputs of N1 + Nb1                      (first-layer pre-activation, |P| < 2^21)
//   c  = min(255, max(0, rne(P 2^-d))), d = e1 - e2 clamped to [-8, 15], e2 from act_range <= 255
//   N2 = rne(W2 2^e3), e3 from max |W2| <= 255; Nb2 = rne(b2 2^(e1 - d + e3)), |Nb2| <= 2^22
//   h2 = max(0, Nb2[v] + sum_u c[u] N2[v][u])                    (exact, |.| < 2^23)
//   logit = (p0 + p1) + (p2 + p3), p_q = fmaf chain over v = 16 t + 4 q + j (t = 0..3, then j = 0..3) of
//           h2[v] W3[v], starting from 0                         (float32; W3 unquantized)
// b3 shifts every logit alike and is not used. act_range is the calibrated upper range of the first hidden
// layer in float units.
struct PrunerQ {
    int32_t n1[PR_H][PR_IN];
    int32_t nb1[PR_H];
    int32_t n2[PR_H][PR_H];
    int32_t nb2[PR_H];
    float w3[PR_H];
    int d;
    int e1, e2, e3;
};

// largest e with m 2^e <= lim (0 when m is 0)
int grid_exponent(double m, double lim) {
    if(!(m > 0.0)) return 0;
    int e = 0;
    while(std::ldexp(m, e) > lim) --e;
    while(std::ldexp(m, e + 1) <= lim) ++e;
    return e;
}

int32_t rne_scaled(double v, int e, int32_t lim) {
    double const r = std::nearbyint(std::ldexp(v, e)); // default rounding mode: nearest, ties to even
    return static_cast<int32_t>(std::min<double>(std::max<double>(r, -lim), lim));
}

void pruner_quantize(float const* w, float actRange, PrunerQ& q) {
    double m1 = 0, m2 = 0;
    for(int i = 0; i < PR_H * PR_IN; ++i) m1 = std::max(m1, std::fabs(static_cast<double>(w[PR_W1 + i])));
    for(int i = 0; i < PR_H * PR_H; ++i) m2 = std::max(m2, std::fabs(static_cast<double>(w[PR_W2 + i])));
    q.e1 = grid_exponent(m1, 255.0);
    q.e2 = grid_exponent(actRange, 255.0);
    q.e3 = grid_exponent(m2, 255.0);
    q.d = std::min(15, std::max(-8, q.e1 - q.e2));
    for(int u = 0; u < PR_H; ++u) {
        for(int k = 0; k < PR_IN; ++k) q.n1[u][k] = rne_scaled(w[PR_W1 + u * PR_IN + k], q.e1, 255);
        q.nb1[u] = rne_scaled(w[PR_B1 + u], q.e1, 1 << 20);
        for(int i = 0; i < PR_H; ++i) q.n2[u][i] = rne_scaled(w[PR_W2 + u * PR_H + i], q.e3, 255);
        q.nb2[u] = rne_scaled(w[PR_B2 + u], q.e1 - q.d + q.e3, 1 << 22);
        q.w3[u] = w[PR_W3 + u];
    }
}

int32_t pruner_code(int32_t p, int d) {
    double const c = std::nearbyint(std::ldexp(static_cast<double>(p), -d)); // exact scaling, then rne
    return static_cast<int32_t>(std::min(255.0, std::max(0.0, c)));
}

float pruner_score(PrunerQ const& q, Board const& b, int piece, int o, int x, int ply) {
    int idx[PR_IN];
    int const n = pruner_inputs(b, piece, o, x, ply, idx);
    int32_t c[PR_H];
    for(int u = 0; u < PR_H; ++u) {
        int32_t p = q.nb1[u];
        for(int j = 0; j < n; ++j) p += q.n1[u][idx[j]];
        c[u] = pruner_code(p, q.d);
    }
    float h2[PR_H];
    for(int v = 0; v < PR_H; ++v) {
        int64_t a = q.nb2[v];
        for(int u = 0; u < PR_H; ++u) a += static_cast<int64_t>(c[u]) * q.n2[v][u];
        h2[v] = static_cast<float>(a > 0 ? a : 0);
    }
    float p[4];
    for(int qq = 0; qq < 4; ++qq) {
        p[qq] = 0.0f;
        for(int t = 0; t < 4; ++t)
            for(int j = 0; j < 4; ++j) {
                int const v = 16 * t + 4 * qq + j;
                p[qq] = std::fmaf(h2[v], q.w3[v], p[qq]);
            }
    }
    return (p[0] + p[1]) + (p[2] + p[3]);
}

struct PrunerRef {
    PrunerQ const* q = nullptr; // quantized network; null = no pruning
    int k = 34;               // placements kept per list at a pruned ply
    unsigned plyMask = 0;     // bit p: prune ply p (1, 2 = middle levels, 3 = leaves)
};

// trace record: parent board (10 words) + meta = piece | o << 3 | (x + 2) << 5 | ply << 9 | label << 11
struct TraceOut {
    std::vector<uint32_t>* out = nullptr;
    bool on = false;
    void add(Board const& b, int piece, int o, int x, int ply, bool label) {
        for(int c = 0; c < WIDTH; ++c) out->push_back(b.c[c]);
        out->push_back(static_cast<uint32_t>(piece) | static_cast<uint32_t>(o) << 3 | static_cast<uint32_t>(x + 2) << 5
                       | static_cast<uint32_t>(ply) << 9 | (label ? 1u << 11 : 0u));
    }
};

// keeps the k best legal placements of one list by pruner score (score descending, index ascending)
void prune_list(PrunerRef const& pr, Board const& parent, int piece, int ply, uint8_t const* legal, int n,
                uint8_t* keep) {
    float sc[MAX_PLACEMENTS];
    int order[MAX_PLACEMENTS];
    int m = 0;
    for(int i = 0; i < n; ++i) {
        keep[i] = 0;
        if(!legal[i]) continue;
        Placement const pl = LUT.nth(piece, i);
        sc[i] = pruner_score(*pr.q, parent, piece, pl.o, pl.x, ply);
        order[m++] = i;
    }
    std::stable_sort(order, order + m, [&](int a, int b) { return sc[a] > sc[b]; });
    for(int j = 0; j < m && j < pr.k; ++j) keep[order[j]] = 1;
}

// Training reference (float64): batch-mean weighted BCE on the logit and its gradient. `rounded` applies the
// inference roundings (bfloat16 W1, W2 and first hidden layer) with straight-through gradients; without them
// the loss is smooth and the gradient can be checked by finite differences.
// sample = parent board (10 words) + meta (trace record layout); weight = posWeight for label 1, else 1.
double pruner_loss_grad(float const* w, uint32_t const* samples, int n, double posWeight, bool rounded, double* grad) {
    auto rw = [&](float v) { return rounded ? static_cast<double>(bf16_round(v)) : static_cast<double>(v); };
    if(grad) std::fill(grad, grad + PR_PARAMS, 0.0);
    double loss = 0.0;
    int idx[PR_IN];
    double pre1[PR_H], h1[PR_H], pre2[PR_H], h2[PR_H], dpre2[PR_H];
    for(int t = 0; t < n; ++t) {
        uint32_t const* rec = samples + 11 * static_cast<size_t>(t);
        Board b;
        std::memcpy(b.c, rec, 40);
        uint32_t const meta = rec[10];
        int const piece = meta & 7, o = (meta >> 3) & 3, x = static_cast<int>((meta >> 5) & 15) - 2,
                  ply = (meta >> 9) & 3;
        double const y = (meta >> 11) & 1u;
        int const m = pruner_inputs(b, piece, o, x, ply, idx);
        for(int u = 0; u < PR_H; ++u) {
            double v = w[PR_B1 + u];
            for(int j = 0; j < m; ++j) v += rw(w[PR_W1 + u * PR_IN + idx[j]]);
            pre1[u] = v;
            double const r = v > 0.0 ? v : 0.0;
            h1[u] = rounded ? static_cast<double>(bf16_round(static_cast<float>(r))) : r;
        }
        double sc = w[PR_B3];
        for(int v = 0; v < PR_H; ++v) {
            double a = w[PR_B2 + v];
            for(int u = 0; u < PR_H; ++u) a += h1[u] * rw(w[PR_W2 + v * PR_H + u]);
            pre2[v] = a;
            h2[v] = a > 0.0 ? a : 0.0;
            sc += h2[v] * w[PR_W3 + v];
        }
        double const wt = y > 0.5 ? posWeight : 1.0;
        double const sp = sc > 0 ? sc + std::log1p(std::exp(-sc)) : std::log1p(std::exp(sc)); // softplus
        loss += wt * (sp - y * sc) / n;
        if(!grad) continue;
        double const sig = 1.0 / (1.0 + std::exp(-sc));
        double const ds = wt * (sig - y) / n;
        grad[PR_B3] += ds;
        for(int v = 0; v < PR_H; ++v) {
            grad[PR_W3 + v] += ds * h2[v];
            dpre2[v] = pre2[v] > 0.0 ? ds * w[PR_W3 + v] : 0.0;
            grad[PR_B2 + v] += dpre2[v];
            for(int u = 0; u < PR_H; ++u) grad[PR_W2 + v * PR_H + u] += dpre2[v] * h1[u];
        }
        for(int u = 0; u < PR_H; ++u) {
            double dh = 0.0;
            for(int v = 0; v < PR_H; ++v) dh += dpre2[v] * rw(w[PR_W2 + v * PR_H + u]);
            double const dp = pre1[u] > 0.0 ? dh : 0.0;
            grad[PR_B1 + u] += dp;
            for(int j = 0; j < m; ++j) grad[PR_W1 + u * PR_IN + idx[j]] += dp;
        }
    }
    return loss;
}

// The lists of one level seen so far, for counting repeats: a list is a repeat when an earlier list of its level
// has the same parent board, parent clear histogram, piece and held-is-I. Its candidates' legality and values
// equal those of the earlier list (they are functions of these and the placement); the search still scores them
// (the reference does), the counters record them (the device search scores each such content once).
struct ListSeen {
    struct Key {
        uint32_t w[WIDTH + 2]; // board columns, clear histogram, piece | heldIsI << 3
        bool operator==(Key const& o) const { return std::memcmp(w, o.w, sizeof(w)) == 0; }
    };
    struct Hash {
        size_t operator()(Key const& k) const {
            uint64_t h = 0x9E3779B97F4A7C15ull;
            for(uint32_t v : k.w) h = (h ^ v) * 0xBF58476D1CE4E5B9ull;
            return static_cast<size_t>(h ^ (h >> 31));
        }
    };
    std::unordered_set<Key, Hash> keys;
    void clear() { keys.clear(); }
    bool repeat(Board const& b, uint32_t clears, uint32_t piece, bool heldIsI) {
        Key k;
        std::memcpy(k.w, b.c, sizeof(b.c));
        k.w[WIDTH] = clears;
        k.w[WIDTH + 1] = piece | (heldIsI ? 8u : 0u);
        return !keys.insert(k).second;
    }
    bool has(Board const& b, uint32_t clears, uint32_t piece, bool heldIsI) const {
        Key k;
        std::memcpy(k.w, b.c, sizeof(b.c));
        k.w[WIDTH] = clears;
        k.w[WIDTH + 1] = piece | (heldIsI ? 8u : 0u);
        return keys.count(k) != 0;
    }
};

// Roots scored by the previous move. When the committed placement cleared no line, the next move's root lists
// (the committed board, an empty clear histogram, a piece and held-is-I) are level-1 lists of the previous search's
// chosen root node whenever piece and held-is-I agree: the same candidates with the same values. The device search
// takes those values from the previous move when every root list has such a match; the search here still scores
// them (the reference does) and the counters record them.
struct RootReuse {
    bool valid = false;     // the last committed placement cleared no line
    uint32_t keys[2] = {};  // the chosen root's level-1 lists: piece | heldIsI << 3 | 16 (0 = no list)
    uint32_t level1[64] = {}; // this search's level-1 lists of kept root k: [2 k + child]
};

// Lists of levels 1 and 2 scored by the previous move (the device's plain 16-32-32-1 kernels keep the scores of
// its level-2 and leaf lists). When RootReuse applies, a root of this search is a level-1 candidate of the chosen
// root; if that candidate was a kept level-1 node of the previous search, each of its level-2 lists with the same
// piece and held-is-I is a level-1 list of this search with the same content (the board follows from the path,
// and the clear histograms agree because the committed placement cleared no line). Likewise one level deeper: a
// kept level-1 node of this search in such a list is a level-2 candidate of the previous search, and if that was
// kept, its leaf lists are this search's level-2 lists. The device takes those lists' scores from the previous
// move when they are not repeats; the search here scores them and the counters record them (with RootReuse's).
struct DeepReuse {
    uint32_t rank = 0;                  // the chosen root's rank: its level-1 lists are 2 rank and 2 rank + 1
    uint32_t rootChild[2] = {};         // this search: child c of the chosen root whose level-1 list root list g is
    uint32_t kept1[32] = {}, n1 = 0;    // slots (list * 34 + placement) of the kept level-1 nodes, in rank order
    uint32_t kept2[32] = {}, n2 = 0;    // and of the kept level-2 nodes
    uint32_t key2[64] = {}, keyL[64] = {}; // level-2 and leaf lists: piece | heldIsI << 3 | 16 (0 = no list)
    int32_t id1[64] = {};               // this search's level-1 list -> the previous search's level-2 list, or -1
};

// Content check of DeepReuse (tests only, `checkDeep`): every list it counts must have the content (board, clear
// histogram, piece, held-is-I) of a list the previous search built one level down (its level-2 lists for this
// level 1, its leaf lists for this level 2), which the structural identities above are meant to guarantee.
struct DeepCheck {
    bool on = false;
    ListSeen prev2, prevL, cur2, curL;
    uint64_t counted = 0, mismatched = 0;
};

struct BeamWorkspace {
    std::vector<uint32_t> rootScores, midScores, topSlots, topScores, topRootSlots, topRootScores;
    std::vector<uint8_t> hasChild;
    Frontier kept[2];
    ListSeen seen;
    RootReuse reuse;
    DeepReuse deep;
    DeepCheck check;
};

// the searches whose device kernels keep level-2 and leaf scores for the next move (DeepReuse): the full-beam
// 16-32-32-1 network at depth 4 (the LDS of V11's and the wide network's kernels has no room for them; pruned
// searches score only part of a pruned level)
inline bool deep_reuse_on(ModelRef const& model, PrunerRef const& pr, int depth) {
    return model.kind == TA3O_MODEL_F16 && pr.q == nullptr && depth == 4;
}

// search_move_beam (:483-570) at runtime depth = ks.size() + 1
SearchResult search_beam(Engine const& g, ModelRef const& model, std::vector<uint32_t> const& ks,
                         BeamWorkspace& ws, Counters& cnt, PrunerRef const& pr = PrunerRef{},
                         TraceOut trace = TraceOut{}) {
    int const depth = static_cast<int>(ks.size()) + 1;
    if(g.gameOver) return {false, false, 0, 0, 0.0f};

    int const MP = MAX_PLACEMENTS;
    int const maxChildMid = TREES.maxChildMid[depth], maxLast = TREES.maxLast[depth];
    uint32_t const k0 = ks[0];
    uint32_t kMax = 0;
    for(auto k : ks) kMax = std::max(kMax, k);
    uint32_t midMax = 0; // mid_slots (:174-179)
    for(int level = 1; level + 1 < depth; ++level) midMax = std::max(midMax, ks[level - 1]);
    size_t const midSlots = midMax == 0 ? 1 : static_cast<size_t>(midMax) * maxChildMid * MP;
    uint32_t const leafOrderCount = ks[depth - 2] * maxLast * MP; // leaf_orders (:188)
    uint32_t const orderCount = leafOrderCount + k0;                // order_span (:196-198)

    // gather_slots (:226-237): [held, current, lookahead 0..depth-1]
    uint32_t slot[MAX_DEPTH + 2];
    slot[0] = g.held;
    slot[1] = g.queue[0];
    for(int i = 0; i < depth; ++i) slot[2 + i] = g.queue[1 + i];
    Tree const& tree = TREES.t[depth][g.held == NO_PIECE ? 1 : 0];
    auto heldIsI = [&](Node const& n) { return n.heldSlot == HELD_NONE ? false : slot[n.heldSlot] == PIECE_I; };
    Board const& board = g.board;

    // stage 0: score_roots (:283-301)
    bool reuseRoots = ws.reuse.valid; // RootReuse: every root list matches a level-1 list of the last chosen root
    for(size_t gi = 0; gi < tree.layers[0].size(); ++gi) {
        Node const& node = tree.layers[0][gi];
        uint32_t const key = slot[node.slot] | (heldIsI(node) ? 8u : 0u) | 16u;
        reuseRoots = reuseRoots && (ws.reuse.keys[0] == key || ws.reuse.keys[1] == key);
        if(gi < 2) ws.deep.rootChild[gi] = ws.reuse.keys[0] == key ? 0u : 1u;
    }
    // DeepReuse: this search takes lists of the previous one (its state is written by every search of such a kind)
    bool const deepOn = deep_reuse_on(model, pr, depth) && maxChildMid == 2 && maxLast == 2;
    bool const deepUse = deepOn && reuseRoots;
    if(ws.check.on) { // the previous search's level-2 and leaf lists
        std::swap(ws.check.prev2.keys, ws.check.cur2.keys);
        std::swap(ws.check.prevL.keys, ws.check.curL.keys);
        ws.check.cur2.clear();
        ws.check.curL.clear();
    }
    std::fill(std::begin(ws.reuse.level1), std::end(ws.reuse.level1), 0u);
    ws.rootScores.assign(static_cast<size_t>(TREES.maxRootGroups[depth]) * MP, 0u);
    for(size_t gi = 0; gi < tree.layers[0].size(); ++gi) {
        Node const& node = tree.layers[0][gi];
        uint32_t const p0 = slot[node.slot];
        bool const hI = heldIsI(node);
        for(int i = 0; i < LUT.count(p0); ++i) {
            Step const s = apply(board, 0, LUT.nth(p0, i));
            ++cnt.applies;
            if(s.legal) ++cnt.legal;
            cnt.reuseApplies += reuseRoots;
            cnt.reuseLegal += reuseRoots && s.legal;
            ws.rootScores[gi * MP + i] = s.legal ? ordered_bits(model.value(s.clears, s.board, hI)) : 0u;
        }
    }
    ws.topRootSlots.assign(k0, 0);
    ws.topRootScores.assign(k0, 0);
    uint32_t const rootCount = select_top(ws.rootScores, k0, ws.topRootSlots.data(), ws.topRootScores.data());
    if(rootCount == 0) return {false, false, 0, 0, 0.0f};

    uint64_t bestKey = 0;
    int bestFallback = 0; // where the best key so far came from (counted, not used): 0 leaf, 1 mid / 2 leaf dead end
    auto consider = [&](uint32_t score, uint32_t order, int fallback = 0) {
        uint64_t const key = pack_key(score, order, orderCount);
        if(key > bestKey) {
            bestKey = key;
            bestFallback = fallback;
        }
    };

    ws.kept[0].resize(kMax);
    ws.kept[1].resize(kMax);
    int cur = 0;
    uint32_t keptCount = rootCount;

    // seed_frontier (:307-329)
    for(uint32_t k = 0; k < rootCount; ++k) {
        uint32_t const gi = ws.topRootSlots[k] / MP;
        Step const s = apply(board, 0, LUT.nth(slot[tree.layers[0][gi].slot], ws.topRootSlots[k] % MP));
        ++cnt.reapplies;
        ws.kept[0].board[k] = s.board;
        ws.kept[0].clears[k] = s.clears;
        ws.kept[0].score[k] = ws.topRootScores[k];
        ws.kept[0].node[k] = static_cast<uint8_t>(gi);
        ws.kept[0].rootRank[k] = static_cast<uint8_t>(k);
    }

    // middle stages (:540-564)
    ws.topSlots.assign(kMax, 0);
    ws.topScores.assign(kMax, 0);
    for(int level = 1; level + 1 < depth; ++level) {
        Frontier const& F = ws.kept[cur];
        // score_mid_level (:337-372); at a pruned ply only the pruner's top-k legal placements of a list are scored
        ws.midScores.assign(midSlots, 0u);
        ws.hasChild.assign(kMax, 0);
        bool const pruneHere = pr.q != nullptr && ((pr.plyMask >> level) & 1u) != 0;
        std::vector<uint8_t> legalMid(midSlots, 0);
        ws.seen.clear();
        DeepReuse& D = ws.deep;
        uint32_t* const newKeys = level == 2 ? D.key2 : nullptr; // this level's list keys, for the next search
        if(deepOn && level == 1) std::fill(std::begin(D.id1), std::end(D.id1), -1);
        if(deepOn && level == 2) std::fill(std::begin(D.key2), std::end(D.key2), 0u);
        for(uint32_t k = 0; k < keptCount; ++k) {
            Node const& pn = tree.layers[level - 1][F.node[k]];
            // DeepReuse: kept node k as a kept node of the previous search one level down (index into its kept list)
            int prevNode = -1;
            if(deepUse) {
                int32_t prevSlot = -1;
                if(level == 1) {
                    uint32_t const rs = ws.topRootSlots[k];
                    prevSlot = static_cast<int32_t>((2 * D.rank + D.rootChild[rs / MP]) * 34 + rs % MP);
                } else if(D.id1[D.kept1[k] / 34] >= 0) {
                    prevSlot = D.id1[D.kept1[k] / 34] * 34 + static_cast<int32_t>(D.kept1[k] % 34);
                }
                uint32_t const* const prevKept = level == 1 ? D.kept1 : D.kept2;
                uint32_t const prevCount = level == 1 ? D.n1 : D.n2;
                for(uint32_t q = 0; q < prevCount && prevSlot >= 0; ++q)
                    if(prevKept[q] == static_cast<uint32_t>(prevSlot)) {
                        prevNode = static_cast<int>(q);
                        break;
                    }
            }
            for(int ci = pn.childBegin; ci < pn.childEnd; ++ci) {
                Node const& cn = tree.layers[level][ci];
                uint32_t const piece = slot[cn.slot];
                bool const hI = heldIsI(cn);
                size_t const base = (static_cast<size_t>(k) * maxChildMid + (ci - pn.childBegin)) * MP;
                int const n = LUT.count(piece);
                bool const rep = ws.seen.repeat(F.board[k], F.clears[k], piece, hI);
                uint32_t const lkey = piece | (hI ? 8u : 0u) | 16u;
                uint32_t const li = 2 * k + static_cast<uint32_t>(ci - pn.childBegin); // the list's index
                int32_t prevList = -1; // the previous search's list with this content (DeepReuse)
                if(prevNode >= 0) {
                    uint32_t const* const prevKeys = level == 1 ? D.key2 : D.keyL;
                    for(uint32_t c2 = 0; c2 < 2 && prevList < 0; ++c2)
                        if(prevKeys[2 * prevNode + c2] == lkey) prevList = static_cast<int32_t>(2 * prevNode + c2);
                }
                if(deepOn && level == 1) D.id1[li] = prevList;
                if(deepOn && newKeys != nullptr) newKeys[li] = lkey;
                bool const deepHit = prevList >= 0 && !rep;
                if(ws.check.on) {
                    if(deepHit) {
                        ++ws.check.counted;
                        ListSeen const& prevLists = level == 1 ? ws.check.prev2 : ws.check.prevL;
                        ws.check.mismatched += prevLists.has(F.board[k], F.clears[k], piece, hI) ? 0u : 1u;
                    }
                    if(level == 2) ws.check.cur2.repeat(F.board[k], F.clears[k], piece, hI);
                }
                if(level == 1 && k < 32 && ci - pn.childBegin < 2)
                    ws.reuse.level1[2 * k + (ci - pn.childBegin)] = piece | (hI ? 8u : 0u) | 16u;
                uint8_t legal[MAX_PLACEMENTS], keep[MAX_PLACEMENTS];
                int dropTo[MAX_PLACEMENTS];
                for(int i = 0; i < n; ++i) {
                    Fit const f = fit(F.board[k], LUT.nth(piece, i).shape);
                    ++cnt.applies;
                    cnt.repeatApplies += rep;
                    cnt.reuseApplies += deepHit;
                    legal[i] = f.legal ? 1 : 0;
                    dropTo[i] = f.drop;
                    keep[i] = legal[i];
                    legalMid[base + i] = legal[i];
                }
                if(pruneHere) prune_list(pr, F.board[k], static_cast<int>(piece), level, legal, n, keep);
                for(int i = 0; i < n; ++i) {
                    if(!legal[i]) continue;
                    ws.hasChild[k] = 1;
                    if(!keep[i]) continue;
                    ++cnt.legal;
                    cnt.repeatLegal += rep;
                    cnt.reuseLegal += deepHit;
                    Step const st = apply(F.board[k], F.clears[k], LUT.nth(piece, i));
                    (void)dropTo;
                    ws.midScores[base + i] = ordered_bits(model.value(st.clears, st.board, hI));
                }
            }
        }
        uint32_t const nextCount = select_top(ws.midScores, ks[level], ws.topSlots.data(), ws.topScores.data());
        if(deepOn) { // this level's kept nodes, for the next level and the next search
            uint32_t* const kept = level == 1 ? D.kept1 : D.kept2;
            for(uint32_t j = 0; j < nextCount; ++j) kept[j] = ws.topSlots[j];
            (level == 1 ? D.n1 : D.n2) = nextCount;
        }
        if(trace.on) { // every legal candidate of the level; label = it survived the level's top-K
            std::vector<uint8_t> survived(midSlots, 0);
            for(uint32_t j = 0; j < nextCount; ++j) survived[ws.topSlots[j]] = 1;
            for(uint32_t k = 0; k < keptCount; ++k) {
                Node const& pn = tree.layers[level - 1][F.node[k]];
                for(int ci = pn.childBegin; ci < pn.childEnd; ++ci) {
                    uint32_t const piece = slot[tree.layers[level][ci].slot];
                    size_t const base = (static_cast<size_t>(k) * maxChildMid + (ci - pn.childBegin)) * MP;
                    for(int i = 0; i < LUT.count(piece); ++i)
                        if(legalMid[base + i]) {
                            Placement const pl = LUT.nth(piece, i);
                            trace.add(F.board[k], static_cast<int>(piece), pl.o, pl.x, level, survived[base + i] != 0);
                        }
                }
            }
        }
        for(uint32_t k = 0; k < keptCount; ++k) // dead-end kept nodes become depth-`level` leaves (:549-551)
            if(!ws.hasChild[k]) {
                ++cnt.deadMid;
                consider(F.score[k], leafOrderCount + F.rootRank[k], 1);
            }
        // advance_frontier (:379-403)
        Frontier& N = ws.kept[cur ^ 1];
        for(uint32_t j = 0; j < nextCount; ++j) {
            uint32_t const s = ws.topSlots[j];
            uint32_t const k = s / (maxChildMid * MP);
            Node const& pn = tree.layers[level - 1][F.node[k]];
            int const ci = pn.childBegin + static_cast<int>((s / MP) % maxChildMid);
            Step const st = apply(F.board[k], F.clears[k], LUT.nth(slot[tree.layers[level][ci].slot], s % MP));
            ++cnt.reapplies;
            N.board[j] = st.board;
            N.clears[j] = st.clears;
            N.score[j] = ws.topScores[j];
            N.node[j] = static_cast<uint8_t>(ci);
            N.rootRank[j] = F.rootRank[k];
        }
        cur ^= 1;
        keptCount = nextCount;
    }

    // leaf stage: expand_leaves (:407-443); at a pruned ply only the pruner's top-k legal placements are scored
    {
        Frontier const& F = ws.kept[cur];
        int const leafPly = depth - 1;
        bool const pruneHere = pr.q != nullptr && ((pr.plyMask >> leafPly) & 1u) != 0;
        ws.seen.clear();
        if(deepOn) std::fill(std::begin(ws.deep.keyL), std::end(ws.deep.keyL), 0u);
        for(uint32_t j = 0; j < keptCount; ++j) {
            Node const& ln = tree.layers[depth - 2][F.node[j]];
            for(int ci = ln.childBegin; ci < ln.childEnd; ++ci) {
                Node const& cn = tree.layers[depth - 1][ci];
                uint32_t const piece = slot[cn.slot];
                bool const hI = heldIsI(cn);
                uint32_t const base = (j * maxLast + (ci - ln.childBegin)) * MP;
                int const n = LUT.count(piece);
                bool const rep = ws.seen.repeat(F.board[j], F.clears[j], piece, hI);
                if(deepOn) ws.deep.keyL[2 * j + (ci - ln.childBegin)] = piece | (hI ? 8u : 0u) | 16u;
                if(ws.check.on) ws.check.curL.repeat(F.board[j], F.clears[j], piece, hI);
                uint8_t legal[MAX_PLACEMENTS], keep[MAX_PLACEMENTS];
                for(int i = 0; i < n; ++i) {
                    legal[i] = fit(F.board[j], LUT.nth(piece, i).shape).legal ? 1 : 0;
                    keep[i] = legal[i];
                    ++cnt.applies;
                    cnt.repeatApplies += rep;
                }
                if(pruneHere) prune_list(pr, F.board[j], static_cast<int>(piece), leafPly, legal, n, keep);
                bool expanded = false;
                uint64_t listBest = 0;
                int listArg = -1;
                for(int i = 0; i < n; ++i) {
                    if(!legal[i]) continue;
                    expanded = true;
                    if(!keep[i]) continue;
                    ++cnt.legal;
                    cnt.repeatLegal += rep;
                    Step const s = apply(F.board[j], F.clears[j], LUT.nth(piece, i));
                    uint32_t const sc = ordered_bits(model.value(s.clears, s.board, hI));
                    consider(sc, base + i);
                    uint64_t const key = pack_key(sc, base + i, orderCount);
                    if(key > listBest) {
                        listBest = key;
                        listArg = i;
                    }
                }
                if(trace.on) // label = the list's best leaf
                    for(int i = 0; i < n; ++i)
                        if(legal[i]) {
                            Placement const pl = LUT.nth(piece, i);
                            trace.add(F.board[j], static_cast<int>(piece), pl.o, pl.x, leafPly, i == listArg);
                        }
                if(!expanded) {
                    ++cnt.deadLeaf;
                    consider(F.score[j], base, 2);
                }
            }
        }
    }

    // decode_move (:450-470)
    Frontier const& F = ws.kept[cur];
    uint32_t const order = key_pos(bestKey, orderCount);
    uint32_t const rootRank = order >= leafOrderCount ? order - leafOrderCount : F.rootRank[order / (maxLast * MP)];
    Node const& rn = tree.layers[0][ws.topRootSlots[rootRank] / MP];
    Placement const pl0 = LUT.nth(slot[rn.slot], ws.topRootSlots[rootRank] % MP);
    ++cnt.moves;
    if(bestFallback == 1) ++cnt.fallbackMidBest;
    if(bestFallback == 2) ++cnt.fallbackLeafBest;
    // the chosen root's level-1 lists, for the next move's RootReuse (rank = kept root index at level 1); a pruned
    // level 1 scores only some of them, and the device search then takes no roots from it
    bool const level1Pruned = pr.q != nullptr && ((pr.plyMask >> 1) & 1u) != 0;
    // the device keeps the level-1 scores of the best 16 kept roots, of the best 6 for the wide network (its LDS
    // budget)
    uint32_t const keptRoots = model.kind == TA3O_MODEL_F32W ? 6u : 16u;
    bool const keep = rootRank < keptRoots && !level1Pruned;
    ws.reuse.keys[0] = keep ? ws.reuse.level1[2 * rootRank] : 0u;
    ws.reuse.keys[1] = keep ? ws.reuse.level1[2 * rootRank + 1] : 0u;
    ws.deep.rank = rootRank; // DeepReuse applies to the next search where RootReuse does
    return {true, rn.rootHold, pl0.o, pl0.x, ordered_bits_inv(static_cast<uint32_t>(bestKey >> 32))};
}

// ---------------------------------------------------------------------------------------------------
// exhaustive search (ta3_ai/incl/ta3/ai/search/search.hpp:117-150, expand_suffix :80-102)
// ---------------------------------------------------------------------------------------------------
struct Seq { uint32_t pieces[MAX_DEPTH]; bool heldIsI[MAX_DEPTH]; bool rootHold; };

struct Exhaustive {
    ModelRef const& model;
    int depth;
    bool any = false;
    float best = 0.0f;
    int bo = 0, bx = 0;
    bool bh = false;
    int ro = 0, rx = 0;
    bool rh = false;

    void consider(float v) {
        if(!any || v > best) {
            any = true;
            best = v;
            bo = ro;
            bx = rx;
            bh = rh;
        }
    }
    bool expand(Seq const& seq, int level, Board const& b, uint32_t clears) {
        uint32_t const p = seq.pieces[level];
        bool anyFit = false;
        for(int i = 0; i < LUT.count(p); ++i) {
            Step const s = apply(b, clears, LUT.nth(p, i));
            if(!s.legal) continue;
            anyFit = true;
            if(level + 1 == depth || !expand(seq, level + 1, s.board, s.clears))
                consider(model.value(s.clears, canonical(s.board), seq.heldIsI[level]));
        }
        return anyFit;
    }
};

SearchResult search_exhaustive(Engine const& g, ModelRef const& model, int depth) {
    if(g.gameOver) return {false, false, 0, 0, 0.0f};
    // generate_sequences_into (variation_sequences.hpp:263-286)
    uint32_t slot[MAX_DEPTH + 2];
    slot[0] = g.held;
    slot[1] = g.queue[0];
    for(int i = 0; i < depth; ++i) slot[2 + i] = g.queue[1 + i];
    auto const plans = build_plans(depth, g.held == NO_PIECE);
    Exhaustive ex{model, depth};
    for(auto const& p : plans) {
        Seq seq{};
        for(int i = 0; i < depth; ++i) {
            seq.pieces[i] = slot[p.slot[i]];
            seq.heldIsI[i] = p.held[i] != HELD_NONE && slot[p.held[i]] == PIECE_I;
        }
        seq.rootHold = p.rootHold;
        uint32_t const p0 = seq.pieces[0];
        for(int i = 0; i < LUT.count(p0); ++i) {
            Placement const pl = LUT.nth(p0, i);
            Step const s = apply(g.board, 0, pl);
            if(!s.legal) continue;
            ex.ro = pl.o;
            ex.rx = pl.x;
            ex.rh = seq.rootHold;
            if(!ex.expand(seq, 1, s.board, s.clears))
                ex.consider(model.value(s.clears, canonical(s.board), seq.heldIsI[0]));
        }
    }
    if(!ex.any) return {false, false, 0, 0, 0.0f};
    return {true, ex.bh, ex.bo, ex.bx, ex.best};
}

// ---------------------------------------------------------------------------------------------------
// stats_v4 / score_v1 (ta3_ai/incl/ta3/ai/models/metrics/tetris_stats.hpp:19-23, tetris_stats_v4.hpp:8-15,
// metrics.hpp:94-127, score_v1.hpp:8-34)
// ---------------------------------------------------------------------------------------------------
struct Stats {
    uint32_t pieces = 0;
    uint64_t lines = 0;
    uint64_t holesPrev = 0, holesCur = 0;
    uint64_t clears[5] = {0, 0, 0, 0, 0};
    uint64_t sumMaxHeight = 0, sumHoleDepths = 0;

    void advance(Board const& b, uint32_t cleared) {
        ++pieces;
        lines += cleared;
        holesPrev = holesCur;
        holesCur = static_cast<uint64_t>(holes(b));
        ++clears[cleared];
        Metrics const m = metrics(b);
        sumMaxHeight += static_cast<uint64_t>(m.maxHeight);
        sumHoleDepths += static_cast<uint64_t>(m.holeDepths);
    }

    double score() const {
        static constexpr double PAYOUTS[] = {0.0, 0.2, 0.8, 2.0, 10.0};
        double clearBonus = 0;
        for(int k = 1; k < 5; ++k) clearBonus += PAYOUTS[k] * static_cast<double>(clears[k]);
        double const survival = 0.02 * static_cast<double>(pieces);
        double const avgMaxHeight = static_cast<double>(sumMaxHeight) / static_cast<double>(pieces);
        double const heightPenalty = std::max(avgMaxHeight - 12.0, 0.) * 1.0;
        double const avgHoleDepths = static_cast<double>(sumHoleDepths) / static_cast<double>(pieces);
        double const holesPenalty = avgHoleDepths * 1.5;
        return clearBonus + survival - heightPenalty - holesPenalty;
    }
};

const std::vector<uint32_t> REF_WIDTHS = {16, 32, 32}; // variation_sequences.hpp:22
uint32_t g_widths[3] = {16, 32, 32}; // widths the game loops search with (set before a run, read by its threads)

// eval_host inner loop for one (model, seed) game (model_eval_host.cpp:33-52)
float play_game(ModelRef const& model, uint64_t seed, uint32_t maxMoves, BeamWorkspace& ws, Counters& cnt,
                uint8_t* moves = nullptr, uint32_t* boards = nullptr, int* nRecorded = nullptr,
                PrunerRef const& pr = PrunerRef{}, std::vector<uint32_t>* traceOut = nullptr, uint32_t traceEvery = 0) {
    Engine g;
    engine_reset(g, seed);
    if(g_garbage) add_garbage_rows(g, seed);
    std::vector<uint32_t> const widths(g_widths, g_widths + 3);
    Stats stats;
    int rec = 0;
    ws.reuse.valid = false; // the first move has no previous search
    ws.check.cur2.clear();
    ws.check.curL.clear();
    for(uint32_t i = 0; i < maxMoves && !g.gameOver; ++i) {
        TraceOut tr;
        tr.out = traceOut;
        // traced moves: index congruent to the game seed modulo traceEvery (spreads them over games)
        tr.on = traceOut != nullptr && traceEvery != 0 && i % traceEvery == seed % traceEvery;
        SearchResult const r = search_beam(g, model, widths, ws, cnt, pr, tr);
        if(!r.some) break;
        if(r.hold) engine_hold(g);
        int const cleared = engine_place(g, r.o, r.x);
        ws.reuse.valid = cleared == 0;
        ++cnt.placements;
        if(moves) {
            moves[rec * 4 + 0] = r.hold ? 1 : 0;
            moves[rec * 4 + 1] = static_cast<uint8_t>(r.o);
            moves[rec * 4 + 2] = static_cast<uint8_t>(r.x + 2);
            moves[rec * 4 + 3] = cleared < 0 ? 255 : static_cast<uint8_t>(cleared);
            if(boards) std::memcpy(boards + rec * 10, g.board.c, 40);
            ++rec;
        }
        if(cleared < 0) break;
        stats.advance(g.board, static_cast<uint32_t>(cleared));
    }
    if(nRecorded) *nRecorded = rec;
    return stats.pieces == 0 ? -1.0e6f : static_cast<float>(stats.score());
}

Engine from_c(ta3o_engine const* e) {
    Engine g;
    std::memcpy(g.board.c, e->cols, 40);
    std::memcpy(g.queue, e->queue, sizeof(g.queue));
    g.held = e->held;
    g.gameOver = e->game_over != 0;
    std::memcpy(g.bag, e->bag, sizeof(g.bag));
    g.bagPos = e->bag_pos;
    std::memcpy(g.rng, e->rng, sizeof(g.rng));
    return g;
}

void to_c(Engine const& g, ta3o_engine* e) {
    std::memcpy(e->cols, g.board.c, 40);
    std::memcpy(e->queue, g.queue, sizeof(g.queue));
    e->held = g.held;
    e->game_over = g.gameOver ? 1u : 0u;
    std::memcpy(e->bag, g.bag, sizeof(g.bag));
    e->bag_pos = g.bagPos;
    std::memcpy(e->rng, g.rng, sizeof(g.rng));
}

void to_counters(Counters const& c, ta3o_counters* out) {
    if(!out) return;
    out->moves += c.moves;
    out->applies += c.applies;
    out->legal += c.legal;
    out->reapplies += c.reapplies;
    out->placements += c.placements;
    out->dead_mid += c.deadMid;
    out->dead_leaf += c.deadLeaf;
    out->fallback_mid_best += c.fallbackMidBest;
    out->fallback_leaf_best += c.fallbackLeafBest;
    out->repeat_applies += c.repeatApplies;
    out->repeat_legal += c.repeatLegal;
    out->reuse_applies += c.reuseApplies;
    out->reuse_legal += c.reuseLegal;
}

Board board_from(uint32_t const cols[10]) {
    Board b;
    std::memcpy(b.c, cols, 40);
    return b;
}

} // namespace

extern "C" {

int ta3o_num_params(int kind) {
    switch(kind) {
        case TA3O_MODEL_V11: return 97;
        case TA3O_MODEL_F16: return 1633;
        case TA3O_MODEL_HEIGHT: return 0;
        case TA3O_MODEL_F32W: return 20865;
        default: return -1;
    }
}

int ta3o_num_features(int kind) {
    switch(kind) {
        case TA3O_MODEL_V11: return 10;
        case TA3O_MODEL_F16: return 16;
        case TA3O_MODEL_F32W: return 32;
        default: return 0;
    }
}

void ta3o_rng_seed(uint64_t seed, uint64_t state[4]) { rng_seed(seed, state); }
uint64_t ta3o_rng_next(uint64_t state[4]) { return rng_next(state); }

void ta3o_expand_seeds(uint64_t seed, uint32_t games, uint64_t* out) {
    uint64_t st[4];
    rng_seed(seed, st);
    for(uint32_t i = 0; i < games; ++i) out[i] = rng_next(st);
}

int ta3o_board_available(const uint32_t cols[10], int type, int o, int x, int y) {
    return available(board_from(cols), type, o, x, y) ? 1 : 0;
}

int ta3o_board_drop_location(const uint32_t cols[10], int type, int o, int x, int y) {
    return y + drop_distance(board_from(cols), shape_at(type, o, x), y); // board2.hpp:277-279
}

void ta3o_board_place(uint32_t cols[10], int type, int o, int x, int y) {
    Board b = board_from(cols);
    place_shape(b, shape_at(type, o, x), y);
    std::memcpy(cols, b.c, 40);
}

int ta3o_board_clear_lines(uint32_t cols[10]) {
    Board b = board_from(cols);
    int const n = clear_lines(b);
    std::memcpy(cols, b.c, 40);
    return n;
}

int ta3o_board_full_lines(const uint32_t cols[10]) { return popc32(full_rows(board_from(cols))); }
int ta3o_board_holes(const uint32_t cols[10]) { return holes(board_from(cols)); }

int ta3o_board_fit(const uint32_t cols[10], int type, int o, int x, int* drop) {
    Fit const f = fit(board_from(cols), shape_at(type, o, x));
    if(drop) *drop = f.drop;
    return f.legal ? 1 : 0;
}

void ta3o_piece_shape(int type, int o, uint32_t* cols_nibbles, int* left) {
    *cols_nibbles = PT.shapeCols[type][o];
    *left = PT.shapeLeft[type][o];
}

void ta3o_engine_reset(ta3o_engine* e, uint64_t seed) {
    Engine g;
    engine_reset(g, seed);
    if(g_garbage) add_garbage_rows(g, seed);
    to_c(g, e);
}

void ta3o_set_garbage(int enabled) { g_garbage = enabled != 0; }

int ta3o_engine_place(ta3o_engine* e, int o, int x) {
    Engine g = from_c(e);
    int const r = engine_place(g, o, x);
    to_c(g, e);
    return r;
}

void ta3o_engine_hold(ta3o_engine* e) {
    Engine g = from_c(e);
    engine_hold(g);
    to_c(g, e);
}

int ta3o_lut_count(int piece) { return LUT.count(piece); }

void ta3o_lut_entry(int piece, int i, int* o, int* x) {
    Placement const p = LUT.nth(piece, i);
    *o = p.o;
    *x = p.x;
}

int ta3o_plans(int depth, int held_empty, uint8_t* slots, uint8_t* held_slots, uint8_t* root_hold) {
    if(depth < 2 || depth > MAX_DEPTH) return -1;
    auto const plans = build_plans(depth, held_empty != 0);
    for(size_t p = 0; p < plans.size(); ++p) {
        for(int i = 0; i < depth; ++i) {
            slots[p * depth + i] = plans[p].slot[i];
            held_slots[p * depth + i] = plans[p].held[i];
        }
        root_hold[p] = plans[p].rootHold ? 1 : 0;
    }
    return static_cast<int>(plans.size());
}

void ta3o_tree(int depth, int held_empty, uint8_t* nodes, uint8_t* counts) {
    Tree const& t = TREES.t[depth][held_empty ? 1 : 0];
    size_t const maxSeq = size_t{1} << depth;
    for(int l = 0; l < depth; ++l) {
        counts[l] = static_cast<uint8_t>(t.layers[l].size());
        for(size_t n = 0; n < t.layers[l].size(); ++n) {
            uint8_t* o = nodes + (l * maxSeq + n) * 6;
            Node const& v = t.layers[l][n];
            o[0] = v.slot;
            o[1] = v.parent;
            o[2] = v.childBegin;
            o[3] = v.childEnd;
            o[4] = v.rootHold ? 1 : 0;
            o[5] = v.heldSlot;
        }
    }
}

uint32_t ta3o_ordered_bits(float v) { return ordered_bits(v); }
float ta3o_ordered_bits_inv(uint32_t s) { return ordered_bits_inv(s); }

void ta3o_features(int kind, uint32_t clears, const uint32_t cols[10], int held_is_i, float* out) {
    Board const b = board_from(cols);
    if(kind == TA3O_MODEL_V11) features_v11(clears, b, held_is_i != 0, out);
    else if(kind == TA3O_MODEL_F16) features_f16(clears, b, held_is_i != 0, out);
    else if(kind == TA3O_MODEL_F32W) features_f32w(clears, b, held_is_i != 0, out);
}

float ta3o_model_value(int kind, const float* w, uint32_t clears, const uint32_t cols[10], int held_is_i) {
    ModelRef const m{kind, w};
    return m.value(clears, board_from(cols), held_is_i != 0);
}

float ta3o_f32w_value_fp32(const float* w, uint32_t clears, const uint32_t cols[10], int held_is_i) {
    float in[32], h1[128], h2[128], out[1];
    features_f32w(clears, board_from(cols), held_is_i != 0, in);
    dense_fma(in, 32, 128, true, w, h1);
    dense_fma(h1, 128, 128, true, w + 4224, h2);
    dense_fma(h2, 128, 1, false, w + 4224 + 16512, out);
    return out[0];
}

int ta3o_search_beam(const ta3o_engine* e, int kind, const float* w, const uint32_t* widths, int nwidths,
                     int* hold, int* o, int* x, float* value, ta3o_counters* counters) {
    if(nwidths < 1 || nwidths + 1 > MAX_DEPTH) return -1;
    std::vector<uint32_t> ks(widths, widths + nwidths);
    for(auto k : ks)
        if(k == 0) return -1;
    Engine const g = from_c(e);
    ModelRef const m{kind, w};
    BeamWorkspace ws;
    Counters c;
    SearchResult const r = search_beam(g, m, ks, ws, c);
    to_counters(c, counters);
    if(!r.some) return 0;
    *hold = r.hold ? 1 : 0;
    *o = r.o;
    *x = r.x;
    *value = r.value;
    return 1;
}

int ta3o_search_beam_pruned(const ta3o_engine* e, int kind, const float* w, const float* pruner_w, float act_range,
                            int k, unsigned ply_mask, int* hold, int* o, int* x, float* value) {
    std::vector<uint32_t> const ks{16u, 32u, 32u};
    Engine const g = from_c(e);
    ModelRef const m{kind, w};
    auto q = std::make_unique<PrunerQ>();
    pruner_quantize(pruner_w, act_range, *q);
    PrunerRef pr;
    pr.q = q.get();
    pr.k = k;
    pr.plyMask = ply_mask;
    BeamWorkspace ws;
    Counters c;
    SearchResult const r = search_beam(g, m, ks, ws, c, pr);
    if(!r.some) return 0;
    *hold = r.hold ? 1 : 0;
    *o = r.o;
    *x = r.x;
    *value = r.value;
    return 1;
}

uint32_t ta3o_select_top(uint32_t* scores, uint32_t n, uint32_t k, uint32_t* out_slots, uint32_t* out_scores) {
    return select_top(scores, n, k, out_slots, out_scores);
}

void ta3o_set_widths(const uint32_t* widths) {
    for(int l = 0; l < 3; ++l) g_widths[l] = widths ? widths[l] : REF_WIDTHS[l];
}

void ta3o_set_select_linear(int enabled) { g_selectLinear.store(enabled ? 1 : 0, std::memory_order_relaxed); }

int ta3o_search_exhaustive(const ta3o_engine* e, int kind, const float* w, int depth, int* hold, int* o, int* x,
                           float* value) {
    if(depth < 2 || depth > MAX_DEPTH) return -1;
    Engine const g = from_c(e);
    ModelRef const m{kind, w};
    SearchResult const r = search_exhaustive(g, m, depth);
    if(!r.some) return 0;
    *hold = r.hold ? 1 : 0;
    *o = r.o;
    *x = r.x;
    *value = r.value;
    return 1;
}

void ta3o_eval_host(int kind, const float* w, const uint64_t* seeds, uint32_t games, uint32_t max_moves, float* out,
                    ta3o_counters* counters) {
    ModelRef const m{kind, w};
    BeamWorkspace ws;
    Counters c;
    for(uint32_t g = 0; g < games; ++g) out[g] = play_game(m, seeds[g], max_moves, ws, c);
    to_counters(c, counters);
}

void ta3o_eval_host_mt(int kind, const float* w, uint32_t num_models, const uint64_t* seeds, uint32_t games,
                       uint32_t max_moves, int threads, float* out, uint32_t* placements_out,
                       ta3o_counters* counters) {
    int const P = ta3o_num_params(kind);
    uint64_t const items = static_cast<uint64_t>(num_models) * games;
    std::atomic<uint64_t> next{0};
    if(threads < 1) threads = 1;
    std::vector<Counters> perThread(static_cast<size_t>(threads));
    auto worker = [&](int tid) {
        BeamWorkspace ws;
        for(;;) {
            uint64_t const item = next.fetch_add(1);
            if(item >= items) break;
            uint64_t const mi = item / games, gi = item % games;
            ModelRef const m{kind, w + mi * static_cast<uint64_t>(P > 0 ? P : 0)};
            uint64_t const before = perThread[tid].placements;
            out[item] = play_game(m, seeds[gi], max_moves, ws, perThread[tid]);
            if(placements_out) placements_out[item] = static_cast<uint32_t>(perThread[tid].placements - before);
        }
    };
    std::vector<std::thread> pool;
    for(int t = 1; t < threads; ++t) pool.emplace_back(worker, t);
    worker(0);
    for(auto& t : pool) t.join();
    Counters total;
    for(auto const& c : perThread) total.add(c);
    to_counters(total, counters);
}

int ta3o_check_deep_reuse(int kind, const float* w, uint64_t seed, uint32_t max_moves, uint64_t* out) {
    ModelRef const m{kind, w};
    BeamWorkspace ws;
    ws.check.on = true;
    Counters c;
    play_game(m, seed, max_moves, ws, c);
    out[0] = ws.check.counted;
    out[1] = ws.check.mismatched;
    out[2] = c.moves;
    out[3] = c.reuseApplies;
    return 0;
}

int ta3o_play_trace(int kind, const float* w, uint64_t seed, uint32_t max_moves, uint8_t* moves, uint32_t* boards,
                    float* fitness) {
    ModelRef const m{kind, w};
    BeamWorkspace ws;
    Counters c;
    int rec = 0;
    float const f = play_game(m, seed, max_moves, ws, c, moves, boards, &rec);
    if(fitness) *fitness = f;
    return rec;
}

int ta3o_pruner_num_params(void) { return PR_PARAMS; }

float ta3o_pruner_score(const float* w, float act_range, const uint32_t cols[10], int piece, int o, int x, int ply) {
    auto q = std::make_unique<PrunerQ>();
    pruner_quantize(w, act_range, *q);
    return pruner_score(*q, board_from(cols), piece, o, x, ply);
}

void ta3o_pruner_scores(const float* w, float act_range, const uint32_t* samples, int n, float* out) {
    auto q = std::make_unique<PrunerQ>();
    pruner_quantize(w, act_range, *q);
    for(int t = 0; t < n; ++t) {
        uint32_t const* rec = samples + 11 * static_cast<size_t>(t);
        uint32_t const meta = rec[10];
        out[t] = pruner_score(*q, board_from(rec), meta & 7, (meta >> 3) & 3, static_cast<int>((meta >> 5) & 15) - 2,
                              (meta >> 9) & 3);
    }
}

void ta3o_pruner_quant_info(const float* w, float act_range, int* info) {
    auto q = std::make_unique<PrunerQ>();
    pruner_quantize(w, act_range, *q);
    info[0] = q->e1;
    info[1] = q->e2;
    info[2] = q->e3;
    info[3] = q->d;
}

double ta3o_pruner_loss_grad(const float* w, const uint32_t* samples, int n, double pos_weight, int rounded,
                             double* grad) {
    return pruner_loss_grad(w, samples, n, pos_weight, rounded != 0, grad);
}

void ta3o_eval_host_pruned_mt(int kind, const float* w, uint32_t num_models, const uint64_t* seeds, uint32_t games,
                              uint32_t max_moves, int threads, const float* pruner_w, float act_range, int k,
                              unsigned ply_mask, float* out, uint32_t* placements_out, ta3o_counters* counters) {
    int const P = ta3o_num_params(kind);
    uint64_t const items = static_cast<uint64_t>(num_models) * games;
    std::atomic<uint64_t> next{0};
    if(threads < 1) threads = 1;
    auto q = std::make_unique<PrunerQ>();
    PrunerRef pr;
    if(pruner_w) {
        pruner_quantize(pruner_w, act_range, *q);
        pr.q = q.get();
    }
    pr.k = k;
    pr.plyMask = ply_mask;
    std::vector<Counters> perThread(static_cast<size_t>(threads));
    auto worker = [&](int tid) {
        BeamWorkspace ws;
        for(;;) {
            uint64_t const item = next.fetch_add(1);
            if(item >= items) break;
            uint64_t const mi = item / games, gi = item % games;
            ModelRef const m{kind, w + mi * static_cast<uint64_t>(P > 0 ? P : 0)};
            uint64_t const before = perThread[tid].placements;
            out[item] = play_game(m, seeds[gi], max_moves, ws, perThread[tid], nullptr, nullptr, nullptr, pr);
            if(placements_out) placements_out[item] = static_cast<uint32_t>(perThread[tid].placements - before);
        }
    };
    std::vector<std::thread> pool;
    for(int t = 1; t < threads; ++t) pool.emplace_back(worker, t);
    worker(0);
    for(auto& t : pool) t.join();
    Counters total;
    for(auto const& c : perThread) total.add(c);
    to_counters(total, counters);
}

uint64_t ta3o_trace_records(int kind, const float* w, uint64_t seed, uint32_t max_moves, uint32_t every,
                            uint32_t* out, uint64_t cap_records) {
    ModelRef const m{kind, w};
    BeamWorkspace ws;
    Counters c;
    std::vector<uint32_t> rec;
    play_game(m, seed, max_moves, ws, c, nullptr, nullptr, nullptr, PrunerRef{}, &rec, every);
    uint64_t const n = rec.size() / 11;
    uint64_t const copy = n < cap_records ? n : cap_records;
    if(out) std::memcpy(out, rec.data(), copy * 11 * sizeof(uint32_t));
    return n;
}

} // extern "C"
