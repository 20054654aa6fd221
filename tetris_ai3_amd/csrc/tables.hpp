// Compile-time game tables of the evaluator: piece shapes, the deduplicated placement list and the
// hold-decision prefix trees. Everything here is constexpr; the device keeps one copy in constant memory.
//
// The tables are alike to the reference's by necessity (they define the game and the search order):
//  - piece block coordinates: ta3_sim/incl/ta3/sim/pieces/piece_offsets.hpp:19-85
//    (set(type, BOTTOM, TOP, LEFT, RIGHT) argument order);
//  - placement dedup and order: ta3_ai/incl/ta3/ai/search/placements.hpp:107-205;
//  - hold plans and prefix trees: ta3_ai/incl/ta3/ai/search/variation_sequences.hpp:216-255,
//    ta3_ai/incl/ta3/ai/search/beam.hpp:69-126.
#pragma once
#include <cstdint>

namespace ta3h {

inline constexpr int WIDTH = 10;
inline constexpr int HEIGHT = 23;
inline constexpr uint32_t FIELD = (1u << HEIGHT) - 1;
inline constexpr uint32_t FLOOR = 1u << HEIGHT;
inline constexpr int NUM_PIECES = 7;
inline constexpr uint32_t NO_PIECE = 7;
inline constexpr uint32_t PIECE_I = 0;
inline constexpr int MAX_PLACEMENTS = 34;
inline constexpr int LUT_TOTAL = 162;
inline constexpr int DEPTH = 4;          // plies searched per move
inline constexpr int NUM_SLOTS = DEPTH + 2;
inline constexpr int TREE_NODES = 16;    // 2^DEPTH plans per hold state
inline constexpr uint8_t HELD_NONE = 7;

struct LutEntry {
    uint16_t cols; // nibble c = occupied rows of board column col0 + c (bit 0 = top row of the piece box)
    int8_t col0;   // first occupied board column
    uint8_t ox;    // orientation in bits 0-1, window x + 2 in bits 2-5
};

struct TreeNode {
    uint8_t slot, parent, childBegin, childEnd, rootHold, heldSlot;
};

struct Tables {
    uint16_t shapeCols[NUM_PIECES][4];
    int8_t shapeLeft[NUM_PIECES][4];
    LutEntry lut[LUT_TOTAL];
    uint8_t lutOff[NUM_PIECES + 1];
    TreeNode tree[2][DEPTH][TREE_NODES]; // [held empty][layer][node]
    uint8_t treeCount[2][DEPTH];
};

namespace detail {

struct XY {
    int x, y;
};

// (type, orientation) -> 4 blocks; orientation TOP=0 RIGHT=1 BOTTOM=2 LEFT=3, type I=0 O=1 T=2 S=3 Z=4 J=5 L=6
constexpr XY block(int t, int o, int b) {
    // rows: one per type, BOTTOM/TOP/LEFT/RIGHT groups of 4 blocks as in the reference table
    constexpr XY data[7][4][4] = {
        /* I */ {{{0, 2}, {1, 2}, {2, 2}, {3, 2}}, {{0, 1}, {1, 1}, {2, 1}, {3, 1}},
                 {{1, 0}, {1, 1}, {1, 2}, {1, 3}}, {{2, 0}, {2, 1}, {2, 2}, {2, 3}}},
        /* O */ {{{1, 0}, {2, 0}, {1, 1}, {2, 1}}, {{1, 0}, {2, 0}, {1, 1}, {2, 1}},
                 {{1, 0}, {2, 0}, {1, 1}, {2, 1}}, {{1, 0}, {2, 0}, {1, 1}, {2, 1}}},
        /* T */ {{{0, 1}, {1, 1}, {2, 1}, {1, 2}}, {{1, 0}, {0, 1}, {1, 1}, {2, 1}},
                 {{1, 0}, {0, 1}, {1, 1}, {1, 2}}, {{1, 0}, {1, 1}, {2, 1}, {1, 2}}},
        /* S */ {{{1, 1}, {2, 1}, {1, 2}, {0, 2}}, {{1, 0}, {2, 0}, {0, 1}, {1, 1}},
                 {{1, 0}, {1, 1}, {2, 1}, {2, 2}}, {{0, 0}, {0, 1}, {1, 1}, {1, 2}}},
        /* Z */ {{{1, 1}, {0, 1}, {1, 2}, {2, 2}}, {{0, 0}, {1, 0}, {1, 1}, {2, 1}},
                 {{1, 0}, {0, 1}, {1, 1}, {0, 2}}, {{2, 0}, {1, 1}, {2, 1}, {1, 2}}},
        /* J */ {{{0, 1}, {1, 1}, {2, 1}, {2, 2}}, {{0, 0}, {0, 1}, {1, 1}, {2, 1}},
                 {{1, 0}, {1, 1}, {0, 2}, {1, 2}}, {{1, 0}, {2, 0}, {1, 1}, {1, 2}}},
        /* L */ {{{0, 1}, {1, 1}, {2, 1}, {0, 2}}, {{2, 0}, {0, 1}, {1, 1}, {2, 1}},
                 {{0, 0}, {1, 0}, {1, 1}, {1, 2}}, {{1, 0}, {1, 1}, {1, 2}, {2, 2}}},
    };
    // data groups are {BOTTOM, TOP, LEFT, RIGHT}; map orientation -> group
    constexpr int group[4] = {1, 3, 0, 2};
    return data[t][group[o]][b];
}

constexpr void box_columns(int t, int o, uint32_t (&box)[4]) {
    for(int c = 0; c < 4; ++c) box[c] = 0;
    for(int b = 0; b < 4; ++b) {
        XY const v = block(t, o, b);
        box[v.x] |= 1u << v.y;
    }
}

constexpr int ctz_c(uint32_t v) {
    int b = 0;
    while(((v >> b) & 1u) == 0) ++b;
    return b;
}

struct Plan {
    uint8_t slot[DEPTH];
    uint8_t held[DEPTH];
    bool rootHold;
};

constexpr void build_plans(bool heldEmpty, Plan (&plans)[TREE_NODES]) {
    for(uint32_t d = 0; d < (1u << DEPTH); ++d) {
        uint8_t cur = 1, hld = 0, k = 1;
        bool empty = heldEmpty;
        Plan p{};
        for(int i = 0; i < DEPTH; ++i) {
            bool didHold = false;
            if(d & (1u << i)) {
                if(empty) {
                    hld = cur;
                    cur = static_cast<uint8_t>(1 + k++);
                    empty = false;
                } else {
                    uint8_t const t = cur;
                    cur = hld;
                    hld = t;
                }
                didHold = true;
            }
            p.slot[i] = cur;
            if(i == 0) p.rootHold = didHold;
            p.held[i] = empty ? HELD_NONE : hld;
            if(i + 1 < DEPTH) cur = static_cast<uint8_t>(1 + k++);
        }
        plans[d] = p;
    }
}

constexpr void build_tree(Tables& t, int he) {
    Plan plans[TREE_NODES]{};
    build_plans(he == 1, plans);
    int planNode[TREE_NODES]{};
    for(int l = 0; l < DEPTH; ++l) t.treeCount[he][l] = 0;
    for(int pi = 0; pi < TREE_NODES; ++pi) {
        int n = 0;
        while(n < t.treeCount[he][0] && !(t.tree[he][0][n].slot == plans[pi].slot[0]
                                           && (t.tree[he][0][n].rootHold != 0) == plans[pi].rootHold))
            ++n;
        if(n == t.treeCount[he][0]) {
            t.tree[he][0][n] = TreeNode{plans[pi].slot[0], 0, 0, 0, static_cast<uint8_t>(plans[pi].rootHold ? 1 : 0),
                                        plans[pi].held[0]};
            ++t.treeCount[he][0];
        }
        planNode[pi] = n;
    }
    for(int level = 1; level < DEPTH; ++level) {
        int nextNode[TREE_NODES]{};
        for(int parent = 0; parent < t.treeCount[he][level - 1]; ++parent) {
            t.tree[he][level - 1][parent].childBegin = t.treeCount[he][level];
            for(int pi = 0; pi < TREE_NODES; ++pi) {
                if(planNode[pi] != parent) continue;
                int n = t.tree[he][level - 1][parent].childBegin;
                while(n < t.treeCount[he][level] && t.tree[he][level][n].slot != plans[pi].slot[level]) ++n;
                if(n == t.treeCount[he][level]) {
                    t.tree[he][level][n] = TreeNode{plans[pi].slot[level], static_cast<uint8_t>(parent), 0, 0, 0,
                                                    plans[pi].held[level]};
                    ++t.treeCount[he][level];
                }
                nextNode[pi] = n;
            }
            t.tree[he][level - 1][parent].childEnd = t.treeCount[he][level];
        }
        for(int pi = 0; pi < TREE_NODES; ++pi) planNode[pi] = nextNode[pi];
    }
}

} // namespace detail

constexpr Tables build_tables() {
    Tables t{};
    for(int p = 0; p < NUM_PIECES; ++p)
        for(int o = 0; o < 4; ++o) {
            uint32_t box[4]{};
            detail::box_columns(p, o, box);
            int left = 0;
            while(left < 4 && box[left] == 0) ++left;
            int right = 4;
            while(right > left && box[right - 1] == 0) --right;
            uint32_t cols = 0;
            for(int c = left; c < right; ++c) cols |= box[c] << (4 * (c - left));
            t.shapeCols[p][o] = static_cast<uint16_t>(cols);
            t.shapeLeft[p][o] = static_cast<int8_t>(left);
        }
    int w = 0;
    for(int p = 0; p < NUM_PIECES; ++p) {
        t.lutOff[p] = static_cast<uint8_t>(w);
        uint32_t seen[4][4]{};
        int seenN[4]{};
        int nSeen = 0;
        for(int o = 0; o < 4; ++o) {
            uint32_t box[4]{};
            detail::box_columns(p, o, box);
            int const left = t.shapeLeft[p][o];
            int n = 0;
            for(int c = 0; c < 4; ++c)
                if((t.shapeCols[p][o] >> (4 * c)) & 0xF) n = c + 1;
            uint32_t sig[4]{};
            int minbit = 32;
            for(int c = 0; c < n; ++c) {
                sig[c] = box[left + c];
                if(sig[c] != 0 && detail::ctz_c(sig[c]) < minbit) minbit = detail::ctz_c(sig[c]);
            }
            if(minbit == 32) minbit = 0;
            for(int c = 0; c < n; ++c) sig[c] >>= minbit;
            bool dup = false;
            for(int j = 0; j < nSeen && !dup; ++j) {
                if(seenN[j] != n) continue;
                bool same = true;
                for(int c = 0; c < n; ++c) same = same && seen[j][c] == sig[c];
                dup = same;
            }
            if(dup) continue;
            for(int c = 0; c < 4; ++c) seen[nSeen][c] = sig[c];
            seenN[nSeen++] = n;
            for(int x = -left; x <= WIDTH - left - n; ++x) {
                t.lut[w].cols = t.shapeCols[p][o];
                t.lut[w].col0 = static_cast<int8_t>(x + left);
                t.lut[w].ox = static_cast<uint8_t>(o | ((x + 2) << 2));
                ++w;
            }
        }
    }
    t.lutOff[NUM_PIECES] = static_cast<uint8_t>(w);
    detail::build_tree(t, 0);
    detail::build_tree(t, 1);
    return t;
}

inline constexpr Tables TABLES = build_tables();

static_assert(TABLES.lutOff[NUM_PIECES] == LUT_TOTAL, "162 deduplicated placements");
static_assert(TABLES.lutOff[1] - TABLES.lutOff[0] == 17 && TABLES.lutOff[2] - TABLES.lutOff[1] == 9
                  && TABLES.lutOff[3] - TABLES.lutOff[2] == 34 && TABLES.lutOff[4] - TABLES.lutOff[3] == 17
                  && TABLES.lutOff[5] - TABLES.lutOff[4] == 17 && TABLES.lutOff[6] - TABLES.lutOff[5] == 34
                  && TABLES.lutOff[7] - TABLES.lutOff[6] == 34,
              "placement counts I/S/Z=17, O=9, T/J/L=34");
static_assert(TABLES.treeCount[0][0] == 2 && TABLES.treeCount[1][0] == 2 && TABLES.treeCount[0][3] == 16
                  && TABLES.treeCount[1][3] == 16,
              "two root groups and sixteen leaves per hold state");

} // namespace ta3h
