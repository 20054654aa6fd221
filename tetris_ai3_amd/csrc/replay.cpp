// Host replay of a device move log, for previews (the reference's AiPreview, ta3_ai/src/preview.cpp:40-94, draws
// the boards its games pass through; here the games are played on the GPU with the move log on, and this replays
// them on the host to recover every board). Not on the search path: the engine rules are the ones the kernel
// plays (ta3_sim/incl/ta3/sim/tetris_engine.hpp:103-161, board2.hpp:306-342), over the same constexpr tables.
#include <cstdint>

#include "ta3h.h"
#include "tables.hpp"

namespace {

using ta3h::FIELD;
using ta3h::FLOOR;
using ta3h::HEIGHT;
using ta3h::TABLES;
using ta3h::WIDTH;

struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed) { // splitmix64 seeding (xoshiro256ss.hpp:20-28)
        for(auto& w : s) {
            seed += 0x9e3779b97f4a7c15ULL;
            uint64_t z = seed;
            z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
            z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
            w = z ^ (z >> 31);
        }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() { // xoshiro256** (:33-45)
        uint64_t const r = rotl(s[1] * 5, 7) * 9;
        uint64_t const t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
};

struct Game {
    uint32_t cols[WIDTH] = {};
    uint32_t queue[6] = {};
    uint32_t held = ta3h::NO_PIECE;
    uint32_t bag[7] = {};
    uint32_t bagPos = 0;
    Rng rng;
    bool over = false;

    explicit Game(uint64_t seed) : rng(seed) {
        for(auto& q : queue) q = draw();
        check();
    }
    uint32_t draw() { // 7-bag consumed from the back (tetris_engine.hpp:116-129)
        if(bagPos > 0) return bag[--bagPos];
        for(uint32_t i = 0; i < 7; ++i) bag[i] = i;
        for(uint32_t i = 6; i > 0; --i) {
            uint32_t const j = static_cast<uint32_t>(rng.next() % (i + 1));
            uint32_t const t = bag[i];
            bag[i] = bag[j];
            bag[j] = t;
        }
        bagPos = 6;
        return bag[6];
    }
    static uint32_t nibble(uint32_t cols, int col0, int x) {
        uint32_t const d = static_cast<uint32_t>(x - col0);
        return d < 4 ? (cols >> (4 * d)) & 0xFu : 0u;
    }
    void check() { // the current piece must fit at the spawn window {3, 0} (tetris_engine.hpp:159-161)
        uint32_t const t = queue[0];
        uint32_t const cols = TABLES.shapeCols[t][0];
        int const col0 = 3 + TABLES.shapeLeft[t][0];
        int const w = (32 - __builtin_clz(cols) + 3) / 4;
        bool ok = col0 >= 0 && col0 + w <= WIDTH;
        for(int x = 0; x < WIDTH && ok; ++x) ok = (cols_at(x) & nibble(cols, col0, x)) == 0;
        over = !ok;
    }
    uint32_t cols_at(int x) const { return cols[x]; }
    void advance() {
        for(int i = 1; i < 6; ++i) queue[i - 1] = queue[i];
        queue[5] = draw();
    }
    void hold() { // tetris_engine.hpp:148-157
        if(held == ta3h::NO_PIECE) {
            held = queue[0];
            advance();
        } else {
            uint32_t const t = held;
            held = queue[0];
            queue[0] = t;
        }
        check();
    }
    int place(int o, int x) { // hard drop, lock, clear (tetris_engine.hpp:137-146); -1 = the piece did not fit
        uint32_t const t = queue[0];
        uint32_t const shape = TABLES.shapeCols[t][o];
        int const col0 = x + TABLES.shapeLeft[t][o];
        int const w = (32 - __builtin_clz(shape) + 3) / 4;
        if(col0 < 0 || col0 + w > WIDTH) return -1; // off the board: no placement of the search
        int minDrop = HEIGHT;
        for(int c = 0; c < WIDTH; ++c) {
            uint32_t const nib = nibble(shape, col0, c);
            if(nib == 0) continue;
            if((cols[c] & nib) != 0) return -1; // collides at the spawn row
            int const low = 31 - __builtin_clz(nib);
            uint32_t const below = cols[c] & ~((1u << (low + 1)) - 1);
            int const drop = __builtin_ctz(below | FLOOR) - low - 1;
            minDrop = drop < minDrop ? drop : minDrop;
        }
        for(int c = 0; c < WIDTH; ++c) cols[c] |= nibble(shape, col0, c) << minDrop;
        uint32_t full = FIELD;
        for(uint32_t c : cols) full &= c;
        int const cleared = __builtin_popcount(full);
        for(auto& col : cols) { // compact each column past its full rows, top-most first (board2.hpp:218-234)
            uint32_t rows = full;
            while(rows) {
                uint32_t const row = rows & (~rows + 1);
                col = ((col & (row - 1)) << 1) | (col & ~((row << 1) - 1));
                rows ^= row;
            }
        }
        advance();
        check();
        return cleared;
    }
    void garbage(uint64_t seed) { // the build's start-position extension (as the kernel's add_garbage_rows)
        Rng r(seed ^ 0x6A09E667F3BCC908ULL);
        uint32_t const n = static_cast<uint32_t>(r.next() % 5u);
        for(uint32_t k = 0; k < n; ++k) {
            uint32_t const hole = static_cast<uint32_t>(r.next() % WIDTH);
            for(uint32_t x = 0; x < WIDTH; ++x)
                if(x != hole) cols[x] |= 1u << (HEIGHT - 1 - k);
        }
        check();
    }
};

} // namespace

extern "C" int ta3h_replay(uint64_t seed, int garbage, const uint8_t* moves, uint32_t n, uint32_t* boards,
                           uint8_t* pieces, uint8_t* cleared) {
    if(n > 0 && moves == nullptr) return -1;
    Game g(seed);
    if(garbage) g.garbage(seed);
    uint32_t i = 0;
    for(; i < n && !g.over; ++i) {
        uint32_t const b = moves[i];
        if(b & 0x80u) g.hold();
        uint32_t const piece = g.queue[0];
        int const c = g.place(static_cast<int>((b >> 5) & 3u), static_cast<int>(b & 31u) - 2);
        if(c < 0) return -1;
        if(boards)
            for(int x = 0; x < WIDTH; ++x) boards[WIDTH * i + x] = g.cols[x];
        if(pieces) pieces[i] = static_cast<uint8_t>(piece);
        if(cleared) cleared[i] = static_cast<uint8_t>(c);
    }
    return static_cast<int>(i);
}
