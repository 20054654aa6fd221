// Instrumented scalar op counts of one candidate of the search (TEST INFRASTRUCTURE ONLY: the op model that
// bench.py prices its roofline with is frozen from these counts).
//
// A restatement of the candidate path in which every scalar operation on the algorithm's data is counted: each
// integer or float add, subtract, multiply, divide, fused multiply-add, and / or / xor / not, shift, compare,
// select, min / max, popcount, count-trailing-zeros / bit-width, square root and int -> float conversion counts
// one. Loop control, addressing, loads and stores count nothing. Common subexpressions are computed once (a
// column's height and popcount feed every metric, as the reference's fused pass intends).
//   fit        Board2::fit (ta3_sim/incl/ta3/sim/board2.hpp:306-320) over the columns the piece occupies: the
//              others hold a zero nibble and change neither the collision nor the drop
//   place      Board2::placed (:322-327) over the occupied columns, then clearLines (:333-342, :218-234) and the
//              clear histogram bump (clear_hist_t::added, ta3_ai/incl/ta3/ai/model_defs.hpp:53-55)
//   features   V11: ModelV11::extractInputs (ta3_ai/incl/ta3/ai/models/models/model_v11.hpp:63-91) over the fused
//              metrics (ta3_ai/incl/ta3/ai/models/metrics/v_board_metrics.hpp:41-56, 130-232); 16 features: the
//              same pass with the five metrics the 16-feature set adds
//   network    dense layers (ta3_ai/incl/ta3/ai/models/models/utility.hpp:29-46): every multiply-add of the
//              network (the device's exact zero skipping is an implementation of the same network)
// The restated results equal the oracle's (tests/test_oracle_opcount.py checks fit, drop, board, clears and
// features against the oracle on every record it counts).
#include <cmath>
#include <cstdint>
#include <cstring>

#include "ta3_oracle.h"

namespace {

constexpr int WIDTH = 10, HEIGHT = 23;
constexpr uint32_t FIELD = (1u << HEIGHT) - 1, FLOOR = 1u << HEIGHT;

thread_local uint64_t g_ops = 0;

// counted 32-bit unsigned / signed integers and floats
struct U { uint32_t v; };
struct I { int32_t v; };
struct F { float v; };

inline U op(uint32_t v) { ++g_ops; return {v}; }
inline I opi(int32_t v) { ++g_ops; return {v}; }
inline F opf(float v) { ++g_ops; return {v}; }
inline bool opb(bool v) { ++g_ops; return v; }

inline U operator&(U a, U b) { return op(a.v & b.v); }
inline U operator|(U a, U b) { return op(a.v | b.v); }
inline U operator^(U a, U b) { return op(a.v ^ b.v); }
inline U operator~(U a) { return op(~a.v); }
inline U operator+(U a, U b) { return op(a.v + b.v); }
inline U operator-(U a, U b) { return op(a.v - b.v); }
inline U operator<<(U a, I s) { return op(a.v << s.v); }
inline U operator>>(U a, I s) { return op(a.v >> s.v); }
inline I operator+(I a, I b) { return opi(a.v + b.v); }
inline I operator-(I a, I b) { return opi(a.v - b.v); }
inline I operator*(I a, I b) { return opi(a.v * b.v); }
inline I operator>>(I a, I s) { return opi(a.v >> s.v); }
inline bool operator<(I a, I b) { return opb(a.v < b.v); }
inline bool operator>=(I a, I b) { return opb(a.v >= b.v); }
inline bool operator!=(U a, U b) { return opb(a.v != b.v); }
inline bool operator==(U a, U b) { return opb(a.v == b.v); }
inline F operator+(F a, F b) { return opf(a.v + b.v); }
inline F operator*(F a, F b) { return opf(a.v * b.v); }
inline F operator/(F a, F b) { return opf(a.v / b.v); }

inline U K(uint32_t v) { return {v}; } // a constant: no operation
inline I Ki(int32_t v) { return {v}; }
inline F Kf(float v) { return {v}; }
inline U sel(bool c, U a, U b) { ++g_ops; return c ? a : b; }
inline I seli(bool c, I a, I b) { ++g_ops; return c ? a : b; }
inline I imin(I a, I b) { return opi(a.v < b.v ? a.v : b.v); }
inline I imax(I a, I b) { return opi(a.v > b.v ? a.v : b.v); }
inline I popc(U a) { return opi(__builtin_popcount(a.v)); }
inline I ctz(U a) { return opi(a.v == 0 ? 32 : __builtin_ctz(a.v)); }
inline I bitw(U a) { return opi(a.v == 0 ? 0 : 32 - __builtin_clz(a.v)); }
inline I as_i(U a) { return {static_cast<int32_t>(a.v)}; } // reinterpretation: no operation
inline F to_f(I a) { return opf(static_cast<float>(a.v)); }
inline F fsqrt(F a) { return opf(a.v <= 0.0f ? 0.0f : std::sqrt(a.v)); } // v_board_metrics.hpp:22-36 (one compare folded)
inline F fma_(F a, F b, F c) { return opf(std::fma(a.v, b.v, c.v)); }
inline F relu(F a) { return opf(0.0f < a.v ? a.v : 0.0f); }

struct Shape { uint32_t cols; int col0; };

// the piece's nibble at board column x (board2.hpp:165-168)
inline U column_of(Shape s, int x) {
    I const d = Ki(x) - Ki(s.col0);
    bool const in = opb(static_cast<uint32_t>(d.v) < 4u);
    U const nib = (K(s.cols) >> (d * Ki(4))) & K(0xFu);
    return sel(in, nib, K(0));
}

// columns [lo, hi] the piece occupies
inline void span_of(Shape s, int& lo, int& hi) {
    lo = s.col0;
    hi = s.col0 + ((32 - __builtin_clz(s.cols)) + 3) / 4 - 1;
}

bool fit(U const (&b)[WIDTH], Shape s, int& dropOut) {
    int lo, hi;
    span_of(s, lo, hi);
    U hit = K(0);
    I minDrop = Ki(HEIGHT);
    for(int x = lo; x <= hi; ++x) {
        U const nib = column_of(s, x);
        hit = hit | (b[x] & nib);
        I const low = bitw(nib) - Ki(1);
        U const mask = ~((K(1) << (low + Ki(1))) - K(1));
        I const drop = ctz((b[x] & mask) | K(FLOOR)) - low - Ki(1);
        minDrop = seli(nib != K(0), imin(minDrop, drop), minDrop);
    }
    dropOut = minDrop.v;
    return hit == K(0);
}

int place_clear(U (&b)[WIDTH], Shape s, int y) {
    int lo, hi;
    span_of(s, lo, hi);
    for(int x = lo; x <= hi; ++x) b[x] = b[x] | (column_of(s, x) << Ki(y));
    U full = K(FIELD);
    for(int x = 0; x < WIDTH; ++x) full = full & b[x];
    if(full == K(0)) return 0;
    I const n = popc(full);
    for(int x = 0; x < WIDTH; ++x) {
        U col = b[x], rows = full;
        for(int i = 0; i < n.v; ++i) {
            U const row = rows & (~rows + K(1));
            U const above = col & (row - K(1));
            U const below = col & ~((row << Ki(1)) - K(1));
            col = (above << Ki(1)) | below;
            rows = rows ^ row;
        }
        b[x] = col;
    }
    return n.v;
}

constexpr uint32_t PLANES[5] = {0x2AAAAAu, 0x4CCCCCu, 0x70F0F0u, 0x00FF00u, 0x7F0000u}; // bit b of the row index

// the fused column pass (v_board_metrics.hpp:130-225); `wide` adds the 16-feature set's metrics
void features(U const (&b)[WIDTH], uint32_t clears, bool heldIsI, bool wide, float* out) {
    I holes = Ki(0), agg = Ki(0), maxH = Ki(0), minH = Ki(HEIGHT), sv = Ki(0), ytr = Ki(0), hd = Ki(0), bump = Ki(0);
    I h[WIDTH];
    for(int x = 0; x < WIDTH; ++x) {
        U const col = b[x];
        h[x] = Ki(HEIGHT) - ctz(col | K(FLOOR));
        I const n = popc(col);
        holes = holes + (h[x] - n);
        maxH = imax(maxH, h[x]);
        if(x > 0) {
            I const d = h[x] - h[x - 1];
            sv = sv + d * d;
            if(wide) {
                I const ad = seli(d < Ki(0), Ki(0) - d, d);
                bump = bump + ad;
            }
        }
        I s = Ki(0); // sum_set_bit_indices (:41-56): one popcount per bit plane of the row index
        for(int p = 0; p < 5; ++p) s = s + (popc(col & K(PLANES[p])) * Ki(1 << p));
        I const pos = Ki(HEIGHT - 1) * n + ((n * (n + Ki(1))) >> Ki(1));
        hd = hd + (pos - (s + n * n));
        if(wide) {
            agg = agg + h[x];
            minH = imin(minH, h[x]);
            U const yc = col | K(FLOOR);
            ytr = ytr + popc((yc ^ (yc >> Ki(1))) & K(FIELD));
        }
    }
    I wells = Ki(0); // wells_t (:204-225), WELL_DEPTH 3
    for(int x = 0; x < WIDTH; ++x) {
        bool well;
        if(x == 0) well = h[1] >= h[0] + Ki(3);
        else if(x == WIDTH - 1) well = h[WIDTH - 2] >= h[WIDTH - 1] + Ki(3);
        else {
            I const lim = h[x] + Ki(3);
            well = opb((h[x - 1] >= lim) & (h[x + 1] >= lim));
        }
        wells = wells + seli(well, Ki(1), Ki(0));
    }
    U const hist = K(clears);
    I cnt[4];
    for(int k = 0; k < 4; ++k) {
        cnt[k] = as_i((hist >> Ki(8 * k)) & K(0xFFu));
        out[k] = (to_f(cnt[k]) * Kf(0.25f)).v;
    }
    out[4] = (fsqrt(to_f(sv)) / Kf(100.0f)).v;
    out[5] = (to_f(holes) / Kf(20.0f)).v;
    out[6] = (to_f(hd) / Kf(200.0f)).v;
    out[7] = to_f(maxH).v;
    out[8] = heldIsI ? 1.0f : 0.0f;
    ++g_ops; // the held-I select
    out[9] = (to_f(wells) / Kf(5.0f)).v;
    if(wide) {
        out[10] = (to_f(agg) / Kf(200.0f)).v;
        out[11] = (to_f(ytr) / Kf(40.0f)).v;
        out[12] = (to_f(bump) / Kf(100.0f)).v;
        I const lines = cnt[0] + cnt[1] * Ki(2) + cnt[2] * Ki(3) + cnt[3] * Ki(4);
        out[13] = (to_f(lines) * Kf(0.1f)).v;
        I const cs = Ki(0) - cnt[0] - cnt[1] * Ki(2) - cnt[2] * Ki(3) + cnt[3] * Ki(10);
        out[14] = (to_f(cs) * Kf(0.1f)).v;
        out[15] = to_f(maxH - minH).v;
    }
}

// the value network: V11 10 -> 8 -> 1 with separately rounded multiply and add; 16 -> 32 -> 32 -> 1 with FMAs
float network(int kind, float const* w, float const* f) {
    if(kind == TA3O_MODEL_V11) {
        F out = Kf(w[96]);
        for(int o = 0; o < 8; ++o) {
            F v = Kf(w[80 + o]);
            for(int i = 0; i < 10; ++i) v = v + Kf(f[i]) * Kf(w[o * 10 + i]);
            out = out + relu(v) * Kf(w[88 + o]);
        }
        return out.v;
    }
    F h1[32], h2[32];
    for(int o = 0; o < 32; ++o) {
        F v = Kf(w[512 + o]);
        for(int i = 0; i < 16; ++i) v = fma_(Kf(f[i]), Kf(w[o * 16 + i]), v);
        h1[o] = relu(v);
    }
    for(int o = 0; o < 32; ++o) {
        F v = Kf(w[544 + 1024 + o]);
        for(int i = 0; i < 32; ++i) v = fma_(h1[i], Kf(w[544 + o * 32 + i]), v);
        h2[o] = relu(v);
    }
    F out = Kf(w[1632]);
    for(int i = 0; i < 32; ++i) out = fma_(h2[i], Kf(w[1600 + i]), out);
    return out.v;
}

} // namespace

extern "C" void ta3o_candidate_ops(int kind, const float* w, const uint32_t* samples, int n, uint64_t* ops_out,
                                   int32_t* legal_out, uint32_t* boards_out, float* features_out, float* values_out) {
    uint64_t sums[4] = {0, 0, 0, 0};
    int const nf = kind == TA3O_MODEL_V11 ? 10 : 16;
    for(int r = 0; r < n; ++r) {
        uint32_t const* rec = samples + 11 * r;
        uint32_t const meta = rec[10];
        int const piece = static_cast<int>(meta & 7u), o = static_cast<int>((meta >> 3) & 3u);
        int const x = static_cast<int>((meta >> 5) & 15u) - 2;
        uint32_t shapeCols = 0;
        int left = 0;
        ta3o_piece_shape(piece, o, &shapeCols, &left);
        Shape const s{shapeCols, x + left};
        U b[WIDTH];
        for(int c = 0; c < WIDTH; ++c) b[c] = K(rec[c]);
        g_ops = 0;
        int drop = 0;
        bool const legal = fit(b, s, drop);
        sums[0] += g_ops;
        legal_out[r] = legal ? drop : -1;
        if(!legal) continue;
        g_ops = 0;
        int const cleared = place_clear(b, s, drop);
        uint32_t const clears = cleared == 0 ? 0u : (1u << (8 * (cleared - 1)));
        ++g_ops; // the histogram bump
        sums[1] += g_ops;
        for(int c = 0; c < WIDTH; ++c) boards_out[WIDTH * r + c] = b[c].v;
        float f[16];
        g_ops = 0;
        features(b, clears, false, kind != TA3O_MODEL_V11, f);
        sums[2] += g_ops;
        std::memcpy(features_out + 16 * r, f, sizeof(float) * nf);
        g_ops = 0;
        values_out[r] = network(kind, w, f);
        sums[3] += g_ops;
    }
    for(int k = 0; k < 4; ++k) ops_out[k] = sums[k];
}
