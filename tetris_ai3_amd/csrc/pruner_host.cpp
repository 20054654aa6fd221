// Host side of the pruner's inference form: quantization of the FP32 master weights onto power-of-two
// integer grids and the device tables built from them (layout and rules in pruner.hpp).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>

#include "pruner.hpp"
#include "tables.hpp"

namespace ta3h {
namespace {

// largest e with m 2^e <= lim; 0 for m = 0
int grid_exp(double m, double lim) {
    if(!(m > 0.0)) return 0;
    int e = static_cast<int>(std::floor(std::log2(lim / m)));
    while(std::ldexp(m, e) > lim) --e;     // guard the floating log against an off-by-one
    while(std::ldexp(m, e + 1) <= lim) ++e;
    return e;
}

// round(v 2^e) to nearest, ties to even, clamped to [-lim, lim]
int32_t to_grid(double v, int e, double lim) {
    double const r = std::nearbyint(std::ldexp(v, e));
    return static_cast<int32_t>(r < -lim ? -lim : (r > lim ? lim : r));
}

uint16_t bf16_bits_exact(int32_t n) { // n is a bfloat16-exact integer (|n| <= 256)
    float const f = static_cast<float>(n);
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return static_cast<uint16_t>(u >> 16);
}

double max_abs(float const* p, int n) {
    double m = 0.0;
    for(int i = 0; i < n; ++i) m = std::max(m, std::fabs(static_cast<double>(p[i])));
    return m;
}

} // namespace

void build_pruner_tables(float const* w, float actRange, PrunerDev& out, int* exps) {
    int const e1 = grid_exp(max_abs(w + PR_W1, PR_H * PR_IN), 255.0);
    int const e2 = grid_exp(static_cast<double>(actRange), 255.0);
    int const e3 = grid_exp(max_abs(w + PR_W2, PR_H * PR_H), 255.0);
    int const d = std::clamp(e1 - e2, -8, 15);
    if(exps) {
        exps[0] = e1;
        exps[1] = e2;
        exps[2] = e3;
        exps[3] = d;
    }
    std::memset(&out, 0, sizeof(out));
    auto n1 = [&](int u, int k) { return to_grid(w[PR_W1 + u * PR_IN + k], e1, 255.0); };
    for(int u = 0; u < PR_H; ++u) {
        for(int k = 0; k < 240; ++k) out.w1b[u][k] = bf16_bits_exact(n1(u, k));
        int32_t const nb1 = to_grid(w[PR_B1 + u], e1, 1048576.0);
        for(int ply = 1; ply <= 3; ++ply)
            for(int piece = 0; piece < 7; ++piece)
                for(int i = 0; i < TABLES.lutOff[piece + 1] - TABLES.lutOff[piece]; ++i) {
                    int const e = TABLES.lutOff[piece] + i;
                    int const o = TABLES.lut[e].ox & 3, xi = TABLES.lut[e].ox >> 2; // xi = x + 2
                    int32_t const p = nb1 + n1(u, 240 + piece) + n1(u, 247 + o) + n1(u, 251 + xi) + n1(u, 262 + ply - 1);
                    out.ptab[ply - 1][e][u] = static_cast<float>(std::ldexp(static_cast<double>(p), -d)); // exact
                }
        for(int i = 0; i < PR_H; ++i) { // input i sits in quarter (i % 32) / 8, half i / 32
            uint16_t const v = bf16_bits_exact(to_grid(w[PR_W2 + u * PR_H + i], e3, 255.0));
            out.w2f[u][(i % 32) / 8][8 * (i / 32) + i % 8] = v;
        }
        int const t = u / 16, q = (u % 16) / 4, j = u % 4; // output u = 16 t + 4 q + j
        out.bw[t][q][j] = static_cast<float>(to_grid(w[PR_B2 + u], e1 - d + e3, 4194304.0));
        out.bw[t][q][4 + j] = w[PR_W3 + u];
    }
    out.projScale = static_cast<float>(std::ldexp(1.0, -d));
}

} // namespace ta3h
