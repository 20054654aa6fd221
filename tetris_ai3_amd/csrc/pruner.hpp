// Placement-proposal network (pruner): layout constants and device entry points (no device code here).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace ta3h {

inline constexpr int PR_IN = 265, PR_H = 64;
inline constexpr int PR_W1 = 0, PR_B1 = PR_W1 + PR_H * PR_IN, PR_W2 = PR_B1 + PR_H, PR_B2 = PR_W2 + PR_H * PR_H,
                     PR_W3 = PR_B2 + PR_H, PR_B3 = PR_W3 + PR_H, PR_PARAMS = PR_B3 + 1;
inline constexpr int PR_SPLIT = 32; // sample ranges of a training batch's gradient sum (partial sums, PR_SPLIT x PR_PARAMS)

// Inference form of a pruner: layers 1 and 2 are an integer network on power-of-two grids, so every sum the
// BF16 MFMAs form is an exact integer below 2^24 and their FP32 accumulation is exact in any order; layer 3 is
// FP32 in a fixed order:
//   N1 = rne(W1 2^e1), e1 the largest exponent with max |W1| 2^e1 <= 255; Nb1 = rne(b1 2^e1), |Nb1| <= 2^20
//   P  = sum of N1 over the active inputs + Nb1                   (first-layer pre-activation, |P| < 2^21)
//   c  = min(255, max(0, rne(P 2^-d))); d = e1 - e2 clamped to [-8, 15], e2 the largest exponent with
//        actRange 2^e2 <= 255 (actRange = calibrated upper range of the first hidden layer)
//   N2 = rne(W2 2^e3), e3 from max |W2| <= 255; Nb2 = rne(b2 2^(e1 - d + e3)), |Nb2| <= 2^22
//   h2 = max(0, Nb2[v] + sum_u c[u] N2[v][u])                     (exact, |.| < 2^23)
//   logit = (p0 + p1) + (p2 + p3), p_q = FMA chain over v = 16 t + 4 q + j (t = 0..3, then j = 0..3) of
//           h2[v] W3[v] from 0 (FP32, W3 unquantized)
// b3 shifts every logit alike and is not used. Tables:
//  w1b[u][k]     N1 board part as bfloat16 (exact: |N1| <= 255), k < 240 (slots 240..255 zero) -- MFMA B
//                operand of the board projection
//  ptab[p][e][u] (Nb1[u] + N1[u][piece] + N1[u][rotation] + N1[u][column] + N1[u][ply p + 1]) 2^-d for
//                placement-list entry e (exact in FP32)
//  projScale     2^-d (the board projection is stored scaled, so P 2^-d = projection + ptab exactly)
//  w2f[v][q]     MFMA A fragments of layer 2 for output row v and lane quarter q: N2[v][8q .. 8q + 7] then
//                N2[v][32 + 8q .. 32 + 8q + 7] as bfloat16 (one 32-byte load per lane)
//  bw[t][q]      Nb2[16t + 4q .. + 3] (the layer-2 accumulator's initial values) then W3[16t + 4q .. + 3]
struct PrunerDev {
    uint16_t w1b[PR_H][256];
    float ptab[3][162][PR_H];
    uint16_t w2f[PR_H][4][16];
    float bw[4][4][8];
    float projScale;
    int32_t pad[15];
};
static_assert(offsetof(PrunerDev, w2f) % 32 == 0 && offsetof(PrunerDev, bw) % 32 == 0, "32-byte vector loads");

// builds the inference tables from the FP32 master weights; exps (optional) gets {e1, e2, e3, d}
void build_pruner_tables(float const* w, float actRange, PrunerDev& out, int* exps);

// training (pruner_train.hip)
hipError_t pruner_fwd_bwd(float const* w, uint32_t const* samples, uint32_t const* order, uint32_t n, float invBatch,
                          float posWeight, float* act, float* delta, float* dscore, float* loss, hipStream_t s);
// partial: PR_SPLIT x PR_PARAMS floats of scratch
hipError_t pruner_grad_adam(float* w, float* m, float* v, float* gradOut, float* partial, uint32_t const* samples,
                            uint32_t const* order, uint32_t n, float const* act, float const* delta,
                            float const* dscore, int adam, float lr, float beta1, float beta2, float eps, float corr1,
                            float corr2, hipStream_t s);
hipError_t pruner_score_samples(float const* w, uint32_t const* samples, uint32_t n, float* out, hipStream_t s);
// first hidden layer (after ReLU, as trained) of n samples: out[n][64]
hipError_t pruner_hidden_samples(float const* w, uint32_t const* samples, uint32_t n, float* out, hipStream_t s);
// integer-grid inference logits of n samples on the BF16 MFMAs (beam_kernel.hip)
hipError_t pruner_infer_samples(PrunerDev const* pr, uint32_t const* samples, uint32_t n, float* out, hipStream_t s);

} // namespace ta3h
