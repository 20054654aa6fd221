// Placement-proposal network (pruner): layout constants and device entry points (no device code here).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace ta3h {

inline constexpr int PR_IN = 265, PR_H = 64;
inline constexpr int PR_W1 = 0, PR_B1 = PR_W1 + PR_H * PR_IN, PR_W2 = PR_B1 + PR_H, PR_B2 = PR_W2 + PR_H * PR_H,
                     PR_W3 = PR_B2 + PR_H, PR_B3 = PR_W3 + PR_H, PR_PARAMS = PR_B3 + 1;

// Inference form of a pruner: an integer network on power-of-two grids, so every sum the device forms is an
// exact integer below 2^24 and the FP32 accumulation of the BF16 MFMAs is exact in any order:
//   N1 = rne(W1 2^e1), e1 the largest exponent with max |W1| 2^e1 <= 255; Nb1 = rne(b1 2^e1), |Nb1| <= 2^20
//   P  = sum of N1 over the active inputs + Nb1                   (first-layer pre-activation, |P| < 2^21)
//   c  = 0 if P <= 0, else min(255, round-half-up(P 2^-d)); d = e1 - e2 clamped to [-8, 24], e2 the largest
//        exponent with actRange 2^e2 <= 255 (actRange = calibrated upper range of the first hidden layer)
//   N2 = rne(W2 2^e3), e3 from max |W2| <= 255; Nb2 = rne(b2 2^(e2 + e3)), |Nb2| <= 2^22
//   h2 = max(0, sum_u c[u] N2[v][u] + Nb2[v])                     (|.| < 2^23)
//   N3 = rne(W3 2^e4), e4 from max |W3| <= 32767
//   logit = float(sum_v h2[v] N3[v])                              (exact below 2^45 in FP64, rounded once)
// b3 shifts every logit alike and is not used. Tables:
//  w1b[u][k]     N1 board part as bfloat16 (exact: |N1| <= 255), k < 240 (slots 240..255 zero) -- MFMA B
//                operand of the board projection
//  ptab[p][e][u] Nb1[u] + N1[u][piece] + N1[u][rotation] + N1[u][column] + N1[u][ply p + 1] for placement-list
//                entry e (exact integers in FP32)
//  w2b[v][u]     N2 as bfloat16 -- MFMA A operand of layer 2
//  b2[v], w3[v]  Nb2 and N3 (exact integers in FP32); actShift = d
struct PrunerDev {
    uint16_t w1b[PR_H][256];
    float ptab[3][162][PR_H];
    uint16_t w2b[PR_H][PR_H];
    float b2[PR_H];
    float w3[PR_H];
    int32_t actShift;
    int32_t pad[15];
};

// builds the inference tables from the FP32 master weights; exps (optional) gets {e1, e2, e3, e4, d}
void build_pruner_tables(float const* w, float actRange, PrunerDev& out, int* exps);

// training (pruner_train.hip)
hipError_t pruner_fwd_bwd(float const* w, uint32_t const* samples, uint32_t const* order, uint32_t n, float invBatch,
                          float posWeight, float* act, float* delta, float* dscore, float* loss, hipStream_t s);
hipError_t pruner_grad_adam(float* w, float* m, float* v, float* gradOut, uint32_t const* samples,
                            uint32_t const* order, uint32_t n, float const* act, float const* delta,
                            float const* dscore, int adam, float lr, float beta1, float beta2, float eps, float corr1,
                            float corr2, hipStream_t s);
hipError_t pruner_score_samples(float const* w, uint32_t const* samples, uint32_t n, float* out, hipStream_t s);
// first hidden layer (after ReLU, as trained) of n samples: out[n][64]
hipError_t pruner_hidden_samples(float const* w, uint32_t const* samples, uint32_t n, float* out, hipStream_t s);
// integer-grid inference logits of n samples on the BF16 MFMAs (beam_kernel.hip)
hipError_t pruner_infer_samples(PrunerDev const* pr, uint32_t const* samples, uint32_t n, float* out, hipStream_t s);

} // namespace ta3h
