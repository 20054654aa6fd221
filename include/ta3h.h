/*
 * ta3h.h -- C-ABI of the MI355X (gfx950) beam-search evaluator.
 *
 * Drop-in boundary for the reference's ta3::gpu::ModelEvaluator
 * (ta3_gpu/incl/ta3/gpu/model_eval.hpp:27-77). Plain pointers and sizes only; no HIP or C++ types.
 * include/ta3h/model_eval.hpp wraps these entry points in the reference's C++ class shape.
 *
 * Threading (as the reference, model_eval.hpp:22-26): one evaluator is NOT thread-safe; use one per
 * host thread. Distinct evaluators run on distinct HIP streams and may be used concurrently.
 * Every entry point returns 0 on success and a nonzero TA3H_E* code on failure; the evaluator's
 * ta3h_last_error() then describes it ("hip: <what>: <runtime error>" or an argument error).
 */
#ifndef TA3H_H
#define TA3H_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ta3h_evaluator ta3h_evaluator;

enum {
    TA3H_OK = 0,
    TA3H_E_ARG = 1,  /* invalid argument (sizes, model kind, launch shape) */
    TA3H_E_HIP = 2,  /* HIP runtime error */
    TA3H_E_STATE = 3 /* call order (e.g. launch before stage) */
};

/* value networks the search scores leaves with */
enum {
    TA3H_MODEL_V11 = 0, /* reference model_t = ModelV11, 97 params (ta3_ai/incl/ta3/ai/model.hpp:44) */
    TA3H_MODEL_F16 = 1  /* 16 features -> 32 -> 32 -> 1, 1633 params */
};

/* Replaces ModelEvaluator::ModelEvaluator() (model_eval.hpp:29; model_eval.cu:92). Allocates nothing
 * on the device until the first evaluation (grow-only buffers, model_eval.cu:54-70). */
int ta3h_create(ta3h_evaluator** out);

/* Replaces ~ModelEvaluator() (model_eval.hpp:30; model_eval.cu:72-83): best effort, never fails. */
void ta3h_destroy(ta3h_evaluator* ev);

/* Last error message of this evaluator ("" if none). The pointer stays valid until the next call. */
const char* ta3h_last_error(const ta3h_evaluator* ev);

/* Parameters per model of a kind (ModelV11::NUM_PARAMS, model_v11.hpp:159-160), or -1. */
int ta3h_num_params(int kind);

/* Selects the value network (default TA3H_MODEL_V11; the reference binds it at compile time,
 * ta3_ai/incl/ta3/ai/model.hpp:44). */
int ta3h_set_model(ta3h_evaluator* ev, int kind);

/* Launch shape: waves cooperating on one game (1, 2 or 4; a workgroup plays one game at a time),
 * persistent work-queue mode (1) or one workgroup per item (0), and the persistent grid in workgroups
 * (0 = fill the device at the occupancy limit). Results never depend on these knobs. */
int ta3h_set_launch(ta3h_evaluator* ev, int waves_per_game, int persistent, uint32_t grid_blocks);

/* Replaces ModelEvaluator::eval(weights, num_models, seeds, max_moves) (model_eval.hpp:46-51;
 * model_eval.cu:97-153): num_models models, model-major weights (model m owns
 * weights[m*num_params, (m+1)*num_params)), each played on the same `games` seeds; out_fitness[m*games+g]
 * gets model m on game g (score_v1, or -1e6 for a game with no placed piece). num_params must equal
 * ta3h_num_params(current model). num_models*games == 0 returns success without launching.
 * One host synchronisation per call. */
int ta3h_eval(ta3h_evaluator* ev, const float* weights, uint32_t num_models, uint32_t num_params,
              const uint64_t* seeds, uint32_t games, uint32_t max_moves, float* out_fitness);

/* The same evaluation split into its device steps, for callers that keep inputs resident and time or
 * overlap the device work: stage copies inputs into device buffers; launch enqueues the search on the
 * evaluator's stream (asynchronous); sync waits for it; fetch copies results back (placements: engine.place
 * calls per item, may be NULL). */
int ta3h_stage(ta3h_evaluator* ev, const float* weights, uint32_t num_models, uint32_t num_params,
               const uint64_t* seeds, uint32_t games, uint32_t max_moves);
int ta3h_launch(ta3h_evaluator* ev);
int ta3h_sync(ta3h_evaluator* ev);
int ta3h_fetch(ta3h_evaluator* ev, float* out_fitness, uint32_t* out_placements);

/* Per-item work counters of the last launch: out[2*item] = candidate placements fitted, out[2*item+1] =
 * candidates that fit and were scored (the algorithmic work of the search; no reference counterpart). */
int ta3h_fetch_counters(ta3h_evaluator* ev, uint64_t* out_counters);

/* Optional per-item trace (no reference counterpart): when enabled before ta3h_stage, each launch records
 * the device wall clock when every (model, game) item starts and ends; fetch writes out[2*item + {0,1}]
 * and the clock rate. Used to measure load balance over uneven game lengths. */
int ta3h_set_timeline(ta3h_evaluator* ev, int enabled);
int ta3h_fetch_timeline(ta3h_evaluator* ev, uint64_t* out, double* ticks_per_second);

/* Device time of the last launch's search kernel in milliseconds (HIP events on the evaluator's stream). */
int ta3h_last_kernel_ms(ta3h_evaluator* ev, float* ms);

/* Workgroups the last launch used, waves per game, and the number of CUs of the device. */
int ta3h_launch_info(ta3h_evaluator* ev, uint32_t* grid_blocks, uint32_t* waves_per_game, uint32_t* num_cus);

/* Replaces ta3::gpu::device_bits_canary() (ta3_gpu/incl/ta3/gpu/device_canary.hpp:16): runs ctz / popcount /
 * bit_width on the device against known answers; *ok = 1 when all match. */
int ta3h_device_bits_canary(int* ok);

/* Device values of the normalised features for integers 0..n-1, computed as the search computes them:
 * out[6*i + {0..5}] = {sqrt(i)/100 (i <= 4761, else 0), i/20, i/200, i/5, i/40, i/100} in float32 (no
 * reference counterpart; lets a test pin the device arithmetic against host IEEE arithmetic). */
int ta3h_feature_norms(float* out, uint32_t n);

/* Host copy of the compile-time tables, for checks without a GPU: the placement list as
 * (orientation, x) pairs per piece (lut_ox[162*2], lut_off[8]) and both hold-plan trees
 * (tree[2][4][16][6] = slot, parent, childBegin, childEnd, rootHold, heldSlot; tree_count[2][4]).
 * No reference counterpart. */
void ta3h_tables(int8_t* lut_ox, uint8_t* lut_off, uint8_t* tree, uint8_t* tree_count);

#ifdef __cplusplus
}
#endif
#endif
