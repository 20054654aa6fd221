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
    TA3H_E_STATE = 3, /* call order (e.g. launch before stage) */
    TA3H_E_SCRATCH = 4 /* ta3h_eval_noalloc: the batch needs more than its static scratch array */
};

/* value networks the search scores leaves with */
enum {
    TA3H_MODEL_V11 = 0, /* reference model_t = ModelV11, 97 params (ta3_ai/incl/ta3/ai/model.hpp:44) */
    TA3H_MODEL_F16 = 1, /* 16 features -> 32 -> 32 -> 1, 1633 params (FP32 fused multiply-adds) */
    TA3H_MODEL_F32W = 3 /* 32 features -> 128 -> 128 -> 1, 20865 params (BF16 MFMA, FP32 accumulate) */
};

/* ta3h_set_phases: first-phase budget chosen from max_moves (the default) */
#define TA3H_PHASES_AUTO 0xFFFFFFFFu

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

/* Two-phase schedule for the following persistent launches (no reference counterpart; results never change):
 * the first phase plays every game for at most first_phase_moves moves and parks the unfinished ones on the
 * device (engine state, counters, move index); the second phase plays only those on to max_moves, so the long
 * games all start together instead of trickling out of the work queue; the second phase takes the games of the
 * models with the most parked games first. 0 = one phase. TA3H_PHASES_AUTO (the default) = max_moves / 32
 * clamped to [64, 512]: 64 at a 2,000-piece cap (a full 16-32-32-1 generation of 200 models x 50 games ran 4%
 * faster than at 128, V11 2%), 156 at 5,000 (the wide workload 1% faster than at 312). Shorter turns keep the
 * games that run to the cap closer together, so fewer end alone. */
int ta3h_set_phases(ta3h_evaluator* ev, uint32_t first_phase_moves);

/* Start every game of the following launches with 0-4 garbage rows at the bottom, each full but for one hole,
 * drawn from xoshiro256** seeded with (game seed ^ 0x6A09E667F3BCC908): n = next() % 5, then for rows 22, 21,
 * ...: hole = next() % 10. The piece sequence is the seed's as without garbage. No reference counterpart (a
 * build-defined start-position extension). */
int ta3h_set_garbage(ta3h_evaluator* ev, int enabled);

/* Beam widths of the following launches: each level keeps its k0 / k1 / k2 best nodes (1..16, 1..32, 1..32).
 * The reference fixes them at compile time to {16, 32, 32} (ta3_ai/incl/ta3/ai/search/variation_sequences.hpp:22),
 * the default here; narrower beams give the moves of the reference's search built with those widths. */
int ta3h_set_beam_widths(ta3h_evaluator* ev, uint32_t k0, uint32_t k1, uint32_t k2);

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

/* Bytes of the static device array that ta3h_eval_noalloc keeps its scratch buffers in (one per GPU, held until
 * the process ends). */
#define TA3H_NOALLOC_SCRATCH_BYTES (256u << 20)

/* No counterpart in the reference. For callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc): ta3h_eval on device buffers the caller owns, enqueued on
 * `stream` (a hipStream_t; NULL = the null stream) without allocating and without synchronising. d_weights: num_models
 * x ta3h_num_params(kind) floats in the host layout ta3h_eval takes; d_seeds: games seeds; d_fitness: num_models x
 * games floats, model-major; d_placements: the same count of engine.place calls (may be NULL). Launch shape and
 * two-phase budget are the defaults. The search's scratch (work-queue counters, the weights repacked into the
 * network's device layout, parked games, the resume list and the second phase's work ring) lives in one static
 * device array of TA3H_NOALLOC_SCRATCH_BYTES per GPU: 256 B + models x the layout's bytes (V11 448, 16-32-32-1
 * 16,832 with its MFMA operand fragments, wide 42,560, each rounded up to 256 B per section) + the parked record per (model, game) (512 B; 18,592 B
 * for 16-32-32-1, whose record carries the previous move's level-2 and leaf scores) + 4 B per (model, game) + 4 B
 * per model + 8 B + 4 B per (model, game) and turn of the second phase (1 + (max_moves - b) / b turns for the
 * budget b = max_moves / 32 clamped to [64, 512]; 31 at 2,000 moves, 32 at 5,000); a larger batch returns
 * TA3H_E_SCRATCH (e.g. 200 x 50 at 16-32-32-1, 2,000 moves: 190.6 MB; 256 x 100 wide, 5,000 moves: 27.4 MB). Calls
 * that overlap on one GPU share that array and race: serialise them (one stream, or events). Returns TA3H_OK,
 * TA3H_E_ARG, TA3H_E_SCRATCH or TA3H_E_HIP (there is no evaluator to hold an error message). */
int ta3h_eval_noalloc(int kind, const float* d_weights, uint32_t num_models, const uint64_t* d_seeds, uint32_t games,
                      uint32_t max_moves, float* d_fitness, uint32_t* d_placements, void* stream);

/* The same evaluation split into its device steps, for callers that keep inputs resident and time or
 * overlap the device work: stage copies inputs into device buffers; launch enqueues the search on the
 * evaluator's stream (asynchronous); sync waits for it; fetch copies results back (placements: engine.place
 * calls per item, may be NULL). */
int ta3h_stage(ta3h_evaluator* ev, const float* weights, uint32_t num_models, uint32_t num_params,
               const uint64_t* seeds, uint32_t games, uint32_t max_moves);
int ta3h_launch(ta3h_evaluator* ev);
int ta3h_sync(ta3h_evaluator* ev);
int ta3h_fetch(ta3h_evaluator* ev, float* out_fitness, uint32_t* out_placements);

/* Per-item work counters of the last launch (no reference counterpart): out[3*item] = candidate placements
 * fitted, out[3*item+1] = candidates that fit and were scored (the work the search executed),
 * out[3*item+2] = lanes the scoring chunks occupied (64 per chunk; legal / lanes = lane utilisation). The search
 * scores a repeated list once (a list of a level whose parent board, clear histogram, piece and held-is-I equal
 * an earlier list's: the same candidates with the same values) and, after a placement that cleared no line,
 * takes the root lists from the previous move's level-1 scores when they match (V11 and 16-32-32-1: any kept
 * root; the wide network: the six best; not in pruned searches that prune level 1, nor in trace and move-log
 * launches, which score every candidate); the full-beam 16-32-32-1 search then also takes the level-1 and
 * level-2 lists that the previous move scored along the committed root's kept nodes (bench.py's 16-32-32-1
 * generation: the reference's search fits 3,476 candidates per placement, this one 2,378). */
int ta3h_fetch_counters(ta3h_evaluator* ev, uint64_t* out_counters);

/* Optional per-item trace (no reference counterpart): when enabled before ta3h_stage, each launch records
 * the device wall clock when every (model, game) item starts and ends; fetch writes out[2*item + {0,1}]
 * and the clock rate. Used to measure load balance over uneven game lengths. */
int ta3h_set_timeline(ta3h_evaluator* ev, int enabled);
int ta3h_fetch_timeline(ta3h_evaluator* ev, uint64_t* out, double* ticks_per_second);

/* Optional move log (no reference counterpart; the per-move record of a game that eval_host's loop plays,
 * model_eval_host.cpp:37-47): when enabled, the following launches record every committed move of every item;
 * fetch writes out[item * max_moves + i] for the item's placements i = 0 .. placements - 1 as one byte,
 * hold << 7 | rotation << 5 | (x + 2) (bytes past an item's placement count are undefined). Used to report the
 * per-move agreement of device games with the CPU search. max_moves bytes per item of device memory. */
int ta3h_set_move_log(ta3h_evaluator* ev, int enabled);
int ta3h_fetch_moves(ta3h_evaluator* ev, uint8_t* out);

/* Device time of the last launch's search kernel in milliseconds (HIP events on the evaluator's stream). */
int ta3h_last_kernel_ms(ta3h_evaluator* ev, float* ms);

/* Workgroups the last launch used, waves per game, and the number of CUs of the device. */
int ta3h_launch_info(ta3h_evaluator* ev, uint32_t* grid_blocks, uint32_t* waves_per_game, uint32_t* num_cus);

/* Replaces ta3::gpu::device_bits_canary() (ta3_gpu/incl/ta3/gpu/device_canary.hpp:16): runs ctz / popcount /
 * bit_width on the device against known answers; *ok = 1 when all match. */
int ta3h_device_bits_canary(int* ok);

/* Leaf values of n boards (boards[10*i .. +10], clear histogram clears[i], held-is-I flag held[i]) as the
 * search scores a candidate board, for one model's weights (host layout): the device counterpart of the
 * reference's net_ref::evaluate (ta3_gpu/incl/ta3/gpu/detail/value_model.hpp:14-21). */
int ta3h_model_values(int kind, const float* weights, const uint32_t* boards, const uint32_t* clears,
                      const uint8_t* held, uint32_t n, float* out);

/* Device values of the normalised features for integers 0..n-1, computed as the search computes them:
 * out[6*i + {0..5}] = {sqrt(i)/100 (i <= 4761, else 0), i/20, i/200, i/5, i/40, i/100} in float32 (no
 * reference counterpart; lets a test pin the device arithmetic against host IEEE arithmetic). */
int ta3h_feature_norms(float* out, uint32_t n);

/* Host copy of the compile-time tables, for checks without a GPU: the placement list as
 * (orientation, x) pairs per piece (lut_ox[162*2], lut_off[8]) and both hold-plan trees
 * (tree[2][4][16][6] = slot, parent, childBegin, childEnd, rootHold, heldSlot; tree_count[2][4]).
 * No reference counterpart. */
void ta3h_tables(int8_t* lut_ox, uint8_t* lut_off, uint8_t* tree, uint8_t* tree_count);

/* ---- one CMA-ES generation as a captured HIP graph --------------------------------------------------------
 * No reference counterpart (the reference launches per island-generation, model_eval.cu:97-153). Captures, on the
 * evaluator's stream: weights (pinned host staging) -> device, queue reset, the search kernel under the current
 * model / launch / pruner / trace settings, fitness and placements -> pinned host. Replays differ only in the
 * weights. Replaces stage + launch + fetch for callers that run the same (num_models, seeds, max_moves) every
 * generation; results equal ta3h_eval's. ta3h_stage / ta3h_eval and every ta3h_set_* call discard the graph. */
int ta3h_graph_prepare(ta3h_evaluator* ev, uint32_t num_models, uint32_t num_params, const uint64_t* seeds,
                       uint32_t games, uint32_t max_moves);

/* Replays the generation graph with new model-major weights (num_models x num_params as prepared); one host
 * synchronisation. out_fitness[m*games+g], out_placements (optional) as ta3h_fetch. ta3h_last_kernel_ms then
 * reports the whole replay (copies included). */
int ta3h_graph_run(ta3h_evaluator* ev, const float* weights, float* out_fitness, uint32_t* out_placements);

/* ---- CMA-ES island loop (the evaluator's caller) and .champ checkpoints -----------------------------------
 * Mirrors ta3_trainer's Trainer2 / TetrisProblem2 (ta3_trainer/src/trainer2.cpp:144-171, 317-366;
 * tetris_problem2.cpp:29-72): `islands` x `population` models per generation, evaluated on the iteration's
 * common seeds (G games, simulate()'s expansion of one seed drawn per iteration from the trainer seed), model
 * fitness = mean of its games (maximised), decision vectors in double narrowed to float weights. Optimiser:
 * separable CMA-ES per island (the reference's sep-CMA-ES option, src/pagmo/sep_cmaes.cpp), island means
 * ~ N(0, init_std^2), step size sigma0, lambda = population, mu = lambda / 2; ring migration after each
 * iteration. All islands' models of a generation run as one batch: one graph replay on the evaluator. */
typedef struct ta3h_trainer ta3h_trainer;

typedef struct {
    int kind;                 /* TA3H_MODEL_* */
    uint32_t islands;         /* 4 */
    uint32_t population;      /* lambda per island, e.g. 50 */
    uint32_t games;           /* games per model and generation, e.g. 50 */
    uint32_t max_moves;       /* game cap, e.g. 2000 */
    uint32_t generations;     /* generations per iteration, e.g. 10 */
    double sigma0;            /* initial step size, e.g. 0.3 */
    double init_std;          /* std of the initial island means, e.g. 0.5 */
    uint64_t seed;            /* trainer seed: islands, sampling and per-iteration game seeds */
} ta3h_train_config;

typedef struct {
    uint64_t iteration;        /* iterations completed */
    double best_fitness;       /* best model fitness of the iteration */
    double last_mean_fitness;  /* mean model fitness of its last generation */
    double champion_fitness;   /* best model fitness so far (the champion's) */
    double device_ms;          /* sum of the generations' graph replay times (0 with a callback) */
    uint64_t iteration_seed;   /* seed whose expansion gave the iteration's games */
} ta3h_iteration_report;

/* Evaluation hook for running the same loop on another evaluator (e.g. a CPU restatement): fill
 * fitness[m * games + g] for num_models x num_params weights on `games` seeds; return 0 on success. */
typedef int (*ta3h_eval_callback)(void* ctx, const float* weights, uint32_t num_models, uint32_t num_params,
                                  const uint64_t* seeds, uint32_t games, uint32_t max_moves, float* fitness);

int ta3h_trainer_create(const ta3h_train_config* cfg, ta3h_trainer** out);
void ta3h_trainer_destroy(ta3h_trainer* t);
const char* ta3h_trainer_last_error(const ta3h_trainer* t);

/* One iteration (cfg.generations generations) evaluated on ev (its model kind is set to cfg.kind; its
 * pruner / garbage / launch settings apply) or, when cb is not NULL, through cb. */
int ta3h_trainer_run_iteration(ta3h_trainer* t, ta3h_evaluator* ev, ta3h_eval_callback cb, void* ctx,
                               ta3h_iteration_report* report);

/* Best model so far (num_params doubles; out may be NULL) and its fitness. */
int ta3h_trainer_champion(const ta3h_trainer* t, double* out, uint64_t cap, double* fitness);

/* Current search mean (num_params doubles) and step size of one island. */
int ta3h_trainer_island_state(const ta3h_trainer* t, uint32_t island, double* mean, double* sigma);

/* Checkpoint of the whole trainer between iterations (replaces Trainer2::saveState's Boost/zstd archive of the
 * optimiser state, trainer2.cpp:213-249, with this library's own versioned binary format): config, iteration
 * count, seed generator, champion and every island's CMA-ES state. A trainer loaded from it (a new handle,
 * destroyed with ta3h_trainer_destroy) continues exactly as the saved one would have. Load fails with
 * TA3H_E_ARG on a missing, truncated, oversized or foreign file. */
int ta3h_trainer_save(const ta3h_trainer* t, const char* path);
int ta3h_trainer_load(const char* path, ta3h_trainer** out);

/* .champ files (trainer2.cpp:251-265): u64 count, then count little-endian doubles. Read with w = NULL to
 * learn the count. */
int ta3h_champ_write(const char* path, const double* w, uint64_t count);
int ta3h_champ_read(const char* path, double* w, uint64_t cap, uint64_t* count);

/* Host replay of a move log (no GPU), for previews: the game of `seed` (with the garbage rows of ta3h_set_garbage
 * when `garbage` is set) played with moves[0 .. n) in the move-log byte format (hold << 7 | rotation << 5 | (x + 2));
 * boards[10 i .. 10 i + 9] = the board after move i, pieces[i] = the piece it placed, cleared[i] = rows it cleared
 * (each output may be NULL). Returns the moves replayed (fewer than n after a top-out) or -1 for a move that does
 * not fit. Replaces the board snapshots the reference's preview renders (ta3_ai/src/preview.cpp:40-94). */
int ta3h_replay(uint64_t seed, int garbage, const uint8_t* moves, uint32_t n, uint32_t* boards, uint8_t* pieces,
                uint8_t* cleared);

/* ---- learned pruner (placement-proposal network) and its training traces -------------------------------
 * No reference counterpart: the reference scores every placement of every kept node (ai/search/beam.hpp:337-372,
 * 407-443). The network: 265 inputs (board slot 24c + r = bit 3 + r of column c, r < 20; one-hots piece 240 + p,
 * rotation 247 + o, column 251 + x + 2, ply 262 + ply - 1) -> 64 ReLU -> 64 ReLU -> 1 logit, FP32 master weights
 * W1[64][265], b1[64], W2[64][64], b2[64], W3[64], b3. Trace record / training sample: 11 words = parent board
 * (10 columns) + meta = piece | o << 3 | (x + 2) << 5 | ply << 9 | label << 11. */

/* Master parameters of the pruner (21,249). */
int ta3h_pruner_num_params(void);

/* Grid exponents {e1, e2, e3, d} of the inference form of weights w with first-hidden-layer range act_range
 * (host only): W1 on 2^-e1, W2 on 2^-e3, first hidden layer codes rne(P 2^-d) with P on 2^-e1. */
int ta3h_pruner_quant_info(const float* w, float act_range, int* info);

/* Pruned beam for the following launches: at plies whose bit is set in ply_mask (bit 1, 2: middle levels, bit 3:
 * leaves) only the k (1..34) best legal placements of every (kept node, piece) list by pruner logit are scored by
 * the value network. w = NULL restores the full beam. k = 34 reproduces the full beam exactly. Pruned launches
 * run 4 waves per game and record no traces. */
int ta3h_set_pruner(ta3h_evaluator* ev, const float* w, float act_range, uint32_t k, uint32_t ply_mask);

/* Full-beam traces for the following launches: every legal candidate of plies 1..3 of every `every`-th move
 * of every game (the moves whose index is congruent to the game's seed modulo `every`), labelled 1 if the level kept it (plies 1, 2) or it is its list's
 * best leaf (ply 3). Up to cap_records records are kept per launch, in no particular order. cap_records = 0
 * turns traces off. Trace launches run 4 waves per game; results are the full beam's. */
int ta3h_set_trace(ta3h_evaluator* ev, uint64_t cap_records, uint32_t every);

/* Copies min(produced, capacity set by ta3h_set_trace, cap_records) records of the last launch to out
 * (NULL: none; *copied gets the number) and reports the records produced (which may exceed the capacity; the
 * excess was dropped). */
int ta3h_fetch_trace(ta3h_evaluator* ev, uint32_t* out, uint64_t cap_records, uint64_t* produced, uint64_t* copied);

/* Batch-mean weighted binary cross-entropy of n samples (weight pos_weight for label 1) and its gradient with
 * respect to the master weights, with bfloat16 W1, W2 and first hidden layer in the forward pass
 * (straight-through). samples = NULL uses the evaluator's trace buffer. */
int ta3h_pruner_grad(ta3h_evaluator* ev, const float* w, const uint32_t* samples, uint32_t n, float pos_weight,
                     float* grad, float* loss);

/* Adam (beta 0.9, 0.999, eps 1e-8, fresh moments) over `epochs` shuffled passes of n samples in mini-batches
 * of `batch`, updating w in place; epoch_loss[e] (optional) = mean training loss of epoch e. samples = NULL
 * trains on the evaluator's trace buffer without a host copy. */
int ta3h_pruner_train(ta3h_evaluator* ev, float* w, const uint32_t* samples, uint64_t n, uint32_t epochs,
                      uint32_t batch, float lr, float pos_weight, uint64_t seed, float* epoch_loss);

/* First hidden layer (after ReLU, bfloat16 as trained) of n samples: out[n][64]; used to calibrate
 * act_range. */
int ta3h_pruner_hidden(ta3h_evaluator* ev, const float* w, const uint32_t* samples, uint32_t n, float* out);

/* Inference logits of n samples as the pruned search computes them (layers 1-2 on integer grids on BF16 MFMA,
 * layer 3 FP32 in a fixed order); NaN for a record whose (rotation, column) is no placement of its piece. */
int ta3h_pruner_logits(ta3h_evaluator* ev, const float* w, float act_range, const uint32_t* samples, uint32_t n,
                       float* out);

#ifdef __cplusplus
}
#endif
#endif
