/*
 * ta3_oracle.h -- C-ABI of the CPU oracle (TEST INFRASTRUCTURE ONLY).
 *
 * A clean scalar C++ restatement of the reference's CPU evaluation path
 * (ta3_gpu/src/common/model_eval_host.cpp:26-54 and everything it calls). It is
 * the checker for the HIP evaluator: only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it. The product library never links it.
 *
 * Piece enum: I=0 O=1 T=2 S=3 Z=4 J=5 L=6, NO_PIECE=7 (ta3_sim/.../piece_defs.hpp:19-31).
 * Orientation enum: TOP=0 RIGHT=1 BOTTOM=2 LEFT=3 (piece_defs.hpp:129-136).
 * Board: 10 u32 columns, bit y = row y, y=0 is the top row, 23 rows (board2.hpp:20-27).
 */
#ifndef TA3_ORACLE_H
#define TA3_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* model kinds (value networks the search can score leaves with) */
enum {
    TA3O_MODEL_V11 = 0,    /* reference default model_t: 10 features -> 8 ReLU -> 1, 97 params */
    TA3O_MODEL_F16 = 1,    /* 16 features -> 32 -> 32 -> 1 (ReLU hidden), 1633 params */
    TA3O_MODEL_HEIGHT = 2, /* test double of the reference search tests, 0 params */
    TA3O_MODEL_F32W = 3    /* 32 features -> 128 -> 128 -> 1 (ReLU hidden), 20865 params, BF16 MFMA arithmetic */
};

/* plain-old-data engine state, field-for-field the reference TetrisEngine (tetris_engine.hpp:81-94) */
typedef struct {
    uint32_t cols[10];
    uint32_t queue[6]; /* [0] current, [1..5] lookahead */
    uint32_t held;     /* 7 = empty */
    uint32_t game_over;
    uint32_t bag[7];
    uint32_t bag_pos;
    uint64_t rng[4];
} ta3o_engine;

/* counters of the algorithmic work, summed over the moves searched */
typedef struct {
    uint64_t moves;       /* search_move_beam calls that returned a move */
    uint64_t applies;     /* candidate fit() calls (legal or not) in scoring stages */
    uint64_t legal;       /* candidates that fit and were scored by the model */
    uint64_t reapplies;   /* frontier re-applies (seed/advance) */
    uint64_t placements;  /* engine.place calls */
    uint64_t dead_mid;    /* kept nodes of a middle level with no legal child in any list (beam.hpp:549-551) */
    uint64_t dead_leaf;   /* leaf lists with no legal placement (beam.hpp:439-440) */
    uint64_t fallback_mid_best;  /* moves whose best key is a middle-level dead-end fallback */
    uint64_t fallback_leaf_best; /* moves whose best key is a leaf-list dead-end fallback */
    /* of applies / legal: those in repeated lists of a middle level or the leaves (an earlier list of the level
       has the same parent board, clear histogram, piece and held-is-I); counted, still scored here */
    uint64_t repeat_applies;
    uint64_t repeat_legal;
    /* of applies / legal: root candidates equal to candidates the previous move's search scored (the committed
       placement cleared no line and every root list is a level-1 list of the chosen root node: same board, empty
       clear histogram, piece and held-is-I); counted, still scored here */
    uint64_t reuse_applies;
    uint64_t reuse_legal;
} ta3o_counters;

int ta3o_num_params(int kind);

/* xoshiro256** (ta3_sim/incl/ta3/sim/utility/xoshiro256ss.hpp:20-48) */
void ta3o_rng_seed(uint64_t seed, uint64_t state[4]);
uint64_t ta3o_rng_next(uint64_t state[4]);
/* gpu::simulate seed expansion (ta3_gpu/incl/ta3/gpu/model_eval.hpp:92-95) */
void ta3o_expand_seeds(uint64_t seed, uint32_t games, uint64_t* out);

/* Board2 operations with (piece, orientation, window offset) arguments (board2.hpp) */
int ta3o_board_available(const uint32_t cols[10], int type, int o, int x, int y);
int ta3o_board_drop_location(const uint32_t cols[10], int type, int o, int x, int y);
void ta3o_board_place(uint32_t cols[10], int type, int o, int x, int y);
int ta3o_board_clear_lines(uint32_t cols[10]);
int ta3o_board_full_lines(const uint32_t cols[10]);
int ta3o_board_holes(const uint32_t cols[10]);
/* fused fit on the nibble-packed shape of (type, o, x): returns legal, writes drop distance */
int ta3o_board_fit(const uint32_t cols[10], int type, int o, int x, int* drop);
/* piece shape table entry (piece_offsets.hpp:186-209) */
void ta3o_piece_shape(int type, int o, uint32_t* cols_nibbles, int* left);

/* process-wide option: start every game (engine_reset, eval_host, traces) with 0-4 garbage rows, each full
 * but for one hole, drawn from xoshiro256** seeded with seed ^ 0x6A09E667F3BCC908: n = next() % 5, then per
 * row y = 22, 21, ...: hole = next() % 10 */
void ta3o_set_garbage(int enabled);

/* process-wide option: beam widths of the games the oracle plays (eval_host*, traces; 3 values; NULL = the reference's
 * {16, 32, 32}, variation_sequences.hpp:22) */
void ta3o_set_widths(const uint32_t* widths);

/* TetrisEngine (tetris_engine.hpp:103-161); place returns cleared rows or -1 for DIED */
void ta3o_engine_reset(ta3o_engine* e, uint64_t seed);
int ta3o_engine_place(ta3o_engine* e, int o, int x);
void ta3o_engine_hold(ta3o_engine* e);

/* placement LUT (ta3_ai/incl/ta3/ai/search/placements.hpp:189-220) */
int ta3o_lut_count(int piece);
void ta3o_lut_entry(int piece, int i, int* o, int* x);

/* hold plans for a search depth (variation_sequences.hpp:216-255); returns the plan count.
 * slots/held_slots are [count][depth], root_hold is [count]. */
int ta3o_plans(int depth, int held_empty, uint8_t* slots, uint8_t* held_slots, uint8_t* root_hold);
/* plan tree (beam.hpp:69-126): nodes[layer][idx] = {slot, parent, childBegin, childEnd, rootHold, heldSlot},
 * out array is [depth][2^depth][6]; counts is [depth] */
void ta3o_tree(int depth, int held_empty, uint8_t* nodes, uint8_t* counts);

/* ordered float keys (beam.hpp:204-211) */
uint32_t ta3o_ordered_bits(float v);
float ta3o_ordered_bits_inv(uint32_t s);

/* model features and value */
int ta3o_num_features(int kind);
void ta3o_features(int kind, uint32_t clears, const uint32_t cols[10], int held_is_i, float* out);
float ta3o_model_value(int kind, const float* w, uint32_t clears, const uint32_t cols[10], int held_is_i);
/* the wide network in plain FP32 (fused multiply-adds in dense() order, no bfloat16 rounding) */
float ta3o_f32w_value_fp32(const float* w, uint32_t clears, const uint32_t cols[10], int held_is_i);

/* select_top (beam.hpp:244-265) over n scores: the k best nonzero scores, highest first, lowest slot on ties;
 * writes their slots and scores, zeroes them in `scores` and returns how many were selected */
uint32_t ta3o_select_top(uint32_t* scores, uint32_t n, uint32_t k, uint32_t* out_slots, uint32_t* out_scores);
/* process-wide option: 1 = select_top always runs the reference's repeated-argmax loop; 0 (default) = beams wider
 * than 64 take an equivalent single sort (only the full-width checks use such beams) */
void ta3o_set_select_linear(int enabled);

/* one beamed search move (beam.hpp:483-570) at runtime depth = nwidths + 1.
 * returns 1 and writes the move, or 0 when none. */
int ta3o_search_beam(const ta3o_engine* e, int kind, const float* w, const uint32_t* widths, int nwidths,
                     int* hold, int* o, int* x, float* value, ta3o_counters* counters);
/* exhaustive depth-`depth` search (search.hpp:117-150) */
int ta3o_search_exhaustive(const ta3o_engine* e, int kind, const float* w, int depth,
                           int* hold, int* o, int* x, float* value);

/* eval_host (model_eval_host.cpp:26-54) for one model: out[g] = fitness of game g */
void ta3o_eval_host(int kind, const float* w, const uint64_t* seeds, uint32_t games, uint32_t max_moves,
                    float* out, ta3o_counters* counters);
/* the same computation over num_models x games items on `threads` host threads; out is model-major.
 * placements_out (optional) gets per-item engine.place counts. */
void ta3o_eval_host_mt(int kind, const float* w, uint32_t num_models, const uint64_t* seeds, uint32_t games,
                       uint32_t max_moves, int threads, float* out, uint32_t* placements_out,
                       ta3o_counters* counters);

/* one game, recording every move: moves[i*4+{0..3}] = {hold, orientation, x+2, cleared (255 = DIED)}
 * and the committed board after it (boards[i*10 .. +10]). returns moves recorded; fitness written. */
int ta3o_play_trace(int kind, const float* w, uint64_t seed, uint32_t max_moves, uint8_t* moves,
                    uint32_t* boards, float* fitness);

/* one game with the content check of the lists of levels 1 and 2 counted as taken from the previous move:
 * out[0] = lists counted, out[1] = those whose content no list of the previous search one level down had,
 * out[2] = moves, out[3] = reuse_applies */
int ta3o_check_deep_reuse(int kind, const float* w, uint64_t seed, uint32_t max_moves, uint64_t* out);

/* placement-proposal network (build-defined): 265 inputs -> 64 -> 64 -> 1, 21,249 parameters */
int ta3o_pruner_num_params(void);
/* inference logit of the integer-grid network (quantized from w with the first hidden layer's calibrated
 * range act_range) for a candidate (parent board, piece, orientation, x, ply 1..3) */
float ta3o_pruner_score(const float* w, float act_range, const uint32_t cols[10], int piece, int o, int x, int ply);
/* the same for n trace records */
void ta3o_pruner_scores(const float* w, float act_range, const uint32_t* samples, int n, float* out);
/* grid exponents of the quantized network: info = {e1, e2, e3, d} */
void ta3o_pruner_quant_info(const float* w, float act_range, int* info);
/* batch-mean weighted BCE and its gradient (float64) over trace records (11 words each);
 * rounded = apply the inference bfloat16 roundings (straight-through) */
double ta3o_pruner_loss_grad(const float* w, const uint32_t* samples, int n, double pos_weight, int rounded,
                             double* grad);
/* one pruned beamed search move at the reference widths {16, 32, 32} (what eval_host_pruned_mt plays per move);
 * returns 1 and writes the move, or 0 when none */
int ta3o_search_beam_pruned(const ta3o_engine* e, int kind, const float* w, const float* pruner_w, float act_range,
                            int k, unsigned ply_mask, int* hold, int* o, int* x, float* value);
/* eval_host_mt with the pruned beam: at plies in ply_mask (bit 1, 2: middle levels, bit 3: leaves) only
 * the pruner's top-k legal placements of every list are scored by the value network */
void ta3o_eval_host_pruned_mt(int kind, const float* w, uint32_t num_models, const uint64_t* seeds, uint32_t games,
                              uint32_t max_moves, int threads, const float* pruner_w, float act_range, int k,
                              unsigned ply_mask, float* out, uint32_t* placements_out, ta3o_counters* counters);
/* trace records of one full-beam game: every legal candidate of plies 1..3 of the moves whose index is
 * congruent to seed modulo `every`:
 * 10 board words + meta (piece | o << 3 | (x+2) << 5 | ply << 9 | label << 11); returns the count
 * (copies at most cap_records) */
uint64_t ta3o_trace_records(int kind, const float* w, uint64_t seed, uint32_t max_moves, uint32_t every,
                            uint32_t* out, uint64_t cap_records);

/* Instrumented op counts of n candidates given as trace records (parent board + piece / rotation / column): ops[0]
 * = scalar operations of fit summed over the records, ops[1] of placing + clearing (legal records), ops[2] of the
 * features, ops[3] of the value network (kind V11 or F16, weights w). legal[r] = drop distance or -1; boards[10 r]
 * = the board after placing and clearing; features[16 r] (V11: first 10) and values[r] as the restatement
 * computes them, for checking against the oracle (oracle/opcount.cpp). */
void ta3o_candidate_ops(int kind, const float* w, const uint32_t* samples, int n, uint64_t* ops, int32_t* legal,
                        uint32_t* boards, float* features, float* values);

#ifdef __cplusplus
}
#endif
#endif
