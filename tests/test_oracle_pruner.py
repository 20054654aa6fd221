"""Pins the oracle's placement-proposal network (pruner), its training gradient and the pruned beam.

The pruner is build-defined (no reference counterpart): 265 inputs (200 board bits in 240 slots, one-hot
piece / rotation / column / ply) -> 64 -> 64 -> 1, bfloat16 W1, W2 and first hidden layer at inference.
"""
import numpy as np
import pytest

import oracle as O

P = 21249


def bf16(a):
    u = np.asarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def random_pruner(seed, scale=0.1):
    rng = np.random.default_rng(seed)
    w = (rng.normal(0, scale, P)).astype(np.float32)
    return w


def inputs(cols, piece, o, x, ply):
    xs = np.zeros(265)
    for c in range(10):
        for r in range(20):
            if (int(cols[c]) >> (3 + r)) & 1:
                xs[24 * c + r] = 1
    xs[240 + piece] = 1
    xs[247 + o] = 1
    xs[251 + x + 2] = 1
    xs[262 + ply - 1] = 1
    return xs


def np_score(w, cols, piece, o, x, ply):
    W1 = bf16(w[:64 * 265]).reshape(64, 265).astype(np.float64)
    b1 = w[64 * 265:64 * 266].astype(np.float64)
    o2 = 64 * 266
    W2 = bf16(w[o2:o2 + 4096]).reshape(64, 64).astype(np.float64)
    b2 = w[o2 + 4096:o2 + 4160].astype(np.float64)
    W3 = w[o2 + 4160:o2 + 4224].astype(np.float64)
    b3 = float(w[o2 + 4224])
    pre1 = b1 + W1 @ inputs(cols, piece, o, x, ply)
    h1 = bf16(np.maximum(pre1.astype(np.float32), 0)).astype(np.float64)
    h2 = np.maximum(b2 + W2 @ h1, 0)
    return b3 + h2 @ W3


def test_param_count_and_layout():
    assert O.pruner_num_params() == P == 64 * 265 + 64 + 64 * 64 + 64 + 64 + 1


def test_score_matches_dense_numpy():
    w = random_pruner(1)
    rng = np.random.default_rng(2)
    for _ in range(20):
        cols = rng.integers(0, 1 << 23, 10).astype(np.uint32)
        piece, o, x, ply = int(rng.integers(0, 7)), int(rng.integers(0, 4)), int(rng.integers(-2, 9)), int(rng.integers(1, 4))
        got = O.pruner_score(w, cols, piece, o, x, ply)
        want = np_score(w, cols, piece, o, x, ply)
        assert abs(got - want) <= 1e-5 * max(1.0, abs(want))


def small_batch(seed=3, n=40):
    rng = np.random.default_rng(seed)
    recs = np.zeros((n, 11), dtype=np.uint32)
    recs[:, :10] = rng.integers(0, 1 << 23, (n, 10))
    piece = rng.integers(0, 7, n)
    o = rng.integers(0, 4, n)
    x = rng.integers(-2, 9, n)
    ply = rng.integers(1, 4, n)
    lab = rng.integers(0, 2, n)
    recs[:, 10] = piece | (o << 3) | ((x + 2) << 5) | (ply << 9) | (lab << 11)
    return recs


def test_gradient_matches_finite_differences():
    """the backward formulas against central differences of the (unrounded, smooth) loss"""
    w = random_pruner(4, 0.2)
    recs = small_batch()
    _, g = O.pruner_loss_grad(w, recs, pos_weight=3.0, rounded=False)
    rng = np.random.default_rng(5)
    regions = [(0, 64 * 265), (64 * 265, 64 * 266), (64 * 266, 64 * 266 + 4096), (64 * 266 + 4096, P)]
    checked = 0
    for lo, hi in regions:
        cand = [i for i in rng.permutation(np.arange(lo, hi))[:400] if abs(g[i]) > 1e-6][:6]
        for i in cand:
            h = 1e-2 * max(abs(float(w[i])), 1e-2)
            wp, wm = w.astype(np.float64).copy(), w.astype(np.float64).copy()
            wp[i] += h
            wm[i] -= h
            # the loss is evaluated from float32 weights: use the actually representable step
            wp32, wm32 = wp.astype(np.float32), wm.astype(np.float32)
            lp, _ = O.pruner_loss_grad(wp32, recs, 3.0, False, with_grad=False)
            lm, _ = O.pruner_loss_grad(wm32, recs, 3.0, False, with_grad=False)
            fd = (lp - lm) / (float(wp32[i]) - float(wm32[i]))
            assert abs(fd - g[i]) <= 2e-3 * abs(g[i]) + 1e-7, (i, fd, g[i])
            checked += 1
    assert checked >= 12


def test_pruned_beam_with_k34_is_the_full_beam():
    rng = np.random.default_rng(6)
    ws = rng.normal(0, 0.5, 2 * 97).astype(np.float32)
    seeds = O.expand_seeds(9, 3)
    full, full_pl = O.eval_host_mt(O.MODEL_V11, ws, 2, seeds, 120)
    pr = random_pruner(7)
    pruned, pl, _ = O.eval_host_pruned_mt(O.MODEL_V11, ws, 2, seeds, 120, pr, 34, 0b1110)
    assert pruned.tobytes() == full.tobytes() and (pl == full_pl).all()


def test_pruning_cuts_value_evaluations():
    rng = np.random.default_rng(8)
    ws = rng.normal(0, 0.5, 97).astype(np.float32)
    seeds = O.expand_seeds(10, 2)
    pr = random_pruner(11)
    cnt_full = O.Counters()
    O.eval_host_mt(O.MODEL_V11, ws, 1, seeds, 60, counters=cnt_full)
    _, _, cnt = O.eval_host_pruned_mt(O.MODEL_V11, ws, 1, seeds, 60, pr, 8, 0b1100)
    per_move_full = cnt_full.legal / cnt_full.moves
    per_move = cnt.legal / cnt.moves
    assert per_move < 0.75 * per_move_full


def test_trace_labels():
    """ply 1/2 labels mark exactly the level's kept top-K (32 per move when enough candidates), ply 3 marks
    one best leaf per list"""
    rng = np.random.default_rng(12)
    w = rng.uniform(-2, 2, 97).astype(np.float32)
    recs = O.trace_records(O.MODEL_V11, w, 21, 30, 10)
    meta = recs[:, 10]
    ply = (meta >> 9) & 3
    lab = (meta >> 11) & 1
    assert set(np.unique(ply)) == {1, 2, 3}
    n_moves = 3  # moves 0, 10, 20
    assert lab[ply == 1].sum() == 32 * n_moves
    assert lab[ply == 2].sum() == 32 * n_moves
    # one positive per leaf list (<= 32 kept nodes x 2 lists per move); a (board, piece) pair can name
    # several lists, since different hold paths reach identical boards, so every pair has >= 1 positive
    leaf = recs[ply == 3]
    keys = {}
    for r in leaf:
        k = (tuple(r[:10]), int(r[10] & 7))
        keys[k] = keys.get(k, 0) + int((r[10] >> 11) & 1)
    assert all(v >= 1 for v in keys.values())
    assert len(keys) <= lab[ply == 3].sum() <= 64 * n_moves
