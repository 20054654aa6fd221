"""Pins the oracle's placement-proposal network (pruner), its training gradient and the pruned beam.

The pruner is build-defined (no reference counterpart): 265 inputs (200 board bits in 240 slots, one-hot
piece / rotation / column / ply) -> 64 -> 64 -> 1. Training rounds W1, W2 and the first hidden layer to
bfloat16; inference runs layers 1-2 as an integer network on power-of-two grids (8-bit W1, W2 and first
hidden layer) and layer 3 in float32 in a fixed order.
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


def np_float_score(w, xs_list):
    """dense float64 forward pass of the unrounded network, one row per input vector"""
    w = np.asarray(w, dtype=np.float64)
    W1 = w[:64 * 265].reshape(64, 265)
    b1 = w[64 * 265:64 * 266]
    o2 = 64 * 266
    W2 = w[o2:o2 + 4096].reshape(64, 64)
    b2 = w[o2 + 4096:o2 + 4160]
    W3 = w[o2 + 4160:o2 + 4224]
    b3 = float(w[o2 + 4224])
    X = np.asarray(xs_list)
    h1 = np.maximum(X @ W1.T + b1, 0)
    h2 = np.maximum(h1 @ W2.T + b2, 0)
    return h2 @ W3 + b3, h1


def test_param_count_and_layout():
    assert O.pruner_num_params() == P == 64 * 265 + 64 + 64 * 64 + 64 + 64 + 1


def random_candidates(seed, n):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        cols = rng.integers(0, 1 << 23, 10).astype(np.uint32)
        out.append((cols, int(rng.integers(0, 7)), int(rng.integers(0, 4)), int(rng.integers(-2, 9)),
                    int(rng.integers(1, 4))))
    return out


def test_integer_inference_tracks_the_float_network():
    """the quantized logit, scaled back by 2^-(e1-d+e3) and shifted by b3, is the dense float64 network up to
    8-bit quantization error, and ranks candidates the same way"""
    w = random_pruner(1)
    cands = random_candidates(2, 300)
    ref, h1 = np_float_score(w, [inputs(*c) for c in cands])
    act = float(np.max(h1)) * 1.01  # no hidden unit clips
    e1, e2, e3, d = O.pruner_quant_info(w, act)
    assert 2.0 ** e1 * np.abs(w[:64 * 265]).max() <= 255 < 2.0 ** (e1 + 1) * np.abs(w[:64 * 265]).max()
    assert 2.0 ** e2 * act <= 255 < 2.0 ** (e2 + 1) * act
    assert d == e1 - e2
    got = np.array([O.pruner_score(w, act, *c) for c in cands], dtype=np.float64)
    approx = got * 2.0 ** -(e1 - d + e3) + float(w[-1])
    err = np.abs(approx - ref)
    # 8-bit grids: steps of max|W|/255 per weight and act/255 per hidden activation
    assert err.max() <= 0.1 * ref.std() and err.mean() <= 0.025 * ref.std(), (err.max(), err.mean(), ref.std())
    rank_ref = np.argsort(np.argsort(ref))
    rank_got = np.argsort(np.argsort(approx))
    assert np.corrcoef(rank_ref, rank_got)[0, 1] > 0.995


def test_integer_inference_is_scale_invariant():
    """power-of-two rescaling of a layer moves its grid exponent and leaves every logit bit-identical"""
    w = random_pruner(3)
    cands = random_candidates(4, 40)
    act = 2.0
    base = [O.pruner_score(w, act, *c) for c in cands]
    w2 = w.copy()
    w2[:64 * 266] *= 4.0                           # W1 and b1: the first hidden layer grows 4x ...
    w2[64 * 266:64 * 266 + 4096] *= 0.25           # ... and W2 takes it back: the same float function
    assert [O.pruner_score(w2, 4 * act, *c) for c in cands] == base
    w3 = w.copy()
    w3[64 * 266 + 4160:64 * 266 + 4224] *= 0.5  # W3: a power of two scales the float32 layer exactly
    assert [O.pruner_score(w3, act, *c) for c in cands] == [0.5 * b for b in base]
    e = O.pruner_quant_info(w2, 4 * act)
    assert e[0] == O.pruner_quant_info(w, act)[0] - 2


def test_batch_scores_match_single():
    w = random_pruner(5)
    recs = small_batch(6, 30)
    got = O.pruner_scores(w, 1.5, recs)
    for t in range(30):
        m = int(recs[t, 10])
        want = O.pruner_score(w, 1.5, recs[t, :10], m & 7, (m >> 3) & 3, ((m >> 5) & 15) - 2, (m >> 9) & 3)
        assert got[t] == want


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
    pruned, pl, _ = O.eval_host_pruned_mt(O.MODEL_V11, ws, 2, seeds, 120, pr, 1.0, 34, 0b1110)
    assert pruned.tobytes() == full.tobytes() and (pl == full_pl).all()


def test_pruning_cuts_value_evaluations():
    rng = np.random.default_rng(8)
    ws = rng.normal(0, 0.5, 97).astype(np.float32)
    seeds = O.expand_seeds(10, 2)
    pr = random_pruner(11)
    cnt_full = O.Counters()
    O.eval_host_mt(O.MODEL_V11, ws, 1, seeds, 60, counters=cnt_full)
    _, _, cnt = O.eval_host_pruned_mt(O.MODEL_V11, ws, 1, seeds, 60, pr, 1.0, 8, 0b1100)
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
    n_moves = 3  # moves 1, 11, 21 (index = seed 21 mod 10)
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
