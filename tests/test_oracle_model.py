"""Pins the oracle's value models, stats and fitness.

- Features: brute-force cell-grid definitions of the reference metrics
  (ta3_ai/incl/ta3/ai/models/metrics/v_board_metrics.hpp:130-232), normalised with numpy float32
  (IEEE correctly rounded division and sqrt), compared bit-exactly.
- MLP: the dense layout of ta3_ai/incl/ta3/ai/models/models/utility.hpp:29-46 evaluated with numpy
  float32 scalars (separately rounded multiply and add), compared bit-exactly.
- score_v1 (ta3_ai/incl/ta3/ai/models/metrics/score_v1.hpp:20-34) recomputed from a recorded game.
- End to end: the reference's shipped v11 champion (models/v11_16_32_32_score_296.champ, mean score 296
  under beam widths 16/32/32) must play near-perfect games on fresh seeds.
"""
import json
import os

import numpy as np
import pytest

import oracle as O

ROWS = 23
HERE = os.path.dirname(os.path.abspath(__file__))
f32 = np.float32


def brute_metrics(cols):
    g = [[(int(cols[x]) >> y) & 1 for y in range(ROWS)] for x in range(10)]
    heights = []
    holes = hole_depths = ytr = 0
    for x in range(10):
        filled = [y for y in range(ROWS) if g[x][y]]
        top = min(filled) if filled else ROWS
        heights.append(ROWS - top)
        holes += sum(1 for y in range(top, ROWS) if not g[x][y])
        for y in filled:  # each filled cell: empty cells below it
            hole_depths += sum(1 for yy in range(y + 1, ROWS) if not g[x][yy])
        cells = g[x] + [1]  # floor counts as filled
        ytr += sum(1 for y in range(ROWS) if cells[y] != cells[y + 1])
    sv = sum((heights[x] - heights[x - 1]) ** 2 for x in range(1, 10))
    bump = sum(abs(heights[x] - heights[x - 1]) for x in range(1, 10))
    wells = 0
    for x in range(10):
        h = heights[x]
        if x == 0:
            wells += heights[1] >= h + 3
        elif x == 9:
            wells += heights[8] >= h + 3
        else:
            wells += heights[x - 1] >= h + 3 and heights[x + 1] >= h + 3
    return dict(heights=heights, holes=holes, hd=hole_depths, ytr=ytr, sv=sv, bump=bump, wells=wells)


def brute_features_v11(clears, cols, held_i):
    m = brute_metrics(cols)
    out = [f32((clears >> (8 * k)) & 0xFF) * f32(0.25) for k in range(4)]
    sq = f32(0) if m["sv"] <= 0 else np.sqrt(f32(m["sv"]))
    out += [sq / f32(100), f32(m["holes"]) / f32(20), f32(m["hd"]) / f32(200), f32(max(m["heights"])),
            f32(1 if held_i else 0), f32(m["wells"]) / f32(5)]
    return np.array(out, dtype=np.float32), m


def brute_features_f16(clears, cols, held_i):
    v11, m = brute_features_v11(clears, cols, held_i)
    cnt = [(clears >> (8 * k)) & 0xFF for k in range(4)]
    lines = sum((k + 1) * cnt[k] for k in range(4))
    score = -cnt[0] - 2 * cnt[1] - 3 * cnt[2] + 10 * cnt[3]
    extra = [f32(sum(m["heights"])) / f32(200), f32(m["ytr"]) / f32(40), f32(m["bump"]) / f32(100),
             f32(lines) * f32(0.1), f32(score) * f32(0.1), f32(max(m["heights"]) - min(m["heights"]))]
    return np.concatenate([v11, np.array(extra, dtype=np.float32)])


def random_boards(n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        cols = np.zeros(10, dtype=np.uint32)
        lo = int(rng.integers(0, 23))
        for x in range(10):
            for y in range(lo, ROWS):
                if rng.random() < rng.uniform(0.2, 0.95):
                    cols[x] |= np.uint32(1 << y)
        out.append(cols)
    out.append(np.zeros(10, dtype=np.uint32))
    out.append(np.full(10, (1 << 23) - 1, dtype=np.uint32))
    return out


def test_v11_features_match_brute_force():
    rng = np.random.default_rng(2)
    for cols in random_boards(300, 7):
        clears = int(rng.integers(0, 5)) | int(rng.integers(0, 5)) << 8 | int(rng.integers(0, 3)) << 16 | \
            int(rng.integers(0, 4)) << 24
        held = bool(rng.integers(0, 2))
        want, _ = brute_features_v11(clears, cols, held)
        got = O.features(O.MODEL_V11, clears, cols, held)
        assert got.tobytes() == want.tobytes()


def test_f16_features_match_brute_force():
    rng = np.random.default_rng(3)
    for cols in random_boards(200, 9):
        clears = int(rng.integers(0, 5)) | int(rng.integers(0, 3)) << 8 | int(rng.integers(0, 4)) << 24
        held = bool(rng.integers(0, 2))
        got = O.features(O.MODEL_F16, clears, cols, held)
        assert got.tobytes() == brute_features_f16(clears, cols, held).tobytes()


def np_dense(x, n_in, n_out, relu, w):
    out = []
    for o in range(n_out):
        v = f32(w[n_in * n_out + o])
        for i in range(n_in):
            v = f32(v + f32(f32(x[i]) * f32(w[o * n_in + i])))
        out.append((v if f32(0) < v else f32(0)) if relu else v)
    return out


def test_v11_value_matches_numpy_float32():
    rng = np.random.default_rng(4)
    for cols in random_boards(60, 1):
        w = rng.uniform(-2, 2, 97).astype(np.float32)
        x = O.features(O.MODEL_V11, 0x00010002, cols, False)
        want = np_dense(np_dense(x, 10, 8, True, w), 8, 1, False, w[88:])[0]
        got = O.model_value(O.MODEL_V11, w, 0x00010002, cols, False)
        assert np.float32(got).tobytes() == want.tobytes()


def fma32(a, b, c):
    """IEEE float32 fused multiply-add by exact rational arithmetic: round(a*b + c) once, to nearest even."""
    from fractions import Fraction
    x = Fraction(float(a)) * Fraction(float(b)) + Fraction(float(c))
    if x == 0:
        return f32(0.0) if not (np.signbit(f32(a) * f32(b)) and np.signbit(f32(c))) else f32(-0.0)
    sign = -1 if x < 0 else 1
    x = abs(x)
    e = x.numerator.bit_length() - x.denominator.bit_length()
    if Fraction(2) ** e > x:
        e -= 1
    e = max(e, -126)                       # subnormal range keeps the exponent of 2^-126
    q = x / Fraction(2) ** (e - 23)        # 24 significant bits before the point
    n = q.numerator // q.denominator
    rem = q - n
    if rem > Fraction(1, 2) or (rem == Fraction(1, 2) and n % 2 == 1):
        n += 1
    return f32(sign * float(Fraction(n) * Fraction(2) ** (e - 23)))


def test_fma32_reference_rounds_once():
    # 1 + 2^-24 is a tie in float32; a*b + c computed in float64 then rounded would lose the fused rounding
    a, b, c = f32(1.0 + 2.0 ** -12), f32(1.0 + 2.0 ** -12), f32(-1.0)
    assert fma32(a, b, c) == f32(2.0 ** -11 + 2.0 ** -24)
    assert fma32(f32(3.0), f32(5.0), f32(-1.0)) == f32(14.0)


def fma_dense(x, n_in, n_out, relu, w):
    out = []
    for o in range(n_out):
        v = f32(w[n_in * n_out + o])
        for i in range(n_in):
            v = fma32(x[i], w[o * n_in + i], v)
        out.append((v if f32(0) < v else f32(0)) if relu else v)
    return out


def test_f16_value_matches_exact_fma_reference():
    """the build-defined 16-32-32-1 network: dense layout of utility.hpp:29-46 with each
    multiply-add fused (one IEEE rounding), checked against exact rational arithmetic"""
    rng = np.random.default_rng(5)
    for cols in random_boards(4, 2):
        w = rng.normal(0, 0.5, 1633).astype(np.float32)
        x = O.features(O.MODEL_F16, 0x01000001, cols, True)
        h1 = fma_dense(x, 16, 32, True, w)
        h2 = fma_dense(h1, 32, 32, True, w[544:])
        want = fma_dense(h2, 32, 1, False, w[1600:])[0]
        got = O.model_value(O.MODEL_F16, w, 0x01000001, cols, True)
        assert np.float32(got).tobytes() == want.tobytes()


def test_layer_layout_biases_after_matrix():
    """weights W[o*In+i] then the Out biases (utility.hpp:40): a zero matrix leaves bias-only outputs"""
    w = np.zeros(97, dtype=np.float32)
    w[80:88] = np.arange(1, 9, dtype=np.float32)  # layer-1 biases
    w[88:96] = 1.0                               # layer-2 weights
    w[96] = -3.0                                 # layer-2 bias
    v = O.model_value(O.MODEL_V11, w, 0, np.zeros(10, dtype=np.uint32), False)
    assert v == float(sum(range(1, 9)) - 3)


# --- stats / fitness -----------------------------------------------------------------------------
def score_from_trace(moves, boards):
    pieces, clears = 0, [0] * 5
    smax = shd = 0
    for (hold, o, x2, cleared), cols in zip(moves, boards):
        if cleared == 255:
            break
        pieces += 1
        clears[cleared] += 1
        m = brute_metrics(cols)
        smax += max(m["heights"])
        shd += m["hd"]
    if pieces == 0:
        return f32(-1e6)
    payouts = [0.0, 0.2, 0.8, 2.0, 10.0]
    bonus = 0.0
    for k in range(1, 5):
        bonus += payouts[k] * float(clears[k])
    s = bonus + 0.02 * pieces - max(smax / pieces - 12.0, 0.0) * 1.0 - (shd / pieces) * 1.5
    return f32(s)


def test_score_v1_closed_form_on_recorded_games():
    rng = np.random.default_rng(6)
    for seed in (1, 2, 3):
        w = rng.uniform(-2, 2, 97).astype(np.float32)
        moves, boards, fit = O.play_trace(O.MODEL_V11, w, seed, 150)
        assert np.float32(fit) == score_from_trace(moves, boards)
        assert fit == O.eval_host(O.MODEL_V11, w, [seed], 150)[0]


def test_zero_moves_gives_sentinel():  # model_eval_host.cpp:51 / SURVEY failure cases
    w = np.zeros(97, dtype=np.float32)
    assert list(O.eval_host(O.MODEL_V11, w, [1, 2], 0)) == [-1e6, -1e6]


def load_champion():
    with open(os.path.join(HERE, "golden", "v11_champion.json")) as f:
        return np.array(json.load(f)["weights_f64"], dtype=np.float64).astype(np.float32)


def test_v11_champion_plays_near_perfect_games():
    """models/v11_16_32_32_score_296.champ: mean score_v1 296 at the 300-move runner cap
    (ta3_trainer/runner/src/main.cpp:29-30); the theoretical max is ~306 (30 tetrises + survival)."""
    w = load_champion()
    seeds = O.expand_seeds(0x3E28DF7B1811145B, 8)
    fit, placements = O.eval_host_mt(O.MODEL_V11, w, 1, seeds, 300)
    assert (placements == 300).all()
    assert fit.mean() > 285.0
    assert (fit > 280.0).sum() >= 6


def test_batched_equals_per_model():  # parity_test.cpp:68-101 pattern, on the oracle itself
    rng = np.random.default_rng(10)
    ws = [rng.uniform(-2, 2, 97).astype(np.float32) for _ in range(3)]
    seeds = [1, 2, 3, 4]
    batched, _ = O.eval_host_mt(O.MODEL_V11, np.concatenate(ws), 3, seeds, 60, threads=3)
    for m in range(3):
        single = O.eval_host(O.MODEL_V11, ws[m], seeds, 60)
        assert batched[m * 4:(m + 1) * 4].tobytes() == single.tobytes()


def test_golden_parity_cases_reproduce():
    with open(os.path.join(HERE, "golden", "parity_cases.json")) as f:
        cases = json.load(f)["cases"]
    for c in cases:
        w = np.array(c["weights"], dtype=np.float32)
        got = O.eval_host(c["kind"], w, c["seeds"], c["max_moves"])
        assert got.tobytes() == np.array(c["fitness"], dtype=np.float32).tobytes()
