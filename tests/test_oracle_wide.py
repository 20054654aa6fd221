"""Pins the oracle's wide value network (32 features -> 128 -> 128 -> 1) and the garbage-row start positions.

- The 16 extra features against brute-force cell-grid definitions (numpy float32 division), bit-exactly.
- The device-arithmetic value (bfloat16 inputs / weights / first hidden layer, exact accumulation, FP32
  output layer in a fixed order) against an independent numpy restatement, and against the plain FP32
  network within the bfloat16 rounding error.
- Garbage rows: 0-4 full rows with one hole each at the bottom, deterministic per seed; the piece sequence is
  the one without garbage.
"""
import numpy as np

import oracle as O
from test_oracle_model import brute_features_f16, brute_metrics, random_boards

ROWS = 23
f32 = np.float32


def brute_features_f32w(clears, cols, held):
    base = brute_features_f16(clears, cols, held)
    m = brute_metrics(cols)
    g = [[(int(cols[x]) >> y) & 1 for y in range(ROWS)] for x in range(10)]
    h = m["heights"]
    hole_rows, hole_cols = set(), 0
    for x in range(10):
        top = ROWS - h[x]
        holes = [y for y in range(top, ROWS) if not g[x][y]]
        hole_rows.update(holes)
        hole_cols += 1 if holes else 0
    cells = sum(sum(c) for c in g)
    row_tr = 0
    for y in range(ROWS):
        line = [1] + [g[x][y] for x in range(10)] + [1]  # walls filled
        row_tr += sum(1 for x in range(11) if line[x] != line[x + 1])
    deepest = 0
    for x in range(10):
        nb = [h[x - 1]] if x == 9 else ([h[x + 1]] if x == 0 else [h[x - 1], h[x + 1]])
        deepest = max(deepest, min(nb) - h[x])
    extra = [f32(v) / f32(20) for v in h]
    extra += [f32(len(hole_rows)) / f32(20), f32(row_tr) / f32(100), f32(deepest) / f32(20), f32(cells) / f32(200),
              f32(hole_cols) / f32(5), f32(min(h)) / f32(20)]
    return np.concatenate([base, np.array(extra, dtype=np.float32)])


def test_f32w_features_match_brute_force():
    rng = np.random.default_rng(4)
    for cols in random_boards(150, 11):
        clears = int(rng.integers(0, 5)) | int(rng.integers(0, 3)) << 8 | int(rng.integers(0, 4)) << 24
        held = bool(rng.integers(0, 2))
        got = O.features(O.MODEL_F32W, clears, cols, held)
        assert got.tobytes() == brute_features_f32w(clears, cols, held).tobytes()


def bf16(a):
    u = np.asarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def np_device_value(w, x):
    """bfloat16 operands, exact (float64) sums rounded once, output layer as four FMA chains"""
    W1 = bf16(w[:4096]).reshape(128, 32).astype(np.float64)
    b1 = w[4096:4224].astype(np.float64)
    W2 = bf16(w[4224:4224 + 16384]).reshape(128, 128).astype(np.float64)
    b2 = w[4224 + 16384:4224 + 16512].astype(np.float64)
    W3 = w[4224 + 16512:4224 + 16640]
    b3 = w[-1]
    pre1 = (b1 + W1 @ bf16(x).astype(np.float64)).astype(np.float32)
    h1 = bf16(np.maximum(pre1, f32(0))).astype(np.float64)
    h2 = np.maximum((b2 + W2 @ h1).astype(np.float32), f32(0))
    parts = []
    for q in range(4):
        p = 0.0
        for t in range(8):
            for j in range(4):
                o = 16 * t + 4 * q + j
                p = float(f32(np.float64(h2[o]) * np.float64(W3[o]) + p))  # exact product, one rounding
        parts.append(f32(p))
    return (parts[0] + parts[1]) + (parts[2] + parts[3]) + b3


def test_f32w_device_value_matches_numpy():
    rng = np.random.default_rng(5)
    w = rng.normal(0, 0.3, 20865).astype(np.float32)
    for cols in random_boards(40, 13):
        clears = int(rng.integers(0, 4)) | int(rng.integers(0, 2)) << 24
        x = O.features(O.MODEL_F32W, clears, cols, False)
        got = O.model_value(O.MODEL_F32W, w, clears, cols, False)
        want = np_device_value(w, x)
        assert abs(got - want) <= 1e-5 * max(1.0, abs(want)), (got, want)


def test_f32w_device_value_tracks_fp32_network():
    """bfloat16 rounding (8 significant bits) of inputs, weights and the first hidden layer moves the value by
    at most 5% of its spread over boards, and keeps the boards' order"""
    rng = np.random.default_rng(6)
    w = rng.normal(0, 0.2, 20865).astype(np.float32)
    dev, ref = [], []
    for cols in random_boards(60, 17):
        dev.append(O.model_value(O.MODEL_F32W, w, 0, cols, False))
        ref.append(O.f32w_value_fp32(w, 0, cols, False))
    dev, ref = np.array(dev), np.array(ref)
    assert np.abs(dev - ref).max() <= 0.05 * ref.std() + 1e-3
    assert np.corrcoef(dev, ref)[0, 1] > 0.999


def garbage_rows(seed):
    """restatement of the garbage-row draw: xoshiro256** on seed ^ 0x6A09E667F3BCC908"""
    st = np.zeros(4, dtype=np.uint64)
    O.lib().ta3o_rng_seed(seed ^ 0x6A09E667F3BCC908, st.ctypes.data_as(O.C.POINTER(O.C.c_uint64)))
    nxt = lambda: O.lib().ta3o_rng_next(st.ctypes.data_as(O.C.POINTER(O.C.c_uint64)))
    n = nxt() % 5
    return [nxt() % 10 for _ in range(n)]


def bf16_bits(v):
    u = np.asarray(v, dtype=np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)


def test_reciprocal_multiply_equals_quotient_after_bf16():
    """the device normalises the wide network's integer features (0..2,600: heights, holes, hole depths, wells,
    transitions, cells) by one multiply with the float32 reciprocal instead of the IEEE quotient the oracle takes;
    the two differ in the last float32 bit for about a fifth of the numerators, and never after rounding to
    bfloat16, the precision the network reads them at"""
    x = np.arange(0, 2601, dtype=f32)
    for c in (5, 20, 40, 100, 200):
        quotient = (x / f32(c)).astype(f32)
        product = (x * f32(1.0 / c)).astype(f32)
        assert (quotient != product).sum() > 0, c # the shortcut is not exact in float32 ...
        assert (bf16_bits(quotient) == bf16_bits(product)).all(), c # ... but is after the rounding


def test_garbage_rows_start_positions():
    counts = set()
    for seed in O.expand_seeds(5, 40):
        seed = int(seed)
        plain = O.engine(seed)
        with O.garbage():
            e = O.engine(seed)
        assert list(e.queue) == list(plain.queue) and list(e.bag) == list(plain.bag)
        holes = garbage_rows(seed)
        counts.add(len(holes))
        cols = [int(c) for c in e.cols]
        for y in range(ROWS):
            row = [(cols[x] >> y) & 1 for x in range(10)]
            k = ROWS - 1 - y
            if k < len(holes):
                assert row == [0 if x == holes[k] else 1 for x in range(10)]
            else:
                assert sum(row) == 0
    assert counts == {0, 1, 2, 3, 4}


def test_garbage_games_differ_and_are_deterministic():
    w = np.random.default_rng(8).normal(0, 0.5, 97).astype(np.float32)
    seeds = O.expand_seeds(9, 6)
    plain, _ = O.eval_host_mt(O.MODEL_V11, w, 1, seeds, 80)
    with O.garbage():
        g1, _ = O.eval_host_mt(O.MODEL_V11, w, 1, seeds, 80)
        g2, _ = O.eval_host_mt(O.MODEL_V11, w, 1, seeds, 80)
    assert g1.tobytes() == g2.tobytes()
    assert g1.tobytes() != plain.tobytes()
