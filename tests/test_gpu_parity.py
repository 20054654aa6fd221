"""GPU parity: the HIP evaluator (through the C-ABI) against the CPU oracle and the golden fixtures.

The bar is bit-exactness: fitness values (float32 of the FP64 score), placement counts and therefore
every move of every game must equal the oracle's (the reference's own acceptance test demands
device == host exactly, ta3_gpu/tests/src/parity_test.cpp:51-101).
"""
import json
import os
import sys

import numpy as np
import pytest

import oracle as O
import tetris_ai3_amd as T

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ev():
    e = T.ModelEvaluator()
    yield e
    e.close()


def test_device_bits_canary():
    say("case: bits canary")
    assert T.device_bits_canary()


def test_feature_norms_match_host_ieee():
    """every normalised feature value the search can produce equals host IEEE arithmetic (numpy float32:
    correctly rounded sqrt and division), over each feature's whole integer domain"""
    say("case: feature norms 0..4761")
    n = 4762
    got = T.feature_norms(n)
    i = np.arange(n, dtype=np.float32)
    sq = np.where(i > 0, np.sqrt(i), np.float32(0)).astype(np.float32) / np.float32(100)
    domains = {0: (sq, n), 1: (i / np.float32(20), 231), 2: (i / np.float32(200), 1321), 3: (i / np.float32(5), 11),
               4: (i / np.float32(40), 241), 5: (i / np.float32(100), n)}
    for col, (want, dom) in domains.items():
        bad = np.nonzero(got[:dom, col].view(np.uint32) != want[:dom].astype(np.float32).view(np.uint32))[0]
        assert bad.size == 0, (col, bad[:10], got[bad[:3], col], want[bad[:3]])


def test_golden_parity_cases(ev):
    with open(os.path.join(HERE, "golden", "parity_cases.json")) as f:
        cases = json.load(f)["cases"]
    for ci, c in enumerate(cases):
        say(f"case: golden {ci} kind={c['kind']} seeds={c['seeds']} cap={c['max_moves']}")
        e = T.ModelEvaluator(model=c["kind"])
        w = np.array(c["weights"], dtype=np.float32)
        got = e.eval(w, 1, c["seeds"], c["max_moves"])
        assert got.tobytes() == np.array(c["fitness"], dtype=np.float32).tobytes(), (ci, got, c["fitness"])
        e.close()


@pytest.mark.parametrize("kind", [T.MODEL_V11, T.MODEL_F16])
def test_single_model_matches_oracle(kind):
    """parity_test.cpp:49-66: three weight vectors x seeds {1,2,3,4} x 200 moves, device == host exactly"""
    e = T.ModelEvaluator(model=kind)
    P = T.num_params(kind)
    seeds = [1, 2, 3, 4]
    for ms in (11, 22, 33):
        say(f"case: single kind={kind} modelSeed={ms}")
        rng = np.random.default_rng(ms)
        w = (rng.uniform(-2, 2, P) if kind == T.MODEL_V11 else rng.normal(0, 0.5, P)).astype(np.float32)
        dev = e.eval(w, seeds, 200)
        host = O.eval_host(kind, w, seeds, 200)
        assert dev.tobytes() == host.tobytes(), (ms, dev, host)
    e.close()


@pytest.mark.parametrize("kind", [T.MODEL_V11, T.MODEL_F16])
def test_zero_skipping_matches_oracle(kind):
    """The value networks skip the steps whose input is zero in every active lane of a wave (exact for finite
    weights); a model with a non-finite weight takes the plain path. One batch against the oracle: a random
    model; one with -0.0 biases (the signed-zero corner the skipping must not change); one with half of its
    first-layer units dead (negative bias, zero weights: most units skip); one whose output bias is +inf
    (plain path; every leaf value is +inf, so ties fall to the lowest slot)."""
    P = T.num_params(kind)
    b1, n1, b3 = (80, 8, 96) if kind == T.MODEL_V11 else (512, 32, 1632)
    fan_in = 10 if kind == T.MODEL_V11 else 16
    rng = np.random.default_rng(71)
    ws = rng.normal(0, 0.5, (4, P)).astype(np.float32)
    ws[1, b1:b1 + n1] = -0.0
    if kind == T.MODEL_F16:
        ws[1, 1568:1600] = -0.0
    ws[1, b3] = -0.0
    dead = np.arange(0, n1, 2)
    for o in dead:
        ws[2, o * fan_in:(o + 1) * fan_in] = 0.0
        ws[2, b1 + o] = -1.0
    ws[3, b3] = np.inf
    seeds = O.expand_seeds(72, 4)
    say(f"case: zero skipping kind={kind}")
    e = T.ModelEvaluator(model=kind)
    e.stage(ws.reshape(-1), 4, seeds, 150)
    e.launch()
    got, pl = e.fetch()
    e.close()
    want, want_pl = O.eval_host_mt(kind, ws.reshape(-1), 4, seeds, 150)
    assert got.tobytes() == want.tobytes() and (pl == want_pl).all(), (got, want)


def test_f16_mfma_rounding_corners():
    """The 4-wave 16-32-32-1 kernel runs the hidden layers of finite-weight models on v_mfma_f32_16x16x4_f32, whose
    four products per step must be added in k order and rounded like the oracle's fmaf chain. One batch, bit-exact
    against the oracle, of models built to expose a different order or a flush to zero:
    - layer-1 weights, biases and layer-2 biases at subnormal scale (both layers' sums below FLT_MIN; output
      weights near 1e30 bring the differences back to normal range);
    - layer-2 weights and biases at subnormal scale, likewise;
    - weights near 1e15 (hidden values near 1e31, no overflow);
    - alternating +-1e8 and +-1 weights (heavy cancellation: any other summation order rounds differently);
    - a random model."""
    kind = T.MODEL_F16
    P = T.num_params(kind)
    rng = np.random.default_rng(73)
    ws = rng.normal(0, 0.5, (5, P)).astype(np.float32)
    ws[0, 0:544] *= np.float32(1e-39)       # W1, b1
    ws[0, 1568:1600] *= np.float32(1e-39)   # b2
    ws[1, 544:1600] *= np.float32(1e-40)    # W2, b2
    ws[0:2, 1600:1633] *= np.float32(1e30)  # W3, b3
    ws[2, 0:1600] *= np.float32(1e15)       # W1, b1, W2, b2
    big = np.where(np.arange(1568 - 544) % 2 == 0, np.float32(1e8), np.float32(1.0))
    ws[3, 544:1568] *= big                  # W2: alternating scales
    ws[3, 0:512] *= np.where(np.arange(512) % 3 == 0, np.float32(1e8), np.float32(1.0))
    assert np.isfinite(ws).all()
    assert (np.abs(ws[0, 0:544]) < np.finfo(np.float32).tiny).mean() > 0.9  # mostly subnormal
    seeds = O.expand_seeds(74, 4)
    say("case: f16 MFMA rounding corners, 5 models x 4 games, cap 150")
    e = T.ModelEvaluator(model=kind)
    e.set_launch(waves_per_game=4)
    e.stage(ws.reshape(-1), 5, seeds, 150)
    e.launch()
    got, pl = e.fetch()
    e.close()
    want, want_pl = O.eval_host_mt(kind, ws.reshape(-1), 5, seeds, 150)
    assert got.tobytes() == want.tobytes() and (pl == want_pl).all(), (got, want)


def test_work_counters():
    """Per-item counters: candidates fitted and legal equal the oracle's less its repeated lists (an earlier list
    of the level with the same parent board, clear histogram, piece and held-is-I), which the device scores once,
    and less the roots the previous move scored (the committed placement cleared no line) and, for this network,
    the lists of levels 1 and 2 the device takes from the previous move's scores; candidate lanes issued
    are whole
    64-lane chunks covering every fitted candidate of a plain beam, with less than one chunk of padding per
    stage (four stages per move)"""
    kind, nm = T.MODEL_F16, 3
    w = np.random.default_rng(81).normal(0, 0.5, nm * T.num_params(kind)).astype(np.float32)
    seeds = O.expand_seeds(82, 4)
    say("case: work counters")
    e = T.ModelEvaluator(model=kind)
    e.stage(w, nm, seeds, 100)
    e.launch()
    _, pl = e.fetch()
    cnt = e.fetch_counters().astype(np.int64)
    e.close()
    P = T.num_params(kind)
    for m in range(nm):
        c = O.Counters()
        O.eval_host(kind, w[m * P:(m + 1) * P], seeds, 100, c)
        assert c.repeat_legal > 0 and c.reuse_legal > 0
        assert int(cnt[m * 4:(m + 1) * 4, 0].sum()) == c.applies - c.repeat_applies - c.reuse_applies
        assert int(cnt[m * 4:(m + 1) * 4, 1].sum()) == c.legal - c.repeat_legal - c.reuse_legal
    applies, legal, lanes = cnt[:, 0], cnt[:, 1], cnt[:, 2]
    assert (lanes % 64 == 0).all() and (legal <= applies).all() and (applies <= lanes).all()
    assert (lanes - applies < 4 * 64 * (pl.astype(np.int64) + 1)).all()


def test_multi_model_batch_matches_single_and_host(ev):
    """parity_test.cpp:68-101"""
    rng = np.random.default_rng(100)
    ws = [rng.uniform(-2, 2, 97).astype(np.float32) for _ in range(3)]
    seeds = [1, 2, 3, 4]
    say("case: batched 3 models x 4 seeds x 200")
    batched = ev.eval(np.concatenate(ws), 3, seeds, 200)
    for m in range(3):
        single = ev.eval(ws[m], seeds, 200)
        assert batched[m * 4:(m + 1) * 4].tobytes() == single.tobytes()
        host = O.eval_host(O.MODEL_V11, ws[m], seeds, 200)
        assert single.tobytes() == host.tobytes()


def test_champion_games_and_placements(ev):
    with open(os.path.join(HERE, "golden", "v11_champion.json")) as f:
        w = np.array(json.load(f)["weights_f64"], dtype=np.float64).astype(np.float32)
    seeds = O.expand_seeds(0x3E28DF7B1811145B, 8)
    say("case: v11 champion 8 games x 300 moves")
    ev.stage(w, 1, seeds, 300)
    ev.launch()
    ev.sync()
    fit, pl = ev.fetch()
    host, host_pl = O.eval_host_mt(O.MODEL_V11, w, 1, seeds, 300)
    assert fit.tobytes() == host.tobytes()
    assert (pl == host_pl).all()
    assert fit.mean() > 285.0


def test_edge_cases(ev):
    w = np.random.default_rng(1).uniform(-2, 2, 97).astype(np.float32)
    say("case: max_moves=0")
    assert list(ev.eval(w, [1, 2, 3], 0)) == [-1e6] * 3
    say("case: max_moves=1")
    assert ev.eval(w, [5, 6], 1).tobytes() == O.eval_host(O.MODEL_V11, w, [5, 6], 1).tobytes()
    say("case: one game, many models")
    ws = np.random.default_rng(2).uniform(-2, 2, (70, 97)).astype(np.float32)
    got = ev.eval(ws.reshape(-1), 70, [9], 40)
    want, _ = O.eval_host_mt(O.MODEL_V11, ws.reshape(-1), 70, [9], 40)
    assert got.tobytes() == want.tobytes()


@pytest.mark.parametrize("wpb,persistent,grid", [(1, False, 0), (2, True, 0), (4, True, 3), (4, False, 0), (1, True, 1),
                                                 (2, False, 0), (4, True, 0), (4, True, 7)])
def test_launch_knobs_do_not_change_results(wpb, persistent, grid):
    rng = np.random.default_rng(7)
    ws = rng.normal(0, 0.5, (6, 97)).astype(np.float32)
    seeds = O.expand_seeds(3, 5)
    e = T.ModelEvaluator()
    e.set_launch(wpb, persistent, grid)
    say(f"case: launch wpb={wpb} persistent={persistent} grid={grid}")
    got = e.eval(ws.reshape(-1), 6, seeds, 120)
    want, _ = O.eval_host_mt(O.MODEL_V11, ws.reshape(-1), 6, seeds, 120)
    assert got.tobytes() == want.tobytes()
    e.close()


def test_config0_cpu_runnable_case_bit_exact():
    """4 models x 8 games, cap 1000, weights N(0, 0.5^2), seeds from simulate(seed=1): both networks"""
    for kind in (T.MODEL_V11, T.MODEL_F16):
        P = T.num_params(kind)
        rng = np.random.Generator(np.random.PCG64(0x7E7215))
        w = rng.normal(0.0, 0.5, 4 * P).astype(np.float32)
        seeds = T.expand_seeds(1, 8)
        e = T.ModelEvaluator(model=kind)
        say(f"case: config0 kind={kind}")
        e.stage(w, 4, seeds, 1000)
        e.launch()
        e.sync()
        fit, pl = e.fetch()
        host, host_pl = O.eval_host_mt(kind, w, 4, seeds, 1000)
        assert fit.tobytes() == host.tobytes()
        assert (pl == host_pl).all()
        e.close()


def test_cpp_dropin_header_caller(tmp_path):
    """a C++ caller of include/ta3h/model_eval.hpp (the reference's class shape) passes the reference's
    parity_test.cpp acceptance pattern against the oracle's fitness"""
    import subprocess
    exe = os.path.join(HERE, "cpp", "model_eval_test")
    assert os.path.exists(exe), "build() compiles tests/cpp/model_eval_test"
    rng = np.random.default_rng(100)
    ws = np.concatenate([rng.uniform(-2, 2, 97).astype(np.float32) for _ in range(3)])
    want, _ = O.eval_host_mt(O.MODEL_V11, ws, 3, [1, 2, 3, 4], 200)
    wf, ef = tmp_path / "w.f32", tmp_path / "e.f32"
    ws.tofile(wf)
    want.astype(np.float32).tofile(ef)
    say("case: C++ drop-in header caller")
    r = subprocess.run([exe, str(wf), str(ef)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "ok", (r.returncode, r.stdout, r.stderr)
